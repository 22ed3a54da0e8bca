// C++ host side of the solver: the reference's Problem / g_ACO structure
// (cvrp_ocl/problem.h:82-140, cvrp_ocl/gpu/g_aco.h:16-37) on top of the C ABI
// of libcvrpaco.so.  A caller of the reference swaps
//     OpenclEnv env(*inst); g_ACO aco(env, *inst);
// for
//     GpuACO aco(*inst);
// and keeps init_aco / run_aco_iteration / exit_aco.
#pragma once

#include <string>
#include <vector>

#include "cvrp_aco.h"

namespace cvrp {

// Instance plus run parameters (problem.h:82-140; set_default_parameters,
// problem.cpp:186-220, with real values instead of compile-time macros).
struct Problem {
    std::string name;
    int num_node = 0;
    std::vector<double> x, y;
    std::vector<int> demand;
    int vehicle_capacity = 0;
    float max_distance = 0.0f;
    float service_time = 0.0f;
    float optimum = 0.0f;
    std::vector<float> distance;  // num_node * num_node, row-major

    int iteration = 0;
    int max_iteration = 250;
    double max_runtime = 600.0;

    float alpha = 1.0f, beta = 2.0f, rho = 0.1f;
    int n_ants = 0;  // 0: num_node, as the reference
    int ras_ranks = 6;
    int k = 0, nn_ls = 0;
    int selection = ACO_SEL_IROULETTE, rng = ACO_RNG_PARK_MILLER;
    int ls_flags = 0;
    int tau0_rule = ACO_TAU0_FIRST_COLONY;
    unsigned long long seed = 1;
    bool use_graph = true;

    // sub-problem decomposition (problem.h:136-139, problem.cpp:216-218):
    // every master_iterations iterations a round over num_subs sub-problems
    // (0: num_node/50), sub_iterations iterations of sub_ants ants each
    bool decompose = false;
    int num_subs = 0;
    int master_iterations = 1;
    int sub_iterations = 75;
    int sub_ants = 64;
    bool pheromone_back = false;

    std::vector<int> best_tour;  // best_so_far_ant
    float best_length = 0.0f;
    float best_time = 0.0f;
};

// read_instance_file (io.cpp:334-517) and compute_distances
// (vrpHelper.cpp:160-184).  dist_mode: ACO_DIST_*, or -1 for the file's
// EDGE_WEIGHT_TYPE with EUC_2D as the reference's unrounded float.
void read_instance_file(Problem &instance, const char *path, int dist_mode);
// check_solution (problem.cpp:226-319)
bool check_solution(const Problem &instance, const std::vector<int> &tour);

class GpuACO {
public:
    explicit GpuACO(Problem &instance);
    ~GpuACO();
    GpuACO(const GpuACO &) = delete;
    GpuACO &operator=(const GpuACO &) = delete;

    void init_aco();
    void exit_aco();
    void run_aco_iteration();

    void construct_solutions();
    void local_search();
    void update_statistics();
    void pheromone_update();
    void pheromone_init(float trail_0);

    void update_best_so_far_from_device();

    // One decomposition round (cvrp_decompose.h, aco_decompose_round) on the
    // best-so-far; returns the merged tour's length, the tour in `merged`.
    float decompose_round(int round, std::vector<int> &merged);
    // run_aco_iteration with `ant` put in as the last ant after local search
    void run_aco_iteration_with(const std::vector<int> &ant);
    const std::vector<aco_bsf_record> &records() const { return records_; }
    aco_counters counters();

private:
    Problem &instance;
    aco_ctx *ctx = nullptr;
    std::vector<aco_bsf_record> records_;
};

}  // namespace cvrp
