// GpuACO: the reference's g_ACO host driver (cvrp_ocl/gpu/g_aco.cpp) as a
// thin C++ layer over the C ABI.  Errors print the library message and exit
// with status 1, as the reference's check_error does (g_aco.cpp:607-614).
#include "g_aco.h"

#include "cvrp_decompose.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

namespace cvrp {

static void check(int rc, const char *what)
{
    if (rc != ACO_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, aco_last_error());
        exit(1);
    }
}

void read_instance_file(Problem &inst, const char *path, int dist_mode)
{
    int32_t n = 0;
    check(aco_read_vrplib(path, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr),
          "read_instance_file");
    inst.num_node = n;
    inst.x.resize(n);
    inst.y.resize(n);
    inst.demand.resize(n);
    int32_t Q = 0, ew = 0;
    float D = 0, svc = 0, opt = 0;
    check(aco_read_vrplib(path, n, &n, inst.x.data(), inst.y.data(), inst.demand.data(), &Q, &D,
                          &svc, &opt, &ew),
          "read_instance_file");
    inst.vehicle_capacity = Q;
    inst.max_distance = D;
    inst.service_time = svc;
    inst.optimum = opt;
    std::string p(path);
    const size_t slash = p.find_last_of('/');
    inst.name = slash == std::string::npos ? p : p.substr(slash + 1);
    if (dist_mode < 0) {
        static const int by_type[4] = {ACO_DIST_FLOAT, ACO_DIST_CEIL, ACO_DIST_GEO, ACO_DIST_ATT};
        dist_mode = by_type[ew & 3];
    }
    inst.distance.resize((size_t)n * n);
    check(aco_compute_distances(n, inst.x.data(), inst.y.data(), dist_mode, inst.distance.data()),
          "compute_distances");
}

bool check_solution(const Problem &inst, const std::vector<int> &tour)
{
    const int rc = aco_check_solution(inst.num_node, inst.distance.data(), inst.demand.data(),
                                      inst.vehicle_capacity, inst.max_distance, inst.service_time,
                                      tour.data(), (int32_t)tour.size());
    if (rc < 0) check(rc, "check_solution");
    return rc == 1;
}

// g_ACO::g_ACO + create_memory_objects (g_aco.cpp:35-43, 640-761)
GpuACO::GpuACO(Problem &inst) : instance(inst)
{
    aco_problem pb{};
    pb.n = inst.num_node;
    pb.dist = inst.distance.data();
    pb.demand = inst.demand.data();
    pb.capacity = inst.vehicle_capacity;
    pb.max_dist = inst.max_distance;
    pb.service_time = inst.service_time;
    aco_params prm;
    aco_default_params(&prm);
    prm.alpha = inst.alpha;
    prm.beta = inst.beta;
    prm.rho = inst.rho;
    prm.n_ants = inst.n_ants;
    prm.ras_ranks = inst.ras_ranks;
    prm.k = inst.k;
    prm.nn_ls = inst.nn_ls;
    prm.selection = inst.selection;
    prm.rng = inst.rng;
    prm.ls_flags = inst.ls_flags;
    prm.tau0_rule = inst.tau0_rule;
    prm.seed = inst.seed;
    prm.use_graph = inst.use_graph ? 1 : 0;
    check(aco_create(&pb, &prm, &ctx), "aco_create");
}

GpuACO::~GpuACO() { aco_destroy(ctx); }

// g_aco.cpp:63-93
void GpuACO::init_aco()
{
    instance.iteration = 0;
    instance.best_length = INFINITY;
    check(aco_init(ctx), "init_aco");
    int32_t it = 0;
    check(aco_get_iteration(ctx, &it), "init_aco");
    instance.iteration = it;
}

// g_aco.cpp:95-117: best-so-far ant and its record trajectory
void GpuACO::exit_aco()
{
    update_best_so_far_from_device();
    int32_t count = 0;
    check(aco_get_bsf(ctx, nullptr, 0, &count), "exit_aco");
    records_.resize(count);
    check(aco_get_bsf(ctx, records_.data(), count, &count), "exit_aco");
    if (count > 0) instance.best_time = records_.back().time;
}

// g_aco.cpp:122-134
void GpuACO::run_aco_iteration() { check(aco_iteration(ctx), "run_aco_iteration"); }

void GpuACO::construct_solutions() { check(aco_construct(ctx), "construct_solutions"); }
void GpuACO::local_search() { check(aco_local_search(ctx), "local_search"); }
void GpuACO::update_statistics() { check(aco_update_statistics(ctx), "update_statistics"); }
void GpuACO::pheromone_update() { check(aco_pheromone_update(ctx), "pheromone_update"); }
void GpuACO::pheromone_init(float trail_0) { check(aco_pheromone_init(ctx, trail_0), "pheromone_init"); }

// g_aco.cpp:527-552
void GpuACO::update_best_so_far_from_device()
{
    std::vector<int> tour(2 * (size_t)instance.num_node);
    int32_t size = 0;
    float len = 0.0f;
    check(aco_get_best(ctx, tour.data(), (int32_t)tour.size(), &size, &len),
          "update_best_so_far_from_device");
    tour.resize(size);
    instance.best_tour = tour;
    instance.best_length = len;
}

float GpuACO::decompose_round(int round, std::vector<int> &merged)
{
    const Problem &in = instance;
    aco_problem pb{};
    pb.n = in.num_node;
    pb.dist = in.distance.data();
    pb.demand = in.demand.data();
    pb.capacity = in.vehicle_capacity;
    pb.max_dist = in.max_distance;
    pb.service_time = in.service_time;
    std::vector<float> x(in.x.begin(), in.x.end()), y(in.y.begin(), in.y.end());   // Point floats
    aco_decompose_params dp{};
    dp.subs = in.num_subs;
    dp.sub_iterations = in.sub_iterations;
    dp.rotation = round;
    dp.pheromone_back = in.pheromone_back ? 1 : 0;
    aco_default_params(&dp.sub_params);
    aco_params &sp = dp.sub_params;
    sp.alpha = in.alpha;
    sp.beta = in.beta;
    sp.rho = in.rho;
    sp.n_ants = in.sub_ants;
    sp.ras_ranks = in.ras_ranks;
    sp.selection = ACO_SEL_KNN;
    sp.rng = ACO_RNG_PHILOX;
    sp.k = in.k > 0 ? in.k : 16;
    sp.ls_flags = ACO_LS_2OPT | ACO_LS_SWAP | ACO_LS_RELOCATE;
    sp.seed = in.seed + 1000003ull * (unsigned long long)(round + 1);
    merged.assign(2 * (size_t)in.num_node + 2, 0);
    int32_t size = 0, subs = 0, replaced = 0;
    float len = 0.0f;
    check(aco_decompose_round(ctx, 0, &pb, x.data(), y.data(), &dp, merged.data(), (int32_t)merged.size(), &size,
                              &len, &subs, &replaced),
          "decompose_round");
    merged.resize(size);
    return len;
}

void GpuACO::run_aco_iteration_with(const std::vector<int> &ant)
{
    construct_solutions();
    if (instance.ls_flags) local_search();
    const int m = instance.n_ants > 0 ? instance.n_ants : instance.num_node;
    check(aco_set_ant_at(ctx, 0, m - 1, ant.data(), (int32_t)ant.size()), "run_aco_iteration_with");
    update_statistics();
    pheromone_update();
}

aco_counters GpuACO::counters()
{
    aco_counters c{};
    check(aco_get_counters(ctx, &c), "counters");
    return c;
}

}  // namespace cvrp
