// cvrp_aco: command-line driver with the reference's trial structure
// (cvrp_ocl/main.cpp:61-94): for each try, load the instance, init, iterate
// until the termination condition, report the best-so-far ant.  The
// reference's compile-time parameters (kernel.cl:9-12, problem.cpp:186-220)
// are flags here, with the same defaults.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>

#include "g_aco.h"

using namespace cvrp;

static void usage(const char *prog)
{
    fprintf(stderr,
            "usage: %s [options] instance.vrp\n"
            "  --alpha A  --beta B  --rho R      pheromone / heuristic weights (1, 2, 0.1)\n"
            "  --ants M                          ants per iteration (default: nodes)\n"
            "  --iterations I                    iteration limit (250)\n"
            "  --time T                          time limit in seconds (600)\n"
            "  --tries N                         independent trials (15)\n"
            "  --ras R                           ranked ants incl. best-so-far (6)\n"
            "  --selection iroulette|knn --k K   next-node rule and candidate width\n"
            "  --rng pm|philox                   random numbers\n"
            "  --ls none|2opt|2opt-random|2opt+swap|full|full-random\n"
            "                                    local search (none); -random: 2-opt in the CPU\n"
            "                                    variant's random scan order\n"
            "  --nn-ls D                         2-opt neighbour depth (nodes-2)\n"
            "  --tau0 first|nn                   initial trail rule (first)\n"
            "  --dist float|nint                 EUC_2D distances (float, as the reference)\n"
            "  --decompose                       sub-problem rounds on the best-so-far\n"
            "  --subs N  --sub-iterations I      sub-problems (nodes/50) and their iterations (75)\n"
            "  --sub-ants M  --master-iterations P  ants per sub-colony (64), iterations per round (1)\n"
            "  --pheromone-back                  sub-colony trails written into the master\n"
            "  --seed S  --no-graph  --report DIR\n"
            "  --sol DIR                         best tour of all tries as DIR/<name>.sol\n",
            prog);
    exit(1);
}

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// The tour in the .sol layout of the reference's dataset (what the dataset's
// .sol files hold): total cost, route count, two solver-metadata lines (0:
// not recorded), then per route "0 1 1 cost load stops 0 c1 .. ck 0", costs
// the route's distances summed in order, 6 significant digits.
static bool write_sol(const Problem &inst, const std::string &file)
{
    const std::vector<int> &t = inst.best_tour;
    const int n = inst.num_node;
    std::vector<std::vector<int>> routes;
    std::vector<int> cur{0};
    for (size_t q = 1; q < t.size(); ++q) {
        cur.push_back(t[q]);
        if (t[q] == 0) {
            if (cur.size() > 2) routes.push_back(cur);
            cur.assign(1, 0);
        }
    }
    std::vector<double> cost(routes.size());
    double total = 0.0;
    for (size_t r = 0; r < routes.size(); ++r) {
        for (size_t q = 0; q + 1 < routes[r].size(); ++q)
            cost[r] += inst.distance[(size_t)routes[r][q] * n + routes[r][q + 1]];
        total += cost[r];
    }
    FILE *fp = fopen(file.c_str(), "w");
    if (!fp) return false;
    fprintf(fp, "%.6g\n%zu\n0\n0\n", total, routes.size());
    for (size_t r = 0; r < routes.size(); ++r) {
        int load = 0;
        for (int v : routes[r]) load += inst.demand[v];
        fprintf(fp, " 0 1 1 %.6g %d  %zu", cost[r], load, routes[r].size());
        for (int v : routes[r]) fprintf(fp, " %d", v);
        fprintf(fp, "\n");
    }
    return fclose(fp) == 0;
}

int main(int argc, char **argv)
{
    Problem defaults;
    int tries = 15, dist_mode = -1;
    const char *path = nullptr, *report_dir = nullptr, *sol_dir = nullptr;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        auto val = [&]() -> const char * {
            if (a + 1 >= argc) usage(argv[0]);
            return argv[++a];
        };
        if (s == "--alpha") defaults.alpha = (float)atof(val());
        else if (s == "--beta") defaults.beta = (float)atof(val());
        else if (s == "--rho") defaults.rho = (float)atof(val());
        else if (s == "--ants") defaults.n_ants = atoi(val());
        else if (s == "--iterations") defaults.max_iteration = atoi(val());
        else if (s == "--time") defaults.max_runtime = atof(val());
        else if (s == "--tries") tries = atoi(val());
        else if (s == "--ras") defaults.ras_ranks = atoi(val());
        else if (s == "--k") defaults.k = atoi(val());
        else if (s == "--nn-ls") defaults.nn_ls = atoi(val());
        else if (s == "--seed") defaults.seed = strtoull(val(), nullptr, 10);
        else if (s == "--no-graph") defaults.use_graph = false;
        else if (s == "--decompose") defaults.decompose = true;
        else if (s == "--subs") defaults.num_subs = atoi(val());
        else if (s == "--sub-iterations") defaults.sub_iterations = atoi(val());
        else if (s == "--sub-ants") defaults.sub_ants = atoi(val());
        else if (s == "--master-iterations") defaults.master_iterations = atoi(val());
        else if (s == "--pheromone-back") defaults.pheromone_back = true;
        else if (s == "--report") report_dir = val();
        else if (s == "--sol") sol_dir = val();
        else if (s == "--selection") {
            const std::string v = val();
            defaults.selection = v == "knn" ? ACO_SEL_KNN : ACO_SEL_IROULETTE;
            if (v == "knn" && defaults.rng == ACO_RNG_PARK_MILLER) defaults.rng = ACO_RNG_PHILOX;
        } else if (s == "--rng") {
            defaults.rng = std::string(val()) == "philox" ? ACO_RNG_PHILOX : ACO_RNG_PARK_MILLER;
        } else if (s == "--ls") {
            const std::string v = val();
            defaults.ls_flags = v == "none" ? 0 : v == "2opt" ? ACO_LS_2OPT
                              : v == "2opt-random" ? ACO_LS_2OPT_RANDOM
                              : v == "2opt+swap" ? (ACO_LS_2OPT | ACO_LS_SWAP)
                              : v == "full" ? (ACO_LS_2OPT | ACO_LS_SWAP | ACO_LS_RELOCATE)
                              : v == "full-random" ? (ACO_LS_2OPT_RANDOM | ACO_LS_SWAP | ACO_LS_RELOCATE) : -1;
            if (defaults.ls_flags < 0) usage(argv[0]);
        } else if (s == "--tau0") {
            defaults.tau0_rule = std::string(val()) == "nn" ? ACO_TAU0_NN_TOUR : ACO_TAU0_FIRST_COLONY;
        } else if (s == "--dist") {
            dist_mode = std::string(val()) == "nint" ? ACO_DIST_NINT : ACO_DIST_FLOAT;
        } else if (s.size() > 1 && s[0] == '-') {
            usage(argv[0]);
        } else {
            path = argv[a];
        }
    }
    if (!path) usage(argv[0]);
    if (defaults.selection == ACO_SEL_KNN && defaults.k <= 0) defaults.k = 20;

    Problem best_inst;   // the try with the shortest best tour
    for (int ntry = 0; ntry < tries; ++ntry) {
        Problem inst = defaults;
        const double t0 = now_s();
        read_instance_file(inst, path, dist_mode);  // main.cpp:73-74
        printf("instance %s: %d nodes, capacity %d, optimum %.2f\n", inst.name.c_str(),
               inst.num_node, inst.vehicle_capacity, inst.optimum);
        printf("Initialization took %.6f seconds\n", now_s() - t0);
        inst.seed = defaults.seed + (unsigned long long)ntry;
        GpuACO aco(inst);
        aco.init_aco();
        // termination_condition (main.cpp:29-34), on the device best-so-far;
        // with --decompose a round after every master_iterations iterations,
        // its merged tour put into the next iteration when it is shorter
        std::vector<int> pending;
        float pending_len = INFINITY;
        int round = 0;
        while (inst.iteration < inst.max_iteration && now_s() - t0 < inst.max_runtime) {
            if (!pending.empty()) {
                aco.run_aco_iteration_with(pending);
                pending.clear();
                pending_len = INFINITY;
            } else {
                aco.run_aco_iteration();
            }
            inst.iteration++;
            if (inst.decompose && inst.iteration % std::max(1, inst.master_iterations) == 0) {
                std::vector<int> merged;
                const float len = aco.decompose_round(round++, merged);
                aco.update_best_so_far_from_device();
                if (len < inst.best_length) {
                    pending = merged;
                    pending_len = len;
                }
            }
            if (inst.optimum > 0.0f) {
                aco.update_best_so_far_from_device();
                const float best = std::min(inst.best_length, pending_len);
                if (fabsf(best - inst.optimum) < 10 * 0.001f) break;
            }
        }
        aco.exit_aco();
        if (!pending.empty() && pending_len < inst.best_length) {
            inst.best_tour = pending;
            inst.best_length = pending_len;
        }
        // exit_report (io.cpp:238-291)
        const bool ok = check_solution(inst, inst.best_tour);
        printf("try %d: Best Length %.2f Iterations %d At time %.3f Tot.time %.3f valid %s\n", ntry,
               inst.best_length, inst.iteration, inst.best_time, now_s() - t0, ok ? "yes" : "NO");
        if (report_dir) {
            std::string f = std::string(report_dir) + "/best_so_far." + inst.name;
            FILE *fp = fopen(f.c_str(), ntry ? "a" : "w");
            if (fp) {
                for (const aco_bsf_record &r : aco.records())
                    fprintf(fp, "try %d iter %d length %.2f time %.4f\n", ntry, r.iter, r.length, r.time);
                fclose(fp);
            }
            f = std::string(report_dir) + "/best." + inst.name;
            fp = fopen(f.c_str(), ntry ? "a" : "w");
            if (fp) {
                fprintf(fp, "try %d length %.2f tour", ntry, inst.best_length);
                for (int v : inst.best_tour) fprintf(fp, " %d", v);
                fprintf(fp, "\n");
                fclose(fp);
            }
        }
        if (!ok) {
            fprintf(stderr, "invalid solution\n");
            exit(1);
        }
        if (best_inst.best_tour.empty() || inst.best_length < best_inst.best_length) best_inst = inst;
    }
    std::string base = best_inst.name;   // the instance file's name without ".vrp"
    if (base.size() > 4 && base.compare(base.size() - 4, 4, ".vrp") == 0) base.resize(base.size() - 4);
    if (sol_dir && !write_sol(best_inst, std::string(sol_dir) + "/" + base + ".sol")) {
        fprintf(stderr, "cannot write the .sol file in %s\n", sol_dir);
        exit(1);
    }
    return 0;
}
