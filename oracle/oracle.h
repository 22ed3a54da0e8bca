/*
 * CPU oracle for the CVRP ant-colony hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, what the
 * reference (i-sunny/cvrp_aco_ocl) computes on its per-iteration path
 * (construct_solutions -> local_search -> update_statistics ->
 * pheromone_update, cvrp_ocl/gpu/g_aco.cpp:122-134, kernels in
 * cvrp_ocl/gpu/kernel.cl).  It is the checker that the HIP path is compared
 * against and the timed CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links or calls it.
 *
 * Arithmetic contract (both sides): IEEE float32, round-to-nearest, no FMA
 * contraction, denormals kept, correctly rounded division.  Every function
 * cites the reference lines it follows; where the reference has no
 * counterpart the comment says so and states the rule that is implemented.
 */
#ifndef CVRP_ORACLE_H
#define CVRP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EPSILON 0.001f          /* utilities.h:19, kernel.cl:6 */
#define ORC_LANES   64              /* work-items per ant (WG/4 on AMD) */
#define ORC_MAX_MOVES_PER_ROUTE (1 << 14)  /* termination cap, 2-opt */
#define ORC_MAX_MOVES_FACTOR 64             /* swap/relocate cap = 64*n */

/* ---------------- random numbers ---------------- */
float    orc_ran01(int32_t *idum);                        /* utilities.cpp:271-288 */
uint64_t orc_splitmix64(uint64_t x);
int32_t  orc_pm_seed(uint64_t seed, int ant, int lane);
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int32_t  orc_ls_seed(uint64_t seed);                      /* scan-order stream seed */
int32_t  orc_pm_jump(int32_t z, uint64_t t);             /* state after t draws */
float    orc_philox_u01(uint64_t seed, uint32_t iter, uint32_t ant, uint32_t step);

/* ---------------- instance preprocessing ---------------- */
/* mode 0: float sqrt as the reference (vrpHelper.cpp:67-81); 1: nint int32 */
void orc_distances(int n, const double *x, const double *y, int mode, float *dist);
void orc_sort2(float *v, int *v2, int left, int right);  /* utilities.cpp:246-267 */
/* tie 0: reference quicksort order (vrpHelper.cpp:188-236); 1: (dist, index) */
void orc_nn_lists(int n, const float *dist, int nn, int tie, int *out);
/* heur = (1/(d+0.1f))^beta, beta==2 closed form y*y (kernel.cl:678,788) */
void orc_heuristic(int n, const float *dist, float beta, const float *prior, float *heur);
/* priority order of the reference's reduction tree (kernel.cl:340-366) */
void orc_tree_rank(int N, int *rank);
int  orc_tree_argmax(int N, const float *val);            /* kernel.cl:340-366 */
int  orc_minreduce_phase0(int n, int grp, const float *val); /* kernel.cl:594-665 */

/* ---------------- problem bundle ---------------- */
typedef struct {
    int n;
    const float *dist;      /* n*n row-major, symmetric */
    const int   *dem;       /* n */
    int          Q;
    float        D;         /* INFINITY when absent */
    float        svc;
    const float *heur;      /* n*n */
    const int   *nnk;       /* n*k candidate rows (roulette_knn) */
    int          k;
    const int   *nnls;      /* n*nn_ls (2-opt) */
    int          nn_ls;
} orc_problem;

/* selection: 0 iroulette_full (kernel.cl:325-373), 1 roulette_knn
 * rng: 0 Park-Miller per-lane streams, 1 Philox counter draws              */
typedef struct {
    int      selection;
    int      rng;
    uint64_t seed;
    uint32_t iteration;
} orc_draw_cfg;

/* one ant; tinfo n*n, tiknn n*k; pm = this ant's 64 lane states (in/out)
 * returns tour size; writes *len and *fallbacks                          */
int orc_construct_ant(const orc_problem *p, const float *tinfo, const float *tiknn,
                      const orc_draw_cfg *cfg, int ant, int32_t *pm,
                      int *tour, float *len, int *fallbacks);

/* ---------------- statistics (kernel.cl:546-741) ---------------- */
/* returns iteration-best ant; flag set when it improves the best-so-far */
int  orc_iteration_best(int m, const float *lens, float bsf_len, int *flag);
void orc_elites(int m, int ras, const float *lens, int *elite);

/* ---------------- pheromone (kernel.cl:672-789) ---------------- */
void orc_pheromone_init(int n, float tau0, const float *heur, float alpha,
                        float *tau, float *tinfo);
/* tours: ras tours (elite order), sizes, lens; evaporation then sequential
 * deposit in elite order, then total_info                               */
void orc_pheromone_update(int n, float rho, float alpha, int ras,
                          const int *const *tours, const int *sizes, const float *lens,
                          const float *heur, float *tau, float *tinfo);
void orc_gather_knn(int n, int k, const int *nnk, const float *tinfo, float *tiknn);

/* ---------------- local search ---------------- */
/* 2-opt of every route, reference GPU order (kernel.cl:126-319).  perm is
 * NULL for the sequential scan order, or a function that returns the scan
 * order of a route of num nodes (CPU variant, localSearch.cpp:97-119,180-294)
 * through the Park-Miller seed *seed.  returns moves applied; *evals counts
 * evaluated moves                                                        */
int   orc_two_opt(const orc_problem *p, int *tour, int size, int random_order,
                  int32_t *seed, long long *evals);
/* first-improvement swap (localSearch.cpp:301-410) */
int   orc_swap(const orc_problem *p, int *tour, int size, long long *evals);
/* first-improvement relocate; deltas neighbourSearch.cpp:271-293,344-351,
 * apply move.cpp:32-51; scan policy is this project's own                 */
int   orc_relocate(const orc_problem *p, int *tour, int size, long long *evals);
/* ls flags: 1 two_opt, 2 swap, 4 relocate; returns new length           */
float orc_local_search(const orc_problem *p, int flags, int *tour, int size,
                       long long *evals);
/* the same with 8: 2-opt in random scan order drawn from *seed           */
float orc_local_search_seeded(const orc_problem *p, int flags, int *tour, int size,
                              long long *evals, int32_t *seed);

/* ---------------- solution checks (problem.cpp:226-319) ---------------- */
float orc_tour_length(int n, const float *dist, const int *tour, int size);
int   orc_check_solution(const orc_problem *p, const int *tour, int size);
/* nearest-neighbour capacity-respecting tour (tau0 = 1/(n*L_nn)) */
int   orc_nn_tour(const orc_problem *p, int *tour, float *len);

/* ---------------- whole loop (g_aco.cpp:63-134) ---------------- */
typedef struct {
    int      m;
    float    alpha, beta, rho;
    int      ras;
    int      selection, rng, ls_flags;
    int      tau0_rule;     /* 0 first colony (g_aco.cpp:79-90), 1 nn tour */
    uint64_t seed;
    int      threads;       /* OpenMP threads over ants (0 = 1) */
} orc_params;

typedef struct { int iter; float length; float time; } orc_bsf_record;

typedef struct orc_colony orc_colony;
orc_colony *orc_colony_create(const orc_problem *p, const orc_params *prm);
void  orc_colony_destroy(orc_colony *c);
void  orc_colony_init(orc_colony *c);
void  orc_colony_construct(orc_colony *c);
void  orc_colony_local_search(orc_colony *c);
void  orc_colony_statistics(orc_colony *c);
void  orc_colony_pheromone_update(orc_colony *c);
void  orc_colony_iteration(orc_colony *c);
void orc_colony_next_iteration(orc_colony *c);   /* caller's iteration++ (main.cpp:80-88) */
int   orc_colony_iter(const orc_colony *c);
float orc_colony_best(const orc_colony *c, int *tour, int *size);
int   orc_colony_ants(const orc_colony *c, int *tours, int stride, int *sizes, float *lens);
int   orc_colony_bsf(const orc_colony *c, orc_bsf_record *rec, int cap);
void  orc_colony_tau(const orc_colony *c, float *tau, float *tinfo);
void  orc_colony_set_tau(orc_colony *c, const float *tau, const float *tinfo);
void  orc_colony_set_ants(orc_colony *c, const int *tours, int stride, const int *sizes,
                          const float *lens);
long long orc_colony_ls_evals(const orc_colony *c);
/* full-row greedy picks of the construction steps so far */
long long orc_colony_fallbacks(const orc_colony *c);
/* scan-order stream of the random-order 2-opt (ls flag 8) */
int32_t orc_colony_ls_seed(const orc_colony *c);
void  orc_colony_set_ls_seed(orc_colony *c, int32_t seed);
/* m*64 Park-Miller lane states */
void  orc_colony_set_pm_state(orc_colony *c, const int32_t *state);
void  orc_colony_pm_state(const orc_colony *c, int32_t *out);


/* ---------------- sub-problem decomposition ---------------- */
/* compute_route_centers (vrpHelper.cpp:274-294): per route r of a
 * depot-separated tour (routes without customers skipped), beg[r] / end[r]
 * = positions of its opening and closing depot, the summed demand and the
 * demand-weighted centre (each term x*d*1.0/D in double, accumulated into a
 * float as the reference's Point).  Returns the number of routes, -1 if cap
 * is too small. */
int orc_route_centers(int n, const float *x, const float *y, const int *dem, const int *tour, int size,
                      int cap, int *beg, int *end, float *cx, float *cy, int *rdem);
/* No reference counterpart (the reference keeps num_subs = n/50 and a
 * per-sub real_nodes list, problem.cpp:218, problem.h:136-139, but never
 * partitions): routes ordered by the angle of their centre about the
 * depot (atan2 in double; a route of zero demand uses the plain mean of
 * its customers), ties by route order; cut circularly from position
 * rotation mod R into G = clamp(subs > 0 ? subs : min(n/50, R/2), 1, R) contiguous
 * groups: a group closes after a route once its running customer count
 * times G reaches (g+1) times the total, or when exactly one route per
 * remaining group is left.  Group g: real_nodes[offsets[g] .. offsets[g+1])
 * = depot, then its routes' customers (routes in cut order, customers in
 * tour order); its routes as a tour of the sub-problem (local numbering)
 * in sub_tours[toffsets[g] .. toffsets[g+1]).  Returns G, -1 on bad
 * input. */
int orc_partition_routes(int n, const float *x, const float *y, const int *dem, const int *tour, int size,
                         int subs, int rotation, int *offsets, int *real_nodes, int *toffsets, int *sub_tours);
/* init_sub_problem (problem.cpp:84-130): sub distances and demands
 * gathered from the master's through real_nodes. */
void orc_sub_problem(int n, const float *dist, const int *dem, int sub_n, const int *real_nodes,
                     float *sub_dist, int *sub_dem);
/* No reference counterpart: sub s's best tour replaces its original routes
 * when its length (in step order over the sub's distances) is strictly
 * smaller than theirs; the merged tour is the groups in order, each the
 * chosen tour mapped to master numbering, depot shared between groups.
 * Returns the merged size; *len in step order; replaced[s] 0/1. */
int orc_merge_sub_tours(int n, const float *dist, int count, const int *offsets, const int *real_nodes,
                        const int *toffsets, const int *sub_tours, const int *boffsets, const int *best_tours,
                        int *out, float *len, int *replaced);
/* Pheromone of a finished sub-colony into the master (the use problem.h:138-139
 * gives best_pheromone; no code of the reference does it): for i != j of
 * the sub, tau[r_i][r_j] = sub_tau[i][j], total_info recomputed there as
 * compute_total_info (kernel.cl:782-789) does: tau^alpha * heur. */
void orc_pheromone_from_sub(int n, float alpha, const float *heur, float *tau, float *tinfo, int sub_n,
                            const int *real_nodes, const float *sub_tau);

#ifdef __cplusplus
}
#endif
#endif
