/*
 * cvrp_aco.h -- C ABI of the MI355X ant-colony CVRP solver (libcvrpaco.so).
 *
 * This is the drop-in boundary for the reference's per-iteration loop
 * (class g_ACO, cvrp_ocl/gpu/g_aco.h:16-37, driven from
 * cvrp_ocl/main.cpp:80-88).  Plain C types only: pointers, sizes, scalars.
 * Every function returns ACO_OK (0) or a negative ACO_ERR_* code; the text
 * of the last error of the calling thread is aco_last_error().  The
 * reference instead calls exit(1) on any device error
 * (g_aco.cpp:607-614); the C++ host wrapper restores that behaviour.
 *
 * A context owns all device memory and one HIP stream.  Calls are
 * asynchronous on that stream; aco_get_* synchronise.  A context is not
 * thread-safe.  Inputs are copied at aco_create; no host pointer is kept.
 */
#ifndef CVRP_ACO_H
#define CVRP_ACO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACO_OK               0
#define ACO_ERR_ARG         -1   /* invalid argument or size               */
#define ACO_ERR_HIP         -2   /* HIP runtime error                      */
#define ACO_ERR_STATE       -3   /* call out of order (e.g. before init)   */
#define ACO_ERR_INFEASIBLE  -4   /* instance has an unservable customer    */
#define ACO_ERR_NODEV       -5   /* no HIP device                          */
#define ACO_ERR_IO          -6   /* file could not be read or parsed       */

/* next-node rule */
#define ACO_SEL_IROULETTE    0   /* I-Roulette over the full row (kernel.cl:325-373) */
#define ACO_SEL_KNN          1   /* roulette over the k-nearest candidate row        */
/* random numbers */
#define ACO_RNG_PARK_MILLER  0   /* per-lane minimal-standard streams (kernel.cl:42-52) */
#define ACO_RNG_PHILOX       1   /* counter-based Philox-4x32-10                     */
/* local search flags, applied in this order */
#define ACO_LS_2OPT          1   /* intra-route 2-opt (kernel.cl:126-319)            */
#define ACO_LS_SWAP          2   /* swap (localSearch.cpp:301-410)                   */
#define ACO_LS_RELOCATE      4   /* relocate (neighbourSearch.cpp:227-365)           */
/* 2-opt in the CPU variant's random scan order (LocalSearch::two_opt_single_route
 * with generate_random_permutation, localSearch.cpp:97-119,180-294): one
 * Park-Miller stream per instance, consumed ant after ant and route after
 * route as sequential calls would; replaces ACO_LS_2OPT (not both). */
#define ACO_LS_2OPT_RANDOM   8
/* initial trail */
#define ACO_TAU0_FIRST_COLONY 0  /* 0.5, one colony, then 1/(rho*L) (g_aco.cpp:79-90) */
#define ACO_TAU0_NN_TOUR      1  /* 1/(n*L_nn) of a nearest-neighbour tour           */

/* Instance (replaces Problem, cvrp_ocl/problem.h:82-140).  Node 0 is the
 * depot.  dist is n*n row-major and symmetric; with integer (nint)
 * distances every cost and move delta is exact. */
typedef struct aco_problem {
    int32_t        n;             /* nodes including the depot                 */
    const float   *dist;          /* n*n                                       */
    const int32_t *demand;        /* n                                         */
    int32_t        capacity;      /* vehicle_capacity                          */
    float          max_dist;      /* max_distance, INFINITY when unlimited     */
    float          service_time;  /* service_time                              */
    const int32_t *nn_list;       /* optional n*nn_width sorted neighbour ids  */
    int32_t        nn_width;      /* >= max(k, nn_ls) when nn_list != NULL     */
    const float   *prior;         /* optional n*n multiplier of eta^beta       */
} aco_problem;

/* Parameters (problem.cpp:186-220 defaults, runtime instead of macros). */
typedef struct aco_params {
    float    alpha, beta, rho;    /* 1, 2, 0.1                                 */
    int32_t  n_ants;              /* m; 0 = n (the reference's default)        */
    int32_t  ras_ranks;           /* 6: best-so-far + 5 iteration elites        */
    int32_t  k;                   /* candidate row width for ACO_SEL_KNN (<=64) */
    int32_t  nn_ls;               /* neighbour depth used by 2-opt; 0 = n-2     */
    int32_t  selection;           /* ACO_SEL_*                                 */
    int32_t  rng;                 /* ACO_RNG_*                                 */
    int32_t  ls_flags;            /* ACO_LS_* bit set                          */
    int32_t  tau0_rule;           /* ACO_TAU0_*                                */
    uint64_t seed;                /* colony seed                               */
    int32_t  bsf_capacity;        /* best-so-far record slots (4096)           */
    int32_t  use_graph;           /* 1: aco_iteration replays a hipGraph        */
} aco_params;

/* BestSolutionInfo (problem.h:62-66); time is seconds since aco_init,
 * taken from the device clock so a captured iteration needs no host. */
typedef struct aco_bsf_record {
    int32_t iter;
    float   length;
    float   time;
} aco_bsf_record;

typedef struct aco_counters {
    int64_t ls_evals;             /* local-search move evaluations            */
    int64_t ls_moves;             /* applied local-search moves               */
    int64_t fallbacks;            /* full-row greedy picks in construction     */
    int64_t iterations;           /* completed colony iterations              */
    int32_t bsf_overflow;         /* records dropped because the log was full */
} aco_counters;

typedef struct aco_ctx aco_ctx;

/* Fills the reference defaults (problem.cpp:186-220, g_aco.cpp:79-90). */
void aco_default_params(aco_params *params);

/* g_ACO::g_ACO + create_memory_objects (g_aco.cpp:35-43, 640-761):
 * validates the instance (every customer servable alone: demand <= Q and
 * d(0,i)+svc+d(i,0) <= max_dist), preprocesses it and allocates device
 * memory. */
int aco_create(const aco_problem *problem, const aco_params *params, aco_ctx **ctx);
/* g_ACO::~g_ACO (g_aco.cpp:45-61) */
int aco_destroy(aco_ctx *ctx);

/* Batched colonies (no reference counterpart; the reference solves one
 * instance per process): count instances of uneven size in one context,
 * each with params->n_ants ants and its own pheromone, best-so-far and
 * records; every launch covers all of them.  Instance b uses the seed
 * params->seed + b.  The *_at accessors take the instance index; the plain
 * accessors refer to instance 0. */
int aco_create_batch(const aco_problem *problems, int32_t count, const aco_params *params,
                     aco_ctx **ctx);
int32_t aco_batch_size(aco_ctx *ctx);
int aco_get_best_at(aco_ctx *ctx, int32_t instance, int32_t *tour, int32_t capacity,
                    int32_t *size, float *length);
int aco_get_ants_at(aco_ctx *ctx, int32_t instance, int32_t *tours, int32_t stride,
                    int32_t *sizes, float *lengths);
int aco_set_ants_at(aco_ctx *ctx, int32_t instance, const int32_t *tours, int32_t stride,
                    const int32_t *sizes, const float *lengths);
int aco_get_bsf_at(aco_ctx *ctx, int32_t instance, aco_bsf_record *records, int32_t capacity,
                   int32_t *count);
int aco_get_pheromone_at(aco_ctx *ctx, int32_t instance, float *tau, float *total_info);
int aco_set_pheromone_at(aco_ctx *ctx, int32_t instance, const float *tau,
                         const float *total_info);

/* The arrays aco_create derived on the device from an instance's distances
 * (init_problem -> compute_distances / compute_nn_lists, vrpHelper.cpp:160-236,
 * and the heuristic of compute_total_info, kernel.cl:782-789): eta^beta
 * (times the prior) n*n, the depot column n, the candidate rows n*k and their
 * distances, the 2-opt neighbour rows n*nn_ls and their distances, nnpos
 * n*n (position of j in i's neighbour row, 0xFFFF if absent) and the 16-bit
 * distances n*n (filled only when every distance is an integer in
 * [0, 65535]).  Any output may be NULL. */
int aco_get_setup_at(aco_ctx *ctx, int32_t instance, float *heur, float *d0, int32_t *nnk, float *nnkd,
                     int32_t *nnls, float *nnd, uint16_t *nnpos, uint16_t *dist16);

/* Replaces the construction candidate rows of an instance (the first k
 * entries of each row of cand, n rows of `width` >= k entries): the roulette
 * over candidates then draws from these nodes.  The reference takes them as
 * the first nn_ants entries of the distance-sorted nn_list (kernel.cl:
 * 382-493 with compute_nn_lists, vrpHelper.cpp:188-236); a build-only
 * addition lets a caller choose another set (e.g. ranked by a learned
 * prior).  The 2-opt neighbour rows stay distance-sorted.  Each row holds k
 * distinct customers other than its own node, else ACO_ERR_ARG.  Before or
 * after aco_init; afterwards the candidates' total_info is gathered again. */
int aco_set_candidates_at(aco_ctx *ctx, int32_t instance, const int32_t *cand, int32_t width);

/* g_ACO::init_aco (g_aco.cpp:63-93) */
int aco_init(aco_ctx *ctx);
/* g_ACO::run_aco_iteration (g_aco.cpp:122-134): construct -> local search
 * -> statistics -> pheromone update; replayed as one hipGraph when
 * params.use_graph */
int aco_iteration(aco_ctx *ctx);
/* iterations x aco_iteration, without host synchronisation */
int aco_run(aco_ctx *ctx, int32_t iterations);

/* the steps of one iteration (g_aco.cpp:141-177, 182-228, 407-417, 233-243) */
int aco_construct(aco_ctx *ctx);
int aco_local_search(aco_ctx *ctx);
int aco_update_statistics(aco_ctx *ctx);
int aco_pheromone_update(aco_ctx *ctx);
/* g_ACO::pheromone_init (g_aco.cpp:245-263) */
int aco_pheromone_init(aco_ctx *ctx, float tau0);

int aco_synchronize(aco_ctx *ctx);
int aco_get_iteration(aco_ctx *ctx, int32_t *iteration);

/* g_ACO::update_best_so_far_from_device (g_aco.cpp:527-552) */
int aco_get_best(aco_ctx *ctx, int32_t *tour, int32_t capacity, int32_t *size, float *length);
/* all m ants and the best-so-far slot m: tours[(m+1)*stride] */
int aco_get_ants(aco_ctx *ctx, int32_t *tours, int32_t stride, int32_t *sizes, float *lengths);
/* loads m+1 tours (slot m = best-so-far; size 0 leaves it empty); used to
 * replay a state.  Each tour starts and ends at the depot and visits every
 * customer exactly once with no empty route, else ACO_ERR_ARG. */
int aco_set_ants(aco_ctx *ctx, const int32_t *tours, int32_t stride, const int32_t *sizes,
                 const float *lengths);
/* best-so-far records (g_aco.cpp:95-117); *count = records written */
int aco_get_bsf(aco_ctx *ctx, aco_bsf_record *records, int32_t capacity, int32_t *count);
/* n*n pheromone and total_info, row-major */
int aco_get_pheromone(aco_ctx *ctx, float *tau, float *total_info);
int aco_set_pheromone(aco_ctx *ctx, const float *tau, const float *total_info);
/* m*64 Park-Miller lane states of instance 0 (the per-ant draw streams,
 * kernel.cl:42-52); set replaces them (each in [1, 2^31-2]), e.g. to replay
 * a recorded state or force a particular draw */
int aco_get_rng_state(aco_ctx *ctx, int32_t *state);
int aco_set_rng_state(aco_ctx *ctx, const int32_t *state);
/* state of instance 0's scan-order stream for ACO_LS_2OPT_RANDOM (the
 * reference's instance->rnd_seed, utilities.cpp:271-288); after aco_init */
int aco_get_ls_seed(aco_ctx *ctx, int32_t *seed);
int aco_set_ls_seed(aco_ctx *ctx, int32_t seed);
int aco_get_counters(aco_ctx *ctx, aco_counters *counters);
/* per-step device time: enable = 1 records a HIP event pair around each
 * launch (iterations then run without the graph) */
int aco_profile_enable(aco_ctx *ctx, int32_t enable);
/* names: "construct", "local_search", "statistics", "pheromone", "gather";
 * returns total milliseconds and launch count since enabling */
int aco_profile_get(aco_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

/* ---- launch shapes and the autotuner (no counterpart in the reference,
 * whose launch geometry is fixed in g_aco.cpp:141-177,279-318) ----
 * Local-search kernel shape: waves per ant (4, 8, 12 or 16), rows in the LDS
 * distance-row ring (8, 16 or 32; >= waves + 2) and prefetch of the next
 * round's rows (needs ring >= 2 waves + 2).  Construction: construct_apb > 0 stages the candidate
 * table in LDS and runs that many ants per block (1, 2 or 4; candidate
 * roulette without a route-length limit, ants per instance divisible by
 * it); 0 reads candidate rows from L2/HBM.  ls_row16 keeps the
 * local-search rows as 16-bit integers (exact; only when every distance is
 * an integer in [0, 65535], e.g. VRPLIB nint).  fused_iteration runs
 * aco_iteration as one launch (after a one-thread queue reset): blocks of
 * ls_waves (4 or 8, dividing the ants per instance) take W ants of one
 * instance at a time, build them (one wave each) and search them, and the
 * block finishing an instance runs its statistics and pheromone update;
 * candidate roulette without a route-length limit only.  The per-phase
 * entry points (aco_construct, ...) are unaffected.  Every valid shape gives
 * identical results. */
typedef struct aco_launch_shape {
    int32_t ls_waves;
    int32_t ls_ring;
    int32_t ls_prefetch;
    int32_t construct_apb;
    int32_t ls_row16;
    int32_t fused_iteration;
} aco_launch_shape;

typedef struct aco_tune_result {
    aco_launch_shape shape;
    int32_t stage;       /* 0: construction candidate, 1: local-search candidate,
                            2: whole iteration, per-phase launches or fused */
    float ms;            /* mean time of that stage, -1 if it does not apply */
} aco_tune_result;

int aco_get_launch_shape(aco_ctx *ctx, aco_launch_shape *shape);
/* validates the shape and its LDS size; the iteration graph is re-captured */
int aco_set_launch_shape(aco_ctx *ctx, const aco_launch_shape *shape);
/* After aco_init: times every construction candidate `reps` times from the
 * current state, then every local-search candidate on the same constructed
 * tours, keeps the fastest of each, then times whole iterations with those
 * (per-phase launches) against the fused launch with up to three shapes,
 * keeps the fastest and restores the state, so the colony's results do not
 * change; the clock of the best-so-far records restarts afterwards.
 * results (optional, capacity entries) receives every candidate's time. */
int aco_autotune(aco_ctx *ctx, int32_t reps, aco_tune_result *results, int32_t capacity,
                 int32_t *count);

const char *aco_last_error(void);

/* ---- host-side instance preparation (no device needed) ---- */

/* compute_distances (vrpHelper.cpp:160-184).  mode: ACO_DIST_* */
#define ACO_DIST_FLOAT  0   /* float sqrt, unrounded: round_distance (vrpHelper.cpp:67-81) */
#define ACO_DIST_NINT   1   /* VRPLIB nint(euclid), the X-set convention                 */
#define ACO_DIST_CEIL   2   /* ceil_distance (vrpHelper.cpp:83-97)                       */
#define ACO_DIST_GEO    3   /* geo_distance (vrpHelper.cpp:99-136)                       */
#define ACO_DIST_ATT    4   /* att_distance (vrpHelper.cpp:138-158)                      */
int aco_compute_distances(int32_t n, const double *x, const double *y, int32_t mode, float *dist);
/* The same distances for `count` instances on the current HIP device:
 * x[b], y[b] (n[b] doubles) into dist[b] (n[b] x n[b]), host arrays.
 * ACO_DIST_FLOAT, _NINT, _CEIL and _ATT only (equal to aco_compute_distances
 * element for element); GEO stays on the host. */
int aco_compute_distances_batch(int32_t count, const int32_t *n, const double *const *x,
                                const double *const *y, int32_t mode, float *const *dist);
/* compute_nn_lists (vrpHelper.cpp:188-236): self and depot excluded;
 * tie 0 = the reference's quicksort order, tie 1 = (distance, index) */
int aco_compute_nn_lists(int32_t n, const float *dist, int32_t nn, int32_t tie, int32_t *out);
/* The same lists for `count` instances at once on the current HIP device
 * (compute_nn_lists, vrpHelper.cpp:188-236; one thread per row, the same
 * sort steps, so the result equals aco_compute_nn_lists instance by
 * instance).  n[b], dist[b] (n[b] x n[b] host array), nn[b] (1..n[b]-2)
 * entries per row into out[b] (n[b] x nn[b] host array).  aco_create uses
 * it for instances given without nn_list. */
int aco_compute_nn_lists_batch(int32_t count, const int32_t *n, const float *const *dist,
                               const int32_t *nn, int32_t tie, int32_t *const *out);
/* The same lists without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * dist[b] (n[b] x n[b]) and out[b] (n[b] x nn[b]) are DEVICE arrays;
 * `stream` is a hipStream_t (NULL: the null stream) and the call only
 * enqueues work on it (nothing is waited for or copied).  Sorting uses a
 * library-owned 64 MiB device array per GPU, held until the process ends;
 * a row of an n-node instance takes 8n bytes of it, so one launch sorts
 * the rows of several instances while they fit (n <= 4096 always fits).
 * Calls that overlap on one GPU share that array and race: serialise them
 * (one stream, or the caller's own ordering).  Results equal
 * aco_compute_nn_lists_batch. */
int aco_compute_nn_lists_batch_noalloc(int32_t count, const int32_t *n, const float *const *dist,
                                       const int32_t *nn, int32_t tie, int32_t *const *out,
                                       void *stream);
/* read_instance_file (io.cpp:334-517).  First call with capacity 0 to get
 * *n; then call with capacity >= *n to fill the arrays.  max_dist is
 * INFINITY and service_time 0 when the file has none; edge_weight_type is
 * 0 EUC_2D, 1 CEIL_2D, 2 GEO, 3 ATT. */
int aco_read_vrplib(const char *path, int32_t capacity, int32_t *n, double *x, double *y,
                    int32_t *demand, int32_t *vehicle_capacity, float *max_dist,
                    float *service_time, float *optimum, int32_t *edge_weight_type);
/* Initial trail of the ACO_TAU0_NN_TOUR rule, tau0 = 1/(n * L_nn).  No
 * reference counterpart: the reference derives tau0 from a first colony
 * (g_aco.cpp:79-90); this is the nearest-neighbour rule of the ACO
 * literature, restated for capacity and route length: from the current
 * node go to the nearest unvisited customer i (ties to the lowest index)
 * with load + demand[i] <= capacity and rdist + (d(cur,i) + service_time) +
 * d(i,0) <= max_dist; when none fits, back to the depot.  *length adds the
 * edges in step order.  tour (optional, tour_capacity >= 2n) receives the
 * node sequence, depot first and last, *size its length.
 * ACO_ERR_INFEASIBLE when some customer cannot be served. */
int aco_nn_tour(int32_t n, const float *dist, const int32_t *demand, int32_t capacity, float max_dist,
                float service_time, int32_t *tour, int32_t tour_capacity, int32_t *size, float *length);
/* check_solution (problem.cpp:226-319); returns 1 valid, 0 invalid */
int aco_check_solution(int32_t n, const float *dist, const int32_t *demand, int32_t capacity,
                       float max_dist, float service_time, const int32_t *tour, int32_t size);

#ifdef __cplusplus
}
#endif
#endif /* CVRP_ACO_H */
