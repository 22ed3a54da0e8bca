/*
 * cvrp_decompose.h -- sub-problem decomposition of a colony's best solution
 * (libcvrpaco.so).
 *
 * The reference carries the pieces of a route-based decomposition without
 * a driver: compute_route_centers (cvrp_ocl/vrpHelper.cpp:274-294), the
 * sub-problem set-up init_sub_problem (cvrp_ocl/problem.cpp:84-130), the
 * sub-problem fields num_subs = num_node/50, real_nodes and best_pheromone
 * (problem.cpp:218, problem.h:136-139), 75 sub-colony iterations per outer
 * loop and 1 master iteration (problem.cpp:216-217), and the report
 * print_problem_decompositon (io.cpp:83-97).  This header exposes those
 * pieces and a driver built from them: the routes of the best solution are
 * grouped into sub-problems, every sub-problem is solved by its own colony
 * (all of them in one batched context on the GPU), and a sub-problem's best
 * tour replaces its routes when it is shorter.  The master colony needs no
 * local search for this, so instances past the local-search LDS limit
 * (n > about 2000) still get improved routes.
 *
 * Same conventions as cvrp_aco.h: ACO_OK or a negative ACO_ERR_* code,
 * aco_last_error() for the text.
 */
#ifndef CVRP_DECOMPOSE_H
#define CVRP_DECOMPOSE_H

#include <stdint.h>

#include "cvrp_aco.h"

#ifdef __cplusplus
extern "C" {
#endif

/* compute_route_centers (vrpHelper.cpp:274-294).  tour: depot-separated,
 * depot first and last (size entries); x, y: node coordinates as floats
 * (Point, problem.h:40-44).  For each route r with at least one customer:
 * beg[r], end[r] = positions of its opening and closing depot, demand[r]
 * its summed demand and (cx[r], cy[r]) the demand-weighted centre, each
 * term x*d*1.0/D evaluated in double and accumulated into a float as the
 * reference does.  capacity = room in each output; *count = routes. */
int aco_route_centers(int32_t n, const float *x, const float *y, const int32_t *demand,
                      const int32_t *tour, int32_t size, int32_t capacity, int32_t *beg, int32_t *end,
                      float *cx, float *cy, int32_t *route_demand, int32_t *count);

/* Partition of a solution's routes into sub-problems.  No counterpart in
 * the reference, which stops at num_subs and real_nodes: routes are ordered
 * by the angle of their centre about the depot (atan2 in double; a route
 * of zero demand uses the plain mean of its customers), ties by their order
 * in the tour, and cut circularly, starting at position rotation mod R, into
 * G = clamp(subs > 0 ? subs : min(n/50, R/2), 1, R) runs of consecutive
 * routes (the default n/50 of problem.cpp:218, at most half the routes so
 * that every sub-problem can move customers between two routes or more); a run
 * closes after a route once its customers times G reach (g+1) times all
 * customers, or when exactly one route per remaining run is left.
 * Sub-problem g: real_nodes[offsets[g] .. offsets[g+1]) = 0 (the depot),
 * then its routes' customers (routes in cut order, customers in tour
 * order); its routes as a tour of the sub-problem (local numbering, node i
 * = real_nodes[offsets[g] + i]) in sub_tours[tour_offsets[g] ..
 * tour_offsets[g+1]).  Room: offsets and tour_offsets size + 1 entries,
 * real_nodes n + size, sub_tours 2 * size.  *count = G. */
int aco_partition_routes(int32_t n, const float *x, const float *y, const int32_t *demand,
                         const int32_t *tour, int32_t size, int32_t subs, int32_t rotation,
                         int32_t *offsets, int32_t *real_nodes, int32_t *tour_offsets,
                         int32_t *sub_tours, int32_t *count);

/* init_sub_problem (problem.cpp:84-130): the distances (sub_n x sub_n) and
 * demands of the sub-problem on real_nodes[0..sub_n) from the master's;
 * real_nodes[0] must be the depot and every entry distinct. */
int aco_sub_problem(int32_t n, const float *dist, const int32_t *demand, int32_t sub_n,
                    const int32_t *real_nodes, float *sub_dist, int32_t *sub_demand);

/* Merge of sub-problem results.  No counterpart in the reference: sub s's
 * best tour (best_tours[best_offsets[s] .. best_offsets[s+1]), local
 * numbering; an empty range keeps the routes) replaces its original routes
 * (sub_tours as aco_partition_routes returns them) when its length, added
 * in step order over the master's distances, is strictly smaller.  The
 * merged tour is the sub-problems in order, each contributing the chosen
 * routes mapped to master numbering (capacity >= 2n); *length adds its
 * edges in step order; replaced[s] = 1 where the best tour was taken. */
int aco_merge_sub_tours(int32_t n, const float *dist, int32_t count, const int32_t *offsets,
                        const int32_t *real_nodes, const int32_t *tour_offsets,
                        const int32_t *sub_tours, const int32_t *best_offsets,
                        const int32_t *best_tours, int32_t *out, int32_t capacity, int32_t *size,
                        float *length, int32_t *replaced);

/* One ant of instance `instance` (0 <= ant < n_ants) set to `tour` (depot
 * first and last, every customer once, no empty route), its length added in
 * step order as construction does.  No counterpart in the
 * reference: between aco_construct / aco_local_search and
 * aco_update_statistics it puts a solution (e.g. a merged decomposition
 * result) into the colony, which then ranks and deposits it like any ant. */
int aco_set_ant_at(aco_ctx *ctx, int32_t instance, int32_t ant, const int32_t *tour, int32_t size);

/* The sub-colonies' pheromone into the master (best_pheromone,
 * problem.h:138-139: the sub-problem's trail, "used to update the master
 * when the sub-problem's iterations end"; the reference never does it):
 * for every instance s of the batched context `subs`, whose nodes are
 * real_nodes[offsets[map[s]] ..] of the master instance, tau[r_i][r_j] =
 * sub tau[i][j] for i != j and total_info recomputed there as
 * compute_total_info (kernel.cl:782-789), tau^alpha * eta^beta; then the
 * candidate rows are gathered again.  Device to device, one launch. */
int aco_pheromone_from_subs(aco_ctx *master, int32_t instance, aco_ctx *subs, const int32_t *offsets,
                            const int32_t *real_nodes, const int32_t *map);

/* Driver of one decomposition round (the outer loop problem.cpp:216-218
 * implies; no reference code): the best-so-far tour of `instance` of the
 * master context is partitioned (aco_partition_routes with subs and
 * rotation), the sub-problems with at least two customers are set up
 * (aco_sub_problem from `problem`, the master instance as given to
 * aco_create, with coordinates x, y) and solved together in one batched
 * context with `sub_params` (n_ants ants each; seed sub_params.seed + s
 * for the s-th solved sub-problem; k clamped to the smallest sub-problem's
 * n - 2) for `sub_iterations` iterations (0: 75), their bests merged
 * (aco_merge_sub_tours) and, with pheromone_back, their trails written
 * into the master (aco_pheromone_from_subs).  out (capacity >= 2n)
 * receives the merged tour, *size and *length; *subs_out the number of
 * sub-problems and *replaced how many were improved.  The master's own
 * best-so-far is left as it is: the caller puts the merged tour into the
 * colony (aco_set_ant_at before the next aco_update_statistics). */
typedef struct aco_decompose_params {
    int32_t    subs;             /* sub-problems; 0 = min(n/50, routes/2)             */
    int32_t    sub_iterations;   /* per round; 0 = 75 (problem.cpp:217)              */
    int32_t    rotation;         /* first route of the circular cut                  */
    int32_t    pheromone_back;   /* 1: aco_pheromone_from_subs after the sub-colonies */
    aco_params sub_params;       /* sub-colony parameters; n_ants > 0                */
} aco_decompose_params;

int aco_decompose_round(aco_ctx *master, int32_t instance, const aco_problem *problem, const float *x,
                        const float *y, const aco_decompose_params *params, int32_t *out,
                        int32_t capacity, int32_t *size, float *length, int32_t *subs_out,
                        int32_t *replaced);

#ifdef __cplusplus
}
#endif
#endif /* CVRP_DECOMPOSE_H */
