// Test-only C shim over the reference's own CPU sources.
//
// oracle/Makefile compiles this file together with the reference's CPU
// translation units exactly where they lie under the reference checkout
// (utilities.cpp, vrpHelper.cpp, problem.cpp, io.cpp, localSearch.cpp,
// move.cpp, unix_timer.cpp) into oracle/_ref/libcvrpref.so.  Nothing here
// re-implements reference code: each entry point builds the reference's own
// data structures and calls its functions, so the tests can pin the oracle's
// restatement against the real thing.  The device kernels (kernel.cl) have
// no CPU build and are not reachable from here.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "io.h"
#include "localSearch.h"
#include "move.h"
#include "problem.h"
#include "utilities.h"
#include "vrpHelper.h"

namespace {

Problem *make_problem(int n, const float *x, const float *y, const int *dem, int Q, float D,
                      float svc, const float *dist)
{
    Problem *p = new Problem(0);
    p->num_node = n;
    p->nodeptr = (Point *)malloc(sizeof(Point) * n);
    for (int i = 0; i < n; ++i) {
        p->nodeptr[i].x = x[i];
        p->nodeptr[i].y = y[i];
        p->nodeptr[i].demand = dem ? dem[i] : 0;
    }
    p->vehicle_capacity = Q;
    p->max_distance = D;
    p->service_time = svc;
    p->dis_type = DIST_EUC_2D;
    init_problem(p);  // problem.cpp:43-56 (distances, nn lists, ants)
    if (dist) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) p->distance[i][j] = dist[(size_t)i * n + j];
        free(p->nn_list);
        p->nn_list = compute_nn_lists(p);  // vrpHelper.cpp:188-236
    }
    return p;
}

}  // namespace

extern "C" {

// ran01 stream (utilities.cpp:271-288)
void ref_ran01(int *seed, int count, float *out)
{
    for (int i = 0; i < count; ++i) out[i] = ran01(seed);
}

// sort2 (utilities.cpp:246-267)
void ref_sort2(float *v, int *v2, int left, int right) { sort2(v, v2, left, right); }

// distances and nn lists of init_problem; nn width is num_node-2
void ref_nn_lists(int n, const float *x, const float *y, const float *dist, float *dist_out,
                  int *nn_out)
{
    Problem *p = make_problem(n, x, y, NULL, 1, INFINITY, 0.0f, dist);
    int nn = p->nn_ls > p->nn_ants ? p->nn_ls : p->nn_ants;
    if (nn >= n) nn = n - 1;
    for (int i = 0; i < n; ++i) {
        if (dist_out)
            for (int j = 0; j < n; ++j) dist_out[(size_t)i * n + j] = p->distance[i][j];
        for (int j = 0; j < nn; ++j) nn_out[(size_t)i * nn + j] = p->nn_list[i][j];
    }
}

// LocalSearch::do_local_search(AntStruct*): random-order 2-opt of every
// route (localSearch.cpp:83-87), seeded through instance->rnd_seed.
float ref_two_opt(int n, const float *x, const float *y, const int *dem, int Q, float D,
                  float svc, const float *dist, int *seed, int *tour, int size)
{
    Problem *p = make_problem(n, x, y, dem, Q, D, svc, dist);
    p->rnd_seed = *seed;
    AntStruct *a = &p->ants[0];
    memcpy(a->tour, tour, sizeof(int) * size);
    a->tour_size = size;
    LocalSearch ls(p);
    ls.do_local_search(a);
    memcpy(tour, a->tour, sizeof(int) * size);
    *seed = p->rnd_seed;
    return a->tour_length;
}

// LocalSearch::do_local_search(): 2-opt then swap then tour length
// (localSearch.cpp:51-78) applied to one ant.
float ref_two_opt_swap(int n, const float *x, const float *y, const int *dem, int Q, float D,
                       float svc, const float *dist, int *seed, int *tour, int size)
{
    Problem *p = make_problem(n, x, y, dem, Q, D, svc, dist);
    p->rnd_seed = *seed;
    p->n_ants = 1;
    AntStruct *a = &p->ants[0];
    memcpy(a->tour, tour, sizeof(int) * size);
    a->tour_size = size;
    LocalSearch ls(p);
    ls.do_local_search();
    memcpy(tour, a->tour, sizeof(int) * size);
    *seed = p->rnd_seed;
    return a->tour_length;
}

// check_solution (problem.cpp:226-319) and compute_tour_length
// (vrpHelper.cpp:244-253)
int ref_check_solution(int n, const float *x, const float *y, const int *dem, int Q, float D,
                       float svc, const float *dist, const int *tour, int size, float *len)
{
    Problem *p = make_problem(n, x, y, dem, Q, D, svc, dist);
    int *t = (int *)malloc(sizeof(int) * size);
    memcpy(t, tour, sizeof(int) * size);
    int ok = check_solution(p, t, size) ? 1 : 0;
    if (len) *len = compute_tour_length(p, t, size);
    free(t);
    return ok;
}

// InsertionMove::apply / ExchangeMove::apply (move.cpp:17-51)
float ref_apply_move(int kind, int *tour, int size, int pos_n1, int pos_n2, float gain,
                     float len)
{
    AntStruct a;
    a.tour = tour;
    a.tour_size = size;
    a.tour_length = len;
    if (kind == 0) {
        InsertionMove mv(&a, true, gain, pos_n1, pos_n2, 0, 0);
        mv.apply();
    } else {
        ExchangeMove mv(&a, true, gain, pos_n1, pos_n2, 0, 0);
        mv.apply();
    }
    return a.tour_length;
}

// read_instance_file (io.cpp:334-517).  Returns num_node; arrays are filled
// when cap >= num_node.
int ref_read_instance(const char *path, int cap, float *x, float *y, int *dem, int *Q, float *D,
                      float *svc, float *optimum)
{
    Problem *p = new Problem(0);
    read_instance_file(p, path);
    int n = p->num_node;
    if (cap >= n) {
        for (int i = 0; i < n; ++i) {
            x[i] = p->nodeptr[i].x;
            y[i] = p->nodeptr[i].y;
            dem[i] = p->nodeptr[i].demand;
        }
        *Q = p->vehicle_capacity;
        *D = p->max_distance;
        *svc = p->service_time;
        *optimum = p->optimum;
    }
    free(p->nodeptr);
    delete p;
    return n;
}


// compute_route_centers (vrpHelper.cpp:274-294) on routes r = 0..R-1
// spanning tour[beg[r]] .. tour[end[r]]; only nodeptr is read.
void ref_route_centers(int n, const float *x, const float *y, const int *dem, const int *tour, int R,
                       const int *beg, const int *end, float *cx, float *cy, int *rdem)
{
    Problem *p = new Problem(0);
    p->num_node = n;
    p->nodeptr = (Point *)malloc(sizeof(Point) * n);
    for (int i = 0; i < n; ++i) {
        p->nodeptr[i].x = x[i];
        p->nodeptr[i].y = y[i];
        p->nodeptr[i].demand = dem[i];
    }
    vector<RouteCenter *> centers;
    for (int r = 0; r < R; ++r) {
        RouteCenter *c = new RouteCenter();
        c->beg = beg[r];
        c->end = end[r];
        centers.push_back(c);
    }
    compute_route_centers(p, const_cast<int *>(tour), centers);
    for (int r = 0; r < R; ++r) {
        cx[r] = centers[r]->coord->x;
        cy[r] = centers[r]->coord->y;
        rdem[r] = centers[r]->coord->demand;
        delete centers[r]->coord;
        delete centers[r];
    }
    free(p->nodeptr);
    delete p;
}

// init_sub_problem (problem.cpp:84-130): the sub-problem of the master
// instance (x, y, dem, Q, D, svc, dist n*n) on real_nodes[0..sub_n); its
// distances, demands and neighbour lists (width sub_n-2, compute_nn_lists
// on the gathered distances).
void ref_sub_problem(int n, const float *x, const float *y, const int *dem, int Q, float D, float svc,
                     const float *dist, int sub_n, const int *real_nodes, float *sub_dist, int *sub_dem,
                     int *sub_nn)
{
    Problem *master = make_problem(n, x, y, dem, Q, D, svc, dist);
    Problem *sub = new Problem(1);
    sub->num_node = sub_n;
    for (int i = 0; i < sub_n; ++i) sub->real_nodes.push_back(real_nodes[i]);
    init_sub_problem(master, sub);
    const int nn = sub_n - 2;
    for (int i = 0; i < sub_n; ++i) {
        for (int j = 0; j < sub_n; ++j) sub_dist[(size_t)i * sub_n + j] = sub->distance[i][j];
        sub_dem[i] = sub->nodeptr[i].demand;
        if (sub_nn)
            for (int j = 0; j < nn; ++j) sub_nn[(size_t)i * nn + j] = sub->nn_list[i][j];
    }
}

}  // extern "C"
