// Internal declarations shared by the host-side translation units of
// libcvrpaco.so (not part of the public C ABI).
#pragma once

#include <math.h>

#include <string>
#include <vector>

#include "cvrp_aco.h"

namespace cvrp {

enum { kDistFloat = 0, kDistNint = 1, kDistCeil = 2, kDistGeo = 3, kDistAtt = 4 };

struct VrpInstance {
    std::string name;
    int n = 0;
    std::vector<double> x, y;
    std::vector<int> demand;
    int capacity = 0;
    float max_dist = INFINITY;
    float service = 0.0f;
    float optimum = 0.0f;
    int edge_type = 0;
};

int set_error(int code, const std::string &msg);
float distance_of(int mode, double xi, double yi, double xj, double yj);
void nn_lists(int n, const float *dist, int nn, int tie, int *out);
int check_solution(int n, const float *dist, const int *dem, int Q, float D, float svc,
                   const int *tour, int size);
int nn_tour(int n, const float *dist, const int *dem, int Q, float D, float svc, int *tour, float *length);
float nn_tour_length(int n, const float *dist, const int *dem, int Q, float D, float svc);
void tree_rank(int N, int *rank);
int read_vrplib(const char *path, VrpInstance &inst, std::string &err);

// Device view of one instance of a context, for aco_prior_apply.
struct CtxInstanceView {
    int n, P, Q;
    const float *dist;         // n x P
    const int *dem;            // n
    const int *nnk;            // n x k_ctx (construction candidates), or null
    int k_ctx;
    const int *nnls;           // n x nnP (2-opt neighbour rows)
    int nnP, nn_ls;
    float *heur;               // n x P
};
// Fails unless the context exists and aco_init has not run yet.
int ctx_prior_view(aco_ctx *c, std::vector<CtxInstanceView> &out, void **stream);

}  // namespace cvrp
