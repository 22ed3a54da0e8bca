// Host-side instance preparation behind the C ABI: VRPLIB reading,
// distance matrices, neighbour lists, solution checking.  These run before
// the loop (call stack 1 of the reference: read_instance_file ->
// init_problem -> compute_distances / compute_nn_lists).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "aco_internal.h"

namespace cvrp {

// ---------------------------------------------------------- distances ----

static float ftrunc(float x) { return (float)(int)x; }

// TSPLIB GEO with the reference's float intermediates (vrpHelper.cpp:99-136).
static float geo_coord(float v)
{
    const float deg = ftrunc(v);
    const float mn = v - deg;
    return (float)(3.14159265358979323846264 * (deg + 5.0 * mn / 3.0) / 180.0);
}

// mode kDistFloat follows round_distance (vrpHelper.cpp:67-81, float
// sqrt, no rounding); kDistNint is VRPLIB's nint; CEIL/GEO/ATT follow
// vrpHelper.cpp:83-158.
float distance_of(int mode, double xi_, double yi_, double xj_, double yj_)
{
    const float xi = (float)xi_, yi = (float)yi_, xj = (float)xj_, yj = (float)yj_;
    switch (mode) {
    case kDistFloat: {
        const float xd = xi - xj, yd = yi - yj;
        return sqrtf(xd * xd + yd * yd);
    }
    case kDistNint: {
        const double xd = xi_ - xj_, yd = yi_ - yj_;
        return (float)(long long)(sqrt(xd * xd + yd * yd) + 0.5);
    }
    case kDistCeil: {
        const float xd = xi - xj, yd = yi - yj;
        return (float)(int)ceil(sqrtf(xd * xd + yd * yd));
    }
    case kDistGeo: {
        const float lati = geo_coord(xi), latj = geo_coord(xj);
        const float loni = geo_coord(yi), lonj = geo_coord(yj);
        const float q1 = (float)cos(loni - lonj), q2 = (float)cos(lati - latj), q3 = (float)cos(lati + latj);
        return (float)(int)(6378.388 * acos(0.5 * ((1.0 + q1) * q2 - (1.0 - q1) * q3)) + 1.0);
    }
    case kDistAtt: {
        const float xd = xi - xj, yd = yi - yj;
        const float r = (float)sqrt((xd * xd + yd * yd) / 10.0);
        const float t = ftrunc(r);
        return (float)(t < r ? (int)t + 1 : (int)t);
    }
    }
    return NAN;
}

// ------------------------------------------------------ neighbour lists ----

// The reference sorts each row with a middle-pivot quicksort whose left
// recursion keeps the pivot slot (utilities.cpp:246-267).  Its tie order is
// the order the 2-opt scans neighbours in, so the same partition steps are
// kept here (iterative, explicit stack).
static void quicksort_pairs(std::vector<float> &v, std::vector<int> &ix, int left, int right)
{
    std::vector<std::pair<int, int>> stack;
    stack.emplace_back(left, right);
    while (!stack.empty()) {
        auto [l, r] = stack.back();
        stack.pop_back();
        if (l >= r) continue;
        const int mid = (l + r) / 2;
        std::swap(v[l], v[mid]);
        std::swap(ix[l], ix[mid]);
        int last = l;
        for (int k = l + 1; k <= r; ++k)
            if (v[k] < v[l]) {
                ++last;
                std::swap(v[last], v[k]);
                std::swap(ix[last], ix[k]);
            }
        std::swap(v[l], v[last]);
        std::swap(ix[l], ix[last]);
        // the reference recurses left first: push right first
        stack.emplace_back(last + 1, r);
        stack.emplace_back(l, last);
    }
}

void nn_lists(int n, const float *dist, int nn, int tie, int *out)
{
    std::vector<float> v(n);
    std::vector<int> ix(n);
    for (int node = 0; node < n; ++node) {
        for (int i = 0; i < n; ++i) { v[i] = dist[(size_t)node * n + i]; ix[i] = i; }
        v[node] = INFINITY;
        v[0] = INFINITY;
        if (tie == 0) {
            quicksort_pairs(v, ix, 0, n - 1);
        } else {
            std::stable_sort(ix.begin(), ix.end(), [&](int a, int b) { return v[a] < v[b]; });
        }
        for (int i = 0; i < nn; ++i) out[(size_t)node * nn + i] = ix[i];
    }
}

// ------------------------------------------------------------- checks ----

int check_solution(int n, const float *dist, const int *dem, int Q, float D, float svc,
                   const int *tour, int size)
{
    // problem.cpp:226-319
    std::vector<char> used(n, 0);
    used[0] = 1;
    int rbeg = 0;
    auto route_ok = [&](int b, int e) {
        if (tour[b] != 0 || tour[e] != 0) return false;
        if (e - b < 2) return false;
        int load = 0;
        for (int i = b + 1; i < e; ++i) load += dem[tour[i]];
        if (load > Q) return false;
        float d = 0.0f;
        for (int i = b; i < e; ++i) d += dist[(size_t)tour[i] * n + tour[i + 1]];
        d += svc * (float)(e - b - 1);
        return !(d > D);
    };
    for (int i = 1; i < size; ++i) {
        const int v = tour[i];
        if (v < 0 || v >= n) return 0;
        if (v != 0) {
            if (used[v]) return 0;
            used[v] = 1;
        } else {
            if (!route_ok(rbeg, i)) return 0;
            rbeg = i;
        }
    }
    for (int i = 0; i < n; ++i)
        if (!used[i]) return 0;
    return 1;
}

// Greedy nearest feasible customer, ties to the lowest index (tau0 rule):
// from the current node take the unvisited customer with the smallest
// d(cur, i) among those that keep load <= Q and rdist + (d(cur,i) + svc) +
// d(i,0) <= D; when none fits, return to the depot.  The length adds the
// edges in step order.  When tour != nullptr the node sequence (depot
// first and last) is written there, 2n entries at most.  Returns the tour
// size, or -1 when some customer can never be served.
int nn_tour(int n, const float *dist, const int *dem, int Q, float D, float svc, int *tour, float *length)
{
    std::vector<char> vis(n, 0);
    vis[0] = 1;
    int cur = 0, load = 0, left = n - 1, size = 1;
    float rdist = 0.0f, len = 0.0f;
    if (tour) tour[0] = 0;
    while (left > 0) {
        int best = -1;
        float bd = INFINITY;
        for (int i = 1; i < n; ++i) {
            if (vis[i] || load + dem[i] > Q) continue;
            float t = rdist + (dist[(size_t)cur * n + i] + svc);
            t = t + dist[(size_t)i * n];
            if (!(t <= D)) continue;
            if (dist[(size_t)cur * n + i] < bd) { bd = dist[(size_t)cur * n + i]; best = i; }
        }
        int w;
        if (best < 0) {
            if (cur == 0) return -1;
            w = 0; load = 0; rdist = 0.0f;
        } else {
            w = best; load += dem[w]; rdist = rdist + (dist[(size_t)cur * n + w] + svc); left--;
        }
        vis[w] = 1;
        len = len + dist[(size_t)cur * n + w];
        if (tour) tour[size] = w;
        ++size;
        cur = w;
    }
    len = len + dist[(size_t)cur * n];
    if (tour) tour[size] = 0;
    *length = len;
    return size + 1;
}

float nn_tour_length(int n, const float *dist, const int *dem, int Q, float D, float svc)
{
    float len = -1.0f;
    return nn_tour(n, dist, dem, Q, D, svc, nullptr, &len) < 0 ? -1.0f : len;
}

// Tournament order of the reference's argmax loop (kernel.cl:340-366):
// slot i absorbs slot i+offset with ties kept by slot i, an odd offset first
// folds slot offset-1 into slot 0.  rank[i] = position of node i in the
// order in which tied leaves win.
void tree_rank(int N, int *rank)
{
    std::vector<int> head(N), tail(N), next(N, -1);
    for (int i = 0; i < N; ++i) head[i] = tail[i] = i;
    auto append = [&](int a, int b) { next[tail[a]] = head[b]; tail[a] = tail[b]; };
    for (int off = N; off > 0;) {
        if ((off & 1) && off - 1 != 0) append(0, off - 1);
        off /= 2;
        for (int i = 0; i < off; ++i) append(i, i + off);
    }
    int r = 0;
    for (int x = head[0]; x >= 0; x = next[x]) rank[x] = r++;
}

// ------------------------------------------------------------- VRPLIB ----

// read_instance_file (io.cpp:334-517): "KEY : value", "KEY: value" and
// "KEY value" headers, NODE_COORD_SECTION and DEMAND_SECTION read
// positionally (node ids ignored, node 0 is the depot).
int read_vrplib(const char *path, VrpInstance &inst, std::string &err)
{
    std::ifstream f(path);
    if (!f) { err = std::string("cannot open ") + path; return -1; }
    std::string tok;
    inst = VrpInstance();
    auto next_tok = [&](std::string &t) { return static_cast<bool>(f >> t); };
    auto strip = [](std::string s) {
        if (!s.empty() && s.back() == ':') s.pop_back();
        return s;
    };
    auto value_tok = [&](std::string &v) {
        if (!next_tok(v)) return false;
        if (v == ":" && !next_tok(v)) return false;
        return true;
    };
    bool have_coords = false;
    while (next_tok(tok)) {
        const bool colon = !tok.empty() && tok.back() == ':';
        const std::string key = strip(tok);
        std::string v;
        if (key == "NODE_COORD_SECTION") {
            if (inst.n <= 0) { err = "DIMENSION missing before NODE_COORD_SECTION"; return -1; }
            inst.x.resize(inst.n);
            inst.y.resize(inst.n);
            for (int i = 0; i < inst.n; ++i) {
                std::string id;
                if (!(f >> id >> inst.x[i] >> inst.y[i])) { err = "truncated NODE_COORD_SECTION"; return -1; }
            }
            have_coords = true;
        } else if (key == "DEMAND_SECTION") {
            inst.demand.resize(inst.n);
            for (int i = 0; i < inst.n; ++i) {
                std::string id;
                if (!(f >> id >> inst.demand[i])) { err = "truncated DEMAND_SECTION"; return -1; }
            }
        } else if (key == "NAME") {
            if (value_tok(v)) inst.name = v;
        } else if (key == "COMMENT") {
            // the reference scans a float right after the keyword token
            if (next_tok(v)) {
                char *end = nullptr;
                const double d = strtod(v.c_str(), &end);
                if (end && *end == '\0' && end != v.c_str()) inst.optimum = (float)d;
            }
        } else if (key == "TYPE") {
            if (!value_tok(v) || v != "CVRP") { err = "not a CVRP instance"; return -1; }
        } else if (key == "DIMENSION") {
            if (value_tok(v)) inst.n = atoi(v.c_str());
        } else if (key == "CAPACITY") {
            if (value_tok(v)) inst.capacity = atoi(v.c_str());
        } else if (key == "DISTANCE") {
            if (value_tok(v)) inst.max_dist = (float)atof(v.c_str());
        } else if (key == "SERVICE_TIME") {
            if (value_tok(v)) inst.service = (float)atof(v.c_str());
        } else if (key == "EDGE_WEIGHT_TYPE") {
            if (!value_tok(v)) { err = "EDGE_WEIGHT_TYPE without value"; return -1; }
            if (v == "EUC_2D") inst.edge_type = 0;
            else if (v == "CEIL_2D") inst.edge_type = 1;
            else if (v == "GEO") inst.edge_type = 2;
            else if (v == "ATT") inst.edge_type = 3;
            else { err = "EDGE_WEIGHT_TYPE " + v + " not implemented"; return -1; }
        } else if (key == "EOF") {
            break;
        }
        (void)colon;
    }
    if (!have_coords || (int)inst.demand.size() != inst.n || inst.n < 2) {
        err = "instance lacks coordinates or demands";
        return -1;
    }
    return 0;
}

}  // namespace cvrp

using namespace cvrp;

extern "C" {

int aco_compute_distances(int32_t n, const double *x, const double *y, int32_t mode, float *dist)
{
    if (n < 1 || !x || !y || !dist || mode < 0 || mode > kDistAtt)
        return set_error(ACO_ERR_ARG, "aco_compute_distances: bad argument");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) dist[(size_t)i * n + j] = distance_of(mode, x[i], y[i], x[j], y[j]);
    return ACO_OK;
}

int aco_compute_nn_lists(int32_t n, const float *dist, int32_t nn, int32_t tie, int32_t *out)
{
    if (n < 3 || !dist || !out || nn < 1 || nn > n - 2 || (tie != 0 && tie != 1))
        return set_error(ACO_ERR_ARG, "aco_compute_nn_lists: need n >= 3 and 1 <= nn <= n-2");
    nn_lists(n, dist, nn, tie, out);
    return ACO_OK;
}

int aco_read_vrplib(const char *path, int32_t capacity, int32_t *n, double *x, double *y,
                    int32_t *demand, int32_t *vehicle_capacity, float *max_dist,
                    float *service_time, float *optimum, int32_t *edge_weight_type)
{
    if (!path || !n) return set_error(ACO_ERR_ARG, "aco_read_vrplib: null argument");
    VrpInstance inst;
    std::string err;
    if (read_vrplib(path, inst, err) != 0) return set_error(ACO_ERR_IO, err);
    *n = inst.n;
    if (capacity < inst.n) return ACO_OK;
    for (int i = 0; i < inst.n; ++i) {
        if (x) x[i] = inst.x[i];
        if (y) y[i] = inst.y[i];
        if (demand) demand[i] = inst.demand[i];
    }
    if (vehicle_capacity) *vehicle_capacity = inst.capacity;
    if (max_dist) *max_dist = inst.max_dist;
    if (service_time) *service_time = inst.service;
    if (optimum) *optimum = inst.optimum;
    if (edge_weight_type) *edge_weight_type = inst.edge_type;
    return ACO_OK;
}

int aco_nn_tour(int32_t n, const float *dist, const int32_t *demand, int32_t capacity, float max_dist,
                float service_time, int32_t *tour, int32_t tour_capacity, int32_t *size, float *length)
{
    if (n < 2 || !dist || !demand || !length || capacity <= 0 || (tour && tour_capacity < 2 * n))
        return set_error(ACO_ERR_ARG, "aco_nn_tour: bad argument (tour needs 2n entries)");
    float len = 0.0f;
    const int sz = nn_tour(n, dist, demand, capacity, max_dist, service_time, tour, &len);
    if (sz < 0) return set_error(ACO_ERR_INFEASIBLE, "aco_nn_tour: a customer cannot be served");
    if (size) *size = sz;
    *length = len;
    return ACO_OK;
}

int aco_check_solution(int32_t n, const float *dist, const int32_t *demand, int32_t capacity,
                       float max_dist, float service_time, const int32_t *tour, int32_t size)
{
    if (n < 2 || !dist || !demand || !tour || size < 1)
        return set_error(ACO_ERR_ARG, "aco_check_solution: bad argument");
    return check_solution(n, dist, demand, capacity, max_dist, service_time, tour, size);
}

}  // extern "C"
