// Sub-problem decomposition of a colony's best solution: route centres
// (compute_route_centers, vrpHelper.cpp:274-294), the partition of routes
// into sub-problems, the sub-problem set-up (init_sub_problem,
// problem.cpp:84-130), the merge of the sub-colonies' best tours and the
// driver of one round, which solves all sub-problems of a round in one
// batched context.  Host code over the public C ABI; the device work is the
// batched sub-colony run and aco_pheromone_from_subs.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "aco_internal.h"
#include "cvrp_aco.h"
#include "cvrp_decompose.h"

namespace {

using cvrp::set_error;

struct Route {
    int beg, end, demand;
    float cx, cy;
};

// Routes of a depot-separated tour with at least one customer, centres as
// compute_route_centers computes them: a float Point accumulating, per node
// of tour[beg .. end) (the opening depot included, demand 0), the double
// term x*d*1.0/D.
std::vector<Route> routes_of(const float *x, const float *y, const int32_t *demand, const int32_t *tour,
                             int size)
{
    std::vector<Route> out;
    int open = 0;
    for (int q = 1; q < size; ++q) {
        if (tour[q] != 0) continue;
        if (q > open + 1) {
            Route r{open, q, 0, 0.0f, 0.0f};
            for (int j = open; j < q; ++j) r.demand += demand[tour[j]];
            const double D = (double)r.demand;
            for (int j = open; j < q; ++j) {
                const int v = tour[j];
                const float wx = x[v] * (float)demand[v], wy = y[v] * (float)demand[v];
                r.cx = (float)((double)r.cx + (double)wx * 1.0 / D);
                r.cy = (float)((double)r.cy + (double)wy * 1.0 / D);
            }
            out.push_back(r);
        }
        open = q;
    }
    return out;
}

int check_tour(int n, const int32_t *tour, int size, const char *who)
{
    if (!tour || size < 3 || tour[0] != 0 || tour[size - 1] != 0)
        return set_error(ACO_ERR_ARG, std::string(who) + ": tour must start and end at the depot");
    for (int q = 0; q < size; ++q)
        if (tour[q] < 0 || tour[q] >= n) return set_error(ACO_ERR_ARG, std::string(who) + ": tour node out of range");
    return ACO_OK;
}

// edges of t (local numbering through real) added in step order
float mapped_length(int n, const float *dist, const int32_t *real, const int32_t *t, int size)
{
    float len = 0.0f;
    for (int q = 1; q < size; ++q) len += dist[(size_t)real[t[q - 1]] * n + real[t[q]]];
    return len;
}

}  // namespace

extern "C" {

int aco_route_centers(int32_t n, const float *x, const float *y, const int32_t *demand, const int32_t *tour,
                      int32_t size, int32_t capacity, int32_t *beg, int32_t *end, float *cx, float *cy,
                      int32_t *route_demand, int32_t *count)
{
    if (n < 2 || !x || !y || !demand || !count) return set_error(ACO_ERR_ARG, "aco_route_centers: bad argument");
    int rc = check_tour(n, tour, size, "aco_route_centers");
    if (rc) return rc;
    const std::vector<Route> rs = routes_of(x, y, demand, tour, size);
    *count = (int32_t)rs.size();
    if ((int)rs.size() > capacity) return set_error(ACO_ERR_ARG, "aco_route_centers: capacity below the route count");
    for (size_t r = 0; r < rs.size(); ++r) {
        if (beg) beg[r] = rs[r].beg;
        if (end) end[r] = rs[r].end;
        if (cx) cx[r] = rs[r].cx;
        if (cy) cy[r] = rs[r].cy;
        if (route_demand) route_demand[r] = rs[r].demand;
    }
    return ACO_OK;
}

int aco_partition_routes(int32_t n, const float *x, const float *y, const int32_t *demand, const int32_t *tour,
                         int32_t size, int32_t subs, int32_t rotation, int32_t *offsets, int32_t *real_nodes,
                         int32_t *tour_offsets, int32_t *sub_tours, int32_t *count)
{
    if (n < 2 || !x || !y || !demand || !offsets || !real_nodes || !tour_offsets || !sub_tours || !count)
        return set_error(ACO_ERR_ARG, "aco_partition_routes: bad argument");
    int rc = check_tour(n, tour, size, "aco_partition_routes");
    if (rc) return rc;
    const std::vector<Route> rs = routes_of(x, y, demand, tour, size);
    const int R = (int)rs.size();
    if (R == 0) return set_error(ACO_ERR_ARG, "aco_partition_routes: tour without customers");
    // circular order of the routes by the angle of their centre
    std::vector<std::pair<double, int>> order(R);
    for (int r = 0; r < R; ++r) {
        double px = rs[r].cx, py = rs[r].cy;
        if (rs[r].demand == 0) {
            double sx = 0.0, sy = 0.0;
            for (int j = rs[r].beg + 1; j < rs[r].end; ++j) {
                sx += x[tour[j]];
                sy += y[tour[j]];
            }
            const double c = (double)(rs[r].end - rs[r].beg - 1);
            px = sx / c;
            py = sy / c;
        }
        order[r] = {atan2(py - (double)y[0], px - (double)x[0]), r};
    }
    std::sort(order.begin(), order.end());
    // default: n/50 (problem.cpp:218), at most R/2 so that every
    // sub-problem holds two routes or more (customers can change routes)
    int G = subs > 0 ? subs : std::min(n / 50, R / 2);
    G = std::max(1, std::min(G, R));
    long long customers = 0;
    for (const Route &r : rs) customers += r.end - r.beg - 1;
    const int first = ((rotation % R) + R) % R;
    std::vector<int> run(R);   // run of the q-th route of the cut
    int g = 0;
    long long acc = 0;
    for (int q = 0; q < R; ++q) {
        const Route &r = rs[order[(first + q) % R].second];
        run[q] = g;
        acc += r.end - r.beg - 1;
        const int routes_left = R - q - 1, runs_left = G - 1 - g;
        if (runs_left > 0 && (acc * G >= (long long)(g + 1) * customers || routes_left == runs_left)) ++g;
    }
    int no = 0, to = 0, q = 0;
    for (int s = 0; s < G; ++s) {
        offsets[s] = no;
        tour_offsets[s] = to;
        real_nodes[no++] = 0;
        sub_tours[to++] = 0;
        int local = 0;
        for (; q < R && run[q] == s; ++q) {
            const Route &r = rs[order[(first + q) % R].second];
            for (int j = r.beg + 1; j < r.end; ++j) {
                real_nodes[no++] = tour[j];
                sub_tours[to++] = ++local;
            }
            sub_tours[to++] = 0;
        }
    }
    offsets[G] = no;
    tour_offsets[G] = to;
    *count = G;
    return ACO_OK;
}

int aco_sub_problem(int32_t n, const float *dist, const int32_t *demand, int32_t sub_n, const int32_t *real_nodes,
                    float *sub_dist, int32_t *sub_demand)
{
    if (n < 2 || !dist || !demand || sub_n < 1 || sub_n > n || !real_nodes || !sub_dist || !sub_demand)
        return set_error(ACO_ERR_ARG, "aco_sub_problem: bad argument");
    if (real_nodes[0] != 0) return set_error(ACO_ERR_ARG, "aco_sub_problem: real_nodes[0] must be the depot");
    std::vector<char> seen(n, 0);
    for (int i = 0; i < sub_n; ++i) {
        const int v = real_nodes[i];
        if (v < 0 || v >= n || seen[v]) return set_error(ACO_ERR_ARG, "aco_sub_problem: real_nodes out of range or repeated");
        seen[v] = 1;
    }
    for (int i = 0; i < sub_n; ++i) {
        const float *row = dist + (size_t)real_nodes[i] * n;
        float *o = sub_dist + (size_t)i * sub_n;
        for (int j = 0; j < sub_n; ++j) o[j] = row[real_nodes[j]];
        sub_demand[i] = demand[real_nodes[i]];
    }
    return ACO_OK;
}

int aco_merge_sub_tours(int32_t n, const float *dist, int32_t count, const int32_t *offsets,
                        const int32_t *real_nodes, const int32_t *tour_offsets, const int32_t *sub_tours,
                        const int32_t *best_offsets, const int32_t *best_tours, int32_t *out, int32_t capacity,
                        int32_t *size, float *length, int32_t *replaced)
{
    if (n < 2 || !dist || count < 1 || !offsets || !real_nodes || !tour_offsets || !sub_tours || !best_offsets ||
        !out || !size || !length)
        return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: bad argument");
    int o = 0;
    auto put = [&](int v) {
        if (o < capacity) out[o] = v;
        ++o;
    };
    put(0);
    for (int s = 0; s < count; ++s) {
        const int sn = offsets[s + 1] - offsets[s];
        const int32_t *real = real_nodes + offsets[s];
        const int32_t *orig = sub_tours + tour_offsets[s];
        const int osz = tour_offsets[s + 1] - tour_offsets[s];
        const int bsz = best_offsets[s + 1] - best_offsets[s];
        const int32_t *best = bsz > 0 ? best_tours + best_offsets[s] : nullptr;
        if (sn < 1 || osz < 2) return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: empty sub-problem");
        if (real[0] != 0) return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: node lists start at the depot");
        for (int i = 1; i < sn; ++i)
            if (real[i] <= 0 || real[i] >= n) return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: node list entry out of range");
        for (int q = 0; q < osz; ++q)
            if (orig[q] < 0 || orig[q] >= sn) return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: sub tour out of range");
        for (int q = 0; q < bsz; ++q)
            if (!best || best[q] < 0 || best[q] >= sn)
                return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: best tour out of range");
        if (bsz > 0 && (bsz < 3 || best[0] != 0 || best[bsz - 1] != 0))
            return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: best tours start and end at the depot");
        const bool take = bsz > 0 && mapped_length(n, dist, real, best, bsz) < mapped_length(n, dist, real, orig, osz);
        if (replaced) replaced[s] = take ? 1 : 0;
        const int32_t *t = take ? best : orig;
        const int sz = take ? bsz : osz;
        for (int q = 1; q < sz; ++q) put(real[t[q]]);
    }
    *size = o;
    if (o > capacity) return set_error(ACO_ERR_ARG, "aco_merge_sub_tours: capacity below the merged size");
    float len = 0.0f;
    for (int q = 1; q < o; ++q) len += dist[(size_t)out[q - 1] * n + out[q]];
    *length = len;
    return ACO_OK;
}

int aco_decompose_round(aco_ctx *master, int32_t b, const aco_problem *pb, const float *x, const float *y,
                        const aco_decompose_params *prm, int32_t *out, int32_t capacity, int32_t *size,
                        float *length, int32_t *subs_out, int32_t *replaced_out)
{
    if (!master || !pb || !pb->dist || !pb->demand || !x || !y || !prm || !out || !size || !length)
        return set_error(ACO_ERR_ARG, "aco_decompose_round: bad argument");
    if (prm->sub_params.n_ants < 1) return set_error(ACO_ERR_ARG, "aco_decompose_round: sub_params.n_ants must be > 0");
    const int n = pb->n;
    // the master's best-so-far
    std::vector<int32_t> bsf(2 * (size_t)n + 2);
    int32_t bsz = 0;
    float blen = 0.0f;
    int rc = aco_get_best_at(master, b, bsf.data(), (int32_t)bsf.size(), &bsz, &blen);
    if (rc) return rc;
    if (bsz < 3) return set_error(ACO_ERR_STATE, "aco_decompose_round: the master has no best-so-far tour yet");
    // partition
    std::vector<int32_t> off(bsz + 1), real((size_t)n + bsz), toff(bsz + 1), st(2 * (size_t)bsz);
    int32_t G = 0;
    rc = aco_partition_routes(n, x, y, pb->demand, bsf.data(), bsz, prm->subs, prm->rotation, off.data(), real.data(),
                              toff.data(), st.data(), &G);
    if (rc) return rc;
    // sub-problems with at least two customers are solved
    std::vector<int> solved;
    for (int s = 0; s < G; ++s)
        if (off[s + 1] - off[s] >= 3) solved.push_back(s);
    std::vector<int32_t> boff(G + 1, 0), best;
    if (!solved.empty()) {
        std::vector<std::vector<float>> sd(solved.size());
        std::vector<std::vector<int32_t>> sdem(solved.size());
        std::vector<aco_problem> subs(solved.size());
        int min_n = n;
        for (size_t i = 0; i < solved.size(); ++i) {
            const int s = solved[i], sn = off[s + 1] - off[s];
            sd[i].resize((size_t)sn * sn);
            sdem[i].resize(sn);
            rc = aco_sub_problem(n, pb->dist, pb->demand, sn, &real[off[s]], sd[i].data(), sdem[i].data());
            if (rc) return rc;
            aco_problem &q = subs[i];
            memset(&q, 0, sizeof(q));
            q.n = sn;
            q.dist = sd[i].data();
            q.demand = sdem[i].data();
            q.capacity = pb->capacity;
            q.max_dist = pb->max_dist;
            q.service_time = pb->service_time;
            min_n = std::min(min_n, sn);
        }
        aco_params sp = prm->sub_params;
        if (sp.selection == ACO_SEL_KNN) sp.k = std::max(1, std::min(sp.k, min_n - 2));
        if (sp.nn_ls > min_n - 2) sp.nn_ls = 0;
        aco_ctx *sc = nullptr;
        rc = aco_create_batch(subs.data(), (int32_t)subs.size(), &sp, &sc);
        if (rc) return rc;
        rc = aco_init(sc);
        if (!rc) rc = aco_run(sc, prm->sub_iterations > 0 ? prm->sub_iterations : 75);
        std::vector<int32_t> map(solved.size());
        for (size_t i = 0; i < solved.size() && !rc; ++i) {
            const int s = solved[i], sn = off[s + 1] - off[s];
            std::vector<int32_t> t(2 * (size_t)sn + 2);
            int32_t tsz = 0;
            float tl = 0.0f;
            rc = aco_get_best_at(sc, (int32_t)i, t.data(), (int32_t)t.size(), &tsz, &tl);
            if (rc) break;
            boff[s + 1] = tsz;   // counts; prefix sums below
            best.insert(best.end(), t.begin(), t.begin() + tsz);
            map[i] = s;
        }
        if (!rc && prm->pheromone_back) rc = aco_pheromone_from_subs(master, b, sc, off.data(), real.data(), map.data());
        aco_destroy(sc);
        if (rc) return rc;
    }
    for (int s = 0; s < G; ++s) boff[s + 1] += boff[s];
    std::vector<int32_t> rep(G, 0);
    rc = aco_merge_sub_tours(n, pb->dist, G, off.data(), real.data(), toff.data(), st.data(), boff.data(),
                             best.empty() ? nullptr : best.data(), out, capacity, size, length, rep.data());
    if (rc) return rc;
    if (subs_out) *subs_out = G;
    if (replaced_out) {
        int r = 0;
        for (int v : rep) r += v;
        *replaced_out = r;
    }
    return ACO_OK;
}

}  // extern "C"
