// C ABI of libcvrpaco.so: context lifetime, validation, device memory,
// launch sequencing (eager or one hipGraph per iteration), readback.
//
// Mirrors the reference host driver g_ACO (cvrp_ocl/gpu/g_aco.cpp); see
// include/cvrp_aco.h for the per-function mapping.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "aco_device.h"
#include "aco_internal.h"

namespace cvrp {

static thread_local std::string g_err;

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

static uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Lane seed: 1 + splitmix64(seed + golden * (1 + 64*ant + lane)) mod (2^31-2).
// (The reference derives lane seeds from time(NULL), g_aco.cpp:715-719.)
static int32_t pm_seed(uint64_t seed, int ant, int lane)
{
    const uint64_t z = splitmix64(seed + 0x9E3779B97F4A7C15ull *
                                  (uint64_t)(1 + (uint64_t)ant * kLanes + (uint64_t)lane));
    return (int32_t)(1 + z % 2147483646ull);
}

}  // namespace cvrp

using namespace cvrp;

struct ProfSlot {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct aco_ctx {
    aco_params prm{};
    int n = 0, P = 0, m = 0, T = 0, k = 0, nn_ls = 0, device = 0;
    DevProblem dp{};
    DevState ds{};
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool inited = false;
    float tick_to_s = 0.0f;
    std::vector<float> h_dist;
    std::vector<int> h_dem;
    int Q = 0;
    float D = INFINITY, svc = 0.0f;
    bool profiling = false;
    std::map<std::string, ProfSlot> prof;
};

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(ACO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static hipError_t dev_alloc(aco_ctx *c, T **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        c->allocs.push_back(q);
        *p = static_cast<T *>(q);
    }
    return e;
}

// upload an n x n host matrix into an n x P device matrix
static hipError_t upload_pitched(aco_ctx *c, float *dst, const float *src)
{
    return hipMemcpy2DAsync(dst, (size_t)c->P * 4, src, (size_t)c->n * 4, (size_t)c->n * 4, c->n,
                            hipMemcpyHostToDevice, c->stream);
}

static hipError_t download_pitched(aco_ctx *c, float *dst, const float *src)
{
    return hipMemcpy2DAsync(dst, (size_t)c->n * 4, src, (size_t)c->P * 4, (size_t)c->n * 4, c->n,
                            hipMemcpyDeviceToHost, c->stream);
}

template <typename F>
static hipError_t timed(aco_ctx *c, const char *name, F &&fn)
{
    if (!c->profiling) return fn();
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) return e;
    (void)hipEventRecord(a, c->stream);
    e = fn();
    (void)hipEventRecord(b, c->stream);
    c->prof[name].pending.emplace_back(a, b);
    return e;
}

static int run_construct(aco_ctx *c)
{
    HIP_TRY(timed(c, "construct", [&] {
        return launch_construct(c->dp, c->ds, c->prm.selection, c->prm.rng, c->prm.seed, c->stream);
    }));
    return ACO_OK;
}

static int run_local_search(aco_ctx *c)
{
    if (!c->prm.ls_flags) return ACO_OK;
    HIP_TRY(timed(c, "local_search",
                  [&] { return launch_local_search(c->dp, c->ds, c->prm.ls_flags, c->stream); }));
    return ACO_OK;
}

static int run_statistics(aco_ctx *c)
{
    HIP_TRY(timed(c, "statistics",
                  [&] { return launch_statistics(c->dp, c->ds, c->tick_to_s, c->stream); }));
    return ACO_OK;
}

static int run_gather(aco_ctx *c)
{
    if (c->k <= 0) return ACO_OK;
    HIP_TRY(timed(c, "gather", [&] { return launch_gather(c->dp, c->ds, c->stream); }));
    return ACO_OK;
}

static int run_pheromone(aco_ctx *c)
{
    HIP_TRY(timed(c, "pheromone", [&] { return launch_pheromone_update(c->dp, c->ds, c->stream); }));
    return run_gather(c);
}

static int run_iteration_steps(aco_ctx *c)
{
    int rc;
    if ((rc = run_construct(c)) != ACO_OK) return rc;
    if ((rc = run_local_search(c)) != ACO_OK) return rc;
    if ((rc = run_statistics(c)) != ACO_OK) return rc;
    return run_pheromone(c);
}

static int need(aco_ctx *c, bool inited)
{
    if (!c) return set_error(ACO_ERR_ARG, "null context");
    if (inited && !c->inited) return set_error(ACO_ERR_STATE, "aco_init has not been called");
    HIP_TRY(hipSetDevice(c->device));
    return ACO_OK;
}

extern "C" {

const char *aco_last_error(void) { return g_err.c_str(); }

void aco_default_params(aco_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->alpha = 1.0f;
    p->beta = 2.0f;
    p->rho = 0.1f;
    p->n_ants = 0;
    p->ras_ranks = 6;
    p->k = 0;
    p->nn_ls = 0;
    p->selection = ACO_SEL_IROULETTE;
    p->rng = ACO_RNG_PARK_MILLER;
    p->ls_flags = 0;
    p->tau0_rule = ACO_TAU0_FIRST_COLONY;
    p->seed = 1;
    p->bsf_capacity = 4096;
    p->use_graph = 1;
}

int aco_create(const aco_problem *pb, const aco_params *prm_in, aco_ctx **out)
{
    if (!pb || !prm_in || !out) return set_error(ACO_ERR_ARG, "aco_create: null argument");
    *out = nullptr;
    aco_params prm = *prm_in;
    const int n = pb->n;
    if (n < 3 || n > kMaxNodes) return set_error(ACO_ERR_ARG, "aco_create: need 3 <= n <= 4096");
    if (!pb->dist || !pb->demand) return set_error(ACO_ERR_ARG, "aco_create: dist and demand required");
    if (prm.n_ants <= 0) prm.n_ants = n;
    if (prm.nn_ls <= 0) prm.nn_ls = n - 2;
    if (prm.bsf_capacity <= 0) prm.bsf_capacity = 4096;
    if (prm.ras_ranks < 1 || prm.ras_ranks > kMaxRas) return set_error(ACO_ERR_ARG, "ras_ranks must be in [1,16]");
    if (!(prm.rho > 0.0f && prm.rho <= 1.0f)) return set_error(ACO_ERR_ARG, "rho must be in (0,1]");
    if (!isfinite(prm.alpha) || !isfinite(prm.beta)) return set_error(ACO_ERR_ARG, "alpha/beta not finite");
    if (prm.selection != ACO_SEL_IROULETTE && prm.selection != ACO_SEL_KNN)
        return set_error(ACO_ERR_ARG, "unknown selection rule");
    if (prm.rng != ACO_RNG_PARK_MILLER && prm.rng != ACO_RNG_PHILOX) return set_error(ACO_ERR_ARG, "unknown rng");
    if (prm.selection == ACO_SEL_IROULETTE && prm.rng != ACO_RNG_PARK_MILLER)
        return set_error(ACO_ERR_ARG, "iroulette draws from Park-Miller lane streams");
    if (prm.selection == ACO_SEL_KNN && (prm.k < 1 || prm.k > kLanes || prm.k > n - 2))
        return set_error(ACO_ERR_ARG, "k must be in [1, min(64, n-2)] for roulette_knn");
    if (prm.selection != ACO_SEL_KNN) prm.k = 0;
    if (prm.nn_ls > n - 2) return set_error(ACO_ERR_ARG, "nn_ls must be <= n-2");
    if (prm.ls_flags & ~7) return set_error(ACO_ERR_ARG, "unknown local search flag");
    if (prm.tau0_rule != ACO_TAU0_FIRST_COLONY && prm.tau0_rule != ACO_TAU0_NN_TOUR)
        return set_error(ACO_ERR_ARG, "unknown tau0 rule");
    if (pb->capacity <= 0) return set_error(ACO_ERR_ARG, "capacity must be positive");
    if (isnan(pb->max_dist) || !(pb->service_time >= 0.0f) || !isfinite(pb->service_time))
        return set_error(ACO_ERR_ARG, "bad max_dist or service_time");

    // instance checks: finite, non-negative, symmetric distances; every
    // customer servable by a vehicle of its own (otherwise construction
    // could never visit it: kernel.cl:438-479 would not terminate)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const float d = pb->dist[(size_t)i * n + j];
            if (!isfinite(d) || d < 0.0f) return set_error(ACO_ERR_ARG, "distances must be finite and >= 0");
            if (d != pb->dist[(size_t)j * n + i]) return set_error(ACO_ERR_ARG, "distance matrix is not symmetric");
        }
    for (int i = 0; i < n; ++i)
        if (pb->demand[i] < 0) return set_error(ACO_ERR_ARG, "negative demand");
    for (int i = 1; i < n; ++i) {
        if (pb->demand[i] > pb->capacity)
            return set_error(ACO_ERR_INFEASIBLE, "customer " + std::to_string(i) + " demand exceeds capacity");
        float t = 0.0f + (pb->dist[i] + pb->service_time);
        t = t + pb->dist[(size_t)i * n];
        if (!(t <= pb->max_dist))
            return set_error(ACO_ERR_INFEASIBLE, "customer " + std::to_string(i) + " beyond max_dist");
    }

    // neighbour rows: given (checked sorted) or built with the reference order
    const int width = std::max(prm.k, prm.nn_ls);
    std::vector<int> nnl((size_t)n * width);
    if (pb->nn_list) {
        if (pb->nn_width < width) return set_error(ACO_ERR_ARG, "nn_width smaller than max(k, nn_ls)");
        for (int i = 0; i < n; ++i)
            for (int h = 0; h < width; ++h) {
                const int v = pb->nn_list[(size_t)i * pb->nn_width + h];
                if (v < 1 || v >= n || v == i) return set_error(ACO_ERR_ARG, "nn_list entry out of range");
                if (h > 0 && pb->dist[(size_t)i * n + v] <
                                 pb->dist[(size_t)i * n + nnl[(size_t)i * width + h - 1]])
                    return set_error(ACO_ERR_ARG, "nn_list rows must be sorted by distance");
                nnl[(size_t)i * width + h] = v;
            }
    } else {
        nn_lists(n, pb->dist, width, 0, nnl.data());
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(ACO_ERR_NODEV, "no HIP device");

    aco_ctx *c = new aco_ctx();
    c->prm = prm;
    c->n = n;
    c->P = (n + 3) & ~3;
    c->m = prm.n_ants;
    c->T = 2 * n;
    c->k = prm.k;
    c->nn_ls = prm.nn_ls;
    c->h_dist.assign(pb->dist, pb->dist + (size_t)n * n);
    c->h_dem.assign(pb->demand, pb->demand + n);
    c->Q = pb->capacity;
    c->D = pb->max_dist;
    c->svc = pb->service_time;
    auto fail = [&](int code) {
        aco_destroy(c);
        return code;
    };
#define CHK(expr)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(set_error(ACO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    CHK(hipGetDevice(&c->device));
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, c->device));
    const size_t lds_need = std::max(construct_lds_bytes(n, prm.selection),
                                     prm.ls_flags ? local_search_lds_bytes(n) : 0);
    if (lds_need > prop.sharedMemPerBlock)
        return fail(set_error(ACO_ERR_ARG, "instance too large for LDS: needs " + std::to_string(lds_need) +
                                               " bytes, device has " + std::to_string(prop.sharedMemPerBlock)));
    int rate_khz = 0;
    CHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
    c->tick_to_s = rate_khz > 0 ? 1.0f / (1000.0f * (float)rate_khz) : 0.0f;
    CHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    const int m = c->m, P = c->P, T = c->T, ras = prm.ras_ranks;
    // host-side derived arrays
    std::vector<float> heur((size_t)n * n), d0(n);
    for (size_t e = 0; e < (size_t)n * n; ++e) {
        const float y = 1.0f / (pb->dist[e] + 0.1f);
        float h = prm.beta == 2.0f ? y * y : prm.beta == 1.0f ? y : powf(y, prm.beta);
        if (pb->prior) h = h * pb->prior[e];
        heur[e] = h;
    }
    for (int i = 0; i < n; ++i) d0[i] = pb->dist[(size_t)i * n];
    std::vector<int> rank(n), r64(kLanes), r64inv(kLanes);
    tree_rank(n, rank.data());
    tree_rank(kLanes, r64.data());
    for (int l = 0; l < kLanes; ++l) r64inv[r64[l]] = l;
    std::vector<int> nnk((size_t)n * std::max(c->k, 1)), nnls((size_t)n * c->nn_ls);
    for (int i = 0; i < n; ++i) {
        for (int h = 0; h < c->k; ++h) nnk[(size_t)i * c->k + h] = nnl[(size_t)i * width + h];
        for (int h = 0; h < c->nn_ls; ++h) nnls[(size_t)i * c->nn_ls + h] = nnl[(size_t)i * width + h];
    }

    DevProblem &dp = c->dp;
    DevState &ds = c->ds;
    float *dist = nullptr, *d0d = nullptr, *heurd = nullptr;
    int *demd = nullptr, *nnkd = nullptr, *nnlsd = nullptr, *rankd = nullptr, *r64d = nullptr, *r64id = nullptr;
    CHK(dev_alloc(c, &dist, (size_t)n * P));
    CHK(dev_alloc(c, &heurd, (size_t)n * P));
    CHK(dev_alloc(c, &d0d, n));
    CHK(dev_alloc(c, &demd, n));
    CHK(dev_alloc(c, &nnkd, (size_t)n * std::max(c->k, 1)));
    CHK(dev_alloc(c, &nnlsd, (size_t)n * c->nn_ls));
    CHK(dev_alloc(c, &rankd, n));
    CHK(dev_alloc(c, &r64d, kLanes));
    CHK(dev_alloc(c, &r64id, kLanes));
    CHK(dev_alloc(c, &ds.tau, (size_t)n * P));
    CHK(dev_alloc(c, &ds.tinfo, (size_t)n * P));
    CHK(dev_alloc(c, &ds.tiknn, (size_t)n * std::max(c->k, 1)));
    CHK(dev_alloc(c, &ds.tours, (size_t)(m + 1) * T));
    CHK(dev_alloc(c, &ds.sizes, m + 1));
    CHK(dev_alloc(c, &ds.lens, m + 1));
    CHK(dev_alloc(c, &ds.pm, (size_t)m * kLanes));
    CHK(dev_alloc(c, &ds.ctrl, kCtrlWords));
    CHK(dev_alloc(c, &ds.bsf, prm.bsf_capacity));
    CHK(dev_alloc(c, &ds.elite, ras));
    CHK(dev_alloc(c, &ds.succ, (size_t)ras * n));
    CHK(dev_alloc(c, &ds.firstf, (size_t)ras * P));
    CHK(dev_alloc(c, &ds.delta, ras));
    CHK(dev_alloc(c, &ds.ant_evals, m));
    CHK(dev_alloc(c, &ds.ant_moves, m));
    CHK(dev_alloc(c, &ds.ant_fb, m));
    CHK(dev_alloc(c, &ds.totals, kTotWords));
    CHK(dev_alloc(c, &ds.clock0, 1));
    ds.bsf_cap = prm.bsf_capacity;

    // zero everything that padding or first reads may touch
    CHK(hipMemsetAsync(dist, 0, (size_t)n * P * 4, c->stream));
    CHK(hipMemsetAsync(heurd, 0, (size_t)n * P * 4, c->stream));
    CHK(hipMemsetAsync(ds.tau, 0, (size_t)n * P * 4, c->stream));
    CHK(hipMemsetAsync(ds.tinfo, 0, (size_t)n * P * 4, c->stream));
    CHK(hipMemsetAsync(ds.tours, 0, (size_t)(m + 1) * T * 4, c->stream));
    CHK(hipMemsetAsync(ds.sizes, 0, (size_t)(m + 1) * 4, c->stream));
    CHK(hipMemsetAsync(ds.succ, 0, (size_t)ras * n * 4, c->stream));
    CHK(hipMemsetAsync(ds.firstf, 0, (size_t)ras * P, c->stream));
    CHK(hipMemsetAsync(ds.ant_evals, 0, (size_t)m * 8, c->stream));
    CHK(hipMemsetAsync(ds.ant_moves, 0, (size_t)m * 4, c->stream));
    CHK(hipMemsetAsync(ds.ant_fb, 0, (size_t)m * 4, c->stream));
    CHK(upload_pitched(c, dist, pb->dist));
    CHK(upload_pitched(c, heurd, heur.data()));
    CHK(hipMemcpyAsync(d0d, d0.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(demd, pb->demand, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(nnkd, nnk.data(), nnk.size() * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(nnlsd, nnls.data(), nnls.size() * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(rankd, rank.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(r64d, r64.data(), kLanes * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(r64id, r64inv.data(), kLanes * 4, hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));

    dp.n = n; dp.P = P; dp.k = c->k; dp.nn_ls = c->nn_ls; dp.m = m; dp.T = T; dp.ras = ras;
    dp.Q = pb->capacity; dp.D = pb->max_dist; dp.svc = pb->service_time;
    dp.alpha = prm.alpha; dp.rho = prm.rho;
    dp.has_d = isfinite(pb->max_dist) ? 1 : 0;
    dp.dist = dist; dp.d0 = d0d; dp.dem = demd; dp.heur = heurd; dp.nnk = nnkd; dp.nnls = nnlsd;
    dp.rank = rankd; dp.rank64 = r64d; dp.rank64inv = r64id;
#undef CHK
    *out = c;
    return ACO_OK;
}

int aco_destroy(aco_ctx *c)
{
    if (!c) return ACO_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->prof)
        for (auto &ev : kv.second.pending) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ACO_OK;
}

int aco_init(aco_ctx *c)
{
    int rc = need(c, false);
    if (rc) return rc;
    const int m = c->m;
    std::vector<int32_t> pm((size_t)m * kLanes);
    for (int a = 0; a < m; ++a)
        for (int l = 0; l < kLanes; ++l) pm[(size_t)a * kLanes + l] = pm_seed(c->prm.seed, a, l);
    static const int zeros[kCtrlWords] = {0};
    static const long long zeros64[kTotWords] = {0};
    const float inf = INFINITY;
    HIP_TRY(hipMemcpyAsync(c->ds.pm, pm.data(), pm.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ds.ctrl, zeros, sizeof(zeros), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ds.totals, zeros64, sizeof(zeros64), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ds.lens + m, &inf, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->ds.sizes + m, 0, 4, c->stream));
    HIP_TRY(launch_stamp_clock(c->ds, c->stream));
    if (c->prm.tau0_rule == ACO_TAU0_FIRST_COLONY) {
        // g_aco.cpp:79-92: colony on tau = 0.5, then tau0 = 1/(rho * L_bsf)
        HIP_TRY(launch_pheromone_init(c->dp, c->ds, 0.5f, 0, c->stream));
        if ((rc = run_gather(c))) return rc;
        if ((rc = run_construct(c))) return rc;
        if ((rc = run_statistics(c))) return rc;
        HIP_TRY(launch_pheromone_init(c->dp, c->ds, 0.0f, 1, c->stream));
        if ((rc = run_gather(c))) return rc;
    } else {
        const float L = nn_tour_length(c->n, c->h_dist.data(), c->h_dem.data(), c->Q, c->D, c->svc);
        if (!(L > 0.0f)) return set_error(ACO_ERR_INFEASIBLE, "nearest-neighbour tour failed");
        HIP_TRY(launch_pheromone_init(c->dp, c->ds, 1.0f / ((float)c->n * L), 0, c->stream));
        if ((rc = run_gather(c))) return rc;
    }
    c->inited = true;
    return ACO_OK;
}

int aco_construct(aco_ctx *c)
{
    int rc = need(c, true);
    return rc ? rc : run_construct(c);
}

int aco_local_search(aco_ctx *c)
{
    int rc = need(c, true);
    return rc ? rc : run_local_search(c);
}

int aco_update_statistics(aco_ctx *c)
{
    int rc = need(c, true);
    return rc ? rc : run_statistics(c);
}

int aco_pheromone_update(aco_ctx *c)
{
    int rc = need(c, true);
    return rc ? rc : run_pheromone(c);
}

int aco_pheromone_init(aco_ctx *c, float tau0)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!(tau0 >= 0.0f) || !isfinite(tau0)) return set_error(ACO_ERR_ARG, "tau0 must be finite and >= 0");
    HIP_TRY(launch_pheromone_init(c->dp, c->ds, tau0, 0, c->stream));
    return run_gather(c);
}

int aco_iteration(aco_ctx *c)
{
    int rc = need(c, true);
    if (rc) return rc;
    if (!c->prm.use_graph || c->profiling) return run_iteration_steps(c);
    if (!c->gexec) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = run_iteration_steps(c);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(ACO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        c->graph = g;
        HIP_TRY(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    return ACO_OK;
}

int aco_run(aco_ctx *c, int32_t iterations)
{
    if (iterations < 0) return set_error(ACO_ERR_ARG, "negative iteration count");
    for (int i = 0; i < iterations; ++i) {
        int rc = aco_iteration(c);
        if (rc) return rc;
    }
    return ACO_OK;
}

int aco_synchronize(aco_ctx *c)
{
    int rc = need(c, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_iteration(aco_ctx *c, int32_t *it)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!it) return set_error(ACO_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(it, c->ds.ctrl + kCtrlIter, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_best(aco_ctx *c, int32_t *tour, int32_t cap, int32_t *size, float *len)
{
    int rc = need(c, true);
    if (rc) return rc;
    int sz = 0;
    float L = 0.0f;
    HIP_TRY(hipMemcpyAsync(&sz, c->ds.sizes + c->m, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&L, c->ds.lens + c->m, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (size) *size = sz;
    if (len) *len = L;
    if (tour) {
        if (cap < sz) return set_error(ACO_ERR_ARG, "tour buffer too small");
        HIP_TRY(hipMemcpyAsync(tour, c->ds.tours + (size_t)c->m * c->T, (size_t)sz * 4,
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ACO_OK;
}

int aco_get_ants(aco_ctx *c, int32_t *tours, int32_t stride, int32_t *sizes, float *lens)
{
    int rc = need(c, false);
    if (rc) return rc;
    const int m1 = c->m + 1;
    if (tours) {
        if (stride < c->T) return set_error(ACO_ERR_ARG, "stride must be >= 2n");
        HIP_TRY(hipMemcpy2DAsync(tours, (size_t)stride * 4, c->ds.tours, (size_t)c->T * 4,
                                 (size_t)c->T * 4, m1, hipMemcpyDeviceToHost, c->stream));
    }
    if (sizes) HIP_TRY(hipMemcpyAsync(sizes, c->ds.sizes, (size_t)m1 * 4, hipMemcpyDeviceToHost, c->stream));
    if (lens) HIP_TRY(hipMemcpyAsync(lens, c->ds.lens, (size_t)m1 * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_ants(aco_ctx *c, const int32_t *tours, int32_t stride, const int32_t *sizes, const float *lens)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!tours || !sizes || !lens || stride < 1) return set_error(ACO_ERR_ARG, "null argument");
    const int m1 = c->m + 1;
    for (int a = 0; a < m1; ++a) {
        const int sz = sizes[a];
        if (a == c->m && sz == 0) continue;  // empty best-so-far slot
        if (sz < 3 || sz > c->T || sz > stride) return set_error(ACO_ERR_ARG, "tour size out of range");
        const int32_t *t = tours + (size_t)a * stride;
        if (t[0] != 0 || t[sz - 1] != 0) return set_error(ACO_ERR_ARG, "tours must start and end at the depot");
        for (int q = 0; q < sz; ++q)
            if (t[q] < 0 || t[q] >= c->n) return set_error(ACO_ERR_ARG, "tour node out of range");
    }
    std::vector<int32_t> buf((size_t)m1 * c->T, 0);
    for (int a = 0; a < m1; ++a)
        memcpy(&buf[(size_t)a * c->T], tours + (size_t)a * stride, (size_t)sizes[a] * 4);
    HIP_TRY(hipMemcpyAsync(c->ds.tours, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ds.sizes, sizes, (size_t)m1 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ds.lens, lens, (size_t)m1 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_bsf(aco_ctx *c, aco_bsf_record *rec, int32_t cap, int32_t *count)
{
    int rc = need(c, false);
    if (rc) return rc;
    int nb = 0;
    HIP_TRY(hipMemcpyAsync(&nb, c->ds.ctrl + kCtrlNbsf, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (count) *count = nb;
    const int k = std::min(nb, cap);
    if (rec && k > 0) {
        static_assert(sizeof(aco_bsf_record) == sizeof(BsfRecord), "record layout");
        HIP_TRY(hipMemcpyAsync(rec, c->ds.bsf, (size_t)k * sizeof(BsfRecord), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ACO_OK;
}

int aco_get_pheromone(aco_ctx *c, float *tau, float *ti)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (tau) HIP_TRY(download_pitched(c, tau, c->ds.tau));
    if (ti) HIP_TRY(download_pitched(c, ti, c->ds.tinfo));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_pheromone(aco_ctx *c, const float *tau, const float *ti)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!tau || !ti) return set_error(ACO_ERR_ARG, "null argument");
    HIP_TRY(upload_pitched(c, c->ds.tau, tau));
    HIP_TRY(upload_pitched(c, c->ds.tinfo, ti));
    if ((rc = run_gather(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_rng_state(aco_ctx *c, int32_t *state)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!state) return set_error(ACO_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(state, c->ds.pm, (size_t)c->m * kLanes * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_counters(aco_ctx *c, aco_counters *out)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!out) return set_error(ACO_ERR_ARG, "null output");
    long long t[kTotWords];
    int ctrl[kCtrlWords];
    HIP_TRY(hipMemcpyAsync(t, c->ds.totals, sizeof(t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(ctrl, c->ds.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->ls_evals = t[kTotEvals];
    out->ls_moves = t[kTotMoves];
    out->fallbacks = t[kTotFallbacks];
    out->iterations = t[kTotIters];
    out->bsf_overflow = ctrl[kCtrlOverflow];
    return ACO_OK;
}

int aco_profile_enable(aco_ctx *c, int32_t enable)
{
    int rc = need(c, false);
    if (rc) return rc;
    c->profiling = enable != 0;
    return ACO_OK;
}

int aco_profile_get(aco_ctx *c, const char *name, double *total_ms, int64_t *launches)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!name) return set_error(ACO_ERR_ARG, "null name");
    HIP_TRY(hipStreamSynchronize(c->stream));
    auto it = c->prof.find(name);
    if (it == c->prof.end()) {
        if (total_ms) *total_ms = 0.0;
        if (launches) *launches = 0;
        return ACO_OK;
    }
    ProfSlot &s = it->second;
    for (auto &ev : s.pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
        s.total_ms += ms;
        s.launches += 1;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    s.pending.clear();
    if (total_ms) *total_ms = s.total_ms;
    if (launches) *launches = s.launches;
    return ACO_OK;
}

}  // extern "C"
