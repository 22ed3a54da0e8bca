// C ABI of libcvrpaco.so: context lifetime, validation, device memory,
// launch sequencing (eager or one hipGraph per iteration), readback.
//
// A context holds one or more instances (aco_create_batch); every launch
// covers all of them.  Mirrors the reference host driver g_ACO
// (cvrp_ocl/gpu/g_aco.cpp); include/cvrp_aco.h maps each entry point.
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
#include "cvrp_decompose.h"

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
// Seed of the random-order 2-opt's scan-order stream (the reference draws
// it from instance->rnd_seed, problem.cpp:213; here it follows the colony
// seed).
static int32_t ls_seed_of(uint64_t seed)
{
    return (int32_t)(1 + splitmix64(seed ^ 0xC2B2AE3D27D4EB4Full) % 2147483646ull);
}

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

// One instance of a context: host copies needed after creation, and the
// device problem/state structures pointing into the instance's arena.
struct InstanceData {
    int n = 0, P = 0, T = 0;
    int Q = 0;
    float D = INFINITY, svc = 0.0f;
    uint64_t seed = 0;
    int32_t ls_seed0 = 1;     // initial state of the scan-order stream
    int setup_beta = 2;       // eta^beta on the device for beta 1 or 2 (0: uploaded)
    bool setup_prior = false, setup_int16 = false, nnl_given = false;
    DevProblem dp{};
    DevState ds{};
    void *arena = nullptr;
};

struct aco_ctx {
    aco_params prm{};
    int m = 0, device = 0;
    std::vector<InstanceData> inst;
    DevBatch bt{};
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    LsStreams lss;   // batches: local-search size classes side by side
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool inited = false;
    float tick_to_s = 0.0f;
    bool profiling = false;
    std::map<std::string, ProfSlot> prof;
    LsShape ls{4, 16, 1, 0, 0};
    size_t lds_limit = 0;
    // static LDS of the kernel families, added to each launch's dynamic LDS
    size_t lds_static_cons = 0, lds_static_ls = 0, lds_static_fused = 0;
    int max_n = 0;
    bool int16_rows = true;   // every distance an integer in [0, 65535]
    void *setup_scratch = nullptr;   // candidate lists during aco_create
};

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(ACO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Host-to-device uploads through two pinned buffers used in turn: a chunk
// is copied into a buffer whose previous transfer has finished and sent
// asynchronously, so uploads run at DMA speed while the host prepares the
// next chunk.  The destructor waits for the transfers before freeing.
struct PinnedStager {
    size_t kBytes = size_t(32) << 20;   // per buffer
    unsigned char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int cur = 0;
    hipStream_t st = nullptr;

    hipError_t init(hipStream_t s, size_t bytes)
    {
        st = s;
        kBytes = bytes;
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipHostMalloc((void **)&buf[i], kBytes, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    ~PinnedStager()
    {
        for (int i = 0; i < 2; ++i) {
            if (used[i]) (void)hipEventSynchronize(done[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    // the current buffer, once its previous transfer has finished
    hipError_t take(unsigned char *&p)
    {
        p = buf[cur];
        return used[cur] ? hipEventSynchronize(done[cur]) : hipSuccess;
    }
    hipError_t sent()
    {
        const hipError_t e = hipEventRecord(done[cur], st);
        used[cur] = true;
        cur ^= 1;
        return e;
    }
    hipError_t put(void *dst, const void *src, size_t bytes)
    {
        unsigned char *d = static_cast<unsigned char *>(dst);
        const unsigned char *s = static_cast<const unsigned char *>(src);
        while (bytes) {
            const size_t c = bytes < kBytes ? bytes : kBytes;
            unsigned char *p = nullptr;
            hipError_t e = take(p);
            if (e != hipSuccess) return e;
            memcpy(p, s, c);
            e = hipMemcpyAsync(d, p, c, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return e;
            if ((e = sent()) != hipSuccess) return e;
            d += c;
            s += c;
            bytes -= c;
        }
        return hipSuccess;
    }
    // `rows` dense rows of `width` bytes into a device array of row pitch dpitch
    hipError_t put2d(void *dst, size_t dpitch, const void *src, size_t width, size_t rows)
    {
        const size_t per = std::max<size_t>(1, kBytes / width);
        for (size_t r0 = 0; r0 < rows; r0 += per) {
            const size_t nr = std::min(per, rows - r0);
            unsigned char *p = nullptr;
            hipError_t e = take(p);
            if (e != hipSuccess) return e;
            memcpy(p, static_cast<const unsigned char *>(src) + r0 * width, nr * width);
            e = hipMemcpy2DAsync(static_cast<unsigned char *>(dst) + r0 * dpitch, dpitch, p, width, width, nr,
                                 hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return e;
            if ((e = sent()) != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

// Bump allocator over one device allocation (256-byte aligned pieces).
struct Carver {
    unsigned char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t count)
    {
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off = (off + std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
};

// Lays out one instance's device buffers; with base == nullptr only sizes.
static size_t carve_instance(InstanceData &I, int m, int k, int nnP, int ras, int bsf_cap,
                             unsigned char *base)
{
    Carver c{base};
    const int n = I.n, P = I.P, T = I.T;
    DevProblem &dp = I.dp;
    DevState &ds = I.ds;
    float *dist = c.take<float>((size_t)n * P);
    float *heur = c.take<float>((size_t)n * P);
    float *d0 = c.take<float>(n);
    int *dem = c.take<int>(n);
    int *nnk = c.take<int>((size_t)n * std::max(k, 1));
    float *nnkd = c.take<float>((size_t)n * std::max(k, 1));
    int *nnls = c.take<int>((size_t)n * nnP);
    float *nnd = c.take<float>((size_t)n * nnP);
    int *rank = c.take<int>(n);
    unsigned short *nnpos = c.take<unsigned short>((size_t)n * n);
    unsigned short *dist16 = c.take<unsigned short>((size_t)n * P);
    int *r64 = c.take<int>(kLanes);
    int *r64i = c.take<int>(kLanes);
    ds.tau = c.take<float>((size_t)n * P);
    ds.tinfo = c.take<float>((size_t)n * P);
    ds.tiknn = c.take<float>((size_t)n * std::max(k, 1));
    ds.tours = c.take<int>((size_t)(m + 1) * T);
    ds.sizes = c.take<int>(m + 1);
    ds.lens = c.take<float>(m + 1);
    ds.pm = c.take<int>((size_t)m * kLanes);
    ds.ctrl = c.take<int>(kCtrlWords);
    ds.bsf = c.take<BsfRecord>(bsf_cap);
    ds.elite = c.take<int>(ras);
    ds.succ = c.take<int>((size_t)ras * n);
    ds.firstf = c.take<unsigned char>((size_t)ras * P);
    ds.delta = c.take<float>(ras);
    ds.ant_evals = c.take<long long>(m);
    ds.ant_moves = c.take<int>(m);
    ds.ant_fb = c.take<int>(m);
    ds.totals = c.take<long long>(kTotWords);
    ds.bsf_cap = bsf_cap;
    dp.dist = dist; dp.heur = heur; dp.d0 = d0; dp.dem = dem; dp.nnk = nnk; dp.nnkd = nnkd;
    dp.nnls = nnls; dp.nnd = nnd; dp.rank = rank; dp.rank64 = r64; dp.rank64inv = r64i;
    dp.nnpos = nnpos;
    dp.dist16 = dist16;
    return c.off;
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
    HIP_TRY(timed(c, "construct",
                  [&] {
                      return launch_construct(c->bt, c->prm.selection, c->prm.rng, c->ls.construct_apb,
                                              c->prm.k, c->stream);
                  }));
    return ACO_OK;
}

static int run_local_search(aco_ctx *c)
{
    if (!c->prm.ls_flags) return ACO_OK;
    HIP_TRY(timed(c, "local_search",
                  [&] { return launch_local_search(c->bt, c->prm.ls_flags, c->ls, c->stream, &c->lss); }));
    return ACO_OK;
}

static int run_statistics(aco_ctx *c)
{
    HIP_TRY(timed(c, "statistics", [&] { return launch_statistics(c->bt, c->tick_to_s, c->stream); }));
    return ACO_OK;
}

static int run_gather(aco_ctx *c)
{
    if (c->bt.max_nk <= 0) return ACO_OK;
    HIP_TRY(timed(c, "gather", [&] { return launch_gather(c->bt, c->stream); }));
    return ACO_OK;
}

// The pheromone pass also writes the candidate rows' total_info (no gather).
static int run_pheromone(aco_ctx *c)
{
    HIP_TRY(timed(c, "pheromone", [&] { return launch_pheromone_update(c->bt, c->stream); }));
    return ACO_OK;
}

static int run_iteration_steps(aco_ctx *c)
{
    int rc;
    if (c->ls.fused) {
        HIP_TRY(timed(c, "iteration_fused", [&] {
            return launch_iteration_fused(c->bt, c->prm.rng, c->prm.ls_flags, c->ls, c->tick_to_s, c->stream);
        }));
        return ACO_OK;
    }
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

static int need_inst(aco_ctx *c, int b, bool inited)
{
    int rc = need(c, inited);
    if (rc) return rc;
    if (b < 0 || b >= (int)c->inst.size()) return set_error(ACO_ERR_ARG, "instance index out of range");
    return ACO_OK;
}

// Parameter checks shared by every instance of a context.
static int check_params(aco_params &prm, int min_n)
{
    if (prm.bsf_capacity <= 0) prm.bsf_capacity = 4096;
    if (prm.ras_ranks < 1 || prm.ras_ranks > kMaxRas) return set_error(ACO_ERR_ARG, "ras_ranks must be in [1,16]");
    if (!(prm.rho > 0.0f && prm.rho <= 1.0f)) return set_error(ACO_ERR_ARG, "rho must be in (0,1]");
    if (!isfinite(prm.alpha) || !isfinite(prm.beta)) return set_error(ACO_ERR_ARG, "alpha/beta not finite");
    if (prm.selection != ACO_SEL_IROULETTE && prm.selection != ACO_SEL_KNN)
        return set_error(ACO_ERR_ARG, "unknown selection rule");
    if (prm.rng != ACO_RNG_PARK_MILLER && prm.rng != ACO_RNG_PHILOX) return set_error(ACO_ERR_ARG, "unknown rng");
    if (prm.selection == ACO_SEL_IROULETTE && prm.rng != ACO_RNG_PARK_MILLER)
        return set_error(ACO_ERR_ARG, "iroulette draws from Park-Miller lane streams");
    if (prm.selection == ACO_SEL_KNN && (prm.k < 1 || prm.k > kLanes || prm.k > min_n - 2))
        return set_error(ACO_ERR_ARG, "k must be in [1, min(64, n-2)] for roulette_knn");
    if (prm.selection != ACO_SEL_KNN) prm.k = 0;
    if (prm.nn_ls < 0 || prm.nn_ls > min_n - 2) return set_error(ACO_ERR_ARG, "nn_ls must be <= n-2");
    if (prm.ls_flags & ~15) return set_error(ACO_ERR_ARG, "unknown local search flag");
    if ((prm.ls_flags & ACO_LS_2OPT) && (prm.ls_flags & ACO_LS_2OPT_RANDOM))
        return set_error(ACO_ERR_ARG, "ACO_LS_2OPT and ACO_LS_2OPT_RANDOM are two orders of one 2-opt: choose one");
    if (prm.tau0_rule != ACO_TAU0_FIRST_COLONY && prm.tau0_rule != ACO_TAU0_NN_TOUR)
        return set_error(ACO_ERR_ARG, "unknown tau0 rule");
    return ACO_OK;
}

// Instance checks: finite, non-negative, symmetric distances; every
// customer servable by a vehicle of its own (otherwise construction could
// never visit it: kernel.cl:438-479 would not terminate).
static int check_instance(const aco_problem *pb, int width, std::vector<int> &nnl, bool &int16)
{
    const int n = pb->n;
    if (n < 3 || n > kMaxNodes) return set_error(ACO_ERR_ARG, "aco_create: need 3 <= n <= 4096");
    if (!pb->dist || !pb->demand) return set_error(ACO_ERR_ARG, "aco_create: dist and demand required");
    if (pb->capacity <= 0) return set_error(ACO_ERR_ARG, "capacity must be positive");
    if (isnan(pb->max_dist) || !(pb->service_time >= 0.0f) || !isfinite(pb->service_time))
        return set_error(ACO_ERR_ARG, "bad max_dist or service_time");
    int16 = true;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const float d = pb->dist[(size_t)i * n + j];
            if (!isfinite(d) || d < 0.0f) return set_error(ACO_ERR_ARG, "distances must be finite and >= 0");
            if (d != pb->dist[(size_t)j * n + i]) return set_error(ACO_ERR_ARG, "distance matrix is not symmetric");
            int16 = int16 && d <= 65535.0f && d == floorf(d);
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
    nnl.clear();
    if (pb->nn_list) {
        nnl.assign((size_t)n * width, 0);
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
    }
    // without nn_list the lists are built on the device in aco_create
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

int aco_create(const aco_problem *pb, const aco_params *prm, aco_ctx **out)
{
    return aco_create_batch(pb, 1, prm, out);
}

int aco_create_batch(const aco_problem *pbs, int32_t count, const aco_params *prm_in, aco_ctx **out)
{
    if (!pbs || !prm_in || !out || count < 1) return set_error(ACO_ERR_ARG, "aco_create: null argument");
    *out = nullptr;
    aco_params prm = *prm_in;
    int min_n = INT32_MAX, max_n = 0;
    for (int b = 0; b < count; ++b) {
        min_n = std::min(min_n, pbs[b].n);
        max_n = std::max(max_n, pbs[b].n);
    }
    if (prm.n_ants <= 0) {
        if (count > 1) return set_error(ACO_ERR_ARG, "n_ants must be given for a batch");
        prm.n_ants = pbs[0].n;
    }
    int rc = check_params(prm, min_n);
    if (rc) return rc;
    std::vector<std::vector<int>> nnl(count);
    std::vector<int> nn_ls(count);
    std::vector<char> int16(count);
    for (int b = 0; b < count; ++b) {
        nn_ls[b] = prm.nn_ls > 0 ? prm.nn_ls : pbs[b].n - 2;
        bool i16 = false;
        rc = check_instance(&pbs[b], std::max(prm.k, nn_ls[b]), nnl[b], i16);
        if (rc) return rc;
        int16[b] = i16;
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(ACO_ERR_NODEV, "no HIP device");

    aco_ctx *c = new aco_ctx();
    c->prm = prm;
    c->m = prm.n_ants;
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
    // default local-search shape: 4 waves, 16-row ring with prefetch when it
    // fits, else an 8-row ring
    c->lds_limit = prop.sharedMemPerBlock;
    c->max_n = max_n;
    CHK(construct_static_lds(c->lds_static_cons));
    CHK(local_search_static_lds(c->lds_static_ls));
    CHK(fused_static_lds(c->lds_static_fused));
    c->ls = LsShape{4, 16, 1, 0, 0, 0};
    if (local_search_lds_bytes(max_n, 16, 0, prm.ls_flags) + c->lds_static_ls > prop.sharedMemPerBlock)
        c->ls = LsShape{4, 8, 0, 0, 0, 0};
    const size_t lds_need =
        std::max(construct_lds_bytes(max_n, prm.selection) + c->lds_static_cons,
                 prm.ls_flags ? local_search_lds_bytes(max_n, c->ls.ring, 0, prm.ls_flags) + c->lds_static_ls : 0);
    if (lds_need > prop.sharedMemPerBlock)
        return fail(set_error(ACO_ERR_ARG, "instance too large for LDS: needs " + std::to_string(lds_need) +
                                               " bytes, device has " + std::to_string(prop.sharedMemPerBlock)));
    int rate_khz = 0;
    CHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
    c->tick_to_s = rate_khz > 0 ? 1.0f / (1000.0f * (float)rate_khz) : 0.0f;
    CHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int b = 0; b < count; ++b) c->int16_rows = c->int16_rows && int16[b];
    // device candidate lists (n x max(k, nn_ls) per instance): given ones
    // are uploaded, the others built on the device from the uploaded
    // distances (compute_nn_lists, vrpHelper.cpp:188-236); freed at the end
    std::vector<int *> d_nnl(count, nullptr);
    std::vector<int> nnl_width(count);
    {
        size_t total = 0;
        for (int b = 0; b < count; ++b) {
            nnl_width[b] = std::max(prm.k, nn_ls[b]);
            total += ((size_t)pbs[b].n * nnl_width[b] + 63) & ~size_t(63);
        }
        void *q = nullptr;
        CHK(hipMalloc(&q, total * 4));
        c->setup_scratch = q;
        size_t off = 0;
        for (int b = 0; b < count; ++b) {
            d_nnl[b] = static_cast<int *>(q) + off;
            off += ((size_t)pbs[b].n * nnl_width[b] + 63) & ~size_t(63);
        }
    }

    const int m = c->m, ras = prm.ras_ranks, k = prm.k;
    unsigned long long *clock0 = nullptr;
    {
        void *q = nullptr;
        CHK(hipMalloc(&q, 256));
        c->allocs.push_back(q);
        clock0 = static_cast<unsigned long long *>(q);
    }
    std::vector<int> r64(kLanes), r64inv(kLanes);
    tree_rank(kLanes, r64.data());
    for (int l = 0; l < kLanes; ++l) r64inv[r64[l]] = l;

    // pinned buffers sized for the largest instance's uploads (about 24 n^2
    // bytes), between 1 and 32 MB each
    size_t stage_bytes = size_t(1) << 20;
    for (int b = 0; b < count; ++b)
        stage_bytes = std::max(stage_bytes, (size_t)24 * pbs[b].n * pbs[b].n + (size_t(64) << 10));
    stage_bytes = std::min(stage_bytes, size_t(32) << 20);
    PinnedStager up;
    CHK(up.init(c->stream, stage_bytes));
    c->inst.resize(count);
    for (int b = 0; b < count; ++b) {
        const aco_problem *pb = &pbs[b];
        InstanceData &I = c->inst[b];
        const int n = pb->n, nl = nn_ls[b];
        const int nnP = (nl + 3) & ~3;
        I.n = n;
        // row pitch: every row of the fp32 and 16-bit matrices starts on a
        // 16-byte boundary (local search copies rows in 16-byte pieces)
        I.P = (n + 7) & ~7;
        I.T = 2 * n;
        I.Q = pb->capacity;
        I.D = pb->max_dist;
        I.svc = pb->service_time;
        I.seed = prm.seed + (uint64_t)b;
        I.ls_seed0 = ls_seed_of(I.seed);
        const int P = I.P;
        const size_t bytes = carve_instance(I, m, k, nnP, ras, prm.bsf_capacity, nullptr);
        CHK(hipMalloc(&I.arena, bytes));
        c->allocs.push_back(I.arena);
        CHK(hipMemsetAsync(I.arena, 0, bytes, c->stream));
        carve_instance(I, m, k, nnP, ras, prm.bsf_capacity, static_cast<unsigned char *>(I.arena));
        I.ds.clock0 = clock0;

        // uploads: the distance matrix, demands, the reduction ranks, and the
        // prior (into heur, multiplied on the device) or, for beta other
        // than 1 and 2, eta^beta itself (host powf, as the CPU oracle)
        DevProblem &dp = I.dp;
        auto h2d = [&](void *dst, const void *src, size_t bytes_) { return up.put(dst, src, bytes_); };
        CHK(up.put2d((void *)dp.dist, (size_t)P * 4, pb->dist, (size_t)n * 4, n));
        const int beta_code = prm.beta == 2.0f ? 2 : prm.beta == 1.0f ? 1 : 0;
        if (beta_code == 0) {
            std::vector<float> heur((size_t)n * n);
            for (size_t e = 0; e < (size_t)n * n; ++e) {
                const float y = 1.0f / (pb->dist[e] + 0.1f);
                float h = powf(y, prm.beta);
                if (pb->prior) h = h * pb->prior[e];
                heur[e] = h;
            }
            CHK(up.put2d((void *)dp.heur, (size_t)P * 4, heur.data(), (size_t)n * 4, n));
        } else if (pb->prior) {
            CHK(up.put2d((void *)dp.heur, (size_t)P * 4, pb->prior, (size_t)n * 4, n));
        }
        CHK(h2d((void *)dp.dem, pb->demand, (size_t)n * 4));
        std::vector<int> rank(n);
        tree_rank(n, rank.data());
        CHK(h2d((void *)dp.rank, rank.data(), (size_t)n * 4));
        CHK(h2d((void *)dp.rank64, r64.data(), kLanes * 4));
        CHK(h2d((void *)dp.rank64inv, r64inv.data(), kLanes * 4));
        if (!nnl[b].empty()) CHK(h2d(d_nnl[b], nnl[b].data(), nnl[b].size() * 4));
        I.setup_beta = beta_code;
        I.setup_prior = pb->prior != nullptr;
        I.setup_int16 = int16[b] != 0;
        I.nnl_given = !nnl[b].empty();
        dp.n = n; dp.P = P; dp.k = k; dp.nn_ls = nl; dp.nnP = nnP; dp.m = m; dp.T = I.T; dp.ras = ras;
        dp.Q = pb->capacity; dp.D = pb->max_dist; dp.svc = pb->service_time;
        dp.alpha = prm.alpha; dp.rho = prm.rho;
        dp.has_d = isfinite(pb->max_dist) ? 1 : 0;
        dp.dem16 = 1;
        for (int i = 0; i < n; ++i)
            if (pb->demand[i] < 0 || pb->demand[i] > 65535) dp.dem16 = 0;
        dp.seed = I.seed;
        dp.tau0_nn = 0.0f;
        if (prm.tau0_rule == ACO_TAU0_NN_TOUR) {
            const float Lnn = nn_tour_length(n, pb->dist, pb->demand, pb->capacity, pb->max_dist,
                                             pb->service_time);
            if (!(Lnn > 0.0f)) return fail(set_error(ACO_ERR_INFEASIBLE, "nearest-neighbour tour failed"));
            dp.tau0_nn = 1.0f / ((float)n * Lnn);
        }
    }

    // candidate lists of the instances given without them, from the device
    // distances, then every derived array of every instance
    {
        std::vector<int> nv, pv, wv;
        std::vector<const float *> dv;
        std::vector<int *> ov;
        for (int b = 0; b < count; ++b) {
            if (c->inst[b].nnl_given) continue;
            nv.push_back(c->inst[b].n);
            pv.push_back(c->inst[b].P);
            wv.push_back(nnl_width[b]);
            dv.push_back(c->inst[b].dp.dist);
            ov.push_back(d_nnl[b]);
        }
        CHK(hipStreamSynchronize(c->stream));   // distances uploaded
        if (!nv.empty()) {
            const int rc2 =
                gpu_nn_lists_dev(c->stream, (int)nv.size(), nv.data(), dv.data(), pv.data(), wv.data(), 0, ov.data());
            if (rc2) return fail(rc2);
        }
        for (int b = 0; b < count; ++b) {
            const InstanceData &I = c->inst[b];
            CHK(launch_setup_rows(I.dp, I.setup_beta, I.setup_prior ? 1 : 0, I.setup_int16 ? 1 : 0, c->stream));
            CHK(launch_setup_nn(I.dp, d_nnl[b], nnl_width[b], c->stream));
        }
        CHK(hipStreamSynchronize(c->stream));
        (void)hipFree(c->setup_scratch);
        c->setup_scratch = nullptr;
    }

    // device tables of the batch
    std::vector<DevProblem> hp(count);
    std::vector<DevState> hs(count);
    std::vector<int> ant_inst((size_t)count * m), ant_base(count);
    DevBatch &bt = c->bt;
    bt.B = count;
    bt.ants = count * m;
    bt.max_n = bt.max_P = bt.max_nk = 0;
    bt.has_d = 0;
    bt.dem16 = 1;
    for (int b = 0; b < count; ++b) {
        hp[b] = c->inst[b].dp;
        hs[b] = c->inst[b].ds;
        ant_base[b] = b * m;
        for (int a = 0; a < m; ++a) ant_inst[(size_t)b * m + a] = b;
        bt.max_n = std::max(bt.max_n, hp[b].n);
        bt.max_P = std::max(bt.max_P, hp[b].P);
        bt.max_nk = std::max(bt.max_nk, hp[b].n * hp[b].k);
        bt.has_d |= hp[b].has_d;
        bt.dem16 &= hp[b].dem16;
    }
    // local-search size classes by n (upper bounds), largest first
    static const int kClassMax[kLsClasses] = {128, 256, 384, 512, 768, 1 << 30};
    std::vector<int> ls_order;
    ls_order.reserve(ant_inst.size());
    bt.ls_classes = 0;
    for (int cl = kLsClasses - 1; cl >= 0; --cl) {
        const int lo = cl > 0 ? kClassMax[cl - 1] : 0;
        const int first = (int)ls_order.size();
        int maxn = 0;
        for (int b = 0; b < count; ++b) {
            if (hp[b].n <= lo || hp[b].n > kClassMax[cl]) continue;
            maxn = std::max(maxn, hp[b].n);
            for (int a = 0; a < m; ++a) ls_order.push_back(b * m + a);
        }
        if ((int)ls_order.size() == first) continue;
        bt.ls_first[bt.ls_classes] = first;
        bt.ls_maxn[bt.ls_classes] = maxn;
        ++bt.ls_classes;
    }
    bt.ls_first[bt.ls_classes] = (int)ls_order.size();
    // streams for the size classes beyond the first (default priority)
    if (bt.ls_classes > 1) {
        CHK(hipEventCreateWithFlags(&c->lss.fork, hipEventDisableTiming));
        for (int i = 0; i < bt.ls_classes - 1; ++i) {
            CHK(hipStreamCreateWithFlags(&c->lss.aux[i], hipStreamNonBlocking));
            c->lss.count = i + 1;
            CHK(hipEventCreateWithFlags(&c->lss.join[i], hipEventDisableTiming));
        }
    }
    void *q = nullptr;
    const size_t tb = sizeof(DevProblem) * count + sizeof(DevState) * count + 4 * ant_inst.size() +
                      4 * ant_base.size() + 4 * ls_order.size() + 1536;
    CHK(hipMalloc(&q, tb));
    c->allocs.push_back(q);
    Carver cv{static_cast<unsigned char *>(q)};
    DevProblem *dprobs = cv.take<DevProblem>(count);
    DevState *dstates = cv.take<DevState>(count);
    int *dinst = cv.take<int>(ant_inst.size());
    int *dbase = cv.take<int>(ant_base.size());
    int *dorder = cv.take<int>(ls_order.size());
    int *dfq = cv.take<int>(1);
    CHK(hipMemcpyAsync(dprobs, hp.data(), sizeof(DevProblem) * count, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(dorder, ls_order.data(), 4 * ls_order.size(), hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(dstates, hs.data(), sizeof(DevState) * count, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(dinst, ant_inst.data(), 4 * ant_inst.size(), hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(dbase, ant_base.data(), 4 * ant_base.size(), hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    bt.cons_lds_budget = c->lds_limit - c->lds_static_cons;
    bt.cus = prop.multiProcessorCount;
    bt.probs = dprobs;
    bt.states = dstates;
    bt.ant_inst = dinst;
    bt.ant_base = dbase;
    bt.ls_order = dorder;
    bt.fq_next = dfq;
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
    if (c->setup_scratch) (void)hipFree(c->setup_scratch);
    for (int i = 0; i < c->lss.count; ++i) {
        (void)hipStreamDestroy(c->lss.aux[i]);
        (void)hipEventDestroy(c->lss.join[i]);
    }
    if (c->lss.fork) (void)hipEventDestroy(c->lss.fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ACO_OK;
}

int32_t aco_batch_size(aco_ctx *c) { return c ? (int32_t)c->inst.size() : 0; }

int aco_init(aco_ctx *c)
{
    int rc = need(c, false);
    if (rc) return rc;
    const int m = c->m;
    static const int zeros[kCtrlWords] = {0};
    static const long long zeros64[kTotWords] = {0};
    static const float inf = INFINITY;
    std::vector<std::vector<int32_t>> pms(c->inst.size());
    for (size_t b = 0; b < c->inst.size(); ++b) {
        InstanceData &I = c->inst[b];
        std::vector<int32_t> &pm = pms[b];
        pm.resize((size_t)m * kLanes);
        for (int a = 0; a < m; ++a)
            for (int l = 0; l < kLanes; ++l) pm[(size_t)a * kLanes + l] = pm_seed(I.seed, a, l);
        HIP_TRY(hipMemcpyAsync(I.ds.pm, pm.data(), pm.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(I.ds.ctrl, zeros, sizeof(zeros), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(I.ds.ctrl + kCtrlLsSeed, &c->inst[b].ls_seed0, 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(I.ds.totals, zeros64, sizeof(zeros64), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(I.ds.lens + m, &inf, 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemsetAsync(I.ds.sizes + m, 0, 4, c->stream));
    }
    HIP_TRY(launch_stamp_clock(c->bt, c->stream));
    if (c->prm.tau0_rule == ACO_TAU0_FIRST_COLONY) {
        // g_aco.cpp:79-92: colony on tau = 0.5, then tau0 = 1/(rho * L_bsf)
        HIP_TRY(launch_pheromone_init(c->bt, kTau0Arg, 0.5f, c->stream));
        if ((rc = run_gather(c))) return rc;
        if ((rc = run_construct(c))) return rc;
        if ((rc = run_statistics(c))) return rc;
        HIP_TRY(launch_pheromone_init(c->bt, kTau0Bsf, 0.0f, c->stream));
        if ((rc = run_gather(c))) return rc;
    } else {
        HIP_TRY(launch_pheromone_init(c->bt, kTau0Nn, 0.0f, c->stream));
        if ((rc = run_gather(c))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));  // pm staging vectors go out of scope
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
    HIP_TRY(launch_pheromone_init(c->bt, kTau0Arg, tau0, c->stream));
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
    HIP_TRY(hipMemcpyAsync(it, c->inst[0].ds.ctrl + kCtrlIter, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_best_at(aco_ctx *c, int32_t b, int32_t *tour, int32_t cap, int32_t *size, float *len)
{
    int rc = need_inst(c, b, true);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    int sz = 0;
    float L = 0.0f;
    HIP_TRY(hipMemcpyAsync(&sz, I.ds.sizes + c->m, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&L, I.ds.lens + c->m, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (size) *size = sz;
    if (len) *len = L;
    if (tour) {
        if (cap < sz) return set_error(ACO_ERR_ARG, "tour buffer too small");
        HIP_TRY(hipMemcpyAsync(tour, I.ds.tours + (size_t)c->m * I.T, (size_t)sz * 4,
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ACO_OK;
}

int aco_get_best(aco_ctx *c, int32_t *tour, int32_t cap, int32_t *size, float *len)
{
    return aco_get_best_at(c, 0, tour, cap, size, len);
}

int aco_get_ants_at(aco_ctx *c, int32_t b, int32_t *tours, int32_t stride, int32_t *sizes, float *lens)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    const int m1 = c->m + 1;
    if (tours) {
        if (stride < I.T) return set_error(ACO_ERR_ARG, "stride must be >= 2n");
        HIP_TRY(hipMemcpy2DAsync(tours, (size_t)stride * 4, I.ds.tours, (size_t)I.T * 4, (size_t)I.T * 4, m1,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    if (sizes) HIP_TRY(hipMemcpyAsync(sizes, I.ds.sizes, (size_t)m1 * 4, hipMemcpyDeviceToHost, c->stream));
    if (lens) HIP_TRY(hipMemcpyAsync(lens, I.ds.lens, (size_t)m1 * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_ants(aco_ctx *c, int32_t *tours, int32_t stride, int32_t *sizes, float *lens)
{
    return aco_get_ants_at(c, 0, tours, stride, sizes, lens);
}

// A tour the kernels accept as an ant's solution: depot first and last,
// every customer exactly once, no empty route (construction never makes
// one).  Returns the reason it is not, or nullptr.
static const char *tour_structure_error(const int32_t *t, int size, int n, std::vector<char> &seen)
{
    if (t[0] != 0 || t[size - 1] != 0) return "tours must start and end at the depot";
    seen.assign(n, 0);
    int customers = 0;
    for (int q = 0; q < size; ++q) {
        const int v = t[q];
        if (v < 0 || v >= n) return "tour node out of range";
        if (v == 0) {
            if (q > 0 && t[q - 1] == 0) return "tour has an empty route";
            continue;
        }
        if (seen[v]) return "tour visits a customer twice";
        seen[v] = 1;
        ++customers;
    }
    return customers == n - 1 ? nullptr : "tour misses a customer";
}

int aco_set_ants_at(aco_ctx *c, int32_t b, const int32_t *tours, int32_t stride, const int32_t *sizes,
                    const float *lens)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    if (!tours || !sizes || !lens || stride < 1) return set_error(ACO_ERR_ARG, "null argument");
    const InstanceData &I = c->inst[b];
    const int m1 = c->m + 1;
    std::vector<char> seen;
    for (int a = 0; a < m1; ++a) {
        const int sz = sizes[a];
        if (a == c->m && sz == 0) continue;  // empty best-so-far slot
        if (sz < 3 || sz > I.T || sz > stride) return set_error(ACO_ERR_ARG, "tour size out of range");
        if (const char *e = tour_structure_error(tours + (size_t)a * stride, sz, I.n, seen))
            return set_error(ACO_ERR_ARG, std::string(e) + " (ant " + std::to_string(a) + ")");
    }
    std::vector<int32_t> buf((size_t)m1 * I.T, 0);
    for (int a = 0; a < m1; ++a) memcpy(&buf[(size_t)a * I.T], tours + (size_t)a * stride, (size_t)sizes[a] * 4);
    HIP_TRY(hipMemcpyAsync(I.ds.tours, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(I.ds.sizes, sizes, (size_t)m1 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(I.ds.lens, lens, (size_t)m1 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_ants(aco_ctx *c, const int32_t *tours, int32_t stride, const int32_t *sizes, const float *lens)
{
    return aco_set_ants_at(c, 0, tours, stride, sizes, lens);
}

int aco_set_ant_at(aco_ctx *c, int32_t b, int32_t ant, const int32_t *tour, int32_t size)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    if (ant < 0 || ant >= c->m) return set_error(ACO_ERR_ARG, "aco_set_ant_at: ant out of range");
    if (!tour || size < 3 || size > I.T) return set_error(ACO_ERR_ARG, "aco_set_ant_at: tour size out of range");
    std::vector<char> seen;
    if (const char *e = tour_structure_error(tour, size, I.n, seen))
        return set_error(ACO_ERR_ARG, std::string("aco_set_ant_at: ") + e);
    HIP_TRY(hipMemcpyAsync(I.ds.tours + (size_t)ant * I.T, tour, (size_t)size * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(I.ds.sizes + ant, &size, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_ant_length(I.dp, I.ds, ant, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_pheromone_from_subs(aco_ctx *c, int32_t b, aco_ctx *sc, const int32_t *offsets, const int32_t *real_nodes,
                            const int32_t *map)
{
    int rc = need_inst(c, b, true);
    if (rc) return rc;
    if (!sc || !offsets || !real_nodes || !map) return set_error(ACO_ERR_ARG, "aco_pheromone_from_subs: null argument");
    if (!sc->inited) return set_error(ACO_ERR_STATE, "aco_pheromone_from_subs: sub-colonies not initialised");
    if (sc->device != c->device) return set_error(ACO_ERR_ARG, "aco_pheromone_from_subs: contexts on different devices");
    const InstanceData &M = c->inst[b];
    const int count = (int)sc->inst.size();
    std::vector<SubTauJob> jobs(count);
    std::vector<int32_t> reals;
    std::vector<char> seen(M.n, 0);
    int rows = 0;
    for (int s = 0; s < count; ++s) {
        const InstanceData &S = sc->inst[s];
        const int g = map[s];
        if (g < 0 || offsets[g + 1] - offsets[g] != S.n)
            return set_error(ACO_ERR_ARG, "aco_pheromone_from_subs: sub-problem " + std::to_string(s) +
                                              " does not match its node list");
        const int32_t *r = real_nodes + offsets[g];
        if (r[0] != 0) return set_error(ACO_ERR_ARG, "aco_pheromone_from_subs: node lists start at the depot");
        for (int i = 1; i < S.n; ++i) {
            if (r[i] <= 0 || r[i] >= M.n || seen[r[i]])
                return set_error(ACO_ERR_ARG, "aco_pheromone_from_subs: customer out of range or in two sub-problems");
            seen[r[i]] = 1;
        }
        jobs[s] = SubTauJob{S.ds.tau, nullptr, S.n, S.P, rows};
        rows += S.n;
        reals.insert(reals.end(), r, r + S.n);
    }
    void *buf = nullptr;
    const size_t jb = (sizeof(SubTauJob) * count + 255) & ~size_t(255);
    HIP_TRY(hipMalloc(&buf, jb + reals.size() * 4));
    int32_t *dreal = reinterpret_cast<int32_t *>(static_cast<unsigned char *>(buf) + jb);
    size_t ro = 0;
    for (int s = 0; s < count; ++s) {
        jobs[s].real = dreal + ro;
        ro += (size_t)jobs[s].n;
    }
    hipError_t e = hipMemcpyAsync(buf, jobs.data(), sizeof(SubTauJob) * count, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dreal, reals.data(), reals.size() * 4, hipMemcpyHostToDevice, c->stream);
    // the sub-colonies' last launches were on their own stream
    if (e == hipSuccess) e = hipStreamSynchronize(sc->stream);
    if (e == hipSuccess)
        e = launch_tau_from_subs(static_cast<const SubTauJob *>(buf), count, rows, M.dp, M.ds, c->stream);
    if (e == hipSuccess && c->bt.max_nk > 0) e = launch_gather(c->bt, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
    (void)hipFree(buf);
    if (e != hipSuccess) return set_error(ACO_ERR_HIP, std::string("aco_pheromone_from_subs: ") + hipGetErrorString(e));
    return ACO_OK;
}

int aco_get_bsf_at(aco_ctx *c, int32_t b, aco_bsf_record *rec, int32_t cap, int32_t *count)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    int nb = 0;
    HIP_TRY(hipMemcpyAsync(&nb, I.ds.ctrl + kCtrlNbsf, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (count) *count = nb;
    const int k = std::min(nb, cap);
    if (rec && k > 0) {
        static_assert(sizeof(aco_bsf_record) == sizeof(BsfRecord), "record layout");
        HIP_TRY(hipMemcpyAsync(rec, I.ds.bsf, (size_t)k * sizeof(BsfRecord), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ACO_OK;
}

int aco_get_bsf(aco_ctx *c, aco_bsf_record *rec, int32_t cap, int32_t *count)
{
    return aco_get_bsf_at(c, 0, rec, cap, count);
}

int aco_get_pheromone_at(aco_ctx *c, int32_t b, float *tau, float *ti)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    if (tau) HIP_TRY(hipMemcpy2DAsync(tau, (size_t)I.n * 4, I.ds.tau, (size_t)I.P * 4, (size_t)I.n * 4, I.n,
                                      hipMemcpyDeviceToHost, c->stream));
    if (ti) HIP_TRY(hipMemcpy2DAsync(ti, (size_t)I.n * 4, I.ds.tinfo, (size_t)I.P * 4, (size_t)I.n * 4, I.n,
                                     hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_pheromone(aco_ctx *c, float *tau, float *ti) { return aco_get_pheromone_at(c, 0, tau, ti); }

int aco_set_pheromone_at(aco_ctx *c, int32_t b, const float *tau, const float *ti)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    if (!tau || !ti) return set_error(ACO_ERR_ARG, "null argument");
    const InstanceData &I = c->inst[b];
    HIP_TRY(hipMemcpy2DAsync(I.ds.tau, (size_t)I.P * 4, tau, (size_t)I.n * 4, (size_t)I.n * 4, I.n,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(I.ds.tinfo, (size_t)I.P * 4, ti, (size_t)I.n * 4, (size_t)I.n * 4, I.n,
                             hipMemcpyHostToDevice, c->stream));
    if ((rc = run_gather(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_pheromone(aco_ctx *c, const float *tau, const float *ti) { return aco_set_pheromone_at(c, 0, tau, ti); }

int aco_get_rng_state(aco_ctx *c, int32_t *state)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!state) return set_error(ACO_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(state, c->inst[0].ds.pm, (size_t)c->m * kLanes * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_setup_at(aco_ctx *c, int32_t b, float *heur, float *d0, int32_t *nnk, float *nnkd, int32_t *nnls,
                     float *nnd, uint16_t *nnpos, uint16_t *dist16)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    const DevProblem &dp = I.dp;
    const int n = I.n, k = c->prm.k, nl = dp.nn_ls;
    if (heur)
        HIP_TRY(hipMemcpy2DAsync(heur, (size_t)n * 4, dp.heur, (size_t)I.P * 4, (size_t)n * 4, n, hipMemcpyDeviceToHost,
                                 c->stream));
    if (d0) HIP_TRY(hipMemcpyAsync(d0, dp.d0, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    if (nnk && k > 0) HIP_TRY(hipMemcpyAsync(nnk, dp.nnk, (size_t)n * k * 4, hipMemcpyDeviceToHost, c->stream));
    if (nnkd && k > 0) HIP_TRY(hipMemcpyAsync(nnkd, dp.nnkd, (size_t)n * k * 4, hipMemcpyDeviceToHost, c->stream));
    if (nnls)
        HIP_TRY(hipMemcpy2DAsync(nnls, (size_t)nl * 4, dp.nnls, (size_t)dp.nnP * 4, (size_t)nl * 4, n,
                                 hipMemcpyDeviceToHost, c->stream));
    if (nnd)
        HIP_TRY(hipMemcpy2DAsync(nnd, (size_t)nl * 4, dp.nnd, (size_t)dp.nnP * 4, (size_t)nl * 4, n,
                                 hipMemcpyDeviceToHost, c->stream));
    if (nnpos) HIP_TRY(hipMemcpyAsync(nnpos, dp.nnpos, (size_t)n * n * 2, hipMemcpyDeviceToHost, c->stream));
    if (dist16 && c->int16_rows)
        HIP_TRY(hipMemcpy2DAsync(dist16, (size_t)n * 2, dp.dist16, (size_t)I.P * 2, (size_t)n * 2, n,
                                 hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_candidates_at(aco_ctx *c, int32_t b, const int32_t *cand, int32_t width)
{
    int rc = need_inst(c, b, false);
    if (rc) return rc;
    const InstanceData &I = c->inst[b];
    const DevProblem &dp = I.dp;
    const int n = I.n, k = dp.k;
    if (k < 1) return set_error(ACO_ERR_STATE, "aco_set_candidates_at: the context has no candidate rows (k = 0)");
    if (!cand || width < k) return set_error(ACO_ERR_ARG, "aco_set_candidates_at: need cand with width >= k");
    std::vector<int> seen(n, -1);
    for (int i = 0; i < n; ++i)
        for (int h = 0; h < k; ++h) {
            const int v = cand[(size_t)i * width + h];
            if (v < 1 || v >= n || v == i || seen[v] == i)
                return set_error(ACO_ERR_ARG, "aco_set_candidates_at: row " + std::to_string(i) +
                                                  " needs k distinct customers other than itself");
            seen[v] = i;
        }
    HIP_TRY(hipMemcpy2DAsync(const_cast<int *>(dp.nnk), (size_t)k * 4, cand, (size_t)width * 4, (size_t)k * 4, n,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_candidate_dists(dp, c->stream));
    // the candidates' total_info, when the pheromone exists already
    if (c->inited && (rc = run_gather(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_rng_state(aco_ctx *c, const int32_t *state)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!state) return set_error(ACO_ERR_ARG, "null state");
    const size_t cnt = (size_t)c->m * kLanes;
    for (size_t q = 0; q < cnt; ++q)
        if (state[q] < 1 || state[q] > 2147483646)
            return set_error(ACO_ERR_ARG, "Park-Miller states must lie in [1, 2^31 - 2]");
    HIP_TRY(hipMemcpyAsync(c->inst[0].ds.pm, state, cnt * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_ls_seed(aco_ctx *c, int32_t *seed)
{
    int rc = need(c, true);
    if (rc) return rc;
    if (!seed) return set_error(ACO_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(seed, c->inst[0].ds.ctrl + kCtrlLsSeed, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_set_ls_seed(aco_ctx *c, int32_t seed)
{
    int rc = need(c, true);
    if (rc) return rc;
    if (seed < 1 || seed > 2147483646) return set_error(ACO_ERR_ARG, "the seed must lie in [1, 2^31 - 2]");
    HIP_TRY(hipMemcpyAsync(c->inst[0].ds.ctrl + kCtrlLsSeed, &seed, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ACO_OK;
}

int aco_get_counters(aco_ctx *c, aco_counters *out)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!out) return set_error(ACO_ERR_ARG, "null output");
    memset(out, 0, sizeof(*out));
    for (size_t b = 0; b < c->inst.size(); ++b) {
        long long t[kTotWords];
        int ctrl[kCtrlWords];
        HIP_TRY(hipMemcpyAsync(t, c->inst[b].ds.totals, sizeof(t), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(ctrl, c->inst[b].ds.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        out->ls_evals += t[kTotEvals];
        out->ls_moves += t[kTotMoves];
        out->fallbacks += t[kTotFallbacks];
        if (b == 0) out->iterations = t[kTotIters];
        out->bsf_overflow += ctrl[kCtrlOverflow];
    }
    return ACO_OK;
}

int aco_profile_enable(aco_ctx *c, int32_t enable)
{
    int rc = need(c, false);
    if (rc) return rc;
    c->profiling = enable != 0;
    return ACO_OK;
}

static void drop_graph(aco_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

static bool shape_fits(const aco_ctx *c, const LsShape &sh)
{
    if (!ls_shape_valid(sh) ||
        local_search_lds_bytes(c->max_n, sh.ring, sh.row16, c->prm.ls_flags) + c->lds_static_ls > c->lds_limit)
        return false;
    if (sh.row16 && !c->int16_rows) return false;
    if (sh.fused && (!fused_iteration_ok(c->bt, c->prm.selection, c->prm.ls_flags, c->m, sh) ||
                     fused_lds_bytes(c->max_n, sh) + c->lds_static_fused > c->lds_limit))
        return false;
    if (!construct_staged_ok(c->bt, c->prm.selection, c->prm.k, c->m, sh.construct_apb)) return false;
    return sh.construct_apb == 0 ||
           construct_staged_rows(c->max_n, c->prm.k, sh.construct_apb, c->bt.cons_lds_budget) >= 0;
}

static aco_launch_shape to_public(const LsShape &sh)
{
    return aco_launch_shape{sh.waves, sh.ring, sh.prefetch, sh.construct_apb, sh.row16, sh.fused};
}

int aco_get_launch_shape(aco_ctx *c, aco_launch_shape *out)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!out) return set_error(ACO_ERR_ARG, "null shape");
    *out = to_public(c->ls);
    return ACO_OK;
}

int aco_set_launch_shape(aco_ctx *c, const aco_launch_shape *in)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (!in) return set_error(ACO_ERR_ARG, "null shape");
    const LsShape sh{in->ls_waves, in->ls_ring, in->ls_prefetch ? 1 : 0, in->construct_apb,
                     in->ls_row16 ? 1 : 0, in->fused_iteration ? 1 : 0};
    if (sh.row16 && !c->int16_rows)
        return set_error(ACO_ERR_ARG, "16-bit rows need every distance to be an integer in [0, 65535]");
    if (!ls_shape_valid(sh))
        return set_error(ACO_ERR_ARG, "invalid shape: waves 4|8|12|16, ring 8|16|32 >= waves+2, "
                                      "prefetch needs ring >= 2*waves+2");
    if (!construct_staged_ok(c->bt, c->prm.selection, c->prm.k, c->m, sh.construct_apb))
        return set_error(ACO_ERR_ARG, "construct_apb must be 0, or 1|2|4 dividing the ants per instance "
                                      "for candidate roulette without a route-length limit");
    if (sh.fused && !fused_iteration_ok(c->bt, c->prm.selection, c->prm.ls_flags, c->m, sh))
        return set_error(ACO_ERR_ARG, "fused_iteration needs candidate roulette, no route-length limit, local "
                                      "search, and 4 or 8 waves dividing the ants per instance");
    if (!shape_fits(c, sh))
        return set_error(ACO_ERR_ARG, "shape needs more LDS than the device's " + std::to_string(c->lds_limit) +
                                          " bytes per block");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ls = sh;
    drop_graph(c);
    return ACO_OK;
}

int aco_autotune(aco_ctx *c, int32_t reps, aco_tune_result *results, int32_t capacity, int32_t *count)
{
    int rc = need(c, true);
    if (rc) return rc;
    if (reps < 1 || reps > 1000) return set_error(ACO_ERR_ARG, "reps must be in [1, 1000]");
    static const int kApb[] = {0, 1, 2, 4};
    static const LsShape kLs[] = {{4, 16, 1, 0, 0},  {4, 16, 0, 0, 0},  {4, 8, 0, 0, 0},   {8, 32, 1, 0, 0},
                                  {8, 16, 0, 0, 0},  {4, 16, 1, 0, 1},  {8, 32, 1, 0, 1},  {8, 16, 0, 0, 1},
                                  {12, 16, 0, 0, 0}, {12, 16, 0, 0, 1}, {12, 32, 1, 0, 1}, {16, 32, 0, 0, 0},
                                  {16, 32, 0, 0, 1}, {4, 8, 0, 0, 1}};
    constexpr int kNc = (int)(sizeof(kApb) / sizeof(kApb[0]));
    constexpr int kNl = (int)(sizeof(kLs) / sizeof(kLs[0]));
    constexpr int kNi = 4;   // stage 2: the per-phase iteration and up to 3 fused shapes
    if (count) *count = kNc + kNl + kNi;
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the mutable per-instance state lies between ds.tau and the end of
    // ds.totals in the instance arena
    struct Span { unsigned char *lo; size_t bytes; };
    std::vector<Span> spans;
    size_t total = 0;
    for (auto &I : c->inst) {
        unsigned char *lo = reinterpret_cast<unsigned char *>(I.ds.tau);
        unsigned char *hi = reinterpret_cast<unsigned char *>(I.ds.totals + kTotWords);
        spans.push_back({lo, (size_t)(hi - lo)});
        total += (size_t)(hi - lo);
    }
    void *orig = nullptr, *built = nullptr;
    HIP_TRY(hipMalloc(&orig, total));
    if (hipMalloc(&built, total) != hipSuccess) {
        (void)hipFree(orig);
        return set_error(ACO_ERR_HIP, "aco_autotune: snapshot allocation failed");
    }
    auto copy_all = [&](void *buf, bool save) -> hipError_t {
        size_t off = 0;
        for (auto &sp : spans) {
            unsigned char *b = static_cast<unsigned char *>(buf) + off;
            hipError_t e = save ? hipMemcpyAsync(b, sp.lo, sp.bytes, hipMemcpyDeviceToDevice, c->stream)
                                : hipMemcpyAsync(sp.lo, b, sp.bytes, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return e;
            off += sp.bytes;
        }
        return hipSuccess;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> tc(kNc, -1.0f), tl(kNl, -1.0f);
    int best_c = -1, best_l = -1;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) err = copy_all(orig, true);
    // mean time of `reps` launches, each from the state saved in `from`
    auto time_it = [&](void *from, auto &&launch, float &out) {
        float sum = 0.0f;
        for (int r = 0; r < reps && err == hipSuccess; ++r) {
            err = copy_all(from, false);
            if (err == hipSuccess) err = hipEventRecord(e0, c->stream);
            if (err == hipSuccess) err = launch();
            if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            sum += ms;
        }
        out = sum / (float)reps;
    };
    // stage 0: construction
    for (int q = 0; q < kNc && err == hipSuccess; ++q) {
        LsShape sh = c->ls;
        sh.construct_apb = kApb[q];
        if (!shape_fits(c, sh)) continue;
        time_it(orig, [&] { return launch_construct(c->bt, c->prm.selection, c->prm.rng, kApb[q], c->prm.k, c->stream); },
                tc[q]);
        if (err == hipSuccess && (best_c < 0 || tc[q] < tc[best_c])) best_c = q;
    }
    // stage 1: local search on the tours of one construction
    if (err == hipSuccess) err = copy_all(orig, false);
    if (err == hipSuccess) err = launch_construct(c->bt, c->prm.selection, c->prm.rng, 0, c->prm.k, c->stream);
    if (err == hipSuccess) err = copy_all(built, true);
    for (int q = 0; q < kNl && err == hipSuccess && c->prm.ls_flags; ++q) {
        LsShape sh = kLs[q];
        sh.construct_apb = 0;
        if (!shape_fits(c, sh)) continue;
        time_it(built, [&] { return launch_local_search(c->bt, c->prm.ls_flags, kLs[q], c->stream, &c->lss); }, tl[q]);
        if (err == hipSuccess && (best_l < 0 || tl[q] < tl[best_l])) best_l = q;
    }
    // stage 2: whole iterations, per-phase launches against the fused launch
    LsShape chosen = c->ls;
    if (best_l >= 0) {
        chosen.waves = kLs[best_l].waves;
        chosen.ring = kLs[best_l].ring;
        chosen.prefetch = kLs[best_l].prefetch;
        chosen.row16 = kLs[best_l].row16;
    }
    if (best_c >= 0) chosen.construct_apb = kApb[best_c];
    chosen.fused = 0;
    LsShape it_sh[kNi];
    float ti[kNi];
    int best_i = -1;
    for (int q = 0; q < kNi; ++q) {
        ti[q] = -1.0f;
        it_sh[q] = chosen;
        it_sh[q].fused = q > 0;
    }
    it_sh[2].waves = 4; it_sh[2].ring = 16; it_sh[2].prefetch = 1;
    it_sh[3].waves = 8; it_sh[3].ring = 16; it_sh[3].prefetch = 0;
    for (int q = 0; q < kNi && err == hipSuccess && c->prm.ls_flags; ++q) {
        const LsShape sh = it_sh[q];
        bool dup = false;
        for (int r = 1; r < q; ++r)
            dup |= it_sh[r].waves == sh.waves && it_sh[r].ring == sh.ring && it_sh[r].prefetch == sh.prefetch;
        if (dup || !shape_fits(c, sh)) continue;
        time_it(orig, [&] {
            if (sh.fused) return launch_iteration_fused(c->bt, c->prm.rng, c->prm.ls_flags, sh, c->tick_to_s, c->stream);
            hipError_t e = launch_construct(c->bt, c->prm.selection, c->prm.rng, sh.construct_apb, c->prm.k, c->stream);
            if (e == hipSuccess) e = launch_local_search(c->bt, c->prm.ls_flags, sh, c->stream, &c->lss);
            if (e == hipSuccess) e = launch_statistics(c->bt, c->tick_to_s, c->stream);
            if (e == hipSuccess) e = launch_pheromone_update(c->bt, c->stream);
            return e;
        }, ti[q]);
        if (err == hipSuccess && (best_i < 0 || ti[q] < ti[best_i])) best_i = q;
    }
    if (best_i > 0) chosen = it_sh[best_i];
    if (err == hipSuccess) err = copy_all(orig, false);
    // the search clock (best-so-far record times) restarts after tuning
    if (err == hipSuccess) err = launch_stamp_clock(c->bt, c->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(orig);
    (void)hipFree(built);
    if (err != hipSuccess) return set_error(ACO_ERR_HIP, std::string("aco_autotune: ") + hipGetErrorString(err));
    c->ls = chosen;
    drop_graph(c);
    int w = 0;
    for (int q = 0; q < kNc && results && w < capacity; ++q, ++w) {
        LsShape sh = chosen;
        sh.construct_apb = kApb[q];
        results[w] = aco_tune_result{to_public(sh), 0, tc[q]};
    }
    for (int q = 0; q < kNl && results && w < capacity; ++q, ++w) {
        LsShape sh = kLs[q];
        sh.construct_apb = chosen.construct_apb;
        results[w] = aco_tune_result{to_public(sh), 1, tl[q]};
    }
    for (int q = 0; q < kNi && results && w < capacity; ++q, ++w)
        results[w] = aco_tune_result{to_public(it_sh[q]), 2, ti[q]};
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

int cvrp::ctx_prior_view(aco_ctx *c, std::vector<CtxInstanceView> &out, void **stream)
{
    int rc = need(c, false);
    if (rc) return rc;
    if (c->inited) return set_error(ACO_ERR_STATE, "the prior must be applied before aco_init");
    out.clear();
    for (auto &I : c->inst) {
        CtxInstanceView v;
        v.n = I.n;
        v.P = I.P;
        v.Q = I.Q;
        v.dist = I.dp.dist;
        v.dem = I.dp.dem;
        v.nnk = c->prm.k > 0 ? I.dp.nnk : nullptr;
        v.k_ctx = c->prm.k;
        v.nnls = I.dp.nnls;
        v.nnP = I.dp.nnP;
        v.nn_ls = I.dp.nn_ls;
        v.heur = const_cast<float *>(I.dp.heur);
        out.push_back(v);
    }
    *stream = (void *)c->stream;
    return ACO_OK;
}
