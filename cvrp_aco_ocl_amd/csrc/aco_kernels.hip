// HIP kernels (gfx950) for one ACO iteration: solution construction,
// iteration statistics, the fused pheromone pass (candidate rows included)
// and the candidate-row gather used where no pass ran.
// Local search lives in aco_local_search.hip.
//
// Arithmetic contract: compiled with -ffp-contract=off, IEEE division and
// denormals kept, so every float result equals the CPU restatement's.
#include "aco_device.h"
#include "aco_ls.h"

namespace cvrp {

using u32 = uint32_t;
using u64 = unsigned long long;

__device__ __forceinline__ u32 f2u(float f) { return __float_as_uint(f); }

// ---------------------------------------------------------------- RNG ----

// Park-Miller minimal standard (kernel.cl:42-52): z <- 16807 z mod (2^31-1)
// by Schrage's method; ran01 = float(z) * 2^-31.
__device__ __forceinline__ float pm_ran01(int &z)
{
    int k = z / 127773;
    z = 16807 * (z - k * 127773) - 2836 * k;
    if (z < 0) z += 2147483647;
    return (float)z * 0x1p-31f;
}

// Philox-4x32-10 (Salmon et al., SC'11), word 0 of ctr=(iter, ant, step, 0),
// top 24 bits -> [0,1).
__device__ __forceinline__ float philox_u01(u64 seed, u32 iter, u32 ant, u32 step)
{
    u32 c0 = iter, c1 = ant, c2 = step, c3 = 0u;
    u32 k0 = (u32)seed, k1 = (u32)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const u32 hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const u32 hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const u32 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return (float)(c0 >> 8) * 0x1p-24f;
}

// Philox draws of 64 consecutive steps at once: lane l holds the draw of
// step first + l (the same counters as one draw per step).
__device__ __forceinline__ float philox_block(u64 seed, u32 iter, u32 ant, u32 first, int lane)
{
    return philox_u01(seed, iter, ant, first + (u32)lane);
}

// ------------------------------------------------------- wave helpers ----

__device__ __forceinline__ u64 wave_max_u64(u64 v) { return wave_reduce_u64<true>(v); }

__device__ __forceinline__ u64 wave_min_u64(u64 v) { return wave_reduce_u64<false>(v); }

// Value of the lane selected by a DPP control (0 where the source lane is
// outside the row or the row is masked off).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_src(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xF, false));
}

// Inclusive wave scan on the DPP network (no LDS traffic): Hillis-Steele
// inside each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's total into
// row 1 and row 2's into row 3 (row_bcast:15), then lane 31's prefix into
// rows 2 and 3 (row_bcast:31).  The CPU restatement adds in the same order.
// Adding 0 to a non-negative value leaves it unchanged.
__device__ __forceinline__ float wave_scan_f32(float x, int lane)
{
    (void)lane;
    x = x + dpp_src<0x111, 0xF>(x);
    x = x + dpp_src<0x112, 0xF>(x);
    x = x + dpp_src<0x114, 0xF>(x);
    x = x + dpp_src<0x118, 0xF>(x);
    x = x + dpp_src<0x142, 0xA>(x);
    x = x + dpp_src<0x143, 0xC>(x);
    return x;
}

__device__ __forceinline__ float lane_f32(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
__device__ __forceinline__ int lane_i32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ------------------------------------------------------- construction ----

// Shared per-ant LDS: demands, depot distances, (I-Roulette) reduction rank
// and the visited flags.
struct AntLds {
    int *dem;
    float *d0;
    int *rank;
    unsigned char *vis;
};

__device__ __forceinline__ AntLds ant_lds(unsigned char *smem, int n, bool with_rank)
{
    AntLds L;
    L.dem = reinterpret_cast<int *>(smem);
    L.d0 = reinterpret_cast<float *>(L.dem + n);
    L.rank = reinterpret_cast<int *>(L.d0 + n);
    L.vis = reinterpret_cast<unsigned char *>(with_rank ? L.rank + n : L.rank);
    return L;
}

// The instance's demands, depot distances and (with rank) reduction ranks
// copied into LDS by one wave, visited flags set to the depot alone:
// kCopyUnroll nodes per lane have their loads issued before their stores
// (a loop storing each trip would wait for that trip's loads in turn).
constexpr int kCopyUnroll = 8;
template <bool RANK>
__device__ __forceinline__ void ant_lds_fill(const DevProblem &p, const AntLds &L, int n, int lane)
{
    for (int i0 = lane; i0 < n; i0 += kCopyUnroll * kLanes) {
        int dm[kCopyUnroll], rk[kCopyUnroll];
        float dz[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const int i = i0 + u * kLanes, ic = i < n ? i : n - 1;
            dm[u] = p.dem[ic];
            dz[u] = p.d0[ic];
            rk[u] = RANK ? p.rank[ic] : 0;
        }
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const int i = i0 + u * kLanes;
            if (i < n) {
                L.dem[i] = dm[u];
                L.d0[i] = dz[u];
                if (RANK) L.rank[i] = rk[u];
                L.vis[i] = i == 0;
            }
        }
    }
}


size_t construct_lds_bytes(int n, int selection)
{
    return (size_t)n * (4 + 4 + (selection == kSelIRoulette ? 4 : 0) + 1) + 16;
}

// The four-wave I-Roulette kernel's exchange of wave maxima (2 x 4 u64) after
// the per-ant arrays, 8-byte aligned.
__host__ __device__ __forceinline__ size_t iroulette4_exchange_offset(int n)
{
    return ((size_t)n * 13 + 16 + 7) & ~size_t(7);
}

// kernel.cl:445-452 mask: not visited, load + d_i <= Q and, with a route
// length limit, rdist + (d(cur,i) + svc) + d(i,0) <= D; every condition
// evaluated, without branches, so the loads feeding it are issued together.
template <bool HAS_D>
__device__ __forceinline__ bool feasible_mask(const DevProblem &p, const AntLds &L, int i, int demi, float dci,
                                              int load, float rdist)
{
    bool f = (L.vis[i] == 0) & (load + demi <= p.Q);
    if (HAS_D) {
        float t = rdist + (dci + p.svc);
        t = t + L.d0[i];
        f = f & (t <= p.D);
    }
    return f;
}

// Full-row greedy pick when the roulette has no positive mass: largest
// total_info among feasible nodes, ties to the lowest index; 0 = none.
// The visited flags and capacity test come from LDS first; a pass whose
// nodes are all infeasible reads nothing from global memory (a route that
// is full returns to the depot without touching the row), otherwise the
// row values (and, with a length limit, distances) are loaded together,
// one memory latency per pass.  Two shapes:
//   WIDE   16 nodes per lane per pass (1024 per pass), every load of the
//          pass issued (clamped index, masked afterwards): fewest round
//          trips, for a device with spare memory bandwidth (few ants)
//   narrow 8 nodes per lane per pass, a chunk of 64 nodes none of which is
//          feasible loads the row's first element instead (one cached
//          line): fewer bytes, for a busy device
template <bool HAS_D, bool WIDE>
__device__ __forceinline__ void greedy_pass(const DevProblem &p, const AntLds &L, const float *tirow,
                                            const float *drow, int load, float rdist, int lane, int base,
                                            u64 &best)
{
    constexpr int CH = WIDE ? 16 : 8;
    unsigned fm = 0;   // bit u: node base + u*64 + lane passes the LDS tests
    unsigned need = 0; // bit u: some lane of chunk u does (uniform)
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int i = base + u * kLanes + lane;
        const int ic = i < p.n ? i : p.n - 1;   // in-row index, masked below
        const bool f = (i < p.n) & (L.vis[ic] == 0) & (load + L.dem[ic] <= p.Q);
        fm |= f ? 1u << u : 0u;
        if (!WIDE) need |= __ballot(f) ? 1u << u : 0u;
    }
    if (WIDE ? !__ballot(fm != 0) : !need) return;
    float tv[CH], dv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int i = base + u * kLanes + lane;
        const int ia = WIDE ? (i < p.n ? i : p.n - 1) : (((need >> u) & 1u) && i < p.n ? i : 0);
        tv[u] = gload(tirow + ia);
        dv[u] = HAS_D ? gload(drow + ia) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int i = base + u * kLanes + lane;
        bool f = (fm >> u) & 1u;
        if (HAS_D) {
            float t = rdist + (dv[u] + p.svc);
            t = t + L.d0[i < p.n ? i : p.n - 1];
            f = f & (t <= p.D);
        }
        if (f) {
            const u64 key = ((u64)(f2u(tv[u]) + 1u) << 32) | (u64)(0xFFFFFFFFu - (u32)i);
            best = key > best ? key : best;
        }
    }
}

template <bool HAS_D, bool WIDE = false, bool PEEL = WIDE>
__device__ __forceinline__ int greedy_pick(const DevProblem &p, const AntLds &L, const float *tirow,
                           const float *drow, int load, float rdist, int lane)
{
    constexpr int PASS = (WIDE ? 16 : 8) * kLanes;
    u64 best = 0;
    // PEEL: the first pass on its own, with constant offsets from the row's
    // start and scalar-based loads (faster for the large instances of the
    // wide shape and the staged kernel; one loop body suits the small
    // instances of a batch better)
    if (PEEL) greedy_pass<HAS_D, WIDE>(p, L, tirow, drow, load, rdist, lane, 0, best);
    for (int base = PEEL ? PASS : 0; base < p.n; base += PASS)
        greedy_pass<HAS_D, WIDE>(p, L, tirow, drow, load, rdist, lane, base, best);
    best = wave_max_u64(best);
    return best ? (int)(0xFFFFFFFFu - (u32)(best & 0xFFFFFFFFull)) : 0;
}

// Tour positions are buffered one per lane (position q lives in lane q&63)
// and written 64 at a time.
__device__ __forceinline__ void tour_put(int *tour, int &tbuf, int step, int w, int lane)
{
    if (lane == (step & 63)) tbuf = w;
    if ((step & 63) == 63) tour[step - 63 + lane] = tbuf;
}

__device__ __forceinline__ void tour_close(int *tour, int &tbuf, int step, int lane)
{
    if (lane == (step & 63)) tbuf = 0;
    if (lane <= (step & 63)) tour[(step & ~63) + lane] = tbuf;
}

// Tour length after construction: the edges d(tour[q-1], tour[q]), q = 1..
// last, added one by one in step order -- the same additions the step loop
// would make -- with the loads issued 64 at a time.
__device__ __forceinline__ float tour_length_in_order(const DevProblem &p, const int *tour, int last,
                                                      int lane)
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");   // this wave's tour stores, fresh L1
    // the loads of kLenTrips chunks of 64 edges are issued together (a
    // position past the last reads edge 1 instead; its value is not added)
    constexpr int kLenTrips = 4;
    float t = 0.0f;
    for (int base0 = 1; base0 <= last; base0 += kLenTrips * kLanes) {
        float d[kLenTrips];
#pragma unroll
        for (int g = 0; g < kLenTrips; ++g) {
            const int q = base0 + g * kLanes + lane, qc = q <= last ? q : 1;
            d[g] = p.dist[(size_t)tour[qc - 1] * p.P + tour[qc]];
        }
#pragma unroll
        for (int g = 0; g < kLenTrips; ++g) {
            const int base = base0 + g * kLanes;
            const int cnt = last - base + 1 < kLanes ? last - base + 1 : kLanes;
            for (int l = 0; l < cnt; ++l) t = t + lane_f32(d[g], l);
        }
    }
    return t;
}

// construct_solution with I-Roulette (kernel.cl:325-373, 382-493): one
// wavefront per ant, lane l draws for nodes l, l+64, ... from its own
// Park-Miller stream; the reduction's tie order is reproduced by taking the
// max of (prob, 4095-rank, node) keys.
// Block -> (instance, ant) of a batched launch.
#define ANT_PROLOGUE                                        \
    const int g_ = blockIdx.x;                              \
    const int inst_ = bt.ant_inst[g_];                      \
    const DevProblem p = bt.probs[inst_];                   \
    const DevState s = bt.states[inst_];                    \
    const int a = g_ - bt.ant_base[inst_];

constexpr int kIrChunk = 16;   // I-Roulette: nodes per lane whose loads are issued together

// C nodes of one lane for an I-Roulette pass: their row values (and
// distances), visited flags, demands, ranks (and depot distances), all
// requested before any is used (clamped in-row indices; nodes >= n are
// skipped by the caller), then the key of kernel.cl's reduction,
// (prob, 4095 - rank, node), with the feasibility mask of kernel.cl:445-452
// evaluated without branches.
template <bool HAS_D, int C>
struct IrNodes {
    float tv[C], dv[C], d0[C];
    int vis[C], dem[C], rank[C];
    __device__ __forceinline__ void load(const DevProblem &p, const AntLds &L, const float *tirow,
                                         const float *drow, int j, int i)
    {
        const int ic = i < p.n ? i : p.n - 1;
        tv[j] = gload(tirow + ic);
        dv[j] = HAS_D ? gload(drow + ic) : 0.0f;
        d0[j] = HAS_D ? L.d0[ic] : 0.0f;
        vis[j] = L.vis[ic];
        dem[j] = L.dem[ic];
        rank[j] = L.rank[ic];
    }
    __device__ __forceinline__ u64 key(const DevProblem &p, int j, int i, float r, int load, float rdist,
                                       u64 best) const
    {
        const float pp = tv[j] * 100.0f;
        bool c = (vis[j] == 0) & (load + dem[j] <= p.Q);
        if (HAS_D) {
            float t = rdist + (dv[j] + p.svc);
            t = t + d0[j];
            c = c & (t <= p.D);
        }
        const float prob = c ? r * pp : 0.0f;
        const u64 k = ((u64)f2u(prob) << 32) | ((u64)(4095 - rank[j]) << 12) | (u64)i;
        return k > best ? k : best;
    }
};

template <bool HAS_D>
__global__ __launch_bounds__(64) void k_construct_iroulette(DevBatch bt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ANT_PROLOGUE
    const int n = p.n, lane = threadIdx.x;
    AntLds L = ant_lds(smem, n, true);
    ant_lds_fill<true>(p, L, n, lane);
    __syncthreads();
    int z = s.pm[a * kLanes + lane];
    int *tour = s.tours + (size_t)a * p.T;
    int cur = 0, load = 0, vcnt = 1, step = 0, fb = 0, tbuf = 0;
    float rdist = 0.0f, tdist = 0.0f;
    while (vcnt < n) {
        const float *tirow = s.tinfo + (size_t)cur * p.P;
        const float *drow = p.dist + (size_t)cur * p.P;
        // the row's values (and distances) for 16 nodes per lane are loaded
        // together with their LDS fields, then the lane draws for them in node
        // order: one memory latency per 1024 nodes instead of one per node
        u64 best = 0;
        for (int base = 0; base < n; base += kIrChunk * kLanes) {
            IrNodes<HAS_D, kIrChunk> F;
#pragma unroll
            for (int u = 0; u < kIrChunk; ++u) F.load(p, L, tirow, drow, u, base + u * kLanes + lane);
#pragma unroll
            for (int u = 0; u < kIrChunk; ++u) {
                const int i = base + u * kLanes + lane;
                if (i < n) {
                    const float r = pm_ran01(z) + 0.001f;
                    best = F.key(p, u, i, r, load, rdist, best);
                }
            }
        }
        best = wave_max_u64(best);
        step++;
        int w;
        if ((best >> 32) == 0) {
            if (cur == 0) {
                w = greedy_pick<HAS_D>(p, L, tirow, drow, load, rdist, lane);
                fb++;
            } else {
                w = 0;
            }
        } else {
            w = (int)(best & 0xFFFull);
        }
        w = uniform(w);
        // without a length limit the loop carries no edge lengths (summed
        // afterwards in the same order), so no step waits for one
        const float dcw = HAS_D ? drow[w] : 0.0f;
        if (w == 0) {
            load = 0;
            rdist = 0.0f;
        } else {
            load += L.dem[w];
            rdist = rdist + (dcw + p.svc);
            vcnt++;
        }
        if (lane == 0) L.vis[w] = 1;
        tdist = tdist + dcw;
        tour_put(tour, tbuf, step, w, lane);
        cur = w;
        __syncthreads();
    }
    tdist = tdist + L.d0[cur];
    step++;
    tour_close(tour, tbuf, step, lane);
    if (!HAS_D) tdist = tour_length_in_order(p, tour, step, lane);
    if (lane == 0) {
        s.sizes[a] = step + 1;
        s.lens[a] = tdist;
        s.ant_fb[a] = fb;
        s.ant_evals[a] = 0;
        s.ant_moves[a] = 0;
    }
    s.pm[a * kLanes + lane] = z;
}

// The same construction with W = 4 or 8 waves per ant, for launches with at
// most one ant per CU (otherwise the single-wave kernel keeps the CU's other
// SIMDs busy with other ants).  A lane's draws stay its own Park-Miller
// stream: node l + 64u uses the stream's (u+1)-th draw of the step,
// z * 16807^(u+1) mod (2^31 - 1), so wave w takes u = w, w+W, ... starting
// from the jump 16807^(w+1) and stepping by 16807^W; after the step every
// wave advances the lane's state by the lane's node count.  The same keys as the
// single-wave kernel; the waves' maxima meet in LDS (one barrier per step,
// two buffers by step parity), and every wave sets the visited flag itself,
// so it reads its own write in the next step.
__device__ __forceinline__ int pm_mulmod(int z, unsigned a)
{
    // z * a mod (2^31 - 1) for z, a in [1, 2^31 - 2]: 2^31 = 1 (mod 2^31 - 1),
    // and the product is never 0 mod the prime, so one subtraction suffices
    const u64 prod = (u64)(u32)z * (u64)a;
    u32 r = (u32)(prod & 0x7FFFFFFFull) + (u32)(prod >> 31);
    r = r >= 0x7FFFFFFFu ? r - 0x7FFFFFFFu : r;
    return (int)r;
}

// Waves per ant: 4, or 8 when the instance has more than 512 nodes (two
// waves per SIMD hide more of each step's latency; on smaller rows most of
// an eighth-wave's lanes would have no node).
constexpr int kIrWavesMax = 8;
__host__ __device__ constexpr int iroulette_waves(int n) { return n > 512 ? 8 : 4; }

template <bool HAS_D, int kIrWaves>
__global__ __launch_bounds__(kIrWaves * 64) void k_construct_iroulette4(DevBatch bt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ANT_PROLOGUE
    const int n = p.n, tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    AntLds L = ant_lds(smem, n, true);
    u64 *sh = reinterpret_cast<u64 *>(smem + iroulette4_exchange_offset(n));   // [2][kIrWaves]
    for (int i = tid; i < n; i += kIrWaves * kLanes) {
        L.dem[i] = p.dem[i];
        L.d0[i] = p.d0[i];
        L.rank[i] = p.rank[i];
        L.vis[i] = i == 0;
    }
    __syncthreads();
    int z = s.pm[a * kLanes + lane];
    const unsigned a4 = (unsigned)pm_jump(1, kIrWaves), aw = (unsigned)pm_jump(1, (unsigned long long)w + 1);
    const int cnt = lane < n ? (n - lane + kLanes - 1) / kLanes : 0;   // the lane's nodes (draws) per step
    const unsigned acnt = (unsigned)pm_jump(1, (unsigned long long)cnt);
    int *tour = s.tours + (size_t)a * p.T;
    int cur = 0, load = 0, vcnt = 1, step = 0, fb = 0, tbuf = 0;
    float rdist = 0.0f, tdist = 0.0f;
    constexpr int C = kIrChunk / kIrWaves;   // a wave's nodes per lane per pass
    while (vcnt < n) {
        const float *tirow = s.tinfo + (size_t)cur * p.P;
        const float *drow = p.dist + (size_t)cur * p.P;
        u64 best = 0;
        int zz = pm_mulmod(z, aw);
        // pass over u = base + w + Wj (node lane + 64u), j < C
        for (int ub = 0; ub * kLanes < n; ub += kIrChunk) {
            IrNodes<HAS_D, C> F;
#pragma unroll
            for (int j = 0; j < C; ++j) F.load(p, L, tirow, drow, j, lane + kLanes * (ub + w + kIrWaves * j));
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const int i = lane + kLanes * (ub + w + kIrWaves * j);
                if (i < n) {
                    const float r = (float)zz * 0x1p-31f + 0.001f;
                    zz = pm_mulmod(zz, a4);
                    best = F.key(p, j, i, r, load, rdist, best);
                }
            }
        }
        best = wave_max_u64(best);
        u64 *ex = sh + (step & 1) * kIrWaves;
        if (lane == 0) ex[w] = best;
        __syncthreads();
#pragma unroll
        for (int v = 0; v < kIrWaves; ++v) {
            const u64 o = ex[v];
            best = o > best ? o : best;
        }
        z = pm_mulmod(z, acnt);
        step++;
        int wn;
        if ((best >> 32) == 0) {
            if (cur == 0) {
                wn = greedy_pick<HAS_D>(p, L, tirow, drow, load, rdist, lane);
                fb++;
            } else {
                wn = 0;
            }
        } else {
            wn = (int)(best & 0xFFFull);
        }
        wn = uniform(wn);
        const float dcw = HAS_D ? drow[wn] : 0.0f;   // without a limit: summed afterwards
        if (wn == 0) {
            load = 0;
            rdist = 0.0f;
        } else {
            load += L.dem[wn];
            rdist = rdist + (dcw + p.svc);
            vcnt++;
        }
        if (lane == 0) L.vis[wn] = 1;
        tdist = tdist + dcw;
        if (w == 0) tour_put(tour, tbuf, step, wn, lane);
        cur = wn;
    }
    tdist = tdist + L.d0[cur];
    step++;
    if (w == 0) {
        tour_close(tour, tbuf, step, lane);
        if (!HAS_D) tdist = tour_length_in_order(p, tour, step, lane);
        if (lane == 0) {
            s.sizes[a] = step + 1;
            s.lens[a] = tdist;
            s.ant_fb[a] = fb;
            s.ant_evals[a] = 0;
            s.ant_moves[a] = 0;
        }
        s.pm[a * kLanes + lane] = z;
    }
}

// Roulette over the k-nearest candidate row: lane c holds candidate c's
// total_info masked by feasibility, a wavefront scan gives the wheel,
// r = u * wheel[k-1] and the first lane whose prefix exceeds r wins (the
// last lane with mass when rounding leaves none).  No mass -> full-row
// greedy pick, and the depot when nothing is feasible.
// One ant by one wavefront (lane = its lane); L holds the instance's
// demands and depot distances and this ant's visited flags (depot set).
// The step loop synchronises the wave only, so the waves of a block may
// build different ants.
template <bool HAS_D, int RNG, bool WIDE = false>
__device__ void construct_knn_ant(const DevProblem &p, const DevState &s, int a, const AntLds &L, int lane)
{
    const int n = p.n, k = p.k;
    const u64 seed = p.seed;
    const u32 iter = (u32)s.ctrl[kCtrlIter];
    int z = RNG == kRngParkMiller ? s.pm[a * kLanes] : 0;  // lane 0's stream, held by all lanes
    int *tour = s.tours + (size_t)a * p.T;
    int cur = 0, load = 0, vcnt = 1, step = 0, fb = 0, tbuf = 0;
    float rdist = 0.0f, tdist = 0.0f;
    float ublk = 0.0f;   // Philox draws of steps step+1 .. step+64, refreshed every 64 steps
    while (vcnt < n) {
        step++;
        if (RNG == kRngPhilox && ((step - 1) & 63) == 0) ublk = philox_block(seed, iter, (u32)a, (u32)step, lane);
        // candidate id, its total_info and (with a length limit) d(cur, id):
        // one contiguous read each, issued together; lanes >= k read the row's
        // last entry and are masked.  total_info is finite and >= 0, so the
        // product by 1 or 0 is exact and keeps its load unconditional.
        const int e = cur * k + (lane < k ? lane : k - 1);
        const int j = gload(p.nnk + e);
        const float tj = gload(s.tiknn + e);
        const float dcj = HAS_D ? gload(p.nnkd + e) : 0.0f;
        const int dj = L.dem[j];   // the chosen lane's value is read out below
        const bool ok = (lane < k) & feasible_mask<HAS_D>(p, L, j, dj, dcj, load, rdist);
        const float pj = tj * (ok ? 1.0f : 0.0f);
        const float x = wave_scan_f32(pj, lane);
        const float total = lane_f32(x, k - 1);
        int w, dw;
        float dcw;
        if (total > 0.0f) {
            const float u = RNG == kRngPhilox ? lane_f32(ublk, (step - 1) & 63) : pm_ran01(z);
            const float r = u * total;
            // only lanes with mass: the scan's additions are grouped by the DPP
            // tree, so a lane without mass can round above its left
            // neighbour's prefix and must not be taken
            const u64 m1 = __ballot(pj > 0.0f && x > r);
            int wc;
            if (m1) {
                wc = __ffsll((unsigned long long)m1) - 1;
            } else {
                const u64 m2 = __ballot(pj > 0.0f);
                wc = 63 - __clzll((long long)m2);
            }
            w = lane_i32(j, wc);
            dw = lane_i32(dj, wc);
            dcw = HAS_D ? lane_f32(dcj, wc) : 0.0f;
        } else {
            const float *drow = p.dist + (size_t)cur * p.P;
            w = greedy_pick<HAS_D, WIDE>(p, L, s.tinfo + (size_t)cur * p.P, drow, load, rdist, lane);
            fb++;
            w = uniform(w);
            // scalar reads keep their waits on this path (see the staged
            // kernel)
            dw = uniform(L.dem[w]);
            dcw = HAS_D ? drow[w] : 0.0f;
        }
        w = uniform(w);
        if (w == 0) {
            load = 0;
            rdist = 0.0f;
        } else {
            load += dw;
            rdist = rdist + (dcw + p.svc);
            vcnt++;
        }
        if (lane == 0) L.vis[w] = 1;
        tdist = tdist + dcw;
        tour_put(tour, tbuf, step, w, lane);
        cur = w;
        // lane 0's flag is read by the other lanes next step: LDS accesses of
        // one wave complete in order
        __builtin_amdgcn_wave_barrier();
    }
    if (HAS_D) tdist = tdist + L.d0[cur];
    step++;
    tour_close(tour, tbuf, step, lane);
    // without a length limit the step loop carries no edge lengths: the same
    // additions in the same order afterwards
    if (!HAS_D) tdist = tour_length_in_order(p, tour, step, lane);
    if (lane == 0) {
        s.sizes[a] = step + 1;
        s.lens[a] = tdist;
        s.ant_fb[a] = fb;
        s.ant_evals[a] = 0;
        s.ant_moves[a] = 0;
        if (RNG == kRngParkMiller) s.pm[a * kLanes] = z;
    }
}

template <bool HAS_D, int RNG, bool WIDE>
__global__ __launch_bounds__(64) void k_construct_knn(DevBatch bt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ANT_PROLOGUE
    const int n = p.n, lane = threadIdx.x;
    AntLds L = ant_lds(smem, n, false);
    ant_lds_fill<false>(p, L, n, lane);
    __syncthreads();
    construct_knn_ant<HAS_D, RNG, WIDE>(p, s, a, L, lane);
}

// The same roulette with the instance's candidate table staged in LDS, for
// instances without a route-length limit and with 16-bit demands: A ants
// per block (one wave each, no block-wide barrier after staging) share the
// candidate entries {total_info, id | demand << 16}, 8 bytes each, so a
// step reads its candidates' values, ids and demands with one LDS load and
// then their visited flags with a second.  When the whole n*k table does not
// fit, the first `rows_cap` rows are staged and a step from a later node
// reads its row from the candidate arrays in global memory (ids and values
// together, then the demands from LDS), as the unstaged kernel does; the
// branch is wave-uniform.  The chosen edge's length is fetched after the
// pick and added one step later (the same additions in the same order).
// LDS: dem[n] | vis[A][pitch] | ent[rows * k].
__device__ __forceinline__ int construct_vis_pitch(int n) { return (n + 15) & ~15; }

template <int RNG, int A>
__global__ __launch_bounds__(64 * A) void k_construct_knn_lds(DevBatch bt, int rows_cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wv = wave_id(), lane = threadIdx.x & 63;
    const int inst = bt.ant_inst[blockIdx.x * A];       // all A ants in one instance
    const DevProblem p = bt.probs[inst];
    const DevState s = bt.states[inst];
    const int a = blockIdx.x * A + wv - bt.ant_base[inst];
    const int n = p.n, k = p.k;
    const int rows = n < rows_cap ? n : rows_cap;       // staged rows 0 .. rows-1
    const int nmax = bt.max_n, pitch = construct_vis_pitch(nmax);
    int *dem = reinterpret_cast<int *>(smem);
    unsigned char *vis_all = reinterpret_cast<unsigned char *>(dem + ((nmax + 3) & ~3));
    unsigned char *vis = vis_all + wv * pitch;
    uint2 *ent = reinterpret_cast<uint2 *>(vis_all + A * pitch);
    for (int i = threadIdx.x; i < n; i += 64 * A) dem[i] = p.dem[i];
    for (int i = lane; i < n; i += kLanes) vis[i] = i == 0;
    __syncthreads();
    // entries kStageUnroll per thread at a time: ids and values loaded
    // together, then the demands of those ids from LDS
    constexpr int kStageUnroll = 8;
    const int nent = rows * k;
    for (int e0 = threadIdx.x; e0 < nent; e0 += 64 * A * kStageUnroll) {
        int id[kStageUnroll];
        unsigned tv[kStageUnroll];
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int e = e0 + u * 64 * A;
            const int ec = e < nent ? e : nent - 1;   // in-range index, store masked below
            id[u] = gload(p.nnk + ec);
            tv[u] = __float_as_uint(gload(s.tiknn + ec));
        }
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int e = e0 + u * 64 * A;
            if (e < nent) ent[e] = make_uint2(tv[u], (unsigned)id[u] | ((unsigned)dem[id[u]] << 16));
        }
    }
    __syncthreads();
    const AntLds L{dem, nullptr, nullptr, vis};
    const u64 seed = p.seed;
    const u32 iter = (u32)s.ctrl[kCtrlIter];
    int z = RNG == kRngParkMiller ? s.pm[a * kLanes] : 0;
    int *tour = s.tours + (size_t)a * p.T;
    int cur = 0, load = 0, vcnt = 1, step = 0, fb = 0, tbuf = 0;
    float ublk = 0.0f;   // Philox draws of steps step+1 .. step+64, refreshed every 64 steps
    const int c = lane < k ? lane : k - 1;   // lanes >= k read the row's last entry and are masked
    while (vcnt < n) {
        step++;
        if (RNG == kRngPhilox && ((step - 1) & 63) == 0) ublk = philox_block(seed, iter, (u32)a, (u32)step, lane);
        int j, dj;
        float tj;
        if (cur < rows) {
            const uint2 en = ent[cur * k + c];
            j = (int)(en.y & 0xFFFFu);
            dj = (int)(en.y >> 16);
            tj = __uint_as_float(en.x);
        } else {
            const int e = cur * k + c;
            j = gload(p.nnk + e);
            tj = gload(s.tiknn + e);
            dj = dem[j];
        }
        const bool ok = (lane < k) & (vis[j] == 0) & (load + dj <= p.Q);   // feasible_mask<false>
        // total_info is finite and >= 0, so x * 1 = x and x * 0 = 0 exactly; the
        // product keeps its load unconditional (one 8-byte LDS read with the id)
        const float pj = tj * (ok ? 1.0f : 0.0f);
        const float x = wave_scan_f32(pj, lane);
        const float total = lane_f32(x, k - 1);
        int w, dw;
        if (total > 0.0f) {
            const float u = RNG == kRngPhilox ? lane_f32(ublk, (step - 1) & 63) : pm_ran01(z);
            const float r = u * total;
            // only lanes with mass: the scan's additions are grouped by the DPP
            // tree, so a lane without mass can round above its left
            // neighbour's prefix and must not be taken
            const u64 m1 = __ballot(pj > 0.0f && x > r);
            int wc;
            if (m1) {
                wc = __ffsll((unsigned long long)m1) - 1;
            } else {
                const u64 m2 = __ballot(pj > 0.0f);
                wc = 63 - __clzll((long long)m2);
            }
            w = lane_i32(j, wc);
            dw = lane_i32(dj, wc);
        } else {
            w = greedy_pick<false, false, true>(p, L, s.tinfo + (size_t)cur * p.P, nullptr, load, 0.0f, lane);
            fb++;
            w = uniform(w);
            // scalar, so the wait for this read stays on this path: a vector
            // result would place it where the paths join, where every step
            // would also wait for the visited-flag store
            dw = uniform(dem[w]);
        }
        w = uniform(w);
        if (w == 0) {
            load = 0;
        } else {
            load += dw;
            vcnt++;
        }
        // lane 0's flag is read by the other lanes next step: LDS accesses of
        // one wave complete in order, and the store and the loads use the
        // same array, so the compiler keeps their order
        if (lane == 0) vis[w] = 1;
        tour_put(tour, tbuf, step, w, lane);
        cur = w;
        __builtin_amdgcn_wave_barrier();
    }
    step++;
    tour_close(tour, tbuf, step, lane);
    const float tdist = tour_length_in_order(p, tour, step, lane);
    if (lane == 0) {
        s.sizes[a] = step + 1;
        s.lens[a] = tdist;
        s.ant_fb[a] = fb;
        s.ant_evals[a] = 0;
        s.ant_moves[a] = 0;
        if (RNG == kRngParkMiller) s.pm[a * kLanes] = z;
    }
}

// LDS of the staged kernel for instances of up to n nodes: the demand and
// visited arrays, then as many candidate rows as fit in `budget` bytes (all
// n when they do).  Returns 0 when not even the fixed part fits.
static size_t construct_fixed_bytes(int n, int apb)
{
    return (size_t)((n + 3) & ~3) * 4 + (size_t)apb * ((n + 15) & ~15) + 16;
}

int construct_staged_rows(int n, int k, int apb, size_t budget)
{
    const size_t fixed = construct_fixed_bytes(n, apb);
    if (k <= 0 || fixed > budget) return -1;
    const size_t rows = (budget - fixed) / ((size_t)k * 8);
    return rows < (size_t)n ? (int)rows : n;
}

size_t construct_lds_staged_bytes(int n, int k, int apb, size_t budget)
{
    const int rows = construct_staged_rows(n, k, apb, budget);
    return rows < 0 ? (size_t)-1 : construct_fixed_bytes(n, apb) + (size_t)rows * k * 8;
}

bool construct_staged_ok(const DevBatch &b, int selection, int k, int ants_per_instance, int apb)
{
    if (apb == 0) return true;
    if (apb != 1 && apb != 2 && apb != 4) return false;
    return selection == kSelKnn && !b.has_d && b.dem16 && k > 0 && ants_per_instance % apb == 0;
}

static hipError_t max_static_lds(const void *const *fns, int count, size_t &bytes)
{
    bytes = 0;
    for (int i = 0; i < count; ++i) {
        hipFuncAttributes a;
        const hipError_t e = hipFuncGetAttributes(&a, fns[i]);
        if (e != hipSuccess) return e;
        bytes = a.sharedSizeBytes > bytes ? a.sharedSizeBytes : bytes;
    }
    return hipSuccess;
}

hipError_t construct_static_lds(size_t &bytes)
{
    const void *fns[] = {(const void *)k_construct_iroulette<false>, (const void *)k_construct_iroulette<true>,
                         (const void *)k_construct_iroulette4<false, 4>, (const void *)k_construct_iroulette4<true, 4>,
                         (const void *)k_construct_iroulette4<false, 8>, (const void *)k_construct_iroulette4<true, 8>,
                         (const void *)k_construct_knn<false, kRngPhilox, true>,
                         (const void *)k_construct_knn<true, kRngPhilox, true>,
                         (const void *)k_construct_knn<false, kRngParkMiller, true>,
                         (const void *)k_construct_knn<true, kRngParkMiller, true>,
                         (const void *)k_construct_knn<false, kRngPhilox, false>,
                         (const void *)k_construct_knn<true, kRngPhilox, false>,
                         (const void *)k_construct_knn<false, kRngParkMiller, false>,
                         (const void *)k_construct_knn<true, kRngParkMiller, false>,
                         (const void *)k_construct_knn_lds<kRngPhilox, 1>, (const void *)k_construct_knn_lds<kRngPhilox, 2>,
                         (const void *)k_construct_knn_lds<kRngPhilox, 4>,
                         (const void *)k_construct_knn_lds<kRngParkMiller, 1>,
                         (const void *)k_construct_knn_lds<kRngParkMiller, 2>,
                         (const void *)k_construct_knn_lds<kRngParkMiller, 4>};
    return max_static_lds(fns, (int)(sizeof(fns) / sizeof(fns[0])), bytes);
}

hipError_t launch_construct(const DevBatch &b, int selection, int rng, int apb, int k, hipStream_t st)
{
    if (apb > 0) {
        const int rows = construct_staged_rows(b.max_n, k, apb, b.cons_lds_budget);
        if (rows < 0) return hipErrorInvalidValue;
        const size_t lds = construct_lds_staged_bytes(b.max_n, k, apb, b.cons_lds_budget);
        const dim3 grid(b.ants / apb), block(kLanes * apb);
#define CKL(R, A) hipLaunchKernelGGL((k_construct_knn_lds<R, A>), grid, block, lds, st, b, rows)
        if (rng == kRngPhilox) {
            if (apb == 1) CKL(kRngPhilox, 1); else if (apb == 2) CKL(kRngPhilox, 2); else CKL(kRngPhilox, 4);
        } else {
            if (apb == 1) CKL(kRngParkMiller, 1); else if (apb == 2) CKL(kRngParkMiller, 2); else CKL(kRngParkMiller, 4);
        }
#undef CKL
        return hipGetLastError();
    }
    const size_t lds = construct_lds_bytes(b.max_n, selection);
    const dim3 grid(b.ants), block(kLanes);
    if (selection == kSelIRoulette) {
        if (b.ants <= b.cus) {
            // at most one ant per CU: four waves per ant
            const size_t lds4 = iroulette4_exchange_offset(b.max_n) + 2 * kIrWavesMax * sizeof(u64);
            const int nw = iroulette_waves(b.max_n);
            const dim3 block4(nw * kLanes);
            if (nw == 8) {
                if (b.has_d) hipLaunchKernelGGL((k_construct_iroulette4<true, 8>), grid, block4, lds4, st, b);
                else hipLaunchKernelGGL((k_construct_iroulette4<false, 8>), grid, block4, lds4, st, b);
            } else {
                if (b.has_d) hipLaunchKernelGGL((k_construct_iroulette4<true, 4>), grid, block4, lds4, st, b);
                else hipLaunchKernelGGL((k_construct_iroulette4<false, 4>), grid, block4, lds4, st, b);
            }
        } else if (b.has_d) {
            hipLaunchKernelGGL(k_construct_iroulette<true>, grid, block, lds, st, b);
        } else {
            hipLaunchKernelGGL(k_construct_iroulette<false>, grid, block, lds, st, b);
        }
    } else {
        // the wide fallback pass while a step waits on its own memory
        // latency (at most one ant's wave per SIMD); with more ants the
        // device is busy and the narrow pass's fewer bytes win
        const bool spec = b.ants <= 4 * b.cus;
#define CK(D, R, S) hipLaunchKernelGGL((k_construct_knn<D, R, S>), grid, block, lds, st, b)
        if (rng == kRngPhilox) {
            if (b.has_d) { if (spec) CK(true, kRngPhilox, true); else CK(true, kRngPhilox, false); }
            else { if (spec) CK(false, kRngPhilox, true); else CK(false, kRngPhilox, false); }
        } else {
            if (b.has_d) { if (spec) CK(true, kRngParkMiller, true); else CK(true, kRngParkMiller, false); }
            else { if (spec) CK(false, kRngParkMiller, true); else CK(false, kRngParkMiller, false); }
        }
#undef CK
    }
    return hipGetLastError();
}

// --------------------------------------------------------- statistics ----

constexpr int kStatThreads = 1024;
constexpr int kStatUnroll = 4;   // tour positions per thread whose loads are issued together

// Block reductions over NT threads (NT / 64 waves); red holds NT/64 + 1.
template <int NT>
__device__ u64 block_min_u64(u64 v, u64 *red)
{
    constexpr int NW = NT / kLanes;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_min_u64(v);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (wv == 0) {
        u64 x = lane < NW ? red[lane] : ~0ull;
        x = wave_min_u64(x);
        if (lane == 0) red[NW] = x;
    }
    __syncthreads();
    const u64 r = red[NW];
    __syncthreads();
    return r;
}

template <int NT>
__device__ long long block_sum_i64(long long v, long long *red)
{
    constexpr int NW = NT / kLanes;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    long long r = 0;
    for (int w = 0; w < NW; ++w) r += red[w];
    __syncthreads();
    return r;
}

// update_statistics (g_aco.cpp:407-522): iteration best by the tie order of
// best_solution_phase_0/1 (kernel.cl:594-665), best-so-far copy and record
// (kernel.cl:546-586) stamped with the device clock; then get_elites
// (kernel.cl:717-741), the deposit amounts (kernel.cl:767-768) and the
// elite edge tables the pheromone pass reads; finally the iteration count.
// One instance by the NT threads of the calling block.
template <int NT>
__device__ void statistics_instance(const DevProblem &p, const DevState &s, float tick_to_s)
{
    constexpr int kStatThreads = NT, kStatWaves = NT / kLanes;
    __shared__ u64 red[kStatWaves + 1];
    __shared__ long long red64[kStatWaves];
    __shared__ int sh_elite[kMaxRas];
    const int tid = threadIdx.x, m = p.m, T = p.T;

    // 1. iteration best: key (length bits, group, rank-in-group).  The
    //    length of the thread's first ant is kept for the elite rounds.
    const float len_own = tid < m ? s.lens[tid] : 0.0f;
    u64 kmin = ~0ull;
    for (int a = tid; a < m; a += kStatThreads) {
        const float len = a == tid ? len_own : s.lens[a];
        const u64 key = ((u64)f2u(len) << 32) | (u64)((a & ~63) | p.rank64[a & 63]);
        kmin = key < kmin ? key : kmin;
    }
    kmin = block_min_u64<NT>(kmin, red);
    const u32 lo = (u32)(kmin & 0xFFFFFFFFull);
    const int best = (int)((lo & ~63u) | (u32)p.rank64inv[lo & 63]);
    const float best_len = s.lens[best];
    const float bsf_len = s.lens[m];
    const bool flag = best_len - bsf_len < -kEps;
    __syncthreads();

    // 2. best-so-far copy and record
    if (flag) {
        const int sz = s.sizes[best];
        const int *src = s.tours + (size_t)best * T;
        int *dst = s.tours + (size_t)m * T;
        // a trip's loads are issued before its stores (source and
        // destination are rows of one buffer, which the compiler cannot tell
        // apart, so it would otherwise wait for each load in turn)
        for (int q0 = tid; q0 < sz; q0 += kStatUnroll * kStatThreads) {
            int v[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                const int q = q0 + u * kStatThreads;
                v[u] = src[q < sz ? q : 0];
            }
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                const int q = q0 + u * kStatThreads;
                if (q < sz) dst[q] = v[u];
            }
        }
        if (tid == 0) {
            s.sizes[m] = sz;
            s.lens[m] = best_len;
            const int pos = s.ctrl[kCtrlNbsf];
            if (pos < s.bsf_cap) {
                BsfRecord r;
                r.iter = s.ctrl[kCtrlIter];
                r.length = best_len;
                r.time = (float)((double)(wall_clock64() - *s.clock0) * (double)tick_to_s);
                s.bsf[pos] = r;
                s.ctrl[kCtrlNbsf] = pos + 1;
            } else {
                s.ctrl[kCtrlOverflow] += 1;
            }
        }
    }
    if (tid == 0) {
        s.ctrl[kCtrlIterBest] = best;
        s.ctrl[kCtrlFlag] = flag ? 1 : 0;
    }

    // 3. elites: best-so-far slot, then ras-1 iteration ranks, ties to the
    //    lowest index (repeated strict '<' scans of get_elites)
    if (tid == 0) sh_elite[0] = m;
    __syncthreads();
    for (int g = 1; g < p.ras; ++g) {
        u64 km = ~0ull;
        for (int a = tid; a < m; a += kStatThreads) {
            bool taken = false;
            for (int q = 1; q < g; ++q) taken |= sh_elite[q] == a;
            const float v = taken ? __int_as_float(0x7f800000) : a == tid ? len_own : s.lens[a];
            const u64 key = ((u64)f2u(v) << 32) | (u64)a;
            km = key < km ? key : km;
        }
        km = block_min_u64<NT>(km, red);
        if (tid == 0) sh_elite[g] = (km >> 32) == 0x7f800000ull ? 0 : (int)(km & 0xFFFFFFFFull);
        __syncthreads();
    }

    // 4. deposit amounts and elite edge tables
    for (int g = tid; g < p.ras; g += kStatThreads) {
        s.elite[g] = sh_elite[g];
        s.delta[g] = (float)(p.ras - g) / s.lens[sh_elite[g]];
    }
    for (int q = tid; q < p.ras * p.P; q += kStatThreads) s.firstf[q] = 0;
    __syncthreads();
    // the elites' tour sizes, requested together
    int szs[kMaxRas];
#pragma unroll
    for (int g = 0; g < kMaxRas; ++g) szs[g] = g < p.ras ? s.sizes[sh_elite[g]] : 0;
    for (int g = 0; g < p.ras; ++g) {
        const int e = sh_elite[g];
        const int sz = szs[g];
        const int *t = s.tours + (size_t)e * T;
        // loads first, then the stores (each node is written once per elite)
        for (int q0 = tid; q0 < sz - 1; q0 += kStatUnroll * kStatThreads) {
            int tu[kStatUnroll], tv[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                const int q = q0 + u * kStatThreads;
                const int qc = q < sz - 1 ? q : 0;
                tu[u] = t[qc];
                tv[u] = t[qc + 1];
            }
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                const int q = q0 + u * kStatThreads;
                if (q < sz - 1) {
                    if (tu[u] != 0) s.succ[g * p.n + tu[u]] = tv[u];
                    else if (tv[u] != 0) s.firstf[(size_t)g * p.P + tv[u]] = 1;
                }
            }
        }
    }

    // 5. counters and iteration
    long long ev = 0, mv = 0, fb = 0;
    for (int a = tid; a < m; a += kStatThreads) {
        ev += s.ant_evals[a];
        mv += s.ant_moves[a];
        fb += s.ant_fb[a];
    }
    ev = block_sum_i64<NT>(ev, red64);
    mv = block_sum_i64<NT>(mv, red64);
    fb = block_sum_i64<NT>(fb, red64);
    if (tid == 0) {
        s.totals[kTotEvals] += ev;
        s.totals[kTotMoves] += mv;
        s.totals[kTotFallbacks] += fb;
        s.totals[kTotIters] += 1;
        s.ctrl[kCtrlIter] += 1;
    }
}

__global__ __launch_bounds__(kStatThreads) void k_statistics(DevBatch bt, float tick_to_s)
{
    // copies in registers: fields read through the global records would be
    // loaded again after every barrier, each load waiting for all
    // outstanding memory operations
    const DevProblem p = bt.probs[blockIdx.x];
    const DevState s = bt.states[blockIdx.x];
    statistics_instance<kStatThreads>(p, s, tick_to_s);
}

hipError_t launch_statistics(const DevBatch &b, float tick_to_s, hipStream_t st)
{
    hipLaunchKernelGGL(k_statistics, dim3(b.B), dim3(kStatThreads), 0, st, b, tick_to_s);
    return hipGetLastError();
}

// --------------------------------------------------------- pheromone ----

// Fused pheromone_evaporation + pheromone_deposit + compute_total_info
// (kernel.cl:685-692, 747-776, 782-789) in one pass over n x P.  Each
// wavefront owns 128 float4 of one row j (two per lane, both loaded before
// either is used), so the row's elite successors (succ[g][j], <= ras values)
// are wave-uniform.  Per element:
//   t = (1-rho)*tau;  for g in order: if edge (j,h) is in elite g: t += d_g
//   tau = t;  total_info = t^alpha * heur
// The elite-order sum is deterministic (the reference's CAS atomics are not).
constexpr int kPhF4 = 128;   // float4 per work item (row segment)

__device__ __forceinline__ int pheromone_items(const DevProblem &p) { return p.n * ((p.P / 4 + kPhF4 - 1) / kPhF4); }

__device__ __forceinline__ float4 pheromone_f4(const DevProblem &p, const DevState &s, int j, int c4, float evap,
                                               const int *sc, const float *dl, float4 t, float4 h, size_t off)
{
    u32 ff[kMaxRas];
#pragma unroll
    for (int g = 0; g < kMaxRas; ++g) {
        ff[g] = 0;
        if (g < p.ras && j == 0) ff[g] = *reinterpret_cast<const u32 *>(s.firstf + (size_t)g * p.P + c4 * 4);
    }
    float v[4] = {t.x, t.y, t.z, t.w};
    const float hv[4] = {h.x, h.y, h.z, h.w};
    float ti[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = c4 * 4 + c;
        float x = evap * v[c];
#pragma unroll
        for (int g = 0; g < kMaxRas; ++g) {
            if (g < p.ras) {
                const bool hit = j != 0 ? sc[g] == col : ((ff[g] >> (8 * c)) & 0xFFu) != 0;
                if (hit) x = x + dl[g];
            }
        }
        v[c] = x;
        ti[c] = (p.alpha == 1.0f ? x : powf(x, p.alpha)) * hv[c];
    }
    *reinterpret_cast<float4 *>(s.tau + off) = make_float4(v[0], v[1], v[2], v[3]);
    const float4 out = make_float4(ti[0], ti[1], ti[2], ti[3]);
    *reinterpret_cast<float4 *>(s.tinfo + off) = out;
    return out;
}


// Work item `wave` (row j, kPhF4 float4 of it) of one instance's pass.
__device__ __forceinline__ void pheromone_item(const DevProblem &p, const DevState &s, int wave, int lane)
{
    const float evap = 1.0f - p.rho;
    const int chunks = (p.P / 4 + kPhF4 - 1) / kPhF4;
    const int j = wave / chunks;
    if (j >= p.n) return;
    const int c4a = (wave - j * chunks) * kPhF4 + lane, c4b = c4a + 64;
    const bool a_in = c4a * 4 < p.P, b_in = c4b * 4 < p.P;
    int sc[kMaxRas];
    float dl[kMaxRas];
#pragma unroll
    for (int g = 0; g < kMaxRas; ++g) {
        sc[g] = -1;
        dl[g] = 0.0f;
        if (g < p.ras) {
            sc[g] = j != 0 ? s.succ[g * p.n + j] : -1;
            dl[g] = s.delta[g];
        }
    }
    const size_t offa = (size_t)j * p.P + (size_t)c4a * 4, offb = offa + 256;
    float4 ta = make_float4(0, 0, 0, 0), ha = ta, tb = ta, hb = ta;
    if (a_in) {
        ta = *reinterpret_cast<const float4 *>(s.tau + offa);
        ha = *reinterpret_cast<const float4 *>(p.heur + offa);
    }
    if (b_in) {
        tb = *reinterpret_cast<const float4 *>(s.tau + offb);
        hb = *reinterpret_cast<const float4 *>(p.heur + offb);
    }
    float4 ia = make_float4(0, 0, 0, 0), ib = ia;
    if (a_in) ia = pheromone_f4(p, s, j, c4a, evap, sc, dl, ta, ha, offa);
    if (b_in) ib = pheromone_f4(p, s, j, c4b, evap, sc, dl, tb, hb, offb);
    // the candidate row's total_info (tiknn[j][c] = total_info[j][nnk[j][c]])
    // for the candidates whose column lies in this segment: lane l holds
    // columns c0 + 4l.. and c0 + 256 + 4l.., so the value comes from lane
    // ((col - c0) mod 256) / 4 by a cross-lane read
    const int c0 = (c4a - lane) * 4;
    for (int cb = 0; cb < p.k; cb += kLanes) {
        const int c = cb + lane;
        const int col = c < p.k ? p.nnk[(size_t)j * p.k + c] : -1;
        const int rel = col - c0;
        const int src = (rel & 255) >> 2;
        // every lane takes part in all eight cross-lane reads (a read inside a
        // conditional would run with the other lanes inactive), then plain
        // selects pick the value (an indexed pick from a float4 would go
        // through scratch memory)
        const float a0 = __shfl(ia.x, src), a1 = __shfl(ia.y, src), a2 = __shfl(ia.z, src),
                    a3 = __shfl(ia.w, src);
        const float b0 = __shfl(ib.x, src), b1 = __shfl(ib.y, src), b2 = __shfl(ib.z, src),
                    b3 = __shfl(ib.w, src);
        const bool hi = rel >= 256;
        const int e = rel & 3;
        const float x0 = hi ? b0 : a0, x1 = hi ? b1 : a1, x2 = hi ? b2 : a2, x3 = hi ? b3 : a3;
        const float x = e == 0 ? x0 : e == 1 ? x1 : e == 2 ? x2 : x3;
        if (c < p.k && rel >= 0 && rel < 4 * kPhF4) s.tiknn[(size_t)j * p.k + c] = x;
    }
}

__global__ __launch_bounds__(256) void k_pheromone_update(DevBatch bt)
{
    pheromone_item(bt.probs[blockIdx.y], bt.states[blockIdx.y], blockIdx.x * 4 + (threadIdx.x >> 6),
                   threadIdx.x & 63);
}

hipError_t launch_pheromone_update(const DevBatch &b, hipStream_t st)
{
    const int chunks = (b.max_P / 4 + kPhF4 - 1) / kPhF4;
    const int waves = b.max_n * chunks;
    hipLaunchKernelGGL(k_pheromone_update, dim3((waves + 3) / 4, b.B), dim3(256), 0, st, b);
    return hipGetLastError();
}

// total_info of each candidate row, contiguous (n x k)
__global__ __launch_bounds__(256) void k_gather_knn(DevBatch bt)
{
    const DevProblem p = bt.probs[blockIdx.y];
    const DevState s = bt.states[blockIdx.y];
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= p.n * p.k) return;
    const int i = idx / p.k;
    s.tiknn[idx] = s.tinfo[(size_t)i * p.P + p.nnk[idx]];
}

hipError_t launch_gather(const DevBatch &b, hipStream_t st)
{
    if (b.max_nk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_knn, dim3((b.max_nk + 255) / 256, b.B), dim3(256), 0, st, b);
    return hipGetLastError();
}

// pheromone_init (kernel.cl:672-679).  mode kTau0Bsf takes tau0 =
// 1/(rho * L_bsf) on the device (g_aco.cpp:89) so initialisation needs no
// host round trip; kTau0Nn the instance's 1/(n * L_nn).
__global__ __launch_bounds__(256) void k_pheromone_init(DevBatch bt, int mode, float tau0)
{
    const DevProblem p = bt.probs[blockIdx.y];
    const DevState s = bt.states[blockIdx.y];
    const float t0 = mode == kTau0Bsf ? 1.0f / (p.rho * s.lens[p.m]) : mode == kTau0Nn ? p.tau0_nn : tau0;
    const float ta = p.alpha == 1.0f ? t0 : powf(t0, p.alpha);
    const size_t total = (size_t)p.n * p.P / 4;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < total; q += (size_t)gridDim.x * 256) {
        const float4 h = reinterpret_cast<const float4 *>(p.heur)[q];
        reinterpret_cast<float4 *>(s.tau)[q] = make_float4(t0, t0, t0, t0);
        reinterpret_cast<float4 *>(s.tinfo)[q] = make_float4(ta * h.x, ta * h.y, ta * h.z, ta * h.w);
    }
}

hipError_t launch_pheromone_init(const DevBatch &b, int mode, float tau0, hipStream_t st)
{
    const size_t total = (size_t)b.max_n * b.max_P / 4;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_pheromone_init, dim3(blocks, b.B), dim3(256), 0, st, b, mode, tau0);
    return hipGetLastError();
}

// --------------------------------------------------- fused iteration ----

// One launch per iteration (candidate roulette, no route-length limit):
// blocks of W waves take work items from a queue -- W consecutive ants of
// one instance, the instances in the local-search size-class order (largest
// first).  The W waves construct the item's ants, one each
// (construct_knn_ant), then search them one after another with the whole
// block (local_search_ant).  The block that finishes an instance's last
// item -- a device-scope counter per instance, with release/acquire fences
// around it -- runs that instance's statistics, pheromone pass and
// candidate gather, the same functions the per-phase kernels run.  No block
// waits for another; every block leaves when the queue is empty.
__global__ void k_fused_reset(DevBatch bt)
{
    if (threadIdx.x == 0) *bt.fq_next = 0;
}

// (a floor of 3 (W = 4) or 4 (W = 8) waves per SIMD: with construction,
// search and the last block's passes in one kernel the compiler would
// otherwise keep enough loop-invariant values live to allow only 2; 3 is
// what LDS allows 4-wave blocks at n = 500 anyway)
template <int W, typename RT, int RNG>
__global__ __launch_bounds__(W * 64, W == 4 ? 3 : 4) void k_iteration_fused(DevBatch bt, int flags, int ring, int prefetch,
                                                           float tick_to_s)
{
    constexpr int NT = W * 64;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int sh_item, sh_last;
    const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
    const int items = bt.ants / W;
    for (;;) {
        if (tid == 0) sh_item = atomicAdd(bt.fq_next, 1);
        __syncthreads();
        // uniform, so the instance's descriptors load into scalar registers
        const int it = __builtin_amdgcn_readfirstlane(sh_item);
        __syncthreads();
        if (it >= items) break;
        const int g0 = bt.ls_order[it * W];
        const int inst = bt.ant_inst[g0];
        const DevProblem p = bt.probs[inst];
        const DevState s = bt.states[inst];
        const int a0 = g0 - bt.ant_base[inst];
        {
            const int n = p.n;
            AntLds L;
            L.dem = reinterpret_cast<int *>(smem);
            L.d0 = reinterpret_cast<float *>(L.dem + n);
            L.rank = nullptr;
            L.vis = reinterpret_cast<unsigned char *>(L.d0 + n) + wv * ((n + 15) & ~15);
            for (int i = tid; i < n; i += NT) {
                L.dem[i] = p.dem[i];
                L.d0[i] = p.d0[i];
            }
            for (int i = lane; i < n; i += kLanes) L.vis[i] = i == 0;
            __syncthreads();
            construct_knn_ant<false, RNG>(p, s, a0 + wv, L, lane);
            __syncthreads();
        }
#pragma unroll 1
        for (int q = 0; q < W; ++q) {
            local_search_ant<W, RT>(p, s, a0 + q, flags, ring, prefetch, smem);
            __syncthreads();
        }
        __threadfence();   // this item's tours, lengths and counters, device-wide
        if (tid == 0) sh_last = atomicAdd(s.ctrl + kCtrlDone, W) + W == p.m;
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(sh_last)) {
            __threadfence();   // every other item of the instance
            statistics_instance<NT>(p, s, tick_to_s);
            __syncthreads();
            const int np = pheromone_items(p);
#pragma unroll 1
            for (int w = wv; w < np; w += W) pheromone_item(p, s, w, lane);   // candidate rows included
            if (tid == 0) s.ctrl[kCtrlDone] = 0;
        }
        __syncthreads();
    }
}

bool fused_iteration_ok(const DevBatch &b, int selection, int ls_flags, int ants_per_instance, const LsShape &sh)
{
    return selection == kSelKnn && !b.has_d && ls_flags != 0 && !(ls_flags & kLs2OptRandom) &&
           (sh.waves == 4 || sh.waves == 8) &&
           ants_per_instance % sh.waves == 0 && ls_shape_valid(sh);
}

size_t fused_lds_bytes(int n, const LsShape &sh)
{
    const size_t cons = (size_t)n * 8 + (size_t)sh.waves * ((n + 15) & ~15) + 16;
    const size_t ls = local_search_lds_bytes(n, sh.ring, sh.row16, 0);
    return cons > ls ? cons : ls;
}

hipError_t fused_static_lds(size_t &bytes)
{
    const void *fns[] = {(const void *)k_iteration_fused<4, float, kRngPhilox>,
                         (const void *)k_iteration_fused<4, unsigned short, kRngPhilox>,
                         (const void *)k_iteration_fused<8, float, kRngPhilox>,
                         (const void *)k_iteration_fused<8, unsigned short, kRngPhilox>,
                         (const void *)k_iteration_fused<4, float, kRngParkMiller>,
                         (const void *)k_iteration_fused<4, unsigned short, kRngParkMiller>,
                         (const void *)k_iteration_fused<8, float, kRngParkMiller>,
                         (const void *)k_iteration_fused<8, unsigned short, kRngParkMiller>};
    return max_static_lds(fns, (int)(sizeof(fns) / sizeof(fns[0])), bytes);
}

hipError_t launch_iteration_fused(const DevBatch &b, int rng, int ls_flags, const LsShape &sh, float tick_to_s,
                                  hipStream_t st)
{
    hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(64), 0, st, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const size_t lds = fused_lds_bytes(b.max_n, sh);
    // enough blocks to fill the device; each leaves when the queue is empty
    const int items = b.ants / sh.waves;
    int blocks = cus * (sh.waves == 4 ? 3 : 2);
    if (blocks > items) blocks = items;
#define KF(W, RT, R) hipLaunchKernelGGL((k_iteration_fused<W, RT, R>), dim3(blocks), dim3(W * kLanes), lds, st, b, \
                                        ls_flags, sh.ring, sh.prefetch, tick_to_s)
    if (rng == kRngPhilox) {
        if (sh.waves == 8) { if (sh.row16) KF(8, unsigned short, kRngPhilox); else KF(8, float, kRngPhilox); }
        else { if (sh.row16) KF(4, unsigned short, kRngPhilox); else KF(4, float, kRngPhilox); }
    } else {
        if (sh.waves == 8) { if (sh.row16) KF(8, unsigned short, kRngParkMiller); else KF(8, float, kRngParkMiller); }
        else { if (sh.row16) KF(4, unsigned short, kRngParkMiller); else KF(4, float, kRngParkMiller); }
    }
#undef KF
    return hipGetLastError();
}

__global__ void k_stamp_clock(DevBatch bt)
{
    if (threadIdx.x == 0) *bt.states[0].clock0 = wall_clock64();
}

hipError_t launch_stamp_clock(const DevBatch &b, hipStream_t st)
{
    hipLaunchKernelGGL(k_stamp_clock, dim3(1), dim3(64), 0, st, b);
    return hipGetLastError();
}

}  // namespace cvrp
