// Device-side data layout shared by the HIP kernels and the C-ABI layer.
//
// One instance lives in HBM as row-major n x P float32 matrices (P = n
// rounded up to a multiple of 4 so rows are 16-byte aligned for float4
// passes): dist, heur (= eta^beta, times an optional prior), tau and
// total_info.  Candidate rows (n x k ids and their total_info) are stored
// contiguously so one construction step reads k*8 bytes.  Tours are
// int32[(m+1) x T], T = 2n, slot m holds the best-so-far tour; sizes and
// lengths are separate arrays (the reference packs the size into the last
// tour slot, g_aco.cpp:39,640-653).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvrp {

constexpr int kLanes = 64;             // wavefront = one ant

// Load through a global-address-space pointer: the device arrays reach the
// kernels through descriptor structs, so the compiler sees generic pointers
// and would emit flat loads, which also occupy the LDS counter (an LDS wait
// then waits for them too) and cost more to issue.
template <typename T>
__device__ __forceinline__ T gload(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)(p);
}

// Index of the calling wave in its block, as a wave-uniform value (the
// compiler cannot prove threadIdx.x >> 6 uniform by itself; without this,
// loops and branches on it are compiled as divergent).
__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// ---- lane exchange on the DPP network (no LDS traffic) ----
// DPP controls: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1]
// (lane ^ 2), row_shl / row_shr by 4 and 8 (lane +- 4, +- 8 within a row of
// 16 lanes).  Out-of-row reads give 0.
enum { kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppShl4 = 0x104, kDppShr4 = 0x114, kDppShl8 = 0x108,
       kDppShr8 = 0x118 };

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v)
{
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)v, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(v >> 32), CTRL, 0xF, 0xF, true);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// value of lane (lane ^ 4) / (lane ^ 8), within the row
__device__ __forceinline__ unsigned long long dpp_xor4_u64(unsigned long long v, int lane)
{
    const unsigned long long a = dpp_u64<kDppShl4>(v), b = dpp_u64<kDppShr4>(v);
    return (lane & 4) ? b : a;
}

__device__ __forceinline__ unsigned long long dpp_xor8_u64(unsigned long long v, int lane)
{
    const unsigned long long a = dpp_u64<kDppShl8>(v), b = dpp_u64<kDppShr8>(v);
    return (lane & 8) ? b : a;
}

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l gets
// lane l-1's (l+1's) value; lane 0 (lane 63), which has no source, gets fill.
__device__ __forceinline__ int wave_prev(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int wave_next(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ float wave_prev(float v, float fill)
{
    return __int_as_float(wave_prev(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ float wave_next(float v, float fill)
{
    return __int_as_float(wave_next(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ float readlane_f32(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}

// Max / min over all 64 lanes (all lanes active), uniform result.
template <bool MAX>
__device__ __forceinline__ unsigned long long wave_reduce_u64(unsigned long long v)
{
    const int lane = (int)(threadIdx.x & 63);
    auto pick = [](unsigned long long a, unsigned long long b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    v = pick(v, dpp_u64<kDppXor1>(v));
    v = pick(v, dpp_u64<kDppXor2>(v));
    v = pick(v, dpp_xor4_u64(v, lane));
    v = pick(v, dpp_xor8_u64(v, lane));
    const unsigned long long r0 = readlane_u64(v, 0), r1 = readlane_u64(v, 16);
    const unsigned long long r2 = readlane_u64(v, 32), r3 = readlane_u64(v, 48);
    return pick(pick(r0, r1), pick(r2, r3));
}
constexpr float kEps = 0.001f;         // utilities.h:19
constexpr int kMaxNodes = 4096;        // 12-bit node ids in reduction keys
constexpr int kMaxRas = 16;
constexpr int kMaxMovesPerRoute = 1 << 14;
constexpr int kMaxMovesFactor = 64;    // swap/relocate cap = 64*n per call

enum { kSelIRoulette = 0, kSelKnn = 1 };
enum { kRngParkMiller = 0, kRngPhilox = 1 };
enum { kLs2Opt = 1, kLsSwap = 2, kLsRelocate = 4, kLs2OptRandom = 8 };

// control words (DevState::ctrl)
enum { kCtrlIter = 0, kCtrlNbsf = 1, kCtrlIterBest = 2, kCtrlFlag = 3, kCtrlOverflow = 4,
       kCtrlDone = 5,   // fused iteration: ants of the instance finished this launch
       kCtrlLsSeed = 6, // Park-Miller state of the random-order 2-opt's scan-order stream
       kCtrlWords = 8 };
// running totals (DevState::totals)
enum { kTotEvals = 0, kTotMoves = 1, kTotFallbacks = 2, kTotIters = 3, kTotWords = 4 };

struct BsfRecord {
    int32_t iter;
    float length;
    float time;
};

struct DevProblem {
    int n, P, k, nn_ls, nnP, m, T, ras, Q;
    float D, svc, alpha, rho;
    float tau0_nn;            // 1/(n * L_nn) for the nearest-neighbour tau0 rule
    int has_d;                // max_dist is finite
    int dem16;                // every demand is in [0, 65535]
    unsigned long long seed;  // Philox key of this instance
    const float *dist;        // n*P
    const float *d0;          // n, distance to the depot (column 0)
    const int *dem;           // n
    const float *heur;        // n*P
    const int *nnk;           // n*k candidate ids
    const float *nnkd;        // n*k: d(i, nnk[i][c]), read with the ids
    const int *nnls;          // n*nnP: 2-opt neighbour ids, rows padded to nnP (multiple of 4)
    const float *nnd;         // n*nnP: d(i, nnls[i][h]), so a row scan reads (id, distance) pairs
    const unsigned short *nnpos;  // n*n: position of j in i's nn row (0xFFFF if absent)
    const unsigned short *dist16; // n*P: distances as 16-bit integers (valid when they all are)
    const int *rank;          // n, priority order of the I-Roulette reduction
    const int *rank64;        // 64, same for best_solution_phase_0 groups
    const int *rank64inv;     // 64
};

struct DevState {
    float *tau, *tinfo, *tiknn;
    int *tours, *sizes;
    float *lens;
    int *pm;                  // m*64 Park-Miller lane states
    int *ctrl;                // kCtrlWords
    BsfRecord *bsf;
    int bsf_cap;
    int *elite;               // ras
    int *succ;                // ras*n: successor of customer j in elite g
    unsigned char *firstf;    // ras*P: h directly follows the depot in elite g
    float *delta;             // ras
    long long *ant_evals;     // m
    int *ant_moves, *ant_fb;  // m
    long long *totals;        // kTotWords
    unsigned long long *clock0;
};

// A launch covers every instance of a context: per-instance problem and
// state live in device tables, construction and local-search blocks map
// to (instance, ant) through ant_inst / ant_base, the per-instance kernels
// use blockIdx.y as the instance.
// Local-search size classes of a batch: the ants of instances with similar
// n are launched together with LDS sized for the class, so ants of small
// instances pack densely on the CUs.
constexpr int kLsClasses = 6;

struct DevBatch {
    const DevProblem *probs;  // B
    const DevState *states;   // B
    const int *ant_inst;      // ants: instance of global ant g
    const int *ant_base;      // B: first global ant of instance b
    int B, ants;
    int max_n, max_P, max_nk;
    int has_d;                // some instance has a route-length limit
    int dem16;                // every instance's demands fit 16 bits
    const int *ls_order;              // global ant ids grouped by size class, largest class first
    int ls_classes;                   // non-empty classes
    int ls_first[kLsClasses + 1];     // class c: ls_order[ls_first[c] .. ls_first[c+1])
    int ls_maxn[kLsClasses];          // largest n in class c
    int *fq_next;                     // fused iteration: next work item (reset before each launch)
    size_t cons_lds_budget;           // dynamic LDS the staged construction may use (host side)
    int cus;                          // compute units of the device (host side)
};

enum { kTau0Arg = 0, kTau0Bsf = 1, kTau0Nn = 2 };

// Local-search launch shape (see aco_launch_shape).
struct LsShape {
    int waves;
    int ring;
    int prefetch;
    int construct_apb;   // construction: ants per block with the candidate table in LDS, 0 = off
    int row16;           // local-search rows as exact 16-bit integers (integer distances <= 65535)
    int fused = 0;       // one fused launch per iteration (k_iteration_fused)
};

// ---- host launchers (aco_kernels.hip, aco_local_search.hip) ----
// apb = ants per block of the LDS-staged candidate roulette (0: rows from
// L2/HBM, one ant per block); k = candidate list width
hipError_t launch_construct(const DevBatch &b, int selection, int rng, int apb, int k, hipStream_t st);
// Staged construction for instances of up to n nodes within `budget` bytes
// of dynamic LDS: candidate rows staged (n when the whole table fits, -1
// when not even the demand and visited arrays fit) and the LDS launched.
int construct_staged_rows(int n, int k, int apb, size_t budget);
size_t construct_lds_staged_bytes(int n, int k, int apb, size_t budget);
bool construct_staged_ok(const DevBatch &b, int selection, int k, int ants_per_instance, int apb);
// Extra streams on which the size classes of a batch run beside the first
// (forked from and joined back into the launch stream; under capture they
// become parallel graph branches).  count = 0: every class on the launch
// stream.
struct LsStreams {
    hipStream_t aux[kLsClasses - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kLsClasses - 1] = {};
    int count = 0;
};
hipError_t launch_local_search(const DevBatch &b, int flags, const LsShape &sh, hipStream_t st,
                               const LsStreams *par = nullptr);
// Park-Miller state after t draws from z (16807^t z mod 2^31-1).
__host__ __device__ inline int pm_jump(int z, unsigned long long t)
{
    const unsigned long long M = 2147483647ull;
    unsigned long long a = 16807ull, r = 1ull;
    while (t) {
        if (t & 1ull) r = r * a % M;
        a = a * a % M;
        t >>= 1;
    }
    return (int)(r * (unsigned long long)(unsigned)z % M);
}
// compute_nn_lists for `count` instances on the device (host arrays in and
// out, one thread per row, the reference's sort steps); ACO_OK or an error.
int gpu_nn_lists(hipStream_t st, int count, const int *n, const float *const *dist, const int *width,
                 int tie, int *const *out);
// the same from device distances (rows of pitch pitch[b]) into device arrays
int gpu_nn_lists_dev(hipStream_t st, int count, const int *n, const float *const *dist, const int *pitch,
                     const int *width, int tie, int *const *out);
// instance set-up on the device (aco_setup.hip): eta^beta (beta_code 1 or 2;
// 0 = uploaded), the prior already in heur multiplied in, d0, 16-bit rows;
// candidate and neighbour tables from the lists nnl (n x width)
hipError_t launch_setup_rows(const DevProblem &p, int beta_code, int has_prior, int int16, hipStream_t st);
hipError_t launch_setup_nn(const DevProblem &p, const int *nnl, int width, hipStream_t st);
hipError_t launch_candidate_dists(const DevProblem &p, hipStream_t st);
// sub-colony pheromone into a master instance (rows of all jobs in one launch)
struct SubTauJob {
    const float *tau;   // sub tau, n rows of pitch P
    const int *real;    // n master node ids, real[0] = depot
    int n, P, first_row;
};
hipError_t launch_tau_from_subs(const SubTauJob *jobs, int njobs, int rows, const DevProblem &master,
                                const DevState &ms, hipStream_t st);
// lens[ant] of an instance = its tour's edges added in step order
hipError_t launch_ant_length(const DevProblem &p, const DevState &s, int ant, hipStream_t st);
hipError_t launch_statistics(const DevBatch &b, float tick_to_s, hipStream_t st);
hipError_t launch_pheromone_update(const DevBatch &b, hipStream_t st);
hipError_t launch_gather(const DevBatch &b, hipStream_t st);
hipError_t launch_pheromone_init(const DevBatch &b, int mode, float tau0, hipStream_t st);
hipError_t launch_stamp_clock(const DevBatch &b, hipStream_t st);
size_t construct_lds_bytes(int n, int selection);
size_t local_search_lds_bytes(int n, int ring, int row16, int flags);
// Fused iteration (aco_kernels.hip): candidate roulette without a route
// length limit, W = 4 or 8 waves dividing the ants per instance.
bool fused_iteration_ok(const DevBatch &b, int selection, int ls_flags, int ants_per_instance, const LsShape &sh);
size_t fused_lds_bytes(int n, const LsShape &sh);
hipError_t launch_iteration_fused(const DevBatch &b, int rng, int ls_flags, const LsShape &sh, float tick_to_s,
                                  hipStream_t st);
bool ls_shape_valid(const LsShape &sh);
// Largest static LDS (__shared__ arrays of the kernel itself) over every
// instantiation of the construction, local-search and fused kernels; the
// dynamic LDS of a launch must fit beside it.
hipError_t construct_static_lds(size_t &bytes);
hipError_t local_search_static_lds(size_t &bytes);
hipError_t fused_static_lds(size_t &bytes);

}  // namespace cvrp
