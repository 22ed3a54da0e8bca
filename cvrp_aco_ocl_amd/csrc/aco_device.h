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
constexpr float kEps = 0.001f;         // utilities.h:19
constexpr int kMaxNodes = 4096;        // 12-bit node ids in reduction keys
constexpr int kMaxRas = 16;
constexpr int kMaxMovesPerRoute = 1 << 14;
constexpr int kMaxMovesFactor = 64;    // swap/relocate cap = 64*n per call

enum { kSelIRoulette = 0, kSelKnn = 1 };
enum { kRngParkMiller = 0, kRngPhilox = 1 };
enum { kLs2Opt = 1, kLsSwap = 2, kLsRelocate = 4 };

// control words (DevState::ctrl)
enum { kCtrlIter = 0, kCtrlNbsf = 1, kCtrlIterBest = 2, kCtrlFlag = 3, kCtrlOverflow = 4,
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
struct DevBatch {
    const DevProblem *probs;  // B
    const DevState *states;   // B
    const int *ant_inst;      // ants: instance of global ant g
    const int *ant_base;      // B: first global ant of instance b
    int B, ants;
    int max_n, max_P, max_nk;
    int has_d;                // some instance has a route-length limit
};

enum { kTau0Arg = 0, kTau0Bsf = 1, kTau0Nn = 2 };

// Local-search launch shape (see aco_launch_shape).
struct LsShape {
    int waves;
    int ring;
    int prefetch;
};

// ---- host launchers (aco_kernels.hip, aco_local_search.hip) ----
hipError_t launch_construct(const DevBatch &b, int selection, int rng, hipStream_t st);
hipError_t launch_local_search(const DevBatch &b, int flags, const LsShape &sh, hipStream_t st);
hipError_t launch_statistics(const DevBatch &b, float tick_to_s, hipStream_t st);
hipError_t launch_pheromone_update(const DevBatch &b, hipStream_t st);
hipError_t launch_gather(const DevBatch &b, hipStream_t st);
hipError_t launch_pheromone_init(const DevBatch &b, int mode, float tau0, hipStream_t st);
hipError_t launch_stamp_clock(const DevBatch &b, hipStream_t st);
size_t construct_lds_bytes(int n, int selection);
size_t local_search_lds_bytes(int n, int ring);
bool ls_shape_valid(const LsShape &sh);

}  // namespace cvrp
