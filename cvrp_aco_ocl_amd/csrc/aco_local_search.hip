// Local search kernel (gfx950): one wavefront per ant, the ant's tour and
// per-route tables resident in LDS.
//
//   2-opt     the reference's GPU 2-opt (kernel.cl:126-319): per route,
//             sequential scan order, first improvement over the nn list
//             with radius pruning, don't-look bits.  Routes are independent
//             (the depot's don't-look bit starts clear on route 0 and set on
//             later routes), so lane r runs route r, r+64, ...
//   swap      LocalSearch::swap (localSearch.cpp:301-410): for each position
//             i in order, the first j > i whose swap gains more than
//             EPSILON and keeps loads and route lengths feasible.  The 64
//             lanes evaluate 64 consecutive j at once against the rows
//             d(pred i,.), d(i,.), d(succ i,.) staged in LDS, and a ballot
//             picks the lowest improving j -- the sequential order exactly.
//   relocate  the same scan for insertion moves (neighbourSearch.cpp:280-293,
//             move.cpp:32-51), rows d(i,.) in LDS.
#include "aco_device.h"

namespace cvrp {

using u64 = unsigned long long;

struct LsLds {
    int *tour, *tmp, *rl;
    float *dtw, *rd, *rowA, *rowB, *rowC;
    short *pos, *rof, *ridx, *zpos, *rc;
    unsigned char *dlb;
};

__host__ __device__ constexpr size_t ls_align16(size_t x) { return (x + 15) & ~size_t(15); }

// LDS byte offsets of the per-ant arrays (T = 2n tour slots).
struct LsOffsets {
    size_t tour, tmp, dtw, rl, rd, ra, rb, rc_row, pos, rof, ridx, zpos, rc, dlb, total;
};

__host__ __device__ inline LsOffsets ls_offsets(int n)
{
    const size_t T = 2 * (size_t)n, N = (size_t)n;
    LsOffsets o;
    size_t at = 0;
    o.tour = at;   at = ls_align16(at + 4 * T);
    o.tmp = at;    at = ls_align16(at + 4 * T);
    o.dtw = at;    at = ls_align16(at + 4 * T);
    o.rl = at;     at = ls_align16(at + 4 * N);
    o.rd = at;     at = ls_align16(at + 4 * N);
    o.ra = at;     at = ls_align16(at + 4 * N);
    o.rb = at;     at = ls_align16(at + 4 * N);
    o.rc_row = at; at = ls_align16(at + 4 * N);
    o.pos = at;    at = ls_align16(at + 2 * N);
    o.rof = at;    at = ls_align16(at + 2 * N);
    o.ridx = at;   at = ls_align16(at + 2 * T);
    o.zpos = at;   at = ls_align16(at + 2 * (N + 1));
    o.rc = at;     at = ls_align16(at + 2 * N);
    o.dlb = at;    at = ls_align16(at + N);
    o.total = at;
    return o;
}

__device__ __forceinline__ LsLds ls_bind(int n, unsigned char *b)
{
    const LsOffsets o = ls_offsets(n);
    LsLds L;
    L.tour = (int *)(b + o.tour); L.tmp = (int *)(b + o.tmp); L.dtw = (float *)(b + o.dtw);
    L.rl = (int *)(b + o.rl); L.rd = (float *)(b + o.rd);
    L.rowA = (float *)(b + o.ra); L.rowB = (float *)(b + o.rb); L.rowC = (float *)(b + o.rc_row);
    L.pos = (short *)(b + o.pos); L.rof = (short *)(b + o.rof); L.ridx = (short *)(b + o.ridx);
    L.zpos = (short *)(b + o.zpos); L.rc = (short *)(b + o.rc); L.dlb = b + o.dlb;
    return L;
}

size_t local_search_lds_bytes(int n) { return ls_offsets(n).total; }

__device__ __forceinline__ float Dd(const DevProblem &p, int a, int b)
{
    return p.dist[(size_t)a * p.P + b];
}

// dtw[q] = d(tour[q], tour[q+1]) for q in [lo, hi]
__device__ __forceinline__ void ls_edges(const DevProblem &p, LsLds &L, int lo, int hi, int lane)
{
    for (int q = lo + lane; q <= hi; q += kLanes) L.dtw[q] = Dd(p, L.tour[q], L.tour[q + 1]);
}

// Depot positions (zpos), customer positions and route ids (pos, rof), route
// index per position (ridx, as localSearch.cpp:319-337 counts routes), and
// the per-route load / distance+service / customer count.  Returns R.
__device__ __forceinline__ int ls_routes(const DevProblem &p, LsLds &L, int S, int lane, bool tables)
{
    int zc = 0;
    for (int base = 0; base < S; base += kLanes) {
        const int q = base + lane;
        const int v = q < S ? L.tour[q] : -1;
        const u64 mz = __ballot(v == 0);
        const int before = zc + __popcll(mz & ((1ull << lane) - 1ull));
        if (q < S) {
            if (v == 0) {
                L.zpos[before] = (short)q;
            } else {
                L.pos[v] = (short)q;
                L.rof[v] = (short)(before - 1);
            }
            L.ridx[q] = (short)(before > 0 ? before - 1 : 0);
        }
        zc += __popcll(mz);
    }
    __syncthreads();
    const int R = zc - 1;
    if (tables) {
        for (int r = lane; r < R; r += kLanes) {
            const int beg = L.zpos[r], end = L.zpos[r + 1];
            int load = 0;
            float dist = 0.0f;
            for (int i = beg + 1; i <= end; ++i) {
                load += p.dem[L.tour[i]];
                dist += L.dtw[i - 1];
            }
            L.rl[r] = load;
            L.rd[r] = dist + p.svc * (float)(end - beg - 1);
            L.rc[r] = (short)(end - beg - 1);
        }
        __syncthreads();
    }
    return R;
}

// ------------------------------------------------------------- 2-opt ----

__device__ __forceinline__ void two_opt_route(const DevProblem &p, LsLds &L, int r, int rbeg, int rend,
                              long long &evals, int &moves_total)
{
    bool dlb0 = r > 0;
    const int num = rend - rbeg + 1, nn = p.nn_ls;
    int moves = 0;
    bool imp = true;
    while (imp) {
        imp = false;
        for (int l = 0; l < num; ++l) {
            const int pos1 = rbeg + l;
            const int n1 = L.tour[pos1];
            if (n1 == 0 ? dlb0 : (L.dlb[n1] != 0)) continue;
            int h1 = 0, h2 = 0, h3 = 0, h4 = 0;
            bool found = false;
            const int *nl = p.nnls + (size_t)n1 * nn;
            const int s1 = pos1 == rend ? 0 : L.tour[pos1 + 1];
            float radius = Dd(p, n1, s1);
            // nn rows are sorted by distance, so once radius - d <= EPS no
            // later neighbour can be evaluated: stopping there is the same
            // as the reference's 'else break' on the next in-route node.
            for (int h = 0; h < nn; ++h) {
                const int n2 = nl[h];
                const float dn = Dd(p, n1, n2);
                if (!(radius - dn > kEps)) break;
                if (L.rof[n2] != r) continue;
                evals++;
                const int pos2 = L.pos[n2];
                const int s2 = pos2 == rend ? 0 : L.tour[pos2 + 1];
                const float gain = -radius + dn + Dd(p, s1, s2) - L.dtw[pos2];
                if (gain < -kEps) { h1 = n1; h2 = s1; h3 = n2; h4 = s2; found = true; break; }
            }
            if (!found) {
                const int p1 = pos1 == rbeg ? L.tour[rend] : L.tour[pos1 - 1];
                radius = Dd(p, p1, n1);
                for (int h = 0; h < nn; ++h) {
                    const int n2 = nl[h];
                    const float dn = Dd(p, n1, n2);
                    if (!(radius - dn > kEps)) break;
                    if (L.rof[n2] != r) continue;
                    const int pos2 = L.pos[n2];
                    const int p2 = pos2 == rbeg ? L.tour[rend] : L.tour[pos2 - 1];
                    if (p2 == n1 || p1 == n2) continue;
                    evals++;
                    const float gain = -radius + dn + Dd(p, p1, p2) - L.dtw[pos2 - 1];
                    if (gain < -kEps) { h1 = p1; h2 = n1; h3 = p2; h4 = n2; found = true; break; }
                }
            }
            if (!found) {
                if (n1 == 0) dlb0 = true;
                else L.dlb[n1] = 1;
                continue;
            }
            imp = true;
            const int hs[4] = {h1, h2, h3, h4};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (hs[q] == 0) dlb0 = false;
                else L.dlb[hs[q]] = 0;
            }
            int ph1 = h1 == 0 ? rbeg : L.pos[h1];
            int ph3 = h3 == 0 ? rbeg : L.pos[h3];
            if (ph3 < ph1) {
                int t = h1; h1 = h3; h3 = t;
                t = h2; h2 = h4; h4 = t;
                t = ph1; ph1 = ph3; ph3 = t;
            }
            int i = ph1, j = ph3;
            L.dtw[i] = Dd(p, h1, h3);
            L.dtw[j] = Dd(p, h2, h4);
            i++; j--;
            while (i < j) {
                const float t = L.dtw[i];
                L.dtw[i] = L.dtw[j];
                L.dtw[j] = t;
                i++; j--;
            }
            i = h2 == 0 ? rbeg : L.pos[h2];
            j = ph3;
            while (i < j) {
                const int a = L.tour[i], b = L.tour[j];
                L.tour[i] = b;
                L.tour[j] = a;
                L.pos[a] = (short)j;
                L.pos[b] = (short)i;
                i++; j--;
            }
            if (++moves >= kMaxMovesPerRoute) { moves_total += moves; return; }
        }
    }
    moves_total += moves;
}

// -------------------------------------------------------------- swap ----

__device__ __forceinline__ void ls_swap(const DevProblem &p, LsLds &L, int S, int lane, long long &ev, int &moves)
{
    const int n = p.n, cap = kMaxMovesFactor * n;
    int done = 0;
    for (int i = 1; i < S; ++i) {
        const int n1 = L.tour[i];
        if (n1 == 0) continue;
        const int pn1 = L.tour[i - 1], sn1 = L.tour[i + 1], r1 = L.ridx[i];
        for (int c = lane; c < n; c += kLanes) {
            L.rowA[c] = Dd(p, pn1, c);
            L.rowB[c] = Dd(p, n1, c);
            L.rowC[c] = Dd(p, sn1, c);
        }
        __syncthreads();
        const float a1 = L.rowA[n1], b1 = L.rowB[sn1];
        const int dn1 = p.dem[n1];
        int found = -1;
        for (int jb = i + 1; jb < S; jb += kLanes) {
            const int j = jb + lane;
            bool ok = false;
            if (j < S) {
                const int n2 = L.tour[j];
                if (n2 != 0) {
                    ev++;
                    const int pn2 = L.tour[j - 1], sn2 = L.tour[j + 1], r2 = L.ridx[j];
                    float gain;
                    if (j == i + 1)
                        gain = -(a1 + L.dtw[j]) + (L.rowA[n2] + L.rowB[sn2]);
                    else
                        gain = -(a1 + b1 + L.dtw[j - 1] + L.dtw[j]) +
                               (L.rowA[n2] + L.rowC[n2] + L.rowB[pn2] + L.rowB[sn2]);
                    if (gain < -kEps) {
                        if (r1 != r2) {
                            const int dn2 = p.dem[n2];
                            const int load1 = L.rl[r1] - dn1 + dn2, load2 = L.rl[r2] - dn2 + dn1;
                            const float dist1 = (L.rd[r1] - (a1 + b1)) + (L.rowA[n2] + L.rowC[n2]);
                            const float dist2 = (L.rd[r2] - (L.dtw[j - 1] + L.dtw[j])) +
                                                (L.rowB[pn2] + L.rowB[sn2]);
                            ok = !(load1 > p.Q || load2 > p.Q || dist1 > p.D || dist2 > p.D);
                        } else {
                            ok = true;
                        }
                    }
                }
            }
            const u64 mk = __ballot(ok);
            if (mk) { found = jb + __ffsll(mk) - 1; break; }
        }
        if (found >= 0) {
            if (lane == 0) {
                const int j = found, n2 = L.tour[j], pn2 = L.tour[j - 1], sn2 = L.tour[j + 1];
                const int r2 = L.ridx[j];
                int load1, load2;
                float dist1, dist2;
                if (r1 != r2) {
                    const int dn2 = p.dem[n2];
                    load1 = L.rl[r1] - dn1 + dn2;
                    load2 = L.rl[r2] - dn2 + dn1;
                    dist1 = (L.rd[r1] - (a1 + b1)) + (L.rowA[n2] + L.rowC[n2]);
                    dist2 = (L.rd[r2] - (L.dtw[j - 1] + L.dtw[j])) + (L.rowB[pn2] + L.rowB[sn2]);
                } else {
                    float gain;
                    if (j == i + 1)
                        gain = -(a1 + L.dtw[j]) + (L.rowA[n2] + L.rowB[sn2]);
                    else
                        gain = -(a1 + b1 + L.dtw[j - 1] + L.dtw[j]) +
                               (L.rowA[n2] + L.rowC[n2] + L.rowB[pn2] + L.rowB[sn2]);
                    load1 = L.rl[r1];
                    load2 = load1;
                    dist1 = L.rd[r1] + gain;
                    dist2 = dist1;
                }
                L.tour[i] = n2;
                L.tour[j] = n1;
                L.rl[r1] = load1; L.rl[r2] = load2;
                L.rd[r1] = dist1; L.rd[r2] = dist2;
                L.dtw[i - 1] = Dd(p, L.tour[i - 1], L.tour[i]);
                L.dtw[i] = Dd(p, L.tour[i], L.tour[i + 1]);
                L.dtw[j - 1] = Dd(p, L.tour[j - 1], L.tour[j]);
                L.dtw[j] = Dd(p, L.tour[j], L.tour[j + 1]);
            }
            i--;
            ++done;
        }
        __syncthreads();
        if (done >= cap) break;
    }
    if (lane == 0) moves += done;
}

// ---------------------------------------------------------- relocate ----

__device__ __forceinline__ void ls_relocate(const DevProblem &p, LsLds &L, int S, int lane, long long &ev,
                            int &moves)
{
    const int n = p.n, cap = kMaxMovesFactor * n;
    int done = 0, i = 1;
    while (i < S - 1 && done < cap) {
        const int n1 = L.tour[i];
        if (n1 == 0 || L.rc[L.ridx[i]] < 2) { i++; continue; }
        const int r1 = L.ridx[i], pn1 = L.tour[i - 1], sn1 = L.tour[i + 1];
        const float a1 = L.dtw[i - 1], b1 = L.dtw[i], c1 = Dd(p, pn1, sn1);
        for (int c = lane; c < n; c += kLanes) L.rowB[c] = Dd(p, n1, c);
        __syncthreads();
        const int dn1 = p.dem[n1];
        int found = -1;
        for (int jb = 1; jb < S - 1; jb += kLanes) {
            const int j = jb + lane;
            bool ok = false;
            if (j < S - 1 && j != i) {
                const int n2 = L.tour[j];
                if (n2 != 0) {
                    ev++;
                    const int pn2 = L.tour[j - 1], sn2 = L.tour[j + 1], r2 = L.ridx[j];
                    float rm, add;
                    if (i > j) { rm = L.dtw[j - 1]; add = L.rowB[n2] + L.rowB[pn2]; }
                    else { rm = L.dtw[j]; add = L.rowB[n2] + L.rowB[sn2]; }
                    const float gain = -(a1 + b1 + rm) + (add + c1);
                    if (gain < -kEps) {
                        if (r1 != r2) {
                            const float dist2 = ((L.rd[r2] + p.svc) - rm) + add;
                            const float dist1 = ((L.rd[r1] - p.svc) - (a1 + b1)) + c1;
                            ok = !(L.rl[r2] + dn1 > p.Q) && !(dist2 > p.D) && !(dist1 > p.D);
                        } else {
                            ok = true;
                        }
                    }
                }
            }
            const u64 mk = __ballot(ok);
            if (mk) { found = jb + __ffsll(mk) - 1; break; }
        }
        if (found < 0) {
            i++;
            __syncthreads();
            continue;
        }
        const int j = found;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        if (lane == 0) {
            const int n2 = L.tour[j], pn2 = L.tour[j - 1], sn2 = L.tour[j + 1], r2 = L.ridx[j];
            float rm, add;
            if (i > j) { rm = L.dtw[j - 1]; add = L.rowB[n2] + L.rowB[pn2]; }
            else { rm = L.dtw[j]; add = L.rowB[n2] + L.rowB[sn2]; }
            const float gain = -(a1 + b1 + rm) + (add + c1);
            if (r1 != r2) {
                L.rl[r1] -= dn1;
                L.rl[r2] += dn1;
                L.rd[r1] = ((L.rd[r1] - p.svc) - (a1 + b1)) + c1;
                L.rd[r2] = ((L.rd[r2] + p.svc) - rm) + add;
                L.rc[r1] = (short)(L.rc[r1] - 1);
                L.rc[r2] = (short)(L.rc[r2] + 1);
            } else {
                L.rd[r1] = L.rd[r1] + gain;
            }
        }
        // shift: copy [lo, hi] aside, then write it back moved by one
        for (int q = lo + lane; q <= hi; q += kLanes) L.tmp[q] = L.tour[q];
        __syncthreads();
        for (int q = lo + lane; q <= hi; q += kLanes) {
            int v;
            if (i < j) v = q == j ? n1 : L.tmp[q + 1];
            else v = q == j ? n1 : L.tmp[q - 1];
            L.tour[q] = v;
        }
        __syncthreads();
        ls_edges(p, L, lo - 1, hi, lane);
        ls_routes(p, L, S, lane, false);  // refreshes ridx (and pos/zpos)
        ++done;
    }
    if (lane == 0) moves += done;
}

// ------------------------------------------------------------ kernel ----

__global__ __launch_bounds__(64) void k_local_search(DevProblem p, DevState s, int flags)
{
    extern __shared__ __align__(16) unsigned char smem[];
    LsLds L = ls_bind(p.n, smem);
    const int a = blockIdx.x, lane = threadIdx.x;
    const int S = s.sizes[a];
    int *gt = s.tours + (size_t)a * p.T;
    for (int q = lane; q < S; q += kLanes) L.tour[q] = gt[q];
    for (int v = lane; v < p.n; v += kLanes) L.dlb[v] = 0;
    __syncthreads();
    long long ev = 0;
    int moves = 0;
    if (flags & kLs2Opt) {
        ls_edges(p, L, 0, S - 2, lane);
        const int R = ls_routes(p, L, S, lane, false);
        for (int r = lane; r < R; r += kLanes)
            two_opt_route(p, L, r, L.zpos[r], L.zpos[r + 1] - 1, ev, moves);
        __syncthreads();
    }
    if (flags & kLsSwap) {
        ls_edges(p, L, 0, S - 2, lane);
        __syncthreads();
        ls_routes(p, L, S, lane, true);
        ls_swap(p, L, S, lane, ev, moves);
    }
    if (flags & kLsRelocate) {
        ls_edges(p, L, 0, S - 2, lane);
        __syncthreads();
        ls_routes(p, L, S, lane, true);
        ls_relocate(p, L, S, lane, ev, moves);
    }
    ls_edges(p, L, 0, S - 2, lane);
    __syncthreads();
    for (int q = lane; q < S; q += kLanes) gt[q] = L.tour[q];
    // tour length in position order (vrpHelper.cpp:244-253)
    if (lane == 0) {
        float len = 0.0f;
        for (int q = 0; q < S - 1; ++q) len += L.dtw[q];
        s.lens[a] = len;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ev += __shfl_xor(ev, o);
        moves += __shfl_xor(moves, o);
    }
    if (lane == 0) {
        s.ant_evals[a] = ev;
        s.ant_moves[a] = moves;
    }
}

hipError_t launch_local_search(const DevProblem &p, const DevState &s, int flags, hipStream_t st)
{
    if (!flags) return hipSuccess;
    hipLaunchKernelGGL(k_local_search, dim3(p.m), dim3(kLanes), local_search_lds_bytes(p.n), st,
                       p, s, flags);
    return hipGetLastError();
}

}  // namespace cvrp
