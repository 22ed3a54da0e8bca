// Local search kernel (gfx950): one 256-thread block (4 wavefronts) per ant,
// the ant's tour, per-route tables and a cache of distance rows in LDS.
//
//   2-opt     the reference's GPU 2-opt (kernel.cl:126-319): per route,
//             sequential scan order, first improvement over the nn list
//             with radius pruning, don't-look bits.  Routes are independent
//             (the depot's don't-look bit starts clear on route 0 and set on
//             later routes), so thread r runs route r, r+256, ...  The nn
//             row is read as aligned (id, distance) batches of four and the
//             radius comes from the LDS edge array, so a candidate costs one
//             global load only when it is evaluated.
//   swap      LocalSearch::swap (localSearch.cpp:301-410): for each position
//             i in order, the first j > i whose swap gains more than
//             EPSILON and keeps loads and route lengths feasible.
//   relocate  the same scan for insertion moves (neighbourSearch.cpp:280-293,
//             move.cpp:32-51).
//
// swap and relocate run speculatively on four consecutive positions: wave w
// scans position i0+w against the current tour (64 j per ballot).  The
// smallest position with a move is exactly the sequential scan's next move,
// because no earlier position moved; results of later positions are
// discarded.  Distance rows d(v,.) of the nodes around i0 are cached in LDS
// by node id and reused across rounds.
#include "aco_device.h"

namespace cvrp {

using u64 = unsigned long long;

constexpr int kLsWaves = 4;
constexpr int kLsThreads = kLsWaves * kLanes;
constexpr int kRowSlots = 16;   // distance-row ring, slot = tour position mod ring size
constexpr int kPrefetch = 4;    // row elements per thread held for the next round
constexpr int kPrefetchMaxN = kLsThreads * kPrefetch;
// LDS words of L.res: [0,4) and [4,8) per-wave scan results of even and odd
// rounds, 8 route count.
constexpr int kResR = 8;

// Ring size: 16 rows when the next round's rows are prefetched (n <= 1024,
// 6 current + 4 next positions), else 8 (6 current positions).
__host__ __device__ constexpr int ls_ring_slots(int n) { return n <= kPrefetchMaxN ? 16 : 8; }

struct LsLds {
    int *tour, *tmp, *rl, *dem, *res;
    float *dtw, *rd, *rows;
    short *pos, *rof, *ridx, *zpos, *rc;
    unsigned char *dlb;
};

__host__ __device__ constexpr size_t ls_align16(size_t x) { return (x + 15) & ~size_t(15); }

struct LsOffsets {
    size_t tour, tmp, dtw, rl, rd, dem, rows, res, pos, rof, ridx, zpos, rc, dlb, total;
};

__host__ __device__ inline LsOffsets ls_offsets(int n)
{
    const size_t T = 2 * (size_t)n, N = (size_t)n;
    LsOffsets o;
    size_t at = 0;
    o.tour = at; at = ls_align16(at + 4 * T);
    o.tmp = at;  at = ls_align16(at + 4 * T);
    o.dtw = at;  at = ls_align16(at + 4 * T);
    o.rl = at;   at = ls_align16(at + 4 * N);
    o.rd = at;   at = ls_align16(at + 4 * N);
    o.dem = at;  at = ls_align16(at + 4 * N);
    o.rows = at; at = ls_align16(at + 4 * N * ls_ring_slots(n));
    o.res = at;  at = ls_align16(at + 4 * 32);
    o.pos = at;  at = ls_align16(at + 2 * N);
    o.rof = at;  at = ls_align16(at + 2 * N);
    o.ridx = at; at = ls_align16(at + 2 * T);
    o.zpos = at; at = ls_align16(at + 2 * (N + 1));
    o.rc = at;   at = ls_align16(at + 2 * N);
    o.dlb = at;  at = ls_align16(at + N);
    o.total = at;
    return o;
}

__device__ __forceinline__ LsLds ls_bind(int n, unsigned char *b)
{
    const LsOffsets o = ls_offsets(n);
    LsLds L;
    L.tour = (int *)(b + o.tour); L.tmp = (int *)(b + o.tmp); L.dtw = (float *)(b + o.dtw);
    L.rl = (int *)(b + o.rl); L.rd = (float *)(b + o.rd); L.dem = (int *)(b + o.dem);
    L.rows = (float *)(b + o.rows); L.res = (int *)(b + o.res);
    L.pos = (short *)(b + o.pos); L.rof = (short *)(b + o.rof); L.ridx = (short *)(b + o.ridx);
    L.zpos = (short *)(b + o.zpos); L.rc = (short *)(b + o.rc); L.dlb = b + o.dlb;
    return L;
}

// Bytes for any instance with at most n nodes: the ring shrinks above
// kPrefetchMaxN, so the largest layout may belong to a smaller instance.
size_t local_search_lds_bytes(int n)
{
    const size_t big = ls_offsets(n).total;
    const size_t mid = ls_offsets(n < kPrefetchMaxN ? n : kPrefetchMaxN).total;
    return big > mid ? big : mid;
}

__device__ __forceinline__ float Dd(const DevProblem &p, int a, int b)
{
    return p.dist[(size_t)a * p.P + b];
}

// dtw[q] = d(tour[q], tour[q+1]) for q in [lo, hi]
__device__ __forceinline__ void ls_edges(const DevProblem &p, LsLds &L, int lo, int hi, int tid)
{
    for (int q = lo + tid; q <= hi; q += kLsThreads) L.dtw[q] = Dd(p, L.tour[q], L.tour[q + 1]);
}

// Wave 0: depot positions (zpos), customer positions and route ids (pos,
// rof) and route index per position (ridx, counted as localSearch.cpp:319-337
// does).  Then, with tables, the per-route load / distance+service /
// customer count, one thread per route.  Returns R.
__device__ __forceinline__ int ls_routes(const DevProblem &p, LsLds &L, int S, int tid, bool tables)
{
    if (tid < kLanes) {
        const int lane = tid;
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
        if (lane == 0) L.res[kResR] = zc - 1;
    }
    __syncthreads();
    const int R = L.res[kResR];
    if (tables) {
        for (int r = tid; r < R; r += kLsThreads) {
            const int beg = L.zpos[r], end = L.zpos[r + 1];
            int load = 0;
            float dist = 0.0f;
            for (int i = beg + 1; i <= end; ++i) {
                load += L.dem[L.tour[i]];
                dist += L.dtw[i - 1];
            }
            L.rl[r] = load;
            L.rd[r] = dist + p.svc * (float)(end - beg - 1);
            L.rc[r] = (short)(end - beg - 1);
        }
    }
    __syncthreads();
    return R;
}

// ------------------------------------------------------------- 2-opt ----

// One scan for n1 (SUCC: partner edges (n1,s1),(n2,s2); otherwise
// (p1,n1),(p2,n2); nb1 is s1 or p1).  The reference walks n1's nn row,
// skips nodes of other routes and stops at the first in-route node with
// radius - d <= EPS (kernel.cl:180-230).  Rows are sorted by distance, so
// the nodes it evaluates are exactly the in-route nodes with
// radius - d > EPS that appear in the row, and the move it applies is the
// improving one earliest in the row.  Here the route's own positions are
// scanned (independent loads, no walk over other routes' neighbours) and
// the improving candidate with the smallest row position nnpos[n1][n2] is
// kept.  Returns true with h1..h4 set on an improving move.
template <bool SUCC>
__device__ __forceinline__ bool two_opt_scan(const DevProblem &p, const LsLds &L, int rbeg,
                                             int rend, int pos1, int n1, int nb1, float radius,
                                             long long &evals, int &h1, int &h2, int &h3, int &h4)
{
    const unsigned short *rowpos = p.nnpos + (size_t)n1 * p.n;
    int best = 0x7fffffff, b2 = 0, bx = 0;
#pragma unroll 4
    for (int q = rbeg + 1; q <= rend; ++q) {
        if (q == pos1) continue;
        const int n2 = L.tour[q];
        const int rp = rowpos[n2];
        const float dn = Dd(p, n1, n2);
        if (rp >= p.nn_ls || !(radius - dn > kEps)) continue;
        if (SUCC) {
            const int s2 = q == rend ? 0 : L.tour[q + 1];
            evals++;
            const float gain = -radius + dn + Dd(p, nb1, s2) - L.dtw[q];
            if (gain < -kEps && rp < best) { best = rp; b2 = n2; bx = s2; }
        } else {
            const int p2 = L.tour[q - 1];
            if (p2 == n1 || nb1 == n2) continue;
            evals++;
            const float gain = -radius + dn + Dd(p, nb1, p2) - L.dtw[q - 1];
            if (gain < -kEps && rp < best) { best = rp; b2 = n2; bx = p2; }
        }
    }
    if (best == 0x7fffffff) return false;
    if (SUCC) { h1 = n1; h2 = nb1; h3 = b2; h4 = bx; }
    else { h1 = nb1; h2 = n1; h3 = bx; h4 = b2; }
    return true;
}

__device__ __forceinline__ void two_opt_route(const DevProblem &p, LsLds &L, int r, int rbeg,
                                              int rend, long long &evals, int &moves_total)
{
    bool dlb0 = r > 0;
    const int num = rend - rbeg + 1;
    int moves = 0;
    bool imp = true;
    while (imp) {
        imp = false;
        for (int l = 0; l < num; ++l) {
            const int pos1 = rbeg + l;
            const int n1 = L.tour[pos1];
            if (n1 == 0 ? dlb0 : (L.dlb[n1] != 0)) continue;
            int h1 = 0, h2 = 0, h3 = 0, h4 = 0;
            const int s1 = pos1 == rend ? 0 : L.tour[pos1 + 1];
            // d(n1, s1) is the edge leaving position pos1
            bool found = two_opt_scan<true>(p, L, rbeg, rend, pos1, n1, s1, L.dtw[pos1], evals,
                                            h1, h2, h3, h4);
            if (!found) {
                const int p1 = pos1 == rbeg ? L.tour[rend] : L.tour[pos1 - 1];
                // d(p1, n1): the edge entering pos1, or the closing edge
                // (tour[rend], depot) when n1 is the depot
                const float rad = pos1 == rbeg ? L.dtw[rend] : L.dtw[pos1 - 1];
                found = two_opt_scan<false>(p, L, rbeg, rend, pos1, n1, p1, rad, evals, h1, h2, h3, h4);
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

// ------------------------------------------------- distance row cache ----

// Rows d(tour[q], .) live in slot q mod R of an LDS ring of R rows.  Every thread keeps
// the same copy of the slot tags (the node whose row a slot holds) in
// registers, so the fetch plan needs no serial thread and no LDS round
// trip.  The ring holds the current round's positions and the next round's
// without collision (at most 10 consecutive positions).
struct RowRing {
    int tag[kRowSlots];
};

__device__ __forceinline__ int ring_get(const RowRing &r, int slot)
{
    int v = -1;
#pragma unroll
    for (int s = 0; s < kRowSlots; ++s) v = s == slot ? r.tag[s] : v;
    return v;
}

__device__ __forceinline__ void ring_set(RowRing &r, int slot, int node)
{
#pragma unroll
    for (int s = 0; s < kRowSlots; ++s) r.tag[s] = s == slot ? node : r.tag[s];
}

__device__ __forceinline__ int ring_slot(const DevProblem &p, int pos)
{
    return pos & (ls_ring_slots(p.n) - 1);
}

__device__ __forceinline__ const float *row_at(const DevProblem &p, const LsLds &L, int pos)
{
    return L.rows + ring_slot(p, pos) * p.n;
}

// Make the rows of positions [first, first+cnt) present (cnt <= 6).  All
// threads load the missing ones; a barrier follows when any was missing
// (the decision is uniform: tags and the tour are the same for all threads).
__device__ __forceinline__ void rows_ensure(const DevProblem &p, LsLds &L, RowRing &ring, int first,
                                            int cnt, int tid)
{
    const int n = p.n;
    bool loaded = false;
#pragma unroll
    for (int q = 0; q < kLsWaves + 2; ++q) {
        if (q >= cnt) break;
        const int pos = first + q;
        const int v = L.tour[pos];
        const int slot = ring_slot(p, pos);
        if (ring_get(ring, slot) != v) {
            float *row = L.rows + slot * n;
            const float *src = p.dist + (size_t)v * p.P;
            for (int c = tid; c < n; c += kLsThreads) row[c] = src[c];
            ring_set(ring, slot, v);
            loaded = true;
        }
    }
    if (loaded) __syncthreads();
}

// Next round's rows (positions [first, first+cnt), cnt <= 4) are loaded
// into registers before a scan and stored into the ring after it, so the
// loads overlap the scan.  Only for n <= 1024 (4 elements per thread).
struct RowPrefetch {
    float v[kLsWaves][kPrefetch];
    int node[kLsWaves];
};

__device__ __forceinline__ void prefetch_issue(const DevProblem &p, const LsLds &L, const RowRing &ring,
                                               RowPrefetch &pf, int first, int cnt, int S, int tid)
{
#pragma unroll
    for (int q = 0; q < kLsWaves; ++q) {
        pf.node[q] = -1;
        const int pos = first + q;
        if (q < cnt && pos < S && p.n <= kPrefetchMaxN) {
            const int v = L.tour[pos];
            if (ring_get(ring, ring_slot(p, pos)) != v) {
                pf.node[q] = v;
                const float *src = p.dist + (size_t)v * p.P;
#pragma unroll
                for (int e = 0; e < kPrefetch; ++e) {
                    const int c = tid + e * kLsThreads;
                    pf.v[q][e] = c < p.n ? src[c] : 0.0f;
                }
            }
        }
    }
}

__device__ __forceinline__ void prefetch_store(const DevProblem &p, LsLds &L, RowRing &ring,
                                               const RowPrefetch &pf, int first, int tid)
{
#pragma unroll
    for (int q = 0; q < kLsWaves; ++q) {
        if (pf.node[q] < 0) continue;
        const int slot = ring_slot(p, first + q);
        float *row = L.rows + slot * p.n;
#pragma unroll
        for (int e = 0; e < kPrefetch; ++e) {
            const int c = tid + e * kLsThreads;
            if (c < p.n) row[c] = pf.v[q][e];
        }
        ring_set(ring, slot, pf.node[q]);
    }
}

// -------------------------------------------------------------- swap ----

// Position i against the current tour; returns the first improving j or -1.
__device__ __forceinline__ int swap_scan(const DevProblem &p, const LsLds &L, int S, int i,
                                         const float *rowA, const float *rowB, const float *rowC,
                                         int lane, long long &ev)
{
    const int n1 = L.tour[i];
    const int sn1 = L.tour[i + 1], r1 = L.ridx[i];
    const float a1 = rowA[n1], b1 = rowB[sn1];
    const int dn1 = L.dem[n1];
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
                    gain = -(a1 + L.dtw[j]) + (rowA[n2] + rowB[sn2]);
                else
                    gain = -(a1 + b1 + L.dtw[j - 1] + L.dtw[j]) +
                           (rowA[n2] + rowC[n2] + rowB[pn2] + rowB[sn2]);
                if (gain < -kEps) {
                    if (r1 != r2) {
                        const int dn2 = L.dem[n2];
                        const int load1 = L.rl[r1] - dn1 + dn2, load2 = L.rl[r2] - dn2 + dn1;
                        const float dist1 = (L.rd[r1] - (a1 + b1)) + (rowA[n2] + rowC[n2]);
                        const float dist2 = (L.rd[r2] - (L.dtw[j - 1] + L.dtw[j])) +
                                            (rowB[pn2] + rowB[sn2]);
                        ok = !(load1 > p.Q || load2 > p.Q || dist1 > p.D || dist2 > p.D);
                    } else {
                        ok = true;
                    }
                }
            }
        }
        const u64 mk = __ballot(ok);
        if (mk) return jb + __ffsll(mk) - 1;
    }
    return -1;
}

__device__ __forceinline__ void ls_swap(const DevProblem &p, LsLds &L, RowRing &ring, int S, int tid,
                                        long long &ev, int &moves)
{
    const int lane = tid & 63, w = tid >> 6;
    const int cap = kMaxMovesFactor * p.n;
    int done = 0, i0 = 1, parity = 0;
    while (i0 < S) {
        // rows of positions i0-1 .. i0+W (pred, self, succ of each scanned i)
        rows_ensure(p, L, ring, i0 - 1, min(kLsWaves + 2, S - i0 + 1), tid);
        RowPrefetch pf;
        prefetch_issue(p, L, ring, pf, i0 + kLsWaves + 1, kLsWaves, S, tid);
        const int i = i0 + w;
        int j = -1;
        if (i < S && L.tour[i] != 0)
            j = swap_scan(p, L, S, i, row_at(p, L, i - 1), row_at(p, L, i), row_at(p, L, i + 1), lane, ev);
        int *res = L.res + 4 * parity;
        parity ^= 1;
        if (lane == 0) res[w] = j;
        prefetch_store(p, L, ring, pf, i0 + kLsWaves + 1, tid);
        __syncthreads();
        int wstar = -1;
        for (int q = 0; q < kLsWaves; ++q)
            if (res[q] >= 0) { wstar = q; break; }
        if (wstar < 0) {
            // no barrier: the next round writes the other result buffer, and
            // rows it loads belong to positions no wave still reads
            i0 += kLsWaves;
            continue;
        }
        const int is = i0 + wstar, js = res[wstar];
        if (tid == 0) {
            const float *rowA = row_at(p, L, is - 1), *rowB = row_at(p, L, is), *rowC = row_at(p, L, is + 1);
            const int n1 = L.tour[is], sn1 = L.tour[is + 1], r1 = L.ridx[is];
            const int n2 = L.tour[js], pn2 = L.tour[js - 1], sn2 = L.tour[js + 1], r2 = L.ridx[js];
            const float a1 = rowA[n1], b1 = rowB[sn1];
            const int dn1 = L.dem[n1];
            int load1, load2;
            float dist1, dist2;
            if (r1 != r2) {
                const int dn2 = L.dem[n2];
                load1 = L.rl[r1] - dn1 + dn2;
                load2 = L.rl[r2] - dn2 + dn1;
                dist1 = (L.rd[r1] - (a1 + b1)) + (rowA[n2] + rowC[n2]);
                dist2 = (L.rd[r2] - (L.dtw[js - 1] + L.dtw[js])) + (rowB[pn2] + rowB[sn2]);
            } else {
                float gain;
                if (js == is + 1)
                    gain = -(a1 + L.dtw[js]) + (rowA[n2] + rowB[sn2]);
                else
                    gain = -(a1 + b1 + L.dtw[js - 1] + L.dtw[js]) +
                           (rowA[n2] + rowC[n2] + rowB[pn2] + rowB[sn2]);
                load1 = L.rl[r1];
                load2 = load1;
                dist1 = L.rd[r1] + gain;
                dist2 = dist1;
            }
            L.tour[is] = n2;
            L.tour[js] = n1;
            L.rl[r1] = load1; L.rl[r2] = load2;
            L.rd[r1] = dist1; L.rd[r2] = dist2;
            L.dtw[is - 1] = Dd(p, L.tour[is - 1], L.tour[is]);
            L.dtw[is] = Dd(p, L.tour[is], L.tour[is + 1]);
            L.dtw[js - 1] = Dd(p, L.tour[js - 1], L.tour[js]);
            L.dtw[js] = Dd(p, L.tour[js], L.tour[js + 1]);
        }
        __syncthreads();
        i0 = is;  // re-examine the position that moved (the reference's i--)
        if (++done >= cap) break;
    }
    if (tid == 0) moves += done;
}

// ---------------------------------------------------------- relocate ----

__device__ __forceinline__ int relocate_scan(const DevProblem &p, const LsLds &L, int S, int i,
                                             const float *rowP, const float *rowB, int lane,
                                             long long &ev)
{
    const int n1 = L.tour[i];
    const int r1 = L.ridx[i], sn1 = L.tour[i + 1];
    const float a1 = L.dtw[i - 1], b1 = L.dtw[i], c1 = rowP[sn1];
    const int dn1 = L.dem[n1];
    for (int jb = 1; jb < S - 1; jb += kLanes) {
        const int j = jb + lane;
        bool ok = false;
        if (j < S - 1 && j != i) {
            const int n2 = L.tour[j];
            if (n2 != 0) {
                ev++;
                const int pn2 = L.tour[j - 1], sn2 = L.tour[j + 1], r2 = L.ridx[j];
                float rm, add;
                if (i > j) { rm = L.dtw[j - 1]; add = rowB[n2] + rowB[pn2]; }
                else { rm = L.dtw[j]; add = rowB[n2] + rowB[sn2]; }
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
        if (mk) return jb + __ffsll(mk) - 1;
    }
    return -1;
}

__device__ __forceinline__ void ls_relocate(const DevProblem &p, LsLds &L, RowRing &ring, int S,
                                            int tid, long long &ev, int &moves)
{
    const int lane = tid & 63, w = tid >> 6;
    const int cap = kMaxMovesFactor * p.n;
    int done = 0, i0 = 1, parity = 0;
    while (i0 < S - 1 && done < cap) {
        // rows of positions i0-1 .. i0+W-1: d(pred, .) gives c1, d(n1, .) the gains
        rows_ensure(p, L, ring, i0 - 1, min(kLsWaves + 1, S - i0), tid);
        RowPrefetch pf;
        prefetch_issue(p, L, ring, pf, i0 + kLsWaves, kLsWaves, S - 1, tid);
        const int i = i0 + w;
        int j = -1;
        if (i < S - 1 && L.tour[i] != 0 && L.rc[L.ridx[i]] >= 2)
            j = relocate_scan(p, L, S, i, row_at(p, L, i - 1), row_at(p, L, i), lane, ev);
        int *res = L.res + 4 * parity;
        parity ^= 1;
        if (lane == 0) res[w] = j;
        prefetch_store(p, L, ring, pf, i0 + kLsWaves, tid);
        __syncthreads();
        int wstar = -1;
        for (int q = 0; q < kLsWaves; ++q)
            if (res[q] >= 0) { wstar = q; break; }
        if (wstar < 0) {
            // no barrier: the next round writes the other result buffer, and
            // rows it loads belong to positions no wave still reads
            i0 += kLsWaves;
            continue;
        }
        const int is = i0 + wstar, js = res[wstar];
        const int n1 = L.tour[is];
        const int lo = is < js ? is : js, hi = is < js ? js : is;
        if (tid == 0) {
            const float *rowP = row_at(p, L, is - 1), *rowB = row_at(p, L, is);
            const int r1 = L.ridx[is], sn1 = L.tour[is + 1];
            const int n2 = L.tour[js], pn2 = L.tour[js - 1], sn2 = L.tour[js + 1], r2 = L.ridx[js];
            const float a1 = L.dtw[is - 1], b1 = L.dtw[is], c1 = rowP[sn1];
            const int dn1 = L.dem[n1];
            float rm, add;
            if (is > js) { rm = L.dtw[js - 1]; add = rowB[n2] + rowB[pn2]; }
            else { rm = L.dtw[js]; add = rowB[n2] + rowB[sn2]; }
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
        for (int q = lo + tid; q <= hi; q += kLsThreads) L.tmp[q] = L.tour[q];
        __syncthreads();
        for (int q = lo + tid; q <= hi; q += kLsThreads) {
            int v;
            if (is < js) v = q == js ? n1 : L.tmp[q + 1];
            else v = q == js ? n1 : L.tmp[q - 1];
            L.tour[q] = v;
        }
        __syncthreads();
        ls_edges(p, L, lo - 1, hi, tid);
        ls_routes(p, L, S, tid, false);  // refreshes ridx (and pos/zpos)
        i0 = is;
        ++done;
    }
    if (tid == 0) moves += done;
}

// ------------------------------------------------------------ kernel ----

__global__ __launch_bounds__(kLsThreads) void k_local_search(DevBatch bt, int flags)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int inst = bt.ant_inst[blockIdx.x];
    const DevProblem p = bt.probs[inst];
    const DevState s = bt.states[inst];
    const int a = blockIdx.x - bt.ant_base[inst], tid = threadIdx.x;
    LsLds L = ls_bind(p.n, smem);
    const int S = s.sizes[a];
    int *gt = s.tours + (size_t)a * p.T;
    for (int q = tid; q < S; q += kLsThreads) L.tour[q] = gt[q];
    for (int v = tid; v < p.n; v += kLsThreads) {
        L.dlb[v] = 0;
        L.dem[v] = p.dem[v];
    }
    __syncthreads();
    RowRing ring;
#pragma unroll
    for (int q = 0; q < kRowSlots; ++q) ring.tag[q] = -1;
    long long ev = 0;
    int moves = 0;
    if (flags & kLs2Opt) {
        ls_edges(p, L, 0, S - 2, tid);
        const int R = ls_routes(p, L, S, tid, false);
        for (int r = tid; r < R; r += kLsThreads)
            two_opt_route(p, L, r, L.zpos[r], L.zpos[r + 1] - 1, ev, moves);
        __syncthreads();
    }
    if (flags & kLsSwap) {
        ls_edges(p, L, 0, S - 2, tid);
        __syncthreads();
        ls_routes(p, L, S, tid, true);
        ls_swap(p, L, ring, S, tid, ev, moves);
    }
    if (flags & kLsRelocate) {
        ls_edges(p, L, 0, S - 2, tid);
        __syncthreads();
        ls_routes(p, L, S, tid, true);
        ls_relocate(p, L, ring, S, tid, ev, moves);
    }
    ls_edges(p, L, 0, S - 2, tid);
    __syncthreads();
    for (int q = tid; q < S; q += kLsThreads) gt[q] = L.tour[q];
    // tour length in position order (vrpHelper.cpp:244-253)
    if (tid == 0) {
        float len = 0.0f;
        for (int q = 0; q < S - 1; ++q) len += L.dtw[q];
        s.lens[a] = len;
    }
    __shared__ long long red_ev[kLsWaves];
    __shared__ int red_mv[kLsWaves];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ev += __shfl_xor(ev, o);
        moves += __shfl_xor(moves, o);
    }
    if ((tid & 63) == 0) {
        red_ev[tid >> 6] = ev;
        red_mv[tid >> 6] = moves;
    }
    __syncthreads();
    if (tid == 0) {
        long long e = 0;
        int mv = 0;
        for (int q = 0; q < kLsWaves; ++q) { e += red_ev[q]; mv += red_mv[q]; }
        s.ant_evals[a] = e;
        s.ant_moves[a] = mv;
    }
}

hipError_t launch_local_search(const DevBatch &b, int flags, hipStream_t st)
{
    if (!flags) return hipSuccess;
    hipLaunchKernelGGL(k_local_search, dim3(b.ants), dim3(kLsThreads), local_search_lds_bytes(b.max_n),
                       st, b, flags);
    return hipGetLastError();
}

}  // namespace cvrp
