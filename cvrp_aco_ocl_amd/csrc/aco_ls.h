// Local search device code (gfx950): one block of W wavefronts per ant; the
// ant's tour, per-route tables and a ring of distance rows in LDS.
//
//   2-opt     the reference's GPU 2-opt (kernel.cl:126-319): per route,
//             sequential scan order, first improvement over the nn list
//             with radius pruning, don't-look bits.  Routes are independent
//             (the depot's don't-look bit starts clear on route 0 and set on
//             later routes); a group of 8 lanes searches one route.
//   swap      LocalSearch::swap (localSearch.cpp:301-410): for each position
//             i in order, the first j > i whose swap gains more than
//             EPSILON and keeps loads and route lengths feasible.
//   relocate  the same scan for insertion moves (neighbourSearch.cpp:280-293,
//             move.cpp:32-51).
//
// swap and relocate run speculatively on W consecutive positions: wave w
// scans position i0+w against the current tour (4 x 64 j per loop trip).
// The smallest position with a move is exactly the sequential scan's next
// move, because no earlier position moved; results of later positions are
// discarded.  Distance rows of the positions around i0 sit in an LDS ring
// and the next round's rows are fetched while the current round scans.
//
// Launch shape (LsShape): W (4, 8, 12 or 16), the ring size, prefetch on/off
// and the row type; every shape gives the same tours and the same
// evaluated-move counts.
#pragma once

#include "aco_device.h"

namespace cvrp {

using u64 = unsigned long long;

// LDS words of L.res for W waves (W <= 16): [0,W) and [W,2W) per-wave scan
// results (accepted j) of even and odd rounds, 4W the route count, 4W+1 the
// lowest wave of the current round with a move (round key).
constexpr int kResWords = 80;

// Candidates of one scan trip, and the padding of the LDS tour and edge
// arrays past the tour's end that lets a trip read without clamping.
// Candidates per lane per scan trip.  Smaller trips waste less work past an
// early move and let waves above a lower move stop sooner; larger ones hide
// more latency.  Measured best: swap 2; relocate 2 with 12 or 16 waves per
// ant (one ant per CU), 4 with fewer (several ants per CU).
constexpr int kSwapUnroll = 2;
template <int W> constexpr int kRelocateUnroll = W >= 12 ? 2 : 4;
constexpr int kScanUnroll = 4;   // the largest of the above
constexpr int kScanPad = kScanUnroll * 64 + 4;

template <int W> struct LsW {
    static_assert(W == 4 || W == 8 || W == 12 || W == 16, "4, 8, 12 or 16 waves per ant");
    static constexpr int threads = W * kLanes;
    static constexpr int res_r = 4 * W;
    static constexpr int res_first = 4 * W + 1;
};

struct LsLds {
    int *tour, *rl, *dem, *res;
    float *dtw, *rd, *rows;
    short *pos, *rof, *ridx, *zpos, *rc, *perm;
    unsigned char *dlb;
};

__host__ __device__ constexpr size_t ls_align16(size_t x) { return (x + 15) & ~size_t(15); }

struct LsOffsets {
    size_t tour, dtw, rl, rd, dem, rows, res, pos, rof, ridx, zpos, rc, perm, dlb, total;
};

// perm (the random scan order of ACO_LS_2OPT_RANDOM) takes LDS only when
// that flag is set.
__host__ __device__ inline LsOffsets ls_offsets(int n, int ring, int rowbytes, bool perm = false)
{
    const size_t T = 2 * (size_t)n, N = (size_t)n;
    LsOffsets o;
    size_t at = 0;
    o.tour = at; at = ls_align16(at + 4 * (T + kScanPad));
    o.dtw = at;  at = ls_align16(at + 4 * (T + kScanPad));
    o.rl = at;   at = ls_align16(at + 4 * N);
    o.rd = at;   at = ls_align16(at + 4 * N);
    o.dem = at;  at = ls_align16(at + 4 * N);
    // row slots of a whole number of 16-byte pieces (the row copy's unit)
    o.rows = at; at = ls_align16(at + (size_t)rowbytes * (size_t)(rowbytes == 4 ? (n + 3) & ~3 : (n + 7) & ~7) * (size_t)ring);
    o.res = at;  at = ls_align16(at + 4 * kResWords);
    o.pos = at;  at = ls_align16(at + 2 * N);
    o.rof = at;  at = ls_align16(at + 2 * N);
    o.ridx = at; at = ls_align16(at + 2 * (T + kScanPad));
    o.zpos = at; at = ls_align16(at + 2 * (N + 1));
    o.rc = at;   at = ls_align16(at + 2 * N);
    o.perm = at; at = ls_align16(at + (perm ? 2 * T : 0));
    o.dlb = at;  at = ls_align16(at + N);
    o.total = at;
    return o;
}

__device__ __forceinline__ LsLds ls_bind(int n, int ring, int rowbytes, bool perm, unsigned char *b)
{
    const LsOffsets o = ls_offsets(n, ring, rowbytes, perm);
    LsLds L;
    L.tour = (int *)(b + o.tour); L.dtw = (float *)(b + o.dtw);
    L.rl = (int *)(b + o.rl); L.rd = (float *)(b + o.rd); L.dem = (int *)(b + o.dem);
    L.rows = (float *)(b + o.rows); L.res = (int *)(b + o.res);
    L.pos = (short *)(b + o.pos); L.rof = (short *)(b + o.rof); L.ridx = (short *)(b + o.ridx);
    L.zpos = (short *)(b + o.zpos); L.rc = (short *)(b + o.rc); L.perm = (short *)(b + o.perm);
    L.dlb = b + o.dlb;
    return L;
}


__device__ __forceinline__ float Dd(const DevProblem &p, int a, int b)
{
    return p.dist[(size_t)a * p.P + b];
}

// dtw[q] = d(tour[q], tour[q+1]) for q in [lo, hi]
template <int W>
__device__ __forceinline__ void ls_edges(const DevProblem &p, LsLds &L, int lo, int hi, int tid)
{
    for (int q = lo + tid; q <= hi; q += LsW<W>::threads) L.dtw[q] = Dd(p, L.tour[q], L.tour[q + 1]);
}

// Wave 0: depot positions (zpos), customer positions and route ids (pos,
// rof) and route index per position (ridx, counted as localSearch.cpp:319-337
// does).  Then, with tables, the per-route load / distance+service /
// customer count, one thread per route.  Returns R.
template <int W>
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
        if (lane == 0) L.res[LsW<W>::res_r] = zc - 1;
    }
    __syncthreads();
    const int R = L.res[LsW<W>::res_r];
    if (tables) {
        for (int r = tid; r < R; r += LsW<W>::threads) {
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

// A route is searched by a group of kOptGroup lanes of one wave: the
// candidate scan is split over the lanes and reduced; moves are applied
// with the reversals split over the lanes.
constexpr int kOptGroup = 8;

// Minimum over the aligned group of 8 lanes (lane ^ 1, ^ 2, ^ 4 stay inside
// the group, all of whose lanes are active).
__device__ __forceinline__ u64 group_min_u64(u64 v)
{
    static_assert(kOptGroup == 8, "the DPP exchanges cover groups of 8 lanes");
    const int lane = (int)(threadIdx.x & 63);
    u64 w = dpp_u64<kDppXor1>(v);
    v = w < v ? w : v;
    w = dpp_u64<kDppXor2>(v);
    v = w < v ? w : v;
    w = dpp_xor4_u64(v, lane);
    return w < v ? w : v;
}

__device__ __forceinline__ void group_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The two scans for n1: the successor scan (partner edges (n1,s1),(n2,s2),
// radius d(n1,s1)) and, when it finds no move, the predecessor scan (edges
// (p1,n1),(p2,n2), radius d(p1,n1)).  The reference walks n1's nn row,
// skips nodes of other routes and stops at the first in-route node with
// radius - d <= EPS (kernel.cl:180-230).  Rows are sorted by distance, so
// the nodes it evaluates are exactly the in-route nodes with
// radius - d > EPS that appear in the row, and the move it applies is the
// improving one earliest in the row.  Here the route's own positions are
// scanned (independent loads, no walk over other routes' neighbours) and
// the improving candidate with the smallest row position nnpos[n1][n2] is
// kept (row positions are distinct, so the group minimum is unique).  Both
// scans take the same candidates n2, with the same row positions and
// d(n1, n2), so one pass evaluates both: one round of memory latency per
// position instead of two.
//
// Candidate q (q may lie past rend: the caller masks it).  Loads are
// unconditional (clamped, in-bounds indices) so a lane's candidates have
// all their loads in flight at once.  ok*: in the row (rp < nn_ls), inside
// the radius and, for the predecessor scan, not adjacent; gains as the
// reference computes them; xs = s2, xp = p2.
__device__ __forceinline__ void two_opt_cand(const DevProblem &p, const LsLds &L, int rend, int pos1, int q,
                                             int n1, int s1, int p1, float rads, float radp,
                                             const unsigned short *rowpos, bool &oks, bool &okp, float &gs,
                                             float &gp, int &rp, int &n2, int &xs, int &xp)
{
    const int qq = q <= rend ? q : rend;
    n2 = L.tour[qq];
    xs = qq == rend ? 0 : L.tour[qq + 1];
    xp = L.tour[qq - 1];
    const float ews = L.dtw[qq], ewp = L.dtw[qq - 1];
    rp = rowpos[n2];
    const float dn = Dd(p, n1, n2);
    const float dxs = Dd(p, s1, xs), dxp = Dd(p, p1, xp);
    const bool base = (q <= rend) & (q != pos1) & (rp < p.nn_ls);
    oks = base & (rads - dn > kEps);
    okp = base & (radp - dn > kEps) & !(xp == n1 || p1 == n2);
    gs = -rads + dn + dxs - ews;
    gp = -radp + dn + dxp - ewp;
}

// Candidates per lane per trip of a 2-opt scan: positions q, q+8, q+16 and
// q+24 of the group's lanes, all loads issued before any is used (routes of
// up to 32 customers take one trip, i.e. one memory latency).
constexpr int kOptUnroll = 4;

// Returns 0 (no move), 1 (successor scan's move) or 2 (predecessor scan's
// move) with h1..h4 set (same in all lanes of the group).  The evaluated
// moves are the sequential pair's: the successor scan's up to its move, or
// all of them plus the predecessor scan's up to its move.
__device__ __forceinline__ int two_opt_scans(const DevProblem &p, const LsLds &L, int rbeg, int rend, int pos1,
                                             int n1, int s1, int p1, float rads, float radp, int gl,
                                             long long &evals, int &h1, int &h2, int &h3, int &h4)
{
    constexpr int U = kOptUnroll;
    const unsigned short *rowpos = p.nnpos + (size_t)n1 * p.n;
    u64 bs = ~0ull, bp = ~0ull;
    int cs = 0, cp = 0;
    // the first trip's candidates, kept for the evaluated-move count
    bool oks_k[U], okp_k[U];
    int rp_k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { oks_k[u] = okp_k[u] = false; rp_k[u] = 0; }
    for (int q0 = rbeg + 1 + gl; q0 <= rend; q0 += U * kOptGroup) {
        bool oks[U], okp[U];
        float gs[U], gp[U];
        int rp[U], a[U], xs[U], xp[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            two_opt_cand(p, L, rend, pos1, q0 + u * kOptGroup, n1, s1, p1, rads, radp, rowpos, oks[u], okp[u],
                         gs[u], gp[u], rp[u], a[u], xs[u], xp[u]);
        const bool first = q0 == rbeg + 1 + gl;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cs += oks[u];
            cp += okp[u];
            const u64 kr = ((u64)rp[u] << 32) | ((u64)a[u] << 16);
            if (oks[u] && gs[u] < -kEps) bs = (kr | (u64)xs[u]) < bs ? (kr | (u64)xs[u]) : bs;
            if (okp[u] && gp[u] < -kEps) bp = (kr | (u64)xp[u]) < bp ? (kr | (u64)xp[u]) : bp;
            if (first) { oks_k[u] = oks[u]; okp_k[u] = okp[u]; rp_k[u] = rp[u]; }
        }
    }
    int kind = 0;
    u64 best = group_min_u64(bs);
    if (best != ~0ull) {
        kind = 1;
    } else {
        best = group_min_u64(bp);
        if (best != ~0ull) kind = 2;
    }
    if (kind == 0) {
        evals += cs + cp;
        return 0;
    }
    // the reference's walk stops at the accepted candidate: it evaluated
    // the candidates up to that row position
    const int rps = (int)(best >> 32);
    int cnt = 0;
    if (rend - rbeg <= U * kOptGroup) {
#pragma unroll
        for (int u = 0; u < U; ++u) cnt += (kind == 1 ? oks_k[u] : okp_k[u]) && rp_k[u] <= rps;
    } else {
        for (int q0 = rbeg + 1 + gl; q0 <= rend; q0 += U * kOptGroup) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                bool oks, okp;
                float gs, gp;
                int rp, a, xs, xp;
                two_opt_cand(p, L, rend, pos1, q0 + u * kOptGroup, n1, s1, p1, rads, radp, rowpos, oks, okp, gs,
                             gp, rp, a, xs, xp);
                cnt += (kind == 1 ? oks : okp) && rp <= rps;
            }
        }
    }
    // the predecessor scan runs after the whole successor scan
    evals += cnt + (kind == 2 ? cs : 0);
    const int b2 = (int)((best >> 16) & 0xFFFF), bx = (int)(best & 0xFFFF);
    if (kind == 1) { h1 = n1; h2 = s1; h3 = b2; h4 = bx; }
    else { h1 = p1; h2 = n1; h3 = bx; h4 = b2; }
    return kind;
}

// Scan order of the CPU local search (generate_random_permutation,
// localSearch.cpp:97-119): offsets 0..num-1 of the route's positions,
// shuffled with one draw each from the Park-Miller state z; written by one
// lane to perm[rbeg .. rbeg+num).  A draw of exactly 1.0 (state >=
// 2147483584) would index past the end, where the reference asserts; it
// takes the last offset instead.
__device__ __forceinline__ void route_permutation(LsLds &L, int rbeg, int num, int z)
{
    short *r = L.perm + rbeg;
    for (int i = 0; i < num; ++i) r[i] = (short)i;
    for (int i = 0; i < num; ++i) {
        int k = z / 127773;
        z = 16807 * (z - k * 127773) - 2836 * k;
        if (z < 0) z += 2147483647;
        const float rnd = (float)z * 0x1p-31f;
        int node = (int)(rnd * (float)(num - i));
        if (i + node >= num) node = num - 1 - i;
        const short h = r[i];
        r[i] = r[i + node];
        r[i + node] = h;
    }
}

// Position of the l-th scan step: in sequence, or the random order.
template <bool RANDOM>
__device__ __forceinline__ int scan_pos(const LsLds &L, int rbeg, int l)
{
    return RANDOM ? rbeg + L.perm[rbeg + l] : rbeg + l;
}

template <bool RANDOM>
__device__ __forceinline__ bool two_opt_skip(const LsLds &L, int rbeg, int l, bool dlb0)
{
    const int v = L.tour[scan_pos<RANDOM>(L, rbeg, l)];
    return v == 0 ? dlb0 : (L.dlb[v] != 0);
}

// First scan index in [l, num) whose position is not skipped, or num: the
// GW lanes of the calling group (gl = 0 .. GW-1, an aligned run of lanes)
// test GW indices at once and take the lowest from a ballot.  No don't-look
// bit changes between the tests of a sequential walk, so this is the index
// that walk stops at.
template <bool RANDOM, int GW>
__device__ __forceinline__ int two_opt_next(const LsLds &L, int rbeg, int l, int num, bool dlb0, int gl)
{
    const int gshift = (int)(threadIdx.x & 63) & ~(GW - 1);
    for (int base = l; base < num; base += GW) {
        const int q = base + gl;
        const bool ok = q < num && !two_opt_skip<RANDOM>(L, rbeg, q < num ? q : base, dlb0);
        const unsigned gm = (unsigned)(__ballot(ok) >> gshift) & ((1u << GW) - 1u);
        if (gm) return base + __builtin_ctz(gm);
    }
    return num;
}

// kernel.cl:126-277 for route r at positions [rbeg, rend] (rbeg holds the
// depot), run by the kOptGroup lanes gl = 0.. of one group; RANDOM scans
// the positions in the order of perm (localSearch.cpp:180-294).
template <bool RANDOM>
__device__ __forceinline__ void two_opt_route(const DevProblem &p, LsLds &L, int r, int rbeg,
                                              int rend, int gl, long long &evals, int &moves_total)
{
    bool dlb0 = r > 0;
    const int num = rend - rbeg + 1;
    int moves = 0;
    bool imp = true;
    while (imp) {
        imp = false;
        for (int l = 0; l < num; ++l) {
            l = two_opt_next<RANDOM, kOptGroup>(L, rbeg, l, num, dlb0, gl);
            if (l >= num) break;
            const int pos1 = scan_pos<RANDOM>(L, rbeg, l);
            const int n1 = L.tour[pos1];
            int h1 = 0, h2 = 0, h3 = 0, h4 = 0;
            const int s1 = pos1 == rend ? 0 : L.tour[pos1 + 1];
            const int p1 = pos1 == rbeg ? L.tour[rend] : L.tour[pos1 - 1];
            // d(n1, s1) is the edge leaving position pos1; d(p1, n1) the edge
            // entering it, or the closing edge (tour[rend], depot) when n1 is
            // the depot
            const float rads = L.dtw[pos1], radp = pos1 == rbeg ? L.dtw[rend] : L.dtw[pos1 - 1];
            const bool found =
                two_opt_scans(p, L, rbeg, rend, pos1, n1, s1, p1, rads, radp, gl, evals, h1, h2, h3, h4) != 0;
            if (!found) {
                if (n1 == 0) dlb0 = true;
                else if (gl == 0) L.dlb[n1] = 1;
                group_sync();
                continue;
            }
            imp = true;
            const int hs[4] = {h1, h2, h3, h4};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (hs[q] == 0) dlb0 = false;
                else if (gl == 0) L.dlb[hs[q]] = 0;
            }
            int ph1 = h1 == 0 ? rbeg : L.pos[h1];
            int ph3 = h3 == 0 ? rbeg : L.pos[h3];
            if (ph3 < ph1) {
                int t = h1; h1 = h3; h3 = t;
                t = h2; h2 = h4; h4 = t;
                t = ph1; ph1 = ph3; ph3 = t;
            }
            const int tb = h2 == 0 ? rbeg : L.pos[h2];
            group_sync();
            // edge weights: the two new edges, the segment between reversed
            if (gl == 0) {
                L.dtw[ph1] = Dd(p, h1, h3);
                L.dtw[ph3] = Dd(p, h2, h4);
            }
            const int ed = (ph3 - ph1 - 1) / 2;
            for (int t = gl; t < ed; t += kOptGroup) {
                const int i = ph1 + 1 + t, j = ph3 - 1 - t;
                const float a = L.dtw[i];
                L.dtw[i] = L.dtw[j];
                L.dtw[j] = a;
            }
            // tour positions [tb, ph3] reversed
            const int tp = (ph3 - tb + 1) / 2;
            for (int t = gl; t < tp; t += kOptGroup) {
                const int i = tb + t, j = ph3 - t;
                const int a = L.tour[i], b = L.tour[j];
                L.tour[i] = b;
                L.tour[j] = a;
                L.pos[a] = (short)j;
                L.pos[b] = (short)i;
            }
            group_sync();
            if (++moves >= kMaxMovesPerRoute) break;
        }
        if (moves >= kMaxMovesPerRoute) break;
    }
    if (gl == 0) moves_total += moves;
}

// kind (2 bits) and the move's four nodes (13 bits each; n <= 4096)
__device__ __forceinline__ u64 two_opt_pack(int kind, int h1, int h2, int h3, int h4)
{
    return ((u64)kind << 52) | ((u64)h1 << 39) | ((u64)h2 << 26) | ((u64)h3 << 13) | (u64)h4;
}

// The same search with two positions at a time (16 lanes per route): the
// halves of the group scan the next two positions whose don't-look bits are
// clear.  The second result stands only when the first position has no
// move: a position without a move changes nothing but its own don't-look
// bit, which the scans do not read, so the second scan saw what the
// sequential search would see.  After a move at the first position the
// search continues right after it and scans the second again.  Halves
// exchange their results with one DPP swap (16-lane groups are DPP rows).
template <bool RANDOM>
__device__ __forceinline__ void two_opt_route_pair(const DevProblem &p, LsLds &L, int r, int rbeg, int rend,
                                                   int gl16, long long &evals, int &moves_total)
{
    constexpr int G = 2 * kOptGroup;
    const int half = gl16 / kOptGroup, hl = gl16 % kOptGroup, lane = (int)(threadIdx.x & 63);
    bool dlb0 = r > 0;
    const int num = rend - rbeg + 1;
    int moves = 0;
    bool imp = true;
    while (imp) {
        imp = false;
        int l = 0;
        for (;;) {
            const int la = two_opt_next<RANDOM, G>(L, rbeg, l, num, dlb0, gl16);
            if (la >= num) break;
            const int lb = two_opt_next<RANDOM, G>(L, rbeg, la + 1, num, dlb0, gl16);
            const int na = L.tour[scan_pos<RANDOM>(L, rbeg, la)];
            const int nb = lb < num ? L.tour[scan_pos<RANDOM>(L, rbeg, lb)] : -1;
            const int lm = half == 0 ? la : lb;
            int kind = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0;
            long long ev = 0;
            if (lm < num) {
                const int pos1 = scan_pos<RANDOM>(L, rbeg, lm);
                const int n1 = L.tour[pos1];
                const int s1 = pos1 == rend ? 0 : L.tour[pos1 + 1];
                const int p1 = pos1 == rbeg ? L.tour[rend] : L.tour[pos1 - 1];
                const float rads = L.dtw[pos1], radp = pos1 == rbeg ? L.dtw[rend] : L.dtw[pos1 - 1];
                kind = two_opt_scans(p, L, rbeg, rend, pos1, n1, s1, p1, rads, radp, hl, ev, h1, h2, h3, h4);
            }
            const u64 mine = two_opt_pack(kind, h1, h2, h3, h4);
            const u64 other = dpp_xor8_u64(mine, lane);
            const u64 first = half == 0 ? mine : other, second = half == 0 ? other : mine;
            const bool k0 = (first >> 52) != 0, k1 = (second >> 52) != 0;
            // the second scan's evaluations count only when it stands
            if (half == 0 || !k0) evals += ev;
            if (!k0) {
                if (na == 0) dlb0 = true;
                else if (gl16 == 0) L.dlb[na] = 1;
                if (!k1 && nb >= 0) {
                    if (nb == 0) dlb0 = true;
                    else if (gl16 == 0) L.dlb[nb] = 1;
                }
            }
            if (!k0 && !k1) {
                group_sync();
                l = lb + 1;
                continue;
            }
            imp = true;
            const u64 win = k0 ? first : second;
            const int msk = (1 << 13) - 1;
            h1 = (int)(win >> 39) & msk;
            h2 = (int)(win >> 26) & msk;
            h3 = (int)(win >> 13) & msk;
            h4 = (int)win & msk;
            const int hs[4] = {h1, h2, h3, h4};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (hs[q] == 0) dlb0 = false;
                else if (gl16 == 0) L.dlb[hs[q]] = 0;
            }
            int ph1 = h1 == 0 ? rbeg : L.pos[h1];
            int ph3 = h3 == 0 ? rbeg : L.pos[h3];
            if (ph3 < ph1) {
                int t = h1; h1 = h3; h3 = t;
                t = h2; h2 = h4; h4 = t;
                t = ph1; ph1 = ph3; ph3 = t;
            }
            const int tb = h2 == 0 ? rbeg : L.pos[h2];
            group_sync();
            if (gl16 == 0) {
                L.dtw[ph1] = Dd(p, h1, h3);
                L.dtw[ph3] = Dd(p, h2, h4);
            }
            const int ed = (ph3 - ph1 - 1) / 2;
            for (int t = gl16; t < ed; t += G) {
                const int i = ph1 + 1 + t, j = ph3 - 1 - t;
                const float a = L.dtw[i];
                L.dtw[i] = L.dtw[j];
                L.dtw[j] = a;
            }
            const int tp = (ph3 - tb + 1) / 2;
            for (int t = gl16; t < tp; t += G) {
                const int i = tb + t, j = ph3 - t;
                const int a = L.tour[i], b = L.tour[j];
                L.tour[i] = b;
                L.tour[j] = a;
                L.pos[a] = (short)j;
                L.pos[b] = (short)i;
            }
            group_sync();
            l = (k0 ? la : lb) + 1;
            if (++moves >= kMaxMovesPerRoute) break;
        }
        if (moves >= kMaxMovesPerRoute) break;
    }
    if (gl16 == 0) moves_total += moves;
}

// ------------------------------------------------- distance row cache ----

// Rows d(tour[q], .) live in slot q mod R of an LDS ring of R rows (R = 8,
// 16 or 32, a launch parameter; a round needs W+2 positions, W+2 more with
// prefetch).  The window of positions [base, base + R) maps one to one onto the
// slots; bit o of a validity mask says that the slot of position base + o
// holds the row of the node now at that position.  Moving the window shifts
// the mask, a move clears the bits of the positions it changed.  All of it
// is uniform (every thread holds the same mask), so checks are a few scalar
// instructions and need no LDS round trip.
struct RowWindow {
    int base;
    unsigned mask;
};

// Rows hold fp32 distances, or (when every distance of the context is an
// integer in [0, 65535]) the same values as exact 16-bit integers, copied
// from the 16-bit distance matrix.
__device__ __forceinline__ float rget(const float *r, int i) { return r[i]; }
__device__ __forceinline__ float rget(const unsigned short *r, int i) { return (float)r[i]; }

template <typename RT> struct RowRing {
    RT *rows;
    int mask;     // ring size - 1
    int stride;   // elements per row slot: n rounded up to a whole number of 16-byte pieces
    __device__ __forceinline__ int slot(int pos) const { return pos & mask; }
    __device__ __forceinline__ RT *at(int pos) const { return rows + slot(pos) * stride; }
};

template <typename RT> __host__ __device__ constexpr int row_stride(int n)
{
    return sizeof(RT) == 4 ? (n + 3) & ~3 : (n + 7) & ~7;
}

// Source rows in global memory: fp32 distances, or their 16-bit copy.
__device__ __forceinline__ const float *row_src(const DevProblem &p, int v, float *) { return p.dist + (size_t)v * p.P; }
__device__ __forceinline__ const unsigned short *row_src(const DevProblem &p, int v, unsigned short *)
{
    return p.dist16 + (size_t)v * p.P;
}

// Copy distance row v into LDS with direct global->LDS loads issued by one
// wave: 16 bytes per lane per instruction (one instruction for a 16-bit
// row of up to 512 nodes), no registers; completes with vmcnt, so readers
// wait for vmcnt(0) and a barrier.  Source rows start on 16-byte
// boundaries (the row pitch P is a multiple of 8 and the matrices are
// 256-byte aligned) and so do the ring slots; the copy may run past n up
// to the next 16-byte boundary, which stays within P and within the slot.
template <typename RT>
__device__ __forceinline__ void row_copy_wave(const DevProblem &p, int v, RT *dst, int lane)
{
    const int nq = (p.n * (int)sizeof(RT) + 15) >> 4;
    const uint4 *src = reinterpret_cast<const uint4 *>(row_src(p, v, (RT *)nullptr));
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (int c = 0; c < nq; c += kLanes) {
        if (c + lane < nq)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + c + lane),
                                             (__attribute__((address_space(3))) void *)(d + c), 16, 0, 0);
    }
}

// Window offsets q < 32 with q mod W == wave.
template <int W> __device__ __forceinline__ unsigned wave_offsets(int wave)
{
    constexpr unsigned every = W == 4 ? 0x11111111u : W == 8 ? 0x01010101u : W == 12 ? 0x01001001u : 0x00010001u;
    return every << wave;
}

// Copy the rows of the window offsets in `need` (bit q: position base + q):
// offset q is issued by wave q mod W, so a run of up to W offsets is one row
// per wave and each wave finds its own with a few scalar instructions.
template <int W, typename RT>
__device__ __forceinline__ void rows_issue(const DevProblem &p, const LsLds &L, const RowRing<RT> &rr,
                                           int base, unsigned need, int wave, int lane)
{
    for (unsigned m = need & wave_offsets<W>(wave); m; m &= m - 1u) {
        const int pos = base + __builtin_ctz(m);
        row_copy_wave<RT>(p, __builtin_amdgcn_readfirstlane(L.tour[pos]), rr.at(pos), lane);
    }
}

// Start the window at position base (never moves backwards within a scan).
__device__ __forceinline__ void window_move(RowWindow &w, int base)
{
    const int d = base - w.base;
    w.mask = d >= 32 ? 0u : (w.mask >> d);
    w.base = base;
}

// Positions [lo, hi] changed: their rows are stale.
__device__ __forceinline__ void window_clear(RowWindow &w, int lo, int hi)
{
    const int a = lo - w.base > 0 ? lo - w.base : 0;
    const int b = hi - w.base < 31 ? hi - w.base : 31;
    if (a > b) return;
    const unsigned span = (b - a + 1) >= 32 ? 0xFFFFFFFFu : ((1u << (b - a + 1)) - 1u);
    w.mask &= ~(span << a);
}

// Make the rows of window offsets [0, cnt) present (cnt <= W+2), spread
// over the waves; a barrier follows when any was missing.
template <int W, typename RT>
__device__ __forceinline__ void rows_ensure(const DevProblem &p, const LsLds &L, const RowRing<RT> &rr,
                                            RowWindow &w, int cnt, int tid)
{
    const unsigned need = ((1u << cnt) - 1u) & ~w.mask;
    w.mask |= (1u << cnt) - 1u;
    if (need) {
        rows_issue<W, RT>(p, L, rr, w.base, need, wave_id(), tid & 63);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
}

// Next round's rows (window offsets [off, off+W), positions < S) are copied
// into their ring slots before the scan (direct loads, no registers) and
// waited for before the round's barrier.  Enabled per launch when the ring
// holds both rounds' positions (2W+2 <= R): the positions of this round and
// the next never share a slot.
template <int W, typename RT>
__device__ __forceinline__ unsigned prefetch_issue(const DevProblem &p, const LsLds &L, const RowRing<RT> &rr,
                                                   bool on, const RowWindow &w, int off, int S, int tid)
{
    if (!on) return 0;
    const int pos0 = w.base + off;
    const int avail = S - pos0 < W ? S - pos0 : W;
    if (avail <= 0) return 0;
    const unsigned want = ((1u << avail) - 1u) & ~(w.mask >> off);
    if (want) rows_issue<W, RT>(p, L, rr, pos0, want, wave_id(), tid & 63);
    return want;
}

// Before the round's barrier: the prefetched rows have landed.
__device__ __forceinline__ void prefetch_wait(RowWindow &w, unsigned want, int off)
{
    if (want) __builtin_amdgcn_s_waitcnt(0);
    w.mask |= want << off;
}

// -------------------------------------------------------------- swap ----

// Early stop of speculative scans: a wave that accepts a move raises the
// round key (round << 5 | 31 - wave) with an LDS max.  A wave whose key is
// exceeded knows that a lower wave of this round has a move, so its own
// result will be discarded, and stops scanning.  Keys of earlier rounds are
// smaller, and no wave reaches the next round before every wave has passed
// this round's barrier, so no reset is needed.
struct ScanStop {
    const unsigned *key;
    unsigned mine;
};
__device__ __forceinline__ unsigned round_key(unsigned round, int w) { return (round << 5) | (unsigned)(31 - w); }

// Fresh read of the round key: a volatile load through an LDS pointer (a
// volatile load through a generic pointer would become a flat load that
// waits on the memory counter).
__device__ __forceinline__ unsigned read_key(const unsigned *key)
{
    return *(volatile const __attribute__((address_space(3))) unsigned *)(key);
}

// After a round's barrier: the first wave (in position order) whose scan
// accepted a move and that move's j (js), or -1.  The highest key of the
// round is that wave's; a key of an earlier round means no wave moved.  Each
// wave counts its own evaluated moves when its position is at or before that
// wave's (all positions when none moved): the moves the sequential scan
// evaluates.
template <int W>
__device__ __forceinline__ int round_pick(const int *res, const unsigned *key, unsigned round, int lane, int &js)
{
    const int rv = lane < W ? res[lane] : -1;
    const unsigned k = __builtin_amdgcn_readfirstlane(read_key(key));
    const int wstar = (k >> 5) == round ? 31 - (int)(k & 31u) : -1;
    js = wstar < 0 ? -1 : __builtin_amdgcn_readlane(rv, wstar);
    return wstar;
}

// Customers at tour positions [1, x], x >= 1: ridx[x] counts the depots at
// positions [1, x-1].  Swaps and the relocate shift move each node's route
// index with it and keep the order of depots, so this stays exact while
// the search runs.  The scans count their evaluated moves with it: the
// candidates a sequential scan evaluates up to position j are the
// customers between its first candidate and j.
__device__ __forceinline__ int customers_upto(const LsLds &L, int x)
{
    return x - L.ridx[x] - (L.tour[x] == 0 ? 1 : 0);
}

// A trip of a scan covers U chunks of 64 consecutive candidates (U =
// kSwapUnroll or kRelocateUnroll<W>; the arrays are padded for the largest),
// one per lane.  Each lane reads its candidate's neighbours in the tour and
// edge arrays directly (past the tour's end the padding reads as depots and
// zero edges), and the gain and the mask of improving candidates are formed
// without branches; only a chunk with an improving candidate takes the
// feasibility checks.
//
// Row entries are summed in their own type before one conversion: fp32 rows
// add in the order the reference's expression does; 16-bit rows hold
// integers, whose sums below 2^24 are exact, so converting the integer sum
// gives the same float as adding the converted entries.
template <typename RT> struct RowSum;
template <> struct RowSum<float> {
    using T = float;
    static __device__ __forceinline__ float f(float x) { return x; }
};
template <> struct RowSum<unsigned short> {
    using T = unsigned;
    static __device__ __forceinline__ float f(unsigned x) { return (float)x; }
};

// Swap of position i with position j (localSearch.cpp:301-410).  n2 =
// tour[j], dm/d0 the edges (pn2, n2) and (n2, sn2), rA/rC = d(pn1, n2),
// d(sn1, n2), rBp/rBs = d(n1, pn2), d(n1, sn2).  The gain is the
// reference's expression: -(a1 + b1 + dm + d0) + (rA + rC + rBp + rBs), or
// -(a1 + d0) + (rA + rBs) for adjacent positions (j = i + 1).
//
// Loads and route lengths stay feasible (checked only for improving moves
// between two routes).
__device__ __forceinline__ bool swap_feasible(const DevProblem &p, const LsLds &L, int j, int r1, int dn1,
                                              float a1, float b1, int n2, float rA, float rC, float dm,
                                              float d0, float rBp, float rBs)
{
    const int r2 = L.ridx[j];
    if (r1 == r2) return true;
    const int dn2 = L.dem[n2];
    const int load1 = L.rl[r1] - dn1 + dn2, load2 = L.rl[r2] - dn2 + dn1;
    const float dist1 = (L.rd[r1] - (a1 + b1)) + (rA + rC);
    const float dist2 = (L.rd[r2] - (dm + d0)) + (rBp + rBs);
    return !(load1 > p.Q || load2 > p.Q || dist1 > p.D || dist2 > p.D);
}

// Position i against the current tour; returns the first accepted j or -1
// and in evc the moves the sequential scan evaluates up to it.
template <typename RT>
__device__ __forceinline__ int swap_scan(const DevProblem &p, const LsLds &L, int S, int i,
                                         const RT *rowA, const RT *rowB, const RT *rowC,
                                         int lane, const ScanStop &stop, int &evc)
{
    using V = typename RowSum<RT>::T;
    constexpr int U = kSwapUnroll;
    const int n1 = L.tour[i];
    const int sn1 = L.tour[i + 1], r1 = L.ridx[i];
    const float a1 = rget(rowA, n1), b1 = rget(rowB, sn1);
    const int dn1 = L.dem[n1];
    for (int jb = i + 1; jb < S; jb += U * kLanes) {
        int tp[U], t[U], tn[U];
        float dm[U], d0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = jb + u * kLanes + lane;
            tp[u] = L.tour[j - 1];
            t[u] = L.tour[j];
            tn[u] = L.tour[j + 1];
            dm[u] = L.dtw[j - 1];
            d0[u] = L.dtw[j];
        }
        V rA[U], rC[U], rBp[U], rBs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rA[u] = rowA[t[u]];
            rC[u] = rowC[t[u]];
            rBp[u] = rowB[tp[u]];
            rBs[u] = rowB[tn[u]];
        }
        // after the row reads, so its latency overlaps theirs
        const unsigned key = read_key(stop.key);
        bool imp[U], any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float g = -(a1 + b1 + dm[u] + d0[u]) + RowSum<RT>::f(rA[u] + rC[u] + rBp[u] + rBs[u]);
            if (u == 0) {
                // j = i + 1 (adjacent) is lane 0 of the first chunk of the first trip
                const float adj = -(a1 + d0[u]) + RowSum<RT>::f(rA[u] + rBs[u]);
                g = jb + lane == i + 1 ? adj : g;
            }
            // depots are no candidates; j >= S reads a padding depot
            imp[u] = (t[u] != 0) & (g < -kEps);
            any = any | imp[u];
        }
        if (__ballot(any)) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = jb + u * kLanes + lane;
                const bool ok = imp[u] && swap_feasible(p, L, j, r1, dn1, a1, b1, t[u], RowSum<RT>::f(rA[u]),
                                                        RowSum<RT>::f(rC[u]), dm[u], d0[u],
                                                        RowSum<RT>::f(rBp[u]), RowSum<RT>::f(rBs[u]));
                const u64 mk = __ballot(ok);
                if (mk) {
                    const int js = jb + u * kLanes + (__ffsll(mk) - 1);
                    evc = customers_upto(L, js) - customers_upto(L, i);
                    return js;
                }
            }
        }
        // stopped early: a lower wave has a move and this result is dropped
        if (key > stop.mine) break;
    }
    evc = customers_upto(L, S - 1) - customers_upto(L, i);
    return -1;
}

template <int W, typename RT>
__device__ __forceinline__ void ls_swap(const DevProblem &p, LsLds &L, const RowRing<RT> &rr, bool pf_on,
                                        int S, int tid, unsigned &round, long long &ev, int &moves)
{
    const int lane = tid & 63, w = wave_id();
    const int cap = kMaxMovesFactor * p.n;
    int done = 0, i0 = 1, parity = 0;
    RowWindow win = {0, 0u};
    while (i0 < S) {
        // rows of positions i0-1 .. i0+W (pred, self, succ of each scanned i)
        window_move(win, i0 - 1);
        rows_ensure<W, RT>(p, L, rr, win, min(W + 2, S - i0 + 1), tid);
        const unsigned pfw = prefetch_issue<W, RT>(p, L, rr, pf_on, win, W + 2, S, tid);
        const int i = i0 + w;
        int j = -1, evc = 0;
        const ScanStop stop = {reinterpret_cast<const unsigned *>(L.res + LsW<W>::res_first),
                               round_key(++round, w)};
        // uniform per wave (scalar branch; padding past S reads as depots)
        if (i < S && __builtin_amdgcn_readfirstlane(L.tour[i]) != 0)
            j = swap_scan(p, L, S, i, rr.at(i - 1), rr.at(i), rr.at(i + 1), lane, stop, evc);
        if (j >= 0 && lane == 0) atomicMax(reinterpret_cast<unsigned *>(L.res + LsW<W>::res_first), stop.mine);
        int *res = L.res + W * parity;
        parity ^= 1;
        if (lane == 0) res[w] = j;
        prefetch_wait(win, pfw, W + 2);
        __syncthreads();
        // the sequential scan evaluates the waves' positions in order and
        // stops at the first accepted move
        int js;
        const int wstar = round_pick<W>(res, stop.key, round, lane, js);
        if (lane == 0 && (wstar < 0 || w <= wstar)) ev += evc;
        if (wstar < 0) {
            // no barrier: the next round writes the other result buffer, and
            // rows it loads belong to positions no wave still reads
            i0 += W;
            continue;
        }
        const int is = i0 + wstar;
        if (tid == 0) {
            // every value read before any write and without branches, so the
            // reads go out in two batches (positions' entries, then rows,
            // demands and route tables); the branches below only select
            const RT *rowA = rr.at(is - 1), *rowB = rr.at(is), *rowC = rr.at(is + 1);
            const int n1 = L.tour[is], sn1 = L.tour[is + 1], r1 = L.ridx[is];
            const int n2 = L.tour[js], pn2 = L.tour[js - 1], sn2 = L.tour[js + 1], r2 = L.ridx[js];
            const float dm = L.dtw[js - 1], d0 = L.dtw[js];
            const float a1 = rget(rowA, n1), b1 = rget(rowB, sn1);
            // row values of the new edges: (pn1, n2), (n2, sn1), (pn2, n1),
            // (n1, sn2) and, for adjacent positions, (n2, n1) (symmetric
            // distances: the same values as the distance matrix)
            const float an2 = rget(rowA, n2), cn2 = rget(rowC, n2), bpn2 = rget(rowB, pn2);
            const float bsn2 = rget(rowB, sn2), bn2 = rget(rowB, n2);
            const int dn1 = L.dem[n1], dn2 = L.dem[n2], rl1 = L.rl[r1], rl2 = L.rl[r2];
            const float rd1 = L.rd[r1], rd2 = L.rd[r2];
            const bool adj = js == is + 1;
            int load1, load2;
            float dist1, dist2;
            if (r1 != r2) {
                load1 = rl1 - dn1 + dn2;
                load2 = rl2 - dn2 + dn1;
                dist1 = (rd1 - (a1 + b1)) + (an2 + cn2);
                dist2 = (rd2 - (dm + d0)) + (bpn2 + bsn2);
            } else {
                const float gain = adj ? -(a1 + d0) + (an2 + bsn2)
                                       : -(a1 + b1 + dm + d0) + (an2 + cn2 + bpn2 + bsn2);
                load1 = rl1;
                load2 = load1;
                dist1 = rd1 + gain;
                dist2 = dist1;
            }
            L.tour[is] = n2;
            L.tour[js] = n1;
            L.rl[r1] = load1; L.rl[r2] = load2;
            L.rd[r1] = dist1; L.rd[r2] = dist2;
            if (adj) {
                L.dtw[is - 1] = an2;   // (pn1, n2)
                L.dtw[is] = bn2;       // (n2, n1)
                L.dtw[js] = bsn2;      // (n1, sn2)
            } else {
                L.dtw[is - 1] = an2;   // (pn1, n2)
                L.dtw[is] = cn2;       // (n2, sn1)
                L.dtw[js - 1] = bpn2;  // (pn2, n1)
                L.dtw[js] = bsn2;      // (n1, sn2)
            }
        }
        __syncthreads();
        window_clear(win, is, is);
        window_clear(win, js, js);
        i0 = is;  // re-examine the position that moved (the reference's i--)
        if (++done >= cap) break;
    }
    if (tid == 0) moves += done;
}

// ---------------------------------------------------------- relocate ----

// Relocate of position i next to position j (neighbourSearch.cpp:280-293,
// move.cpp:32-51).  n2 = tour[j]; rm is the edge the insertion removes,
// (pn2, n2) before j or (n2, sn2) after it; add = d(n1, n2) + d(n1, x)
// with x = pn2 or sn2.
__device__ __forceinline__ float relocate_gain(float a1, float b1, float c1, float rm, float add)
{
    return -(a1 + b1 + rm) + (add + c1);
}

// Load and route lengths stay feasible (checked only for improving moves
// between two routes).
__device__ __forceinline__ bool relocate_feasible(const DevProblem &p, const LsLds &L, int j, int r1, int dn1,
                                                  float a1, float b1, float c1, float rm, float add)
{
    const int r2 = L.ridx[j];
    if (r1 == r2) return true;
    const float dist2 = ((L.rd[r2] + p.svc) - rm) + add;
    const float dist1 = ((L.rd[r1] - p.svc) - (a1 + b1)) + c1;
    return !(L.rl[r2] + dn1 > p.Q) && !(dist2 > p.D) && !(dist1 > p.D);
}

// Candidates are taken by edge: inserting n1 on edge q = (tour[q],
// tour[q+1]) is candidate j = q + 1 (before tour[j], rm = edge j - 1) when
// q < i and candidate j = q (after tour[j], rm = edge j) when q > i; edges
// i - 1 and i touch n1 itself (j = i).  j grows with q, so edges are taken
// in the sequential scan's order, and every candidate's gain has the same
// form: add = d(n1, tour[q]) + d(n1, tour[q+1]), rm = the edge's weight.
//
// Without a route-length limit a move of n1 is feasible exactly when its
// target route is n1's own or has room for dn1 (R <= 64 routes: one bit per
// route, from one ballot); the trip then tests each candidate's route bit
// with the other masks, and only a chunk with an improving feasible
// candidate leaves the common path.  Otherwise every route bit is set and
// the full check runs for improving candidates.
template <int U, typename RT>
__device__ __forceinline__ int relocate_scan(const DevProblem &p, const LsLds &L, int S, int R, int i,
                                             const RT *rowP, const RT *rowB, int lane,
                                             const ScanStop &stop, int &evc)
{
    using V = typename RowSum<RT>::T;
    const int n1 = L.tour[i];
    const int r1 = L.ridx[i], sn1 = L.tour[i + 1];
    const float a1 = L.dtw[i - 1], b1 = L.dtw[i], c1 = rget(rowP, sn1);
    const int dn1 = L.dem[n1];
    u64 room = ~0ull;
    if (!p.has_d && R <= kLanes) {
        const int rl = lane < R ? L.rl[lane] : 0;
        room = __ballot(lane < R && (lane == r1 || rl + dn1 <= p.Q));
    }
    for (int qb = 0; qb < S - 1; qb += U * kLanes) {
        int t0[U], t1[U], r2[U];
        float rm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = qb + u * kLanes + lane;
            t0[u] = L.tour[q];
            t1[u] = L.tour[q + 1];
            rm[u] = L.dtw[q];
            // route of the candidate position j (q + 1 before i, q after)
            r2[u] = L.ridx[q < i ? q + 1 : q];
        }
        V v0[U], v1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v0[u] = rowB[t0[u]];
            v1[u] = rowB[t1[u]];
        }
        // after the row reads, so its latency overlaps theirs
        const unsigned key = read_key(stop.key);
        float add[U];
        bool imp[U], any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = qb + u * kLanes + lane;
            // the node at j; q >= S - 1 reads the closing depot or padding
            const int tj = q < i ? t1[u] : t0[u];
            add[u] = RowSum<RT>::f(v0[u] + v1[u]);
            imp[u] = (tj != 0) & (tj != n1) & (relocate_gain(a1, b1, c1, rm[u], add[u]) < -kEps) &
                     (((room >> (r2[u] & 63)) & 1ull) != 0);
            any = any | imp[u];
        }
        if (__ballot(any)) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = qb + u * kLanes + lane;
                const int j = q < i ? q + 1 : q;
                const bool ok = imp[u] && relocate_feasible(p, L, j, r1, dn1, a1, b1, c1, rm[u], add[u]);
                const u64 mk = __ballot(ok);
                if (mk) {
                    const int qf = qb + u * kLanes + (__ffsll(mk) - 1);
                    const int js = qf < i ? qf + 1 : qf;
                    // customers in [1, js] other than n1 itself
                    evc = customers_upto(L, js) - (i <= js ? 1 : 0);
                    return js;
                }
            }
        }
        // stopped early: a lower wave has a move and this result is dropped
        if (key > stop.mine) break;
    }
    // every customer position in [1, S-2] but i
    evc = customers_upto(L, S - 2) - 1;
    return -1;
}

template <int W, typename RT>
__device__ __forceinline__ void ls_relocate(const DevProblem &p, LsLds &L, const RowRing<RT> &rr,
                                            bool pf_on, int S, int tid, unsigned &round, long long &ev,
                                            int &moves)
{
    const int lane = tid & 63, w = wave_id();
    const int cap = kMaxMovesFactor * p.n;
    // relocation never empties a route, so the route count stays
    const int R = __builtin_amdgcn_readfirstlane(L.res[LsW<W>::res_r]);
    int done = 0, i0 = 1, parity = 0;
    RowWindow win = {0, 0u};
    while (i0 < S - 1 && done < cap) {
        // rows of positions i0-1 .. i0+W-1: d(pred, .) gives c1, d(n1, .) the gains
        window_move(win, i0 - 1);
        rows_ensure<W, RT>(p, L, rr, win, min(W + 1, S - i0), tid);
        const unsigned pfw = prefetch_issue<W, RT>(p, L, rr, pf_on, win, W + 1, S - 1, tid);
        const int i = i0 + w;
        int j = -1, evc = 0;
        const ScanStop stop = {reinterpret_cast<const unsigned *>(L.res + LsW<W>::res_first),
                               round_key(++round, w)};
        // uniform per wave (scalar branches); the route must keep a customer
        bool scan = i < S - 1;
        if (scan) scan = __builtin_amdgcn_readfirstlane(L.tour[i]) != 0;
        if (scan) scan = __builtin_amdgcn_readfirstlane(L.rc[L.ridx[i]]) >= 2;
        if (scan)
            j = relocate_scan<kRelocateUnroll<W>>(p, L, S, R, i, rr.at(i - 1), rr.at(i), lane, stop, evc);
        if (j >= 0 && lane == 0) atomicMax(reinterpret_cast<unsigned *>(L.res + LsW<W>::res_first), stop.mine);
        int *res = L.res + W * parity;
        parity ^= 1;
        if (lane == 0) res[w] = j;
        prefetch_wait(win, pfw, W + 1);
        __syncthreads();
        // the sequential scan evaluates the waves' positions in order and
        // stops at the first accepted move
        int js;
        const int wstar = round_pick<W>(res, stop.key, round, lane, js);
        if (lane == 0 && (wstar < 0 || w <= wstar)) ev += evc;
        if (wstar < 0) {
            // no barrier: the next round writes the other result buffer, and
            // rows it loads belong to positions no wave still reads
            i0 += W;
            continue;
        }
        const int is = i0 + wstar;
        const int lo = is < js ? is : js, hi = is < js ? js : is;
        // Every value the move needs, read by every thread before any write
        // and without branches, so the reads go out in two batches: the
        // positions' entries, then the rows, the demand and the route
        // tables.  x is the insertion's other neighbour (pn2 before j, sn2
        // after it); the new edges are (pn1, sn1) = c1s and (n1, n2), (n1,
        // x) = e_n2, e_x.
        const int n1 = L.tour[is], sn1 = L.tour[is + 1], n2 = L.tour[js];
        const int xs = L.tour[is > js ? js - 1 : js + 1];
        const int r1 = L.ridx[is], r2 = L.ridx[js];
        const float a1 = L.dtw[is - 1], b1 = L.dtw[is], rm = L.dtw[is > js ? js - 1 : js];
        const RT *rowPs = rr.at(is - 1), *rowBs = rr.at(is);
        const float c1s = rget(rowPs, sn1), e_n2 = rget(rowBs, n2), e_x = rget(rowBs, xs);
        const int dn1 = L.dem[n1], rl1 = L.rl[r1], rl2 = L.rl[r2];
        const float rd1 = L.rd[r1], rd2 = L.rd[r2];
        const int rc1 = L.rc[r1], rc2 = L.rc[r2];
        if (tid == 0) {
            const float add = e_n2 + e_x;
            const float gain = -(a1 + b1 + rm) + (add + c1s);
            if (r1 != r2) {
                L.rl[r1] = rl1 - dn1;
                L.rl[r2] = rl2 + dn1;
                L.rd[r1] = ((rd1 - p.svc) - (a1 + b1)) + c1s;
                L.rd[r2] = ((rd2 + p.svc) - rm) + add;
                L.rc[r1] = (short)(rc1 - 1);
                L.rc[r2] = (short)(rc2 + 1);
            } else {
                L.rd[r1] = rd1 + gain;
            }
        }
        // shift, in place: positions [lo, hi] move by one towards is, n1
        // lands at js (in the route of the node that was at js); route
        // indices and the edge weights of [lo-1, hi] move with them (no
        // reload).  Chunks of NT positions are read into registers, then
        // written; they are taken in the direction of the move (upwards when
        // is < js, downwards otherwise), so a source position is read before
        // its own chunk is written.
        {
            constexpr int NT = LsW<W>::threads;
            const bool up = is < js;
            const int first = lo - 1, last = hi;
            const int nch = (last - first + NT) / NT;
            for (int c = 0; c < nch; ++c) {
                const int q = (up ? first + c * NT : last - (c + 1) * NT + 1) + tid;
                const bool in_t = q >= lo && q <= hi && q >= first, in_d = q >= first && q <= last;
                int v = 0, r = 0;
                float d = 0.0f;
                if (in_t) {
                    const int src = q == js ? js : up ? q + 1 : q - 1;
                    v = q == js ? n1 : L.tour[src];
                    r = L.ridx[src];
                }
                if (in_d) {
                    if (up) d = q == is - 1 ? c1s : q == js - 1 ? e_n2 : q == js ? e_x : L.dtw[q + 1];
                    else d = q == js - 1 ? e_x : q == js ? e_n2 : q == is ? c1s : L.dtw[q - 1];
                }
                __syncthreads();
                if (in_t) {
                    L.tour[q] = v;
                    L.ridx[q] = (short)r;
                }
                if (in_d) L.dtw[q] = d;
                __syncthreads();
            }
        }
        window_clear(win, lo, hi);
        i0 = is;
        ++done;
    }
    if (tid == 0) moves += done;
}

// ---------------------------------------------------------- one ant ----

// Local search of ant a of one instance by the W waves of the calling block,
// LDS at smem (local_search_lds_bytes).  The caller synchronises the block
// before reusing smem.
// R2: the instantiation with the random-order 2-opt (ACO_LS_2OPT_RANDOM);
// kept out of the others, whose code it would otherwise enlarge.
// Issue priority of wave w of W by quarters: the lowest quarter 3, the
// highest 0 (s_setprio takes an immediate).
__device__ __forceinline__ void scan_priority(int w, int W)
{
    const int q = __builtin_amdgcn_readfirstlane((w * 4) / W);
    if (q == 0) __builtin_amdgcn_s_setprio(3);
    else if (q == 1) __builtin_amdgcn_s_setprio(2);
    else if (q == 2) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

template <int W, typename RT, bool R2 = false>
__device__ void local_search_ant(const DevProblem &p, const DevState &s, int a, int flags, int ring,
                                 int prefetch, unsigned char *smem)
{
    constexpr int NT = LsW<W>::threads;
    const int tid = threadIdx.x;
    LsLds L = ls_bind(p.n, ring, (int)sizeof(RT), R2, smem);
    const RowRing<RT> rr = {reinterpret_cast<RT *>(L.rows), ring - 1, row_stride<RT>(p.n)};
    const bool pf_on = prefetch != 0;
    // uniform: loop bounds, row windows and masks stay in scalar registers
    const int S = __builtin_amdgcn_readfirstlane(s.sizes[a]);
    int *gt = s.tours + (size_t)a * p.T;
    for (int q = tid; q < S; q += NT) L.tour[q] = gt[q];
    // padding past the tour's end: depots and zero edges (never written
    // again; moves stay inside [1, S-2])
    for (int q = tid; q < kScanPad; q += NT) {
        L.tour[S + q] = 0;
        L.dtw[S - 1 + q] = 0.0f;
    }
    for (int v = tid; v < p.n; v += NT) {
        L.dlb[v] = 0;
        L.dem[v] = p.dem[v];
    }
    if (tid == 0) L.res[LsW<W>::res_first] = 0;
    __syncthreads();
    long long ev = 0;
    int moves = 0;
    unsigned round = 0;
    if (flags & kLs2Opt) {
        ls_edges<W>(p, L, 0, S - 2, tid);
        const int R = ls_routes<W>(p, L, S, tid, false);
        if (R <= NT / (2 * kOptGroup)) {
            // a group of 16 lanes per route: two positions at a time
            const int grp = tid / (2 * kOptGroup), gl = tid % (2 * kOptGroup);
            if (grp < R) two_opt_route_pair<false>(p, L, grp, L.zpos[grp], L.zpos[grp + 1] - 1, gl, ev, moves);
        } else {
            const int grp = tid / kOptGroup, gl = tid % kOptGroup;
            for (int r = grp; r < R; r += NT / kOptGroup)
                two_opt_route<false>(p, L, r, L.zpos[r], L.zpos[r + 1] - 1, gl, ev, moves);
        }
        __syncthreads();
    }
    if constexpr (R2) {
        // CPU variant's scan order: one Park-Miller stream per instance,
        // consumed ant after ant and route after route (one draw per route
        // position, size - 1 per ant) as sequential calls would; ant a
        // starts after the draws of ants 0..a-1, route r after those of the
        // routes before it (zpos[r] positions), reached by jump-ahead
        __shared__ long long red_off[W];
        long long off = 0;
        for (int b = tid; b < a; b += NT) off += s.sizes[b] - 1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
        if ((tid & 63) == 0) red_off[tid >> 6] = off;
        ls_edges<W>(p, L, 0, S - 2, tid);
        const int R = ls_routes<W>(p, L, S, tid, false);   // its barriers publish red_off
        off = 0;
        for (int q = 0; q < W; ++q) off += red_off[q];
        const int zant = pm_jump(s.ctrl[kCtrlLsSeed], (unsigned long long)off);
        if (R <= NT / (2 * kOptGroup)) {
            const int grp = tid / (2 * kOptGroup), gl = tid % (2 * kOptGroup);
            if (grp < R) {
                const int rbeg = L.zpos[grp], rend = L.zpos[grp + 1] - 1;
                if (gl == 0) route_permutation(L, rbeg, rend - rbeg + 1, pm_jump(zant, (unsigned long long)rbeg));
                group_sync();
                two_opt_route_pair<true>(p, L, grp, rbeg, rend, gl, ev, moves);
            }
        } else {
            const int grp = tid / kOptGroup, gl = tid % kOptGroup;
            for (int r = grp; r < R; r += NT / kOptGroup) {
                const int rbeg = L.zpos[r], rend = L.zpos[r + 1] - 1;
                if (gl == 0) route_permutation(L, rbeg, rend - rbeg + 1, pm_jump(zant, (unsigned long long)rbeg));
                group_sync();
                two_opt_route<true>(p, L, r, rbeg, rend, gl, ev, moves);
            }
        }
        __syncthreads();
    }
    // swap and relocate: wave w scans position i0 + w of a round, and the
    // lowest position with a move decides it; the waves above it mostly do
    // work that is discarded.  Lower waves issue first (timing only).
    if (flags & (kLsSwap | kLsRelocate)) scan_priority(tid >> 6, W);
    if (flags & kLsSwap) {
        ls_edges<W>(p, L, 0, S - 2, tid);
        __syncthreads();
        ls_routes<W>(p, L, S, tid, true);
        ls_swap<W, RT>(p, L, rr, pf_on, S, tid, round, ev, moves);
    }
    if (flags & kLsRelocate) {
        ls_edges<W>(p, L, 0, S - 2, tid);
        __syncthreads();
        ls_routes<W>(p, L, S, tid, true);
        ls_relocate<W, RT>(p, L, rr, pf_on, S, tid, round, ev, moves);
    }
    __builtin_amdgcn_s_setprio(0);
    ls_edges<W>(p, L, 0, S - 2, tid);
    __syncthreads();
    for (int q = tid; q < S; q += NT) gt[q] = L.tour[q];
    // tour length in position order (vrpHelper.cpp:244-253)
    if (tid == 0) {
        float len = 0.0f;
        for (int q = 0; q < S - 1; ++q) len += L.dtw[q];
        s.lens[a] = len;
    }
    __shared__ long long red_ev[W];
    __shared__ int red_mv[W];
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
        for (int q = 0; q < W; ++q) { e += red_ev[q]; mv += red_mv[q]; }
        s.ant_evals[a] = e;
        s.ant_moves[a] = mv;
    }
}

}  // namespace cvrp
