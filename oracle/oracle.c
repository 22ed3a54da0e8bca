/*
 * CPU oracle for the CVRP ant-colony hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference's per-iteration semantics (see
 * oracle.h).  Loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg only; it is the checker, never the product.
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off -fno-fast-math (oracle/Makefile).
 *
 * Parity pins (what fixes each function independently of this file):
 *   orc_ran01          Park-Miller known answer z_10000 = 1043618065 from
 *                      z_0 = 1 (Park & Miller, CACM 31(10), 1988); reference
 *                      build oracle/_ref (utilities.cpp ran01)
 *   orc_philox4x32_10  Random123 known-answer vectors (Salmon et al., SC'11)
 *   orc_sort2/nn_lists reference build oracle/_ref (compute_nn_lists)
 *   orc_distances      X-set .sol route costs (exact), CMT .sol totals
 *   orc_tree_*         brute-force comparison with a literal simulation of
 *                      the reduction loop on tied inputs
 *   orc_construct_ant  invariants (feasibility, each customer once), special
 *                      cases (one feasible candidate, alpha=beta=0 draws
 *                      argmax, k >= n-2 candidate row == full row)
 *   statistics         brute force argmin / top-k
 *   pheromone          closed form tau' = (1-rho) tau + sum_g (ras-g)/L_g
 *   local search       reference build oracle/_ref (LocalSearch 2-opt+swap),
 *                      brute force delta == cost(apply) - cost exactly
 *   check_solution     reference build oracle/_ref, X-set .sol files
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* =========================================================================
 * Random numbers
 * ========================================================================= */

/* Park-Miller minimal standard via Schrage; AM = 1.0f/IM rounds to 2^-31.
 * utilities.cpp:271-288, kernel.cl:42-52 */
float orc_ran01(int32_t *idum)
{
    const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    const float AM = 1.0f / (float)IM;
    int32_t k = (*idum) / IQ;
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    return AM * (float)(*idum);
}

/* SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014). */
uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Per-lane Park-Miller seed.  The reference seeds lanes from time(NULL)
 * (g_aco.cpp:715-719); this project seeds them explicitly so runs repeat:
 * seed_l = 1 + splitmix64(seed + golden*(1 + 64*ant + lane)) mod (2^31-2). */
int32_t orc_pm_seed(uint64_t seed, int ant, int lane)
{
    uint64_t z = orc_splitmix64(seed + 0x9E3779B97F4A7C15ull *
                                (uint64_t)(1 + (uint64_t)ant * ORC_LANES + (uint64_t)lane));
    return (int32_t)(1 + z % 2147483646ull);
}

/* Seed of the local search's scan-order stream (the reference draws it
 * from instance->rnd_seed, problem.cpp:213; here it is derived from the
 * colony seed). */
int32_t orc_ls_seed(uint64_t seed)
{
    return (int32_t)(1 + orc_splitmix64(seed ^ 0xC2B2AE3D27D4EB4Full) % 2147483646ull);
}

/* 16807^t * z mod (2^31 - 1): the Park-Miller state after t draws. */
int32_t orc_pm_jump(int32_t z, uint64_t t)
{
    const uint64_t M = 2147483647ull;
    uint64_t a = 16807, r = 1;
    while (t) {
        if (t & 1) r = r * a % M;
        a = a * a % M;
        t >>= 1;
    }
    return (int32_t)(r * (uint64_t)z % M);
}

/* Philox-4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One uniform in [0,1): word 0 of Philox(key=seed, ctr=(iter, ant, step, 0)),
 * top 24 bits scaled by 2^-24 (exact in float). */
float orc_philox_u01(uint64_t seed, uint32_t iter, uint32_t ant, uint32_t step)
{
    uint32_t ctr[4] = {iter, ant, step, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return (float)(o[0] >> 8) * (1.0f / 16777216.0f);
}

/* =========================================================================
 * Instance preprocessing
 * ========================================================================= */

/* mode 0 follows round_distance (vrpHelper.cpp:67-81): float coordinates,
 * float sqrt, no rounding.  mode 1 is VRPLIB nint (the X-set convention). */
void orc_distances(int n, const double *x, const double *y, int mode, float *dist)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            float d;
            if (mode == 0) {
                float xd = (float)x[i] - (float)x[j];
                float yd = (float)y[i] - (float)y[j];
                d = sqrtf(xd * xd + yd * yd);
            } else {
                double xd = x[i] - x[j], yd = y[i] - y[j];
                d = (float)(long)(sqrt(xd * xd + yd * yd) + 0.5);
            }
            dist[(size_t)i * n + j] = d;
        }
}

static void swap2(float *v, int *v2, int i, int j)
{
    float t = v[i]; v[i] = v[j]; v[j] = t;
    int t2 = v2[i]; v2[i] = v2[j]; v2[j] = t2;
}

/* Paired quicksort with middle pivot whose left recursion keeps the pivot
 * slot (utilities.cpp:246-267).  Its tie order is what compute_nn_lists
 * yields, so it is restated step for step. */
void orc_sort2(float *v, int *v2, int left, int right)
{
    if (left >= right) return;
    swap2(v, v2, left, (left + right) / 2);
    int last = left;
    for (int k = left + 1; k <= right; ++k)
        if (v[k] < v[left]) swap2(v, v2, ++last, k);
    swap2(v, v2, left, last);
    orc_sort2(v, v2, left, last);
    orc_sort2(v, v2, last + 1, right);
}

typedef struct { float d; int i; } dkey;
static int cmp_dkey(const void *a, const void *b)
{
    const dkey *x = (const dkey *)a, *y = (const dkey *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* compute_nn_lists (vrpHelper.cpp:188-236): per node, self and depot set to
 * INFINITY, sorted, first nn entries kept. */
void orc_nn_lists(int n, const float *dist, int nn, int tie, int *out)
{
    float *v = (float *)malloc(sizeof(float) * n);
    int *ix = (int *)malloc(sizeof(int) * n);
    dkey *kv = (dkey *)malloc(sizeof(dkey) * n);
    for (int node = 0; node < n; ++node) {
        for (int i = 0; i < n; ++i) { v[i] = dist[(size_t)node * n + i]; ix[i] = i; }
        v[node] = INFINITY;
        v[0] = INFINITY;
        if (tie == 0) {
            orc_sort2(v, ix, 0, n - 1);
        } else {
            for (int i = 0; i < n; ++i) { kv[i].d = v[i]; kv[i].i = i; }
            qsort(kv, n, sizeof(dkey), cmp_dkey);
            for (int i = 0; i < n; ++i) ix[i] = kv[i].i;
        }
        for (int i = 0; i < nn; ++i) out[(size_t)node * nn + i] = ix[i];
    }
    free(v); free(ix); free(kv);
}

/* eta^beta with eta = 1/(d+0.1f) (kernel.cl:678,788).  beta == 2 uses the
 * closed form y*y (a correctly rounded pow gives the same value); an
 * optional learned prior multiplies it. */
void orc_heuristic(int n, const float *dist, float beta, const float *prior, float *heur)
{
    for (size_t e = 0; e < (size_t)n * n; ++e) {
        float y = 1.0f / (dist[e] + 0.1f);
        float h = (beta == 2.0f) ? y * y : (beta == 1.0f) ? y : powf(y, beta);
        if (prior) h = h * prior[e];
        heur[e] = h;
    }
}

/* Priority order of the reference's argmax loop (kernel.cl:340-366): slot i
 * absorbs slot i+offset (strict '>', ties keep slot i) and an odd offset
 * first folds slot offset-1 into slot 0.  A tournament whose ties go to the
 * left slot has one total order on its leaves; rank[i] is node i's place in
 * it (0 = wins all ties).  argmax by (value, -rank) == the loop's result. */
void orc_tree_rank(int N, int *rank)
{
    int *head = (int *)malloc(sizeof(int) * N), *tail = (int *)malloc(sizeof(int) * N);
    int *next = (int *)malloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) { head[i] = tail[i] = i; next[i] = -1; }
#define APPEND(a, b) do { next[tail[a]] = head[b]; tail[a] = tail[b]; } while (0)
    int off = N;
    while (off > 0) {
        if ((off & 1) && off - 1 != 0) APPEND(0, off - 1);
        off /= 2;
        for (int i = 0; i < off; ++i) APPEND(i, i + off);
    }
#undef APPEND
    int r = 0;
    for (int x = head[0]; x >= 0; x = next[x]) rank[x] = r++;
    free(head); free(tail); free(next);
}

/* Literal restatement of independent_roulette's reduction (kernel.cl:340-366). */
int orc_tree_argmax(int N, const float *val)
{
    float *p = (float *)malloc(sizeof(float) * N);
    int *ix = (int *)malloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) { p[i] = val[i]; ix[i] = i; }
    int offset = N;
    while (offset > 0) {
        if (offset % 2 != 0) {
            float mine = p[0], other = p[offset - 1];
            if (other > mine) { p[0] = other; ix[0] = ix[offset - 1]; }
        }
        offset /= 2;
        for (int i = 0; i < offset; ++i) {
            float other = p[i + offset], mine = p[i];
            if (other > mine) { p[i] = other; ix[i] = ix[i + offset]; }
        }
    }
    int r = ix[0];
    free(p); free(ix);
    return r;
}

/* best_solution_phase_0 / phase_1 (kernel.cl:594-665): per-group halving
 * tree with strict '<' then a sequential scan over groups. */
int orc_minreduce_phase0(int n, int grp, const float *val)
{
    int ngrp = (n + grp - 1) / grp;
    float *s = (float *)malloc(sizeof(float) * grp);
    int *ix = (int *)malloc(sizeof(int) * grp);
    float bestv = 0.0f; int besti = 0;
    for (int g = 0; g < ngrp; ++g) {
        for (int l = 0; l < grp; ++l) {
            int gid = g * grp + l;
            s[l] = gid < n ? val[gid] : INFINITY;
            ix[l] = gid;
        }
        for (int off = grp / 2; off > 0; off /= 2)
            for (int l = 0; l < off; ++l) {
                float other = s[l + off], mine = s[l];
                if (other < mine) { s[l] = other; ix[l] = ix[l + off]; }
            }
        if (g == 0 || s[0] < bestv) { bestv = s[0]; besti = ix[0]; }
    }
    free(s); free(ix);
    return besti;
}

/* =========================================================================
 * Construction
 * ========================================================================= */

/* feasibility mask of kernel.cl:445-452 */
static inline int feasible(const orc_problem *p, const unsigned char *visited, int i,
                           int cur, int load, float rdist)
{
    if (visited[i]) return 0;
    if (load + p->dem[i] > p->Q) return 0;
    const int n = p->n;
    float t = rdist + (p->dist[(size_t)cur * n + i] + p->svc);
    t = t + p->dist[(size_t)i * n];
    return t <= p->D;
}

/* Full-row greedy pick used when the roulette has no positive mass: the
 * feasible unvisited node with the largest total_info, ties to the lowest
 * index.  The reference has no such rule (it appends the depot again, and
 * loops forever when that happens at the depot, kernel.cl:461-465).
 * Returns 0 when no node is feasible. */
static int greedy_pick(const orc_problem *p, const float *tinfo, const unsigned char *visited,
                       int cur, int load, float rdist)
{
    const int n = p->n;
    uint64_t best = 0;
    int w = 0;
    for (int i = 0; i < n; ++i) {
        if (!feasible(p, visited, i, cur, load, rdist)) continue;
        uint64_t key = (uint64_t)f2u(tinfo[(size_t)cur * n + i]) + 1u;
        if (key > best) { best = key; w = i; }
    }
    return w;
}

/* construct_solution for one ant (kernel.cl:382-493).
 * selection 0 (iroulette_full): lane l of the ant's 64 draws for nodes
 *   l, l+64, ... from its own Park-Miller stream; p_i = (r+0.001)*(ti*100)
 *   masked; argmax by the reduction tree (kernel.cl:325-373); zero mass
 *   sends the ant to the depot (kernel.cl:461-465) unless it is already
 *   there, where the greedy pick above applies.
 * selection 1 (roulette_knn, this project's rule): the k candidate nodes of
 *   the current row, masked, inclusive scan in the wavefront's DPP order
 *   (Hillis-Steele with offsets 1,2,4,8 inside each row of 16 lanes, then
 *   row 0's total into row 1 and row 2's into row 3, then lane 31's prefix
 *   into rows 2 and 3), r = u * prefix[k-1],
 *   first candidate with positive mass and prefix > r (else the last with
 *   positive mass: the tree's groupings round differently, so a masked
 *   lane's prefix can exceed its left neighbour's and must not be taken);
 *   zero mass falls back to the greedy pick over the full row, and to the
 *   depot when no node is feasible. */
int orc_construct_ant(const orc_problem *p, const float *tinfo, const float *tiknn,
                      const orc_draw_cfg *cfg, int ant, int32_t *pm,
                      int *tour, float *len, int *fallbacks)
{
    const int n = p->n, k = p->k;
    unsigned char *visited = (unsigned char *)calloc(n, 1);
    float *prob = (float *)malloc(sizeof(float) * n);
    int step = 0, vcnt = 1, cur = 0, load = 0, fb = 0;
    float rdist = 0.0f, tdist = 0.0f;
    visited[0] = 1;
    tour[0] = 0;
    while (vcnt < n) {
        int w;
        step++;
        if (cfg->selection == 0) {
            for (int lane = 0; lane < ORC_LANES; ++lane)
                for (int i = lane; i < n; i += ORC_LANES) {
                    float r = orc_ran01(&pm[lane]) + 0.001f;
                    float pp = tinfo[(size_t)cur * n + i] * 100.0f;
                    int c = feasible(p, visited, i, cur, load, rdist);
                    prob[i] = c ? r * pp : 0.0f;
                }
            w = orc_tree_argmax(n, prob);
            if (prob[w] == 0.0f) {
                if (cur == 0) { w = greedy_pick(p, tinfo, visited, cur, load, rdist); fb++; }
                else w = 0;
            }
        } else {
            float pv[ORC_LANES], x[ORC_LANES];
            int cj[ORC_LANES];
            for (int c = 0; c < ORC_LANES; ++c) {
                pv[c] = 0.0f; cj[c] = 0;
                if (c < k) {
                    int j = p->nnk[(size_t)cur * k + c];
                    cj[c] = j;
                    if (feasible(p, visited, j, cur, load, rdist))
                        pv[c] = tiknn[(size_t)cur * k + c];
                }
                x[c] = pv[c];
            }
            for (int d = 1; d < 16; d <<= 1)
                for (int c = ORC_LANES - 1; c >= 0; --c)
                    if ((c & 15) >= d) x[c] = x[c] + x[c - d];
            {
                const float t15 = x[15], t47 = x[47];
                for (int c = 16; c < 32; ++c) x[c] = x[c] + t15;
                for (int c = 48; c < 64; ++c) x[c] = x[c] + t47;
            }
            {
                const float t31 = x[31];
                for (int c = 32; c < 64; ++c) x[c] = x[c] + t31;
            }
            float total = x[k - 1];
            if (total > 0.0f) {
                float u = cfg->rng == 1 ? orc_philox_u01(cfg->seed, cfg->iteration, (uint32_t)ant,
                                                         (uint32_t)step)
                                        : orc_ran01(&pm[0]);
                float r = u * total;
                int wc = -1;
                for (int c = 0; c < k; ++c) if (pv[c] > 0.0f && x[c] > r) { wc = c; break; }
                if (wc < 0)
                    for (int c = k - 1; c >= 0; --c) if (pv[c] > 0.0f) { wc = c; break; }
                w = cj[wc];
            } else {
                w = greedy_pick(p, tinfo, visited, cur, load, rdist);
                fb++;
            }
        }
        float dcw = p->dist[(size_t)cur * n + w];
        if (w == 0) { load = 0; rdist = 0.0f; }
        else { load += p->dem[w]; rdist = rdist + (dcw + p->svc); vcnt++; }
        visited[w] = 1;
        tour[step] = w;
        tdist = tdist + dcw;
        cur = w;
    }
    tdist = tdist + p->dist[(size_t)tour[step] * n];
    step++;
    tour[step] = 0;
    free(visited); free(prob);
    *len = tdist;
    if (fallbacks) *fallbacks += fb;
    return step + 1;
}

/* =========================================================================
 * Statistics (kernel.cl:594-665, 546-586, 717-741)
 * ========================================================================= */

int orc_iteration_best(int m, const float *lens, float bsf_len, int *flag)
{
    int b = orc_minreduce_phase0(m, ORC_LANES, lens);
    *flag = (lens[b] - bsf_len < -ORC_EPSILON) ? 1 : 0;
    return b;
}

void orc_elites(int m, int ras, const float *lens, int *elite)
{
    float *h = (float *)malloc(sizeof(float) * m);
    memcpy(h, lens, sizeof(float) * m);
    elite[0] = m;
    for (int i = 0; i < ras - 1; ++i) {
        float b = h[0];
        int target = 0;
        for (int k = 0; k < m; ++k)
            if (h[k] < b) { b = h[k]; target = k; }
        h[target] = INFINITY;
        elite[i + 1] = target;
    }
    free(h);
}

/* =========================================================================
 * Pheromone (kernel.cl:672-692, 747-789)
 * ========================================================================= */

static inline float total_info(float tau, float alpha, float heur)
{
    return (alpha == 1.0f ? tau : powf(tau, alpha)) * heur;
}

void orc_pheromone_init(int n, float tau0, const float *heur, float alpha, float *tau, float *tinfo)
{
    for (size_t e = 0; e < (size_t)n * n; ++e) {
        tau[e] = tau0;
        tinfo[e] = total_info(tau0, alpha, heur[e]);
    }
}

void orc_pheromone_update(int n, float rho, float alpha, int ras,
                          const int *const *tours, const int *sizes, const float *lens,
                          const float *heur, float *tau, float *tinfo)
{
    const float evap = 1.0f - rho;
    for (size_t e = 0; e < (size_t)n * n; ++e) tau[e] = evap * tau[e];
    for (int g = 0; g < ras; ++g) {
        float d_tau = (float)(ras - g) / lens[g];
        const int *t = tours[g];
        for (int i = 0; i < sizes[g] - 1; ++i) {
            size_t e = (size_t)t[i] * n + t[i + 1];
            tau[e] = tau[e] + d_tau;
        }
    }
    for (size_t e = 0; e < (size_t)n * n; ++e) tinfo[e] = total_info(tau[e], alpha, heur[e]);
}

void orc_gather_knn(int n, int k, const int *nnk, const float *tinfo, float *tiknn)
{
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < k; ++c)
            tiknn[(size_t)i * k + c] = tinfo[(size_t)i * n + nnk[(size_t)i * k + c]];
}

/* =========================================================================
 * Local search
 * ========================================================================= */

typedef struct {
    const orc_problem *p;
    int *tour;
    unsigned char *dlb, *inroute;
    int *pos;
    float *dtw;
    long long evals;
} two_opt_ctx;

static inline float D_(const orc_problem *p, int a, int b) { return p->dist[(size_t)a * p->n + b]; }

/* two_opt_single_route (kernel.cl:126-277; CPU order variant
 * localSearch.cpp:180-294).  perm == NULL scans positions in sequence. */
static int two_opt_route(two_opt_ctx *c, int rbeg, int rend, const int *perm)
{
    const orc_problem *p = c->p;
    int *tour = c->tour, *pos = c->pos;
    unsigned char *dlb = c->dlb;
    float *dtw = c->dtw;
    const int num = rend - rbeg + 1, nn = p->nn_ls;
    int moves = 0, imp = 1;
    while (imp) {
        imp = 0;
        for (int l = 0; l < num; ++l) {
            int pos1 = rbeg + (perm ? perm[l] : l);
            int n1 = tour[pos1];
            if (dlb[n1]) continue;
            int h1 = 0, h2 = 0, h3 = 0, h4 = 0, found = 0;
            int s1 = pos1 == rend ? tour[rbeg] : tour[pos1 + 1];
            float radius = D_(p, n1, s1);
            for (int h = 0; h < nn; ++h) {
                int n2 = p->nnls[(size_t)n1 * nn + h];
                if (!c->inroute[n2]) continue;
                if (radius - D_(p, n1, n2) > ORC_EPSILON) {
                    c->evals++;
                    int pos2 = pos[n2];
                    int s2 = pos2 == rend ? tour[rbeg] : tour[pos2 + 1];
                    float gain = -radius + D_(p, n1, n2) + D_(p, s1, s2) - dtw[pos2];
                    if (gain < -ORC_EPSILON) { h1 = n1; h2 = s1; h3 = n2; h4 = s2; found = 1; break; }
                } else break;
            }
            if (!found) {
                int p1 = pos1 == rbeg ? tour[rend] : tour[pos1 - 1];
                radius = D_(p, p1, n1);
                for (int h = 0; h < nn; ++h) {
                    int n2 = p->nnls[(size_t)n1 * nn + h];
                    if (!c->inroute[n2]) continue;
                    if (radius - D_(p, n1, n2) > ORC_EPSILON) {
                        int pos2 = pos[n2];
                        int p2 = pos2 == rbeg ? tour[rend] : tour[pos2 - 1];
                        if (p2 == n1 || p1 == n2) continue;
                        c->evals++;
                        float gain = -radius + D_(p, n1, n2) + D_(p, p1, p2) - dtw[pos2 - 1];
                        if (gain < -ORC_EPSILON) { h1 = p1; h2 = n1; h3 = p2; h4 = n2; found = 1; break; }
                    } else break;
                }
            }
            if (!found) { dlb[n1] = 1; continue; }
            imp = 1;
            dlb[h1] = 0; dlb[h2] = 0; dlb[h3] = 0; dlb[h4] = 0;
            if (pos[h3] < pos[h1]) {
                int t = h1; h1 = h3; h3 = t;
                t = h2; h2 = h4; h4 = t;
            }
            int i = pos[h1], j = pos[h3];
            dtw[i] = D_(p, h1, h3);
            dtw[j] = D_(p, h2, h4);
            i++; j--;
            while (i < j) { float t = dtw[i]; dtw[i] = dtw[j]; dtw[j] = t; i++; j--; }
            i = pos[h2]; j = pos[h3];
            while (i < j) {
                int a = tour[i], b = tour[j];
                tour[i] = b; tour[j] = a;
                pos[a] = j; pos[b] = i;
                i++; j--;
            }
            if (++moves >= ORC_MAX_MOVES_PER_ROUTE) return moves;
        }
    }
    return moves;
}

/* generate_random_permutation (localSearch.cpp:97-119) */
static void random_permutation(int n, int32_t *seed, int *r)
{
    for (int i = 0; i < n; ++i) r[i] = i;
    int assigned = 0;
    for (int i = 0; i < n; ++i) {
        float rnd = orc_ran01(seed);
        int node = (int)(rnd * (float)(n - assigned));
        if (i + node >= n) node = n - 1 - i;   /* the reference asserts here */
        int help = r[i]; r[i] = r[i + node]; r[i + node] = help;
        assigned++;
    }
}

/* two_opt_solution (kernel.cl:279-319; localSearch.cpp:127-169) */
int orc_two_opt(const orc_problem *p, int *tour, int size, int random_order,
                int32_t *seed, long long *evals)
{
    const int n = p->n;
    two_opt_ctx c;
    c.p = p; c.tour = tour; c.evals = 0;
    c.dlb = (unsigned char *)calloc(n, 1);
    c.inroute = (unsigned char *)calloc(n, 1);
    c.pos = (int *)malloc(sizeof(int) * n);
    c.dtw = (float *)malloc(sizeof(float) * size);
    int *perm = random_order ? (int *)malloc(sizeof(int) * size) : NULL;
    for (int j = 0; j < size; ++j) c.pos[tour[j]] = j;
    for (int i = 0; i < size - 1; ++i) c.dtw[i] = D_(p, tour[i], tour[i + 1]);
    int rbeg = 0, moves = 0, route = 0;
    for (int i = 1; i < size; ++i) {
        if (tour[i] == 0) {
            c.pos[0] = rbeg;
            /* The depot is n1 at l = 0 of every route and its don't-look bit
             * is shared.  A route's last pass finds no move, so it leaves the
             * bit set: route 0 starts with it clear, every later route with
             * it set.  Stating that here keeps routes independent, and it
             * equals the reference whenever the move cap is not reached. */
            c.dlb[0] = route > 0;
            route++;
            if (perm) random_permutation(i - 1 - rbeg + 1, seed, perm);
            moves += two_opt_route(&c, rbeg, i - 1, perm);
            for (int j = 0; j < n; ++j) c.inroute[j] = 0;
            rbeg = i;
        } else {
            c.inroute[tour[i]] = 1;
        }
    }
    if (evals) *evals += c.evals;
    free(c.dlb); free(c.inroute); free(c.pos); free(c.dtw); free(perm);
    return moves;
}

/* per-route load and distance incl. service (localSearch.cpp:319-337) */
static int route_tables(const orc_problem *p, const int *tour, int size,
                        int *rl, float *rd, int *rc, int *ridx)
{
    int k = 0, beg = 0, load = 0;
    float dist = 0.0f;
    ridx[0] = 0;
    for (int i = 1; i < size; ++i) {
        load += p->dem[tour[i]];
        dist += D_(p, tour[i - 1], tour[i]);
        ridx[i] = k;
        if (tour[i] == 0) {
            rl[k] = load;
            rd[k] = dist + p->svc * (float)(i - beg - 1);
            if (rc) rc[k] = i - beg - 1;
            k++; load = 0; dist = 0.0f; beg = i;
        }
    }
    return k;
}

/* LocalSearch::swap (localSearch.cpp:301-410): first improvement over
 * position pairs i < j, '-EPSILON' acceptance, load and distance limits on
 * inter-route swaps, re-examine position i after a move. */
int orc_swap(const orc_problem *p, int *tour, int size, long long *evals)
{
    const int n = p->n;
    int *rl = (int *)malloc(sizeof(int) * n);
    float *rd = (float *)malloc(sizeof(float) * n);
    int *ridx = (int *)malloc(sizeof(int) * size);
    route_tables(p, tour, size, rl, rd, NULL, ridx);
    const int cap = ORC_MAX_MOVES_FACTOR * n;
    int moves = 0, p1 = 0;
    long long ev = 0;
    for (int i = 1; i < size; ++i) {
        int n1 = tour[i];
        if (n1 == 0) { p1++; continue; }
        int pn1 = tour[i - 1], sn1 = tour[i + 1];
        int p2 = p1;
        for (int j = i + 1; j < size; ++j) {
            int n2 = tour[j];
            if (n2 == 0) { p2++; continue; }
            int pn2 = tour[j - 1], sn2 = tour[j + 1];
            float gain;
            ev++;
            if (j == i + 1)
                gain = -(D_(p, pn1, n1) + D_(p, n2, sn2)) + (D_(p, pn1, n2) + D_(p, n1, sn2));
            else
                gain = -(D_(p, pn1, n1) + D_(p, n1, sn1) + D_(p, pn2, n2) + D_(p, n2, sn2))
                       + (D_(p, pn1, n2) + D_(p, n2, sn1) + D_(p, pn2, n1) + D_(p, n1, sn2));
            if (gain < -ORC_EPSILON) {
                int load1, load2;
                float dist1, dist2;
                if (p1 != p2) {
                    load1 = rl[p1] - p->dem[n1] + p->dem[n2];
                    load2 = rl[p2] - p->dem[n2] + p->dem[n1];
                    dist1 = rd[p1] - (D_(p, pn1, n1) + D_(p, n1, sn1)) + (D_(p, pn1, n2) + D_(p, n2, sn1));
                    dist2 = rd[p2] - (D_(p, pn2, n2) + D_(p, n2, sn2)) + (D_(p, pn2, n1) + D_(p, n1, sn2));
                    if (load1 > p->Q || load2 > p->Q || dist1 > p->D || dist2 > p->D) continue;
                } else {
                    load1 = rl[p1]; load2 = load1;
                    dist1 = rd[p1] + gain; dist2 = dist1;
                }
                tour[i] = n2; tour[j] = n1;
                rl[p1] = load1; rl[p2] = load2;
                rd[p1] = dist1; rd[p2] = dist2;
                i--;
                moves++;
                break;
            }
        }
        if (moves >= cap) break;
    }
    if (evals) *evals += ev;
    free(rl); free(rd); free(ridx);
    return moves;
}

/* Relocate: move customer at position i before (i > j) or after (i < j) the
 * customer at position j (InsertionMove::apply, move.cpp:32-51), gain
 * formulas of neighbourSearch.cpp:280-293.  Policy (this project's): first
 * improvement over j = 1..size-2 in order for each i in order; a route
 * keeps at least one customer; inter-route moves need load2 <= Q and both
 * new route distances <= D; position i is re-examined after a move. */
int orc_relocate(const orc_problem *p, int *tour, int size, long long *evals)
{
    const int n = p->n;
    int *rl = (int *)malloc(sizeof(int) * n), *rc = (int *)malloc(sizeof(int) * n);
    float *rd = (float *)malloc(sizeof(float) * n);
    int *ridx = (int *)malloc(sizeof(int) * size);
    route_tables(p, tour, size, rl, rd, rc, ridx);
    const int cap = ORC_MAX_MOVES_FACTOR * n;
    int moves = 0, i = 1;
    long long ev = 0;
    while (i < size - 1 && moves < cap) {
        int n1 = tour[i];
        if (n1 == 0 || rc[ridx[i]] < 2) { i++; continue; }
        int r1 = ridx[i], pn1 = tour[i - 1], sn1 = tour[i + 1];
        float a = D_(p, pn1, n1), b = D_(p, n1, sn1), c = D_(p, pn1, sn1);
        int found = 0;
        for (int j = 1; j < size - 1; ++j) {
            if (j == i) continue;
            int n2 = tour[j];
            if (n2 == 0) continue;
            int pn2 = tour[j - 1], sn2 = tour[j + 1], r2 = ridx[j];
            float gain, ins_rm, ins_add;
            ev++;
            if (i > j) {
                ins_rm = D_(p, pn2, n2);
                ins_add = D_(p, n1, n2) + D_(p, pn2, n1);
            } else {
                ins_rm = D_(p, n2, sn2);
                ins_add = D_(p, n2, n1) + D_(p, n1, sn2);
            }
            gain = -(a + b + ins_rm) + (ins_add + c);
            if (!(gain < -ORC_EPSILON)) continue;
            if (r1 != r2) {
                int load2 = rl[r2] + p->dem[n1];
                if (load2 > p->Q) continue;
                float dist2 = ((rd[r2] + p->svc) - ins_rm) + ins_add;
                float dist1 = ((rd[r1] - p->svc) - (a + b)) + c;
                if (dist2 > p->D || dist1 > p->D) continue;
                rl[r1] -= p->dem[n1]; rl[r2] = load2;
                rd[r1] = dist1; rd[r2] = dist2;
                rc[r1]--; rc[r2]++;
            } else {
                rd[r1] = rd[r1] + gain;
            }
            int tmp = tour[i];
            if (i < j) { for (int x = i; x < j; ++x) tour[x] = tour[x + 1]; }
            else { for (int x = i; x > j; --x) tour[x] = tour[x - 1]; }
            tour[j] = tmp;
            int rr = 0;
            for (int x = 1; x < size; ++x) { ridx[x] = rr; if (tour[x] == 0) rr++; }
            found = 1;
            moves++;
            break;
        }
        if (!found) i++;
    }
    if (evals) *evals += ev;
    free(rl); free(rc); free(rd); free(ridx);
    return moves;
}

float orc_local_search(const orc_problem *p, int flags, int *tour, int size, long long *evals)
{
    return orc_local_search_seeded(p, flags, tour, size, evals, NULL);
}

/* flags 8: 2-opt in the CPU variant's random scan order, drawing from *seed
 * (one draw per route position, localSearch.cpp:97-119,127-169) */
float orc_local_search_seeded(const orc_problem *p, int flags, int *tour, int size, long long *evals,
                              int32_t *seed)
{
    if (flags & 1) orc_two_opt(p, tour, size, 0, NULL, evals);
    if (flags & 8) orc_two_opt(p, tour, size, 1, seed, evals);
    if (flags & 2) orc_swap(p, tour, size, evals);
    if (flags & 4) orc_relocate(p, tour, size, evals);
    return orc_tour_length(p->n, p->dist, tour, size);
}

/* =========================================================================
 * Solution checks (vrpHelper.cpp:244-267, problem.cpp:226-319)
 * ========================================================================= */

float orc_tour_length(int n, const float *dist, const int *tour, int size)
{
    float len = 0.0f;
    for (int i = 0; i < size - 1; ++i) len += dist[(size_t)tour[i] * n + tour[i + 1]];
    return len;
}

static int check_route(const orc_problem *p, const int *tour, int rbeg, int rend)
{
    if (tour[rbeg] != 0 || tour[rend] != 0) return 0;
    if (rend - rbeg < 2) return 0;
    int load = 0;
    for (int i = rbeg + 1; i < rend; ++i) load += p->dem[tour[i]];
    if (load > p->Q) return 0;
    float d = orc_tour_length(p->n, p->dist, tour + rbeg, rend - rbeg + 1);
    d += p->svc * (float)(rend - rbeg - 1);
    return !(d > p->D);
}

int orc_check_solution(const orc_problem *p, const int *tour, int size)
{
    const int n = p->n;
    int *used = (int *)calloc(n, sizeof(int));
    int ok = 1, rbeg = 0;
    used[0] = 1;
    for (int i = 1; i < size && ok; ++i) {
        if (tour[i] < 0 || tour[i] >= n) { ok = 0; break; }
        if (tour[i] != 0) {
            if (used[tour[i]]) ok = 0;
            else used[tour[i]] = 1;
        }
        if (ok && tour[i] == 0) {
            if (!check_route(p, tour, rbeg, i)) ok = 0;
            rbeg = i;
        }
    }
    for (int i = 0; i < n && ok; ++i) if (!used[i]) ok = 0;
    free(used);
    return ok;
}

/* Nearest feasible customer, ties to the lowest index; back to the depot
 * when none fits.  No reference counterpart (tau0 = 1/(n*L_nn)). */
int orc_nn_tour(const orc_problem *p, int *tour, float *len)
{
    const int n = p->n;
    unsigned char *visited = (unsigned char *)calloc(n, 1);
    int cur = 0, load = 0, vcnt = 1, s = 0;
    float rdist = 0.0f, L = 0.0f;
    visited[0] = 1; tour[0] = 0;
    while (vcnt < n) {
        int best = -1;
        float bd = INFINITY;
        for (int i = 1; i < n; ++i)
            if (feasible(p, visited, i, cur, load, rdist) && D_(p, cur, i) < bd) { bd = D_(p, cur, i); best = i; }
        int w;
        if (best < 0) {
            if (cur == 0) { free(visited); return -1; }
            w = 0; load = 0; rdist = 0.0f;
        } else {
            w = best; load += p->dem[w]; rdist = rdist + (D_(p, cur, w) + p->svc); vcnt++;
        }
        visited[w] = 1;
        tour[++s] = w;
        L = L + D_(p, cur, w);
        cur = w;
    }
    L = L + D_(p, cur, 0);
    tour[++s] = 0;
    free(visited);
    *len = L;
    return s + 1;
}

/* =========================================================================
 * Whole colony loop (g_aco.cpp:63-134, 233-243, 407-417)
 * ========================================================================= */

struct orc_colony {
    orc_problem p;
    orc_params prm;
    int n, m, T;
    float *tau, *tinfo, *tiknn;
    int *tours, *sizes;
    float *lens;
    int32_t *pm;
    int iteration;
    orc_bsf_record *bsf;
    int nbsf, bsf_cap;
    long long ls_evals;
    long long fallbacks;
    int32_t ls_seed;        /* scan-order stream of the random-order 2-opt */
    double t0;
};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

orc_colony *orc_colony_create(const orc_problem *p, const orc_params *prm)
{
    orc_colony *c = (orc_colony *)calloc(1, sizeof(orc_colony));
    c->p = *p; c->prm = *prm;
    c->n = p->n; c->m = prm->m; c->T = 2 * p->n;
    size_t nn = (size_t)p->n * p->n;
    c->tau = (float *)malloc(sizeof(float) * nn);
    c->tinfo = (float *)malloc(sizeof(float) * nn);
    c->tiknn = (float *)malloc(sizeof(float) * (size_t)p->n * (p->k > 0 ? p->k : 1));
    c->tours = (int *)calloc((size_t)(c->m + 1) * c->T, sizeof(int));
    c->sizes = (int *)calloc(c->m + 1, sizeof(int));
    c->lens = (float *)calloc(c->m + 1, sizeof(float));
    c->pm = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->m * ORC_LANES);
    for (int a = 0; a < c->m; ++a)
        for (int l = 0; l < ORC_LANES; ++l) c->pm[(size_t)a * ORC_LANES + l] = orc_pm_seed(prm->seed, a, l);
    c->bsf_cap = 4096;
    c->ls_seed = orc_ls_seed(prm->seed);
    c->bsf = (orc_bsf_record *)calloc(c->bsf_cap, sizeof(orc_bsf_record));
    c->lens[c->m] = INFINITY;
    return c;
}

void orc_colony_destroy(orc_colony *c)
{
    if (!c) return;
    free(c->tau); free(c->tinfo); free(c->tiknn); free(c->tours); free(c->sizes);
    free(c->lens); free(c->pm); free(c->bsf); free(c);
}

static void colony_gather(orc_colony *c)
{
    if (c->p.k > 0) orc_gather_knn(c->n, c->p.k, c->p.nnk, c->tinfo, c->tiknn);
}

static void colony_tau_init(orc_colony *c, float tau0)
{
    orc_pheromone_init(c->n, tau0, c->p.heur, c->prm.alpha, c->tau, c->tinfo);
    colony_gather(c);
}

void orc_colony_construct(orc_colony *c)
{
    orc_draw_cfg cfg = {c->prm.selection, c->prm.rng, c->prm.seed, (uint32_t)c->iteration};
    int nthreads = c->prm.threads > 0 ? c->prm.threads : 1;
    (void)nthreads;
    long long fbs = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : fbs)
#endif
    for (int a = 0; a < c->m; ++a) {
        float len;
        int fb = 0;
        int sz = orc_construct_ant(&c->p, c->tinfo, c->tiknn, &cfg, a, c->pm + (size_t)a * ORC_LANES,
                                   c->tours + (size_t)a * c->T, &len, &fb);
        c->sizes[a] = sz;
        c->lens[a] = len;
        fbs += fb;
    }
    c->fallbacks += fbs;
}

void orc_colony_local_search(orc_colony *c)
{
    if (!c->prm.ls_flags) return;
    if (c->prm.ls_flags & 8) {
        /* one scan-order stream, consumed ant after ant as sequential calls
         * of the CPU local search would */
        long long ev = 0;
        for (int a = 0; a < c->m; ++a)
            c->lens[a] = orc_local_search_seeded(&c->p, c->prm.ls_flags, c->tours + (size_t)a * c->T,
                                                 c->sizes[a], &ev, &c->ls_seed);
        c->ls_evals += ev;
        return;
    }
    int nthreads = c->prm.threads > 0 ? c->prm.threads : 1;
    long long ev = 0;
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : ev)
#endif
    for (int a = 0; a < c->m; ++a) {
        long long e = 0;
        c->lens[a] = orc_local_search(&c->p, c->prm.ls_flags, c->tours + (size_t)a * c->T,
                                      c->sizes[a], &e);
        ev += e;
    }
    c->ls_evals += ev;
}

void orc_colony_statistics(orc_colony *c)
{
    int flag;
    int b = orc_iteration_best(c->m, c->lens, c->lens[c->m], &flag);
    if (flag) {
        memcpy(c->tours + (size_t)c->m * c->T, c->tours + (size_t)b * c->T, sizeof(int) * c->sizes[b]);
        c->sizes[c->m] = c->sizes[b];
        c->lens[c->m] = c->lens[b];
        if (c->nbsf < c->bsf_cap) {
            c->bsf[c->nbsf].iter = c->iteration;
            c->bsf[c->nbsf].length = c->lens[b];
            c->bsf[c->nbsf].time = (float)(now_s() - c->t0);
            c->nbsf++;
        }
    }
}

void orc_colony_pheromone_update(orc_colony *c)
{
    int ras = c->prm.ras;
    int *elite = (int *)malloc(sizeof(int) * ras);
    const int **tours = (const int **)malloc(sizeof(int *) * ras);
    int *sizes = (int *)malloc(sizeof(int) * ras);
    float *lens = (float *)malloc(sizeof(float) * ras);
    orc_elites(c->m, ras, c->lens, elite);
    for (int g = 0; g < ras; ++g) {
        tours[g] = c->tours + (size_t)elite[g] * c->T;
        sizes[g] = c->sizes[elite[g]];
        lens[g] = c->lens[elite[g]];
    }
    orc_pheromone_update(c->n, c->prm.rho, c->prm.alpha, ras, tours, sizes, lens, c->p.heur,
                         c->tau, c->tinfo);
    colony_gather(c);
    free(elite); free(tours); free(sizes); free(lens);
}

void orc_colony_init(orc_colony *c)
{
    c->t0 = now_s();
    c->iteration = 0;
    c->lens[c->m] = INFINITY;
    c->nbsf = 0;
    if (c->prm.tau0_rule == 0) {
        colony_tau_init(c, 0.5f);
        orc_colony_construct(c);
        orc_colony_statistics(c);
        float tau0 = 1.0f / (c->prm.rho * c->lens[c->m]);
        colony_tau_init(c, tau0);
        c->iteration = 1;
    } else {
        int *t = (int *)malloc(sizeof(int) * 2 * c->n);
        float L = 0.0f;
        orc_nn_tour(&c->p, t, &L);
        free(t);
        colony_tau_init(c, 1.0f / ((float)c->n * L));
        c->iteration = 0;
    }
}

void orc_colony_iteration(orc_colony *c)
{
    orc_colony_construct(c);
    orc_colony_local_search(c);
    orc_colony_statistics(c);
    orc_colony_pheromone_update(c);
    c->iteration++;
}

/* The caller's inst->iteration++ after an iteration (main.cpp:80-88), for
 * replays that call the four phases one by one (orc_colony_iteration does
 * it itself). */
void orc_colony_next_iteration(orc_colony *c) { c->iteration++; }

int orc_colony_iter(const orc_colony *c) { return c->iteration; }

float orc_colony_best(const orc_colony *c, int *tour, int *size)
{
    if (tour) memcpy(tour, c->tours + (size_t)c->m * c->T, sizeof(int) * c->sizes[c->m]);
    if (size) *size = c->sizes[c->m];
    return c->lens[c->m];
}

int orc_colony_ants(const orc_colony *c, int *tours, int stride, int *sizes, float *lens)
{
    for (int a = 0; a <= c->m; ++a) {
        if (tours) memcpy(tours + (size_t)a * stride, c->tours + (size_t)a * c->T, sizeof(int) * c->sizes[a]);
        if (sizes) sizes[a] = c->sizes[a];
        if (lens) lens[a] = c->lens[a];
    }
    return c->m;
}

int orc_colony_bsf(const orc_colony *c, orc_bsf_record *rec, int cap)
{
    int k = c->nbsf < cap ? c->nbsf : cap;
    if (rec) memcpy(rec, c->bsf, sizeof(orc_bsf_record) * k);
    return c->nbsf;
}

void orc_colony_tau(const orc_colony *c, float *tau, float *tinfo)
{
    size_t nn = (size_t)c->n * c->n;
    if (tau) memcpy(tau, c->tau, sizeof(float) * nn);
    if (tinfo) memcpy(tinfo, c->tinfo, sizeof(float) * nn);
}

void orc_colony_set_tau(orc_colony *c, const float *tau, const float *tinfo)
{
    size_t nn = (size_t)c->n * c->n;
    memcpy(c->tau, tau, sizeof(float) * nn);
    memcpy(c->tinfo, tinfo, sizeof(float) * nn);
    colony_gather(c);
}

void orc_colony_set_ants(orc_colony *c, const int *tours, int stride, const int *sizes, const float *lens)
{
    for (int a = 0; a <= c->m; ++a) {
        memcpy(c->tours + (size_t)a * c->T, tours + (size_t)a * stride, sizeof(int) * sizes[a]);
        c->sizes[a] = sizes[a];
        c->lens[a] = lens[a];
    }
}

long long orc_colony_fallbacks(const orc_colony *c) { return c->fallbacks; }
int32_t orc_colony_ls_seed(const orc_colony *c) { return c->ls_seed; }
void orc_colony_set_ls_seed(orc_colony *c, int32_t seed) { c->ls_seed = seed; }
void orc_colony_set_pm_state(orc_colony *c, const int32_t *state)
{
    memcpy(c->pm, state, sizeof(int32_t) * (size_t)c->m * ORC_LANES);
}

long long orc_colony_ls_evals(const orc_colony *c) { return c->ls_evals; }

void orc_colony_pm_state(const orc_colony *c, int32_t *out)
{
    memcpy(out, c->pm, sizeof(int32_t) * (size_t)c->m * ORC_LANES);
}

/* =========================================================================
 * Sub-problem decomposition (vrpHelper.cpp:274-294, problem.cpp:84-130,
 * problem.h:136-139; partition and merge rules have no reference
 * counterpart, see oracle.h)
 * ========================================================================= */

int orc_route_centers(int n, const float *x, const float *y, const int *dem, const int *tour, int size,
                      int cap, int *beg, int *end, float *cx, float *cy, int *rdem)
{
    (void)n;
    int R = 0, p = 0;
    for (int q = 1; q < size; ++q) {
        if (tour[q] != 0) continue;
        if (q - p >= 2) {
            if (R >= cap) return -1;
            int d = 0;
            for (int j = p; j < q; ++j) d += dem[tour[j]];
            float fx = 0.0f, fy = 0.0f;
            for (int j = p; j < q; ++j) {
                const int t = tour[j];
                fx = (float)((double)fx + (double)(x[t] * (float)dem[t]) * 1.0 / (double)d);
                fy = (float)((double)fy + (double)(y[t] * (float)dem[t]) * 1.0 / (double)d);
            }
            beg[R] = p;
            end[R] = q;
            cx[R] = fx;
            cy[R] = fy;
            rdem[R] = d;
            ++R;
        }
        p = q;
    }
    return R;
}

typedef struct {
    double ang;
    int r;
} orc_route_key;

static int route_key_cmp(const void *a, const void *b)
{
    const orc_route_key *u = (const orc_route_key *)a, *v = (const orc_route_key *)b;
    if (u->ang < v->ang) return -1;
    if (u->ang > v->ang) return 1;
    return (u->r > v->r) - (u->r < v->r);
}

int orc_partition_routes(int n, const float *x, const float *y, const int *dem, const int *tour, int size,
                         int subs, int rotation, int *offsets, int *real_nodes, int *toffsets, int *sub_tours)
{
    const int cap = size;
    int *beg = (int *)malloc(sizeof(int) * cap * 3);
    int *end = beg + cap, *rdem = end + cap;
    float *cx = (float *)malloc(sizeof(float) * cap * 2), *cy = cx + cap;
    orc_route_key *key = (orc_route_key *)malloc(sizeof(orc_route_key) * cap);
    int *group = (int *)malloc(sizeof(int) * cap);
    const int R = orc_route_centers(n, x, y, dem, tour, size, cap, beg, end, cx, cy, rdem);
    if (R < 1) {
        free(beg); free(cx); free(key); free(group);
        return -1;
    }
    for (int r = 0; r < R; ++r) {
        double px = cx[r], py = cy[r];
        if (rdem[r] == 0) {
            px = 0.0;
            py = 0.0;
            for (int j = beg[r] + 1; j < end[r]; ++j) {
                px += x[tour[j]];
                py += y[tour[j]];
            }
            px /= (double)(end[r] - beg[r] - 1);
            py /= (double)(end[r] - beg[r] - 1);
        }
        key[r].ang = atan2(py - (double)y[0], px - (double)x[0]);
        key[r].r = r;
    }
    qsort(key, R, sizeof(orc_route_key), route_key_cmp);
    int G = subs > 0 ? subs : (n / 50 < R / 2 ? n / 50 : R / 2);
    if (G < 1) G = 1;
    if (G > R) G = R;
    int total = 0;
    for (int r = 0; r < R; ++r) total += end[r] - beg[r] - 1;
    const int start = ((rotation % R) + R) % R;
    int g = 0, acc = 0;
    for (int q = 0; q < R; ++q) {
        const int r = key[(start + q) % R].r;
        group[q] = g;
        acc += end[r] - beg[r] - 1;
        const int left_routes = R - q - 1, left_groups = G - 1 - g;
        if (left_groups > 0 && ((long long)acc * G >= (long long)(g + 1) * total || left_routes == left_groups)) ++g;
    }
    int no = 0, to = 0, q = 0;
    for (int gg = 0; gg < G; ++gg) {
        offsets[gg] = no;
        toffsets[gg] = to;
        real_nodes[no++] = 0;
        sub_tours[to++] = 0;
        int local = 1;
        for (; q < R && group[q] == gg; ++q) {
            const int r = key[(start + q) % R].r;
            for (int j = beg[r] + 1; j < end[r]; ++j) {
                real_nodes[no++] = tour[j];
                sub_tours[to++] = local++;
            }
            sub_tours[to++] = 0;
        }
    }
    offsets[G] = no;
    toffsets[G] = to;
    free(beg); free(cx); free(key); free(group);
    return G;
}

void orc_sub_problem(int n, const float *dist, const int *dem, int sub_n, const int *real_nodes,
                     float *sub_dist, int *sub_dem)
{
    for (int i = 0; i < sub_n; ++i) {
        const int ri = real_nodes[i];
        for (int j = 0; j < sub_n; ++j) sub_dist[(size_t)i * sub_n + j] = dist[(size_t)ri * n + real_nodes[j]];
        sub_dem[i] = dem[ri];
    }
}

static float mapped_length(int n, const float *dist, const int *real, const int *t, int size)
{
    float len = 0.0f;
    for (int q = 0; q + 1 < size; ++q) len += dist[(size_t)real[t[q]] * n + real[t[q + 1]]];
    return len;
}

int orc_merge_sub_tours(int n, const float *dist, int count, const int *offsets, const int *real_nodes,
                        const int *toffsets, const int *sub_tours, const int *boffsets, const int *best_tours,
                        int *out, float *len, int *replaced)
{
    int o = 0;
    out[o++] = 0;
    for (int s = 0; s < count; ++s) {
        const int *real = real_nodes + offsets[s];
        const int *orig = sub_tours + toffsets[s];
        const int osz = toffsets[s + 1] - toffsets[s];
        const int *best = best_tours + boffsets[s];
        const int bsz = boffsets[s + 1] - boffsets[s];
        const int take = bsz > 0 && mapped_length(n, dist, real, best, bsz) < mapped_length(n, dist, real, orig, osz);
        replaced[s] = take;
        const int *t = take ? best : orig;
        const int sz = take ? bsz : osz;
        for (int q = 1; q < sz; ++q) out[o++] = real[t[q]];
    }
    *len = orc_tour_length(n, dist, out, o);
    return o;
}

void orc_pheromone_from_sub(int n, float alpha, const float *heur, float *tau, float *tinfo, int sub_n,
                            const int *real_nodes, const float *sub_tau)
{
    for (int i = 0; i < sub_n; ++i)
        for (int j = 0; j < sub_n; ++j) {
            if (i == j) continue;
            const size_t e = (size_t)real_nodes[i] * n + real_nodes[j];
            const float t = sub_tau[(size_t)i * sub_n + j];
            tau[e] = t;
            tinfo[e] = total_info(t, alpha, heur[e]);
        }
}
