// Learned edge prior: features, MLP forward / backward on the matrix cores,
// Adam, minibatch training and the dense prior for aco_problem.prior.
//
// No counterpart in the reference (its eta is 1/(d+0.1)^beta from the
// distance alone, kernel.cl:672-679,782-789).
//
// MLP 12 -> 64 -> 64 -> 1 with ReLU.  BF16 operands, FP32 accumulation on
// the matrix cores.  One wave processes 16 edges per tile with the edges on
// the N dimension of every product: the accumulator layout of one layer
// (lane l holds rows 4*(l/16)+j, column l%16) is exactly the B-operand
// layout of the next, so the forward pass needs no data movement between
// layers.  The weight gradients contract over edges, which needs the
// transposed layout: 16x16 tiles go through a per-wave LDS buffer.
//
// Layer 1 (K = 12 features, padded to 16) and the weight gradients (K = the
// 16 edges of a tile) use v_mfma_f32_16x16x16_bf16.  The 64-deep products
// (layer 2 and dH1 = W2^T dz2) use gfx950's v_mfma_f32_16x16x32_bf16, two
// per 16-row block: its A and B operands are two 4-element halves put side
// by side, each half one 16-deep slice in the layout above.  Element i of a
// lane's A operand and element i of its B operand carry the same index of
// the contraction, so the product sums exactly the slices' terms, whatever
// order the unit adds them in.
//
// Determinism: per-wave partial gradients (fixed tile assignment) are summed
// by the reduce kernel in a fixed order; there are no atomics.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "aco_internal.h"
#include "cvrp_aco.h"
#include "cvrp_prior.h"

namespace cvrp {
namespace prior {

constexpr int NF = 12, NFP = 16, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = H * NF, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NP = OFF_B3 + 1;
static_assert(NP == ACO_PRIOR_PARAMS, "parameter layout");
constexpr int NPL = NP + 1;         // partial row: gradient, then the loss sum
constexpr int kThreads = 256, kWaves = kThreads / 64;
constexpr int kTilesPerWave = 4;    // training: 64 edges per wave per launch
constexpr int kMaxBatch = 65536;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline unsigned short f2bf(float f)
{
    // round to nearest even (finite inputs)
    unsigned u;
    memcpy(&u, &f, 4);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ v4s ld4(const unsigned short *p) { return *reinterpret_cast<const v4s *>(p); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// c + [a0 a1] x [b0; b1]: two 16-deep slices in one 32-deep product.
__device__ __forceinline__ v4f mfma2(v4s a0, v4s a1, v4s b0, v4s b1, v4f c)
{
    const v8s a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
    const v8s b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c,
                                                    0, 0, 0);
}

// Block-resident network: bf16 weights (W2 also transposed, for dH1),
// fp32 biases, w3 rounded to bf16 and widened, and per-wave 16x16 bf16
// transpose buffers.
struct NetLds {
    alignas(16) unsigned short w1[H][NFP];
    alignas(16) unsigned short w2[H][H];
    alignas(16) unsigned short w2t[H][H];
    alignas(16) float b1[H];
    alignas(16) float b2[H];
    alignas(16) float w3[H];
    float b3;
    alignas(16) unsigned short tr[kWaves][9][16][16];
};

__device__ void load_net(NetLds &S, const unsigned short *wb, const float *pf)
{
    const int t = threadIdx.x;
    for (int e = t; e < H * NFP; e += kThreads) {
        const int o = e / NFP, f = e % NFP;
        S.w1[o][f] = f < NF ? wb[OFF_W1 + o * NF + f] : (unsigned short)0;
    }
    for (int e = t; e < H * H; e += kThreads) {
        const int o = e / H, i = e % H;
        const unsigned short w = wb[OFF_W2 + e];
        S.w2[o][i] = w;
        S.w2t[i][o] = w;
    }
    for (int e = t; e < H; e += kThreads) {
        S.b1[e] = pf[OFF_B1 + e];
        S.b2[e] = pf[OFF_B2 + e];
        S.w3[e] = bf2f(wb[OFF_W3 + e]);
    }
    if (t == 0) S.b3 = pf[OFF_B3];
    __syncthreads();
}

struct Fwd {
    v4f z1[4];
    v4s h1[4];
    v4f z2[4];
    float z;
};

// The weights one lane feeds to the matrix cores (row r = l%16, slice
// g = l/16 of every 16-deep operand), read from the block's LDS copy at
// each use ...
struct LdsWeights {
    const NetLds &S;
    int r, g;
    __device__ v4s w1(int t) const { return ld4(&S.w1[16 * t + r][4 * g]); }
    __device__ v4s w2(int mt, int kt) const { return ld4(&S.w2[16 * mt + r][16 * kt + 4 * g]); }
    __device__ v4f b1(int t) const { return *reinterpret_cast<const v4f *>(&S.b1[16 * t + 4 * g]); }
    __device__ v4f b2(int mt) const { return *reinterpret_cast<const v4f *>(&S.b2[16 * mt + 4 * g]); }
    __device__ float w3(int mt, int j) const { return S.w3[16 * mt + 4 * g + j]; }
    __device__ float b3() const { return S.b3; }
};

// ... or from the parameter buffers in global memory (bf16 weights, W1
// rows of 12 features zero-padded to 16 here; fp32 biases; w3 widened from
// its bf16 copy) ...
struct GlobalWeights {
    const unsigned short *wb;
    const float *pf;
    int r, g;
    __device__ v4s w1(int t) const
    {
        if (g * 4 >= NF) return v4s{0, 0, 0, 0};
        return ld4(wb + OFF_W1 + (16 * t + r) * NF + 4 * g);
    }
    __device__ v4s w2(int mt, int kt) const { return ld4(wb + OFF_W2 + (16 * mt + r) * H + 16 * kt + 4 * g); }
    __device__ v4f b1(int t) const { return *reinterpret_cast<const v4f *>(pf + OFF_B1 + 16 * t + 4 * g); }
    __device__ v4f b2(int mt) const { return *reinterpret_cast<const v4f *>(pf + OFF_B2 + 16 * mt + 4 * g); }
    __device__ float w3(int mt, int j) const { return bf2f(wb[OFF_W3 + 16 * mt + 4 * g + j]); }
    __device__ float b3() const { return pf[OFF_B3]; }
};

// ... or held in registers for a whole loop over tiles (about 90 VGPRs: the
// inference kernel, whose waves run many tiles each).
struct RegWeights {
    v4s w1_[4], w2_[4][4];
    v4f b1_[4], b2_[4];
    float w3_[16], b3_;
    template <typename Src>
    __device__ explicit RegWeights(const Src &L)
    {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            w1_[t] = L.w1(t);
            b1_[t] = L.b1(t);
            b2_[t] = L.b2(t);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w2_[t][k] = L.w2(t, k);
                w3_[4 * t + k] = L.w3(t, k);
            }
        }
        b3_ = L.b3();
    }
    __device__ v4s w1(int t) const { return w1_[t]; }
    __device__ v4s w2(int mt, int kt) const { return w2_[mt][kt]; }
    __device__ v4f b1(int t) const { return b1_[t]; }
    __device__ v4f b2(int mt) const { return b2_[mt]; }
    __device__ float w3(int mt, int j) const { return w3_[4 * mt + j]; }
    __device__ float b3() const { return b3_; }
};

// x: B operand of layer 1 (lane: features 4*(l/16)+j of edge l%16).
template <typename Wt>
__device__ __forceinline__ void fwd_tile(const Wt &W, v4s x, Fwd &F)
{
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        F.z1[t] = mfma(W.w1(t), x, W.b1(t));
        v4s h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = F.z1[t][j];
            h[j] = (short)f2bf(z > 0.0f ? z : 0.0f);
        }
        F.h1[t] = h;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        v4f c = mfma2(W.w2(mt, 0), W.w2(mt, 1), F.h1[0], F.h1[1], W.b2(mt));
        F.z2[mt] = mfma2(W.w2(mt, 2), W.w2(mt, 3), F.h1[2], F.h1[3], c);
    }
    float s = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float h = F.z2[mt][j] > 0.0f ? F.z2[mt][j] : 0.0f;
            s += W.w3(mt, j) * h;
        }
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    F.z = s + W.b3();
}

// Accumulator-layout tile (rows 4g+j, column r) -> LDS, read back as the
// operand of a product contracting over the former columns: lane gets
// row r, columns 4g..4g+3.
__device__ __forceinline__ v4s transpose_tile(unsigned short (*T)[16], v4s d, int lane)
{
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) T[4 * g + j][r] = (unsigned short)d[j];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const v4s out = ld4(&T[r][4 * g]);
    return out;
}

__device__ __forceinline__ v4s pack_bf16(v4f v)
{
    v4s h;
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = (short)f2bf(v[j]);
    return h;
}

__device__ __forceinline__ v4s load_x(const unsigned short *xb, long long e, bool valid, int g)
{
    if (!valid) return v4s{0, 0, 0, 0};
    return ld4(xb + (size_t)e * NFP + 4 * g);
}

// ---------------------------------------------------------------- kernels --

// Inference: each wave takes tiles tile, tile + waves, ... with its weight
// fragments in registers (read once from global memory: no LDS, no block
// barrier) and the next tile's features requested before the current
// tile's products.
__global__ __launch_bounds__(kThreads) void k_forward(const unsigned short *xb, long long E,
                                                      const unsigned short *wb, const float *pf,
                                                      float *logits)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const long long tiles = (E + 15) / 16;
    const long long waves = (long long)gridDim.x * kWaves;
    long long tile = (long long)blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (tile >= tiles) return;   // wave-uniform
    const RegWeights W(GlobalWeights{wb, pf, r, g});
    v4s x = load_x(xb, tile * 16 + r, tile * 16 + r < E, g);
    for (; tile < tiles; tile += waves) {
        const long long e = tile * 16 + r;
        const long long en = (tile + waves) * 16 + r;
        const v4s xn = load_x(xb, en, en < E, g);
        Fwd F;
        fwd_tile(W, x, F);
        if (g == 0 && e < E) logits[e] = F.z;
        x = xn;
    }
}

// Fused forward + backward over kTilesPerWave tiles of the batch
// (positions [0, Eb) of idx); writes one partial row per wave.
__global__ __launch_bounds__(kThreads) void k_train(const unsigned short *xb, const float *label,
                                                    const int *idx, int Eb,
                                                    const unsigned short *wb, const float *pf,
                                                    float pos_weight, float invE, float *partials)
{
    __shared__ NetLds S;
    load_net(S, wb, pf);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
    const int gw = blockIdx.x * kWaves + w;
    const int tiles = (Eb + 15) / 16;
    unsigned short(*T)[16][16] = S.tr[w];
    const v4f zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
    v4f dW2[4][4], dW1[4], db1[4], db2[4], dw3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        dW1[a] = db1[a] = db2[a] = dw3[a] = zero4;
#pragma unroll
        for (int b = 0; b < 4; ++b) dW2[a][b] = zero4;
    }
    float db3 = 0.0f, loss = 0.0f;
    for (int tt = 0; tt < kTilesPerWave; ++tt) {
        const int tile = gw * kTilesPerWave + tt;
        if (tile >= tiles) break;
        const int pos = tile * 16 + r;
        const bool valid = pos < Eb;
        const int e = valid ? idx[pos] : 0;
        const v4s x = load_x(xb, e, valid, g);
        const float y = valid ? label[e] : 0.0f;
        Fwd F;
        fwd_tile(LdsWeights{S, r, g}, x, F);
        const float we = valid ? (y > 0.5f ? pos_weight : 1.0f) : 0.0f;
        const float sig = 1.0f / (1.0f + expf(-F.z));
        const float ge = we * (sig - y) * invE;
        if (g == 0 && valid) {
            loss += we * ((fmaxf(F.z, 0.0f) - y * F.z) + log1pf(expf(-fabsf(F.z))));
            db3 += ge;
        }
        // layer 3 and the ReLU of layer 2
        v4s dz2b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f d;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool on = F.z2[mt][j] > 0.0f;
                dw3[mt][j] += ge * (on ? F.z2[mt][j] : 0.0f);
                d[j] = on ? S.w3[16 * mt + 4 * g + j] * ge : 0.0f;
                db2[mt][j] += d[j];
            }
            dz2b[mt] = pack_bf16(d);
        }
        // dH1 = W2^T dz2 (dz2 is already the B operand), ReLU of layer 1
        v4s dz1b[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            v4f c = mfma2(ld4(&S.w2t[16 * it + r][4 * g]), ld4(&S.w2t[16 * it + r][16 + 4 * g]), dz2b[0],
                          dz2b[1], zero4);
            c = mfma2(ld4(&S.w2t[16 * it + r][32 + 4 * g]), ld4(&S.w2t[16 * it + r][48 + 4 * g]), dz2b[2],
                      dz2b[3], c);
            v4f d;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                d[j] = F.z1[it][j] > 0.0f ? c[j] : 0.0f;
                db1[it][j] += d[j];
            }
            dz1b[it] = pack_bf16(d);
        }
        // weight gradients contract over the 16 edges of the tile
        v4s aT[4], hT[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            aT[q] = transpose_tile(T[q], dz2b[q], lane);
            hT[q] = transpose_tile(T[4 + q], F.h1[q], lane);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int it = 0; it < 4; ++it) dW2[mt][it] = mfma(aT[mt], hT[it], dW2[mt][it]);
        const v4s xT = transpose_tile(T[8], x, lane);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            const v4s a1 = transpose_tile(T[ot], dz1b[ot], lane);
            dW1[ot] = mfma(a1, xT, dW1[ot]);
        }
    }
    // per-wave partial row
    float *P = partials + (size_t)gw * NPL;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int j = 0; j < 4; ++j) P[OFF_W2 + (16 * mt + 4 * g + j) * H + 16 * it + r] = dW2[mt][it][j];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (r < NF) P[OFF_W1 + (16 * ot + 4 * g + j) * NF + r] = dW1[ot][j];
    // bias / w3 gradients: sum over the 16 edge lanes of each row group
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float a = db1[t][j], b = db2[t][j], c = dw3[t][j];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a += __shfl_xor(a, o);
                b += __shfl_xor(b, o);
                c += __shfl_xor(c, o);
            }
            if (r == 0) {
                P[OFF_B1 + 16 * t + 4 * g + j] = a;
                P[OFF_B2 + 16 * t + 4 * g + j] = b;
                P[OFF_W3 + 16 * t + 4 * g + j] = c;
            }
        }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        db3 += __shfl_xor(db3, o);
        loss += __shfl_xor(loss, o);
    }
    if (lane == 0) {
        P[OFF_B3] = db3;
        P[NP] = loss * invE;
    }
}

// Per parameter: the sum of the partial rows in a fixed order (runs of
// consecutive waves, then the runs in order), then (update != 0) Adam in
// fp32.  Parameter NP is the loss.
struct AdamArgs {
    float lr, beta1, om_beta1, beta2, om_beta2, c1, c2, eps;
};

// A block takes kReduceParams parameters; each is summed by kReduceGroups
// threads over consecutive runs of the partial rows, then the runs' sums are
// added in run order (a fixed order: runs are deterministic).
constexpr int kReduceParams = 32, kReduceGroups = 8;

__global__ __launch_bounds__(kReduceParams * kReduceGroups) void k_reduce_adam(
    const float *partials, int W, float *pf, float *m, float *v, unsigned short *wb, float *grad_out,
    float *loss_out, int update, AdamArgs a)
{
    __shared__ float red[kReduceGroups][kReduceParams];
    const int i = threadIdx.x % kReduceParams, q = threadIdx.x / kReduceParams;
    const int p = blockIdx.x * kReduceParams + i;
    const int run = (W + kReduceGroups - 1) / kReduceGroups;
    const int w0 = q * run, w1 = min(W, w0 + run);
    float part = 0.0f;
    if (p <= NP)
        for (int w = w0; w < w1; ++w) part += partials[(size_t)w * NPL + p];
    red[q][i] = part;
    __syncthreads();
    if (q != 0 || p > NP) return;
    float s = red[0][i];
    for (int g = 1; g < kReduceGroups && g * run < W; ++g) s += red[g][i];
    if (p == NP) {
        if (loss_out) *loss_out = s;
        return;
    }
    if (grad_out) grad_out[p] = s;
    if (!update) return;
    const float mm = a.beta1 * m[p] + a.om_beta1 * s;
    const float vv = a.beta2 * v[p] + a.om_beta2 * (s * s);
    const float mh = mm / a.c1, vh = vv / a.c2;
    const float np = pf[p] - a.lr * (mh / (sqrtf(vh) + a.eps));
    m[p] = mm;
    v[p] = vv;
    pf[p] = np;
    wb[p] = f2bf(np);
}

__global__ void k_to_bf16(const float *pf, unsigned short *wb, int count)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) wb[p] = f2bf(pf[p]);
}

// Features -> bf16 rows of 16 (features 12..15 zero).
__global__ void k_pack_features(const float *feat, long long E, unsigned short *xb)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * NFP) return;
    const long long e = t / NFP;
    const int f = (int)(t % NFP);
    xb[t] = f < NF ? f2bf(feat[e * NF + f]) : (unsigned short)0;
}

// dist rows have pitch P (>= n)
__global__ void k_node_stats(int n, int P, const float *x, const float *y, const float *dist,
                             const int *nn, int nnw, int k, float *rad, float *ang, float *kd)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float dx = x[v] - x[0], dy = y[v] - y[0];
    rad[v] = sqrtf(dx * dx + dy * dy);
    ang[v] = atan2f(dy, dx) / 3.14159265358979323846f;
    kd[v] = dist[(size_t)v * P + nn[(size_t)v * nnw + k - 1]];
}

// One block: R = max radius (1 if 0), mkd = mean k-th candidate distance
// (summed in double, 1 if 0).
__global__ void k_node_reduce(int n, const float *rad, const float *kd, float *out)
{
    __shared__ double s[256];
    __shared__ float mx[256];
    const int t = threadIdx.x;
    double a = 0.0;
    float m = 0.0f;
    for (int v = t; v < n; v += 256) {
        a += (double)kd[v];
        m = fmaxf(m, rad[v]);
    }
    s[t] = a;
    mx[t] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            s[t] += s[t + o];
            mx[t] = fmaxf(mx[t], mx[t + o]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const float R = mx[0] > 0.0f ? mx[0] : 1.0f;
        const float mkd = (float)(s[0] / (double)n);
        out[0] = R;
        out[1] = mkd > 0.0f ? mkd : 1.0f;
    }
}

__global__ void k_edge_features(int n, int P, int k, const float *dist, const int *nn, int nnw,
                                const int *dem, float Q, const float *rad, const float *ang,
                                const float *kd, const float *Rm, float *feat, unsigned short *xb)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    const int i = (int)(t / k), r = (int)(t % k);
    const int j = nn[(size_t)i * nnw + r];
    const float R = Rm[0], mkd = Rm[1];
    const float dij = dist[(size_t)i * P + j];
    float f[NF];
    f[0] = dij / mkd;
    f[1] = (float)r / (float)k;
    f[2] = (float)dem[i] / Q;
    f[3] = (float)dem[j] / Q;
    f[4] = ang[i];
    f[5] = ang[j];
    f[6] = rad[i] / R;
    f[7] = rad[j] / R;
    f[8] = kd[i] / mkd;
    f[9] = kd[j] / mkd;
    f[10] = ((dist[i] + dist[j]) - dij) / R;
    bool mutual = false;
    for (int q = 0; q < k; ++q) mutual |= nn[(size_t)j * nnw + q] == i;
    f[11] = mutual ? 1.0f : 0.0f;
    if (feat)
        for (int q = 0; q < NF; ++q) feat[t * NF + q] = f[q];
    if (xb)
        for (int q = 0; q < NFP; ++q) xb[t * NFP + q] = q < NF ? f2bf(f[q]) : (unsigned short)0;
}

__device__ __forceinline__ float prior_pow(float p, float ex)
{
    if (ex == 1.0f) return p;
    if (ex == 2.0f) return p * p;
    return powf(p, ex);
}

__global__ void k_prior_fill(long long cnt, float val, float *out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) out[t] = val;
}

__global__ void k_prior_scatter(int n, int k, const int *nn, int nnw, const float *logits, float ex,
                                float *out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    const int i = (int)(t / k), r = (int)(t % k);
    const int j = nn[(size_t)i * nnw + r];
    const float p = 1.0f / (1.0f + expf(-logits[t]));
    out[(size_t)i * n + j] = prior_pow(p, ex);
}

// heur[i][j] (pitch P) *= prior[i][j] (pitch n), j < n
__global__ void k_heur_mul(int n, int P, const float *prior, float *heur)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    const size_t e = (size_t)i * P + j;
    heur[e] = heur[e] * prior[t];
}

// ------------------------------------------------------------------- host --

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

static hipError_t dalloc(DevBuf &b, size_t bytes)
{
    return hipMalloc(&b.p, bytes ? bytes : 16);
}

}  // namespace prior
}  // namespace cvrp

using namespace cvrp;
using namespace cvrp::prior;

struct aco_prior {
    hipStream_t stream = nullptr;
    float *pf = nullptr, *m = nullptr, *v = nullptr;
    unsigned short *wb = nullptr;
    int step = 0;
};

#define PCHK(expr)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_error(ACO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// SplitMix64 (Steele, Lea & Flood, 2014): state += golden; output mix(state)
static uint64_t sm_next(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
}

static void glorot_init(uint64_t seed, std::vector<float> &p)
{
    p.assign(NP, 0.0f);
    uint64_t s = seed;
    const int offs[3] = {OFF_W1, OFF_W2, OFF_W3}, cnts[3] = {H * NF, H * H, H};
    const int fin[3] = {NF, H, H}, fout[3] = {H, H, 1};
    for (int q = 0; q < 3; ++q) {
        const float bound = sqrtf(6.0f / (float)(fin[q] + fout[q]));
        for (int t = 0; t < cnts[q]; ++t) {
            const float u = (float)(sm_next(s) >> 40) * (1.0f / 16777216.0f);
            p[offs[q] + t] = (2.0f * u - 1.0f) * bound;
        }
    }
}

static int upload_params(aco_prior *net, const float *params)
{
    PCHK(hipMemcpyAsync(net->pf, params, NP * sizeof(float), hipMemcpyHostToDevice, net->stream));
    PCHK(hipMemsetAsync(net->m, 0, NP * sizeof(float), net->stream));
    PCHK(hipMemsetAsync(net->v, 0, NP * sizeof(float), net->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((NP + 255) / 256), dim3(256), 0, net->stream, net->pf, net->wb, NP);
    PCHK(hipGetLastError());
    PCHK(hipStreamSynchronize(net->stream));
    net->step = 0;
    return ACO_OK;
}

static AdamArgs adam_args(int step, float lr, float b1, float b2, float eps)
{
    // bias corrections and (1 - beta) in double from the fp32 betas, rounded once
    AdamArgs a;
    a.lr = lr;
    a.beta1 = b1;
    a.beta2 = b2;
    a.om_beta1 = (float)(1.0 - (double)b1);
    a.om_beta2 = (float)(1.0 - (double)b2);
    a.c1 = (float)(1.0 - pow((double)b1, (double)step));
    a.c2 = (float)(1.0 - pow((double)b2, (double)step));
    a.eps = eps;
    return a;
}

// bf16 features of `count` edges (host fp32, count x 12) into xb.
static int stage_features(aco_prior *net, const float *feat, long long count, DevBuf &xb)
{
    DevBuf f32;
    PCHK(dalloc(f32, (size_t)count * NF * sizeof(float)));
    PCHK(dalloc(xb, (size_t)count * NFP * sizeof(unsigned short)));
    PCHK(hipMemcpyAsync(f32.p, feat, (size_t)count * NF * sizeof(float), hipMemcpyHostToDevice, net->stream));
    const long long tot = count * NFP;
    hipLaunchKernelGGL(k_pack_features, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, net->stream,
                       f32.as<float>(), count, xb.as<unsigned short>());
    PCHK(hipGetLastError());
    PCHK(hipStreamSynchronize(net->stream));
    return ACO_OK;
}

static int launch_forward(aco_prior *net, const unsigned short *xb, long long count, float *logits,
                          hipStream_t st = nullptr)
{
    const long long tiles = (count + 15) / 16;
    long long blocks = (tiles + kWaves - 1) / kWaves;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_forward, dim3((unsigned)blocks), dim3(kThreads), 0, st ? st : net->stream, xb,
                       count, net->wb, net->pf, logits);
    PCHK(hipGetLastError());
    return ACO_OK;
}

static float floor_pow(float floor, float exponent)
{
    return exponent == 1.0f ? floor : exponent == 2.0f ? floor * floor : powf(floor, exponent);
}

// One batch: fused forward/backward, then the ordered reduce (+ Adam).
static int launch_batch(aco_prior *net, const unsigned short *xb, const float *label, const int *idx,
                        int Eb, float pos_weight, float *partials, float *grad_out, float *loss_out,
                        bool update, const AdamArgs &aa)
{
    const int tiles = (Eb + 15) / 16;
    const int waves = (tiles + kTilesPerWave - 1) / kTilesPerWave;
    const int blocks = (waves + kWaves - 1) / kWaves;
    hipLaunchKernelGGL(k_train, dim3(blocks), dim3(kThreads), 0, net->stream, xb, label, idx, Eb,
                       net->wb, net->pf, pos_weight, 1.0f / (float)Eb, partials);
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPL + kReduceParams - 1) / kReduceParams),
                       dim3(kReduceParams * kReduceGroups), 0, net->stream, partials,
                       blocks * kWaves, net->pf, net->m, net->v, net->wb, grad_out, loss_out,
                       update ? 1 : 0, aa);
    PCHK(hipGetLastError());
    return ACO_OK;
}

static size_t partial_bytes(int Eb)
{
    const int tiles = (Eb + 15) / 16;
    const int waves = (tiles + kTilesPerWave - 1) / kTilesPerWave;
    const int blocks = (waves + kWaves - 1) / kWaves;
    return (size_t)blocks * kWaves * NPL * sizeof(float);
}

extern "C" {

void aco_prior_default_train_params(aco_prior_train_params *tp)
{
    if (!tp) return;
    tp->epochs = 4;
    tp->batch = 16384;
    tp->lr = 3e-3f;
    tp->lr_final = 3e-4f;
    tp->pos_weight = 4.0f;
    tp->beta1 = 0.9f;
    tp->beta2 = 0.999f;
    tp->eps = 1e-8f;
    tp->shuffle_seed = 1;
}

int aco_prior_create(uint64_t seed, aco_prior **out)
{
    if (!out) return set_error(ACO_ERR_ARG, "null output handle");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(ACO_ERR_NODEV, "no HIP device");
    aco_prior *net = new aco_prior();
    auto fail = [&](int rc) {
        aco_prior_destroy(net);
        return rc;
    };
    if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&net->pf, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&net->m, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&net->v, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&net->wb, NP * sizeof(unsigned short)) != hipSuccess)
        return fail(set_error(ACO_ERR_HIP, "aco_prior_create: device allocation failed"));
    std::vector<float> p;
    glorot_init(seed, p);
    const int rc = upload_params(net, p.data());
    if (rc != ACO_OK) return fail(rc);
    *out = net;
    return ACO_OK;
}

int aco_prior_destroy(aco_prior *net)
{
    if (!net) return ACO_OK;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    if (net->pf) (void)hipFree(net->pf);
    if (net->m) (void)hipFree(net->m);
    if (net->v) (void)hipFree(net->v);
    if (net->wb) (void)hipFree(net->wb);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return ACO_OK;
}

int aco_prior_get_params(aco_prior *net, float *params, int32_t count)
{
    if (!net || !params) return set_error(ACO_ERR_ARG, "null argument");
    if (count != NP) return set_error(ACO_ERR_ARG, "count must be ACO_PRIOR_PARAMS");
    PCHK(hipMemcpyAsync(params, net->pf, NP * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    PCHK(hipStreamSynchronize(net->stream));
    return ACO_OK;
}

int aco_prior_set_params(aco_prior *net, const float *params, int32_t count)
{
    if (!net || !params) return set_error(ACO_ERR_ARG, "null argument");
    if (count != NP) return set_error(ACO_ERR_ARG, "count must be ACO_PRIOR_PARAMS");
    for (int i = 0; i < NP; ++i)
        if (!std::isfinite(params[i])) return set_error(ACO_ERR_ARG, "non-finite parameter");
    return upload_params(net, params);
}

static int check_instance_args(int32_t n, const float *x, const float *y, const int32_t *demand,
                               int32_t capacity, const float *dist, const int32_t *nn,
                               int32_t nn_width, int32_t k)
{
    if (n < 2 || n > 65536) return set_error(ACO_ERR_ARG, "n must be in [2, 65536]");
    if (!x || !y || !demand || !dist || !nn) return set_error(ACO_ERR_ARG, "null array");
    if (capacity <= 0) return set_error(ACO_ERR_ARG, "capacity must be positive");
    if (k < 1 || k > nn_width || k > n - 1) return set_error(ACO_ERR_ARG, "need 1 <= k <= nn_width, k < n");
    for (long long t = 0; t < (long long)n * nn_width; ++t)
        if (nn[t] < 0 || nn[t] >= n) return set_error(ACO_ERR_ARG, "nn list entry out of range");
    return ACO_OK;
}

// Device features of one instance: fp32 (optional) and bf16 rows.
static int device_features(hipStream_t st, int32_t n, const float *x, const float *y,
                           const int32_t *demand, int32_t capacity, const float *dist,
                           const int32_t *nn, int32_t nnw, int32_t k, DevBuf &feat_d, DevBuf &xb,
                           DevBuf &nn_d, bool want_f32)
{
    DevBuf xd, yd, dd, demd, rad, ang, kd, Rm;
    const size_t N = (size_t)n, E = N * (size_t)k;
    PCHK(dalloc(xd, N * 4));
    PCHK(dalloc(yd, N * 4));
    PCHK(dalloc(dd, N * N * 4));
    PCHK(dalloc(demd, N * 4));
    PCHK(dalloc(nn_d, N * nnw * 4));
    PCHK(dalloc(rad, N * 4));
    PCHK(dalloc(ang, N * 4));
    PCHK(dalloc(kd, N * 4));
    PCHK(dalloc(Rm, 16));
    PCHK(dalloc(xb, E * NFP * 2));
    if (want_f32) PCHK(dalloc(feat_d, E * NF * 4));
    PCHK(hipMemcpyAsync(xd.p, x, N * 4, hipMemcpyHostToDevice, st));
    PCHK(hipMemcpyAsync(yd.p, y, N * 4, hipMemcpyHostToDevice, st));
    PCHK(hipMemcpyAsync(dd.p, dist, N * N * 4, hipMemcpyHostToDevice, st));
    PCHK(hipMemcpyAsync(demd.p, demand, N * 4, hipMemcpyHostToDevice, st));
    PCHK(hipMemcpyAsync(nn_d.p, nn, N * nnw * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_node_stats, dim3((n + 255) / 256), dim3(256), 0, st, n, n, xd.as<float>(),
                       yd.as<float>(), dd.as<float>(), nn_d.as<int>(), nnw, k, rad.as<float>(),
                       ang.as<float>(), kd.as<float>());
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(k_node_reduce, dim3(1), dim3(256), 0, st, n, rad.as<float>(), kd.as<float>(),
                       Rm.as<float>());
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(k_edge_features, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, st, n, n, k,
                       dd.as<float>(), nn_d.as<int>(), nnw, demd.as<int>(), (float)capacity,
                       rad.as<float>(), ang.as<float>(), kd.as<float>(), Rm.as<float>(),
                       want_f32 ? feat_d.as<float>() : nullptr, xb.as<unsigned short>());
    PCHK(hipGetLastError());
    PCHK(hipStreamSynchronize(st));
    return ACO_OK;
}

int aco_prior_features(int32_t n, const float *x, const float *y, const int32_t *demand,
                       int32_t capacity, const float *dist, const int32_t *nn, int32_t nn_width,
                       int32_t k, float *feat)
{
    int rc = check_instance_args(n, x, y, demand, capacity, dist, nn, nn_width, k);
    if (rc != ACO_OK) return rc;
    if (!feat) return set_error(ACO_ERR_ARG, "null feature buffer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(ACO_ERR_NODEV, "no HIP device");
    hipStream_t st;
    PCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    {
        DevBuf fd, xb, nnd;
        rc = device_features(st, n, x, y, demand, capacity, dist, nn, nn_width, k, fd, xb, nnd, true);
        if (rc == ACO_OK) {
            const hipError_t e = hipMemcpy(feat, fd.p, (size_t)n * k * NF * 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = set_error(ACO_ERR_HIP, std::string("feature readback: ") + hipGetErrorString(e));
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

int aco_prior_forward(aco_prior *net, const float *feat, int64_t count, float *logits)
{
    if (!net || !feat || !logits) return set_error(ACO_ERR_ARG, "null argument");
    if (count < 1 || count > (1ll << 31)) return set_error(ACO_ERR_ARG, "count out of range");
    DevBuf xb, out;
    int rc = stage_features(net, feat, count, xb);
    if (rc != ACO_OK) return rc;
    PCHK(dalloc(out, (size_t)count * 4));
    rc = launch_forward(net, xb.as<unsigned short>(), count, out.as<float>());
    if (rc != ACO_OK) return rc;
    PCHK(hipMemcpyAsync(logits, out.p, (size_t)count * 4, hipMemcpyDeviceToHost, net->stream));
    PCHK(hipStreamSynchronize(net->stream));
    return ACO_OK;
}

int aco_prior_gradient(aco_prior *net, const float *feat, const float *label, int64_t count,
                       float pos_weight, float *grad, float *loss)
{
    if (!net || !feat || !label || !grad) return set_error(ACO_ERR_ARG, "null argument");
    if (count < 1 || count > kMaxBatch) return set_error(ACO_ERR_ARG, "count must be in [1, 65536]");
    if (!(pos_weight > 0.0f) || !std::isfinite(pos_weight)) return set_error(ACO_ERR_ARG, "pos_weight must be positive");
    DevBuf xb, lab, idx, part, g, l;
    int rc = stage_features(net, feat, count, xb);
    if (rc != ACO_OK) return rc;
    std::vector<int> id((size_t)count);
    for (int i = 0; i < (int)count; ++i) id[i] = i;
    PCHK(dalloc(lab, (size_t)count * 4));
    PCHK(dalloc(idx, (size_t)count * 4));
    PCHK(dalloc(part, partial_bytes((int)count)));
    PCHK(dalloc(g, NP * 4));
    PCHK(dalloc(l, 16));
    PCHK(hipMemcpyAsync(lab.p, label, (size_t)count * 4, hipMemcpyHostToDevice, net->stream));
    PCHK(hipMemcpyAsync(idx.p, id.data(), (size_t)count * 4, hipMemcpyHostToDevice, net->stream));
    rc = launch_batch(net, xb.as<unsigned short>(), lab.as<float>(), idx.as<int>(), (int)count, pos_weight,
                      part.as<float>(), g.as<float>(), l.as<float>(), false, adam_args(1, 0, 0.9f, 0.999f, 1e-8f));
    if (rc != ACO_OK) return rc;
    PCHK(hipMemcpyAsync(grad, g.p, NP * 4, hipMemcpyDeviceToHost, net->stream));
    float lv = 0.0f;
    PCHK(hipMemcpyAsync(&lv, l.p, 4, hipMemcpyDeviceToHost, net->stream));
    PCHK(hipStreamSynchronize(net->stream));
    if (loss) *loss = lv;
    return ACO_OK;
}

int aco_prior_adam_step(aco_prior *net, const float *grad, float lr, float beta1, float beta2, float eps)
{
    if (!net || !grad) return set_error(ACO_ERR_ARG, "null argument");
    // a single partial row holding the gradient (loss slot zero)
    DevBuf part;
    PCHK(dalloc(part, NPL * 4));
    PCHK(hipMemsetAsync(part.p, 0, NPL * 4, net->stream));
    PCHK(hipMemcpyAsync(part.p, grad, NP * 4, hipMemcpyHostToDevice, net->stream));
    net->step += 1;
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPL + kReduceParams - 1) / kReduceParams),
                       dim3(kReduceParams * kReduceGroups), 0, net->stream, part.as<float>(), 1,
                       net->pf, net->m, net->v, net->wb, nullptr, nullptr, 1,
                       adam_args(net->step, lr, beta1, beta2, eps));
    PCHK(hipGetLastError());
    PCHK(hipStreamSynchronize(net->stream));
    return ACO_OK;
}

int aco_prior_train(aco_prior *net, const float *feat, const float *label, int64_t count,
                    const aco_prior_train_params *tp, float *epoch_loss)
{
    if (!net || !feat || !label || !tp) return set_error(ACO_ERR_ARG, "null argument");
    if (count < 1 || count >= (1ll << 31)) return set_error(ACO_ERR_ARG, "count out of range");
    if (tp->epochs < 1 || tp->batch < 16 || tp->batch > kMaxBatch)
        return set_error(ACO_ERR_ARG, "epochs >= 1 and batch in [16, 65536] required");
    if (!(tp->lr > 0.0f) || !(tp->lr_final >= 0.0f) || !(tp->pos_weight > 0.0f))
        return set_error(ACO_ERR_ARG, "lr, lr_final, pos_weight must be positive");
    DevBuf xb, lab, idx, part, losses;
    int rc = stage_features(net, feat, count, xb);
    if (rc != ACO_OK) return rc;
    const long long nb = (count + tp->batch - 1) / tp->batch;
    const long long total = nb * tp->epochs;
    PCHK(dalloc(lab, (size_t)count * 4));
    PCHK(dalloc(idx, (size_t)count * 4));
    PCHK(dalloc(part, partial_bytes(tp->batch)));
    PCHK(dalloc(losses, (size_t)total * 4));
    PCHK(hipMemcpyAsync(lab.p, label, (size_t)count * 4, hipMemcpyHostToDevice, net->stream));
    std::vector<int> perm((size_t)count);
    uint64_t s = tp->shuffle_seed;
    long long step = 0;
    for (int ep = 0; ep < tp->epochs; ++ep) {
        for (long long i = 0; i < count; ++i) perm[(size_t)i] = (int)i;
        for (long long i = count - 1; i > 0; --i) {
            const long long j = (long long)(sm_next(s) % (uint64_t)(i + 1));
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
        PCHK(hipStreamSynchronize(net->stream));  // previous epoch done with idx
        PCHK(hipMemcpyAsync(idx.p, perm.data(), (size_t)count * 4, hipMemcpyHostToDevice, net->stream));
        for (long long b = 0; b < nb; ++b, ++step) {
            const long long lo = b * tp->batch;
            const int Eb = (int)std::min<long long>(tp->batch, count - lo);
            const double frac = total > 1 ? (double)step / (double)(total - 1) : 0.0;
            const float lr = (float)(tp->lr_final + 0.5 * (tp->lr - tp->lr_final) * (1.0 + cos(M_PI * frac)));
            net->step += 1;
            rc = launch_batch(net, xb.as<unsigned short>(), lab.as<float>(), idx.as<int>() + lo, Eb,
                              tp->pos_weight, part.as<float>(), nullptr, losses.as<float>() + step, true,
                              adam_args(net->step, lr, tp->beta1, tp->beta2, tp->eps));
            if (rc != ACO_OK) return rc;
        }
    }
    std::vector<float> lh((size_t)total);
    PCHK(hipMemcpyAsync(lh.data(), losses.p, (size_t)total * 4, hipMemcpyDeviceToHost, net->stream));
    PCHK(hipStreamSynchronize(net->stream));
    if (epoch_loss)
        for (int ep = 0; ep < tp->epochs; ++ep) {
            double a = 0.0;
            for (long long b = 0; b < nb; ++b) a += lh[(size_t)(ep * nb + b)];
            epoch_loss[ep] = (float)(a / (double)nb);
        }
    return ACO_OK;
}

int aco_prior_build(aco_prior *net, int32_t n, const float *x, const float *y,
                    const int32_t *demand, int32_t capacity, const float *dist, const int32_t *nn,
                    int32_t nn_width, int32_t k, float exponent, float floor, float *prior)
{
    if (!net || !prior) return set_error(ACO_ERR_ARG, "null argument");
    int rc = check_instance_args(n, x, y, demand, capacity, dist, nn, nn_width, k);
    if (rc != ACO_OK) return rc;
    if (!(exponent >= 0.0f) || !std::isfinite(exponent)) return set_error(ACO_ERR_ARG, "exponent must be >= 0");
    if (!(floor > 0.0f && floor <= 1.0f)) return set_error(ACO_ERR_ARG, "floor must be in (0, 1]");
    DevBuf fd, xb, nnd, lg, out;
    rc = device_features(net->stream, n, x, y, demand, capacity, dist, nn, nn_width, k, fd, xb, nnd, false);
    if (rc != ACO_OK) return rc;
    const long long E = (long long)n * k, NN = (long long)n * n;
    PCHK(dalloc(lg, (size_t)E * 4));
    PCHK(dalloc(out, (size_t)NN * 4));
    rc = launch_forward(net, xb.as<unsigned short>(), E, lg.as<float>());
    if (rc != ACO_OK) return rc;
    const float fl = floor_pow(floor, exponent);
    hipLaunchKernelGGL(k_prior_fill, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, net->stream, NN, fl,
                       out.as<float>());
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(k_prior_scatter, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, net->stream, n, k,
                       nnd.as<int>(), nn_width, lg.as<float>(), exponent, out.as<float>());
    PCHK(hipGetLastError());
    PCHK(hipMemcpyAsync(prior, out.p, (size_t)NN * 4, hipMemcpyDeviceToHost, net->stream));
    PCHK(hipStreamSynchronize(net->stream));
    return ACO_OK;
}

int aco_prior_apply(aco_ctx *ctx, aco_prior *net, const float *const *x, const float *const *y,
                    int32_t k, float exponent, float floor)
{
    if (!ctx || !net || !x || !y) return set_error(ACO_ERR_ARG, "null argument");
    if (!(exponent >= 0.0f) || !std::isfinite(exponent)) return set_error(ACO_ERR_ARG, "exponent must be >= 0");
    if (!(floor > 0.0f && floor <= 1.0f)) return set_error(ACO_ERR_ARG, "floor must be in (0, 1]");
    std::vector<CtxInstanceView> views;
    void *sv = nullptr;
    int rc = ctx_prior_view(ctx, views, &sv);
    if (rc != ACO_OK) return rc;
    hipStream_t st = (hipStream_t)sv;
    int max_n = 0;
    for (size_t b = 0; b < views.size(); ++b) {
        const CtxInstanceView &v = views[b];
        if (!x[b] || !y[b]) return set_error(ACO_ERR_ARG, "null coordinates for instance " + std::to_string(b));
        if (k < 1 || k > v.n - 1) return set_error(ACO_ERR_ARG, "need 1 <= k < n");
        if (!((v.nnk && k <= v.k_ctx) || k <= v.nn_ls))
            return set_error(ACO_ERR_ARG, "k exceeds the context's candidate and neighbour lists");
        max_n = std::max(max_n, v.n);
    }
    const size_t N = (size_t)max_n, E = N * (size_t)k;
    DevBuf xd, yd, rad, ang, kd, Rm, xb, lg, dense;
    PCHK(dalloc(xd, N * 4));
    PCHK(dalloc(yd, N * 4));
    PCHK(dalloc(rad, N * 4));
    PCHK(dalloc(ang, N * 4));
    PCHK(dalloc(kd, N * 4));
    PCHK(dalloc(Rm, 16));
    PCHK(dalloc(xb, E * NFP * 2));
    PCHK(dalloc(lg, E * 4));
    PCHK(dalloc(dense, N * N * 4));
    const float fl = floor_pow(floor, exponent);
    for (size_t b = 0; b < views.size(); ++b) {
        const CtxInstanceView &v = views[b];
        const int n = v.n;
        const bool use_k = v.nnk && k <= v.k_ctx;
        const int *nn = use_k ? v.nnk : v.nnls;
        const int nnw = use_k ? v.k_ctx : v.nnP;
        const long long En = (long long)n * k, NN = (long long)n * n;
        PCHK(hipMemcpyAsync(xd.p, x[b], (size_t)n * 4, hipMemcpyHostToDevice, st));
        PCHK(hipMemcpyAsync(yd.p, y[b], (size_t)n * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_node_stats, dim3((n + 255) / 256), dim3(256), 0, st, n, v.P, xd.as<float>(),
                           yd.as<float>(), v.dist, nn, nnw, k, rad.as<float>(), ang.as<float>(), kd.as<float>());
        PCHK(hipGetLastError());
        hipLaunchKernelGGL(k_node_reduce, dim3(1), dim3(256), 0, st, n, rad.as<float>(), kd.as<float>(),
                           Rm.as<float>());
        PCHK(hipGetLastError());
        hipLaunchKernelGGL(k_edge_features, dim3((unsigned)((En + 255) / 256)), dim3(256), 0, st, n, v.P, k,
                           v.dist, nn, nnw, v.dem, (float)v.Q, rad.as<float>(), ang.as<float>(),
                           kd.as<float>(), Rm.as<float>(), nullptr, xb.as<unsigned short>());
        PCHK(hipGetLastError());
        rc = launch_forward(net, xb.as<unsigned short>(), En, lg.as<float>(), st);
        if (rc != ACO_OK) return rc;
        hipLaunchKernelGGL(k_prior_fill, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, st, NN, fl,
                           dense.as<float>());
        PCHK(hipGetLastError());
        hipLaunchKernelGGL(k_prior_scatter, dim3((unsigned)((En + 255) / 256)), dim3(256), 0, st, n, k, nn,
                           nnw, lg.as<float>(), exponent, dense.as<float>());
        PCHK(hipGetLastError());
        hipLaunchKernelGGL(k_heur_mul, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, st, n, v.P,
                           dense.as<float>(), v.heur);
        PCHK(hipGetLastError());
    }
    PCHK(hipStreamSynchronize(st));
    return ACO_OK;
}

}  // extern "C"
