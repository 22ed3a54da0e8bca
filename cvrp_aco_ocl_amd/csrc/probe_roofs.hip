// Microbenchmarks of two more roofs of this device, built into
// libcvrpprobe.so beside the LDS probe: the HBM copy bandwidth (the roof
// of the streaming pheromone pass) and the dense BF16 matrix-core rate
// (the roof of the learned prior's MLP).  bench.py reports them next to
// the published MI355X figures.
//
//   k_copy       grid-stride copy of 16-byte elements from one buffer to
//                another (each byte read once and written once), 4 blocks
//                of 256 threads per CU; bytes = 2 x buffer size per launch.
//                Of the variants tried (1, 2 or 4 elements per thread per
//                trip, plain or streaming accesses, 4 to 32 blocks per CU,
//                and hipMemcpy device to device) this one moved the most.
//   k_mfma_bf16  every wave runs `trips` trips of 4 independent MFMA chains
//                on register operands (no memory traffic in the loop):
//                shape 0 = v_mfma_f32_16x16x32_bf16 (16384 FLOP each),
//                shape 1 = v_mfma_f32_32x32x16_bf16 (32768 FLOP each).
//                The accumulators are folded into one value per lane and
//                written out, so no product is dropped.
// Times come from HIP events around each launch; the best of `reps`
// launches after a warm-up launch is kept.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void k_copy(const v4u *src, v4u *dst, size_t count)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

// Operands from the thread index (small BF16 values, so the sums stay finite).
__device__ __forceinline__ v8bf operand(int seed)
{
    v8s v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (short)(0x3C00 + ((seed + 3 * j) & 63));   // about 0.0078 .. 0.0154
    return __builtin_bit_cast(v8bf, v);
}

template <int SHAPE>
__global__ __launch_bounds__(kThreads) void k_mfma_bf16(float *out, int trips)
{
    const int tid = threadIdx.x;
    const v8bf a = operand(tid), b = operand(tid * 7 + 1);
    float fold = 0.0f;
    // trips >= 1 (checked by the host), so the loops are do-while
    if constexpr (SHAPE == 0) {
        // starting values from the thread index: with all-zero ones the
        // compiler splits the accumulators and moves them every trip
        const float z0 = (float)(tid & 1);
        v4f c0 = {z0, z0, z0, z0}, c1 = c0 + 1.0f, c2 = c0 + 2.0f, c3 = c0 + 3.0f;
        int t = 0;
        do {
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c3, 0, 0, 0);
        } while (++t < trips);
        const v4f s = (c0 + c1) + (c2 + c3);
        fold = (s[0] + s[1]) + (s[2] + s[3]);
    } else {
        v16f c0 = {}, c1 = {}, c2 = {}, c3 = {};
        int t = 0;
        do {
            c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c3, 0, 0, 0);
        } while (++t < trips);
        const v16f s = (c0 + c1) + (c2 + c3);
#pragma unroll
        for (int j = 0; j < 16; ++j) fold += s[j];
    }
    out[(size_t)blockIdx.x * kThreads + tid] = fold;
}

// Best-of-`reps` time of launch() after one warm-up launch, in ms; <0 on a
// HIP error.
template <typename F>
float best_ms(int reps, F &&launch)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0f;
    }
    float best = 1e30f;
    for (int r = 0; r <= reps; ++r) {
        (void)hipEventRecord(a, 0);
        launch();
        (void)hipEventRecord(b, 0);
        if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess) {
            best = -1.0f;
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (r > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

int device_cus(int &cus)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -2;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
    return 0;
}

}  // namespace

// HBM copy bandwidth in GB/s (read + write bytes over the best launch) for
// two buffers of `mib` MiB on the current device; 0 on success.
extern "C" int probe_hbm_copy(int reps, int mib, double *gbps, double *ms_best)
{
    if (reps < 1 || mib < 1 || mib > 16384 || !gbps) return -1;
    int cus = 0;
    if (device_cus(cus)) return -2;
    const size_t bytes = (size_t)mib << 20, count = bytes / sizeof(v4u);
    v4u *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess) return -2;
    if (hipMalloc(&dst, bytes) != hipSuccess) {
        (void)hipFree(src);
        return -2;
    }
    int rc = hipMemset(src, 1, bytes) == hipSuccess ? 0 : -2;
    const dim3 grid(cus * 4), block(kThreads);
    const float ms = rc ? -1.0f : best_ms(reps, [&] { hipLaunchKernelGGL(k_copy, grid, block, 0, 0, src, dst, count); });
    (void)hipFree(src);
    (void)hipFree(dst);
    if (ms < 0.0f) return -2;
    *gbps = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    if (ms_best) *ms_best = ms;
    return 0;
}

// Dense BF16 matrix-core rate in TFLOP/s (shape 0: 16x16x32, 1: 32x32x16)
// with `blocks_per_cu` blocks of 4 waves per CU; 0 on success.
extern "C" int probe_mfma_bf16(int reps, int trips, int shape, int blocks_per_cu, double *tflops, double *ms_best)
{
    if (reps < 1 || trips < 1 || (shape != 0 && shape != 1) || blocks_per_cu < 1 || blocks_per_cu > 8 || !tflops)
        return -1;
    int cus = 0;
    if (device_cus(cus)) return -2;
    const int blocks = cus * blocks_per_cu;
    float *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * kThreads * sizeof(float)) != hipSuccess) return -2;
    const dim3 grid(blocks), block(kThreads);
    const float ms = best_ms(reps, [&] {
        if (shape == 0) hipLaunchKernelGGL(k_mfma_bf16<0>, grid, block, 0, 0, out, trips);
        else hipLaunchKernelGGL(k_mfma_bf16<1>, grid, block, 0, 0, out, trips);
    });
    (void)hipFree(out);
    if (ms < 0.0f) return -2;
    const double per_wave_trip = shape == 0 ? 4.0 * 16384.0 : 4.0 * 32768.0;
    const double flop = (double)blocks * (kThreads / 64) * (double)trips * per_wave_trip;
    *tflops = flop / (ms * 1e-3) / 1e12;
    if (ms_best) *ms_best = ms;
    return 0;
}
