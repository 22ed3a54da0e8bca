// LDS read-bandwidth microbenchmark (gfx950), the measured roof of the
// LDS-bound local search.  Built as libcvrpprobe.so, separate from the
// solver library; bench.py loads it for the roofline's peak.
//
// Every CU runs `blocks_per_cu` blocks of 256 threads (4 waves); a block
// fills 64 KB of LDS, then each thread issues 16-byte LDS reads
// (ds_read_b128) in trips of 16: at trip t the lanes of wave w read chunks
// lane + 1024 ((t + w) mod 4) + 64 k, k = 0..15, so each read of a wave
// covers 64 consecutive 16-byte chunks (1 KB, every bank equally: the
// minimum for 16-byte reads), the 16 reads of a trip differ only by an
// immediate offset (no address arithmetic) and the window moves every
// trip.  The values are XOR-folded (two 3-input bit operations per read)
// into one word per thread and written out so no read can be dropped.  Bytes = blocks x 256 x 16 x trips x 16;
// time from HIP events around the launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;
constexpr int kChunks = 4096;   // 64 KB of uint4

__global__ __launch_bounds__(kThreads) void k_lds_read(uint32_t *out, int trips)
{
    __shared__ uint4 buf[kChunks];
    const int tid = threadIdx.x;
    for (int c = tid; c < kChunks; c += kThreads)
        buf[c] = make_uint4(c * 4u + blockIdx.x, c * 4u + 1u, c * 4u + 2u, c * 4u + 3u);
    __syncthreads();
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint32_t acc = 0;
    for (int t = 0; t < trips; ++t) {
        const uint4 *win = buf + lane + 1024 * ((t + wave) & 3);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint4 v = win[64 * k];
            // two 3-input XORs (v_bitop3_b32, table 0x96) per read, so the
            // VALU work per read stays below the LDS's own rate
            acc = __builtin_amdgcn_bitop3_b32(acc, v.x, v.y, 0x96);
            acc = __builtin_amdgcn_bitop3_b32(acc, v.z, v.w, 0x96);
        }
    }
    out[(size_t)blockIdx.x * kThreads + tid] = acc;
}

}  // namespace

// Aggregate LDS read bandwidth in GB/s (best of `reps` launches, each
// thread making `trips` trips of 16 reads) on the current device; 0 on
// success.
extern "C" int probe_lds_bandwidth(int reps, int trips, int blocks_per_cu, double *gbps, double *ms_best)
{
    if (reps < 1 || trips < 1 || blocks_per_cu < 1 || !gbps) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -2;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
    const int blocks = cus * blocks_per_cu;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * kThreads * 4) != hipSuccess) return -2;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    float best = 1e30f;
    int rc = 0;
    for (int r = 0; r <= reps && rc == 0; ++r) {   // first launch warms up
        (void)hipEventRecord(a, 0);
        hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(kThreads), 0, 0, out, trips);
        (void)hipEventRecord(b, 0);
        if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess) rc = -2;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (r > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(out);
    if (rc) return rc;
    const double bytes = (double)blocks * kThreads * (double)trips * 16.0 * 16.0;
    *gbps = bytes / (best * 1e-3) / 1e9;
    if (ms_best) *ms_best = best;
    return 0;
}
