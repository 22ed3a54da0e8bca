// Local search kernel (gfx950): one block of W wavefronts per ant; the
// search itself is local_search_ant (aco_ls.h).  Batches launch one grid
// per instance-size class, each with LDS for its own largest n.
#include "aco_ls.h"

namespace cvrp {

size_t local_search_lds_bytes(int n, int ring, int row16, int flags)
{
    return ls_offsets(n, ring, row16 ? 2 : 4, (flags & kLs2OptRandom) != 0).total;
}

// With 4 or 8 waves per ant several ants share a CU: registers are held to
// 128 per lane so that four waves fit each SIMD (not for the random-order
// 2-opt instantiations, which would spill).
template <int W, typename RT, bool R2>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(W <= 8 && !R2 ? 4 : 1)))
void k_local_search(DevBatch bt, int first, int flags, int ring, int prefetch)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int g = bt.ls_order[first + blockIdx.x];   // this launch's size class
    const int inst = bt.ant_inst[g];
    // copies in registers, read once before the first barrier: fields read
    // through the global records would be loaded again after every barrier,
    // and each such load waits for all outstanding memory operations,
    // in-flight row copies included
    const DevProblem p = bt.probs[inst];
    const DevState s = bt.states[inst];
    local_search_ant<W, RT, R2>(p, s, g - bt.ant_base[inst], flags, ring, prefetch, smem);
}

// Largest static LDS (the kernel's own __shared__ arrays, beside the
// dynamic allocation) of the local-search kernels.
hipError_t local_search_static_lds(size_t &bytes)
{
    const void *fns[] = {(const void *)k_local_search<4, float, false>,
                         (const void *)k_local_search<4, unsigned short, false>,
                         (const void *)k_local_search<8, float, false>,
                         (const void *)k_local_search<8, unsigned short, false>,
                         (const void *)k_local_search<12, float, false>,
                         (const void *)k_local_search<12, unsigned short, false>,
                         (const void *)k_local_search<16, float, false>,
                         (const void *)k_local_search<16, unsigned short, false>,
                         (const void *)k_local_search<4, float, true>,
                         (const void *)k_local_search<4, unsigned short, true>,
                         (const void *)k_local_search<8, float, true>,
                         (const void *)k_local_search<8, unsigned short, true>,
                         (const void *)k_local_search<12, float, true>,
                         (const void *)k_local_search<12, unsigned short, true>,
                         (const void *)k_local_search<16, float, true>,
                         (const void *)k_local_search<16, unsigned short, true>};
    bytes = 0;
    for (const void *f : fns) {
        hipFuncAttributes a;
        const hipError_t e = hipFuncGetAttributes(&a, f);
        if (e != hipSuccess) return e;
        bytes = a.sharedSizeBytes > bytes ? a.sharedSizeBytes : bytes;
    }
    return hipSuccess;
}

// After a local search with the random-order 2-opt: each instance's
// scan-order stream advances past the draws its ants made (size - 1 each).
__global__ __launch_bounds__(256) void k_ls_seed_advance(DevBatch bt)
{
    __shared__ long long red[4];
    const DevProblem p = bt.probs[blockIdx.x];
    const DevState s = bt.states[blockIdx.x];
    long long t = 0;
    for (int a = threadIdx.x; a < p.m; a += 256) t += s.sizes[a] - 1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0)
        s.ctrl[kCtrlLsSeed] = pm_jump(s.ctrl[kCtrlLsSeed], (unsigned long long)(red[0] + red[1] + red[2] + red[3]));
}

bool ls_shape_valid(const LsShape &sh)
{
    if (sh.waves != 4 && sh.waves != 8 && sh.waves != 12 && sh.waves != 16) return false;
    if (sh.ring != 8 && sh.ring != 16 && sh.ring != 32) return false;
    if (sh.ring < sh.waves + 2) return false;
    if (sh.prefetch && sh.ring < 2 * sh.waves + 2) return false;
    return true;
}

hipError_t launch_local_search(const DevBatch &b, int flags, const LsShape &sh, hipStream_t st,
                               const LsStreams *par)
{
    if (!flags) return hipSuccess;
    if (!ls_shape_valid(sh)) return hipErrorInvalidValue;
    // one launch per size class, largest instances first, each with LDS for
    // its own largest n.  The classes search disjoint ants, so with extra
    // streams they run side by side (one launch's tail overlaps the next
    // class) and join the launch stream afterwards.
    const bool fork = par && par->count >= b.ls_classes - 1 && b.ls_classes > 1;
    if (fork) {
        const hipError_t e = hipEventRecord(par->fork, st);
        if (e != hipSuccess) return e;
    }
    for (int cl = 0; cl < b.ls_classes; ++cl) {
        const int first = b.ls_first[cl], count = b.ls_first[cl + 1] - first;
        const size_t lds = local_search_lds_bytes(b.ls_maxn[cl], sh.ring, sh.row16, flags);
        const hipStream_t s = fork && cl > 0 ? par->aux[cl - 1] : st;
        if (s != st) {
            const hipError_t e = hipStreamWaitEvent(s, par->fork, 0);
            if (e != hipSuccess) return e;
        }
#define KLS(W, RT, R2) hipLaunchKernelGGL((k_local_search<W, RT, R2>), dim3(count), dim3(W * kLanes), lds, s, b, \
                                          first, flags, sh.ring, sh.prefetch)
#define KLS2(W, RT) do { if (flags & kLs2OptRandom) KLS(W, RT, true); else KLS(W, RT, false); } while (0)
        switch (sh.waves) {
        case 16: if (sh.row16) KLS2(16, unsigned short); else KLS2(16, float); break;
        case 12: if (sh.row16) KLS2(12, unsigned short); else KLS2(12, float); break;
        case 8: if (sh.row16) KLS2(8, unsigned short); else KLS2(8, float); break;
        default: if (sh.row16) KLS2(4, unsigned short); else KLS2(4, float); break;
        }
#undef KLS2
#undef KLS
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (s != st) {
            e = hipEventRecord(par->join[cl - 1], s);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, par->join[cl - 1], 0);
            if (e != hipSuccess) return e;
        }
    }
    if (flags & kLs2OptRandom) {
        hipLaunchKernelGGL(k_ls_seed_advance, dim3(b.B), dim3(256), 0, st, b);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace cvrp
