// Candidate (nearest-neighbour) lists of many instances on the GPU:
// compute_nn_lists (vrpHelper.cpp:188-236), one thread per row.
//
// Each row is sorted with the reference's own steps, sort2
// (utilities.cpp:246-267): the middle element is swapped to the front as
// pivot, smaller elements are swapped forward one by one, the pivot is
// swapped to the boundary, and the two parts [left, last] and
// [last + 1, right] are sorted in turn.  The parts are disjoint, so the
// order they are sorted in does not change the result: here the smaller
// part goes next and the larger waits on a stack of at most log2(n) + 1
// entries.  Rows with equal distances therefore come out in the
// reference's order (tie 0), or in index order when the index breaks ties
// (tie 1), exactly as the host builder aco_compute_nn_lists.
#include <vector>

#include "aco_device.h"
#include "aco_internal.h"

namespace cvrp {

struct NnJob {
    const float *dist;   // n x n, row-major
    float *v;            // n x n sort keys
    int *ix;             // n x n node ids
    int *out;            // n x width
    int n, width, first_row;
};

constexpr int kNnStack = 32;   // > log2(n) + 1 for every n the library accepts

__device__ __forceinline__ bool nn_less(float a, int ia, float b, int ib, int tie)
{
    return a < b || (tie != 0 && a == b && ia < ib);
}

__device__ void nn_sort_row(float *v, int *ix, int n, int tie)
{
    int sl[kNnStack], sr[kNnStack];
    int sp = 0, l = 0, r = n - 1;
    for (;;) {
        if (l >= r) {
            if (sp == 0) return;
            --sp;
            l = sl[sp];
            r = sr[sp];
            continue;
        }
        const int mid = (l + r) / 2;
        const float pv = v[mid];
        const int pi = ix[mid];
        v[mid] = v[l];
        ix[mid] = ix[l];
        v[l] = pv;
        ix[l] = pi;
        int last = l;
        for (int k = l + 1; k <= r; ++k) {
            const float vk = v[k];
            const int ik = ix[k];
            if (nn_less(vk, ik, pv, pi, tie)) {
                ++last;
                v[k] = v[last];
                ix[k] = ix[last];
                v[last] = vk;
                ix[last] = ik;
            }
        }
        v[l] = v[last];
        ix[l] = ix[last];
        v[last] = pv;
        ix[last] = pi;
        if (last - l + 1 < r - last) {
            sl[sp] = last + 1;
            sr[sp] = r;
            r = last;
        } else {
            sl[sp] = l;
            sr[sp] = last;
            l = last + 1;
        }
        ++sp;
    }
}

__global__ __launch_bounds__(256) void k_nn_lists(const NnJob *jobs, int njobs, int rows, int tie)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].first_row <= row) lo = mid;
        else hi = mid - 1;
    }
    const NnJob J = jobs[lo];
    const int node = row - J.first_row, n = J.n;
    const float *d = J.dist + (size_t)node * n;
    float *v = J.v + (size_t)node * n;
    int *ix = J.ix + (size_t)node * n;
    for (int i = 0; i < n; ++i) {
        v[i] = d[i];
        ix[i] = i;
    }
    v[node] = INFINITY;   // a node is not its own neighbour
    v[0] = INFINITY;      // nor is the depot anyone's
    nn_sort_row(v, ix, n, tie);
    for (int h = 0; h < J.width; ++h) J.out[(size_t)node * J.width + h] = ix[h];
}

// Instances are processed in chunks of at most this many matrix elements.
constexpr size_t kNnChunkElems = size_t(1) << 26;

int gpu_nn_lists(hipStream_t st, int count, const int *n, const float *const *dist, const int *width,
                 int tie, int *const *out)
{
    for (int b = 0; b < count; ++b)
        if (n[b] < 3 || n[b] > kMaxNodes || width[b] < 1 || width[b] > n[b] - 2)
            return set_error(ACO_ERR_ARG, "nn lists: need 3 <= n <= 4096 and 1 <= width <= n-2");
    int b0 = 0;
    while (b0 < count) {
        // one chunk: instances b0 .. b1-1
        size_t elems = 0, outs = 0;
        int b1 = b0, rows = 0;
        while (b1 < count && (b1 == b0 || elems + (size_t)n[b1] * n[b1] <= kNnChunkElems)) {
            elems += (size_t)n[b1] * n[b1];
            outs += (size_t)n[b1] * width[b1];
            rows += n[b1];
            ++b1;
        }
        const int nj = b1 - b0;
        const size_t bytes = elems * 12 + outs * 4 + (size_t)nj * sizeof(NnJob) + 1024;
        unsigned char *buf = nullptr;
        if (hipMalloc(&buf, bytes) != hipSuccess)
            return set_error(ACO_ERR_HIP, "nn lists: device allocation of " + std::to_string(bytes) + " bytes failed");
        float *ddist = reinterpret_cast<float *>(buf);
        float *dv = ddist + elems;
        int *dix = reinterpret_cast<int *>(dv + elems);
        int *dout = dix + elems;
        NnJob *djobs = reinterpret_cast<NnJob *>(
            (reinterpret_cast<uintptr_t>(dout + outs) + 255) & ~uintptr_t(255));
        std::vector<NnJob> jobs(nj);
        size_t eo = 0, oo = 0;
        int ro = 0;
        hipError_t err = hipSuccess;
        for (int j = 0; j < nj && err == hipSuccess; ++j) {
            const int b = b0 + j;
            jobs[j] = NnJob{ddist + eo, dv + eo, dix + eo, dout + oo, n[b], width[b], ro};
            err = hipMemcpyAsync(ddist + eo, dist[b], (size_t)n[b] * n[b] * 4, hipMemcpyHostToDevice, st);
            eo += (size_t)n[b] * n[b];
            oo += (size_t)n[b] * width[b];
            ro += n[b];
        }
        if (err == hipSuccess)
            err = hipMemcpyAsync(djobs, jobs.data(), sizeof(NnJob) * nj, hipMemcpyHostToDevice, st);
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_nn_lists, dim3((rows + 255) / 256), dim3(256), 0, st, djobs, nj, rows, tie);
            err = hipGetLastError();
        }
        oo = 0;
        for (int j = 0; j < nj && err == hipSuccess; ++j) {
            const int b = b0 + j;
            err = hipMemcpyAsync(out[b], dout + oo, (size_t)n[b] * width[b] * 4, hipMemcpyDeviceToHost, st);
            oo += (size_t)n[b] * width[b];
        }
        const hipError_t es = hipStreamSynchronize(st);
        if (err == hipSuccess) err = es;
        (void)hipFree(buf);
        if (err != hipSuccess) return set_error(ACO_ERR_HIP, std::string("nn lists: ") + hipGetErrorString(err));
        b0 = b1;
    }
    return ACO_OK;
}

}  // namespace cvrp

extern "C" int aco_compute_nn_lists_batch(int32_t count, const int32_t *n, const float *const *dist,
                                          const int32_t *nn, int32_t tie, int32_t *const *out)
{
    using namespace cvrp;
    if (count < 0 || (count > 0 && (!n || !dist || !nn || !out)) || (tie != 0 && tie != 1))
        return set_error(ACO_ERR_ARG, "aco_compute_nn_lists_batch: bad argument");
    for (int b = 0; b < count; ++b) {
        if (n[b] < 3 || n[b] > kMaxNodes || nn[b] < 1 || nn[b] > n[b] - 2 || !dist[b] || !out[b])
            return set_error(ACO_ERR_ARG, "aco_compute_nn_lists_batch: instance " + std::to_string(b) +
                                              " needs 3 <= n <= 4096 and 1 <= nn <= n-2");
    }
    if (count == 0) return ACO_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(ACO_ERR_NODEV, "no HIP device");
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return set_error(ACO_ERR_HIP, "aco_compute_nn_lists_batch: stream creation failed");
    const int rc = gpu_nn_lists(st, count, n, dist, nn, tie, out);
    (void)hipStreamDestroy(st);
    return rc;
}
