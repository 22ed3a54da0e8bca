// Candidate (nearest-neighbour) lists of many instances on the GPU:
// compute_nn_lists (vrpHelper.cpp:188-236), one thread per row, the row
// sorted in LDS (k_nn_lists; at most 60 KB and 64 rows per block) or,
// for the entry point that allocates nothing, in a global scratch array.
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
#include <mutex>
#include <string>
#include <vector>

#include "aco_device.h"
#include "aco_internal.h"

namespace cvrp {

struct NnJob {
    const float *dist;   // n rows of pitch dpitch, row-major
    int *out;            // n x width
    int n, width, first_row, dpitch;
};

constexpr int kNnStack = 32;   // > log2(n) + 1 for every n the library accepts

__device__ __forceinline__ bool nn_less(float a, int ia, float b, int ib, int tie)
{
    return a < b || (tie != 0 && a == b && ia < ib);
}

// Sorts keys v[0..n) with node ids ix[0..n) (global or LDS arrays).
template <typename I>
__device__ void nn_sort_row(float *v, I *ix, int n, int tie)
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
        ix[l] = (I)pi;
        int last = l;
        for (int k = l + 1; k <= r; ++k) {
            const float vk = v[k];
            const int ik = ix[k];
            if (nn_less(vk, ik, pv, pi, tie)) {
                ++last;
                v[k] = v[last];
                ix[k] = ix[last];
                v[last] = vk;
                ix[last] = (I)ik;
            }
        }
        v[l] = v[last];
        ix[l] = ix[last];
        v[last] = pv;
        ix[last] = (I)pi;
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

// One row per thread, sorted in LDS: keys as floats, node ids as 16-bit
// values (n <= 4096), row r of the block at r * nmax.  The rows of a block
// may belong to different instances.
__global__ __launch_bounds__(64) void k_nn_lists(const NnJob *jobs, int njobs, int rows, int tie, int nmax)
{
    extern __shared__ __align__(16) unsigned char smem[];
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
    const float *d = J.dist + (size_t)node * J.dpitch;
    float *v = reinterpret_cast<float *>(smem) + (size_t)threadIdx.x * nmax;
    unsigned short *ix = reinterpret_cast<unsigned short *>(smem + (size_t)blockDim.x * nmax * 4) +
                         (size_t)threadIdx.x * nmax;
    for (int i = 0; i < n; ++i) {
        v[i] = d[i];
        ix[i] = (unsigned short)i;
    }
    v[node] = INFINITY;   // a node is not its own neighbour
    v[0] = INFINITY;      // nor is the depot anyone's
    nn_sort_row(v, ix, n, tie);
    for (int h = 0; h < J.width; ++h) J.out[(size_t)node * J.width + h] = ix[h];
}

// Rows per block of k_nn_lists for `rows` rows of at most nmax nodes: the
// rows of a wave run their sorts side by side (divergent), so a row gets a
// wave of its own while there are fewer rows than SIMDs (4 per CU), and
// rows are packed up to 64 per wave beyond that, within 60 KB of LDS per
// block (6 bytes per row element).
static int nn_rows_per_block(int rows, int nmax, int cus)
{
    const int simds = 4 * (cus > 0 ? cus : 1);
    int r = (rows + simds - 1) / simds;
    const int cap = (60 << 10) / (6 * nmax);
    r = r > cap ? cap : r;
    return r < 1 ? 1 : (r > 64 ? 64 : r);
}

static hipError_t launch_nn_lists(const NnJob *djobs, int nj, int rows, int tie, int nmax, hipStream_t st)
{
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const int rpb = nn_rows_per_block(rows, nmax, cus);
    const size_t lds = (size_t)rpb * nmax * 6;
    hipLaunchKernelGGL(k_nn_lists, dim3((rows + rpb - 1) / rpb), dim3(rpb), lds, st, djobs, nj, rows, tie, nmax);
    return hipGetLastError();
}

// ---- the same lists without device allocation ----
//
// aco_compute_nn_lists_batch_noalloc sorts in `scratch`, a file-scope device
// array of kNnScratchBytes that every call on a GPU shares (each GPU holds
// its own copy until the process ends).  A launch takes up to kNnRowJobs
// row ranges, passed by value, so nothing is copied to the device first;
// each row of an n-node instance uses 8n bytes of scratch (sort keys and
// node ids).  Every scratch word a row reads is written by that row first,
// so nothing has to start at zero.  Calls that overlap on one GPU race on
// the array.
constexpr size_t kNnScratchBytes = size_t(64) << 20;
constexpr int kNnRowJobs = 32;

static __device__ __align__(256) unsigned char scratch[kNnScratchBytes];

struct NnRowJob {
    const float *dist;   // n x n device array
    int *out;            // n x width device array
    float *v;            // rows x n scratch keys
    int *ix;             // rows x n scratch ids
    int n, width, row0, first;   // first: first thread of this job in the launch
};

struct NnRowJobs {
    NnRowJob j[kNnRowJobs];
    int count, threads;
};

__global__ __launch_bounds__(256) void k_nn_rows(NnRowJobs J, int tie)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= J.threads) return;
    int q = 0;
    while (q + 1 < J.count && J.j[q + 1].first <= t) ++q;
    const NnRowJob R = J.j[q];
    const int r = t - R.first, node = R.row0 + r, n = R.n;
    const float *d = R.dist + (size_t)node * n;
    float *v = R.v + (size_t)r * n;
    int *ix = R.ix + (size_t)r * n;
    for (int i = 0; i < n; ++i) {
        v[i] = d[i];
        ix[i] = i;
    }
    v[node] = INFINITY;
    v[0] = INFINITY;
    nn_sort_row(v, ix, n, tie);
    for (int h = 0; h < R.width; ++h) R.out[(size_t)node * R.width + h] = ix[h];
}

// Per-device address of `scratch` (the array's name is a host-side stand-in,
// not a device address), looked up once per device with that device current.
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_address(void **ptr)
{
    int dev = 0;
    const hipError_t ed = hipGetDevice(&dev);
    if (ed != hipSuccess || dev < 0 || dev >= 64)
        return set_error(ACO_ERR_NODEV, std::string("aco_compute_nn_lists_batch_noalloc: no current device: ") +
                                            hipGetErrorString(ed));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(ACO_ERR_HIP, std::string("aco_compute_nn_lists_batch_noalloc: hipGetSymbolAddress: ") +
                                              hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return ACO_OK;
}

// Instances are processed in chunks of at most this many matrix elements
// (the host path stages the distance matrices on the device).
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
        int b1 = b0, rows = 0, nmax = 0;
        while (b1 < count && (b1 == b0 || elems + (size_t)n[b1] * n[b1] <= kNnChunkElems)) {
            elems += (size_t)n[b1] * n[b1];
            outs += (size_t)n[b1] * width[b1];
            rows += n[b1];
            nmax = n[b1] > nmax ? n[b1] : nmax;
            ++b1;
        }
        const int nj = b1 - b0;
        const size_t bytes = elems * 4 + outs * 4 + (size_t)nj * sizeof(NnJob) + 1024;
        unsigned char *buf = nullptr;
        if (hipMalloc(&buf, bytes) != hipSuccess)
            return set_error(ACO_ERR_HIP, "nn lists: device allocation of " + std::to_string(bytes) + " bytes failed");
        float *ddist = reinterpret_cast<float *>(buf);
        int *dout = reinterpret_cast<int *>(ddist + elems);
        NnJob *djobs = reinterpret_cast<NnJob *>(
            (reinterpret_cast<uintptr_t>(dout + outs) + 255) & ~uintptr_t(255));
        std::vector<NnJob> jobs(nj);
        size_t eo = 0, oo = 0;
        int ro = 0;
        hipError_t err = hipSuccess;
        for (int j = 0; j < nj && err == hipSuccess; ++j) {
            const int b = b0 + j;
            jobs[j] = NnJob{ddist + eo, dout + oo, n[b], width[b], ro, n[b]};
            err = hipMemcpyAsync(ddist + eo, dist[b], (size_t)n[b] * n[b] * 4, hipMemcpyHostToDevice, st);
            eo += (size_t)n[b] * n[b];
            oo += (size_t)n[b] * width[b];
            ro += n[b];
        }
        if (err == hipSuccess)
            err = hipMemcpyAsync(djobs, jobs.data(), sizeof(NnJob) * nj, hipMemcpyHostToDevice, st);
        if (err == hipSuccess) err = launch_nn_lists(djobs, nj, rows, tie, nmax, st);
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

// The same lists from distance matrices already on the device (n[b] rows
// of pitch pitch[b]) into device arrays out[b] (n[b] x width[b]).  Enqueued
// on st and waited for.
int gpu_nn_lists_dev(hipStream_t st, int count, const int *n, const float *const *dist, const int *pitch,
                     const int *width, int tie, int *const *out)
{
    if (count == 0) return ACO_OK;
    int rows = 0, nmax = 0;
    std::vector<NnJob> jobs(count);
    for (int b = 0; b < count; ++b) {
        jobs[b] = NnJob{dist[b], out[b], n[b], width[b], rows, pitch[b]};
        rows += n[b];
        nmax = n[b] > nmax ? n[b] : nmax;
    }
    NnJob *djobs = nullptr;
    if (hipMalloc(&djobs, sizeof(NnJob) * count) != hipSuccess)
        return set_error(ACO_ERR_HIP, "nn lists: device allocation failed");
    hipError_t err = hipMemcpyAsync(djobs, jobs.data(), sizeof(NnJob) * count, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = launch_nn_lists(djobs, count, rows, tie, nmax, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (err == hipSuccess) err = es;
    (void)hipFree(djobs);
    if (err != hipSuccess) return set_error(ACO_ERR_HIP, std::string("nn lists: ") + hipGetErrorString(err));
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

extern "C" int aco_compute_nn_lists_batch_noalloc(int32_t count, const int32_t *n, const float *const *dist,
                                                  const int32_t *nn, int32_t tie, int32_t *const *out, void *stream)
{
    using namespace cvrp;
    if (count < 0 || (count > 0 && (!n || !dist || !nn || !out)) || (tie != 0 && tie != 1))
        return set_error(ACO_ERR_ARG, "aco_compute_nn_lists_batch_noalloc: bad argument");
    for (int b = 0; b < count; ++b) {
        if (n[b] < 3 || n[b] > kMaxNodes || nn[b] < 1 || nn[b] > n[b] - 2 || !dist[b] || !out[b])
            return set_error(ACO_ERR_ARG, "aco_compute_nn_lists_batch_noalloc: instance " + std::to_string(b) +
                                              " needs 3 <= n <= 4096 and 1 <= nn <= n-2");
        if ((size_t)n[b] * 8 > kNnScratchBytes)
            return set_error(ACO_ERR_ARG, "aco_compute_nn_lists_batch_noalloc: a row of instance " +
                                              std::to_string(b) + " exceeds the " +
                                              std::to_string(kNnScratchBytes) + "-byte scratch");
    }
    if (count == 0) return ACO_OK;
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char *const sbase = static_cast<unsigned char *>(base);
    // fill launches: row ranges of consecutive instances while scratch and
    // the job table last
    NnRowJobs J;
    size_t used = 0;
    J.count = 0;
    J.threads = 0;
    auto flush = [&]() -> int {
        if (J.count == 0) return ACO_OK;
        hipLaunchKernelGGL(k_nn_rows, dim3((J.threads + 255) / 256), dim3(256), 0, st, J, (int)tie);
        const hipError_t e = hipGetLastError();
        J.count = 0;
        J.threads = 0;
        used = 0;
        if (e != hipSuccess)
            return set_error(ACO_ERR_HIP, std::string("aco_compute_nn_lists_batch_noalloc: ") + hipGetErrorString(e));
        return ACO_OK;
    };
    for (int b = 0; b < count; ++b) {
        const size_t row_bytes = (size_t)n[b] * 8;
        int row = 0;
        while (row < n[b]) {
            if (J.count == kNnRowJobs || used + row_bytes > kNnScratchBytes) {
                if ((rc = flush())) return rc;
            }
            const int rows = (int)std::min<size_t>((size_t)(n[b] - row), (kNnScratchBytes - used) / row_bytes);
            NnRowJob &R = J.j[J.count++];
            R.dist = dist[b];
            R.out = out[b];
            R.v = reinterpret_cast<float *>(sbase + used);
            R.ix = reinterpret_cast<int *>(sbase + used + (size_t)rows * n[b] * 4);
            R.n = n[b];
            R.width = nn[b];
            R.row0 = row;
            R.first = J.threads;
            J.threads += rows;
            used += (size_t)rows * row_bytes;
            used = (used + 255) & ~size_t(255);
            row += rows;
        }
    }
    return flush();
}
