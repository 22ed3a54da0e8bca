// Instance set-up on the device (the step before the loop: init_problem ->
// compute_distances / compute_nn_lists, vrpHelper.cpp:160-236): the
// distance matrix and the candidate lists are uploaded or built once, and
// every derived array is filled here instead of on the host.
//
//   k_setup_rows  eta^beta (beta 1 or 2, times an optional prior already
//                 uploaded into heur), the depot column d0 and the 16-bit
//                 copy of the distances
//   k_setup_nn    the candidate rows (nnk and their distances), the 2-opt
//                 neighbour rows (nnls / nnd, padded), and nnpos (position
//                 of j in i's neighbour row, 0xFFFF if absent)
//   k_candidate_dists
//                 distances of caller-given construction candidate rows
//   k_distances   distance matrices from coordinates (aco_compute_distances
//                 for the EUC_2D, nint, CEIL_2D and ATT conventions)
//   k_tau_from_subs, k_ant_length
//                 write-backs of the sub-problem decomposition: sub-colony
//                 pheromone into the master, the length of an injected ant
//
// Each value is the same float expression the host preparation and the
// CPU oracle evaluate (no contraction, IEEE division), so the arrays are
// bit-identical to the host-built ones.
#include <string>
#include <vector>

#include "aco_device.h"
#include "aco_internal.h"

namespace cvrp {

__global__ __launch_bounds__(256) void k_setup_rows(DevProblem p, int beta_code, int has_prior, int int16)
{
    const int i = blockIdx.y;
    float *heur = const_cast<float *>(p.heur) + (size_t)i * p.P;
    unsigned short *d16 = const_cast<unsigned short *>(p.dist16) + (size_t)i * p.P;
    const float *drow = p.dist + (size_t)i * p.P;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < p.n; j += gridDim.x * 256) {
        const float d = drow[j];
        if (beta_code) {
            const float y = 1.0f / (d + 0.1f);
            float h = beta_code == 2 ? y * y : y;
            if (has_prior) h = h * heur[j];
            heur[j] = h;
        }
        if (int16) d16[j] = (unsigned short)d;
        if (j == 0) const_cast<float *>(p.d0)[i] = d;
    }
}

__global__ __launch_bounds__(256) void k_setup_nn(DevProblem p, const int *nnl, int width)
{
    const int i = blockIdx.x, n = p.n;
    const int *row = nnl + (size_t)i * width;
    unsigned short *pos = const_cast<unsigned short *>(p.nnpos) + (size_t)i * n;
    for (int j = threadIdx.x; j < n; j += 256) pos[j] = 0xFFFF;
    __syncthreads();
    const float *drow = p.dist + (size_t)i * p.P;
    int *ls = const_cast<int *>(p.nnls) + (size_t)i * p.nnP;
    float *ld = const_cast<float *>(p.nnd) + (size_t)i * p.nnP;
    for (int h = threadIdx.x; h < p.nnP; h += 256) {
        if (h < p.nn_ls) {
            const int v = row[h];
            ls[h] = v;
            ld[h] = drow[v];
            pos[v] = (unsigned short)h;
        } else {
            ls[h] = 0;
            ld[h] = INFINITY;
        }
    }
    int *kk = const_cast<int *>(p.nnk) + (size_t)i * (p.k > 0 ? p.k : 1);
    float *kd = const_cast<float *>(p.nnkd) + (size_t)i * (p.k > 0 ? p.k : 1);
    for (int h = threadIdx.x; h < p.k; h += 256) {
        const int v = row[h];
        kk[h] = v;
        kd[h] = drow[v];
    }
}

hipError_t launch_setup_rows(const DevProblem &p, int beta_code, int has_prior, int int16, hipStream_t st)
{
    const dim3 grid((p.n + 255) / 256, p.n);
    hipLaunchKernelGGL(k_setup_rows, grid, dim3(256), 0, st, p, beta_code, has_prior, int16);
    return hipGetLastError();
}

hipError_t launch_setup_nn(const DevProblem &p, const int *nnl, int width, hipStream_t st)
{
    hipLaunchKernelGGL(k_setup_nn, dim3(p.n), dim3(256), 0, st, p, nnl, width);
    return hipGetLastError();
}

// Distances of the construction candidate rows after the caller replaced
// their ids (aco_set_candidates_at): nnkd[i][c] = d(i, nnk[i][c]); the 2-opt
// rows and nnpos are not touched.
__global__ __launch_bounds__(256) void k_candidate_dists(DevProblem p)
{
    const int i = blockIdx.x;
    const float *drow = p.dist + (size_t)i * p.P;
    const int *kk = p.nnk + (size_t)i * p.k;
    float *kd = const_cast<float *>(p.nnkd) + (size_t)i * p.k;
    for (int c = threadIdx.x; c < p.k; c += 256) kd[c] = drow[kk[c]];
}

hipError_t launch_candidate_dists(const DevProblem &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_candidate_dists, dim3(p.n), dim3(256), 0, st, p);
    return hipGetLastError();
}

// ------------------------------------------- decomposition write-backs ----

// Pheromone of finished sub-colonies into the master instance: one block per
// sub-problem row i, tau[r_i][r_j] = sub tau[i][j] (j != i) and total_info
// there as the pheromone pass computes it, tau^alpha * heur.
__global__ __launch_bounds__(256) void k_tau_from_subs(const SubTauJob *jobs, int njobs, DevProblem p, DevState s)
{
    const int row = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].first_row <= row) lo = mid;
        else hi = mid - 1;
    }
    const SubTauJob J = jobs[lo];
    const int i = row - J.first_row;
    const size_t base = (size_t)J.real[i] * p.P;
    const float *src = J.tau + (size_t)i * J.P;
    for (int j = threadIdx.x; j < J.n; j += 256) {
        if (j == i) continue;
        const size_t e = base + J.real[j];
        const float t = src[j];
        const float ta = p.alpha == 1.0f ? t : powf(t, p.alpha);
        s.tau[e] = t;
        s.tinfo[e] = ta * p.heur[e];
    }
}

hipError_t launch_tau_from_subs(const SubTauJob *jobs, int njobs, int rows, const DevProblem &master,
                                const DevState &ms, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tau_from_subs, dim3(rows), dim3(256), 0, st, jobs, njobs, master, ms);
    return hipGetLastError();
}

// Step-order length of one ant's tour (one thread: the additions are
// sequential, as in construction).
__global__ void k_ant_length(DevProblem p, DevState s, int ant)
{
    if (threadIdx.x != 0) return;
    const int *t = s.tours + (size_t)ant * p.T;
    const int size = s.sizes[ant];
    float len = 0.0f;
    for (int q = 1; q < size; ++q) len = len + p.dist[(size_t)t[q - 1] * p.P + t[q]];
    s.lens[ant] = len;
}

hipError_t launch_ant_length(const DevProblem &p, const DevState &s, int ant, hipStream_t st)
{
    hipLaunchKernelGGL(k_ant_length, dim3(1), dim3(64), 0, st, p, s, ant);
    return hipGetLastError();
}

// ------------------------------------------------------------ distances ----

struct DistJob {
    const double *x, *y;
    float *dist;     // n x n
    int n, first;    // first: first row of this job in the launch
};

// compute_distances (vrpHelper.cpp:160-184) with the conventions of
// distance_of (aco_host.cpp): float sqrt unrounded, nint of the double
// distance, ceil of the float sqrt, ATT's pseudo-Euclidean rounding.  One
// thread per (row, column).
__global__ __launch_bounds__(256) void k_distances(const DistJob *jobs, int njobs, int rows, int mode)
{
    const int row = blockIdx.y * 65535 + blockIdx.x;
    if (row >= rows) return;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].first <= row) lo = mid;
        else hi = mid - 1;
    }
    const DistJob J = jobs[lo];
    const int i = row - J.first, n = J.n;
    const double xi_ = J.x[i], yi_ = J.y[i];
    const float xi = (float)xi_, yi = (float)yi_;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double xj_ = J.x[j], yj_ = J.y[j];
        const float xj = (float)xj_, yj = (float)yj_;
        float d;
        if (mode == kDistNint) {
            const double xd = xi_ - xj_, yd = yi_ - yj_;
            d = (float)(long long)(sqrt(xd * xd + yd * yd) + 0.5);
        } else if (mode == kDistCeil) {
            const float xd = xi - xj, yd = yi - yj;
            d = (float)(int)ceilf(sqrtf(xd * xd + yd * yd));
        } else if (mode == kDistAtt) {
            const float xd = xi - xj, yd = yi - yj;
            const float r = (float)sqrt((double)(xd * xd + yd * yd) / 10.0);
            const float t = (float)(int)r;
            d = (float)(t < r ? (int)t + 1 : (int)t);
        } else {
            const float xd = xi - xj, yd = yi - yj;
            d = sqrtf(xd * xd + yd * yd);
        }
        J.dist[(size_t)i * n + j] = d;
    }
}

}  // namespace cvrp

extern "C" int aco_compute_distances_batch(int32_t count, const int32_t *n, const double *const *x,
                                           const double *const *y, int32_t mode, float *const *dist)
{
    using namespace cvrp;
    if (count < 0 || (count > 0 && (!n || !x || !y || !dist)))
        return set_error(ACO_ERR_ARG, "aco_compute_distances_batch: bad argument");
    if (mode != kDistFloat && mode != kDistNint && mode != kDistCeil && mode != kDistAtt)
        return set_error(ACO_ERR_ARG, "aco_compute_distances_batch: EUC_2D, nint, CEIL_2D or ATT "
                                      "(GEO distances: aco_compute_distances)");
    size_t cells = 0, coords = 0;
    int rows = 0;
    for (int b = 0; b < count; ++b) {
        if (n[b] < 1 || n[b] > kMaxNodes || !x[b] || !y[b] || !dist[b])
            return set_error(ACO_ERR_ARG, "aco_compute_distances_batch: instance " + std::to_string(b) +
                                              " needs 1 <= n <= 4096 and its arrays");
        cells += (size_t)n[b] * n[b];
        coords += (size_t)n[b];
        rows += n[b];
    }
    if (count == 0) return ACO_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(ACO_ERR_NODEV, "no HIP device");
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return set_error(ACO_ERR_HIP, "aco_compute_distances_batch: stream creation failed");
    const size_t bytes = cells * 4 + coords * 16 + (size_t)count * sizeof(DistJob) + 1024;
    unsigned char *buf = nullptr;
    hipError_t err = hipMalloc(&buf, bytes);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(st);
        return set_error(ACO_ERR_HIP, "aco_compute_distances_batch: device allocation failed");
    }
    float *dd = reinterpret_cast<float *>(buf);
    double *dx = reinterpret_cast<double *>((reinterpret_cast<uintptr_t>(dd + cells) + 15) & ~uintptr_t(15));
    double *dy = dx + coords;
    DistJob *djobs = reinterpret_cast<DistJob *>((reinterpret_cast<uintptr_t>(dy + coords) + 255) & ~uintptr_t(255));
    std::vector<DistJob> jobs(count);
    size_t co = 0, cc = 0;
    int ro = 0;
    for (int b = 0; b < count && err == hipSuccess; ++b) {
        jobs[b] = DistJob{dx + co, dy + co, dd + cc, n[b], ro};
        err = hipMemcpyAsync(dx + co, x[b], (size_t)n[b] * 8, hipMemcpyHostToDevice, st);
        if (err == hipSuccess) err = hipMemcpyAsync(dy + co, y[b], (size_t)n[b] * 8, hipMemcpyHostToDevice, st);
        co += (size_t)n[b];
        cc += (size_t)n[b] * n[b];
        ro += n[b];
    }
    if (err == hipSuccess) err = hipMemcpyAsync(djobs, jobs.data(), sizeof(DistJob) * count, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) {
        const dim3 grid(rows < 65535 ? rows : 65535, (rows + 65534) / 65535);
        hipLaunchKernelGGL(k_distances, grid, dim3(256), 0, st, djobs, count, rows, (int)mode);
        err = hipGetLastError();
    }
    cc = 0;
    for (int b = 0; b < count && err == hipSuccess; ++b) {
        err = hipMemcpyAsync(dist[b], dd + cc, (size_t)n[b] * n[b] * 4, hipMemcpyDeviceToHost, st);
        cc += (size_t)n[b] * n[b];
    }
    const hipError_t es = hipStreamSynchronize(st);
    if (err == hipSuccess) err = es;
    (void)hipFree(buf);
    (void)hipStreamDestroy(st);
    if (err != hipSuccess)
        return set_error(ACO_ERR_HIP, std::string("aco_compute_distances_batch: ") + hipGetErrorString(err));
    return ACO_OK;
}
