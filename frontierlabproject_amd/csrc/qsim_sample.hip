// Sampling measurement outcomes from |psi|^2 in two levels, all in FP64:
// (1) per-block probability sums over blocks of 2^kSampleB amplitudes;
// (2) the host draws each shot's block from the block sums and a residual
//     inside it; a block with shots builds its own prefix sums in LDS and
//     binary-searches every residual.
#include "qsim_internal.h"
#include "qsim_sample.h"

namespace qsim {

template <typename R>
__global__ __launch_bounds__(kTPB) void k_block_prob(const cx<R> *__restrict__ psi, uint64_t dim,
                                                     double *__restrict__ out)
{
    __shared__ double sh[kWaves];
    const uint64_t b0 = (uint64_t)blockIdx.x << kSampleB;
    const uint64_t len = dim - b0 < (1ull << kSampleB) ? dim - b0 : (1ull << kSampleB);
    NSum acc;
    for (uint64_t i = threadIdx.x; i < len; i += kTPB) {
        const cx<R> v = psi[b0 + i];
        acc.add((double)v.re * v.re + (double)v.im * v.im);
    }
    const double t = block_sum(acc.value(), sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// One workgroup per block that received shots.  Inclusive prefix sums of the
// block's probabilities (FP64, in LDS), then for shot s of this block the
// first index whose prefix sum exceeds residual[s].
template <typename R>
__global__ __launch_bounds__(kTPB) void k_sample_blocks(const cx<R> *__restrict__ psi, uint64_t dim,
                                                        const SampleGroup *__restrict__ groups,
                                                        const double *__restrict__ residual,
                                                        uint64_t *__restrict__ index_out)
{
    constexpr int B = 1 << kSampleB;
    constexpr int PER = B / kTPB;  // consecutive amplitudes per thread
    __shared__ double cdf[B];
    __shared__ double part[kTPB];
    const SampleGroup g = groups[blockIdx.x];
    const uint64_t b0 = g.block << kSampleB;
    const int len = (int)(dim - b0 < (uint64_t)B ? dim - b0 : (uint64_t)B);
    // per-thread running sums over PER consecutive amplitudes
    double run = 0.0;
    const int e0 = threadIdx.x * PER;
    for (int k = 0; k < PER; ++k) {
        const int e = e0 + k;
        double p = 0.0;
        if (e < len) {
            const cx<R> v = psi[b0 + e];
            p = (double)v.re * v.re + (double)v.im * v.im;
        }
        run += p;
        cdf[e] = run;
    }
    part[threadIdx.x] = run;
    __syncthreads();
    // exclusive scan of the per-thread totals (Hillis-Steele over kTPB)
    for (int off = 1; off < kTPB; off <<= 1) {
        const double add = threadIdx.x >= off ? part[threadIdx.x - off] : 0.0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    const double before = threadIdx.x ? part[threadIdx.x - 1] : 0.0;
    for (int k = 0; k < PER; ++k) cdf[e0 + k] += before;
    __syncthreads();
    const double top = cdf[len - 1];
    for (uint32_t s = threadIdx.x; s < g.count; s += kTPB) {
        const double u = residual[g.first + s];
        // first e with cdf[e] > u; when rounding leaves none, the first e with
        // cdf[e] == top, the last amplitude with nonzero probability
        const bool over = !(u < top);
        int lo = 0, hi = len - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (over ? cdf[mid] >= top : cdf[mid] > u) hi = mid;
            else lo = mid + 1;
        }
        index_out[g.first + s] = b0 + (uint64_t)lo;
    }
}

template <typename R>
hipError_t launch_block_prob(const cx<R> *psi, uint64_t dim, double *out, hipStream_t s)
{
    const uint64_t nb = (dim + (1ull << kSampleB) - 1) >> kSampleB;
    hipLaunchKernelGGL(k_block_prob<R>, dim3((unsigned)nb), dim3(kTPB), 0, s, psi, dim, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_sample_blocks(const cx<R> *psi, uint64_t dim, const SampleGroup *groups, int ngroups,
                                const double *residual, uint64_t *index_out, hipStream_t s)
{
    if (ngroups < 1) return hipSuccess;
    hipLaunchKernelGGL(k_sample_blocks<R>, dim3((unsigned)ngroups), dim3(kTPB), 0, s, psi, dim, groups, residual,
                       index_out);
    return hipGetLastError();
}

template hipError_t launch_block_prob<float>(const cx<float> *, uint64_t, double *, hipStream_t);
template hipError_t launch_block_prob<double>(const cx<double> *, uint64_t, double *, hipStream_t);
template hipError_t launch_sample_blocks<float>(const cx<float> *, uint64_t, const SampleGroup *, int,
                                                const double *, uint64_t *, hipStream_t);
template hipError_t launch_sample_blocks<double>(const cx<double> *, uint64_t, const SampleGroup *, int,
                                                 const double *, uint64_t *, hipStream_t);

}  // namespace qsim
