// Persistent batch kernel (see qsim_batch.h).
#include "qsim_batch.h"

namespace qsim {

namespace {

__device__ __forceinline__ float zsign(uint64_t idx, uint64_t zmask)
{
    return (__popcll(idx & zmask) & 1) ? -1.f : 1.f;
}

// Sum of the kWaves per-wave partials of each probe of pass pd into its
// output slot (this workgroup owns the circuit: plain stores).
__device__ __forceinline__ void flush_probes(const PassDev *pd, const int32_t *__restrict__ probe_ids,
                                             const double *pacc, double *__restrict__ out)
{
    for (int q = threadIdx.x; q < pd->nprobe; q += kTPB) {
        double v = 0.0;
        for (int w = 0; w < kWaves; ++w) v += pacc[q * kWaves + w];
        out[probe_ids[pd->probe_off + q]] = v;
    }
}

__device__ __forceinline__ void zero_pacc(double *pacc, int nprobe)
{
    for (int t = threadIdx.x; t < nprobe * kWaves; t += kTPB) pacc[t] = 0.0;
}

__device__ __forceinline__ void row_offsets(const PassDev *pd, uint64_t *rowoff)
{
    const int cf = pd->cfast, nrow = 1 << (pd->K - cf);
    for (int t = threadIdx.x; t < nrow; t += kTPB) {
        uint64_t o = 0;
        for (int k = cf; k < pd->K; ++k)
            if ((t >> (k - cf)) & 1) o |= 1ull << pd->pos[k];
        rowoff[t] = o;
    }
}

// Whole state in LDS: element e of the chunk is state index e (K == n, the
// pass covers every bit in ascending order).
__device__ double run_resident(const BatchCirc &C, const BatchArgs &A, cx<float> *c0, cx<float> *c1,
                               double *pacc, double *sh)
{
    const int E = 1 << C.n;
    // slots split as swz(e - threadIdx.x) ^ swz(threadIdx.x): kTPB is a power of two
    const int st = swz((int)threadIdx.x);
    for (int e = threadIdx.x; e < E; e += kTPB) c0[swz(e - threadIdx.x) ^ st] = cx<float>{e == 0 ? 1.f : 0.f, 0.f};
    __syncthreads();
    for (int p = 0; p < C.nfwd; ++p) run_steps<float, 1>(c0, c1, A.passes + C.fwd0 + p, A.steps, A.mats, pacc);
    NSum acc;
    for (int e = threadIdx.x; e < E; e += kTPB) {
        const int sl = swz(e - threadIdx.x) ^ st;
        const cx<float> v = c0[sl];
        const float sg = zsign((uint64_t)e, C.zmask);
        acc.add((double)sg * ((double)v.re * v.re + (double)v.im * v.im));
        c1[sl] = cx<float>{sg * v.re, sg * v.im};
    }
    const double ez = block_sum(acc.value(), sh);
    __syncthreads();
    for (int p = 0; p < C.nrev; ++p) {
        const PassDev *pd = A.passes + C.rev0 + p;
        zero_pacc(pacc, pd->nprobe);
        __syncthreads();
        run_steps<float, 2>(c0, c1, pd, A.steps, A.mats, pacc);
        flush_probes(pd, A.probe_ids, pacc, A.probes + C.probe0);
        __syncthreads();
    }
    return ez;
}

// State in this workgroup's HBM slot, streamed through LDS chunk by chunk.
__device__ double run_streamed(const BatchCirc &C, const BatchArgs &A, cx<float> *psi, cx<float> *lam,
                               cx<float> *c0, cx<float> *c1, uint64_t *rowoff, double *pacc, double *sh)
{
    const uint64_t dim = 1ull << C.n;
    if (C.nfwd == 0) {  // no forward pass to generate |0>: write it
        for (uint64_t e = threadIdx.x; e < dim; e += kTPB) psi[e] = cx<float>{e == 0 ? 1.f : 0.f, 0.f};
        __syncthreads();
    }
    for (int p = 0; p < C.nfwd; ++p) {
        const PassDev *pd = A.passes + C.fwd0 + p;
        const int K = pd->K, cf = pd->cfast, E = 1 << K;
        row_offsets(pd, rowoff);
        __syncthreads();
        const uint64_t nchunks = 1ull << pd->ncomp;
        for (uint64_t c = 0; c < nchunks; ++c) {
            const uint64_t base = chunk_base(c, pd);
            if (p == 0) {  // the input is |0>: generate the chunk instead of reading it
                const int st = swz((int)threadIdx.x);
                for (int e = threadIdx.x; e < E; e += kTPB)
                    c0[swz(e - threadIdx.x) ^ st] = cx<float>{(base == 0 && e == 0) ? 1.f : 0.f, 0.f};
            } else {
                chunk_copy<float, true>(psi, c0, base, rowoff, E, cf);
            }
            __syncthreads();
            run_steps<float, 1>(c0, c1, pd, A.steps, A.mats, pacc);
            chunk_copy<float, false>(psi, c0, base, rowoff, E, cf);
            __syncthreads();
        }
    }
    NSum acc;
    for (uint64_t e = threadIdx.x; e < dim; e += kTPB) {
        const cx<float> v = psi[e];
        const float sg = zsign(e, C.zmask);
        acc.add((double)sg * ((double)v.re * v.re + (double)v.im * v.im));
        lam[e] = cx<float>{sg * v.re, sg * v.im};
    }
    const double ez = block_sum(acc.value(), sh);
    __syncthreads();
    for (int p = 0; p < C.nrev; ++p) {
        const PassDev *pd = A.passes + C.rev0 + p;
        const int K = pd->K, cf = pd->cfast, E = 1 << K;
        row_offsets(pd, rowoff);
        zero_pacc(pacc, pd->nprobe);
        __syncthreads();
        const uint64_t nchunks = 1ull << pd->ncomp;
        for (uint64_t c = 0; c < nchunks; ++c) {
            const uint64_t base = chunk_base(c, pd);
            chunk_copy<float, true>(psi, c0, base, rowoff, E, cf);
            chunk_copy<float, true>(lam, c1, base, rowoff, E, cf);
            __syncthreads();
            run_steps<float, 2>(c0, c1, pd, A.steps, A.mats, pacc);
            chunk_copy<float, false>(psi, c0, base, rowoff, E, cf);
            chunk_copy<float, false>(lam, c1, base, rowoff, E, cf);
            __syncthreads();
        }
        flush_probes(pd, A.probe_ids, pacc, A.probes + C.probe0);
        __syncthreads();
    }
    return ez;
}

}  // namespace

__global__ __launch_bounds__(kTPB) void k_batch_reset(unsigned *counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) counter[0] = 0u;
}

// Every workgroup loops: take the next queue position, run that circuit,
// record its result and wall-clock interval; exit once the queue is empty.
// Two waves per SIMD: the LDS (two chunks, ~68 KB) allows two workgroups per
// CU, and the register cap that allows it (256, with ~110 registers spilled
// in the two-vector steps) still lifts the 4096-circuit batch from 660 to
// 1047 circuits/s (measured) by overlapping one workgroup's chunk copies
// with the other's steps.
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(2))) void k_batch(BatchArgs A)
{
    __shared__ __align__(16) cx<float> c0[1 << kBatchK];
    __shared__ __align__(16) cx<float> c1[1 << kBatchK];
    __shared__ uint64_t rowoff[kBatchRows];
    __shared__ double pacc[kBatchProbes * kWaves];
    __shared__ double sh[kWaves];
    __shared__ int s_next;
    cx<float> *psi = A.scratch + (uint64_t)blockIdx.x * 2 * A.slot_elems;
    cx<float> *lam = psi + A.slot_elems;
    for (;;) {
        if (threadIdx.x == 0) s_next = (int)atomicAdd(A.counter, 1u);
        __syncthreads();
        const int i = s_next;
        __syncthreads();
        if (i >= A.ncirc) break;
        const int ci = A.order[i];
        const BatchCirc C = A.circs[ci];
        const uint64_t t0 = wall_clock64();
        const double ez = (C.K == C.n) ? run_resident(C, A, c0, c1, pacc, sh)
                                       : run_streamed(C, A, psi, lam, c0, c1, rowoff, pacc, sh);
        if (threadIdx.x == 0) {
            A.expect[ci] = ez;
            A.times[2 * ci] = t0;
            A.times[2 * ci + 1] = wall_clock64();
        }
        __syncthreads();
    }
}

hipError_t launch_batch_reset(unsigned *counter, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counter);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &a, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch, dim3((unsigned)blocks), dim3(kTPB), 0, s, a);
    return hipGetLastError();
}

int batch_blocks_per_cu_max()
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, kTPB, 0) != hipSuccess || per_cu < 1) return 1;
    return per_cu;
}

}  // namespace qsim
