// Persistent batch kernel (see qsim_batch.h).
#include "qsim_batch.h"

namespace qsim {

namespace {

__device__ __forceinline__ float zsign(uint64_t idx, uint64_t zmask)
{
    return (__popcll(idx & zmask) & 1) ? -1.f : 1.f;
}

// A gang's place in a task: its rank, size, barrier counter and the barriers
// passed so far.
struct Gang {
    int rank, size;
    unsigned *bar;
    unsigned passed;
};

// Barrier of all ranks of a gang.  Each workgroup's stores are complete
// (__syncthreads) and released device-wide (thread 0's agent-scope fence)
// before its arrival counts; after the count reaches passed * size every rank
// acquires.  The wait sleeps between polls.
__device__ __forceinline__ void gang_sync(Gang &g)
{
    __syncthreads();
    if (g.size > 1) {
        if (threadIdx.x == 0) {
            __threadfence();
            atomicAdd(g.bar, 1u);
            const unsigned target = (g.passed + 1) * (unsigned)g.size;
            while (__hip_atomic_load(g.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
                __builtin_amdgcn_s_sleep(8);
            __threadfence();
        }
        ++g.passed;
        __syncthreads();
    }
}

// [lo, hi) of `count` items for this rank: contiguous shares in rank order.
__device__ __forceinline__ void rank_share(uint64_t count, const Gang &g, uint64_t &lo, uint64_t &hi)
{
    lo = count * (uint64_t)g.rank / (uint64_t)g.size;
    hi = count * (uint64_t)(g.rank + 1) / (uint64_t)g.size;
}

// Sum of the kWaves per-wave partials of each probe of pass pd into its
// output slot: the circuit's probe output when one workgroup runs it, else
// this rank's partial slot (added by rank 0 at the end).
template <typename PP>
__device__ __forceinline__ void flush_probes(PP pd, const int32_t *__restrict__ probe_ids,
                                             const double *pacc, double *__restrict__ out, int stride, int rank)
{
    for (int q = threadIdx.x; q < pd->nprobe; q += kTPB) {
        double v = 0.0;
        for (int w = 0; w < kWaves; ++w) v += pacc[q * kWaves + w];
        out[(int64_t)probe_ids[pd->probe_off + q] * stride + rank] = v;
    }
}

__device__ __forceinline__ void zero_pacc(double *pacc, int nprobe)
{
    for (int t = threadIdx.x; t < nprobe * kWaves; t += kTPB) pacc[t] = 0.0;
}

template <typename PP>
__device__ __forceinline__ void row_offsets(PP pd, uint64_t *rowoff)
{
    const int cf = pd->cfast, nrow = 1 << (pd->K - cf);
    for (int t = threadIdx.x; t < nrow; t += kTPB) {
        uint64_t o = 0;
        for (int k = cf; k < pd->K; ++k)
            if ((t >> (k - cf)) & 1) o |= 1ull << pd->pos[k];
        rowoff[t] = o;
    }
}

// Brings a pass's step records and matrix entries into the L2 cache with one
// sweep of vector loads, all in flight together, so the scalar loads of the
// steps that follow (one record and its matrices per step, read once per
// chunk) hit in the cache.  Returns a value that depends on every load.
template <typename PP>
__device__ __forceinline__ uint32_t warm_pass(PP pd, const StepDev *steps, const cx<float> *mats)
{
    const char *s0 = reinterpret_cast<const char *>(steps + pd->step_off);
    const char *m0 = reinterpret_cast<const char *>(mats + pd->mat_lo);
    const int sbytes = pd->nsteps * (int)sizeof(StepDev), mbytes = (pd->mat_hi - pd->mat_lo) * 8;
    uint32_t acc = 0;
    for (int o = threadIdx.x * 128; o < sbytes; o += kTPB * 128) acc ^= *reinterpret_cast<const uint32_t *>(s0 + o);
    for (int o = threadIdx.x * 128; o < mbytes; o += kTPB * 128) acc ^= *reinterpret_cast<const uint32_t *>(m0 + o);
    return acc;
}

__device__ __forceinline__ void keep(uint32_t v) { asm volatile("" ::"v"(v)); }

// The passes [first, first + count) of a circuit: NV = 1, the forward sweep
// on psi; NV = 2, the reverse sweep on (psi, lam) with its probes.  With the
// state in LDS (resident) each pass is one chunk that never leaves LDS;
// otherwise this rank's share of every pass's chunks streams through LDS,
// with a gang barrier after each pass whose output another rank reads.  One
// inlined copy per NV serves both kinds of circuit: the kernel's code then
// fits the instruction cache.
template <int NV>
__device__ __forceinline__ void sweep(const BatchCirc &C, const BatchArgs &A, Gang &g, bool resident,
                                      cx<float> *psi, cx<float> *lam, cx<float> *c0, cx<float> *c1,
                                      uint64_t *rowoff, double *pacc, double *pout, int first, int count)
{
    for (int p = 0; p < count; ++p) {
        const QSIM_CONST PassDev *pd = as_const(A.passes + first + p);
        keep(warm_pass(pd, A.steps, A.mats));
        const int K = pd->K, cf = pd->cfast, E = 1 << K;
        if (NV == 2) zero_pacc(pacc, pd->nprobe);
        uint64_t lo = 0, hi = 1;
        if (!resident) {
            row_offsets(pd, rowoff);
            rank_share(1ull << pd->ncomp, g, lo, hi);
        }
        __syncthreads();
        for (uint64_t c = lo; c < hi; ++c) {
            const uint64_t base = resident ? 0 : chunk_base(c, pd);
            if (!resident) {
                if (NV == 1 && p == 0) {  // the input is |0>: generate the chunk instead of reading it
                    const int st = swz((int)threadIdx.x);
                    for (int e = threadIdx.x; e < E; e += kTPB)
                        c0[swz(e - threadIdx.x) ^ st] = cx<float>{(base == 0 && e == 0) ? 1.f : 0.f, 0.f};
                } else {
                    chunk_copy<float, true>(psi, c0, base, rowoff, E, cf);
                    if (NV == 2) chunk_copy<float, true>(lam, c1, base, rowoff, E, cf);
                }
                __syncthreads();
            }
            run_steps<float, NV, true>(c0, c1, pd, as_const(A.steps), as_const(A.mats), pacc);
            if (!resident) {
                chunk_copy<float, false>(psi, c0, base, rowoff, E, cf);
                if (NV == 2) chunk_copy<float, false>(lam, c1, base, rowoff, E, cf);
                __syncthreads();
            }
        }
        if (NV == 2) {
            flush_probes(pd, A.probe_ids, pacc, pout, g.size, g.rank);
            __syncthreads();
        }
        // the reverse sweep's last pass is not read again: only the partials
        // need the barrier that follows it
        if (!resident && (NV == 1 || p + 1 < count)) gang_sync(g);
    }
}

// Whole state in LDS (element e of the chunk is state index e: K == n, the
// pass covers every bit in ascending order) or in an HBM slot; with a gang,
// this rank's share.  Returns <Z> (the full value in rank 0; other ranks'
// return values are unused).
__device__ __forceinline__ double run_circuit(const BatchCirc &C, const BatchArgs &A, Gang &g, cx<float> *psi,
                                              cx<float> *lam, cx<float> *c0, cx<float> *c1, uint64_t *rowoff,
                                              double *pacc, double *sh)
{
    const bool resident = C.K == C.n;
    double *part = g.size > 1 ? A.gang_part + C.part0 : nullptr;
    uint64_t lo, hi;
    if (resident) {
        const int E = 1 << C.n;
        // slots split as swz(e - threadIdx.x) ^ swz(threadIdx.x): kTPB is a power of two
        const int st = swz((int)threadIdx.x);
        for (int e = threadIdx.x; e < E; e += kTPB) c0[swz(e - threadIdx.x) ^ st] = cx<float>{e == 0 ? 1.f : 0.f, 0.f};
        __syncthreads();
    } else if (C.nfwd == 0) {  // no forward pass to generate |0>: write it
        rank_share(1ull << C.n, g, lo, hi);
        for (uint64_t e = lo + threadIdx.x; e < hi; e += kTPB) psi[e] = cx<float>{e == 0 ? 1.f : 0.f, 0.f};
        gang_sync(g);
    }
    sweep<1>(C, A, g, resident, psi, lam, c0, c1, rowoff, pacc, nullptr, C.fwd0, C.nfwd);
    // <Z> in FP64 and lam = Z psi
    NSum acc;
    if (resident) {
        const int E = 1 << C.n, st = swz((int)threadIdx.x);
        for (int e = threadIdx.x; e < E; e += kTPB) {
            const int sl = swz(e - threadIdx.x) ^ st;
            const cx<float> v = c0[sl];
            const float sg = zsign((uint64_t)e, C.zmask);
            acc.add((double)sg * ((double)v.re * v.re + (double)v.im * v.im));
            c1[sl] = cx<float>{sg * v.re, sg * v.im};
        }
    } else {
        rank_share(1ull << C.n, g, lo, hi);
        for (uint64_t e = lo + threadIdx.x; e < hi; e += kTPB) {
            const cx<float> v = psi[e];
            const float sg = zsign(e, C.zmask);
            acc.add((double)sg * ((double)v.re * v.re + (double)v.im * v.im));
            lam[e] = cx<float>{sg * v.re, sg * v.im};
        }
    }
    double ez = block_sum(acc.value(), sh);
    if (part && threadIdx.x == 0) part[g.rank] = ez;
    if (resident)
        __syncthreads();
    else
        gang_sync(g);
    double *pout = part ? part + g.size : A.probes + C.probe0;
    sweep<2>(C, A, g, resident, psi, lam, c0, c1, rowoff, pacc, pout, C.rev0, C.nrev);
    if (part) {
        gang_sync(g);  // every rank's partials are written
        if (g.rank == 0) {
            for (int q = threadIdx.x; q < C.nprobe; q += kTPB) {
                double v = 0.0;
                for (int r = 0; r < g.size; ++r) v += pout[(int64_t)q * g.size + r];
                A.probes[C.probe0 + q] = v;
            }
            ez = 0.0;
            for (int r = 0; r < g.size; ++r) ez += part[r];
        }
    }
    return ez;
}

}  // namespace

// Clears the queue head, the gang barriers, the gang partials and the probe
// output (slots no probe op writes read 0), and sets every circuit's start
// time to all ones (lowered by its ranks).
__global__ __launch_bounds__(kTPB) void k_batch_reset(BatchArgs A)
{
    const int64_t t = (int64_t)blockIdx.x * kTPB + threadIdx.x, stride = (int64_t)gridDim.x * kTPB;
    if (t == 0) A.counter[0] = 0u;
    for (int64_t i = t; i < A.ngangs; i += stride) A.gang_bar[i] = 0u;
    for (int64_t i = t; i < A.ncirc; i += stride) A.times[2 * i] = ~0ull;
    for (int64_t i = t; i < A.part_total; i += stride) A.gang_part[i] = 0.0;
    for (int64_t i = t; i < A.probe_total; i += stride) A.probes[i] = 0.0;
}

// Every workgroup loops: take the next task, run that circuit (or its rank's
// share), record its result and wall-clock interval; exit once the queue is
// empty.  Two waves per SIMD: the LDS (two chunks, ~68 KB) allows two
// workgroups per CU, and the register cap that allows it (256, with ~110
// registers spilled in the two-vector steps) overlaps one workgroup's chunk
// copies with the other's steps.
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(QSIM_BATCH_WAVES))) void k_batch(BatchArgs A)
{
    __shared__ __align__(16) cx<float> c0[1 << kBatchK];
    __shared__ __align__(16) cx<float> c1[1 << kBatchK];
    __shared__ uint64_t rowoff[kBatchRows];
    __shared__ double pacc[kBatchProbes * kWaves];
    __shared__ double sh[kWaves];
    __shared__ int s_next, s_owner;
    for (;;) {
        if (threadIdx.x == 0) s_next = (int)atomicAdd(A.counter, 1u);
        __syncthreads();
        const int i = s_next;
        __syncthreads();
        if (i >= A.ntasks) break;
        const BatchTask t = A.tasks[i];
        const int ci = t.circ;
        const BatchCirc C = A.circs[ci];
        if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long *>(A.times + 2 * ci),
                                        (unsigned long long)wall_clock64());
        Gang g{t.rank, C.gang, C.gang > 1 ? A.gang_bar + C.gslot : nullptr, 0u};
        int owner = blockIdx.x;
        if (g.size > 1) {  // the state lives in rank 0's slot
            if (g.rank == 0 && threadIdx.x == 0) A.gang_owner[C.gslot] = owner;
            gang_sync(g);
            if (threadIdx.x == 0) s_owner = __hip_atomic_load(A.gang_owner + C.gslot, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            owner = s_owner;
        }
        cx<float> *psi = A.scratch + (uint64_t)owner * 2 * A.slot_elems;
        const double ez = run_circuit(C, A, g, psi, psi + A.slot_elems, c0, c1, rowoff, pacc, sh);
        if (threadIdx.x == 0 && t.rank == 0) {
            A.expect[ci] = ez;
            A.times[2 * ci + 1] = wall_clock64();
        }
        __syncthreads();
    }
}

hipError_t launch_batch_reset(const BatchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_reset, dim3(256), dim3(kTPB), 0, s, a);
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
