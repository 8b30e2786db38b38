// HIP kernels for gfx950 (CDNA4): gate streams, bit permutation, chunk passes,
// Pauli chunk passes and FP64 reductions.
//
// Index convention: amplitude index i = sum_q b_q 2^q; a 2-qubit matrix acts
// little-endian over (q0, q1): row/col = b_q0 + 2 b_q1
// (qgpusim_cuda/src/kernels.cu:54-55).
#include <algorithm>

#include "qsim_internal.h"

namespace qsim {

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ cx<float> lo2(float4 v) { return {v.x, v.y}; }
__device__ __forceinline__ cx<float> hi2(float4 v) { return {v.z, v.w}; }
__device__ __forceinline__ float4 pack2(cx<float> a, cx<float> b) { return make_float4(a.re, a.im, b.re, b.im); }


static inline unsigned grid_for(uint64_t units, int per_block)
{
    uint64_t b = (units + per_block - 1) / per_block;
    const uint64_t cap = 1ull << 20;
    if (b > cap) b = cap;
    if (b == 0) b = 1;
    return (unsigned)b;
}

// ----------------------------------------------------------------------------
// |basis> initialisation (a fill kernel: safe inside captured graphs)
// ----------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kTPB) void k_init_basis(cx<R> *__restrict__ psi, uint64_t dim, uint64_t index)
{
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t i = (uint64_t)blockIdx.x * kTPB + threadIdx.x; i < dim; i += stride)
        psi[i] = cx<R>{i == index ? R(1) : R(0), R(0)};
}

template <typename R>
hipError_t launch_init_basis(cx<R> *psi, uint64_t dim, uint64_t index, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_basis<R>, dim3(grid_for(dim, kTPB * 4)), dim3(kTPB), 0, s, psi, dim, index);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// 1-qubit streams.  complex64: float4 = 2 consecutive amplitudes.
// ----------------------------------------------------------------------------
constexpr int kUPT = 2;

// q >= 1: unit u = 2 consecutive pairs; vec index insert_zero(u, q-1).
__global__ __launch_bounds__(kTPB) void k1q_f_hi(float4 *__restrict__ v, uint64_t nunits, int qm1, Mat1<float> U)
{
    const uint64_t half = 1ull << qm1;
    const uint64_t stride = (uint64_t)gridDim.x * (kTPB * kUPT);
    for (uint64_t u0 = (uint64_t)blockIdx.x * (kTPB * kUPT) + threadIdx.x; u0 < nunits; u0 += stride) {
        float4 A[kUPT], B[kUPT];
        uint64_t idx[kUPT];
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            const uint64_t u = u0 + (uint64_t)j * kTPB;
            idx[j] = insert_zero(u < nunits ? u : 0, qm1);
            A[j] = v[idx[j]];
            B[j] = v[idx[j] + half];
        }
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            cx<float> a0 = lo2(A[j]), a1 = hi2(A[j]), b0 = lo2(B[j]), b1 = hi2(B[j]);
            pair_apply(U.u, a0, b0);
            pair_apply(U.u, a1, b1);
            if (u0 + (uint64_t)j * kTPB < nunits) {
                v[idx[j]] = pack2(a0, a1);
                v[idx[j] + half] = pack2(b0, b1);
            }
        }
    }
}

// q == 0: the pair sits inside one float4.
__global__ __launch_bounds__(kTPB) void k1q_f_lo(float4 *__restrict__ v, uint64_t nunits, Mat1<float> U)
{
    const uint64_t stride = (uint64_t)gridDim.x * (kTPB * kUPT);
    for (uint64_t u0 = (uint64_t)blockIdx.x * (kTPB * kUPT) + threadIdx.x; u0 < nunits; u0 += stride) {
        float4 A[kUPT];
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            const uint64_t u = u0 + (uint64_t)j * kTPB;
            A[j] = v[u < nunits ? u : 0];
        }
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            cx<float> a = lo2(A[j]), b = hi2(A[j]);
            pair_apply(U.u, a, b);
            const uint64_t u = u0 + (uint64_t)j * kTPB;
            if (u < nunits) v[u] = pack2(a, b);
        }
    }
}

// complex128: one amplitude per 16-byte load.
__global__ __launch_bounds__(kTPB) void k1q_d(double2 *__restrict__ v, uint64_t nunits, int q, Mat1<double> U)
{
    const uint64_t half = 1ull << q;
    const uint64_t stride = (uint64_t)gridDim.x * (kTPB * kUPT);
    for (uint64_t u0 = (uint64_t)blockIdx.x * (kTPB * kUPT) + threadIdx.x; u0 < nunits; u0 += stride) {
        double2 A[kUPT], B[kUPT];
        uint64_t idx[kUPT];
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            const uint64_t u = u0 + (uint64_t)j * kTPB;
            idx[j] = insert_zero(u < nunits ? u : 0, q);
            A[j] = v[idx[j]];
            B[j] = v[idx[j] + half];
        }
#pragma unroll
        for (int j = 0; j < kUPT; ++j) {
            cx<double> a{A[j].x, A[j].y}, b{B[j].x, B[j].y};
            pair_apply(U.u, a, b);
            if (u0 + (uint64_t)j * kTPB < nunits) {
                v[idx[j]] = make_double2(a.re, a.im);
                v[idx[j] + half] = make_double2(b.re, b.im);
            }
        }
    }
}

template <>
hipError_t launch_1q<float>(cx<float> *psi, int n, int q, const Mat1<float> &U, hipStream_t s)
{
    const uint64_t dim = 1ull << n;
    float4 *v = reinterpret_cast<float4 *>(psi);
    if (q >= 1) {
        const uint64_t nu = dim / 4;
        hipLaunchKernelGGL(k1q_f_hi, dim3(grid_for(nu, kTPB * kUPT)), dim3(kTPB), 0, s, v, nu, q - 1, U);
    } else {
        const uint64_t nu = dim / 2;
        hipLaunchKernelGGL(k1q_f_lo, dim3(grid_for(nu, kTPB * kUPT)), dim3(kTPB), 0, s, v, nu, U);
    }
    return hipGetLastError();
}

template <>
hipError_t launch_1q<double>(cx<double> *psi, int n, int q, const Mat1<double> &U, hipStream_t s)
{
    const uint64_t nu = (1ull << n) / 2;
    hipLaunchKernelGGL(k1q_d, dim3(grid_for(nu, kTPB * kUPT)), dim3(kTPB), 0, s,
                       reinterpret_cast<double2 *>(psi), nu, q, U);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// 2-qubit streams
// ----------------------------------------------------------------------------
// complex64, lo >= 1: unit = 2 consecutive quads; 4 float4 loads.
__global__ __launch_bounds__(kTPB) void k2q_f_hi(float4 *__restrict__ v, uint64_t nunits, int lo, int hi,
                                                 int q0, int q1, Mat2<float> U)
{
    const uint64_t o1 = 1ull << (q0 - 1), o2 = 1ull << (q1 - 1);
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t u = (uint64_t)blockIdx.x * kTPB + threadIdx.x; u < nunits; u += stride) {
        const uint64_t b = insert_zero(insert_zero(u, lo - 1), hi - 1);
        float4 V0 = v[b], V1 = v[b + o1], V2 = v[b + o2], V3 = v[b + o1 + o2];
        cx<float> a0 = lo2(V0), a1 = lo2(V1), a2 = lo2(V2), a3 = lo2(V3);
        cx<float> c0 = hi2(V0), c1 = hi2(V1), c2 = hi2(V2), c3 = hi2(V3);
        quad_apply(U.u, a0, a1, a2, a3);
        quad_apply(U.u, c0, c1, c2, c3);
        v[b] = pack2(a0, c0);
        v[b + o1] = pack2(a1, c1);
        v[b + o2] = pack2(a2, c2);
        v[b + o1 + o2] = pack2(a3, c3);
    }
}

// complex64, lo == 0: float4 at the base holds bit lo = 0/1.
__global__ __launch_bounds__(kTPB) void k2q_f_lo(float4 *__restrict__ v, uint64_t nunits, int hi,
                                                 int q0_is_low, Mat2<float> U)
{
    const uint64_t oh = 1ull << (hi - 1);
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t u = (uint64_t)blockIdx.x * kTPB + threadIdx.x; u < nunits; u += stride) {
        const uint64_t b = insert_zero(u, hi - 1);
        float4 A = v[b], B = v[b + oh];
        cx<float> x00 = lo2(A), x01 = hi2(A), x10 = lo2(B), x11 = hi2(B);  // x[hi][lo]
        if (q0_is_low) {
            quad_apply(U.u, x00, x01, x10, x11);
        } else {
            quad_apply(U.u, x00, x10, x01, x11);
        }
        v[b] = pack2(x00, x01);
        v[b + oh] = pack2(x10, x11);
    }
}

__global__ __launch_bounds__(kTPB) void k2q_d(double2 *__restrict__ v, uint64_t nunits, int lo, int hi,
                                              int q0, int q1, Mat2<double> U)
{
    const uint64_t o1 = 1ull << q0, o2 = 1ull << q1;
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t u = (uint64_t)blockIdx.x * kTPB + threadIdx.x; u < nunits; u += stride) {
        const uint64_t b = insert_zero(insert_zero(u, lo), hi);
        double2 V0 = v[b], V1 = v[b + o1], V2 = v[b + o2], V3 = v[b + o1 + o2];
        cx<double> a0{V0.x, V0.y}, a1{V1.x, V1.y}, a2{V2.x, V2.y}, a3{V3.x, V3.y};
        quad_apply(U.u, a0, a1, a2, a3);
        v[b] = make_double2(a0.re, a0.im);
        v[b + o1] = make_double2(a1.re, a1.im);
        v[b + o2] = make_double2(a2.re, a2.im);
        v[b + o1 + o2] = make_double2(a3.re, a3.im);
    }
}

template <>
hipError_t launch_2q<float>(cx<float> *psi, int n, int q0, int q1, const Mat2<float> &U, hipStream_t s)
{
    const uint64_t dim = 1ull << n;
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    float4 *v = reinterpret_cast<float4 *>(psi);
    if (lo >= 1) {
        const uint64_t nu = dim / 8;
        hipLaunchKernelGGL(k2q_f_hi, dim3(grid_for(nu, kTPB)), dim3(kTPB), 0, s, v, nu, lo, hi, q0, q1, U);
    } else {
        const uint64_t nu = dim / 4;
        hipLaunchKernelGGL(k2q_f_lo, dim3(grid_for(nu, kTPB)), dim3(kTPB), 0, s, v, nu, hi, (int)(q0 == lo), U);
    }
    return hipGetLastError();
}

template <>
hipError_t launch_2q<double>(cx<double> *psi, int n, int q0, int q1, const Mat2<double> &U, hipStream_t s)
{
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    const uint64_t nu = (1ull << n) / 4;
    hipLaunchKernelGGL(k2q_d, dim3(grid_for(nu, kTPB)), dim3(kTPB), 0, s,
                       reinterpret_cast<double2 *>(psi), nu, lo, hi, q0, q1, U);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// diagonal 1-qubit gate: psi[i] *= bit_q(i) ? p1 : p0
// ----------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kTPB) void k_diag(cx<R> *__restrict__ psi, uint64_t dim, int q, cx<R> p0, cx<R> p1)
{
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t i = (uint64_t)blockIdx.x * kTPB + threadIdx.x; i < dim; i += stride)
        psi[i] = cmul(psi[i], ((i >> q) & 1) ? p1 : p0);
}

template <typename R>
hipError_t launch_diag1q(cx<R> *psi, int n, int q, cx<R> p0, cx<R> p1, hipStream_t s)
{
    const uint64_t dim = 1ull << n;
    hipLaunchKernelGGL(k_diag<R>, dim3(grid_for(dim, kTPB * 4)), dim3(kTPB), 0, s, psi, dim, q, p0, p1);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Bit permutation out[j(i)] = in[i] through an LDS tile.
//
// A block owns the 2^s amplitudes that vary over the tile bit set S = (input
// bits 0..a-1) u (input bits landing on output bits 0..b-1).  It reads them as
// runs of 2^a contiguous input amplitudes, and writes them as runs of 2^b
// contiguous output amplitudes.  Tile slot r (input order) is stored at
// r ^ ((r >> a) & (2^a - 1)) so that the output-order reads spread over banks.
// ----------------------------------------------------------------------------
template <typename R, bool NT, int TPB>
__global__ __launch_bounds__(TPB) void k_permute(const cx<R> *__restrict__ in, cx<R> *__restrict__ out, PermArgs P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int s = P.s, a = P.a, b = P.b;
    const int S = 1 << s;
    cx<R> *tile = reinterpret_cast<cx<R> *>(smem);
    uint64_t *row_in = reinterpret_cast<uint64_t *>(tile + S);
    uint64_t *row_out = row_in + (1 << (s - a));
    uint32_t *w2r_lo = reinterpret_cast<uint32_t *>(row_out + (1 << (s - b)));
    uint32_t *w2r_hi = w2r_lo + (1 << b);

    for (int t = threadIdx.x; t < (1 << (s - a)); t += TPB) {
        uint64_t o = 0;
        for (int k = 0; k < s - a; ++k)
            if ((t >> k) & 1) o |= 1ull << P.sin_hi[k];
        row_in[t] = o;
    }
    for (int t = threadIdx.x; t < (1 << (s - b)); t += TPB) {
        uint64_t o = 0;
        uint32_t r = 0;
        for (int k = 0; k < s - b; ++k)
            if ((t >> k) & 1) {
                o |= 1ull << P.sout_hi[k];
                r |= 1u << P.w2r[b + k];
            }
        row_out[t] = o;
        w2r_hi[t] = r;
    }
    for (int t = threadIdx.x; t < (1 << b); t += TPB) {
        uint32_t r = 0;
        for (int k = 0; k < b; ++k)
            if ((t >> k) & 1) r |= 1u << P.w2r[k];
        w2r_lo[t] = r;
    }
    const uint64_t g = xcd_slab(blockIdx.x, P.ncomp);
    uint64_t base_in = 0, base_out = 0;
    for (int m = 0; m < P.ncomp; ++m)
        if ((g >> m) & 1) {
            base_in |= 1ull << P.comp_in[m];
            base_out |= 1ull << P.comp_out[m];
        }
    __syncthreads();

    const int amask = (1 << a) - 1, bmask = (1 << b) - 1;
    // tile slot of r: bits 1..a-1 xor the low bits of r >> a (bit 0 kept, so
    // (r, r^1) stay adjacent); spreads the output-order reads over banks.
    auto slot = [=](uint32_t r) { return r ^ (((r >> a) << 1) & amask); };
    if (sizeof(cx<R>) == 8 && a >= 1 && b >= 1) {
        const int nv = S >> 1;
        for (int v0 = 0; v0 < nv; v0 += TPB * kBatch) {
            v4f buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * TPB + threadIdx.x;
                const int r = v << 1;
                if (v < nv) buf[j] = gload<NT>(reinterpret_cast<const v4f *>(in + base_in + row_in[r >> a] + (r & amask)));
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * TPB + threadIdx.x;
                if (v < nv) *reinterpret_cast<v4f *>(tile + slot(v << 1)) = buf[j];
            }
        }
        __syncthreads();
        for (int v0 = 0; v0 < nv; v0 += TPB * kBatch) {
            v4f buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * TPB + threadIdx.x;
                const int w = v << 1;
                if (v < nv) {
                    const uint32_t r0 = w2r_lo[w & bmask] | w2r_hi[w >> b];
                    const uint32_t r1 = w2r_lo[(w + 1) & bmask] | w2r_hi[w >> b];
                    const cx<R> x0 = tile[slot(r0)], x1 = tile[slot(r1)];
                    buf[j] = v4f{(float)x0.re, (float)x0.im, (float)x1.re, (float)x1.im};
                }
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * TPB + threadIdx.x;
                const int w = v << 1;
                if (v < nv) gstore<NT>(reinterpret_cast<v4f *>(out + base_out + row_out[w >> b] + (w & bmask)), buf[j]);
            }
        }
    } else {
        typedef typename RawAmp<R>::type raw;
        const raw *ir = reinterpret_cast<const raw *>(in);
        raw *ow = reinterpret_cast<raw *>(out);
        raw *tr = reinterpret_cast<raw *>(tile);
        for (int r0 = 0; r0 < S; r0 += TPB * kBatch) {
            raw buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int r = r0 + j * TPB + threadIdx.x;
                if (r < S) buf[j] = gload<NT>(ir + base_in + row_in[r >> a] + (r & amask));
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int r = r0 + j * TPB + threadIdx.x;
                if (r < S) tr[slot(r)] = buf[j];
            }
        }
        __syncthreads();
        for (int w0 = 0; w0 < S; w0 += TPB * kBatch) {
            raw buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int w = w0 + j * TPB + threadIdx.x;
                if (w < S) buf[j] = tr[slot(w2r_lo[w & bmask] | w2r_hi[w >> b])];
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int w = w0 + j * TPB + threadIdx.x;
                if (w < S) gstore<NT>(ow + base_out + row_out[w >> b] + (w & bmask), buf[j]);
            }
        }
    }
}

template <typename R>
static size_t perm_smem(const PermArgs &P)
{
    return (sizeof(cx<R>) << P.s) + 8u * ((1u << (P.s - P.a)) + (1u << (P.s - P.b))) +
           4u * ((1u << P.b) + (1u << (P.s - P.b)));
}

template <typename R>
hipError_t launch_permute(const cx<R> *in, cx<R> *out, const PermArgs &P, hipStream_t s)
{
    const size_t sm = perm_smem<R>(P);
    const uint64_t blocks = 1ull << (P.n - P.s);
    hipLaunchKernelGGL((k_permute<R, true, kTPB>), dim3((unsigned)blocks), dim3(kTPB), sm, s, in, out, P);
    return hipGetLastError();
}



// One chunk per block.  complex64 forward passes ask for 4 waves per SIMD
// (at most 128 VGPRs): the copy batches otherwise push occupancy to 3.
// k-qubit dense gate (k = 3, 4) as one coalesced stream: thread t updates
// the 2^k amplitudes of group t (zeros inserted at the op's qubits in
// ascending order; member m at sum_i b_i(m) 2^{q_i}) with the row-major
// matrix U (index sum_i b_{q_i} 2^i), read through the scalar cache.
struct GateK {
    int32_t q[4];        // the op's qubits, matrix-index order
    int32_t sorted[4];   // the same, ascending
};

template <typename R, int K>
__global__ __launch_bounds__(kTPB) void k_gate_k(cx<R> *__restrict__ psi, uint64_t groups, GateK g,
                                                 const cx<R> *__restrict__ U)
{
    const uint64_t t = (uint64_t)blockIdx.x * kTPB + threadIdx.x;
    if (t >= groups) return;
    uint64_t base = t;
#pragma unroll
    for (int j = 0; j < K; ++j) base = insert_zero(base, g.sorted[j]);
    constexpr int M = 1 << K;
    cx<R> a[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        uint64_t off = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) off |= (uint64_t)((m >> i) & 1) << g.q[i];
        a[m] = psi[base + off];
    }
#pragma unroll
    for (int r = 0; r < M; ++r) {
        cx<R> acc = cmul(U[M * r], a[0]);
#pragma unroll
        for (int c = 1; c < M; ++c) acc = cmad(U[M * r + c], a[c], acc);
        uint64_t off = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) off |= (uint64_t)((r >> i) & 1) << g.q[i];
        psi[base + off] = acc;
    }
}

template <typename R>
hipError_t launch_gate_k(cx<R> *psi, int n, const int *q, int k, const cx<R> *d_U, hipStream_t s)
{
    GateK g;
    for (int j = 0; j < 4; ++j) g.q[j] = g.sorted[j] = j < k ? q[j] : 0;
    std::sort(g.sorted, g.sorted + k);
    const uint64_t groups = 1ull << (n - k);
    const unsigned blocks = (unsigned)((groups + kTPB - 1) / kTPB);
    if (k == 3)
        hipLaunchKernelGGL((k_gate_k<R, 3>), dim3(blocks), dim3(kTPB), 0, s, psi, groups, g, d_U);
    else
        hipLaunchKernelGGL((k_gate_k<R, 4>), dim3(blocks), dim3(kTPB), 0, s, psi, groups, g, d_U);
    return hipGetLastError();
}

template <typename R, int NV, bool GENERIC>
__global__ __launch_bounds__(kTPB)
__attribute__((amdgpu_waves_per_eu((NV == 1 && sizeof(R) == 4) ? 4 : 1))) void k_tile_pass(
    cx<R> *__restrict__ psi, cx<R> *__restrict__ lam, const PassDev *__restrict__ pd,
    const StepDev *__restrict__ steps, const cx<R> *__restrict__ mats, double *__restrict__ probe_partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int K = pd->K, cf = pd->cfast;
    const int E = 1 << K;
    cx<R> *c0 = reinterpret_cast<cx<R> *>(smem);
    cx<R> *c1 = c0 + E;  // only with NV == 2
    uint64_t *rowoff = reinterpret_cast<uint64_t *>(c0 + NV * E);
    const int nrow = 1 << (K - cf);
    double *pacc = reinterpret_cast<double *>(rowoff + nrow);
    const int nprobe = pd->nprobe;
    const uint64_t nchunks = 1ull << pd->ncomp;

    for (int t = threadIdx.x; t < nrow; t += kTPB) {
        uint64_t o = 0;
        for (int k = cf; k < K; ++k)
            if ((t >> (k - cf)) & 1) o |= 1ull << pd->pos[k];
        rowoff[t] = o;
    }
    __syncthreads();

    const uint64_t c = xcd_slab(blockIdx.x, pd->ncomp);
    if (c >= nchunks) return;
    const uint64_t base = chunk_base(c, pd);
    chunk_copy<R, true, true>(psi, c0, base, rowoff, E, cf);
    if (NV == 2) chunk_copy<R, true, true>(lam, c1, base, rowoff, E, cf);
    if (NV == 2)
        for (int t = threadIdx.x; t < nprobe * kWaves; t += kTPB) pacc[t] = 0.0;
    __syncthreads();
    run_steps<R, NV, GENERIC>(c0, c1, pd, steps, mats, pacc);
    chunk_copy<R, false, true>(psi, c0, base, rowoff, E, cf);
    if (NV == 2) {
        chunk_copy<R, false, true>(lam, c1, base, rowoff, E, cf);
        for (int p = threadIdx.x; p < nprobe; p += kTPB) {
            double v = 0.0;
            for (int w = 0; w < kWaves; ++w) v += pacc[p * kWaves + w];
            probe_partial[c * nprobe + p] = v;
        }
    }
}

template <typename R, int NV, bool GENERIC>
static hipError_t tile_pass_launch(cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                                   const StepDev *d_steps, const cx<R> *d_mats, double *d_probe_partial,
                                   hipStream_t s)
{
    const size_t sm = (size_t)NV * (sizeof(cx<R>) << hp.K) + 8u * (1u << (hp.K - hp.cfast)) +
                      (NV == 2 ? 8u * kWaves * (size_t)hp.nprobe : 0u);
    const uint64_t blocks = 1ull << (n - hp.K);
    hipLaunchKernelGGL((k_tile_pass<R, NV, GENERIC>), dim3((unsigned)blocks), dim3(kTPB), sm, s, psi, lam, d_pass, d_steps,
                       d_mats, d_probe_partial);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_tile_pass(cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                            const StepDev *d_steps, const cx<R> *d_mats, double *d_probe_partial, bool generic,
                            hipStream_t s)
{
    if (generic) {
        if (lam) return tile_pass_launch<R, 2, true>(psi, lam, n, hp, d_pass, d_steps, d_mats, d_probe_partial, s);
        return tile_pass_launch<R, 1, true>(psi, nullptr, n, hp, d_pass, d_steps, d_mats, nullptr, s);
    }
    if (lam) return tile_pass_launch<R, 2, false>(psi, lam, n, hp, d_pass, d_steps, d_mats, d_probe_partial, s);
    return tile_pass_launch<R, 1, false>(psi, nullptr, n, hp, d_pass, d_steps, d_mats, nullptr, s);
}

// ----------------------------------------------------------------------------
// FP64 reductions with Neumaier-compensated per-thread sums; fixed order,
// so results are deterministic.
// ----------------------------------------------------------------------------
// partial[p * nblocks + b] layout -> out[ids[p]] = sum_b
__global__ __launch_bounds__(kTPB) void k_probe_reduce(const double *__restrict__ partial, int nblocks, int nprobe,
                                                       const int32_t *__restrict__ ids, double *__restrict__ out)
{
    __shared__ double sh[kWaves];
    const int p = blockIdx.x;
    NSum acc;
    for (int b = threadIdx.x; b < nblocks; b += kTPB) acc.add(partial[(uint64_t)b * nprobe + p]);
    const double t = block_sum(acc.value(), sh);
    if (threadIdx.x == 0) out[ids[p]] = t;
}

hipError_t launch_probe_reduce(const double *partial, int nblocks, int nprobe, const int32_t *d_ids, double *out,
                               hipStream_t s)
{
    if (nprobe <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_reduce, dim3(nprobe), dim3(kTPB), 0, s, partial, nblocks, nprobe, d_ids, out);
    return hipGetLastError();
}

constexpr int kDotBlocks = 2048;
int max_dot_blocks() { return kDotBlocks; }

// partial[b*2 + {0,1}] = sum conj(a_i) b_i over the block's grid-stride share;
// the two vectors may differ in precision (every product is taken in FP64)
template <typename RA, typename RB>
__global__ __launch_bounds__(kTPB) void k_dot(const cx<RA> *__restrict__ a, const cx<RB> *__restrict__ b, uint64_t dim,
                                              double *__restrict__ partial)
{
    __shared__ double sh[kWaves];
    NSum re, im;
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t i = (uint64_t)blockIdx.x * kTPB + threadIdx.x; i < dim; i += stride) {
        const cx<RA> x = a[i];
        const cx<RB> y = b[i];
        const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
        re.add(xr * yr + xi * yi);
        im.add(xr * yi - xi * yr);
    }
    double t = block_sum(re.value(), sh);
    if (threadIdx.x == 0) partial[2 * blockIdx.x] = t;
    t = block_sum(im.value(), sh);
    if (threadIdx.x == 0) partial[2 * blockIdx.x + 1] = t;
}

template <typename RA, typename RB>
hipError_t launch_dot(const cx<RA> *a, const cx<RB> *b, uint64_t dim, double *d_partial, int *nblocks_out,
                      hipStream_t s)
{
    uint64_t nb = (dim + kTPB * 8 - 1) / (kTPB * 8);
    if (nb > kDotBlocks) nb = kDotBlocks;
    if (nb < 1) nb = 1;
    *nblocks_out = (int)nb;
    hipLaunchKernelGGL((k_dot<RA, RB>), dim3((unsigned)nb), dim3(kTPB), 0, s, a, b, dim, d_partial);
    return hipGetLastError();
}

template hipError_t launch_dot<float, float>(const cx<float> *, const cx<float> *, uint64_t, double *, int *,
                                             hipStream_t);
template hipError_t launch_dot<double, double>(const cx<double> *, const cx<double> *, uint64_t, double *, int *,
                                               hipStream_t);
template hipError_t launch_dot<float, double>(const cx<float> *, const cx<double> *, uint64_t, double *, int *,
                                              hipStream_t);
template hipError_t launch_dot<double, float>(const cx<double> *, const cx<float> *, uint64_t, double *, int *,
                                              hipStream_t);

// out[w] = sum_b partial[b * width + w], one block per w, fixed order.
__global__ __launch_bounds__(kTPB) void k_sum_partials(const double *__restrict__ partial, int nblocks, int width,
                                                       double *__restrict__ out)
{
    __shared__ double sh[kWaves];
    const int w = blockIdx.x;
    NSum acc;
    for (int b = threadIdx.x; b < nblocks; b += kTPB) acc.add(partial[(uint64_t)b * width + w]);
    const double t = block_sum(acc.value(), sh);
    if (threadIdx.x == 0) out[w] = t;
}

hipError_t launch_sum_partials(const double *partial, int nblocks, int width, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(width), dim3(kTPB), 0, s, partial, nblocks, width, out);
    return hipGetLastError();
}

template <typename R>
__global__ __launch_bounds__(kTPB) void k_scale(cx<R> *__restrict__ psi, uint64_t dim, R f)
{
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t i = (uint64_t)blockIdx.x * kTPB + threadIdx.x; i < dim; i += stride) {
        cx<R> v = psi[i];
        psi[i] = cx<R>{v.re * f, v.im * f};
    }
}

template <typename R>
hipError_t launch_scale(cx<R> *psi, uint64_t dim, double factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale<R>, dim3(grid_for(dim, kTPB * 4)), dim3(kTPB), 0, s, psi, dim, (R)factor);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Pauli chunk pass: for the terms whose X support lies in the chunk,
//   t_e = sum_t c_t i^{|x&z|} (-1)^{popc((i ^ x) & z)} psi[i ^ x]   (i = state index of e)
// then lam = t (WRITE) or lam += t (ACC), and/or the block's partial of
// sum_e conj(psi_e) t_e (EXPECT, complex, 2 doubles per block).
// ----------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kTPB) void k_pauli_chunk(const cx<R> *__restrict__ psi, cx<R> *__restrict__ lam,
                                                      const PassDev *__restrict__ pd,
                                                      const PauliDev *__restrict__ terms, int nterms, int mode,
                                                      double *__restrict__ partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double sh[kWaves];
    const int K = pd->K, cf = pd->cfast;
    const int E = 1 << K;
    cx<R> *c0 = reinterpret_cast<cx<R> *>(smem);
    uint64_t *rowoff = reinterpret_cast<uint64_t *>(c0 + E);
    const int nrow = 1 << (K - cf);
    for (int t = threadIdx.x; t < nrow; t += kTPB) {
        uint64_t o = 0;
        for (int k = cf; k < K; ++k)
            if ((t >> (k - cf)) & 1) o |= 1ull << pd->pos[k];
        rowoff[t] = o;
    }
    const uint64_t g = xcd_slab(blockIdx.x, pd->ncomp);
    uint64_t base = 0;
    for (int m = 0; m < pd->ncomp; ++m)
        if ((g >> m) & 1) base |= 1ull << pd->comp[m];
    __syncthreads();
    const int lowmask = (1 << cf) - 1;
    chunk_copy<R, true, true>(const_cast<cx<R> *>(psi), c0, base, rowoff, E, cf);
    __syncthreads();

    NSum er, ei;
    const int st = swzr<R>((int)threadIdx.x);  // slots split like e = (e - threadIdx.x) ^ threadIdx.x
    for (int e = threadIdx.x; e < E; e += kTPB) {
        const uint64_t addr = base + rowoff[e >> cf] + (e & lowmask);
        const int se = swzr<R>(e - threadIdx.x) ^ st;
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k < nterms; ++k) {
            const PauliDev T = terms[k];
            const cx<R> v = c0[se ^ swzr<R>(T.xloc)];
            const uint64_t src = addr ^ T.xmask;
            // i^{|x&z|} * (-1)^{popc(src & z)}
            const int ph = (__popcll(T.xmask & T.zmask) + 2 * __popcll(src & T.zmask)) & 3;
            double vr = v.re, vi = v.im, pr, pi;
            switch (ph) {
            case 0: pr = vr; pi = vi; break;
            case 1: pr = -vi; pi = vr; break;
            case 2: pr = -vr; pi = -vi; break;
            default: pr = vi; pi = -vr; break;
            }
            tr += T.cre * pr - T.cim * pi;
            ti += T.cre * pi + T.cim * pr;
        }
        if (mode & (PAULI_WRITE | PAULI_ACC)) {
            cx<R> o{(R)tr, (R)ti};
            if (mode & PAULI_ACC) {
                const cx<R> old = lam[addr];
                o = cx<R>{(R)(tr + (double)old.re), (R)(ti + (double)old.im)};
            }
            lam[addr] = o;
        }
        if (mode & PAULI_EXPECT) {
            const cx<R> p = c0[se];
            const double pr = p.re, pi = p.im;
            er.add(pr * tr + pi * ti);
            ei.add(pr * ti - pi * tr);
        }
    }
    if (mode & PAULI_EXPECT) {
        double t = block_sum(er.value(), sh);
        if (threadIdx.x == 0) partial[2 * blockIdx.x] = t;
        t = block_sum(ei.value(), sh);
        if (threadIdx.x == 0) partial[2 * blockIdx.x + 1] = t;
    }
}

template <typename R>
hipError_t launch_pauli_chunk(const cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                              const PauliDev *d_terms, int nterms, int mode, double *d_partial, hipStream_t s)
{
    const size_t sm = (sizeof(cx<R>) << hp.K) + 8u * (1u << (hp.K - hp.cfast));
    const uint64_t blocks = 1ull << (n - hp.K);
    hipLaunchKernelGGL(k_pauli_chunk<R>, dim3((unsigned)blocks), dim3(kTPB), sm, s, psi, lam, d_pass, d_terms, nterms,
                       mode, d_partial);
    return hipGetLastError();
}

// Raise the dynamic-LDS limit of every kernel that stages chunks/tiles to the
// device's per-block maximum (160 KB on gfx950).
hipError_t prepare_kernels()
{
    int dev = 0, lds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const void *fns[] = {
        reinterpret_cast<const void *>(k_tile_pass<float, 1, false>),
        reinterpret_cast<const void *>(k_tile_pass<float, 2, false>),
        reinterpret_cast<const void *>(k_tile_pass<double, 1, false>),
        reinterpret_cast<const void *>(k_tile_pass<double, 2, false>),
        reinterpret_cast<const void *>(k_tile_pass<float, 1, true>),
        reinterpret_cast<const void *>(k_tile_pass<float, 2, true>),
        reinterpret_cast<const void *>(k_tile_pass<double, 1, true>),
        reinterpret_cast<const void *>(k_tile_pass<double, 2, true>),
        reinterpret_cast<const void *>(k_permute<float, true, kTPB>),
        reinterpret_cast<const void *>(k_permute<double, true, kTPB>),
        reinterpret_cast<const void *>(k_pauli_chunk<float>), reinterpret_cast<const void *>(k_pauli_chunk<double>),
    };
    for (const void *f : fns) hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipGetLastError();
    return hipSuccess;
}

// explicit instantiations
#define INST(R)                                                                                              \
    template hipError_t launch_init_basis<R>(cx<R> *, uint64_t, uint64_t, hipStream_t);                     \
    template hipError_t launch_diag1q<R>(cx<R> *, int, int, cx<R>, cx<R>, hipStream_t);                     \
    template hipError_t launch_permute<R>(const cx<R> *, cx<R> *, const PermArgs &, hipStream_t);          \
    template hipError_t launch_tile_pass<R>(cx<R> *, cx<R> *, int, const PassDev &, const PassDev *,        \
                                            const StepDev *, const cx<R> *, double *, bool, hipStream_t);   \
    template hipError_t launch_pauli_chunk<R>(const cx<R> *, cx<R> *, int, const PassDev &,                 \
                                              const PassDev *, const PauliDev *, int, int, double *,        \
                                              hipStream_t);                                                 \
    template hipError_t launch_scale<R>(cx<R> *, uint64_t, double, hipStream_t);                           \
    template hipError_t launch_gate_k<R>(cx<R> *, int, const int *, int, const cx<R> *, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace qsim
