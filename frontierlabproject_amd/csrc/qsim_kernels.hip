// HIP kernels for gfx950 (CDNA4): gate streams, bit permutation, chunk passes,
// Pauli chunk passes and FP64 reductions.
//
// Index convention: amplitude index i = sum_q b_q 2^q; a 2-qubit matrix acts
// little-endian over (q0, q1): row/col = b_q0 + 2 b_q1
// (qgpusim_cuda/src/kernels.cu:54-55).
#include "qsim_internal.h"

namespace qsim {

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint64_t insert_zero(uint64_t x, int p)
{
    const uint64_t lo = x & ((1ull << p) - 1);
    return ((x >> p) << (p + 1)) | lo;
}

__device__ __forceinline__ uint32_t ins0(uint32_t x, int p)
{
    return ((x >> p) << (p + 1)) | (x & ((1u << p) - 1u));
}

template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename R>
__device__ __forceinline__ cx<R> cmad(cx<R> a, cx<R> b, cx<R> c)  // a*b + c
{
    return {fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}

template <typename R>
__device__ __forceinline__ void pair_apply(const cx<R> *u, cx<R> &x, cx<R> &y)
{
    const cx<R> x0 = x, y0 = y;
    x = cmad(u[0], x0, cmul(u[1], y0));
    y = cmad(u[2], x0, cmul(u[3], y0));
}

template <typename R>
__device__ __forceinline__ void quad_apply(const cx<R> *u, cx<R> &v0, cx<R> &v1, cx<R> &v2, cx<R> &v3)
{
    const cx<R> a0 = v0, a1 = v1, a2 = v2, a3 = v3;
    v0 = cmad(u[0], a0, cmad(u[1], a1, cmad(u[2], a2, cmul(u[3], a3))));
    v1 = cmad(u[4], a0, cmad(u[5], a1, cmad(u[6], a2, cmul(u[7], a3))));
    v2 = cmad(u[8], a0, cmad(u[9], a1, cmad(u[10], a2, cmul(u[11], a3))));
    v3 = cmad(u[12], a0, cmad(u[13], a1, cmad(u[14], a2, cmul(u[15], a3))));
}

__device__ __forceinline__ cx<float> lo2(float4 v) { return {v.x, v.y}; }
__device__ __forceinline__ cx<float> hi2(float4 v) { return {v.z, v.w}; }
__device__ __forceinline__ float4 pack2(cx<float> a, cx<float> b) { return make_float4(a.re, a.im, b.re, b.im); }

typedef float v4f __attribute__((ext_vector_type(4)));  // native vectors: plain loads/stores
typedef float v2f __attribute__((ext_vector_type(2)));
// one amplitude as raw bits (struct copies would go through memcpy)
template <typename R> struct RawAmp { typedef v2f type; };
template <> struct RawAmp<double> { typedef v4f type; };

constexpr int kBatch = 8;  // global accesses in flight per thread in the copy loops

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
template <typename R>
__global__ __launch_bounds__(kTPB) void k_permute(const cx<R> *__restrict__ in, cx<R> *__restrict__ out, PermArgs P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int s = P.s, a = P.a, b = P.b;
    const int S = 1 << s;
    cx<R> *tile = reinterpret_cast<cx<R> *>(smem);
    uint64_t *row_in = reinterpret_cast<uint64_t *>(tile + S);
    uint64_t *row_out = row_in + (1 << (s - a));
    uint32_t *w2r_lo = reinterpret_cast<uint32_t *>(row_out + (1 << (s - b)));
    uint32_t *w2r_hi = w2r_lo + (1 << b);

    for (int t = threadIdx.x; t < (1 << (s - a)); t += kTPB) {
        uint64_t o = 0;
        for (int k = 0; k < s - a; ++k)
            if ((t >> k) & 1) o |= 1ull << P.sin_hi[k];
        row_in[t] = o;
    }
    for (int t = threadIdx.x; t < (1 << (s - b)); t += kTPB) {
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
    for (int t = threadIdx.x; t < (1 << b); t += kTPB) {
        uint32_t r = 0;
        for (int k = 0; k < b; ++k)
            if ((t >> k) & 1) r |= 1u << P.w2r[k];
        w2r_lo[t] = r;
    }
    const uint64_t g = blockIdx.x;
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
        for (int v0 = 0; v0 < nv; v0 += kTPB * kBatch) {
            v4f buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * kTPB + threadIdx.x;
                const int r = v << 1;
                if (v < nv) buf[j] = *reinterpret_cast<const v4f *>(in + base_in + row_in[r >> a] + (r & amask));
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * kTPB + threadIdx.x;
                if (v < nv) *reinterpret_cast<v4f *>(tile + slot(v << 1)) = buf[j];
            }
        }
        __syncthreads();
        for (int v0 = 0; v0 < nv; v0 += kTPB * kBatch) {
            v4f buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int v = v0 + j * kTPB + threadIdx.x;
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
                const int v = v0 + j * kTPB + threadIdx.x;
                const int w = v << 1;
                if (v < nv) *reinterpret_cast<v4f *>(out + base_out + row_out[w >> b] + (w & bmask)) = buf[j];
            }
        }
    } else {
        typedef typename RawAmp<R>::type raw;
        const raw *ir = reinterpret_cast<const raw *>(in);
        raw *ow = reinterpret_cast<raw *>(out);
        raw *tr = reinterpret_cast<raw *>(tile);
        for (int r0 = 0; r0 < S; r0 += kTPB * kBatch) {
            raw buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int r = r0 + j * kTPB + threadIdx.x;
                if (r < S) buf[j] = ir[base_in + row_in[r >> a] + (r & amask)];
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int r = r0 + j * kTPB + threadIdx.x;
                if (r < S) tr[slot(r)] = buf[j];
            }
        }
        __syncthreads();
        for (int w0 = 0; w0 < S; w0 += kTPB * kBatch) {
            raw buf[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int w = w0 + j * kTPB + threadIdx.x;
                if (w < S) buf[j] = tr[slot(w2r_lo[w & bmask] | w2r_hi[w >> b])];
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int w = w0 + j * kTPB + threadIdx.x;
                if (w < S) ow[base_out + row_out[w >> b] + (w & bmask)] = buf[j];
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
    hipLaunchKernelGGL(k_permute<R>, dim3((unsigned)blocks), dim3(kTPB), sm, s, in, out, P);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Register-step building blocks: 16 amplitudes per thread over 4 register bits;
// every register index is a compile-time constant.
// ----------------------------------------------------------------------------
template <typename R, int B>
__device__ __forceinline__ void reg1(cx<R> *a, const cx<R> *u)
{
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (!(m & (1 << B))) pair_apply(u, a[m], a[m | (1 << B)]);
}

// 4x4 on register bits (B0, B1), matrix index b_B0 + 2 b_B1
template <typename R, int B0, int B1>
__device__ __forceinline__ void reg2(cx<R> *a, const cx<R> *u)
{
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (!(m & ((1 << B0) | (1 << B1))))
            quad_apply(u, a[m], a[m | (1 << B0)], a[m | (1 << B1)], a[m | (1 << B0) | (1 << B1)]);
}

// Im( sum conj(l) * (G psi) ) over the pairs of register bit B.
template <typename R, int B>
__device__ __forceinline__ double probe1(const cx<R> *a, const cx<R> *l, const cx<R> *g)
{
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (!(m & (1 << B))) {
            const int m1 = m | (1 << B);
            const cx<R> x = a[m], y = a[m1];
            const cx<R> gx = cmad(g[0], x, cmul(g[1], y));
            const cx<R> gy = cmad(g[2], x, cmul(g[3], y));
            acc += (double)(l[m].re * gx.im - l[m].im * gx.re);
            acc += (double)(l[m1].re * gy.im - l[m1].im * gy.re);
        }
    return acc;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// 1q slot arithmetic on register bit B (applied to one vector).
template <typename R, int B>
__device__ __forceinline__ void apply1(cx<R> *a, int kind, const cx<R> *__restrict__ mp)
{
    switch (kind) {
    case S1_GEN: {
        const cx<R> u[4] = {mp[0], mp[1], mp[2], mp[3]};
        reg1<R, B>(a, u);
        break;
    }
    case S1_DIAG: {
        const cx<R> d0 = mp[0], d1 = mp[1];
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] = cmul(a[m], ((m >> B) & 1) ? d1 : d0);
        break;
    }
    case S1_PHASE: {
        const cx<R> d1 = mp[0];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if ((m >> B) & 1) a[m] = cmul(a[m], d1);
        break;
    }
    case S1_REAL: {
        const cx<R> r01 = mp[0], r23 = mp[1];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{fma(r01.re, x.re, r01.im * y.re), fma(r01.re, x.im, r01.im * y.im)};
                a[m1] = cx<R>{fma(r23.re, x.re, r23.im * y.re), fma(r23.re, x.im, r23.im * y.im)};
            }
        break;
    }
    case S1_RI: {  // [[p, iq], [ir, t]]
        const cx<R> pq = mp[0], rt = mp[1];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{fma(pq.re, x.re, -pq.im * y.im), fma(pq.re, x.im, pq.im * y.re)};
                a[m1] = cx<R>{fma(rt.im, y.re, -rt.re * x.im), fma(rt.im, y.im, rt.re * x.re)};
            }
        break;
    }
    case S1_X: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const cx<R> t = a[m];
                a[m] = a[m | (1 << B)];
                a[m | (1 << B)] = t;
            }
        break;
    }
    case S1_Y: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{y.im, -y.re};
                a[m1] = cx<R>{-x.im, x.re};
            }
        break;
    }
    case S1_BFLY: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{x.re + y.re, x.im + y.im};
                a[m1] = cx<R>{x.re - y.re, x.im - y.im};
            }
        break;
    }
    case S1_Z: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if ((m >> B) & 1) a[m] = cx<R>{-a[m].re, -a[m].im};
        break;
    }
    case S1_PI: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if ((m >> B) & 1) a[m] = cx<R>{-a[m].im, a[m].re};
        break;
    }
    case S1_MI: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if ((m >> B) & 1) a[m] = cx<R>{a[m].im, -a[m].re};
        break;
    }
    default:
        break;
    }
}

// One 1q slot on register bit B.
template <typename R, int NV, int B>
__device__ __forceinline__ void slot1(cx<R> *a, cx<R> *l, int kind, const cx<R> *__restrict__ mp, int probe,
                                      double *pacc, bool act)
{
    if (kind == S1_SKIP) return;
    if (kind == S1_PROBE || kind == S1_PROBE_GEN) {
        if (NV == 2) {
            const cx<R> g[4] = {mp[0], mp[1], mp[2], mp[3]};
            const double v = wave_sum(act ? probe1<R, B>(a, l, g) : 0.0);
            if ((threadIdx.x & 63) == 0) pacc[probe * kWaves + (threadIdx.x >> 6)] += v;
            if (kind == S1_PROBE_GEN) {
                const cx<R> u[4] = {mp[4], mp[5], mp[6], mp[7]};
                reg1<R, B>(a, u);
                reg1<R, B>(l, u);
            }
        }
        return;
    }
    apply1<R, B>(a, kind, mp);
    if (NV == 2) apply1<R, B>(l, kind, mp);
}

// pair slot arithmetic on register bits (F, S), F < S, index b_F + 2 b_S.
template <typename R, int F, int S>
__device__ __forceinline__ void apply2(cx<R> *a, int kind, const cx<R> *__restrict__ mp)
{
    switch (kind) {
    case S2_CPHASE: {
        const cx<R> d3 = mp[0];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (((m >> F) & 1) && ((m >> S) & 1)) a[m] = cmul(a[m], d3);
        break;
    }
    case S2_CX_FIRST: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (((m >> F) & 1) && !((m >> S) & 1)) {
                const cx<R> t = a[m];
                a[m] = a[m | (1 << S)];
                a[m | (1 << S)] = t;
            }
        break;
    }
    case S2_CX_SECOND: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (((m >> S) & 1) && !((m >> F) & 1)) {
                const cx<R> t = a[m];
                a[m] = a[m | (1 << F)];
                a[m | (1 << F)] = t;
            }
        break;
    }
    case S2_SWAP: {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (((m >> F) & 1) && !((m >> S) & 1)) {
                const int o = (m & ~(1 << F)) | (1 << S);
                const cx<R> t = a[m];
                a[m] = a[o];
                a[o] = t;
            }
        break;
    }
    default:
        break;
    }
}

// Dense 16x16 step: a <- M a (M row-major over register index m).
template <typename R>
__device__ __forceinline__ void dense16(cx<R> *a, const cx<R> *__restrict__ M)
{
    cx<R> o[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        cx<R> acc = cmul(M[16 * r], a[0]);
#pragma unroll
        for (int c = 1; c < 16; ++c) acc = cmad(M[16 * r + c], a[c], acc);
        o[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) a[r] = o[r];
}

// One pair slot on register bits (F, S), F < S.
template <typename R, int NV, int F, int S>
__device__ __forceinline__ void slot2(cx<R> *a, cx<R> *l, int kind, const cx<R> *__restrict__ mp)
{
    if (kind == S2_SKIP) return;
    apply2<R, F, S>(a, kind, mp);
    if (NV == 2) apply2<R, F, S>(l, kind, mp);
}

// ----------------------------------------------------------------------------
// Chunk pass.  Each block gathers one chunk (2^K amplitudes over the pass's
// bit set) into LDS, runs the register steps, writes it back in place.
// NV = 1: forward on psi.  NV = 2: reverse sweep on (psi, lam) with gradient
// probes Im<lam|G|psi> reduced per block.
// ----------------------------------------------------------------------------
// LDS slot of chunk element e: bits 6..10 are folded into bits 1..5, which
// spreads the strided step gathers over banks and keeps (e, e^1) adjacent so
// amplitude pairs move as one 16-byte access.
__device__ __forceinline__ int swz(int e) { return e ^ (((e >> 6) & 31) << 1); }

// Chunk <-> global copies.  Element e sits at base + rowoff[e >> cf] + (e & low).
// complex64 with cf >= 1 moves two amplitudes per 16-byte access.  Each thread
// issues kBatch loads before it writes any of them (memory-level parallelism).
template <typename R, bool LOAD>
__device__ __forceinline__ void chunk_copy(cx<R> *__restrict__ g, cx<R> *__restrict__ lds, uint64_t base,
                                           const uint64_t *__restrict__ rowoff, int E, int cf)
{
    const int lowmask = (1 << cf) - 1;
    if (sizeof(cx<R>) == 8 && cf >= 1) {
        const int nv = E >> 1;
        for (int v0 = 0; v0 < nv; v0 += kTPB * kBatch) {
            v4f buf[kBatch];
            if (LOAD) {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * kTPB + threadIdx.x;
                    const int e = v << 1;
                    if (v < nv) buf[j] = *reinterpret_cast<const v4f *>(g + base + rowoff[e >> cf] + (e & lowmask));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * kTPB + threadIdx.x;
                    if (v < nv) *reinterpret_cast<v4f *>(lds + swz(v << 1)) = buf[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * kTPB + threadIdx.x;
                    if (v < nv) buf[j] = *reinterpret_cast<const v4f *>(lds + swz(v << 1));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * kTPB + threadIdx.x;
                    const int e = v << 1;
                    if (v < nv) *reinterpret_cast<v4f *>(g + base + rowoff[e >> cf] + (e & lowmask)) = buf[j];
                }
            }
        }
    } else {
        typedef typename RawAmp<R>::type raw;
        const raw *gr = reinterpret_cast<const raw *>(g);
        raw *gw = reinterpret_cast<raw *>(g);
        raw *lr = reinterpret_cast<raw *>(lds);
        for (int e0 = 0; e0 < E; e0 += kTPB * kBatch) {
            raw buf[kBatch];
            if (LOAD) {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * kTPB + threadIdx.x;
                    if (e < E) buf[j] = gr[base + rowoff[e >> cf] + (e & lowmask)];
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * kTPB + threadIdx.x;
                    if (e < E) lr[swz(e)] = buf[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * kTPB + threadIdx.x;
                    if (e < E) buf[j] = lr[swz(e)];
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * kTPB + threadIdx.x;
                    if (e < E) gw[base + rowoff[e >> cf] + (e & lowmask)] = buf[j];
                }
            }
        }
    }
}

// The register steps of one pass over the chunk staged in LDS (c0, and c1 for
// the second vector).  Ends with a barrier after every step.
template <typename R, int NV>
__device__ __forceinline__ void run_steps(cx<R> *c0, cx<R> *c1, const PassDev *__restrict__ pd,
                                          const StepDev *__restrict__ steps, const cx<R> *__restrict__ mats,
                                          double *pacc)
{
    const int E = 1 << pd->K;
    const int ns = pd->nsteps;
    const StepDev *st = steps + pd->step_off;
    const int ngroups = E >> kG;
    for (int si = 0; si < ns; ++si) {
        const StepDev *sd = st + si;
        const int r0 = sd->bits[0], r1 = sd->bits[1], r2 = sd->bits[2], r3 = sd->bits[3];
        const int s0 = sd->sbits[0], s1 = sd->sbits[1], s2 = sd->sbits[2], s3 = sd->sbits[3];
        const int nr = sd->nrounds, dn = sd->dense, sc = sd->scale;
        // swz is XOR-linear and the group base and register offsets occupy
        // disjoint bits, so each element's LDS slot is swz(base) ^ swz(offset):
        // one XOR per element with a wave-uniform constant.
        int moff[16];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            moff[m] = swz(((m & 1) << r0) | (((m >> 1) & 1) << r1) | (((m >> 2) & 1) << r2) |
                          (((m >> 3) & 1) << r3));
        // wave-uniform trip count: probes reduce across all 64 lanes
        for (int gb = 0; gb < ngroups; gb += kTPB) {
            const int gi = gb + threadIdx.x;
            const bool act = gi < ngroups;
            // idle lanes read a valid group and never store
            const uint32_t g32 = (uint32_t)(act ? gi : ngroups - 1);
            const int sb = swz((int)ins0(ins0(ins0(ins0(g32, s0), s1), s2), s3));
            cx<R> a[16], l[NV == 2 ? 16 : 1];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                a[m] = c0[sb ^ moff[m]];
                if (NV == 2) l[m] = c1[sb ^ moff[m]];
            }
            if (dn >= 0) {
                // rows go straight back to LDS: no second 16-amplitude array
                const cx<R> f = sc >= 0 ? mats[sc] : cx<R>{R(1), R(0)};
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const cx<R> *M = mats + dn + 16 * r;
                    cx<R> acc = cmul(M[0], a[0]);
#pragma unroll
                    for (int c = 1; c < 16; ++c) acc = cmad(M[c], a[c], acc);
                    if (sc >= 0) acc = cmul(acc, f);
                    if (act) c0[sb ^ moff[r]] = acc;
                }
                if (NV == 2) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const cx<R> *M = mats + dn + 16 * r;
                        cx<R> acc = cmul(M[0], l[0]);
#pragma unroll
                        for (int c = 1; c < 16; ++c) acc = cmad(M[c], l[c], acc);
                        if (sc >= 0) acc = cmul(acc, f);
                        if (act) c1[sb ^ moff[r]] = acc;
                    }
                }
                continue;
            } else {
                // one copy of the round code, run nrounds times
#pragma unroll 1
                for (int r = 0; r < nr; ++r) {
                    const int b = r * kSlotsPerRound;
                    slot1<R, NV, 0>(a, l, sd->kind[b + 0], mats + sd->mat[b + 0], sd->probe[4 * r + 0], pacc, act);
                    slot1<R, NV, 1>(a, l, sd->kind[b + 1], mats + sd->mat[b + 1], sd->probe[4 * r + 1], pacc, act);
                    slot1<R, NV, 2>(a, l, sd->kind[b + 2], mats + sd->mat[b + 2], sd->probe[4 * r + 2], pacc, act);
                    slot1<R, NV, 3>(a, l, sd->kind[b + 3], mats + sd->mat[b + 3], sd->probe[4 * r + 3], pacc, act);
                    slot2<R, NV, 0, 1>(a, l, sd->kind[b + 4], mats + sd->mat[b + 4]);
                    slot2<R, NV, 2, 3>(a, l, sd->kind[b + 5], mats + sd->mat[b + 5]);
                    slot2<R, NV, 0, 2>(a, l, sd->kind[b + 6], mats + sd->mat[b + 6]);
                    slot2<R, NV, 1, 3>(a, l, sd->kind[b + 7], mats + sd->mat[b + 7]);
                    slot2<R, NV, 0, 3>(a, l, sd->kind[b + 8], mats + sd->mat[b + 8]);
                    slot2<R, NV, 1, 2>(a, l, sd->kind[b + 9], mats + sd->mat[b + 9]);
                }
            }
            if (sc >= 0) {  // scalar factored out of earlier gates
                const cx<R> f = mats[sc];
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    a[m] = cmul(a[m], f);
                    if (NV == 2) l[m] = cmul(l[m], f);
                }
            }
            if (act) {
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    c0[sb ^ moff[m]] = a[m];
                    if (NV == 2) c1[sb ^ moff[m]] = l[m];
                }
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t chunk_base(uint64_t c, const PassDev *__restrict__ pd)
{
    uint64_t base = 0;
    for (int m = 0; m < pd->ncomp; ++m)
        if ((c >> m) & 1) base |= 1ull << pd->comp[m];
    return base;
}

// PF 16-byte vectors per thread of one chunk: element(s) of vector v = tid + j*kTPB.
template <int PF>
struct Prefetch {
    v4f v[PF > 0 ? PF : 1];

    template <typename R>
    __device__ __forceinline__ void load(const cx<R> *__restrict__ g, uint64_t base,
                                         const uint64_t *__restrict__ rowoff, int cf)
    {
        const int lowmask = (1 << cf) - 1;
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const int t = threadIdx.x + j * kTPB;
            const int e = sizeof(cx<R>) == 8 ? (t << 1) : t;
            v[j] = *reinterpret_cast<const v4f *>(g + base + rowoff[e >> cf] + (e & lowmask));
        }
    }

    template <typename R>
    __device__ __forceinline__ void to_lds(cx<R> *lds) const
    {
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const int t = threadIdx.x + j * kTPB;
            const int e = sizeof(cx<R>) == 8 ? (t << 1) : t;
            *reinterpret_cast<v4f *>(lds + swz(e)) = v[j];
        }
    }
};

// PF > 0: persistent blocks walk the chunks with stride gridDim.x and load the
// next chunk into registers while the current one is in the register steps.
// PF == 0: one chunk per block.
#ifdef QSIM_TILE_WAVES
#define QSIM_TILE_ATTR __attribute__((amdgpu_waves_per_eu(QSIM_TILE_WAVES)))
#else
#define QSIM_TILE_ATTR
#endif
#ifndef QSIM_PREFETCH
#define QSIM_PREFETCH 0
#endif

template <typename R, int NV, int PF>
__global__ __launch_bounds__(kTPB)
__attribute__((amdgpu_waves_per_eu((NV == 1 && sizeof(R) == 4) ? 4 : 1))) QSIM_TILE_ATTR void k_tile_pass(cx<R> *__restrict__ psi, cx<R> *__restrict__ lam,
                                                    const PassDev *__restrict__ pd,
                                                    const StepDev *__restrict__ steps,
                                                    const cx<R> *__restrict__ mats,
                                                    double *__restrict__ probe_partial)
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

    Prefetch<PF> pf0;
    Prefetch<NV == 2 ? PF : 0> pf1;
    uint64_t c = blockIdx.x;
    if (PF > 0 && c < nchunks) {
        const uint64_t b0 = chunk_base(c, pd);
        pf0.load(psi, b0, rowoff, cf);
        if (NV == 2) pf1.load(lam, b0, rowoff, cf);
    }
    for (; c < nchunks; c += (PF > 0 ? gridDim.x : nchunks)) {
        const uint64_t base = chunk_base(c, pd);
        if (PF > 0) {
            pf0.to_lds(c0);
            if (NV == 2) pf1.to_lds(c1);
        } else {
            chunk_copy<R, true>(psi, c0, base, rowoff, E, cf);
            if (NV == 2) chunk_copy<R, true>(lam, c1, base, rowoff, E, cf);
        }
        if (NV == 2)
            for (int t = threadIdx.x; t < nprobe * kWaves; t += kTPB) pacc[t] = 0.0;
        __syncthreads();
        if (PF > 0) {
            const uint64_t nx = c + gridDim.x;
            if (nx < nchunks) {
                const uint64_t bn = chunk_base(nx, pd);
                pf0.load(psi, bn, rowoff, cf);
                if (NV == 2) pf1.load(lam, bn, rowoff, cf);
            }
        }
        run_steps<R, NV>(c0, c1, pd, steps, mats, pacc);
        chunk_copy<R, false>(psi, c0, base, rowoff, E, cf);
        if (NV == 2) {
            chunk_copy<R, false>(lam, c1, base, rowoff, E, cf);
            for (int p = threadIdx.x; p < nprobe; p += kTPB) {
                double v = 0.0;
                for (int w = 0; w < kWaves; ++w) v += pacc[p * kWaves + w];
                probe_partial[c * nprobe + p] = v;
            }
        }
        __syncthreads();
    }
}

constexpr int kPF = 8;  // prefetch vectors per thread: complex64 K=12, complex128 K=11

template <typename R, int NV, int PF>
static hipError_t tile_pass_launch(cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                                   const StepDev *d_steps, const cx<R> *d_mats, double *d_probe_partial,
                                   hipStream_t s)
{
    const size_t sm = (size_t)NV * (sizeof(cx<R>) << hp.K) + 8u * (1u << (hp.K - hp.cfast)) +
                      (NV == 2 ? 8u * kWaves * (size_t)hp.nprobe : 0u);
    const uint64_t nchunks = 1ull << (n - hp.K);
    uint64_t blocks = nchunks;
    if (PF > 0) {
        // enough persistent blocks to fill every CU at the kernel's occupancy
        static int cus = 0;
        if (!cus) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                cus = 256;
        }
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_pass<R, NV, PF>, kTPB, sm) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const uint64_t resident = (uint64_t)cus * per_cu;
        if (blocks > resident) blocks = resident;
    }
    hipLaunchKernelGGL((k_tile_pass<R, NV, PF>), dim3((unsigned)blocks), dim3(kTPB), sm, s, psi, lam, d_pass, d_steps,
                       d_mats, d_probe_partial);
    return hipGetLastError();
}

template <typename R, int NV>
static hipError_t tile_pass_nv(cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                               const StepDev *d_steps, const cx<R> *d_mats, double *d_probe_partial, hipStream_t s)
{
    // the prefetch path needs PF whole 16-byte vectors per thread
    const bool pf = QSIM_PREFETCH && (sizeof(cx<R>) == 8 ? (hp.K == 12 && hp.cfast >= 1) : (hp.K == 11));
    if (pf) return tile_pass_launch<R, NV, kPF>(psi, lam, n, hp, d_pass, d_steps, d_mats, d_probe_partial, s);
    return tile_pass_launch<R, NV, 0>(psi, lam, n, hp, d_pass, d_steps, d_mats, d_probe_partial, s);
}

template <typename R>
hipError_t launch_tile_pass(cx<R> *psi, cx<R> *lam, int n, const PassDev &hp, const PassDev *d_pass,
                            const StepDev *d_steps, const cx<R> *d_mats, double *d_probe_partial, hipStream_t s)
{
    if (lam) return tile_pass_nv<R, 2>(psi, lam, n, hp, d_pass, d_steps, d_mats, d_probe_partial, s);
    return tile_pass_nv<R, 1>(psi, nullptr, n, hp, d_pass, d_steps, d_mats, nullptr, s);
}

// ----------------------------------------------------------------------------
// FP64 reductions with Neumaier-compensated per-thread sums; fixed order,
// so results are deterministic.
// ----------------------------------------------------------------------------
struct NSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x)
    {
        const double t = s + x;
        c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

__device__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kWaves; ++w) t += sh[w];
    return t;  // valid in thread 0
}

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

// partial[b*2 + {0,1}] = sum conj(a_i) b_i over the block's grid-stride share
template <typename R>
__global__ __launch_bounds__(kTPB) void k_dot(const cx<R> *__restrict__ a, const cx<R> *__restrict__ b, uint64_t dim,
                                              double *__restrict__ partial)
{
    __shared__ double sh[kWaves];
    NSum re, im;
    const uint64_t stride = (uint64_t)gridDim.x * kTPB;
    for (uint64_t i = (uint64_t)blockIdx.x * kTPB + threadIdx.x; i < dim; i += stride) {
        const cx<R> x = a[i], y = b[i];
        const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
        re.add(xr * yr + xi * yi);
        im.add(xr * yi - xi * yr);
    }
    double t = block_sum(re.value(), sh);
    if (threadIdx.x == 0) partial[2 * blockIdx.x] = t;
    t = block_sum(im.value(), sh);
    if (threadIdx.x == 0) partial[2 * blockIdx.x + 1] = t;
}

template <typename R>
hipError_t launch_dot(const cx<R> *a, const cx<R> *b, uint64_t dim, double *d_partial, int *nblocks_out, hipStream_t s)
{
    uint64_t nb = (dim + kTPB * 8 - 1) / (kTPB * 8);
    if (nb > kDotBlocks) nb = kDotBlocks;
    if (nb < 1) nb = 1;
    *nblocks_out = (int)nb;
    hipLaunchKernelGGL(k_dot<R>, dim3((unsigned)nb), dim3(kTPB), 0, s, a, b, dim, d_partial);
    return hipGetLastError();
}

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
    const uint64_t g = blockIdx.x;
    uint64_t base = 0;
    for (int m = 0; m < pd->ncomp; ++m)
        if ((g >> m) & 1) base |= 1ull << pd->comp[m];
    __syncthreads();
    const int lowmask = (1 << cf) - 1;
    chunk_copy<R, true>(const_cast<cx<R> *>(psi), c0, base, rowoff, E, cf);
    __syncthreads();

    NSum er, ei;
    for (int e = threadIdx.x; e < E; e += kTPB) {
        const uint64_t addr = base + rowoff[e >> cf] + (e & lowmask);
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k < nterms; ++k) {
            const PauliDev T = terms[k];
            const cx<R> v = c0[swz(e ^ T.xloc)];
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
            const cx<R> p = c0[swz(e)];
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
        reinterpret_cast<const void *>(k_tile_pass<float, 1, 0>), reinterpret_cast<const void *>(k_tile_pass<float, 2, 0>),
        reinterpret_cast<const void *>(k_tile_pass<double, 1, 0>), reinterpret_cast<const void *>(k_tile_pass<double, 2, 0>),
        reinterpret_cast<const void *>(k_tile_pass<float, 1, kPF>), reinterpret_cast<const void *>(k_tile_pass<float, 2, kPF>),
        reinterpret_cast<const void *>(k_tile_pass<double, 1, kPF>), reinterpret_cast<const void *>(k_tile_pass<double, 2, kPF>),
        reinterpret_cast<const void *>(k_permute<float>), reinterpret_cast<const void *>(k_permute<double>),
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
                                            const StepDev *, const cx<R> *, double *, hipStream_t);         \
    template hipError_t launch_pauli_chunk<R>(const cx<R> *, cx<R> *, int, const PassDev &,                 \
                                              const PassDev *, const PauliDev *, int, int, double *,        \
                                              hipStream_t);                                                 \
    template hipError_t launch_dot<R>(const cx<R> *, const cx<R> *, uint64_t, double *, int *, hipStream_t); \
    template hipError_t launch_scale<R>(cx<R> *, uint64_t, double, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace qsim
