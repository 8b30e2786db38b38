// Device helpers shared by the compiled kernels (qsim_kernels.hip) and the
// chunk-pass kernels generated at run time (qsim_jit.cpp): complex
// arithmetic, register-step gate slots, LDS swizzle and chunk copies.
// Self-contained apart from qsim_dev_types.h, so hiprtc can compile it.
#pragma once
#include "qsim_dev_types.h"

namespace qsim {

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

typedef float v4f __attribute__((ext_vector_type(4)));  // native vectors: plain loads/stores
typedef float v2f __attribute__((ext_vector_type(2)));
// one amplitude as raw bits (struct copies would go through memcpy)
template <typename R> struct RawAmp { typedef v2f type; };
template <> struct RawAmp<double> { typedef v4f type; };

constexpr int kBatch = 8;  // global accesses in flight per thread in the copy loops

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
// Blocks are dispatched to MI355X's 8 XCDs round-robin (block b -> XCD b % 8).
// Rotating the block index right by 3 bits over the nbits-bit chunk index
// hands each XCD one contiguous eighth of the chunks, walked in order; this
// lifts the chunk passes and the permutation by 5-20 % over the identity
// order on 33-qubit states (measured).
__device__ __forceinline__ uint64_t xcd_slab(uint64_t b, int nbits)
{
    if (nbits <= 3) return b;
    return ((b & 7ull) << (nbits - 3)) | (b >> 3);
}

// Global accesses of the copy loops; NT marks them non-temporal (streamed
// once: no reuse expected in the caches).
template <bool NT, typename V>
__device__ __forceinline__ V gload(const V *p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <bool NT, typename V>
__device__ __forceinline__ void gstore(V *p, V v)
{
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <typename R, bool LOAD, bool NT = false>
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
                    if (v < nv) buf[j] = gload<NT>(reinterpret_cast<const v4f *>(g + base + rowoff[e >> cf] + (e & lowmask)));
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
                    if (v < nv) gstore<NT>(reinterpret_cast<v4f *>(g + base + rowoff[e >> cf] + (e & lowmask)), buf[j]);
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
                    if (e < E) buf[j] = gload<NT>(gr + base + rowoff[e >> cf] + (e & lowmask));
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
                    if (e < E) gstore<NT>(gw + base + rowoff[e >> cf] + (e & lowmask), buf[j]);
                }
            }
        }
    }
}

__device__ __forceinline__ uint64_t chunk_base(uint64_t c, const PassDev *__restrict__ pd)
{
    uint64_t base = 0;
    for (int m = 0; m < pd->ncomp; ++m)
        if ((c >> m) & 1) base |= 1ull << pd->comp[m];
    return base;
}

}  // namespace qsim
