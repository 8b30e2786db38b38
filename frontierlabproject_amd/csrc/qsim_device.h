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

// exp(2 pi i t) for a phase t in turns held as a 32-bit fraction (t = u / 2^32,
// so sums of phases wrap exactly): sin and cos of pi x, x = 2t in [-1, 1).
__device__ __forceinline__ cx<float> cis_turns(uint32_t u)
{
    float s, c;
    sincospif((float)(int32_t)u * 0x1p-31f, &s, &c);
    return {c, s};
}

template <typename R>
__device__ __forceinline__ cx<R> cmad(cx<R> a, cx<R> b, cx<R> c)  // a*b + c
{
    return {fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}

// Read-only device data typed in the constant address space: a uniform
// address is then read with a scalar load (s_load) into scalar registers.
#define QSIM_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const QSIM_CONST T *as_const(const T *p)
{
    return (const QSIM_CONST T *)p;
}

// Complex element i of a matrix pointer of any address space (members are
// read one by one: an address-space-qualified struct cannot bind to the
// generic reference a struct copy takes).
template <typename P>
__device__ __forceinline__ auto ldc(P p, int i) -> cx<decltype(+p[0].re)>
{
    return {p[i].re, p[i].im};
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

#ifndef QSIM_KBATCH
#define QSIM_KBATCH 8
#endif
constexpr int kBatch = QSIM_KBATCH;  // global accesses in flight per thread in the copy loops

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

// One DPP move of a double (both halves) with row-local control CTRL; every
// lane active, every source lane inside its row.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the wave's 64 lanes without LDS round trips: within each row of
// 16 lanes by DPP (exchange with lane ^ 1, then ^ 2, then add the row
// rotated by 4 and by 8), then the four rows' sums, read from their first
// lanes, added in row order.  The order of every addition is fixed, so the
// result is deterministic; it is wave-uniform.  Needs every lane active.
__device__ __forceinline__ double wave_sum_rows(double v)
{
    v += dpp_f64<0xb1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4e>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x124>(v);  // row_ror:4
    v += dpp_f64<0x128>(v);  // row_ror:8
    return ((readlane_f64(v, 0) + readlane_f64(v, 16)) + readlane_f64(v, 32)) + readlane_f64(v, 48);
}

// A wave-uniform 64-bit value pinned to scalar registers, so that a lane
// offset added to a pointer plus it becomes the scalar-base + 32-bit-offset
// form of a global access (no 64-bit address arithmetic per access on the
// vector unit).
__device__ __forceinline__ uint64_t uniform_u64(uint64_t a)
{
    // an empty asm statement (no instruction) that takes and returns the
    // value in scalar registers: the optimiser can no longer split the sum
    // it came from across the vector address arithmetic
    asm volatile("" : "+s"(a));
    return a;
}

// LDS element at byte offset `byte` of `buf` (callers keep offsets in bytes
// so that xors with constant offsets need no scaling).
template <typename R>
__device__ __forceinline__ cx<R> &lds_at(cx<R> *buf, int byte)
{
    return *reinterpret_cast<cx<R> *>(reinterpret_cast<char *>(buf) + byte);
}

// 1q slot arithmetic on register bit B (applied to one vector).
template <typename R, int B, typename MP>
__device__ __forceinline__ void apply1(cx<R> *a, int kind, MP mp)
{
    if (kind & S1_SWAP) {  // U = V X: exchange each pair, then apply V
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const cx<R> t = a[m];
                a[m] = a[m | (1 << B)];
                a[m | (1 << B)] = t;
            }
        kind &= ~S1_SWAP;
    }
    switch (kind) {
    case S1_GEN: {
        const cx<R> u[4] = {ldc(mp, 0), ldc(mp, 1), ldc(mp, 2), ldc(mp, 3)};
        reg1<R, B>(a, u);
        break;
    }
    case S1_DIAG: {
        const cx<R> d0 = ldc(mp, 0), d1 = ldc(mp, 1);
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] = cmul(a[m], ((m >> B) & 1) ? d1 : d0);
        break;
    }
    case S1_PHASE: {
        const cx<R> d1 = ldc(mp, 0);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if ((m >> B) & 1) a[m] = cmul(a[m], d1);
        break;
    }
    case S1_REAL: {
        const cx<R> r01 = ldc(mp, 0), r23 = ldc(mp, 1);
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
        const cx<R> pq = ldc(mp, 0), rt = ldc(mp, 1);
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
    case S1_REAL1: {  // x' = x + a y, y' = b x + y
        const cx<R> ab = ldc(mp, 0);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{fma(ab.re, y.re, x.re), fma(ab.re, y.im, x.im)};
                a[m1] = cx<R>{fma(ab.im, x.re, y.re), fma(ab.im, x.im, y.im)};
            }
        break;
    }
    case S1_RI1: {  // x' = x + i q y, y' = i r x + y
        const cx<R> qr = ldc(mp, 0);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{fma(-qr.re, y.im, x.re), fma(qr.re, y.re, x.im)};
                a[m1] = cx<R>{fma(-qr.im, x.im, y.re), fma(qr.im, x.re, y.im)};
            }
        break;
    }
    case S1_REALG: {  // x' = x + a y, y' = b x + d y (a, b, d real)
        const cx<R> ab = ldc(mp, 0), d0 = ldc(mp, 1);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cx<R>{fma(ab.re, y.re, x.re), fma(ab.re, y.im, x.im)};
                a[m1] = cx<R>{fma(ab.im, x.re, d0.re * y.re), fma(ab.im, x.im, d0.re * y.im)};
            }
        break;
    }
    case S1_GEN1: {  // x' = x + b y, y' = c x + d y
        const cx<R> b = ldc(mp, 0), c = ldc(mp, 1), d = ldc(mp, 2);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (!((m >> B) & 1)) {
                const int m1 = m | (1 << B);
                const cx<R> x = a[m], y = a[m1];
                a[m] = cmad(b, y, x);
                a[m1] = cmad(c, x, cmul(d, y));
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

// Generic 1q forms (see S2_CU_FIRST): general 2x2 or diagonal.
template <typename R, int B, typename MP>
__device__ __forceinline__ void apply1g(cx<R> *a, int kind, MP mp)
{
    if (kind == S1_GEN) {
        const cx<R> u[4] = {ldc(mp, 0), ldc(mp, 1), ldc(mp, 2), ldc(mp, 3)};
        reg1<R, B>(a, u);
    } else if (kind == S1_DIAG) {
        const cx<R> d0 = ldc(mp, 0), d1 = ldc(mp, 1);
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] = cmul(a[m], ((m >> B) & 1) ? d1 : d0);
    }
}

// One 1q slot on register bit B.  u: the slot matrix's first four entries,
// already read; mp: the matrix itself (a probe's second matrix follows G).
template <typename R, int NV, int B, bool GENERIC, typename UP, typename MP>
__device__ __forceinline__ void slot1(cx<R> *a, cx<R> *l, int kind, UP u, MP mp, int probe, double *pacc, bool act)
{
    if (kind == S1_SKIP) return;
    if (kind == S1_PROBE || kind == S1_PROBE_GEN) {
        if (NV == 2) {
            const cx<R> g[4] = {ldc(u, 0), ldc(u, 1), ldc(u, 2), ldc(u, 3)};
            const double v = wave_sum_rows(act ? probe1<R, B>(a, l, g) : 0.0);
            if ((threadIdx.x & 63) == 0) pacc[probe * kWaves + (threadIdx.x >> 6)] += v;
            if (kind == S1_PROBE_GEN) {
                const cx<R> w[4] = {ldc(mp, 4), ldc(mp, 5), ldc(mp, 6), ldc(mp, 7)};
                reg1<R, B>(a, w);
                reg1<R, B>(l, w);
            }
        }
        return;
    }
    if (GENERIC) {
        apply1g<R, B>(a, kind, u);
        if (NV == 2) apply1g<R, B>(l, kind, u);
    } else {
        apply1<R, B>(a, kind, u);
        if (NV == 2) apply1<R, B>(l, kind, u);
    }
}

// 2x2 U (row-major) on register bit T over the register groups whose bit C
// is 1.
template <typename R, int C, int T, typename MP>
__device__ __forceinline__ void cu2(cx<R> *a, MP mp)
{
    const cx<R> u0 = ldc(mp, 0), u1 = ldc(mp, 1), u2 = ldc(mp, 2), u3 = ldc(mp, 3);
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (((m >> C) & 1) && !((m >> T) & 1)) {
            const int m1 = m | (1 << T);
            const cx<R> x = a[m], y = a[m1];
            a[m] = cmad(u0, x, cmul(u1, y));
            a[m1] = cmad(u2, x, cmul(u3, y));
        }
}

// pair slot arithmetic on register bits (F, S), F < S, index b_F + 2 b_S.
template <typename R, int F, int S, typename MP>
__device__ __forceinline__ void apply2(cx<R> *a, int kind, MP mp)
{
    switch (kind) {
    case S2_CPHASE: {
        const cx<R> d3 = ldc(mp, 0);
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
    case S2_CU_FIRST:
        cu2<R, F, S>(a, mp);
        break;
    case S2_CU_SECOND:
        cu2<R, S, F>(a, mp);
        break;
    default:
        break;
    }
}

// Generic pair forms (see S2_CU_FIRST): controlled phase or controlled 2x2.
template <typename R, int F, int S, typename MP>
__device__ __forceinline__ void apply2g(cx<R> *a, int kind, MP mp)
{
    if (kind == S2_CPHASE) {
        const cx<R> d3 = ldc(mp, 0);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            if (((m >> F) & 1) && ((m >> S) & 1)) a[m] = cmul(a[m], d3);
    } else if (kind == S2_CU_FIRST) {
        cu2<R, F, S>(a, mp);
    } else if (kind == S2_CU_SECOND) {
        cu2<R, S, F>(a, mp);
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
template <typename R, int NV, int F, int S, bool GENERIC, typename MP>
__device__ __forceinline__ void slot2(cx<R> *a, cx<R> *l, int kind, MP mp)
{
    if (kind == S2_SKIP) return;
    if (GENERIC) {
        apply2g<R, F, S>(a, kind, mp);
        if (NV == 2) apply2g<R, F, S>(l, kind, mp);
    } else {
        apply2<R, F, S>(a, kind, mp);
        if (NV == 2) apply2<R, F, S>(l, kind, mp);
    }
}

// ----------------------------------------------------------------------------
// Chunk pass.  Each block gathers one chunk (2^K amplitudes over the pass's
// bit set) into LDS, runs the register steps, writes it back in place.
// NV = 1: forward on psi.  NV = 2: reverse sweep on (psi, lam) with gradient
// probes Im<lam|G|psi> reduced per block.
// ----------------------------------------------------------------------------

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

// TPB is a power of two, so element (v0 + j * TPB) + threadIdx.x splits into
// disjoint bit fields and its slot is swz(uniform part) ^ swz(thread part).
template <typename R, bool LOAD, bool NT = false, int TPB = kTPB>
__device__ __forceinline__ void chunk_copy(cx<R> *__restrict__ g, cx<R> *__restrict__ lds, uint64_t base,
                                           const uint64_t *__restrict__ rowoff, int E, int cf)
{
    const int lowmask = (1 << cf) - 1;
    if (sizeof(cx<R>) == 8 && cf >= 1) {
        const int nv = E >> 1;
        const int st = swzr<R>((int)threadIdx.x << 1);
        for (int v0 = 0; v0 < nv; v0 += TPB * kBatch) {
            v4f buf[kBatch];
            if (LOAD) {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * TPB + threadIdx.x;
                    const int e = v << 1;
                    if (v < nv) buf[j] = gload<NT>(reinterpret_cast<const v4f *>(g + base + rowoff[e >> cf] + (e & lowmask)));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * TPB + threadIdx.x;
                    if (v < nv) *reinterpret_cast<v4f *>(lds + (st ^ swzr<R>((v0 + j * TPB) << 1))) = buf[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * TPB + threadIdx.x;
                    if (v < nv) buf[j] = *reinterpret_cast<const v4f *>(lds + (st ^ swzr<R>((v0 + j * TPB) << 1)));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int v = v0 + j * TPB + threadIdx.x;
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
        const int st = swzr<R>((int)threadIdx.x);
        for (int e0 = 0; e0 < E; e0 += TPB * kBatch) {
            raw buf[kBatch];
            if (LOAD) {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * TPB + threadIdx.x;
                    if (e < E) buf[j] = gload<NT>(gr + base + rowoff[e >> cf] + (e & lowmask));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * TPB + threadIdx.x;
                    if (e < E) lr[st ^ swzr<R>(e0 + j * TPB)] = buf[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * TPB + threadIdx.x;
                    if (e < E) buf[j] = lr[st ^ swzr<R>(e0 + j * TPB)];
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int e = e0 + j * TPB + threadIdx.x;
                    if (e < E) gstore<NT>(gw + base + rowoff[e >> cf] + (e & lowmask), buf[j]);
                }
            }
        }
    }
}

template <typename PP>
__device__ __forceinline__ uint64_t chunk_base(uint64_t c, PP pd)
{
    uint64_t base = 0;
    for (int m = 0; m < pd->ncomp; ++m)
        if ((c >> m) & 1) base |= 1ull << pd->comp[m];
    return base;
}

// Copy-in of a complex64 chunk whose source is in another layout (a layout
// change folded into the pass).  Read index r enumerates the chunk in source
// order: rows of 2^cf contiguous source amplitudes at base + rowoff[r >> cf];
// sigma(r) is the amplitude's chunk index in the pass layout, and bit jb0 of
// it is source bit 0, so the two amplitudes of one 16-byte read land at
// sigma(r) and sigma(r) | 2^jb0.  Needs cf >= 1.
template <int TPB, typename Sigma>
__device__ __forceinline__ void chunk_copy_in_mapped(const cx<float> *__restrict__ g, cx<float> *__restrict__ lds,
                                                     uint64_t base, const uint64_t *__restrict__ rowoff, int E,
                                                     int cf, Sigma sigma, int jb0)
{
    const int lowmask = (1 << cf) - 1, nv = E >> 1;
    // sigma and swz are linear: the slot of a read index splits like its bits
    const int st = swz(sigma((int)threadIdx.x << 1)), s1 = swz(1 << jb0);
    for (int v0 = 0; v0 < nv; v0 += TPB * kBatch) {
        v4f buf[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int v = v0 + j * TPB + threadIdx.x;
            const int r = v << 1;
            if (v < nv) buf[j] = gload<true>(reinterpret_cast<const v4f *>(g + base + rowoff[r >> cf] + (r & lowmask)));
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int v = v0 + j * TPB + threadIdx.x;
            if (v < nv) {
                const int sl = st ^ swz(sigma((v0 + j * TPB) << 1));
                lds[sl] = cx<float>{buf[j].x, buf[j].y};
                lds[sl ^ s1] = cx<float>{buf[j].z, buf[j].w};
            }
        }
    }
}

// complex128 form: one 16-byte amplitude per read, so any source order works;
// read index r sits at base + rowoff[r >> cf] + (r & (2^cf - 1)) and lands at
// chunk index sigma(r).
template <int TPB, typename Sigma>
__device__ __forceinline__ void chunk_copy_in_mapped_d(const cx<double> *__restrict__ g, cx<double> *__restrict__ lds,
                                                       uint64_t base, const uint64_t *__restrict__ rowoff, int E,
                                                       int cf, Sigma sigma)
{
    typedef typename RawAmp<double>::type raw;
    const raw *gr = reinterpret_cast<const raw *>(g);
    raw *lr = reinterpret_cast<raw *>(lds);
    const int lowmask = (1 << cf) - 1;
    const int st = swz_d(sigma((int)threadIdx.x));
    for (int r0 = 0; r0 < E; r0 += TPB * kBatch) {
        raw buf[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int r = r0 + j * TPB + threadIdx.x;
            if (r < E) buf[j] = gload<true>(gr + base + rowoff[r >> cf] + (r & lowmask));
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int r = r0 + j * TPB + threadIdx.x;
            if (r < E) lr[st ^ swz_d(sigma(r0 + j * TPB))] = buf[j];
        }
    }
}

// The register steps of one pass over the chunk staged in LDS (c0, and c1 for
// the second vector).  Ends with a barrier after every step.  The pass, step
// and matrix pointers may be generic or typed in the constant address space
// (QSIM_CONST: the descriptors are then read with scalar loads).  GENERIC:
// the plan uses only the generic slot forms (see S2_CU_FIRST).
template <typename R, int NV, bool GENERIC, typename PP, typename SP, typename MP>
__device__ __forceinline__ void run_steps(cx<R> *c0, cx<R> *c1, PP pd, SP steps, MP mats, double *pacc)
{
    const int E = 1 << pd->K;
    const int ns = pd->nsteps;
    const SP st = steps + pd->step_off;
    const int ngroups = E >> kG;
    for (int si = 0; si < ns; ++si) {
        const SP sd = st + si;
        const int r0 = sd->bits[0], r1 = sd->bits[1], r2 = sd->bits[2], r3 = sd->bits[3];
        const int s0 = sd->sbits[0], s1 = sd->sbits[1], s2 = sd->sbits[2], s3 = sd->sbits[3];
        const int nr = sd->nrounds, dn = sd->dense, sc = sd->scale;
        // swz is XOR-linear and the group base and register offsets occupy
        // disjoint bits, so each element's LDS slot is swzr<R>(base) ^ swzr<R>(offset):
        // one XOR per element with a wave-uniform constant.
        // (the 16 offsets are XOR combinations of the four register bits' slots)
        int moff[16];
        const int b0 = swzr_bit<R>(r0), b1 = swzr_bit<R>(r1), b2 = swzr_bit<R>(r2), b3 = swzr_bit<R>(r3);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            moff[m] = ((m & 1) ? b0 : 0) ^ ((m & 2) ? b1 : 0) ^ ((m & 4) ? b2 : 0) ^ ((m & 8) ? b3 : 0);
        if constexpr (GENERIC && NV == 1 && sizeof(R) == 4) {
            // fused step of a complex64 one-vector pass: the matrix cores
            if (dn >= 0) {
                dense16_mfma<kTPB>(c0, mats + dn, moff, s0, s1, s2, s3, ngroups, sc >= 0 ? mats + sc : MP(nullptr));
                __syncthreads();
                continue;
            }
        }
        // wave-uniform trip count: probes reduce across all 64 lanes
        for (int gb = 0; gb < ngroups; gb += kTPB) {
            const int gi = gb + threadIdx.x;
            const bool act = gi < ngroups;
            // idle lanes read a valid group and never store
            const uint32_t g32 = (uint32_t)(act ? gi : ngroups - 1);
            const int sb = swzr<R>((int)ins0(ins0(ins0(ins0(g32, s0), s1), s2), s3));
            cx<R> a[16], l[NV == 2 ? 16 : 1];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                a[m] = c0[sb ^ moff[m]];
                if (NV == 2) l[m] = c1[sb ^ moff[m]];
            }
            if (dn >= 0) {
                // rows go straight back to LDS: no second 16-amplitude array
                const cx<R> f = sc >= 0 ? ldc(mats, sc) : cx<R>{R(1), R(0)};
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const MP M = mats + dn + 16 * r;
                    cx<R> acc = cmul(ldc(M, 0), a[0]);
#pragma unroll
                    for (int c = 1; c < 16; ++c) acc = cmad(ldc(M, c), a[c], acc);
                    if (sc >= 0) acc = cmul(acc, f);
                    if (act) c0[sb ^ moff[r]] = acc;
                }
                if (NV == 2) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const MP M = mats + dn + 16 * r;
                        cx<R> acc = cmul(ldc(M, 0), l[0]);
#pragma unroll
                        for (int c = 1; c < 16; ++c) acc = cmad(ldc(M, c), l[c], acc);
                        if (sc >= 0) acc = cmul(acc, f);
                        if (act) c1[sb ^ moff[r]] = acc;
                    }
                }
                continue;
            } else {
                // one copy of the round code, run nrounds times.  Every
                // slot's kind and matrix pointer, and the first four entries
                // of each 1q slot's matrix, are read before any slot runs, so
                // the loads are in flight together (matrix arenas end with at
                // least 256 bytes of padding).
#pragma unroll 1
                for (int r = 0; r < nr; ++r) {
                    const int b = r * kSlotsPerRound;
                    int kd[kSlotsPerRound];
                    MP mp[kSlotsPerRound];
                    cx<R> u[4][4];
#pragma unroll
                    for (int j = 0; j < kSlotsPerRound; ++j) {
                        kd[j] = sd->kind[b + j];
                        mp[j] = mats + sd->mat[b + j];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int i = 0; i < 4; ++i) u[j][i] = ldc(mp[j], i);
                    slot1<R, NV, 0, GENERIC>(a, l, kd[0], u[0], mp[0], sd->probe[4 * r + 0], pacc, act);
                    slot1<R, NV, 1, GENERIC>(a, l, kd[1], u[1], mp[1], sd->probe[4 * r + 1], pacc, act);
                    slot1<R, NV, 2, GENERIC>(a, l, kd[2], u[2], mp[2], sd->probe[4 * r + 2], pacc, act);
                    slot1<R, NV, 3, GENERIC>(a, l, kd[3], u[3], mp[3], sd->probe[4 * r + 3], pacc, act);
                    slot2<R, NV, 0, 1, GENERIC>(a, l, kd[4], mp[4]);
                    slot2<R, NV, 2, 3, GENERIC>(a, l, kd[5], mp[5]);
                    slot2<R, NV, 0, 2, GENERIC>(a, l, kd[6], mp[6]);
                    slot2<R, NV, 1, 3, GENERIC>(a, l, kd[7], mp[7]);
                    slot2<R, NV, 0, 3, GENERIC>(a, l, kd[8], mp[8]);
                    slot2<R, NV, 1, 2, GENERIC>(a, l, kd[9], mp[9]);
                }
            }
            if (sc >= 0) {  // scalar factored out of earlier gates
                const cx<R> f = ldc(mats, sc);
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


// Dense 16x16 step of a complex64 chunk on the matrix cores.  The wave's 64
// register groups are the columns of X (16 x 64, row = register index m), and
// X <- M X runs as 16-column blocks of three real 16x16x4 FP32 MFMA products
// each (3M): P1 = Mr Xr, P2 = Mi Xi, P3 = (Mr + Mi)(Xr + Xi);
// Re = P1 - P2, Im = P3 - P1 - P2, FP32 accumulation.
// v_mfma_f32_16x16x4f32 lane layout: A[i][k] from lane i + 16 k, B[k][j] from
// lane j + 16 k, D[4 q + r][j] in lane j + 16 q, element r.
// Lane j + 16 q therefore reads rows 4 kb + q (kb = 0..3) of group
// (block base + j) and writes rows 4 q + r of the same group: every group is
// read and written by one lane set of one wave, in program order.
// moff[m]: LDS slot offset of register index m; s0..s3: the step's register
// bits ascending; ng: register groups in the chunk; sc: optional factor.
template <int TPB = kTPB, typename MP = const cx<float> *>
__device__ __forceinline__ void dense16_mfma(cx<float> *c0, MP M, const int *moff, int s0, int s1, int s2, int s3,
                                             int ng, MP sc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, q = lane >> 4;
    float ar[4], ai[4], as[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const cx<float> m = ldc(M, 16 * j + 4 * kb + q);
        ar[kb] = m.re;
        ai[kb] = m.im;
        as[kb] = m.re + m.im;
    }
    const cx<float> f = sc ? ldc(sc, 0) : cx<float>{1.f, 0.f};
    for (int gb = wave * 64; gb < ng; gb += TPB) {  // wave-uniform trip count
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int g = gb + 16 * nb + j;
            const bool act = g < ng;
            const int sb = swz((int)ins0(ins0(ins0(ins0((uint32_t)(act ? g : ng - 1), s0), s1), s2), s3));
            float xr[4], xi[4];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const cx<float> v = c0[sb ^ moff[4 * kb + q]];
                xr[kb] = v.re;
                xi[kb] = v.im;
            }
            v4f p1 = {0.f, 0.f, 0.f, 0.f}, p2 = p1, p3 = p1;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[kb], xr[kb], p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[kb], xi[kb], p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(as[kb], xr[kb] + xi[kb], p3, 0, 0, 0);
            }
            if (act) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cx<float> o{p1[r] - p2[r], p3[r] - p1[r] - p2[r]};
                    if (sc) o = cmul(o, f);
                    c0[sb ^ moff[4 * q + r]] = o;
                }
            }
        }
    }
}

// Dense 32x32 step (five register bits) of a complex64 chunk on the matrix
// cores: X (32 x groups) <- M X as two 16-row halves, each the sum of eight
// 16x16x4 k-blocks of the three 3M products.  Lane j + 16 q reads rows
// 4 kb + q (kb = 0..7) of group (block base + j) and writes rows
// 16 h + 4 q + r of it, all reads before any write.  Waves take blocks of 16
// groups in turn, so a chunk of 2^12 amplitudes (128 groups) keeps every wave
// of a 256-thread workgroup busy.  moff[m]: LDS slot offset of register index
// m (0..31); s0..s4: the register bits ascending; ng: groups in the chunk.
template <int TPB = kTPB, typename MP = const cx<float> *>
__device__ __forceinline__ void dense32_mfma(cx<float> *c0, MP M, const int *moff, int s0, int s1, int s2, int s3,
                                             int s4, int ng, MP sc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, q = lane >> 4;
    float ar[2][8], ai[2][8], as[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const cx<float> m = ldc(M, 32 * (16 * h + j) + 4 * kb + q);
            ar[h][kb] = m.re;
            ai[h][kb] = m.im;
            as[h][kb] = m.re + m.im;
        }
    const cx<float> f = sc ? ldc(sc, 0) : cx<float>{1.f, 0.f};
    for (int gb = wave * 16; gb < ng; gb += TPB / 4) {  // wave-uniform trip count
        const int g = gb + j;
        const bool act = g < ng;
        const int sb =
            swz((int)ins0(ins0(ins0(ins0(ins0((uint32_t)(act ? g : ng - 1), s0), s1), s2), s3), s4));
        float xr[8], xi[8], xs[8];
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const cx<float> v = c0[sb ^ moff[4 * kb + q]];
            xr[kb] = v.re;
            xi[kb] = v.im;
            xs[kb] = v.re + v.im;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            v4f p1 = {0.f, 0.f, 0.f, 0.f}, p2 = p1, p3 = p1;
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[h][kb], xr[kb], p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[h][kb], xi[kb], p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(as[h][kb], xs[kb], p3, 0, 0, 0);
            }
            if (act) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cx<float> o{p1[r] - p2[r], p3[r] - p1[r] - p2[r]};
                    if (sc) o = cmul(o, f);
                    c0[sb ^ moff[16 * h + 4 * q + r]] = o;
                }
            }
        }
    }
}

// FP64 reductions with Neumaier-compensated per-thread sums; fixed order,
// so results are deterministic.
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

__device__ __forceinline__ double block_sum(double v, double *sh)
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

}  // namespace qsim
