/*
 * Plain-C per-gate statevector simulator -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 variant (suffix _d): the checker.  FP32 variant (suffix _f): the CPU
 * port timed as the CPU baseline.  Amplitudes are interleaved (re, im).
 * Restates the index math of the reference kernels:
 *   1q pair  i0 = (k / 2^q) * 2^(q+1) + k % 2^q, i1 = i0 + 2^q
 *            (qgpusim_cuda/src/kernels.cu:33-38)
 *   2q quad  base = zero bits inserted at low and high positions, matrix index
 *            b_q0 + 2*b_q1 (kernels.cu:86-114, 132-136)
 *   permute  j = sum_p bit_p(i) << map[p]; out[j] = in[i] (kernels.cu:176-184)
 * Parallelised with OpenMP over pairs/quads/amplitudes; the tiled run
 * (svref_run_tiled_*) over blocks of the state.
 */
#include <stdint.h>
#include <string.h>

typedef struct { int32_t kind, q0, q1, mat_offset; } svref_op;

#define DEFINE_VARIANT(T, SUF)                                                 \
void svref_apply_1q_##SUF(T *psi, int n, int q, const T *U)                   \
{                                                                              \
    const uint64_t stride = 1ULL << q, pairs = (1ULL << n) >> 1;               \
    const T u0r = U[0], u0i = U[1], u1r = U[2], u1i = U[3];                    \
    const T u2r = U[4], u2i = U[5], u3r = U[6], u3i = U[7];                    \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int64_t k = 0; k < (int64_t)pairs; ++k) {                             \
        uint64_t i0 = ((uint64_t)k / stride) * 2 * stride + (uint64_t)k % stride; \
        uint64_t i1 = i0 + stride;                                             \
        T ar = psi[2 * i0], ai = psi[2 * i0 + 1];                              \
        T br = psi[2 * i1], bi = psi[2 * i1 + 1];                              \
        psi[2 * i0]     = u0r * ar - u0i * ai + u1r * br - u1i * bi;           \
        psi[2 * i0 + 1] = u0r * ai + u0i * ar + u1r * bi + u1i * br;           \
        psi[2 * i1]     = u2r * ar - u2i * ai + u3r * br - u3i * bi;           \
        psi[2 * i1 + 1] = u2r * ai + u2i * ar + u3r * bi + u3i * br;           \
    }                                                                          \
}                                                                              \
                                                                               \
void svref_apply_2q_##SUF(T *psi, int n, int q0, int q1, const T *U)          \
{                                                                              \
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;                  \
    const uint64_t mlo = 1ULL << lo, mhi = 1ULL << hi;                         \
    const uint64_t quads = (1ULL << n) >> 2;                                   \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int64_t kk = 0; kk < (int64_t)quads; ++kk) {                          \
        uint64_t k = (uint64_t)kk;                                             \
        uint64_t base = (k & (mlo - 1))                                        \
                      | (((k >> lo) & ((1ULL << (hi - lo - 1)) - 1)) << (lo + 1)) \
                      | ((k >> (hi - 1)) << (hi + 1));                         \
        /* idx[m] = memory index of matrix basis m = b_q0 + 2 b_q1 */          \
        uint64_t m0 = 1ULL << q0, m1 = 1ULL << q1;                             \
        uint64_t idx[4] = {base, base | m0, base | m1, base | m0 | m1};        \
        T vr[4], vi[4];                                                        \
        for (int m = 0; m < 4; ++m) { vr[m] = psi[2 * idx[m]]; vi[m] = psi[2 * idx[m] + 1]; } \
        for (int r = 0; r < 4; ++r) {                                          \
            T sr = 0, si = 0;                                                  \
            for (int c = 0; c < 4; ++c) {                                      \
                T ur = U[2 * (4 * r + c)], ui = U[2 * (4 * r + c) + 1];        \
                sr += ur * vr[c] - ui * vi[c];                                 \
                si += ur * vi[c] + ui * vr[c];                                 \
            }                                                                  \
            psi[2 * idx[r]] = sr; psi[2 * idx[r] + 1] = si;                    \
        }                                                                      \
    }                                                                          \
}                                                                              \
                                                                               \
void svref_permute_##SUF(const T *in, T *out, int n, const int32_t *map)      \
{                                                                              \
    const uint64_t dim = 1ULL << n;                                            \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int64_t ii = 0; ii < (int64_t)dim; ++ii) {                            \
        uint64_t i = (uint64_t)ii, j = 0;                                      \
        for (int p = 0; p < n; ++p) j |= ((i >> p) & 1ULL) << map[p];          \
        out[2 * j] = in[2 * i]; out[2 * j + 1] = in[2 * i + 1];                \
    }                                                                          \
}                                                                              \
                                                                               \
/* Run an op list (kind 1 = 1q on q0, kind 2 = 2q on (q0, q1)); matrices are   \
 * complex elements at mats + 2 * mat_offset. Returns the number of ops run. */ \
int svref_run_##SUF(T *psi, int n, const svref_op *ops, int nops, const T *mats) \
{                                                                              \
    for (int i = 0; i < nops; ++i) {                                           \
        const T *U = mats + 2 * (int64_t)ops[i].mat_offset;                    \
        if (ops[i].kind == 1) svref_apply_1q_##SUF(psi, n, ops[i].q0, U);      \
        else if (ops[i].kind == 2) svref_apply_2q_##SUF(psi, n, ops[i].q0, ops[i].q1, U); \
        else return -1;                                                        \
    }                                                                          \
    return nops;                                                               \
}                                                                              \
                                                                               \
/* One block of 2^b amplitudes (serial): 1q pair and 2q quad updates as      \
 * above, with every qubit below b. */                                         \
static void block_1q_##SUF(T *psi, int b, int q, const T *U)                   \
{                                                                              \
    const uint64_t stride = 1ULL << q, pairs = (1ULL << b) >> 1;               \
    for (uint64_t k = 0; k < pairs; ++k) {                                     \
        uint64_t i0 = (k / stride) * 2 * stride + k % stride, i1 = i0 + stride; \
        T ar = psi[2 * i0], ai = psi[2 * i0 + 1];                              \
        T br = psi[2 * i1], bi = psi[2 * i1 + 1];                              \
        psi[2 * i0]     = U[0] * ar - U[1] * ai + U[2] * br - U[3] * bi;       \
        psi[2 * i0 + 1] = U[0] * ai + U[1] * ar + U[2] * bi + U[3] * br;       \
        psi[2 * i1]     = U[4] * ar - U[5] * ai + U[6] * br - U[7] * bi;       \
        psi[2 * i1 + 1] = U[4] * ai + U[5] * ar + U[6] * bi + U[7] * br;       \
    }                                                                          \
}                                                                              \
                                                                               \
static void block_2q_##SUF(T *psi, int b, int q0, int q1, const T *U)          \
{                                                                              \
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;                  \
    const uint64_t mlo = 1ULL << lo, quads = (1ULL << b) >> 2;                 \
    const uint64_t m0 = 1ULL << q0, m1 = 1ULL << q1;                           \
    for (uint64_t k = 0; k < quads; ++k) {                                     \
        uint64_t base = (k & (mlo - 1))                                        \
                      | (((k >> lo) & ((1ULL << (hi - lo - 1)) - 1)) << (lo + 1)) \
                      | ((k >> (hi - 1)) << (hi + 1));                         \
        uint64_t idx[4] = {base, base | m0, base | m1, base | m0 | m1};        \
        T vr[4], vi[4];                                                        \
        for (int m = 0; m < 4; ++m) { vr[m] = psi[2 * idx[m]]; vi[m] = psi[2 * idx[m] + 1]; } \
        for (int r = 0; r < 4; ++r) {                                          \
            T sr = 0, si = 0;                                                  \
            for (int c = 0; c < 4; ++c) {                                      \
                T ur = U[2 * (4 * r + c)], ui = U[2 * (4 * r + c) + 1];        \
                sr += ur * vr[c] - ui * vi[c];                                 \
                si += ur * vi[c] + ui * vr[c];                                 \
            }                                                                  \
            psi[2 * idx[r]] = sr; psi[2 * idx[r] + 1] = si;                    \
        }                                                                      \
    }                                                                          \
}                                                                              \
                                                                               \
/* Tiled run (the CPU counterpart of an LQR tile whose local qubits sit in   \
 * the low bits): every op's qubits are below block_bits, so the state splits \
 * into blocks of 2^block_bits contiguous amplitudes, and each block runs the  \
 * whole op list while it is in the cache (one pass over memory for the list; \
 * blocks in parallel).  Returns nops, or -1 when an op reaches a qubit at or \
 * above block_bits or block_bits exceeds n. */                                \
int svref_run_tiled_##SUF(T *psi, int n, const svref_op *ops, int nops, const T *mats, int block_bits) \
{                                                                              \
    if (block_bits > n || block_bits < 2) return -1;                           \
    for (int i = 0; i < nops; ++i) {                                           \
        if (ops[i].kind != 1 && ops[i].kind != 2) return -1;                   \
        if (ops[i].q0 < 0 || ops[i].q0 >= block_bits) return -1;               \
        if (ops[i].kind == 2 && (ops[i].q1 < 0 || ops[i].q1 >= block_bits)) return -1; \
    }                                                                          \
    const int64_t nblocks = (int64_t)1 << (n - block_bits);                    \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int64_t blk = 0; blk < nblocks; ++blk) {                              \
        T *p = psi + 2 * ((uint64_t)blk << block_bits);                        \
        for (int i = 0; i < nops; ++i) {                                       \
            const T *U = mats + 2 * (int64_t)ops[i].mat_offset;                \
            if (ops[i].kind == 1) block_1q_##SUF(p, block_bits, ops[i].q0, U);  \
            else block_2q_##SUF(p, block_bits, ops[i].q0, ops[i].q1, U);       \
        }                                                                      \
    }                                                                          \
    return nops;                                                               \
}

DEFINE_VARIANT(double, d)
DEFINE_VARIANT(float, f)

int svref_num_threads(void)
{
    int t = 1;
#ifdef _OPENMP
    #pragma omp parallel
    {
        #pragma omp master
        t = omp_get_num_threads();
    }
#endif
    return t;
}
