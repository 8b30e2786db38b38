/*
 * qgpusim -- C ABI of the MI355X (gfx950) statevector backend.
 *
 * This is the drop-in boundary that replaces the reference's pybind11 module
 * `qgpusim_cuda` (qgpusim/backends/custom/qgpusim_cuda/src/bindings.cpp:1004-1092).
 * Every entry point returns 0 on success and a non-zero status on failure;
 * qsim_last_error() then holds the message (the text the reference raised as
 * RuntimeError).  Entry points that take a qsim_state run asynchronously on the
 * state's own HIP stream unless stated otherwise.  A state is not thread-safe:
 * use it from one thread at a time.
 *
 * Amplitude layout: index i = sum_q b_q 2^q, complex64 (interleaved float re,
 * im) or complex128 (interleaved double).  Matrices are row-major in the
 * state's precision; a 4x4 matrix on (q0, q1) is little-endian: row/column
 * index b_q0 + 2*b_q1 (kernels.cu:54-55).
 */
#ifndef QGPUSIM_H
#define QGPUSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QSIM_C64 = 0, QSIM_C128 = 1 };

/* tile_mode of run_custom_backend (qgpusim/backends/custom/backend.py:460-473) */
enum { QSIM_SEQUENTIAL = 0, QSIM_GRAPHED = 1, QSIM_COOPERATIVE = 2 };

enum { QSIM_OP_1Q = 1, QSIM_OP_2Q = 2, QSIM_OP_PROBE = 3, QSIM_OP_3Q = 4, QSIM_OP_4Q = 5 };

typedef struct qsim_state qsim_state;

/* One op of a tile.  Replaces the ('1q', q0, U_dev) / ('2q', q0, q1, U_dev)
 * tuples parsed at bindings.cpp:182-220: the matrix is `mat_offset` complex
 * elements into the matrix arena passed beside the ops.  QSIM_OP_PROBE (only
 * for qsim_adjoint) uses q1 as its output slot.  QSIM_OP_3Q / QSIM_OP_4Q (no
 * reference counterpart: the reference drops gates on three or more qubits,
 * qgpusim/backends/custom/backend.py:448-451) apply an 8x8 / 16x16 matrix to
 * qubits (q0, q1 & 255, (q1 >> 8) & 255[, (q1 >> 16) & 255]), row and column
 * index sum_i b_{q_i} 2^i; in the chunked modes each runs as one dense
 * register step. */
typedef struct {
    int32_t kind, q0, q1, mat_offset;
} qsim_op;

/* Pauli term coef * i^{popcount(x&z)} X^x Z^z (Y where x and z both set). */
typedef struct {
    uint64_t xmask, zmask;
    double coef_re, coef_im;
} qsim_pauli;

/* Message of the last failed call on this thread ("" if none). */
const char *qsim_last_error(void);

/* ABI version (incremented on incompatible changes). */
int qsim_abi_version(void);

/* Statevector(num_qubits) / StatevectorF32(num_qubits):
 * bindings.cpp:92-110 and :645-658.  Allocates the amplitudes on the current
 * HIP device and sets |0...0>. */
int qsim_create(int num_qubits, int dtype, qsim_state **out);

/* ~Statevector (bindings.cpp:112-122). */
int qsim_destroy(qsim_state *st);

/* num_qubits / dim properties (bindings.cpp:622-623). */
int qsim_num_qubits(const qsim_state *st);
uint64_t qsim_dim(const qsim_state *st);
int qsim_dtype(const qsim_state *st);

/* Device address of the amplitudes (changes after qsim_permute). */
int qsim_device_ptr(qsim_state *st, void **out);

/* reset() -> |0...0> (bindings.cpp:577-582).  Written lazily: the next op
 * list in a chunked mode with specialised kernels generates |0...0> in its
 * first pass instead of reading it; any other use of the state writes it
 * first.  qsim_create starts in this state. */
int qsim_reset(qsim_state *st);

/* from_numpy / to_numpy (bindings.cpp:547-574).  `count` is the number of
 * amplitudes and must equal the state's dimension.  Both synchronise. */
int qsim_from_host(qsim_state *st, const void *src, uint64_t count);
int qsim_to_host(qsim_state *st, void *dst, uint64_t count);

/* One amplitude, copied to the host after the state's pending work.  No
 * reference counterpart (used for fidelity checks without a full copy). */
int qsim_get_amplitude(qsim_state *st, uint64_t index, double *re, double *im);

/* synchronize() (bindings.cpp:585-587): applies a pending layout change,
 * then waits for the state's stream. */
int qsim_synchronize(qsim_state *st);

/* Host only, no device work: map[p_old] = new position of the qubit held at
 * p_old, with the validation order and messages of bindings.cpp:435-464. */
int qsim_permute_map(int num_qubits, const int32_t *layout_old, const int32_t *layout_new,
                     int32_t *map_out);

/* permute(layout_old, layout_new) (bindings.cpp:434-483), with the same map
 * and errors.  No data moves: the layout change is composed into the state's
 * stored bit order (which may differ from the logical layout).  An op list
 * in a chunked mode with specialised kernels runs on the stored order, its
 * qubits mapped through it; when some of them are stored high, its first
 * chunk pass reads the old stored order and writes one where they sit low
 * (only they and the qubits they displace move) into the second buffer.  Any
 * other use of the state (reads, reductions, single gates, the sequential
 * mode, kernels compiled ahead of time) first restores the logical layout
 * with one LDS-tiled gather/scatter pass into the second buffer.  Results are
 * those of an immediate permutation. */
int qsim_permute(qsim_state *st, const int32_t *layout_old, const int32_t *layout_new, int num_qubits);

/* Host only: the integer checks of parse_tile_ops (bindings.cpp:196-216). */
int qsim_validate_ops(int num_qubits, const qsim_op *ops, int nops, uint64_t mat_count);

/* apply_tile_dev / apply_tile_graphed / apply_tile_cooperative
 * (bindings.cpp:223-240, 244-306, 312-373).
 *   QSIM_SEQUENTIAL : one coalesced 1q/2q stream kernel per op;
 *   QSIM_COOPERATIVE: the ops are planned into chunk passes; each pass stages
 *                     2^K-amplitude chunks in LDS and runs all its ops there;
 *   QSIM_GRAPHED    : the chunk passes captured into a hipGraph, cached on the
 *                     (buffer, ops, matrices) signature and replayed on a hit.
 * `mats` holds `mat_count` complex elements in the state's precision. */
int qsim_apply_ops(qsim_state *st, int mode, const qsim_op *ops, int nops, const void *mats,
                   uint64_t mat_count);

/* apply_tile_* followed by permute(layout_old, layout_new): no reference
 * counterpart as one call; it replaces the pair of calls the tile loop of
 * run_custom_backend makes for its last tile with ops and the layout changes
 * after it (qgpusim/backends/custom/backend.py:417-473; bindings.cpp:434-483).
 * Results, errors and the state afterwards are those of qsim_apply_ops then
 * qsim_permute.  In a chunked mode with specialised kernels the last chunk
 * pass writes its output in the new layout into the second buffer (no
 * separate permutation pass), when that pass has room for the chunk bits
 * that become the output's contiguous rows; otherwise the change is recorded
 * as by qsim_permute. */
int qsim_apply_ops_permute(qsim_state *st, int mode, const qsim_op *ops, int nops, const void *mats,
                           uint64_t mat_count, const int32_t *layout_old, const int32_t *layout_new, int num_qubits);

/* apply_1q / apply_2q with a host matrix (bindings.cpp:499-544). */
int qsim_apply_1q(qsim_state *st, int target_qubit, const void *U);
int qsim_apply_2q(qsim_state *st, int q0, int q1, const void *U);

/* apply_diagonal_1q (bindings.cpp:611-619). */
int qsim_apply_diagonal_1q(qsim_state *st, int target_qubit, double p0_re, double p0_im,
                           double p1_re, double p1_im);

/* Module functions apply_1q_gate / apply_2q_gate on a caller-owned device
 * vector (bindings.cpp:931-1001).  `stream` may be NULL (default stream).
 * The call returns once the kernel has completed. */
int qsim_apply_1q_raw(void *psi, int dtype, int num_qubits, int target_qubit, const void *U,
                      void *stream);
int qsim_apply_2q_raw(void *psi, int dtype, int num_qubits, int q0, int q1, const void *U,
                      void *stream);

/* ---- No counterpart in the reference: required by the benchmark workloads. ---- */

/* <psi|psi> in FP64 (Neumaier-compensated, fixed order).  Synchronises.  A
 * pending layout change stays pending (the norm does not depend on it): the
 * call reads the state once and writes nothing. */
int qsim_norm2(qsim_state *st, double *out);

/* <a|b> in FP64.  Both states need the same size; they may differ in
 * precision (a complex64 result judged against its complex128 run).
 * Synchronises. */
int qsim_overlap(qsim_state *a, qsim_state *b, double *re, double *im);

/* Scale every amplitude by `factor` (periodic renormalisation). */
int qsim_scale(qsim_state *st, double factor);

/* <psi| sum_t P_t |psi> in FP64 over chunk passes.  Synchronises. */
int qsim_expect_pauli(qsim_state *st, const qsim_pauli *terms, int nterms, double *re, double *im);

/* Adjoint gradient.  lam = H psi (with <psi|H|psi> in FP64), then the reverse
 * op list (inverse gates, each parameterised gate preceded by a
 * QSIM_OP_PROBE carrying its generator G) runs over (psi, lam) in chunk
 * passes; probe slot k receives Im<lam|G|psi> at its position, reduced in
 * FP64.  psi is left at the circuit's input state.  Uses the state's second
 * buffer for lam.  Synchronises. */
int qsim_adjoint(qsim_state *st, const qsim_pauli *terms, int nterms, const qsim_op *rev_ops,
                 int nops, const void *mats, uint64_t mat_count, int nprobes, double *probes_out,
                 double *expect_out);

/* Chunk-pass geometry: K chunk bits, cmin always-present low bits.  Defaults:
 * K = 13, cmin = 4 for complex64; K = 11, cmin = 3 for complex128.  Ranges:
 * 6 <= K <= 13 (12 for complex128), 0 <= cmin <= K - 2.  Two-vector adjoint
 * passes use at most K = 12 (complex64) or 11 (complex128). */
int qsim_set_tile_params(qsim_state *st, int K, int cmin);

/* No reference counterpart (the reference has one fixed kernel per gate).
 * 1 (default; environment QSIM_JIT=0 switches it off at qsim_create):
 * forward chunk passes run kernels specialised per pass, with the pass's
 * register bits, LDS offsets and gate kinds compiled in by hiprtc on first
 * use and cached per device for the life of the process.  0: every pass runs
 * the compiled slot-program kernel, on a plan whose gates take generic
 * arithmetic forms (general or diagonal 2x2, controlled phase or controlled
 * 2x2).  Both apply the same gates in the same order; results agree to
 * rounding. */
int qsim_set_jit(qsim_state *st, int enable);

/* Host only, no reference counterpart: plan `ops` as qsim_apply_ops would (as
 * the adjoint's reverse sweep would, on psi and lam, when `ops` hold
 * QSIM_OP_PROBE entries; with a pending layout change in_map[p] = new
 * position of bit p, forward only, or NULL; with a layout change after the
 * list out_map[p] = output position of bit p, as qsim_apply_ops_permute
 * would, or NULL) and compile, without loading, every specialised pass
 * kernel for gfx950.  *out_folded (may be NULL) tells whether the last pass
 * writes the out_map order.  Returns the number of passes, or -1 with the
 * compiler log. */
int qsim_jit_compile_check(int num_qubits, int dtype, const qsim_op *ops, int nops,
                           const void *mats, uint64_t mat_count, int K, int cmin, int fuse,
                           const int32_t *in_map, const int32_t *out_map, int *out_folded);

/* Host only, no reference counterpart: the stored bit order an op list would
 * run on (the layout change its first chunk pass folds in).  perm_map[p] is
 * the logical bit held by stored bit p; dtype sets the low region (2^18
 * complex64 or 2^17 complex128 amplitudes, 2 MiB); cmin is the row bits.
 * Writes tgt_out[q] = stored position of logical bit q: the op qubits end
 * below max(#op qubits, region), the row bits holding op qubits first; only
 * the incoming qubits and those they displace move.  Returns 0, or -1. */
int qsim_choose_layout(int num_qubits, int dtype, int cmin, const int32_t *perm_map, const qsim_op *ops, int nops,
                       int32_t *tgt_out);

/* No reference counterpart.  Fusion mode of forward chunk passes.  A fused
 * register step applies one 4-qubit (16x16) unitary, the complex128 product
 * of its gates rounded once, as a complex GEMM on the FP32 matrix cores
 * (three real 16x16x4 MFMA products per 16 columns, FP32 accumulation) when
 * the pass kernel is specialised and the state is complex64, else by the
 * vector units.
 * 0 (default): slot programs (each gate by its own cheap arithmetic); a step
 *    is fused when its gates cost more than a dense product on the vector
 *    units (64 multiply-adds per amplitude);
 * 1: every register step is fused;
 * 2: a step is fused when its gates cost more than the product on the matrix
 *    cores (32 vector multiply-adds per amplitude);
 * 3: every register step is fused over five register bits: one 5-qubit
 *    (32x32) unitary per step, two 16-row halves of eight 16x16x4 k-blocks
 *    each on the matrix cores.  Specialised complex64 kernels only: with the
 *    run-time compiler off (qsim_set_jit) or a complex128 state, mode 3 runs
 *    as mode 1;
 * 4: as 2, except that each step is first grown as one 32x32 unitary over
 *    five register bits and kept when the ops it takes would cost more than
 *    bar5 (qsim_set_fusion_bars) on the vector units; complex64 specialised
 *    kernels only, else as 2. */
int qsim_set_fusion(qsim_state *st, int mode);

/* No reference counterpart.  Cost bars of fusion modes 2 and 4, in real
 * multiply-adds per amplitude (0: the defaults, 32 and 64): bar, the slot
 * arithmetic above which a 4-qubit step is fused; bar5, the vector work of
 * the ops (four per nonzero of an average matrix row, summed over the ops)
 * above which mode 4 keeps a 32x32 step.  Each in [0, 4096]. */
int qsim_set_fusion_bars(qsim_state *st, double bar, double bar5);

/* No reference counterpart (its custom backend has no sampling task):
 * `shots` computational-basis outcomes drawn from |psi|^2 / <psi|psi>, as
 * amplitude indices in the state's current bit order, grouped by index
 * block (not in draw order).  Two levels, all in FP64: per-block
 * probability sums over blocks of 2^13 amplitudes on the GPU; each shot's
 * block and residual from a splitmix64 stream keyed by `seed` on the host;
 * the exact index by a binary search over the block's prefix sums in LDS.
 * Applies a pending layout change first.  At most 2^31 shots per call. */
int qsim_sample(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out);

/* Device scratch held by the library for qsim_sample_noalloc, per GPU. */
#define QSIM_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  qsim_sample with the same arguments and the same
 * results, but it calls no hipMalloc or hipFree: its device buffers (block
 * sums, shot groups, residuals, indices) come from a fixed array of
 * QSIM_NOALLOC_SCRATCH_BYTES in the library, one per GPU, held until the
 * process ends.  Needs 8 * ceil(2^n / 2^13) + 16 * min(shots, blocks) +
 * 16 * shots bytes (each part rounded up to 256); larger requests are
 * refused before any work.  A pending layout change is applied first, which
 * needs the state's second buffer: a state without one is refused rather
 * than given one.  Every call on a GPU shares that GPU's array, which keeps
 * the last call's contents: the library runs the calls on one GPU one at a
 * time (each returns after its kernels end); overlapping them on one GPU by
 * any other route, e.g. another copy of the library in the same process,
 * would race. */
int qsim_sample_noalloc(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out);

/* No reference counterpart (the reference reads NVML): free and total memory
 * of the current device (hipMemGetInfo). */
int qsim_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

/* Per-kernel-class timing with HIP events recorded on the state's stream around
 * each launch (not inside captured graphs).  Classes: 0 chunk pass,
 * 1 permutation, 2 gate stream, 3 Pauli chunk pass. */
int qsim_profile(qsim_state *st, int enable);

/* Synchronises, then writes the summed milliseconds and launch counts per class
 * recorded since the last read; returns the number of classes written. */
int qsim_profile_read(qsim_state *st, double *ms_out, int64_t *count_out, int nclasses);

/* Counters: [0] kernel launches, [1] graph launches, [2] graph captures,
 * [3] chunk passes, [4] permutation passes run, [5] chunk passes run by
 * specialised kernels, [6] layout changes folded into a chunk pass.
 * Launches are counted when enqueued or captured: a graph replay adds to [1]
 * (and to [6] when it replays a folded layout change) only.
 * Returns the number written. */
int qsim_get_counters(const qsim_state *st, int64_t *out, int n);

/* Host only: plan `ops` into chunk passes without a device, after a pending
 * layout change in_map (stored bit p moves to in_map[p]; null for none) and
 * before one out_map (null for none), as the chunked modes plan them.
 * Writes, per pass, [K, cfast, nsteps, pos[0..K-1]] followed, per step, by
 * [register bits 0..3, rounds, dense] into `desc` (capacity desc_cap ints),
 * and the op order of execution into `order` (nops entries; each op index
 * once).  Returns the number of passes, or -1. */
int qsim_plan_describe(int num_qubits, int dtype, const qsim_op *ops, int nops, const void *mats,
                       uint64_t mat_count, int K, int cmin, const int32_t *in_map, const int32_t *out_map,
                       int32_t *desc, int desc_cap, int32_t *order);

/* ---- Batched circuits (no reference counterpart: the reference runs one
 * circuit per call).  Many independent circuits run in one persistent
 * launch: each workgroup takes circuits from a queue, largest first, and
 * for each one runs its forward op list from |0...0>, reduces <Z_mask> in
 * FP64, sets lam = Z_mask psi and runs its reverse list (the adjoint sweep,
 * as for qsim_adjoint), writing every probe.  States of at most 12 qubits
 * stay in LDS; larger ones stream through a per-workgroup HBM slot.
 * complex64 only. */
typedef struct qsim_batch qsim_batch;

int qsim_batch_create(qsim_batch **out);
int qsim_batch_destroy(qsim_batch *b);

/* Chunk bits for streamed circuits by qubit count (K_by_n[n], n < nk, each in
 * [6, 12]) and workgroups per CU (0: the occupancy maximum).  Applies to the
 * next qsim_batch_load. */
int qsim_batch_set_tuning(qsim_batch *b, const int32_t *K_by_n, int nk, int blocks_per_cu);

/* Workgroups per streamed circuit by qubit count (gang_by_n[n], n < nk, each
 * in [1, 64]; default 2^(n-20) above 20 qubits, else 1): a circuit run by a
 * gang of G workgroups has each rank take a contiguous 1/G of every pass's
 * chunks, with a device-wide barrier between passes; results are those of
 * one workgroup up to the order of the FP64 partial sums (per rank, then
 * added in rank order).  States held in LDS always use one.  The
 * launch caps G at a quarter of its workgroups.  Applies to the next
 * qsim_batch_load. */
int qsim_batch_set_gang(qsim_batch *b, const int32_t *gang_by_n, int nk);
/* Forward-sweep form by qubit count (fuse_by_n[n], n < nk, each 0 or 1;
 * default 1): 1 fuses every register step of the forward sweep into one
 * 16x16 unitary (the complex128 product of its gates, rounded once) applied
 * as a complex GEMM on the FP32 matrix cores; 0 runs each gate by its own
 * arithmetic.  The reverse sweep (gradient probes between its gates) always
 * runs gate by gate.  Applies to the next qsim_batch_load. */
int qsim_batch_set_fusion(qsim_batch *b, const int32_t *fuse_by_n, int nk);
/* Circuit i has num_qubits[i] (6..34) qubits, observable Z on zmask[i],
 * forward ops fwd_ops[fwd_off[i] .. fwd_off[i+1]), reverse ops (QSIM_OP_PROBE
 * allowed, q1 = probe slot < nprobes[i]) rev_ops[rev_off[i] .. rev_off[i+1]);
 * mat_offset indexes the shared complex64 arena `mats` (mat_count elements).
 * Plans every circuit and uploads one descriptor image. */
int qsim_batch_load(qsim_batch *b, int ncirc, const int32_t *num_qubits, const uint64_t *zmask,
                    const qsim_op *fwd_ops, const int64_t *fwd_off, const qsim_op *rev_ops,
                    const int64_t *rev_off, const void *mats, uint64_t mat_count,
                    const int32_t *nprobes);

/* Runs the loaded batch (use_graph: replay a hipGraph captured on first use)
 * and synchronises.  Outputs (each may be NULL): expect_out[ncirc],
 * probes_out[sum nprobes] (circuit order), exec_ms_out[ncirc] (each
 * circuit's time on its workgroup), *total_ms_out (the launch, HIP events). */
int qsim_batch_run(qsim_batch *b, int use_graph, double *expect_out, double *probes_out,
                   double *exec_ms_out, double *total_ms_out);

/* [0] circuits, [1] probes, [2] passes, [3] workgroups, [4] HBM slot bytes,
 * [5] queue tasks (one per circuit, or per rank of a gang), [6] gangs.
 * Returns the number written. */
int qsim_batch_info(const qsim_batch *b, int64_t *out, int n);

#ifdef __cplusplus
}
#endif

#endif /* QGPUSIM_H */
