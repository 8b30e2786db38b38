// Chunk-pass planner: splits an op list on bit positions into HBM passes.
//
// A pass gathers, for every setting of the bits outside a K-bit set P, the 2^K
// amplitudes that vary over P (a "chunk"), stages them in LDS, runs a sequence
// of register steps over them and writes them back in place.  P always holds
// the lowest `cmin` bit positions, so every chunk is read and written as runs
// of 2^cmin contiguous amplitudes.
//
// A step gathers 16 amplitudes per thread over 4 chunk-local bits (register
// bits 0..3) and either
//  * runs up to kRounds rounds of slots [1q on register bit 0, 1, 2, 3; pair
//    ops on (0,1) (2,3) (0,2) (1,3) (0,3) (1,2)], each slot enabled and typed by
//    a flag, so every register index in the kernel is a compile-time constant
//    (ops go to the earliest slot after every earlier op on a shared bit); or
//  * applies one dense 16x16 matrix, the complex128 product of its ops, for
//    pair ops without a cheap slot form (general 4x4, diagonal, controlled-U).
//
// Ops may move earlier than ops they share no bit with (they commute): once an
// op is left behind, its bits are blocked for the rest of the scan, the same
// look-ahead rule as the tile construction of
// qgpusim/transpiler/passes.py:89-128.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "qsim_dev_types.h"

namespace qsim {

enum OpKind : int32_t {
    OP_1Q = 1,      // 2x2 matrix on bit q0
    OP_2Q = 2,      // 4x4 matrix on (q0, q1), index b_q0 + 2 b_q1
    OP_PROBE = 3,   // accumulate Im<lam| G |psi> for the 2x2 matrix G on q0
    OP_3Q = 4,      // 8x8 matrix on (q0, q1, q2), index b_q0 + 2 b_q1 + 4 b_q2 (a dense step)
    OP_4Q = 5,      // 16x16 matrix on (q0, q1, q2, q3) (a dense step)
};

struct PlanOp {
    int32_t kind;
    int32_t q0, q1;      // bit positions in the state
    int32_t mat_off;     // complex-element offset of the matrix in the caller's arena
    int32_t probe;       // OP_PROBE: index of the output slot, else -1
    int32_t q2 = -1, q3 = -1;  // third and fourth qubits of OP_3Q / OP_4Q
};

// Qubits an op acts on (1 to 4), in its matrix's index order.
inline int op_qubits(const PlanOp &op, int *q)
{
    q[0] = op.q0;
    if (op.kind == OP_1Q || op.kind == OP_PROBE) return 1;
    q[1] = op.q1;
    if (op.kind == OP_2Q) return 2;
    q[2] = op.q2;
    if (op.kind == OP_3Q) return 3;
    q[3] = op.q3;
    return 4;
}

inline uint64_t op_mask(const PlanOp &op)
{
    int q[4];
    const int k = op_qubits(op, q);
    uint64_t m = 0;
    for (int i = 0; i < k; ++i) m |= 1ull << q[i];
    return m;
}

// A diagonal gate applied after its register step's slot rounds, with at
// least one qubit on a chunk bit outside the step's register bits: each
// thread multiplies its amplitudes by phase factors chosen by its group's
// bits (specialised kernels only).  Its bits outside the registers stay
// outside them for the rest of the step, and no later non-diagonal gate of
// the step touches its bits, so applying it last keeps the gate order.
// A qubit outside the pass's chunk (diagonal gates need no chunk bits) is
// fixed per block: its value is bit m of the block's chunk index (the m-th
// PassDev::comp position), encoded as -2 - m.
struct FreeDiag {
    int32_t a, b;  // the op's qubits in matrix-index order: chunk-local bit, or -2 - m (b = -1: 1q)
    int32_t mo;    // arena offset of its diagonal (2 or 4 complex entries)
    int32_t ro;    // complex64 unit-modulus diagonals: offset of its phases in Plan::raw, else -1
};

struct Plan {
    std::vector<PassDev> passes;
    std::vector<StepDev> steps;
    std::vector<int32_t> probe_ids;  // per pass, slots in order (probe_off indexes here)
    std::vector<double> mats;        // re, im pairs (converted to the state precision on upload)
    std::vector<int32_t> order;      // op indices in execution order
    // When set (a layout change folded into the first pass): in_map[p] is the
    // position in the pass layout of input bit position p.  The first pass
    // then reads the input in the old layout and writes its output, in the
    // pass layout, to another buffer.
    std::vector<int32_t> in_map;
    // When set (a layout change folded into the last pass): out_map[p] is the
    // position in the output of pass-layout bit position p.  The last pass
    // then writes its output in that order to another buffer; it holds the
    // preimages of the output's lowest cmin positions, so its writes run in
    // rows.  Left empty by plan_passes when the last pass has no room for them.
    std::vector<int32_t> out_map;
    // When set (the input is |0...0>, not yet written): the first pass
    // generates its chunks instead of reading them.
    bool zero_input = false;
    // Only the generic slot forms (see S2_CU_FIRST): the plan runs on the
    // compiled slot-program kernels.
    bool generic = false;
    // Free diagonals of step s: fdiag[fd_off[s] .. fd_off[s + 1]); both
    // empty when the plan has none.
    std::vector<FreeDiag> fdiag;
    std::vector<int32_t> fd_off;
    // 32-bit words uploaded after the matrices (at byte offset raw_offset()
    // from the arena): the phases of complex64 free diagonals in turns, as
    // fractions of 2^32, so that kernels add them exactly modulo a turn.
    std::vector<uint32_t> raw;
    size_t raw_offset(size_t amp_bytes) const { return (mats.size() / 2 * amp_bytes + 255) & ~size_t(255); }
    int step_fdiag(int s, const FreeDiag **fd) const
    {
        if (fd_off.empty()) return 0;
        *fd = fdiag.data() + fd_off[s];
        return fd_off[s + 1] - fd_off[s];
    }
};

// Scalar factors (a diagonal gate's first entry, H's 1/sqrt2) are taken out of
// the gates and applied once, after a later step; no amplitude leaves
// [2^-24, 2^24] times its exact value in between.  Plans with probes keep the
// factors unit-modulus (probes are invariant only under phases).
//
// Plan `ops` (in order) for an n-bit state.  `mats` holds the ops' matrices
// as complex elements of `elem_bytes` (8: complex64, 16: complex128).  K is
// clamped to n; cmin to K-2.  max_probes bounds the probe slots of one pass.
// Returns false and sets err on invalid input.
// fuse (plans without probes only): 0, a slot step becomes one dense 16x16
// matrix (the product of all its ops) when its slot arithmetic exceeds the
// product's 64 multiply-adds per amplitude on the vector units; 1, every
// register step becomes dense; 2, as 0 with the bar at 32 multiply-adds, the
// cost of the product on the FP32 matrix cores (48 real multiply-adds per
// amplitude at twice the vector rate); 3 (complex64, not generic), every
// register step becomes one dense 32x32 matrix over five register bits
// (StepDev::bit4 holds the fifth), else as 1; 4, as 2 with 32x32 steps
// where they take enough work (below; complex64, not generic).
// in_map (n entries, or null): the input is in another layout (see
// Plan::in_map); the first pass also holds the images of the lowest cmin
// input bit positions, so it reads runs of 2^cmin contiguous amplitudes.
// out_map (n entries, or null): the output goes to another layout (see
// Plan::out_map), folded into the last pass when it has room.
// generic: only the generic slot forms, no scalars taken out, and no step
// made dense for its cost (plans for the compiled slot-program kernels, where
// the kinds are only known at run time and code size is what limits speed).
// Plans that are not generic, in fusion modes 0, 2 and 4, run diagonal 1q/2q
// gates as free diagonals (FreeDiag): such a gate needs
// no chunk bits, so the chunk search and the in-order rule place only the
// non-diagonal gates, and within a pass it costs no register step.
// bar (fuse 2 and 4; 0: kFuseBar): the slot cost above which a 4-bit step
// becomes dense.  bar5 (fuse 4; 0: kFuseBar5): a step is first grown as one
// 32x32 matrix over five register bits, and kept when the vector work of the
// ops it takes (four real multiply-adds per nonzero of an average matrix
// row, summed) exceeds bar5; else it is built as for fuse 2.
constexpr double kFuseBar = 32.0, kFuseBar5 = 64.0;
bool plan_passes(int n, const PlanOp *ops, int nops, const void *mats, int elem_bytes, int K,
                 int cmin, int max_probes, Plan &out, std::string &err, int fuse = 0,
                 const int32_t *in_map = nullptr, const int32_t *out_map = nullptr, bool generic = false,
                 double bar = 0.0, double bar5 = 0.0);

// Pauli chunk passes: each term's X support must lie inside its pass's chunk.
struct PauliPass {
    PassDev pd;
    std::vector<int> terms;        // term indices in this pass
    std::vector<int32_t> xloc;     // their X masks in chunk-local bits
};

bool plan_pauli(int n, const uint64_t *xmasks, int nterms, int K, int cmin,
                std::vector<PauliPass> &out, std::string &err);

}  // namespace qsim
