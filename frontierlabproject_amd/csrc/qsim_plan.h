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

namespace qsim {

constexpr int kMaxK = 14;      // chunk bits
constexpr int kG = 4;          // register-step bits
constexpr int kMaxComp = 50;   // bits outside a chunk
#ifndef QSIM_ROUNDS
#define QSIM_ROUNDS 4
#endif
constexpr int kRounds = QSIM_ROUNDS;  // maximum slot rounds per step
constexpr int kSlotsPerRound = 10;
constexpr int kSlots = kRounds * kSlotsPerRound;
constexpr int k1qSlots = kRounds * 4;

enum OpKind : int32_t {
    OP_1Q = 1,      // 2x2 matrix on bit q0
    OP_2Q = 2,      // 4x4 matrix on (q0, q1), index b_q0 + 2 b_q1
    OP_PROBE = 3,   // accumulate Im<lam| G |psi> for the 2x2 matrix G on q0
};

// Slot flags.  Each gate is classified exactly (entries compared with 0, 1, i)
// so the kernel runs the cheapest arithmetic for it.  1q slots (x: bit 0
// amplitude, y: bit 1 amplitude of each pair):
enum : int32_t {
    S1_SKIP = 0,
    S1_GEN = 1,        // [[u0,u1],[u2,u3]]                     4 cx
    S1_DIAG = 2,       // diag(d0, d1)                           2 cx
    S1_PROBE = 3,      // gradient probe with G                  8 cx (G, then identity)
    S1_PROBE_GEN = 4,  // probe with G, then apply U             8 cx (G, U)
    S1_PHASE = 5,      // diag(1, d1)                            1 cx
    S1_REAL = 6,       // real matrix                            2 cx (r0,r1),(r2,r3)
    S1_RI = 7,         // [[a, ib],[ic, d]], a b c d real        2 cx (a,b),(c,d)
    S1_X = 8,          // swap x, y
    S1_Y = 9,          // x' = -i y, y' = i x
    S1_BFLY = 10,      // x' = x + y, y' = x - y (H without its 1/sqrt2)
    S1_Z = 11,         // y' = -y
    S1_PI = 12,        // y' = i y   (S)
    S1_MI = 13,        // y' = -i y  (S^dag)
};
// pair slots (matrix index b_first + 2 b_second, first < second register bit);
// other pair ops make the step dense:
enum : int32_t {
    S2_SKIP = 0,
    S2_CPHASE = 5,       // diag(1, 1, 1, d3)                     1 cx
    S2_CX_FIRST = 6,     // flip second where first = 1
    S2_CX_SECOND = 7,    // flip first where second = 1
    S2_SWAP = 8,         // exchange index 1 and 2
};

struct PlanOp {
    int32_t kind;
    int32_t q0, q1;      // bit positions in the state
    int32_t mat_off;     // complex-element offset of the matrix in the caller's arena
    int32_t probe;       // OP_PROBE: index of the output slot, else -1
};

// Device-side records (plain int32 so one upload carries them).
struct PassDev {
    int32_t K, cfast, ncomp, nsteps;
    int32_t step_off, nprobe, probe_off, pad;
    int32_t pos[kMaxK];          // chunk bit k -> state bit position (ascending)
    int32_t comp[kMaxComp];      // block-index bit m -> state bit position
};

struct StepDev {
    int32_t bits[kG];            // register bit t -> chunk-local bit
    int32_t sbits[kG];           // the same bits, ascending
    int32_t nrounds;             // slot rounds in use
    int32_t dense;               // arena offset of a 16x16 matrix (dense step) or -1
    int32_t scale;               // arena offset of a factor applied after the step, or -1
    int32_t pad;
    int32_t kind[kSlots];        // slot flags
    int32_t mat[kSlots];         // complex-element offsets into the plan's arena
    int32_t probe[k1qSlots];     // probe slot within the pass, per 1q slot
};

struct Plan {
    std::vector<PassDev> passes;
    std::vector<StepDev> steps;
    std::vector<int32_t> probe_ids;  // per pass, slots in order (probe_off indexes here)
    std::vector<double> mats;        // re, im pairs (converted to the state precision on upload)
    std::vector<int32_t> order;      // op indices in execution order
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
bool plan_passes(int n, const PlanOp *ops, int nops, const void *mats, int elem_bytes, int K,
                 int cmin, int max_probes, Plan &out, std::string &err);

// Pauli chunk passes: each term's X support must lie inside its pass's chunk.
struct PauliPass {
    PassDev pd;
    std::vector<int> terms;        // term indices in this pass
    std::vector<int32_t> xloc;     // their X masks in chunk-local bits
};

bool plan_pauli(int n, const uint64_t *xmasks, int nterms, int K, int cmin,
                std::vector<PauliPass> &out, std::string &err);

}  // namespace qsim
