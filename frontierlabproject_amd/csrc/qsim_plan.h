// Chunk-pass planner: splits an op list on bit positions into HBM passes.
//
// A pass gathers, for every setting of the bits outside a K-bit set P, the 2^K
// amplitudes that vary over P (a "chunk"), stages them in LDS, runs a sequence
// of register steps over them and writes them back in place.  P always holds
// the lowest `cmin` bit positions, so every chunk is read and written as runs
// of 2^cmin contiguous amplitudes.
//
// A step gathers 16 amplitudes per thread over 4 chunk-local bits (register
// bits 0..3) and runs a fixed program of slots on them: kRounds rounds of
// [1q on register bit 0, 1, 2, 3; pair ops on (0,1) (2,3) (0,2) (1,3) (0,3)
// (1,2)].  Each slot is enabled and typed by a flag, so every register index
// in the kernel is a compile-time constant.  Ops are placed into the earliest
// slot that follows every earlier op on a shared bit.
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
constexpr int kRounds = 3;     // slot rounds per step
constexpr int kSlotsPerRound = 10;
constexpr int kSlots = kRounds * kSlotsPerRound;
constexpr int k1qSlots = kRounds * 4;

enum OpKind : int32_t {
    OP_1Q = 1,      // 2x2 matrix on bit q0
    OP_2Q = 2,      // 4x4 matrix on (q0, q1), index b_q0 + 2 b_q1
    OP_PROBE = 3,   // accumulate Im<lam| G |psi> for the 2x2 matrix G on q0
};

// Slot flags.  1q slots:
enum : int32_t { S1_SKIP = 0, S1_GEN = 1, S1_DIAG = 2, S1_PROBE = 3, S1_PROBE_GEN = 4 };
// pair slots (matrix index b_first + 2 b_second, first < second register bit):
enum : int32_t { S2_SKIP = 0, S2_GEN = 1, S2_DIAG = 2, S2_CTRL_FIRST = 3, S2_CTRL_SECOND = 4 };

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
