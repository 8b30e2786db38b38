// Types and constants shared by the host planner, the compiled kernels and the
// chunk-pass kernels compiled at run time.  Uses no standard header when built
// by hiprtc (__HIPCC_RTC__), which supplies the fixed-width integer types.
#pragma once
#ifdef __HIPCC_RTC__
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int32_t int32_t;
#else
#include <cstdint>
#endif

namespace qsim {

template <typename R> struct cx { R re, im; };

constexpr int kTPB = 256;              // threads per block for every kernel
constexpr int kWaves = kTPB / 64;

constexpr int kMaxK = 14;      // chunk bits
constexpr int kMaxKFold = 13;  // chunk bits of a pass that folds a layout change (complex64 LDS)
constexpr int kG = 4;          // register-step bits

// LDS slot of chunk element e: a fixed linear map over GF(2) that XORs a
// pattern into bits 1..6 for each set bit 4..12 of e.  Bit 0 is untouched, so
// an amplitude pair (e, e^1) stays one aligned 16-byte slot pair; a bit only
// feeds lower bits, so slots stay below 2^K for e < 2^K and the map is its
// own kind of triangular bijection.  Linearity lets callers split e into
// disjoint parts and XOR their images.  The patterns were picked by timing a
// register step (16 reads and writes of 8 bytes per lane) for every 4-bit
// register set on MI355X: the step averages 1660 cycles per 512-thread block,
// against 2270 for XORing bits 6..10 into 1..5 and 1590 for the best map per set.
#ifdef __HIPCC_RTC__
#define QSIM_HD __host__ __device__
#elif defined(__HIPCC__)
#define QSIM_HD __host__ __device__
#else
#define QSIM_HD
#endif
QSIM_HD inline constexpr int swz_pattern(int k)
{
    return k == 4 ? 0x0a : k == 5 ? 0x16 : k == 6 ? 0x3c : k == 7 ? 0x1e : k == 8 ? 0x1c : k == 9 ? 0x7c
         : k == 10 ? 0x76 : k == 11 ? 0x7a : k == 12 ? 0x7a : 0;
}
QSIM_HD inline constexpr int swz(int e)
{
    int s = e;
    for (int k = 4; k <= 12; ++k)
        if ((e >> k) & 1) s ^= swz_pattern(k);
    return s;
}
// Slot of the single element 2^k.
QSIM_HD inline constexpr int swz_bit(int k) { return (1 << k) ^ swz_pattern(k); }
// complex128 chunks: one amplitude per 16-byte slot, so a bank line of 128
// (256) bytes holds 8 (16) slots and a wave's 16-byte accesses spread over
// the banks when its lanes' slots differ in bits 0..2 (0..3).  A register
// step's lanes differ in the lowest non-register chunk bits; these patterns
// (bits 3..13 into bits 0..3, chosen by a search over every 4-bit register
// set of a 12-bit chunk) keep those bits' images independent for most sets:
// the average lane-group multiplicity drops from 1.75 (1.46) to 1.19 (1.19).
// Measured on the 33-qubit complex128 circuit: LDS bank-conflict cycles from
// 28.9 % to 24.8 % of the chunk passes' time, passes 55.2 -> 54.3 ms.
QSIM_HD inline constexpr int swz_pattern_d(int k)
{
    return k == 3 ? 0x3 : k == 4 ? 0x7 : k == 5 ? 0xe : k == 6 ? 0xd : k == 7 ? 0x8 : k == 8 ? 0xd : k == 9 ? 0x3
         : k == 10 ? 0x5 : k == 11 ? 0x9 : k == 12 ? 0x3 : k == 13 ? 0xa : 0;
}
QSIM_HD inline constexpr int swz_d(int e)
{
    int s = e;
    for (int k = 3; k <= 13; ++k)
        if ((e >> k) & 1) s ^= swz_pattern_d(k);
    return s;
}
// The swizzle of a chunk of cx<R>: every kernel templated on the precision
// addresses its LDS through this one.
template <typename R>
QSIM_HD inline constexpr int swzr(int e)
{
    return sizeof(R) == 8 ? swz_d(e) : swz(e);
}
template <typename R>
QSIM_HD inline constexpr int swzr_bit(int k)
{
    return swzr<R>(1 << k);
}
constexpr int kMaxComp = 50;   // bits outside a chunk
#ifndef QSIM_ROUNDS
#define QSIM_ROUNDS 4
#endif
constexpr int kRounds = QSIM_ROUNDS;  // maximum slot rounds per step
constexpr int kSlotsPerRound = 10;
constexpr int kSlots = kRounds * kSlotsPerRound;
constexpr int k1qSlots = kRounds * 4;

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
    // a real factor u0 taken out (plans without probes, |u0| >= the off-diagonal):
    S1_REAL1 = 14,     // [[1, a],[b, 1]], a b real               1 cx (a,b)
    S1_RI1 = 15,       // [[1, iq],[ir, 1]], q r real             1 cx (q,r)
    S1_GEN1 = 16,      // [[1, b],[c, d]]  (u0 complex, taken out) 3 cx
    S1_NOP = 17,       // identity (a diagonal gate reduced to its factor)
    S1_REALG = 18,     // [[1, a],[b, d]], a b d real (u0 taken out) 2 cx (a,b),(d,0)
    // flag: exchange the pair (x, y) first, then apply the kind (U = V X with
    // V = U X; a register rename in specialised kernels)
    S1_SWAP = 64,
};
// pair slots (matrix index b_first + 2 b_second, first < second register bit);
// other pair ops make the step dense:
enum : int32_t {
    S2_SKIP = 0,
    S2_CPHASE = 5,       // diag(1, 1, 1, d3)                     1 cx
    S2_CX_FIRST = 6,     // flip second where first = 1
    S2_CX_SECOND = 7,    // flip first where second = 1
    S2_SWAP = 8,         // exchange index 1 and 2
    S2_CU_FIRST = 9,     // 2x2 U on the second bit where the first is 1    4 cx
    S2_CU_SECOND = 10,   // 2x2 U on the first bit where the second is 1    4 cx
};
// Generic slot forms (plans for the compiled slot-program kernels): 1q slots
// are S1_GEN, S1_DIAG or probes, pair slots S2_CPHASE, S2_CU_FIRST or
// S2_CU_SECOND.  Every form is arithmetic on the registers it changes (no
// register permutations), so the kernels compile without the register-file
// copies that permutation forms cost where the kinds are only known at run time.

// Device-side records (plain int32 so one upload carries them).
struct PassDev {
    int32_t K, cfast, ncomp, nsteps;
    int32_t step_off, nprobe, probe_off, pad;
    int32_t pos[kMaxK];          // chunk bit k -> state bit position (ascending)
    int32_t comp[kMaxComp];      // block-index bit m -> state bit position
    int32_t mat_lo, mat_hi;      // the pass's matrix entries: [mat_lo, mat_hi) of the arena
};

struct StepDev {
    int32_t bits[kG];            // register bit t -> chunk-local bit
    int32_t sbits[kG];           // the same bits, ascending
    int32_t nrounds;             // slot rounds in use
    int32_t dense;               // arena offset of a 16x16 matrix (dense step) or -1
    int32_t scale;               // arena offset of a factor applied after the step, or -1
    int32_t bit4;                // fifth register bit of a 32x32 dense step (register index 4), or -1
    int32_t kind[kSlots];        // slot flags
    int32_t mat[kSlots];         // complex-element offsets into the plan's arena
    int32_t probe[k1qSlots];     // probe slot within the pass, per 1q slot
};

}  // namespace qsim
