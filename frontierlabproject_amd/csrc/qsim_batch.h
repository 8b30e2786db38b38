// Batched circuits: many independent small states in one persistent launch.
//
// Each workgroup takes tasks from a queue (largest circuits first) and runs,
// for each one, the forward chunk passes, <Z_mask> in FP64, lam = Z_mask psi
// and the adjoint reverse sweep with its gradient probes.  A circuit whose
// state fits the workgroup's LDS (K == n) never touches HBM after its
// descriptors; larger ones stream through an HBM slot in chunks of 2^K
// amplitudes.  A large circuit can be one task per rank of a gang of G
// workgroups: each rank runs a contiguous 1/G of every pass's chunks (and of
// the <Z> and lam sweep), in the slot of the rank-0 workgroup, with a
// device-wide barrier on the gang's counter between passes.  Partial sums go
// to per-rank slots that rank 0 adds in rank order, so results do not depend
// on timing.  The ranks of a gang sit next to each other in the queue, which
// hands tasks out in order to running workgroups, and a gang never has more
// ranks than the launch has workgroups: every rank a barrier waits for is
// taken by a workgroup that is running or will take it once its current task
// (never a wait on a later gang) ends, so every barrier is reached.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsim_internal.h"

namespace qsim {

#ifndef QSIM_BATCH_K
#define QSIM_BATCH_K 12
#endif
#ifndef QSIM_BATCH_WAVES
#define QSIM_BATCH_WAVES 2
#endif
constexpr int kBatchK = QSIM_BATCH_K;            // chunk bits held in LDS (two complex64 vectors)
constexpr int kBatchCmin = 4;                    // row length of streamed chunks: >= 2^4 amplitudes
constexpr int kBatchRows = 1 << (kBatchK - kBatchCmin);
constexpr int kBatchProbes = 64;                 // probe slots per pass (LDS partials)

constexpr int kBatchMaxGang = 64;                // workgroups per circuit at most

struct BatchCirc {
    int32_t n, K;            // qubits; chunk bits (K == n: the state lives in LDS)
    int32_t fwd0, nfwd;      // forward passes [fwd0, fwd0 + nfwd)
    int32_t rev0, nrev;      // reverse-sweep passes
    int32_t probe0, nprobe;  // this circuit's gradient slots in the probe output
    uint64_t zmask;          // observable: product of Z on these qubits
    int32_t gang;            // workgroups running the circuit (1: one, no barriers)
    int32_t gslot;           // gang index (barrier counter, owner), or -1
    int64_t part0;           // gang partials: [0, gang) <Z>, then probe q rank r at gang + q * gang + r
};

struct BatchTask {
    int32_t circ, rank;
};

struct BatchArgs {
    const BatchCirc *circs;
    const BatchTask *tasks;   // queue position -> (circuit, rank)
    int32_t ncirc, ntasks;
    int32_t ngangs, pad;
    const PassDev *passes;
    const StepDev *steps;
    const int32_t *probe_ids;
    const cx<float> *mats;
    cx<float> *scratch;       // per workgroup: psi and lam slots of slot_elems amplitudes
    uint64_t slot_elems;
    unsigned *counter;        // queue head (zeroed by the reset kernel)
    unsigned *gang_bar;       // [ngangs] barrier arrivals (zeroed by the reset kernel)
    int32_t *gang_owner;      // [ngangs] workgroup whose slot holds the gang's state
    double *gang_part;        // gang partial sums (see BatchCirc::part0)
    int64_t part_total;       // doubles in gang_part
    int64_t probe_total;      // doubles in probes
    double *expect;           // [ncirc]
    double *probes;           // [total probes]
    uint64_t *times;          // [2 * ncirc]: wall-clock ticks at first start (reset to
                              // all ones, lowered by each rank) and at the end
};

hipError_t launch_batch_reset(const BatchArgs &a, hipStream_t s);
hipError_t launch_batch(const BatchArgs &a, int blocks, hipStream_t s);
int batch_blocks_per_cu_max();

}  // namespace qsim
