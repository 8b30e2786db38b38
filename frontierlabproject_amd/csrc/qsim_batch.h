// Batched circuits: many independent small states in one persistent launch.
//
// Each workgroup takes circuits from a queue (largest first) and runs, for
// each one, the forward chunk passes, <Z_mask> in FP64, lam = Z_mask psi and
// the adjoint reverse sweep with its gradient probes.  A circuit whose state
// fits the workgroup's LDS (K == n) never touches HBM after its descriptors;
// larger ones stream through a private HBM slot in chunks of 2^K amplitudes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsim_internal.h"

namespace qsim {

constexpr int kBatchK = 12;                      // chunk bits held in LDS (two complex64 vectors)
constexpr int kBatchCmin = 4;                    // row length of streamed chunks: >= 2^4 amplitudes
constexpr int kBatchRows = 1 << (kBatchK - kBatchCmin);
constexpr int kBatchProbes = 64;                 // probe slots per pass (LDS partials)

struct BatchCirc {
    int32_t n, K;            // qubits; chunk bits (K == n: the state lives in LDS)
    int32_t fwd0, nfwd;      // forward passes [fwd0, fwd0 + nfwd)
    int32_t rev0, nrev;      // reverse-sweep passes
    int32_t probe0, nprobe;  // this circuit's gradient slots in the probe output
    uint64_t zmask;          // observable: product of Z on these qubits
};

struct BatchArgs {
    const BatchCirc *circs;
    const int32_t *order;     // queue position -> circuit
    int32_t ncirc, pad;
    const PassDev *passes;
    const StepDev *steps;
    const int32_t *probe_ids;
    const cx<float> *mats;
    cx<float> *scratch;       // per workgroup: psi and lam slots of slot_elems amplitudes
    uint64_t slot_elems;
    unsigned *counter;        // queue head (zeroed by the reset kernel)
    double *expect;           // [ncirc]
    double *probes;           // [total probes]
    uint64_t *times;          // [2 * ncirc]: wall-clock ticks at start and end
};

hipError_t launch_batch_reset(unsigned *counter, hipStream_t s);
hipError_t launch_batch(const BatchArgs &a, int blocks, hipStream_t s);
int batch_blocks_per_cu_max();

}  // namespace qsim
