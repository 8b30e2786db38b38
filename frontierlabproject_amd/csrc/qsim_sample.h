// Two-level FP64 sampling of measurement outcomes (qsim_sample.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsim_internal.h"

namespace qsim {

constexpr int kSampleB = 13;  // amplitudes per sampling block: 2^13 (64 KB of FP64 prefix sums in LDS)

// The shots [first, first + count) fall into amplitude block `block`.
struct SampleGroup {
    uint64_t block;
    uint32_t first, count;
};

// out[b] = sum of |psi_i|^2 over block b (FP64, compensated).
template <typename R>
hipError_t launch_block_prob(const cx<R> *psi, uint64_t dim, double *out, hipStream_t s);

// index_out[first + k] = the amplitude index of shot first + k of each group,
// given its residual probability inside the block.
template <typename R>
hipError_t launch_sample_blocks(const cx<R> *psi, uint64_t dim, const SampleGroup *groups, int ngroups,
                                const double *residual, uint64_t *index_out, hipStream_t s);

}  // namespace qsim
