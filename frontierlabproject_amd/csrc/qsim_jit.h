// Chunk-pass kernels specialised per pass and compiled at run time (hiprtc).
//
// The compiled kernel k_tile_pass reads each step's slot program from device
// memory and branches on every slot's gate kind, so each slot ends in a join
// where the 16 amplitudes are copied back into fixed registers; it runs plans
// in the generic slot forms, which keep those copies to one per changed
// amplitude.  A specialised kernel has the pass's register bits, LDS offsets
// and gate kinds as constants: the slots become straight-line code, register
// exchanges (X, CX, SWAP) become renames, and the plan uses the cheapest form
// of each gate.  Both apply the same gates in the same order and include
// qsim_device.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "qsim_plan.h"

namespace qsim {

// One function per pass of `plan` for `nv` vectors (1: psi, forward;
// 2: psi and lam with gradient probes, the adjoint's reverse sweep).  nullptr
// marks a pass left to k_tile_pass (probes in a one-vector call).  Kernels
// are cached per device and source for the life of the process; missing ones
// are compiled in parallel.  `arch` is the device's gcnArchName.  Returns
// false with `err`.
//   nv 1: kernel(psi, mats)
//   nv 2: kernel(psi, lam, probe_partial, mats), probe_partial[block * nprobe + k]
bool jit_functions(const Plan &plan, bool dbl, int nv, int device, const std::string &arch,
                   std::vector<hipFunction_t> &out, std::string &err);

// Source of the specialised kernel for pass `p` (kernel name `name`).
std::string jit_pass_source(const Plan &plan, size_t p, bool dbl, int nv, const std::string &name);

// Compile every pass of `plan` for `arch` without loading it (no device
// needed).  Returns false with the compiler log in `err`.
bool jit_compile_only(const Plan &plan, bool dbl, int nv, const std::string &arch, std::string &err);

// Kernels compiled so far in this process (all devices).
long jit_compiled_count();

// Threads per block of the specialised kernel of a pass with K chunk bits.
int jit_block_threads(int K, int nv);

// Whether the first pass of `plan` folds its layout change (Plan::in_map):
// its kernel is then kernel(psi_in, psi_out, mats), reading the source layout
// and writing the pass layout to the other buffer (complex64, one vector).
bool jit_mapped_first(const Plan &plan, bool dbl, int nv);

// Whether the last pass of `plan` writes the output order (Plan::out_map):
// its kernel is then kernel(psi_in, psi_out, mats), writing the other buffer
// in that order (reading the source layout too when it is also the first
// pass of a plan with Plan::in_map; one vector).
bool jit_mapped_last(const Plan &plan, bool dbl, int nv);

}  // namespace qsim
