// Internal declarations shared by the kernels and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "qsim_plan.h"
#include "qsim_device.h"

struct qsim_state;

namespace qsim {

template <typename R> struct Mat1 { cx<R> u[4]; };
template <typename R> struct Mat2 { cx<R> u[16]; };

constexpr int kMaxProbes = 256;        // probe slots per chunk pass
constexpr int kPermMaxS = 16;

// Permutation launch description (passed by value).
struct PermArgs {
    int32_t n, s, a, b;                 // tile bits, input-/output-contiguous low bits
    int32_t ncomp, pad[3];
    int32_t sin_hi[kPermMaxS];          // r bit a+k -> input position
    int32_t sout_hi[kPermMaxS];         // w bit b+k -> output position
    int32_t w2r[kPermMaxS];             // w bit k -> r bit
    int32_t comp_in[64];                // block-index bit m -> input position
    int32_t comp_out[64];               // ... -> output position
};

// Pauli term on a chunk: coef * i^{|x&z|} X^x Z^z  (Y where x and z both set)
struct PauliDev {
    uint64_t xmask, zmask;     // state bit masks
    int32_t xloc, pad;         // x mask in chunk-local bits
    double cre, cim;           // coefficient (complex)
};

enum PauliMode : int32_t {
    PAULI_WRITE = 1,   // lam = sum of terms  (overwrite)
    PAULI_ACC = 2,     // lam += sum of terms
    PAULI_EXPECT = 4,  // partial[block] = sum conj(psi) * (terms psi)
};

// Launchers (qsim_kernels.hip).  All take the stream last and return the
// launch status.
template <typename R>
hipError_t launch_init_basis(cx<R> *psi, uint64_t dim, uint64_t index, hipStream_t s);
template <typename R>
hipError_t launch_1q(cx<R> *psi, int n, int q, const Mat1<R> &U, hipStream_t s);
template <typename R>
hipError_t launch_2q(cx<R> *psi, int n, int q0, int q1, const Mat2<R> &U, hipStream_t s);
template <typename R>
hipError_t launch_diag1q(cx<R> *psi, int n, int q, cx<R> p0, cx<R> p1, hipStream_t s);
template <typename R>
hipError_t launch_permute(const cx<R> *in, cx<R> *out, const PermArgs &a, hipStream_t s);
template <typename R>
hipError_t launch_tile_pass(cx<R> *psi, cx<R> *lam, int n, const PassDev &host_pass,
                            const PassDev *d_pass, const StepDev *d_steps, const cx<R> *d_mats,
                            double *d_probe_partial, bool generic, hipStream_t s);
hipError_t launch_probe_reduce(const double *partial, int nblocks, int nprobe,
                               const int32_t *d_ids, double *out, hipStream_t s);
template <typename R>
hipError_t launch_pauli_chunk(const cx<R> *psi, cx<R> *lam, int n, const PassDev &host_pass,
                              const PassDev *d_pass, const PauliDev *d_terms, int nterms,
                              int mode, double *d_partial, hipStream_t s);
template <typename RA, typename RB>
hipError_t launch_dot(const cx<RA> *a, const cx<RB> *b, uint64_t dim, double *d_partial,
                      int *nblocks_out, hipStream_t s);
hipError_t launch_sum_partials(const double *partial, int nblocks, int width, double *out,
                               hipStream_t s);
template <typename R>
hipError_t launch_scale(cx<R> *psi, uint64_t dim, double factor, hipStream_t s);
// Dense gate on k = 3 or 4 qubits q[0..k-1] (matrix index sum_i b_{q_i} 2^i),
// matrix in device memory; n >= k.
template <typename R>
hipError_t launch_gate_k(cx<R> *psi, int n, const int *q, int k, const cx<R> *d_U, hipStream_t s);

int max_dot_blocks();
hipError_t prepare_kernels();

// Sets the thread's last-error string (qsim_last_error) and returns -1.
int set_error(const std::string &msg);

// qsim_sample's work; with `scratch` non-null every device buffer it needs
// comes from [scratch, scratch + scratch_bytes) instead of hipMalloc (refused
// beforehand when too small).
int sample_impl(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out, void *scratch,
                size_t scratch_bytes);
// The device a state was created on (-1 for null).
int state_device(const qsim_state *st);

}  // namespace qsim
