"""FP64 dense statevector simulator.  Test infrastructure.

Index convention: amplitude index ``i = sum_q b_q 2^q`` (qubit q is bit q), as in
qgpusim_cuda/src/kernels.cu:33-38.  A k-qubit matrix acts little-endian over its
qargs: row/column index ``sum_t b_{qargs[t]} 2^t`` (kernels.cu:54-55).
"""
import numpy as np


def zero_state(n):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def apply(psi, U, qargs, n):
    """Return U applied to ``psi`` on ``qargs`` (tensor contraction)."""
    k = len(qargs)
    U = np.asarray(U, dtype=np.complex128)
    assert U.shape == (1 << k, 1 << k)
    t = psi.reshape((2,) * n)
    # C-order reshape: axis a holds bit (n-1-a).  The matrix reshaped to
    # (2,)*2k has out-axes for qargs[k-1],...,qargs[0] then in-axes likewise.
    Ut = U.reshape((2,) * (2 * k))
    psi_axes = [n - 1 - qargs[k - 1 - j] for j in range(k)]
    out = np.tensordot(Ut, t, axes=(list(range(k, 2 * k)), psi_axes))
    # out axes: U's out axes (qargs[k-1]..qargs[0]) then the remaining psi axes
    rest = [a for a in range(n) if a not in psi_axes]
    src = list(range(k)) + list(range(k, n))
    dst = psi_axes + rest
    out = np.moveaxis(out, src, dst)
    return np.ascontiguousarray(out).reshape(-1)


def full_operator(U, qargs, n):
    """Brute-force 2^n x 2^n operator of U on ``qargs``, built index by index."""
    dim = 1 << n
    k = len(qargs)
    mask = 0
    for q in qargs:
        mask |= 1 << q
    M = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        c = sum(((col >> q) & 1) << t for t, q in enumerate(qargs))
        for r in range(1 << k):
            row = col & ~mask
            for t, q in enumerate(qargs):
                row |= ((r >> t) & 1) << q
            M[row, col] += U[r, c]
    return M


def simulate(gates, n, psi0=None):
    """Run a gate list ``[(matrix, qargs), ...]`` in program order."""
    psi = zero_state(n) if psi0 is None else np.array(psi0, dtype=np.complex128)
    for U, qargs in gates:
        psi = apply(psi, U, list(qargs), n)
    return psi


def fidelity(a, b):
    """|<a|b>|^2 / (<a|a><b|b>) accumulated in FP64 (math.fsum-free, chunked)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    ov = np.vdot(a, b)
    na = np.vdot(a, a).real
    nb = np.vdot(b, b).real
    return float(abs(ov) ** 2 / (na * nb))


def max_abs_err_phase_aligned(ref, cand):
    """Max |ref - e^{i phi} cand| with phi aligning the global phase."""
    ref = np.asarray(ref, dtype=np.complex128)
    cand = np.asarray(cand, dtype=np.complex128)
    ov = np.vdot(cand, ref)
    ph = ov / abs(ov) if abs(ov) > 0 else 1.0
    return float(np.max(np.abs(ref - ph * cand)))
