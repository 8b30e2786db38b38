"""Layout permutation restatement.  Test infrastructure.

``layout[pos]`` is the global qubit held at bit position ``pos``
(qgpusim/transpiler/passes.py:21-23).  Changing layouts moves every amplitude
``i`` to ``j = sum_p bit_p(i) << map[p]`` with
``map[p_old] = new_pos[layout_old[p_old]]`` (qgpusim_cuda/src/bindings.cpp:447-464,
qgpusim_cuda/src/kernels.cu:176-184): ``out[j] = in[i]`` (a scatter).
"""
import numpy as np


class LayoutError(RuntimeError):
    pass


def perm_map(layout_old, layout_new):
    """Bit-position map with the reference's validation order and messages
    (bindings.cpp:435-463)."""
    n = len(layout_old)
    if len(layout_new) != n:
        raise LayoutError("layout_old/layout_new must have length num_qubits")
    new_pos = [-1] * n
    for p, gq in enumerate(layout_new):
        if gq < 0 or gq >= n:
            raise LayoutError("layout_new contains invalid qubit id")
        if new_pos[gq] != -1:
            raise LayoutError("layout_new is not a permutation")
        new_pos[gq] = p
    out = [-1] * n
    for p_old, gq in enumerate(layout_old):
        if gq < 0 or gq >= n:
            raise LayoutError("layout_old contains invalid qubit id")
        p_new = new_pos[gq]
        if p_new < 0:
            raise LayoutError("layout_old contains qubit not in layout_new")
        out[p_old] = p_new
    return out


def scatter(psi, bitmap):
    """out[j] = in[i], j = sum_p bit_p(i) << bitmap[p]  (numpy, vectorised)."""
    n = len(bitmap)
    i = np.arange(1 << n, dtype=np.uint64)
    j = np.zeros_like(i)
    for p in range(n):
        j |= ((i >> np.uint64(p)) & np.uint64(1)) << np.uint64(bitmap[p])
    out = np.empty_like(psi)
    out[j] = psi
    return out


def permute(psi, layout_old, layout_new):
    return scatter(psi, perm_map(layout_old, layout_new))
