"""State and distribution agreement (qgpusim/metrics/correctness.py:1-53).

``statevector_overlap`` normalises both vectors and returns (|<a|b>|,
||a - b||_2) as the reference does; every sum here is taken in complex128
(the reference sums each chunk in the inputs' precision), so complex64 inputs
of 2^30+ amplitudes keep FP64 accuracy.  ``statevector_overlap_device`` gives
the same numbers for two device states without copying them to the host
(FP64 compensated sums on the GPU).
"""
from typing import Dict, List, Tuple, Union

import numpy as np


def _stable_vdot(a: np.ndarray, b: np.ndarray, chunk_size: int = 2 ** 24) -> complex:
    """sum conj(a) * b in chunks of chunk_size, accumulated in complex128
    (correctness.py:5-18)."""
    total = np.complex128(0)
    for i in range(0, len(a), chunk_size):
        total += np.vdot(np.asarray(a[i:i + chunk_size], dtype=np.complex128),
                         np.asarray(b[i:i + chunk_size], dtype=np.complex128))
    return complex(total)


def _norm(a: np.ndarray, chunk_size: int = 2 ** 24) -> float:
    return float(np.sqrt(max(_stable_vdot(a, a, chunk_size).real, 0.0)))


def statevector_overlap(cpu_vec: Union[np.ndarray, List], gpu_vec: Union[np.ndarray, List],
                        chunk_size: int = 2 ** 24) -> Tuple[float, float]:
    """(|<cpu|gpu>|, ||cpu - gpu||_2) of the two vectors normalised
    (correctness.py:22-37).  The L2 distance is phase-sensitive."""
    a = np.asarray(cpu_vec)
    b = np.asarray(gpu_vec)
    if a.shape != b.shape:
        raise ValueError("statevector_overlap: vectors differ in length")
    na, nb = _norm(a, chunk_size), _norm(b, chunk_size)
    ov = abs(_stable_vdot(a, b, chunk_size)) / (na * nb)
    # ||a/na - b/nb||^2 = 2 - 2 Re<a,b>/(na nb)
    d2 = 0.0
    for i in range(0, len(a), chunk_size):
        d = (np.asarray(a[i:i + chunk_size], dtype=np.complex128) / na
             - np.asarray(b[i:i + chunk_size], dtype=np.complex128) / nb)
        d2 += float(np.vdot(d, d).real)
    return float(ov), float(np.sqrt(d2))


def statevector_overlap_device(sv_a, sv_b) -> Tuple[float, float]:
    """statevector_overlap of two states resident on the GPU (same size and
    precision), from FP64 norms and overlap computed there."""
    na, nb = sv_a.norm_squared(), sv_b.norm_squared()
    ov = sv_a.overlap(sv_b)
    s = ov / np.sqrt(na * nb)
    return float(abs(s)), float(np.sqrt(max(2.0 - 2.0 * s.real, 0.0)))


def total_variation_distance(counts_a: Dict[str, int], counts_b: Dict[str, int]) -> float:
    """0.5 * sum_k |p_a(k) - p_b(k)| of two count dictionaries
    (correctness.py:39-53)."""
    keys = set(counts_a) | set(counts_b)
    shots_a = sum(counts_a.values()) or 1
    shots_b = sum(counts_b.values()) or 1
    return 0.5 * sum(abs(counts_a.get(k, 0) / shots_a - counts_b.get(k, 0) / shots_b) for k in keys)
