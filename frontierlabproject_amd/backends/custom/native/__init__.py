"""Loader for the HIP backend (replaces qgpusim/backends/custom/native/__init__.py).

Re-raises if the shared library cannot be loaded (the reference does the same,
native/__init__.py:4-13): there is no CPU or CuPy fallback.
"""
from . import qgpusim_hip
from .qgpusim_hip import (Statevector, StatevectorF32, apply_1q_gate, apply_2q_gate)

# the reference reaches the compiled module as `native.qgpusim_cuda`
qgpusim_cuda = qgpusim_hip

__all__ = ["Statevector", "StatevectorF32", "apply_1q_gate", "apply_2q_gate",
           "qgpusim_hip", "qgpusim_cuda"]
