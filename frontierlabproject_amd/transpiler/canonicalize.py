"""Single-qubit run merging: the part of the harness's ``transpile()`` step
(qgpusim/runner.py:45) that the reference's published rows pin.

The custom path transpiles the circuit for the Aer target before the LQR
pass.  On the reference's random circuits that step turns each qubit's
rz·rx pair of a layer into one single-qubit gate: 944 gates with 320 rz and
320 rx become 624 gates with none (results/result.csv:2 against :6, 20
qubits, depth 16; :9 against :13 at 25 qubits), and the depth goes from 64
to 48.  ``merge_1q_runs`` does exactly that merge and nothing else: every
maximal run of single-qubit unitary gates on one qubit becomes one
``unitary`` gate holding the run's matrix product (exact, so no global phase
moves), and all other instructions pass through in order.  The rest of
Qiskit's optimisation (basis translation, commutation, two-qubit resynthesis)
is not reproduced; Qiskit is absent, and nothing the reference ships pins it.
"""
import numpy as np

from ..circuit import UnitaryGate


def _is_1q_unitary(ci):
    op = ci.operation
    return (len(ci.qubits) == 1 and not ci.clbits and hasattr(op, "to_matrix")
            and not getattr(op, "_directive", False))


def merge_1q_runs(qc):
    """A copy of ``qc`` with each maximal run of single-qubit gates on a qubit
    replaced by one gate.  A run of one gate is kept as it is.  A merged gate
    is emitted where its run ends: just before the next instruction on its
    qubit, or at the end of the circuit (in qubit order)."""
    out = qc.copy()
    out.data = []
    pending = {}

    def flush(q):
        run = pending.pop(q, None)
        if not run:
            return
        if len(run) == 1:
            out.data.append(run[0])
            return
        m = np.eye(2, dtype=np.complex128)
        for ci in run:
            m = np.asarray(ci.operation.to_matrix(), dtype=np.complex128) @ m
        out.append(UnitaryGate(m), [q])

    for ci in qc.data:
        if _is_1q_unitary(ci):
            pending.setdefault(ci.qubits[0], []).append(ci)
            continue
        for q in ci.qubits:
            flush(q)
        out.data.append(ci)
    for q in qc.qubits:
        flush(q)
    return out
