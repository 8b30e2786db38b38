"""Counts from sampled amplitude indices.

Indices come from the GPU sampler (``Statevector.sample``) in the state's
bit order, where bit position p holds qubit ``layout[p]``.  A measurement of
qubit q into classical bit c sets bit c of the outcome; counts are keyed by
bitstrings with classical bit 0 rightmost, as Qiskit prints them.
"""
import numpy as np


def measurements_of(qc):
    """[(qubit index, clbit index)] of the circuit's measurements; every
    qubit into its own bit when it has none."""
    out = []
    for ci in qc.data:
        if ci.operation.name == "measure":
            out.append((qc.find_bit(ci.qubits[0]).index, int(ci.clbits[0])))
    if not out:
        out = [(q, q) for q in range(qc.num_qubits)]
    return out


def counts_from_indices(indices, layout, measurements, num_clbits):
    """{bitstring: count} of sampled indices (see the module docstring)."""
    pos = {int(q): p for p, q in enumerate(layout)}
    idx = np.asarray(indices, dtype=np.uint64)
    out = np.zeros_like(idx)
    for q, c in measurements:
        out |= ((idx >> np.uint64(pos[q])) & np.uint64(1)) << np.uint64(c)
    keys, cnt = np.unique(out, return_counts=True)
    return {format(int(k), "0%db" % num_clbits): int(v) for k, v in zip(keys, cnt)}
