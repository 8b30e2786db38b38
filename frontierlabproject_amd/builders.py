"""The harness's test circuits (qgpusim/circuits/builders.py:10-57).

``build_circuit(circuit, n_qubits, depth, seed)`` with the reference's
circuit names and argument meaning:

* "ghz": H on qubit 0, then CX(0, i) for i = 1..n-1;
* "random": ``max(1, depth)`` layers; per layer RZ then RX on every qubit,
  with angles 2π·u from Python's ``random.Random(seed)`` drawn in that order
  (the same generator and draw order as the reference, so the same seed gives
  the same angles), then CX on the even pairs (i, i+1) and on the odd pairs;
* "qft": the QFT without final swaps (H, then controlled phases), the
  circuit the reference takes from Qiskit's ``QFT(n, do_swaps=False)``.
  Its unitary is fixed: the DFT with bit-reversed output.  The order in which
  Qiskit's decomposition lists the commuting controlled phases is not pinned
  here (Qiskit is absent), and it only matters to the tile schedule.
"""
import random

import numpy as np

from .circuit import QuantumCircuit
from .workloads import qft_circuit


def build_circuit(circuit: str, n_qubits: int, depth: int, seed: int) -> QuantumCircuit:
    qc = QuantumCircuit(n_qubits)
    if circuit == "ghz":
        qc.h(0)
        for i in range(1, n_qubits):
            qc.cx(0, i)
        return qc
    if circuit == "random":
        rng = random.Random(seed)
        for _ in range(max(1, depth)):
            for i in range(n_qubits):
                qc.rz(2 * np.pi * rng.random(), i)
                qc.rx(2 * np.pi * rng.random(), i)
            for i in range(0, n_qubits - 1, 2):
                qc.cx(i, i + 1)
            for i in range(1, n_qubits - 1, 2):
                qc.cx(i, i + 1)
        return qc
    if circuit == "qft":
        return qft_circuit(n_qubits, swaps=False)
    raise ValueError(f"Unknown circuit type: {circuit}")
