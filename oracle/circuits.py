"""Seeded generators for the benchmark workloads (oracle side).  Test infrastructure.

A circuit is a list of ``Op(name, qubits, params)``; ``unitary`` is False only for
the terminal all-qubit statevector marker (qgpusim/runner.py:38-40).  Seeds:
``numpy.random.Generator(PCG64(1000 + config_index))`` unless given.
"""
import numpy as np

from . import gates as G


class Op:
    __slots__ = ("name", "qubits", "params", "unitary")

    def __init__(self, name, qubits, params=(), unitary=True):
        self.name = name
        self.qubits = tuple(int(q) for q in qubits)
        self.params = tuple(float(p) for p in params)
        self.unitary = unitary

    def matrix(self):
        return G.matrix(self.name, self.params)

    def __repr__(self):
        return f"Op({self.name},{self.qubits},{self.params})"


RANDOM_GATESET = ("h", "x", "y", "z", "s", "t", "rx", "ry", "rz")


def random_circuit(n, depth, seed):
    """Per layer: one gate from RANDOM_GATESET on every qubit (angle U[0,2pi)
    for rotations), then CX on the pairs (perm[2i], perm[2i+1]) of a seeded
    permutation, each with probability 0.5 (control = first)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = []
    for _ in range(depth):
        for q in range(n):
            name = RANDOM_GATESET[int(rng.integers(0, len(RANDOM_GATESET)))]
            if name in ("rx", "ry", "rz"):
                ops.append(Op(name, (q,), (float(rng.uniform(0.0, 2 * np.pi)),)))
            else:
                ops.append(Op(name, (q,)))
        perm = rng.permutation(n)
        for i in range(n // 2):
            if rng.random() < 0.5:
                ops.append(Op("cx", (perm[2 * i], perm[2 * i + 1])))
    return ops


def reference_random_circuit(n, depth):
    """Gate-structure twin of the reference's transpiled "random" circuit
    (qgpusim/circuits/builders.py:23-35 after transpile merges rz.rx into one
    1q gate): per layer one 1q gate per qubit, CX on even pairs, then odd pairs.
    Angles do not affect the schedule; a fixed 'h' stands in for the merged gate."""
    ops = []
    for _ in range(depth):
        for q in range(n):
            ops.append(Op("h", (q,)))
        for i in range(0, n - 1, 2):
            ops.append(Op("cx", (i, i + 1)))
        for i in range(1, n - 1, 2):
            ops.append(Op("cx", (i, i + 1)))
    return ops


def qft_circuit(n, swaps=True):
    """QFT: for j = n-1..0: H(j), then CP(pi/2^(j-k)) on (k, j) for k = j-1..0;
    then SWAP(i, n-1-i).  With swaps it equals sqrt(N) * ifft (Qiskit's sign)."""
    ops = []
    for j in reversed(range(n)):
        ops.append(Op("h", (j,)))
        for k in reversed(range(j)):
            ops.append(Op("cp", (k, j), (np.pi / float(2 ** (j - k)),)))
    if swaps:
        for i in range(n // 2):
            ops.append(Op("swap", (i, n - 1 - i)))
    return ops


def ansatz_circuit(n, layers, theta):
    """Hardware-efficient ansatz: per layer RY(theta), RZ(theta) on every qubit
    (parameters interleaved per qubit), then CZ on the ring (i, i+1 mod n)."""
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.shape == (layers * 2 * n,)
    ops = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            ops.append(Op("ry", (q,), (theta[k],)))
            ops.append(Op("rz", (q,), (theta[k + 1],)))
            k += 2
        for q in range(n):
            ops.append(Op("cz", (q, (q + 1) % n)))
    return ops


def ansatz_params(n, layers, seed=1003):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-np.pi, np.pi, size=layers * 2 * n)


def random_state(n, seed):
    """i.i.d. complex Gaussian amplitudes, normalised in FP64: the interleaved
    (re, im) vector of 2^(n+1) standard normals, its k-th of 16 equal slices
    drawn by the k-th PCG64 stream spawned from SeedSequence(seed)."""
    flat = np.empty(2 << n, dtype=np.float64)
    streams = min(16, flat.size)
    step = flat.size // streams
    for k, ss in enumerate(np.random.SeedSequence(seed).spawn(streams)):
        flat[k * step:(k + 1) * step] = np.random.Generator(np.random.PCG64(ss)).standard_normal(step)
    psi = flat.view(np.complex128)
    return psi / np.sqrt(np.vdot(psi, psi).real)


def marker(n):
    """Terminal all-qubit statevector marker (no matrix)."""
    return Op("save_statevector", tuple(range(n)), unitary=False)


def to_dense_gates(ops):
    return [(op.matrix(), op.qubits) for op in ops if op.unitary]
