"""Circuit generators for the benchmark workloads.

* ``random_circuit(n, depth, seed)``: per layer one gate drawn uniformly from
  {H, X, Y, Z, S, T, RX, RY, RZ} on every qubit (angles U[0, 2pi)), then CX
  (control first) on the pairs (perm[2i], perm[2i+1]) of a seeded permutation,
  each with probability 0.5.  RNG: ``numpy.random.Generator(PCG64(seed))``.
* ``qft_circuit(n)``: H and CP(pi/2^k) ladder, then the n/2 final SWAPs; the
  CP angles are computed in FP64.
* ``ansatz_circuit(n, layers, theta)``: RY(theta), RZ(theta) on every qubit
  (parameters interleaved per qubit), then CZ on the ring (i, i+1 mod n).
* ``ring_observable(n)``: H = sum_i Z_i Z_{i+1 mod n} + 0.5 sum_i X_i as
  (xmask, zmask, coef) terms.
"""
import math
import os

import numpy as np

from .circuit import QuantumCircuit

RANDOM_GATESET = ("h", "x", "y", "z", "s", "t", "rx", "ry", "rz")


def random_gates(n, depth, seed):
    """Gate stream of the seeded random circuit: (name, qubits, angle or None).
    Per layer a gate from RANDOM_GATESET on every qubit (angles ~ U[0, 2 pi)),
    then CX(control = perm[2i], target = perm[2i+1]) with p = 0.5 on the pairs
    of a random permutation."""
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(depth):
        for q in range(n):
            name = RANDOM_GATESET[int(rng.integers(0, len(RANDOM_GATESET)))]
            if name in ("rx", "ry", "rz"):
                yield name, (q,), float(rng.uniform(0.0, 2 * math.pi))
            else:
                yield name, (q,), None
        perm = rng.permutation(n)
        for i in range(n // 2):
            if rng.random() < 0.5:
                yield "cx", (int(perm[2 * i]), int(perm[2 * i + 1])), None


def random_circuit(n, depth, seed, measure_marker=False):
    qc = QuantumCircuit(n)
    for name, qs, th in random_gates(n, depth, seed):
        if th is None:
            getattr(qc, name)(*qs)
        else:
            getattr(qc, name)(th, *qs)
    if measure_marker:
        qc.save_statevector()
    return qc


def batch_spec(count=4096, seed=1004, nmin=12, nmax=22, dmin=10, dmax=40):
    """The batched-circuits workload: (num_qubits, depth, circuit seed) per
    circuit, qubit counts uniform in [nmin, nmax], depths uniform in
    [dmin, dmax], circuit seeds seed * 100000 + index."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, size=count)
    ds = rng.integers(dmin, dmax + 1, size=count)
    return [(int(ns[i]), int(ds[i]), seed * 100000 + i) for i in range(count)]


def qft_circuit(n, swaps=True):
    qc = QuantumCircuit(n)
    for j in reversed(range(n)):
        qc.h(j)
        for k in reversed(range(j)):
            qc.cp(math.pi / float(2 ** (j - k)), k, j)
    if swaps:
        for i in range(n // 2):
            qc.swap(i, n - 1 - i)
    return qc


def ansatz_params(n, layers, seed=1003):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-math.pi, math.pi, size=layers * 2 * n)


def ansatz_spec(n, layers):
    """Parameterised op list: (name, qubits, param_index or None)."""
    spec = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            spec.append(("ry", (q,), k))
            spec.append(("rz", (q,), k + 1))
            k += 2
        for q in range(n):
            spec.append(("cz", (q, (q + 1) % n), None))
    return spec


def ansatz_circuit(n, layers, theta):
    qc = QuantumCircuit(n)
    for name, qs, k in ansatz_spec(n, layers):
        if k is None:
            getattr(qc, name)(*qs)
        else:
            getattr(qc, name)(float(theta[k]), *qs)
    return qc


def ring_observable(n):
    terms = []
    for i in range(n):
        terms.append((0, (1 << i) | (1 << ((i + 1) % n)), 1.0))
    for i in range(n):
        terms.append((1 << i, 0, 0.5))
    return terms


RANDOM_STATE_STREAMS = 16


def random_state(n, seed, workers=None):
    """Seeded state with i.i.d. complex Gaussian amplitudes (real and
    imaginary parts standard normal), normalised in FP64, as complex128.

    The 2^(n+1) normals are drawn by RANDOM_STATE_STREAMS independent PCG64
    streams spawned from SeedSequence(seed), stream k filling the k-th equal
    slice of the interleaved (re, im) vector, so the state depends on the
    seed only, not on how many threads (``workers``, default all host cores,
    at most 16) draw it; at 30 qubits this takes seconds instead of a
    minute."""
    from concurrent.futures import ThreadPoolExecutor

    psi = np.empty(1 << n, dtype=np.complex128)
    flat = psi.view(np.float64)
    nstreams = min(RANDOM_STATE_STREAMS, flat.size)
    seqs = np.random.SeedSequence(seed).spawn(nstreams)
    step = flat.size // nstreams

    def draw(k):
        np.random.Generator(np.random.PCG64(seqs[k])).standard_normal(out=flat[k * step:(k + 1) * step])

    workers = workers or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        list(ex.map(draw, range(nstreams)))
    psi /= math.sqrt(np.vdot(psi, psi).real)
    return psi
