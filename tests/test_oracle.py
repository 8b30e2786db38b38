"""Pins of the CPU oracle: every function is checked against something it does
not compute itself (textbook identities, brute force, numpy.transpose, the
FFT, and the reference's own published schedule counts)."""
import numpy as np
import pytest

from oracle import circuits as C
from oracle import dense, gates as G, lqr
from oracle import permute as P
from oracle import svref

from conftest import result_csv_rows


def test_gate_identities():
    I2 = np.eye(2)
    assert np.allclose(G.h() @ G.h(), I2)
    assert np.allclose(G.s() @ G.s(), G.z())
    assert np.allclose(G.t() @ G.t(), G.s())
    assert np.allclose(G.rx(np.pi), -1j * G.x())
    assert np.allclose(G.ry(np.pi), -1j * G.y())
    assert np.allclose(G.rz(np.pi), -1j * G.z())
    assert np.allclose(G.p(0.3), np.exp(0.15j) * G.rz(0.3))
    assert np.allclose(G.x() @ G.y(), 1j * G.z())
    for m in (G.cx(), G.cz(), G.cp(0.7), G.swap()):
        assert np.allclose(m @ m.conj().T, np.eye(4))


def test_little_endian_cx():
    # CX(control=q0, target=q1) on |q1 q0> = |0 1> (index 1) gives |1 1> (index 3)
    v = np.zeros(4)
    v[1] = 1
    assert np.argmax(abs(G.cx() @ v)) == 3
    v = np.zeros(4)
    v[2] = 1  # control clear: unchanged
    assert np.argmax(abs(G.cx() @ v)) == 2


@pytest.mark.parametrize("qargs", [[0], [3], [1, 4], [4, 1], [0, 5], [2, 3], [5, 0, 3], [1, 2, 0],
                                   [4, 0, 5, 2], [0, 1, 2, 3]])
def test_dense_apply_matches_brute_force(qargs):
    rng = np.random.default_rng(7)
    n = 6
    k = len(qargs)
    U = np.linalg.qr(rng.standard_normal((2 ** k, 2 ** k)) + 1j * rng.standard_normal((2 ** k, 2 ** k)))[0]
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    assert np.allclose(dense.apply(psi, U, qargs, n), dense.full_operator(U, qargs, n) @ psi, atol=1e-13)


def test_bit_order_is_little_endian():
    # X on qubit q flips bit q of the index
    n = 5
    for q in range(n):
        out = dense.apply(dense.zero_state(n), G.x(), [q], n)
        assert np.argmax(abs(out)) == 1 << q


def test_c_port_matches_dense():
    n = 9
    ops = C.random_circuit(n, 10, 11) + C.qft_circuit(n)
    g = C.to_dense_gates(ops)
    ref = dense.simulate(g, n)
    psi = dense.zero_state(n)
    svref.run(psi, n, g)
    assert np.max(abs(psi - ref)) < 1e-12
    psi32 = dense.zero_state(n).astype(np.complex64)
    svref.run(psi32, n, g)
    assert 1 - dense.fidelity(ref, psi32) < 1e-5


@pytest.mark.parametrize("seed", range(4))
def test_permute_matches_transpose(seed):
    rng = np.random.default_rng(seed)
    n = 8
    lo = list(rng.permutation(n))
    ln = list(rng.permutation(n))
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    bm = P.perm_map(lo, ln)
    t = psi.reshape((2,) * n)
    axes = [0] * n
    for p in range(n):
        axes[n - 1 - bm[p]] = n - 1 - p
    expect = np.transpose(t, axes).reshape(-1)
    assert np.array_equal(P.scatter(psi, bm), expect)
    assert np.array_equal(svref.permute(psi, n, bm), expect)


def test_permute_semantics_moves_qubits():
    # a qubit's bit follows it to its new position
    n = 4
    psi = dense.apply(dense.zero_state(n), G.x(), [0], n)  # global qubit 0 set
    out = P.permute(psi, [0, 1, 2, 3], [2, 3, 0, 1])       # qubit 0 now at position 2
    assert np.argmax(abs(out)) == 1 << 2


@pytest.mark.parametrize("lo,ln,msg", [
    ([0, 1, 2], [0, 1, 3], "layout_new contains invalid qubit id"),
    ([0, 1, 2], [0, 0, 1], "layout_new is not a permutation"),
    ([0, 1, 5], [0, 1, 2], "layout_old contains invalid qubit id"),
    ([0, 1], [0, 1, 2], "layout_old/layout_new must have length num_qubits"),
])
def test_permute_errors(lo, ln, msg):
    with pytest.raises(P.LayoutError, match=msg):
        P.perm_map(lo, ln)


@pytest.mark.parametrize("row", result_csv_rows(), ids=lambda r: f"csv_line{r['csv_line']}")
def test_lqr_reference_known_answers(row):
    """The reference's published custom-backend rows (results/result.csv):
    the gate-structure twin of its transpiled random circuit has the row's
    shape (qubits, gates, CX, depth with the save marker taking no layer),
    and the schedule restatement gives the row's tiles, reorders and ops."""
    n, d = row["nqubits"], row["depth"]
    body = C.reference_random_circuit(n, d)
    circ = body + [C.marker(n)]
    level = [0] * n
    for g in body:
        top = max(level[q] for q in g.qubits) + 1
        for q in g.qubits:
            level[q] = top
    assert (n, len(body), max(level)) == (row["nqubits_t"], row["total_gates"], row["depth_t"])
    assert sum(g.name == "cx" for g in body) == row["cx_count"] == row["twoq_count"]
    plan, _, _ = lqr.lqr(circ, n, row["nL"])
    assert lqr.summarize(plan) == {"num_tiles": row["num_tiles"], "num_reorders": row["num_reorders"],
                                   "num_gates": row["num_gates_custom"]}


@pytest.mark.parametrize("seed", range(6))
def test_lqr_invariants(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 12))
    nL = int(rng.integers(2, n + 1))
    circ = C.random_circuit(n, int(rng.integers(2, 8)), seed) + [C.marker(n)]
    plan, sched, M = lqr.lqr(circ, n, nL)
    flat = [g for t in plan for g in t.gates]
    assert len(flat) == len(circ) and {id(g) for g in flat} == {id(g) for g in circ}
    for t in plan:
        assert sorted(t.layout) == list(range(n))
        assert t.layout[:len(t.logical_qubits)] == t.logical_qubits
        for g in t.gates:
            assert set(g.qubits) <= set(t.logical_qubits)
        if not any(len(g.qubits) > nL for g in t.gates):
            assert len(t.logical_qubits) <= max(nL, 1)
    # per-qubit program order preserved
    pos = {id(g): i for i, g in enumerate(flat)}
    for q in range(n):
        seq = [pos[id(g)] for g in circ if q in g.qubits]
        assert seq == sorted(seq)


@pytest.mark.parametrize("seed", range(3))
def test_lqr_tiled_run_equals_program_order(seed):
    n, nL = 9, 4
    circ = C.random_circuit(n, 6, 100 + seed) + [C.marker(n)]
    plan, _, _ = lqr.lqr(circ, n, nL)
    ref = dense.simulate(C.to_dense_gates(circ), n)
    psi = dense.zero_state(n)
    layout = list(range(n))
    for t in plan:
        if t.layout != layout:
            psi = P.permute(psi, layout, t.layout)
            layout = t.layout
        for g in t.gates:
            if g.unitary:
                psi = dense.apply(psi, g.matrix(), [t.global_to_local_map[q] for q in g.qubits], n)
    assert layout == list(range(n))
    assert np.max(abs(psi - ref)) < 1e-12


@pytest.mark.parametrize("n", [3, 6, 8])
def test_qft_equals_fft(n):
    psi0 = C.random_state(n, 5)
    out = dense.simulate(C.to_dense_gates(C.qft_circuit(n)), n, psi0)
    expect = np.fft.ifft(psi0) * np.sqrt(1 << n)
    assert np.max(abs(out - expect)) < 1e-12


def test_ansatz_shape_and_observable_structure():
    n, L = 6, 3
    th = C.ansatz_params(n, L)
    ops = C.ansatz_circuit(n, L, th)
    assert len(th) == 2 * n * L
    assert sum(1 for o in ops if o.name == "cz") == n * L


def test_random_generator_counts():
    ops = C.random_circuit(20, 20, 1000)
    one = [o for o in ops if len(o.qubits) == 1]
    two = [o for o in ops if len(o.qubits) == 2]
    assert len(one) == 400
    assert 60 < len(two) < 140


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_tiled_c_port_matches_dense(dtype, tol):
    """The C port's tiled run (blocks of 2^9 amplitudes each running the whole
    list) equals program-order FP64 simulation; an op outside the block is
    refused."""
    n, b = 13, 9
    ops = [(U, q) for U, q in C.to_dense_gates(C.random_circuit(n, 8, 31)) if max(q) < b]
    psi0 = C.random_state(n, 2)
    ref = dense.simulate(ops, n, psi0)
    got = psi0.astype(dtype)
    svref.run_tiled(got, n, ops, b)
    assert np.max(np.abs(got - ref)) < tol
    with pytest.raises(ValueError):
        svref.run_tiled(psi0.astype(dtype), n, [(G.x(), [b])], b)
