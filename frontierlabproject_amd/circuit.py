"""Minimal Qiskit-compatible circuit IR (Qiskit is not a dependency).

Provides the small part of Qiskit's API that the tile path touches:

* ``QuantumCircuit`` with ``num_qubits``, ``qubits``, ``data``, ``find_bit``,
  ``global_phase`` and the usual gate methods, plus ``save_statevector`` (the
  all-qubit terminal marker the reference appends, qgpusim/runner.py:38-40);
* gate objects with ``name``, ``params``, ``num_qubits`` and ``to_matrix()``
  in Qiskit's little-endian convention (a 2-qubit matrix on qargs (q0, q1) is
  indexed by ``b_q0 + 2*b_q1``);
* ``DAGCircuit``/``DAGOpNode`` with ``qubits``, ``topological_op_nodes()``,
  ``copy_empty_like()`` and ``apply_operation_back()``, which is what
  ``LazyQubitReordering.run`` uses (qgpusim/transpiler/passes.py:359-420).

``topological_op_nodes()`` returns program order, which is a valid
topological order.  Qiskit's own order (a lexicographic topological sort) is
not reproduced; a real Qiskit DAG can be passed to the pass instead.
"""
import math
from typing import List, Sequence

import numpy as np

_R2 = 1.0 / math.sqrt(2.0)


class Qubit:
    """A circuit qubit (identity-hashed, like Qiskit's bits)."""
    __slots__ = ("_index",)

    def __init__(self, index):
        self._index = index

    def __repr__(self):
        return f"Qubit({self._index})"


class BitLocations:
    __slots__ = ("index", "registers")

    def __init__(self, index):
        self.index = index
        self.registers = []


class Instruction:
    """Base of every operation: ``name``, ``num_qubits``, ``params``.
    Directives (barriers, the save marker) take no layer in ``depth()``."""

    _directive = False

    def __init__(self, name, num_qubits, params=()):
        self.name = name
        self.num_qubits = num_qubits
        self.num_clbits = 0
        self.params = list(params)

    def __repr__(self):
        p = ",".join(f"{x:.6g}" if isinstance(x, float) else str(x) for x in self.params)
        return f"{self.name}({p})" if p else self.name


class Gate(Instruction):
    """A unitary operation with a closed-form matrix."""

    def to_matrix(self):
        return _MATRIX[self.name](*self.params)

    def inverse(self):
        inv = _INVERSE.get(self.name)
        if inv is not None:
            name, params = inv(self.params)
            return Gate(name, self.num_qubits, params)
        return UnitaryGate(np.conj(self.to_matrix()).T)


class UnitaryGate(Gate):
    """Arbitrary unitary given as a matrix (little-endian over its qargs)."""

    def __init__(self, matrix):
        m = np.asarray(matrix, dtype=np.complex128)
        k = int(round(math.log2(m.shape[0])))
        if m.shape != (1 << k, 1 << k):
            raise ValueError("unitary must be a square 2^k matrix")
        super().__init__("unitary", k, [])
        self._m = m

    def to_matrix(self):
        return self._m.copy()

    def inverse(self):
        return UnitaryGate(np.conj(self._m).T)


class SaveStatevector(Instruction):
    """Terminal all-qubit marker without a matrix (skipped by the backend).
    A directive, as Aer's save instructions are."""

    _directive = True

    def __init__(self, num_qubits):
        super().__init__("save_statevector", num_qubits)


class Measure(Instruction):
    """Computational-basis measurement of one qubit into one classical bit."""

    def __init__(self):
        super().__init__("measure", 1)


class Barrier(Instruction):
    _directive = True

    def __init__(self, num_qubits):
        super().__init__("barrier", num_qubits)


def _rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def _rz(t):
    return np.array([[complex(math.cos(t / 2), -math.sin(t / 2)), 0],
                     [0, complex(math.cos(t / 2), math.sin(t / 2))]], dtype=np.complex128)


def _phase(lam):
    return np.array([[1, 0], [0, complex(math.cos(lam), math.sin(lam))]], dtype=np.complex128)


def _u(theta, phi, lam):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lam) * s],
                     [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]], dtype=np.complex128)


def _controlled(u):
    """Control on qargs[0] (bit 0 of the index), target qargs[1]."""
    m = np.eye(4, dtype=np.complex128)
    m[np.ix_([1, 3], [1, 3])] = u
    return m


_MATRIX = {
    "id": lambda: np.eye(2, dtype=np.complex128),
    "h": lambda: np.array([[_R2, _R2], [_R2, -_R2]], dtype=np.complex128),
    "x": lambda: np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "y": lambda: np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "z": lambda: np.array([[1, 0], [0, -1]], dtype=np.complex128),
    "s": lambda: np.array([[1, 0], [0, 1j]], dtype=np.complex128),
    "sdg": lambda: np.array([[1, 0], [0, -1j]], dtype=np.complex128),
    "t": lambda: _phase(math.pi / 4),
    "tdg": lambda: _phase(-math.pi / 4),
    "sx": lambda: 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], dtype=np.complex128),
    "rx": _rx,
    "ry": _ry,
    "rz": _rz,
    "p": _phase,
    "u": _u,
    "cx": lambda: _controlled(np.array([[0, 1], [1, 0]], dtype=np.complex128)),
    "cy": lambda: _controlled(np.array([[0, -1j], [1j, 0]], dtype=np.complex128)),
    "cz": lambda: np.diag([1, 1, 1, -1]).astype(np.complex128),
    "cp": lambda lam: np.diag([1, 1, 1, complex(math.cos(lam), math.sin(lam))]).astype(np.complex128),
    "crz": lambda t: _controlled(_rz(t)),
    "swap": lambda: np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                             dtype=np.complex128),
}

def _perm3(a, b):
    """3-qubit permutation matrix exchanging basis states a and b."""
    m = np.eye(8, dtype=np.complex128)
    m[[a, b]] = m[[b, a]]
    return m


# Three-qubit gates, index b_q0 + 2 b_q1 + 4 b_q2 (little-endian like the rest):
# ccx flips qargs[2] when qargs[0] and qargs[1] are 1 (|011> <-> |111>), ccz
# negates |111>, cswap exchanges qargs[1] and qargs[2] when qargs[0] is 1
# (|011> <-> |101>).
_MATRIX.update({
    "ccx": lambda: _perm3(3, 7),
    "ccz": lambda: np.diag([1, 1, 1, 1, 1, 1, 1, -1]).astype(np.complex128),
    "cswap": lambda: _perm3(3, 5),
})

_NUM_QUBITS = {name: (2 if name in ("cx", "cy", "cz", "cp", "crz", "swap") else 3 if name in ("ccx", "ccz", "cswap")
                      else 1) for name in _MATRIX}

# Exact 1q/2q networks of the three-qubit gates (no global phase): the
# Toffoli network of six CNOTs, T/T^dag and two Hadamards (Nielsen & Chuang,
# Fig. 4.9); CCZ = H_t CCX H_t; CSWAP = CX(b, a) CCX(c, a, b) CX(b, a).
# Entries are (gate name, positions into the gate's qargs).
_CCX_NET = [("h", (2,)), ("cx", (1, 2)), ("tdg", (2,)), ("cx", (0, 2)), ("t", (2,)), ("cx", (1, 2)),
            ("tdg", (2,)), ("cx", (0, 2)), ("t", (1,)), ("t", (2,)), ("h", (2,)), ("cx", (0, 1)),
            ("t", (0,)), ("tdg", (1,)), ("cx", (0, 1))]
DECOMPOSITIONS = {
    "ccx": _CCX_NET,
    "ccz": [("h", (2,))] + _CCX_NET + [("h", (2,))],
    "cswap": [("cx", (2, 1))] + _CCX_NET + [("cx", (2, 1))],
}


def decompose(name, qargs):
    """[(gate, qubits)] of the 1q/2q network for a three-qubit gate on qargs."""
    return [(gate(g), [qargs[i] for i in q]) for g, q in DECOMPOSITIONS[name]]


_SELF_INVERSE = ("id", "h", "x", "y", "z", "cx", "cy", "cz", "swap", "ccx", "ccz", "cswap")
_INVERSE = {name: (lambda p, n=name: (n, p)) for name in _SELF_INVERSE}
_INVERSE.update({
    "s": lambda p: ("sdg", []), "sdg": lambda p: ("s", []),
    "t": lambda p: ("tdg", []), "tdg": lambda p: ("t", []),
    "rx": lambda p: ("rx", [-p[0]]), "ry": lambda p: ("ry", [-p[0]]), "rz": lambda p: ("rz", [-p[0]]),
    "p": lambda p: ("p", [-p[0]]), "cp": lambda p: ("cp", [-p[0]]), "crz": lambda p: ("crz", [-p[0]]),
    "u": lambda p: ("u", [-p[0], -p[2], -p[1]]),
})


def gate(name, *params):
    """Gate object by Qiskit name."""
    if name not in _MATRIX:
        raise KeyError(f"unknown gate {name!r}")
    return Gate(name, _NUM_QUBITS[name], [float(x) for x in params])


class CircuitInstruction:
    __slots__ = ("operation", "qubits", "clbits")

    def __init__(self, operation, qubits, clbits=()):
        self.operation = operation
        self.qubits = tuple(qubits)
        self.clbits = tuple(clbits)

    def __iter__(self):  # legacy (op, qargs, cargs) unpacking
        return iter((self.operation, list(self.qubits), list(self.clbits)))


class QuantumCircuit:
    def __init__(self, num_qubits, global_phase=0.0):
        self.qubits: List[Qubit] = [Qubit(i) for i in range(num_qubits)]
        self.clbits: list = []
        self.data: List[CircuitInstruction] = []
        self.global_phase = float(global_phase)
        self._index = {q: i for i, q in enumerate(self.qubits)}
        self._nclbits = 0

    @property
    def num_qubits(self):
        return len(self.qubits)

    def find_bit(self, q):
        return BitLocations(self._index[q])

    def _q(self, q):
        return q if isinstance(q, Qubit) else self.qubits[int(q)]

    def append(self, op, qargs: Sequence, cargs: Sequence = ()):
        qs = tuple(self._q(q) for q in qargs)
        if len(set(qs)) != len(qs):
            raise ValueError("duplicate qubit arguments")
        self.data.append(CircuitInstruction(op, qs, cargs))
        return self

    def _g(self, name, qargs, *params):
        return self.append(gate(name, *params), qargs)

    # 1q
    def id(self, q): return self._g("id", [q])
    def h(self, q): return self._g("h", [q])
    def x(self, q): return self._g("x", [q])
    def y(self, q): return self._g("y", [q])
    def z(self, q): return self._g("z", [q])
    def s(self, q): return self._g("s", [q])
    def sdg(self, q): return self._g("sdg", [q])
    def t(self, q): return self._g("t", [q])
    def tdg(self, q): return self._g("tdg", [q])
    def sx(self, q): return self._g("sx", [q])
    def rx(self, theta, q): return self._g("rx", [q], theta)
    def ry(self, theta, q): return self._g("ry", [q], theta)
    def rz(self, phi, q): return self._g("rz", [q], phi)
    def p(self, lam, q): return self._g("p", [q], lam)
    def u(self, theta, phi, lam, q): return self._g("u", [q], theta, phi, lam)
    # 2q
    def cx(self, c, t): return self._g("cx", [c, t])
    def cy(self, c, t): return self._g("cy", [c, t])
    def cz(self, c, t): return self._g("cz", [c, t])
    def cp(self, lam, c, t): return self._g("cp", [c, t], lam)
    def crz(self, theta, c, t): return self._g("crz", [c, t], theta)
    def swap(self, a, b): return self._g("swap", [a, b])
    def ccx(self, c0, c1, t): return self._g("ccx", [c0, c1, t])
    def ccz(self, a, b, c): return self._g("ccz", [a, b, c])
    def cswap(self, c, a, b): return self._g("cswap", [c, a, b])

    def unitary(self, matrix, qargs):
        return self.append(UnitaryGate(matrix), qargs)

    def save_statevector(self):
        return self.append(SaveStatevector(self.num_qubits), self.qubits)

    def copy(self):
        c = QuantumCircuit(self.num_qubits, self.global_phase)
        c.qubits = self.qubits
        c._index = self._index
        c.data = list(self.data)
        c._nclbits = self._nclbits
        return c

    def inverse(self):
        """Adjoint circuit (unitary ops only)."""
        c = QuantumCircuit(self.num_qubits, -self.global_phase)
        for ci in reversed(self.data):
            if not hasattr(ci.operation, "to_matrix"):
                continue
            c.append(ci.operation.inverse(), [self._index[q] for q in ci.qubits])
        return c

    def count_ops(self):
        out = {}
        for ci in self.data:
            out[ci.operation.name] = out.get(ci.operation.name, 0) + 1
        return out

    def depth(self):
        """Layers of the circuit: each instruction sits one layer above the
        latest earlier instruction sharing a qubit or classical bit with it.
        Directives (barriers, the save marker) take no layer but align the
        wires they span (Qiskit's default filter).  The reference's published
        rows rely on this: depth_t = 48 for its 16-layer random circuits with
        the save marker present (results/result.csv:6,13)."""
        level = {}
        d = 0
        for ci in self.data:
            wires = [("q", self._index[q]) for q in ci.qubits] + [("c", c) for c in ci.clbits]
            top = max((level.get(w, 0) for w in wires), default=0)
            if getattr(ci.operation, "_directive", False):
                for w in wires:
                    level[w] = top
                continue
            for w in wires:
                level[w] = top + 1
            d = max(d, top + 1)
        return d

    @property
    def num_clbits(self):
        return self._nclbits

    def measure_all(self):
        """Barrier, then a measurement of every qubit into a new classical
        bit (bit i <- qubit i), as Qiskit's measure_all."""
        base = self._nclbits
        self._nclbits += self.num_qubits
        self.append(Barrier(self.num_qubits), self.qubits)
        for i, q in enumerate(self.qubits):
            self.data.append(CircuitInstruction(Measure(), (q,), (base + i,)))

    def __len__(self):
        return len(self.data)


class DAGOpNode:
    __slots__ = ("op", "qargs", "cargs")

    def __init__(self, op, qargs, cargs=()):
        self.op = op
        self.qargs = tuple(qargs)
        self.cargs = tuple(cargs)

    @property
    def name(self):
        return self.op.name

    def __repr__(self):
        return f"DAGOpNode({self.op!r}, {self.qargs})"


class DAGCircuit:
    def __init__(self):
        self.qubits: List[Qubit] = []
        self.clbits: list = []
        self.global_phase = 0.0
        self._nodes: List[DAGOpNode] = []

    def topological_op_nodes(self):
        return iter(list(self._nodes))

    def op_nodes(self):
        return list(self._nodes)

    def copy_empty_like(self):
        d = DAGCircuit()
        d.qubits = list(self.qubits)
        d.clbits = list(self.clbits)
        d.global_phase = self.global_phase
        d.num_clbits = getattr(self, "num_clbits", 0)
        return d

    def apply_operation_back(self, op, qargs=(), cargs=()):
        node = DAGOpNode(op, qargs, cargs)
        self._nodes.append(node)
        return node

    def size(self):
        return len(self._nodes)


def circuit_to_dag(qc: QuantumCircuit) -> DAGCircuit:
    d = DAGCircuit()
    d.qubits = list(qc.qubits)
    d.global_phase = qc.global_phase
    for ci in qc.data:
        d.apply_operation_back(ci.operation, ci.qubits, ci.clbits)
    d.num_clbits = qc.num_clbits
    return d


def dag_to_circuit(dag: DAGCircuit) -> QuantumCircuit:
    qc = QuantumCircuit(0, dag.global_phase)
    qc.qubits = list(dag.qubits)
    qc._index = {q: i for i, q in enumerate(qc.qubits)}
    for node in dag.topological_op_nodes():
        qc.data.append(CircuitInstruction(node.op, node.qargs, node.cargs))
    used = [int(c) for node in dag.topological_op_nodes() for c in node.cargs]
    qc._nclbits = max(getattr(dag, "num_clbits", 0), 1 + max(used) if used else 0)
    return qc
