"""Lazy Qubit Reordering (flat time-space tiling) and the ``Tile`` plan.

Same algorithms, outputs and ``property_set`` keys as
qgpusim/transpiler/passes.py:83-420, so the tile schedule is identical for the
same gate order.  The sets of the reference are held as integer bitmasks here
(bit q = qubit q), so set tests, unions and the "lowest free qubit" fill are
single integer operations, and a tile's gates leave the remaining list in one
filter instead of one list rebuild per gate (passes.py:236), which makes the
schedule O(G * tiles) instead of O(G^2).

Works on this package's ``DAGCircuit`` and, by duck typing, on a Qiskit DAG
(``dag.qubits``, ``dag.topological_op_nodes()``, ``node.qargs``, ``node.op``).
"""
from dataclasses import dataclass, field
from typing import Dict, List


@dataclass
class Tile:
    """One tile of the schedule (passes.py:11-26)."""
    tile_idx: int
    # global qubit ids that are local in this tile (Q_L), ascending
    logical_qubits: List[int]
    # global qubit -> local bit index (0..|Q_L|-1)
    global_to_local_map: Dict[int, int]
    # layout[pos] = global qubit at bit position pos
    layout: List[int]
    # gates of this tile, in execution order
    gates: list = field(default_factory=list)


def _build_tile_layout(logical_qubits: List[int], num_qubits: int) -> List[int]:
    """Q_L at the lowest positions, then the other qubits ascending
    (passes.py:29-37)."""
    inq = set(logical_qubits)
    return list(logical_qubits) + [q for q in range(num_qubits) if q not in inq]


def _bits(mask: int) -> List[int]:
    out = []
    q = 0
    while mask:
        if mask & 1:
            out.append(q)
        mask >>= 1
        q += 1
    return out


def _fill_lowest(ql: int, num_qubits: int, nL: int) -> int:
    """Add the lowest qubits not in ql until it holds nL (passes.py:168-173)."""
    have = bin(ql).count("1")
    q = 0
    while have < nL and q < num_qubits:
        if not (ql >> q) & 1:
            ql |= 1 << q
            have += 1
        q += 1
    return ql


class BasePass:
    def __init__(self):
        self.property_set = {}


class LazyQubitReordering(BasePass):
    """Flat time-space tiling with lazy qubit reordering (passes.py:83-420)."""

    def __init__(self, nL: int):
        super().__init__()
        self.nL = nL

    # Algorithm 2 (passes.py:89-128)
    @staticmethod
    def tile_construction(QL: int, Cremain: list, T: Dict[int, int]) -> list:
        avail = QL
        G = []
        for U in Cremain:
            t = T[id(U)]
            if t & ~avail == 0:
                G.append(U)
            else:
                avail &= ~t
                if not avail:
                    break
        return G

    # Algorithm 3 (passes.py:131-176)
    @staticmethod
    def greedy_qubit_mapping(num_qubits: int, Cremain: list, T: Dict[int, int], nL: int) -> int:
        avail = (1 << num_qubits) - 1
        QL = 0
        for U in Cremain:
            t = T[id(U)]
            if t & ~avail == 0 and bin(QL | t).count("1") <= nL:
                QL |= t
            else:
                avail &= ~t
                if not avail:
                    break
        if not QL:
            return 0
        return _fill_lowest(QL, num_qubits, nL)

    qubit_mapping = greedy_qubit_mapping

    # Algorithm 1 (passes.py:182-243)
    def flat_tiling(self, num_qubits: int, C: list, T: Dict[int, int], nL: int):
        QL = (1 << min(nL, num_qubits)) - 1
        Cremain = list(C)
        sched = []
        M = {}
        while Cremain:
            G = self.tile_construction(QL, Cremain, T)
            if not G:
                # force the first remaining gate into its own tile (passes.py:213-227)
                forced = Cremain[0]
                QL = _fill_lowest(T[id(forced)], num_qubits, nL)
                G = [forced]
            for g in G:
                sched.append(g)
                M[id(g)] = QL
            taken = {id(g) for g in G}
            Cremain = [g for g in Cremain if id(g) not in taken]
            if Cremain:
                QL = self.qubit_mapping(num_qubits, Cremain, T, nL)
        return sched, M

    @staticmethod
    def extract_tiles(sched: list, M: Dict[int, int]):
        """Group consecutive gates with equal Q_L (passes.py:247-272)."""
        tiles = []
        if not sched:
            return tiles
        cur = M[id(sched[0])]
        gates = [sched[0]]
        for node in sched[1:]:
            if M[id(node)] == cur:
                gates.append(node)
            else:
                tiles.append((cur, gates))
                cur = M[id(node)]
                gates = [node]
        tiles.append((cur, gates))
        return tiles

    @staticmethod
    def summarize_tiling(sched: list, M: Dict[int, int]):
        """Tile statistics (passes.py:274-313)."""
        tiles = LazyQubitReordering.extract_tiles(sched, M)
        if not tiles:
            return {"num_tiles": 0, "avg_tile_size": 0.0, "max_tile_size": 0,
                    "avg_Q_L_size": 0.0, "max_Q_L_size": 0}
        sizes = [len(g) for _, g in tiles]
        qls = [bin(q).count("1") for q, _ in tiles]
        return {"num_tiles": len(tiles), "avg_tile_size": sum(sizes) / len(tiles),
                "max_tile_size": max(sizes), "avg_Q_L_size": sum(qls) / len(tiles),
                "max_Q_L_size": max(qls)}

    @staticmethod
    def build_tiling_plan(sched: list, M: Dict[int, int], num_qubits: int) -> List[Tile]:
        """passes.py:315-357."""
        plan = []
        for idx, (ql, gates) in enumerate(LazyQubitReordering.extract_tiles(sched, M)):
            qs = _bits(ql)
            plan.append(Tile(tile_idx=idx, logical_qubits=qs,
                             global_to_local_map={g: i for i, g in enumerate(qs)},
                             layout=_build_tile_layout(qs, num_qubits), gates=gates))
        return plan

    def schedule(self, nodes: list, qindex: dict, num_qubits: int):
        """Depth sort + flat tiling on nodes in topological order
        (passes.py:366-392).  Returns (schedule, M) with M keyed by id(node)."""
        last = [-1] * num_qubits
        depth = {}
        T = {}
        for node in nodes:
            qs = [qindex[q] for q in node.qargs]
            d = max(last[i] for i in qs) + 1 if qs else 0
            depth[id(node)] = d
            for i in qs:
                last[i] = d
            m = 0
            for i in qs:
                m |= 1 << i
            T[id(node)] = m
        C = sorted(nodes, key=lambda nd: depth[id(nd)])  # stable
        return self.flat_tiling(num_qubits, C, T, self.nL)

    def run(self, dag):
        """passes.py:359-420."""
        num_qubits = len(dag.qubits)
        qindex = {q: i for i, q in enumerate(dag.qubits)}
        nodes = list(dag.topological_op_nodes())
        sched, M = self.schedule(nodes, qindex, num_qubits)
        byid = {id(nd): nd for nd in nodes}
        self.property_set["flat_tiling_schedule"] = sched
        self.property_set["flat_tiling_mapping"] = {byid[k]: set(_bits(v)) for k, v in M.items()}
        self.property_set["flat_tiling_plan"] = self.build_tiling_plan(sched, M, num_qubits)
        new_dag = dag.copy_empty_like()
        for node in sched:
            new_dag.apply_operation_back(node.op, qargs=node.qargs, cargs=node.cargs)
        return new_dag
