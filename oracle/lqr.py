"""Literal restatement of the flat time-space tiling (Lazy Qubit Reordering).

Test infrastructure.  Follows qgpusim/transpiler/passes.py step for step on a
Qiskit-free node IR: a node is any object with a ``qubits`` tuple; identity
(``is``) distinguishes nodes, as the reference's ``g is not gate`` removal does
(passes.py:236).  Python sets are used exactly where the reference uses them so
that fill order (``sorted(Q - QL)``) and set equality match.

Gate order: the reference takes ``dag.topological_op_nodes()`` (passes.py:366),
a Qiskit routine that is not importable here; this restatement takes the input
list as the topological order (program order) -- parity with Qiskit's
within-depth ordering is unpinned.
"""
from dataclasses import dataclass, field
from typing import Dict, List


@dataclass
class Tile:
    """passes.py:11-26."""
    tile_idx: int
    logical_qubits: List[int]
    global_to_local_map: Dict[int, int]
    layout: List[int]
    gates: list = field(default_factory=list)


def build_tile_layout(logical_qubits, num_qubits):
    """passes.py:29-37: QL first, then the remaining qubits ascending."""
    ql = list(logical_qubits)
    s = set(ql)
    return ql + [q for q in range(num_qubits) if q not in s]


def tile_construction(QL, Cremain, T_map):
    """Algorithm 2, passes.py:89-128."""
    avail = set(QL)
    G = []
    for U in list(Cremain):
        tg = T_map[id(U)]
        if tg.issubset(avail):
            G.append(U)
        else:
            avail -= tg
            if not avail:
                break
    return G


def qubit_mapping(Q, Cremain, T_map, nL):
    """Algorithm 3, passes.py:131-176."""
    avail = set(Q)
    QL = set()
    for U in list(Cremain):
        tg = T_map[id(U)]
        if tg.issubset(avail) and len(QL | tg) <= nL:
            QL |= tg
        else:
            avail -= tg
            if not avail:
                break
    if not QL:
        return set()
    if len(QL) < nL:
        for q in sorted(Q - QL):
            if len(QL) >= nL:
                break
            QL.add(q)
    return QL


def flat_tiling(Q, C, T_map, nL):
    """Algorithm 1, passes.py:182-243.  Returns (schedule, M) with M keyed by id."""
    QL = set(range(min(nL, len(Q))))
    Cremain = list(C)
    sched = []
    M = {}
    while Cremain:
        G = tile_construction(QL, Cremain, T_map)
        if not G:
            forced = Cremain[0]
            QL = set(T_map[id(forced)])
            if len(QL) < nL:
                for q in sorted(Q - QL):
                    if len(QL) >= nL:
                        break
                    QL.add(q)
            G = [forced]
        for gate in G:
            sched.append(gate)
            M[id(gate)] = set(QL)
            Cremain = [g for g in Cremain if g is not gate]
        if Cremain:
            QL = qubit_mapping(Q, Cremain, T_map, nL)
    return sched, M


def extract_tiles(sched, M):
    """passes.py:247-272: group consecutive gates with equal Q_L."""
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


def build_tiling_plan(sched, M, num_qubits):
    """passes.py:315-357."""
    plan = []
    for idx, (ql_set, gates) in enumerate(extract_tiles(sched, M)):
        ql = sorted(ql_set)
        plan.append(Tile(tile_idx=idx, logical_qubits=ql,
                         global_to_local_map={g: i for i, g in enumerate(ql)},
                         layout=build_tile_layout(ql, num_qubits), gates=gates))
    return plan


def depth_sort(nodes, num_qubits):
    """passes.py:369-383: ASAP depth per node, then a stable sort by depth."""
    last = {i: -1 for i in range(num_qubits)}
    depth = {}
    for node in nodes:
        qs = list(node.qubits)
        d = max(last[i] for i in qs) + 1 if qs else 0
        depth[id(node)] = d
        for i in qs:
            last[i] = d
    return sorted(nodes, key=lambda nd: depth[id(nd)])


def lqr(nodes, num_qubits, nL):
    """LazyQubitReordering.run (passes.py:359-398) on a program-order node list.

    Returns (tile_plan, schedule, M)."""
    Q = set(range(num_qubits))
    C = depth_sort(nodes, num_qubits)
    T_map = {id(nd): set(nd.qubits) for nd in C}
    sched, M = flat_tiling(Q, C, T_map, nL)
    return build_tiling_plan(sched, M, num_qubits), sched, M


def summarize(plan):
    """Reference-style counts: tiles, reorders (layout changes starting from the
    identity layout, backend.py:414-428) and unitary ops (markers excluded,
    backend.py:434-436)."""
    layout = list(range(len(plan[0].layout))) if plan else []
    reorders = 0
    ops = 0
    for t in plan:
        if t.layout != layout:
            reorders += 1
            layout = t.layout
        ops += sum(1 for g in t.gates if getattr(g, "unitary", True))
    return {"num_tiles": len(plan), "num_reorders": reorders, "num_gates": ops}
