"""Host driver of the tile path: ``run_custom_backend`` and its helpers.

Mirrors qgpusim/backends/custom/backend.py:16-522 (signature, tile loop,
fusion rule, metrics keys), with the native calls going to the HIP backend.

Differences from the reference, each deliberate:

* the gate-matrix cache is keyed on (name, params, dtype); the reference's key
  omits the dtype (backend.py:30-35);
* three-qubit gates with exact 1q/2q networks (ccx, ccz, cswap) run as those
  networks inside their tile (all their qubits are local there); any other
  op on three or four qubits runs as one dense 16x16 register step in the
  chunked modes (one 3- or 4-qubit stream kernel in the sequential mode);
  ops on more qubits raise ``ValueError``; the reference drops them all
  silently (backend.py:448-451);
* fused matrices are multiplied in complex128 and then cast;
* ``qc.global_phase`` is applied (the reference never reads it): folded into
  the matrix of the first op the circuit applies, so it costs no pass;
* ``use_cuda_graphs=False`` is honoured (the reference accepts and ignores
  it, backend.py:336): ``tile_mode="graphed"`` then runs the same chunk passes
  without capturing a graph, and the metrics report "cooperative";
* the state is returned in the circuit's qubit order also when the plan has
  no final all-qubit tile (the reference then returns it in the last tile's
  layout, passes.py:213-227);
* the layout changes after the last tile with ops (the final all-qubit
  tile's, and the return to the circuit's order) are written by that tile's
  last chunk pass when it has room for them (``apply_tile_permute``), else
  recorded and applied by one permutation pass when the state is read;
* SWAP gates run as relabelling of stored positions (``absorb_swaps``): no
  data moves until the next layout change or the final order is written; the
  metrics still count them;
* for states of 20 qubits or more, consecutive tiles run as one op list when
  the chunk planner needs fewer passes for them together (``tile_program``,
  ``merge_tiles``); the metrics still count per plan tile;
* there is no CuPy fallback simulator: the HIP library is required.

``tile_mode``:
  "sequential"  -> one coalesced HBM stream kernel per op (the plain path);
  "cooperative" -> the tile's ops planned into LDS chunk passes;
  "graphed"     -> the same chunk passes, captured once into a hipGraph per
                   distinct tile and replayed on later runs.
"""
import time
from collections import OrderedDict
from typing import List

import numpy as np

from ...circuit import DECOMPOSITIONS, decompose
from ...transpiler.passes import Tile
from .native import qgpusim_hip as qgpusim_cuda
from .native.qgpusim_hip import COOPERATIVE, GRAPHED, SEQUENTIAL

_TILE_MODE_CODES = {"cooperative": COOPERATIVE, "graphed": GRAPHED, "sequential": SEQUENTIAL}

USE_NATIVE = True

MAX_CACHE_SIZE = 512
# Periodic renormalisation (complex64 runs): every RENORM_EVERY tiles with
# ops, the state's norm is reduced in FP64 on the GPU (one read of the state);
# when |<psi|psi> - 1| exceeds RENORM_TOL, 1/||psi|| is folded into the first
# matrix of the next tile with ops, so the scaling costs no pass.
RENORM_EVERY = 16
RENORM_TOL = 1e-6
GATE_CACHE: "OrderedDict" = OrderedDict()
KERNEL_LAUNCH_OVERHEAD_US = 10.0


def _gate_key(op, dtype):
    """Cache key (name, float params, dtype), or None when the matrix must not
    be cached: a unitary given by its matrix, or any parameter that is not a
    real scalar (an array, a string, a symbolic expression), since two such
    gates of one name can hold different matrices."""
    name = getattr(op, "name", "")
    if name == "unitary":
        return None
    params = []
    for p in getattr(op, "params", []):
        if isinstance(p, (str, bytes)) or np.ndim(p) != 0:
            return None
        try:
            params.append(float(p))
        except (TypeError, ValueError):
            return None
    return (name, tuple(params), np.dtype(dtype).str)


def get_gate_matrix_host(op, dtype=np.complex128):
    """Matrix of ``op`` as a C-contiguous array of ``dtype`` (LRU-cached)."""
    key = _gate_key(op, dtype)
    if key is not None:
        hit = GATE_CACHE.get(key)
        if hit is not None:
            GATE_CACHE.move_to_end(key)
            return hit
    U = np.ascontiguousarray(np.asarray(op.to_matrix(), dtype=np.complex128).astype(dtype))
    if key is not None:
        GATE_CACHE[key] = U
        while len(GATE_CACHE) > MAX_CACHE_SIZE:
            GATE_CACHE.popitem(last=False)
    return U


# The backend copies matrices into its own descriptor arena, so the "device"
# variant returns the same host matrix (kept for API compatibility).
get_gate_matrix_dev = get_gate_matrix_host


def clear_gate_cache():
    GATE_CACHE.clear()


def fuse_tile_ops(tile_ops, dtype=np.complex128):
    """Merge consecutive 1q ops on one qubit and 2q ops on the same ordered
    pair with nothing in between on their qubits (backend.py:88-200).
    Products are taken in complex128 (``U_new @ U_prev``)."""
    if len(tile_ops) <= 1:
        return tile_ops
    result = []
    last_gate_idx = {}

    def touches(o, qs):  # o = (kind, qubits..., U)
        return any(q in qs for q in o[1:-1])

    for op in tile_ops:
        if op[0] == "1q":
            q, U = op[1], op[2]
            idx = last_gate_idx.get(q)
            if idx is not None:
                prev = result[idx]
                if prev[0] == "1q" and prev[1] == q and not any(
                        touches(result[i], (q,)) for i in range(idx + 1, len(result))):
                    fused = np.asarray(U, dtype=np.complex128) @ np.asarray(prev[2], dtype=np.complex128)
                    result[idx] = ("1q", q, fused.astype(dtype))
                    continue
            last_gate_idx[q] = len(result)
            result.append(("1q", q, U))
        elif op[0] == "2q":
            q0, q1, U = op[1], op[2], op[3]
            prev_idx = None
            for i in range(len(result) - 1, -1, -1):
                p = result[i]
                if p[0] == "2q" and p[1] == q0 and p[2] == q1:
                    prev_idx = i
                    break
                if touches(p, (q0, q1)):
                    break
            if prev_idx is not None and not any(
                    touches(result[i], (q0, q1)) for i in range(prev_idx + 1, len(result))):
                fused = np.asarray(U, dtype=np.complex128) @ np.asarray(result[prev_idx][3], dtype=np.complex128)
                result[prev_idx] = ("2q", q0, q1, fused.astype(dtype))
                last_gate_idx[q0] = prev_idx
                last_gate_idx[q1] = prev_idx
                continue
            last_gate_idx[q0] = len(result)
            last_gate_idx[q1] = len(result)
            result.append(("2q", q0, q1, U))
        else:
            result.append(op)
    return result


def fold_global_phase(tile_ops, phase, dtype, scale=1.0):
    """``tile_ops`` with scale * exp(i * phase) folded into the matrix of its
    first op (product in complex128, then cast); unchanged when the factor is
    1 or the list is empty."""
    if (not phase and scale == 1.0) or not tile_ops:
        return tile_ops
    first = tile_ops[0]
    U = (np.asarray(first[-1], dtype=np.complex128) * (scale * np.exp(1j * phase))).astype(dtype)
    return [tuple(first[:-1]) + (U,)] + list(tile_ops[1:])


def _global_phase(qc):
    try:
        return float(getattr(qc, "global_phase", 0.0) or 0.0)
    except TypeError:  # an unbound parameter expression
        raise ValueError("run_custom_backend: the circuit's global phase has unbound parameters")


def build_tile_ops(tile: Tile, qubit_map, np_dtype):
    """Ops of one tile with local qubit indices (backend.py:431-451)."""
    ops = []
    ngates = 0
    for node in tile.gates:
        op = node.op
        if not hasattr(op, "to_matrix"):
            continue
        ngates += 1
        U = get_gate_matrix_host(op, np_dtype)
        local = [tile.global_to_local_map[qubit_map[q]] for q in node.qargs]
        if len(local) == 1:
            ops.append(("1q", local[0], U))
        elif len(local) == 2:
            ops.append(("2q", local[0], local[1], U))
        elif op.name in DECOMPOSITIONS:
            for g, qs in decompose(op.name, local):
                Ug = get_gate_matrix_host(g, np_dtype)
                ops.append(("1q", qs[0], Ug) if len(qs) == 1 else ("2q", qs[0], qs[1], Ug))
        elif len(local) in (3, 4):
            # one dense register step inside the tile (all its qubits are local)
            ops.append((f"{len(local)}q", *local, U))
        else:
            raise ValueError(f"gate {op.name!r} acts on {len(local)} qubits; "
                             "the tile path supports gates on at most 4 qubits")
    return ops, ngates


_SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


def absorb_swaps(tile_ops):
    """Run SWAP gates as relabelling instead of data movement.  A SWAP of
    positions (a, b) exchanges which stored positions hold the tile's local
    qubits a and b: the ops after it address the exchanged positions and the
    op is dropped.  Returns (ops, where) with where[l] the stored position
    holding the tile's local qubit l after the ops; a SWAP merged with another
    gate by fusion is no longer exact and stays an op."""
    where = None
    out = []
    for o in tile_ops:
        if o[0] == "2q" and np.array_equal(np.asarray(o[3]), _SWAP):
            if where is None:
                where = list(range(1 + max(max(x[1:-1]) for x in tile_ops)))
            where[o[1]], where[o[2]] = where[o[2]], where[o[1]]
            continue
        if where is None:
            out.append(o)
        else:
            out.append((o[0], *[where[q] for q in o[1:-1]], o[-1]))
    return out, where


def _relabelled(actual, layout, where):
    """The stored arrangement (position -> qubit) after a tile's absorbed
    SWAPs, from `actual`: the tile's local qubit l is qubit layout[l], and
    after the SWAPs it sits where the qubit at local position where[l] sat
    (no data moves)."""
    if where is None:
        return list(actual)
    pos = {q: p for p, q in enumerate(actual)}
    out = list(actual)
    for l, w in enumerate(where):
        out[pos[layout[w]]] = layout[l]
    return out


def _chunk_passes(ops, num_qubits, np_dtype):
    """Chunk passes the planner makes for an op list (host only, default
    chunk geometry: 13 chunk bits and 4 row bits for complex64, 11 and 3 for
    complex128)."""
    if not ops:
        return 0
    from .native import qgpusim_hip as H
    arr, arena = H.pack_ops(ops, num_qubits, np_dtype)
    K, cmin = (13, 4) if np_dtype == np.complex64 else (11, 3)
    return len(H.plan_describe(num_qubits, arr, len(ops), arena, K, cmin)[0])


# Merged op lists stay bounded: at most this many ops, and on average at
# most this many per chunk pass (each pass becomes one specialised kernel, whose
# size grows with its ops; a state that fits one chunk would otherwise take a
# whole circuit into one kernel).
MERGE_MAX_OPS = 2048
MERGE_MAX_OPS_PER_PASS = 192
# Below this many qubits a chunk pass costs less than compiling a kernel for a
# merged op list (which, unlike per-tile lists, rarely repeats): no merging.
MERGE_MIN_QUBITS = 20


def tile_program(tile_plan, qubit_map, np_dtype, num_qubits, enable_fusion=True, merge_tiles=True):
    """The calls that run a tile plan: a list of
    ("perm", old, new)                         layout change (no data moves),
    ("tile", ops, tiles)                        one op list,
    ("tile_perm", ops, old, new, changes)       the last op list, its last pass
                                                writing the circuit's order,
    ops in stored positions of the arrangement the call starts from, `tiles`
    the plan tiles the call runs, and per-tile (gates, fused ops, ops run).

    Consecutive tiles run as one op list when the planner needs fewer chunk
    passes for them together than apart (a later tile's ops are addressed
    through the first one's layout: the chunk planner, not the tile plan,
    decides what moves where); SWAPs run as relabelling (absorb_swaps)."""
    n = num_qubits
    identity = list(range(n))
    merge_tiles = merge_tiles and n >= MERGE_MIN_QUBITS
    per_tile = []  # (layout, fused ops on logical qubits, gates, fused ops)
    for tile in tile_plan:
        ops, ng = build_tile_ops(tile, qubit_map, np_dtype)
        if ops and enable_fusion:
            ops = fuse_tile_ops(ops, np_dtype)
        lops = [(o[0], *[tile.layout[q] for q in o[1:-1]], o[-1]) for o in ops]
        per_tile.append((list(tile.layout), lops, ng, len(ops)))

    def in_frame(lops, frame):
        pos = {q: p for p, q in enumerate(frame)}
        return [(o[0], *[pos[q] for q in o[1:-1]], o[-1]) for o in lops]

    # groups of consecutive tiles: (frame layout, ops in frame positions, tiles)
    groups = []
    cur_passes = 0
    for i, (layout, lops, _, _) in enumerate(per_tile):
        if not lops:
            if groups:
                groups[-1][2].append(i)
            continue
        if groups and merge_tiles and len(groups[-1][1]) + len(lops) <= MERGE_MAX_OPS:
            frame, gops, gt = groups[-1]
            cand = gops + in_frame(lops, frame)
            p_cand = _chunk_passes(absorb_swaps(cand)[0], n, np_dtype)
            p_own = _chunk_passes(absorb_swaps(in_frame(lops, layout))[0], n, np_dtype)
            if p_cand < cur_passes + p_own and len(cand) <= MERGE_MAX_OPS_PER_PASS * p_cand:
                groups[-1] = (frame, cand, gt + [i])
                cur_passes = p_cand
                continue
        groups.append((layout, in_frame(lops, layout), [i]))
        if merge_tiles:
            cur_passes = _chunk_passes(absorb_swaps(groups[-1][1])[0], n, np_dtype)
    runs = [absorb_swaps(g[1]) for g in groups]
    last = max([k for k, r in enumerate(runs) if r[0]], default=-1)
    prog = []
    actual = list(identity)
    for k, (frame, _, gt) in enumerate(groups):
        run, where = runs[k]
        if run and actual != frame:
            prog.append(("perm", list(actual), list(frame)))
            actual = list(frame)
        after = actual if k > last >= 0 else _relabelled(actual, frame, where)
        if run:
            if k == last:
                # later SWAPs only relabel: the last pass writes the circuit's
                # order from the arrangement after all of them
                for j in range(k + 1, len(groups)):
                    after = _relabelled(after, groups[j][0], runs[j][1])
            if k == last and after != identity:
                changes = sum(1 for j in range(gt[-1] + 1, len(tile_plan))
                              if tile_plan[j].layout != tile_plan[j - 1].layout)
                prog.append(("tile_perm", run, list(after), list(identity), max(1, changes), gt))
                after = list(identity)
            else:
                prog.append(("tile", run, gt))
        actual = after
    if actual != identity:
        prog.append(("perm", list(actual), list(identity)))
    stats = [(ng, nf, sum(1 for o in absorb_swaps(in_frame(lops, layout))[0]))
             for layout, lops, ng, nf in per_tile]
    return prog, stats


def run_custom_backend(qc, tile_plan: List[Tile], task: str, shots: int = None,
                       use_float32: bool = False, use_cuda_graphs: bool = True,
                       tile_mode: str = "cooperative", enable_fusion: bool = True, sample_seed: int = 0,
                       renormalize_every: int = None, renormalize_tol: float = RENORM_TOL,
                       merge_tiles: bool = True):
    """Simulate ``qc`` along ``tile_plan`` on the GPU (backend.py:336-522).

    Returns ``(statevector, metrics)``; ``metrics`` has the reference's keys:
    num_tiles, num_gates, num_fused_ops, fusion_ratio, num_reorders,
    num_kernel_launches, num_graph_replays, avoided_launches, t_perm_s,
    t_apply_s, t_sync_s, t_total_s, t_launch_overhead_saved_us, tile_mode.
    ``permute`` moves no data (it composes the state's stored bit order; the
    next tile's first chunk pass moves only the qubits that tile needs low,
    and the last layout is restored when the state is read), so t_perm_s
    holds only its host side and the data movement is in t_apply_s and
    t_sync_s.  With ``task="sampling"`` and ``shots``, ``metrics["counts"]``
    holds outcomes sampled on the GPU from the final state (seed
    ``sample_seed``) over the circuit's measurements, or all qubits.
    ``renormalize_every`` (default RENORM_EVERY for complex64, 0 = never for
    complex128): tiles with ops between FP64 norm checks; a norm off by more
    than ``renormalize_tol`` is corrected in the next tile's first matrix
    (``metrics["renorm_checks"]``, ``metrics["renormalizations"]``).
    ``merge_tiles`` (default True): consecutive tiles run as one op list when
    the chunk planner needs fewer passes for them together (tile_program);
    the metrics still count per plan tile.
    """
    if tile_mode not in ("cooperative", "graphed", "sequential"):
        raise ValueError(f"unknown tile_mode {tile_mode!r}")
    if tile_mode == "graphed" and not use_cuda_graphs:
        tile_mode = "cooperative"  # the same chunk passes, no graph capture
    num_qubits = qc.num_qubits
    np_dtype = np.complex64 if use_float32 else np.complex128
    qubit_map = {q: qc.find_bit(q).index for q in qc.qubits}

    t_start = time.perf_counter()
    num_tiles = len(tile_plan)
    num_reorders = num_gates = num_fused_ops = 0
    num_kernel_launches = num_graph_replays = 0
    t_perm = t_apply = 0.0

    phase = _global_phase(qc)  # applied with the first op (or to the result when nothing runs)
    if renormalize_every is None:
        renormalize_every = RENORM_EVERY if use_float32 else 0
    scale, since_check, renorm_checks, renorms = 1.0, 0, 0, 0
    sv = qgpusim_cuda.StatevectorF32(num_qubits) if use_float32 else qgpusim_cuda.Statevector(num_qubits)
    identity = list(range(num_qubits))
    t0 = time.perf_counter()
    prog, stats = tile_program(tile_plan, qubit_map, np_dtype, num_qubits, enable_fusion, merge_tiles)
    t_apply += time.perf_counter() - t0
    # metrics per plan tile, as the reference counts them
    layout = list(identity)
    for i, tile in enumerate(tile_plan):
        if tile.layout != layout:
            num_reorders += 1
            layout = tile.layout
        ng, nf, nrun = stats[i]
        num_gates += ng
        num_fused_ops += nf
        if nrun:
            if tile_mode == "cooperative":
                num_kernel_launches += 1
            elif tile_mode == "graphed":
                num_graph_replays += 1
            else:
                num_kernel_launches += nrun
    mode_code = _TILE_MODE_CODES[tile_mode]
    last_call = max([k for k, e in enumerate(prog) if e[0] != "perm"], default=-1)
    for k, entry in enumerate(prog):
        t0 = time.perf_counter()
        if entry[0] == "perm":
            sv.permute(entry[1], entry[2])
            t_perm += time.perf_counter() - t0
            continue
        tile_ops = entry[1]
        tile_ops, phase, scale = fold_global_phase(tile_ops, phase, np_dtype, scale), 0.0, 1.0
        if entry[0] == "tile_perm":
            sv.apply_tile_permute(mode_code, tile_ops, entry[2], entry[3])
        elif tile_mode == "cooperative":
            sv.apply_tile_cooperative(tile_ops)
        elif tile_mode == "graphed":
            sv.apply_tile_graphed(tile_ops)
        else:
            sv.apply_tile_dev(tile_ops)
        since_check += sum(1 for i in entry[-1] if stats[i][2])
        if renormalize_every and since_check >= renormalize_every and k < last_call:
            since_check = 0
            renorm_checks += 1
            n2 = sv.norm_squared()
            if abs(n2 - 1.0) > renormalize_tol:
                scale = 1.0 / np.sqrt(n2)
                renorms += 1
        t_apply += time.perf_counter() - t0
    layout = identity
    t_sync0 = time.perf_counter()
    sv.synchronize()
    t_sync = time.perf_counter() - t_sync0
    counts = None
    if task == "sampling" and shots:
        from ...sampling import counts_from_indices, measurements_of
        meas = measurements_of(qc)
        nbits = max(getattr(qc, "num_clbits", 0) or 0, 1 + max(c for _, c in meas))
        counts = counts_from_indices(sv.sample(int(shots), seed=sample_seed), layout, meas, nbits)
    result = sv.to_numpy()
    if phase:
        result *= np.exp(1j * phase).astype(np_dtype)
    counters = sv.counters()
    del sv

    t_total = time.perf_counter() - t_start
    fusion_ratio = num_gates / max(1, num_fused_ops) if enable_fusion else 1.0
    if tile_mode == "graphed":
        avoided = num_fused_ops - num_graph_replays
    elif tile_mode == "cooperative":
        avoided = num_fused_ops - num_tiles
    else:
        avoided = 0
    metrics = {
        "num_tiles": num_tiles,
        "num_gates": num_gates,
        "num_fused_ops": num_fused_ops,
        "fusion_ratio": fusion_ratio,
        "num_reorders": num_reorders,
        "num_kernel_launches": num_kernel_launches,
        "num_graph_replays": num_graph_replays,
        "avoided_launches": avoided,
        "t_perm_s": t_perm,
        "t_apply_s": t_apply,
        "t_sync_s": t_sync,
        "t_total_s": t_total,
        "t_launch_overhead_saved_us": avoided * KERNEL_LAUNCH_OVERHEAD_US,
        "tile_mode": tile_mode,
        # measured by the backend (not in the reference)
        "device_kernel_launches": counters["kernel_launches"],
        "device_graph_launches": counters["graph_launches"],
        "device_chunk_passes": counters["chunk_passes"],
        "device_folded_permutations": counters["folded_permutations"],
        "device_permutation_passes": counters["permutations"],
        "renorm_checks": renorm_checks,
        "renormalizations": renorms,
    }
    if counts is not None:
        metrics["counts"] = counts
    return result, metrics
