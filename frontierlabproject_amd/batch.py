"""Batched circuits: thousands of small independent circuits, each returning
<Z_mask> and its adjoint gradient, from one persistent launch.

The circuits are packed into flat op lists over one shared complex64 arena
(fixed gates, their inverses and the rotation generators stored once; each
rotation angle contributes U(theta) and U(theta)^dag).  The C layer plans every
circuit into chunk passes and a persistent kernel runs them, largest first,
each circuit on one workgroup or, above a size threshold, on a gang of
workgroups; states of at most 12 qubits stay in LDS.  The launch is captured
once as a hipGraph and replayed.

Gradients are with respect to the rotation angles (RX, RY, RZ) in gate order,
by the adjoint method of ``adjoint.py`` (probes Im<lam|G|psi>, G = X, Y, Z).

``autotune`` picks, per qubit count, the forward-sweep form (fused dense
steps on the matrix cores or gate-by-gate slot steps) and the chunk bits of
streamed circuits, then the gang threshold (how many workgroups share each
large circuit) on the whole batch, and caches the choice in a JSON file so
later runs reproduce it.
"""
import ctypes
import json
import math
import os
import time

import numpy as np

from . import workloads as W
from .backends.custom.native import qgpusim_hip as hip
from .circuit import gate

OP_DTYPE = np.dtype([("kind", "<i4"), ("q0", "<i4"), ("q1", "<i4"), ("mat_offset", "<i4")])
_ROT = ("rx", "ry", "rz")
_FIXED = ("h", "x", "y", "z", "s", "t", "cx")
_GEN = {"rx": np.array([[0, 1], [1, 0]]), "ry": np.array([[0, -1j], [1j, 0]]), "rz": np.array([[1, 0], [0, -1]])}


def _shared_arena():
    """Fixed gates U and U^dag, and the generators, with their offsets."""
    mats, off = [], {}
    pos = 0
    for name in _FIXED:
        U = gate(name).to_matrix()
        for key, M in ((name, U), (name + "^dag", U.conj().T)):
            off[key] = pos
            mats.append(np.asarray(M, dtype=np.complex128).reshape(-1))
            pos += M.size
    for name in _ROT:
        off["G" + name] = pos
        mats.append(np.asarray(_GEN[name], dtype=np.complex128).reshape(-1))
        pos += 4
    return np.concatenate(mats), off


def _rotations(names, thetas):
    """2x2 matrices (Qiskit conventions) for arrays of axes and angles."""
    c, s = np.cos(thetas / 2), np.sin(thetas / 2)
    out = np.zeros((len(thetas), 4), dtype=np.complex128)
    isx, isy, isz = names == 0, names == 1, names == 2
    out[isx] = np.stack([c, -1j * s, -1j * s, c], axis=1)[isx]
    out[isy] = np.stack([c, -s, s, c], axis=1)[isy]
    ez = np.exp(-0.5j * thetas)
    out[isz] = np.stack([ez, 0 * ez, 0 * ez, np.conj(ez)], axis=1)[isz]
    return out


def pack(circuits):
    """circuits: list of (num_qubits, [(name, qubits, angle or None), ...]).
    Returns the flat arrays qsim_batch_load takes, and each circuit's angles."""
    shared, off = _shared_arena()
    base = shared.size
    fwd, rev, fwd_off, rev_off = [], [], [0], [0]
    rot_axis, rot_theta, nprobe, thetas = [], [], [], []
    for n, gates in circuits:
        k0 = len(rot_theta)
        f = []
        for name, qs, th in gates:
            if th is None:
                if name not in off:
                    raise ValueError(f"gate {name!r} is not in the batched gate set")
                f.append((2 if len(qs) == 2 else 1, qs[0], qs[1] if len(qs) == 2 else -1, off[name], name, None))
            else:
                if name not in _ROT:
                    raise ValueError(f"gate {name!r} cannot carry an angle here")
                r = len(rot_theta)
                rot_axis.append(_ROT.index(name))
                rot_theta.append(th)
                f.append((1, qs[0], -1, base + 8 * r, name, r - k0))
        fwd.extend((k, a, b, o) for k, a, b, o, _, _ in f)
        fwd_off.append(len(fwd))
        for k, a, b, o, name, slot in reversed(f):
            if slot is not None:
                rev.append((hip.OP_PROBE, a, slot, off["G" + name]))
                rev.append((1, a, -1, o + 4))
            else:
                rev.append((k, a, b, off[name + "^dag"]))
        rev_off.append(len(rev))
        nprobe.append(len(rot_theta) - k0)
        thetas.append(np.asarray(rot_theta[k0:], dtype=np.float64))
    rot = _rotations(np.asarray(rot_axis, dtype=np.int64), np.asarray(rot_theta, dtype=np.float64))
    rot_dag = rot[:, [0, 2, 1, 3]].conj()
    arena = np.concatenate([shared, np.stack([rot, rot_dag], axis=1).reshape(-1)]).astype(np.complex64)
    return {
        "nq": np.asarray([c[0] for c in circuits], dtype=np.int32),
        "fwd": np.asarray(fwd, dtype=OP_DTYPE), "fwd_off": np.asarray(fwd_off, dtype=np.int64),
        "rev": np.asarray(rev, dtype=OP_DTYPE), "rev_off": np.asarray(rev_off, dtype=np.int64),
        "mats": arena, "nprobe": np.asarray(nprobe, dtype=np.int32), "thetas": thetas,
    }


def random_batch(spec):
    """[(n, depth, seed)] -> circuits of the seeded random-circuit family."""
    return [(n, list(W.random_gates(n, d, s))) for n, d, s in spec]


class BatchSimulator:
    """<Z_mask> and the adjoint gradient of every circuit, one persistent launch.

    ``zmasks``: per-circuit observable masks (default Z on qubit 0).
    ``tuning``: {"K_by_n": {n: K}, "gang_by_n": {n: G}, "fuse_by_n": {n: 0 or 1},
    "blocks_per_cu": int} (see ``autotune``)."""

    def __init__(self, circuits, zmasks=None, tuning=None, packed=None):
        self.packed = packed if packed is not None else pack(circuits)
        p = self.packed
        self.ncirc = len(p["nq"])
        self.zmask = np.asarray(zmasks if zmasks is not None else np.ones(self.ncirc), dtype=np.uint64)
        h = hip._P()
        hip._check(hip.lib.qsim_batch_create(ctypes.byref(h)))
        self._h = h
        if tuning:
            kbn = np.full(41, 12, dtype=np.int32)
            for n, k in tuning.get("K_by_n", {}).items():
                kbn[int(n)] = int(k)
            hip._check(hip.lib.qsim_batch_set_tuning(self._h, kbn.ctypes.data, 41, int(tuning.get("blocks_per_cu", 0))))
            if "gang_by_n" in tuning:
                gbn = np.asarray([default_gang(n) for n in range(41)], dtype=np.int32)
                for n, g in tuning["gang_by_n"].items():
                    gbn[int(n)] = int(g)
                hip._check(hip.lib.qsim_batch_set_gang(self._h, gbn.ctypes.data, 41))
            if "fuse_by_n" in tuning:
                fbn = np.zeros(41, dtype=np.int32)
                for n, f in tuning["fuse_by_n"].items():
                    fbn[int(n)] = int(f)
                hip._check(hip.lib.qsim_batch_set_fusion(self._h, fbn.ctypes.data, 41))
        hip._check(hip.lib.qsim_batch_load(
            self._h, self.ncirc, p["nq"].ctypes.data, self.zmask.ctypes.data, p["fwd"].ctypes.data,
            p["fwd_off"].ctypes.data, p["rev"].ctypes.data, p["rev_off"].ctypes.data, p["mats"].ctypes.data,
            p["mats"].size, p["nprobe"].ctypes.data))
        self.probe_off = np.concatenate([[0], np.cumsum(p["nprobe"])]).astype(np.int64)

    def run(self, use_graph=True):
        """-> (expect[ncirc], [grad per circuit], exec_ms[ncirc], total_ms)."""
        e = np.zeros(self.ncirc, dtype=np.float64)
        g = np.zeros(max(1, int(self.probe_off[-1])), dtype=np.float64)
        lat = np.zeros(self.ncirc, dtype=np.float64)
        tot = ctypes.c_double()
        hip._check(hip.lib.qsim_batch_run(self._h, 1 if use_graph else 0, e.ctypes.data, g.ctypes.data,
                                          lat.ctypes.data, ctypes.addressof(tot)))
        grads = [g[self.probe_off[i]:self.probe_off[i + 1]].copy() for i in range(self.ncirc)]
        return e, grads, lat, tot.value

    def info(self):
        buf = (ctypes.c_int64 * 7)()
        hip.lib.qsim_batch_info(self._h, ctypes.addressof(buf), 7)
        return dict(zip(("circuits", "probes", "passes", "workgroups", "hbm_slot_bytes", "tasks", "gangs"), list(buf)))

    def close(self):
        if getattr(self, "_h", None):
            hip.lib.qsim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_gang(n):
    """The library's default workgroups per circuit: 2^(n-20) above 20 qubits."""
    return min(64, 1 << (n - 20)) if n > 20 else 1


def default_cache_path():
    return os.environ.get("QSIM_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "qgpusim", "batch_autotune.json"))


def gang_policy(threshold, gmax=32):
    """Workgroups per circuit for a gang threshold T: 2^(n - T) above T qubits
    (at most gmax), else 1."""
    return {n: (min(gmax, 1 << (n - threshold)) if n > threshold else 1) for n in range(6, 41)}


def _time_batch(packed, tuning, reps=2):
    """Milliseconds of the launch (best of `reps` after one warm-up run)."""
    sim = BatchSimulator(None, tuning=tuning, packed=packed)
    try:
        sim.run(use_graph=True)
        return min(sim.run(use_graph=True)[3] for _ in range(reps))
    finally:
        sim.close()


def autotune(spec, k_candidates=(11, 12), fuse_candidates=(0, 1), thresholds=(16, 17, 18, 19, 20, 23),
             per_n=48, cache_path=None, log=None):
    """Tuning for the batch `spec` ([(n, depth, seed)]), in two stages.

    1. Per qubit count n present: the forward-sweep form (fused dense steps on
       the matrix cores, or gate-by-gate slot steps) and, for streamed circuits
       (n > 12), the chunk bits K, timed on a sample of up to `per_n` circuits
       of that size run alone (gangs as the library's default).  K is at most
       12: two complex64 chunks of 2^12 amplitudes are what the kernel's LDS
       holds, so larger chunks are not candidates.
    2. Whole batch: the gang threshold T (circuits above T qubits run on
       2^(n - T) workgroups, at most 32; T = 23 runs every circuit of this
       family on one workgroup), timed on the full batch with stage 1's
       choices.

    Results are cached under the workload signature (qubit counts, depths,
    seeds and the candidate sets); a cached entry is returned as is."""
    cache_path = cache_path or default_cache_path()
    key = autotune_key(spec, k_candidates, fuse_candidates, thresholds, per_n)
    cache = {}
    if os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    if key in cache:
        return cache[key]
    ns = sorted({n for n, _, _ in spec})
    k_by_n, fuse_by_n, stage1 = {}, {}, {}
    for n in ns:
        sample = [s for s in spec if s[0] == n][:per_n]
        packed = pack(random_batch(sample))
        best, stage1[n] = None, {}
        for K in (k_candidates if n > 12 else (n,)):
            for fuse in fuse_candidates:
                tun = {"K_by_n": {n: min(K, 12)}, "fuse_by_n": {n: fuse}}
                t = _time_batch(packed, tun)
                stage1[n][f"K{K}_fuse{fuse}"] = round(t, 4)
                if best is None or t < best[2]:
                    best = (K, fuse, t)
        k_by_n[n], fuse_by_n[n] = min(best[0], 12), best[1]
        if log:
            log(f"autotune n={n}: {stage1[n]} -> K={k_by_n[n]}, fused={fuse_by_n[n]}")
    packed = pack(random_batch(spec))
    stage2, best = {}, None
    for T in thresholds:
        tun = {"K_by_n": k_by_n, "fuse_by_n": fuse_by_n, "gang_by_n": gang_policy(T)}
        t = _time_batch(packed, tun, reps=1)
        stage2[T] = round(t, 2)
        if best is None or t < best[1]:
            best = (T, t)
        if log:
            log(f"autotune gang threshold {T}: {t:.1f} ms")
    entry = {"K_by_n": {str(n): k for n, k in k_by_n.items()},
             "fuse_by_n": {str(n): f for n, f in fuse_by_n.items()},
             "gang_threshold": best[0],
             "gang_by_n": {str(n): g for n, g in gang_policy(best[0]).items() if n in ns},
             "blocks_per_cu": 0,
             "timings_ms": {"per_n": {str(n): t for n, t in stage1.items()},
                            "gang_threshold": {str(T): t for T, t in stage2.items()}}}
    cache[key] = entry
    os.makedirs(os.path.dirname(cache_path) or ".", exist_ok=True)
    with open(cache_path, "w") as f:
        json.dump(cache, f, indent=1)
    return entry


def autotune_key(spec, k_candidates=(11, 12), fuse_candidates=(0, 1), thresholds=(16, 17, 18, 19, 20, 23),
                 per_n=48):
    """The cache key of an autotune call: the workload's digest and the
    candidate sets."""
    sig = {"spec_sha": _spec_digest(spec), "k": list(k_candidates), "fuse": list(fuse_candidates),
           "thresholds": list(thresholds), "per_n": per_n}
    return json.dumps(sig, sort_keys=True)


def _spec_digest(spec):
    import hashlib
    h = hashlib.sha256()
    for n, d, seed in spec:
        h.update(f"{n},{d},{seed};".encode())
    return h.hexdigest()[:16]
