"""Result rows in the reference's CSV schema (benches/run_benchmarks.py:551-566,
benches/custom_logging.py), so its plotting reads this package's results
unchanged.

``FIELDNAMES`` is the 49-column header in the reference's order.
``result_row`` builds one candidate row from ``run_once``'s outputs with the
reference's formats: times "%.6f"; overlap "%.12f", l2 "%.3e" and
passed = overlap > 1 - 1e-3 for statevectors; tvd and kl_div "%.6f" and
passed = tvd < 0.02 for counts (run_benchmarks.py:256-303).  ``log_run``
runs one circuit and appends its row.
"""
import csv
import datetime
import os
import platform
from typing import Dict, Optional

import numpy as np

from .metrics.correctness import statevector_overlap, total_variation_distance

FIELDNAMES = [
    "timestamp", "host", "os", "python",
    "mode", "device", "task", "circuit", "nqubits", "depth", "shots", "repeat_idx",
    "transpile_s", "simulate_s", "total_s",
    "nqubits_t", "depth_t", "total_gates", "twoq_count", "cx_count", "cz_count", "swap_count", "rz_count", "rx_count",
    "rss_mb", "gpu_mem_mb", "gpu_util",
    "overlap", "l2", "tvd", "kl_div", "passed",
    "num_tiles", "num_gates_custom", "num_fused_ops", "fusion_ratio", "num_reorders",
    "num_kernel_launches", "num_graph_replays", "avoided_launches",
    "t_perm_s", "t_apply_s", "t_sync_s", "t_launch_overhead_saved_us", "tile_mode",
    "notes", "transpiler", "nL", "backend",
]

_STATS = ["nqubits_t", "depth_t", "total_gates", "twoq_count", "cx_count", "cz_count", "swap_count", "rz_count",
          "rx_count"]
_USAGE = ["rss_mb", "gpu_mem_mb", "gpu_util"]
_KERNEL = ["num_tiles", "num_gates_custom", "num_fused_ops", "fusion_ratio", "num_reorders", "num_kernel_launches",
           "num_graph_replays", "avoided_launches", "t_perm_s", "t_apply_s", "t_sync_s",
           "t_launch_overhead_saved_us", "tile_mode"]


def env_info() -> Dict[str, str]:
    return {"host": platform.node(), "os": platform.platform(), "python": platform.python_version()}


def ensure_csv(path: str) -> None:
    """Create the file with the header if it does not exist."""
    if os.path.dirname(path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
    if not os.path.exists(path):
        with open(path, "w", newline="") as f:
            csv.DictWriter(f, fieldnames=FIELDNAMES).writeheader()


def append_csv(path: str, row: Dict[str, object]) -> None:
    with open(path, "a", newline="") as f:
        csv.DictWriter(f, fieldnames=FIELDNAMES).writerow({k: v for k, v in row.items() if k in FIELDNAMES})


def kl_div(counts_p: Dict[str, int], counts_q: Dict[str, int], eps: float = 1e-12) -> float:
    """KL(p || q) of two count dictionaries with eps smoothing
    (run_benchmarks.py:112-122)."""
    keys = set(counts_p) | set(counts_q)
    sp = sum(counts_p.values()) or 1
    sq = sum(counts_q.values()) or 1
    k = len(keys) or 1
    kl = 0.0
    for key in keys:
        p = (counts_p.get(key, 0) + eps) / (sp + eps * k)
        q = (counts_q.get(key, 0) + eps) / (sq + eps * k)
        kl += p * np.log(p / q)
    return float(kl)


def result_row(combo: Dict[str, object], t_trans: float, t_sim: float, stats: Dict[str, object],
               res_usage: Dict[str, object], mode: str = "gpu", device: str = "GPU", repeat_idx: int = 1,
               overlap_l2=None, tvd_kl=None, notes: str = "", env: Optional[Dict[str, str]] = None):
    """One candidate row.  combo: task ("statevector" | "sampling"), circuit,
    nqubits, depth, shots, transpiler, nL, backend.  overlap_l2 = (overlap,
    l2) or tvd_kl = (tvd, kl or None) fill the correctness columns."""
    env = env or env_info()
    row = {
        "timestamp": datetime.datetime.now().isoformat(timespec="seconds"),
        "host": env.get("host", ""), "os": env.get("os", ""), "python": env.get("python", ""),
        "mode": mode, "device": device, "task": combo["task"], "circuit": combo["circuit"],
        "nqubits": combo["nqubits"], "depth": combo["depth"],
        "shots": combo.get("shots", 0) if combo["task"] == "sampling" else 0,
        "repeat_idx": repeat_idx,
        "transpile_s": f"{t_trans:.6f}", "simulate_s": f"{t_sim:.6f}", "total_s": f"{t_trans + t_sim:.6f}",
        "notes": notes, "transpiler": combo.get("transpiler", "custom"), "nL": combo.get("nL", ""),
        "backend": combo.get("backend", "custom"),
    }
    for k in _STATS + _KERNEL:
        row[k] = (stats or {}).get(k, "")
    for k in _USAGE:
        row[k] = (res_usage or {}).get(k, "")
    if overlap_l2 is not None:
        ov, l2 = overlap_l2
        passed = ov > 1 - 1e-3
        row.update({"overlap": f"{ov:.12f}", "l2": f"{l2:.3e}", "passed": str(passed)})
        if not notes:
            row["notes"] = f"overlap={ov:.12f}; l2={l2:.3e}; pass={passed}"
    elif tvd_kl is not None:
        tvd, kl = tvd_kl
        passed = tvd < 0.02
        row.update({"tvd": f"{tvd:.6f}", "passed": str(passed)})
        if kl is not None:
            row["kl_div"] = f"{kl:.6f}"
        if not notes:
            row["notes"] = f"TVD={tvd:.6f}; " + (f"KL={kl:.6f}; " if kl is not None else "") + f"pass={passed}"
    return row


def log_run(csv_path: str, qc, combo: Dict[str, object], reference=None, repeat_idx: int = 1, sample_seed: int = 0):
    """Run ``qc`` with ``runner.run_once`` (custom transpiler and backend) and
    append its row.  ``reference``: a statevector (statevector task) or a
    count dictionary (sampling task) to check the result against, or None."""
    from .runner import run_once
    task = combo["task"]
    t_tr, t_sim, extra, stats, usage = run_once(None, qc, task, combo.get("shots"), task == "sampling",
                                                transpiler="custom", num_local_qubits=combo["nL"],
                                                backend="custom", device_used="GPU", sample_seed=sample_seed)
    ov_l2 = tvd_kl = None
    if reference is not None:
        if task == "statevector":
            ov_l2 = statevector_overlap(reference, extra["statevector"])
        else:
            tvd_kl = (total_variation_distance(reference, extra["counts"]), kl_div(reference, extra["counts"]))
    ensure_csv(csv_path)
    row = result_row(combo, t_tr, t_sim, stats, usage, repeat_idx=repeat_idx, overlap_l2=ov_l2, tvd_kl=tvd_kl)
    append_csv(csv_path, row)
    return row, extra
