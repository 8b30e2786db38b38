"""Circuit statistics and resource snapshots (qgpusim/metrics/resources.py).

Same keys and string formats as the reference, so result rows line up.  The
reference reads this process's VRAM and the device's utilisation through NVML
(resources.py:39-86).  On ROCm the same figures come from the kernel driver:
the process's VRAM from the KFD's per-process counters
(/sys/class/kfd/kfd/proc/<pid>/vram_<gpu id>, bytes, summed over GPUs) and
the utilisation from the first GPU's gpu_busy_percent in sysfs.  Without the
KFD counters the device-wide memory in use (hipMemGetInfo) stands in; without
the busy counter gpu_util stays empty, as the reference leaves it without NVML.
"""
import glob
import os
from typing import Dict

_NON_GATE = {"barrier", "measure", "reset", "save_statevector", "snapshot"}


def _tqc_stats(tqc) -> Dict[str, int]:
    """Gate counts and depth of a circuit (resources.py:21-36)."""
    ops = tqc.count_ops()
    twoq = int(ops.get("cx", 0) + ops.get("cz", 0) + ops.get("swap", 0))
    return {
        "nqubits_t": tqc.num_qubits,
        "depth_t": tqc.depth(),
        "cx_count": int(ops.get("cx", 0)),
        "cz_count": int(ops.get("cz", 0)),
        "swap_count": int(ops.get("swap", 0)),
        "rz_count": int(ops.get("rz", 0)),
        "rx_count": int(ops.get("rx", 0)),
        "twoq_count": twoq,
        "total_gates": sum(v for k, v in ops.items() if k not in _NON_GATE),
    }


def _process_vram_bytes():
    """VRAM held by this process on every GPU (KFD counters), or None."""
    files = glob.glob(f"/sys/class/kfd/kfd/proc/{os.getpid()}/vram_*")
    if not files:
        return None
    total = 0
    for path in files:
        with open(path) as f:
            total += int(f.read().strip() or 0)
    return total


def _gpu_busy_percent():
    """Instantaneous utilisation of the first GPU (sysfs), or None."""
    for path in sorted(glob.glob("/sys/class/drm/card*/device/gpu_busy_percent")):
        try:
            with open(path) as f:
                return int(f.read().strip())
        except (OSError, ValueError):
            continue
    return None


def _mem_snapshot() -> Dict[str, str]:
    """Process RSS, this process's VRAM (MiB) and GPU utilisation (%)."""
    out = {"rss_mb": "", "gpu_mem_mb": "", "gpu_util": ""}
    try:
        import psutil
        out["rss_mb"] = f"{psutil.Process(os.getpid()).memory_info().rss / (1024 ** 2):.2f}"
    except Exception:
        pass
    try:
        used = _process_vram_bytes()
        if used is None:
            from ..backends.custom.native import qgpusim_hip as hip
            free, total = hip.device_memory()
            used = total - free
        out["gpu_mem_mb"] = f"{used / (1024 ** 2):.2f}"
    except Exception:
        pass
    busy = _gpu_busy_percent()
    if busy is not None:
        out["gpu_util"] = str(busy)
    return out


# CSV form of each usage field: memory in MiB with two decimals, utilisation
# as a whole percentage (fraction dropped).
_USAGE_FORMAT = {"rss_mb": lambda v: f"{v:.2f}", "gpu_mem_mb": lambda v: f"{v:.2f}",
                 "gpu_util": lambda v: str(int(v))}


def _number(text):
    """The field's value as a float, or None when it holds no number."""
    try:
        return float(text)
    except (TypeError, ValueError):
        return None


def _max_usage(a, b):
    """Peak of two usage snapshots, field by field.

    Rule: when both fields hold numbers the result is the larger, written in
    the field's CSV format; when only one does, that side's text is kept as it
    is; a field that no number fills on the first side takes the second
    side's text (or stays as it was when the second side lacks the field).
    A missing first snapshot counts as all fields empty."""
    merged = {k: "" for k in _USAGE_FORMAT}
    if a:
        merged.update(a)
    for key, fmt in _USAGE_FORMAT.items():
        mine, theirs = merged.get(key, ""), b.get(key, merged.get(key, ""))
        x, y = _number(mine), _number(theirs)
        if x is None:
            merged[key] = theirs
        elif y is not None:
            merged[key] = fmt(max(x, y))
    return merged
