"""Circuit statistics and resource snapshots (qgpusim/metrics/resources.py).

Same keys and string formats as the reference, so result rows line up.  The
reference reads GPU memory and utilisation through NVML; here device memory
in use comes from hipMemGetInfo (device-wide, in MiB) and utilisation is not
sampled (left empty, as the reference leaves it when NVML is missing).
"""
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


def _mem_snapshot() -> Dict[str, str]:
    """Process RSS and device memory in use, in MiB (resources.py:39-86)."""
    out = {"rss_mb": "", "gpu_mem_mb": "", "gpu_util": ""}
    try:
        import psutil
        out["rss_mb"] = f"{psutil.Process(os.getpid()).memory_info().rss / (1024 ** 2):.2f}"
    except Exception:
        pass
    try:
        from ..backends.custom.native import qgpusim_hip as hip
        free, total = hip.device_memory()
        out["gpu_mem_mb"] = f"{(total - free) / (1024 ** 2):.2f}"
    except Exception:
        pass
    return out


def _to_float(x):
    try:
        return float(x)
    except Exception:
        return None


def _max_usage(a, b):
    """Field-wise maximum of two snapshots, kept as CSV-style strings
    (resources.py:96-127)."""
    out = dict(a) if a else {"rss_mb": "", "gpu_mem_mb": "", "gpu_util": ""}
    for k in ("rss_mb", "gpu_mem_mb"):
        av, bv = _to_float(out.get(k, "")), _to_float(b.get(k, ""))
        if av is None:
            out[k] = b.get(k, out.get(k, ""))
        elif bv is not None:
            out[k] = f"{max(av, bv):.2f}"
    av, bv = _to_float(out.get("gpu_util", "")), _to_float(b.get("gpu_util", ""))
    if av is None:
        out["gpu_util"] = b.get("gpu_util", out.get("gpu_util", ""))
    elif bv is not None:
        out["gpu_util"] = str(int(max(av, bv)))
    return out
