"""One measured run of a circuit (qgpusim/runner.py:12-178).

``run_once`` keeps the reference's signature and return value
``(transpile_s, simulate_s, extra, stats, res_usage)`` and its stats keys.  It
runs the path this package implements: ``transpiler="custom"`` with
``backend="custom"`` (the save_statevector marker, the LQR pass, then
``run_custom_backend`` in graphed mode, complex64, fusion on).  In place of
the Qiskit ``transpile()`` call before the LQR pass (runner.py:45) the circuit
goes through ``merge_1q_runs``: each maximal run of single-qubit gates on a
qubit becomes one gate, which is what that call does to the reference's
circuits according to its published rows (gate counts, depth and tiles are
then the same); Qiskit is not a dependency here.
The reference's Aer paths (``transpiler="baseline"`` or ``backend="aer"``)
need qiskit-aer and raise a RuntimeError saying so.

``task="sampling"`` with ``measure=True`` adds ``measure_all`` and returns
``extra["counts"]`` sampled on the GPU from the final state (the reference's
custom backend has no sampling path).
"""
import time

from .backends.custom.backend import run_custom_backend
from .metrics.resources import _max_usage, _mem_snapshot, _tqc_stats
from .transpiler.canonicalize import merge_1q_runs
from .transpiler.pm import make_custom_pm


def run_once(sim, qc, task: str, shots: int, measure: bool, transpiler: str = "baseline",
             num_local_qubits: int = 4, backend: str = "aer", device_used: str = "CPU",
             save_sv: bool = True, sample_seed: int = 0):
    """Transpile, simulate and measure one circuit; see the module docstring.
    Returns (transpile_s, simulate_s, extra, stats, res_usage)."""
    if transpiler != "custom" or backend != "custom":
        raise RuntimeError("run_once: the Aer paths (transpiler='baseline' or backend='aer') need qiskit-aer, "
                           "which this package does not use; pass transpiler='custom', backend='custom'")
    if task not in ("statevector", "sampling"):
        raise ValueError(f"run_once: unknown task {task!r}")
    qc_run = qc
    if task == "sampling" and measure:
        qc_run = qc.copy()
        qc_run.measure_all()
    res_peak = {"rss_mb": "", "gpu_mem_mb": "", "gpu_util": ""}

    t0 = time.perf_counter()
    qc_sv = qc_run.copy()
    if task == "statevector" and save_sv:
        qc_sv.save_statevector()
    if task == "statevector":
        qc_sv = merge_1q_runs(qc_sv)  # runner.py:45 transpiles the statevector task only
    pm, lqr_pass = make_custom_pm(num_local_qubits)
    tqc = pm.run(qc_sv)
    tiling_plan = lqr_pass.property_set.get("flat_tiling_plan", None)
    transpile_s = time.perf_counter() - t0
    if tiling_plan is None:
        raise ValueError("backend='custom' requires transpiler='custom' and tiling_plan")
    res_peak = _max_usage(res_peak, _mem_snapshot())

    t0 = time.perf_counter()
    psi, custom_metrics = run_custom_backend(tqc, tiling_plan, task, shots, tile_mode="graphed",
                                             use_float32=True, enable_fusion=True, sample_seed=sample_seed)
    extra = {}
    if task == "statevector":
        extra["statevector"] = psi
    elif "counts" in custom_metrics:
        extra["counts"] = custom_metrics.pop("counts")
    simulate_s = time.perf_counter() - t0
    res_peak = _max_usage(res_peak, _mem_snapshot())

    stats = _tqc_stats(tqc)
    stats.update({
        "num_tiles": custom_metrics.get("num_tiles", ""),
        "num_gates_custom": custom_metrics.get("num_gates", ""),
        "num_fused_ops": custom_metrics.get("num_fused_ops", ""),
        "fusion_ratio": f"{custom_metrics.get('fusion_ratio', 1.0):.2f}",
        "num_reorders": custom_metrics.get("num_reorders", ""),
        "num_kernel_launches": custom_metrics.get("num_kernel_launches", ""),
        "num_graph_replays": custom_metrics.get("num_graph_replays", ""),
        "avoided_launches": custom_metrics.get("avoided_launches", ""),
        "t_perm_s": f"{custom_metrics.get('t_perm_s', 0.0):.6f}",
        "t_apply_s": f"{custom_metrics.get('t_apply_s', 0.0):.6f}",
        "t_sync_s": f"{custom_metrics.get('t_sync_s', 0.0):.6f}",
        "t_launch_overhead_saved_us": f"{custom_metrics.get('t_launch_overhead_saved_us', 0.0):.1f}",
        "tile_mode": custom_metrics.get("tile_mode", ""),
    })
    return transpile_s, simulate_s, extra, stats, res_peak
