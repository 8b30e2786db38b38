"""Correctness and resource metrics of a run (qgpusim/metrics/)."""
from .correctness import _stable_vdot, statevector_overlap, total_variation_distance  # noqa: F401
from .resources import _max_usage, _mem_snapshot, _tqc_stats  # noqa: F401
