from .canonicalize import merge_1q_runs
from .passes import LazyQubitReordering, Tile
from .pm import PassManager, make_custom_pm

__all__ = ["LazyQubitReordering", "Tile", "PassManager", "make_custom_pm", "merge_1q_runs"]
