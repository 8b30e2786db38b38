from .backend import run_custom_backend, fuse_tile_ops

__all__ = ["run_custom_backend", "fuse_tile_ops"]
