"""ctypes wrapper around the oracle's plain-C simulator.  Test infrastructure."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsvref.so")
_lib = None


class _Op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q0", ctypes.c_int32),
                ("q1", ctypes.c_int32), ("mat_offset", ctypes.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for suf in ("d", "f"):
            getattr(_lib, f"svref_apply_1q_{suf}").argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
            getattr(_lib, f"svref_apply_2q_{suf}").argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
            getattr(_lib, f"svref_permute_{suf}").argtypes = [
                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            getattr(_lib, f"svref_run_{suf}").argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            getattr(_lib, f"svref_run_{suf}").restype = ctypes.c_int
            getattr(_lib, f"svref_run_tiled_{suf}").argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
            getattr(_lib, f"svref_run_tiled_{suf}").restype = ctypes.c_int
    return _lib


def _suffix(psi):
    if psi.dtype == np.complex128:
        return "d"
    if psi.dtype == np.complex64:
        return "f"
    raise TypeError(psi.dtype)


def pack(ops, dtype):
    """ops: [(matrix, qargs)] -> (ctypes op array, matrix arena)."""
    arr = (_Op * max(1, len(ops)))()
    mats = []
    off = 0
    for i, (U, qargs) in enumerate(ops):
        U = np.ascontiguousarray(U, dtype=dtype).reshape(-1)
        arr[i].kind = len(qargs)
        arr[i].q0 = qargs[0]
        arr[i].q1 = qargs[1] if len(qargs) > 1 else -1
        arr[i].mat_offset = off
        mats.append(U)
        off += U.size
    arena = np.concatenate(mats) if mats else np.zeros(1, dtype=dtype)
    return arr, np.ascontiguousarray(arena, dtype=dtype)


def run(psi, n, ops):
    """Apply ``[(matrix, qargs)]`` in place (1q/2q only)."""
    suf = _suffix(psi)
    arr, arena = pack(ops, psi.dtype)
    r = getattr(lib(), f"svref_run_{suf}")(psi.ctypes.data, n, ctypes.addressof(arr),
                                            len(ops), arena.ctypes.data)
    if r < 0:
        raise ValueError("svref: op with more than two qubits")
    return psi


def run_tiled(psi, n, ops, block_bits):
    """Apply ``[(matrix, qargs)]`` in place, every qarg below ``block_bits``:
    blocks of 2^block_bits contiguous amplitudes each run the whole list."""
    suf = _suffix(psi)
    arr, arena = pack(ops, psi.dtype)
    r = getattr(lib(), f"svref_run_tiled_{suf}")(psi.ctypes.data, n, ctypes.addressof(arr), len(ops),
                                                  arena.ctypes.data, block_bits)
    if r < 0:
        raise ValueError("svref: an op reaches a qubit outside the block, or has more than two qubits")
    return psi


def permute(psi, n, bitmap):
    suf = _suffix(psi)
    m = np.ascontiguousarray(bitmap, dtype=np.int32)
    out = np.empty_like(psi)
    getattr(lib(), f"svref_permute_{suf}")(psi.ctypes.data, out.ctypes.data, n, m.ctypes.data)
    return out


def num_threads():
    return lib().svref_num_threads()
