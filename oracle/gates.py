"""Closed-form gate matrices (FP64), Qiskit conventions.  Test infrastructure.

Qiskit's ``Operator``/``to_matrix`` convention is little-endian: for a gate on
qargs (q0, q1) the matrix row/column index is ``b_q0 + 2*b_q1``
(qgpusim_cuda/src/kernels.cu:54-55, 108-114).  So CX(control=q0, target=q1)
maps index 1 (q0=1, q1=0) to index 3.
"""
import numpy as np

_S2 = 1.0 / np.sqrt(2.0)


def h():
    return np.array([[_S2, _S2], [_S2, -_S2]], dtype=np.complex128)


def x():
    return np.array([[0, 1], [1, 0]], dtype=np.complex128)


def y():
    return np.array([[0, -1j], [1j, 0]], dtype=np.complex128)


def z():
    return np.array([[1, 0], [0, -1]], dtype=np.complex128)


def s():
    return np.array([[1, 0], [0, 1j]], dtype=np.complex128)


def t():
    return np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=np.complex128)


def rx(theta):
    c, s_ = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -1j * s_], [-1j * s_, c]], dtype=np.complex128)


def ry(theta):
    c, s_ = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -s_], [s_, c]], dtype=np.complex128)


def rz(theta):
    return np.array([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]],
                    dtype=np.complex128)


def p(lam):
    return np.array([[1, 0], [0, np.exp(1j * lam)]], dtype=np.complex128)


def cx():
    # control = qargs[0] (bit 0 of the matrix index), target = qargs[1]
    m = np.zeros((4, 4), dtype=np.complex128)
    for col in range(4):
        c0, c1 = col & 1, (col >> 1) & 1
        r1 = c1 ^ c0
        m[c0 | (r1 << 1), col] = 1
    return m


def cz():
    return np.diag([1, 1, 1, -1]).astype(np.complex128)


def cp(lam):
    return np.diag([1, 1, 1, np.exp(1j * lam)]).astype(np.complex128)


def swap():
    m = np.zeros((4, 4), dtype=np.complex128)
    for col in range(4):
        m[((col & 1) << 1) | (col >> 1), col] = 1
    return m


ONE_Q = {"h": h, "x": x, "y": y, "z": z, "s": s, "t": t}
ROT = {"rx": rx, "ry": ry, "rz": rz, "p": p}
TWO_Q = {"cx": cx, "cz": cz, "swap": swap}


def matrix(name, params=()):
    """Matrix of gate ``name`` with parameters ``params``."""
    if name in ONE_Q:
        return ONE_Q[name]()
    if name in ROT:
        return ROT[name](*params)
    if name in TWO_Q:
        return TWO_Q[name]()
    if name == "cp":
        return cp(*params)
    raise KeyError(name)
