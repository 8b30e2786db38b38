"""Adjoint differentiation of <psi(theta)|H|psi(theta)> on the GPU.

Forward: the parameterised circuit runs as LDS chunk passes on one state;
<H> is reduced in FP64 over Pauli chunk passes.

Backward (Jones & Gacon, arXiv:2009.02823): lam = H psi, then for the gates in
reverse order, each parameterised gate U_k contributes
    dE/dtheta_k = 2 Re <lam_k| (dU_k/dtheta) U_k^dag |psi_k> = Im <lam_k| G_k |psi_k>,
with G = 2i (dU/dtheta) U^dag (G = X, Y, Z for RX, RY, RZ; diag(0, -2) for
P), after which U_k^dag is applied to both psi and lam.  Nothing is stored
per gate: psi is un-computed on the way back, so memory stays at two
statevectors.  The whole sweep runs in the same chunk passes as the forward
(two vectors per chunk), and each probe is reduced per block in FP64.

``AdjointExpectation`` wraps this in a ``torch.autograd.Function``.
"""
import ctypes

import numpy as np

from .backends.custom.native import qgpusim_hip as hip
from .circuit import gate

_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)

# G = 2i (dU/dtheta) U^dag for gates with one angle
GENERATORS = {
    "rx": _X,
    "ry": _Y,
    "rz": _Z,
    "p": np.diag([0.0, -2.0]).astype(np.complex128),
}


def _matrix(name, theta, fixed):
    params = ([theta] if theta is not None else []) + list(fixed)
    return gate(name, *params).to_matrix()


class _OpBuffer:
    """ctypes op array + matrix arena built op by op."""

    def __init__(self, nops, dtype):
        self.arr = (hip.QsimOp * max(1, nops))()
        self.n = 0
        self.mats = []
        self.off = 0
        self.dtype = dtype

    def add(self, kind, q0, q1, m):
        o = self.arr[self.n]
        o.kind, o.q0, o.q1, o.mat_offset = kind, q0, q1, self.off
        flat = np.ascontiguousarray(m, dtype=self.dtype).reshape(-1)
        self.mats.append(flat)
        self.off += flat.size
        self.n += 1

    def arena(self):
        if not self.mats:
            return np.zeros(16, dtype=self.dtype)
        return np.ascontiguousarray(np.concatenate(self.mats), dtype=self.dtype)


def _norm_spec(spec):
    return [tuple(s) + ((),) if len(s) == 3 else tuple(s) for s in spec]


def forward_ops(spec, theta, dtype):
    """Packed op list of the circuit at parameters theta."""
    spec = _norm_spec(spec)
    buf = _OpBuffer(len(spec), dtype)
    for name, qs, k, fixed in spec:
        m = _matrix(name, None if k is None else float(theta[k]), fixed)
        if len(qs) == 1:
            buf.add(hip.OP_1Q, qs[0], -1, m)
        else:
            buf.add(hip.OP_2Q, qs[0], qs[1], m)
    return buf


def reverse_ops(spec, theta, dtype):
    """Packed reverse sweep: inverse gates in reverse order, each parameterised
    gate preceded by a probe of its generator in the parameter's slot."""
    spec = _norm_spec(spec)
    nprobe = sum(1 for s in spec if s[2] is not None)
    buf = _OpBuffer(len(spec) + nprobe, dtype)
    slot = nprobe
    for name, qs, k, fixed in reversed(spec):
        m = _matrix(name, None if k is None else float(theta[k]), fixed)
        if k is not None:
            slot -= 1
            buf.add(hip.OP_PROBE, qs[0], slot, GENERATORS[name])
        mdag = np.conj(m).T
        if len(qs) == 1:
            buf.add(hip.OP_1Q, qs[0], -1, mdag)
        else:
            buf.add(hip.OP_2Q, qs[0], qs[1], mdag)
    return buf, nprobe


class AdjointSimulator:
    """Expectation and adjoint gradient of a parameterised circuit.

    ``spec``: list of ``(name, qubits, param_index or None[, fixed_params])``;
    ``observable``: list of ``(xmask, zmask, coef)`` Pauli terms.
    """

    def __init__(self, num_qubits, spec, observable, use_float32=True):
        self.n = num_qubits
        self.spec = [tuple(s) + ((),) if len(s) == 3 else tuple(s) for s in spec]
        for name, qs, k, _ in self.spec:
            if k is not None and name not in GENERATORS:
                raise ValueError(f"gate {name!r} cannot carry a parameter")
            if len(qs) not in (1, 2):
                raise ValueError("only 1- and 2-qubit gates are supported")
        self.observable = list(observable)
        self.dtype = np.complex64 if use_float32 else np.complex128
        self.sv = hip.StatevectorF32(num_qubits) if use_float32 else hip.Statevector(num_qubits)
        self.nparams = 1 + max([k for _, _, k, _ in self.spec if k is not None], default=-1)
        self._probe_param = [k for _, _, k, _ in self.spec if k is not None]
        # Bumped whenever the state changes: a caller that saw generation g
        # after prepare() knows the state still holds that psi(theta) while
        # the generation is g.
        self.generation = 0
        self._theta = None

    def _forward_ops(self, theta):
        return forward_ops(self.spec, theta, self.dtype)

    def _reverse_ops(self, theta):
        return reverse_ops(self.spec, theta, self.dtype)

    def prepare(self, theta, psi0=None):
        """Run the circuit: the state holds psi(theta) afterwards."""
        theta = np.asarray(theta, dtype=np.float64)
        if psi0 is None:
            self.sv.reset()
        else:
            self.sv.from_numpy(psi0)
        buf = self._forward_ops(theta)
        self.generation += 1
        self.sv.apply_packed(hip.COOPERATIVE, buf.arr, buf.n, buf.arena())
        self._theta = theta

    def expectation(self):
        return self.sv.expectation(self.observable).real

    def forward(self, theta, psi0=None):
        """<psi(theta)|H|psi(theta)> (FP64 reduction)."""
        self.prepare(theta, psi0)
        return self.expectation()

    def gradient_from_state(self):
        """Adjoint sweep from the state left by prepare()/forward().
        Returns (E, grad); the state is un-computed back to the input."""
        if self._theta is None:
            raise RuntimeError("gradient_from_state: no prepared state (call prepare or forward first)")
        buf, nprobe = self._reverse_ops(self._theta)
        self.generation += 1  # the sweep un-computes the state
        probes, e = self.sv.adjoint(self.observable, buf.arr, buf.n, buf.arena(), nprobe)
        self._theta = None
        grad = np.zeros(self.nparams, dtype=np.float64)
        np.add.at(grad, np.asarray(self._probe_param, dtype=np.int64), probes)
        return e, grad

    def value_and_grad(self, theta, psi0=None):
        self.prepare(theta, psi0)
        return self.gradient_from_state()


def _make_function():
    import torch

    class AdjointExpectation(torch.autograd.Function):
        """E(theta) = <psi(theta)|H|psi(theta)> with the adjoint backward."""

        @staticmethod
        def forward(ctx, theta, sim):
            th = theta.detach().cpu().double().numpy().copy()
            e = sim.forward(th)
            ctx.sim = sim
            ctx.theta = th
            ctx.generation = sim.generation
            return torch.tensor(e, dtype=torch.float64, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            sim = ctx.sim
            if sim.generation == ctx.generation:
                # the state is still this forward's psi(theta): sweep it
                _, g = sim.gradient_from_state()
            else:
                # another forward (or a sweep) ran on the simulator since:
                # prepare this forward's state again
                _, g = sim.value_and_grad(ctx.theta)
            gt = torch.as_tensor(g, dtype=torch.float64, device=grad_out.device)
            return grad_out * gt, None

    return AdjointExpectation


_FN = None


def adjoint_expectation(theta, sim):
    """torch: E(theta) differentiable through the adjoint kernel."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(theta, sim)
