"""Gradients of parametrised circuits
(utils/cpu/cpuVariationalQuantumCircuitGradients.jl), with the adjoint
method on the device replacing the reference's Enzyme default.

`VQEEnergy` binds an ansatz tape, a Pauli-sum Hamiltonian and a device
state; it evaluates E(θ) (optionally gate by gate with the one-gate kernels,
the "plain kernels" parameter shift is timed with) and the adjoint gradient
for one parameter vector or a batch of them.  The finite-difference,
parameter-shift and verification helpers take any cost function, as the
reference's do.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import call
from .circuits import Tape
from .state import GPUState


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class VQEEnergy:
    def __init__(self, tape: Tape, xmask, zmask, coeff, batch: int = 1, dtype="c128", ctx=None):
        self.tape = tape
        self.n_params = tape.n_params
        self.ops = tape.to_ctypes()
        self.xm = np.ascontiguousarray(xmask, np.uint64)
        self.zm = np.ascontiguousarray(zmask, np.uint64)
        self.c = np.ascontiguousarray(coeff, np.float64)
        self.psi = GPUState(tape.n, dtype, batch, ctx)
        self.batch = batch

    def _theta(self, theta):
        th = np.ascontiguousarray(np.asarray(theta, np.float64).reshape(self.batch, self.n_params))
        return th

    def energy(self, theta, one_gate_kernels: bool = False) -> np.ndarray:
        th = self._theta(theta)
        out = np.empty(self.batch, np.float64)
        call("smoq_tape_energy", self.psi.handle, self.ops, len(self.tape.ops), _p(th),
             self.n_params, self.xm.size, _p(self.xm), _p(self.zm), _p(self.c),
             1 if one_gate_kernels else 0, _p(out))
        return out

    def energy_and_grad(self, theta):
        """Adjoint gradient: (E[batch], grad[batch, n_params])."""
        th = self._theta(theta)
        e = np.empty(self.batch, np.float64)
        g = np.empty((self.batch, self.n_params), np.float64)
        call("smoq_adjoint_gradient", self.psi.handle, self.ops, len(self.tape.ops), _p(th),
             self.n_params, self.xm.size, _p(self.xm), _p(self.zm), _p(self.c), _p(e), _p(g))
        return e, g

    def cost_fn(self, one_gate_kernels: bool = False):
        """θ -> E(θ) for a single parameter vector (batch 1)."""
        if self.batch != 1:
            raise ValueError("cost_fn needs batch 1")
        return lambda th: float(self.energy(th, one_gate_kernels)[0])

    def free(self):
        self.psi.free()


def adjoint_gradient_batch(items, dtype="c128", ctx=None):
    """Adjoint gradients of several parameter sets, each with its own qubit
    count, in one persistent launch (smoq_adjoint_gradient_batch).

    items: sequence of (tape, theta, xmask, zmask, coeff) — the ansatz tape
    (its own n), one parameter vector and H = Σ coeff_t P_t.  Returns
    (energies[n_items], [grad_i]) with grad_i of length tape_i.n_params."""
    from ._lib import smoq_vqe_item
    from .state import _DTYPES, default_context

    ctx = ctx or default_context()
    items = list(items)
    arr = (smoq_vqe_item * max(1, len(items)))()
    keep = []
    energies = np.zeros(len(items), np.float64)
    grads = []
    for i, (tape, theta, xm, zm, c) in enumerate(items):
        th = np.ascontiguousarray(theta, np.float64).reshape(-1)
        if th.size != tape.n_params:
            raise ValueError(f"item {i}: {th.size} angles for {tape.n_params} parameters")
        xm = np.ascontiguousarray(xm, np.uint64)
        zm = np.ascontiguousarray(zm, np.uint64)
        c = np.ascontiguousarray(c, np.float64)
        g = np.zeros(max(1, tape.n_params), np.float64)
        ops = tape.to_ctypes()
        keep += [th, xm, zm, c, g, ops]
        grads.append(g[:tape.n_params])
        r = arr[i]
        r.n_qubits = tape.n
        r.n_ops = len(tape.ops)
        r.ops = ctypes.addressof(ops)
        r.n_params = tape.n_params
        r.n_terms = c.size
        r.theta = th.ctypes.data
        r.xmask = xm.ctypes.data
        r.zmask = zm.ctypes.data
        r.coeff = c.ctypes.data
        r.energy = energies.ctypes.data + 8 * i
        r.grad = g.ctypes.data
    call("smoq_adjoint_gradient_batch", ctx.handle, _DTYPES[dtype], arr, len(items))
    return energies, grads


def calculate_gradients_parameter_shift(cost_fn, theta, shift: float = math.pi / 2):
    """calculate_gradients_parameter_shift (…Gradients.jl:138-163)."""
    theta = np.asarray(theta, np.float64)
    grad = np.zeros_like(theta)
    tp, tm = theta.copy(), theta.copy()
    for k in range(theta.size):
        tp[k] = theta[k] + shift
        tm[k] = theta[k] - shift
        grad[k] = (cost_fn(tp) - cost_fn(tm)) / (2 * math.sin(shift))
        tp[k] = theta[k]
        tm[k] = theta[k]
    return grad


def calculate_gradients_finite_difference(cost_fn, theta, eps: float = 1e-5):
    """calculate_gradients_finite_difference (…Gradients.jl:250-269)."""
    theta = np.asarray(theta, np.float64)
    grad = np.zeros_like(theta)
    w = theta.copy()
    for k in range(theta.size):
        w[k] = theta[k] + eps
        cp = cost_fn(w)
        w[k] = theta[k] - eps
        cm = cost_fn(w)
        grad[k] = (cp - cm) / (2 * eps)
        w[k] = theta[k]
    return grad


def verify_gradients(grad_fn, cost_fn, theta, atol: float = 1e-4):
    """verify_gradients (…Gradients.jl:284-296) with the adjoint gradient in
    place of Enzyme: (match, max_diff) against parameter shift."""
    g = np.asarray(grad_fn(theta))
    ps = calculate_gradients_parameter_shift(cost_fn, theta)
    d = float(np.max(np.abs(g - ps)))
    return d < atol, d
