"""Trotter evolution with the reference's names
(utils/cpu/cpuQuantumChannelUnitaryEvolutionTrotter.jl), on the device.

Local unitaries are computed on the host (scipy's expm in place of Julia's
exp, to ~1e-16) and handed to the fused-pass kernels as a tape: 2-qubit
gates keep the reference's convention of apply_two_qubit_gate_bitwise_cpu!
(indices sorted ascending, U not permuted), 1-qubit gates are generic 2x2.
The reference integrator is first order; `second_order_gates` composes the
symmetric step from the same gate list.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.linalg import expm

from .circuits import Plan, Tape
from .hamiltonian import HamiltonianParams

_SX = np.array([[0, 1], [1, 0]], np.complex128)
_SY = np.array([[0, -1j], [1j, 0]], np.complex128)
_SZ = np.array([[1, 0], [0, -1]], np.complex128)


@dataclass
class FastTrotterGate:
    """…UnitaryEvolutionTrotter.jl:62-65 (indices 1-based)."""
    indices: list
    U: np.ndarray


def precompute_trotter_gates_bitwise_cpu(params: HamiltonianParams, dt: float):
    """…UnitaryEvolutionTrotter.jl:322-354: exp(-i(Jx XX + Jy YY + Jz ZZ)dt)
    per bond (x bonds then y bonds), then exp(-i(hx X + hy Y + hz Z)dt) per
    site with any |h| > 1e-14."""
    gates = []
    XX, YY, ZZ = np.kron(_SX, _SX), np.kron(_SY, _SY), np.kron(_SZ, _SZ)
    for c in list(params.x_bonds) + list(params.y_bonds):
        H = c.Jxx * XX + c.Jyy * YY + c.Jzz * ZZ
        gates.append(FastTrotterGate([c.i, c.j], expm(-1j * H * dt)))
    for f in params.fields:
        if abs(f.hx) > 1e-14 or abs(f.hy) > 1e-14 or abs(f.hz) > 1e-14:
            H = f.hx * _SX + f.hy * _SY + f.hz * _SZ
            gates.append(FastTrotterGate([f.idx], expm(-1j * H * dt)))
    return gates


def second_order_gates(params: HamiltonianParams, dt: float, merge_middle: bool = True):
    """Symmetric step U2(dt) = Π_{g} g(dt/2) · Π_{g reversed} g(dt/2) built from
    precompute_trotter_gates_bitwise_cpu(params, dt/2); the two copies of the
    last gate meet in the middle and are multiplied into one."""
    half = precompute_trotter_gates_bitwise_cpu(params, dt / 2)
    seq = half + half[::-1]
    if merge_middle and half:
        m = len(half) - 1
        g = half[m]
        seq = half[:m] + [FastTrotterGate(list(g.indices), g.U @ g.U)] + half[::-1][1:]
    return seq


def trotter_tape(gates, N: int) -> Tape:
    t = Tape(N)
    for g in gates:
        if len(g.indices) == 1:
            t.add("U1", [g.indices[0] - 1], m=g.U)
        elif len(g.indices) == 2:
            lo, hi = sorted(g.indices)
            t.add("U2", [lo - 1, hi - 1], m=g.U)
        else:
            raise ValueError("Gates with more than 2 qubits not supported")
    return t


_plans: dict = {}


def _plan_for(gates, N, psi) -> Plan:
    key = (id(gates), len(gates), N, psi.dtype)
    p = _plans.get(key)
    if p is None or p[0] is not gates:
        plan = trotter_tape(gates, N).compile(psi.dtype, ctx=psi.ctx)
        _plans[key] = (gates, plan)
        return plan
    return p[1]


def apply_fast_trotter_step_cpu(psi, gates, N: int):
    """apply_fast_trotter_step_cpu! (…UnitaryEvolutionTrotter.jl:150-162)."""
    if psi.n != N:
        raise ValueError(f"state has {psi.n} qubits, N = {N}")
    _plan_for(gates, N, psi).run(psi)
    return psi


def evolve_trotter_psi_cpu(psi, gates, N: int, n_substeps: int = 1):
    """evolve_trotter_psi_cpu! (…UnitaryEvolutionTrotter.jl:250-256)."""
    if psi.n != N:
        raise ValueError(f"state has {psi.n} qubits, N = {N}")
    _plan_for(gates, N, psi).run(psi, n_substeps)
    return psi


def solve_trotter_bitwise_cpu(psi0, gates, N: int, t_list, dt: float):
    """solve_trotter_bitwise_cpu (…UnitaryEvolutionTrotter.jl:293-311):
    returns (t_list, states) with the states copied to host arrays."""
    psi = psi0.clone()
    plan = _plan_for(gates, N, psi)
    states = [psi.to_numpy()]
    t_cur = t_list[0]
    for t_target in t_list[1:]:
        steps = 0
        while t_cur < t_target - dt / 2:
            steps += 1
            t_cur += dt
        plan.run(psi, steps)
        states.append(psi.to_numpy())
    return list(t_list), states
