"""Drop-in gate functions with the reference's names and argument order.

Each `apply_*_psi(ψ, k, [θ,] N)` mirrors `apply_*_psi!` of
utils/cpu/cpuQuantumChannelGates.jl (1-based qubits, in-place, returns ψ);
ψ is a GPUState.  The work runs in the gfx950 kernels of libsmoq_amd.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import G, call
from .state import GPUState


def _check(psi: GPUState, N: int, *qubits: int) -> None:
    if psi.n != N:
        raise ValueError(f"state has {psi.n} qubits, N = {N}")
    for q in qubits:
        if not 1 <= q <= N:
            raise ValueError(f"qubit index {q} must be in 1..{N}")
    if len(set(qubits)) != len(qubits):
        raise ValueError(f"qubits {qubits} must be distinct")


def _named(psi: GPUState, gate: str, qubits, theta: float, N: int) -> GPUState:
    _check(psi, N, *qubits)
    q = (ctypes.c_int * 3)(*([k - 1 for k in qubits] + [-1] * (3 - len(qubits))))
    call("smoq_apply_named", psi.handle, G[gate], q, float(theta))
    return psi


# ---- single-qubit (cpuQuantumChannelGates.jl:142-390)
def apply_rx_psi(psi, k, theta, N):
    return _named(psi, "RX", [k], theta, N)


def apply_ry_psi(psi, k, theta, N):
    return _named(psi, "RY", [k], theta, N)


def apply_rz_psi(psi, k, theta, N):
    return _named(psi, "RZ", [k], theta, N)


def apply_hadamard_psi(psi, k, N):
    return _named(psi, "H", [k], 0.0, N)


def apply_pauli_x_psi(psi, k, N):
    return _named(psi, "X", [k], 0.0, N)


def apply_pauli_y_psi(psi, k, N):
    return _named(psi, "Y", [k], 0.0, N)


def apply_pauli_z_psi(psi, k, N):
    return _named(psi, "Z", [k], 0.0, N)


def apply_s_psi(psi, k, N):
    return _named(psi, "S", [k], 0.0, N)


def apply_sdagger_psi(psi, k, N):
    return _named(psi, "SDG", [k], 0.0, N)


def apply_t_psi(psi, k, N):
    return _named(psi, "T", [k], 0.0, N)


def apply_tdagger_psi(psi, k, N):
    return _named(psi, "TDG", [k], 0.0, N)


# ---- three-qubit (cpuQuantumChannelGates.jl:448-487)
def apply_ccz_psi(psi, q1, q2, q3, N):
    return _named(psi, "CCZ", [q1, q2, q3], 0.0, N)


def apply_toffoli_psi(psi, c1, c2, target, N):
    return _named(psi, "TOFFOLI", [c1, c2, target], 0.0, N)


# ---- jump operators (cpuQuantumChannelGates.jl:739-771)
def apply_sigma_minus_psi(psi, k, N):
    return _named(psi, "SIGMA_MINUS", [k], 0.0, N)


def apply_sigma_plus_psi(psi, k, N):
    return _named(psi, "SIGMA_PLUS", [k], 0.0, N)


# ---- two-qubit (cpuQuantumChannelGates.jl:870-1100)
def apply_cnot_psi(psi, control, target, N):
    return _named(psi, "CNOT", [control, target], 0.0, N)


def apply_cz_psi(psi, qi, qj, N):
    return _named(psi, "CZ", [qi, qj], 0.0, N)


def apply_cry_psi(psi, control, target, theta, N):
    return _named(psi, "CRY", [control, target], theta, N)


def apply_rxx_psi(psi, qi, qj, theta, N):
    return _named(psi, "RXX", [qi, qj], theta, N)


def apply_ryy_psi(psi, qi, qj, theta, N):
    return _named(psi, "RYY", [qi, qj], theta, N)


def apply_rzz_psi(psi, qi, qj, theta, N):
    return _named(psi, "RZZ", [qi, qj], theta, N)


# ---- generic matrices
def _mat(U, n):
    u = np.ascontiguousarray(np.asarray(U, dtype=np.complex128).reshape(n, n))
    return u, u.ctypes.data_as(ctypes.c_void_p)


def apply_gate(psi, U, qubit, N):
    """apply_gate! (utils/cpu/cpuQuantumChannelRandomUnitaries.jl:230)."""
    _check(psi, N, qubit)
    u, p = _mat(U, 2)
    call("smoq_apply_1q", psi.handle, qubit - 1, p)
    return psi


def apply_one_qubit_gate_bitwise_cpu(psi, U, qubit, N):
    """apply_one_qubit_gate_bitwise_cpu! (…UnitaryEvolutionTrotter.jl:77-99)."""
    return apply_gate(psi, U, qubit, N)


def apply_2qubit_gate(psi, U, q1, q2, N):
    """apply_2qubit_gate! (cpuQuantumChannelRandomUnitaries.jl:317-370):
    basis |00>,|10>,|01>,|11> with q1 as the first bit, for any q1, q2."""
    _check(psi, N, q1, q2)
    u, p = _mat(U, 4)
    call("smoq_apply_2q", psi.handle, q1 - 1, q2 - 1, p)
    return psi


def apply_two_qubit_gate_bitwise_cpu(psi, U, q1, q2, N):
    """apply_two_qubit_gate_bitwise_cpu! (…UnitaryEvolutionTrotter.jl:107-142).

    As in the reference, (q1, q2) are put in ascending order and U is NOT
    permuted: U's first basis bit is always the lower-numbered qubit."""
    lo, hi = (q1, q2) if q1 < q2 else (q2, q1)
    return apply_2qubit_gate(psi, U, lo, hi, N)
