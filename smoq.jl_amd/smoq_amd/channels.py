"""Single-qubit noise channels with the reference's names and dispatch
(utils/cpu/cpuQuantumChannelKrausOperators.jl `apply_channel_*!`).

On a state vector (GPUState, batch = trajectories) each channel is the
reference's stochastic trajectory branch, decided on the device from one
uniform per (qubit, item) — smoq_qubit_channel.  The reference draws those
with `rand()`; here they are an argument (or drawn from `rng`), so a run can
be replayed exactly.

On a density matrix (DensityMatrix, vec(ρ) as 2N qubits) each channel is the
reference's exact Kraus update of ρ, which acts on the four elements
ρ(i, j) with the same other bits as a 4×4 map on the pair (row bit k,
column bit k) — basis index b(row) + 2 b(col) — and so runs as one 2-qubit
gate of the fused passes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import QC, call
from .circuits import Tape
from .density import DensityMatrix
from .state import GPUState


def _qubit_list(qubits, N: int) -> list[int]:
    qs = [int(q) for q in np.atleast_1d(qubits)]
    for q in qs:
        if not 1 <= q <= N:
            raise ValueError(f"All qubit indices must be in 1..{N}, got {q}")
    return qs


def uniform_block(batch: int, n_qubits: int, uniforms=None, rng=None) -> np.ndarray:
    """The (n_qubits, batch) C-order block the device reads: u[k, b] for the
    k-th listed qubit of item b.  `uniforms` is given per item as
    uniforms[b, k] (a flat length-n_qubits array when batch = 1); without it,
    `rng` (default: a fresh numpy Generator) draws them."""
    if uniforms is None:
        g = rng if rng is not None else np.random.default_rng()
        return np.ascontiguousarray(g.random((n_qubits, batch)))
    u = np.asarray(uniforms, np.float64)
    if u.size != n_qubits * batch:
        raise ValueError(f"need {n_qubits * batch} uniforms (batch {batch} x {n_qubits} qubits), "
                         f"got {u.size}")
    return np.ascontiguousarray(u.reshape(batch, n_qubits).T)


def _records(out: np.ndarray, batch: int) -> np.ndarray:
    r = out.reshape(-1, batch).T.astype(np.int64)
    return r[0] if batch == 1 else r


def qubit_channel(psi: GPUState, kind: str, p: float, qubits, uniforms=None, *,
                  rng=None) -> np.ndarray:
    """smoq_qubit_channel for the 1-based `qubits` in order; returns the
    per-qubit records (shape (n_qubits,), or (batch, n_qubits)):
    measure_z → outcome; the flip channels → 0 none, 1 X, 2 Y, 3 Z;
    amplitude_damping → 4 on a jump else 0; force_zero / force_one → the bit."""
    qs = _qubit_list(qubits, psi.n)
    if not qs:
        return np.zeros((0,) if psi.batch == 1 else (psi.batch, 0), np.int64)
    q0 = np.ascontiguousarray([q - 1 for q in qs], np.int32)
    k = QC[kind]
    u = None
    if kind not in ("force_zero", "force_one"):
        u = uniform_block(psi.batch, len(qs), uniforms, rng)
    out = np.empty(len(qs) * psi.batch, np.int32)
    call("smoq_qubit_channel", psi.handle, len(qs), q0.ctypes.data_as(ctypes.c_void_p), k,
         ctypes.c_double(float(p)), None if u is None else u.ctypes.data_as(ctypes.c_void_p),
         out.ctypes.data_as(ctypes.c_void_p))
    return _records(out, psi.batch)


# ---- density-matrix maps ---------------------------------------------------------
def kraus_pair_map(kind: str, p: float) -> tuple[str, np.ndarray]:
    """The reference's Kraus update of one qubit (…KrausOperators.jl:86-235)
    as a gate on (row bit, column bit): ("DIAG2", diagonal) or ("U2", 4x4)."""
    p = float(p)
    if kind == "dephasing":  # coherences times sqrt(1-p)
        d = np.sqrt(1 - p)
        return "DIAG2", np.array([1.0, d, d, 1.0], np.complex128)
    if kind == "phase_flip":  # coherences times 1 - 2p
        c = 1 - 2 * p
        return "DIAG2", np.array([1.0, c, c, 1.0], np.complex128)
    M = np.zeros((4, 4), np.complex128)
    if kind == "depolarizing":
        # same bits: (1-p) v + p/3 (2 v_flip + v); different bits: (1-p) v - p/3 v
        p3 = p / 3.0
        M[0, 0] = M[3, 3] = 1 - p + p3
        M[0, 3] = M[3, 0] = 2 * p3
        M[1, 1] = M[2, 2] = 1 - p - p3
    elif kind == "bit_flip":  # (1-p) ρ + p XρX
        M[0, 0] = M[1, 1] = M[2, 2] = M[3, 3] = 1 - p
        M[0, 3] = M[3, 0] = M[1, 2] = M[2, 1] = p
    elif kind == "amplitude_damping":
        M[0, 0], M[0, 3] = 1.0, p
        M[3, 3] = 1 - p
        M[1, 1] = M[2, 2] = np.sqrt(1 - p)
    else:
        raise ValueError(f"unknown channel {kind!r}")
    return "U2", M


def _apply(kind: str, state, p: float, qubits, N: int, uniforms, rng):
    if isinstance(state, DensityMatrix):
        if state.N != N:
            raise ValueError(f"density matrix has {state.N} qubits, N = {N}")
        if not 0 <= p <= 1:
            raise ValueError(f"channel probability must be in [0,1], got {p}")
        qs = _qubit_list(qubits, N)
        if p == 0.0 or not qs:
            return state
        gate, m = kraus_pair_map(kind, p)
        t = Tape(2 * N)
        for q in qs:
            t.add(gate, [q - 1, q - 1 + N], m=m)
        t.apply(state.vec)
        return state
    if state.n != N:
        raise ValueError(f"state has {state.n} qubits, N = {N}")
    qubit_channel(state, kind, p, qubits, uniforms, rng=rng)
    return state


def apply_channel_dephasing(state, p: float, qubits, N: int, uniforms=None, *, rng=None):
    """apply_channel_dephasing! (…KrausOperators.jl:416-431): ρ coherences of
    each listed qubit times sqrt(1-p); ψ gets Z with probability
    (1 - sqrt(1-p))/2 per qubit (_apply_dephasing_single!, :262-282; ρ :86-103)."""
    return _apply("dephasing", state, p, qubits, N, uniforms, rng)


def apply_channel_depolarizing(state, p: float, qubits, N: int, uniforms=None, *, rng=None):
    """apply_channel_depolarizing! (…KrausOperators.jl:448-463): ρ → (1-p)ρ +
    (p/3)(XρX + YρY + ZρZ); ψ gets X, Y, Z when u < p/3, 2p/3, p (:285-318; ρ :107-141)."""
    return _apply("depolarizing", state, p, qubits, N, uniforms, rng)


def apply_channel_amplitude_damping(state, gamma: float, qubits, N: int, uniforms=None, *,
                                    rng=None):
    """apply_channel_amplitude_damping! (…KrausOperators.jl:482-497): ρ00 +=
    γρ11, ρ11 *= 1-γ, coherences times sqrt(1-γ); ψ jumps (|1> amplitudes
    added onto |0>) when u < γ P(1), else |1> times sqrt(1-γ), then
    renormalised (:350-385; ρ :145-181)."""
    return _apply("amplitude_damping", state, gamma, qubits, N, uniforms, rng)


def apply_channel_bit_flip(state, p: float, qubits, N: int, uniforms=None, *, rng=None):
    """apply_channel_bit_flip! (…KrausOperators.jl:514-529): (1-p)ρ + pXρX;
    ψ gets X when u < p (:321-334; ρ :184-215)."""
    return _apply("bit_flip", state, p, qubits, N, uniforms, rng)


def apply_channel_phase_flip(state, p: float, qubits, N: int, uniforms=None, *, rng=None):
    """apply_channel_phase_flip! (…KrausOperators.jl:546-561): (1-p)ρ + pZρZ;
    ψ gets Z when u < p (:337-347; ρ :218-235)."""
    return _apply("phase_flip", state, p, qubits, N, uniforms, rng)


__all__ = [
    "apply_channel_amplitude_damping", "apply_channel_bit_flip", "apply_channel_dephasing",
    "apply_channel_depolarizing", "apply_channel_phase_flip", "kraus_pair_map", "qubit_channel",
    "uniform_block",
]
