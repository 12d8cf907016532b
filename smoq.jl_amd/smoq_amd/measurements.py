"""Measurement sampling with the reference's decision rules
(utils/cpu/cpuQuantumStateMeasurements.jl and the random-circuit benchmark).

Uniforms are supplied by the caller (see streams.uniforms) instead of the
reference's global `rand()`, so that draws are reproducible bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import QC, SAMPLE_CUMSUM_SEARCH, SAMPLE_SEQUENTIAL, SAMPLE_SEQUENTIAL_DIAG, call
from .channels import _qubit_list, _records, qubit_channel, uniform_block
from .circuits import Tape
from .density import DensityMatrix
from .state import GPUState

_SEM = {"cumsum": SAMPLE_CUMSUM_SEARCH, "sequential": SAMPLE_SEQUENTIAL,
        "sequential_diag": SAMPLE_SEQUENTIAL_DIAG}


def sample_indices(psi: GPUState, uniforms, semantics: str = "cumsum") -> np.ndarray:
    """0-based basis indices, one per uniform.

    "cumsum": searchsortedfirst(cumsum(abs2.(ψ)), r) - 1 as in
    run_circuit_sampling (run_benchmark_smoq_Float64.jl:424-429); 2^N when r
    exceeds the last prefix (the reference's bitstring is then all zeros).
    "sequential": sample_from_probabilities (cpuQuantumStateMeasurements.jl:203-213).
    """
    r = np.ascontiguousarray(uniforms, np.float64)
    out = np.empty(r.size, np.uint64)
    call("smoq_sample", psi.handle, r.ctypes.data_as(ctypes.c_void_p), r.size, _SEM[semantics],
         out.ctypes.data_as(ctypes.c_void_p))
    return out


def bitstrings(indices, N: int) -> np.ndarray:
    """bits[s, q-1] = (idx >> (q-1)) & 1 — the reference's per-shot vectors."""
    idx = np.asarray(indices, np.uint64)
    sh = np.arange(N, dtype=np.uint64)
    return ((idx[:, None] >> sh[None, :]) & np.uint64(1)).astype(np.int64)


def _one_item(state, fn: str) -> None:
    b = state.vec.batch if isinstance(state, DensityMatrix) else state.batch
    if b != 1:
        raise ValueError(f"{fn} draws from one state; got a batch of {b}")


def _check_n(state, N: int) -> None:
    n = state.N if isinstance(state, DensityMatrix) else state.n
    if n != N:
        raise ValueError(f"state has {n} qubits, N = {N}")


def _draw(r, rng) -> float:
    if r is None:
        return float((rng if rng is not None else np.random.default_rng()).random())
    return float(r)


def projective_measurement_all(state, N: int, r: float | None = None, *, rng=None):
    """projective_measurement_all! (cpuQuantumStateMeasurements.jl:484-502,
    505-520): one draw with sample_from_probabilities' rule over |ψ_i|^2 (or
    real(ρ_ii)), then the state collapses to the drawn basis state (|i><i|).
    Returns (bitstring, state); batch must be 1."""
    _check_n(state, N)
    _one_item(state, "projective_measurement_all")
    u = _draw(r, rng)
    if isinstance(state, DensityMatrix):
        idx = int(sample_indices(state.vec, [u], "sequential_diag")[0])
        state.vec.set_basis(idx | (idx << N))
    else:
        idx = int(sample_indices(state, [u], "sequential")[0])
        state.set_basis(idx)
    return [(idx >> (k - 1)) & 1 for k in range(1, N + 1)], state


def _bases(basis, n: int) -> list[str]:
    if isinstance(basis, str):
        return [basis.lstrip(":").lower()] * n
    b = [str(x).lstrip(":").lower() for x in basis]
    if len(b) != n:
        raise ValueError("Must specify basis for each qubit")
    return b


def _dm_measure(rho: DensityMatrix, kind: str, qs, uniforms, rng) -> np.ndarray:
    q0 = np.ascontiguousarray([q - 1 for q in qs], np.int32)
    B = rho.vec.batch
    u = uniform_block(B, len(qs), uniforms, rng) if kind == "measure_z" else None
    out = np.empty(len(qs) * B, np.int32)
    call("smoq_dm_qubit_channel", rho.vec.handle, len(qs), q0.ctypes.data_as(ctypes.c_void_p),
         QC[kind], None if u is None else u.ctypes.data_as(ctypes.c_void_p),
         out.ctypes.data_as(ctypes.c_void_p))
    return _records(out, B)


def projective_measurement(state, qubits, basis, N: int, uniforms=None, *, rng=None):
    """projective_measurement! (cpuQuantumStateMeasurements.jl:319-345, 348-369,
    401-424): each listed qubit (1-based, in order) measured in its basis with
    collapse — :z directly (_measure_z_single!, :224-256 / :259-295), :x as
    H, Z-measure, H, :y as S†, H, Z-measure, H, S.  One uniform per (item,
    qubit): uniforms[b, k], or a flat list for batch 1.  Density matrices take
    :z only, as in the reference.  Returns (outcomes, state), outcomes of
    shape (n_qubits,) or (batch, n_qubits)."""
    _check_n(state, N)
    qs = _qubit_list(qubits, N)
    bases = _bases(basis, len(qs))
    for b in bases:
        if b not in ("z", "x", "y"):
            raise ValueError(f"Unknown basis: :{b}. Use :z, :x, :y.")
    if isinstance(state, DensityMatrix):
        if any(b != "z" for b in bases):
            raise NotImplementedError("X/Y basis for density matrix not implemented "
                                      "(as in the reference). Use :z or pure states.")
        return _dm_measure(state, "measure_z", qs, uniforms, rng), state
    B = state.batch
    u = uniform_block(B, len(qs), uniforms, rng)  # (n_qubits, batch)
    if all(b == "z" for b in bases):
        return qubit_channel(state, "measure_z", 0.0, qs, u.T), state
    outs = []
    for k, (q, b) in enumerate(zip(qs, bases)):
        if b == "x":
            Tape(N).add("H", [q - 1]).apply(state)
        elif b == "y":
            Tape(N).add("SDG", [q - 1]).add("H", [q - 1]).apply(state)
        outs.append(qubit_channel(state, "measure_z", 0.0, [q], u[k]))
        if b == "x":
            Tape(N).add("H", [q - 1]).apply(state)
        elif b == "y":
            Tape(N).add("H", [q - 1]).add("S", [q - 1]).apply(state)
    return np.concatenate(outs, axis=-1), state


def sample_state(state, qubits, basis, N: int, uniforms=None, *, rng=None):
    """sample_state (cpuQuantumStateMeasurements.jl:427-446): outcomes without
    collapse — a measured copy of ψ; for ρ one draw from the populations
    (sample_from_probabilities, :z only) with its bits at the listed qubits."""
    _check_n(state, N)
    qs = _qubit_list(qubits, N)
    if isinstance(state, DensityMatrix):
        if _bases(basis, len(qs)) != ["z"] * len(qs):
            raise NotImplementedError("Only Z-basis sampling implemented for density matrices")
        _one_item(state, "sample_state")
        idx = int(sample_indices(state.vec, [_draw(uniforms, rng)], "sequential_diag")[0])
        return [(idx >> (k - 1)) & 1 for k in qs]
    cp = state.clone()
    try:
        out, _ = projective_measurement(cp, qs, basis, N, uniforms, rng=rng)
    finally:
        cp.free()
    return out


_FORCE = {"zero": ("force_zero", False), "one": ("force_one", False),
          "plus": ("force_zero", True), "minus": ("force_one", True)}


def reset_state(state, qubits, targets=None, N: int | None = None, uniforms=None, *, rng=None):
    """reset_state! (cpuQuantumStateMeasurements.jl:555-581, 638-718).

    ψ: for each listed qubit (1-based) in order, a Z measurement with collapse
    (one uniform per item: uniforms[b, k]) and then _force_qubit_state!
    (:584-626): the other half's amplitudes are added onto the target half
    and ψ renormalised; :plus / :minus force |0> / |1> and apply H.
    ρ: _force_qubit_rho! (:686-708), P_b ρ P_b / trace, :zero and :one only.
    `targets` is one symbol or one per qubit; reset_state(state, target, N
    [, uniforms]) resets every qubit."""
    if isinstance(qubits, str):  # reset_state!(state, target::Symbol, N)
        target, n_all, extra = qubits, int(targets), N
        qubits, targets, N = list(range(1, n_all + 1)), target, n_all
        if uniforms is None:
            uniforms = extra
    _check_n(state, N)
    qs = _qubit_list(qubits, N)
    tg = [targets] * len(qs) if isinstance(targets, str) else list(targets)
    if len(tg) != len(qs):
        raise ValueError("Must specify target for each qubit")
    tg = [str(t).lstrip(":").lower() for t in tg]
    for t in tg:
        if t not in _FORCE:
            raise ValueError(f"Unknown target state: :{t}. Use :zero, :one, :plus, :minus.")
    if isinstance(state, DensityMatrix):
        for q, t in zip(qs, tg):
            if t not in ("zero", "one"):
                raise ValueError("For density matrices, only :zero and :one targets implemented")
            _dm_measure(state, "force_" + t, [q], None, None)
        return state
    u = uniform_block(state.batch, len(qs), uniforms, rng)
    for k, (q, t) in enumerate(zip(qs, tg)):
        qubit_channel(state, "measure_z", 0.0, [q], u[k])
        kind, had = _FORCE[t]
        qubit_channel(state, kind, 0.0, [q])
        if had:
            Tape(N).add("H", [q - 1]).apply(state)
    return state


reset_qubit = reset_state  # reset_qubit! (…Measurements.jl:745-760), an alias
