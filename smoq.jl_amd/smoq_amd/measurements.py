"""Measurement sampling with the reference's decision rules
(utils/cpu/cpuQuantumStateMeasurements.jl and the random-circuit benchmark).

Uniforms are supplied by the caller (see streams.uniforms) instead of the
reference's global `rand()`, so that draws are reproducible bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import SAMPLE_CUMSUM_SEARCH, SAMPLE_SEQUENTIAL, call
from .state import GPUState

_SEM = {"cumsum": SAMPLE_CUMSUM_SEARCH, "sequential": SAMPLE_SEQUENTIAL}


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


def projective_measurement_all(psi: GPUState, N: int, r: float):
    """projective_measurement_all! (cpuQuantumStateMeasurements.jl:484-502):
    one draw with sample_from_probabilities' rule, then ψ collapses to the
    drawn basis state.  Returns (bitstring, ψ)."""
    if psi.n != N:
        raise ValueError(f"state has {psi.n} qubits, N = {N}")
    idx = int(sample_indices(psi, [r], "sequential")[0])
    psi.set_basis(idx)
    return [(idx >> (k - 1)) & 1 for k in range(1, N + 1)], psi
