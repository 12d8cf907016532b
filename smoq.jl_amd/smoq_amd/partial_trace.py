"""Reduced density matrices on the device, with the reference's names
(utils/cpu/cpuQuantumStatePartialTrace.jl: `partial_trace`,
`find_connected_regions`, `partial_trace_regions`).

partial_trace returns a DensityMatrix (vec(ρ_A) on the device, batch as the
input's) instead of a host Matrix, so reduced states stay resident for the
density-matrix read-outs; `.to_numpy()` gives the reference's matrix.  The
kept qubits index ρ_A in increasing order, as the reference's keep_indices.
"""
from __future__ import annotations

from ._lib import call
from .density import DensityMatrix
from .state import GPUState


def _keep_mask(trace_qubits, N: int) -> tuple[int, int]:
    tr = {int(q) for q in trace_qubits}
    for q in tr:
        if not 1 <= q <= N:
            raise ValueError(f"trace qubit {q} outside 1..{N}")
    keep = [q for q in range(1, N + 1) if q not in tr]
    if not keep:
        raise ValueError("Cannot trace out all sites - no subsystem remains")
    return sum(1 << (q - 1) for q in keep), len(keep)


def partial_trace(state, trace_qubits, N: int) -> DensityMatrix:
    """partial_trace(ψ, trace_qubits, N) (…PartialTrace.jl:68-127) and
    partial_trace(ρ, trace_qubits, N) (:141-188): the reduced density matrix
    of the qubits not in `trace_qubits` (1-based), for every batch item."""
    mask, k = _keep_mask(trace_qubits, N)
    if isinstance(state, DensityMatrix):
        if state.N != N:
            raise ValueError(f"density matrix has {state.N} qubits, N = {N}")
        src = state.vec
        out = DensityMatrix(k, src.dtype, src.batch, ctx=src.ctx)
        call("smoq_dm_partial_trace", src.handle, mask, out.vec.handle)
        return out
    if not isinstance(state, GPUState):
        raise TypeError("partial_trace takes a GPUState or a DensityMatrix")
    if state.n != N:
        raise ValueError(f"state has {state.n} qubits, N = {N}")
    out = DensityMatrix(k, state.dtype, state.batch, ctx=state.ctx)
    call("smoq_partial_trace", state.handle, mask, out.vec.handle)
    return out


def find_connected_regions(sites) -> list[list[int]]:
    """find_connected_regions (…PartialTrace.jl:267-291): consecutive runs of
    the sorted sites."""
    s = sorted(int(x) for x in sites)
    if not s:
        return []
    regions, cur = [], [s[0]]
    for a, b in zip(s, s[1:]):
        if b == a + 1:
            cur.append(b)
        else:
            regions.append(cur)
            cur = [b]
    regions.append(cur)
    return regions


def partial_trace_regions(state, trace_sites, N: int) -> tuple[DensityMatrix, ...]:
    """partial_trace_regions (…PartialTrace.jl:353-390, 439-476): one reduced
    density matrix per connected region of the kept sites, each with every
    other site traced out; ordered by the regions' first site."""
    _keep_mask(trace_sites, N)  # validates, and rejects tracing out everything
    tr = {int(q) for q in trace_sites}
    regions = find_connected_regions([q for q in range(1, N + 1) if q not in tr])
    if len(regions) == 1:
        return (partial_trace(state, trace_sites, N),)
    out = []
    for reg in regions:
        rs = set(reg)
        out.append(partial_trace(state, [q for q in range(1, N + 1) if q not in rs], N))
    return tuple(out)


__all__ = ["find_connected_regions", "partial_trace", "partial_trace_regions"]
