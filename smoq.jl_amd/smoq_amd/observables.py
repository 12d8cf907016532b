"""Observables with the reference's names (cpuQuantumStateObservables.jl).

All reductions run on the device (deterministic two-stage FP64 sums).
Pauli strings are given as (x, z) bit masks: X on qubit q sets x bit q-1,
Z sets z bit q-1, Y sets both (P = i^{|x&z|} X^x Z^z).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import call
from .state import GPUState


def expect_pauli(psi: GPUState, xmask, zmask) -> np.ndarray:
    """True <P_t> for every term; shape (n_terms,) or (batch, n_terms)."""
    xm = np.ascontiguousarray(np.atleast_1d(np.asarray(xmask, dtype=np.uint64)))
    zm = np.ascontiguousarray(np.atleast_1d(np.asarray(zmask, dtype=np.uint64)))
    if xm.shape != zm.shape:
        raise ValueError("xmask and zmask must have the same length")
    out = np.empty((psi.batch, xm.size), np.float64)
    call("smoq_expect_pauli", psi.handle, xm.size, xm.ctypes.data_as(ctypes.c_void_p),
         zm.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return out[0] if psi.batch == 1 else out


def pauli_masks(ops, qubits):
    """(x, z) masks of a Pauli string given as symbols and 1-based qubits."""
    x = z = 0
    for op, q in zip(ops, qubits):
        o = str(op).lstrip(":").upper()
        b = 1 << (q - 1)
        if o == "X":
            x |= b
        elif o == "Y":
            x |= b
            z |= b
        elif o == "Z":
            z |= b
        else:
            raise ValueError(f"Unknown Pauli operator: {op}")
    return x, z


def _sym(p) -> str:
    return str(p).lstrip(":").lower()


def expect_local(psi: GPUState, k: int, N: int, pauli) -> float:
    """expect_local (cpuQuantumStateObservables.jl:175-185)."""
    p = _sym(pauli)
    if p not in ("x", "y", "z"):
        raise ValueError(f"Unknown Pauli operator: {pauli}. Use :x, :y, or :z")
    if psi.n != N or not 1 <= k <= N:
        raise ValueError("qubit index out of range")
    x, z = pauli_masks([p], [k])
    return float(expect_pauli(psi, [x], [z])[0])


def expect_corr(psi: GPUState, i: int, j: int, N: int, pauli_pair) -> float:
    """expect_corr (cpuQuantumStateObservables.jl:306-316)."""
    p = _sym(pauli_pair)
    if p not in ("zz", "xx", "yy"):
        raise ValueError(f"Unknown Pauli pair: {pauli_pair}. Use :zz, :xx, or :yy")
    if psi.n != N or not (1 <= i <= N and 1 <= j <= N) or i == j:
        raise ValueError("qubit indices out of range")
    x, z = pauli_masks([p[0], p[1]], [i, j])
    return float(expect_pauli(psi, [x], [z])[0])


def get_expectation_pauli_string(psi: GPUState, operators, qubits, N: int) -> float:
    """get_expectation_pauli_string (cpuQuantumStateObservables.jl:840-894).

    Returns what the reference returns.  For strings containing X or Y the
    reference keeps only bras whose X/Y bits are all 0 paired with the ket
    whose X/Y bits are all 1 (its :966-1044): <P>/2 for one X/Y operator, a
    smaller partial sum for more.  This mirror keeps that value for drop-in
    parity; `expect_pauli` gives the true <P>."""
    if len(operators) != len(qubits):
        raise ValueError("operators and qubits must have same length")
    if not all(1 <= q <= N for q in qubits):
        raise ValueError("qubit indices must be in range [1, N]")
    if len(set(qubits)) != len(qubits):
        raise ValueError("qubits must be unique")
    if len(operators) == 0:
        return 1.0
    x, z = pauli_masks(operators, qubits)
    xm = np.array([x], np.uint64)
    zm = np.array([z], np.uint64)
    out = np.empty(psi.batch, np.float64)
    call("smoq_pauli_string_reference", psi.handle, 1, xm.ctypes.data_as(ctypes.c_void_p),
         zm.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return float(out[0])


def population_excited_psi(psi: GPUState, k: int, N: int) -> float:
    """population_excited_psi (cpuQuantumChannelGates.jl:784-796):
    sum of |ψ_i|^2 over indices with bit k-1 set = (<I> - <Z_k>)/2."""
    x, z = pauli_masks(["Z"], [k])
    v = expect_pauli(psi, [0, 0], [0, z])
    return float(0.5 * (v[0] - v[1]))


def xxz_terms(N: int, Jx: float, Jy: float, Jz: float, hx: float):
    """(x, z, coeff) of the open-chain H = Σ Jx XX + Jy YY + Jz ZZ + hx Σ X."""
    xs, zs, cs = [], [], []
    for i in range(1, N):
        b = (1 << (i - 1)) | (1 << i)
        for (x, z, c) in ((b, 0, Jx), (b, b, Jy), (0, b, Jz)):
            if c != 0.0:
                xs.append(x)
                zs.append(z)
                cs.append(c)
    if hx != 0.0:
        for i in range(N):
            xs.append(1 << i)
            zs.append(0)
            cs.append(hx)
    return xs, zs, cs


def compute_energy(psi: GPUState, N: int, Jx: float, Jy: float, Jz: float, hx: float) -> float:
    """compute_energy (scripts_variational_quantum_circuits/
    demo_variational_quantum_eigensolver.jl:251-289): <H> of the open XXZ
    chain with a transverse field."""
    xs, zs, cs = xxz_terms(N, Jx, Jy, Jz, hx)
    if not xs:
        return 0.0
    v = expect_pauli(psi, xs, zs)
    return float(np.dot(np.asarray(cs), v))


def staggered_magnetisation(psi: GPUState) -> float:
    """Σ_k (-1)^(k-1) <Z_k> / N (k 1-based)."""
    N = psi.n
    v = expect_pauli(psi, [0] * N, [1 << k for k in range(N)])
    sgn = np.array([(-1) ** k for k in range(N)], np.float64)
    return float(np.dot(sgn, v) / N)


def qrc_observable_masks(L: int, n_rails: int):
    """(xmask, zmask) of the observable list of
    measure_all_observables_state_vector_cpu! (cpuQuantumStateObservables.jl:1483-1519),
    in its order: X, Y, Z on every qubit k = 1..N (N = L·n_rails), then XX,
    YY, ZZ on every intra-rail bond (site, site+1), rail by rail, then on
    every rung (rail r site s, rail r+1 site s)."""
    if L < 1 or n_rails < 1:
        raise ValueError(f"L = {L}, n_rails = {n_rails}: both must be >= 1")
    N = L * n_rails
    xs, zs = [], []
    for k in range(N):
        b = 1 << k
        xs += [b, b, 0]
        zs += [0, b, b]
    bonds = [(r * L + s, r * L + s + 1) for r in range(n_rails) for s in range(L - 1)]
    bonds += [(r * L + s, (r + 1) * L + s) for r in range(n_rails - 1) for s in range(L)]
    for i, j in bonds:
        b = (1 << i) | (1 << j)
        xs += [b, b, 0]
        zs += [0, b, b]
    return xs, zs


def measure_all_observables_state_vector(psi: GPUState, L: int, n_rails: int) -> np.ndarray:
    """measure_all_observables_state_vector_cpu! (cpuQuantumStateObservables.jl:1483-1519):
    3N local values, then 3 per intra-rail bond, then 3 per rung, in the
    reference's order; one device call evaluates all of them (the tiled
    reduction reads the state once per group of tile bits).  Shape (n_ops,)
    or (batch, n_ops)."""
    if psi.n != L * n_rails:
        raise ValueError(f"state has {psi.n} qubits, L·n_rails = {L * n_rails}")
    xs, zs = qrc_observable_masks(L, n_rails)
    return expect_pauli(psi, xs, zs)


def measure_all_observables_one_pass(psi: GPUState, L: int, n_rails: int) -> np.ndarray:
    """measure_all_observables_one_pass (cpuQuantumStateObservables.jl:1532-1638):
    the same values in the same order as measure_all_observables_state_vector."""
    return measure_all_observables_state_vector(psi, L, n_rails)


def collective_spin_fast(psi: GPUState, N: int):
    """collective_spin_fast (cpuQuantumStateObservables.jl:1888-1908):
    (<Jx>, <Jy>, <Jz>) with Jα = ½ Σ_k σα_k, from one term list."""
    if psi.n != N:
        raise ValueError(f"state has {psi.n} qubits, N = {N}")
    xs = [1 << k for k in range(N)] * 2 + [0] * N
    zs = [0] * N + [1 << k for k in range(N)] * 2
    v = np.asarray(expect_pauli(psi, xs, zs))
    sums = [v[..., a * N:(a + 1) * N].sum(-1) / 2.0 for a in range(3)]
    if psi.batch == 1:
        return tuple(float(s) for s in sums)
    return tuple(sums)
