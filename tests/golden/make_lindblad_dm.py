"""Writes lindblad_tfim_damping.json: <Z_i>(t) of the exact density-matrix
Lindblad equation for the damped transverse-field Ising chain,

    dρ/dt = -i[H, ρ] + γ Σ_k (σ⁻_k ρ σ⁺_k - ½{σ⁺_k σ⁻_k, ρ}),
    H = J Σ_k Z_k Z_{k+1} + h Σ_k X_k   (open chain),

from ρ(0) = |1…1><1…1| (every qubit excited; σ⁻ = |0><1| lowers bit k of
the index), at N = 8 and N = 10, J = 1, h = 0.8, γ = 0.1, t = 1, 2, …, 10.
The Liouvillian acts on vec(ρ) (column-major) as a sparse matrix and the
time series comes from scipy's expm_multiply.  These are the reference
values the trajectory average of the batched MCWF kernel is checked
against.  Run: python tests/golden/make_lindblad_dm.py (about a minute).
"""
from __future__ import annotations

import json
import os

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "lindblad_tfim_damping.json")
PARAMS = {"J": 1.0, "h": 0.8, "gamma": 0.1, "times": [float(t) for t in range(1, 11)],
          "sizes": [8, 10]}


def hamiltonian(N: int, J: float, h: float) -> sp.csr_matrix:
    D = 1 << N
    idx = np.arange(D)
    zs = 1.0 - 2.0 * ((idx[:, None] >> np.arange(N)[None, :]) & 1)
    diag = J * sum(zs[:, k] * zs[:, k + 1] for k in range(N - 1)) if N > 1 else np.zeros(D)
    rows, cols, vals = [idx], [idx], [diag.astype(np.complex128)]
    for k in range(N):
        rows.append(idx)
        cols.append(idx ^ (1 << k))
        vals.append(np.full(D, h, np.complex128))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(D, D))


def lowering(N: int, k: int) -> sp.csr_matrix:
    D = 1 << N
    src = np.arange(D)[(np.arange(D) >> k) & 1 == 1]
    return sp.csr_matrix((np.ones(src.size, np.complex128), (src ^ (1 << k), src)), shape=(D, D))


def liouvillian(N: int, J: float, h: float, gamma: float) -> sp.csr_matrix:
    """L with vec(A ρ B) = (B^T ⊗ A) vec(ρ), vec column-major."""
    D = 1 << N
    H = hamiltonian(N, J, h)
    I = sp.identity(D, dtype=np.complex128, format="csr")
    L = -1j * (sp.kron(I, H) - sp.kron(H.T, I))
    for k in range(N):
        c = lowering(N, k)
        cdc = (c.conj().T @ c).tocsr()
        L = L + gamma * (sp.kron(c.conj(), c) - 0.5 * sp.kron(I, cdc) - 0.5 * sp.kron(cdc.T, I))
    return L.tocsr()


def z_history(N: int, J: float, h: float, gamma: float, times) -> np.ndarray:
    """[len(times), N] array of <Z_i>(t)."""
    D = 1 << N
    rho0 = np.zeros(D * D, np.complex128)
    rho0[(D - 1) * D + (D - 1)] = 1.0
    L = liouvillian(N, J, h, gamma)
    times = np.asarray(times, np.float64)
    grid = np.concatenate([[0.0], times])
    out = expm_multiply(L, rho0, start=0.0, stop=float(times[-1]), num=len(times) + 1,
                        endpoint=True)
    # expm_multiply returns the uniform grid start..stop; pick the requested times
    ts = np.linspace(0.0, float(times[-1]), len(times) + 1)
    assert np.allclose(ts, grid), "times must be uniform from 0"
    idx = np.arange(D)
    zs = 1.0 - 2.0 * ((idx[:, None] >> np.arange(N)[None, :]) & 1)
    res = []
    for v in out[1:]:
        pop = v.reshape(D, D, order="F").diagonal().real
        res.append(pop @ zs)
    return np.array(res)


def main() -> None:
    data = dict(PARAMS)
    data["z"] = {}
    for N in PARAMS["sizes"]:
        z = z_history(N, PARAMS["J"], PARAMS["h"], PARAMS["gamma"], PARAMS["times"])
        data["z"][str(N)] = z.tolist()
        print(f"N={N}: <Z_1>(t) = {np.round(z[:, 0], 6).tolist()}", flush=True)
    with open(FIXTURE, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", FIXTURE)


if __name__ == "__main__":
    main()
