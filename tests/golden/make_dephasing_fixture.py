"""Writes mcwf_dephasing_xxz.json: the σ_z-dephasing comparison of the
reference's demo (scripts_quantum_simulations/
demo_lidbladian_dm_exact_trotter_mcwf_exact_trotter_comparison.jl:64-300)
at N = 2, 4, 6, 8:

* `reference`: the FINAL VALUES (t = 10) tables of its recorded run,
  scripts_quantum_simulations/demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/
  output.log (columns DM+Exact, DM+Trotter, MCWF+Exact, MCWF+Trotter;
  rows mean <X>, <Y>, <Z> over sites and mean <XX>, <YY>, <ZZ> over
  neighbour pairs; 4 decimals);
* `exact`: the same six numbers from the exact Lindblad equation
  dρ/dt = -i[H, ρ] + γ Σ_k (Z_k ρ Z_k - ρ), H = Σ (X X + Y Y + 0.5 Z Z) over
  the open chain + Σ X_k, ρ(0) = |0…0><0…0|, γ = 0.1, t = 10 (sparse
  Liouvillian, scipy expm_multiply).

The reference file is read here, once, to extract its numbers; the tests
read only the JSON.  Run: python tests/golden/make_dephasing_fixture.py
"""
from __future__ import annotations

import json
import os
import re

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "mcwf_dephasing_xxz.json")
LOG = ("/root/reference/scripts_quantum_simulations/"
       "demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/output.log")
SIZES = (2, 4, 6, 8)
ROWS = ("X", "Y", "Z", "XX", "YY", "ZZ")
COLUMNS = ("DM+Exact", "DM+Trotter", "MCWF+Exact", "MCWF+Trotter")
PARAMS = {"J": [1.0, 1.0, 0.5], "hx": 1.0, "gamma": 0.1, "dt": 0.05, "n_steps": 200,
          "t": 10.0, "initial": "all zero", "n_traj_reference": 1000}


def parse_log(path: str) -> dict:
    text = open(path, encoding="utf-8").read()
    out = {}
    for block in re.split(r"\n={10,}\n\s+N = ", text)[1:]:
        n = int(block.split()[0])
        if n not in SIZES:
            continue
        table = block.split("FINAL VALUES (t = 10)")[1].split("DEVIATIONS")[0]
        vals = {}
        for row in ROWS:
            m = re.search(r"⟨" + row + r"⟩\s*\|([^\n]+)", table)
            vals[row] = [float(v) for v in m.group(1).split("|")]
        out[str(n)] = {c: [vals[r][i] for r in ROWS] for i, c in enumerate(COLUMNS)}
    return out


def paulis(N: int):
    D = 1 << N
    idx = np.arange(D)
    X = lambda k: sp.csr_matrix((np.ones(D), (idx ^ (1 << k), idx)), shape=(D, D))
    Z = lambda k: sp.diags(1.0 - 2.0 * ((idx >> k) & 1)).tocsr()
    # Y|b> = i (-1)^b |b ^ 1>
    Y = lambda k: sp.csr_matrix((1j * (1.0 - 2.0 * ((idx >> k) & 1)), (idx ^ (1 << k), idx)),
                                shape=(D, D))
    return X, Y, Z


def exact_values(N: int, J=(1.0, 1.0, 0.5), hx=1.0, gamma=0.1, t=10.0) -> list:
    """The six mean observables at time t."""
    D = 1 << N
    X, Y, Z = paulis(N)
    H = sum(J[0] * X(k) @ X(k + 1) + J[1] * Y(k) @ Y(k + 1) + J[2] * Z(k) @ Z(k + 1)
            for k in range(N - 1)) + sum(hx * X(k) for k in range(N))
    H = sp.csr_matrix(H, dtype=np.complex128)
    I = sp.identity(D, dtype=np.complex128, format="csr")
    L = -1j * (sp.kron(I, H) - sp.kron(H.T, I))
    for k in range(N):
        Zk = sp.csr_matrix(Z(k), dtype=np.complex128)
        L = L + gamma * (sp.kron(Zk.T, Zk) - sp.kron(I, I))
    rho0 = np.zeros(D * D, np.complex128)
    rho0[0] = 1.0
    rho = expm_multiply(L.tocsr(), rho0, start=0.0, stop=t, num=2, endpoint=True)[-1]
    rho = rho.reshape(D, D, order="F")
    ex = lambda O: float(np.real(O.multiply(rho.T).sum()))
    site = lambda f: float(np.mean([ex(f(k)) for k in range(N)]))
    pair = lambda f: float(np.mean([ex(f(k) @ f(k + 1)) for k in range(N - 1)]))
    return [site(X), site(Y), site(Z), pair(X), pair(Y), pair(Z)]


def main() -> None:
    data = {"params": PARAMS, "rows": list(ROWS), "columns": list(COLUMNS),
            "reference": parse_log(LOG), "exact": {}}
    for N in SIZES:
        data["exact"][str(N)] = exact_values(N)
        print(N, np.round(data["exact"][str(N)], 4), data["reference"][str(N)]["DM+Exact"],
              flush=True)
    with open(FIXTURE, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", FIXTURE)


if __name__ == "__main__":
    main()
