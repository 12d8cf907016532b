"""Pins the oracle's density-matrix restatements (lindblad_dm_step[_trotter]!,
_apply_dissipator_dm!, the _expect_*_dm read-outs) against dense matrices and
the first-order Lindblad generator, and against the reference's own recorded
run: the DM+Trotter and DM+Exact columns of the σ_z-dephasing demo's
output.log (tests/golden/mcwf_dephasing_xxz.json).  CPU only."""
import json
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
SM = np.array([[0, 1], [0, 0]], np.complex128)  # σ⁻ = |0><1|
SP = SM.T.copy()
X = np.array([[0, 1], [1, 0]], np.complex128)
Z = np.diag([1.0, -1.0]).astype(np.complex128)


def rand_rho(N, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(1 << N, 1 << N)) + 1j * rng.normal(size=(1 << N, 1 << N))
    rho = A @ A.conj().T
    return rho / np.trace(rho).real


def dense_2q(U, lo, hi, N):
    """Dense matrix of apply_2q_trotter (sorted qubits, U basis [00, lo, hi, both])."""
    D = 1 << N
    M = np.zeros((D, D), np.complex128)
    for c in range(D):
        e = np.zeros(D, np.complex128)
        e[c] = 1.0
        orc.apply_2q_trotter(e, U, lo, hi, N)
        M[:, c] = e
    return M


def random_gates(N, n, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(n):
        if rng.uniform() < 0.5:
            q = int(rng.integers(1, N + 1))
            U = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
            gates.append(([q], U))
        else:
            a, b = (int(v) + 1 for v in rng.choice(N, 2, replace=False))
            U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
            gates.append(([a, b], U))
    return gates


def test_dm_gate_list_is_u_rho_u_dagger():
    N = 3
    gates = random_gates(N, 8, 1)
    U = np.eye(1 << N, dtype=np.complex128)
    for idx, G in gates:
        M = orc.dense_1q(G, idx[0] - 1, N) if len(idx) == 1 else dense_2q(G, min(idx), max(idx), N)
        U = M @ U
    rho = rand_rho(N, 2)
    want = U @ rho @ U.conj().T
    got = orc.dm_apply_gate_list(rho.copy(), gates, N)
    assert np.max(np.abs(got - want)) < 1e-13


@pytest.mark.parametrize("kind,L,pair", [("sigma_minus", SM, False), ("sigma_plus", SP, False),
                                         ("sigma_z", Z, True), ("sigma_x", X, True)])
def test_dm_dissipator_is_first_order_lindblad(kind, L, pair):
    """ρ + γdt (L ρ L† − ½{L†L, ρ}) on one site, dense."""
    N, gdt = 3, 0.07
    for k in (1, 2, 3):
        rho = rand_rho(N, 10 + k)
        Lk = orc.dense_1q(L, k - 1, N)
        LdL = Lk.conj().T @ Lk
        want = rho + gdt * (Lk @ rho @ Lk.conj().T - 0.5 * (LdL @ rho + rho @ LdL))
        got = orc.dm_dissipator(rho.copy(), kind, k, gdt, N)
        assert np.max(np.abs(got - want)) < 1e-15, (kind, k)


def test_dm_expectations_are_traces():
    N = 4
    rho = rand_rho(N, 5)
    for k in range(1, N + 1):
        for p, (x, z) in {"x": (1, 0), "y": (1, 1), "z": (0, 1)}.items():
            P = orc.pauli_matrix(x << (k - 1), z << (k - 1), N)
            assert abs(orc.dm_expect_local(rho, k, N, p) - np.trace(P @ rho).real) < 1e-14
    for i in range(1, N):
        for p, (x, z) in {"xx": (3, 0), "yy": (3, 3), "zz": (0, 3)}.items():
            P = orc.pauli_matrix(x << (i - 1), z << (i - 1), N)
            assert abs(orc.dm_expect_corr(rho, i, i + 1, N, p) - np.trace(P @ rho).real) < 1e-14


def demo_means(rho, N):
    sites = [np.mean([orc.dm_expect_local(rho, k, N, p) for k in range(1, N + 1)]) for p in "xyz"]
    pairs = [np.mean([orc.dm_expect_corr(rho, i, i + 1, N, p) for i in range(1, N)])
             for p in ("xx", "yy", "zz")]
    return np.array(sites + pairs)


def demo_gates(N, dt):
    from smoq_amd.mcwf import xxz_field_trotter_gates

    return [(list(g.indices), g.U) for g in xxz_field_trotter_gates(N, dt)]


def demo_hamiltonian(N):
    from smoq_amd.hamiltonian import build_hamiltonian_parameters, pauli_sum

    params = build_hamiltonian_parameters(N, 1, J_x_direction=(1.0, 1.0, 0.5),
                                          J_y_direction=(0.0, 0.0, 0.0), h_field=(1.0, 0.0, 0.0))
    xs, zs, cs = pauli_sum(params)
    return sum(c * orc.pauli_matrix(x, z, N) for x, z, c in zip(xs, zs, cs))


@pytest.mark.parametrize("N", [2, 4, 6])
def test_oracle_dm_reproduces_the_reference_demo_columns(N):
    """200 steps of lindblad_dm_step_trotter! and of lindblad_dm_step! from
    |0…0><0…0| (XXZ J = (1, 1, 0.5), h_x = 1, σ_z dephasing γ = 0.1,
    dt = 0.05): the six means at t = 10 against the reference's recorded
    DM+Trotter and DM+Exact values (printed with 4 decimals)."""
    with open(os.path.join(HERE, "golden", "mcwf_dephasing_xxz.json")) as f:
        ref = json.load(f)
    P = ref["params"]
    gamma, dt, steps = P["gamma"], P["dt"], P["n_steps"]
    rho0 = np.zeros((1 << N, 1 << N), np.complex128)
    rho0[0, 0] = 1.0
    gates = demo_gates(N, dt)
    rho = rho0.copy()
    for _ in range(steps):
        orc.dm_trotter_step(rho, gates, N, gamma, dt)
    got = demo_means(rho, N)
    want = np.array(ref["reference"][str(N)]["DM+Trotter"])
    assert np.max(np.abs(got - want)) < 6e-5, (got, want)
    U = expm(-1j * demo_hamiltonian(N) * dt)
    rho = rho0.copy()
    for _ in range(steps):
        orc.dm_exact_step(rho, U, N, gamma, dt)
    got = demo_means(rho, N)
    want = np.array(ref["reference"][str(N)]["DM+Exact"])
    assert np.max(np.abs(got - want)) < 6e-5, (got, want)
