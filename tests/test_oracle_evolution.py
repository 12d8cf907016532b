"""Pins the oracle's evolution pieces (Trotter composition, Chebyshev series)
against closed forms and scipy's expm / expm_multiply.  CPU only."""
import math

import numpy as np
import pytest
from scipy.linalg import expm
from scipy.sparse.linalg import expm_multiply

from oracle import oracle as orc

X = np.array([[0, 1], [1, 0]], np.complex128)
Y = np.array([[0, -1j], [1j, 0]], np.complex128)
Z = np.diag([1.0, -1.0]).astype(np.complex128)


def op_on(mats, N):
    M = np.array([[1.0 + 0j]])
    for b in range(N - 1, -1, -1):
        M = np.kron(M, mats.get(b, np.eye(2)))
    return M


def xxz_dense(N, J, delta, h):
    H = np.zeros((1 << N, 1 << N), complex)
    for i in range(N - 1):
        H += J * (op_on({i: X, i + 1: X}, N) + op_on({i: Y, i + 1: Y}, N)) \
            + J * delta * op_on({i: Z, i + 1: Z}, N)
    for i in range(N):
        H += h[i] * op_on({i: Z}, N)
    return H


def xxz_gates(N, J, delta, h, dt):
    XX, YY, ZZ = np.kron(X, X), np.kron(Y, Y), np.kron(Z, Z)
    gates = [([i, i + 1], expm(-1j * (J * XX + J * YY + J * delta * ZZ) * dt))
             for i in range(1, N)]
    gates += [([i + 1], expm(-1j * h[i] * Z * dt)) for i in range(N)]
    return gates


def neel(N):
    psi = np.zeros(1 << N, np.complex128)
    psi[sum(1 << (k - 1) for k in range(2, N + 1, 2))] = 1.0
    return psi


def test_commuting_terms_are_exact():
    # ZZ bonds and Z fields commute: one first-order step is exact
    N = 6
    rng = np.random.default_rng(0)
    h = rng.uniform(-1, 1, N)
    H = xxz_dense(N, 0.0, 1.0, h)  # J = 0 leaves only the diagonal part
    ZZ = np.kron(Z, Z)
    gates = [([i, i + 1], expm(-1j * ZZ * 0.0)) for i in range(1, N)]
    gates += [([i + 1], expm(-1j * h[i] * Z * 0.37)) for i in range(N)]
    psi = (rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N))
    psi /= np.linalg.norm(psi)
    got = orc.trotter_step(psi.copy(), gates, N)
    want = expm(-1j * H * 0.37) @ psi
    assert np.linalg.norm(got - want) < 1e-13


def test_second_order_error_is_order_dt_squared():
    N, J, delta = 6, 1.0, 0.5
    h = np.random.default_rng(1).uniform(-1, 1, N)
    H = xxz_dense(N, J, delta, h)
    psi0 = neel(N)
    T = 0.4
    errs = []
    for steps in (8, 16):
        dt = T / steps
        half = xxz_gates(N, J, delta, h, dt / 2)
        seq = half + half[::-1]
        psi = psi0.copy()
        for _ in range(steps):
            orc.trotter_step(psi, seq, N)
        errs.append(np.linalg.norm(psi - expm(-1j * H * T) @ psi0))
    ratio = errs[0] / errs[1]
    assert 3.5 < ratio < 4.5, (errs, ratio)


def test_oracle_chebyshev_matches_expm_multiply():
    N, J, delta = 8, 1.0, 0.5
    h = np.random.default_rng(2).uniform(-1, 1, N)
    bonds = [(i, i + 1, J, J, J * delta) for i in range(1, N)]
    fields = np.array([[0.0, 0.0, h[i]] for i in range(N)])
    B = sum(abs(b[2]) + abs(b[3]) + abs(b[4]) for b in bonds) + float(np.sum(np.abs(h)))
    psi0 = neel(N)
    got, m = orc.chebyshev_evolve(psi0, N, bonds, fields, 2.0, -B, B, tol=1e-14)
    want = expm_multiply(-1j * 2.0 * xxz_dense(N, J, delta, h), psi0)
    assert np.linalg.norm(got - want) < 1e-11
    assert m > 10


# ------------------------------------------------------------------ MCWF --
SM = np.array([[0, 1], [0, 0]], np.complex128)  # σ⁻ = |0><1|


def dense_decay_jump(psi, N, gdt, r):
    """One decay/decision/jump/renormalise of the batched MCWF step with
    dense operators: exp(-γdt/2 · n_exc) as a diagonal, σ⁻ as a matrix."""
    idx = np.arange(1 << N)
    pc = np.array([bin(i).count("1") for i in idx])
    psi = psi * np.sqrt(1.0 - gdt) ** pc
    nsq = np.vdot(psi, psi).real
    dp = 1.0 - nsq
    site = 0
    if r < dp and dp > 0:
        pops = np.array([np.vdot(psi, op_on({k: SM.conj().T @ SM}, N) @ psi).real
                         for k in range(N)])
        hits = np.nonzero(r < np.cumsum(gdt * pops / nsq))[0]
        k = int(hits[0]) if hits.size else 0
        psi = op_on({k: SM}, N) @ psi
        site = k + 1
    return psi / np.linalg.norm(psi), site, dp


def test_mcwf_decay_jump_known_cases():
    """|11>, γdt = 0.2: dp = 1 - 0.8² = 0.36, cumulative site weights
    0.2 and 0.4; r picks no jump / site 1 / site 2 / (none reachable)."""
    dec = np.sqrt(0.8) ** np.arange(3)
    for r, want_site, want_idx in [(0.5, 0, 3), (0.1, 1, 2), (0.3, 2, 1)]:
        psi = np.zeros(4, np.complex128)
        psi[3] = 1.0
        site, dp = orc.mcwf_decay_jump(psi, 2, 0.2, dec, r)
        assert site == want_site and abs(dp - 0.36) < 1e-15
        assert abs(psi[want_idx] - 1.0) < 1e-15 and np.sum(np.abs(psi) ** 2) == pytest.approx(1)


def test_mcwf_decay_jump_matches_dense_operators():
    rng = np.random.default_rng(8)
    N, gdt = 4, 0.3
    dec = np.sqrt(1.0 - gdt) ** np.arange(N + 1)
    sites = set()
    for _ in range(200):
        psi = rng.normal(size=16) + 1j * rng.normal(size=16)
        psi /= np.linalg.norm(psi)
        r = rng.uniform(0, 0.5)
        want, wsite, wdp = dense_decay_jump(psi.copy(), N, gdt, r)
        got = psi.copy()
        site, dp = orc.mcwf_decay_jump(got, N, gdt, dec, r)
        assert site == wsite and abs(dp - wdp) < 1e-14
        assert np.max(np.abs(got - want)) < 1e-13
        sites.add(site)
    assert sites == {0, 1, 2, 3, 4}


def test_mcwf_fallback_site():
    """(|0> + |1>)/√2, γdt = 0.5: dp = 0.25 but the site weight is only
    γdt·pop/||ψ||² = 1/6, so r = 0.2 jumps through the fallback (site 1)."""
    dec = np.sqrt(0.5) ** np.arange(2)
    for r in (0.1, 0.2):
        psi = np.array([1.0, 1.0], np.complex128) / np.sqrt(2.0)
        site, dp = orc.mcwf_decay_jump(psi, 1, 0.5, dec, r)
        assert site == 1 and abs(dp - 0.25) < 1e-15
        assert np.max(np.abs(psi - [1.0, 0.0])) < 1e-15


def test_mcwf_trajectory_matches_dense_step():
    """Ten-step trajectories of the oracle (gate list + decay/jump) against
    dense matrices: bond diagonals, site rotations, decay, σ⁻."""
    N, gamma_dt = 3, 0.25
    rng = np.random.default_rng(2)
    dt = 0.1
    gates = [([1, 2], expm(-1j * 0.7 * np.kron(Z, Z) * dt)),
             ([2, 3], expm(-1j * 1.3 * np.kron(Z, Z) * dt))]
    gates += [([k + 1], expm(-1j * (0.8 * X + 0.3 * Z) * dt)) for k in range(N)]
    # dense step: the 2-qubit gates are diagonal, applied as one diagonal
    idx = np.arange(8)
    diag = np.ones(8, np.complex128)
    for q, G in gates[:2]:
        e = ((idx >> (q[0] - 1)) & 1) | (((idx >> (q[1] - 1)) & 1) << 1)
        diag *= np.diag(G)[e]
    U = np.diag(diag)
    for q, G in gates[2:]:
        U = op_on({q[0] - 1: G}, N) @ U
    dec = np.sqrt(1.0 - gamma_dt) ** np.arange(N + 1)
    for traj in range(20):
        psi0 = rng.normal(size=8) + 1j * rng.normal(size=8)
        psi0 /= np.linalg.norm(psi0)
        rs = rng.uniform(0, 1, 10) * 0.6
        got = psi0.copy()
        jumps, _ = orc.mcwf_trajectory(got, gates, N, gamma_dt, dec, rs)
        ref = psi0.copy()
        want = []
        for r in rs:
            ref, site, _ = dense_decay_jump(U @ ref, N, gamma_dt, r)
            want.append(site - 1)
        np.testing.assert_array_equal(jumps, want)
        assert np.max(np.abs(got - ref)) < 1e-12


def test_mcwf_dephasing_step_matches_dense_operators():
    """lindblad_mcwf_step_trotter!'s dissipative half with σ_z jumps: the
    uniform decay sqrt(1 - Nγdt), the decision, σ_z on site floor(u2·N) as a
    dense matrix, renormalisation."""
    rng = np.random.default_rng(4)
    N, gdt = 3, 0.2
    decay = np.sqrt(1.0 - N * gdt)
    sites = set()
    for _ in range(100):
        psi = rng.normal(size=8) + 1j * rng.normal(size=8)
        psi /= np.linalg.norm(psi)
        r, u2 = rng.uniform(0, 1), rng.uniform(0, 1)
        got = psi.copy()
        site, dp = orc.mcwf_dephasing(got, N, decay, r, u2)
        want = psi * decay
        wdp = 1.0 - np.vdot(want, want).real
        wsite = 0
        if r < wdp:
            k = min(N - 1, int(u2 * N))
            want = op_on({k: Z}, N) @ want
            wsite = k + 1
        want /= np.linalg.norm(want)
        assert site == wsite and abs(dp - wdp) < 1e-15
        assert np.max(np.abs(got - want)) < 1e-14
        sites.add(site)
    assert sites == {0, 1, 2, 3}


@pytest.mark.parametrize("N,k_iter", [(6, 40), (3, 30)])
def test_oracle_lanczos_matches_dense_spectrum(N, k_iter):
    """The oracle's restatement of estimate_spectral_range_lanczos: extreme
    Ritz values equal the dense extreme eigenvalues (XXZ chain with z fields);
    with 2^N < k_iter the recurrence stops on beta < 1e-12."""
    rng = np.random.default_rng(3)
    h = rng.uniform(-1, 1, N)
    bonds = [(i, i + 1, 1.0, 1.0, 0.5) for i in range(1, N)]
    fields = np.array([[0.0, 0.0, hi] for hi in h])
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    lo, hi, steps = orc.lanczos_range(v, N, bonds, fields, k_iter)
    ev = np.linalg.eigvalsh(xxz_dense(N, 1.0, 0.5, h))
    assert abs(lo - ev[0]) < 1e-9 and abs(hi - ev[-1]) < 1e-9
    if (1 << N) < k_iter:
        assert steps <= (1 << N) + 1
