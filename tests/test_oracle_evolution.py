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
