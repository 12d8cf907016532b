"""Pins the oracle's classical-shadow snapshots (CPU only): the outcome as a
function of the uniform is the inverse CDF of the dense Born probabilities
|<s| U |ψ>|^2 with U the Kronecker product of the per-qubit rotations
(checked at the midpoints between consecutive cumulative sums), eigenstates
of the measured Paulis give their fixed outcome, and the Clifford table
built like the reference's SINGLE_QUBIT_CLIFFORDS has 24 distinct elements
that map Paulis to signed Paulis."""
import numpy as np
import pytest

from oracle import oracle as orc
from smoq_amd.shadows import SINGLE_QUBIT_CLIFFORDS as CL

H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)
SDG = np.diag([1, -1j])
ROT = {1: H, 2: H @ SDG, 3: np.eye(2)}
X = np.array([[0, 1], [1, 0]], np.complex128)
Y = np.array([[0, -1j], [1j, 0]], np.complex128)
Z = np.diag([1.0, -1.0]).astype(np.complex128)


def kron_rot(mats):
    U = np.array(mats[-1])
    for m in reversed(mats[:-1]):
        U = np.kron(U, m)
    return U


def rand_psi(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_pauli_snapshot_is_the_inverse_cdf(N):
    rng = np.random.default_rng(N)
    psi = rand_psi(N, N)
    for _ in range(6):
        bases = rng.integers(1, 4, N)
        p = np.abs(kron_rot([ROT[int(b)] for b in bases]) @ psi) ** 2
        c = np.cumsum(p)
        mids = np.concatenate([[c[0] / 2], (c[:-1] + c[1:]) / 2])
        for i, r in enumerate(mids):
            if p[i] < 1e-9:
                continue
            assert orc.pauli_shadow_snapshot(psi, N, bases, r) == i


def test_pauli_eigenstates():
    plus = np.array([1, 1], np.complex128) / np.sqrt(2)
    plus_i = np.array([1, 1j], np.complex128) / np.sqrt(2)
    minus = np.array([1, -1], np.complex128) / np.sqrt(2)
    # r > 0: at r = 0 the rule r <= cumsum picks index 0 even when P(0) = 0,
    # as the reference's does
    for r in (0.1, 0.3, 0.999):
        assert orc.pauli_shadow_snapshot(plus, 1, [1], r) == 0
        assert orc.pauli_shadow_snapshot(minus, 1, [1], r) == 1
        assert orc.pauli_shadow_snapshot(plus_i, 1, [2], r) == 0
        assert orc.pauli_shadow_snapshot(np.array([0, 1], np.complex128), 1, [3], r) == 1
    # |+> (x) |1> (qubit 1 = |+>), bases X then Z: outcome bits (0, 1) = index 2
    psi = np.kron(np.array([0, 1]), plus).astype(np.complex128)
    assert orc.pauli_shadow_snapshot(psi, 2, [1, 3], 0.5) == 2


def test_clifford_table():
    assert CL.shape == (24, 2, 2)
    for i in range(24):
        assert np.abs(CL[i].conj().T @ CL[i] - np.eye(2)).max() < 1e-9
        for j in range(i):
            # distinct up to a global phase
            ov = abs(np.trace(CL[i].conj().T @ CL[j])) / 2
            assert ov < 1 - 1e-6
        for P in (X, Y, Z):
            Q = CL[i] @ P @ CL[i].conj().T
            assert min(np.abs(Q - s * R).max() for R in (X, Y, Z) for s in (1, -1)) < 1e-9


def test_clifford_snapshot_inverse_cdf():
    N = 3
    rng = np.random.default_rng(7)
    psi = rand_psi(N, 7)
    for _ in range(5):
        idx = rng.integers(1, 25, N)
        U = kron_rot([CL[i - 1] for i in idx])
        p = np.abs(U @ psi) ** 2
        c = np.cumsum(p)
        mids = np.concatenate([[c[0] / 2], (c[:-1] + c[1:]) / 2])
        for i, r in enumerate(mids):
            if p[i] < 1e-9:
                continue
            assert orc.clifford_shadow_snapshot(psi, N, idx, r, CL) == i


def test_rho_snapshot_matches_pure_state_form():
    """On |ψ><ψ| the ρ form draws what the ψ form draws (away from ties)."""
    N = 3
    psi = rand_psi(N, 11)
    rho = np.outer(psi, psi.conj())
    rng = np.random.default_rng(3)
    for _ in range(40):
        bases = rng.integers(1, 4, N)
        r = rng.random()
        a = orc.pauli_shadow_snapshot(psi, N, bases, r)
        b = orc.rho_shadow_snapshot(rho, N, [ROT[int(x)] for x in bases], r)
        assert a == b


def test_global_clifford_unitary():
    """The reference's unitary: H/S on qubit q is the q-th Kronecker factor
    from the left (bit N - q), CNOT acts on bits c-1 -> t-1; the product maps
    Paulis to signed Paulis (a Clifford)."""
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    U = orc.global_clifford_unitary(2, [("H", 1)])
    assert np.allclose(U, np.kron(H, np.eye(2)))  # acts on bit 1 = qubit 2's position
    U = orc.global_clifford_unitary(2, [("CNOT", 1, 2)])
    assert np.allclose(U @ np.array([0, 1, 0, 0]), [0, 0, 0, 1])  # bit 0 set flips bit 1
    rng = np.random.default_rng(4)
    gates = []
    for _ in range(12):
        r = rng.integers(0, 3)
        if r == 2:
            c, t = rng.choice(3, 2, replace=False) + 1
            gates.append(("CNOT", int(c), int(t)))
        else:
            gates.append(("H" if r == 0 else "S", int(rng.integers(1, 4))))
    U = orc.global_clifford_unitary(3, gates)
    assert np.abs(U.conj().T @ U - np.eye(8)).max() < 1e-12
    for P in (X, Y, Z):
        Q = U @ np.kron(np.eye(4), P) @ U.conj().T
        # a signed Pauli string has one nonzero of modulus 1 per row
        assert np.allclose(np.sort(np.abs(Q), axis=1)[:, -1], 1.0)
        assert np.allclose(np.sum(np.abs(Q) > 1e-9, axis=1), 1)
