"""Pins the oracle's noise-channel and measurement restatements (CPU only).

Density-matrix channels against textbook Kraus sums sum_k K rho K^dagger;
the state-vector trajectory branches against those channels by exact
ensemble averages over the uniform (each branch weighted by the length of
its interval of u); the amplitude-damping branch's additive jump against
hand-computed vectors; measurement against the Born rule, idempotence and
the dephasing identity sum_b P_b rho P_b; the reference demo's own asserted
values (scripts_quantum_simulations/demo_quantum_channels.jl:259-368)."""
import numpy as np
import pytest

from oracle import oracle as orc

I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], np.complex128)
Y = np.array([[0, -1j], [1j, 0]], np.complex128)
Z = np.diag([1.0, -1.0]).astype(np.complex128)
H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)


def embed(K, k, N):
    """K on qubit k (1-based, bit k-1) of N qubits."""
    out = np.eye(1, dtype=np.complex128)
    for q in range(N, 0, -1):  # kron puts the highest qubit first
        out = np.kron(out, K if q == k else I2)
    return out


def kraus_ops(kind, p):
    if kind == "dephasing":
        return [np.diag([1.0, np.sqrt(1 - p)]), np.diag([0.0, np.sqrt(p)])]
    if kind == "depolarizing":
        return [np.sqrt(1 - p) * I2] + [np.sqrt(p / 3) * P for P in (X, Y, Z)]
    if kind == "amplitude_damping":
        return [np.diag([1.0, np.sqrt(1 - p)]), np.array([[0, np.sqrt(p)], [0, 0]])]
    if kind == "bit_flip":
        return [np.sqrt(1 - p) * I2, np.sqrt(p) * X]
    if kind == "phase_flip":
        return [np.sqrt(1 - p) * I2, np.sqrt(p) * Z]
    raise ValueError(kind)


def rand_psi(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return v / np.linalg.norm(v)


def rand_rho(N, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(1 << N, 1 << N)) + 1j * rng.normal(size=(1 << N, 1 << N))
    r = A @ A.conj().T
    return r / np.trace(r).real


KINDS = ["dephasing", "depolarizing", "amplitude_damping", "bit_flip", "phase_flip"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("p", [0.0, 0.13, 0.5, 1.0])
def test_dm_channel_is_the_kraus_sum(kind, p):
    N = 3
    for k in (1, 2, 3):
        rho = rand_rho(N, 7 + k)
        want = sum(embed(K, k, N) @ rho @ embed(K, k, N).conj().T for K in kraus_ops(kind, p))
        got = orc.dm_channel(rho.copy(), kind, p, k, N)
        assert np.abs(got - want).max() < 1e-14
        assert abs(np.trace(got).real - 1.0) < 1e-14


def _branches(kind, p, pop1=None):
    """(u, weight) covering [0, 1) by the branch intervals of the decision."""
    if kind == "dephasing":
        q = (1 - np.sqrt(1 - p)) / 2
        cuts = [q]
    elif kind == "depolarizing":
        cuts = [p / 3, 2 * p / 3, p]
    elif kind in ("bit_flip", "phase_flip"):
        cuts = [p]
    else:
        cuts = [p * pop1]
    edges = [0.0] + cuts + [1.0]
    return [((a + b) / 2, b - a) for a, b in zip(edges[:-1], edges[1:]) if b > a]


@pytest.mark.parametrize("kind", ["dephasing", "depolarizing", "bit_flip", "phase_flip"])
@pytest.mark.parametrize("p", [0.07, 0.4, 0.9])
def test_trajectory_average_is_the_channel(kind, p):
    N = 3
    for k in (1, 3):
        psi = rand_psi(N, 11 + k)
        rho = np.outer(psi, psi.conj())
        avg = np.zeros_like(rho)
        for u, w in _branches(kind, p):
            v = psi.copy()
            orc.psi_channel(v, kind, p, k, N, u)
            avg += w * np.outer(v, v.conj())
        want = orc.dm_channel(rho.copy(), kind, p, k, N)
        assert np.abs(avg - want).max() < 1e-14


@pytest.mark.parametrize("gamma", [0.1, 0.4, 0.95])
def test_amplitude_damping_average_on_one_qubit(gamma):
    # on one qubit the jump branch is |0> up to a phase, so the average of the
    # two renormalised branches is exactly the channel
    psi = np.array([0.6, 0.8j * np.exp(0.3j)], np.complex128)
    pop1 = abs(psi[1]) ** 2
    avg = np.zeros((2, 2), np.complex128)
    for u, w in _branches("amplitude_damping", gamma, pop1):
        v = psi.copy()
        orc.psi_channel(v, "amplitude_damping", gamma, 1, 1, u)
        avg += w * np.outer(v, v.conj())
    want = orc.dm_channel(np.outer(psi, psi.conj()), "amplitude_damping", gamma, 1, 1)
    assert np.abs(avg - want).max() < 1e-14


def test_amplitude_damping_jump_adds_amplitudes():
    # Bell-like (|00> + |11>)/sqrt2 (indices 0 and 3); a jump on qubit 1 adds
    # psi[1] onto psi[0] and psi[3] onto psi[2]: (|00> + |01>... ) = indices 0, 2
    psi = np.array([1, 0, 0, 1], np.complex128) / np.sqrt(2)
    v = psi.copy()
    assert orc.psi_channel(v, "amplitude_damping", 0.8, 1, 2, 0.1) == 4
    assert np.allclose(v, np.array([1, 0, 1, 0]) / np.sqrt(2), atol=1e-15)
    # |+>: the jump sums to 2/sqrt2 on |0>, renormalised to |0>
    v = np.array([1, 1], np.complex128) / np.sqrt(2)
    assert orc.psi_channel(v, "amplitude_damping", 0.8, 1, 1, 0.3) == 4
    assert np.allclose(v, [1, 0], atol=1e-15)
    # no jump: (1, sqrt(1-g)) / sqrt(2 - g)
    v = np.array([1, 1], np.complex128) / np.sqrt(2)
    assert orc.psi_channel(v, "amplitude_damping", 0.8, 1, 1, 0.5) == 0
    assert np.allclose(v, np.array([1, np.sqrt(0.2)]) / np.sqrt(1.2), atol=1e-15)
    # |0> never jumps and is left alone
    v = np.array([1, 0], np.complex128)
    assert orc.psi_channel(v, "amplitude_damping", 1.0, 1, 1, 0.0) == 0
    assert np.array_equal(v, [1, 0])


def test_demo_asserted_values():
    """The assertions of demo_quantum_channels.jl TEST 1.1-1.5 (:259-368)."""
    N = 2
    plus = orc.product_ket([(1 / np.sqrt(2), 1 / np.sqrt(2))] * N)
    rho = np.outer(plus, plus.conj())
    d0, c0 = np.diag(rho).copy(), abs(rho[0, 1])
    orc.dm_channel(rho, "dephasing", 0.5, 1, N)
    assert np.linalg.norm(np.diag(rho) - d0) < 1e-12
    assert abs(abs(rho[0, 1]) / c0 - np.sqrt(0.5)) < 1e-10
    assert abs(np.trace(rho) - 1) < 1e-12
    rho = np.outer(plus, plus.conj())
    for k in (1, 2):
        orc.dm_channel(rho, "dephasing", 1.0, k, N)
    assert np.abs(rho).sum() - np.abs(np.diag(rho)).sum() < 1e-10
    one = np.diag([0.0, 1.0]).astype(np.complex128)
    r = orc.dm_channel(one.copy(), "amplitude_damping", 0.5, 1, 1)
    assert abs(r[0, 0] - 0.5) < 1e-12 and abs(r[1, 1] - 0.5) < 1e-12
    zero = np.diag([1.0, 0.0]).astype(np.complex128)
    assert np.linalg.norm(orc.dm_channel(zero.copy(), "amplitude_damping", 0.9, 1, 1) - zero) < 1e-12
    r = orc.dm_channel(zero.copy(), "bit_flip", 0.5, 1, 1)
    assert abs(r[0, 0] - 0.5) < 1e-12 and abs(r[1, 1] - 0.5) < 1e-12
    r = orc.dm_channel(zero.copy(), "bit_flip", 1.0, 1, 1)
    assert abs(r[0, 0]) < 1e-12 and abs(r[1, 1] - 1) < 1e-12
    r = orc.dm_channel(np.full((2, 2), 0.5, np.complex128), "phase_flip", 0.25, 1, 1)
    assert abs(abs(r[0, 1]) - 0.5 * 0.5) < 1e-12
    assert abs(r[0, 0] - 0.5) < 1e-12 and abs(r[1, 1] - 0.5) < 1e-12
    r = orc.dm_channel(zero.copy(), "depolarizing", 0.75, 1, 1)
    assert r[0, 0].real < 1 and r[1, 1].real > 0 and abs(np.trace(r) - 1) < 1e-12
    assert np.allclose(r, np.eye(2) / 2, atol=1e-15)  # p = 3/4 is fully depolarizing


def test_measure_z_born_rule_and_idempotence():
    N = 4
    psi = rand_psi(N, 3)
    for k in range(1, N + 1):
        lo, hi = orc._halves(N, k)
        p0 = float(np.sum(np.abs(psi[lo]) ** 2))
        for u, want in ((p0 * 0.999, 0), (p0 * 1.001 + 1e-9, 1)):
            v = psi.copy()
            o = orc.measure_z(v, k, N, u)
            assert o == want
            keep = lo if o == 0 else hi
            ref = np.zeros_like(psi)
            ref[keep] = psi[keep] / np.sqrt(p0 if o == 0 else 1 - p0)
            assert np.abs(v - ref).max() < 1e-15
            w = v.copy()
            assert orc.measure_z(w, k, N, 0.999999 if o == 0 else 0.0) == o
            assert np.abs(w - v).max() < 1e-15


def test_measure_ensemble_is_z_dephasing_and_matches_dm():
    N = 3
    psi = rand_psi(N, 5)
    rho = np.outer(psi, psi.conj())
    for k in range(1, N + 1):
        lo, hi = orc._halves(N, k)
        p0 = float(np.sum(np.abs(psi[lo]) ** 2))
        avg = np.zeros_like(rho)
        for u, w in ((p0 / 2, p0), ((1 + p0) / 2, 1 - p0)):
            v = psi.copy()
            o = orc.measure_z(v, k, N, u)
            avg += w * np.outer(v, v.conj())
            r = rho.copy()
            assert orc.dm_measure_z(r, k, N, u) == o
            assert np.abs(r - np.outer(v, v.conj())).max() < 1e-14
        P0 = embed(np.diag([1.0, 0.0]), k, N)
        P1 = embed(np.diag([0.0, 1.0]), k, N)
        assert np.abs(avg - (P0 @ rho @ P0 + P1 @ rho @ P1)).max() < 1e-14


def test_force_and_reset_known_states():
    # Bell state (|00> + |11>)/sqrt2: force qubit 1 to |0> adds psi[1] onto
    # psi[0] and psi[3] onto psi[2]
    bell = np.array([1, 0, 0, 1], np.complex128) / np.sqrt(2)
    v = orc.force_qubit(bell.copy(), 1, "zero", 2)
    assert np.allclose(v, np.array([1, 0, 1, 0]) / np.sqrt(2), atol=1e-15)
    v = orc.force_qubit(bell.copy(), 1, "one", 2)
    assert np.allclose(v, np.array([0, 1, 0, 1]) / np.sqrt(2), atol=1e-15)
    # reset after a measurement: outcome 1 (u = 0.9), then |0>, then |+>
    v = orc.reset_state(bell.copy(), [1], ["zero"], 2, [0.9])
    assert np.allclose(v, [0, 0, 1, 0], atol=1e-15)
    v = orc.reset_state(bell.copy(), [1], ["plus"], 2, [0.1])
    assert np.allclose(v, np.array([1, 1, 0, 0]) / np.sqrt(2), atol=1e-15)
    v = orc.reset_state(bell.copy(), [2], ["minus"], 2, [0.1])
    assert np.allclose(v, np.array([1, 0, -1, 0]) / np.sqrt(2), atol=1e-15)


def test_dm_force_reference_docstring_example():
    """_force_qubit_rho! docstring (…Measurements.jl:665-684): Bell |Phi+><Phi+|
    forced to |0> on qubit 1 is |00><00|."""
    bell = np.array([1, 0, 0, 1], np.complex128) / np.sqrt(2)
    r = orc.dm_force(np.outer(bell, bell.conj()), 1, 0, 2)
    want = np.zeros((4, 4))
    want[0, 0] = 1
    assert np.abs(r - want).max() < 1e-15
    rho = rand_rho(3, 9)
    for k in (1, 2, 3):
        for b in (0, 1):
            P = embed(np.diag([1.0 - b, float(b)]), k, 3)
            want = P @ rho @ P
            want /= np.trace(want).real
            assert np.abs(orc.dm_force(rho.copy(), k, b, 3) - want).max() < 1e-14


def test_sample_from_probabilities_brute_force():
    rng = np.random.default_rng(0)
    p = rng.random(37)
    p /= p.sum()
    c = np.cumsum(p)
    for r in list(rng.random(200)) + [0.0, 1.0, 1.5]:
        want = int(np.searchsorted(c, r, side="left"))
        got = orc.sample_from_probabilities(p, r)
        # the running sum of the loop equals numpy's sequential cumsum
        assert got == min(want, len(p) - 1)
    psi = rand_psi(5, 2)
    for r in rng.random(50):
        assert orc.sample_from_probabilities(np.abs(psi) ** 2, r) == \
            int(orc.sample_sequential(psi, 5, r)[0])
