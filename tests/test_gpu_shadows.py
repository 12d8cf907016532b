"""Classical-shadow snapshots on the device (smoq_shadow_snapshots) against
the oracle's restatement of classical_shadow_snapshot / clifford_shadow_snapshot
on the same bases and uniforms: Pauli outcomes bit-exact (the device rotates
with the reference's arithmetic), local-Clifford outcomes equal (the
reference multiplies by the dense Kronecker product, so only a draw within
rounding of a cumulative boundary could differ), and a product state at
N = 22 whose Z-measured qubits have known outcomes, across several chunks of
rotated states."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rand_psi(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("N,M,dtype", [(1, 200, "c128"), (3, 5000, "c128"), (6, 2000, "c128"),
                                       (10, 300, "c128"), (14, 60, "c128"), (6, 2000, "c64")])
def test_pauli_snapshots_bit_exact(ctx, N, M, dtype):
    from smoq_amd import GPUState
    from smoq_amd.shadows import collect_shadows

    print(f"[case] pauli N={N} M={M} {dtype}", flush=True)
    psi = rand_psi(N, N + M)
    if dtype == "c64":
        psi = psi.astype(np.complex64).astype(np.complex128)
    s = GPUState.from_numpy(psi.astype(np.complex64) if dtype == "c64" else psi)
    rng = np.random.default_rng(M)
    bases = rng.integers(1, 4, (M, N))
    u = rng.random(M)
    sh = collect_shadows(s, N, M, bases=bases, uniforms=u)
    assert np.array_equal(sh.codes, bases)
    for m in range(M):
        i = orc.pauli_shadow_snapshot(psi, N, bases[m], u[m])
        assert list(sh.outcomes[m]) == [(i >> k) & 1 for k in range(N)], m
    snap = sh[0]
    assert list(snap.bases) == list(bases[0])
    s.free()


def test_walk_path_many_draws_per_string(ctx):
    """From 2^16 amplitudes one wave walks each rotated state's prefix and
    answers its draws in ascending order: 400 snapshots over 12 strings."""
    from smoq_amd import GPUState
    from smoq_amd.shadows import collect_shadows

    N, M = 16, 400
    psi = rand_psi(N, 16)
    s = GPUState.from_numpy(psi)
    rng = np.random.default_rng(16)
    strings = rng.integers(1, 4, (12, N))
    bases = strings[rng.integers(0, 12, M)]
    u = rng.random(M)
    u[:3] = [0.0, 1.0 - 2.0 ** -53, u[3]]  # both ends and a repeated uniform
    sh = collect_shadows(s, N, M, bases=bases, uniforms=u)
    for m in range(M):
        i = orc.pauli_shadow_snapshot(psi, N, bases[m], u[m])
        assert list(sh.outcomes[m]) == [(i >> k) & 1 for k in range(N)], m
    s.free()


@pytest.mark.parametrize("N,M", [(2, 3000), (5, 800)])
def test_local_clifford_snapshots(ctx, N, M):
    from smoq_amd import GPUState
    from smoq_amd.shadows import SINGLE_QUBIT_CLIFFORDS, collect_clifford_shadows

    psi = rand_psi(N, 3 * N)
    s = GPUState.from_numpy(psi)
    rng = np.random.default_rng(N)
    idx = rng.integers(1, 25, (M, N))
    u = rng.random(M)
    sh = collect_clifford_shadows(s, N, M, indices=idx, uniforms=u)
    for m in range(M):
        i = orc.clifford_shadow_snapshot(psi, N, idx[m], u[m], SINGLE_QUBIT_CLIFFORDS)
        assert list(sh.outcomes[m]) == [(i >> k) & 1 for k in range(N)], m
    s.free()


def test_chunks_of_large_states(ctx):
    """N = 22: 100 distinct strings of 96 MiB rotated states and prefixes run
    in two chunks; qubits measured in Z of |0...0> read 0, X and Y outcomes
    are fair coins; three snapshots against the oracle."""
    from smoq_amd import GPUState
    from smoq_amd.shadows import collect_shadows

    N, M = 22, 100
    s = GPUState.basis(N, 0, ctx=ctx)
    rng = np.random.default_rng(22)
    bases = rng.integers(1, 4, (M, N))
    u = rng.random(M)
    sh = collect_shadows(s, N, M, bases=bases, uniforms=u)
    z = bases == 3
    assert np.all(sh.outcomes[z] == 0)
    ones = sh.outcomes[~z].mean()
    assert abs(ones - 0.5) < 5 * np.sqrt(0.25 / (~z).sum())
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1.0
    for m in (0, 57, 99):
        i = orc.pauli_shadow_snapshot(psi, N, bases[m], u[m])
        assert list(sh.outcomes[m]) == [(i >> k) & 1 for k in range(N)]
    s.free()


def test_arguments_are_checked(ctx):
    from smoq_amd import GPUState, SmoqError
    from smoq_amd._lib import SMOQ_EINVAL
    from smoq_amd.shadows import SHADOW_PAULI, ShadowConfig, collect_shadows, shadow_outcomes

    s = GPUState.basis(3, 0, ctx=ctx)
    with pytest.raises(SmoqError) as e:
        shadow_outcomes(s, np.full((2, 3), 4), [0.1, 0.2], SHADOW_PAULI)
    assert e.value.code == SMOQ_EINVAL
    assert len(collect_shadows(s, ShadowConfig(3, 10, "global_clifford"),
                               rng=np.random.default_rng(2))) == 10
    with pytest.raises(ValueError):
        ShadowConfig(3, 10, "bogus")
    sh = collect_shadows(s, ShadowConfig(3, 10, "clifford_group"), rng=np.random.default_rng(1))
    assert sh.kind == "local_clifford" and len(sh) == 10
    s.free()


@pytest.mark.parametrize("group", ["pauli", "local_clifford"])
def test_density_matrix_snapshots(ctx, group):
    from smoq_amd.density import DensityMatrix
    from smoq_amd.shadows import SINGLE_QUBIT_CLIFFORDS, ShadowConfig, collect_shadows

    N, M = 3, 600
    rng = np.random.default_rng(5)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = A @ A.conj().T
    rho /= np.trace(rho).real
    d = DensityMatrix.from_numpy(rho)
    hi = 4 if group == "pauli" else 25
    bases = rng.integers(1, hi, (M, N))
    u = rng.random(M)
    sh = collect_shadows(d, ShadowConfig(N, M, group), bases=bases, uniforms=u)
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    rot = {1: H, 2: H @ np.diag([1, -1j]), 3: np.eye(2)}
    for m in range(M):
        mats = [rot[int(c)] if group == "pauli" else SINGLE_QUBIT_CLIFFORDS[int(c) - 1]
                for c in bases[m]]
        i = orc.rho_shadow_snapshot(rho, N, mats, u[m])
        assert list(sh.outcomes[m]) == [(i >> k) & 1 for k in range(N)], m
    d.free()


@pytest.mark.parametrize("N,M", [(1, 200), (3, 400), (6, 300), (8, 10)])
def test_global_clifford_snapshots(ctx, N, M):
    from smoq_amd import GPUState
    from smoq_amd.shadows import collect_global_clifford_shadows

    psi = rand_psi(N, 5 * N)
    s = GPUState.from_numpy(psi)
    rng = np.random.default_rng(N)
    drawn = collect_global_clifford_shadows(s, N, M, rng=rng)  # the vectorised generator
    circuits = [sn.circuit for sn in drawn]
    for c in circuits[:5]:  # its circuits have sample_random_global_clifford's shape
        assert sum(g[0] == "CNOT" for g in c.gates) == (2 * N - 1) * (N // 2)
    u = np.random.default_rng(99).random(M)
    snaps = collect_global_clifford_shadows(s, N, M, circuits=circuits, uniforms=u)
    for m in range(M):
        i = orc.global_clifford_snapshot(psi, N, circuits[m].gates, u[m])
        assert list(snaps[m].outcomes) == [(i >> k) & 1 for k in range(N)], m
    s.free()


def test_global_clifford_density_matrix(ctx):
    from smoq_amd.density import DensityMatrix
    from smoq_amd.shadows import ShadowConfig, collect_shadows

    N, M = 3, 60
    rng = np.random.default_rng(8)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = A @ A.conj().T
    rho /= np.trace(rho).real
    d = DensityMatrix.from_numpy(rho)
    u = rng.random(M)
    snaps = collect_shadows(d, ShadowConfig(N, M, "global_clifford"), uniforms=u,
                            rng=np.random.default_rng(9))
    for m, sn in enumerate(snaps):
        i = orc.global_clifford_snapshot(rho, N, sn.circuit.gates, u[m])
        assert list(sn.outcomes) == [(i >> k) & 1 for k in range(N)], m
    d.free()
