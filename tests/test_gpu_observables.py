"""Observable and sampling parity on the device against the CPU oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rand_state(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return (v / np.linalg.norm(v)).astype(np.complex128)


def test_local_and_corr_parity(ctx):
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    for N in (2, 6, 13):
        psi = rand_state(N, N)
        s = GPUState.from_numpy(psi)
        for k in range(1, N + 1):
            for p in "xyz":
                assert abs(ob.expect_local(s, k, N, p) - orc.expect_local(psi, k, N, p)) < 1e-12
        for i in range(1, N):
            for pair in ("zz", "xx", "yy"):
                j = (i % N) + 1
                a = ob.expect_corr(s, i, j, N, pair)
                b = orc.expect_corr(psi, i, j, N, pair)
                assert abs(a - b) < 1e-12


def test_pauli_strings_true_and_reference(ctx):
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    N = 9
    psi = rand_state(N, 4)
    s = GPUState.from_numpy(psi)
    rng = np.random.default_rng(2)
    xs, zs, want = [], [], []
    for _ in range(60):
        k = int(rng.integers(1, 5))
        qs = [int(x) + 1 for x in rng.choice(N, k, replace=False)]
        ops = [str(rng.choice(list("XYZ"))) for _ in range(k)]
        x, z = ob.pauli_masks(ops, qs)
        xs.append(x)
        zs.append(z)
        want.append(np.vdot(psi, orc.pauli_matrix(x, z, N) @ psi).real)
        ref = orc.pauli_string_ref(psi, ops, qs, N)
        got = ob.get_expectation_pauli_string(s, ops, qs, N)
        assert abs(got - ref) < 1e-12, (ops, qs)
    got = ob.expect_pauli(s, xs, zs)
    assert np.max(np.abs(got - np.array(want))) < 1e-12


def test_compute_energy_parity(ctx):
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    N = 10
    psi = rand_state(N, 5)
    s = GPUState.from_numpy(psi)
    for (Jx, Jy, Jz, hx) in [(1, 1, 1, 0.5), (0, 0, 1, 1), (-1, 0.3, 0.2, -0.7)]:
        a = ob.compute_energy(s, N, Jx, Jy, Jz, hx)
        b = orc.compute_energy(psi, N, Jx, Jy, Jz, hx)
        assert abs(a - b) < 1e-11 * max(1.0, abs(b))


def test_layered_circuit_z_and_zz(ctx):
    # Plain circuit at N=20: 10 layers of Ry, Rz on all qubits + CZ bricks;
    # outputs <Z_i> and <Z_i Z_{i+1}> against the oracle's gate-by-gate run.
    from smoq_amd import GPUState
    from smoq_amd import observables as ob
    from smoq_amd.circuits import layered_ry_rz_cz

    N = 20
    tape = layered_ry_rz_cz(N, 10, seed=2024)
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1
    for gate, qs, th, par, m in tape.ops:
        orc.apply_named(ref, gate.lower(), [q + 1 for q in qs], th, N)
    s = GPUState.basis(N, 0)
    npass = tape.apply(s)
    assert npass < len(tape)
    xs = [0] * (2 * N - 1)
    zs = [1 << i for i in range(N)] + [(3 << i) for i in range(N - 1)]
    got = ob.expect_pauli(s, xs, zs)
    want = [orc.expect_local(ref, i + 1, N, "z") for i in range(N)] + \
           [orc.expect_corr(ref, i + 1, i + 2, N, "zz") for i in range(N - 1)]
    assert np.max(np.abs(got - np.array(want))) < 1e-12
    assert np.linalg.norm(s.to_numpy() - ref) < 1e-12


@pytest.mark.parametrize("N", [1, 3, 7, 8, 12, 17])
def test_sampling_bit_exact(ctx, N):
    from smoq_amd import GPUState
    from smoq_amd import measurements as me
    from smoq_amd import streams

    psi = rand_state(N, 40 + N)
    s = GPUState.from_numpy(psi)
    r = streams.uniforms(77 + N, 4000)
    r[:4] = [0.0, 1.0 - 2.0 ** -53, 0.5, 1e-300]
    got = me.sample_indices(s, r, "cumsum").astype(np.int64)
    want = orc.sample_cumsum_search(psi, N, r)
    assert np.array_equal(got, want)
    got = me.sample_indices(s, r[:500], "sequential").astype(np.int64)
    want = orc.sample_sequential(psi, N, r[:500])
    assert np.array_equal(got, want)


def test_sampling_edge_cases(ctx):
    from smoq_amd import GPUState
    from smoq_amd import measurements as me

    # all mass on one basis state; r above the last prefix (unnormalised
    # state) gives 2^N for the cumsum rule and the last index for the
    # sequential rule, as in the reference
    N = 6
    psi = np.zeros(1 << N, np.complex128)
    psi[37] = 0.5
    s = GPUState.from_numpy(psi)
    r = np.array([0.0, 0.1, 0.25, 0.3, 0.99])
    assert me.sample_indices(s, r, "cumsum").tolist() == [0, 37, 37, 64, 64]
    assert me.sample_indices(s, r, "sequential").tolist() == [0, 37, 37, 63, 63]
    bits, st = me.projective_measurement_all(s, N, 0.2)
    assert bits == [(37 >> k) & 1 for k in range(N)]
    col = st.to_numpy()
    assert col[37] == 1 and np.count_nonzero(col) == 1


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 1e-5)])
def test_measure_all_observables_parity(ctx, dtype, tol):
    """measure_all_observables_* on rails (L, n_rails) through the tiled
    multi-mask reduction, against the oracle's per-observable loops
    (_expect_X/Y/Z, _expect_XX/YY/ZZ), in the reference's order."""
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    for L, rails in [(1, 1), (2, 1), (3, 2), (5, 1), (4, 3), (13, 1), (8, 2), (6, 3)]:
        N = L * rails
        psi = rand_state(N, 100 + N)
        s = GPUState.from_numpy(psi.astype(np.complex64) if dtype == "c64" else psi)
        ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "c64" else psi
        want = []
        for k in range(1, N + 1):
            want += [orc.expect_local(ref, k, N, p) for p in "xyz"]
        bonds = [(r * L + q + 1, r * L + q + 2) for r in range(rails) for q in range(L - 1)]
        bonds += [(r * L + q + 1, (r + 1) * L + q + 1) for r in range(rails - 1) for q in range(L)]
        for i, j in bonds:
            want += [orc.expect_corr(ref, i, j, N, p) for p in ("xx", "yy", "zz")]
        got = ob.measure_all_observables_state_vector(s, L, rails)
        assert got.shape == (3 * N + 3 * len(bonds),)
        assert np.max(np.abs(got - np.array(want))) < tol, (L, rails)
        one = ob.measure_all_observables_one_pass(s, L, rails)
        np.testing.assert_array_equal(one, got)
        jx, jy, jz = ob.collective_spin_fast(s, N)
        assert abs(jx - 0.5 * sum(want[0:3 * N:3])) < tol
        assert abs(jy - 0.5 * sum(want[1:3 * N:3])) < tol
        assert abs(jz - 0.5 * sum(want[2:3 * N:3])) < tol
        s.free()


def test_tiled_reduction_many_masks_batched(ctx):
    """Long term lists with many x masks (local, neighbour, random strings of
    any weight) on batched states at N up to 20: every term against the
    dense <ψ|P|ψ> (N <= 10) or the oracle's loops, each batch item alone."""
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    rng = np.random.default_rng(7)
    for N, B in [(4, 3), (10, 2), (14, 3), (17, 2), (20, 1)]:
        xs, zs = ob.qrc_observable_masks(N, 1)
        for _ in range(30):
            xs.append(int(rng.integers(0, 1 << N)))
            zs.append(int(rng.integers(0, 1 << N)))
        psis = np.stack([rand_state(N, 1000 * N + b) for b in range(B)])
        s = GPUState.from_numpy(psis)
        got = ob.expect_pauli(s, xs, zs)
        if B == 1:
            got = got[None, :]
        for b in range(B):
            if N <= 10:
                want = [np.vdot(psis[b], orc.pauli_matrix(x, z, N) @ psis[b]).real
                        for x, z in zip(xs, zs)]
                free = [pauli_dense_free(psis[b], x, z) for x, z in zip(xs, zs)]
                assert np.max(np.abs(np.array(free) - want)) < 1e-13
            else:
                want = []
                for k in range(1, N + 1):
                    want += [orc.expect_local(psis[b], k, N, p) for p in "xyz"]
                for i in range(1, N):
                    want += [orc.expect_corr(psis[b], i, i + 1, N, p) for p in ("xx", "yy", "zz")]
                want += [pauli_dense_free(psis[b], x, z) for x, z in zip(xs[len(want):],
                                                                         zs[len(want):])]
            assert np.max(np.abs(got[b] - np.array(want))) < 1e-12, (N, b)
        s.free()


def pauli_dense_free(psi, x, z):
    """<ψ| i^{|x&z|} X^x Z^z |ψ> by index arithmetic (no dense matrix)."""
    idx = np.arange(psi.size, dtype=np.int64)
    sgn = 1.0 - 2.0 * (np.bitwise_count(idx & z) & 1)
    ph = 1j ** (bin(x & z).count("1") & 3)
    return float(np.real(np.vdot(psi[idx ^ x], ph * sgn * psi)))
