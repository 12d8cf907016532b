"""Density matrices on the device (vec(ρ) as a 2N-qubit state): gates as
U ρ U†, the four dissipators, Tr(ρ P) read-outs, the dense step and the
reference demo's DM columns, against the oracle's restatements."""
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rand_rho(N, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(1 << N, 1 << N)) + 1j * rng.normal(size=(1 << N, 1 << N))
    rho = A @ A.conj().T
    return rho / np.trace(rho).real


def dense_tape_unitary(tape):
    """The tape's N-qubit operator, column by column through the oracle."""
    from test_gpu_gates import oracle_run_tape

    D = 1 << tape.n
    U = np.zeros((D, D), np.complex128)
    for c in range(D):
        e = np.zeros(D, np.complex128)
        e[c] = 1.0
        U[:, c] = oracle_run_tape(tape, e)
    return U


@pytest.mark.parametrize("N,dtype,tol", [(3, "c128", 1e-12), (5, "c128", 1e-12), (5, "c64", 2e-6)])
def test_gates_on_rho_are_u_rho_u_dagger(ctx, N, dtype, tol):
    from smoq_amd.density import DensityMatrix, apply_tape_rho
    from test_gpu_gates import random_tape

    tape = random_tape(N, 40, 100 + N)
    U = dense_tape_unitary(tape)
    rho = rand_rho(N, N)
    want = U @ rho @ U.conj().T
    d = DensityMatrix.from_numpy(rho.astype(np.complex64) if dtype == "c64" else rho)
    apply_tape_rho(d, tape)
    assert np.max(np.abs(d.to_numpy() - want)) < tol
    assert abs(d.trace() - 1.0) < 10 * tol
    d.free()


def test_named_rho_mirrors(ctx):
    from smoq_amd import density as dm

    N = 3
    rho = rand_rho(N, 1)
    d = dm.DensityMatrix.from_numpy(rho)
    ops = [("rx", 1, 0.3), ("ry", 2, 1.1), ("rz", 3, -0.7), ("h", 1, None), ("x", 2, None),
           ("z", 3, None), ("t", 1, None), ("tdagger", 2, None)]
    ref = rho.copy()
    for name, k, th in ops:
        fn = {"h": dm.apply_hadamard_rho, "x": dm.apply_pauli_x_rho, "z": dm.apply_pauli_z_rho,
              "t": dm.apply_t_rho, "tdagger": dm.apply_tdagger_rho}.get(name)
        if th is not None:
            getattr(dm, f"apply_{name}_rho")(d, k, th, N)
        else:
            fn(d, k, N)
        gname = {"tdagger": "tdg"}.get(name, name)
        G = np.eye(1 << N, dtype=np.complex128)
        cols = []
        for c in range(1 << N):
            e = G[:, c].copy()
            orc.apply_named(e, gname, [k], th or 0.0, N)
            cols.append(e)
        U = np.stack(cols, 1)
        ref = U @ ref @ U.conj().T
    for name, a, b, th in [("cnot", 1, 3, None), ("cz", 2, 3, None), ("cry", 3, 1, 0.9),
                           ("rxx", 1, 2, 0.4), ("ryy", 2, 3, -1.3), ("rzz", 1, 3, 2.2)]:
        if th is None:
            getattr(dm, f"apply_{name}_rho")(d, a, b, N)
        else:
            getattr(dm, f"apply_{name}_rho")(d, a, b, th, N)
        cols = []
        for c in range(1 << N):
            e = np.zeros(1 << N, np.complex128)
            e[c] = 1.0
            orc.apply_named(e, name, [a, b], th or 0.0, N)
            cols.append(e)
        U = np.stack(cols, 1)
        ref = U @ ref @ U.conj().T
    assert np.max(np.abs(d.to_numpy() - ref)) < 1e-12
    d.free()


@pytest.mark.parametrize("kind", ["sigma_minus", "sigma_plus", "sigma_z", "sigma_x"])
def test_dissipators_match_oracle(ctx, kind):
    from smoq_amd.density import DensityMatrix, apply_dissipator

    N, gdt = 4, 0.09
    rho = rand_rho(N, 7)
    d = DensityMatrix.from_numpy(rho)
    ref = rho.copy()
    for k in (1, 3, 4):
        apply_dissipator(d, kind, k, gdt)
        orc.dm_dissipator(ref, kind, k, gdt, N)
    assert np.max(np.abs(d.to_numpy() - ref)) < 1e-14
    d.free()


def test_dm_readouts(ctx):
    from smoq_amd import density as dm

    N = 5
    rho = rand_rho(N, 9)
    d = dm.DensityMatrix.from_numpy(np.stack([rho, rho.conj()]))  # batch of 2
    rng = np.random.default_rng(3)
    xs = [int(v) for v in rng.integers(0, 1 << N, 20)]
    zs = [int(v) for v in rng.integers(0, 1 << N, 20)]
    got = d.expect_pauli(xs, zs)
    for b, r in enumerate((rho, rho.conj())):
        want = [np.trace(orc.pauli_matrix(x, z, N) @ r).real for x, z in zip(xs, zs)]
        assert np.max(np.abs(got[b] - want)) < 1e-14
    d.free()
    d = dm.DensityMatrix.from_numpy(rho)
    for k in range(1, N + 1):
        for p in "xyz":
            assert abs(dm.expect_local_dm(d, k, N, p) - orc.dm_expect_local(rho, k, N, p)) < 1e-14
        pop = sum(rho[i, i].real for i in range(1 << N) if (i >> (k - 1)) & 1)
        assert abs(dm.population_excited_rho(d, k, N) - pop) < 1e-14
    for i in range(1, N):
        for p in ("xx", "yy", "zz"):
            assert abs(dm.expect_corr_dm(d, i, i + 1, N, p)
                       - orc.dm_expect_corr(rho, i, i + 1, N, p)) < 1e-14
    assert abs(d.purity() - np.sum(np.abs(rho) ** 2)) < 1e-14
    allobs = dm.measure_all_observables_dm(d, N, 1)
    assert allobs.shape == (3 * N + 3 * (N - 1),)
    d.free()
    r = 1 / math.sqrt(2)
    p = dm.DensityMatrix.product(["zero", "plus", "one"])
    ket = orc.product_ket([(1, 0), (r, r), (0, 1)])
    assert np.max(np.abs(p.to_numpy() - np.outer(ket, ket.conj()))) < 1e-15
    p.free()
    b = dm.DensityMatrix.basis(3, 5)
    v = b.to_numpy()
    assert v[5, 5] == 1.0 and np.count_nonzero(v) == 1
    b.free()


def test_trotter_and_dense_steps_match_oracle(ctx):
    from smoq_amd import density as dm
    from smoq_amd.mcwf import xxz_field_trotter_gates

    N, gamma, dt = 4, 0.3, 0.05
    rho = rand_rho(N, 4)
    fg = xxz_field_trotter_gates(N, dt)
    gates = [(list(g.indices), g.U) for g in fg]
    d = dm.DensityMatrix.from_numpy(rho)
    ref = rho.copy()
    for _ in range(5):
        dm.lindblad_dm_step_trotter(d, fg, N, gamma, dt)
        orc.dm_trotter_step(ref, gates, N, gamma, dt)
    assert np.max(np.abs(d.to_numpy() - ref)) < 1e-13
    d.free()
    rng = np.random.default_rng(1)
    H = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    U = expm(-1j * (H + H.conj().T) * dt)
    Ud = dm.dense_operator(U)
    d = dm.DensityMatrix.from_numpy(rho)
    ref = rho.copy()
    for _ in range(3):
        dm.lindblad_dm_step(d, Ud, N, gamma, dt)
        orc.dm_exact_step(ref, U, N, gamma, dt)
    assert np.max(np.abs(d.to_numpy() - ref)) < 1e-13
    d.free()
    Ud.free()


def test_dm_demo_columns(ctx):
    """The σ_z-dephasing demo's DM+Trotter and DM+Exact values at t = 10
    (N = 2, 4, 6, 8; output.log, 4 decimals) from evolve_lindblad_dm."""
    from smoq_amd import density as dm
    from smoq_amd.hamiltonian import build_hamiltonian_parameters, pauli_sum
    from smoq_amd.mcwf import xxz_field_trotter_gates
    from smoq_amd.observables import qrc_observable_masks

    with open(os.path.join(HERE, "golden", "mcwf_dephasing_xxz.json")) as f:
        ref = json.load(f)
    P = ref["params"]
    gamma, dt, steps = P["gamma"], P["dt"], P["n_steps"]
    for N in (2, 4, 6, 8):
        xs, zs = qrc_observable_masks(N, 1)
        kinds = np.array([0, 1, 2] * N + [3, 4, 5] * (N - 1))

        def means(v):
            return np.array([v[kinds == c].mean() for c in range(6)])

        d = dm.DensityMatrix.basis(N, 0)
        rec = dm.evolve_lindblad_dm(d, xxz_field_trotter_gates(N, dt), N, gamma, dt, steps,
                                    obs_every=steps, xmask=xs, zmask=zs)
        assert rec.shape == (2, len(xs))
        assert np.max(np.abs(means(rec[-1]) - ref["reference"][str(N)]["DM+Trotter"])) < 6e-5
        d.free()
        params = build_hamiltonian_parameters(N, 1, J_x_direction=(1.0, 1.0, 0.5),
                                              J_y_direction=(0.0, 0.0, 0.0),
                                              h_field=(1.0, 0.0, 0.0))
        px, pz, pc = pauli_sum(params)
        H = sum(c * orc.pauli_matrix(x, z, N) for x, z, c in zip(px, pz, pc))
        U = dm.dense_operator(expm(-1j * H * dt))
        d = dm.DensityMatrix.basis(N, 0)
        rec = dm.evolve_lindblad_dm(d, None, N, gamma, dt, steps, obs_every=steps, xmask=xs,
                                    zmask=zs, U=U)
        assert np.max(np.abs(means(rec[-1]) - ref["reference"][str(N)]["DM+Exact"])) < 6e-5
        d.free()
        U.free()
