"""Trotter steps, compiled plans (graph replay), H·ψ and Chebyshev
propagation on the device against the CPU oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b)


def neel(N):
    psi = np.zeros(1 << N, np.complex128)
    psi[sum(1 << (k - 1) for k in range(2, N + 1, 2))] = 1.0
    return psi


def xxz(N, seed, delta=0.5):
    from smoq_amd.hamiltonian import xxz_chain

    h = np.random.default_rng(seed).uniform(-1, 1, N)
    return xxz_chain(N, 1.0, delta, h), h


def as_oracle(gates):
    return [(list(g.indices), g.U) for g in gates]


def test_first_order_trotter_step_parity(ctx):
    from smoq_amd import GPUState
    from smoq_amd.trotter import apply_fast_trotter_step_cpu, precompute_trotter_gates_bitwise_cpu

    N = 11
    params, _ = xxz(N, 3)
    gates = precompute_trotter_gates_bitwise_cpu(params, 0.05)
    psi = GPUState.from_numpy(neel(N))
    ref = neel(N)
    for _ in range(5):
        apply_fast_trotter_step_cpu(psi, gates, N)
        orc.trotter_step(ref, as_oracle(gates), N)
    assert rel(psi.to_numpy(), ref) < 1e-12


@pytest.mark.parametrize("N,dtype,tol", [(12, "c128", 1e-12), (13, "c128", 1e-12),
                                         (12, "c64", 1e-5)])
def test_second_order_evolution_and_observables(ctx, N, dtype, tol):
    from smoq_amd import GPUState
    from smoq_amd.trotter import second_order_gates, trotter_tape

    params, _ = xxz(N, 5)
    seq = second_order_gates(params, 0.01)
    assert len(seq) == 2 * (2 * N - 1) - 1
    plan = trotter_tape(seq, N).compile(dtype)
    psi = GPUState.from_numpy(neel(N).astype(np.complex64 if dtype == "c64" else np.complex128))
    zs = [1 << k for k in range(N)]
    rec = plan.evolve_observe(psi, 40, 10, [0] * N, zs)
    ref = neel(N)
    want = []
    for step in range(1, 41):
        orc.trotter_step(ref, as_oracle(seq), N)
        if step % 10 == 0:
            want.append([orc.expect_local(ref, k + 1, N, "z") for k in range(N)])
    assert rec.shape == (4, N)
    assert np.max(np.abs(rec - np.array(want))) < (1e-12 if dtype == "c128" else 1e-5)
    assert rel(psi.to_numpy(), ref) < tol


def test_plan_replay_matches_tape(ctx):
    # the captured graph is replayed; interleave eager work between replays
    from smoq_amd import GPUState
    from smoq_amd.circuits import random_rotation_brick

    N = 14
    tape = random_rotation_brick(N, 6, seed=9)
    plan = tape.compile()
    a = GPUState.basis(N, 0)
    b = GPUState.basis(N, 0)
    for r in range(3):
        plan.run(a, 2)
        tape.apply(b)
        tape.apply(b)
        assert rel(a.to_numpy(), b.to_numpy()) < 1e-13, r
    # a second state of the same shape re-captures
    c = GPUState.basis(N, 0)
    plan.run(c, 6)
    assert rel(c.to_numpy(), a.to_numpy()) < 1e-13


def test_hpsi_parity_with_field_y(ctx):
    from smoq_amd import GPUState
    from smoq_amd.chebyshev import apply_hamiltonian_bitwise
    from smoq_amd.hamiltonian import build_hamiltonian_parameters

    N = 9
    params = build_hamiltonian_parameters(3, 3, J_x_direction=(0.7, 0.4, 1.1),
                                          J_y_direction=(0.2, 0.9, -0.5),
                                          h_field=(0.3, -0.8, 0.6))
    rng = np.random.default_rng(4)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    v /= np.linalg.norm(v)
    bonds = [(b.i, b.j, b.Jxx, b.Jyy, b.Jzz) for b in params.x_bonds + params.y_bonds]
    fields = np.array([[f.hx, f.hy, f.hz] for f in params.fields])
    ref = orc.hpsi(v, N, bonds, fields)
    a = GPUState.from_numpy(v)
    out = GPUState(N)
    apply_hamiltonian_bitwise(out, a, params)
    assert np.max(np.abs(out.to_numpy() - ref)) < 1e-13


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-5)])
def test_chebyshev_parity(ctx, dtype, tol):
    from smoq_amd import GPUState
    from smoq_amd.chebyshev import chebyshev_evolve_psi
    from smoq_amd.hamiltonian import spectral_bound

    N = 10
    params, h = xxz(N, 6)
    B = spectral_bound(params)
    bonds = [(b.i, b.j, b.Jxx, b.Jyy, b.Jzz) for b in params.x_bonds]
    fields = np.array([[0.0, 0.0, hi] for hi in h])
    ref, m_ref = orc.chebyshev_evolve(neel(N), N, bonds, fields, 3.0, -B, B)
    psi = GPUState.from_numpy(neel(N).astype(np.complex64 if dtype == "c64" else np.complex128))
    m, a, b = chebyshev_evolve_psi(psi, params, 3.0, spectral_bounds=(-B, B))
    assert m == m_ref
    assert rel(psi.to_numpy(), ref) < tol


def test_trotter_converges_to_chebyshev(ctx):
    # physics check: second-order Trotter at dt=0.01 vs Chebyshev at t=1
    from smoq_amd import GPUState
    from smoq_amd.chebyshev import chebyshev_evolve_psi
    from smoq_amd.trotter import second_order_gates, trotter_tape

    N = 12
    params, _ = xxz(N, 7)
    b = GPUState.from_numpy(neel(N))
    chebyshev_evolve_psi(b, params, 1.0)
    exact = b.to_numpy()
    errs = []
    for dt, steps in ((0.02, 50), (0.01, 100)):
        plan = trotter_tape(second_order_gates(params, dt), N).compile()
        a = GPUState.from_numpy(neel(N))
        plan.run(a, steps)
        errs.append(rel(a.to_numpy(), exact))
    assert errs[1] < 1e-3
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


@pytest.mark.parametrize("Nx,Ny,k_iter", [(8, 1, 30), (3, 3, 40), (12, 1, 30), (2, 1, 30)])
def test_lanczos_range_matches_oracle(ctx, Nx, Ny, k_iter):
    """estimate_spectral_range_lanczos on the device against the oracle's
    restatement from the same start vector (2D lattice with all field
    components, and a 2-qubit case that stops early on beta < 1e-12); at
    N <= 9 both against the dense spectrum."""
    from smoq_amd import GPUState
    from smoq_amd.chebyshev import estimate_spectral_range_lanczos
    from smoq_amd.hamiltonian import build_hamiltonian_parameters, pauli_sum

    N = Nx * Ny
    params = build_hamiltonian_parameters(Nx, Ny, J_x_direction=(0.7, 0.4, 1.1),
                                          J_y_direction=(0.2, 0.9, -0.5),
                                          h_field=(0.3, -0.8, 0.6))
    rng = np.random.default_rng(N)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    bonds = [(b.i, b.j, b.Jxx, b.Jyy, b.Jzz) for b in params.x_bonds + params.y_bonds]
    fields = np.array([[f.hx, f.hy, f.hz] for f in params.fields])
    lo_ref, hi_ref, steps = orc.lanczos_range(v, N, bonds, fields, k_iter)
    start = GPUState.from_numpy(v)
    lo, hi = estimate_spectral_range_lanczos(params, k_iter, start=start)
    assert np.max(np.abs(start.to_numpy() - v)) == 0.0  # start left unchanged
    start.free()
    assert abs(lo - lo_ref) < 1e-9 * max(1.0, abs(lo_ref)) and abs(hi - hi_ref) < 1e-9 * max(
        1.0, abs(hi_ref)), (lo, lo_ref, hi, hi_ref)
    if N <= 9:
        xs, zs, cs = pauli_sum(params)
        H = sum(c * orc.pauli_matrix(x, z, N) for x, z, c in zip(xs, zs, cs))
        ev = np.linalg.eigvalsh(H)
        assert abs(lo - ev[0]) < 1e-8 and abs(hi - ev[-1]) < 1e-8
    if N == 2:
        assert steps < k_iter
