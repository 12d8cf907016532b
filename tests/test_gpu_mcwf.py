"""Batched MCWF trajectories (one persistent launch) against the CPU oracle:
jump sites bit-exact on shared uniforms, <Z_i> records and final states to
1e-12 (complex128) / 1e-5 (complex64); the trajectory mean against closed
forms (pure decay) and the exact Lindblad solution (small N)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def excited(N, T):
    psi = np.zeros((T, 1 << N), np.complex128)
    psi[:, -1] = 1.0
    return psi


def random_states(N, T, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=(T, 1 << N)) + 1j * rng.normal(size=(T, 1 << N))
    return psi / np.linalg.norm(psi, axis=1, keepdims=True)


def chain_gates(N, dt, seed=None, J=1.0, h=0.7):
    """TFIM chain gates; with a seed, per-bond J and per-site (hx, hz) are
    random (the generic, non-uniform bond path)."""
    from smoq_amd.hamiltonian import CouplingTerm, FieldTerm, HamiltonianParams
    from smoq_amd.trotter import precompute_trotter_gates_bitwise_cpu

    if seed is None:
        Js, hx, hz = [J] * (N - 1), [h] * N, [0.0] * N
    else:
        rng = np.random.default_rng(seed)
        Js, hx, hz = rng.uniform(0.5, 1.5, N - 1), rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    params = HamiltonianParams(N, [CouplingTerm(i, i + 1, 0.0, 0.0, float(Js[i - 1]))
                                   for i in range(1, N)], [],
                               [FieldTerm(i, float(hx[i - 1]), 0.0, float(hz[i - 1]))
                                for i in range(1, N + 1)])
    return precompute_trotter_gates_bitwise_cpu(params, dt)


def oracle_run(psi0, gates, N, gamma, dt, U, obs_every):
    from smoq_amd.mcwf import decay_powers

    dec = decay_powers(N, gamma, dt)
    gl = [(list(g.indices), g.U) for g in gates]
    out_psi, jumps, zs = [], [], []
    for t in range(psi0.shape[0]):
        p = psi0[t].astype(np.complex128).copy()
        j, z = orc.mcwf_trajectory(p, gl, N, gamma * dt, dec, U[t], obs_every)
        out_psi.append(p)
        jumps.append(j)
        zs.append(z)
    return np.array(out_psi), np.array(jumps), np.array(zs)


def check_group(res, psi_dev, ref, tol, N):
    ref_psi, ref_jumps, ref_z = ref
    np.testing.assert_array_equal(res.jumps, ref_jumps)
    got = psi_dev.to_numpy().reshape(ref_psi.shape)
    assert np.max(np.abs(got - ref_psi)) < tol
    if res.z is not None:
        assert res.z.shape == ref_z.shape
        assert np.max(np.abs(res.z - ref_z)) < tol


@pytest.mark.parametrize("N,T,steps,seed", [(6, 16, 150, None), (14, 4, 25, 3)])
def test_mcwf_trajectories_match_oracle(ctx, N, T, steps, seed):
    """N=6: state resident in LDS, uniform-chain bond table; N=14: two
    tiled passes per step over HBM, generic per-bond diagonals."""
    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_batched_trotter
    from smoq_amd.streams import uniforms

    gamma, dt = 0.8, 0.05
    gates = chain_gates(N, dt, seed)
    psi0 = excited(N, T) if seed is None else random_states(N, T, seed)
    U = uniforms(100 + N, T * steps).reshape(T, steps)
    dev = GPUState.from_numpy(psi0)
    res = evolve_mcwf_batched_trotter(dev, gates, N, T, steps, gamma, dt, uniforms=U,
                                      obs_every=5)
    assert (res.jumps >= 0).sum() > 0  # the run exercises jumps
    check_group(res, dev, oracle_run(psi0, gates, N, gamma, dt, U, 5), 1e-12, N)


def test_mcwf_mixed_sizes_in_one_launch(ctx):
    """Groups of 1, 2, 5, 12, 13 and 20 qubits (padded tile, LDS-resident,
    two-pass with h = 1 and h = 8) share one launch."""
    from smoq_amd import GPUState
    from smoq_amd.mcwf import MCWFGroup, evolve_mcwf_groups, mcwf_gates
    from smoq_amd.streams import uniforms

    gamma, dt, steps, obs = 1.5, 0.04, 8, 2
    cases = [(1, 5), (2, 6), (5, 7), (12, 3), (13, 2), (20, 2)]
    groups, refs, devs = [], [], []
    for i, (N, T) in enumerate(cases):
        gates = chain_gates(N, dt, seed=7 + i) if N > 1 else chain_gates(N, dt)
        psi0 = random_states(N, T, 40 + i)
        U = uniforms(900 + i, T * steps).reshape(T, steps)
        dev = GPUState.from_numpy(psi0)
        devs.append(dev)
        groups.append(MCWFGroup(dev, mcwf_gates(gates, N), gamma, dt, U))
        refs.append(oracle_run(psi0, gates, N, gamma, dt, U, obs))
    results = evolve_mcwf_groups(groups, gamma, dt, steps, obs)
    for (N, _), res, dev, ref in zip(cases, results, devs, refs):
        check_group(res, dev, ref, 1e-12, N)


def test_mcwf_complex64(ctx):
    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_batched_trotter
    from smoq_amd.streams import uniforms

    N, T, steps, gamma, dt = 10, 8, 60, 0.8, 0.05
    gates = chain_gates(N, dt)
    psi0 = excited(N, T)
    U = uniforms(77, T * steps).reshape(T, steps)
    dev = GPUState.from_numpy(psi0.astype(np.complex64))
    res = evolve_mcwf_batched_trotter(dev, gates, N, T, steps, gamma, dt, uniforms=U,
                                      obs_every=10)
    check_group(res, dev, oracle_run(psi0, gates, N, gamma, dt, U, 10), 1e-5, N)


def test_mcwf_single_step_api(ctx):
    from smoq_amd import GPUState
    from smoq_amd.mcwf import lindblad_mcwf_batched_step

    N, T, gamma, dt = 4, 64, 2.0, 0.1
    gates = chain_gates(N, dt)
    psi0 = excited(N, T)
    U = np.linspace(0.0, 0.99, T)
    dev = GPUState.from_numpy(psi0)
    jumps = lindblad_mcwf_batched_step(dev, gates, N, T, gamma, dt, uniforms=U)
    ref_psi, ref_jumps, _ = oracle_run(psi0, gates, N, gamma, dt, U.reshape(T, 1), 0)
    np.testing.assert_array_equal(jumps, ref_jumps[:, 0])
    assert np.max(np.abs(dev.to_numpy() - ref_psi)) < 1e-12


def test_mcwf_pure_decay_closed_form(ctx):
    """One qubit, no Hamiltonian: the jump step of each trajectory is the
    first step with r < 1 - (sqrt(1-γdt))², and the excited fraction after
    k steps has mean (1-γdt)^k."""
    from smoq_amd import GPUState
    from smoq_amd.mcwf import MCWFGates, evolve_mcwf_batched_trotter
    from smoq_amd.streams import uniforms

    T, steps, gamma, dt = 20000, 100, 1.0, 0.01
    g = MCWFGates(1, np.zeros((0, 2), np.int32), np.zeros((0, 4), np.complex128),
                  np.eye(2, dtype=np.complex128).reshape(1, 2, 2))
    U = uniforms(5, T * steps).reshape(T, steps)
    dev = GPUState.from_numpy(excited(1, T))
    res = evolve_mcwf_batched_trotter(dev, g, 1, T, steps, gamma, dt, uniforms=U, obs_every=1)
    d = np.sqrt(1.0 - gamma * dt)
    dp = 1.0 - d * d
    hit = U < dp
    first = np.where(hit.any(1), hit.argmax(1), steps)
    want = np.full((T, steps), -1, np.int32)
    want[first < steps, first[first < steps]] = 0
    np.testing.assert_array_equal(res.jumps, want)
    excited_frac = (1.0 - res.z[:, :, 0].mean(0)) / 2.0
    k = np.arange(1, steps + 1)
    p = (1.0 - gamma * dt) ** k
    assert np.all(np.abs(excited_frac - p) < 5 * np.sqrt(p * (1 - p) / T) + 1e-12)


def test_mcwf_mean_matches_lindblad(ctx):
    """TFIM chain with amplitude damping, N = 3: the mean <Z_i> of 16384
    trajectories at t = 2 against the exact Lindblad solution (Liouvillian
    exponential); statistical error ~3e-3, first-order splitting ~2e-3."""
    from scipy.linalg import expm

    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_batched_trotter, tfim_mcwf_gates

    N, J, h, gamma, dt, steps, T = 3, 1.0, 0.7, 0.3, 0.01, 200, 16384
    dev = GPUState.from_numpy(excited(N, T))
    res = evolve_mcwf_batched_trotter(dev, tfim_mcwf_gates(N, J, h, dt), N, T, steps, gamma,
                                      dt, seed=11, obs_every=steps)
    z_mc = res.z[:, 0, :].mean(0)
    D = 1 << N
    X = np.array([[0, 1], [1, 0]])
    Z = np.diag([1.0, -1.0])
    sm = np.array([[0, 1], [0, 0]])

    def op(o, k):
        return np.kron(np.eye(1 << (N - 1 - k)), np.kron(o, np.eye(1 << k)))

    H = sum(J * op(Z, k) @ op(Z, k + 1) for k in range(N - 1)) + sum(h * op(X, k)
                                                                     for k in range(N))
    I = np.eye(D)
    L = -1j * (np.kron(I, H) - np.kron(H.T, I))
    for k in range(N):
        c = op(sm, k)
        cdc = c.conj().T @ c
        L = L + gamma * (np.kron(c.conj(), c) - 0.5 * np.kron(I, cdc) - 0.5 * np.kron(cdc.T, I))
    rho0 = np.zeros((D, D))
    rho0[-1, -1] = 1.0
    rho = (expm(L * dt * steps) @ rho0.reshape(-1, order="F")).reshape(D, D, order="F")
    z_ex = np.array([np.trace(rho @ op(Z, k)).real for k in range(N)])
    assert np.max(np.abs(z_mc - z_ex)) < 0.02


def test_mcwf_rejects_bad_input(ctx):
    from smoq_amd import GPUState, SmoqError
    from smoq_amd.mcwf import MCWFGates, MCWFGroup, evolve_mcwf_groups

    dev = GPUState(4, batch=2)
    g = MCWFGates(4, np.array([[0, 4]], np.int32), np.ones((1, 4), np.complex128),
                  np.tile(np.eye(2, dtype=np.complex128), (4, 1, 1)))
    with pytest.raises(SmoqError):
        evolve_mcwf_groups([MCWFGroup(dev, g, 0.1, 0.1, np.zeros((2, 3)))], 0.1, 0.1, 3)
    g.bond_q[0, 1] = 1
    with pytest.raises(ValueError):  # γ dt > 1
        evolve_mcwf_groups([MCWFGroup(dev, g, 20.0, 0.1, np.zeros((2, 3)))], 20.0, 0.1, 3)
    with pytest.raises(ValueError):
        evolve_mcwf_groups([MCWFGroup(dev, g, 0.1, 0.1, np.zeros((2, 2)))], 0.1, 0.1, 3)
    big = GPUState(21)
    g21 = MCWFGates(21, np.zeros((0, 2), np.int32), np.zeros((0, 4), np.complex128),
                    np.tile(np.eye(2, dtype=np.complex128), (21, 1, 1)))
    with pytest.raises(SmoqError):
        evolve_mcwf_groups([MCWFGroup(big, g21, 0.1, 0.1, np.zeros((1, 1)))], 0.1, 0.1, 1)


def test_mcwf_tfim_damping_vs_density_matrix(ctx):
    """The damped TFIM chain (J = 1, h = 0.8, γ = 0.1, dt = 0.005, 2000
    steps) at N = 8 and N = 10, 4096 trajectories each in one launch: the
    trajectory mean of <Z_i>(t) at t = 1..10 against the exact
    density-matrix solution (tests/golden/lindblad_tfim_damping.json),
    within 5 standard errors + 0.01 (first-order splitting)."""
    import json
    import os

    from smoq_amd import GPUState
    from smoq_amd.mcwf import MCWFGroup, evolve_mcwf_groups, tfim_mcwf_gates
    from smoq_amd.streams import uniforms

    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "lindblad_tfim_damping.json")) as f:
        ref = json.load(f)
    J, h, gamma = ref["J"], ref["h"], ref["gamma"]
    dt, steps, every, T = 0.005, 2000, 200, 4096
    assert ref["times"] == [every * dt * (r + 1) for r in range(steps // every)]
    sizes = ref["sizes"]
    groups = []
    for N in sizes:
        s = GPUState.basis(N, (1 << N) - 1, batch=T)
        U = uniforms(31 + N, T * steps).reshape(T, steps)
        groups.append(MCWFGroup(s, tfim_mcwf_gates(N, J, h, dt), gamma, dt, U))
    res = evolve_mcwf_groups(groups, gamma, dt, steps, every)
    for N, r in zip(sizes, res):
        z_dm = np.array(ref["z"][str(N)])
        z_mc = r.z.mean(0)
        se = r.z.std(0) / np.sqrt(T)
        assert z_mc.shape == z_dm.shape == (steps // every, N)
        assert np.all(np.abs(z_mc - z_dm) < 5 * se + 0.01), (N, np.max(np.abs(z_mc - z_dm)))


def test_mcwf_zero_norm_is_reported(ctx):
    """(|00> + |10>)/√2 with γdt = 0.5 and r = 0.2: the decision jumps but the
    site weights (0, 1/6) stay below r, so the fallback site 0 is chosen,
    whose population is 0: σ⁻ leaves a zero vector (NaN in the reference
    loop).  The call reports SMOQ_ENUMERIC naming the trajectory and step;
    the other trajectory completes."""
    from smoq_amd import GPUState, SmoqError
    from smoq_amd import _lib
    from smoq_amd.mcwf import MCWFGates, MCWFGroup, evolve_mcwf_groups

    psi0 = np.zeros((2, 4), np.complex128)
    psi0[0, 0] = 1.0
    psi0[1, 0] = psi0[1, 2] = 1.0 / np.sqrt(2.0)
    dev = GPUState.from_numpy(psi0)
    g = MCWFGates(2, np.zeros((0, 2), np.int32), np.zeros((0, 4), np.complex128),
                  np.tile(np.eye(2, dtype=np.complex128), (2, 1, 1)))
    grp = MCWFGroup(dev, g, 5.0, 0.1, np.full((2, 3), 0.2))
    with pytest.raises(SmoqError) as e:
        evolve_mcwf_groups([grp], 5.0, 0.1, 3)
    assert e.value.code == _lib.SMOQ_ENUMERIC
    assert "step 0 of trajectory 1" in str(e.value)
    assert grp.jumps[0].tolist() == [-1, -1, -1] and grp.jumps[1, 0] == 0


# ---------------------------------------- any coherent step: plan + jump --
def xxz_x_gates(N, dt, J=(1.0, 1.0, 0.5), hx=1.0):
    """The Trotter gates of the reference's σ_z-dephasing demo: XXZ bonds
    (non-diagonal) and x fields, from precompute_trotter_gates_bitwise_cpu."""
    from smoq_amd.hamiltonian import build_hamiltonian_parameters
    from smoq_amd.trotter import precompute_trotter_gates_bitwise_cpu

    params = build_hamiltonian_parameters(N, 1, J_x_direction=J, J_y_direction=(0.0, 0.0, 0.0),
                                          h_field=(hx, 0.0, 0.0))
    return precompute_trotter_gates_bitwise_cpu(params, dt)


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 1e-5)])
def test_mcwf_dephasing_matches_oracle(ctx, dtype, tol):
    """lindblad_mcwf_step_trotter! with σ_z jumps: compiled plan for the XXZ
    gate list, then the jump kernel; jump sites bit-exact, states to tol."""
    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_trotter
    from smoq_amd.streams import uniforms

    N, T, steps, gamma, dt = 5, 6, 40, 0.5, 0.05
    gates = xxz_x_gates(N, dt)
    psi0 = random_states(N, T, 9)
    U = uniforms(55, T * steps * 2).reshape(T, steps, 2)
    dev = GPUState.from_numpy(psi0.astype(np.complex64 if dtype == "c64" else np.complex128))
    jumps, _ = evolve_mcwf_trotter(dev, gates, N, T, steps, gamma, dt, jump="sigma_z",
                                   uniforms=U)
    gl = [(list(g.indices), g.U) for g in gates]
    got = dev.to_numpy()
    for t in range(T):
        p = psi0[t].copy()
        want = orc.mcwf_dephasing_trajectory(p, gl, N, gamma * dt, U[t, :, 0], U[t, :, 1])
        np.testing.assert_array_equal(jumps[t], want)
        assert np.max(np.abs(got[t] - p)) < tol
    assert (jumps >= 0).sum() > 10


def test_mcwf_plan_path_sigma_minus_equals_persistent_kernel(ctx):
    """The same σ⁻ trajectories through the persistent kernel and through
    compiled plan + jump kernel: identical jump sites, states to 1e-12."""
    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_batched_trotter, evolve_mcwf_trotter
    from smoq_amd.streams import uniforms

    N, T, steps, gamma, dt = 10, 8, 50, 0.8, 0.05
    gates = chain_gates(N, dt, seed=5)
    psi0 = random_states(N, T, 12)
    U = uniforms(77, T * steps * 2).reshape(T, steps, 2)
    a = GPUState.from_numpy(psi0)
    ra = evolve_mcwf_batched_trotter(a, gates, N, T, steps, gamma, dt,
                                     uniforms=np.ascontiguousarray(U[:, :, 0]))
    b = GPUState.from_numpy(psi0)
    jb, _ = evolve_mcwf_trotter(b, gates, N, T, steps, gamma, dt, jump="sigma_minus", uniforms=U)
    np.testing.assert_array_equal(ra.jumps, jb)
    assert np.max(np.abs(a.to_numpy() - b.to_numpy())) < 1e-12


def test_mcwf_dephasing_reproduces_the_reference_demo(ctx):
    """The reference's σ_z-dephasing demo (XXZ chain J = (1, 1, 0.5), h_x = 1,
    γ = 0.1, dt = 0.05, 200 steps from |0…0>): trajectory means of the six
    observables at t = 10 from 8192 trajectories against the exact Lindblad
    values, within 5 standard errors + 0.006 (first-order stepping; the
    reference's own DM+Exact column differs from the exact values by up to
    0.0026), and the exact values against that recorded column
    (tests/golden/mcwf_dephasing_xxz.json)."""
    import json
    import os

    from smoq_amd import GPUState
    from smoq_amd.mcwf import evolve_mcwf_trotter

    with open(os.path.join(os.path.dirname(__file__), "golden", "mcwf_dephasing_xxz.json")) as f:
        ref = json.load(f)
    P = ref["params"]
    T, steps = 8192, P["n_steps"]
    for N in (2, 4, 6, 8):
        xm, zm, kinds = [], [], []
        for k in range(N):
            xm += [1 << k, 1 << k, 0]
            zm += [0, 1 << k, 1 << k]
            kinds += [0, 1, 2]
        for k in range(N - 1):
            pr = 3 << k
            xm += [pr, pr, 0]
            zm += [0, pr, pr]
            kinds += [3, 4, 5]
        psi = GPUState.basis(N, 0, batch=T)
        _, out = evolve_mcwf_trotter(psi, xxz_x_gates(N, P["dt"], tuple(P["J"]), P["hx"]), N, T,
                                     steps, P["gamma"], P["dt"], jump="sigma_z", seed=100 + N,
                                     obs_every=steps, xmask=xm, zmask=zm)
        kinds = np.array(kinds)
        per_traj = np.stack([out[0][:, kinds == c].mean(1) for c in range(6)], 1)  # [traj, 6]
        mean, se = per_traj.mean(0), per_traj.std(0) / np.sqrt(T)
        exact = np.array(ref["exact"][str(N)])
        assert np.all(np.abs(mean - exact) < 5 * se + 0.006), (N, mean, exact)
