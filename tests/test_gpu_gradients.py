"""Adjoint gradients on the device against the oracle's parameter-shift
gradient (exact for exp(-iθP/2) with P a Pauli string), central finite
differences and the batched / complex64 variants."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def tfim_terms(N, J=1.0, h=1.0):
    from smoq_amd.observables import xxz_terms

    return xxz_terms(N, 0.0, 0.0, J, h)


def oracle_energy(tape, theta, N, J=1.0, h=1.0):
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1.0
    for gate, qs, th, par, m in tape.ops:
        if par >= 0:
            th = theta[par]
        orc.apply_named(psi, gate.lower(), [q + 1 for q in qs], th, N)
    return orc.compute_energy(psi, N, 0.0, 0.0, J, h)


def oracle_param_shift(tape, theta, N):
    g = np.zeros_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        g[k] = (oracle_energy(tape, tp, N) - oracle_energy(tape, tm, N)) / 2.0
    return g


@pytest.fixture(params=[0, 2], ids=["per_gate_sweep", "fused_sweep"])
def jit_mode(request, ctx):
    # 0: the gate-by-gate reverse sweep; 2: fused passes with overlap gates
    # (compiled passes forced, so a compilation failure is an error)
    ctx.set_jit(request.param)
    yield request.param
    ctx.set_jit(1)


def test_adjoint_matches_oracle_parameter_shift(ctx, jit_mode):
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy

    N, L = 7, 3
    tape = hardware_efficient_ansatz(N, L)
    theta = np.random.default_rng(1).uniform(0, 2 * math.pi, tape.n_params)
    xs, zs, cs = tfim_terms(N)
    v = VQEEnergy(tape, xs, zs, cs)
    e, g = v.energy_and_grad(theta)
    assert abs(e[0] - oracle_energy(tape, theta, N)) < 1e-12
    want = oracle_param_shift(tape, theta, N)
    assert np.max(np.abs(g[0] - want)) < 1e-10


def test_all_rotation_kinds_against_parameter_shift_and_fd(ctx, jit_mode):
    from smoq_amd.circuits import Tape
    from smoq_amd.gradients import VQEEnergy, calculate_gradients_finite_difference

    N = 6
    rng = np.random.default_rng(3)
    t = Tape(N)
    p = 0
    for rep in range(2):
        for q in range(N):
            t.add("H", [q]) if rep == 0 else None
            t.add(["RX", "RY", "RZ"][(q + rep) % 3], [q], param=p)
            p += 1
        for q in range(0, N - 1):
            t.add(["RXX", "RYY", "RZZ"][(q + rep) % 3], [q, q + 1], param=p)
            p += 1
        t.add("CZ", [0, N - 1]).add("T", [2])
    n_ps = p
    t.add("CRY", [1, 4], param=p)
    theta = rng.uniform(0, 2 * math.pi, t.n_params)
    xs, zs, cs = tfim_terms(N, 0.7, 1.3)
    v = VQEEnergy(t, xs, zs, cs)
    e, g = v.energy_and_grad(theta)
    # parameter shift is exact for every Pauli-string rotation (all but CRY)
    ps = np.zeros(n_ps)
    for k in range(n_ps):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps[k] = (oracle_energy(t, tp, N, 0.7, 1.3) - oracle_energy(t, tm, N, 0.7, 1.3)) / 2
    assert np.max(np.abs(g[0][:n_ps] - ps)) < 1e-10
    fd = calculate_gradients_finite_difference(v.cost_fn(), theta, 1e-6)
    assert np.max(np.abs(g[0] - fd)) < 1e-7


def test_batched_adjoint_equals_single(ctx):
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy

    N, L, B = 9, 2, 3
    tape = hardware_efficient_ansatz(N, L)
    thetas = np.random.default_rng(5).uniform(0, 2 * math.pi, (B, tape.n_params))
    xs, zs, cs = tfim_terms(N)
    vb = VQEEnergy(tape, xs, zs, cs, batch=B)
    eb, gb = vb.energy_and_grad(thetas)
    v1 = VQEEnergy(tape, xs, zs, cs)
    for b in range(B):
        e1, g1 = v1.energy_and_grad(thetas[b])
        assert abs(eb[b] - e1[0]) < 1e-12
        assert np.max(np.abs(gb[b] - g1[0])) < 1e-12
    # plain one-gate-kernel energies agree with the fused forward pass
    assert abs(v1.energy(thetas[0], one_gate_kernels=True)[0] - v1.energy(thetas[0])[0]) < 1e-12


@pytest.mark.parametrize("N,L,B,dtype", [(12, 3, 4, "c128"), (13, 2, 5, "c64"),
                                         (16, 2, 3, "c128")])
def test_batched_fused_adjoint_matches_single_sets(ctx, N, L, B, dtype):
    """A batch of parameter sets runs each pass once for every item (each
    item its own copy of the compiled programs): energies and gradients
    equal the single-set sweep's, item by item, one item with all angles
    zero included."""
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy

    tape = hardware_efficient_ansatz(N, L)
    thetas = np.random.default_rng(50 + N).uniform(0, 2 * math.pi, (B, tape.n_params))
    thetas[B - 1] = 0.0
    xs, zs, cs = tfim_terms(N)
    print(f"[adjoint batch] N={N} L={L} B={B} {dtype}", flush=True)
    eb, gb = VQEEnergy(tape, xs, zs, cs, batch=B, dtype=dtype).energy_and_grad(thetas)
    v1 = VQEEnergy(tape, xs, zs, cs, dtype=dtype)
    tol = 1e-11 if dtype == "c128" else 2e-5
    for b in range(B):
        e1, g1 = v1.energy_and_grad(thetas[b])
        assert abs(eb[b] - e1[0]) < tol, b
        assert np.max(np.abs(gb[b] - g1[0])) < tol, b


def test_adjoint_complex64(ctx, jit_mode):
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy

    N, L = 10, 3
    tape = hardware_efficient_ansatz(N, L)
    theta = np.random.default_rng(7).uniform(0, 2 * math.pi, tape.n_params)
    xs, zs, cs = tfim_terms(N)
    e128, g128 = VQEEnergy(tape, xs, zs, cs).energy_and_grad(theta)
    e64, g64 = VQEEnergy(tape, xs, zs, cs, dtype="c64").energy_and_grad(theta)
    assert abs(e64[0] - e128[0]) < 1e-5
    assert np.max(np.abs(g64[0] - g128[0])) < 1e-5


def test_fused_adjoint_many_passes_against_oracle(ctx, jit_mode):
    # 14 qubits: the reverse sweep's 15-qubit buffer spans several tiles and
    # passes; every rotation kind, controls and fixed gates in between
    from smoq_amd.circuits import Tape
    from smoq_amd.gradients import VQEEnergy

    N = 14
    rng = np.random.default_rng(11)
    t = Tape(N)
    p = 0
    for layer in range(3):
        for q in range(N):
            t.add(["RY", "RZ", "RX"][(q + layer) % 3], [q], param=p)
            p += 1
        for q in range(layer % 2, N - 1, 2):
            t.add("CZ", [q, q + 1])
        t.add(["RXX", "RYY", "RZZ"][layer], [layer, N - 1 - layer], param=p)
        p += 1
        t.add("H", [5]).add("CNOT", [3, 11]).add("T", [13])
    theta = rng.uniform(0, 2 * math.pi, t.n_params)
    xs, zs, cs = tfim_terms(N, 0.9, 1.1)
    v = VQEEnergy(t, xs, zs, cs)
    e, g = v.energy_and_grad(theta)
    assert abs(e[0] - oracle_energy(t, theta, N, 0.9, 1.1)) < 1e-11
    ps = np.zeros(t.n_params)
    for k in range(t.n_params):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps[k] = (oracle_energy(t, tp, N, 0.9, 1.1) - oracle_energy(t, tm, N, 0.9, 1.1)) / 2.0
    assert np.max(np.abs(g[0] - ps)) < 1e-10
    # the state keeps U(θ)|0>, and a second call gives the same gradient
    e2, g2 = v.energy_and_grad(theta)
    assert np.array_equal(g2, g) and e2[0] == e[0]
    v.free()
