"""Batched adjoint gradients over parameter sets of different qubit counts in
one persistent launch (smoq_adjoint_gradient_batch): every item against the
oracle's parameter-shift gradient and energy (≤1e-10 / 1e-12), against the
single-state adjoint (smoq_adjoint_gradient) and complex64 against
complex128 (≤1e-5)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def tfim_terms(N, J=1.0, h=1.0):
    from smoq_amd.observables import xxz_terms

    return xxz_terms(N, 0.0, 0.0, J, h)


def oracle_energy(tape, theta, N):
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1.0
    for gate, qs, th, par, m in tape.ops:
        if par >= 0:
            th = theta[par]
        orc.apply_named(psi, gate.lower(), [q + 1 for q in qs], th, N)
    return orc.compute_energy(psi, N, 0.0, 0.0, 1.0, 1.0)


def oracle_param_shift(tape, theta, N):
    g = np.zeros_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        g[k] = (oracle_energy(tape, tp, N) - oracle_energy(tape, tm, N)) / 2.0
    return g


def mixed_tape(N, seed):
    """Every rotation kind (RX RY RZ RXX RYY RZZ CRY) plus fixed gates."""
    from smoq_amd.circuits import Tape

    rng = np.random.default_rng(seed)
    t = Tape(N)
    p = 0
    for q in range(N):
        t.add("H", [q])
    for rep in range(2):
        for q in range(N):
            t.add(["RX", "RY", "RZ"][(q + rep) % 3], [q], param=p)
            p += 1
        for q in range(N - 1):
            t.add(["RXX", "RYY", "RZZ", "CRY"][(q + rep) % 4], [q, q + 1], param=p)
            p += 1
        t.add("CZ", [0, N - 1]).add("T", [1]).add("CNOT", [N - 1, 0])
        t.add("RY", [int(rng.integers(N))], theta=float(rng.normal()))
    return t


def test_uneven_sizes_match_oracle_parameter_shift(ctx):
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import adjoint_gradient_batch

    sizes = [4, 7, 5, 9, 3, 6]
    rng = np.random.default_rng(11)
    items = []
    for N in sizes:
        tape = hardware_efficient_ansatz(N, 2)
        th = rng.uniform(0, 2 * math.pi, tape.n_params)
        items.append((tape, th) + tuple(tfim_terms(N)))
    print(f"[gpu] vqe batch sizes={sizes}", flush=True)
    E, G = adjoint_gradient_batch(items)
    for (tape, th, *_), e, g, N in zip(items, E, G, sizes):
        assert abs(e - oracle_energy(tape, th, N)) < 1e-12
        assert np.max(np.abs(g - oracle_param_shift(tape, th, N))) < 1e-10


@pytest.mark.parametrize("tile,route", [(0, "one_launch"), (12, "one_launch"), (0, "auto"),
                                        (0, "sweep")])
def test_batch_matches_single_state_adjoint(ctx, tile, route):
    """One launch for every item (at the default tile and at 2^12), the
    automatic split between the launch and the whole-device sweep, and the
    sweep alone, against the single-state adjoint."""
    from smoq_amd.gradients import VQEEnergy, adjoint_gradient_batch

    sizes = [10, 13, 11, 16, 12, 10, 14, 18]
    items = []
    for i, N in enumerate(sizes):
        tape = mixed_tape(N, i)
        th = np.random.default_rng(100 + i).uniform(0, 2 * math.pi, tape.n_params)
        items.append((tape, th) + tuple(tfim_terms(N)))
    print(f"[gpu] vqe batch vs single sizes={sizes} tile={tile} route={route}", flush=True)
    ctx.set_fusion(tile if tile else -1, 0)
    if route != "auto":
        os.environ["SMOQ_VQE_ROUTE"] = route
    try:
        E, G = adjoint_gradient_batch(items)
    finally:
        ctx.set_fusion(-1, 0)
        os.environ.pop("SMOQ_VQE_ROUTE", None)
    for (tape, th, xs, zs, cs), e, g in zip(items, E, G):
        v = VQEEnergy(tape, xs, zs, cs)
        e1, g1 = v.energy_and_grad(th)
        v.free()
        assert abs(e - e1[0]) < 1e-11
        assert np.max(np.abs(g - g1[0])) < 1e-10


@pytest.mark.parametrize("route", ["sweep", "auto"])
def test_sweep_groups_items_of_one_tape(ctx, route):
    """Large items that share a tape and Hamiltonian go through the sweep as
    one batch of parameter sets; interleaved with items of other tapes (a
    copy of the tape in its own arrays counts as the same problem), each
    against the single-state adjoint."""
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy, adjoint_gradient_batch

    N = 16
    tape = mixed_tape(N, 3)
    other = mixed_tape(N, 4)
    small = hardware_efficient_ansatz(9, 2)
    rng = np.random.default_rng(21)
    items = []
    for k in range(7):
        t = other if k == 2 else (small if k == 5 else (tape if k != 4 else mixed_tape(N, 3)))
        n = 9 if t is small else N
        terms = tfim_terms(n)
        items.append((t, rng.uniform(0, 2 * math.pi, t.n_params)) + tuple(terms))
    print(f"[gpu] vqe batch grouped sweep N={N} route={route}", flush=True)
    if route != "auto":
        os.environ["SMOQ_VQE_ROUTE"] = route
    try:
        E, G = adjoint_gradient_batch(items)
    finally:
        os.environ.pop("SMOQ_VQE_ROUTE", None)
    for (t, th, xs, zs, cs), e, g in zip(items, E, G):
        v = VQEEnergy(t, xs, zs, cs)
        e1, g1 = v.energy_and_grad(th)
        v.free()
        assert abs(e - e1[0]) < 1e-11
        assert np.max(np.abs(g - g1[0])) < 1e-10


def test_complex64_batch_against_complex128(ctx):
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import adjoint_gradient_batch

    sizes = [8, 12, 6, 15]
    items = []
    for i, N in enumerate(sizes):
        tape = hardware_efficient_ansatz(N, 3)
        th = np.random.default_rng(7 + i).uniform(0, 2 * math.pi, tape.n_params)
        items.append((tape, th) + tuple(tfim_terms(N)))
    E64, G64 = adjoint_gradient_batch(items, dtype="c64")
    E128, G128 = adjoint_gradient_batch(items, dtype="c128")
    for e64, e128, g64, g128 in zip(E64, E128, G64, G128):
        assert abs(e64 - e128) < 1e-5
        assert np.max(np.abs(g64 - g128)) < 1e-5


def test_many_identical_structures_and_repeat_calls(ctx):
    """64 parameter sets of one N=12 ansatz (the batched form of a VQE step),
    twice in a row: the second call reuses the work buffers and must give the
    same numbers bit for bit."""
    from smoq_amd.circuits import hardware_efficient_ansatz
    from smoq_amd.gradients import VQEEnergy, adjoint_gradient_batch

    N = 12
    tape = hardware_efficient_ansatz(N, 4)
    xs, zs, cs = tfim_terms(N)
    rng = np.random.default_rng(5)
    thetas = rng.uniform(0, 2 * math.pi, (64, tape.n_params))
    items = [(tape, th, xs, zs, cs) for th in thetas]
    E1, G1 = adjoint_gradient_batch(items)
    E2, G2 = adjoint_gradient_batch(items)
    assert np.array_equal(E1, E2)
    assert all(np.array_equal(a, b) for a, b in zip(G1, G2))
    v = VQEEnergy(tape, xs, zs, cs)
    for k in (0, 31, 63):
        e, g = v.energy_and_grad(thetas[k])
        assert abs(E1[k] - e[0]) < 1e-11
        assert np.max(np.abs(G1[k] - g[0])) < 1e-10
    v.free()


def test_invalid_items_rejected(ctx):
    from smoq_amd._lib import SmoqError
    from smoq_amd.circuits import Tape
    from smoq_amd.gradients import adjoint_gradient_batch

    t = Tape(2).add("RY", [0], param=0)
    with pytest.raises(SmoqError):
        adjoint_gradient_batch([(t, [0.1], [0], [1], [1.0])])  # 2 qubits < 3
    t = Tape(4).add("RY", [0], param=0)
    with pytest.raises(SmoqError):
        adjoint_gradient_batch([(t, [0.1], [0], [1 << 5], [1.0])])  # term outside 4 qubits
