"""The density-matrix Lindblad fixture (tests/golden/lindblad_tfim_damping.json)
and its generator, pinned by an independent matrix-form RK4 integrator and
by the closed form of pure amplitude damping."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def gen():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "make_lindblad_dm", os.path.join(HERE, "golden", "make_lindblad_dm.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def rk4_z(N, J, h, gamma, dt, n_steps):
    """ρ as a dense matrix: -i[H, ρ] - γ/2 (popcount(i) + popcount(j)) ρ_ij
    + γ Σ_k ρ_{i|m, j|m} on the rows/columns with bit k clear."""
    D = 1 << N
    idx = np.arange(D)
    zs = 1.0 - 2.0 * ((idx[:, None] >> np.arange(N)[None, :]) & 1)
    H = np.diag(J * sum(zs[:, k] * zs[:, k + 1] for k in range(N - 1))).astype(np.complex128)
    for k in range(N):
        H[idx, idx ^ (1 << k)] += h
    pc = np.array([bin(i).count("1") for i in range(D)], np.float64)
    anti = 0.5 * gamma * (pc[:, None] + pc[None, :])
    lows = [idx[(idx >> k) & 1 == 0] for k in range(N)]

    def L(r):
        out = -1j * (H @ r - r @ H) - anti * r
        for k, lo in enumerate(lows):
            hi = lo | (1 << k)
            out[np.ix_(lo, lo)] += gamma * r[np.ix_(hi, hi)]
        return out

    rho = np.zeros((D, D), np.complex128)
    rho[-1, -1] = 1.0
    for _ in range(n_steps):
        k1 = L(rho)
        k2 = L(rho + 0.5 * dt * k1)
        k3 = L(rho + 0.5 * dt * k2)
        k4 = L(rho + dt * k3)
        rho = rho + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return rho.diagonal().real @ zs


def test_generator_matches_matrix_rk4_small_chain():
    z = gen().z_history(3, 1.0, 0.7, 0.3, [0.5, 1.0])
    assert np.max(np.abs(z[1] - rk4_z(3, 1.0, 0.7, 0.3, 0.005, 200))) < 1e-9


def test_generator_pure_damping_closed_form():
    t = np.array([1.0, 2.0, 3.0])
    z = gen().z_history(2, 0.0, 0.0, 0.4, t)
    want = 1.0 - 2.0 * np.exp(-0.4 * t)
    assert np.max(np.abs(z - want[:, None])) < 1e-12


def test_fixture_first_record_matches_rk4():
    with open(os.path.join(HERE, "golden", "lindblad_tfim_damping.json")) as f:
        d = json.load(f)
    assert d["sizes"] == [8, 10] and d["times"][0] == 1.0
    z8 = np.array(d["z"]["8"])
    assert z8.shape == (10, 8)
    z = rk4_z(8, d["J"], d["h"], d["gamma"], 0.01, 100)
    assert np.max(np.abs(z8[0] - z)) < 1e-8
    # the chain's mirror symmetry
    for key in ("8", "10"):
        zz = np.array(d["z"][key])
        assert np.max(np.abs(zz - zz[:, ::-1])) < 1e-10


def test_dephasing_fixture_matches_the_reference_run():
    """The exact σ_z-dephasing Lindblad values (generator) against the
    reference's recorded DM+Exact column (output.log, 4 decimals): they
    differ only by the reference's first-order time stepping."""
    import importlib.util

    with open(os.path.join(HERE, "golden", "mcwf_dephasing_xxz.json")) as f:
        d = json.load(f)
    assert d["rows"] == ["X", "Y", "Z", "XX", "YY", "ZZ"]
    for N in ("2", "4", "6", "8"):
        exact = np.array(d["exact"][N])
        dm = np.array(d["reference"][N]["DM+Exact"])
        assert np.max(np.abs(exact - dm)) < 3e-3, (N, exact - dm)
    spec = importlib.util.spec_from_file_location(
        "make_dephasing_fixture", os.path.join(HERE, "golden", "make_dephasing_fixture.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    assert np.max(np.abs(np.array(m.exact_values(2)) - np.array(d["exact"]["2"]))) < 1e-12
