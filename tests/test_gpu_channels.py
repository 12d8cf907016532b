"""Noise channels, Z measurement, forced reset and sampling on the device
(smoq_qubit_channel, smoq_dm_qubit_channel, SMOQ_SAMPLE_SEQUENTIAL_DIAG)
against the oracle's restatements, item by item on the same uniforms:
records equal, amplitudes within 1e-12 (complex128) / 2e-5 (complex64)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KINDS = ["dephasing", "depolarizing", "amplitude_damping", "bit_flip", "phase_flip"]
H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)
S = np.diag([1, 1j])
SDG = np.diag([1, -1j])


def rand_batch(N, B, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(B, 1 << N)) + 1j * rng.normal(size=(B, 1 << N))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def rand_rho(N, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(1 << N, 1 << N)) + 1j * rng.normal(size=(1 << N, 1 << N))
    r = A @ A.conj().T
    return r / np.trace(r).real


def spread(B, n, seed):
    """Uniforms that hit every branch interval: a low-discrepancy spread."""
    i = np.arange(B * n, dtype=np.float64)
    return ((i * 0.6180339887498949 + 0.013 * seed) % 1.0).reshape(B, n)


def upload(psi, dtype):
    from smoq_amd import GPUState

    return GPUState.from_numpy(psi.astype(np.complex64) if dtype == "c64" else psi)


# N = 14 with batch 2 runs four chunks per item, so qubits 13 and 14 pair
# amplitudes across chunks
CASES = [(6, 3, "c128", 1e-12), (14, 2, "c128", 1e-12), (14, 2, "c64", 2e-5)]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("N,B,dtype,tol", CASES)
def test_psi_channels_match_oracle(ctx, kind, N, B, dtype, tol):
    from smoq_amd.channels import qubit_channel

    print(f"[case] {kind} N={N} B={B} {dtype}", flush=True)
    qubits = [1, N, 2, N - 1, 3]
    p = 0.35
    psi0 = rand_batch(N, B, N + B)
    if dtype == "c64":
        psi0 = psi0.astype(np.complex64).astype(np.complex128)
    u = spread(B, len(qubits), 1)
    s = upload(psi0, dtype)
    rec = np.atleast_2d(qubit_channel(s, kind, p, qubits, u))
    got = s.to_numpy().reshape(B, -1)
    for b in range(B):
        v = psi0[b].copy()
        want_rec = [orc.psi_channel(v, kind, p, q, N, u[b, k]) for k, q in enumerate(qubits)]
        assert list(rec[b]) == want_rec
        assert np.max(np.abs(got[b] - v)) < tol
    s.free()


@pytest.mark.parametrize("N,B,dtype,tol", CASES)
def test_measure_z_matches_oracle(ctx, N, B, dtype, tol):
    from smoq_amd.measurements import projective_measurement

    qubits = [N, 1, 3]
    psi0 = rand_batch(N, B, 40 + N)
    if dtype == "c64":
        psi0 = psi0.astype(np.complex64).astype(np.complex128)
    u = spread(B, len(qubits), 2)
    s = upload(psi0, dtype)
    out, _ = projective_measurement(s, qubits, "z", N, u if B > 1 else u[0])
    out = np.atleast_2d(out)
    got = s.to_numpy().reshape(B, -1)
    for b in range(B):
        v = psi0[b].copy()
        want = [orc.measure_z(v, q, N, u[b, k]) for k, q in enumerate(qubits)]
        assert list(out[b]) == want
        assert np.max(np.abs(got[b] - v)) < tol
    s.free()


def test_measure_in_x_and_y_bases(ctx):
    from smoq_amd.measurements import projective_measurement, sample_state

    N, B = 5, 4
    qubits, bases = [2, 5, 1, 4], ["x", "y", "z", "y"]
    psi0 = rand_batch(N, B, 3)
    u = spread(B, 4, 3)
    s = upload(psi0, "c128")
    before = s.to_numpy()
    peek = sample_state(s, qubits, bases, N, u)
    assert np.array_equal(s.to_numpy(), before)  # sample_state leaves ψ alone
    out, _ = projective_measurement(s, qubits, bases, N, u)
    assert np.array_equal(peek, out)
    got = s.to_numpy()
    for b in range(B):
        v = psi0[b].copy()
        for k, (q, ba) in enumerate(zip(qubits, bases)):
            pre = {"x": [H], "y": [SDG, H], "z": []}[ba]
            post = {"x": [H], "y": [H, S], "z": []}[ba]
            for U in pre:
                orc.apply_1q(v, U, q, N)
            assert out[b, k] == orc.measure_z(v, q, N, u[b, k])
            for U in post:
                orc.apply_1q(v, U, q, N)
        assert np.max(np.abs(got[b] - v)) < 1e-12
    s.free()


@pytest.mark.parametrize("N,B,dtype,tol", CASES)
def test_reset_state_matches_oracle(ctx, N, B, dtype, tol):
    from smoq_amd.measurements import reset_state

    qubits, targets = [1, N, 2, 4], ["zero", "one", "plus", "minus"]
    psi0 = rand_batch(N, B, 70 + N)
    if dtype == "c64":
        psi0 = psi0.astype(np.complex64).astype(np.complex128)
    u = spread(B, 4, 4)
    s = upload(psi0, dtype)
    reset_state(s, qubits, targets, N, u if B > 1 else u[0])
    got = s.to_numpy().reshape(B, -1)
    for b in range(B):
        want = orc.reset_state(psi0[b].copy(), qubits, targets, N, u[b])
        assert np.max(np.abs(got[b] - want)) < tol
    s.free()


def test_force_records_and_all_qubit_reset(ctx):
    from smoq_amd.channels import qubit_channel
    from smoq_amd.measurements import reset_state

    N = 4
    psi0 = rand_batch(N, 1, 5)[0]
    s = upload(psi0, "c128")
    assert list(qubit_channel(s, "force_one", 0.0, [2, 3])) == [1, 1]
    v = psi0.copy()
    for q in (2, 3):
        lo, hi = orc._halves(N, q)
        v[hi] += v[lo]
        v[lo] = 0
        v /= np.sqrt(np.sum(np.abs(v) ** 2))
    assert np.max(np.abs(s.to_numpy() - v)) < 1e-12
    u = [0.5, 0.1, 0.9, 0.3]
    reset_state(s, "zero", N, u)  # reset_state!(ψ, :zero, N)
    want = orc.reset_state(v, [1, 2, 3, 4], ["zero"] * N, N, u)
    assert abs(abs(want[0]) - 1) < 1e-12  # |0000> up to the kept amplitude's phase
    assert np.max(np.abs(s.to_numpy() - want)) < 1e-12
    s.free()


def test_projective_measurement_all_psi_and_rho(ctx):
    from smoq_amd import GPUState
    from smoq_amd.density import DensityMatrix
    from smoq_amd.measurements import projective_measurement_all, sample_state

    N = 5
    psi = rand_batch(N, 1, 9)[0]
    rho = rand_rho(N, 9)
    for r in (0.0, 0.137, 0.5, 0.9, 0.999999):
        s = GPUState.from_numpy(psi)
        bits, _ = projective_measurement_all(s, N, r)
        i = orc.sample_from_probabilities(psi.real ** 2 + psi.imag ** 2, r)
        assert bits == [(i >> k) & 1 for k in range(N)]
        e = np.zeros(1 << N)
        e[i] = 1
        assert np.array_equal(s.to_numpy(), e)
        s.free()
        d = DensityMatrix.from_numpy(rho)
        j = orc.sample_from_probabilities(np.diag(rho).real, r)
        assert sample_state(d, [1, 3, 5], "z", N, r) == [(j >> (k - 1)) & 1 for k in (1, 3, 5)]
        bits, _ = projective_measurement_all(d, N, r)
        assert bits == [(j >> k) & 1 for k in range(N)]
        E = np.zeros((1 << N, 1 << N))
        E[j, j] = 1
        assert np.array_equal(d.to_numpy(), E)
        d.free()


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("N,dtype,tol", [(3, "c128", 1e-12), (5, "c128", 1e-12),
                                         (5, "c64", 2e-6)])
def test_rho_channels_match_oracle(ctx, kind, N, dtype, tol):
    from smoq_amd import channels
    from smoq_amd.density import DensityMatrix

    rho = rand_rho(N, 20 + N)
    if dtype == "c64":
        rho = rho.astype(np.complex64).astype(np.complex128)
    qubits = [1, N, 2]
    d = DensityMatrix.from_numpy(rho.astype(np.complex64) if dtype == "c64" else rho)
    getattr(channels, f"apply_channel_{kind}")(d, 0.3, qubits, N)
    want = rho.copy()
    for q in qubits:
        orc.dm_channel(want, kind, 0.3, q, N)
    assert np.max(np.abs(d.to_numpy() - want)) < tol
    assert abs(d.trace() - 1.0) < 10 * tol
    d.free()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_rho_measure_and_force_match_oracle(ctx, dtype, tol):
    from smoq_amd.density import DensityMatrix
    from smoq_amd.measurements import projective_measurement, reset_state

    N, B = 4, 3
    rhos = np.stack([rand_rho(N, 30 + b) for b in range(B)])
    if dtype == "c64":
        rhos = rhos.astype(np.complex64).astype(np.complex128)
    qubits = [2, 4, 1]
    u = spread(B, 3, 5)
    d = DensityMatrix.from_numpy(rhos.astype(np.complex64) if dtype == "c64" else rhos)
    out, _ = projective_measurement(d, qubits, "z", N, u)
    got = d.to_numpy()
    for b in range(B):
        r = rhos[b].copy()
        assert list(out[b]) == [orc.dm_measure_z(r, q, N, u[b, k]) for k, q in enumerate(qubits)]
        assert np.max(np.abs(got[b] - r)) < tol
    d.free()
    d = DensityMatrix.from_numpy(rhos.astype(np.complex64) if dtype == "c64" else rhos)
    reset_state(d, [3, 1], ["one", "zero"], N)
    got = d.to_numpy()
    for b in range(B):
        r = orc.dm_force(orc.dm_force(rhos[b].copy(), 3, 1, N), 1, 0, N)
        assert np.max(np.abs(got[b] - r)) < tol
    with pytest.raises(ValueError):
        reset_state(d, [1], ["plus"], N)
    with pytest.raises(NotImplementedError):
        projective_measurement(d, [1], "x", N, [0.5] * B)
    d.free()


def test_demo_channel_assertions_on_device(ctx):
    """demo_quantum_channels.jl TEST 1.1-1.5 (:259-368) through the device."""
    from smoq_amd import channels as ch
    from smoq_amd.density import DensityMatrix

    d = DensityMatrix.product(["plus", "plus"])
    r0 = d.to_numpy()
    ch.apply_channel_dephasing(d, 0.5, [1], 2)
    r = d.to_numpy()
    assert np.linalg.norm(np.diag(r) - np.diag(r0)) < 1e-12
    assert abs(abs(r[0, 1]) / abs(r0[0, 1]) - np.sqrt(0.5)) < 1e-10
    assert abs(d.trace() - 1) < 1e-12
    ch.apply_channel_dephasing(d, 1.0, [1, 2], 2)
    r = d.to_numpy()
    assert np.abs(r).sum() - np.abs(np.diag(r)).sum() < 1e-10
    d.free()
    d = DensityMatrix.product(["one"])
    ch.apply_channel_amplitude_damping(d, 0.5, [1], 1)
    r = d.to_numpy()
    assert abs(r[0, 0] - 0.5) < 1e-12 and abs(r[1, 1] - 0.5) < 1e-12
    d.free()
    for kind, p, want in (("amplitude_damping", 0.9, [1, 0]), ("bit_flip", 0.5, [0.5, 0.5]),
                          ("bit_flip", 1.0, [0, 1]), ("depolarizing", 0.75, [0.5, 0.5])):
        d = DensityMatrix.product(["zero"])
        getattr(ch, f"apply_channel_{kind}")(d, p, [1], 1)
        r = d.to_numpy()
        assert np.max(np.abs(np.diag(r) - want)) < 1e-12, kind
        d.free()
    d = DensityMatrix.product(["plus"])
    ch.apply_channel_phase_flip(d, 0.25, [1], 1)
    r = d.to_numpy()
    assert abs(abs(r[0, 1]) - 0.25) < 1e-12 and abs(r[0, 0] - 0.5) < 1e-12
    d.free()


def test_demo_trajectory_tables_match_density_matrix(ctx):
    """demo_quantum_channels.jl's trajectory-vs-ρ tables: |+>^6 with p = 0.2 on
    qubit 1, 100 trajectories per channel (each trajectory equal to the
    oracle's on the same uniform), and |0>/|1> single-qubit bit flip /
    amplitude damping with 10000 trajectories, whose <Z> averages lie within
    5 standard errors of Tr(ρ Z)."""
    from smoq_amd import channels as ch
    from smoq_amd import make_product_ket, streams
    from smoq_amd.density import DensityMatrix
    from smoq_amd.observables import expect_pauli

    N, T, p = 6, 100, 0.2
    plus = orc.product_ket([(1 / np.sqrt(2), 1 / np.sqrt(2))] * N)
    for kind in KINDS:
        s = make_product_ket(["plus"] * N, batch=T)
        u = streams.uniforms(7, T).reshape(T, 1)
        rec = ch.qubit_channel(s, kind, p, [1], u)
        got = s.to_numpy()
        for t in range(0, T, 9):
            v = plus.copy()
            assert rec[t, 0] == orc.psi_channel(v, kind, p, 1, N, u[t, 0])
            assert np.max(np.abs(got[t] - v)) < 1e-12
        d = DensityMatrix.product(["plus"] * N)
        getattr(ch, f"apply_channel_{kind}")(d, p, [1], N)
        xm, zm = [1, 0, 1 << (N - 1)], [0, 1, 0]
        traj = expect_pauli(s, xm, zm).mean(axis=0)
        exact = d.expect_pauli(xm, zm)
        sd = np.sqrt(np.maximum(1 - exact ** 2, 1e-3) / T)
        assert np.all(np.abs(traj - exact) < 5 * sd + 1e-12), kind
        s.free()
        d.free()
    for kind, start, p in (("bit_flip", "zero", 0.3), ("amplitude_damping", "one", 0.4)):
        T = 10000
        s = make_product_ket([start], batch=T)
        ch.qubit_channel(s, kind, p, [1], streams.uniforms(11, T).reshape(T, 1))
        z = expect_pauli(s, [0], [1])[:, 0]
        d = DensityMatrix.product([start])
        getattr(ch, f"apply_channel_{kind}")(d, p, [1], 1)
        zx = float(d.expect_pauli([0], [1])[0])
        assert abs(z.mean() - zx) < 5 * np.sqrt((1 - zx ** 2) / T), kind
        s.free()
        d.free()


def test_errors_are_reported(ctx):
    from smoq_amd import GPUState, SmoqError
    from smoq_amd._lib import SMOQ_EINVAL, SMOQ_ENUMERIC
    from smoq_amd.channels import qubit_channel

    s = GPUState.from_numpy(np.zeros(8, np.complex128))
    with pytest.raises(SmoqError) as e:  # an all-zero state has no outcome to keep
        qubit_channel(s, "measure_z", 0.0, [1], [0.5])
    assert e.value.code == SMOQ_ENUMERIC
    s.free()
    s = GPUState.from_numpy(rand_batch(3, 1, 0)[0])
    with pytest.raises(SmoqError) as e:
        qubit_channel(s, "bit_flip", 1.5, [1], [0.5])
    assert e.value.code == SMOQ_EINVAL
    with pytest.raises(ValueError):
        qubit_channel(s, "bit_flip", 0.5, [4], [0.5])
    s.free()
