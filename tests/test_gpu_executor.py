"""smoq_apply_circuit (apply_circuit!) on the device against the oracle's
executor restatement on the same uniforms: outcomes and uniform counts
equal, amplitudes within 1e-12 (complex128) / 2e-5 (complex64); density
matrices within 1e-12; teleportation's invariant; the error paths."""
import numpy as np
import pytest

from oracle import oracle as orc
from smoq_amd import builder as B
from smoq_amd._lib import SmoqError

pytestmark = pytest.mark.gpu

GATES1 = ["rx", "ry", "rz", "h", "x", "y", "z", "s", "t"]
GATES2 = ["cz", "cnot", "cx", "rxx", "ryy", "rzz"]
NOISE = ["depolarizing", "amplitude_damping", "dephasing", "bit_flip", "phase_damping",
         "thermal_relaxation"]


def random_circuit(N, seed, n_ops=60, conditions=True, dm=False):
    """Every operation kind in random order; if_outcome on some gates,
    resets and noise (also pointing past the outcomes recorded so far)."""
    rng = np.random.default_rng(seed)
    ops, p, n_meas = [], 1, 0
    for _ in range(n_ops):
        r = rng.random()
        cond = int(rng.integers(1, n_meas + 3)) if conditions and rng.random() < 0.25 else 0
        qs = [int(q) for q in rng.permutation(N)[:2] + 1]
        if r < 0.45:
            g = str(rng.choice(GATES1))
            if g in ("rx", "ry", "rz") and rng.random() < 0.6:
                ops.append(B.gate(g, [qs[0]], param_idx=p, if_outcome=cond))
                p += 1
            else:
                ops.append(B.gate(g, [qs[0]], angle=float(rng.normal()), if_outcome=cond))
        elif r < 0.7:
            g = str(rng.choice(GATES2))
            if g.startswith("r") and rng.random() < 0.5:
                ops.append(B.gate(g, qs, param_idx=p, if_outcome=cond))
                p += 1
            else:
                ops.append(B.gate(g, qs, angle=float(rng.normal()), if_outcome=cond))
        elif r < 0.82:
            k = int(rng.integers(1, 3))
            ops.append(B.noise(str(rng.choice(NOISE)), qs[:k], p=float(rng.uniform(0.05, 0.6)),
                               if_outcome=cond))
        elif r < 0.92:
            k = int(rng.integers(1, 3))
            ops.append(B.measure(qs[:k], basis=str(rng.choice(["z", "x", "y"])) if not dm
                                 else "z"))
            n_meas += k
        else:
            ops.append(B.reset([qs[0]], target_state=str(rng.choice(["zero", "one", "plus",
                                                                         "minus"])),
                               if_outcome=cond))
    c = B.ParameterizedCircuit.from_operations(N, ops)
    return c, rng.normal(size=max(c.n_params, 1))


def zero_state(N):
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1
    return psi


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_circuit_matches_oracle(ctx, seed, dtype):
    from smoq_amd import GPUState
    from smoq_amd.executor import apply_circuit

    N = 6
    print(f"[gpu] apply_circuit N={N} seed={seed} {dtype}", flush=True)
    c, th = random_circuit(N, seed)
    u = np.random.default_rng(100 + seed).random(400)
    psi = GPUState.basis(N, 0, dtype=dtype)
    res = apply_circuit(psi, c, th, uniforms=u)
    ref = zero_state(N)
    out, used = orc.circuit_psi(ref, c.operations, th, N, u)
    assert res.outcomes == out
    assert res.uniforms_used == used
    tol = 1e-12 if dtype == "c128" else 2e-5
    assert np.abs(psi.to_numpy().astype(np.complex128) - ref).max() < tol
    psi.free()


def test_trajectory_batch_matches_oracle_per_item(ctx):
    from smoq_amd import GPUState
    from smoq_amd.executor import apply_circuit

    N, Bn = 5, 3
    c, th = random_circuit(N, 11, conditions=False)
    per_op = [len(op.qubits) for op in c.operations
              if op.type not in B.GATE_TYPES and op.type not in ("thermal_relaxation",)]
    total = sum(per_op)
    rng = np.random.default_rng(7)
    per_item = rng.random((Bn, total))
    # device layout: per op, batch values for each listed qubit, qubit-major
    flat, off = [], 0
    for k in per_op:
        flat.append(per_item[:, off:off + k].T.reshape(-1))
        off += k
    flat = np.concatenate(flat)
    psi = GPUState(N, "c128", batch=Bn)
    psi.set_basis(0)
    res = apply_circuit(psi, c, th, uniforms=flat)
    got = psi.to_numpy()
    assert res.uniforms_used == flat.size
    for b in range(Bn):
        ref = zero_state(N)
        out, _ = orc.circuit_psi(ref, c.operations, th, N, per_item[b])
        assert list(res.outcomes[b]) == out
        assert np.abs(got[b] - ref).max() < 1e-12
    psi.free()


@pytest.mark.parametrize("seed", [21, 22, 23])
def test_density_matrix_circuit_matches_oracle(ctx, seed):
    from smoq_amd.density import DensityMatrix
    from smoq_amd.executor import apply_circuit

    N = 4
    print(f"[gpu] apply_circuit rho N={N} seed={seed}", flush=True)
    c, th = random_circuit(N, seed, n_ops=40, dm=True)
    rho = DensityMatrix.basis(N, 0)
    res = apply_circuit(rho, c, th)
    assert res.outcomes == []
    ref = np.zeros((1 << N, 1 << N), np.complex128)
    ref[0, 0] = 1
    orc.circuit_rho(ref, c.operations, th, N)
    assert np.abs(rho.to_numpy() - ref).max() < 1e-12
    rho.free()


def test_noisy_ansatz_density_matrix_and_energy(ctx):
    """The demo's noisy hardware-efficient ansatz (demo_vqe_dm_vs_mcwf.jl:353-361)
    on ρ, compared with the oracle element by element."""
    from smoq_amd.density import DensityMatrix
    from smoq_amd.executor import apply_circuit

    N = 4
    c = B.hardware_efficient_ansatz(N, 2, entangler="cz", topology="chain",
                                    rotations=("ry", "rz"), noise_type="depolarizing",
                                    noise_p=0.05)
    th = B.initialize_parameters(c, init="small_random", rng=np.random.default_rng(3))
    rho = DensityMatrix.basis(N, 0)
    apply_circuit(rho, c, th)
    ref = np.zeros((1 << N, 1 << N), np.complex128)
    ref[0, 0] = 1
    orc.circuit_rho(ref, c.operations, th, N)
    assert np.abs(rho.to_numpy() - ref).max() < 1e-12
    rho.free()


def test_teleportation_on_device(ctx):
    from smoq_amd import GPUState
    from smoq_amd.executor import apply_circuit

    alpha, beta = 0.8, 2.1
    ops = [B.gate("ry", [1], angle=alpha), B.gate("rz", [1], angle=beta),
           B.gate("h", [2]), B.gate("cnot", [2, 3]), B.gate("cnot", [1, 2]), B.gate("h", [1]),
           B.measure([1]), B.measure([2]),
           B.gate("x", [3], if_outcome=2), B.gate("z", [3], if_outcome=1)]
    c = B.ParameterizedCircuit.from_operations(3, ops)
    phi = np.array([np.cos(alpha / 2) * np.exp(-0.5j * beta), np.sin(alpha / 2) *
                    np.exp(0.5j * beta)])
    for u in ([0.2, 0.3], [0.2, 0.7], [0.7, 0.3], [0.9, 0.9]):
        psi = GPUState.basis(3, 0)
        res = apply_circuit(psi, c, None, uniforms=u)
        m1, m2 = res.outcomes
        assert (m1, m2) == (int(u[0] > 0.5), int(u[1] > 0.5))
        v = psi.to_numpy()
        assert abs(abs(np.vdot(phi, v[[m1 + 2 * m2, m1 + 2 * m2 + 4]])) - 1) < 1e-12
        psi.free()


def test_error_paths(ctx):
    from smoq_amd import GPUState
    from smoq_amd.density import DensityMatrix
    from smoq_amd.executor import apply_circuit

    c = B.ParameterizedCircuit.from_operations(2, [B.trace_out([1])])
    rho = DensityMatrix.basis(2, 0)
    with pytest.raises(SmoqError):
        apply_circuit(rho, c)
    rho.free()
    cond = B.ParameterizedCircuit.from_operations(2, [B.measure([1]),
                                                     B.gate("x", [2], if_outcome=1)])
    psi = GPUState(2, "c128", batch=2)
    psi.set_basis(0)
    with pytest.raises(SmoqError):
        apply_circuit(psi, cond, uniforms=[0.5, 0.5])
    psi.free()
    psi = GPUState.basis(2, 0)
    m = B.ParameterizedCircuit.from_operations(2, [B.measure([1, 2])])
    with pytest.raises(SmoqError):
        apply_circuit(psi, m, uniforms=[0.5])
    with pytest.raises(ValueError):
        apply_circuit(psi, B.hardware_efficient_ansatz(2, 1), [0.1])
    with pytest.warns(UserWarning):
        apply_circuit(psi, B.ParameterizedCircuit.from_operations(2, [B.trace_out([2])]))
    psi.free()
