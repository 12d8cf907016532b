"""The circuit IR (smoq_amd.builder) against the reference's builder layout,
the oracle's executor restatement pinned by invariants (teleportation with
if_outcome corrections, ρ of a noiseless circuit = |ψ><ψ|, ρ of a noisy
circuit = the exact probability-weighted sum over the trajectory branches,
closed forms of the single-qubit maps), and the lowering into
smoq_circuit_op records.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from smoq_amd import builder as B
from smoq_amd._lib import CK
from smoq_amd.executor import lower_circuit


def test_hardware_efficient_ansatz_layout():
    # cpuVariationalQuantumCircuitBuilder.jl:413-437: per layer the rotations
    # (each kind over qubits 1..N, one parameter each), the noise layer, the
    # entanglers; then a final rotation layer
    c = B.hardware_efficient_ansatz(3, 2, rotations=("ry", "rz"), noise_type="depolarizing",
                                    noise_p=0.1)
    exp = []
    p = 1
    for _ in range(2):
        for g in ("ry", "rz"):
            for q in (1, 2, 3):
                exp.append((g, (q,), p))
                p += 1
        exp += [("depolarizing", (q,), 0) for q in (1, 2, 3)]
        exp += [("cz", (1, 2), 0), ("cz", (2, 3), 0)]
    for g in ("ry", "rz"):
        for q in (1, 2, 3):
            exp.append((g, (q,), p))
            p += 1
    got = [(op.type, op.qubits, op.param_idx) for op in c.operations]
    assert got == exp
    assert c.n_params == 18 and B.get_parameter_count(c) == 18
    assert all(op.p == 0.1 for op in c.operations if op.type == "depolarizing")
    assert len(c.layers) == 3 * 2 + 1
    info = B.get_parameter_info(c)
    assert [d["param_idx"] for d in info] == list(range(1, 19))
    assert info[0]["op_index"] == 1 and info[6]["op_index"] == 12


def test_entangling_topologies_and_strong_ansatz():
    pairs = lambda layer: [op.qubits for op in layer.operations]  # noqa: E731
    assert pairs(B.entangling_layer(4, topology="chain")) == [(1, 2), (2, 3), (3, 4)]
    assert pairs(B.entangling_layer(3, topology="ring")) == [(1, 2), (2, 3), (3, 1)]
    assert pairs(B.entangling_layer(2, topology="ring")) == [(1, 2)]
    assert pairs(B.entangling_layer(3, topology="all")) == [(1, 2), (1, 3), (2, 3)]
    with pytest.raises(ValueError):
        B.entangling_layer(3, topology="star")
    c = B.strong_entangling_ansatz(3, 1)
    kinds = [op.type for op in c.operations]
    assert kinds[:9] == ["rx"] * 3 + ["ry"] * 3 + ["rz"] * 3
    assert kinds[9:12] == ["cz"] * 3 and c.n_params == 18
    # a single qubit: no entangling layer
    assert [op.type for op in B.hardware_efficient_ansatz(1, 2).operations] == ["ry"] * 3


def test_repeat_operations_and_constructors():
    layer = [B.gate("ry", [1], param_idx=1), B.gate("ry", [2], param_idx=2), B.gate("cz", [1, 2]),
             B.noise("depolarizing", [1, 2], p=0.01)]
    ops = B.repeat_operations(layer, 3)
    assert [op.param_idx for op in ops] == [1, 2, 0, 0, 3, 4, 0, 0, 5, 6, 0, 0]
    assert [op.param_idx for op in B.repeat_operations(layer, 2, renumber_params=False)] == \
        [1, 2, 0, 0] * 2
    m = B.measure([1, 2], basis=":x")
    assert m.type == "measure_x" and m.basis == "x"
    r = B.reset(2, target_state=":plus", if_outcome=1)
    assert r.type == "reset" and r.target_state == "plus" and r.if_outcome == 1
    with pytest.raises(ValueError):
        B.gate("u3", [1])
    with pytest.raises(ValueError):
        B.noise("leakage", [1])
    c = B.ParameterizedCircuit.from_operations(2, ops)
    assert c.n_params == 6
    assert B.count_operations(c) == {"ry": 6, "cz": 3, "depolarizing": 3}
    assert "2 qubits, 12 operations, 6 parameters" in B.describe_circuit(c)
    th = B.initialize_parameters(c, init="random", rng=np.random.default_rng(1))
    assert th.shape == (6,) and np.all(np.abs(th) <= np.pi)
    assert np.all(B.initialize_parameters(c, init=0.5) == 0.5)
    assert B.extract_gate_angles(c, np.arange(1.0, 7.0))[5] == 3.0


def _teleport(alpha, beta):
    ops = [B.gate("ry", [1], angle=alpha), B.gate("rz", [1], angle=beta),
           B.gate("h", [2]), B.gate("cnot", [2, 3]),
           B.gate("cnot", [1, 2]), B.gate("h", [1]),
           B.measure([1]), B.measure([2]),
           B.gate("x", [3], if_outcome=2), B.gate("z", [3], if_outcome=1)]
    return B.ParameterizedCircuit.from_operations(3, ops)


def test_oracle_teleportation_invariant():
    """if_outcome semantics (executor :103-108): after the conditioned X and Z
    the third qubit holds the first qubit's input state, for all four
    measurement records."""
    alpha, beta = 1.1, -0.7
    phi = orc.apply_named(orc.apply_named(np.array([1, 0], np.complex128), "ry", [1], alpha, 1),
                          "rz", [1], beta, 1)
    c = _teleport(alpha, beta)
    for u1, u2 in itertools.product((0.25, 0.75), repeat=2):
        psi = np.zeros(8, np.complex128)
        psi[0] = 1
        out, used = orc.circuit_psi(psi, c.operations, [], 3, [u1, u2])
        assert out == [int(u1 > 0.5), int(u2 > 0.5)] and used == 2
        m1, m2 = out
        third = psi[[m1 + 2 * m2, m1 + 2 * m2 + 4]]
        assert abs(abs(np.vdot(phi, third)) - 1) < 1e-12
        assert np.linalg.norm(psi) == pytest.approx(1.0, abs=1e-12)
    # without the corrections the outcome-(1, 1) state differs from phi
    psi = np.zeros(8, np.complex128)
    psi[0] = 1
    orc.circuit_psi(psi, c.operations[:-2], [], 3, [0.75, 0.75])
    assert abs(abs(np.vdot(phi, psi[[3, 7]])) - 1) > 1e-3


def _noiseless(N, seed):
    rng = np.random.default_rng(seed)
    ops, p = [], 1
    for _ in range(3):
        for q in range(1, N + 1):
            g = rng.choice(["rx", "ry", "rz", "h", "x", "z"])
            if g in ("rx", "ry", "rz"):
                ops.append(B.gate(g, [q], param_idx=p))
                p += 1
            else:
                ops.append(B.gate(g, [q]))
        for q in range(1, N):
            g = rng.choice(["cz", "cnot", "rxx", "ryy", "rzz"])
            a, b = (q, q + 1) if rng.random() < 0.5 else (q + 1, q)
            ops.append(B.gate(g, [a, b], angle=float(rng.normal())))
    return B.ParameterizedCircuit.from_operations(N, ops), rng.normal(size=p - 1)


def test_oracle_rho_of_noiseless_circuit_is_projector():
    N = 3
    c, th = _noiseless(N, 5)
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1
    orc.circuit_psi(psi, c.operations, th, N, [])
    rho = np.zeros((1 << N, 1 << N), np.complex128)
    rho[0, 0] = 1
    orc.circuit_rho(rho, c.operations, th, N)
    assert np.abs(rho - np.outer(psi, psi.conj())).max() < 1e-13


def test_oracle_rho_noise_is_branch_average():
    """Depolarizing and bit flip on ρ (executor :346-375) equal the exact
    probability-weighted sum over the pure-state branches of
    _apply_depolarizing_single! / _apply_bit_flip_single!, whose branch
    probabilities do not depend on the state."""
    N, pd, pb = 2, 0.2, 0.3
    ops = [B.gate("ry", [1], angle=0.9), B.gate("cnot", [1, 2]),
           B.noise("depolarizing", [1, 2], p=pd), B.gate("rx", [2], angle=-0.4),
           B.noise("bit_flip", [1], p=pb), B.gate("rzz", [1, 2], angle=0.3)]
    dep = [(pd / 6, pd / 3), (pd / 2, pd / 3), (5 * pd / 6, pd / 3), ((1 + pd) / 2, 1 - pd)]
    bit = [(pb / 2, pb), ((1 + pb) / 2, 1 - pb)]
    avg = np.zeros((4, 4), np.complex128)
    for (u1, w1), (u2, w2), (u3, w3) in itertools.product(dep, dep, bit):
        psi = np.zeros(4, np.complex128)
        psi[0] = 1
        orc.circuit_psi(psi, ops, [], N, [u1, u2, u3])
        avg += w1 * w2 * w3 * np.outer(psi, psi.conj())
    rho = np.zeros((4, 4), np.complex128)
    rho[0, 0] = 1
    orc.circuit_rho(rho, ops, [], N)
    assert np.abs(rho - avg).max() < 1e-13


def test_oracle_rho_single_qubit_closed_forms():
    plus = np.full((2, 2), 0.5, np.complex128)
    one = np.diag([0.0, 1.0]).astype(np.complex128)
    p = 0.3
    r = orc.circuit_rho(plus.copy(), [B.noise("dephasing", [1], p=p)], [], 1)
    assert np.allclose(r, [[0.5, 0.5 * (1 - p)], [0.5 * (1 - p), 0.5]], atol=1e-15)
    r = orc.circuit_rho(one.copy(), [B.noise("amplitude_damping", [1], p=p)], [], 1)
    assert np.allclose(r, np.diag([p, 1 - p]), atol=1e-15)
    r = orc.circuit_rho(plus.copy(), [B.noise("amplitude_damping", [1], p=p)], [], 1)
    c = 0.5 * (1 - p / 2)
    assert np.allclose(r, [[0.5 + 0.5 * p, c], [c, 0.5 * (1 - p)]], atol=1e-15)
    for op in (B.measure([1]), B.reset([1], target_state="one"), B.noise("phase_damping", [1],
                                                                           p=p)):
        r = orc.circuit_rho(plus.copy(), [op], [], 1)
        exp = np.diag([0.5, 0.5]) if op.type != "phase_damping" else plus
        assert np.allclose(r, exp, atol=1e-15)
    # y, s, t have no ρ branch in the reference's executor
    r = orc.circuit_rho(plus.copy(), [B.gate("y", [1]), B.gate("s", [1]), B.gate("t", [1])], [], 1)
    assert np.array_equal(r, plus)
    with pytest.raises(ValueError):
        orc.circuit_rho(plus.copy(), [B.trace_out([1])], [], 1)


def test_lowering_records():
    ops = [B.gate("rx", [2], param_idx=3), B.gate("cx", [1, 3]), B.noise("phase_damping", [1, 3],
                                                                        p=0.25),
           B.measure([3, 1], basis="y"), B.reset([2], target_state="minus", if_outcome=2),
           B.gate("rz", [1], angle=0.5), B.trace_out([2])]
    c = B.ParameterizedCircuit.from_operations(3, ops)
    arr, q, nq, n_meas = lower_circuit(c)
    assert nq == 10 and n_meas == 2
    assert list(q[:nq]) == [1, 0, 2, 0, 2, 2, 0, 1, 0, 1]
    rec = [(arr[i].kind, arr[i].param, arr[i].q_first, arr[i].q_count, arr[i].if_outcome,
            arr[i].arg) for i in range(len(ops))]
    assert rec == [(CK["rx"], 2, 0, 1, 0, 0), (CK["cnot"], -1, 1, 2, 0, 0),
                   (CK["phase_damping"], -1, 3, 2, 0, 0), (CK["measure"], -1, 5, 2, 0, 2),
                   (CK["reset"], -1, 7, 1, 2, 3), (CK["rz"], -1, 8, 1, 0, 0),
                   (CK["trace_out"], -1, 9, 1, 0, 0)]
    assert arr[2].p == 0.25 and arr[5].angle == 0.5
