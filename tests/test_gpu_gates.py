"""Gate parity: the gfx950 kernels against the CPU oracle on identical
seeded states (complex128 to 1e-12 relative norm-wise, complex64 to 1e-5)."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 1e-5


def rand_state(N, seed, batch=None):
    rng = np.random.default_rng(seed)
    shape = (1 << N,) if batch is None else (batch, 1 << N)
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return v.astype(np.complex128)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b)


NAMED = [("rx", 1, "apply_rx_psi", True), ("ry", 1, "apply_ry_psi", True),
         ("rz", 1, "apply_rz_psi", True), ("h", 1, "apply_hadamard_psi", False),
         ("x", 1, "apply_pauli_x_psi", False), ("y", 1, "apply_pauli_y_psi", False),
         ("z", 1, "apply_pauli_z_psi", False), ("s", 1, "apply_s_psi", False),
         ("sdg", 1, "apply_sdagger_psi", False), ("t", 1, "apply_t_psi", False),
         ("tdg", 1, "apply_tdagger_psi", False),
         ("sigma_minus", 1, "apply_sigma_minus_psi", False),
         ("sigma_plus", 1, "apply_sigma_plus_psi", False),
         ("cnot", 2, "apply_cnot_psi", False), ("cz", 2, "apply_cz_psi", False),
         ("cry", 2, "apply_cry_psi", True), ("rxx", 2, "apply_rxx_psi", True),
         ("ryy", 2, "apply_ryy_psi", True), ("rzz", 2, "apply_rzz_psi", True),
         ("ccz", 3, "apply_ccz_psi", False), ("toffoli", 3, "apply_toffoli_psi", False)]


@pytest.mark.parametrize("name,nq,fn,has_theta", NAMED)
def test_named_gate_parity(ctx, name, nq, fn, has_theta):
    import smoq_amd.gates as g
    from smoq_amd import GPUState

    rng = np.random.default_rng(100 + NAMED.index((name, nq, fn, has_theta)))
    for N in (3, 7, 11):
        for trial in range(3):
            q = [int(x) + 1 for x in rng.choice(N, nq, replace=False)]
            theta = float(rng.uniform(0, 2 * math.pi))
            psi = rand_state(N, trial)
            ref = psi.copy()
            orc.apply_named(ref, name, q, theta, N)
            s = GPUState.from_numpy(psi)
            args = [s, *q] + ([theta] if has_theta else []) + [N]
            getattr(g, fn)(*args)
            got = s.to_numpy()
            assert rel(got, ref) < TOL128, (name, N, q)


def test_generic_1q_2q_parity(ctx):
    import smoq_amd.gates as g
    from smoq_amd import GPUState

    rng = np.random.default_rng(7)
    N = 9
    for trial in range(6):
        U1 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        U2 = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        q = int(rng.integers(1, N + 1))
        q1, q2 = [int(x) + 1 for x in rng.choice(N, 2, replace=False)]
        psi = rand_state(N, trial)
        ref = psi.copy()
        orc.apply_1q(ref, U1, q, N)
        orc.apply_2q_trotter(ref, U2, q1, q2, N)
        s = GPUState.from_numpy(psi)
        g.apply_one_qubit_gate_bitwise_cpu(s, U1, q, N)
        g.apply_two_qubit_gate_bitwise_cpu(s, U2, q1, q2, N)
        assert rel(s.to_numpy(), ref) < TOL128


def test_apply_2qubit_gate_permutes_u_when_q1_above_q2(ctx):
    # apply_2qubit_gate! swaps U with SWAP*U*SWAP when q1 > q2: the result
    # equals the ascending call with the permuted matrix.
    import smoq_amd.gates as g
    from smoq_amd import GPUState

    rng = np.random.default_rng(8)
    N = 6
    U = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    SW = np.eye(4)[[0, 2, 1, 3]]
    psi = rand_state(N, 3)
    a = GPUState.from_numpy(psi)
    b = GPUState.from_numpy(psi)
    g.apply_2qubit_gate(a, U, 5, 2, N)
    g.apply_2qubit_gate(b, SW @ U @ SW, 2, 5, N)
    assert rel(a.to_numpy(), b.to_numpy()) < 1e-14


def test_reference_kats_on_device(ctx):
    import smoq_amd.gates as g
    from smoq_amd import GPUState

    R2 = 1 / math.sqrt(2)
    cases = [("apply_ry_psi", math.pi / 2, [R2, R2]), ("apply_ry_psi", math.pi, [0, 1]),
             ("apply_rx_psi", math.pi / 2, [R2, -1j * R2]), ("apply_rx_psi", math.pi, [0, -1j]),
             ("apply_rz_psi", math.pi / 2, [np.exp(-1j * math.pi / 4), 0])]
    for fn, th, want in cases:
        s = GPUState.basis(1, 0)
        getattr(g, fn)(s, 1, th, 1)
        assert np.max(np.abs(s.to_numpy() - np.array(want))) < 1e-10


def random_tape(N, n_ops, seed):
    from smoq_amd.circuits import Tape

    rng = np.random.default_rng(seed)
    t = Tape(N)
    kinds = ["H", "X", "Y", "Z", "S", "SDG", "T", "TDG", "RX", "RY", "RZ", "CNOT", "CZ", "CRY",
             "RXX", "RYY", "RZZ", "CCZ", "TOFFOLI", "U1", "U2"]
    for _ in range(n_ops):
        k = str(rng.choice(kinds))
        nq = {"CNOT": 2, "CZ": 2, "CRY": 2, "RXX": 2, "RYY": 2, "RZZ": 2, "U2": 2,
              "CCZ": 3, "TOFFOLI": 3}.get(k, 1)
        q = [int(x) for x in rng.choice(N, nq, replace=False)]
        m = None
        if k == "U1":
            m = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
        if k == "U2":
            m = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
        t.add(k, q, float(rng.uniform(0, 2 * math.pi)), m=m)
    return t


ORACLE_NAME = {"H": "h", "X": "x", "Y": "y", "Z": "z", "S": "s", "SDG": "sdg", "T": "t",
               "TDG": "tdg", "RX": "rx", "RY": "ry", "RZ": "rz", "CNOT": "cnot", "CZ": "cz",
               "CRY": "cry", "RXX": "rxx", "RYY": "ryy", "RZZ": "rzz", "CCZ": "ccz",
               "TOFFOLI": "toffoli"}


def oracle_run_tape(tape, psi, theta=None):
    N = tape.n
    for gate, qs, th, par, m in tape.ops:
        if par >= 0:
            th = theta[par]
        if gate == "U1":
            orc.apply_1q(psi, m, qs[0] + 1, N)
        elif gate == "U2":
            # basis b(q0) + 2 b(q1): permute into the ascending convention
            U = np.asarray(m)
            if qs[0] > qs[1]:
                SW = np.eye(4)[[0, 2, 1, 3]]
                U = SW @ U @ SW
            orc.apply_2q_trotter(psi, U, min(qs) + 1, max(qs) + 1, N)
        else:
            orc.apply_named(psi, ORACLE_NAME[gate], [q + 1 for q in qs], th, N)
    return psi


@pytest.mark.parametrize("N,tile,low", [(4, 6, 3), (5, 8, 3), (9, 8, 3), (12, 8, 4),
                                        (14, 10, 5), (16, 11, 5), (16, 11, 6), (18, 10, 4)])
def test_fused_tape_parity(ctx, N, tile, low):
    from smoq_amd import GPUState

    ctx.set_fusion(tile, low)
    try:
        tape = random_tape(N, 120, seed=N * 31 + tile)
        psi = rand_state(N, N)
        ref = oracle_run_tape(tape, psi.copy())
        s = GPUState.from_numpy(psi)
        npass = tape.apply(s)
        assert 1 <= npass <= len(tape)
        assert rel(s.to_numpy(), ref) < TOL128
    finally:
        ctx.set_fusion(-1, 4)


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("N,tile", [(9, 6), (13, 10), (14, 12), (16, 11)])
def test_fused_kernels_precompiled_and_run_time(ctx, N, tile, jit):
    """The precompiled pass kernel (jit 0) and the pass kernels compiled for
    the circuit structure (jit 2) against the oracle on the same tape:
    rotations, CZ bricks, controlled and 2-qubit gates, diagonals."""
    from smoq_amd import GPUState

    ctx.set_fusion(tile, 4)
    ctx.set_jit(jit)
    try:
        tape = random_tape(N, 150, seed=7 * N + tile + jit)
        psi = rand_state(N, N + 1)
        ref = oracle_run_tape(tape, psi.copy())
        s = GPUState.from_numpy(psi)
        tape.apply(s)
        assert rel(s.to_numpy(), ref) < TOL128
        s64 = GPUState.from_numpy(psi.astype(np.complex64))
        tape.apply(s64)
        assert rel(s64.to_numpy(), ref) < 1e-5
    finally:
        ctx.set_fusion(-1, 4)
        ctx.set_jit(1)


@pytest.mark.parametrize("N,tile,low", [(10, 6, 3), (14, 8, 4), (15, 10, 5), (17, 11, 5),
                                        (16, 12, 5), (18, 12, 4)])
def test_fused_brick_circuit_parity(ctx, N, tile, low):
    # rotation layers + CZ bricks: exercises merged CZ signs between register
    # slots, between a slot and a lane qubit and between lane qubits, and
    # merged register-only / lane-only diagonal phases
    from smoq_amd import GPUState
    from smoq_amd.circuits import random_rotation_brick

    ctx.set_fusion(tile, low)
    try:
        tape = random_rotation_brick(N, 8, seed=N + tile)
        tape.add("CZ", [0, N - 1]).add("RZZ", [1, N - 2], 0.7).add("CCZ", [0, 2, N - 1])
        tape.add("S", [N - 1]).add("T", [3]).add("RZ", [N - 3], 1.3).add("CZ", [2, 3])
        ref = np.zeros(1 << N, np.complex128)
        ref[0] = 1
        oracle_run_tape(tape, ref)
        s = GPUState.basis(N, 0)
        tape.apply(s)
        assert rel(s.to_numpy(), ref) < TOL128
    finally:
        ctx.set_fusion(-1, 4)


def test_fused_batch_and_c64(ctx):
    from smoq_amd import GPUState

    N = 13
    tape = random_tape(N, 80, seed=5)
    psi = rand_state(N, 2, batch=3)
    refs = [oracle_run_tape(tape, p.copy()) for p in psi]
    s = GPUState.from_numpy(psi)
    tape.apply(s)
    got = s.to_numpy()
    for b in range(3):
        assert rel(got[b], refs[b]) < TOL128
    s64 = GPUState.from_numpy(psi.astype(np.complex64))
    tape.apply(s64)
    got64 = s64.to_numpy()
    for b in range(3):
        assert rel(got64[b], refs[b]) < TOL64


def test_parametrised_tape_uses_theta_array(ctx):
    from smoq_amd import GPUState
    from smoq_amd.circuits import hardware_efficient_ansatz

    N, L = 8, 3
    t = hardware_efficient_ansatz(N, L)
    theta = np.random.default_rng(1).uniform(0, 2 * math.pi, t.n_params)
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1
    oracle_run_tape(t, ref, theta)
    s = GPUState.basis(N, 0)
    t.apply(s, theta)
    assert rel(s.to_numpy(), ref) < TOL128


def test_invalid_arguments_raise(ctx):
    import smoq_amd.gates as g
    from smoq_amd import GPUState, SmoqError
    from smoq_amd.circuits import Tape

    s = GPUState.basis(4, 0)
    with pytest.raises(ValueError):
        g.apply_rx_psi(s, 5, 0.1, 4)
    with pytest.raises(ValueError):
        g.apply_cz_psi(s, 2, 2, 4)
    t = Tape(4).add("RX", [0], param=3)
    with pytest.raises(SmoqError):
        t.apply(s, theta=[0.0])
    with pytest.raises(SmoqError):
        s.set_basis(16)


def test_product_states_and_initial_psi(ctx):
    """make_product_ket / smoq_state_set_product against the oracle's
    restatement for random kets (batch 2, complex128 and complex64), the
    reference's string form, and make_initial_psi."""
    from smoq_amd import GPUState, make_initial_psi, make_product_ket

    rng = np.random.default_rng(5)
    for N in (1, 2, 7, 13):
        kets = [rng.normal(size=2) + 1j * rng.normal(size=2) for _ in range(N)]
        want = orc.product_ket(kets)
        for dtype, tol in (("c128", 1e-14), ("c64", 1e-5)):
            s = GPUState(N, dtype, batch=2)
            s.set_product(kets)
            got = s.to_numpy()
            assert np.max(np.abs(got - want[None, :])) < tol * max(1.0, np.max(np.abs(want)))
            s.free()
    r = 1 / np.sqrt(2)
    s = make_product_ket("0+-1")
    want = orc.product_ket([(1, 0), (r, r), (r, -r), (0, 1)])
    assert np.max(np.abs(s.to_numpy() - want)) < 1e-15
    s.free()
    s = make_product_ket(["one", "zero", "plus"])
    assert np.max(np.abs(s.to_numpy() - orc.product_ket([(0, 1), (1, 0), (r, r)]))) < 1e-15
    s.free()
    s = make_initial_psi(5, [1, 0, 1, 1, 0])
    v = s.to_numpy()
    assert v[0b01101] == 1.0 and np.count_nonzero(v) == 1
    s.free()
    with pytest.raises(ValueError):
        make_product_ket(["zero", "up"])


def structured_tape(N, n_ops, seed):
    """Rotations, CZs and matrices with exact zeros: diagonal 2x2 and 4x4
    (phase gates, lowered without a register slot), number-conserving 4x4
    (an XXZ bond: phases on |00>, |11> and a block on |01>, |10>) and dense
    4x4."""
    from scipy.linalg import expm

    from smoq_amd.circuits import Tape

    rng = np.random.default_rng(seed)
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    t = Tape(N)
    for _ in range(n_ops):
        k = int(rng.integers(6))
        if k == 0:
            t.add("U1", [int(rng.integers(N))],
                  m=np.diag(np.exp(1j * rng.uniform(0, 6.3, 2))))
        elif k == 1:
            q = [int(x) for x in rng.choice(N, 2, replace=False)]
            t.add("U2", q, m=np.diag(np.exp(1j * rng.uniform(0, 6.3, 4))))
        elif k == 2:
            q = [int(x) for x in rng.choice(N, 2, replace=False)]
            H = np.kron(sx, sx) + np.kron(sy, sy) + rng.uniform(-1, 1) * np.kron(sz, sz)
            t.add("U2", q, m=expm(-1j * rng.uniform(0.01, 0.5) * H))
        elif k == 3:
            q = [int(x) for x in rng.choice(N, 2, replace=False)]
            t.add("U2", q, m=np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0])
        elif k == 4:
            t.add(str(rng.choice(["RX", "RY", "RZ"])), [int(rng.integers(N))],
                  float(rng.uniform(0, 2 * math.pi)))
        else:
            q = [int(x) for x in rng.choice(N, 2, replace=False)]
            t.add("CZ", q)
    return t


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("N,tile", [(14, 10), (16, 12)])
def test_structured_matrices_in_fused_passes(ctx, N, tile, jit):
    """Gates with exactly-zero entries (phase gates given as matrices, XXZ
    bonds): lowered and compiled with those zeros as structure, the result
    is the oracle's, in complex128 and complex64."""
    from smoq_amd import GPUState

    ctx.set_fusion(tile, 4)
    ctx.set_jit(jit)
    try:
        tape = structured_tape(N, 160, seed=5 * N + tile + jit)
        psi = rand_state(N, N + 3)
        ref = oracle_run_tape(tape, psi.copy())
        s = GPUState.from_numpy(psi)
        tape.apply(s)
        assert rel(s.to_numpy(), ref) < TOL128
        s64 = GPUState.from_numpy(psi.astype(np.complex64))
        tape.apply(s64)
        assert rel(s64.to_numpy(), ref) < 1e-5
        # a compiled plan (fixed angles: zero entries as structure)
        plan = tape.compile()
        s.upload(psi)
        plan.run(s)
        assert rel(s.to_numpy(), ref) < TOL128
        plan.free()
    finally:
        ctx.set_fusion(-1, 4)
        ctx.set_jit(1)
