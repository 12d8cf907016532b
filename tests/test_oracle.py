"""Pins the CPU oracle (oracle/smoq_oracle.c) before it is trusted as the
checker: reference known-answer tests, dense-matrix brute force, closed
forms and library routines.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

R2 = 1 / math.sqrt(2)


def rand_state(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return (v / np.linalg.norm(v)).astype(np.complex128)


# ---- reference KATs: demo_parametrized_gates.jl:139-153 (recorded pass log
# scripts_variational_quantum_circuits/demo_parametrized_gates/test_summary.txt)
@pytest.mark.parametrize("gate,theta,expected", [
    ("ry", math.pi / 2, [R2, R2]),
    ("ry", math.pi, [0, 1]),
    ("rx", math.pi / 2, [R2, -1j * R2]),
    ("rx", math.pi, [0, -1j]),
    ("rz", math.pi / 2, [np.exp(-1j * math.pi / 4), 0]),
])
def test_reference_rotation_kats(gate, theta, expected):
    psi = np.array([1, 0], np.complex128)
    orc.apply_named(psi, gate, [1], theta, 1)
    assert np.max(np.abs(psi - np.array(expected, np.complex128))) < 1e-10


def test_hadamard_all_uniform():
    # demo_parametrized_gates.jl:221-249: H^{⊗N}|0> is uniform
    N = 6
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1
    for k in range(1, N + 1):
        orc.apply_named(psi, "h", [k], 0.0, N)
    assert np.allclose(psi, np.full(1 << N, 2 ** (-N / 2)), atol=1e-14)


# ---- dense brute force: every named gate equals its kron-built matrix
I2 = np.eye(2)
X = np.array([[0, 1], [1, 0]], np.complex128)
Y = np.array([[0, -1j], [1j, 0]], np.complex128)
Z = np.diag([1, -1]).astype(np.complex128)


def op_on(mats, N):
    """kron of per-qubit 2x2 matrices; mats: dict 0-based qubit -> matrix."""
    M = np.array([[1.0 + 0j]])
    for b in range(N - 1, -1, -1):
        M = np.kron(M, mats.get(b, I2))
    return M


def proj(b, v):
    return np.diag([1 - v, v]).astype(np.complex128) if b is not None else I2


def dense_gate(name, q, theta, N):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    one = {
        "rx": np.array([[c, -1j * s], [-1j * s, c]]),
        "ry": np.array([[c, -s], [s, c]]),
        "rz": np.diag([np.exp(-1j * theta / 2), np.exp(1j * theta / 2)]),
        "h": np.array([[R2, R2], [R2, -R2]]),
        "x": X, "y": Y, "z": Z,
        "s": np.diag([1, 1j]), "sdg": np.diag([1, -1j]),
        "t": np.diag([1, np.exp(1j * math.pi / 4)]), "tdg": np.diag([1, np.exp(-1j * math.pi / 4)]),
        "sigma_minus": np.array([[0, 1], [0, 0]]), "sigma_plus": np.array([[0, 0], [1, 0]]),
    }
    a = [x - 1 for x in q]
    if name in one:
        return op_on({a[0]: np.asarray(one[name], np.complex128)}, N)
    P0, P1 = np.diag([1, 0]).astype(complex), np.diag([0, 1]).astype(complex)
    D = 1 << N
    if name == "cnot":
        return op_on({a[0]: P0}, N) + op_on({a[0]: P1, a[1]: X}, N)
    if name == "cry":
        return op_on({a[0]: P0}, N) + op_on({a[0]: P1, a[1]: np.array([[c, -s], [s, c]])}, N)
    if name == "cz":
        return np.eye(D) - 2 * op_on({a[0]: P1, a[1]: P1}, N)
    if name == "ccz":
        return np.eye(D) - 2 * op_on({a[0]: P1, a[1]: P1, a[2]: P1}, N)
    if name == "toffoli":
        return np.eye(D) - op_on({a[0]: P1, a[1]: P1}, N) + op_on({a[0]: P1, a[1]: P1, a[2]: X}, N)
    if name in ("rxx", "ryy", "rzz"):
        P = {"rxx": X, "ryy": Y, "rzz": Z}[name]
        PP = op_on({a[0]: P, a[1]: P}, N)
        return math.cos(theta / 2) * np.eye(D) - 1j * math.sin(theta / 2) * PP
    raise ValueError(name)


GATES = [("rx", 1), ("ry", 1), ("rz", 1), ("h", 1), ("x", 1), ("y", 1), ("z", 1), ("s", 1),
         ("sdg", 1), ("t", 1), ("tdg", 1), ("sigma_minus", 1), ("sigma_plus", 1),
         ("cnot", 2), ("cz", 2), ("cry", 2), ("rxx", 2), ("ryy", 2), ("rzz", 2),
         ("ccz", 3), ("toffoli", 3)]


@pytest.mark.parametrize("name,nq", GATES)
def test_oracle_gate_matches_dense(name, nq):
    N = 5
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    for trial in range(4):
        q = [int(x) + 1 for x in rng.choice(N, nq, replace=False)]
        theta = float(rng.uniform(0, 2 * math.pi))
        psi = rand_state(N, trial)
        ref = dense_gate(name, q, theta, N) @ psi
        orc.apply_named(psi, name, q, theta, N)
        assert np.max(np.abs(psi - ref)) < 1e-13, (name, q)


def test_oracle_two_qubit_trotter_quirk():
    # apply_two_qubit_gate_bitwise_cpu! sorts (q1, q2) without permuting U:
    # U's first basis bit is the lower qubit whatever the argument order.
    N = 4
    rng = np.random.default_rng(3)
    U = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    a = rand_state(N, 1)
    b = a.copy()
    orc.apply_2q_trotter(a, U, 3, 1, N)
    orc.apply_2q_trotter(b, U, 1, 3, N)
    assert np.array_equal(a, b)
    # dense: basis e = b(q_lo) + 2 b(q_hi) with q_lo = 1, q_hi = 3
    D = 1 << N
    M = np.zeros((D, D), complex)
    for i in range(D):
        for j in range(D):
            if (i & ~0b101) == (j & ~0b101):
                ei = (i & 1) | (((i >> 2) & 1) << 1)
                ej = (j & 1) | (((j >> 2) & 1) << 1)
                M[i, j] = U[ei, ej]
    c = rand_state(N, 1)
    assert np.max(np.abs(a - M @ c)) < 1e-13


# ---- observables: dense brute force and closed forms
def test_oracle_observables_dense():
    N = 5
    psi = rand_state(N, 7)
    for k in range(1, N + 1):
        for p, P in (("x", X), ("y", Y), ("z", Z)):
            ref = np.vdot(psi, op_on({k - 1: P}, N) @ psi).real
            assert abs(orc.expect_local(psi, k, N, p) - ref) < 1e-13
    for i in range(1, N):
        for pair, P in (("zz", Z), ("xx", X), ("yy", Y)):
            ref = np.vdot(psi, op_on({i - 1: P, i: P}, N) @ psi).real
            assert abs(orc.expect_corr(psi, i, i + 1, N, pair) - ref) < 1e-13


def test_bell_and_ghz_closed_forms():
    bell = np.array([R2, 0, 0, R2], np.complex128)
    # cpuQuantumStateObservables.jl:297-304: Bell <ZZ> = <XX> = 1, <YY> = -1
    assert abs(orc.expect_corr(bell, 1, 2, 2, "zz") - 1) < 1e-15
    assert abs(orc.expect_corr(bell, 1, 2, 2, "xx") - 1) < 1e-15
    assert abs(orc.expect_corr(bell, 1, 2, 2, "yy") + 1) < 1e-15
    ghz = np.zeros(8, np.complex128)
    ghz[0] = ghz[7] = R2
    # <Z1Z2> = 1 and <Z1Z2Z3> = 0 on GHZ (the docstring's 1.0 for ZZZ,
    # cpuQuantumStateObservables.jl:829-831, is not what the loop computes)
    assert abs(orc.pauli_string_ref(ghz, "ZZ", [1, 2], 3) - 1) < 1e-15
    assert abs(orc.pauli_string_ref(ghz, "ZZZ", [1, 2, 3], 3)) < 1e-15
    # the reference returns half of the true <XXX> = 1 (its :1044)
    assert abs(orc.pauli_string_ref(ghz, "XXX", [1, 2, 3], 3) - 0.5) < 1e-15


def test_pauli_string_reference_value():
    # The reference keeps only bra states with every X/Y bit 0 and the ket
    # with every X/Y bit 1 (the single all-flipped pattern with non-zero
    # coefficient), taking Re(coeff * zsign * conj(ψ_bra) ψ_ket).  With one
    # X/Y operator that is exactly <P>/2; with more it drops further pairs.
    N = 5
    psi = rand_state(N, 11)
    rng = np.random.default_rng(5)
    for _ in range(40):
        k = int(rng.integers(1, 4))
        qs = [int(x) + 1 for x in rng.choice(N, k, replace=False)]
        ops = [str(rng.choice(list("XYZ"))) for _ in range(k)]
        x = z = 0
        ny = 0
        for o, q in zip(ops, qs):
            if o in "XY":
                x |= 1 << (q - 1)
            if o in "YZ":
                z |= 1 << (q - 1)
            ny += o == "Y"
        true = np.vdot(psi, orc.pauli_matrix(x, z, N) @ psi).real
        got = orc.pauli_string_ref(psi, ops, qs, N)
        if x == 0:
            assert abs(got - true) < 1e-13
            continue
        zo = z & ~x
        idx = np.arange(1 << N)
        base = idx[(idx & x) == 0]
        sgn = np.array([(-1) ** bin(int(b) & zo).count("1") for b in base])
        want = np.real((-1j) ** ny * np.sum(sgn * np.conj(psi[base]) * psi[base | x]))
        assert abs(got - want) < 1e-13
        if bin(x).count("1") == 1:
            assert abs(got - true / 2) < 1e-13


def test_compute_energy_dense():
    N = 5
    psi = rand_state(N, 3)
    Jx, Jy, Jz, hx = 0.7, -0.3, 1.1, 0.9
    H = np.zeros((1 << N, 1 << N), complex)
    for i in range(N - 1):
        H += Jx * op_on({i: X, i + 1: X}, N) + Jy * op_on({i: Y, i + 1: Y}, N) \
            + Jz * op_on({i: Z, i + 1: Z}, N)
    for i in range(N):
        H += hx * op_on({i: X}, N)
    assert abs(orc.compute_energy(psi, N, Jx, Jy, Jz, hx) - np.vdot(psi, H @ psi).real) < 1e-12


# ---- Julia cumsum (accumulate_pairwise!) and searchsortedfirst
def test_pairwise_cumsum_leaf_rule_known_answer():
    # c[1] = v1; inside the root leaf c[i] = v1 + (v2 + ... + vi).
    u = 2.0 ** -53
    v = np.array([1.0, u, u], np.float64)
    c = orc.cumsum_pairwise(v)
    # 1 + u rounds to 1 (tie to even); 1 + (u + u) = 1 + 2^-52 exactly
    assert c.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -52]
    # a sequential running sum never leaves 1.0
    assert orc.cumsum_sequential(v).tolist() == [1.0, 1.0, 1.0]


def test_pairwise_cumsum_split_known_answer():
    # n = 257: root covers v[2..257] (256 elements) -> halves of 128 -> leaves
    # of 64.  With v1 = 1 and every other v = 2^-53, a leaf's local sum is
    # exact (k 2^-53) and c = carry + local rounds to a multiple of 2^-52.
    # Leaf carries: 1, 1 + 64u, 1 + 128u, 1 + 192u (exact, even multiples of u).
    u = 2.0 ** -53
    v = np.full(257, u)
    v[0] = 1.0
    c = orc.cumsum_pairwise(v)

    def rnd(k):  # 1 + k*u rounded to nearest multiple of 2u, ties to even
        q, r = divmod(k, 2)
        if r == 0:
            return 1.0 + k * u
        return 1.0 + (q + (q % 2)) * 2 * u

    want = [1.0]
    for leaf in range(4):
        carry = 64 * leaf
        for j in range(1, 65):
            want.append(rnd(carry + j))
    assert c.tolist() == want
    # the last entry equals the exact total (1 + 256u) only because the
    # pairwise tree keeps every partial sum exact
    assert c[-1] == 1.0 + 256 * u


def test_pairwise_cumsum_exact_on_integers():
    rng = np.random.default_rng(0)
    for n in (1, 2, 127, 128, 129, 1000, 4097):
        v = rng.integers(0, 1000, size=n).astype(np.float64)
        assert np.array_equal(orc.cumsum_pairwise(v), np.cumsum(v))


def test_searchsortedfirst_matches_numpy():
    rng = np.random.default_rng(1)
    v = np.sort(rng.random(1000))
    v[100:110] = v[100]  # duplicates
    for x in list(rng.random(200)) + [0.0, v[100], v[-1], 1.5, -1.0]:
        assert orc.searchsortedfirst(v, x) == np.searchsorted(v, x, side="left") + 1


def test_samplers_against_numpy():
    N = 8
    psi = rand_state(N, 2)
    r = np.random.default_rng(9).random(500)
    r[:3] = [0.0, 1.0 - 2 ** -53, 0.5]
    seq = orc.sample_sequential(psi, N, r)
    c = np.cumsum(np.abs(psi) ** 2)
    for d, x in enumerate(r):
        hits = np.nonzero(x <= orc.cumsum_sequential(orc.probs(psi)))[0]
        assert seq[d] == (hits[0] if hits.size else (1 << N) - 1)
    cs = orc.sample_cumsum_search(psi, N, r)
    cp = orc.cumsum_pairwise(orc.probs(psi))
    assert np.array_equal(cs, np.searchsorted(cp, r, side="left"))
    # both agree with the probability mass to within rounding
    assert np.max(np.abs(cp - c)) < 1e-14


# ---- H psi pull loop vs dense Hamiltonian
def test_hpsi_dense():
    N = 5
    rng = np.random.default_rng(4)
    bonds = [(i, i + 1, 0.8, 0.6, 0.4) for i in range(1, N)]
    fields = rng.uniform(-1, 1, size=(N, 3))
    H = np.zeros((1 << N, 1 << N), complex)
    for (i, j, a, b, c) in bonds:
        H += a * op_on({i - 1: X, j - 1: X}, N) + b * op_on({i - 1: Y, j - 1: Y}, N) \
            + c * op_on({i - 1: Z, j - 1: Z}, N)
    # The pull loop's field-Y term uses +i for a 0 bit (…Chebyshev.jl:139-142),
    # i.e. it applies -hy*Y; the bond YY term is the true Y⊗Y.
    for f in range(N):
        H += fields[f, 0] * op_on({f: X}, N) - fields[f, 1] * op_on({f: Y}, N) \
            + fields[f, 2] * op_on({f: Z}, N)
    psi = rand_state(N, 8)
    assert np.max(np.abs(orc.hpsi(psi, N, bonds, fields) - H @ psi)) < 1e-13


def test_product_ket_matches_kron_chain():
    """make_product_ket restated: little-endian, ket 1 on the lowest bit;
    against numpy's Kronecker product taken in the opposite order."""
    rng = np.random.default_rng(11)
    kets = [rng.normal(size=2) + 1j * rng.normal(size=2) for _ in range(7)]
    want = np.array([1.0 + 0j])
    for k in kets:
        want = np.kron(k, want)
    got = orc.product_ket(kets)
    assert np.max(np.abs(got - want)) < 1e-14
    r = 1 / np.sqrt(2)
    # |0>|+>|1> (qubit 1 = |0>): amplitudes at indices with bit0 = 0, bit2 = 1
    v = orc.product_ket([(1, 0), (r, r), (0, 1)])
    assert np.allclose(v, [0, 0, 0, 0, r, 0, r, 0])


def test_sequential_rule_as_running_sum_and_search():
    """The oracle's per-draw sequential loop (sample_from_probabilities) is
    the first index of its running sum at or above the draw, the last index
    when none is: the form the full-size check uses."""
    rng = np.random.default_rng(5)
    N = 10
    psi = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    psi /= np.linalg.norm(psi)
    c = orc.cumsum_sequential(orc.probs(psi))
    r = np.concatenate([rng.random(300), c[::37], np.nextafter(c[::41], 2.0), [0.0, 2.0]])
    want = orc.sample_sequential(psi, N, r)
    got = np.minimum(np.searchsorted(c, r, side="left"), c.size - 1)
    assert np.array_equal(got, want)
