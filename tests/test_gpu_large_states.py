"""States whose launches outgrow one grid row, checked through properties
that do not depend on size: the fills (product, basis) of 2^32 complex64
amplitudes, whose one-thread-per-amplitude grid would need 2^32 work-items,
and fused passes at N = 28 with two batch items (2^17 tiles: two grid rows
of 2^16), against <Z_k> = cos(theta_k) after RY, CZ and RZ layers and
against a circuit followed by its inverse."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fills_of_two_to_the_32_amplitudes(ctx):
    from smoq_amd import GPUState
    from smoq_amd.observables import expect_pauli

    N = 32
    print(f"[case] product/basis fill N={N} c64", flush=True)
    s = GPUState(N, "c64", 1, ctx=ctx)
    try:
        c, d = math.cos(0.3), math.sin(0.3)
        kets = [(c, d)] * N
        kets[N - 1] = (1 / math.sqrt(2), 1 / math.sqrt(2))
        s.set_product(kets)
        assert abs(s.norm2()[0] - 1.0) < 1e-4
        hi = 1 << (N - 1)
        v = expect_pauli(s, [0, 0, hi, 1], [1, hi, 0, 0])
        assert np.allclose(v, [c * c - d * d, 0.0, 1.0, 2 * c * d], atol=1e-4)
        idx = (1 << 31) | 5
        s.set_basis(idx)
        assert abs(s.norm2()[0] - 1.0) < 1e-6
        z = expect_pauli(s, [0] * 4, [1, 2, 4, hi])
        assert np.array_equal(z, [-1.0, 1.0, -1.0, -1.0])
    finally:
        s.free()


def _inverse(tape):
    from smoq_amd.circuits import Tape

    swap = {"S": "SDG", "SDG": "S", "T": "TDG", "TDG": "T"}
    neg = {"RX", "RY", "RZ", "CRY", "RXX", "RYY", "RZZ"}
    inv = Tape(tape.n)
    for gate, qubits, theta, param, m in reversed(tape.ops):
        if gate in swap:
            inv.add(swap[gate], qubits)
        elif gate in neg:
            inv.add(gate, qubits, -theta)
        elif gate in ("U1", "U2"):
            inv.add(gate, qubits, m=np.conj(np.asarray(m)).T)
        elif gate == "Y":
            inv.add("Y", qubits)
        else:
            inv.add(gate, qubits, theta)
    return inv


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-10), ("c64", 2e-4)])
def test_fused_passes_on_two_grid_rows(ctx, dtype, tol):
    from smoq_amd import GPUState
    from smoq_amd.circuits import Tape
    from smoq_amd.observables import expect_pauli
    from test_gpu_gates import random_tape

    N, B = 28, 2
    print(f"[case] two-row passes N={N} B={B} {dtype}", flush=True)
    rng = np.random.default_rng(28)
    th = rng.uniform(0, math.pi, N)
    t = Tape(N)
    for k in range(N):
        t.add("RY", [k], th[k])
    for k in range(N - 1):
        t.add("CZ", [k, k + 1])
    for k in range(N):
        t.add("RZ", [k], 0.7 * k)
    s = GPUState(N, dtype, B, ctx=ctx)
    try:
        s.set_basis(0)
        t.apply(s)
        z = expect_pauli(s, [0] * N, [1 << k for k in range(N)])
        assert np.max(np.abs(z - np.cos(th)[None, :])) < tol
        # a random circuit and its inverse return every item to |0...0>
        r = random_tape(N, 120, 5)
        s.set_basis(0)
        r.apply(s)
        _inverse(r).apply(s)
        z = expect_pauli(s, [0] * N, [1 << k for k in range(N)])
        assert np.max(np.abs(z - 1.0)) < tol
        assert np.max(np.abs(s.norm2() - 1.0)) < tol
    finally:
        s.free()
