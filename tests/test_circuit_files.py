"""The reference's shared random circuits (tests/golden/circuits, copied
unchanged from SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/circuits)
through the package loader and the oracle's restatement of the benchmark
(load_circuit_from_file, fuse_gates, run_fused_circuit!, run_circuit_sampling),
pinned by brute force: dense 2^N unitaries, unfused gate-by-gate runs, an
independent count of fused sweeps and numpy's inverse-CDF search.  CPU only."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
CIRCUITS = os.path.join(HERE, "golden", "circuits")
SIZES = (2, 4, 6, 8, 10, 12, 14, 16, 18, 20)


def path(N):
    return os.path.join(CIRCUITS, f"circuit_N{N:02d}.txt")


def dense_gate(name, q1, q2, th, N):
    """Full 2^N matrix of one benchmark gate (1-based qubits)."""
    D = 1 << N
    idx = np.arange(D)
    if name == "CZ":
        b = ((idx >> (q1 - 1)) & 1) & ((idx >> (q2 - 1)) & 1)
        return np.diag(1.0 - 2.0 * b).astype(np.complex128)
    if name == "CNOT":
        tgt = np.where((idx >> (q1 - 1)) & 1, idx ^ (1 << (q2 - 1)), idx)
        M = np.zeros((D, D), np.complex128)
        M[tgt, idx] = 1.0
        return M
    return orc.dense_1q(orc.bench_gate_matrix(name, th), q1 - 1, N)


@pytest.mark.parametrize("N", SIZES)
def test_loader_reads_every_gate(N):
    from smoq_amd.circuits import load_circuit_file

    ops = orc.read_circuit_file(path(N))
    header = open(path(N)).readline()
    assert f"N={N} " in header and "n_gates=1000" in header
    assert len(ops) == 1000
    tape = load_circuit_file(path(N))
    assert tape.n == N and len(tape) == 1000
    for (name, q1, q2, th), (g, qs, theta, _, _) in zip(ops, tape.ops):
        assert g == name
        assert qs == ((q1 - 1,) if q2 == 0 else (q1 - 1, q2 - 1))
        assert theta == th
    two = sum(1 for o in ops if o[2])
    assert 0.2 < two / 1000 < 0.4  # generated with single_ratio 0.7


def independent_sweep_count(ops, N):
    """Fused sweeps counted per qubit: every 2-qubit gate, plus one sweep per
    maximal run of 1-qubit gates on a qubit not broken by a 2-qubit gate
    touching that qubit."""
    count = 0
    for q in range(1, N + 1):
        open_run = False
        for name, q1, q2, _ in ops:
            if q2 == 0 and q1 == q:
                open_run = True
            elif q2 != 0 and q in (q1, q2):
                count += open_run
                open_run = False
        count += open_run
    return count + sum(1 for o in ops if o[2])


@pytest.mark.parametrize("N", SIZES)
def test_fuse_gates_sweep_count(N):
    ops = orc.read_circuit_file(path(N))
    fused = orc.fuse_gates(ops)
    assert len(fused) == independent_sweep_count(ops, N)
    if N == 20:
        assert len(fused) == 621


@pytest.mark.parametrize("N", (2, 4, 6))
def test_fused_run_matches_dense_unitary(N):
    ops = orc.read_circuit_file(path(N))
    U = np.eye(1 << N, dtype=np.complex128)
    for name, q1, q2, th in ops:
        U = dense_gate(name, q1, q2, th, N) @ U
    psi, _ = orc.run_circuit_sampling(N, ops, [0.5])
    assert np.max(np.abs(psi - U[:, 0])) < 1e-12


@pytest.mark.parametrize("N", (8, 10, 12))
def test_fused_run_matches_unfused_gate_by_gate(N):
    ops = orc.read_circuit_file(path(N))
    psi, _ = orc.run_circuit_sampling(N, ops, [0.5])
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1.0
    for name, q1, q2, th in ops:
        qs = [q1] if q2 == 0 else [q1, q2]
        orc.apply_named(ref, name.lower(), qs, th, N)
    assert np.max(np.abs(psi - ref)) < 1e-12
    assert abs(np.vdot(psi, psi).real - 1.0) < 1e-12


def test_sampling_matches_numpy_search():
    from smoq_amd.streams import uniforms

    N = 10
    ops = orc.read_circuit_file(path(N))
    r = uniforms(42, 100)
    psi, idx = orc.run_circuit_sampling(N, ops, r)
    c = np.cumsum(np.abs(psi) ** 2)
    want = np.searchsorted(c, r, side="left")
    # the pairwise and sequential prefixes may differ in the last ulp: a
    # draw may move only where r sits within rounding of a prefix value
    moved = idx != want
    assert moved.sum() == 0 or np.all(np.abs(c[np.minimum(want[moved], c.size - 1)] - r[moved])
                                      < 1e-14)


def test_expectation_benchmark_circuit_gate_counts():
    from smoq_amd.circuits import expectation_benchmark_circuit

    for N, gates in [(12, 188), (14, 220), (16, 252), (18, 284), (20, 316), (22, 348), (24, 380)]:
        assert len(expectation_benchmark_circuit(N)) == gates
    with pytest.raises(ValueError):
        expectation_benchmark_circuit(25)
