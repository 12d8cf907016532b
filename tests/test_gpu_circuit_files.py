"""The reference's shared random circuits (N = 2..20, 1000 gates each) and
its expectation-value benchmark circuit on the device, against the oracle's
restatement of those benchmarks: amplitudes, the 100-shot Born samples of
run_circuit_sampling (bit-exact for the same uniforms) and the 3N + 3(N-1)
observables."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def path(N):
    return os.path.join(HERE, "golden", "circuits", f"circuit_N{N:02d}.txt")


@pytest.mark.parametrize("N", (2, 4, 6, 8, 10, 12, 14, 16, 18, 20))
def test_shared_circuit_amplitudes_and_samples(ctx, N):
    from smoq_amd import GPUState
    from smoq_amd.circuits import load_circuit_file, run_circuit_sampling
    from smoq_amd.streams import uniforms

    tape = load_circuit_file(path(N))
    r = uniforms(4200 + N, 100)
    want_psi, want_idx = orc.run_circuit_sampling(N, orc.read_circuit_file(path(N)), r)
    # precompiled passes; compiled passes without and with lane exchanges
    for jit, swaps in ((0, 0), (2, 0), (2, 1)):
        ctx.set_jit(jit)
        ctx.set_lane_swaps(swaps)
        s = GPUState.basis(N, 0)
        tape.apply(s)
        got = s.to_numpy()
        s.free()
        assert np.max(np.abs(got - want_psi)) < 1e-12, (N, jit, swaps)
        idx, bits = run_circuit_sampling(tape, r)
        np.testing.assert_array_equal(idx, want_idx)
        assert bits.shape == (100, N)
        np.testing.assert_array_equal(bits @ (1 << np.arange(N)), want_idx)
    ctx.set_jit(1)
    ctx.set_lane_swaps(1)


def test_shared_circuit_complex64(ctx):
    from smoq_amd import GPUState
    from smoq_amd.circuits import load_circuit_file

    for N in (12, 20):
        want, _ = orc.run_circuit_sampling(N, orc.read_circuit_file(path(N)), [0.5])
        s = GPUState.basis(N, 0, "c64")
        load_circuit_file(path(N)).apply(s)
        assert np.max(np.abs(s.to_numpy() - want)) < 1e-5
        s.free()


@pytest.mark.parametrize("N", (4, 12, 17, 20))
def test_expectation_benchmark_values(ctx, N):
    """bench_smoq.jl's circuit and compute_observables: the device tape and
    measure_all_observables (one rail) against oracle gates and loops."""
    from smoq_amd import GPUState
    from smoq_amd import observables as ob
    from smoq_amd.circuits import expectation_benchmark_circuit

    tape = expectation_benchmark_circuit(N)
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1.0
    for gate, qs, th, _, _ in tape.ops:
        orc.apply_named(ref, gate.lower(), [q + 1 for q in qs], th, N)
    want = []
    for k in range(1, N + 1):
        want += [orc.expect_local(ref, k, N, p) for p in "xyz"]
    for i in range(1, N):
        want += [orc.expect_corr(ref, i, i + 1, N, p) for p in ("xx", "yy", "zz")]
    s = GPUState.basis(N, 0)
    tape.apply(s)
    got = ob.measure_all_observables_state_vector(s, N, 1)
    s.free()
    assert np.max(np.abs(got - np.array(want))) < 1e-12
