"""Reduced density matrices on the device (smoq_partial_trace,
smoq_dm_partial_trace) against the oracle's restatement of partial_trace:
the LDS-tiled kernel (1..6 kept qubits), the split-K GEMM path (7 and more,
kept qubits lowest or gathered), density matrices, partial_trace_regions,
and large product states against the closed form (the Kronecker product of
the kept qubits' |k><k|).  complex128 within 1e-12, complex64 within 2e-5."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rand_batch(N, B, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(B, 1 << N)) + 1j * rng.normal(size=(B, 1 << N))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def keep_to_trace(keep, N):
    return [q for q in range(1, N + 1) if q not in set(keep)]


TILE = [(6, [1]), (6, [6]), (6, [2, 5]), (12, [1, 2, 3]), (12, [12, 1, 7, 4]),
        (12, [3, 4, 5, 6, 7]), (12, [1, 3, 5, 7, 9, 11]), (16, [16, 8]), (16, [1, 2, 3, 4, 5, 6]),
        (5, [1, 2, 3, 4, 5])]


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-5)])
@pytest.mark.parametrize("N,keep", TILE)
def test_tile_path_matches_oracle(ctx, N, keep, dtype, tol):
    from smoq_amd import GPUState
    from smoq_amd.partial_trace import partial_trace

    print(f"[case] tile N={N} keep={keep} {dtype}", flush=True)
    B = 3
    psi = rand_batch(N, B, N * 7 + len(keep))
    if dtype == "c64":
        psi = psi.astype(np.complex64).astype(np.complex128)
    s = GPUState.from_numpy(psi.astype(np.complex64) if dtype == "c64" else psi)
    tr = keep_to_trace(keep, N)
    r = partial_trace(s, tr, N)
    got = r.to_numpy()
    for b in range(B):
        want = orc.partial_trace_psi(psi[b], tr, N)
        assert np.max(np.abs(got[b] - want)) < tol
    r.free()
    s.free()


GEMM = [(12, list(range(1, 8))), (12, [1, 2, 4, 5, 7, 9, 10, 12]), (14, list(range(1, 11))),
        (10, list(range(1, 11))), (13, [2, 3, 5, 7, 11, 13, 1, 8])]


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-5)])
@pytest.mark.parametrize("N,keep", GEMM)
def test_gemm_path_matches_oracle(ctx, N, keep, dtype, tol):
    from smoq_amd import GPUState
    from smoq_amd.partial_trace import partial_trace

    print(f"[case] gemm N={N} keep={keep} {dtype}", flush=True)
    B = 2
    psi = rand_batch(N, B, N * 11 + len(keep))
    if dtype == "c64":
        psi = psi.astype(np.complex64).astype(np.complex128)
    s = GPUState.from_numpy(psi.astype(np.complex64) if dtype == "c64" else psi)
    tr = keep_to_trace(keep, N)
    r = partial_trace(s, tr, N)
    got = r.to_numpy()
    for b in range(B):
        want = orc.partial_trace_psi(psi[b], tr, N)
        assert np.max(np.abs(got[b] - want)) < tol
    r.free()
    s.free()


@pytest.mark.parametrize("N,keep", [(26, [3, 26]), (24, [1, 5, 9, 13, 17, 21]),
                                    (24, list(range(1, 11))), (22, [2, 4, 6, 8, 10, 12, 14, 16])])
def test_large_product_states_closed_form(ctx, N, keep):
    from smoq_amd import GPUState
    from smoq_amd.partial_trace import partial_trace

    print(f"[case] product N={N} keep={keep}", flush=True)
    rng = np.random.default_rng(N)
    a = rng.uniform(0, np.pi, N)
    ph = rng.uniform(0, 2 * np.pi, N)
    kets = [np.array([np.cos(a[q]), np.exp(1j * ph[q]) * np.sin(a[q])]) for q in range(N)]
    s = GPUState(N, "c128", ctx=ctx)
    s.set_product(kets)
    r = partial_trace(s, keep_to_trace(keep, N), N)
    want = np.ones((1, 1), np.complex128)
    for q in sorted(keep, reverse=True):  # highest kept qubit first in the Kronecker product
        want = np.kron(want, np.outer(kets[q - 1], kets[q - 1].conj()))
    assert np.max(np.abs(r.to_numpy() - want)) < 1e-12
    r.free()
    s.free()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-5)])
def test_density_matrix_form(ctx, dtype, tol):
    from smoq_amd.density import DensityMatrix
    from smoq_amd.partial_trace import partial_trace

    N = 6
    rng = np.random.default_rng(3)
    rhos = []
    for b in range(2):
        A = rng.normal(size=(1 << N, 1 << N)) + 1j * rng.normal(size=(1 << N, 1 << N))
        r = A @ A.conj().T
        rhos.append(r / np.trace(r).real)
    rhos = np.stack(rhos)
    if dtype == "c64":
        rhos = rhos.astype(np.complex64).astype(np.complex128)
    d = DensityMatrix.from_numpy(rhos.astype(np.complex64) if dtype == "c64" else rhos)
    for keep in ([1], [6], [2, 5], [1, 2, 3], [1, 2, 3, 4, 5, 6], [6, 4, 2, 1]):
        tr = keep_to_trace(keep, N)
        r = partial_trace(d, tr, N)
        got = r.to_numpy()
        for b in range(2):
            assert np.max(np.abs(got[b] - orc.partial_trace_rho(rhos[b], tr, N))) < tol
        r.free()
    d.free()


def test_regions_and_errors(ctx):
    from smoq_amd import GPUState, SmoqError
    from smoq_amd._lib import SMOQ_EINVAL, call
    from smoq_amd.partial_trace import (find_connected_regions, partial_trace,
                                        partial_trace_regions)

    # the reference docstring's examples (…PartialTrace.jl:260-265)
    assert find_connected_regions([1, 2, 3, 5, 6, 8, 10]) == [[1, 2, 3], [5, 6], [8], [10]]
    assert find_connected_regions([5, 3, 4, 1, 2]) == [[1, 2, 3, 4, 5]]
    assert find_connected_regions([1, 3, 5, 7]) == [[1], [3], [5], [7]]
    N = 10
    psi = rand_batch(N, 1, 4)[0]
    s = GPUState.from_numpy(psi)
    regs = partial_trace_regions(s, [4, 7, 9], N)
    assert len(regs) == 4
    for reg, r in zip([[1, 2, 3], [5, 6], [8], [10]], regs):
        want = orc.partial_trace_psi(psi, keep_to_trace(reg, N), N)
        assert np.max(np.abs(r.to_numpy() - want)) < 1e-12
        r.free()
    with pytest.raises(ValueError):
        partial_trace(s, list(range(1, N + 1)), N)
    bad = GPUState(3, "c128", ctx=ctx)  # 3 is not 2 x (kept qubits)
    with pytest.raises(SmoqError) as e:
        call("smoq_partial_trace", s.handle, 0b11, bad.handle)
    assert e.value.code == SMOQ_EINVAL
    bad.free()
    s.free()
