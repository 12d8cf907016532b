"""Pauli reductions compiled for the term list (smoq_set_obs_jit): every term
against index arithmetic on the host (numpy, no shared code with the
device path) and against the precompiled tiled passes, for chain, rail and
random term lists (x masks of any weight, z bits inside and outside the
tile bits), single states and batches, complex128 and complex64."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rand_states(N, B, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(B, 1 << N)) + 1j * rng.normal(size=(B, 1 << N))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def pauli_free(psi, x, z):
    """<psi| i^{|x&z|} X^x Z^z |psi> by index arithmetic."""
    idx = np.arange(psi.size, dtype=np.int64)
    sgn = 1.0 - 2.0 * (np.bitwise_count(idx & z) & 1)
    ph = 1j ** (bin(x & z).count("1") & 3)
    return float(np.real(np.vdot(psi[idx ^ x], ph * sgn * psi)))


def term_lists(N, rng):
    from smoq_amd import observables as ob

    xs, zs = ob.qrc_observable_masks(N, 1)
    yield "chain", xs, zs
    if N % 2 == 0:
        xs, zs = ob.qrc_observable_masks(N // 2, 2)
        yield "two rails", xs, zs
    xs, zs = [], []
    for _ in range(40):
        w = int(rng.integers(1, 7))
        x = 0
        for q in rng.choice(N, w, replace=False):
            x |= 1 << int(q)
        xs.append(x if rng.random() < 0.85 else 0)
        zs.append(int(rng.integers(0, 1 << N)))
    yield "random strings", xs, zs


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("N,B", [(12, 1), (13, 3), (16, 2), (20, 1)])
def test_compiled_reduction_parity(ctx, N, B, dtype):
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    rng = np.random.default_rng(900 + N)
    psis = rand_states(N, B, 77 + N)
    if dtype == "c64":
        psis = psis.astype(np.complex64)
    ref_states = psis.astype(np.complex128)
    s = GPUState.from_numpy(psis if B > 1 else psis[0])
    tol = 1e-12 if dtype == "c128" else 1e-10
    try:
        for name, xs, zs in term_lists(N, rng):
            print(f"[obs-jit] N={N} B={B} {dtype} {name}: {len(xs)} terms", flush=True)
            ctx.set_obs_jit(2)
            got = np.atleast_2d(ob.expect_pauli(s, xs, zs))
            ctx.set_obs_jit(0)
            pre = np.atleast_2d(ob.expect_pauli(s, xs, zs))
            for b in range(B):
                want = np.array([pauli_free(ref_states[b], x, z) for x, z in zip(xs, zs)])
                assert np.max(np.abs(got[b] - want)) < tol, (name, b)
                assert np.max(np.abs(pre[b] - want)) < tol, (name, b)
    finally:
        ctx.set_obs_jit(1)
        s.free()


def test_compiled_reduction_measure_all_observables(ctx):
    """measure_all_observables_* on rails through the compiled passes, in
    the reference's order, against the oracle's per-observable loops."""
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    ctx.set_obs_jit(2)
    try:
        for L, rails in [(13, 1), (7, 2), (4, 3), (14, 1)]:
            N = L * rails
            psi = rand_states(N, 1, 300 + N)[0]
            s = GPUState.from_numpy(psi)
            want = []
            for k in range(1, N + 1):
                want += [orc.expect_local(psi, k, N, p) for p in "xyz"]
            bonds = [(r * L + q + 1, r * L + q + 2) for r in range(rails) for q in range(L - 1)]
            bonds += [(r * L + q + 1, (r + 1) * L + q + 1)
                      for r in range(rails - 1) for q in range(L)]
            for i, j in bonds:
                want += [orc.expect_corr(psi, i, j, N, p) for p in ("xx", "yy", "zz")]
            got = ob.measure_all_observables_state_vector(s, L, rails)
            assert np.max(np.abs(got - np.array(want))) < 1e-12, (L, rails)
            s.free()
    finally:
        ctx.set_obs_jit(1)


def test_compiled_reduction_repeat_calls_identical(ctx):
    """The cached plan and module give bit-identical results call after call
    (fixed summation order), and a state change is seen."""
    from smoq_amd import GPUState
    from smoq_amd import observables as ob

    N = 18
    xs, zs = ob.qrc_observable_masks(N, 1)
    psi = rand_states(N, 2, 5)
    s = GPUState.from_numpy(psi)
    ctx.set_obs_jit(2)
    try:
        a = ob.expect_pauli(s, xs, zs)
        b = ob.expect_pauli(s, xs, zs)
        assert np.array_equal(a, b)
        s2 = GPUState.from_numpy(psi[::-1].copy())
        c = ob.expect_pauli(s2, xs, zs)
        assert np.array_equal(c[0], a[1]) and np.array_equal(c[1], a[0])
        s2.free()
    finally:
        ctx.set_obs_jit(1)
        s.free()
