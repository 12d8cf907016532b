"""smoq_expect_pauli_noalloc (partial sums in a fixed device region, no
allocation) against smoq_expect_pauli, through the C ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _terms(N):
    # X, Y, Z on every qubit, XX, YY, ZZ on neighbours, a long Z string and
    # a mask spanning the register (per-mask kernel path)
    xs, zs = [], []
    for k in range(N):
        b = 1 << k
        xs += [b, b, 0]
        zs += [0, b, b]
    for k in range(N - 1):
        b = 3 << k
        xs += [b, b, 0]
        zs += [0, b, b]
    xs += [0, (1 << (N - 1)) | 1]
    zs += [(1 << N) - 1, 1 << (N // 2)]
    return np.ascontiguousarray(xs, np.uint64), np.ascontiguousarray(zs, np.uint64)


def _both(s, xs, zs):
    from smoq_amd import _lib

    L = _lib.lib()
    a = np.full((s.batch, xs.size), np.nan)
    b = np.full((s.batch, xs.size), np.nan)
    _lib.check(L.smoq_expect_pauli(s.handle, xs.size, _p(xs), _p(zs), _p(a)))
    _lib.check(L.smoq_expect_pauli_noalloc(s.handle, xs.size, _p(xs), _p(zs), _p(b)))
    return a, b


def _random_state(N, batch, seed, dtype):
    from smoq_amd import GPUState

    rng = np.random.default_rng(seed)
    v = rng.normal(size=(batch, 1 << N)) + 1j * rng.normal(size=(batch, 1 << N))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return GPUState.from_numpy(v.astype(np.complex64 if dtype == "c64" else np.complex128))


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_matches_expect_pauli_two_calls(ctx, dtype):
    # first call on the larger input, then a smaller one: the fixed region
    # keeps the first call's contents, which must not leak into the second
    for N, batch, seed in ((18, 2, 1), (9, 1, 2)):
        s = _random_state(N, batch, seed, dtype)
        xs, zs = _terms(N)
        a, b = _both(s, xs, zs)
        s.free()
        assert np.all(np.isfinite(a)), (N, dtype)
        assert np.array_equal(a, b), (N, dtype, np.max(np.abs(a - b)))


def test_noalloc_rejects_requests_beyond_its_region(ctx):
    from smoq_amd import _lib

    # 80 items need 80 x 256 KiB of per-mask partial sums: beyond 16 MiB
    s = _random_state(10, 80, 3, "c128")
    xs, zs = _terms(10)
    out = np.zeros((80, xs.size))
    rc = _lib.lib().smoq_expect_pauli_noalloc(s.handle, xs.size, _p(xs), _p(zs), _p(out))
    s.free()
    assert rc == _lib.SMOQ_ENOMEM
    assert b"fixed region" in _lib.lib().smoq_last_error()
