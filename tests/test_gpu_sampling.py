"""Born sampling of long vectors (more than 2^16 amplitudes: leaf sums staged
through LDS, leaf-end search) against the oracle's full pairwise cumsum and
searchsortedfirst, bit for bit.  Structured states (sparse, wide dynamic
range, zero tails) make the reference's prefix array step down across leaf
boundaries; draws are placed inside those steps and at their ends, where
the reference's search result depends on its probe path."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def structured_state(kind, N, seed):
    rng = np.random.default_rng(seed)
    d = 1 << N
    if kind == "gauss":
        v = rng.normal(size=d) + 1j * rng.normal(size=d)
    elif kind == "sparse":
        v = np.zeros(d, np.complex128)
        idx = rng.choice(d, 2000, replace=False)
        v[idx] = rng.normal(size=2000) + 1j * rng.normal(size=2000)
    elif kind == "lognormal":
        v = np.exp(rng.normal(scale=6, size=d)) * np.exp(2j * np.pi * rng.random(d))
    elif kind == "blocks":
        v = rng.normal(size=d) + 1j * rng.normal(size=d)
        v[rng.random(d) < 0.7] *= 1e-12
    elif kind == "zero_chunks":  # every third 4096-term chunk all zero, and a zero tail
        v = rng.normal(size=d) + 1j * rng.normal(size=d)
        v[(np.arange(d) // 4096) % 3 == 1] = 0
        v[d - d // 5:] = 0
    else:  # zero tail
        v = rng.normal(size=d) + 1j * rng.normal(size=d)
        v[d // 3:] = 0
    return v / np.linalg.norm(v)


def draws_at_steps(psi, rng, n_random=2000, max_steps=60):
    c = orc.cumsum_pairwise(orc.probs(psi))
    steps = np.nonzero(np.diff(c) < 0)[0]
    r = list(rng.random(n_random))
    r += [0.0, 1.0, 1.0 - 2.0 ** -53, 1e-300, c[-1], np.nextafter(c[-1], 2.0), c[0]]
    for i in steps[:max_steps]:
        r += [c[i], np.nextafter(c[i + 1], 2.0), c[i + 1]]
    return np.array(r), len(steps)


@pytest.mark.parametrize("N", [17, 19, 21])
@pytest.mark.parametrize("kind", ["gauss", "sparse", "lognormal", "blocks", "zero_tail"])
def test_long_vector_sampling_bit_exact(ctx, N, kind):
    from smoq_amd import GPUState
    from smoq_amd import measurements as me

    psi = structured_state(kind, N, 1000 * N + len(kind))
    r, n_steps = draws_at_steps(psi, np.random.default_rng(N))
    if kind in ("sparse", "lognormal", "blocks"):
        assert n_steps > 0  # the step path is exercised
    s = GPUState.from_numpy(psi)
    got = me.sample_indices(s, r, "cumsum").astype(np.int64)
    s.free()
    want = orc.sample_cumsum_search(psi, N, r)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:10]


def test_long_vector_sampling_c64_and_unnormalised(ctx):
    from smoq_amd import GPUState
    from smoq_amd import measurements as me

    N = 18
    psi = structured_state("lognormal", N, 5).astype(np.complex64)
    s = GPUState.from_numpy(psi)
    wide = psi.astype(np.complex128)  # the oracle squares the same float values in FP64
    r, _ = draws_at_steps(wide, np.random.default_rng(3))
    got = me.sample_indices(s, r, "cumsum").astype(np.int64)
    assert np.array_equal(got, orc.sample_cumsum_search(wide, N, r))
    s.free()
    # half the norm: draws above the last prefix give 2^N, as the reference
    half = (structured_state("gauss", N, 6) * np.sqrt(0.5)).astype(np.complex128)
    s = GPUState.from_numpy(half)
    r = np.random.default_rng(4).random(3000)
    got = me.sample_indices(s, r, "cumsum").astype(np.int64)
    want = orc.sample_cumsum_search(half, N, r)
    assert np.array_equal(got, want)
    assert np.count_nonzero(got == 1 << N) > 1000
    s.free()


def test_long_vector_sampling_repeats_and_batch_item_zero(ctx):
    # the same state sampled twice (cached buffers reused), and a batched
    # state samples its first item
    from smoq_amd import GPUState
    from smoq_amd import measurements as me

    N = 17
    a = structured_state("blocks", N, 8)
    b = structured_state("sparse", N, 9)
    s = GPUState.from_numpy(np.stack([a, b]))
    r, _ = draws_at_steps(a, np.random.default_rng(5))
    first = me.sample_indices(s, r, "cumsum").astype(np.int64)
    again = me.sample_indices(s, r, "cumsum").astype(np.int64)
    s.free()
    want = orc.sample_cumsum_search(a, N, r)
    assert np.array_equal(first, want)
    assert np.array_equal(again, want)


def tie_state(N):
    """Every |psi_i|^2 = 2^-20 + 2^-54 (amplitude 2^-10 + 2^-45): the running
    sum is exact below 1/2 and every step above 1/2 rounds a tie."""
    return np.full(1 << N, 2.0 ** -10 + 2.0 ** -45, np.complex128)


@pytest.mark.parametrize("kind,N", [(k, n) for n in (19, 20) for k in
                                    ("gauss", "lognormal", "blocks", "tail", "zero_chunks", "ties")]
                         + [("lognormal", 26), ("zero_chunks", 26)])
def test_sequential_prefix_in_chunks_bit_exact(ctx, kind, N):
    """The reference's sequential running sum (sample_from_probabilities),
    computed in parallel chunks whose starts are shifted from a guess, is
    the oracle's chain bit for bit: draws placed exactly on the oracle's
    prefix values and one grid step either side of them land where the
    oracle's do.  States cross many binades (lognormal, blocks), end in
    zeros (tail), hold whole chunks of zeros (zero_chunks) or round a tie at
    every step above 1/2 (ties).  At N = 26 the 16384 chunks take two rounds
    of the walk over the chunk ends."""
    from smoq_amd import GPUState
    from smoq_amd import measurements as me

    psi = tie_state(N) if kind == "ties" else structured_state(kind, N, 90 + N)
    c = orc.cumsum_sequential(orc.probs(psi))
    rng = np.random.default_rng(N)
    pos = np.concatenate([rng.integers(0, psi.size, 600), np.arange(0, psi.size, 4096),
                          np.arange(4095, psi.size, 4096)])
    pos = pos[pos < psi.size]
    r = np.concatenate([c[pos], np.nextafter(c[pos], 0.0), np.nextafter(c[pos], 2.0),
                        [0.0, 1.0 - 2.0 ** -53]])
    r = r[(r >= 0.0) & (r < 1.0)]
    s = GPUState.from_numpy(psi)
    got = me.sample_indices(s, r, "sequential").astype(np.int64)
    if N <= 20:
        want = orc.sample_sequential(psi, N, r)
    else:
        # the same rule (first i with r <= c_i) on the oracle's sequential
        # prefix: its per-draw walk is O(2^N)
        want = np.minimum(np.searchsorted(c, r, side="left"), psi.size - 1)
    assert np.array_equal(got, want)
