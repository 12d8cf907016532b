"""Context and state lifecycle rules of the C ABI, the state pool's reuse
contract, and the bounded rotation factoring of compiled passes."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_fini_refuses_while_states_or_plans_live():
    from smoq_amd import GPUState
    from smoq_amd._lib import SMOQ_EINVAL, SmoqError
    from smoq_amd.circuits import Tape
    from smoq_amd.state import Context

    c = Context(0)
    s = GPUState.basis(6, 3, ctx=c)
    plan = Tape(6).add("H", [0]).compile(ctx=c)
    with pytest.raises(SmoqError) as e:
        c.close()
    assert e.value.code == SMOQ_EINVAL
    plan.free()
    with pytest.raises(SmoqError):
        c.close()
    # the context is still usable after a refused fini
    assert s.to_numpy()[3] == 1.0
    s.free()
    c.close()
    assert c.handle is None


def test_pooled_buffer_is_zero_filled_for_another_shape(ctx):
    from smoq_amd import GPUState
    import smoq_amd.gates as g

    a = GPUState.basis(10, 5, "c128")
    for q in range(10):
        g.apply_hadamard_psi(a, q + 1, 10)
    assert np.count_nonzero(a.to_numpy()) == 1 << 10
    a.free()
    # same byte size, other qubit count and dtype: the pooled buffer comes back
    b = GPUState(11, "c64")
    assert b.nbytes == 16 << 10
    assert not np.any(b.to_numpy())
    b.free()
    # same byte size, other batch
    c = GPUState.basis(9, 7, "c128", batch=2)
    c.free()
    d = GPUState(10, "c128")
    assert not np.any(d.to_numpy())
    d.free()


def test_evolve_observe_rejects_a_state_of_another_context(ctx):
    from smoq_amd import GPUState
    from smoq_amd._lib import SMOQ_EINVAL, SmoqError
    from smoq_amd.circuits import Tape
    from smoq_amd.state import Context

    other = Context(0)
    s = GPUState.basis(6, 0, ctx=other)
    plan = Tape(6).add("RY", [2], 0.3).compile(ctx=ctx)
    with pytest.raises(SmoqError) as e:
        plan.evolve_observe(s, 2, 1, [0], [1])
    assert e.value.code == SMOQ_EINVAL
    plan.free()
    s.free()
    other.close()


def test_shadow_snapshots_reject_batched_states(ctx):
    from smoq_amd import GPUState
    from smoq_amd._lib import SMOQ_EINVAL, SmoqError
    from smoq_amd.shadows import collect_shadows

    s = GPUState.basis(4, 0, batch=2)
    with pytest.raises(SmoqError) as e:
        collect_shadows(s, 4, 3, bases=np.ones((3, 4), np.uint8), uniforms=np.full(3, 0.5))
    assert e.value.code == SMOQ_EINVAL
    s.free()


def test_long_factored_rotation_pass_stays_finite_in_c64(ctx):
    # ~300 Rx(pi/2) in one compiled pass of a 12-qubit state (one tile): each
    # factored Rx(pi/2) scales the tile by sqrt 2, so unbounded factoring
    # would leave the float range; factoring stops at 2^40 in complex64
    from smoq_amd import GPUState
    from smoq_amd.circuits import Tape

    N = 12
    t = Tape(N)
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1.0
    n_rx = 0
    for layer in range(26):
        for q in range(N):
            t.add("RX", [q], math.pi / 2)
            orc.apply_named(ref, "rx", [q + 1], math.pi / 2, N)
            n_rx += 1
        for q in range(layer % 2, N - 1, 2):
            t.add("CZ", [q, q + 1])
            orc.apply_named(ref, "cz", [q + 1, q + 2], 0.0, N)
    assert n_rx >= 300
    ctx.set_jit(2)
    try:
        for dtype, tol in (("c64", 1e-5), ("c128", 1e-12)):
            s = GPUState.basis(N, 0, dtype)
            t.apply(s)
            got = s.to_numpy().astype(np.complex128)
            s.free()
            assert np.all(np.isfinite(got)), dtype
            assert np.linalg.norm(got - ref) < tol, dtype
    finally:
        ctx.set_jit(1)
