"""The fused-pass autotuner (smoq_autotune), its saved table, and the
on-disk cache of compiled pass modules: tuned and cached runs give the same
amplitudes as the default plan and the oracle."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def oracle_state(tape, N):
    ref = np.zeros(1 << N, np.complex128)
    ref[0] = 1.0
    for gate, qs, th, _, _ in tape.ops:
        orc.apply_named(ref, gate.lower(), [q + 1 for q in qs], th, N)
    return ref


def test_autotune_choice_keeps_results(ctx, tmp_path):
    from smoq_amd import GPUState
    from smoq_amd.circuits import random_rotation_brick
    from smoq_amd.state import Context

    N = 20
    tape = random_rotation_brick(N, 12, seed=41)
    ref = oracle_state(tape, N)
    c = Context(0)
    s = GPUState(N, ctx=c)
    s.set_basis(0)
    tape.apply(s)
    before = s.to_numpy()
    assert np.linalg.norm(before - ref) < 1e-12
    res = c.autotune(tape, s)
    assert res["candidates"] >= 3
    assert res["ms"] > 0 and res["ms"] <= res["default_ms"]
    assert res["tile_qubits"] in (11, 12) and res["low_qubits"] in (4, 5)
    assert res["pass_max_gates"] in (0, 28, 36, 44)
    assert res["lane_swaps"] in (0, 1) and res["searched"] in (0, 1)
    s.set_basis(0)
    n_pass = tape.apply(s)
    assert n_pass == res["n_passes"]
    assert np.linalg.norm(s.to_numpy() - ref) < 1e-12
    # the table survives a save / load into a new context
    path = tmp_path / "tune.txt"
    c.tune_save(path)
    s.free()
    c.close()
    c2 = Context(0)
    assert c2.tune_load(path) >= 1
    s2 = GPUState.basis(N, 0, ctx=c2)
    assert tape.apply(s2) == res["n_passes"]
    assert np.linalg.norm(s2.to_numpy() - ref) < 1e-12
    s2.free()
    c2.close()


def test_jit_disk_cache_round_trip(monkeypatch, tmp_path):
    from smoq_amd import GPUState
    from smoq_amd.circuits import random_rotation_brick
    from smoq_amd.state import Context

    monkeypatch.setenv("SMOQ_JIT_CACHE_DIR", str(tmp_path))
    N = 16
    tape = random_rotation_brick(N, 8, seed=43)
    ref = oracle_state(tape, N)
    got = []
    for _ in range(2):  # compile and store, then load in a fresh context
        c = Context(0)
        c.set_jit(2)
        s = GPUState.basis(N, 0, ctx=c)
        tape.apply(s)
        got.append(s.to_numpy())
        s.free()
        c.close()
        files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".co"))
        assert len(files) >= 1
    assert np.array_equal(got[0], got[1])
    assert np.linalg.norm(got[0] - ref) < 1e-12


def test_pass_hbm_orders_keep_results(tmp_path):
    """Per-pass tile orders and load mappings (the tuned HBM order of each
    pass, set here through a saved table) move no amplitude: the result is
    bit-identical to the default order and matches the oracle."""
    from smoq_amd import GPUState
    from smoq_amd.circuits import random_rotation_brick
    from smoq_amd.state import Context

    N = 20
    tape = random_rotation_brick(N, 12, seed=47)
    ref = oracle_state(tape, N)
    c = Context(0)
    s = GPUState(N, ctx=c)
    res = c.autotune(tape, s)
    s.set_basis(0)
    tape.apply(s)
    plain = s.to_numpy()
    assert np.linalg.norm(plain - ref) < 1e-12
    path = tmp_path / "tune.txt"
    c.tune_save(path)
    s.free()
    c.close()
    line = path.read_text().split("\n")[0].split()
    # arch n dtype batch hash tile low cap swaps, then the tuned local sets
    # when the search won (kept), without the tuned orders
    base = line[:9]
    if "loc" in line:
        i = line.index("loc")
        base += line[i:i + 2 + int(line[i + 1])]
    n_pass = res["n_passes"]
    # rotations, reversals (32) and registers above the low qubits (64)
    codes = [3, 32 | 5, 64, 64 | 32 | 7, 6, 32, 64 | 2, 1]
    io = [codes[p % len(codes)] for p in range(n_pass)]
    path.write_text(" ".join(base + ["io", str(n_pass)] + [str(x) for x in io]) + "\n")
    c2 = Context(0)
    assert c2.tune_load(path) == 1
    s2 = GPUState.basis(N, 0, ctx=c2)
    assert tape.apply(s2) == n_pass
    got = s2.to_numpy()
    assert np.array_equal(got, plain)
    # the same through a compiled plan
    plan = tape.compile(ctx=c2)
    s2.set_basis(0)
    plan.run(s2)
    assert np.array_equal(s2.to_numpy(), plain)
    plan.free()
    s2.free()
    c2.close()
