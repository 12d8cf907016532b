"""Gate tapes (the op list the C ABI consumes) and the seeded workloads.

A Tape is the device-facing form of the reference's circuit IR
(`ParameterizedCircuit` / `CircuitOperation`,
utils/cpu/cpuVariationalQuantumCircuitBuilder.jl:120-132,295-326) and of the
benchmark's `GateOp` lists (run_benchmark_smoq_Float64.jl:114-119).  Qubits
in a Tape are 0-based.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import streams
from ._lib import G, ROTATIONS, call, smoq_op

TWO_PI = 2.0 * math.pi
_NQ = {"CNOT": 2, "CZ": 2, "CRY": 2, "RXX": 2, "RYY": 2, "RZZ": 2, "U2": 2, "DIAG2": 2,
       "CCZ": 3, "TOFFOLI": 3}


class Tape:
    """A gate list for smoq_apply_tape / smoq_plan_create.  Change it through
    add(): the C array handed to the library is built once and reused until
    the next add()."""

    def __init__(self, n_qubits: int):
        self.n = n_qubits
        self.ops: list[tuple] = []  # (gate, qubits, theta, param, matrix)
        self.n_params = 0
        self._ct = None

    def add(self, gate: str, qubits, theta: float = 0.0, param: int = -1, m=None) -> "Tape":
        gate = gate.upper()
        qubits = tuple(int(q) for q in qubits)
        if len(qubits) != _NQ.get(gate, 1):
            raise ValueError(f"{gate} takes {_NQ.get(gate, 1)} qubits, got {qubits}")
        if any(not 0 <= q < self.n for q in qubits):
            raise ValueError(f"qubits {qubits} outside 0..{self.n - 1}")
        if param >= 0:
            self.n_params = max(self.n_params, param + 1)
        self.ops.append((gate, qubits, float(theta), int(param), m))
        self._ct = None
        return self

    def __len__(self):
        return len(self.ops)

    def to_ctypes(self):
        """The smoq_op array of the tape (cached; read-only for callers)."""
        if self._ct is not None and self._ct[0] == len(self.ops):
            return self._ct[1]
        arr = self._build_ctypes()
        self._ct = (len(self.ops), arr)
        return arr

    def _build_ctypes(self):
        arr = (smoq_op * max(1, len(self.ops)))()
        for i, (gate, qs, th, par, m) in enumerate(self.ops):
            o = arr[i]
            o.gate = G[gate]
            for j in range(3):
                o.q[j] = qs[j] if j < len(qs) else -1
            o.param = par
            o.flags = 0
            o.theta = th
            if m is not None:
                flat = np.asarray(m, np.complex128).reshape(-1)
                buf = np.empty(2 * flat.size)
                buf[0::2] = flat.real
                buf[1::2] = flat.imag
                for j, v in enumerate(buf):
                    o.m[j] = v
        return arr

    def apply(self, psi, theta=None) -> int:
        """Applies the tape to a GPUState with fused passes; returns the
        number of HBM passes."""
        if psi.n != self.n:
            raise ValueError(f"tape is for {self.n} qubits, state has {psi.n}")
        arr = self.to_ctypes()
        npass = ctypes.c_int(0)
        if theta is None:
            th_p, nth = None, 0
        else:
            th = np.ascontiguousarray(theta, np.float64)
            th_p, nth = th.ctypes.data_as(ctypes.c_void_p), th.size
            self._theta_keepalive = th
        call("smoq_apply_tape", psi.handle, arr, len(self.ops), th_p, nth, ctypes.byref(npass))
        return npass.value

    def compile(self, dtype="c128", theta=None, ctx=None) -> "Plan":
        return Plan(self, dtype, theta, ctx)

    def logical_gates(self) -> int:
        return len(self.ops)

    def rotation_params(self) -> list:
        return [(i, op) for i, op in enumerate(self.ops) if G[op[0]] in ROTATIONS and op[3] >= 0]


class Plan:
    """A tape planned once (smoq_plan_*): device-resident fused-pass program,
    launches replayed through a hipGraph on every run."""

    def __init__(self, tape: Tape, dtype="c128", theta=None, ctx=None):
        from .state import _DTYPES, default_context

        self.ctx = ctx or default_context()
        self.n = tape.n
        self.dtype = _DTYPES[dtype]
        arr = tape.to_ctypes()
        th = None if theta is None else np.ascontiguousarray(theta, np.float64)
        h = ctypes.c_void_p()
        call("smoq_plan_create", self.ctx.handle, tape.n, self.dtype, arr, len(tape.ops),
             None if th is None else th.ctypes.data_as(ctypes.c_void_p),
             0 if th is None else th.size, ctypes.byref(h))
        self.handle = h
        self.gates = len(tape.ops)

    @property
    def passes(self) -> int:
        n = ctypes.c_int()
        call("smoq_plan_passes", self.handle, ctypes.byref(n))
        return n.value

    def run(self, psi, repeat: int = 1):
        call("smoq_plan_run", self.handle, psi.handle, int(repeat))
        return psi

    def evolve_observe(self, psi, n_steps: int, obs_every: int, xmask, zmask) -> np.ndarray:
        """Runs n_steps applications, recording <P_t> after every obs_every-th:
        array (n_steps // obs_every, [batch,] n_terms)."""
        xm = np.ascontiguousarray(np.asarray(xmask, np.uint64))
        zm = np.ascontiguousarray(np.asarray(zmask, np.uint64))
        rec = n_steps // obs_every
        out = np.zeros((rec, psi.batch, xm.size), np.float64)
        call("smoq_evolve_observe", self.handle, psi.handle, int(n_steps), int(obs_every),
             xm.size, xm.ctypes.data_as(ctypes.c_void_p), zm.ctypes.data_as(ctypes.c_void_p),
             out.ctypes.data_as(ctypes.c_void_p))
        return out[:, 0, :] if psi.batch == 1 else out

    def free(self):
        if getattr(self, "handle", None):
            from . import _lib
            _lib.lib().smoq_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------------- workloads --
def layered_ry_rz_cz(N: int, layers: int, seed: int) -> Tape:
    """Plain circuit on |0…0>: per layer Ry(θ) on every qubit, Rz(φ) on every
    qubit, then CZ on (i, i+1) for even i, then for odd i.  Angles are
    U[0, 2π) from the SplitMix64 stream: per layer the N θ's, then the N φ's."""
    u = streams.uniforms(seed, 2 * N * layers)
    t = Tape(N)
    for l in range(layers):
        for q in range(N):
            t.add("RY", [q], TWO_PI * u[2 * N * l + q])
        for q in range(N):
            t.add("RZ", [q], TWO_PI * u[2 * N * l + N + q])
        for q in range(0, N - 1, 2):
            t.add("CZ", [q, q + 1])
        for q in range(1, N - 1, 2):
            t.add("CZ", [q, q + 1])
    return t


def random_rotation_brick(N: int, layers: int, seed: int) -> Tape:
    """Random-circuit workload: per layer a rotation on every qubit with the
    axis uniform in {X, Y, Z} and the angle uniform in [0, 2π) (draw order:
    per qubit axis then angle), then CZ on brick pairs — (0,1),(2,3),… on
    even layers, (1,2),(3,4),… on odd layers."""
    u = streams.uniforms(seed, 2 * N * layers)
    axes = ("RX", "RY", "RZ")
    t = Tape(N)
    for l in range(layers):
        for q in range(N):
            a = min(2, int(u[2 * (N * l + q)] * 3.0))
            t.add(axes[a], [q], TWO_PI * u[2 * (N * l + q) + 1])
        for q in range(l % 2, N - 1, 2):
            t.add("CZ", [q, q + 1])
    return t


def hardware_efficient_ansatz(N: int, layers: int, entangler: str = "CZ") -> Tape:
    """Parametrised ansatz: per layer Ry(θ_p), Rz(θ_{p+1}) on each qubit
    (interleaved per qubit as build_ansatz_spec,
    scripts_variational_quantum_circuits/demo_variational_quantum_eigensolver.jl:181-186),
    then a CZ ladder (q, q+1), q = 0..N-2.  2·N·layers parameters, no final
    rotation layer."""
    t = Tape(N)
    p = 0
    for l in range(layers):
        for q in range(N):
            t.add("RY", [q], param=p)
            t.add("RZ", [q], param=p + 1)
            p += 2
        for q in range(N - 1):
            t.add(entangler, [q, q + 1])
    return t


def gate_matrix(gate: str, theta: float = 0.0) -> np.ndarray:
    """2x2 matrices of the benchmark's gate set (run_benchmark_smoq_Float64.jl:86-98)."""
    g = gate.upper()
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    r = 1 / math.sqrt(2)
    return {
        "H": np.array([[r, r], [r, -r]], np.complex128),
        "X": np.array([[0, 1], [1, 0]], np.complex128),
        "Y": np.array([[0, -1j], [1j, 0]], np.complex128),
        "Z": np.array([[1, 0], [0, -1]], np.complex128),
        "S": np.array([[1, 0], [0, 1j]], np.complex128),
        "T": np.array([[1, 0], [0, complex(math.cos(math.pi / 4), math.sin(math.pi / 4))]]),
        "RX": np.array([[c, -1j * s], [-1j * s, c]], np.complex128),
        "RY": np.array([[c, -s], [s, c]], np.complex128),
        "RZ": np.array([[complex(c, -s), 0], [0, complex(c, s)]], np.complex128),
    }[g]


# ------------------------------------------------ shared benchmark circuits --
BENCH_GATES = ("H", "X", "Y", "Z", "S", "T", "RX", "RY", "RZ", "CNOT", "CZ")


def load_circuit_file(path: str) -> Tape:
    """The shared random-circuit format read by load_circuit_from_file
    (SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:384-404):
    a `# N=<n> …` header, then one gate per line, `NAME q [theta]` or
    `CNOT c t` / `CZ a b`, 0-based qubits.  Gates outside H X Y Z S T RX RY
    RZ CNOT CZ raise ValueError (the reference's map lookup fails too)."""
    n, rows = None, []
    with open(path, encoding="utf-8") as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("#"):
                for tok in line[1:].split():
                    if tok.startswith("N=") and n is None:
                        n = int(tok[2:])
                continue
            rows.append(line.split())
    if n is None:
        raise ValueError(f"{path}: no '# N=…' header")
    t = Tape(n)
    for parts in rows:
        name = parts[0].upper()
        if name not in BENCH_GATES:
            raise ValueError(f"{path}: unknown gate {parts[0]!r}")
        if name in ("CNOT", "CZ"):
            t.add(name, [int(parts[1]), int(parts[2])])
        elif name in ("RX", "RY", "RZ"):
            t.add(name, [int(parts[1])], float(parts[2]))
        else:
            t.add(name, [int(parts[1])])
    return t


def run_circuit_sampling(tape: Tape, uniforms, dtype: str = "c128", ctx=None):
    """run_circuit_sampling (run_benchmark_smoq_Float64.jl:414-437) on the
    device: |0…0>, the tape in fused passes, then one Born draw per uniform
    with searchsortedfirst(cumsum(abs2.(ψ)), r) - 1.  The uniforms replace
    the reference's rand() per shot.  Returns (0-based indices, bitstrings
    [shot][qubit]); the state is freed."""
    from .measurements import bitstrings, sample_indices
    from .state import GPUState

    psi = GPUState.basis(tape.n, 0, dtype, ctx=ctx)
    try:
        tape.apply(psi)
        idx = sample_indices(psi, uniforms, "cumsum")
    finally:
        psi.free()
    return idx, bitstrings(idx, tape.n)


# angle tables of the expectation-value benchmark (bench_smoq.jl:18-21)
_EXP_RY = (4, 3, 6, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26)
_EXP_RX = (5, 7, 4, 9, 3, 11, 6, 13, 8, 10, 12, 14, 15, 17, 16, 19, 18, 20, 21, 22, 23, 24, 25,
           26)


def expectation_benchmark_circuit(N: int, layers: int = 4) -> Tape:
    """run_circuit of the expectation-value benchmark
    (SmoQ.jl_benchmarks/benchmark_expectation_values/bench_smoq.jl:24-37), 0-based:
    per layer H on odd qubits, Ry(π/a_q) on all, CNOT(q, q+1) for even q,
    S on even qubits, Rx(π/b_q) on all, CZ(q, q+1) for odd q; a and b are the
    benchmark's angle tables (N ≤ 24)."""
    if not 1 <= N <= len(_EXP_RY):
        raise ValueError(f"the benchmark's angle tables cover N = 1..{len(_EXP_RY)}")
    t = Tape(N)
    for _ in range(layers):
        for q in range(1, N, 2):
            t.add("H", [q])
        for q in range(N):
            t.add("RY", [q], math.pi / _EXP_RY[q])
        for q in range(0, N - 1, 2):
            t.add("CNOT", [q, q + 1])
        for q in range(0, N, 2):
            t.add("S", [q])
        for q in range(N):
            t.add("RX", [q], math.pi / _EXP_RX[q])
        for q in range(1, N - 1, 2):
            t.add("CZ", [q, q + 1])
    return t
