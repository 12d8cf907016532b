"""Classical-shadow snapshot collection on the device, with the reference's
names (utils/cpu/cpuClassicalShadows.jl: ShadowConfig, PauliSnapshot,
CliffordSnapshot, sample_pauli_bases, classical_shadow_snapshot,
collect_shadows, SINGLE_QUBIT_CLIFFORDS, clifford_shadow_snapshot,
collect_clifford_shadows).

The reference draws the bases and the outcome uniform with `rand()` inside
each snapshot.  Here both are arrays (given, or drawn from `rng`), and all
snapshots go to the device in one call (smoq_shadow_snapshots): snapshots
with the same rotation string share one rotated state, and each draw is a
search of that state's sequential prefix.  Estimators built on the snapshots
(reconstruction, expectation values, purity) are host post-processing of
these integer arrays and are not part of this module.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import call
from .state import GPUState

SHADOW_PAULI = 0
SHADOW_LOCAL_CLIFFORD = 1


@dataclass
class PauliSnapshot:
    """PauliSnapshot (…ClassicalShadows.jl:416-419): bases 1 = X, 2 = Y,
    3 = Z and outcomes 0/1, one per qubit (qubit 1 first)."""
    bases: np.ndarray
    outcomes: np.ndarray


Snapshot = PauliSnapshot


@dataclass
class CliffordSnapshot:
    """CliffordSnapshot (…ClassicalShadows.jl:566-569): 1-based indices into
    SINGLE_QUBIT_CLIFFORDS and outcomes, one per qubit."""
    clifford_indices: np.ndarray
    outcomes: np.ndarray


@dataclass
class ShadowConfig:
    """ShadowConfig (…ClassicalShadows.jl:366-392) with its group aliases."""
    n_qubits: int
    n_shots: int
    measurement_group: str = "pauli"

    def __post_init__(self):
        if self.n_qubits <= 0 or self.n_shots <= 0:
            raise ValueError("n_qubits and n_shots must be positive")
        g = str(self.measurement_group).lstrip(":")
        g = {"pauli_group": "pauli", "clifford_group": "local_clifford",
             "clifford": "local_clifford"}.get(g, g)
        if g not in ("pauli", "local_clifford", "global_clifford"):
            raise ValueError("measurement_group must be :pauli, :local_clifford, or "
                             ":global_clifford")
        self.measurement_group = g


def _single_qubit_cliffords() -> np.ndarray:
    """SINGLE_QUBIT_CLIFFORDS (…ClassicalShadows.jl:457-505), built the same
    way: breadth-first products G·C with G in (H, S), global phase fixed by the
    first entry of magnitude > 0.1 in column order, entries rounded to 10
    digits, duplicates (Frobenius distance <= 1e-8) dropped."""
    H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)
    S = np.array([[1, 0], [0, 1j]], np.complex128)

    def normalize(M):
        for v in M.T.reshape(-1):  # column-major order
            if abs(v) > 0.1:
                return M / (v / abs(v))
        return M

    def rnd(M):
        return np.round(M.real, 10) + 1j * np.round(M.imag, 10)

    cliffs = [np.eye(2, dtype=np.complex128)]
    explore = [cliffs[0]]
    while True:
        new = []
        for C in explore:
            for G in (H, S):
                N = rnd(normalize(G @ C))
                if not any(np.linalg.norm(N - E) <= 1e-8 for E in cliffs):
                    new.append(N)
                    cliffs.append(N)
        if not new:
            break
        explore = new
    if len(cliffs) != 24:
        raise RuntimeError(f"Expected 24 Cliffords, got {len(cliffs)}")
    return np.array(cliffs)


SINGLE_QUBIT_CLIFFORDS = _single_qubit_cliffords()


def get_clifford_gate(idx: int) -> np.ndarray:
    """get_clifford_gate (…ClassicalShadows.jl:528-530), 1-based."""
    return SINGLE_QUBIT_CLIFFORDS[int(idx) - 1]


def sample_pauli_bases(N: int, rng=None) -> np.ndarray:
    """sample_pauli_bases (…ClassicalShadows.jl:988-990): N draws from 1:3."""
    g = rng if rng is not None else np.random.default_rng()
    return g.integers(1, 4, N).astype(np.uint8)


def sample_local_clifford_indices(N: int, rng=None) -> np.ndarray:
    """sample_local_clifford_indices (…ClassicalShadows.jl:519-521): N draws from 1:24."""
    g = rng if rng is not None else np.random.default_rng()
    return g.integers(1, 25, N).astype(np.uint8)


def _outcomes(idx: np.ndarray, N: int) -> np.ndarray:
    return ((idx[:, None] >> np.arange(N, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(
        np.uint8)


def shadow_outcomes(psi: GPUState, codes, uniforms, kind: int = SHADOW_PAULI) -> np.ndarray:
    """smoq_shadow_snapshots: the 0-based outcome index of every snapshot.
    codes: (M, N) array (Pauli 1..3, or Clifford 0-based 0..23)."""
    c = np.ascontiguousarray(np.asarray(codes, np.uint8).reshape(-1, psi.n))
    u = np.ascontiguousarray(np.asarray(uniforms, np.float64).reshape(-1))
    if u.size != c.shape[0]:
        raise ValueError(f"{c.shape[0]} snapshots but {u.size} uniforms")
    out = np.empty(u.size, np.uint64)
    tab = None
    if kind == SHADOW_LOCAL_CLIFFORD:
        tab = np.ascontiguousarray(SINGLE_QUBIT_CLIFFORDS.view(np.float64).reshape(-1))
    call("smoq_shadow_snapshots", psi.handle, u.size, kind, c.ctypes.data_as(ctypes.c_void_p),
         None if tab is None else tab.ctypes.data_as(ctypes.c_void_p),
         u.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return out


class ShadowSet:
    """M snapshots as two (M, N) uint8 arrays; indexing gives PauliSnapshot /
    CliffordSnapshot objects like the reference's Vector of snapshots."""

    def __init__(self, codes: np.ndarray, outcomes: np.ndarray, kind: str):
        self.codes, self.outcomes, self.kind = codes, outcomes, kind

    def __len__(self):
        return self.codes.shape[0]

    def __getitem__(self, m):
        if self.kind == "pauli":
            return PauliSnapshot(self.codes[m], self.outcomes[m])
        return CliffordSnapshot(self.codes[m], self.outcomes[m])


def collect_shadows(state, N, M: int | None = None, *, bases=None, uniforms=None,
                    rng=None) -> ShadowSet:
    """collect_shadows(ψ, N, M) (…ClassicalShadows.jl:1098-1106) and
    collect_shadows(ψ, config) (:1135-1160, :pauli and :local_clifford):
    M snapshots of a state holding one vector (a batched state is rejected).
    `bases` (M, N) holds 1..3 for Pauli or
    1..24 for local Clifford; drawn from `rng` when absent, as are the
    outcome uniforms."""
    group = "pauli"
    if isinstance(N, ShadowConfig):
        cfg = N
        N, M, group = cfg.n_qubits, cfg.n_shots, cfg.measurement_group
        if group == "global_clifford":
            return collect_global_clifford_shadows(state, N, M, uniforms=uniforms, rng=rng)
    from .density import DensityMatrix

    if not isinstance(state, (GPUState, DensityMatrix)):
        raise TypeError("snapshots are taken of a GPUState or a DensityMatrix")
    if (state.N if isinstance(state, DensityMatrix) else state.n) != N:
        raise ValueError(f"state has the wrong number of qubits for N = {N}")
    g = rng if rng is not None else np.random.default_rng()
    hi = 4 if group == "pauli" else 25
    b = (g.integers(1, hi, (M, N)) if bases is None else np.asarray(bases)).astype(np.uint8)
    b = b.reshape(M, N)
    u = g.random(M) if uniforms is None else np.asarray(uniforms, np.float64)
    if isinstance(state, DensityMatrix):
        return ShadowSet(b, _outcomes(_rho_outcomes(state, b, u, group), N), group)
    if group == "pauli":
        idx = shadow_outcomes(state, b, u, SHADOW_PAULI)
    else:
        idx = shadow_outcomes(state, b - 1, u, SHADOW_LOCAL_CLIFFORD)
    return ShadowSet(b, _outcomes(idx, N), group)


def _rho_outcomes(rho, codes: np.ndarray, u: np.ndarray, group: str) -> np.ndarray:
    """Snapshots of a density matrix (classical_shadow_snapshot(ρ, N),
    …ClassicalShadows.jl:1039-1070; clifford_shadow_snapshot(ρ, N), :608-632):
    for each distinct rotation string, U ρ U† on a copy of vec(ρ) (V on row
    qubit q, conj(V) on column qubit N + q, through the fused passes), then
    every draw of that string from the populations real(ρ_ii) with
    sample_from_probabilities' rule.  The reference clamps the populations at
    0 and divides by their sum first; both change them only at rounding
    level, so a draw can differ only within rounding of a cumulative sum."""
    from .circuits import Tape
    from .density import dm_tape
    from .measurements import sample_indices

    if rho.vec.batch != 1:
        raise ValueError("snapshots are taken of one density matrix (batch 1)")
    N = rho.N
    keys, inv = np.unique(codes, axis=0, return_inverse=True)
    out = np.empty(codes.shape[0], np.uint64)
    work = rho.vec.clone()
    try:
        for k, key in enumerate(keys):
            t = Tape(N)
            for q in range(N):
                c = int(key[q])
                if group == "pauli":
                    if c == 1:
                        t.add("H", [q])
                    elif c == 2:
                        t.add("SDG", [q]).add("H", [q])
                else:
                    t.add("U1", [q], m=SINGLE_QUBIT_CLIFFORDS[c - 1])
            work.copy_from(rho.vec)
            if len(t):
                dm_tape(t).apply(work)
            sel = np.nonzero(inv.reshape(-1) == k)[0]
            out[sel] = sample_indices(work, u[sel], "sequential_diag")
    finally:
        work.free()
    return out


@dataclass
class GlobalCliffordCircuit:
    """GlobalCliffordCircuit (…ClassicalShadows.jl:771-779): gates as
    ("H", q), ("S", q), ("CNOT", c, t), 1-based qubits."""
    n_qubits: int
    gates: list


@dataclass
class GlobalCliffordSnapshot:
    """GlobalCliffordSnapshot (…ClassicalShadows.jl:786-789)."""
    circuit: GlobalCliffordCircuit
    outcomes: np.ndarray


def sample_random_global_clifford(N: int, depth: int | None = None,
                                  rng=None) -> GlobalCliffordCircuit:
    """sample_random_global_clifford (…ClassicalShadows.jl:802-826): per layer
    and qubit one of H, S, H then S, or nothing (1:4 uniform); between layers
    CNOTs on consecutive pairs of a shuffled qubit order."""
    g = rng if rng is not None else np.random.default_rng()
    depth = 2 * N if depth is None else int(depth)
    gates = []
    for layer in range(1, depth + 1):
        for q in range(1, N + 1):
            r = int(g.integers(1, 5))
            if r == 1:
                gates.append(("H", q))
            elif r == 2:
                gates.append(("S", q))
            elif r == 3:
                gates += [("H", q), ("S", q)]
        if layer < depth and N >= 2:
            order = g.permutation(np.arange(1, N + 1))
            for i in range(0, N - 1, 2):
                gates.append(("CNOT", int(order[i]), int(order[i + 1])))
    return GlobalCliffordCircuit(N, gates)


def _gc_bits(circuit: GlobalCliffordCircuit):
    """Device triples of a circuit.  The reference's unitary places H and S
    on qubit q as the q-th Kronecker factor from the left (bit N - q) and
    CNOT on bits c - 1, t - 1 (…ClassicalShadows.jl:835-870); kept."""
    N = circuit.n_qubits
    out = []
    for gt in circuit.gates:
        if gt[0] == "H":
            out.append((0, N - gt[1], 0))
        elif gt[0] == "S":
            out.append((1, N - gt[1], 0))
        else:
            out.append((2, gt[1] - 1, gt[2] - 1))
    return out


def build_global_clifford_unitary(circuit: GlobalCliffordCircuit) -> np.ndarray:
    """build_global_clifford_unitary (…ClassicalShadows.jl:835-870), dense
    (small N): U = G_L ... G_1 with H/S as kron(G_1, ..., G_N) and CNOT on
    bits c - 1 -> t - 1."""
    N = circuit.n_qubits
    dim = 1 << N
    H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)
    S = np.array([[1, 0], [0, 1j]], np.complex128)
    I2 = np.eye(2, dtype=np.complex128)
    U = np.eye(dim, dtype=np.complex128)
    for gt in circuit.gates:
        if gt[0] in ("H", "S"):
            P = H if gt[0] == "H" else S
            G = P if gt[1] == 1 else I2
            for j in range(2, N + 1):
                G = np.kron(G, P if j == gt[1] else I2)
        else:
            c, t = gt[1], gt[2]
            G = np.zeros((dim, dim), np.complex128)
            k = np.arange(dim)
            dst = np.where((k >> (c - 1)) & 1, k ^ (1 << (t - 1)), k)
            G[dst, k] = 1
        U = G @ U
    return U


def _random_gc_triples(N: int, M: int, depth: int, g):
    """M circuits of sample_random_global_clifford's structure drawn at once,
    as the device's flat triples (type, bit a, bit b) and per-circuit offsets."""
    R = g.integers(1, 5, (M, depth, N))
    q = np.arange(1, N + 1)
    # per (circuit, layer): H and S slots of qubits 1..N in order, then CNOTs
    qt = np.broadcast_to(np.array([0, 1], np.uint8), (M, depth, N, 2))
    qa = np.broadcast_to((N - q).astype(np.uint8)[:, None], (M, depth, N, 2))
    qv = np.stack([(R == 1) | (R == 3), (R == 2) | (R == 3)], axis=-1)
    P = N // 2 if N >= 2 else 0
    perm = np.argsort(g.random((M, depth, N)), axis=-1).astype(np.uint8)  # 0-based qubits
    ct = np.full((M, depth, P), 2, np.uint8)
    ca, cb = perm[..., 0:2 * P:2], perm[..., 1:2 * P:2]
    cv = np.ones((M, depth, P), bool)
    cv[:, depth - 1, :] = False  # no CNOT layer after the last
    T = np.concatenate([qt.reshape(M, depth, 2 * N), ct], axis=2).reshape(M, -1)
    A = np.concatenate([qa.reshape(M, depth, 2 * N), ca], axis=2).reshape(M, -1)
    B = np.concatenate([np.zeros((M, depth, 2 * N), np.uint8), cb], axis=2).reshape(M, -1)
    V = np.concatenate([qv.reshape(M, depth, 2 * N), cv], axis=2).reshape(M, -1)
    off = np.zeros(M + 1, np.int64)
    off[1:] = np.cumsum(V.sum(axis=1))
    flat = np.stack([T[V], A[V], B[V]], axis=1).astype(np.uint8)
    return flat, off


def _circuit_of(N: int, trip: np.ndarray) -> GlobalCliffordCircuit:
    gates = []
    for typ, a, b in trip:
        if typ == 0:
            gates.append(("H", N - int(a)))
        elif typ == 1:
            gates.append(("S", N - int(a)))
        else:
            gates.append(("CNOT", int(a) + 1, int(b) + 1))
    return GlobalCliffordCircuit(N, gates)


class GlobalCliffordShadowSet:
    """M global-Clifford snapshots: the circuits as flat device triples with
    offsets, the outcomes as an (M, N) array; indexing gives a
    GlobalCliffordSnapshot."""

    def __init__(self, N, triples, offsets, outcomes):
        self.N, self.triples, self.offsets, self.outcomes = N, triples, offsets, outcomes

    def __len__(self):
        return self.outcomes.shape[0]

    def circuit(self, m) -> GlobalCliffordCircuit:
        return _circuit_of(self.N, self.triples[self.offsets[m]:self.offsets[m + 1]])

    def __getitem__(self, m):
        return GlobalCliffordSnapshot(self.circuit(m), self.outcomes[m])

    def __iter__(self):
        return (self[m] for m in range(len(self)))


def collect_global_clifford_shadows(state, N: int, M: int, *, circuits=None, uniforms=None,
                                    rng=None, depth: int | None = None):
    """collect_global_clifford_shadows(ψ, N, M) (…ClassicalShadows.jl:926-933)
    and (ρ, N, M) (:939-946): one random Clifford circuit per snapshot
    (given, or drawn from `rng` with sample_random_global_clifford's
    structure).  ψ: every snapshot runs in one workgroup on the device
    (smoq_global_clifford_snapshots, N <= 12); ρ: U ρ U† through the fused
    passes per snapshot, the draw from the populations."""
    from .density import DensityMatrix

    g = rng if rng is not None else np.random.default_rng()
    if circuits is not None:
        circs = list(circuits)
        trips = [np.array(_gc_bits(c) or np.zeros((0, 3)), np.uint8).reshape(-1, 3) for c in circs]
        off = np.zeros(len(circs) + 1, np.int64)
        off[1:] = np.cumsum([len(t) for t in trips])
        flat = np.concatenate(trips) if trips else np.zeros((0, 3), np.uint8)
    else:
        flat, off = _random_gc_triples(N, M, 2 * N if depth is None else int(depth), g)
    u = g.random(M) if uniforms is None else np.asarray(uniforms, np.float64).reshape(-1)
    if len(off) != M + 1 or u.size != M:
        raise ValueError(f"need {M} circuits and {M} uniforms")
    if isinstance(state, DensityMatrix):
        from .circuits import Tape
        from .density import dm_tape
        from .measurements import sample_indices

        if state.N != N or state.vec.batch != 1:
            raise ValueError("one density matrix of N qubits")
        idx = np.empty(M, np.uint64)
        work = state.vec.clone()
        try:
            for m in range(M):
                t = Tape(N)
                for typ, a, b in flat[off[m]:off[m + 1]]:
                    if typ == 0:
                        t.add("H", [int(a)])
                    elif typ == 1:
                        t.add("S", [int(a)])
                    else:
                        t.add("CNOT", [int(a), int(b)])
                work.copy_from(state.vec)
                if len(t):
                    dm_tape(t).apply(work)
                idx[m] = sample_indices(work, [u[m]], "sequential_diag")[0]
        finally:
            work.free()
    else:
        if state.n != N:
            raise ValueError(f"state has {state.n} qubits, N = {N}")
        fl = np.ascontiguousarray(flat if flat.size else np.zeros((1, 3), np.uint8))
        u = np.ascontiguousarray(u)
        idx = np.empty(M, np.uint64)
        call("smoq_global_clifford_snapshots", state.handle, M,
             fl.ctypes.data_as(ctypes.c_void_p), off.ctypes.data_as(ctypes.c_void_p),
             u.ctypes.data_as(ctypes.c_void_p), idx.ctypes.data_as(ctypes.c_void_p))
    return GlobalCliffordShadowSet(N, flat, off, _outcomes(idx, N))


def global_clifford_shadow_snapshot(state, N: int, circuit=None, r=None, *, rng=None):
    """global_clifford_shadow_snapshot(ψ or ρ, N) (…ClassicalShadows.jl:884-921)."""
    return collect_global_clifford_shadows(
        state, N, 1, circuits=None if circuit is None else [circuit],
        uniforms=None if r is None else [float(r)], rng=rng)[0]


def classical_shadow_snapshot(state, N: int, bases=None, r=None, *, rng=None) -> PauliSnapshot:
    """classical_shadow_snapshot(ψ, N) (…ClassicalShadows.jl:1001-1037)."""
    b = None if bases is None else np.asarray(bases).reshape(1, N)
    u = None if r is None else [float(r)]
    return collect_shadows(state, N, 1, bases=b, uniforms=u, rng=rng)[0]


def collect_clifford_shadows(state, N: int, M: int, *, indices=None, uniforms=None,
                             rng=None) -> ShadowSet:
    """collect_clifford_shadows(ψ, N, M) (…ClassicalShadows.jl:636-641)."""
    return collect_shadows(state, ShadowConfig(N, M, "local_clifford"), bases=indices,
                           uniforms=uniforms, rng=rng)


def clifford_shadow_snapshot(state, N: int, indices=None, r=None, *, rng=None):
    """clifford_shadow_snapshot(ψ, N) (…ClassicalShadows.jl:580-603)."""
    i = None if indices is None else np.asarray(indices).reshape(1, N)
    u = None if r is None else [float(r)]
    return collect_clifford_shadows(state, N, 1, indices=i, uniforms=u, rng=rng)[0]


__all__ = [
    "GlobalCliffordCircuit", "GlobalCliffordShadowSet", "GlobalCliffordSnapshot",
    "build_global_clifford_unitary",
    "collect_global_clifford_shadows", "global_clifford_shadow_snapshot",
    "sample_random_global_clifford",
    "CliffordSnapshot", "PauliSnapshot", "SINGLE_QUBIT_CLIFFORDS", "ShadowConfig", "ShadowSet",
    "Snapshot", "classical_shadow_snapshot", "clifford_shadow_snapshot",
    "collect_clifford_shadows", "collect_shadows", "get_clifford_gate",
    "sample_local_clifford_indices", "sample_pauli_bases", "shadow_outcomes",
]
