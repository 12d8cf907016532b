"""Density matrices on the device, with the reference's names
(utils/cpu/cpuQuantumChannelGates.jl `apply_*_rho!`,
utils/cpu/cpuQuantumChannelLindbladianEvolutionDensityMatrix.jl,
utils/cpu/cpuQuantumChannelLindbladianEvolution.jl `_apply_dissipator_dm!`,
the `_expect_*_dm` read-outs of utils/cpu/cpuQuantumStateObservables.jl).

ρ of N qubits is stored as vec(ρ): a 2N-qubit state with element (i, j) at
index i + 2^N j — the column-major memory order of the reference's
Matrix{ComplexF64}.  A gate U on qubit q acts as U on qubit q and conj(U) on
qubit N + q (ρ → U ρ U†); a Lindblad dissipator step on site k is a 4×4 map
on the pair (k, N + k).  Both are ordinary gates of the 2N-qubit tape, so
they run on the fused passes (and compiled plans) of the state-vector path;
Tr(ρ P) and the dense U ρ U† have their own entry points
(smoq_dm_expect_pauli, smoq_dm_apply_dense).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import call
from .circuits import Tape
from .state import KETS, Context, GPUState

# conj(U) of the named gates: real ones are themselves; the phase gates swap
# with their adjoints; the rotations change the sign of their angle
_CONJ_NAME = {"S": "SDG", "SDG": "S", "T": "TDG", "TDG": "T"}
_REAL = {"H", "X", "Z", "RY", "CNOT", "CZ", "CRY", "CCZ", "TOFFOLI", "SIGMA_MINUS", "SIGMA_PLUS"}
_NEG_ANGLE = {"RX", "RZ", "RXX", "RYY", "RZZ"}
_Y = np.array([[0, -1j], [1j, 0]], np.complex128)


class DensityMatrix:
    """ρ (batch of them) on the device as vec(ρ), a 2N-qubit GPUState."""

    def __init__(self, N: int, dtype="c128", batch: int = 1, ctx: Context | None = None,
                 vec: GPUState | None = None):
        self.N = int(N)
        self.vec = vec if vec is not None else GPUState(2 * self.N, dtype, batch, ctx=ctx)

    # ---- construction ------------------------------------------------------
    @classmethod
    def from_numpy(cls, rho, dtype=None, ctx: Context | None = None) -> "DensityMatrix":
        r = np.asarray(rho)
        mats = r[None] if r.ndim == 2 else r
        dim = mats.shape[-1]
        N = dim.bit_length() - 1
        if mats.shape[-2:] != (dim, dim) or (1 << N) != dim:
            raise ValueError(f"rho must be 2^N x 2^N, got {r.shape}")
        flat = np.stack([m.reshape(-1, order="F") for m in mats])
        dt = dtype or ("c64" if r.dtype == np.complex64 else "c128")
        return cls(N, vec=GPUState.from_numpy(flat if r.ndim == 3 else flat[0],
                                              dtype=dt, ctx=ctx))

    @classmethod
    def basis(cls, N: int, index: int = 0, dtype="c128", batch: int = 1,
              ctx: Context | None = None) -> "DensityMatrix":
        """|index><index| (make_initial_psi(N, bits) * its adjoint)."""
        return cls(N, vec=GPUState.basis(2 * N, index | (index << N), dtype, batch, ctx=ctx))

    @classmethod
    def product(cls, states, dtype="c128", batch: int = 1,
                ctx: Context | None = None) -> "DensityMatrix":
        """make_product_rho (utils/cpu/cpuQuantumStatePreparation.jl:1199-1202):
        |ψ><ψ| of the product ket, as the product of |s_k> on row qubit k and
        conj(s_k) on column qubit N + k."""
        names = [str(s).lstrip(":") for s in states]
        kets = [KETS[s] for s in names]
        N = len(kets)
        rho = cls(N, dtype, batch, ctx=ctx)
        rho.vec.set_product(list(kets) + [tuple(np.conj(k)) for k in kets])
        return rho

    def to_numpy(self) -> np.ndarray:
        v = self.vec.to_numpy()
        dim = 1 << self.N
        if self.vec.batch == 1:
            return v.reshape(dim, dim, order="F")
        return np.stack([x.reshape(dim, dim, order="F") for x in v])

    def free(self) -> None:
        self.vec.free()

    # ---- read-outs -----------------------------------------------------------
    def expect_pauli(self, xmask, zmask) -> np.ndarray:
        """Re Tr(ρ P_t) for Pauli strings (x, z masks over the N qubits);
        shape (n_terms,) or (batch, n_terms)."""
        xm = np.ascontiguousarray(np.atleast_1d(np.asarray(xmask, np.uint64)))
        zm = np.ascontiguousarray(np.atleast_1d(np.asarray(zmask, np.uint64)))
        if xm.shape != zm.shape:
            raise ValueError("xmask and zmask must have the same length")
        out = np.empty((self.vec.batch, xm.size), np.float64)
        call("smoq_dm_expect_pauli", self.vec.handle, xm.size, xm.ctypes.data_as(ctypes.c_void_p),
             zm.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
        return out[0] if self.vec.batch == 1 else out

    def trace(self):
        v = self.expect_pauli([0], [0])
        return float(v[0]) if self.vec.batch == 1 else v[:, 0]

    def purity(self):
        """Tr(ρ²) = Σ |ρ_ij|² for Hermitian ρ."""
        v = self.vec.norm2()
        return float(v[0]) if self.vec.batch == 1 else v


# ---- gates -------------------------------------------------------------------
def dm_tape(tape: Tape) -> Tape:
    """The 2N-qubit tape of ρ → U ρ U† for every gate of an N-qubit tape:
    each gate on its qubits, then conj(gate) on the column copies (qubit
    N + q).  A parametrised gate's column copy takes parameter p + n_params
    (pass the angles as concatenate([θ, -θ]) for RX/RZ/RXX/RYY/RZZ,
    [θ, θ] for the real ones): see dm_theta."""
    N = tape.n
    out = Tape(2 * N)
    P = tape.n_params
    for gate, qs, th, par, m in tape.ops:
        out.add(gate, qs, th, par, m)
        cq = [q + N for q in qs]
        cpar = par + P if par >= 0 else -1
        if gate in _REAL:
            out.add(gate, cq, th, cpar, m)
        elif gate in _CONJ_NAME:
            out.add(_CONJ_NAME[gate], cq)
        elif gate in _NEG_ANGLE:
            out.add(gate, cq, -th, cpar)
        elif gate == "Y":
            out.add("U1", cq, m=np.conj(_Y))
        elif gate in ("U1", "U2", "DIAG1", "DIAG2"):
            out.add(gate, cq, m=np.conj(np.asarray(m, np.complex128)))
        else:
            raise ValueError(f"no conjugate rule for {gate}")
    if P:
        out.n_params = 2 * P
    return out


def dm_theta(tape: Tape, theta) -> np.ndarray:
    """Angles for dm_tape(tape): θ for the row gates, then the column copies'
    angles (−θ where the conjugate gate negates its angle)."""
    th = np.asarray(theta, np.float64)
    col = th.copy()
    for gate, _, _, par, _ in tape.ops:
        if par >= 0 and gate in _NEG_ANGLE:
            col[par] = -th[par]
    return np.concatenate([th, col])


def apply_tape_rho(rho: DensityMatrix, tape: Tape, theta=None) -> DensityMatrix:
    """ρ ← U ρ U† for the tape's circuit U (fused passes on vec(ρ))."""
    if tape.n != rho.N:
        raise ValueError(f"tape is for {tape.n} qubits, rho has {rho.N}")
    t = dm_tape(tape)
    t.apply(rho.vec, None if theta is None else dm_theta(tape, theta))
    return rho


def _one(rho, gate, qubits, theta=0.0):
    t = Tape(rho.N)
    t.add(gate, [q - 1 for q in qubits], theta)
    return apply_tape_rho(rho, t)


def apply_rx_rho(rho, k, theta, N):
    """apply_rx_rho! (cpuQuantumChannelGates.jl:503-539)."""
    return _one(rho, "RX", [k], theta)


def apply_ry_rho(rho, k, theta, N):
    """apply_ry_rho! (cpuQuantumChannelGates.jl:546-582)."""
    return _one(rho, "RY", [k], theta)


def apply_rz_rho(rho, k, theta, N):
    """apply_rz_rho! (cpuQuantumChannelGates.jl:589-606)."""
    return _one(rho, "RZ", [k], theta)


def apply_hadamard_rho(rho, k, N):
    """apply_hadamard_rho! (cpuQuantumChannelGates.jl:613-645)."""
    return _one(rho, "H", [k])


def apply_pauli_x_rho(rho, k, N):
    """apply_pauli_x_rho! (cpuQuantumChannelGates.jl:652-677)."""
    return _one(rho, "X", [k])


def apply_pauli_z_rho(rho, k, N):
    """apply_pauli_z_rho! (cpuQuantumChannelGates.jl:684-698)."""
    return _one(rho, "Z", [k])


def apply_t_rho(rho, k, N):
    """apply_t_rho! (cpuQuantumChannelGates.jl:397-410)."""
    return _one(rho, "T", [k])


def apply_tdagger_rho(rho, k, N):
    """apply_tdagger_rho! (cpuQuantumChannelGates.jl:417-430)."""
    return _one(rho, "TDG", [k])


def apply_cnot_rho(rho, control, target, N):
    """apply_cnot_rho! (cpuQuantumChannelGates.jl:1111-1137)."""
    return _one(rho, "CNOT", [control, target])


def apply_cz_rho(rho, qi, qj, N):
    """apply_cz_rho! (cpuQuantumChannelGates.jl:1144-1156)."""
    return _one(rho, "CZ", [qi, qj])


def apply_cry_rho(rho, control, target, theta, N):
    """apply_cry_rho! (cpuQuantumChannelGates.jl:945-979)."""
    return _one(rho, "CRY", [control, target], theta)


def apply_rxx_rho(rho, qi, qj, theta, N):
    """apply_rxx_rho! (cpuQuantumChannelGates.jl:1163-1228)."""
    return _one(rho, "RXX", [qi, qj], theta)


def apply_ryy_rho(rho, qi, qj, theta, N):
    """apply_ryy_rho! (cpuQuantumChannelGates.jl:1235-1299)."""
    return _one(rho, "RYY", [qi, qj], theta)


def apply_rzz_rho(rho, qi, qj, theta, N):
    """apply_rzz_rho! (cpuQuantumChannelGates.jl:1306-1326)."""
    return _one(rho, "RZZ", [qi, qj], theta)


# ---- dissipators ---------------------------------------------------------------
def dissipator_map(kind: str, gdt: float) -> tuple[str, np.ndarray]:
    """The first-order dissipator step of _apply_dissipator_dm!
    (utils/cpu/cpuQuantumChannelLindbladianEvolution.jl:377-454) on one site,
    as a gate on the pair (row bit, column bit), basis index
    b(row) + 2 b(col): ("DIAG2", diagonal) for sigma_z, ("U2", 4x4) else."""
    f = float(gdt)
    if kind == "sigma_z":
        return "DIAG2", np.array([1.0, 1 - 2 * f, 1 - 2 * f, 1.0], np.complex128)
    M = np.zeros((4, 4), np.complex128)
    if kind == "sigma_minus":  # |11> feeds |00>, decays; coherences decay by half
        M[0, 0], M[0, 3] = 1.0, f
        M[1, 1] = M[2, 2] = 1 - 0.5 * f
        M[3, 3] = 1 - f
    elif kind == "sigma_plus":
        M[3, 3], M[3, 0] = 1.0, f
        M[1, 1] = M[2, 2] = 1 - 0.5 * f
        M[0, 0] = 1 - f
    elif kind == "sigma_x":  # ρ_ij += f (ρ_{i^m, j^m} - ρ_ij)
        M[0, 0] = M[3, 3] = M[1, 1] = M[2, 2] = 1 - f
        M[0, 3] = M[3, 0] = M[1, 2] = M[2, 1] = f
    else:
        raise ValueError(f"unknown jump operator {kind!r}: sigma_minus, sigma_plus, sigma_z, "
                         "sigma_x")
    return "U2", M


def add_dissipator(tape: Tape, N: int, kind: str, k: int, gdt: float) -> Tape:
    """Appends the dissipator step on site k (1-based) to a 2N-qubit tape."""
    gate, m = dissipator_map(kind, gdt)
    tape.add(gate, [k - 1, k - 1 + N], m=m)
    return tape


def apply_dissipator(rho: DensityMatrix, kind: str, k: int, gdt: float) -> DensityMatrix:
    """_apply_dissipator_dm! for one jump operator on site k (1-based)."""
    t = Tape(2 * rho.N)
    add_dissipator(t, rho.N, kind, k, gdt)
    t.apply(rho.vec)
    return rho


def apply_sigma_z_dissipator(rho, k, gdt, N):
    """apply_sigma_z_dissipator! (…DensityMatrix.jl:53-69)."""
    return apply_dissipator(rho, "sigma_z", k, gdt)


def apply_sigma_minus_dissipator_rho(rho, k, gdt, N):
    """apply_sigma_minus_dissipator_rho! (cpuQuantumChannelGates.jl:833-854)."""
    return apply_dissipator(rho, "sigma_minus", k, gdt)


# ---- read-outs ---------------------------------------------------------------
def expect_local_dm(rho: DensityMatrix, k: int, N: int, pauli) -> float:
    """expect_local on a density matrix: _expect_X/Y/Z_dm (…Observables.jl:1271-1340)."""
    p = str(pauli).lstrip(":").upper()
    b = 1 << (k - 1)
    x, z = {"X": (b, 0), "Y": (b, b), "Z": (0, b)}[p]
    return float(rho.expect_pauli([x], [z])[0])


def expect_corr_dm(rho: DensityMatrix, i: int, j: int, N: int, pair) -> float:
    """expect_corr on a density matrix: _expect_ZZ/XX/YY_dm (…Observables.jl:1357-1415)."""
    p = str(pair).lstrip(":").upper()
    b = (1 << (i - 1)) | (1 << (j - 1))
    x, z = {"XX": (b, 0), "YY": (b, b), "ZZ": (0, b)}[p]
    return float(rho.expect_pauli([x], [z])[0])


def population_excited_rho(rho: DensityMatrix, k: int, N: int) -> float:
    """population_excited_rho (…DensityMatrix.jl:76-86): Σ_{bit k = 1} ρ_ii."""
    b = 1 << (k - 1)
    v = rho.expect_pauli([0, 0], [0, b])
    return float(0.5 * (v[0] - v[1]))


def measure_all_observables_dm(rho: DensityMatrix, L: int, n_rails: int) -> np.ndarray:
    """measure_all_observables_state_vector_cpu(ρ::Matrix, L, n_rails)
    (…Observables.jl:1651-1686): the same list and order as the pure-state
    form."""
    from .observables import qrc_observable_masks

    if rho.N != L * n_rails:
        raise ValueError(f"rho has {rho.N} qubits, L·n_rails = {L * n_rails}")
    xs, zs = qrc_observable_masks(L, n_rails)
    return rho.expect_pauli(xs, zs)


# ---- Lindblad steps ------------------------------------------------------------
def dm_step_tape(gates, N: int, gamma: float, dt: float, jump: str = "sigma_z") -> Tape:
    """The 2N-qubit tape of lindblad_dm_step_trotter!
    (…DensityMatrix.jl:129-243): the Trotter gate list (FastTrotterGate list
    or N-qubit Tape) on rows and columns, then the dissipator on every site
    with γ·dt (the reference's step uses sigma_z)."""
    from .trotter import trotter_tape

    t = gates if isinstance(gates, Tape) else trotter_tape(gates, N)
    out = dm_tape(t)
    for k in range(1, N + 1):
        add_dissipator(out, N, jump, k, gamma * dt)
    return out


def lindblad_dm_step_trotter(rho: DensityMatrix, gates, N: int, gamma: float, dt: float,
                             jump: str = "sigma_z") -> DensityMatrix:
    """lindblad_dm_step_trotter! (…DensityMatrix.jl:129-243): one step."""
    dm_step_tape(gates, N, gamma, dt, jump).apply(rho.vec)
    return rho


def dense_operator(U, dtype="c128", ctx: Context | None = None) -> GPUState:
    """A dense 2^N x 2^N operator on the device (column-major), for
    lindblad_dm_step."""
    u = np.asarray(U)
    return GPUState.from_numpy(u.reshape(-1, order="F"),
                               dtype=dtype, ctx=ctx)


def lindblad_dm_step(rho: DensityMatrix, U: GPUState, N: int, gamma: float, dt: float,
                     jump: str = "sigma_z") -> DensityMatrix:
    """lindblad_dm_step! (…DensityMatrix.jl:105-115): ρ ← U ρ U† with the dense
    propagator (rocBLAS), then the dissipator on every site."""
    call("smoq_dm_apply_dense", rho.vec.handle, U.handle)
    t = Tape(2 * N)
    for k in range(1, N + 1):
        add_dissipator(t, N, jump, k, gamma * dt)
    t.apply(rho.vec)
    return rho


def evolve_lindblad_dm(rho: DensityMatrix, gates, N: int, gamma: float, dt: float, n_steps: int,
                       obs_every: int = 1, xmask=(), zmask=(), jump: str = "sigma_z",
                       U: GPUState | None = None) -> np.ndarray:
    """n_steps Lindblad steps — the Trotter form (compiled once into a plan)
    or, with U given, the dense form — recording Re Tr(ρ P_t) before the
    first step and after every obs_every-th: array (records, [batch,] terms),
    records = n_steps // obs_every + 1 (the demo records step 0 too)."""
    plan = None
    if U is None:
        plan = dm_step_tape(gates, N, gamma, dt, jump).compile(
            "c64" if rho.vec.np_dtype == np.complex64 else "c128", ctx=rho.vec.ctx)
    recs = [rho.expect_pauli(xmask, zmask)] if len(xmask) else []
    diss = None
    if U is not None:
        diss = Tape(2 * N)
        for k in range(1, N + 1):
            add_dissipator(diss, N, jump, k, gamma * dt)
    for s in range(n_steps):
        if plan is not None:
            plan.run(rho.vec)
        else:
            call("smoq_dm_apply_dense", rho.vec.handle, U.handle)
            diss.apply(rho.vec)
        if len(xmask) and obs_every > 0 and (s + 1) % obs_every == 0:
            recs.append(rho.expect_pauli(xmask, zmask))
    if plan is not None:
        plan.free()
    return np.array(recs)


__all__ = [
    "DensityMatrix", "add_dissipator", "apply_cnot_rho", "apply_cry_rho", "apply_cz_rho",
    "apply_dissipator", "apply_hadamard_rho", "apply_pauli_x_rho", "apply_pauli_z_rho",
    "apply_rx_rho", "apply_rxx_rho", "apply_ry_rho", "apply_ryy_rho", "apply_rz_rho",
    "apply_rzz_rho", "apply_sigma_minus_dissipator_rho", "apply_sigma_z_dissipator",
    "apply_t_rho", "apply_tape_rho", "apply_tdagger_rho", "dense_operator", "dissipator_map",
    "dm_step_tape", "dm_tape", "dm_theta", "evolve_lindblad_dm", "expect_corr_dm",
    "expect_local_dm", "lindblad_dm_step", "lindblad_dm_step_trotter",
    "measure_all_observables_dm", "population_excited_rho",
]
