"""Hamiltonian parameters with the reference's structs and lattice builder
(utils/cpu/cpuHamiltonianBuilder.jl), plus their Pauli-sum form for the
device kernels.

H = Σ_bonds (Jxx XX + Jyy YY + Jzz ZZ) + Σ_sites (hx X + hy Y + hz Z), Pauli
operators with explicit positive signs; qubits 1-based in the structs.
"""
from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class CouplingTerm:
    """cpuHamiltonianBuilder.jl:42-49"""
    i: int
    j: int
    Jxx: float
    Jyy: float
    Jzz: float


@dataclass
class FieldTerm:
    """cpuHamiltonianBuilder.jl:51-56"""
    idx: int
    hx: float
    hy: float
    hz: float


@dataclass
class HamiltonianParams:
    """cpuHamiltonianBuilder.jl:69-74"""
    N: int
    x_bonds: list = field(default_factory=list)
    y_bonds: list = field(default_factory=list)
    fields: list = field(default_factory=list)


def get_hamiltonian_coupling(ham_type: str, J: float = 1.0, h: float = 1.0):
    """get_hamiltonian_coupling (cpuHamiltonianBuilder.jl:112-133)."""
    t = str(ham_type).lstrip(":")
    if t in ("heisenberg", "heisenberg_X"):
        return (J, J, J), (J, J, J), (h, 0.0, 0.0)
    if t == "TFIM_ZZ_X":
        return (0.0, 0.0, J), (0.0, 0.0, J), (h, 0.0, 0.0)
    if t in ("heisenberg_no_field", "XXX_0"):
        return (J, J, J), (J, J, J), (0.0, 0.0, 0.0)
    if t == "XX_X":
        return (J, J, 0.0), (J, J, 0.0), (h, 0.0, 0.0)
    raise ValueError(f"Unknown Hamiltonian type: {ham_type}. Supported: :heisenberg_X, "
                     ":TFIM_ZZ_X, :XXX_0, :XX_X")


def build_hamiltonian_parameters(Nx: int, Ny: int, ham_type=None, *, J=1.0, h=1.0,
                                 J_x_direction=None, J_y_direction=None, h_field=None):
    """build_hamiltonian_parameters (cpuHamiltonianBuilder.jl:161-193): x bonds
    within rows, y bonds between adjacent rows, the same field on every site."""
    if ham_type is not None:
        J_x_direction, J_y_direction, h_field = get_hamiltonian_coupling(ham_type, J, h)
    N = Nx * Ny
    x_bonds = [CouplingTerm((r - 1) * Nx + i, (r - 1) * Nx + i + 1, *J_x_direction)
               for r in range(1, Ny + 1) for i in range(1, Nx)]
    y_bonds = []
    if Ny >= 2:
        for i in range(1, Nx + 1):
            y_bonds.append(CouplingTerm(i, Nx + i, *J_y_direction))
        for r in range(2, Ny):
            for i in range(1, Nx + 1):
                y_bonds.append(CouplingTerm((r - 1) * Nx + i, r * Nx + i, *J_y_direction))
    fields = [FieldTerm(i, *h_field) for i in range(1, N + 1)]
    return HamiltonianParams(N, x_bonds, y_bonds, fields)


def xxz_chain(N: int, J: float, delta: float, hz) -> HamiltonianParams:
    """Open XXZ chain H = J Σ (XX + YY + Δ ZZ) + Σ h_i Z_i: bonds (J, J, JΔ),
    fields FieldTerm(i, 0, 0, h_i)."""
    bonds = [CouplingTerm(i, i + 1, J, J, J * delta) for i in range(1, N)]
    fields = [FieldTerm(i + 1, 0.0, 0.0, float(hz[i])) for i in range(N)]
    return HamiltonianParams(N, bonds, [], fields)


def pauli_sum(params: HamiltonianParams, reference_field_y: bool = True):
    """(x, z, c) masks and coefficients of H in the order of the pull loop of
    apply_hamiltonian_bitwise! (…UnitaryEvolutionChebyshev.jl:67-146):
    per bond ZZ, XX, YY; then per site Z, X, Y; zero couplings skipped.

    That loop takes +i for a 0 bit in its field-Y term (:139-142), i.e. it
    applies -hy*Y; reference_field_y keeps that sign for drop-in parity (the
    bond YY terms are the true Y⊗Y)."""
    xs, zs, cs = [], [], []

    def add(x, z, c):
        xs.append(x)
        zs.append(z)
        cs.append(float(c))

    for b in list(params.x_bonds) + list(params.y_bonds):
        m = (1 << (b.i - 1)) | (1 << (b.j - 1))
        if b.Jzz != 0.0:
            add(0, m, b.Jzz)
        if b.Jxx != 0.0:
            add(m, 0, b.Jxx)
        if b.Jyy != 0.0:
            add(m, m, b.Jyy)
    for f in params.fields:
        m = 1 << (f.idx - 1)
        if f.hz != 0.0:
            add(0, m, f.hz)
        if f.hx != 0.0:
            add(m, 0, f.hx)
        if f.hy != 0.0:
            add(m, m, -f.hy if reference_field_y else f.hy)
    return xs, zs, cs


def spectral_bound(params: HamiltonianParams) -> float:
    """||H|| <= Σ |c_t| (each Pauli string has norm 1)."""
    return float(sum(abs(c) for c in pauli_sum(params)[2]))
