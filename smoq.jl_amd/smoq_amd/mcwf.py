"""Monte Carlo wave-function trajectories with the reference's names
(utils/cpu/cpuQuantumChannelLindbladianEvolutionMonteCarloWaveFunction.jl),
on the device, in two forms:

* σ⁻ jumps with a diagonal-bond Trotter step: every step of every trajectory
  inside one persistent kernel launch (smoq_mcwf_evolve): the first-order
  Trotter step (diagonal bond gates, then site gates), the no-jump decay,
  the jump decision, σ⁻ and the renormalisation;
* any coherent step (a gate list compiled into a plan) with σ⁻ or σ_z
  jumps (smoq_mcwf_trotter_evolve / smoq_mcwf_jump): per step the plan's
  passes over all trajectories, then the dissipative half in two launches,
  with observables recorded by the tiled Pauli reduction.  The reference's `rand()` per trajectory and step is
replaced by a caller-supplied uniform (SplitMix64 by default), so a run is
reproducible and can be replayed on the CPU.

Groups of different qubit counts share one launch (`evolve_mcwf_groups`),
the largest states starting first.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import call, smoq_mcwf_group
from .hamiltonian import build_hamiltonian_parameters
from .state import GPUState
from .streams import uniforms as splitmix_uniforms
from .trotter import FastTrotterGate, precompute_trotter_gates_bitwise_cpu


@dataclass
class MCWFGates:
    """A first-order Trotter step in the form the trajectory kernel takes:
    bond_q (B x 2, 0-based, ascending), bond_diag (B x 4, the diagonal in
    the basis b(q0) + 2 b(q1)), site_u (N x 2 x 2, identity where a site
    has no gate)."""
    N: int
    bond_q: np.ndarray
    bond_diag: np.ndarray
    site_u: np.ndarray


def mcwf_gates(gates, N: int) -> MCWFGates:
    """Splits a gate list (FastTrotterGate, 1-based indices) into the kernel's
    form.  The kernel applies every 2-qubit gate before every 1-qubit gate,
    which is the order precompute_trotter_gates_bitwise_cpu emits
    (…UnitaryEvolutionTrotter.jl:330-351); 2-qubit gates must be diagonal
    (ZZ couplings: the transverse-field Ising chain of the MCWF studies)."""
    bq, bd = [], []
    site = np.tile(np.eye(2, dtype=np.complex128), (N, 1, 1))
    seen_site = False
    for g in gates:
        idx = list(g.indices)
        U = np.asarray(g.U, np.complex128)
        if len(idx) == 2:
            if seen_site:
                raise ValueError("mcwf_gates: a 2-qubit gate follows a 1-qubit gate")
            if np.any(U - np.diag(np.diag(U))):
                raise ValueError(f"mcwf_gates: bond {idx} is not diagonal")
            lo, hi = sorted(idx)
            bq.append((lo - 1, hi - 1))
            bd.append(np.diag(U))
        elif len(idx) == 1:
            seen_site = True
            k = idx[0] - 1
            site[k] = U @ site[k]
        else:
            raise ValueError("Gates with more than 2 qubits not supported")
    return MCWFGates(N, np.asarray(bq, np.int32).reshape(-1, 2),
                     np.asarray(bd, np.complex128).reshape(-1, 4), site)


def decay_powers(N: int, gamma: float, dt: float) -> np.ndarray:
    """decay_factor^m for m = 0..N, decay_factor = sqrt(1 - γ dt)
    (…MonteCarloWaveFunction.jl:344, 352-357)."""
    if not 0.0 <= gamma * dt <= 1.0:
        raise ValueError(f"gamma*dt = {gamma * dt} outside [0, 1]")
    f = np.sqrt(1.0 - gamma * dt)
    return np.array([f ** m for m in range(N + 1)], np.float64)


@dataclass
class MCWFResult:
    jumps: np.ndarray    # [traj, step] jump site (0-based) or -1
    z: np.ndarray | None  # [traj, record, site] <Z_i> after every obs_every-th step


class MCWFGroup:
    """The trajectories of one qubit count: the device batch, the step's
    gates, the uniforms and the output arrays of one launch."""

    def __init__(self, psi: GPUState, gates: MCWFGates, gamma: float, dt: float,
                 uniforms: np.ndarray):
        if gates.N != psi.n:
            raise ValueError(f"gates for {gates.N} qubits, state has {psi.n}")
        self.psi = psi
        self.gates = gates
        self.decay = decay_powers(psi.n, gamma, dt)
        self.uniforms = np.ascontiguousarray(uniforms, np.float64)
        self._bq = np.ascontiguousarray(gates.bond_q, np.int32)
        self._bd = np.ascontiguousarray(gates.bond_diag, np.complex128)
        self._su = np.ascontiguousarray(gates.site_u, np.complex128)

    def _struct(self, n_steps: int, obs_every: int):
        T = self.psi.batch
        u = self.uniforms
        if u.ndim != 2 or u.shape[0] != T or u.shape[1] < n_steps:
            raise ValueError(f"uniforms of shape {u.shape}, want ({T}, >= {n_steps})")
        # the first n_steps columns, contiguous; kept alive until the call returns
        self._u = u if u.shape[1] == n_steps else np.ascontiguousarray(u[:, :n_steps])
        self.jumps = np.empty((T, n_steps), np.int32)
        rec = n_steps // obs_every if obs_every > 0 else 0
        self.z = np.empty((T, rec, self.psi.n), np.float64) if rec > 0 else None
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
        return smoq_mcwf_group(self.psi.handle, len(self._bq), p(self._bq), p(self._bd),
                               p(self._su), p(self.decay), p(self._u), p(self.jumps),
                               p(self.z))


def evolve_mcwf_groups(groups, gamma: float, dt: float, n_steps: int, obs_every: int = 0):
    """Runs n_steps of every trajectory of every group in one launch and
    returns one MCWFResult per group."""
    groups = list(groups)
    arr = (smoq_mcwf_group * len(groups))(*[g._struct(n_steps, obs_every) for g in groups])
    call("smoq_mcwf_evolve", groups[0].psi.ctx.handle, arr, len(groups), gamma * dt, n_steps,
         obs_every)
    return [MCWFResult(g.jumps, g.z) for g in groups]


def evolve_mcwf_batched_trotter(psi: GPUState, gates, N: int, n_traj: int, n_steps: int,
                                gamma: float, dt: float, *, uniforms=None, seed: int = 0,
                                obs_every: int = 0) -> MCWFResult:
    """evolve_mcwf_batched_trotter! (…MonteCarloWaveFunction.jl:412-423) with
    apply_trotter! fixed to the gate list applied in order, rand() replaced
    by uniforms[traj, step] (SplitMix64 stream `seed` when None) and the
    observables callback by <Z_i> per trajectory every obs_every steps."""
    if psi.n != N or psi.batch != n_traj:
        raise ValueError(f"state holds {psi.batch} x {psi.n} qubits, want {n_traj} x {N}")
    g = gates if isinstance(gates, MCWFGates) else mcwf_gates(gates, N)
    if uniforms is None:
        uniforms = splitmix_uniforms(seed, n_traj * n_steps).reshape(n_traj, n_steps)
    return evolve_mcwf_groups([MCWFGroup(psi, g, gamma, dt, uniforms)], gamma, dt, n_steps,
                              obs_every)[0]


def lindblad_mcwf_batched_step(psi: GPUState, gates, N: int, n_traj: int, gamma: float,
                               dt: float, *, uniforms=None, seed: int = 0) -> np.ndarray:
    """lindblad_mcwf_batched_step! (…MonteCarloWaveFunction.jl:340-390): one
    step of every trajectory; returns the jump sites (0-based, -1 for none)."""
    u = None if uniforms is None else np.asarray(uniforms, np.float64).reshape(n_traj, 1)
    return evolve_mcwf_batched_trotter(psi, gates, N, n_traj, 1, gamma, dt, uniforms=u,
                                       seed=seed).jumps[:, 0]


# ------------------------------------------- any coherent step (plans) --
JUMP_KINDS = {"sigma_minus": 0, "sigma_z": 1}


def apply_sigma_z(psi: GPUState, k: int, N: int) -> GPUState:
    """apply_sigma_z! (…MonteCarloWaveFunction.jl:170-180): the phase flip of
    the σ_z jump, the same operation as apply_pauli_z_psi."""
    from .gates import apply_pauli_z_psi

    return apply_pauli_z_psi(psi, k, N)


def jump_decay_table(kind: str, N: int, gamma: float, dt: float) -> np.ndarray:
    """σ⁻: sqrt(1-γdt)^m per excitation count m; σ_z: the uniform
    sqrt(1 - Nγdt) of lindblad_mcwf_step_trotter! (…:286-287)."""
    if kind == "sigma_z":
        if not 0.0 <= N * gamma * dt <= 1.0:
            raise ValueError(f"N*gamma*dt = {N * gamma * dt} outside [0, 1]")
        return np.full(N + 1, np.sqrt(1.0 - N * gamma * dt))
    return decay_powers(N, gamma, dt)


def mcwf_jump(psi: GPUState, kind: str, gamma: float, dt: float, uniforms) -> np.ndarray:
    """The dissipative half of one MCWF step for every trajectory (batch
    item) of psi after its coherent step: decay, decision, jump, renormalise.
    uniforms: [traj, 2] (decision, site).  Returns 0-based sites or -1."""
    k = JUMP_KINDS[kind]
    u = np.ascontiguousarray(np.asarray(uniforms, np.float64).reshape(psi.batch, 2))
    dec = np.ascontiguousarray(jump_decay_table(kind, psi.n, gamma, dt))
    jumps = np.empty(psi.batch, np.int32)
    call("smoq_mcwf_jump", psi.handle, k, gamma * dt, dec.ctypes.data_as(ctypes.c_void_p),
         u.ctypes.data_as(ctypes.c_void_p), jumps.ctypes.data_as(ctypes.c_void_p))
    return jumps


def lindblad_mcwf_step_trotter(psi: GPUState, gates, N: int, gamma: float, dt: float, *,
                               uniforms) -> np.ndarray:
    """lindblad_mcwf_step_trotter! (…MonteCarloWaveFunction.jl:279-306), σ_z
    dephasing: the gate list (apply_fast_trotter_step_cpu), then
    mcwf_jump(…, "sigma_z", …); the two rand() draws are supplied as
    uniforms [traj, 2].  Runs every batch item of psi."""
    from .trotter import apply_fast_trotter_step_cpu

    apply_fast_trotter_step_cpu(psi, gates, N)
    return mcwf_jump(psi, "sigma_z", gamma, dt, uniforms)


def evolve_mcwf_trotter(psi: GPUState, gates, N: int, n_traj: int, n_steps: int, gamma: float,
                        dt: float, *, jump: str = "sigma_z", uniforms=None, seed: int = 0,
                        obs_every: int = 0, xmask=(), zmask=()):
    """MCWF with the coherent step of any gate list (FastTrotterGate list or
    Tape), compiled once into a plan: n_steps of [plan, mcwf_jump] on all
    trajectories, <P_t> per trajectory every obs_every steps.  uniforms:
    [traj, step, 2] (SplitMix64 stream `seed` when None).  Returns
    (jumps [traj, step], obs [record, traj, term])."""
    from .circuits import Tape
    from .trotter import trotter_tape

    if psi.n != N or psi.batch != n_traj:
        raise ValueError(f"state holds {psi.batch} x {psi.n} qubits, want {n_traj} x {N}")
    tape = gates if isinstance(gates, Tape) else trotter_tape(gates, N)
    plan = tape.compile(psi.dtype, ctx=psi.ctx)
    if uniforms is None:
        uniforms = splitmix_uniforms(seed, n_traj * n_steps * 2)
    u = np.ascontiguousarray(np.asarray(uniforms, np.float64).reshape(n_traj, n_steps, 2))
    dec = np.ascontiguousarray(jump_decay_table(jump, N, gamma, dt))
    xm = np.ascontiguousarray(np.asarray(xmask, np.uint64))
    zm = np.ascontiguousarray(np.asarray(zmask, np.uint64))
    rec = n_steps // obs_every if obs_every > 0 and xm.size else 0
    out = np.zeros((rec, n_traj, xm.size), np.float64)
    jumps = np.empty((n_traj, n_steps), np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    call("smoq_mcwf_trotter_evolve", plan.handle, psi.handle, JUMP_KINDS[jump], gamma * dt,
         p(dec), int(n_steps), p(u), p(jumps), int(obs_every) if rec else 0, int(xm.size),
         p(xm), p(zm), p(out))
    return jumps, out


def xxz_field_trotter_gates(N: int, dt: float, J=(1.0, 1.0, 0.5), hx: float = 1.0):
    """The first-order step of the σ_z-dephasing demo
    (scripts_quantum_simulations/demo_lidbladian_dm_exact_trotter_mcwf_exact_trotter_comparison.jl:64-300):
    open XXZ chain with couplings J = (Jxx, Jyy, Jzz) and a field hx on X,
    from precompute_trotter_gates_bitwise_cpu (bonds, then sites)."""
    params = build_hamiltonian_parameters(N, 1, J_x_direction=tuple(J),
                                          J_y_direction=(0.0, 0.0, 0.0), h_field=(hx, 0.0, 0.0))
    return precompute_trotter_gates_bitwise_cpu(params, dt)


def tfim_mcwf_gates(N: int, J: float, h: float, dt: float) -> MCWFGates:
    """The transverse-field Ising chain H = J Σ Z_i Z_{i+1} + h Σ X_i (open
    boundary; build_hamiltonian_parameters(N, 1, :TFIM_ZZ_X)) as one
    first-order step: exp(-i J ZZ dt) per bond, then exp(-i h X dt) per site,
    from precompute_trotter_gates_bitwise_cpu."""
    params = build_hamiltonian_parameters(N, 1, "TFIM_ZZ_X", J=J, h=h)
    return mcwf_gates(precompute_trotter_gates_bitwise_cpu(params, dt), N)


__all__ = ["MCWFGates", "MCWFGroup", "MCWFResult", "FastTrotterGate", "apply_sigma_z",
           "decay_powers", "evolve_mcwf_batched_trotter", "evolve_mcwf_groups",
           "evolve_mcwf_trotter", "jump_decay_table", "lindblad_mcwf_batched_step",
           "lindblad_mcwf_step_trotter", "mcwf_gates", "mcwf_jump", "tfim_mcwf_gates",
           "xxz_field_trotter_gates"]
