"""CPU oracle for the SmoQ.jl state-vector path — TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``liboracle.so`` (built from ``smoq_oracle.c`` by the
Makefile next to this file) plus small numpy helpers.  Only the tests, the
smoke check and bench.py's CPU-baseline leg use this module, and only as the
checker: the product never imports it.

States are complex128 numpy arrays; qubit arguments are 1-based like the
reference functions each wrapper restates (see the C file for the
file:line each loop follows).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_P = ctypes.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_apply_rx": [_P, _i, _d, _i],
            "or_apply_ry": [_P, _i, _d, _i],
            "or_apply_rz": [_P, _i, _d, _i],
            "or_apply_h": [_P, _i, _i],
            "or_apply_x": [_P, _i, _i],
            "or_apply_y": [_P, _i, _i],
            "or_apply_phase1": [_P, _i, _d, _d, _i],
            "or_apply_cnot": [_P, _i, _i, _i],
            "or_apply_cz": [_P, _i, _i, _i],
            "or_apply_cry": [_P, _i, _i, _d, _i],
            "or_apply_rxx_ryy": [_P, _i, _i, _d, _i, _i],
            "or_apply_rzz": [_P, _i, _i, _d, _i],
            "or_apply_ccz": [_P, _i, _i, _i, _i],
            "or_apply_toffoli": [_P, _i, _i, _i, _i],
            "or_apply_sigma": [_P, _i, _i, _i],
            "or_apply_1q": [_P, _P, _i, _i],
            "or_apply_2q_trotter": [_P, _P, _i, _i, _i],
            "or_apply_fused_1q": [_P, _P, _i, _i],
            "or_expect_local": [_P, _i, _i, ctypes.c_char],
            "or_expect_corr": [_P, _i, _i, _i, ctypes.c_char],
            "or_pauli_string_ref": [_P, ctypes.c_char_p, _P, _i, _i],
            "or_compute_energy": [_P, _i, _d, _d, _d, _d],
            "or_probs": [_P, _i64, _P],
            "or_cumsum_pairwise": [_P, _i64, _P],
            "or_cumsum_sequential": [_P, _i64, _P],
            "or_searchsortedfirst": [_P, _i64, _d],
            "or_sample_cumsum_search": [_P, _i, _P, _i64, _P],
            "or_sample_sequential": [_P, _i, _P, _i64, _P],
            "or_hpsi": [_P, _P, _i, _i, _P, _P, _P, _P],
            "or_mcwf_decay_jump": [_P, _i, _d, _P, _d, _P],
            "or_mcwf_dephasing": [_P, _i, _d, _d, _d, _P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        for name in ("or_expect_local", "or_expect_corr", "or_pauli_string_ref",
                     "or_compute_energy"):
            getattr(L, name).restype = _d
        L.or_searchsortedfirst.restype = _i64
        L.or_mcwf_decay_jump.restype = _i
        L.or_mcwf_dephasing.restype = _i
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(psi):
    assert psi.dtype == np.complex128 and psi.flags["C_CONTIGUOUS"]
    return psi


# ------------------------------------------------------------------ gates --
def apply_named(psi, gate: str, qubits, theta: float, N: int):
    """In-place gate on a complex128 state; qubits 1-based."""
    L, p = lib(), _p(_c128(psi))
    q = list(qubits)
    if gate == "rx":
        L.or_apply_rx(p, q[0], theta, N)
    elif gate == "ry":
        L.or_apply_ry(p, q[0], theta, N)
    elif gate == "rz":
        L.or_apply_rz(p, q[0], theta, N)
    elif gate == "h":
        L.or_apply_h(p, q[0], N)
    elif gate == "x":
        L.or_apply_x(p, q[0], N)
    elif gate == "y":
        L.or_apply_y(p, q[0], N)
    elif gate == "z":
        L.or_apply_phase1(p, q[0], -1.0, 0.0, N)
    elif gate == "s":
        L.or_apply_phase1(p, q[0], 0.0, 1.0, N)
    elif gate == "sdg":
        L.or_apply_phase1(p, q[0], 0.0, -1.0, N)
    elif gate == "t":
        L.or_apply_phase1(p, q[0], np.cos(np.pi / 4), np.sin(np.pi / 4), N)
    elif gate == "tdg":
        L.or_apply_phase1(p, q[0], np.cos(np.pi / 4), -np.sin(np.pi / 4), N)
    elif gate == "cnot":
        L.or_apply_cnot(p, q[0], q[1], N)
    elif gate == "cz":
        L.or_apply_cz(p, q[0], q[1], N)
    elif gate == "cry":
        L.or_apply_cry(p, q[0], q[1], theta, N)
    elif gate == "rxx":
        L.or_apply_rxx_ryy(p, q[0], q[1], theta, 0, N)
    elif gate == "ryy":
        L.or_apply_rxx_ryy(p, q[0], q[1], theta, 1, N)
    elif gate == "rzz":
        L.or_apply_rzz(p, q[0], q[1], theta, N)
    elif gate == "ccz":
        L.or_apply_ccz(p, q[0], q[1], q[2], N)
    elif gate == "toffoli":
        L.or_apply_toffoli(p, q[0], q[1], q[2], N)
    elif gate == "sigma_minus":
        L.or_apply_sigma(p, q[0], 0, N)
    elif gate == "sigma_plus":
        L.or_apply_sigma(p, q[0], 1, N)
    else:
        raise ValueError(gate)
    return psi


def apply_1q(psi, U, q: int, N: int):
    u = np.ascontiguousarray(np.asarray(U, np.complex128).reshape(2, 2))
    lib().or_apply_1q(_p(_c128(psi)), _p(u), q, N)
    return psi


def apply_2q_trotter(psi, U, q1: int, q2: int, N: int):
    u = np.ascontiguousarray(np.asarray(U, np.complex128).reshape(4, 4))
    lib().or_apply_2q_trotter(_p(_c128(psi)), _p(u), q1, q2, N)
    return psi


def apply_fused_1q(psi, U, q: int, N: int):
    u = np.ascontiguousarray(np.asarray(U, np.complex128).reshape(2, 2))
    lib().or_apply_fused_1q(_p(_c128(psi)), _p(u), q, N)
    return psi


# ------------------------------------------------------------ observables --
def expect_local(psi, k: int, N: int, pauli: str) -> float:
    return lib().or_expect_local(_p(_c128(psi)), k, N, pauli.encode())


def expect_corr(psi, i: int, j: int, N: int, pair: str) -> float:
    return lib().or_expect_corr(_p(_c128(psi)), i, j, N, pair[0].encode())


def pauli_string_ref(psi, ops, qubits, N: int) -> float:
    q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
    s = "".join(ops).encode()
    return lib().or_pauli_string_ref(_p(_c128(psi)), s, _p(q), len(ops), N)


def compute_energy(psi, N, Jx, Jy, Jz, hx) -> float:
    return lib().or_compute_energy(_p(_c128(psi)), N, Jx, Jy, Jz, hx)


# --------------------------------------------------------------- sampling --
def probs(psi):
    out = np.empty(psi.size, np.float64)
    lib().or_probs(_p(_c128(psi)), psi.size, _p(out))
    return out


def cumsum_pairwise(v):
    v = np.ascontiguousarray(v, np.float64)
    c = np.empty_like(v)
    lib().or_cumsum_pairwise(_p(v), v.size, _p(c))
    return c


def cumsum_sequential(v):
    v = np.ascontiguousarray(v, np.float64)
    c = np.empty_like(v)
    lib().or_cumsum_sequential(_p(v), v.size, _p(c))
    return c


def searchsortedfirst(v, x: float) -> int:
    v = np.ascontiguousarray(v, np.float64)
    return lib().or_searchsortedfirst(_p(v), v.size, float(x))


def sample_cumsum_search(psi, N, r):
    r = np.ascontiguousarray(r, np.float64)
    out = np.empty(r.size, np.int64)
    lib().or_sample_cumsum_search(_p(_c128(psi)), N, _p(r), r.size, _p(out))
    return out


def sample_sequential(psi, N, r):
    r = np.ascontiguousarray(r, np.float64)
    out = np.empty(r.size, np.int64)
    lib().or_sample_sequential(_p(_c128(psi)), N, _p(r), r.size, _p(out))
    return out


# ----------------------------------------------------------- H psi, MCWF --
def hpsi(psi, N, bonds, fields):
    """bonds: list of (i, j, Jxx, Jyy, Jzz) 1-based; fields: N x (hx, hy, hz)."""
    bi = np.ascontiguousarray([b[0] for b in bonds], np.int32)
    bj = np.ascontiguousarray([b[1] for b in bonds], np.int32)
    J = np.ascontiguousarray([b[2:5] for b in bonds], np.float64).reshape(-1)
    h = np.ascontiguousarray(fields, np.float64).reshape(-1)
    out = np.empty_like(psi)
    lib().or_hpsi(_p(out), _p(_c128(psi)), N, len(bonds), _p(bi), _p(bj), _p(J), _p(h))
    return out


def mcwf_decay_jump(psi, N, gamma_dt, decay_pow, r):
    dp = np.zeros(1)
    d = np.ascontiguousarray(decay_pow, np.float64)
    site = lib().or_mcwf_decay_jump(_p(_c128(psi)), N, gamma_dt, _p(d), float(r), _p(dp))
    return site, float(dp[0])


def mcwf_dephasing(psi, N, decay, r, u2):
    dp = np.zeros(1)
    site = lib().or_mcwf_dephasing(_p(_c128(psi)), N, float(decay), float(r), float(u2), _p(dp))
    return site, float(dp[0])


def mcwf_dephasing_trajectory(psi, gates, N, gamma_dt, r_seq, u2_seq):
    """lindblad_mcwf_step_trotter! (…MonteCarloWaveFunction.jl:279-306)
    repeated: the gate list in order, then or_mcwf_dephasing with decay
    sqrt(1 - N γdt) and the supplied uniforms.  Returns 0-based sites or -1."""
    decay = np.sqrt(1.0 - N * gamma_dt)
    jumps = []
    for r, u2 in zip(r_seq, u2_seq):
        trotter_step(psi, gates, N)
        site, _ = mcwf_dephasing(psi, N, decay, r, u2)
        jumps.append(site - 1)
    return np.array(jumps, np.int32)


def mcwf_trajectory(psi, gates, N, gamma_dt, decay_pow, r_seq, obs_every=0):
    """One trajectory of evolve_mcwf_batched_trotter!
    (…MonteCarloWaveFunction.jl:412-423): per step the gate list in order
    (apply_fast_trotter_step_cpu!), then or_mcwf_decay_jump with r_seq[step]
    in place of rand().  Returns (jump sites 0-based or -1, <Z_i> after
    every obs_every-th step as [record][i])."""
    idx = np.arange(1 << N)
    jumps, zs = [], []
    for s, r in enumerate(r_seq):
        trotter_step(psi, gates, N)
        site, _ = mcwf_decay_jump(psi, N, gamma_dt, decay_pow, float(r))
        jumps.append(site - 1)
        if obs_every and (s + 1) % obs_every == 0:
            p = np.abs(psi) ** 2
            zs.append([float(np.sum(p * (1 - 2 * ((idx >> k) & 1)))) for k in range(N)])
    return np.array(jumps, np.int32), np.array(zs, np.float64).reshape(-1, N)


def trotter_step(psi, gates, N):
    """apply_fast_trotter_step_cpu! (utils/cpu/cpuQuantumChannelUnitaryEvolutionTrotter.jl:150-162):
    gates = [(indices 1-based, U)] applied in order."""
    for idx, U in gates:
        if len(idx) == 1:
            apply_1q(psi, U, idx[0], N)
        else:
            apply_2q_trotter(psi, U, idx[0], idx[1], N)
    return psi


def lanczos_range(psi0, N, bonds, fields, k_iter=30):
    """estimate_spectral_range_lanczos (utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl:166-220)
    from the given start vector (the reference draws randn): H·v by the C
    pull loop (or_hpsi), the same recurrence and break rule, eigenvalues of
    SymTridiagonal(alphas, betas) by scipy.  Returns (E_min, E_max, steps)."""
    from scipy.linalg import eigh_tridiagonal

    v = np.asarray(psi0, np.complex128) / np.linalg.norm(psi0)
    v_prev = np.zeros_like(v)
    alphas, betas = [], []
    for j in range(k_iter):
        w = hpsi(v, N, bonds, fields)
        if j > 0:
            w = w - betas[-1] * v_prev
        alpha = float(np.vdot(v, w).real)
        alphas.append(alpha)
        w = w - alpha * v
        beta = float(np.linalg.norm(w))
        if beta < 1e-12:
            break
        betas.append(beta)
        v_prev = v
        v = w / beta
    ev = eigh_tridiagonal(np.array(alphas), np.array(betas[:len(alphas) - 1]), eigvals_only=True)
    return float(ev.min()), float(ev.max()), len(alphas)


def chebyshev_evolve(psi, N, bonds, fields, t, emin, emax, tol=1e-12, max_order=1000):
    """chebyshev_evolve_psi! (utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl:262-346)
    with supplied bounds; H·ψ by the C pull loop (or_hpsi), Bessel values
    from scipy.special.jv (the reference calls SpecialFunctions.besselj).
    Returns (psi_out, M_used)."""
    from scipy.special import jv

    width = emax - emin
    lo, hi = emin - 0.01 * abs(width), emax + 0.01 * abs(width)
    a, b = (hi - lo) / 2.0, (hi + lo) / 2.0
    tau = a * t
    psi_in = psi.copy()
    T_prev = psi_in.copy()
    H = hpsi(psi_in, N, bonds, fields)
    T_curr = (H - b * psi_in) / a
    out = jv(0, tau) * T_prev
    out = out + (-2j * jv(1, tau)) * T_curr
    M_used = 2
    for k in range(2, max_order + 1):
        H = hpsi(T_curr, N, bonds, fields)
        T_next = 2 * ((H - b * T_curr) / a) - T_prev
        out = out + (2 * (1, -1j, -1, 1j)[k % 4] * jv(k, tau)) * T_next
        M_used = k + 1
        if abs(jv(k, tau)) < tol:
            break
        T_prev, T_curr = T_curr, T_next
    return out * np.exp(-1j * b * t), M_used



# ------------------------------------------------ random-circuit benchmark --
_R2 = 1.0 / np.sqrt(2.0)
BENCH_FIXED = {
    # run_benchmark_smoq_Float64.jl:83-88
    "H": np.array([[_R2, _R2], [_R2, -_R2]], np.complex128),
    "X": np.array([[0, 1], [1, 0]], np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], np.complex128),
    "Z": np.array([[1, 0], [0, -1]], np.complex128),
    "S": np.array([[1, 0], [0, 1j]], np.complex128),
    "T": np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], np.complex128),
}


def bench_gate_matrix(name: str, theta: float):
    """get_gate_matrix (run_benchmark_smoq_Float64.jl:91-93,146-157)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    if name == "RX":
        return np.array([[c, -1j * s], [-1j * s, c]], np.complex128)
    if name == "RY":
        return np.array([[c, -s], [s, c]], np.complex128)
    if name == "RZ":
        return np.array([[np.exp(-1j * theta / 2), 0], [0, np.exp(1j * theta / 2)]],
                        np.complex128)
    return BENCH_FIXED[name]


def read_circuit_file(path: str):
    """load_circuit_from_file (run_benchmark_smoq_Float64.jl:384-404):
    returns [(name, q1, q2, theta)] with 1-based qubits, q2 = 0 for 1-qubit
    gates; for CNOT/CZ q1 is the first (control) qubit of the line."""
    ops = []
    for line in open(path, encoding="utf-8"):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        parts = line.split()
        name, q1 = parts[0], int(parts[1]) + 1
        if name in ("CNOT", "CZ"):
            ops.append((name, q1, int(parts[2]) + 1, 0.0))
        elif name in ("RX", "RY", "RZ"):
            ops.append((name, q1, 0, float(parts[2])))
        else:
            ops.append((name, q1, 0, 0.0))
    return ops


def fuse_gates(ops):
    """fuse_gates (run_benchmark_smoq_Float64.jl:167-200): per-qubit pending
    2x2 products U_new · U_acc, flushed for both qubits of a 2-qubit gate and
    at the end in ascending qubit order.  Returns [(is_single, qubit,
    control, U, name)]."""
    fused, pending = [], {}

    def flush(q):
        if q in pending:
            fused.append((True, q, 0, pending.pop(q), "U"))

    for name, q1, q2, th in ops:
        if q2 == 0:
            U = bench_gate_matrix(name, th)
            pending[q1] = U @ pending[q1] if q1 in pending else U.copy()
        else:
            flush(q1)
            flush(q2)
            fused.append((False, q2, q1, None, name))
    for q in sorted(pending):
        flush(q)
    return fused


def run_fused_circuit(psi, fused, N: int):
    """run_fused_circuit! (run_benchmark_smoq_Float64.jl:356-369)."""
    for single, q, c, U, name in fused:
        if single:
            apply_fused_1q(psi, U, q, N)
        elif name == "CNOT":
            lib().or_apply_cnot(_p(_c128(psi)), c, q, N)
        else:
            lib().or_apply_cz(_p(_c128(psi)), c, q, N)
    return psi


def run_circuit_sampling(N: int, ops, r):
    """run_circuit_sampling (run_benchmark_smoq_Float64.jl:414-437) with the
    uniforms r in place of rand(): fuse, simulate |0…0>, then
    searchsortedfirst(cumsum(abs2.(ψ)), r) - 1 per shot.  Returns (ψ, 0-based
    indices; 2^N when r exceeds the last prefix)."""
    psi = np.zeros(1 << N, np.complex128)
    psi[0] = 1.0
    run_fused_circuit(psi, fuse_gates(ops), N)
    return psi, sample_cumsum_search(psi, N, r)

# ------------------------------------------------------- density matrices --
def dm_apply_gate_list(rho, gates, N):
    """The coherent half of lindblad_dm_step_trotter!
    (utils/cpu/cpuQuantumChannelLindbladianEvolutionDensityMatrix.jl:129-238):
    every gate on every column (U on the row index, 2-qubit gates with
    sorted qubits and U unpermuted), then conj(U) on every row."""
    dim = 1 << N
    for idx, U in gates:
        for c in range(dim):
            col = np.ascontiguousarray(rho[:, c])
            if len(idx) == 1:
                apply_1q(col, U, idx[0], N)
            else:
                apply_2q_trotter(col, U, min(idx), max(idx), N)
            rho[:, c] = col
    for idx, U in gates:
        Uc = np.conj(np.asarray(U))
        for r in range(dim):
            row = np.ascontiguousarray(rho[r, :])
            if len(idx) == 1:
                apply_1q(row, Uc, idx[0], N)
            else:
                apply_2q_trotter(row, Uc, min(idx), max(idx), N)
            rho[r, :] = row
    return rho


def dm_dissipator(rho, kind, k, gdt, N):
    """_apply_dissipator_dm! (utils/cpu/cpuQuantumChannelLindbladianEvolution.jl:377-454)
    for one site k (1-based), first-order step with factor gdt; sigma_z is
    also apply_sigma_z_dissipator! (…DensityMatrix.jl:53-69) and sigma_minus
    apply_sigma_minus_dissipator_rho! (cpuQuantumChannelGates.jl:833-854)."""
    dim = 1 << N
    m = 1 << (k - 1)
    idx = np.arange(dim)
    bi = ((idx >> (k - 1)) & 1)[:, None]
    bj = ((idx >> (k - 1)) & 1)[None, :]
    old = rho.copy()
    if kind == "sigma_z":
        rho[bi != bj] *= (1 - 2 * gdt)
    elif kind in ("sigma_minus", "sigma_plus"):
        src = 1 if kind == "sigma_minus" else 0
        both = (bi == src) & (bj == src)
        one = ((bi == src) | (bj == src)) & ~both
        I, J = np.nonzero(both)
        rho[I ^ m, J ^ m] += gdt * old[I, J]
        rho[both] = old[both] * (1 - gdt)
        rho[one] = old[one] * (1 - 0.5 * gdt)
    elif kind == "sigma_x":
        I, J = np.meshgrid(idx, idx, indexing="ij")
        rho[:, :] = old + gdt * (old[I ^ m, J ^ m] - old)
    else:
        raise ValueError(kind)
    return rho


def dm_trotter_step(rho, gates, N, gamma, dt):
    """lindblad_dm_step_trotter! (…DensityMatrix.jl:129-243): gates, then the
    sigma_z dissipator on every site with gamma*dt."""
    dm_apply_gate_list(rho, gates, N)
    for k in range(1, N + 1):
        dm_dissipator(rho, "sigma_z", k, gamma * dt, N)
    return rho


def dm_exact_step(rho, U, N, gamma, dt):
    """lindblad_dm_step! (…DensityMatrix.jl:105-115): rho = U rho U', then
    the sigma_z dissipator on every site."""
    rho[:, :] = U @ rho @ np.conj(U).T
    for k in range(1, N + 1):
        dm_dissipator(rho, "sigma_z", k, gamma * dt, N)
    return rho


def dm_expect_local(rho, k, N, pauli):
    """_expect_Z_dm / _expect_X_dm / _expect_Y_dm
    (utils/cpu/cpuQuantumStateObservables.jl:1271-1340)."""
    dim = 1 << N
    idx = np.arange(dim)
    low = idx[((idx >> (k - 1)) & 1) == 0]
    if pauli == "z":
        return float(np.sum(rho[idx, idx].real * (1 - 2 * ((idx >> (k - 1)) & 1))))
    step = 1 << (k - 1)
    if pauli == "x":
        return float(np.sum(2 * rho[low, low + step].real))
    return float(np.sum(2 * rho[low + step, low].imag))


def dm_expect_corr(rho, i, j, N, pair):
    """_expect_ZZ_dm / _expect_XX_dm / _expect_YY_dm (…Observables.jl:1357-1415)."""
    dim = 1 << N
    idx = np.arange(dim)
    bi, bj = (idx >> (i - 1)) & 1, (idx >> (j - 1)) & 1
    if pair == "zz":
        return float(np.sum(rho[idx, idx].real * (1 - 2 * (bi ^ bj))))
    s = idx[(bi == 0) & (bj == 0)]
    si, sj = 1 << (i - 1), 1 << (j - 1)
    s00, s01, s10, s11 = s, s + sj, s + si, s + si + sj
    if pair == "xx":
        return float(np.sum(rho[s00, s11].real + rho[s11, s00].real + rho[s01, s10].real
                            + rho[s10, s01].real))
    return float(np.sum(-2 * rho[s00, s11].real + 2 * rho[s01, s10].real))


# --------------------------------------- noise channels and measurement --
def _halves(N, k):
    idx = np.arange(1 << N)
    lo = idx[(idx & (1 << (k - 1))) == 0]
    return lo, lo | (1 << (k - 1))


def _abs2(v):
    return v.real * v.real + v.imag * v.imag


def psi_channel(psi, kind, p, k, N, u):
    """The state-vector (trajectory) branch of one channel on qubit k (1-based)
    with rand() = u; returns the record (0 none, 1 X, 2 Y, 3 Z, 4 jump).
    utils/cpu/cpuQuantumChannelKrausOperators.jl: _apply_dephasing_single!
    :262-282, _apply_depolarizing_single! :285-318, _apply_bit_flip_single!
    :321-334, _apply_phase_flip_single! :337-347,
    _apply_amplitude_damping_single! :350-385."""
    lo, hi = _halves(N, k)

    def x():
        psi[lo], psi[hi] = psi[hi].copy(), psi[lo].copy()

    def y():
        a, b = psi[lo].copy(), psi[hi].copy()
        psi[lo], psi[hi] = -1j * b, 1j * a

    def z():
        psi[hi] = -psi[hi]

    if kind == "dephasing":
        if p == 0.0:
            return 0
        if u < (1.0 - np.sqrt(1.0 - p)) / 2.0:
            z()
            return 3
        return 0
    if kind == "depolarizing":
        if u < p / 3:
            x()
            return 1
        if u < 2 * p / 3:
            y()
            return 2
        if u < p:
            z()
            return 3
        return 0
    if kind == "bit_flip":
        if u < p:
            x()
            return 1
        return 0
    if kind == "phase_flip":
        if u < p:
            z()
            return 3
        return 0
    if kind == "amplitude_damping":
        pop1 = float(np.sum(_abs2(psi[hi])))
        rec = 0
        if u < p * pop1 and pop1 > 1e-12:
            psi[lo] += psi[hi]
            psi[hi] = 0
            rec = 4
        else:
            psi[hi] *= np.sqrt(1 - p)
        nrm = np.sqrt(np.sum(_abs2(psi)))
        if nrm > 1e-12:
            psi /= nrm
        return rec
    raise ValueError(kind)


def dm_channel(rho, kind, p, k, N):
    """The density-matrix update of one channel on qubit k (1-based),
    …KrausOperators.jl: dephasing :86-103, depolarizing :107-141, amplitude
    damping :145-181, bit flip :184-215, phase flip :218-235."""
    if p == 0.0:
        return rho
    dim = 1 << N
    m = 1 << (k - 1)
    idx = np.arange(dim)
    bi = ((idx >> (k - 1)) & 1)[:, None]
    bj = ((idx >> (k - 1)) & 1)[None, :]
    diff = np.broadcast_to(bi != bj, rho.shape)
    I, J = np.meshgrid(idx, idx, indexing="ij")
    old = rho.copy()
    flip = old[I ^ m, J ^ m]
    if kind == "dephasing":
        rho[diff] *= np.sqrt(1 - p)
    elif kind == "phase_flip":
        rho[diff] *= 1 - 2 * p
    elif kind == "depolarizing":
        p3 = p / 3.0
        same = ~diff
        rho[same] = (1.0 - p) * old[same] + p3 * (2 * flip[same] + old[same])
        rho[diff] = (1.0 - p) * old[diff] - p3 * old[diff]
    elif kind == "bit_flip":
        rho[:, :] = (1 - p) * old + p * flip
    elif kind == "amplitude_damping":
        rho[diff] *= np.sqrt(1 - p)
        both = (bi == 1) & (bj == 1)
        Ib, Jb = np.nonzero(both)
        rho[Ib ^ m, Jb ^ m] += p * old[Ib, Jb]
        rho[Ib, Jb] = old[Ib, Jb] * (1 - p)
    else:
        raise ValueError(kind)
    return rho


def measure_z(psi, k, N, u):
    """_measure_z_single!(ψ) (utils/cpu/cpuQuantumStateMeasurements.jl:224-256):
    outcome = u < P(0) ? 0 : 1, collapse, ψ *= 1/sqrt(kept norm^2)."""
    lo, hi = _halves(N, k)
    p0 = float(np.sum(_abs2(psi[lo])))
    o = 0 if u < p0 else 1
    kill, keep = (hi, lo) if o == 0 else (lo, hi)
    psi[kill] = 0
    psi *= 1.0 / np.sqrt(float(np.sum(_abs2(psi[keep]))))
    return o


def force_qubit(psi, k, target, N):
    """_force_qubit_state! (…Measurements.jl:584-626) with _normalize!
    (:629-635): the other half's amplitudes added onto the target half, ψ
    divided by its norm when norm^2 > 1e-15; :plus / :minus then H."""
    lo, hi = _halves(N, k)
    b = 0 if target in ("zero", "plus") else 1
    src, dst = (hi, lo) if b == 0 else (lo, hi)
    psi[dst] += psi[src]
    psi[src] = 0
    n2 = float(np.sum(_abs2(psi)))
    if n2 > 1e-15:
        psi /= np.sqrt(n2)
    if target in ("plus", "minus"):
        a, c = psi[lo].copy(), psi[hi].copy()
        psi[lo], psi[hi] = (a + c) / np.sqrt(2), (a - c) / np.sqrt(2)
    return psi


def reset_state(psi, qubits, targets, N, uniforms):
    """reset_state!(ψ) (…Measurements.jl:555-567): per qubit a Z measurement
    (one uniform) then the forced target."""
    for k, t, u in zip(qubits, targets, uniforms):
        measure_z(psi, k, N, u)
        force_qubit(psi, k, t, N)
    return psi


def dm_measure_z(rho, k, N, u):
    """_measure_z_single!(ρ) (…Measurements.jl:259-295): P(0) from the diagonal,
    outcome = u < P(0) ? 0 : 1, rows and columns with the other bit zeroed,
    ρ /= prob (prob = P(0) or 1 - P(0)) when prob > 1e-15."""
    idx = np.arange(1 << N)
    bit = (idx >> (k - 1)) & 1
    p0 = float(np.sum(rho[idx, idx].real[bit == 0]))
    o = 0 if u < p0 else 1
    prob = p0 if o == 0 else 1.0 - p0
    rho[bit != o, :] = 0
    rho[:, bit != o] = 0
    if prob > 1e-15:
        rho /= prob
    return o


def dm_force(rho, k, b, N):
    """_force_qubit_rho! (…Measurements.jl:686-708): P_b ρ P_b / Tr when the
    trace exceeds 1e-15."""
    idx = np.arange(1 << N)
    bit = (idx >> (k - 1)) & 1
    rho[bit != b, :] = 0
    rho[:, bit != b] = 0
    tr = float(np.sum(rho[idx, idx].real))
    if tr > 1e-15:
        rho /= tr
    return rho


def sample_from_probabilities(probs, r):
    """sample_from_probabilities (…Measurements.jl:203-213), 0-based: the first
    i with r <= running sum, else the last index."""
    s = 0.0
    for i, v in enumerate(probs):
        s = s + float(v)
        if r <= s:
            return i
    return len(probs) - 1


# --------------------------------------------------------- partial trace --
def _bit_map(bits, count):
    """index -> full index with the bits of the index deposited at `bits`
    (0-based positions, in increasing order): the reference's keep_to_full /
    trace_to_full tables."""
    out = np.zeros(1 << count, np.int64)
    for b_idx, full_bit in enumerate(bits):
        out |= ((np.arange(1 << count) >> b_idx) & 1) << full_bit
    return out


def partial_trace_psi(psi, trace_qubits, N):
    """partial_trace(ψ, trace_qubits, N) (utils/cpu/cpuQuantumStatePartialTrace.jl:68-127):
    rho[i, j] = sum_t ψ[keep_to_full[i] | trace_to_full[t]] conj(ψ[... j ...])."""
    tr = sorted(set(int(q) for q in trace_qubits))
    keep = [q for q in range(1, N + 1) if q not in tr]
    k2f = _bit_map([q - 1 for q in keep], len(keep))
    t2f = _bit_map([q - 1 for q in tr], len(tr))
    rho = np.zeros((len(k2f), len(k2f)), np.complex128)
    for i, ib in enumerate(k2f):
        a = psi[ib | t2f]
        for j, jb in enumerate(k2f):
            rho[i, j] = np.sum(a * np.conj(psi[jb | t2f]))
    return rho


def partial_trace_rho(rho, trace_qubits, N):
    """partial_trace(ρ, trace_qubits, N) (…PartialTrace.jl:141-188):
    rho_A[i, j] = sum_t ρ[i|t, j|t]."""
    tr = sorted(set(int(q) for q in trace_qubits))
    keep = [q for q in range(1, N + 1) if q not in tr]
    k2f = _bit_map([q - 1 for q in keep], len(keep))
    t2f = _bit_map([q - 1 for q in tr], len(tr))
    out = np.zeros((len(k2f), len(k2f)), np.complex128)
    for i, ib in enumerate(k2f):
        for j, jb in enumerate(k2f):
            out[i, j] = np.sum(rho[ib | t2f, jb | t2f])
    return out


# ---------------------------------------------------- classical shadows --
def pauli_shadow_snapshot(psi, N, bases, r):
    """classical_shadow_snapshot(ψ, N) (utils/cpu/cpuClassicalShadows.jl:1001-1037)
    with the bases (1 = X, 2 = Y, 3 = Z) and rand() given: a copy rotated
    qubit by qubit (H; S† then H), then sample_from_probabilities
    (…Measurements.jl:203-213) on |ψ|^2; returns the 0-based outcome index."""
    w = np.array(psi, np.complex128)
    for k in range(1, N + 1):
        if bases[k - 1] == 1:
            apply_named(w, "h", [k], 0.0, N)
        elif bases[k - 1] == 2:
            apply_named(w, "sdg", [k], 0.0, N)
            apply_named(w, "h", [k], 0.0, N)
    return int(sample_sequential(w, N, [r])[0])


def clifford_shadow_snapshot(psi, N, indices, r, table):
    """clifford_shadow_snapshot(ψ, N) (…ClassicalShadows.jl:580-603): ψ times
    the dense kron(C_N, ..., C_1) (build_local_clifford_unitary, :539-546) of
    the 1-based `indices` into `table`, then the same draw."""
    C = np.asarray(table[indices[N - 1] - 1], np.complex128)
    for j in range(N - 2, -1, -1):
        C = np.kron(C, table[indices[j] - 1])
    return int(sample_sequential(C @ np.asarray(psi, np.complex128), N, [r])[0])


def rho_shadow_snapshot(rho, N, mats, r):
    """classical_shadow_snapshot(ρ, N) / clifford_shadow_snapshot(ρ, N)
    (…ClassicalShadows.jl:608-632, 1039-1070): U ρ U' with U the Kronecker
    product of the per-qubit `mats` (qubit 1 lowest), populations clamped at
    0 and divided by their sum, then sample_from_probabilities."""
    U = np.asarray(mats[N - 1], np.complex128)
    for j in range(N - 2, -1, -1):
        U = np.kron(U, mats[j])
    p = np.maximum(np.real(np.diag(U @ rho @ U.conj().T)), 0.0)
    p = p / np.sum(p)
    return sample_from_probabilities(p, r)


def global_clifford_unitary(N, gates):
    """build_global_clifford_unitary (utils/cpu/cpuClassicalShadows.jl:835-870):
    U = G_L ... G_1; ("H", q) and ("S", q) as kron(G_1, ..., G_N) with the
    gate at factor q (bit N - q), ("CNOT", c, t) on bits c-1 -> t-1."""
    dim = 1 << N
    H = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2)
    S = np.array([[1, 0], [0, 1j]], np.complex128)
    U = np.eye(dim, dtype=np.complex128)
    for gt in gates:
        if gt[0] in ("H", "S"):
            P = H if gt[0] == "H" else S
            G = np.ones((1, 1), np.complex128)
            for j in range(1, N + 1):
                G = np.kron(G, P if j == gt[1] else np.eye(2))
        else:
            G = np.zeros((dim, dim), np.complex128)
            for k in range(dim):
                if (k >> (gt[1] - 1)) & 1:
                    G[k ^ (1 << (gt[2] - 1)), k] = 1
                else:
                    G[k, k] = 1
        U = G @ U
    return U


def global_clifford_snapshot(state, N, gates, r):
    """global_clifford_shadow_snapshot (…ClassicalShadows.jl:884-921): U ψ (or
    the diagonal of U ρ U'), then sample_from_probabilities; 0-based index."""
    U = global_clifford_unitary(N, gates)
    st = np.asarray(state, np.complex128)
    if st.ndim == 1:
        p = np.abs(U @ st) ** 2
    else:
        p = np.real(np.diag(U @ st @ U.conj().T))
    return sample_from_probabilities(p, r)


# ---------------------------------------------------------- preparation --
def product_ket(kets):
    """make_product_ket (utils/cpu/cpuQuantumStatePreparation.jl:1182-1191):
    ψ = ket_1, then tensor_product_ket(ψ, ket_k, k-1, 1) (:1108-1123), i.e.
    ψ_total[i_A + (i_B << N_A)] = ψ_A[i_A] ψ_B[i_B]."""
    kets = [np.asarray(k, np.complex128) for k in kets]
    psi = kets[0].copy()
    for k in kets[1:]:
        NA = psi.size
        out = np.empty(NA * 2, np.complex128)
        for b in range(2):
            out[b * NA:(b + 1) * NA] = psi * k[b]
        psi = out
    return psi


# ------------------------------------------------------ dense references --
def dense_1q(U, q: int, N: int):
    """Full 2^N matrix of a 1-qubit gate on 0-based qubit q (little-endian)."""
    mats = [np.eye(2)] * N
    mats = list(mats)
    mats[q] = np.asarray(U, np.complex128)
    M = np.array([[1.0 + 0j]])
    for b in range(N - 1, -1, -1):
        M = np.kron(M, mats[b])
    return M


def pauli_matrix(xmask: int, zmask: int, N: int):
    """Dense i^{|x&z|} X^x Z^z (0-based bit masks)."""
    X = np.array([[0, 1], [1, 0]], np.complex128)
    Z = np.array([[1, 0], [0, -1]], np.complex128)
    I = np.eye(2, dtype=np.complex128)
    M = np.array([[1.0 + 0j]])
    for b in range(N - 1, -1, -1):
        m = I
        if (xmask >> b) & 1 and (zmask >> b) & 1:
            m = 1j * X @ Z
        elif (xmask >> b) & 1:
            m = X
        elif (zmask >> b) & 1:
            m = Z
        M = np.kron(M, m)
    return M


# ------------------------------------------------ circuit executor (VQC) --
# The op records read here have the fields of the reference's
# CircuitOperation (type, qubits (1-based), param_idx, fixed_angle, p,
# target_state, basis, if_outcome), as attributes.
_EXEC_GATES = {"rx", "ry", "rz", "h", "x", "y", "z", "s", "t", "cz", "cnot", "cx", "rxx", "ryy",
               "rzz"}
_EXEC_NOISE = {"depolarizing", "amplitude_damping", "dephasing", "bit_flip", "phase_damping",
               "thermal_relaxation"}
_EXEC_MEASURE = {"measure_z", "measure_x", "measure_y", "measure"}


def _exec_angle(op, theta):
    return float(theta[op.param_idx - 1]) if op.param_idx > 0 else float(op.fixed_angle)


def measure_basis(psi, k, basis, N, u):
    """One qubit of projective_measurement!(ψ, qubits, basis::Symbol)
    (utils/cpu/cpuQuantumStateMeasurements.jl:319-345): :z directly, :x as
    H, Z-measure, H, :y as S†, H, Z-measure, H, S."""
    if basis == "z":
        return measure_z(psi, k, N, u)
    if basis == "x":
        apply_named(psi, "h", [k], 0.0, N)
        o = measure_z(psi, k, N, u)
        apply_named(psi, "h", [k], 0.0, N)
        return o
    if basis == "y":
        apply_named(psi, "sdg", [k], 0.0, N)
        apply_named(psi, "h", [k], 0.0, N)
        o = measure_z(psi, k, N, u)
        apply_named(psi, "h", [k], 0.0, N)
        apply_named(psi, "s", [k], 0.0, N)
        return o
    raise ValueError(basis)


def circuit_psi(psi, ops, theta, N, uniforms):
    """apply_circuit!(ψ, circuit, θ) (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl:95-117)
    with apply_operation! (:124-140): an op with if_outcome = k > 0 runs only
    when outcomes[k] == 1; gates by apply_gate_psi! (:146-181); noise by
    apply_noise_mcwf! (:187-204: phase_damping as dephasing, thermal
    relaxation nothing) with one uniform per listed qubit; measure by
    projective_measurement! with op.basis, outcomes appended; reset by
    reset_state! (one uniform per qubit); trace_out and lindblad_step do
    nothing.  `uniforms` is read in op order.  Returns (outcomes, used)."""
    outcomes, cur = [], 0
    for op in ops:
        if op.if_outcome > 0 and (op.if_outcome > len(outcomes)
                                  or outcomes[op.if_outcome - 1] != 1):
            continue
        t, qs = op.type, list(op.qubits)
        if t in _EXEC_GATES:
            apply_named(psi, "cnot" if t == "cx" else t, qs, _exec_angle(op, theta), N)
        elif t in _EXEC_NOISE:
            kind = {"phase_damping": "dephasing"}.get(t, t)
            if kind == "thermal_relaxation":
                continue
            for q in qs:
                psi_channel(psi, kind, op.p, q, N, float(uniforms[cur]))
                cur += 1
        elif t in _EXEC_MEASURE:
            for q in qs:
                outcomes.append(measure_basis(psi, q, op.basis, N, float(uniforms[cur])))
                cur += 1
        elif t == "reset":
            for q in qs:
                measure_z(psi, q, N, float(uniforms[cur]))
                force_qubit(psi, q, op.target_state, N)
                cur += 1
        elif t in ("trace_out", "lindblad_step"):
            continue
        else:
            raise ValueError(t)
    return outcomes, cur


def gate_unitary(gate, qubits, theta, N):
    """Dense 2^N matrix of a named gate: apply_named on every basis vector."""
    dim = 1 << N
    U = np.zeros((dim, dim), np.complex128)
    for j in range(dim):
        e = np.zeros(dim, np.complex128)
        e[j] = 1.0
        U[:, j] = apply_named(e, gate, qubits, theta, N)
    return U


def _pauli_on(P, k, N):
    return dense_1q(np.asarray(P, np.complex128), k - 1, N)


def circuit_rho(rho, ops, theta, N):
    """apply_circuit!(ρ, circuit, θ) (…Executor.jl:252-262), no if_outcome
    check: gates U ρ U† for the branches of apply_gate_rho! (:290-318; y, s
    and t have none); noise per qubit (:327-343): depolarizing (1-p)ρ +
    (p/3)(XρX + (XZ)ρ(XZ)† + ZρZ) (:346-368), bit flip (1-p)ρ + pXρX
    (:371-375), dephasing (1-p/2)ρ + (p/2)ZρZ (:378-382), amplitude damping
    by apply_sigma_minus_dissipator_rho! (cpuQuantumChannelGates.jl:833-854);
    measure and reset zero the qubit's coherences (:387-427); lindblad_step
    nothing; trace_out is an unknown operation."""
    X = np.array([[0, 1], [1, 0]], np.complex128)
    Z = np.array([[1, 0], [0, -1]], np.complex128)
    idx = np.arange(1 << N)
    for op in ops:
        t, qs = op.type, list(op.qubits)
        if t in _EXEC_GATES:
            if t in ("y", "s", "t"):
                continue
            U = gate_unitary("cnot" if t == "cx" else t, qs, _exec_angle(op, theta), N)
            rho[:, :] = U @ rho @ U.conj().T
        elif t in _EXEC_NOISE:
            for k in qs:
                p = op.p
                if t == "depolarizing":
                    Xk, Zk = _pauli_on(X, k, N), _pauli_on(Z, k, N)
                    XZ = Xk @ Zk
                    rho[:, :] = ((1 - p) * rho + (p / 3) * (Xk @ rho @ Xk)
                                 + (p / 3) * (XZ @ rho @ XZ.conj().T) + (p / 3) * (Zk @ rho @ Zk))
                elif t == "bit_flip":
                    Xk = _pauli_on(X, k, N)
                    rho[:, :] = (1 - p) * rho + p * (Xk @ rho @ Xk)
                elif t == "dephasing":
                    Zk = _pauli_on(Z, k, N)
                    rho[:, :] = (1 - p / 2) * rho + (p / 2) * (Zk @ rho @ Zk)
                elif t == "amplitude_damping":
                    m = 1 << (k - 1)
                    for j in range(1 << N):
                        for i in range(1 << N):
                            bi, bj = (i >> (k - 1)) & 1, (j >> (k - 1)) & 1
                            if bi == 1 and bj == 1:
                                rho[i ^ m, j ^ m] += p * rho[i, j]
                                rho[i, j] *= 1 - p
                            elif bi == 1 or bj == 1:
                                rho[i, j] *= 1 - 0.5 * p
        elif t in _EXEC_MEASURE or t == "reset":
            for k in qs:
                b = (idx >> (k - 1)) & 1
                rho[b[:, None] != b[None, :]] = 0
        elif t == "lindblad_step":
            continue
        else:
            raise ValueError(f"Unknown operation type: {t}")
    return rho
