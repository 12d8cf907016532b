"""Matrix-free H·ψ and Chebyshev propagation with the reference's names
(utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl), on the device."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import call
from .hamiltonian import HamiltonianParams, pauli_sum, spectral_bound


def _arrays(params: HamiltonianParams):
    xs, zs, cs = pauli_sum(params)
    xm = np.ascontiguousarray(xs, np.uint64)
    zm = np.ascontiguousarray(zs, np.uint64)
    c = np.ascontiguousarray(cs, np.float64)
    return xm, zm, c


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def apply_hamiltonian_bitwise(psi_out, psi_in, params: HamiltonianParams):
    """apply_hamiltonian_bitwise! (…Chebyshev.jl:55-152): psi_out = H psi_in."""
    xm, zm, c = _arrays(params)
    call("smoq_hpsi", psi_in.handle, psi_out.handle, xm.size, _p(xm), _p(zm), _p(c))
    return psi_out


def get_rescaling_factors(E_min: float, E_max: float, safety_margin: float = 0.01):
    """get_rescaling_factors (…Chebyshev.jl:232-243)."""
    width = E_max - E_min
    lo = E_min - safety_margin * abs(width)
    hi = E_max + safety_margin * abs(width)
    return (hi - lo) / 2.0, (hi + lo) / 2.0


def estimate_spectral_range_lanczos(params: HamiltonianParams, k_iter: int = 30, *, seed: int = 0,
                                    start=None, dtype: str = "c128", ctx=None):
    """estimate_spectral_range_lanczos (…Chebyshev.jl:166-220) on the device:
    k_iter Lanczos steps of H, then the extreme Ritz values (E_min, E_max).

    The reference starts from an unseeded randn vector; here the start is
    `start` (a GPUState, left unchanged) or, by default, a complex normal
    vector drawn from numpy's generator with `seed`."""
    from .state import GPUState

    own = start is None
    if own:
        rng = np.random.default_rng(seed)
        dim = 1 << params.N
        v = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
        start = GPUState.from_numpy(v.astype(np.complex64 if dtype == "c64" else np.complex128),
                                    ctx=ctx)
    xm, zm, c = _arrays(params)
    lo, hi, used = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    try:
        call("smoq_lanczos_range", start.handle, xm.size, _p(xm), _p(zm), _p(c), int(k_iter),
             ctypes.byref(lo), ctypes.byref(hi), ctypes.byref(used))
    finally:
        if own:
            start.free()
    return lo.value, hi.value


def chebyshev_evolve_psi(psi, params: HamiltonianParams, t: float, tol: float = 1e-12,
                         max_order: int = 1000, spectral_bounds=None):
    """chebyshev_evolve_psi! (…Chebyshev.jl:262-346): psi <- exp(-iHt) psi.

    Returns (M_used, a, b).  The reference estimates the spectral range with
    an unseeded Lanczos run when no bounds are given; here the default is the
    deterministic bound ±Σ|c_t|."""
    if spectral_bounds is None:
        B = spectral_bound(params)
        spectral_bounds = (-B, B)
    emin, emax = spectral_bounds
    xm, zm, c = _arrays(params)
    used = ctypes.c_int(0)
    call("smoq_chebyshev_evolve", psi.handle, xm.size, _p(xm), _p(zm), _p(c), float(t),
         float(emin), float(emax), float(tol), int(max_order), ctypes.byref(used))
    a, b = get_rescaling_factors(emin, emax)
    return used.value, a, b
