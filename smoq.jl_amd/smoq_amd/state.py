"""Device context and state vectors (GPUState) over the C ABI."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import C64, C128, call

_DTYPES = {"c128": C128, "complex128": C128, C128: C128, np.complex128: C128,
           "c64": C64, "complex64": C64, C64: C64, np.complex64: C64}
_NP = {C128: np.complex128, C64: np.complex64}


class Context:
    """One device and one HIP stream (smoq_init / smoq_fini)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        call("smoq_init", ctypes.byref(h), device)
        self.handle = h
        self.device = device

    def sync(self) -> None:
        call("smoq_sync", self.handle)

    def set_fusion(self, tile_qubits: int = 0, low_qubits: int = 0) -> None:
        call("smoq_set_fusion", self.handle, tile_qubits, low_qubits)

    def set_jit(self, mode: int) -> None:
        """Run-time compiled fused passes: 0 never, 1 automatic, 2 always."""
        call("smoq_set_jit", self.handle, int(mode))

    def set_obs_jit(self, mode: int) -> None:
        """Pauli reductions compiled for the term list: 0 off, 1 automatic
        (large states), 2 for every state of at least 12 qubits."""
        call("smoq_set_obs_jit", self.handle, int(mode))

    def set_lane_swaps(self, on: bool) -> None:
        """Stages of compiled passes trading register and lane qubits."""
        call("smoq_set_lane_swaps", self.handle, int(bool(on)))

    def pass_timing(self, enable: bool):
        """Returns (mean fused-pass kernel ms, launches) since the last call."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        call("smoq_pass_timing", self.handle, int(enable), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def autotune(self, tape, psi, theta=None) -> dict:
        """smoq_autotune: times candidate fusion settings for the tape on psi
        (overwritten) and keeps the fastest for its structure; returns the
        chosen settings and timings."""
        from ._lib import smoq_tune_result

        res = smoq_tune_result()
        if theta is None:
            th_p, nth = None, 0
        else:
            th = np.ascontiguousarray(theta, np.float64)
            th_p, nth = th.ctypes.data_as(ctypes.c_void_p), th.size
        call("smoq_autotune", psi.handle, tape.to_ctypes(), len(tape.ops), th_p, nth,
             ctypes.byref(res))
        return {f: getattr(res, f) for f, _ in res._fields_}

    def tune_save(self, path: str) -> None:
        call("smoq_tune_save", self.handle, str(path).encode())

    def tune_load(self, path: str) -> int:
        n = ctypes.c_int(0)
        call("smoq_tune_load", self.handle, str(path).encode(), ctypes.byref(n))
        return n.value

    def close(self) -> None:
        if self.handle:
            call("smoq_fini", self.handle)
            self.handle = None


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class GPUState:
    """`batch` state vectors of n qubits resident in device memory.

    Mirrors the reference's `Vector{ComplexF64}` (batch = 1) and its
    dim x n_traj trajectory matrices (batch > 1, one item per column).
    """

    def __init__(self, n_qubits: int, dtype="c128", batch: int = 1,
                 ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.dtype = _DTYPES[dtype]
        self.n = int(n_qubits)
        self.batch = int(batch)
        h = ctypes.c_void_p()
        call("smoq_state_alloc", self.ctx.handle, self.n, self.dtype, self.batch,
             ctypes.byref(h))
        self.handle = h

    # ---- construction helpers
    @classmethod
    def from_numpy(cls, psi: np.ndarray, dtype=None, ctx: Context | None = None) -> "GPUState":
        psi = np.asarray(psi)
        if dtype is None:
            dtype = "c64" if psi.dtype == np.complex64 else "c128"
        dt = _DTYPES[dtype]
        arr = psi.reshape(-1, psi.shape[-1]) if psi.ndim > 1 else psi.reshape(1, -1)
        dim = arr.shape[1]
        n = dim.bit_length() - 1
        if 1 << n != dim:
            raise ValueError(f"length {dim} is not a power of two")
        s = cls(n, dt, arr.shape[0], ctx)
        s.upload(arr)
        return s

    @classmethod
    def basis(cls, n_qubits: int, index: int = 0, dtype="c128", batch: int = 1,
              ctx: Context | None = None) -> "GPUState":
        s = cls(n_qubits, dtype, batch, ctx)
        s.set_basis(index)
        return s

    # ---- properties
    @property
    def N(self) -> int:
        return self.n

    @property
    def dim(self) -> int:
        return 1 << self.n

    @property
    def np_dtype(self):
        return _NP[self.dtype]

    @property
    def nbytes(self) -> int:
        return self.dim * self.batch * np.dtype(self.np_dtype).itemsize

    # ---- data movement
    def set_basis(self, index: int) -> None:
        call("smoq_state_set_basis", self.handle, ctypes.c_uint64(index))

    def set_product(self, kets) -> None:
        """Every batch item becomes the product state of the per-qubit kets
        (kets[q] = (<0|q>, <1|q>), qubit q is bit q): smoq_state_set_product."""
        k = np.asarray(kets, np.complex128).reshape(-1, 2)
        if k.shape[0] != self.n:
            raise ValueError(f"{k.shape[0]} kets for a {self.n}-qubit state")
        buf = np.ascontiguousarray(k.view(np.float64).reshape(-1))
        call("smoq_state_set_product", self.handle, buf.ctypes.data_as(ctypes.c_void_p))

    def upload(self, psi: np.ndarray) -> None:
        a = np.ascontiguousarray(np.asarray(psi, dtype=self.np_dtype)).reshape(-1)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte state")
        call("smoq_state_upload", self.handle, a.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_size_t(a.nbytes))

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.dim * self.batch, dtype=self.np_dtype)
        call("smoq_state_download", self.handle, out.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_size_t(out.nbytes))
        return out if self.batch == 1 else out.reshape(self.batch, self.dim)

    def copy_from(self, other: "GPUState") -> None:
        call("smoq_state_copy", self.handle, other.handle)

    def clone(self) -> "GPUState":
        s = GPUState(self.n, self.dtype, self.batch, self.ctx)
        s.copy_from(self)
        return s

    def norm2(self) -> np.ndarray:
        out = np.empty(self.batch, np.float64)
        call("smoq_norm2", self.handle, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def free(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().smoq_state_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_R2 = 1.0 / np.sqrt(2.0)
# make_ket(::Symbol) and _char_to_state_symbol
# (utils/cpu/cpuQuantumStatePreparation.jl:133-147,250-262)
KETS = {"zero": (1.0, 0.0), "one": (0.0, 1.0), "plus": (_R2, _R2), "minus": (_R2, -_R2)}
_CHARS = {"0": "zero", "1": "one", "+": "plus", "-": "minus"}


def make_product_ket(states, dtype="c128", batch: int = 1, ctx: Context | None = None) -> GPUState:
    """make_product_ket (cpuQuantumStatePreparation.jl:1182-1191) on the device:
    states[k] in {"zero", "one", "plus", "minus"} (or the characters 0 1 + -,
    as make_ket(::String)) for qubit k+1 = bit k."""
    if isinstance(states, str):
        clean = states.replace("|", "").replace(">", "").replace("\u27e9", "").strip()
        states = [_CHARS[c] if c in _CHARS else c for c in clean]
    names = [str(s).lstrip(":") for s in states]
    bad = [s for s in names if s not in KETS]
    if not names or bad:
        raise ValueError(f"unknown state(s) {bad or names}: use zero, one, plus, minus")
    psi = GPUState(len(names), dtype, batch, ctx=ctx)
    psi.set_product([KETS[s] for s in names])
    return psi


def make_initial_psi(N: int, init_bits=(), dtype="c128", batch: int = 1,
                     ctx: Context | None = None) -> GPUState:
    """make_initial_psi (cpuQuantumStatePreparation.jl:1211-1229): the basis
    state with qubit k in init_bits[k-1] (little-endian), |0…0> when empty."""
    bits = list(init_bits)
    if bits and len(bits) != N:
        raise ValueError("init_bits must have length N")
    idx = sum(1 << k for k, b in enumerate(bits) if b == 1)
    return GPUState.basis(N, idx, dtype, batch, ctx=ctx)
