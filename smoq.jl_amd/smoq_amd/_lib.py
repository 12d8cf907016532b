"""ctypes binding of libsmoq_amd.so (the C ABI declared in include/smoq.h).

The library is loaded from ``smoq.jl_amd/lib/libsmoq_amd.so`` inside this
repository.  There is no CPU fallback: if the library is missing or no
gfx950 device is visible, every call raises ``SmoqError``.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SMOQ_LIB") or os.path.normpath(
    os.path.join(_PKG, "..", "lib", "libsmoq_amd.so"))

SMOQ_OK = 0
SMOQ_EINVAL = -1
SMOQ_ENOMEM = -2
SMOQ_EDEVICE = -3
SMOQ_ENUMERIC = -4
SMOQ_ENOTSUP = -5

C128 = 0
C64 = 1

# smoq_gate_id
G = dict(
    U1=0, U2=1, H=2, X=3, Y=4, Z=5, S=6, SDG=7, T=8, TDG=9, RX=10, RY=11, RZ=12,
    CNOT=13, CZ=14, CRY=15, RXX=16, RYY=17, RZZ=18, CCZ=19, TOFFOLI=20,
    SIGMA_MINUS=21, SIGMA_PLUS=22, DIAG1=23, DIAG2=24,
)
ROTATIONS = {G["RX"], G["RY"], G["RZ"], G["CRY"], G["RXX"], G["RYY"], G["RZZ"]}

SAMPLE_CUMSUM_SEARCH = 0
SAMPLE_SEQUENTIAL = 1
SAMPLE_SEQUENTIAL_DIAG = 2

# smoq_qubit_channel_kind
QC = dict(measure_z=0, dephasing=1, depolarizing=2, bit_flip=3, phase_flip=4,
          amplitude_damping=5, force_zero=6, force_one=7)


class SmoqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"smoq error {code}: {msg}")
        self.code = code


class smoq_op(ctypes.Structure):
    _fields_ = [
        ("gate", ctypes.c_int32),
        ("q", ctypes.c_int32 * 3),
        ("param", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("theta", ctypes.c_double),
        ("m", ctypes.c_double * 32),
    ]


# smoq_circuit_kind: the reference's CircuitOperation types
CK = dict(rx=0, ry=1, rz=2, h=3, x=4, y=5, z=6, s=7, t=8, cz=9, cnot=10, cx=10, rxx=11,
          ryy=12, rzz=13, depolarizing=16, amplitude_damping=17, dephasing=18, bit_flip=19,
          phase_damping=20, thermal_relaxation=21, measure=32, reset=33, trace_out=34,
          lindblad_step=35)


class smoq_circuit_op(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("param", ctypes.c_int32),
        ("q_first", ctypes.c_int32),
        ("q_count", ctypes.c_int32),
        ("if_outcome", ctypes.c_int32),
        ("arg", ctypes.c_int32),
        ("angle", ctypes.c_double),
        ("p", ctypes.c_double),
    ]


class smoq_vqe_item(ctypes.Structure):
    _fields_ = [
        ("n_qubits", ctypes.c_int32),
        ("n_ops", ctypes.c_int32),
        ("ops", ctypes.c_void_p),
        ("n_params", ctypes.c_int32),
        ("n_terms", ctypes.c_int32),
        ("theta", ctypes.c_void_p),
        ("xmask", ctypes.c_void_p),
        ("zmask", ctypes.c_void_p),
        ("coeff", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
        ("grad", ctypes.c_void_p),
    ]


class smoq_tune_result(ctypes.Structure):
    _fields_ = [
        ("tile_qubits", ctypes.c_int),
        ("low_qubits", ctypes.c_int),
        ("pass_max_gates", ctypes.c_int),
        ("lane_swaps", ctypes.c_int),
        ("searched", ctypes.c_int),
        ("n_passes", ctypes.c_int),
        ("candidates", ctypes.c_int),
        ("ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("io_passes", ctypes.c_int),
        ("io_ms", ctypes.c_double),
    ]


class smoq_mcwf_group(ctypes.Structure):
    _fields_ = [
        ("psi", ctypes.c_void_p),
        ("n_bonds", ctypes.c_int),
        ("bond_q", ctypes.c_void_p),
        ("bond_diag", ctypes.c_void_p),
        ("site_u", ctypes.c_void_p),
        ("decay_pow", ctypes.c_void_p),
        ("uniforms", ctypes.c_void_p),
        ("jumps", ctypes.c_void_p),
        ("z_out", ctypes.c_void_p),
    ]


_P = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> argtypes; every function returns int status unless listed in _RET
SIGNATURES = {
    "smoq_last_error": [],
    "smoq_version": [],
    "smoq_init": [ctypes.POINTER(_P), _i],
    "smoq_fini": [_P],
    "smoq_sync": [_P],
    "smoq_set_fusion": [_P, _i, _i],
    "smoq_pass_timing": [_P, _i, ctypes.POINTER(_d), ctypes.POINTER(_i64)],
    "smoq_state_alloc": [_P, _i, _i, _i, ctypes.POINTER(_P)],
    "smoq_state_free": [_P],
    "smoq_state_info": [_P, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                        ctypes.POINTER(_P)],
    "smoq_state_set_basis": [_P, _u64],
    "smoq_state_set_product": [_P, _P],
    "smoq_state_upload": [_P, _P, _sz],
    "smoq_state_download": [_P, _P, _sz],
    "smoq_state_copy": [_P, _P],
    "smoq_apply_1q": [_P, _i, _P],
    "smoq_apply_2q": [_P, _i, _i, _P],
    "smoq_apply_named": [_P, _i, _P, _d],
    "smoq_apply_tape": [_P, _P, _i, _P, _i, ctypes.POINTER(_i)],
    "smoq_expect_pauli": [_P, _i, _P, _P, _P],
    "smoq_pauli_string_reference": [_P, _i, _P, _P, _P],
    "smoq_plan_create": [_P, _i, _i, _P, _i, _P, _i, ctypes.POINTER(_P)],
    "smoq_plan_passes": [_P, ctypes.POINTER(_i)],
    "smoq_plan_run": [_P, _P, _i],
    "smoq_plan_free": [_P],
    "smoq_evolve_observe": [_P, _P, _i, _i, _i, _P, _P, _P],
    "smoq_hpsi": [_P, _P, _i, _P, _P, _P],
    "smoq_adjoint_gradient": [_P, _P, _i, _P, _i, _i, _P, _P, _P, _P, _P],
    "smoq_tape_energy": [_P, _P, _i, _P, _i, _i, _P, _P, _P, _i, _P],
    "smoq_chebyshev_evolve": [_P, _i, _P, _P, _P, _d, _d, _d, _d, _i, ctypes.POINTER(_i)],
    "smoq_lanczos_range": [_P, _i, _P, _P, _P, _i, _P, _P, _P],
    "smoq_dm_expect_pauli": [_P, _i, _P, _P, _P],
    "smoq_dm_apply_dense": [_P, _P],
    "smoq_qubit_channel": [_P, _i, _P, _i, _d, _P, _P],
    "smoq_dm_qubit_channel": [_P, _i, _P, _i, _P, _P],
    "smoq_partial_trace": [_P, ctypes.c_uint64, _P],
    "smoq_dm_partial_trace": [_P, ctypes.c_uint64, _P],
    "smoq_shadow_snapshots": [_P, _i64, _i, _P, _P, _P, _P],
    "smoq_global_clifford_snapshots": [_P, _i64, _P, _P, _P, _P],
    "smoq_norm2": [_P, _P],
    "smoq_expect_pauli_noalloc": [_P, _i, _P, _P, _P],
    "smoq_autotune": [_P, _P, _i, _P, _i, _P],
    "smoq_adjoint_gradient_batch": [_P, _i, _P, _i],
    "smoq_apply_circuit": [_P, _i, _P, _i, _P, _i, _P, _i, _P, ctypes.c_int64, _P, _i, _P, _P],
    "smoq_tune_save": [_P, ctypes.c_char_p],
    "smoq_tune_load": [_P, ctypes.c_char_p, _P],
    "smoq_sample": [_P, _P, _i64, _i, _P],
    "smoq_mcwf_evolve": [_P, _P, _i, _d, _i, _i],
    "smoq_set_jit": [_P, _i],
    "smoq_set_lane_swaps": [_P, _i],
    "smoq_set_obs_jit": [_P, _i],
    "smoq_mcwf_jump": [_P, _i, _d, _P, _P, _P],
    "smoq_mcwf_trotter_evolve": [_P, _P, _i, _d, _P, _i, _P, _P, _i, _i, _P, _P, _P],
}
_RET = {"smoq_last_error": ctypes.c_char_p, "smoq_version": _i}

_lib = None


def lib():
    """Loads the HIP library (raises SmoqError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SmoqError(SMOQ_ENOTSUP,
                            f"{LIB_PATH} is not built; run `make -C smoq.jl_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RET.get(name, _i)
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SMOQ_OK:
        msg = lib().smoq_last_error()
        raise SmoqError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
