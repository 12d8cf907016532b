"""apply_circuit: a ParameterizedCircuit run on the device, with the
reference's dispatch (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl).

The circuit's CircuitOperation list becomes smoq_circuit_op records and one
smoq_apply_circuit call, which runs it natively:

* pure states (GPUState; apply_circuit!(ψ, …), executor :95-117): each
  maximal run of gates is one set of fused passes; noise channels,
  measurements and resets between runs are device calls whose uniforms
  replace the reference's `rand()` (one per listed qubit and item, in
  operation order); `if_outcome` is decided when the operation is reached,
  from the outcomes measured so far, as in the reference;
* density matrices (DensityMatrix; apply_circuit!(ρ, …), executor :252-262):
  every operation is a fixed linear map of vec(ρ), so the whole circuit is
  one tape.

The reference's executor is followed where it differs from the standalone
channel modules: on ρ, dephasing is (1 - p/2) ρ + (p/2) ZρZ, amplitude
damping is the first-order σ⁻ dissipator step, phase damping and thermal
relaxation do nothing, the y / s / t gates are not applied (its
apply_gate_rho! has no branch for them), reset only decoheres the qubit, and
`if_outcome` is not consulted; on ψ, phase damping acts as dephasing and
thermal relaxation does nothing.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from ._lib import CK, call, smoq_circuit_op
from .builder import CircuitResult, MEASURE_TYPES, ParameterizedCircuit
from .density import DensityMatrix
from .state import GPUState

_BASIS = {"z": 0, "x": 1, "y": 2}
_TARGET = {"zero": 0, "one": 1, "plus": 2, "minus": 3}


def lower_circuit(circuit: ParameterizedCircuit):
    """(records, qubit array, measured-qubit count): the smoq_circuit_op
    array of the circuit with its 0-based qubit list."""
    ops = circuit.operations
    arr = (smoq_circuit_op * max(1, len(ops)))()
    qubits: list[int] = []
    n_meas = 0
    for i, op in enumerate(ops):
        t = op.type
        kind = CK["measure"] if t in MEASURE_TYPES else CK.get(t)
        if kind is None:
            raise ValueError(f"Unknown operation type: {t}")
        r = arr[i]
        r.kind = kind
        r.param = op.param_idx - 1 if op.param_idx > 0 else -1
        r.q_first = len(qubits)
        r.q_count = len(op.qubits)
        r.if_outcome = int(op.if_outcome)
        if t in MEASURE_TYPES:
            b = str(op.basis).lstrip(":").lower()
            if b not in _BASIS:
                raise ValueError(f"Unknown basis: :{b}. Use :z, :x, :y.")
            r.arg = _BASIS[b]
            n_meas += len(op.qubits)
        elif t == "reset":
            tg = str(op.target_state).lstrip(":").lower()
            if tg not in _TARGET:
                raise ValueError(f"Unknown target state: :{tg}. Use :zero, :one, :plus, :minus.")
            r.arg = _TARGET[tg]
        r.angle = float(op.fixed_angle)
        r.p = float(op.p)
        qubits.extend(int(q) - 1 for q in op.qubits)
    q = np.ascontiguousarray(qubits if qubits else [0], np.int32)
    return arr, q, len(qubits), n_meas


def apply_circuit(state, circuit: ParameterizedCircuit, theta=None, *, uniforms=None,
                  rng=None) -> CircuitResult:
    """apply_circuit!(state, circuit, θ) (executor :95-117, :252-262).

    θ[k-1] is the reference's θ[k].  For a GPUState the noise, measurement
    and reset operations take their uniforms from `uniforms` (a flat array
    read in operation order: per operation, `batch` values for each listed
    qubit, qubit-major) or, without it, from `rng` (enough for every such
    operation); outcomes come back one per measured qubit (as a
    (batch, n_outcomes) array for a batch of trajectories, which admits no
    if_outcome).  For a DensityMatrix the map is exact and draws nothing."""
    dm = isinstance(state, DensityMatrix)
    if not dm and not isinstance(state, GPUState):
        raise TypeError(f"apply_circuit takes a GPUState or DensityMatrix, "
                        f"got {type(state).__name__}")
    n = state.N if dm else state.n
    if n != circuit.N:
        raise ValueError(f"State dimension mismatch: state has {n} qubits, circuit {circuit.N}")
    th = np.zeros(0) if theta is None else np.ascontiguousarray(theta, np.float64).reshape(-1)
    if th.size < circuit.n_params:
        raise ValueError(f"Not enough parameters: {th.size} < {circuit.n_params}")
    for op in circuit.operations:
        if op.type == "trace_out" and not dm:
            warnings.warn("trace_out in pure state mode: use partial_trace "
                          "(smoq_amd.partial_trace)", stacklevel=2)
        elif op.type == "lindblad_step":
            warnings.warn("Lindbladian step " + ("for DM " if dm else "")
                          + "not yet integrated - skipping", stacklevel=2)
    arr, q, nq, n_meas = lower_circuit(circuit)
    target = state.vec if dm else state
    B = target.batch
    if dm:
        u = np.zeros(1)
    elif uniforms is not None:
        u = np.ascontiguousarray(uniforms, np.float64).reshape(-1)
    else:
        need = sum(len(op.qubits) for op in circuit.operations
                   if op.type not in ("rx", "ry", "rz", "h", "x", "y", "z", "s", "t", "cz",
                                      "cnot", "cx", "rxx", "ryy", "rzz")) * B
        u = (rng if rng is not None else np.random.default_rng()).random(max(need, 1))
    out = np.zeros(max(1, n_meas * B), np.int32)
    n_out = ctypes.c_int(0)
    used = ctypes.c_int64(0)
    call("smoq_apply_circuit", target.handle, int(dm), arr, len(circuit.operations),
         q.ctypes.data_as(ctypes.c_void_p), nq,
         th.ctypes.data_as(ctypes.c_void_p) if th.size else None, th.size,
         u.ctypes.data_as(ctypes.c_void_p), 0 if dm else u.size,
         out.ctypes.data_as(ctypes.c_void_p), n_meas, ctypes.byref(n_out), ctypes.byref(used))
    k = n_out.value
    if B == 1:
        outcomes = [int(v) for v in out[:k]]
    else:
        outcomes = out[:k * B].reshape(k, B).T.astype(np.int64)
    return CircuitResult(state, outcomes, int(used.value))


def apply_circuit_pure(psi: GPUState, circuit: ParameterizedCircuit, theta=None, *,
                       uniforms=None, rng=None) -> CircuitResult:
    return apply_circuit(psi, circuit, theta, uniforms=uniforms, rng=rng)


def apply_circuit_dm(rho: DensityMatrix, circuit: ParameterizedCircuit, theta=None):
    return apply_circuit(rho, circuit, theta)


__all__ = ["apply_circuit", "apply_circuit_dm", "apply_circuit_pure", "lower_circuit"]
