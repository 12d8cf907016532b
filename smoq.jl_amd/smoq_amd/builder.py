"""The variational-circuit IR with the reference's names and fields
(utils/cpu/cpuVariationalQuantumCircuitBuilder.jl): CircuitOperation,
CircuitLayer, ParameterizedCircuit, the constructors gate / noise / measure /
reset / trace_out / lindblad_step, the layer builders and the ansätze.

Operation types are the reference's symbols as strings ("ry", "cz",
"depolarizing", "measure_z", …); a leading ':' is accepted and dropped.
Qubits are 1-based and param_idx counts from 1 (0 = fixed angle), exactly as
in the reference, so a circuit written for it reads the same here.
executor.apply_circuit runs one on a GPUState or DensityMatrix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

# builder :80-95
GATE_TYPES = frozenset({"rx", "ry", "rz", "h", "x", "y", "z", "s", "t",
                        "cz", "cnot", "cx", "rxx", "ryy", "rzz"})
NOISE_TYPES = frozenset({"depolarizing", "amplitude_damping", "dephasing", "bit_flip",
                         "phase_damping", "thermal_relaxation"})
MEASURE_TYPES = frozenset({"measure_z", "measure_x", "measure_y", "measure"})
RESET_TYPES = frozenset({"reset"})
TRACE_TYPES = frozenset({"trace_out"})
ANALOG_TYPES = frozenset({"lindblad_step", "trotter_step"})


def _sym(s) -> str:
    return str(s).lstrip(":").lower()


def _qubits(qubits) -> list[int]:
    return [int(q) for q in np.atleast_1d(qubits)]


@dataclass(frozen=True)
class CircuitOperation:
    """One operation (builder :118-132): type, 1-based qubits, param_idx
    (0 = use fixed_angle, k > 0 = θ[k]), the noise probability p, the reset
    target, the measurement basis, if_outcome (0 = always, k > 0 = only when
    the k-th recorded outcome is 1), and the Lindblad block's H, L_ops, dt."""
    type: str
    qubits: tuple
    param_idx: int = 0
    fixed_angle: float = 0.0
    p: float = 0.0
    target_state: str = "none"
    basis: str = "z"
    if_outcome: int = 0
    H: object = None
    L_ops: tuple = ()
    dt: float = 0.0


@dataclass
class CircuitResult:
    """Final state and the measurement outcomes in execution order (builder
    :153-156); `uniforms_used` counts the uniforms the pure-state run drew."""
    state: object
    outcomes: list = field(default_factory=list)
    uniforms_used: int = 0


def gate(type, qubits, *, param_idx: int = 0, angle: float = 0.0, if_outcome: int = 0):
    """gate(type, qubits; param_idx, angle, if_outcome) (builder :174-181)."""
    t = _sym(type)
    if t not in GATE_TYPES:
        raise ValueError(f"Unknown gate type: {t}")
    return CircuitOperation(t, tuple(_qubits(qubits)), int(param_idx), float(angle), 0.0,
                            "none", "z", int(if_outcome))


def noise(type, qubits, *, p: float = 0.0, if_outcome: int = 0):
    """noise(type, qubits; p) (builder :198-204)."""
    t = _sym(type)
    if t not in NOISE_TYPES:
        raise ValueError(f"Unknown noise type: {t}")
    return CircuitOperation(t, tuple(_qubits(qubits)), 0, 0.0, float(p), "none", "z",
                            int(if_outcome))


def measure(qubits, *, basis="z"):
    """measure(qubits; basis) (builder :219-224): type measure_<basis>."""
    b = _sym(basis)
    return CircuitOperation("measure_" + b, tuple(_qubits(qubits)), basis=b)


def reset(qubits, *, target_state="zero", if_outcome: int = 0):
    """reset(qubits; target_state, if_outcome) (builder :240-244)."""
    return CircuitOperation("reset", tuple(_qubits(qubits)), target_state=_sym(target_state),
                            if_outcome=int(if_outcome))


def lindblad_step(qubits, *, H=None, L_ops=(), dt: float = 0.0):
    """lindblad_step(qubits; H, L_ops, dt) (builder :259-263)."""
    return CircuitOperation("lindblad_step", tuple(_qubits(qubits)), H=H, L_ops=tuple(L_ops),
                            dt=float(dt))


def trace_out(qubits):
    """trace_out(qubits) (builder :279-283)."""
    return CircuitOperation("trace_out", tuple(_qubits(qubits)))


@dataclass
class CircuitLayer:
    """Operations on disjoint qubits (builder :293-295)."""
    operations: list


@dataclass
class ParameterizedCircuit:
    """N qubits, the flat operation list, optional layers and n_params = the
    largest param_idx (builder :311-346)."""
    N: int
    operations: list
    layers: list = field(default_factory=list)
    n_params: int = 0

    @classmethod
    def from_operations(cls, N: int, operations) -> "ParameterizedCircuit":
        ops = list(operations)
        return cls(int(N), ops, [], max([op.param_idx for op in ops if op.param_idx > 0],
                                        default=0))

    @classmethod
    def from_layers(cls, N: int, layers) -> "ParameterizedCircuit":
        layers = list(layers)
        ops = [op for layer in layers for op in layer.operations]
        return cls(int(N), ops, layers,
                   max([op.param_idx for op in ops if op.param_idx > 0], default=0))


def single_qubit_rotation_layer(N: int, param_offset: int, *, gates=("ry",)):
    """(layer, next param index): each listed rotation on qubits 1..N in turn,
    one fresh parameter per gate (builder :356-369)."""
    ops, idx = [], int(param_offset)
    for g in gates:
        for q in range(1, N + 1):
            ops.append(gate(g, [q], param_idx=idx))
            idx += 1
    return CircuitLayer(ops), idx


def entangling_layer(N: int, *, topology="chain", gate_type="cz"):
    """Chain (q, q+1), ring (chain plus (N, 1) for N > 2) or all pairs
    (builder :376-396)."""
    topo = _sym(topology)
    if topo == "chain":
        pairs = [(q, q + 1) for q in range(1, N)]
    elif topo == "ring":
        pairs = [(q, q + 1) for q in range(1, N)] + ([(N, 1)] if N > 2 else [])
    elif topo == "all":
        pairs = [(i, j) for i in range(1, N + 1) for j in range(i + 1, N + 1)]
    else:
        raise ValueError(f"Unknown topology: {topo}")
    return CircuitLayer([gate(gate_type, [a, b]) for a, b in pairs])


def noise_layer(N: int, noise_type, *, p: float = 0.0):
    """The noise channel on every qubit (builder :403-406)."""
    return CircuitLayer([noise(noise_type, [q], p=p) for q in range(1, N + 1)])


def hardware_efficient_ansatz(N: int, n_layers: int, *, entangler="cz", topology="chain",
                              rotations=("ry",), noise_type=None, noise_p: float = 0.0):
    """Per layer: rotations, optional noise on every qubit, entanglers (N > 1);
    then a final rotation layer (builder :413-437)."""
    layers, idx = [], 1
    for _ in range(n_layers):
        rot, idx = single_qubit_rotation_layer(N, idx, gates=rotations)
        layers.append(rot)
        if noise_type is not None:
            layers.append(noise_layer(N, noise_type, p=noise_p))
        if N > 1:
            layers.append(entangling_layer(N, topology=topology, gate_type=entangler))
    rot, idx = single_qubit_rotation_layer(N, idx, gates=rotations)
    layers.append(rot)
    return ParameterizedCircuit.from_layers(N, layers)


def strong_entangling_ansatz(N: int, n_layers: int, *, topology="ring", noise_type=None,
                             noise_p: float = 0.0):
    """Rx, Ry, Rz rotations with CZ entanglers (builder :444-454)."""
    return hardware_efficient_ansatz(N, n_layers, entangler="cz", topology=topology,
                                     rotations=("rx", "ry", "rz"), noise_type=noise_type,
                                     noise_p=noise_p)


def get_parameter_count(circuit: ParameterizedCircuit) -> int:
    return circuit.n_params


def get_parameter_info(circuit: ParameterizedCircuit) -> list[dict]:
    """(param_idx, gate_type, qubits, op_index) per parametrised op, sorted by
    param_idx; op_index is 1-based (builder :474-485)."""
    info = [dict(param_idx=op.param_idx, gate_type=op.type, qubits=list(op.qubits),
                 op_index=i + 1)
            for i, op in enumerate(circuit.operations) if op.param_idx > 0]
    return sorted(info, key=lambda d: d["param_idx"])


def initialize_parameters(circuit: ParameterizedCircuit, *, init="zeros", rng=None):
    """zeros, random (uniform in [-π, π)), small_random ([-0.1, 0.1)) or a
    constant (builder :499-513)."""
    n = circuit.n_params
    if isinstance(init, (int, float)) and not isinstance(init, bool):
        return np.full(n, float(init))
    g = rng if rng is not None else np.random.default_rng()
    kind = _sym(init)
    if kind == "zeros":
        return np.zeros(n)
    if kind == "random":
        return 2 * math.pi * g.random(n) - math.pi
    if kind == "small_random":
        return 0.2 * g.random(n) - 0.1
    raise ValueError(f"Unknown init: {kind}")


def extract_gate_angles(circuit: ParameterizedCircuit, theta) -> dict:
    """1-based op index → angle for parametrised ops and nonzero fixed angles
    (builder :527-541)."""
    out = {}
    for i, op in enumerate(circuit.operations):
        if op.param_idx > 0:
            out[i + 1] = float(theta[op.param_idx - 1])
        elif op.fixed_angle != 0.0:
            out[i + 1] = op.fixed_angle
    return out


def count_operations(circuit: ParameterizedCircuit) -> dict:
    counts: dict = {}
    for op in circuit.operations:
        counts[op.type] = counts.get(op.type, 0) + 1
    return counts


def describe_circuit(circuit: ParameterizedCircuit) -> str:
    """The summary the reference prints (builder :569-583), returned as text."""
    c = count_operations(circuit)
    lines = [f"ParameterizedCircuit: {circuit.N} qubits, {len(circuit.operations)} operations, "
             f"{circuit.n_params} parameters", "-" * 60]
    for label, kinds in (("Gates", GATE_TYPES), ("Noise", NOISE_TYPES),
                         ("Meas/Reset", MEASURE_TYPES | RESET_TYPES)):
        part = [f"{k}×{v}" for k, v in c.items() if k in kinds]
        if part:
            lines.append(f"  {label}: " + ", ".join(part))
    return "\n".join(lines)


def repeat_layer(layer_fn, L: int, N: int, *, param_offset: int = 1):
    """(operations, next param index) of L calls layer_fn(N, offset) ->
    (ops, next) (builder :614-624)."""
    ops, idx = [], int(param_offset)
    for _ in range(L):
        part, idx = layer_fn(N, idx)
        ops.extend(part.operations if isinstance(part, CircuitLayer) else part)
    return ops, idx


def repeat_operations(ops, L: int, *, renumber_params: bool = True) -> list:
    """ops repeated L times; with renumbering each copy's param_idx > 0 is
    shifted by the largest index of one copy (builder :647-674)."""
    ops = list(ops)
    if not renumber_params:
        return ops * L
    step = max([op.param_idx for op in ops], default=0)
    out = []
    for r in range(L):
        for op in ops:
            out.append(replace(op, param_idx=op.param_idx + r * step) if op.param_idx > 0 else op)
    return out


__all__ = [
    "ANALOG_TYPES", "CircuitLayer", "CircuitOperation", "CircuitResult", "GATE_TYPES",
    "MEASURE_TYPES", "NOISE_TYPES", "ParameterizedCircuit", "RESET_TYPES", "TRACE_TYPES",
    "count_operations", "describe_circuit", "entangling_layer", "extract_gate_angles", "gate",
    "get_parameter_count", "get_parameter_info", "hardware_efficient_ansatz",
    "initialize_parameters", "lindblad_step", "measure", "noise", "noise_layer", "repeat_layer",
    "repeat_operations", "reset", "single_qubit_rotation_layer", "strong_entangling_ansatz",
    "trace_out",
]
