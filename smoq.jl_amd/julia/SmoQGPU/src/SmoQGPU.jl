# SmoQGPU: the ccall layer that binds SmoQ.jl's state-vector functions to
# libsmoq_amd.so (the C ABI of include/smoq.h).  Function names and argument
# order are the reference's (utils/cpu/*.jl); a device-resident GPUState takes
# the place of Vector{ComplexF64}.  Qubits are 1-based here and 0-based in the
# ABI; every non-zero status becomes error(...) with smoq_last_error().
module SmoQGPU

using LinearAlgebra: diag
using ChainRulesCore

const LIB = get(ENV, "SMOQ_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libsmoq_amd.so"))

check(rc::Integer) = rc == 0 ? nothing :
    error("smoq error $rc: " * unsafe_string(ccall((:smoq_last_error, LIB), Cstring, ())))

# ---- context and states --------------------------------------------------
mutable struct Context
    h::Ptr{Cvoid}
end
function Context(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smoq_init, LIB), Cint, (Ptr{Ptr{Cvoid}}, Cint), r, device))
    c = Context(r[])
    finalizer(c -> ccall((:smoq_fini, LIB), Cint, (Ptr{Cvoid},), c.h), c)
end
const CTX = Ref{Context}()
ctx() = isassigned(CTX) ? CTX[] : (CTX[] = Context(0))

const C128, C64 = Cint(0), Cint(1)

mutable struct GPUState            # replaces Vector{ComplexF64} / Matrix (dim × n_traj)
    h::Ptr{Cvoid}
    N::Int
    batch::Int
end
function GPUState(N::Integer; batch::Integer = 1, dtype = C128)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smoq_state_alloc, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Ptr{Cvoid}}),
                ctx().h, N, dtype, batch, r))
    s = GPUState(r[], N, batch)
    finalizer(s -> ccall((:smoq_state_free, LIB), Cint, (Ptr{Cvoid},), s.h), s)
end
function GPUState(ψ::AbstractVecOrMat{ComplexF64})
    dim = size(ψ, 1); N = trailing_zeros(dim)
    s = GPUState(N; batch = size(ψ, 2))
    check(ccall((:smoq_state_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Csize_t),
                s.h, ψ, sizeof(ψ)))
    s
end
function Base.Array(s::GPUState)
    out = Matrix{ComplexF64}(undef, 1 << s.N, s.batch)
    check(ccall((:smoq_state_download, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Csize_t),
                s.h, out, sizeof(out)))
    s.batch == 1 ? vec(out) : out
end
make_ket_gpu(N, index = 0) = (s = GPUState(N);
    check(ccall((:smoq_state_set_basis, LIB), Cint, (Ptr{Cvoid}, UInt64), s.h, index)); s)

# ---- gates: cpuQuantumChannelGates.jl names, GPUState in place of ψ -------
const G = (H=2, X=3, Y=4, Z=5, S=6, SDG=7, T=8, TDG=9, RX=10, RY=11, RZ=12, CNOT=13, CZ=14,
           CRY=15, RXX=16, RYY=17, RZZ=18, CCZ=19, TOFFOLI=20, SIGMA_MINUS=21, SIGMA_PLUS=22)
function named!(ψ::GPUState, gate, qs, θ = 0.0)
    all(1 .<= qs .<= ψ.N) || error("qubit out of range")
    q = Cint.(collect(qs) .- 1)
    check(ccall((:smoq_apply_named, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Cdouble),
                ψ.h, gate, q, θ))
    ψ
end
apply_rx_psi!(ψ::GPUState, k, θ, N) = named!(ψ, G.RX, (k,), θ)
apply_ry_psi!(ψ::GPUState, k, θ, N) = named!(ψ, G.RY, (k,), θ)
apply_rz_psi!(ψ::GPUState, k, θ, N) = named!(ψ, G.RZ, (k,), θ)
apply_hadamard_psi!(ψ::GPUState, k, N) = named!(ψ, G.H, (k,))
apply_pauli_x_psi!(ψ::GPUState, k, N) = named!(ψ, G.X, (k,))
apply_cz_psi!(ψ::GPUState, i, j, N) = named!(ψ, G.CZ, (i, j))
apply_cnot_psi!(ψ::GPUState, c, t, N) = named!(ψ, G.CNOT, (c, t))
apply_rzz_psi!(ψ::GPUState, i, j, θ, N) = named!(ψ, G.RZZ, (i, j), θ)
# … the other apply_*_psi! follow the same one-line pattern

# generic matrices (apply_gate!, apply_2qubit_gate!, …Trotter.jl bitwise forms)
interleave(U) = reinterpret(Float64, vec(permutedims(ComplexF64.(U))))   # row-major (re, im)
apply_one_qubit_gate_bitwise_cpu!(ψ::GPUState, U, k, N) =
    (check(ccall((:smoq_apply_1q, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), ψ.h, k - 1,
                 interleave(U))); ψ)
# Q4 kept: sorted qubits, U not permuted (…Trotter.jl:107-142)
apply_two_qubit_gate_bitwise_cpu!(ψ::GPUState, U, i, j, N) =
    (check(ccall((:smoq_apply_2q, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}), ψ.h,
                 min(i, j) - 1, max(i, j) - 1, interleave(U))); ψ)

# ---- observables -----------------------------------------------------------
function expect_pauli(ψ::GPUState, xmask::Vector{UInt64}, zmask::Vector{UInt64})
    out = zeros(Float64, length(xmask) * ψ.batch)
    check(ccall((:smoq_expect_pauli, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{UInt64}, Ptr{UInt64}, Ptr{Float64}),
                ψ.h, length(xmask), xmask, zmask, out))
    out
end
const XZ = Dict(:X => (true, false), :Y => (true, true), :Z => (false, true))
function masks(ops, qubits)
    x = z = UInt64(0)
    for (p, q) in zip(ops, qubits)
        bx, bz = XZ[p]
        bx && (x |= UInt64(1) << (q - 1)); bz && (z |= UInt64(1) << (q - 1))
    end
    x, z
end
expect_local(ψ::GPUState, k, N, p::Symbol) = (m = masks((p,), (k,)); expect_pauli(ψ, [m[1]], [m[2]])[1])
expect_corr(ψ::GPUState, i, j, N, pp::Symbol) =
    (s = String(pp); m = masks((Symbol(s[1]), Symbol(s[2])), (i, j)); expect_pauli(ψ, [m[1]], [m[2]])[1])
# the value the reference returns (Q1), for drop-in parity
function get_expectation_pauli_string(ψ::GPUState, ops, qubits, N)
    x, z = masks(ops, qubits); out = zeros(1)
    check(ccall((:smoq_pauli_string_reference, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{UInt64}, Ptr{UInt64}, Ptr{Float64}), ψ.h, 1, [x], [z], out))
    out[1]
end

# ---- sampling (run_circuit_sampling's rule, uniforms supplied) -------------
function sample_indices(ψ::GPUState, r::Vector{Float64}; semantics = 0)
    out = Vector{UInt64}(undef, length(r))
    check(ccall((:smoq_sample, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Cint, Ptr{UInt64}),
                ψ.h, r, length(r), semantics, out))
    out                                    # 0-based, as searchsortedfirst(...) - 1
end

# ---- gate tapes and compiled plans ----------------------------------------
struct Op                                   # layout of smoq_op
    gate::Int32; q::NTuple{3,Int32}; param::Int32; flags::Int32; theta::Float64
    m::NTuple{32,Float64}
end
Op(gate, qs, θ = 0.0; param = -1) =
    Op(gate, ntuple(i -> i <= length(qs) ? Int32(qs[i] - 1) : Int32(-1), 3), param, 0, θ,
       ntuple(_ -> 0.0, 32))
function apply_tape!(ψ::GPUState, ops::Vector{Op}, θ::Vector{Float64} = Float64[])
    np = Ref{Cint}(0)
    check(ccall((:smoq_apply_tape, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Op}, Cint, Ptr{Float64}, Cint, Ref{Cint}),
                ψ.h, ops, length(ops), θ, length(θ), np))
    ψ
end

# ---- MCWF (lindblad_mcwf_batched_step! / evolve_mcwf_batched_trotter!) -----
struct MCWFGroup                             # layout of smoq_mcwf_group
    psi::Ptr{Cvoid}; n_bonds::Cint; bond_q::Ptr{Int32}; bond_diag::Ptr{Float64}
    site_u::Ptr{Float64}; decay_pow::Ptr{Float64}; uniforms::Ptr{Float64}
    jumps::Ptr{Int32}; z_out::Ptr{Float64}
end
# gates: the FastTrotterGate list of precompute_trotter_gates_bitwise_cpu (diagonal
# bonds first, then 1-qubit fields); uniforms[traj, step] replaces rand()
function evolve_mcwf_batched_trotter!(Ψ::GPUState, gates, apply_trotter!::Function, N, n_traj,
                                      n_steps, γ, dt; uniforms::Matrix{Float64},
                                      obs_every::Integer = 0)
    apply_trotter! === apply_fast_trotter_step_cpu! ||
        error("the device MCWF step applies the gate list in order")
    bq = Int32[]; bd = Float64[]; su = repeat(Float64[1, 0, 0, 0, 0, 0, 1, 0], N)
    for g in gates
        if length(g.indices) == 2
            i, j = minmax(g.indices...)
            append!(bq, (i - 1, j - 1)); append!(bd, reinterpret(Float64, diag(g.U)))
        else
            k = g.indices[1]; su[8k-7:8k] = interleave(g.U * reshape(reinterpret(ComplexF64, su[8k-7:8k]), 2, 2))
        end
    end
    decay = [sqrt(1 - γ * dt)^m for m in 0:N]
    u = Matrix(transpose(uniforms))          # [traj][step] row-major
    jumps = Matrix{Int32}(undef, n_steps, n_traj)
    rec = obs_every > 0 ? n_steps ÷ obs_every : 0
    z = rec > 0 ? Array{Float64}(undef, N, rec, n_traj) : Float64[]
    GC.@preserve bq bd su decay u jumps z begin
        grp = Ref(MCWFGroup(Ψ.h, length(bq) ÷ 2, pointer(bq), pointer(bd), pointer(su),
                            pointer(decay), pointer(u), pointer(jumps),
                            rec > 0 ? pointer(z) : C_NULL))
        check(ccall((:smoq_mcwf_evolve, LIB), Cint, (Ptr{Cvoid}, Ref{MCWFGroup}, Cint, Cdouble, Cint, Cint),
                    ctx().h, grp, 1, γ * dt, n_steps, obs_every))
    end
    jumps, z                                  # jumps[step, traj]: 0-based site or -1
end

# ---- adjoint gradient (ChainRulesCore rrule over the energy) ---------------
function energy_and_grad(ψ::GPUState, ops::Vector{Op}, θ::Vector{Float64},
                         xmask::Vector{UInt64}, zmask::Vector{UInt64}, coeff::Vector{Float64})
    E = zeros(ψ.batch); g = zeros(length(θ))
    check(ccall((:smoq_adjoint_gradient, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Op}, Cint, Ptr{Float64}, Cint, Cint, Ptr{UInt64}, Ptr{UInt64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ψ.h, ops, length(ops), θ, length(θ) ÷ ψ.batch, length(coeff), xmask, zmask, coeff, E, g))
    E, g
end

# ---- VQE energy with a ChainRulesCore rule ----------------------------------
# replaces calculate_gradients (Enzyme reverse mode over apply_circuit! and
# compute_energy, utils/cpu/cpuVariationalQuantumCircuitGradients.jl:88-110):
# Zygote / any ChainRules-aware AD differentiates vqe_energy through the
# device's adjoint sweep instead of tracing the state-vector loops.
struct PauliSum                             # H = sum_t coeff[t] P_t (0-based bit masks)
    xmask::Vector{UInt64}
    zmask::Vector{UInt64}
    coeff::Vector{Float64}
end
function vqe_energy(ops::Vector{Op}, θ::Vector{Float64}, H::PauliSum, N::Int)
    ψ = make_ket_gpu(N); E = zeros(1)
    check(ccall((:smoq_tape_energy, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Op}, Cint, Ptr{Float64}, Cint, Cint, Ptr{UInt64}, Ptr{UInt64},
                 Ptr{Float64}, Cint, Ptr{Float64}),
                ψ.h, ops, length(ops), θ, length(θ), length(H.coeff), H.xmask, H.zmask, H.coeff,
                0, E))
    E[1]
end
function ChainRulesCore.rrule(::typeof(vqe_energy), ops::Vector{Op}, θ::Vector{Float64},
                              H::PauliSum, N::Int)
    E, g = energy_and_grad(make_ket_gpu(N), ops, θ, H.xmask, H.zmask, H.coeff)
    vqe_energy_pullback(Ē) = (NoTangent(), NoTangent(), unthunk(Ē) .* g, NoTangent(), NoTangent())
    return E[1], vqe_energy_pullback
end
# calculate_gradients(cost_fn, θ) for a cost that is a device VQE energy
struct VQECost
    ops::Vector{Op}
    H::PauliSum
    N::Int
end
(c::VQECost)(θ::Vector{Float64}) = vqe_energy(c.ops, θ, c.H, c.N)
calculate_gradients(c::VQECost, θ::Vector{Float64}) =
    energy_and_grad(make_ket_gpu(c.N), c.ops, θ, c.H.xmask, c.H.zmask, c.H.coeff)[2]

# ---- preparation, Lanczos, any-step MCWF, batched read-outs ----------------
# make_product_ket(states::Vector{Symbol}) (cpuQuantumStatePreparation.jl:1182)
const KET = Dict(:zero => (1.0, 0.0), :one => (0.0, 1.0),
                 :plus => (1/sqrt(2), 1/sqrt(2)), :minus => (1/sqrt(2), -1/sqrt(2)))
function make_product_ket_gpu(states::Vector{Symbol}; batch = 1, dtype = C128)
    ψ = GPUState(length(states); batch = batch, dtype = dtype)
    k = Float64[]
    for s in states
        a, b = KET[s]; append!(k, (a, 0.0, b, 0.0))      # (re, im) of <0|q>, <1|q>
    end
    check(ccall((:smoq_state_set_product, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), ψ.h, k)); ψ
end

# estimate_spectral_range_lanczos(params; k_iter) with the start vector passed in
function lanczos_range(start::GPUState, xmask, zmask, coeff; k_iter = 30)
    lo = Ref(0.0); hi = Ref(0.0); used = Ref{Cint}(0)
    check(ccall((:smoq_lanczos_range, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{UInt64}, Ptr{UInt64}, Ptr{Float64}, Cint,
                 Ref{Float64}, Ref{Float64}, Ref{Cint}),
                start.h, length(coeff), xmask, zmask, coeff, k_iter, lo, hi, used))
    lo[], hi[]
end

# lindblad_mcwf_step_trotter! / σ_z dephasing with any gate list: compile the
# step once (smoq_plan_create over the Op list), then n_steps of [plan, jump]
function evolve_mcwf_trotter!(Ψ::GPUState, plan::Ptr{Cvoid}, kind::Integer, γdt, decay::Vector{Float64},
                              n_steps, uniforms::Array{Float64,3};  # [2, step, traj]
                              obs_every = 0, xmask = UInt64[], zmask = UInt64[])
    jumps = Matrix{Int32}(undef, n_steps, Ψ.batch)
    rec = obs_every > 0 ? n_steps ÷ obs_every : 0
    out = zeros(length(xmask), Ψ.batch, rec)
    check(ccall((:smoq_mcwf_trotter_evolve, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Int32},
                 Cint, Cint, Ptr{UInt64}, Ptr{UInt64}, Ptr{Float64}),
                plan, Ψ.h, kind, γdt, decay, n_steps, uniforms, jumps, obs_every, length(xmask),
                xmask, zmask, out))
    jumps, out                               # kind: 0 σ⁻, 1 σ_z; decay: table of smoq_mcwf_jump
end

# ---- noise channels, measurement, reset (one uniform per qubit and item) ----
const QC = Dict(:measure_z => 0, :dephasing => 1, :depolarizing => 2, :bit_flip => 3,
                :phase_flip => 4, :amplitude_damping => 5, :force_zero => 6, :force_one => 7)
# records[b, k] (column-major, the ABI's out[k * batch + b]): the outcome /
# 0 none, 1 X, 2 Y, 3 Z, 4 jump / the forced bit
function qubit_channel!(ψ::GPUState, kind::Symbol, p, qubits::Vector{Int};
                        uniforms = rand(ψ.batch, length(qubits)))   # [b, k]: item b, qubit k
    rec = Matrix{Int32}(undef, ψ.batch, length(qubits))
    check(ccall((:smoq_qubit_channel, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Cdouble, Ptr{Float64}, Ptr{Int32}),
                ψ.h, length(qubits), Cint.(qubits .- 1), QC[kind], p, uniforms, rec))
    rec
end
# apply_channel_dephasing!(ψ, p, qubits, N) and the other four, same arguments;
# on ρ the same names append one U2 / DIAG2 op per qubit on (k-1, k-1+N) (the
# channel's 4×4 pair map, smoq_amd/channels.py: kraus_pair_map) to a tape
for ch in (:dephasing, :depolarizing, :amplitude_damping, :bit_flip, :phase_flip)
    f = Symbol(:apply_channel_, ch, :!)
    @eval $f(ψ::GPUState, p::Float64, qubits::Vector{Int}, N::Int; kw...) =
        (qubit_channel!(ψ, $(QuoteNode(ch)), p, qubits; kw...); ψ)
end
# projective_measurement!(ψ, qubits, :z, N) -> (outcomes, ψ); :x / :y wrap H, S† H as the reference
projective_measurement_z!(ψ::GPUState, qubits::Vector{Int}; kw...) =
    (vec(qubit_channel!(ψ, :measure_z, 0.0, qubits; kw...)), ψ)
# reset_state!(ψ, qubits, targets, N): measure, then force (then H for :plus / :minus)
function reset_state!(ψ::GPUState, qubits::Vector{Int}, targets::Vector{Symbol}, N::Int;
                      uniforms = rand(ψ.batch, length(qubits)))
    for (j, (k, t)) in enumerate(zip(qubits, targets))
        qubit_channel!(ψ, :measure_z, 0.0, [k]; uniforms = uniforms[:, j:j])
        qubit_channel!(ψ, t in (:zero, :plus) ? :force_zero : :force_one, 0.0, [k];
                       uniforms = zeros(ψ.batch, 1))
        t in (:plus, :minus) && apply_hadamard_psi!(ψ, k, N)
    end
    ψ
end
# density matrices (vec(ρ) as a 2N-qubit GPUState): Z measurement and _force_qubit_rho!
function dm_qubit_channel!(ρ::GPUState, kind::Symbol, qubits::Vector{Int};
                           uniforms = rand(ρ.batch, length(qubits)))
    rec = Matrix{Int32}(undef, ρ.batch, length(qubits))
    check(ccall((:smoq_dm_qubit_channel, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Ptr{Float64}, Ptr{Int32}),
                ρ.h, length(qubits), Cint.(qubits .- 1), QC[kind], uniforms, rec))
    rec
end
# projective_measurement_all!(ρ, N) / sample_state(ρ, …): smoq_sample with
# semantics 2 (SMOQ_SAMPLE_SEQUENTIAL_DIAG) over real(diag(ρ)), then |i><i|

# ---- reduced states and classical shadows ---------------------------------
# partial_trace(ψ, trace_qubits, N): a device density matrix (vec(ρ_A) as a
# 2K-qubit GPUState); Matrix(…) of it is the reference's return value
function partial_trace(ψ::GPUState, trace_qubits, N::Int)
    keep = [q for q in 1:N if !(q in trace_qubits)]
    isempty(keep) && error("Cannot trace out all sites - no subsystem remains")
    ρA = GPUState(2 * length(keep); batch = ψ.batch)
    check(ccall((:smoq_partial_trace, LIB), Cint, (Ptr{Cvoid}, UInt64, Ptr{Cvoid}),
                ψ.h, sum(UInt64(1) << (q - 1) for q in keep), ρA.h))
    ρA                                   # reshape(Array(ρA), 2^K, 2^K) is rho_A
end
# collect_shadows(ψ, N, M) with bases (M × N, 1:3) and uniforms given
function collect_shadows(ψ::GPUState, N::Int, M::Int; bases = rand(UInt8(1):UInt8(3), N, M),
                         uniforms = rand(M))           # bases[:, m]: snapshot m
    idx = Vector{UInt64}(undef, M)
    check(ccall((:smoq_shadow_snapshots, LIB), Cint,
                (Ptr{Cvoid}, Int64, Cint, Ptr{UInt8}, Ptr{Float64}, Ptr{Float64}, Ptr{UInt64}),
                ψ.h, M, 0, bases, C_NULL, uniforms, idx))
    [PauliSnapshot(Int.(bases[:, m]), [Int((idx[m] >> (k - 1)) & 1) for k in 1:N]) for m in 1:M]
end
# collect_clifford_shadows: kind 1, codes = indices .- 1 (0-based), and the
# 24 matrices of SINGLE_QUBIT_CLIFFORDS as 192 doubles (row-major re, im)
# collect_global_clifford_shadows: smoq_global_clifford_snapshots with each
# circuit's gates as UInt8 triples (0 = H, 1 = S: bit N - q, as the
# reference's kron order; 2 = CNOT: bits c - 1, t - 1) and Int64 offsets

# ---- apply_circuit! over the reference's ParameterizedCircuit --------------
# (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl:95-117, 252-262): the
# CircuitOperation list becomes smoq_circuit_op records and one
# smoq_apply_circuit call; gates between two other ops run as one set of
# fused passes, noise / measure / reset are device calls that take the
# uniforms the reference draws with rand(), if_outcome is decided per op.
struct CircuitOp                             # layout of smoq_circuit_op
    kind::Int32; param::Int32; q_first::Int32; q_count::Int32
    if_outcome::Int32; arg::Int32; angle::Float64; p::Float64
end
const CK = Dict(:rx => 0, :ry => 1, :rz => 2, :h => 3, :x => 4, :y => 5, :z => 6, :s => 7,
                :t => 8, :cz => 9, :cnot => 10, :cx => 10, :rxx => 11, :ryy => 12, :rzz => 13,
                :depolarizing => 16, :amplitude_damping => 17, :dephasing => 18,
                :bit_flip => 19, :phase_damping => 20, :thermal_relaxation => 21,
                :measure_z => 32, :measure_x => 32, :measure_y => 32, :measure => 32,
                :reset => 33, :trace_out => 34, :lindblad_step => 35)
const BASIS = Dict(:z => 0, :x => 1, :y => 2)
const TARGET = Dict(:zero => 0, :one => 1, :plus => 2, :minus => 3)
function lower_circuit(circuit)
    recs = CircuitOp[]; qs = Int32[]; n_meas = 0; n_draw = 0
    for op in circuit.operations
        kind = CK[op.type]
        arg = kind == 32 ? BASIS[op.basis] : kind == 33 ? TARGET[op.target_state] : 0
        push!(recs, CircuitOp(kind, op.param_idx > 0 ? op.param_idx - 1 : -1, length(qs),
                              length(op.qubits), op.if_outcome, arg, op.fixed_angle, op.p))
        append!(qs, Int32.(op.qubits .- 1))
        kind == 32 && (n_meas += length(op.qubits))
        (16 <= kind <= 20 || kind == 32 || kind == 33) && (n_draw += length(op.qubits))
    end
    recs, qs, n_meas, n_draw
end
struct GPUDensity                            # ρ of N qubits as vec(ρ) on 2N qubits
    vec::GPUState
    N::Int
end
function run_circuit!(s::GPUState, dm::Bool, circuit, θ::Vector{Float64}, uniforms::Vector{Float64})
    recs, qs, n_meas, _ = lower_circuit(circuit)
    outcomes = Vector{Int32}(undef, max(1, n_meas * s.batch))
    n_out = Ref{Cint}(0); used = Ref{Int64}(0)
    check(ccall((:smoq_apply_circuit, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{CircuitOp}, Cint, Ptr{Int32}, Cint, Ptr{Float64}, Cint,
                 Ptr{Float64}, Int64, Ptr{Int32}, Cint, Ref{Cint}, Ref{Int64}),
                s.h, dm, recs, length(recs), qs, length(qs), θ, length(θ), uniforms,
                length(uniforms), outcomes, n_meas, n_out, used))
    k = Int(n_out[])
    s.batch == 1 ? Int.(outcomes[1:k]) : permutedims(reshape(Int.(outcomes[1:k*s.batch]), s.batch, k))
end
# returns (state, outcomes), the fields of the reference's CircuitResult; uniforms in op
# order, s.batch values per listed qubit (qubit-major), default rand()
function apply_circuit!(ψ::GPUState, circuit, θ::Vector{Float64};
                        uniforms::Vector{Float64} = rand(ψ.batch * lower_circuit(circuit)[4]))
    ψ.N == circuit.N || error("State dimension mismatch")
    length(θ) >= circuit.n_params || error("Not enough parameters")
    any(op.type == :trace_out for op in circuit.operations) &&
        @warn "trace_out in pure state mode: use partial_trace from cpuQuantumStatePartialTrace.jl"
    (ψ, run_circuit!(ψ, false, circuit, θ, uniforms))
end
function apply_circuit!(ρ::GPUDensity, circuit, θ::Vector{Float64})
    ρ.N == circuit.N || error("State dimension mismatch")
    length(θ) >= circuit.n_params || error("Not enough parameters")
    run_circuit!(ρ.vec, true, circuit, θ, Float64[])
    nothing
end

end # module
