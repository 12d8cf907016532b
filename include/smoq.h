/*
 * smoq.h — C ABI of the MI355X (gfx950) state-vector kernels.
 *
 * One shared library, libsmoq_amd.so, exports only the functions below.
 * Every function returns an int status (SMOQ_OK = 0, negative on error);
 * smoq_last_error() returns a thread-local message for the last failure on
 * the calling thread.  No C++ exception crosses this boundary.
 *
 * Conventions (fixed by the reference, SmoQ.jl):
 *  - Qubits are 0-based here; qubit q is bit q of the amplitude index
 *    (little-endian; the reference's 1-based qubit k is bit k-1,
 *    utils/cpu/cpuQuantumChannelGates.jl:48-51).
 *  - A state holds `batch` independent vectors of 2^n amplitudes, stored
 *    back to back (item b at offset b*2^n), interleaved (re, im): complex128
 *    = 16 B, complex64 = 8 B per amplitude.  This is the column-major
 *    dim x n_traj layout of the reference's MCWF batches
 *    (utils/cpu/cpuQuantumChannelLindbladianEvolutionMonteCarloWaveFunction.jl:326-345).
 *  - Matrices are row-major complex, (re, im) pairs.  A 2-qubit matrix acts
 *    on the basis index b(q[0]) + 2*b(q[1]) — the |00>,|10>,|01>,|11> order of
 *    apply_2qubit_gate! (utils/cpu/cpuQuantumChannelRandomUnitaries.jl:300-302).
 *  - Kernels run asynchronously on the context's stream; only calls that
 *    return host data (reductions, samples, downloads) synchronise.  Calls on
 *    one context must be serialised by the caller.
 */
#ifndef SMOQ_H
#define SMOQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMOQ_ABI_VERSION 1

enum smoq_status {
  SMOQ_OK = 0,
  SMOQ_EINVAL = -1,    /* bad argument (qubit out of range, size mismatch ...) */
  SMOQ_ENOMEM = -2,    /* device allocation failed or request too large */
  SMOQ_EDEVICE = -3,   /* HIP runtime error */
  SMOQ_ENUMERIC = -4,  /* non-finite or zero norm where a positive one is needed */
  SMOQ_ENOTSUP = -5    /* valid request this build does not handle */
};

enum smoq_dtype { SMOQ_C128 = 0, SMOQ_C64 = 1 };

/* Gate identifiers for smoq_apply_named and smoq_op.gate.  Matrices follow
 * utils/cpu/cpuQuantumChannelGates.jl:53-71 and the per-gate docstrings. */
enum smoq_gate_id {
  SMOQ_G_U1 = 0,         /* generic 2x2 on q[0], m[0..7] */
  SMOQ_G_U2 = 1,         /* generic 4x4 on (q[0], q[1]), m[0..31] */
  SMOQ_G_H = 2,
  SMOQ_G_X = 3,
  SMOQ_G_Y = 4,
  SMOQ_G_Z = 5,
  SMOQ_G_S = 6,
  SMOQ_G_SDG = 7,
  SMOQ_G_T = 8,
  SMOQ_G_TDG = 9,
  SMOQ_G_RX = 10,        /* exp(-i theta X/2) */
  SMOQ_G_RY = 11,
  SMOQ_G_RZ = 12,
  SMOQ_G_CNOT = 13,      /* q[0] control, q[1] target */
  SMOQ_G_CZ = 14,
  SMOQ_G_CRY = 15,       /* q[0] control, q[1] target */
  SMOQ_G_RXX = 16,       /* exp(-i theta XX/2) */
  SMOQ_G_RYY = 17,
  SMOQ_G_RZZ = 18,
  SMOQ_G_CCZ = 19,
  SMOQ_G_TOFFOLI = 20,   /* q[0], q[1] controls, q[2] target */
  SMOQ_G_SIGMA_MINUS = 21, /* |0><1| (non-unitary jump) */
  SMOQ_G_SIGMA_PLUS = 22,  /* |1><0| */
  SMOQ_G_DIAG1 = 23,     /* diag(m[0..1], m[2..3]) on q[0] */
  SMOQ_G_DIAG2 = 24,     /* diagonal, entry b(q0)+2b(q1) = m[2e..2e+1] */
  SMOQ_G_COUNT = 25
};

/* One operation of a gate tape (a circuit).  For the rotation gates
 * (RX, RY, RZ, CRY, RXX, RYY, RZZ) the angle is theta[param] when
 * param >= 0 and the `theta` field otherwise. */
typedef struct smoq_op {
  int32_t gate;
  int32_t q[3];
  int32_t param;
  int32_t flags;         /* reserved, must be 0 */
  double theta;
  double m[32];
} smoq_op;

typedef struct smoq_ctx_s smoq_ctx;
typedef struct smoq_state_s smoq_state;

/* ---- library / context ------------------------------------------------ */
/* No reference counterpart: the reference runs in-process on host arrays. */
const char* smoq_last_error(void);
int smoq_version(void);
int smoq_init(smoq_ctx** out, int device);
int smoq_fini(smoq_ctx* ctx);
int smoq_sync(smoq_ctx* ctx);
/* Fused-pass geometry: tile_qubits (amplitudes per LDS tile = 2^tile_qubits,
 * 6..12) and low_qubits (contiguous low bits kept in every tile, 3..6).
 * 0 keeps the current value; tile_qubits < 0 returns to the automatic
 * choice (the default): 2^10 tiles for the precompiled pass kernel, the
 * largest tile that leaves >= 2048 workgroups for compiled passes. */
int smoq_set_fusion(smoq_ctx* ctx, int tile_qubits, int low_qubits);
/* Fused passes compiled at run time for one circuit structure (hiprtc,
 * cached per context by structure; gate angles stay run-time data):
 * mode 0 never, 1 automatically for states of >= 16 qubits with >= 16
 * gates (the default; the precompiled pass kernel is kept if compilation
 * fails), 2 always (compilation failure is an error).  The SMOQ_JIT
 * environment variable sets the initial mode of new contexts. */
int smoq_set_jit(smoq_ctx* ctx, int mode);
/* Compiled passes may keep a stage's gates on five qubits (three register
 * slots and lane bits 4 and 5, traded by v_permlane16/32_swap) instead of
 * three, with fewer LDS round trips for more VALU work: on (1, the default;
 * SMOQ_LANE_SWAPS=0 in the environment sets new contexts off) or off (0).
 * A structure tuned by smoq_autotune keeps its tuned choice. */
int smoq_set_lane_swaps(smoq_ctx* ctx, int on);
/* Pauli term lists (smoq_expect_pauli, measure_all_observables) through
 * reduction passes compiled for the term list: 0 off (precompiled tiled
 * passes), 1 automatic (states of at least 12 qubits and 2^22 amplitudes
 * over the batch; the default, or SMOQ_OBS_JIT), 2 for every state of at
 * least 12 qubits.  No reference counterpart (tuning). */
int smoq_set_obs_jit(smoq_ctx* ctx, int mode);
/* Mean kernel time (ms) and launch count of the fused-pass kernel since the
 * last call, measured with HIP events on the context stream when timing is
 * enabled (enable != 0).  Used to report the per-launch roofline. */
int smoq_pass_timing(smoq_ctx* ctx, int enable, double* mean_ms, int64_t* launches);

/* ---- state lifecycle --------------------------------------------------
 * Replaces: zeros(ComplexF64, 2^N); ψ[1] = 1 and make_ket / make_initial_psi
 * (utils/cpu/cpuQuantumStatePreparation.jl:208-243,1211-1229) for basis kets. */
int smoq_state_alloc(smoq_ctx* ctx, int n_qubits, int dtype, int batch, smoq_state** out);
/* States of at most 64 MiB go back to a pool of eight buffers on their
 * context (reused, stream-ordered and zero-filled, by a later allocation of
 * the same byte size; released by smoq_fini, or earlier when an allocation
 * would not fit otherwise); larger ones are freed after the context's stream
 * drains.  Free every state and plan before smoq_fini of its context:
 * smoq_fini returns SMOQ_EINVAL, and keeps the context, while any remain. */
int smoq_state_free(smoq_state* s);
int smoq_state_info(const smoq_state* s, int* n_qubits, int* dtype, int* batch, void** device_ptr);
int smoq_state_set_basis(smoq_state* s, uint64_t index);
/* Product state in every batch item: kets[4q .. 4q+3] = (re, im) of <0|q>
 * and <1|q> for qubit q (0-based, qubit q is bit q); amplitude i is the
 * product over qubits multiplied in qubit order, as make_product_ket
 * (utils/cpu/cpuQuantumStatePreparation.jl:1182-1191) builds it with
 * tensor_product_ket (:1108-1123).  Not normalised by the call. */
int smoq_state_set_product(smoq_state* s, const double* kets);
int smoq_state_upload(smoq_state* s, const void* host, size_t bytes);
int smoq_state_download(const smoq_state* s, void* host, size_t bytes);
int smoq_state_copy(smoq_state* dst, const smoq_state* src);

/* ---- gates --------------------------------------------------------------
 * smoq_apply_1q replaces apply_one_qubit_gate_bitwise_cpu!
 *   (utils/cpu/cpuQuantumChannelUnitaryEvolutionTrotter.jl:77-99) and
 *   apply_gate! (utils/cpu/cpuQuantumChannelRandomUnitaries.jl:230).
 * smoq_apply_2q replaces apply_2qubit_gate! (cpuQuantumChannelRandomUnitaries.jl:317-370);
 *   apply_two_qubit_gate_bitwise_cpu! (…Trotter.jl:107-142) is the call with
 *   (min(q1,q2), max(q1,q2)).
 * smoq_apply_named replaces every apply_*_psi! of
 *   utils/cpu/cpuQuantumChannelGates.jl:142-1100 (rx, ry, rz, hadamard,
 *   pauli_x/y/z, s, sdagger, t, tdagger, ccz, toffoli, sigma_minus/plus,
 *   cnot, cz, cry, rxx, ryy, rzz). */
int smoq_apply_1q(smoq_state* s, int q, const double u[8]);
int smoq_apply_2q(smoq_state* s, int qa, int qb, const double u[32]);
int smoq_apply_named(smoq_state* s, int gate, const int* qubits, double theta);
/* Fused application of a whole gate tape: gates are grouped into HBM passes
 * over 2^tile_qubits-amplitude LDS tiles.  Replaces apply_circuit!
 * (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl:95-117) and
 * fuse_gates + run_fused_circuit!
 * (SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:167-200,356-369).
 * theta may be NULL when no op has param >= 0.  Returns the number of HBM
 * passes used in *n_passes when it is not NULL. */
int smoq_apply_tape(smoq_state* s, const smoq_op* ops, int n_ops,
                    const double* theta, int n_theta, int* n_passes);

/* ---- autotuning of the fused passes ---------------------------------------
 * The reference fuses gates by a fixed greedy rule (fuse_gates,
 * SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:167-200).
 * smoq_autotune plans the tape under candidate settings (tile qubits 11 or
 * 12, 4 or 5 contiguous low qubits, a cap of 0 / 28 / 36 / 44 non-diagonal
 * gates per pass, stages with or without register/lane exchanges, local
 * qubit sets from a beam search over plan prefixes) besides the context's
 * own, compiles their pass kernels, times each candidate on s (its
 * contents are overwritten), then times HBM orders (tile-index rotations and
 * reversals, load maps) pass by pass for the winner, and records the
 * fastest for the tape's structure (gate kinds and qubits), qubit count,
 * dtype and batch: later smoq_apply_tape / smoq_plan_create calls with that
 * structure use it, unless smoq_set_fusion fixed a tile size.  No
 * reference counterpart (tooling). */
typedef struct smoq_tune_result {
  int tile_qubits, low_qubits, pass_max_gates; /* the chosen settings */
  int lane_swaps;     /* stages trade register and lane qubits (1) or not */
  int searched;       /* passes' local qubit sets by search (1) or greedy (0) */
  int n_passes;       /* HBM passes per application with them */
  int candidates;     /* settings timed (0: nothing to choose) */
  double ms;          /* one application with the chosen settings */
  double default_ms;  /* one application with the context's settings */
  /* per-pass HBM orders (states of 2^12 or more tiles per pass): passes given
   * a non-default tile order, and the sum of the passes' times with them
   * (0: not searched) */
  int io_passes;
  double io_ms;
} smoq_tune_result;
int smoq_autotune(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                  int n_theta, smoq_tune_result* out);
/* The context's tuned settings to / from a text file (one line per
 * structure); loading merges into the context's table. */
int smoq_tune_save(smoq_ctx* ctx, const char* path);
int smoq_tune_load(smoq_ctx* ctx, const char* path, int* n_entries);

/* ---- compiled tapes ------------------------------------------------------
 * A plan is a tape planned once (fused passes, device-resident program) for
 * states of one qubit count and dtype; every run replays the same launches,
 * captured in a hipGraph after the first run on a given state.  It replaces
 * the per-step gate loop of apply_fast_trotter_step_cpu! / evolve_trotter_psi_cpu!
 * (utils/cpu/cpuQuantumChannelUnitaryEvolutionTrotter.jl:150-162,250-256). */
typedef struct smoq_plan_s smoq_plan;
int smoq_plan_create(smoq_ctx* ctx, int n_qubits, int dtype, const smoq_op* ops, int n_ops,
                     const double* theta, int n_theta, smoq_plan** out);
int smoq_plan_passes(const smoq_plan* p, int* n_passes);
/* applies the tape `repeat` times */
int smoq_plan_run(smoq_plan* p, smoq_state* s, int repeat);
int smoq_plan_free(smoq_plan* p);
/* Runs n_steps applications of the plan and, after every obs_every-th one,
 * records <P_t> for the n_terms Pauli strings: out[r][item][t] with
 * r = 0 .. n_steps/obs_every - 1.  solve_trotter_bitwise_cpu's time loop
 * (…UnitaryEvolutionTrotter.jl:293-311) with the observables taken on the
 * device instead of storing states. */
int smoq_evolve_observe(smoq_plan* p, smoq_state* s, int n_steps, int obs_every, int n_terms,
                        const uint64_t* xmask, const uint64_t* zmask, double* out);

/* ---- Hamiltonian action and Chebyshev propagation -------------------------
 * H = sum_t c_t P_t with real c_t and Pauli strings P_t given as (x, z)
 * masks as in smoq_expect_pauli.
 * smoq_hpsi: out = H in (pull form: one write per amplitude), replacing
 *   apply_hamiltonian_bitwise! (utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl:55-152).
 * smoq_chebyshev_evolve: psi <- exp(-i H t) psi by the Chebyshev-Bessel
 *   series on the spectrum rescaled from [emin, emax] (widened by 1% each
 *   side), stopping after the first order k >= 2 with |J_k(a t)| < tol or at
 *   max_order; replaces chebyshev_evolve_psi! (…Chebyshev.jl:262-346) with
 *   the spectral bounds supplied.  *orders_used receives the reference's
 *   M_used (last order + 1).  Needs two scratch states of the same shape. */
int smoq_hpsi(const smoq_state* in, smoq_state* out, int n_terms, const uint64_t* xmask,
              const uint64_t* zmask, const double* coeff);
int smoq_chebyshev_evolve(smoq_state* s, int n_terms, const uint64_t* xmask,
                          const uint64_t* zmask, const double* coeff, double t, double emin,
                          double emax, double tol, int max_order, int* orders_used);
/* estimate_spectral_range_lanczos (…Chebyshev.jl:166-220) with the start
 * vector supplied (the reference draws an unseeded randn vector): k_iter
 * Lanczos steps of H on start / ||start|| (start itself is not changed),
 * stopping early when beta < 1e-12; *e_min, *e_max are the extreme
 * eigenvalues of the tridiagonal matrix of the alphas and betas,
 * *iters_used the number of alphas.  One state (batch 1); allocates three
 * vectors of its size. */
int smoq_lanczos_range(const smoq_state* start, int n_terms, const uint64_t* xmask,
                       const uint64_t* zmask, const double* coeff, int k_iter, double* e_min,
                       double* e_max, int* iters_used);

/* ---- reduced density matrices ---------------------------------------------
 * partial_trace(psi, trace_qubits, N) (utils/cpu/cpuQuantumStatePartialTrace.jl:68-127)
 * for every batch item: the qubits set in keep_mask (bit q = qubit q, 0-based,
 * at least one) are kept and the others traced out.  out is a separate state
 * of 2K qubits (K = popcount(keep_mask)), same batch, dtype and context: it
 * receives vec(rho_A), element (i, j) at i + 2^K j with i, j over the kept
 * qubits in increasing order (the reference's keep_indices). */
int smoq_partial_trace(const smoq_state* psi, uint64_t keep_mask, smoq_state* out);
/* partial_trace(rho, trace_qubits, N) (…PartialTrace.jl:141-188) of a density
 * matrix stored as vec(rho) (2N vec qubits); keep_mask over its N qubits. */
int smoq_dm_partial_trace(const smoq_state* rho, uint64_t keep_mask, smoq_state* out);

/* ---- classical-shadow snapshots -------------------------------------------
 * psi must hold one vector (batch 1; a batched state is rejected with
 * SMOQ_EINVAL).  For snapshot m (0 <= m < n_snap): a copy of psi rotated
 * qubit by qubit in increasing order by codes[m * n + q], then one draw with
 * sample_from_probabilities' rule (utils/cpu/cpuQuantumStateMeasurements.jl:203-213)
 * on uniforms[m]; out_index[m] is the 0-based outcome (bit q = the outcome of
 * qubit q).  psi is not changed.
 *   SMOQ_SHADOW_PAULI: codes 1 = X (H), 2 = Y (S† then H), 3 = Z (none), with
 *     the reference's arithmetic: classical_shadow_snapshot
 *     (utils/cpu/cpuClassicalShadows.jl:1001-1037), bit-exact outcomes.
 *   SMOQ_SHADOW_LOCAL_CLIFFORD: codes 0..23 index clifford_table (24 row-major
 *     2x2 complex matrices as re, im pairs, 192 doubles): clifford_shadow_snapshot
 *     (…ClassicalShadows.jl:580-603), which applies the same product as one
 *     dense matrix.
 * Snapshots with equal code strings share one rotated state.  At most 32
 * qubits. */
enum smoq_shadow_kind { SMOQ_SHADOW_PAULI = 0, SMOQ_SHADOW_LOCAL_CLIFFORD = 1 };
int smoq_shadow_snapshots(const smoq_state* psi, int64_t n_snap, int kind, const uint8_t* codes,
                          const double* clifford_table, const double* uniforms,
                          uint64_t* out_index);
/* Global-Clifford snapshots of psi (batch 1) (global_clifford_shadow_snapshot,
 * utils/cpu/cpuClassicalShadows.jl:884-903): snapshot m applies gates
 * offsets[m] .. offsets[m+1]-1 of `gates` (triples type, a, b: 0 = H on bit a,
 * 1 = S on bit a, 2 = CNOT control bit a, target bit b) to a copy of psi and
 * draws with sample_from_probabilities' rule on uniforms[m].  The reference
 * multiplies by the circuit's dense unitary instead.  At most 12 qubits. */
int smoq_global_clifford_snapshots(const smoq_state* psi, int64_t n_snap, const uint8_t* gates,
                                   const int64_t* offsets, const double* uniforms,
                                   uint64_t* out_index);

/* ---- single-qubit stochastic operations ------------------------------------
 * For k = 0 .. n_qubits-1 in order and every batch item b, with the caller's
 * uniform u[k * batch + b] in place of the reference's rand() and the record
 * out[k * batch + b] (out may be NULL); qubits are 0-based:
 *   SMOQ_QC_MEASURE_Z: _measure_z_single! (utils/cpu/cpuQuantumStateMeasurements.jl:224-256):
 *     outcome 0 if u < P(0) else 1, collapse, multiply by 1/sqrt(kept
 *     probability); record = outcome.  A zero-probability outcome (only
 *     possible for an unnormalised state) returns SMOQ_ENUMERIC.
 *   SMOQ_QC_DEPHASING / DEPOLARIZING / BIT_FLIP / PHASE_FLIP /
 *   AMPLITUDE_DAMPING: the state-vector (trajectory) branches of
 *     apply_channel_*! (utils/cpu/cpuQuantumChannelKrausOperators.jl:262-385):
 *     Z when u < (1 - sqrt(1-p))/2 (none at p == 0); X, Y, Z when u < p/3,
 *     2p/3, p; X when u < p; Z when u < p; a jump when u < p P(1) and
 *     P(1) > 1e-12, which adds the |1> amplitudes onto the |0> ones as the
 *     reference does, else the |1> amplitudes times sqrt(1-p), then division
 *     by the norm when it exceeds 1e-12.  record = 0 none, 1 X, 2 Y, 3 Z,
 *     4 jump.  p must lie in [0, 1].
 *   SMOQ_QC_FORCE_ZERO / FORCE_ONE: _force_qubit_state! for :zero / :one
 *     (…Measurements.jl:584-603, 629-635): the other half's amplitudes are
 *     added onto the target half, then division by the norm when norm^2
 *     exceeds 1e-15; uniforms may be NULL, record = the forced bit. */
enum smoq_qubit_channel_kind {
  SMOQ_QC_MEASURE_Z = 0,
  SMOQ_QC_DEPHASING = 1,
  SMOQ_QC_DEPOLARIZING = 2,
  SMOQ_QC_BIT_FLIP = 3,
  SMOQ_QC_PHASE_FLIP = 4,
  SMOQ_QC_AMPLITUDE_DAMPING = 5,
  SMOQ_QC_FORCE_ZERO = 6,
  SMOQ_QC_FORCE_ONE = 7
};
int smoq_qubit_channel(smoq_state* s, int n_qubits, const int* qubits, int kind, double p,
                       const double* uniforms, int32_t* out);

/* ---- density matrices -----------------------------------------------------
 * rho of N qubits is a state of 2N qubits holding vec(rho): element (i, j)
 * at index i + 2^N j (column-major, the memory order of the reference's
 * Matrix{ComplexF64}).  Gates act as U on qubit q and conj(U) on qubit
 * N + q; the Lindblad dissipators of the reference are 4x4 maps on the pair
 * (q, N + q); both run through smoq_apply_tape / plans.
 * smoq_dm_expect_pauli: out[b * n_terms + t] = Re Tr(rho_b P_t) for Pauli
 *   strings over the N qubits (masks as smoq_expect_pauli), replacing
 *   _expect_X/Y/Z_dm, _expect_XX/YY/ZZ_dm and the Pauli-string forms
 *   (utils/cpu/cpuQuantumStateObservables.jl:1138-1420); Tr(rho) is the
 *   term with both masks 0.
 * smoq_dm_apply_dense: rho_b <- U rho_b U^H for every item, U a single
 *   2N-qubit state holding the dense 2^N x 2^N operator column-major, by two
 *   rocBLAS GEMMs; the coherent half of lindblad_dm_step!
 *   (utils/cpu/cpuQuantumChannelLindbladianEvolutionDensityMatrix.jl:105-115). */
int smoq_dm_expect_pauli(const smoq_state* rho, int n_terms, const uint64_t* xmask,
                         const uint64_t* zmask, double* out);
int smoq_dm_apply_dense(smoq_state* rho, const smoq_state* u);
/* Z measurement / forced projection of a density matrix (vec(rho) with
 * 2N vec qubits; qubits 0-based below N), same list and uniform layout as
 * smoq_qubit_channel:
 *   SMOQ_QC_MEASURE_Z: _measure_z_single!(rho) (…Measurements.jl:259-295):
 *     b = u < P(0) ? 0 : 1 with P(0) the sum of Re rho_ii over bit k = 0,
 *     rho -> P_b rho P_b / prob with prob = P(0) or 1 - P(0), divided only
 *     when prob > 1e-15; record = b.
 *   SMOQ_QC_FORCE_ZERO / FORCE_ONE: _force_qubit_rho! (…Measurements.jl:686-708):
 *     P_b rho P_b divided by its trace when that exceeds 1e-15. */
int smoq_dm_qubit_channel(smoq_state* rho, int n_qubits, const int* qubits, int kind,
                          const double* uniforms, int32_t* out);

/* ---- circuits of gates, noise, measurements and resets ---------------------
 * smoq_apply_circuit replaces apply_circuit!(ψ, circuit, θ) and
 * apply_circuit!(ρ, circuit, θ) (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl:95-117,
 * 252-262) over the reference's CircuitOperation list
 * (utils/cpu/cpuVariationalQuantumCircuitBuilder.jl:118-132), one record per
 * operation; the op's qubits are qubits[q_first .. q_first + q_count - 1],
 * 0-based.
 * State vector (density_matrix = 0, any batch): every maximal run of gates
 * is one smoq_apply_tape; noise (apply_noise_mcwf!, :187-204; phase damping
 * acts as dephasing, thermal relaxation does nothing), measurement
 * (projective_measurement! with the op's basis, outcomes recorded) and reset
 * (reset_state!) are smoq_qubit_channel calls.  Each of those ops takes
 * q_count * batch uniforms from `uniforms`, in op order, laid out like
 * smoq_qubit_channel's ([k * batch + b]); running out returns SMOQ_EINVAL
 * with the ops before it applied.  Outcome i (0-based, one per measured
 * qubit) of item b goes to outcomes[i * batch + b]; if_outcome = k > 0 runs
 * the op only when outcome k (1-based) is 1, as the reference does, and
 * needs batch 1.  trace_out and lindblad_step are skipped (the reference
 * warns and skips).
 * Density matrix (density_matrix = 1, s = vec(ρ) of 2N qubits): the whole
 * circuit is one tape of maps on vec(ρ), with the reference's ρ executor:
 * gates as U ρ U† except Y, S and T, which its apply_gate_rho! leaves out;
 * per qubit depolarizing (1-p)ρ + (p/3)(XρX + YρY + ZρZ), bit flip
 * (1-p)ρ + pXρX, dephasing (1-p/2)ρ + (p/2)ZρZ, amplitude damping by the
 * first-order σ⁻ dissipator (apply_sigma_minus_dissipator_rho!); phase
 * damping and thermal relaxation nothing; measure and reset zero the
 * qubit's coherences; no uniforms, no outcomes, if_outcome not consulted;
 * trace_out is an unknown operation (SMOQ_EINVAL, nothing applied). */
enum smoq_circuit_kind {
  SMOQ_CK_RX = 0,
  SMOQ_CK_RY = 1,
  SMOQ_CK_RZ = 2,
  SMOQ_CK_H = 3,
  SMOQ_CK_X = 4,
  SMOQ_CK_Y = 5,
  SMOQ_CK_Z = 6,
  SMOQ_CK_S = 7,
  SMOQ_CK_T = 8,
  SMOQ_CK_CZ = 9,
  SMOQ_CK_CNOT = 10, /* also :cx; qubits (control, target) */
  SMOQ_CK_RXX = 11,
  SMOQ_CK_RYY = 12,
  SMOQ_CK_RZZ = 13,
  SMOQ_CK_DEPOLARIZING = 16,
  SMOQ_CK_AMPLITUDE_DAMPING = 17,
  SMOQ_CK_DEPHASING = 18,
  SMOQ_CK_BIT_FLIP = 19,
  SMOQ_CK_PHASE_DAMPING = 20,
  SMOQ_CK_THERMAL_RELAXATION = 21,
  SMOQ_CK_MEASURE = 32,   /* arg: basis 0 z, 1 x, 2 y */
  SMOQ_CK_RESET = 33,     /* arg: target 0 zero, 1 one, 2 plus, 3 minus */
  SMOQ_CK_TRACE_OUT = 34,
  SMOQ_CK_LINDBLAD_STEP = 35
};
typedef struct smoq_circuit_op {
  int32_t kind;
  int32_t param;       /* 0-based index into theta (the reference's param_idx - 1), -1: angle */
  int32_t q_first;
  int32_t q_count;
  int32_t if_outcome;  /* 0: always */
  int32_t arg;         /* measurement basis or reset target */
  double angle;        /* fixed_angle */
  double p;            /* noise probability */
} smoq_circuit_op;
int smoq_apply_circuit(smoq_state* s, int density_matrix, const smoq_circuit_op* ops, int n_ops,
                       const int32_t* qubits, int n_qubit_entries, const double* theta,
                       int n_theta, const double* uniforms, int64_t n_uniforms,
                       int32_t* outcomes, int max_outcomes, int* n_outcomes,
                       int64_t* n_uniforms_used);

/* ---- Monte Carlo wave-function trajectories (σ⁻ jumps) -------------------
 * One group = the trajectories of one qubit count: `psi` holds them as batch
 * items.  Each step of lindblad_mcwf_batched_step!
 * (utils/cpu/cpuQuantumChannelLindbladianEvolutionMonteCarloWaveFunction.jl:340-390)
 * is applied to every trajectory:
 *   1. the first-order Trotter step: the diagonal 2-qubit bond gates in
 *      order, then the 2x2 site gates (the TFIM gate list of
 *      precompute_trotter_gates_bitwise_cpu; bonds must be diagonal);
 *   2. ψ_i *= decay_pow[popcount(i)]   (decay_pow[m] = sqrt(1-γdt)^m);
 *   3. dp = 1 - ||ψ||²; with r = uniforms[traj*n_steps + step]: if r < dp
 *      and dp > 0, site k is the first with r < Σ_{j<=k} γdt pop_j/||ψ||²
 *      (site 0 when none), σ⁻_k is applied and jumps[...] = k, else -1;
 *   4. ψ /= ||ψ||.
 * After every obs_every-th step z_out[(traj*(n_steps/obs_every) + r)*n + i]
 * receives <Z_i>.  All groups run in one persistent launch, one workgroup
 * per trajectory, largest qubit count first.  n <= 20. */
typedef struct smoq_mcwf_group {
  smoq_state* psi;
  int n_bonds;
  const int32_t* bond_q;      /* 2 per bond, 0-based; basis b(q0) + 2 b(q1) */
  const double* bond_diag;    /* 4 complex per bond: the bond gate's diagonal */
  const double* site_u;       /* n x 2x2 complex (row-major), one per site */
  const double* decay_pow;    /* n + 1 values */
  const double* uniforms;     /* [traj][n_steps] */
  int32_t* jumps;             /* out [traj][n_steps] */
  double* z_out;              /* out [traj][n_steps/obs_every][n], may be NULL */
} smoq_mcwf_group;
int smoq_mcwf_evolve(smoq_ctx* ctx, const smoq_mcwf_group* groups, int n_groups, double gamma_dt,
                     int n_steps, int obs_every);

/* MCWF with any coherent step: the dissipative half of one step, for every
 * batch item of s (one trajectory each), after the coherent step has been
 * applied (e.g. a compiled plan).  Jump operators on every site:
 *  SMOQ_JUMP_SIGMA_MINUS — lindblad_mcwf_batched_step!
 *    (…MonteCarloWaveFunction.jl:352-386): ψ_i *= decay_pow[popcount(i)]
 *    (decay_pow[m] = sqrt(1-γdt)^m), site = the first k with
 *    r < Σ_{j<=k} γdt pop_j/||ψ||² (0 when none), σ⁻_k;
 *  SMOQ_JUMP_SIGMA_Z — lindblad_mcwf_step_trotter! (…:279-306): ψ *=
 *    decay_pow[0] (= sqrt(1-nγdt)), site = min(n-1, floor(u2·n)) (the
 *    reference draws rand(1:N)), σ_z_k flips the sign where bit k is set.
 * Both: dp = 1 - ||ψ||², jump iff r < dp and dp > 0, then ψ /= ||ψ||.
 * uniforms[2b] = r and uniforms[2b+1] = u2 for item b; jumps[b] = the
 * 0-based site or -1.  A zero or non-finite norm returns SMOQ_ENUMERIC. */
enum smoq_jump_kind { SMOQ_JUMP_SIGMA_MINUS = 0, SMOQ_JUMP_SIGMA_Z = 1 };
int smoq_mcwf_jump(smoq_state* s, int kind, double gamma_dt, const double* decay_pow,
                   const double* uniforms, int32_t* jumps);
/* n_steps of [the plan, then smoq_mcwf_jump] on every trajectory of s;
 * after every obs_every-th step <P_t> per trajectory is recorded as in
 * smoq_evolve_observe: out[r][item][t].  uniforms[item][step][2],
 * jumps[item][step].  This is evolve_mcwf_batched_trotter!
 * (…:412-423) with apply_trotter! = the plan. */
int smoq_mcwf_trotter_evolve(smoq_plan* p, smoq_state* s, int kind, double gamma_dt,
                             const double* decay_pow, int n_steps, const double* uniforms,
                             int32_t* jumps, int obs_every, int n_terms, const uint64_t* xmask,
                             const uint64_t* zmask, double* out);

/* ---- parametrised circuits: adjoint gradient -----------------------------
 * E_b(θ) = <0|U(θ_b)† H U(θ_b)|0> for every batch item b of `s`, with
 * H = sum_t coeff_t P_t (Pauli masks as in smoq_expect_pauli) and the tape's
 * rotation gates (RX, RY, RZ, RXX, RYY, RZZ, CRY) taking theta[b*n_params +
 * op.param] when op.param >= 0 (R(θ) = exp(-iθG/2)).  One forward sweep,
 * λ = Hψ, then one reverse sweep that un-applies each gate to ψ and λ and,
 * for a parametrised gate, accumulates dE/dθ = Im<λ|G|ψ> in the same pass.
 * On return s holds U(θ_b)|0> per item; energy[b], grad[b*n_params + p].
 * Replaces calculate_gradients (Enzyme) and calculate_gradients_parameter_shift
 * (utils/cpu/cpuVariationalQuantumCircuitGradients.jl:88-110,138-163). */
int smoq_adjoint_gradient(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                          int n_params, int n_terms, const uint64_t* xmask,
                          const uint64_t* zmask, const double* coeff, double* energy,
                          double* grad);
/* <H> after the tape for every item: the cost a parameter-shift or finite
 * difference evaluates (per-item theta as above; s receives U(θ_b)|0>).
 * flags & SMOQ_ENERGY_ONE_GATE_KERNELS: apply the tape gate by gate with the
 * one-gate kernels instead of fused passes. */
#define SMOQ_ENERGY_ONE_GATE_KERNELS 1
int smoq_tape_energy(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                     int n_params, int n_terms, const uint64_t* xmask, const uint64_t* zmask,
                     const double* coeff, int flags, double* energy);

/* Adjoint gradients of many parameter sets in one persistent launch; each
 * item has its own qubit count, tape, angles and Hamiltonian, so VQE runs of
 * different sizes (the parameter loop of
 * scripts_variational_quantum_circuits/demo_variational_quantum_eigensolver.jl:175-289
 * with calculate_gradients, utils/cpu/cpuVariationalQuantumCircuitGradients.jl:88-163)
 * share the device.  Per item, from |0…0> of n_qubits: the forward sweep,
 * W = (ψ, Hψ) interleaved, the reverse sweep with its overlaps, the
 * per-parameter sums, all by one workgroup that owns the item (the items
 * are dealt to the workgroups largest first); one launch runs every item.
 * Large items go instead, by a cost model, to the whole-device sweep
 * (smoq_adjoint_gradient) on a second stream while the launch runs; large
 * items with equal tape and Hamiltonian (compared by contents) run there as
 * one batch of parameter sets.  SMOQ_VQE_ROUTE=one_launch / sweep forces
 * either path.
 * Outputs as smoq_adjoint_gradient for one item: *energy and grad[p].
 * n_qubits 3 .. 30; the states live in the context's work buffers. */
typedef struct smoq_vqe_item {
  int32_t n_qubits;
  int32_t n_ops;
  const smoq_op* ops;
  int32_t n_params;
  int32_t n_terms;
  const double* theta;
  const uint64_t* xmask;
  const uint64_t* zmask;
  const double* coeff;
  double* energy;
  double* grad;
} smoq_vqe_item;
int smoq_adjoint_gradient_batch(smoq_ctx* ctx, int dtype, const smoq_vqe_item* items,
                                int n_items);

/* ---- observables --------------------------------------------------------
 * <P_t> for Pauli strings P_t = i^{|x_t & z_t|} X^{x_t} Z^{z_t} (x/z bit
 * masks; Y on qubit q sets bit q in both).  out has batch*n_terms entries
 * (item-major).  True expectation values: for strings with X/Y this is twice
 * what get_expectation_pauli_string returns (that function omits the
 * factor 2, utils/cpu/cpuQuantumStateObservables.jl:1044).
 * Replaces expect_local / expect_corr / _expect_pauli_Z_string /
 * _expect_pauli_mixed_string (cpuQuantumStateObservables.jl:175-445,840-1049)
 * and compute_energy term by term
 * (scripts_variational_quantum_circuits/demo_variational_quantum_eigensolver.jl:251-289). */
int smoq_expect_pauli(smoq_state* s, int n_terms, const uint64_t* xmask,
                      const uint64_t* zmask, double* out);
/* The value get_expectation_pauli_string returns
 * (utils/cpu/cpuQuantumStateObservables.jl:840-1049): for x_t = 0 the true
 * <Z^z>; for x_t != 0 only bras with every X/Y bit 0 paired with the ket
 * that has every X/Y bit 1,
 *   Re[(-i)^{|x&z|} sum_{i & x = 0} (-1)^{|i & z|} conj(psi_i) psi_{i|x}]
 * (= <P>/2 for a single X/Y operator; its :1044).  Kept for drop-in parity. */
int smoq_pauli_string_reference(smoq_state* s, int n_terms, const uint64_t* xmask,
                                const uint64_t* zmask, double* out);
/* sum_i |psi_i|^2 per batch item (FP64 accumulation). */
int smoq_norm2(smoq_state* s, double* out);
/* smoq_expect_pauli without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and results as smoq_expect_pauli; the partial sums live in a
 * fixed 16 MiB region that each device holds for the life of the process
 * (about batch * 256 KiB + batch * tiles * terms * 8 B are needed; larger
 * requests return SMOQ_ENOMEM without running).  Calls that overlap on one
 * device share that region and race: serialise them. */
int smoq_expect_pauli_noalloc(smoq_state* s, int n_terms, const uint64_t* xmask,
                              const uint64_t* zmask, double* out);
/* ---- measurement sampling ---------------------------------------------- */
enum smoq_sample_semantics {
  /* idx = searchsortedfirst(cumsum(abs2.(ψ)), r) - 1 with Julia's pairwise
   * cumsum, run_benchmark_smoq_Float64.jl:424-429.  When r exceeds the last
   * prefix the reference gets idx = 2^N, whose N-bit string is all zeros;
   * out_index then holds 2^N. */
  SMOQ_SAMPLE_CUMSUM_SEARCH = 0,
  /* first i with r <= sequential prefix, else the last index:
   * sample_from_probabilities, utils/cpu/cpuQuantumStateMeasurements.jl:203-213
   * (0-based result). */
  SMOQ_SAMPLE_SEQUENTIAL = 1,
  /* the same rule over the populations real(rho_ii) of a density matrix
   * stored as vec(rho) (2N vec qubits): projective_measurement_all!(rho) and
   * sample_state(rho) (…Measurements.jl:435-446, 505-520); out_index in [0, 2^N). */
  SMOQ_SAMPLE_SEQUENTIAL_DIAG = 2
};
/* One draw per uniform r[d] from batch item 0; out_index[d] 0-based. */
int smoq_sample(smoq_state* s, const double* uniforms, int64_t n_draws,
                int semantics, uint64_t* out_index);

#ifdef __cplusplus
}
#endif

#endif /* SMOQ_H */
