// Circuits of gates, noise channels, measurements and resets: the executor
// of the reference's variational circuits
// (utils/cpu/cpuVariationalQuantumCircuitExecutor.jl) over the C ABI's own
// entry points.
//
// On a state vector the gates between two non-gate operations form one tape
// (one set of fused passes); every noise, measurement and reset operation is
// a smoq_qubit_channel call, which returns its records to the host, so a
// later if_outcome condition is decided with the outcomes the reference
// would have at that point.  On a density matrix every operation is a fixed
// linear map of vec(ρ), so the whole circuit is one tape.
#include <cstring>
#include <vector>

#include "smoq_internal.h"

namespace {

using namespace smoq;

bool ck_gate(int k) { return k >= SMOQ_CK_RX && k <= SMOQ_CK_RZZ; }
bool ck_noise(int k) { return k >= SMOQ_CK_DEPOLARIZING && k <= SMOQ_CK_THERMAL_RELAXATION; }
bool ck_rotation(int k) {
  return k == SMOQ_CK_RX || k == SMOQ_CK_RY || k == SMOQ_CK_RZ || k == SMOQ_CK_RXX ||
         k == SMOQ_CK_RYY || k == SMOQ_CK_RZZ;
}
int ck_arity(int k) {
  return (k == SMOQ_CK_CZ || k == SMOQ_CK_CNOT || k == SMOQ_CK_RXX || k == SMOQ_CK_RYY ||
          k == SMOQ_CK_RZZ)
             ? 2
             : 1;
}
bool ck_known(int k) {
  return ck_gate(k) || ck_noise(k) || (k >= SMOQ_CK_MEASURE && k <= SMOQ_CK_LINDBLAD_STEP);
}

int ck_gate_code(int k) {
  static const int code[] = {SMOQ_G_RX, SMOQ_G_RY,   SMOQ_G_RZ,  SMOQ_G_H,   SMOQ_G_X,
                             SMOQ_G_Y,  SMOQ_G_Z,    SMOQ_G_S,   SMOQ_G_T,   SMOQ_G_CZ,
                             SMOQ_G_CNOT, SMOQ_G_RXX, SMOQ_G_RYY, SMOQ_G_RZZ};
  return code[k];
}

smoq_op make_op(int gate, int q0, int q1 = -1, int param = -1, double theta = 0.0) {
  smoq_op o;
  std::memset(&o, 0, sizeof o);
  o.gate = gate;
  o.q[0] = q0;
  o.q[1] = q1;
  o.q[2] = -1;
  o.param = param;
  o.theta = theta;
  return o;
}

// the 4x4 map on (row bit, column bit) of vec(ρ), basis index b(row) + 2 b(col)
smoq_op pair_u2(int q, int qc, const double (&M)[4][4]) {
  smoq_op o = make_op(SMOQ_G_U2, q, qc);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o.m[2 * (r * 4 + c)] = M[r][c];
  return o;
}

smoq_op pair_diag(int q, int qc, double coherence) {
  smoq_op o = make_op(SMOQ_G_DIAG2, q, qc);
  const double d[4] = {1.0, coherence, coherence, 1.0};
  for (int e = 0; e < 4; ++e) o.m[2 * e] = d[e];
  return o;
}

int validate(const smoq_state* s, int dm, const smoq_circuit_op* ops, int n_ops,
             const int32_t* qubits, int n_qe, int n_theta) {
  if (n_ops < 0 || (n_ops > 0 && !ops))
    return set_error(SMOQ_EINVAL, "smoq_apply_circuit: bad operation list");
  if (n_qe < 0 || (n_qe > 0 && !qubits))
    return set_error(SMOQ_EINVAL, "smoq_apply_circuit: bad qubit list");
  if (dm && s->n % 2)
    return set_error(SMOQ_EINVAL, "smoq_apply_circuit: a density matrix has an even number of "
                                  "vec qubits, got %d", s->n);
  const int N = dm ? s->n / 2 : s->n;
  for (int i = 0; i < n_ops; ++i) {
    const smoq_circuit_op& op = ops[i];
    const int k = op.kind;
    if (!ck_known(k))
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d has unknown kind %d", i, k);
    if (dm && k == SMOQ_CK_TRACE_OUT)
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: Unknown operation type: trace_out "
                                    "(op %d, density matrix)", i);
    if (op.q_first < 0 || op.q_count < 0 || (int64_t)op.q_first + op.q_count > n_qe)
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d qubit range outside the list", i);
    for (int j = 0; j < op.q_count; ++j) {
      const int q = qubits[op.q_first + j];
      if (q < 0 || q >= N)
        return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d qubit %d outside 0..%d", i, q,
                         N - 1);
    }
    if (ck_gate(k)) {
      if (op.q_count != ck_arity(k))
        return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d gate takes %d qubits, got %d", i,
                         ck_arity(k), op.q_count);
      if (op.q_count == 2 && qubits[op.q_first] == qubits[op.q_first + 1])
        return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d acts twice on qubit %d", i,
                         qubits[op.q_first]);
      if (ck_rotation(k) && (op.param < -1 || op.param >= n_theta))
        return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d parameter %d outside 0..%d", i,
                         op.param, n_theta - 1);
    }
    if (ck_noise(k) && !(op.p >= 0.0 && op.p <= 1.0))
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d probability %g outside [0, 1]", i,
                       op.p);
    if (k == SMOQ_CK_MEASURE && (op.arg < 0 || op.arg > 2))
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d unknown basis %d", i, op.arg);
    if (k == SMOQ_CK_RESET && (op.arg < 0 || op.arg > 3))
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d unknown reset target %d", i, op.arg);
    if (op.if_outcome < 0)
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: op %d if_outcome %d < 0", i,
                       op.if_outcome);
    if (!dm && op.if_outcome > 0 && s->batch != 1)
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: if_outcome needs batch 1, got %d",
                       s->batch);
  }
  return SMOQ_OK;
}

int run_gates(smoq_state* s, std::vector<smoq_op>& run, const double* theta, int n_theta) {
  if (run.empty()) return SMOQ_OK;
  const int rc = smoq_apply_tape(s, run.data(), (int)run.size(), theta, n_theta, nullptr);
  run.clear();
  return rc;
}

// apply_circuit!(ψ, circuit, θ) (executor :95-117) with apply_operation! (:124-140)
int circuit_psi(smoq_state* s, const smoq_circuit_op* ops, int n_ops, const int32_t* qubits,
                const double* theta, int n_theta, const double* uniforms, int64_t n_uni,
                int32_t* outcomes, int max_out, int& n_out, int64_t& used) {
  const int64_t B = s->batch;
  std::vector<smoq_op> run;
  int rc;
  auto take = [&](int count, const double** u) {
    const int64_t m = (int64_t)count * B;
    if (m > 0 && (!uniforms || used + m > n_uni))
      return set_error(SMOQ_EINVAL, "smoq_apply_circuit: uniforms exhausted (%lld given, %lld "
                                    "needed so far)", (long long)n_uni, (long long)(used + m));
    *u = uniforms ? uniforms + used : nullptr;
    used += m;
    return (int)SMOQ_OK;
  };
  auto one_gate = [&](int gate, int q) {
    smoq_op o = make_op(gate, q);
    return smoq_apply_tape(s, &o, 1, nullptr, 0, nullptr);
  };
  for (int i = 0; i < n_ops; ++i) {
    const smoq_circuit_op& op = ops[i];
    if (op.if_outcome > 0 && (op.if_outcome > n_out || outcomes[op.if_outcome - 1] != 1))
      continue;
    const int32_t* q = qubits + op.q_first;
    const int k = op.kind;
    if (ck_gate(k)) {
      // apply_gate_psi! (:146-181): θ[param_idx] or the fixed angle
      const bool rot = ck_rotation(k);
      run.push_back(make_op(ck_gate_code(k), q[0], ck_arity(k) == 2 ? q[1] : -1,
                            rot ? op.param : -1, rot ? op.angle : 0.0));
      continue;
    }
    if ((rc = run_gates(s, run, theta, n_theta))) return rc;
    const double* u = nullptr;
    if (ck_noise(k)) {
      // apply_noise_mcwf! (:187-204)
      int qc;
      switch (k) {
        case SMOQ_CK_DEPOLARIZING: qc = SMOQ_QC_DEPOLARIZING; break;
        case SMOQ_CK_AMPLITUDE_DAMPING: qc = SMOQ_QC_AMPLITUDE_DAMPING; break;
        case SMOQ_CK_BIT_FLIP: qc = SMOQ_QC_BIT_FLIP; break;
        case SMOQ_CK_DEPHASING:
        case SMOQ_CK_PHASE_DAMPING: qc = SMOQ_QC_DEPHASING; break;
        default: continue;  // thermal relaxation has no branch
      }
      if ((rc = take(op.q_count, &u))) return rc;
      if ((rc = smoq_qubit_channel(s, op.q_count, q, qc, op.p, u, nullptr))) return rc;
    } else if (k == SMOQ_CK_MEASURE) {
      // apply_measure_psi! -> projective_measurement! (…Measurements.jl:319-345)
      if (n_out + op.q_count > max_out || (op.q_count > 0 && !outcomes))
        return set_error(SMOQ_EINVAL, "smoq_apply_circuit: more than %d outcomes", max_out);
      if ((rc = take(op.q_count, &u))) return rc;
      int32_t* rec = outcomes + (int64_t)n_out * B;
      if (op.arg == 0) {
        if ((rc = smoq_qubit_channel(s, op.q_count, q, SMOQ_QC_MEASURE_Z, 0.0, u, rec))) return rc;
      } else {
        for (int j = 0; j < op.q_count; ++j) {
          if (op.arg == 2 && (rc = one_gate(SMOQ_G_SDG, q[j]))) return rc;
          if ((rc = one_gate(SMOQ_G_H, q[j]))) return rc;
          if ((rc = smoq_qubit_channel(s, 1, q + j, SMOQ_QC_MEASURE_Z, 0.0, u + j * B,
                                       rec + j * B)))
            return rc;
          if ((rc = one_gate(SMOQ_G_H, q[j]))) return rc;
          if (op.arg == 2 && (rc = one_gate(SMOQ_G_S, q[j]))) return rc;
        }
      }
      n_out += op.q_count;
    } else if (k == SMOQ_CK_RESET) {
      // apply_reset_psi! -> reset_state! (…Measurements.jl:555-626): per
      // qubit a Z measurement, the forced bit, H for :plus / :minus
      if ((rc = take(op.q_count, &u))) return rc;
      for (int j = 0; j < op.q_count; ++j) {
        if ((rc = smoq_qubit_channel(s, 1, q + j, SMOQ_QC_MEASURE_Z, 0.0, u + j * B, nullptr)))
          return rc;
        const int force = (op.arg == 0 || op.arg == 2) ? SMOQ_QC_FORCE_ZERO : SMOQ_QC_FORCE_ONE;
        if ((rc = smoq_qubit_channel(s, 1, q + j, force, 0.0, nullptr, nullptr))) return rc;
        if (op.arg >= 2 && (rc = one_gate(SMOQ_G_H, q[j]))) return rc;
      }
    }
    // trace_out and lindblad_step: skipped (:230-246)
  }
  return run_gates(s, run, theta, n_theta);
}

// apply_circuit!(ρ, circuit, θ) (executor :252-262, 267-427) as one tape on vec(ρ)
int circuit_rho(smoq_state* s, const smoq_circuit_op* ops, int n_ops, const int32_t* qubits,
                const double* theta, int n_theta) {
  const int N = s->n / 2;
  std::vector<smoq_op> t;
  for (int i = 0; i < n_ops; ++i) {
    const smoq_circuit_op& op = ops[i];
    const int32_t* q = qubits + op.q_first;
    const int k = op.kind;
    if (ck_gate(k)) {
      // apply_gate_rho! (:290-318): no branch for y, s, t
      if (k == SMOQ_CK_Y || k == SMOQ_CK_S || k == SMOQ_CK_T) continue;
      const int g = ck_gate_code(k);
      const bool two = ck_arity(k) == 2;
      const bool rot = ck_rotation(k);
      t.push_back(make_op(g, q[0], two ? q[1] : -1, rot ? op.param : -1, rot ? op.angle : 0.0));
      // conj(U) on the column qubits: RY and the angle-free gates are real;
      // the others negate their angle (parameter n_theta + p holds -θ_p)
      const bool neg = rot && k != SMOQ_CK_RY;
      int cpar = rot ? op.param : -1;
      if (neg && cpar >= 0) cpar += n_theta;
      t.push_back(make_op(g, q[0] + N, two ? q[1] + N : -1, cpar,
                          rot ? (neg ? -op.angle : op.angle) : 0.0));
    } else if (ck_noise(k)) {
      // apply_noise_kraus! (:327-343) on each listed qubit
      const double p = op.p;
      for (int j = 0; j < op.q_count; ++j) {
        const int a = q[j], c = q[j] + N;
        if (k == SMOQ_CK_DEPOLARIZING) {  // (1-p)ρ + (p/3)(XρX + YρY + ZρZ), :346-368
          const double p3 = p / 3.0;
          const double M[4][4] = {{1 - p + p3, 0, 0, 2 * p3},
                                  {0, 1 - p - p3, 0, 0},
                                  {0, 0, 1 - p - p3, 0},
                                  {2 * p3, 0, 0, 1 - p + p3}};
          t.push_back(pair_u2(a, c, M));
        } else if (k == SMOQ_CK_BIT_FLIP) {  // (1-p)ρ + pXρX, :371-375
          const double M[4][4] = {{1 - p, 0, 0, p}, {0, 1 - p, p, 0}, {0, p, 1 - p, 0},
                                  {p, 0, 0, 1 - p}};
          t.push_back(pair_u2(a, c, M));
        } else if (k == SMOQ_CK_DEPHASING) {  // (1-p/2)ρ + (p/2)ZρZ, :378-382
          t.push_back(pair_diag(a, c, 1.0 - p));
        } else if (k == SMOQ_CK_AMPLITUDE_DAMPING) {
          // apply_sigma_minus_dissipator_rho! (cpuQuantumChannelGates.jl:833-854)
          const double M[4][4] = {{1, 0, 0, p}, {0, 1 - 0.5 * p, 0, 0}, {0, 0, 1 - 0.5 * p, 0},
                                  {0, 0, 0, 1 - p}};
          t.push_back(pair_u2(a, c, M));
        }
      }
    } else if (k == SMOQ_CK_MEASURE || k == SMOQ_CK_RESET) {
      // _decohere_qubit_rho! (:396-408); _reset_qubit_rho! only decoheres (:418-425)
      for (int j = 0; j < op.q_count; ++j) t.push_back(pair_diag(q[j], q[j] + N, 0.0));
    }
  }
  if (t.empty()) return SMOQ_OK;
  std::vector<double> th2(2 * (size_t)n_theta);
  for (int i = 0; i < n_theta; ++i) {
    th2[i] = theta[i];
    th2[n_theta + i] = -theta[i];
  }
  return smoq_apply_tape(s, t.data(), (int)t.size(), n_theta ? th2.data() : nullptr, 2 * n_theta,
                         nullptr);
}

}  // namespace

extern "C" int smoq_apply_circuit(smoq_state* s, int density_matrix, const smoq_circuit_op* ops,
                                  int n_ops, const int32_t* qubits, int n_qubit_entries,
                                  const double* theta, int n_theta, const double* uniforms,
                                  int64_t n_uniforms, int32_t* outcomes, int max_outcomes,
                                  int* n_outcomes, int64_t* n_uniforms_used) {
  int n_out = 0;
  int64_t used = 0;
  auto finish = [&](int code) {
    if (n_outcomes) *n_outcomes = n_out;
    if (n_uniforms_used) *n_uniforms_used = used;
    return code;
  };
  int rc = check_state(s);
  if (rc) return finish(rc);
  if (n_theta < 0 || (n_theta > 0 && !theta))
    return finish(set_error(SMOQ_EINVAL, "smoq_apply_circuit: bad parameter array"));
  if (n_uniforms < 0 || max_outcomes < 0)
    return finish(set_error(SMOQ_EINVAL, "smoq_apply_circuit: negative buffer size"));
  if ((rc = validate(s, density_matrix != 0, ops, n_ops, qubits, n_qubit_entries, n_theta)))
    return finish(rc);
  if (density_matrix) return finish(circuit_rho(s, ops, n_ops, qubits, theta, n_theta));
  return finish(circuit_psi(s, ops, n_ops, qubits, theta, n_theta, uniforms, n_uniforms, outcomes,
                            max_outcomes, n_out, used));
}
