// Compiled tapes: planned once, program resident on the device, launches
// replayed through a hipGraph.
#include "device_common.h"
#include "smoq_internal.h"

struct smoq_plan_s {
  smoq_ctx* ctx = nullptr;
  int n = 0;
  int dtype = SMOQ_C128;
  smoq::Program prog;
  smoq::ProgLayout layout;
  void* dev = nullptr;  // device image of prog
  std::vector<smoq::IGate> small;  // n < kRegQubits: replayed gate by gate
  std::shared_ptr<smoq::JitModule> jit;  // compiled pass kernels, when used
  // captured launches of one application, valid for one (state, batch)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* graph_state = nullptr;
  int graph_batch = 0;
};

namespace smoq {

static void drop_graph(smoq_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  p->graph_state = nullptr;
}

static int run_once(smoq_plan* p, smoq_state* s) {
  if (!p->small.empty()) {
    for (auto& g : p->small) {
      int rc = launch_gate(s, g);
      if (rc) return rc;
    }
    return SMOQ_OK;
  }
  return launch_program(s, p->prog, p->layout, p->dev, p->jit.get());
}

// capture one application of the plan on state s into an executable graph
static int capture(smoq_plan* p, smoq_state* s) {
  drop_graph(p);
  hipStream_t st = p->ctx->stream;
  SMOQ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = run_once(p, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  p->graph = g;
  SMOQ_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
  p->graph_state = s->d;
  p->graph_batch = s->batch;
  return SMOQ_OK;
}

static int plan_run(smoq_plan* p, smoq_state* s, int repeat) {
  if (repeat <= 0) return SMOQ_OK;
  // pass timing records events between launches, which a graph cannot hold
  if (p->ctx->timing) {
    for (int r = 0; r < repeat; ++r) {
      int rc = run_once(p, s);
      if (rc) return rc;
    }
    return SMOQ_OK;
  }
  if (!p->exec || p->graph_state != s->d || p->graph_batch != s->batch) {
    int rc = capture(p, s);
    if (rc) return rc;
  }
  for (int r = 0; r < repeat; ++r) SMOQ_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" {

int smoq_plan_create(smoq_ctx* ctx, int n_qubits, int dtype, const smoq_op* ops, int n_ops,
                     const double* theta, int n_theta, smoq_plan** out) {
  if (!ctx || !out) return set_error(SMOQ_EINVAL, "smoq_plan_create: null argument");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 36)
    return set_error(SMOQ_EINVAL, "smoq_plan_create: n_qubits %d outside [1, 36]", n_qubits);
  if (dtype != SMOQ_C128 && dtype != SMOQ_C64)
    return set_error(SMOQ_EINVAL, "smoq_plan_create: unknown dtype %d", dtype);
  std::vector<IGate> gs;
  int rc = lower_tape(ops, n_ops, theta, n_theta, n_qubits, gs);
  if (rc) return rc;
  smoq_plan* p = new smoq_plan();
  p->ctx = ctx;
  p->n = n_qubits;
  p->dtype = dtype;
  if (n_qubits < kRegQubits) {
    p->small = gs;
    ctx->live_objects++;
    *out = p;
    return SMOQ_OK;
  }
  const Fusion fu = fusion_for(ctx, n_qubits, dtype, 1, gs);
  plan_program(gs, n_qubits, fu.tile, fu.low,
               dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits, p->prog, fu.cap,
               fu.swaps != 0,
               fu.locals.empty() ? nullptr : &fu.locals);
  apply_pass_io(p->prog, fu.io, fu.low);
  p->layout = program_layout(p->prog, dtype);
  if (jit_wanted(ctx, p->prog, n_qubits)) {
    Program factored = p->prog;
    factor_rotations(factored, dtype);
    rc = jit_get(ctx, factored, dtype, p->jit);
    // automatic mode keeps the precompiled kernel when compilation fails,
    // unless the plan exchanges lanes (compiled passes only)
    if (rc && (ctx->jit_mode >= 2 || program_has_swaps(p->prog))) {
      delete p;
      return rc;
    }
    if (!rc) p->prog = std::move(factored);
  }
  hipError_t e = hipMalloc(&p->dev, p->layout.total);
  if (e != hipSuccess) {
    delete p;
    return hip_error(e, "hipMalloc(plan)");
  }
  if ((rc = upload_program(ctx, p->prog, dtype, p->layout, p->dev))) {
    (void)hipFree(p->dev);
    delete p;
    return rc;
  }
  ctx->live_objects++;
  *out = p;
  return SMOQ_OK;
}

int smoq_plan_passes(const smoq_plan* p, int* n_passes) {
  if (!p || !n_passes) return set_error(SMOQ_EINVAL, "smoq_plan_passes: null argument");
  *n_passes = p->small.empty() ? (int)p->prog.passes.size() : (int)p->small.size();
  return SMOQ_OK;
}

int smoq_plan_run(smoq_plan* p, smoq_state* s, int repeat) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!p) return set_error(SMOQ_EINVAL, "smoq_plan_run: null plan");
  if (s->n != p->n || s->dtype != p->dtype || s->ctx != p->ctx)
    return set_error(SMOQ_EINVAL, "smoq_plan_run: plan is for %d qubits dtype %d", p->n,
                     p->dtype);
  return plan_run(p, s, repeat);
}

int smoq_plan_free(smoq_plan* p) {
  if (!p) return SMOQ_OK;
  p->ctx->live_objects--;
  (void)hipStreamSynchronize(p->ctx->stream);
  drop_graph(p);
  if (p->dev) (void)hipFree(p->dev);
  delete p;
  return SMOQ_OK;
}

int smoq_evolve_observe(smoq_plan* p, smoq_state* s, int n_steps, int obs_every, int n_terms,
                        const uint64_t* xmask, const uint64_t* zmask, double* out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!p) return set_error(SMOQ_EINVAL, "smoq_evolve_observe: null plan");
  if (n_steps < 0 || obs_every < 1)
    return set_error(SMOQ_EINVAL, "smoq_evolve_observe: n_steps %d, obs_every %d", n_steps,
                     obs_every);
  if (s->n != p->n || s->dtype != p->dtype || s->ctx != p->ctx)
    return set_error(SMOQ_EINVAL,
                     "smoq_evolve_observe: plan/state shape or context mismatch");
  const int records = n_steps / obs_every;
  for (int r = 0; r < records; ++r) {
    if ((rc = plan_run(p, s, obs_every))) return rc;
    if (n_terms > 0 &&
        (rc = reduce_pauli(s, n_terms, xmask, zmask, out + (size_t)r * s->batch * n_terms, 0)))
      return rc;
  }
  return plan_run(p, s, n_steps - records * obs_every);
}

int smoq_mcwf_trotter_evolve(smoq_plan* p, smoq_state* s, int kind, double gamma_dt,
                             const double* decay_pow, int n_steps, const double* uniforms,
                             int32_t* jumps, int obs_every, int n_terms, const uint64_t* xmask,
                             const uint64_t* zmask, double* out) {
  int rc = mcwf_jump_check(s, kind, gamma_dt);
  if (rc) return rc;
  if (!p) return set_error(SMOQ_EINVAL, "smoq_mcwf_trotter_evolve: null plan");
  if (s->n != p->n || s->dtype != p->dtype || s->ctx != p->ctx)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_trotter_evolve: plan is for %d qubits dtype %d",
                     p->n, p->dtype);
  if (n_steps < 0 || obs_every < 0)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_trotter_evolve: n_steps %d obs_every %d", n_steps,
                     obs_every);
  if (n_steps > 0 && (!decay_pow || !uniforms || !jumps))
    return set_error(SMOQ_EINVAL, "smoq_mcwf_trotter_evolve: null array");
  const bool record = obs_every > 0 && n_terms > 0;
  if (record && (!xmask || !zmask || !out))
    return set_error(SMOQ_EINVAL, "smoq_mcwf_trotter_evolve: null observable arrays");
  if (n_steps == 0) return SMOQ_OK;
  smoq_ctx* ctx = s->ctx;
  const size_t nb = (size_t)s->batch, ns = (size_t)n_steps;
  const size_t off_u = ((size_t)(s->n + 1) * sizeof(double) + 255) & ~(size_t)255;
  const size_t off_j = off_u + ((2 * nb * ns * sizeof(double) + 255) & ~(size_t)255);
  const size_t off_f = off_j + ((nb * ns * sizeof(int32_t) + 255) & ~(size_t)255);
  const size_t off_p = off_f + 256;
  unsigned char* dev = nullptr;
  SMOQ_HIP(hipMalloc(&dev, off_p + mcwf_jump_partial_bytes(s)));
  const unsigned long long no_fault = ~0ull;
  unsigned long long fault = no_fault;
  hipError_t e = hipMemcpyAsync(dev, decay_pow, (size_t)(s->n + 1) * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev + off_u, uniforms, 2 * nb * ns * sizeof(double),
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev + off_f, &no_fault, sizeof(no_fault), hipMemcpyHostToDevice,
                       ctx->stream);
  for (int step = 0; e == hipSuccess && rc == SMOQ_OK && step < n_steps; ++step) {
    if ((rc = plan_run(p, s, 1))) break;
    if ((rc = mcwf_jump_launch(s, kind, gamma_dt, (const double*)dev,
                               (const double*)(dev + off_u) + 2 * step, 2 * n_steps,
                               (int32_t*)(dev + off_j) + step, n_steps,
                               (unsigned long long*)(dev + off_f), n_steps, step,
                               (double*)(dev + off_p))))
      break;
    if (record && (step + 1) % obs_every == 0) {
      const size_t r = (size_t)(step + 1) / obs_every - 1;
      rc = reduce_pauli(s, n_terms, xmask, zmask, out + r * nb * n_terms, 0);
    }
  }
  if (e == hipSuccess && rc == SMOQ_OK)
    e = hipMemcpyAsync(jumps, dev + off_j, nb * ns * sizeof(int32_t), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess && rc == SMOQ_OK)
    e = hipMemcpyAsync(&fault, dev + off_f, sizeof(fault), hipMemcpyDeviceToHost, ctx->stream);
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (rc) return rc;
  if (e != hipSuccess) return hip_error(e, "smoq_mcwf_trotter_evolve: copy");
  if (se != hipSuccess) return hip_error(se, "smoq_mcwf_trotter_evolve");
  if (fault != no_fault)
    return set_error(SMOQ_ENUMERIC,
                     "smoq_mcwf_trotter_evolve: zero or non-finite norm at step %llu of "
                     "trajectory %llu",
                     fault % ns, fault / ns);
  return SMOQ_OK;
}

}  // extern "C"
