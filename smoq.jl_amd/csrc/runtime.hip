// Context, state lifecycle, error plumbing and gate lowering of libsmoq_amd.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  int code = (e == hipErrorOutOfMemory) ? SMOQ_ENOMEM : SMOQ_EDEVICE;
  return set_error(code, "%s failed: %s", what, hipGetErrorString(e));
}

int ensure_scratch(smoq_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->scratch_bytes) return SMOQ_OK;
  if (ctx->scratch) {
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));
    SMOQ_HIP(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t want = bytes + bytes / 8 + 4096;
  hipError_t e = hipMalloc(&ctx->scratch, want);
  if (e != hipSuccess) {
    ctx->scratch = nullptr;
    return set_error(SMOQ_ENOMEM, "scratch allocation of %zu bytes failed: %s", want,
                     hipGetErrorString(e));
  }
  ctx->scratch_bytes = want;
  return SMOQ_OK;
}

int ensure_work(smoq_ctx* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= smoq_ctx::kWorkSlots)
    return set_error(SMOQ_EINVAL, "work buffer slot %d", slot);
  if (bytes > ctx->work_bytes[slot]) {
    if (ctx->work[slot]) {
      SMOQ_HIP(hipStreamSynchronize(ctx->stream));
      SMOQ_HIP(hipFree(ctx->work[slot]));
      ctx->work[slot] = nullptr;
      ctx->work_bytes[slot] = 0;
    }
    const size_t want = bytes + 4096;
    hipError_t e = hipMalloc(&ctx->work[slot], want);
    if (e != hipSuccess && !ctx->state_pool.empty()) {
      (void)hipGetLastError();
      release_state_pool(ctx);
      e = hipMalloc(&ctx->work[slot], want);
    }
    if (e != hipSuccess) {
      ctx->work[slot] = nullptr;
      return set_error(SMOQ_ENOMEM, "work buffer of %zu bytes: %s", want, hipGetErrorString(e));
    }
    ctx->work_bytes[slot] = want;
  }
  *out = ctx->work[slot];
  return SMOQ_OK;
}

void release_state_pool(smoq_ctx* ctx) {
  if (ctx->state_pool.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->state_pool) (void)hipFree(b.second);
  ctx->state_pool.clear();
}

int ensure_prog(smoq_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->prog_bytes) return SMOQ_OK;
  if (ctx->prog) {
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));
    SMOQ_HIP(hipFree(ctx->prog));
    ctx->prog = nullptr;
    ctx->prog_bytes = 0;
  }
  size_t want = bytes * 2 + 65536;
  SMOQ_HIP(hipMalloc(&ctx->prog, want));
  ctx->prog_bytes = want;
  return SMOQ_OK;
}

int ensure_pinned(smoq_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->host_pinned_bytes) return SMOQ_OK;
  if (ctx->host_pinned) {
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));
    SMOQ_HIP(hipHostFree(ctx->host_pinned));
    ctx->host_pinned = nullptr;
    ctx->host_pinned_bytes = 0;
  }
  size_t want = bytes + 65536;
  SMOQ_HIP(hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault));
  ctx->host_pinned_bytes = want;
  return SMOQ_OK;
}

int time_mark(smoq_ctx* ctx) {
  if (ctx->ev_used == ctx->ev_pool.size()) {
    if (ctx->ev_pool.size() >= 16384)
      return set_error(SMOQ_EINVAL, "pass timing: more than 8192 launches between reads");
    hipEvent_t e;
    SMOQ_HIP(hipEventCreate(&e));
    ctx->ev_pool.push_back(e);
  }
  SMOQ_HIP(hipEventRecord(ctx->ev_pool[ctx->ev_used++], ctx->stream));
  return SMOQ_OK;
}

int check_state(const smoq_state* s) {
  if (!s || !s->ctx || !s->d) return set_error(SMOQ_EINVAL, "null or freed state");
  return SMOQ_OK;
}

int check_qubit(const smoq_state* s, int q, const char* what) {
  if (q < 0 || q >= s->n)
    return set_error(SMOQ_EINVAL, "%s: qubit %d out of range [0, %d)", what, q, s->n);
  return SMOQ_OK;
}

// ---------------------------------------------------------------- lowering --
static void set_u1(IGate& g, int t, double a_re, double a_im, double b_re, double b_im,
                   double c_re, double c_im, double d_re, double d_im) {
  g.kind = IK_U1;
  g.t0 = t;
  double m[8] = {a_re, a_im, b_re, b_im, c_re, c_im, d_re, d_im};
  std::memcpy(g.m, m, sizeof(m));
}

static void set_diag(IGate& g, int nd, const int* q, const double* ph) {
  g.kind = IK_DIAG;
  g.nd = nd;
  for (int i = 0; i < nd; ++i) g.dq[i] = q[i];
  std::memcpy(g.m, ph, sizeof(double) * 2 * (1 << nd));
}

int lower_op(const smoq_op& op, double theta, int n, std::vector<IGate>& out) {
  auto qok = [&](int i) { return op.q[i] >= 0 && op.q[i] < n; };
  auto need = [&](int k) -> int {
    for (int i = 0; i < k; ++i)
      if (!qok(i))
        return set_error(SMOQ_EINVAL, "gate %d: qubit q[%d]=%d out of range [0, %d)", op.gate, i,
                         op.q[i], n);
    for (int i = 0; i < k; ++i)
      for (int j = i + 1; j < k; ++j)
        if (op.q[i] == op.q[j])
          return set_error(SMOQ_EINVAL, "gate %d: repeated qubit %d", op.gate, op.q[i]);
    return SMOQ_OK;
  };
  IGate g;
  const double h = 0.5 * theta;
  const double c = std::cos(h), sn = std::sin(h);
  const double r2 = 1.0 / std::sqrt(2.0);
  int rc;
  switch (op.gate) {
    case SMOQ_G_U1:
      if ((rc = need(1))) return rc;
      if (op.m[2] == 0.0 && op.m[3] == 0.0 && op.m[4] == 0.0 && op.m[5] == 0.0) {
        // a diagonal matrix is a phase gate: it needs no register slot
        const double ph[4] = {op.m[0], op.m[1], op.m[6], op.m[7]};
        set_diag(g, 1, op.q, ph);
        break;
      }
      g.kind = IK_U1;
      g.t0 = op.q[0];
      std::memcpy(g.m, op.m, sizeof(double) * 8);
      break;
    case SMOQ_G_U2: {
      if ((rc = need(2))) return rc;
      bool diag = true;
      for (int r = 0; r < 4; ++r)
        for (int c2 = 0; c2 < 4; ++c2)
          if (r != c2 && (op.m[2 * (4 * r + c2)] != 0.0 || op.m[2 * (4 * r + c2) + 1] != 0.0))
            diag = false;
      if (diag) {
        double ph[8];
        for (int e = 0; e < 4; ++e) {
          ph[2 * e] = op.m[2 * (5 * e)];
          ph[2 * e + 1] = op.m[2 * (5 * e) + 1];
        }
        set_diag(g, 2, op.q, ph);
        break;
      }
      g.kind = IK_U2;
      g.t0 = op.q[0];
      g.t1 = op.q[1];
      std::memcpy(g.m, op.m, sizeof(double) * 32);
      break;
    }
    case SMOQ_G_H:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], r2, 0, r2, 0, r2, 0, -r2, 0);
      break;
    case SMOQ_G_X:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], 0, 0, 1, 0, 1, 0, 0, 0);
      break;
    case SMOQ_G_Y:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], 0, 0, 0, -1, 0, 1, 0, 0);
      break;
    case SMOQ_G_Z:
    case SMOQ_G_S:
    case SMOQ_G_SDG:
    case SMOQ_G_T:
    case SMOQ_G_TDG:
    case SMOQ_G_RZ:
    case SMOQ_G_DIAG1: {
      if ((rc = need(1))) return rc;
      double ph[4] = {1, 0, 1, 0};
      const double t4 = std::sqrt(0.5);  // cos(pi/4) = sin(pi/4)
      switch (op.gate) {
        case SMOQ_G_Z: ph[2] = -1; break;
        case SMOQ_G_S: ph[2] = 0; ph[3] = 1; break;
        case SMOQ_G_SDG: ph[2] = 0; ph[3] = -1; break;
        case SMOQ_G_T: ph[2] = std::cos(M_PI / 4); ph[3] = std::sin(M_PI / 4); break;
        case SMOQ_G_TDG: ph[2] = std::cos(M_PI / 4); ph[3] = -std::sin(M_PI / 4); break;
        case SMOQ_G_RZ: ph[0] = c; ph[1] = -sn; ph[2] = c; ph[3] = sn; break;
        case SMOQ_G_DIAG1: std::memcpy(ph, op.m, sizeof(ph)); break;
      }
      (void)t4;
      set_diag(g, 1, op.q, ph);
      break;
    }
    case SMOQ_G_RX:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], c, 0, 0, -sn, 0, -sn, c, 0);
      break;
    case SMOQ_G_RY:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], c, 0, -sn, 0, sn, 0, c, 0);
      break;
    case SMOQ_G_CNOT:
      if ((rc = need(2))) return rc;
      set_u1(g, op.q[1], 0, 0, 1, 0, 1, 0, 0, 0);
      g.ctrl = 1ull << op.q[0];
      break;
    case SMOQ_G_CRY:
      if ((rc = need(2))) return rc;
      set_u1(g, op.q[1], c, 0, -sn, 0, sn, 0, c, 0);
      g.ctrl = 1ull << op.q[0];
      break;
    case SMOQ_G_CZ: {
      if ((rc = need(2))) return rc;
      double ph[8] = {1, 0, 1, 0, 1, 0, -1, 0};
      set_diag(g, 2, op.q, ph);
      break;
    }
    case SMOQ_G_RZZ: {
      if ((rc = need(2))) return rc;
      double ph[8] = {c, -sn, c, sn, c, sn, c, -sn};
      set_diag(g, 2, op.q, ph);
      break;
    }
    case SMOQ_G_DIAG2: {
      if ((rc = need(2))) return rc;
      set_diag(g, 2, op.q, op.m);
      break;
    }
    case SMOQ_G_RXX:
    case SMOQ_G_RYY: {
      if ((rc = need(2))) return rc;
      // basis e = b(q0) + 2 b(q1); XX and YY connect 0<->3 and 1<->2.
      // RXX: off-diagonal -i s everywhere (…Gates.jl:992-1026).
      // RYY: 0<->3 carries +i s, 1<->2 carries -i s (…Gates.jl:1039-1071).
      g.kind = IK_U2;
      g.t0 = op.q[0];
      g.t1 = op.q[1];
      double far = (op.gate == SMOQ_G_RXX) ? -sn : sn;
      auto setm = [&](int r, int col, double re, double im) {
        g.m[2 * (4 * r + col)] = re;
        g.m[2 * (4 * r + col) + 1] = im;
      };
      for (int i = 0; i < 4; ++i) setm(i, i, c, 0);
      setm(0, 3, 0, far);
      setm(3, 0, 0, far);
      setm(1, 2, 0, -sn);
      setm(2, 1, 0, -sn);
      break;
    }
    case SMOQ_G_CCZ: {
      if ((rc = need(3))) return rc;
      double ph[16];
      for (int i = 0; i < 8; ++i) { ph[2 * i] = 1; ph[2 * i + 1] = 0; }
      ph[14] = -1;
      set_diag(g, 3, op.q, ph);
      break;
    }
    case SMOQ_G_TOFFOLI:
      if ((rc = need(3))) return rc;
      set_u1(g, op.q[2], 0, 0, 1, 0, 1, 0, 0, 0);
      g.ctrl = (1ull << op.q[0]) | (1ull << op.q[1]);
      break;
    case SMOQ_G_SIGMA_MINUS:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], 0, 0, 1, 0, 0, 0, 0, 0);
      break;
    case SMOQ_G_SIGMA_PLUS:
      if ((rc = need(1))) return rc;
      set_u1(g, op.q[0], 0, 0, 0, 0, 1, 0, 0, 0);
      break;
    default:
      return set_error(SMOQ_EINVAL, "unknown gate id %d", op.gate);
  }
  out.push_back(g);
  return SMOQ_OK;
}

static bool is_rotation(int gate) {
  return gate == SMOQ_G_RX || gate == SMOQ_G_RY || gate == SMOQ_G_RZ || gate == SMOQ_G_CRY ||
         gate == SMOQ_G_RXX || gate == SMOQ_G_RYY || gate == SMOQ_G_RZZ;
}

int lower_tape(const smoq_op* ops, int n_ops, const double* theta, int n_theta, int n,
               std::vector<IGate>& gs) {
  if (n_ops < 0 || (n_ops > 0 && !ops)) return set_error(SMOQ_EINVAL, "bad tape");
  gs.reserve(gs.size() + n_ops);
  for (int i = 0; i < n_ops; ++i) {
    const smoq_op& op = ops[i];
    double th = op.theta;
    if (is_rotation(op.gate) && op.param >= 0) {
      if (!theta || op.param >= n_theta)
        return set_error(SMOQ_EINVAL, "tape op %d uses theta[%d], %d given", i, op.param,
                         n_theta);
      th = theta[op.param];
    }
    int rc = lower_op(op, th, n, gs);
    if (rc) return rc;
  }
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" {

const char* smoq_last_error(void) { return g_last_error.c_str(); }

int smoq_version(void) { return SMOQ_ABI_VERSION; }

int smoq_init(smoq_ctx** out, int device) {
  if (!out) return set_error(SMOQ_EINVAL, "smoq_init: null output pointer");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return set_error(SMOQ_EDEVICE, "smoq_init: no HIP device available (%s)",
                     e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (device < 0 || device >= count)
    return set_error(SMOQ_EINVAL, "smoq_init: device %d out of range [0, %d)", device, count);
  SMOQ_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  SMOQ_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(SMOQ_ENOTSUP, "smoq_init: device %d is %s; this library is built for gfx950",
                     device, prop.gcnArchName);
  if (int rc = noalloc_scratch_init(device)) return rc;
  smoq_ctx* ctx = new smoq_ctx();
  ctx->device = device;
  if (const char* j = std::getenv("SMOQ_JIT")) {
    const int m = std::atoi(j);
    if (m >= 0 && m <= 2) ctx->jit_mode = m;
  }
  if (const char* o = std::getenv("SMOQ_OBS_TILES")) ctx->obs_tiles = std::atoi(o) != 0;
  if (const char* o = std::getenv("SMOQ_OBS_JIT")) {
    const int m = std::atoi(o);
    if (m >= 0 && m <= 2) ctx->obs_jit = m;
  }
  if (const char* g = std::getenv("SMOQ_PASS_MAX_GATES")) ctx->pass_max_gates = std::max(0, std::atoi(g));
  if (const char* w = std::getenv("SMOQ_LANE_SWAPS")) ctx->lane_swaps = std::atoi(w) != 0;
  hipError_t se = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (se != hipSuccess) {
    delete ctx;
    return hip_error(se, "hipStreamCreateWithFlags");
  }
  *out = ctx;
  return SMOQ_OK;
}

int smoq_fini(smoq_ctx* ctx) {
  if (!ctx) return SMOQ_OK;
  if (ctx->live_objects > 0)
    return set_error(SMOQ_EINVAL,
                     "smoq_fini: %lld states or plans of this context are still allocated; "
                     "free them first",
                     (long long)ctx->live_objects);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->prog) (void)hipFree(ctx->prog);
  for (auto& b : ctx->state_pool) (void)hipFree(b.second);
  if (ctx->tree_dev) (void)hipFree(ctx->tree_dev);
  for (int i = 0; i < smoq_ctx::kWorkSlots; ++i)
    if (ctx->work[i]) (void)hipFree(ctx->work[i]);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->aux_stream) {
    (void)hipStreamSynchronize(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->aux_stream);
  }
  if (ctx->res_pinned) (void)hipHostFree(ctx->res_pinned);
  blas_release(ctx);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SMOQ_OK;
}

int smoq_sync(smoq_ctx* ctx) {
  if (!ctx) return set_error(SMOQ_EINVAL, "smoq_sync: null context");
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

int smoq_set_fusion(smoq_ctx* ctx, int tile_qubits, int low_qubits) {
  if (!ctx) return set_error(SMOQ_EINVAL, "smoq_set_fusion: null context");
  // a tile of 2^k amplitudes runs 2^(k - kRegQubits) threads
  const int kmax = kMaxTileQubits;
  if (tile_qubits > 0 && (tile_qubits < 6 || tile_qubits > kmax))
    return set_error(SMOQ_EINVAL, "tile_qubits %d outside [6, %d]", tile_qubits, kmax);
  if (low_qubits != 0 && (low_qubits < 3 || low_qubits > 6))
    return set_error(SMOQ_EINVAL, "low_qubits %d outside [3, 6]", low_qubits);
  if (tile_qubits > 0) ctx->tile_qubits = tile_qubits;
  if (tile_qubits < 0) ctx->tile_qubits = 0;  // automatic again
  if (low_qubits) ctx->low_qubits = low_qubits;
  return SMOQ_OK;
}

int smoq_pass_timing(smoq_ctx* ctx, int enable, double* mean_ms, int64_t* launches) {
  if (!ctx) return set_error(SMOQ_EINVAL, "smoq_pass_timing: null context");
  if (ctx->ev_used) {
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i + 1 < ctx->ev_used; i += 2) {
      float ms = 0.f;
      SMOQ_HIP(hipEventElapsedTime(&ms, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
      ctx->timed_ms += ms;
      ctx->timed_launches += 1;
    }
    ctx->ev_used = 0;
  }
  if (mean_ms) *mean_ms = ctx->timed_launches ? ctx->timed_ms / ctx->timed_launches : 0.0;
  if (launches) *launches = ctx->timed_launches;
  ctx->timed_ms = 0.0;
  ctx->timed_launches = 0;
  ctx->timing = enable != 0;
  return SMOQ_OK;
}

int smoq_state_alloc(smoq_ctx* ctx, int n_qubits, int dtype, int batch, smoq_state** out) {
  if (!ctx || !out) return set_error(SMOQ_EINVAL, "smoq_state_alloc: null argument");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 36)
    return set_error(SMOQ_EINVAL, "smoq_state_alloc: n_qubits %d outside [1, 36]", n_qubits);
  if (dtype != SMOQ_C128 && dtype != SMOQ_C64)
    return set_error(SMOQ_EINVAL, "smoq_state_alloc: unknown dtype %d", dtype);
  if (batch < 1) return set_error(SMOQ_EINVAL, "smoq_state_alloc: batch %d < 1", batch);
  smoq_state* s = new smoq_state();
  s->ctx = ctx;
  s->n = n_qubits;
  s->dtype = dtype;
  s->batch = batch;
  s->dim = 1ull << n_qubits;
  s->bytes = s->dim * s->amp_bytes() * (size_t)batch;
  // a small buffer freed on this context is reused: its earlier work is on
  // the same stream, so the zero fill below runs after it
  for (size_t i = 0; i < ctx->state_pool.size(); ++i) {
    if (ctx->state_pool[i].first != s->bytes) continue;
    s->d = ctx->state_pool[i].second;
    ctx->state_pool.erase(ctx->state_pool.begin() + (ptrdiff_t)i);
    const hipError_t me = hipMemsetAsync(s->d, 0, s->bytes, ctx->stream);
    if (me != hipSuccess) {
      ctx->state_pool.emplace_back(s->bytes, s->d);
      delete s;
      return hip_error(me, "hipMemsetAsync(state)");
    }
    ctx->live_objects++;
    *out = s;
    return SMOQ_OK;
  }
  size_t free_b = 0, total_b = 0;
  SMOQ_HIP(hipSetDevice(ctx->device));
  SMOQ_HIP(hipMemGetInfo(&free_b, &total_b));
  // pooled buffers are allocated memory that hipMemGetInfo does not count
  // as free: give them back before refusing a request
  if (s->bytes > free_b && !ctx->state_pool.empty()) {
    release_state_pool(ctx);
    SMOQ_HIP(hipMemGetInfo(&free_b, &total_b));
  }
  if (s->bytes > free_b) {
    size_t want = s->bytes;
    delete s;
    return set_error(SMOQ_ENOMEM, "smoq_state_alloc: %zu bytes requested, %zu free on device",
                     want, free_b);
  }
  hipError_t e = hipMalloc(&s->d, s->bytes);
  if (e != hipSuccess && !ctx->state_pool.empty()) {
    (void)hipGetLastError();
    release_state_pool(ctx);
    e = hipMalloc(&s->d, s->bytes);
  }
  if (e != hipSuccess) {
    delete s;
    return set_error(SMOQ_ENOMEM, "smoq_state_alloc: hipMalloc failed: %s", hipGetErrorString(e));
  }
  const hipError_t me = hipMemsetAsync(s->d, 0, s->bytes, ctx->stream);
  if (me != hipSuccess) {
    (void)hipFree(s->d);
    delete s;
    return hip_error(me, "hipMemsetAsync(state)");
  }
  ctx->live_objects++;
  *out = s;
  return SMOQ_OK;
}

int smoq_state_free(smoq_state* s) {
  if (!s) return SMOQ_OK;
  smoq_ctx* ctx = s->ctx;
  ctx->live_objects--;
  if (s->d && s->bytes <= kStatePoolMaxBytes && ctx->state_pool.size() < kStatePoolMax) {
    ctx->state_pool.emplace_back(s->bytes, s->d);  // no sync, no hipFree
  } else if (s->d) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(s->d);
  }
  s->d = nullptr;
  delete s;
  return SMOQ_OK;
}

int smoq_state_info(const smoq_state* s, int* n_qubits, int* dtype, int* batch, void** ptr) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n_qubits) *n_qubits = s->n;
  if (dtype) *dtype = s->dtype;
  if (batch) *batch = s->batch;
  if (ptr) *ptr = s->d;
  return SMOQ_OK;
}

int smoq_state_upload(smoq_state* s, const void* host, size_t bytes) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!host || bytes != s->bytes)
    return set_error(SMOQ_EINVAL, "smoq_state_upload: %zu bytes given, state holds %zu", bytes,
                     s->bytes);
  SMOQ_HIP(hipMemcpyAsync(s->d, host, bytes, hipMemcpyHostToDevice, s->ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(s->ctx->stream));
  return SMOQ_OK;
}

int smoq_state_download(const smoq_state* s, void* host, size_t bytes) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!host || bytes != s->bytes)
    return set_error(SMOQ_EINVAL, "smoq_state_download: %zu bytes given, state holds %zu", bytes,
                     s->bytes);
  SMOQ_HIP(hipMemcpyAsync(host, s->d, bytes, hipMemcpyDeviceToHost, s->ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(s->ctx->stream));
  return SMOQ_OK;
}

int smoq_state_copy(smoq_state* dst, const smoq_state* src) {
  int rc = check_state(dst);
  if (rc) return rc;
  if ((rc = check_state(src))) return rc;
  if (dst->bytes != src->bytes || dst->dtype != src->dtype)
    return set_error(SMOQ_EINVAL, "smoq_state_copy: shape/dtype mismatch");
  SMOQ_HIP(hipMemcpyAsync(dst->d, src->d, src->bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
  return SMOQ_OK;
}

int smoq_apply_1q(smoq_state* s, int q, const double u[8]) {
  int rc = check_state(s);
  if (rc) return rc;
  if ((rc = check_qubit(s, q, "smoq_apply_1q"))) return rc;
  if (!u) return set_error(SMOQ_EINVAL, "smoq_apply_1q: null matrix");
  IGate g;
  g.kind = IK_U1;
  g.t0 = q;
  std::memcpy(g.m, u, sizeof(double) * 8);
  return launch_gate(s, g);
}

int smoq_apply_2q(smoq_state* s, int qa, int qb, const double u[32]) {
  int rc = check_state(s);
  if (rc) return rc;
  if ((rc = check_qubit(s, qa, "smoq_apply_2q"))) return rc;
  if ((rc = check_qubit(s, qb, "smoq_apply_2q"))) return rc;
  if (qa == qb) return set_error(SMOQ_EINVAL, "smoq_apply_2q: qa == qb == %d", qa);
  if (!u) return set_error(SMOQ_EINVAL, "smoq_apply_2q: null matrix");
  IGate g;
  g.kind = IK_U2;
  g.t0 = qa;
  g.t1 = qb;
  std::memcpy(g.m, u, sizeof(double) * 32);
  return launch_gate(s, g);
}

int smoq_apply_named(smoq_state* s, int gate, const int* qubits, double theta) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!qubits) return set_error(SMOQ_EINVAL, "smoq_apply_named: null qubit array");
  if (gate == SMOQ_G_U1 || gate == SMOQ_G_U2 || gate == SMOQ_G_DIAG1 || gate == SMOQ_G_DIAG2)
    return set_error(SMOQ_EINVAL, "smoq_apply_named: gate %d needs a matrix; use smoq_apply_tape",
                     gate);
  smoq_op op;
  std::memset(&op, 0, sizeof(op));
  op.gate = gate;
  int nq = (gate == SMOQ_G_CCZ || gate == SMOQ_G_TOFFOLI) ? 3
           : (gate >= SMOQ_G_CNOT && gate <= SMOQ_G_RZZ) ? 2 : 1;
  for (int i = 0; i < 3; ++i) op.q[i] = i < nq ? qubits[i] : -1;
  op.param = -1;
  op.theta = theta;
  std::vector<IGate> gs;
  if ((rc = lower_op(op, theta, s->n, gs))) return rc;
  for (auto& g : gs)
    if ((rc = launch_gate(s, g))) return rc;
  return SMOQ_OK;
}

int smoq_apply_tape(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                    int n_theta, int* n_passes) {
  int rc = check_state(s);
  if (rc) return rc;
  std::vector<IGate>& gs = s->ctx->tape_gates;
  gs.clear();
  if ((rc = lower_tape(ops, n_ops, theta, n_theta, s->n, gs))) return rc;
  Program& prog = s->ctx->tape_prog;
  const Fusion fu = fusion_for(s->ctx, s->n, s->dtype, s->batch, gs);
  plan_program(gs, s->n, fu.tile, fu.low,
               s->dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits, prog, fu.cap,
               fu.swaps != 0,
               fu.locals.empty() ? nullptr : &fu.locals);
  apply_pass_io(prog, fu.io, fu.low);
  if (n_passes) *n_passes = (int)prog.passes.size();
  if (gs.empty()) return SMOQ_OK;
  if (s->n < kRegQubits) {
    // too small for a register stage: one streaming kernel per gate
    if (n_passes) *n_passes = (int)gs.size();
    for (auto& g : gs)
      if ((rc = launch_gate(s, g))) return rc;
    return SMOQ_OK;
  }
  bool fixed = true;
  for (int i = 0; i < n_ops; ++i) fixed &= ops[i].param < 0;
  return run_program(s, prog, fixed);
}

}  // extern "C"
