// Gate kernels for gfx950: one-gate streaming kernels and the fused
// LDS-tile pass kernel that applies a whole stage program per HBM pass.
#include "smoq_internal.h"
#include "device_common.h"

namespace smoq {

// ----------------------------------------------------- one-gate kernels --
// Each thread owns one amplitude pair (quad for 2-qubit gates).  Pair index
// t enumerates the indices with the target bit cleared; inserting a zero at
// the target keeps lanes on consecutive addresses for both halves of the
// pair, so every load and store is a coalesced 16 B (c128) / 8 B (c64)
// access.  Batch items are contiguous, so the bits above n carry the item.

template <typename T>
__global__ void __launch_bounds__(256)
    k_u1(typename CT<T>::V* __restrict__ psi, uint64_t npairs, int q, uint64_t ctrl, Mat2<T> u) {
  using V = typename CT<T>::V;
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  uint64_t lo = insert_zero(t, q);
  if ((lo & ctrl) != ctrl) return;
  uint64_t hi = lo | (1ull << q);
  V a = psi[lo], b = psi[hi];
  psi[lo] = cfma2(u.m[0], a, u.m[1], b);
  psi[hi] = cfma2(u.m[2], a, u.m[3], b);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_u2(typename CT<T>::V* __restrict__ psi, uint64_t nquads, int qa, int qb, uint64_t ctrl,
         Mat4<T> u) {
  using V = typename CT<T>::V;
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nquads) return;
  int lo_b = qa < qb ? qa : qb, hi_b = qa < qb ? qb : qa;
  uint64_t i0 = insert_zero(insert_zero(t, lo_b), hi_b);
  if ((i0 & ctrl) != ctrl) return;
  uint64_t ma = 1ull << qa, mb = 1ull << qb;
  uint64_t idx[4] = {i0, i0 | ma, i0 | mb, i0 | ma | mb};
  V x[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) x[e] = psi[idx[e]];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    V acc = cmul(u.m[4 * r], x[0]);
#pragma unroll
    for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(u.m[4 * r + c], x[c]));
    psi[idx[r]] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_diag(typename CT<T>::V* __restrict__ psi, uint64_t total, uint64_t dim_mask, int nd, int q0,
           int q1, int q2, Diag8<T> ph) {
  using V = typename CT<T>::V;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  uint64_t g = i & dim_mask;
  int e = (int)((g >> q0) & 1);
  if (nd > 1) e |= (int)((g >> q1) & 1) << 1;
  if (nd > 2) e |= (int)((g >> q2) & 1) << 2;
  psi[i] = cmul(ph.p[e], psi[i]);
}

template <typename T>
static int launch_gate_t(smoq_state* s, const IGate& g) {
  using V = typename CT<T>::V;
  V* psi = (V*)s->d;
  hipStream_t st = s->ctx->stream;
  const uint64_t total = s->dim * (uint64_t)s->batch;
  const int B = 256;
  if (g.kind == IK_U1) {
    Mat2<T> u;
    for (int i = 0; i < 4; ++i) u.m[i] = make_c<T>(g.m[2 * i], g.m[2 * i + 1]);
    uint64_t np = total >> 1;
    hipLaunchKernelGGL(k_u1<T>, dim3((unsigned)((np + B - 1) / B)), dim3(B), 0, st, psi, np, g.t0,
                       g.ctrl, u);
  } else if (g.kind == IK_U2) {
    Mat4<T> u;
    for (int i = 0; i < 16; ++i) u.m[i] = make_c<T>(g.m[2 * i], g.m[2 * i + 1]);
    uint64_t nq = total >> 2;
    hipLaunchKernelGGL(k_u2<T>, dim3((unsigned)((nq + B - 1) / B)), dim3(B), 0, st, psi, nq, g.t0,
                       g.t1, g.ctrl, u);
  } else {
    Diag8<T> ph;
    for (int i = 0; i < (1 << g.nd); ++i) ph.p[i] = make_c<T>(g.m[2 * i], g.m[2 * i + 1]);
    hipLaunchKernelGGL(k_diag<T>, dim3((unsigned)((total + B - 1) / B)), dim3(B), 0, st, psi, total,
                       s->dim - 1, g.nd, g.dq[0], g.dq[1] < 0 ? 0 : g.dq[1],
                       g.dq[2] < 0 ? 0 : g.dq[2], ph);
  }
  SMOQ_CHECK_LAUNCH("one-gate kernel");
  return SMOQ_OK;
}

int launch_gate(smoq_state* s, const IGate& g) {
  return s->dtype == SMOQ_C128 ? launch_gate_t<double>(s, g) : launch_gate_t<float>(s, g);
}

// |index> in every batch item; a fill kernel (not a memset) so that the call
// can sit inside a captured graph.
template <typename T>
__global__ void __launch_bounds__(256)
    k_set_basis(typename CT<T>::V* __restrict__ psi, uint64_t total, uint64_t dim_mask,
                uint64_t index) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  psi[i] = make_c<T>((i & dim_mask) == index ? 1.0 : 0.0, 0.0);
}

}  // namespace smoq

extern "C" int smoq_state_set_basis(smoq_state* s, uint64_t index) {
  using namespace smoq;
  int rc = check_state(s);
  if (rc) return rc;
  if (index >= s->dim)
    return set_error(SMOQ_EINVAL, "smoq_state_set_basis: index %llu >= 2^%d",
                     (unsigned long long)index, s->n);
  const uint64_t total = s->dim * (uint64_t)s->batch;
  const unsigned nb = (unsigned)((total + 255) / 256);
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_set_basis<double>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (double2*)s->d, total, s->dim - 1, index);
  else
    hipLaunchKernelGGL(k_set_basis<float>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (float2*)s->d, total, s->dim - 1, index);
  SMOQ_CHECK_LAUNCH("k_set_basis");
  return SMOQ_OK;
}

namespace smoq {

// ------------------------------------------------------ fused pass kernel --
// One workgroup owns one tile: the 2^k amplitudes whose non-local bits equal
// the tile id.  The tile is read from HBM into LDS once (coalesced runs of
// 2^m amplitudes), every stage then moves 2^R amplitudes per thread from LDS
// into registers, applies its gates there and writes them back, and the
// tile is finally written to HBM once.  LDS rows are XOR-swizzled so that
// the strided per-stage gathers stay free of bank conflicts.

template <typename T, int A>
__device__ inline void reg_u1(typename CT<T>::V (&x)[1 << kRegQubits], const Mat2<T>& u,
                              const uint64_t (&gi)[1 << kRegQubits], uint64_t ctrl) {
  using V = typename CT<T>::V;
#pragma unroll
  for (int j = 0; j < (1 << kRegQubits); ++j) {
    if (j & (1 << A)) continue;
    const int h = j | (1 << A);
    V a = x[j], b = x[h];
    V na = cfma2(u.m[0], a, u.m[1], b);
    V nb = cfma2(u.m[2], a, u.m[3], b);
    bool on = (gi[j] & ctrl) == ctrl;
    x[j] = on ? na : a;
    x[h] = on ? nb : b;
  }
}

template <typename T, int A, int B>
__device__ inline void reg_u2(typename CT<T>::V (&x)[1 << kRegQubits], const Mat4<T>& u,
                              const uint64_t (&gi)[1 << kRegQubits], uint64_t ctrl) {
  using V = typename CT<T>::V;
#pragma unroll
  for (int j = 0; j < (1 << kRegQubits); ++j) {
    if (j & ((1 << A) | (1 << B))) continue;
    const int id[4] = {j, j | (1 << A), j | (1 << B), j | (1 << A) | (1 << B)};
    V in[4] = {x[id[0]], x[id[1]], x[id[2]], x[id[3]]};
    bool on = (gi[j] & ctrl) == ctrl;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      V acc = cmul(u.m[4 * r], in[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(u.m[4 * r + c], in[c]));
      x[id[r]] = on ? acc : in[r];
    }
  }
}

template <typename T>
__device__ inline void apply_gate_regs(typename CT<T>::V (&x)[1 << kRegQubits],
                                       const GateDev& g, const uint64_t (&gi)[1 << kRegQubits]) {
  using V = typename CT<T>::V;
  if (g.kind == IK_U1) {
    Mat2<T> u;
#pragma unroll
    for (int i = 0; i < 4; ++i) u.m[i] = make_c<T>(g.m[2 * i], g.m[2 * i + 1]);
    switch (g.a) {
      case 0: reg_u1<T, 0>(x, u, gi, g.ctrl); break;
      case 1: reg_u1<T, 1>(x, u, gi, g.ctrl); break;
      case 2: reg_u1<T, 2>(x, u, gi, g.ctrl); break;
      default: reg_u1<T, 3>(x, u, gi, g.ctrl); break;
    }
  } else if (g.kind == IK_U2) {
    Mat4<T> u;
#pragma unroll
    for (int i = 0; i < 16; ++i) u.m[i] = make_c<T>(g.m[2 * i], g.m[2 * i + 1]);
    switch (g.a * 4 + g.b) {
      case 1: reg_u2<T, 0, 1>(x, u, gi, g.ctrl); break;
      case 2: reg_u2<T, 0, 2>(x, u, gi, g.ctrl); break;
      case 3: reg_u2<T, 0, 3>(x, u, gi, g.ctrl); break;
      case 4: reg_u2<T, 1, 0>(x, u, gi, g.ctrl); break;
      case 6: reg_u2<T, 1, 2>(x, u, gi, g.ctrl); break;
      case 7: reg_u2<T, 1, 3>(x, u, gi, g.ctrl); break;
      case 8: reg_u2<T, 2, 0>(x, u, gi, g.ctrl); break;
      case 9: reg_u2<T, 2, 1>(x, u, gi, g.ctrl); break;
      case 11: reg_u2<T, 2, 3>(x, u, gi, g.ctrl); break;
      case 12: reg_u2<T, 3, 0>(x, u, gi, g.ctrl); break;
      case 13: reg_u2<T, 3, 1>(x, u, gi, g.ctrl); break;
      default: reg_u2<T, 3, 2>(x, u, gi, g.ctrl); break;
    }
  } else {
    V p[8];
    const int ne = 1 << g.nd;
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i] = make_c<T>(i < ne ? g.m[2 * i] : 1.0, i < ne ? g.m[2 * i + 1] : 0.0);
    const int d0 = g.dq[0], d1 = g.nd > 1 ? g.dq[1] : 63, d2 = g.nd > 2 ? g.dq[2] : 63;
#pragma unroll
    for (int j = 0; j < (1 << kRegQubits); ++j) {
      const uint64_t G = gi[j];
      const int e = (int)((G >> d0) & 1) | ((int)((G >> d1) & 1) << 1) | ((int)((G >> d2) & 1) << 2);
      V ph = p[0];
#pragma unroll
      for (int c = 1; c < 8; ++c) ph = (e == c) ? p[c] : ph;
      x[j] = cmul(ph, x[j]);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_fused_pass(typename CT<T>::V* __restrict__ psi, int n, const PassDev* __restrict__ pass,
                 const StageDev* __restrict__ stages, const GateDev* __restrict__ gates) {
  using V = typename CT<T>::V;
  constexpr int R = kRegQubits;
  constexpr int NR = 1 << R;
  constexpr int WB = CT<T>::kRowBits;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);

  const int k = pass->k;
  const int NT = 1 << (k - R);
  const int tid = threadIdx.x;
  const int nn = pass->n_nonlocal;
  const uint64_t tile = blockIdx.x;
  const uint64_t item = tile >> nn;
  const uint64_t tt = tile & ((1ull << nn) - 1);
  uint64_t gbase = 0;
  for (int i = 0; i < nn; ++i) gbase |= ((tt >> i) & 1ull) << pass->nonlocal[i];
  V* base = psi + (item << n);

  // thread bits cover local positions 0..k-R-1, the loop index the top R
  uint64_t off_tid = 0;
  for (int j = 0; j < k - R; ++j) off_tid |= (uint64_t)((tid >> j) & 1) << pass->map[j];
  uint64_t off_hi[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) o |= (uint64_t)((i >> j) & 1) << pass->map[k - R + j];
    off_hi[i] = o;
  }
  V x[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) x[i] = base[gbase | off_tid | off_hi[i]];
#pragma unroll
  for (int i = 0; i < NR; ++i) lds[swz<WB>((uint32_t)(tid + i * NT))] = x[i];
  __syncthreads();

  const uint32_t kmask = (1u << k) - 1u;
  for (int s = pass->s0; s < pass->s1; ++s) {
    const StageDev& st = stages[s];
    uint32_t qm = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) qm |= 1u << st.q[j];
    const uint32_t lb = pdep32((uint32_t)tid, kmask & ~qm);
    uint32_t lo[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      uint32_t o = lb;
#pragma unroll
      for (int j = 0; j < R; ++j) o |= (uint32_t)((i >> j) & 1) << st.q[j];
      lo[i] = o;
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) x[i] = lds[swz<WB>(lo[i])];
    // global indices (diagonal phases and controls read them)
    uint64_t gl = gbase;
    for (int b = 0; b < k; ++b) gl |= (uint64_t)((lb >> b) & 1u) << pass->map[b];
    uint64_t gq[R];
#pragma unroll
    for (int j = 0; j < R; ++j) gq[j] = 1ull << pass->map[st.q[j]];
    uint64_t gi[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      uint64_t G = gl;
#pragma unroll
      for (int j = 0; j < R; ++j) G |= ((i >> j) & 1) ? gq[j] : 0ull;
      gi[i] = G;
    }
    for (int g = st.g0; g < st.g1; ++g) apply_gate_regs<T>(x, gates[g], gi);
#pragma unroll
    for (int i = 0; i < NR; ++i) lds[swz<WB>(lo[i])] = x[i];
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < NR; ++i) x[i] = lds[swz<WB>((uint32_t)(tid + i * NT))];
#pragma unroll
  for (int i = 0; i < NR; ++i) base[gbase | off_tid | off_hi[i]] = x[i];
}

template <typename T>
static int run_program_t(smoq_state* s, const Program& prog) {
  using V = typename CT<T>::V;
  smoq_ctx* ctx = s->ctx;
  const size_t pb = prog.passes.size() * sizeof(PassDev);
  const size_t sb = prog.stages.size() * sizeof(StageDev);
  const size_t gb = prog.gates.size() * sizeof(GateDev);
  const size_t total = pb + sb + gb;
  int rc = ensure_prog(ctx, total);
  if (rc) return rc;
  if ((rc = ensure_pinned(ctx, total))) return rc;
  // the pinned staging buffer may still feed an earlier copy
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  std::memcpy(h, prog.passes.data(), pb);
  std::memcpy(h + pb, prog.stages.data(), sb);
  std::memcpy(h + pb + sb, prog.gates.data(), gb);
  SMOQ_HIP(hipMemcpyAsync(ctx->prog, h, total, hipMemcpyHostToDevice, ctx->stream));
  const PassDev* dp = (const PassDev*)ctx->prog;
  const StageDev* ds = (const StageDev*)((unsigned char*)ctx->prog + pb);
  const GateDev* dg = (const GateDev*)((unsigned char*)ctx->prog + pb + sb);
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    const PassDev& pd = prog.passes[p];
    const uint64_t tiles = (uint64_t)s->batch << pd.n_nonlocal;
    const int nt = 1 << (pd.k - kRegQubits);
    const size_t shm = sizeof(V) << pd.k;
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
    hipLaunchKernelGGL(k_fused_pass<T>, dim3((unsigned)tiles), dim3(nt), shm, ctx->stream,
                       (V*)s->d, s->n, dp + p, ds, dg);
    SMOQ_CHECK_LAUNCH("k_fused_pass");
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
  }
  return SMOQ_OK;
}

int run_program(smoq_state* s, const Program& prog) {
  return s->dtype == SMOQ_C128 ? run_program_t<double>(s, prog) : run_program_t<float>(s, prog);
}

}  // namespace smoq
