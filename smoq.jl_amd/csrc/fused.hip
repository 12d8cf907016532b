// Fused gate passes: one HBM read and one HBM write of the state per pass,
// with every gate of the pass applied to an LDS-resident tile.
//
// A workgroup owns one tile: the 2^k amplitudes whose non-local bits equal
// the tile id.  The tile is read from HBM into LDS once (coalesced runs of
// 2^m consecutive amplitudes), every register stage then moves 2^R
// amplitudes per thread from LDS into registers, applies its gates there and
// writes them back, and the tile is finally written to HBM once.  LDS rows
// are XOR-swizzled so that the strided per-stage gathers stay free of bank
// conflicts.  Control and diagonal qubits never need to be local: their bits
// come from the thread's base index (global part) or the register slot.
#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

template <typename T>
using VT = typename CT<T>::V;
constexpr int kNR = 1 << kRegQubits;

template <typename T>
__device__ __forceinline__ VT<T> ld_c(const double* m, int e) {
  return make_c<T>(m[2 * e], m[2 * e + 1]);
}

// u*a for the entry kinds of the specialised 2x2 forms
template <typename T, int PT>
__device__ __forceinline__ VT<T> u1_row(const double* m, int r, VT<T> a, VT<T> b) {
  VT<T> o;
  if (PT == GK_U1_REAL) {
    const T p = (T)m[4 * r], q = (T)m[4 * r + 2];
    o.x = p * a.x + q * b.x;
    o.y = p * a.y + q * b.y;
  } else if (PT == GK_U1_RXT) {
    if (r == 0) {  // (u00 real) a + (i u01i) b
      const T p = (T)m[0], q = (T)m[3];
      o.x = p * a.x - q * b.y;
      o.y = p * a.y + q * b.x;
    } else {  // (i u10i) a + (u11 real) b
      const T p = (T)m[5], q = (T)m[6];
      o.x = q * b.x - p * a.y;
      o.y = q * b.y + p * a.x;
    }
  } else {
    o = cfma2(ld_c<T>(m, 2 * r), a, ld_c<T>(m, 2 * r + 1), b);
  }
  return o;
}

// 2x2 on register slot A for every pair (j, j | 1<<A)
template <typename T, int PT, int A, bool CTRL>
__device__ __forceinline__ void reg_u1(VT<T> (&x)[kNR], const GateDev& g, bool gon) {
  const int creg = g.creg;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & (1 << A)) continue;
    const int h = j | (1 << A);
    const VT<T> a = x[j], b = x[h];
    const VT<T> na = u1_row<T, PT>(g.m, 0, a, b);
    const VT<T> nb = u1_row<T, PT>(g.m, 1, a, b);
    if (CTRL) {
      const bool on = gon && ((j & creg) == creg);
      x[j] = csel(on, na, a);
      x[h] = csel(on, nb, b);
    } else {
      x[j] = na;
      x[h] = nb;
    }
  }
}

template <typename T, int PT, bool CTRL>
__device__ __forceinline__ void reg_u1_slot(VT<T> (&x)[kNR], const GateDev& g, bool gon) {
  switch (g.a) {
    case 0: reg_u1<T, PT, 0, CTRL>(x, g, gon); break;
    case 1: reg_u1<T, PT, 1, CTRL>(x, g, gon); break;
    case 2: reg_u1<T, PT, 2, CTRL>(x, g, gon); break;
    default: reg_u1<T, PT, 3, CTRL>(x, g, gon); break;
  }
}

template <typename T, int PT>
__device__ __forceinline__ void reg_u1_any(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl) {
  if ((g.creg | g.cglob) == 0) {
    reg_u1_slot<T, PT, false>(x, g, true);
  } else {
    reg_u1_slot<T, PT, true>(x, g, (gl & g.cglob) == g.cglob);
  }
}

template <typename T>
__device__ __forceinline__ VT<T> u2_row(const double* m, int r, VT<T> a0, VT<T> a1, VT<T> a2,
                                        VT<T> a3) {
  VT<T> acc = cmul(ld_c<T>(m, 4 * r), a0);
  acc = cadd(acc, cmul(ld_c<T>(m, 4 * r + 1), a1));
  acc = cadd(acc, cmul(ld_c<T>(m, 4 * r + 2), a2));
  return cadd(acc, cmul(ld_c<T>(m, 4 * r + 3), a3));
}

// 4x4 on slots (A low role, B high role)
template <typename T, int A, int B>
__device__ __forceinline__ void reg_u2(VT<T> (&x)[kNR], const GateDev& g, bool gon) {
  const int creg = g.creg;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & ((1 << A) | (1 << B))) continue;
    const int id0 = j, id1 = j | (1 << A), id2 = j | (1 << B), id3 = j | (1 << A) | (1 << B);
    const VT<T> in0 = x[id0], in1 = x[id1], in2 = x[id2], in3 = x[id3];
    const bool on = gon && ((j & creg) == creg);
    const VT<T> o0 = u2_row<T>(g.m, 0, in0, in1, in2, in3);
    const VT<T> o1 = u2_row<T>(g.m, 1, in0, in1, in2, in3);
    const VT<T> o2 = u2_row<T>(g.m, 2, in0, in1, in2, in3);
    const VT<T> o3 = u2_row<T>(g.m, 3, in0, in1, in2, in3);
    x[id0] = csel(on, o0, in0);
    x[id1] = csel(on, o1, in1);
    x[id2] = csel(on, o2, in2);
    x[id3] = csel(on, o3, in3);
  }
}

// merged diagonal gates: x[j] *= table[row][j], row from non-register bits
template <typename T>
__device__ __forceinline__ void reg_diagtab(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl,
                                            const VT<T>* __restrict__ tabs) {
  int row = 0;
  for (int t = 0; t < g.nu; ++t) row |= (int)((gl >> g.ubits[t]) & 1ull) << t;
  const VT<T>* p = tabs + g.tab + row * kNR;
#pragma unroll
  for (int j = 0; j < kNR; ++j) x[j] = cmul(p[j], x[j]);
}

// merged CZs: negate where sum_i popc(G & (G >> s_i) & M_i) is odd
template <typename T>
__device__ __forceinline__ void reg_czset(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl,
                                          const uint64_t (&gq)[kRegQubits]) {
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    uint64_t G = gl;
#pragma unroll
    for (int s = 0; s < kRegQubits; ++s) G |= ((j >> s) & 1) ? gq[s] : 0ull;
    int par = 0;
    for (int i = 0; i < g.ncz; ++i) par ^= __popcll(G & (G >> g.czs[i]) & g.czm[i]);
    if (par & 1) {
      x[j].x = -x[j].x;
      x[j].y = -x[j].y;
    }
  }
}

template <typename T>
__device__ __forceinline__ void apply_gate(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl,
                                           const uint64_t (&gq)[kRegQubits],
                                           const VT<T>* __restrict__ tabs) {
  switch (g.kind) {
    case GK_U1_REAL: reg_u1_any<T, GK_U1_REAL>(x, g, gl); break;
    case GK_U1_RXT: reg_u1_any<T, GK_U1_RXT>(x, g, gl); break;
    case GK_U1_GEN: reg_u1_any<T, GK_U1_GEN>(x, g, gl); break;
    case GK_U2: {
      const bool gon = (gl & g.cglob) == g.cglob;
      switch (g.a * 4 + g.b) {
        case 1: reg_u2<T, 0, 1>(x, g, gon); break;
        case 2: reg_u2<T, 0, 2>(x, g, gon); break;
        case 3: reg_u2<T, 0, 3>(x, g, gon); break;
        case 4: reg_u2<T, 1, 0>(x, g, gon); break;
        case 6: reg_u2<T, 1, 2>(x, g, gon); break;
        case 7: reg_u2<T, 1, 3>(x, g, gon); break;
        case 8: reg_u2<T, 2, 0>(x, g, gon); break;
        case 9: reg_u2<T, 2, 1>(x, g, gon); break;
        case 11: reg_u2<T, 2, 3>(x, g, gon); break;
        case 12: reg_u2<T, 3, 0>(x, g, gon); break;
        case 13: reg_u2<T, 3, 1>(x, g, gon); break;
        default: reg_u2<T, 3, 2>(x, g, gon); break;
      }
      break;
    }
    case GK_DIAGTAB: reg_diagtab<T>(x, g, gl, tabs); break;
    default: reg_czset<T>(x, g, gl, gq); break;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_fused_pass(VT<T>* __restrict__ psi, int n, const PassDev* __restrict__ pass,
                 const StageDev* __restrict__ stages, const GateDev* __restrict__ gates,
                 const VT<T>* __restrict__ tabs) {
  constexpr int R = kRegQubits;
  constexpr int WB = CT<T>::kRowBits;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  VT<T>* lds = reinterpret_cast<VT<T>*>(smem_raw);

  const int k = pass->k;
  const int kr = k - R;
  const int NT = 1 << kr;
  const int tid = threadIdx.x;
  const int nn = pass->n_nonlocal;
  const uint64_t tile = blockIdx.x;
  const uint64_t item = tile >> nn;
  const uint64_t tt = tile & ((1ull << nn) - 1ull);
  uint64_t gbase = 0;
  for (int i = 0; i < nn; ++i) gbase |= ((tt >> i) & 1ull) << pass->nonlocal[i];
  VT<T>* __restrict__ base = psi + (item << n);

  // HBM -> LDS: the thread index covers local bits 0..k-R-1, the unrolled
  // loop index the top R local bits
  uint64_t gt = gbase;
  for (int j = 0; j < kr; ++j) gt |= (uint64_t)((tid >> j) & 1) << pass->map[j];
  uint64_t hb[R];
#pragma unroll
  for (int j = 0; j < R; ++j) hb[j] = 1ull << pass->map[kr + j];
  {
    VT<T> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint64_t o = gt;
#pragma unroll
      for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? hb[j] : 0ull;
      x[i] = base[o];
    }
#pragma unroll
    for (int i = 0; i < kNR; ++i) lds[swz<WB>((uint32_t)(tid + i * NT))] = x[i];
  }
  __syncthreads();

  const uint32_t kmask = (1u << k) - 1u;
  for (int s = pass->s0; s < pass->s1; ++s) {
    const StageDev& st = stages[s];
    uint32_t b[R];
    uint64_t gq[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      b[j] = 1u << st.q[j];
      gq[j] = 1ull << pass->map[st.q[j]];
    }
    const uint32_t lb = pdep32((uint32_t)tid, kmask & ~(b[0] | b[1] | b[2] | b[3]));
    VT<T> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint32_t l = lb;
#pragma unroll
      for (int j = 0; j < R; ++j) l |= ((i >> j) & 1) ? b[j] : 0u;
      x[i] = lds[swz<WB>(l)];
    }
    uint64_t gl = gbase;
    for (int bb = 0; bb < k; ++bb) gl |= (uint64_t)((lb >> bb) & 1u) << pass->map[bb];
    for (int g = st.g0; g < st.g1; ++g) apply_gate<T>(x, gates[g], gl, gq, tabs);
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint32_t l = lb;
#pragma unroll
      for (int j = 0; j < R; ++j) l |= ((i >> j) & 1) ? b[j] : 0u;
      lds[swz<WB>(l)] = x[i];
    }
    __syncthreads();
  }

  {
    VT<T> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) x[i] = lds[swz<WB>((uint32_t)(tid + i * NT))];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint64_t o = gt;
#pragma unroll
      for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? hb[j] : 0ull;
      base[o] = x[i];
    }
  }
}

template <typename T>
static int run_program_t(smoq_state* s, const Program& prog) {
  smoq_ctx* ctx = s->ctx;
  const size_t pb = prog.passes.size() * sizeof(PassDev);
  const size_t sb = prog.stages.size() * sizeof(StageDev);
  const size_t gb = prog.gates.size() * sizeof(GateDev);
  const size_t ntab = prog.tabs.size() / 2;
  const size_t tb = (ntab + 1) * sizeof(VT<T>);
  const size_t off_t = (pb + sb + gb + 255) & ~(size_t)255;
  const size_t total = off_t + tb;
  int rc = ensure_prog(ctx, total);
  if (rc) return rc;
  if ((rc = ensure_pinned(ctx, total))) return rc;
  // the pinned staging buffer may still feed an earlier copy
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  std::memcpy(h, prog.passes.data(), pb);
  std::memcpy(h + pb, prog.stages.data(), sb);
  std::memcpy(h + pb + sb, prog.gates.data(), gb);
  VT<T>* ht = (VT<T>*)(h + off_t);
  for (size_t i = 0; i < ntab; ++i) ht[i] = make_c<T>(prog.tabs[2 * i], prog.tabs[2 * i + 1]);
  SMOQ_HIP(hipMemcpyAsync(ctx->prog, h, total, hipMemcpyHostToDevice, ctx->stream));
  unsigned char* d = (unsigned char*)ctx->prog;
  const PassDev* dp = (const PassDev*)d;
  const StageDev* ds = (const StageDev*)(d + pb);
  const GateDev* dg = (const GateDev*)(d + pb + sb);
  const VT<T>* dt = (const VT<T>*)(d + off_t);
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    const PassDev& pd = prog.passes[p];
    const uint64_t tiles = (uint64_t)s->batch << pd.n_nonlocal;
    const int nt = 1 << (pd.k - kRegQubits);
    const size_t shm = sizeof(VT<T>) << pd.k;
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
    hipLaunchKernelGGL(k_fused_pass<T>, dim3((unsigned)tiles), dim3(nt), shm, ctx->stream,
                       (VT<T>*)s->d, s->n, dp + p, ds, dg, dt);
    SMOQ_CHECK_LAUNCH("k_fused_pass");
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
  }
  return SMOQ_OK;
}

int run_program(smoq_state* s, const Program& prog) {
  return s->dtype == SMOQ_C128 ? run_program_t<double>(s, prog) : run_program_t<float>(s, prog);
}

}  // namespace smoq
