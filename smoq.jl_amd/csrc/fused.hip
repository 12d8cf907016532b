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
constexpr uint32_t kAllPatterns = (1u << kNR) - 1u;
// bit j set for every register pattern j with slot t set
__host__ __device__ constexpr uint32_t slot_patterns(int t) {
  uint32_t m = 0;
  for (int j = 0; j < kNR; ++j)
    if ((j >> t) & 1) m |= 1u << j;
  return m;
}

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
  if constexpr (A >= kRegQubits) return;
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
  if constexpr (A >= kRegQubits || B >= kRegQubits) return;
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

// merged CZs: the lane builds a 16-bit mask of the slot patterns to negate
// (lane-lane pairs, register-register pairs, register-lane pairs)
template <typename T>
__device__ __forceinline__ void reg_czset(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl) {
  int par = 0;
  for (int i = 0; i < g.ncz; ++i) par ^= __popcll(gl & (gl >> g.czs[i]) & g.czm[i]);
  uint32_t neg = (par & 1) ? kAllPatterns : 0u;
  neg ^= (uint32_t)g.czreg;
#pragma unroll
  for (int t = 0; t < kRegQubits; ++t)
    neg ^= (__popcll(gl & g.czcm[t]) & 1) ? slot_patterns(t) : 0u;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    const bool f = (neg >> j) & 1u;
    x[j].x = f ? -x[j].x : x[j].x;
    x[j].y = f ? -x[j].y : x[j].y;
  }
}

// diagonal on register qubits only: uniform phase per slot pattern
template <typename T>
__device__ __forceinline__ void reg_diagreg(VT<T> (&x)[kNR], const GateDev& g) {
#pragma unroll
  for (int j = 0; j < kNR; ++j) x[j] = cmul(ld_c<T>(g.m, j), x[j]);
}

// diagonal on non-register qubits only: one phase per lane
template <typename T>
__device__ __forceinline__ void reg_diaglane(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl) {
  int row = 0;
  for (int t = 0; t < g.nu; ++t) row |= (int)((gl >> g.ubits[t]) & 1ull) << t;
  VT<T> ph = ld_c<T>(g.m, 0);
#pragma unroll
  for (int e = 1; e < (1 << kMaxTabBits); ++e) ph = csel(row == e, ld_c<T>(g.m, e), ph);
#pragma unroll
  for (int j = 0; j < kNR; ++j) x[j] = cmul(ph, x[j]);
}

template <typename T>
__device__ __forceinline__ void apply_gate(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl,
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
    case GK_CZSET: reg_czset<T>(x, g, gl); break;
    case GK_DIAGREG: reg_diagreg<T>(x, g); break;
    default: reg_diaglane<T>(x, g, gl); break;
  }
}

// T is the storage type of the state in HBM and LDS; in registers the
// amplitudes are complex128 whatever T is, so a complex64 state is rounded
// once per register stage instead of after every gate.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fused_pass(VT<T>* __restrict__ psi, int n, const PassDev* __restrict__ pass,
                 const StageDev* __restrict__ stages, const GateDev* __restrict__ gates,
                 const double2* __restrict__ tabs) {
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
  // tile id -> global bits outside the tile (uniform), thread index -> global
  // bits of local positions 0..k-R-1; the unrolled register index covers the
  // top R local positions
  const uint64_t gbase = pdep64(tt, pass->nl_mask);
  VT<T>* __restrict__ base = psi + (item << n);
  const uint64_t gt = gbase | pdep64((uint64_t)tid, pass->tid_mask);
  VT<T>* __restrict__ pt = base + gt;
  uint64_t hb[R];
#pragma unroll
  for (int j = 0; j < R; ++j) hb[j] = pass->hb[j];
  // The swizzle is linear over GF(2): swz(a | b) = swz(a) ^ swz(b) for
  // disjoint a, b, so every LDS address is a per-lane part XOR a uniform part.
  const uint32_t st_lane = swz<WB>((uint32_t)tid);
  {
    VT<T> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint64_t o = 0;  // uniform; disjoint from gt, so | is +
#pragma unroll
      for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? hb[j] : 0ull;
      x[i] = pt[o];
    }
#pragma unroll
    for (int i = 0; i < kNR; ++i) lds[st_lane ^ swz<WB>((uint32_t)(i * NT))] = x[i];
  }
  __syncthreads();

  for (int s = pass->s0; s < pass->s1; ++s) {
    const StageDev& st = stages[s];
    // lane part of the swizzled LDS index and of the global index, from the
    // host-precomputed images of each thread-index bit
    uint32_t sl = 0;
    uint64_t gl = gbase;
#pragma unroll
    for (int i = 0; i < kMaxTileQubits - R; ++i) {
      if (i < st.ntb && ((tid >> i) & 1)) {
        sl ^= st.sw_tid[i];
        gl |= st.g_tid[i];
      }
    }
    uint32_t sr[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint32_t r = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) r ^= ((i >> j) & 1) ? st.sw_reg[j] : 0u;
      sr[i] = r;
    }
    VT<double> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) x[i] = to_d2(lds[sl ^ sr[i]]);
    for (int g = st.g0; g < st.g1; ++g) apply_gate<double>(x, gates[g], gl, tabs);
#pragma unroll
    for (int i = 0; i < kNR; ++i) lds[sl ^ sr[i]] = make_c<T>(x[i].x, x[i].y);
    __syncthreads();
  }

  {
    VT<T> x[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) x[i] = lds[st_lane ^ swz<WB>((uint32_t)(i * NT))];
#pragma unroll
    for (int i = 0; i < kNR; ++i) {
      uint64_t o = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? hb[j] : 0ull;
      pt[o] = x[i];
    }
  }
}

ProgLayout program_layout(const Program& prog, int dtype) {
  ProgLayout L;
  L.pb = prog.passes.size() * sizeof(PassDev);
  L.sb = prog.stages.size() * sizeof(StageDev);
  L.gb = prog.gates.size() * sizeof(GateDev);
  L.ntab = prog.tabs.size() / 2;
  (void)dtype;  // phase tables are complex128 for every state dtype
  L.off_t = (L.pb + L.sb + L.gb + 255) & ~(size_t)255;
  L.total = L.off_t + (L.ntab + 1) * sizeof(double2);
  return L;
}

// copies the program into device buffer d (L.total bytes) through the
// context's pinned staging buffer, ordered on the context stream
int upload_program(smoq_ctx* ctx, const Program& prog, int dtype, const ProgLayout& L, void* d) {
  int rc = ensure_pinned(ctx, L.total);
  if (rc) return rc;
  // the pinned staging buffer may still feed an earlier copy
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  std::memcpy(h, prog.passes.data(), L.pb);
  std::memcpy(h + L.pb, prog.stages.data(), L.sb);
  std::memcpy(h + L.pb + L.sb, prog.gates.data(), L.gb);
  (void)dtype;
  if (L.ntab) std::memcpy(h + L.off_t, prog.tabs.data(), L.ntab * sizeof(double2));
  SMOQ_HIP(hipMemcpyAsync(d, h, L.total, hipMemcpyHostToDevice, ctx->stream));
  return SMOQ_OK;
}

template <typename T>
static int launch_program_t(smoq_state* s, const Program& prog, const ProgLayout& L,
                            const void* dbuf) {
  smoq_ctx* ctx = s->ctx;
  int rc = SMOQ_OK;
  const unsigned char* d = (const unsigned char*)dbuf;
  const PassDev* dp = (const PassDev*)d;
  const StageDev* ds = (const StageDev*)(d + L.pb);
  const GateDev* dg = (const GateDev*)(d + L.pb + L.sb);
  const double2* dt = (const double2*)(d + L.off_t);
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    const PassDev& pd = prog.passes[p];
    const uint64_t tiles = (uint64_t)s->batch << pd.n_nonlocal;
    const int nt = 1 << (pd.k - kRegQubits);
    if (nt > 256 || pd.k > kMaxTileQubits)
      return set_error(SMOQ_EINVAL, "fused pass: tile of 2^%d needs %d threads (max 256)", pd.k,
                       nt);
    const size_t shm = sizeof(VT<T>) << pd.k;
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
    hipLaunchKernelGGL(k_fused_pass<T>, dim3((unsigned)tiles), dim3(nt), shm, ctx->stream,
                       (VT<T>*)s->d, s->n, dp + p, ds, dg, dt);
    SMOQ_CHECK_LAUNCH("k_fused_pass");
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
  }
  return SMOQ_OK;
}

int launch_program(smoq_state* s, const Program& prog, const ProgLayout& L, const void* dbuf) {
  return s->dtype == SMOQ_C128 ? launch_program_t<double>(s, prog, L, dbuf)
                               : launch_program_t<float>(s, prog, L, dbuf);
}

int run_program(smoq_state* s, const Program& prog) {
  const ProgLayout L = program_layout(prog, s->dtype);
  int rc = ensure_prog(s->ctx, L.total);
  if (rc) return rc;
  if ((rc = upload_program(s->ctx, prog, s->dtype, L, s->ctx->prog))) return rc;
  return launch_program(s, prog, L, s->ctx->prog);
}

}  // namespace smoq
