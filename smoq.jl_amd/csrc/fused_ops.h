// Device code of the fused gate passes: the register-stage gate helpers and
// the tile/stage frame shared by the precompiled pass kernel (fused.hip) and
// the pass kernels compiled at run time for one circuit structure (jit.hip),
// which see this file as embedded text.
#pragma once

#include "device_common.h"
#include "fused_types.h"

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

// Factored rotations (GK_U1_ROT, GK_U1_RXR) on slot A without controls:
// one fused multiply-add per output component; the factor is applied when
// the pass stores its tile.  FORM 0: unit diagonal, 1: unit anti-diagonal.
template <typename T, int PT, int A, int FORM>
__device__ __forceinline__ void reg_u1s(VT<T> (&x)[kNR], const GateDev& g) {
  if constexpr (A >= kRegQubits) return;
  const T t = (T)g.m[0];
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & (1 << A)) continue;
    const int h = j | (1 << A);
    const VT<T> a = x[j], b = x[h];
    if constexpr (FORM == 0 && PT == GK_U1_ROT) {  // (a - t b, t a + b)
      x[j].x = fma(-t, b.x, a.x);
      x[j].y = fma(-t, b.y, a.y);
      x[h].x = fma(t, a.x, b.x);
      x[h].y = fma(t, a.y, b.y);
    } else if constexpr (FORM == 0) {  // (a + i t b, i t a + b)
      x[j].x = fma(-t, b.y, a.x);
      x[j].y = fma(t, b.x, a.y);
      x[h].x = fma(-t, a.y, b.x);
      x[h].y = fma(t, a.x, b.y);
    } else if constexpr (PT == GK_U1_ROT) {  // (t a - b, a + t b)
      x[j].x = fma(t, a.x, -b.x);
      x[j].y = fma(t, a.y, -b.y);
      x[h].x = fma(t, b.x, a.x);
      x[h].y = fma(t, b.y, a.y);
    } else {  // (t a + i b, i a + t b)
      x[j].x = fma(t, a.x, -b.y);
      x[j].y = fma(t, a.y, b.x);
      x[h].x = fma(t, b.x, -a.y);
      x[h].y = fma(t, b.y, a.x);
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

// row r of a 4x4 with the exactly-zero entries of Z left out
template <typename T, uint32_t Z>
__device__ __forceinline__ VT<T> u2_row_z(const double* m, int r, VT<T> a0, VT<T> a1, VT<T> a2,
                                          VT<T> a3) {
  const VT<T> in[4] = {a0, a1, a2, a3};
  VT<T> acc = make_c<T>(0.0, 0.0);
  bool first = true;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if ((Z >> (4 * r + c)) & 1u) continue;
    const VT<T> p = cmul(ld_c<T>(m, 4 * r + c), in[c]);
    acc = first ? p : cadd(acc, p);
    first = false;
  }
  return acc;
}

// 4x4 on slots (A low role, B high role) whose zero entries Z are known
// when the kernel is compiled (a two-qubit XXZ bond conserves the number of
// ones: six entries of sixteen)
template <typename T, int A, int B, uint32_t Z, bool CTRL = true>
__device__ __forceinline__ void reg_u2z(VT<T> (&x)[kNR], const GateDev& g, bool gon) {
  if constexpr (A >= kRegQubits || B >= kRegQubits) return;
  const int creg = CTRL ? g.creg : 0;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & ((1 << A) | (1 << B))) continue;
    const int id0 = j, id1 = j | (1 << A), id2 = j | (1 << B), id3 = j | (1 << A) | (1 << B);
    const VT<T> in0 = x[id0], in1 = x[id1], in2 = x[id2], in3 = x[id3];
    const VT<T> o0 = u2_row_z<T, Z>(g.m, 0, in0, in1, in2, in3);
    const VT<T> o1 = u2_row_z<T, Z>(g.m, 1, in0, in1, in2, in3);
    const VT<T> o2 = u2_row_z<T, Z>(g.m, 2, in0, in1, in2, in3);
    const VT<T> o3 = u2_row_z<T, Z>(g.m, 3, in0, in1, in2, in3);
    if constexpr (CTRL) {
      const bool on = gon && ((j & creg) == creg);
      x[id0] = csel(on, o0, in0);
      x[id1] = csel(on, o1, in1);
      x[id2] = csel(on, o2, in2);
      x[id3] = csel(on, o3, in3);
    } else {
      x[id0] = o0;
      x[id1] = o1;
      x[id2] = o2;
      x[id3] = o3;
    }
  }
}

// 4x4 on slots (A low role, B high role)
// CTRL false: no controls (known when the kernel is compiled), no selects
template <typename T, int A, int B, bool CTRL = true>
__device__ __forceinline__ void reg_u2(VT<T> (&x)[kNR], const GateDev& g, bool gon) {
  if constexpr (A >= kRegQubits || B >= kRegQubits) return;
  const int creg = CTRL ? g.creg : 0;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & ((1 << A) | (1 << B))) continue;
    const int id0 = j, id1 = j | (1 << A), id2 = j | (1 << B), id3 = j | (1 << A) | (1 << B);
    const VT<T> in0 = x[id0], in1 = x[id1], in2 = x[id2], in3 = x[id3];
    const VT<T> o0 = u2_row<T>(g.m, 0, in0, in1, in2, in3);
    const VT<T> o1 = u2_row<T>(g.m, 1, in0, in1, in2, in3);
    const VT<T> o2 = u2_row<T>(g.m, 2, in0, in1, in2, in3);
    const VT<T> o3 = u2_row<T>(g.m, 3, in0, in1, in2, in3);
    if constexpr (CTRL) {
      const bool on = gon && ((j & creg) == creg);
      x[id0] = csel(on, o0, in0);
      x[id1] = csel(on, o1, in1);
      x[id2] = csel(on, o2, in2);
      x[id3] = csel(on, o3, in3);
    } else {
      x[id0] = o0;
      x[id1] = o1;
      x[id2] = o2;
      x[id3] = o3;
    }
  }
}

// merged diagonal gates: x[j] *= table[row][j], row from non-register bits
// (NU >= 0: the number of those bits, known when the kernel is compiled)
template <typename T, int NU = -1>
__device__ __forceinline__ void reg_diagtab(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl,
                                            const VT<T>* __restrict__ tabs) {
  const int nu = NU >= 0 ? NU : g.nu;
  int row = 0;
  for (int t = 0; t < nu; ++t) row |= (int)((gl >> g.ubits[t]) & 1ull) << t;
  const VT<T>* p = tabs + g.tab + row * kNR;
#pragma unroll
  for (int j = 0; j < kNR; ++j) x[j] = cmul(p[j], x[j]);
}

// negates x[j] for every register pattern j set in neg: the sign bit of
// both components flips by an integer XOR of the high word
template <typename T>
__device__ __forceinline__ void czset_apply(VT<T> (&x)[kNR], uint32_t neg) {
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    const int s = (int)((neg >> j) & 1u) << 31;
    x[j].x = __hiloint2double(__double2hiint(x[j].x) ^ s, __double2loint(x[j].x));
    x[j].y = __hiloint2double(__double2hiint(x[j].y) ^ s, __double2loint(x[j].y));
  }
}

// XOR of the sign bits of both components of a register amplitude with s
// (0 or 0x80000000)
__device__ __forceinline__ void czsign(VT<double>& a, uint32_t s) {
  a.x = __hiloint2double(__double2hiint(a.x) ^ (int)s, __double2loint(a.x));
  a.y = __hiloint2double(__double2hiint(a.y) ^ (int)s, __double2loint(a.y));
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
  czset_apply<T>(x, neg);
}

// diagonal on register qubits only: uniform phase per slot pattern
template <typename T>
__device__ __forceinline__ void reg_diagreg(VT<T> (&x)[kNR], const GateDev& g) {
#pragma unroll
  for (int j = 0; j < kNR; ++j) x[j] = cmul(ld_c<T>(g.m, j), x[j]);
}

// register diagonal with pattern 0 normalised to 1: patterns in MASK only
template <typename T, uint32_t MASK>
__device__ __forceinline__ void reg_diagreg_m(VT<T> (&x)[kNR], const GateDev& g) {
#pragma unroll
  for (int j = 0; j < kNR; ++j)
    if ((MASK >> j) & 1u) x[j] = cmul(ld_c<T>(g.m, j), x[j]);
}

// diagonal on non-register qubits only: one phase per lane (NU as above)
template <typename T, int NU = -1>
__device__ __forceinline__ void reg_diaglane(VT<T> (&x)[kNR], const GateDev& g, uint64_t gl) {
  const int nu = NU >= 0 ? NU : g.nu;
  constexpr int kRows = NU >= 0 ? (1 << NU) : (1 << kMaxTabBits);
  int row = 0;
  for (int t = 0; t < nu; ++t) row |= (int)((gl >> g.ubits[t]) & 1ull) << t;
  VT<T> ph = ld_c<T>(g.m, 0);
#pragma unroll
  for (int e = 1; e < kRows; ++e) ph = csel(row == e, ld_c<T>(g.m, e), ph);
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

// Adjoint overlap (GK_OVL): the register patterns with the selector slot
// SEL clear hold ψ, those with it set λ at the same index.  For every ψ
// pattern j (index k) the thread adds Im[conj(λ_k) (Gψ)_k] with
// (Gψ)_k = i^CNT (-1)^{|(k ^ x) & z|} ψ_{k ^ x}, x on slots A, B (-1: none),
// for k inside the controls.  With ROWS each 16-lane row's sum goes to
// acc[row of the workgroup] (four DPP steps, no cross-row reads), else each
// wave's sum to acc[wave]; written once per tile, the pass adds the slots in
// order at its end.  NT: threads per tile; below a wave, the sum is a
// butterfly over the tile's lanes into acc[0].
template <int A, int B, int SEL, int CNT, bool CTRL, int NT, bool ROWS>
__device__ __forceinline__ void reg_ovl(const VT<double> (&x)[kNR], const GateDev& g, uint64_t gl,
                                        double* acc) {
  constexpr int xr = (A >= 0 ? 1 << A : 0) | (B >= 0 ? 1 << B : 0);
  constexpr int sm = 1 << SEL;
  const int zr = g.ozreg;
  const int zp = __popcll(gl & g.ozglob) & 1;
  const bool gon = CTRL ? (gl & g.cglob) == g.cglob : true;
  double ov = 0.0;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & sm) continue;
    const int jp = j ^ xr;
    const VT<double> l = x[j | sm];
    const VT<double> w = x[jp];
    double ux, uy;
    if (CNT == 0) {
      ux = w.x;
      uy = w.y;
    } else if (CNT == 1) {
      ux = -w.y;
      uy = w.x;
    } else if (CNT == 2) {
      ux = -w.x;
      uy = -w.y;
    } else {
      ux = w.y;
      uy = -w.x;
    }
    const double c = l.x * uy - l.y * ux;
    // the pattern's sign, and with controls whether it counts, are the same
    // on every lane: a uniform factor of +-1 or 0
    double sj = (__popc(jp & zr) & 1) ? -1.0 : 1.0;
    if (CTRL) sj = ((j & g.creg) == g.creg) ? sj : 0.0;
    ov = fma(c, sj, ov);
  }
  // the lane's own sign and controls (negation commutes with rounding)
  if (zp) ov = -ov;
  if (CTRL && !gon) ov = 0.0;
  if constexpr (NT >= 64 && ROWS) {
    ov += dpp_move<0xB1>(ov);   // quad_perm [1,0,3,2]
    ov += dpp_move<0x4E>(ov);   // quad_perm [2,3,0,1]
    ov += dpp_move<0x141>(ov);  // row_half_mirror
    ov += dpp_move<0x140>(ov);  // row_mirror
    if ((threadIdx.x & 15) == 0) acc[threadIdx.x >> 4] = ov;
  } else if constexpr (NT >= 64) {
    ov = wave_total(ov);
    if ((threadIdx.x & 63) == 0) acc[threadIdx.x >> 6] = ov;
  } else {
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) ov += __shfl_xor(ov, off, NT);
    if (threadIdx.x == 0) acc[0] = ov;
  }
}

// ---------------------------------------------------------------- frames --
// One workgroup, one tile of 2^k amplitudes: the tile id picks the batch
// item and the global bits outside the tile (uniform), the thread index the
// global bits of local positions 0..k-R-1; the unrolled register index
// covers the top R local positions (uniform offsets).
template <typename T>
struct PassFrame {
  VT<T>* pt;         // amplitude of register index 0 of this thread
  uint64_t tile;     // tile index over the whole launch (batch items included)
  uint64_t gbase;    // global bits outside the tile
  uint64_t off[kNR]; // offsets of register index i
  uint32_t st_lane;  // swizzled LDS index of the thread's lane part
  uint32_t sreg[kNR];  // swizzled LDS index of register index i's part
};

// the tile of launch index tt (nn bits): rotated left by tord & 31 after a
// bit reversal when tord & 32 (a bijection on the nn-bit range)
__device__ __forceinline__ uint64_t tile_order(uint64_t tt, int nn, int tord) {
  if (nn <= 0 || tord == 0) return tt;
  if (tord & 32) tt = __brevll(tt) >> (64 - nn);
  const int r = (tord & 31) % nn;
  if (r) tt = ((tt << r) | (tt >> (nn - r))) & ((1ull << nn) - 1ull);
  return tt;
}

// Tile loads and stores of a pass touch every amplitude once: nontemporal
// (streaming) accesses keep them from displacing cached lines.
template <typename T>
__device__ __forceinline__ VT<T> ld_stream(const VT<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  VT<T> r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
template <typename T>
__device__ __forceinline__ void st_stream(VT<T>* p, VT<T> a) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = a.x;
  v.y = a.y;
  __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}

// HBM -> LDS for the workgroup's tile; ends with a barrier
template <typename T>
__device__ __forceinline__ void pass_load(VT<T>* lds, VT<T>* __restrict__ psi, int n,
                                          const PassDev* __restrict__ pass, PassFrame<T>& f) {
  constexpr int R = kRegQubits;
  constexpr int WB = CT<T>::kRowBits;
  const int nn = pass->n_nonlocal;
  // tiles beyond one grid row continue in blockIdx.y (states of 2^32 threads
  // and more per pass)
  const uint64_t tile = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const uint64_t item = tile >> nn;
  const uint64_t tt = tile_order(tile & ((1ull << nn) - 1ull), nn, pass->tord);
  f.tile = (item << nn) | tt;
  f.gbase = pdep64(tt, pass->nl_mask);
  f.pt = psi + (item << n) + (f.gbase | pdep64((uint64_t)threadIdx.x, pass->tid_mask));
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? pass->hb[j] : 0ull;
    f.off[i] = o;
  }
  // The swizzle is linear over GF(2): swz(a | b) = swz(a) ^ swz(b) for
  // disjoint a, b, so every LDS address is a per-lane part XOR a uniform part.
  const uint32_t ltm = pass->ltm;
  const uint32_t lrm = ((1u << pass->k) - 1u) & ~ltm;
  f.st_lane = swz<WB>(pdep32((uint32_t)threadIdx.x, ltm));
#pragma unroll
  for (int i = 0; i < kNR; ++i) f.sreg[i] = swz<WB>(pdep32((uint32_t)i, lrm));
  VT<T> x[kNR];
#pragma unroll
  for (int i = 0; i < kNR; ++i) x[i] = ld_stream<T>(f.pt + f.off[i]);
#pragma unroll
  for (int i = 0; i < kNR; ++i) lds[f.st_lane ^ f.sreg[i]] = x[i];
  __syncthreads();
}

// LDS -> HBM for the tile (after the last stage's barrier)
template <typename T>
__device__ __forceinline__ void pass_store(const VT<T>* lds, const PassFrame<T>& f) {
  VT<T> x[kNR];
#pragma unroll
  for (int i = 0; i < kNR; ++i) x[i] = lds[f.st_lane ^ f.sreg[i]];
#pragma unroll
  for (int i = 0; i < kNR; ++i) st_stream<T>(f.pt + f.off[i], x[i]);
}

// LDS -> HBM with the pass's factored-out scalar applied
template <typename T>
__device__ __forceinline__ void pass_store_scaled(const VT<T>* lds, const PassFrame<T>& f,
                                                  double scale) {
  const T sc = (T)scale;
  VT<T> x[kNR];
#pragma unroll
  for (int i = 0; i < kNR; ++i) x[i] = lds[f.st_lane ^ f.sreg[i]];
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    x[i].x *= sc;
    x[i].y *= sc;
    st_stream<T>(f.pt + f.off[i], x[i]);
  }
}

// LDS -> HBM with the pass's complex factor (rotation scalars and the
// phases taken out of its register diagonals) applied
template <typename T>
__device__ __forceinline__ void pass_store_cscaled(const VT<T>* lds, const PassFrame<T>& f,
                                                   double sre, double sim) {
  VT<T> x[kNR];
#pragma unroll
  for (int i = 0; i < kNR; ++i) x[i] = lds[f.st_lane ^ f.sreg[i]];
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    VT<double> v;
    v.x = (double)x[i].x * sre - (double)x[i].y * sim;
    v.y = (double)x[i].x * sim + (double)x[i].y * sre;
    st_stream<T>(f.pt + f.off[i], make_c<T>(v.x, v.y));
  }
}

// One register stage: 2^R amplitudes of the thread (its lane part plus the
// stage's register qubits) move from LDS into complex128 registers.
// LDS addresses are kept in bytes: an access is then one XOR of the lane
// part with a uniform (mostly compile-time) part.
struct StageFrame {
  uint32_t sl;       // lane part of the swizzled LDS address (bytes)
  uint64_t gl;       // global index of the thread's register index 0
  uint32_t sr[kNR];  // swizzled LDS offsets of the register patterns (bytes)
};

template <typename T>
__device__ __forceinline__ const VT<T>& lds_at(const VT<T>* lds, uint32_t byte_off) {
  return *reinterpret_cast<const VT<T>*>(reinterpret_cast<const unsigned char*>(lds) + byte_off);
}
template <typename T>
__device__ __forceinline__ VT<T>& lds_at(VT<T>* lds, uint32_t byte_off) {
  return *reinterpret_cast<VT<T>*>(reinterpret_cast<unsigned char*>(lds) + byte_off);
}

// NTB >= 0: the stage's thread-index bits, known when the kernel is compiled
template <typename T, int NTB = -1>
__device__ __forceinline__ void stage_load(const VT<T>* lds, const StageDev& st, uint64_t gbase,
                                           VT<double> (&x)[kNR], StageFrame& sf) {
  constexpr int R = kRegQubits;
  const uint32_t tid = threadIdx.x;
  const int ntb = NTB >= 0 ? NTB : st.ntb;
  // lane part of the swizzled LDS index and of the global index, from the
  // host-precomputed images of each thread-index bit (masks, no branches)
  sf.sl = 0;
  sf.gl = gbase;
#pragma unroll
  for (int i = 0; i < kMaxTileQubits - R; ++i) {
    if (i < ntb) {
      const uint32_t b = (tid >> i) & 1u;
      sf.sl ^= st.sw_tid[i] & (0u - b);
      sf.gl |= st.g_tid[i] & (0ull - (uint64_t)b);
    }
  }
  sf.sl *= (uint32_t)sizeof(VT<T>);
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) r ^= ((i >> j) & 1) ? st.sw_reg[j] : 0u;
    sf.sr[i] = r * (uint32_t)sizeof(VT<T>);
  }
#pragma unroll
  for (int i = 0; i < kNR; ++i) x[i] = to_d2(lds_at<T>(lds, sf.sl ^ sf.sr[i]));
}

// The same with the stage's addressing known when the kernel is compiled:
// l_lane is the thread's local index (register positions zero), gl its
// global index, sr the swizzled offsets of the register patterns.
template <typename T>
__device__ __forceinline__ void stage_load_c(const VT<T>* lds, uint32_t l_lane, uint64_t gl,
                                             const uint32_t (&sr)[kNR], VT<double> (&x)[kNR],
                                             StageFrame& sf) {
  sf.sl = swz<CT<T>::kRowBits>(l_lane) * (uint32_t)sizeof(VT<T>);
  sf.gl = gl;
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    sf.sr[i] = sr[i] * (uint32_t)sizeof(VT<T>);
    x[i] = to_d2(lds_at<T>(lds, sf.sl ^ sf.sr[i]));
  }
}

// The stage's mapping after a lane exchange: the swizzled LDS index of the
// thread's lane part, its global index and the register patterns' swizzled
// offsets, all fixed when the kernel is compiled.
template <typename T>
__device__ __forceinline__ void stage_remap(uint32_t sl, uint64_t gl, const uint32_t (&sr)[kNR],
                                            StageFrame& sf) {
  sf.sl = sl * (uint32_t)sizeof(VT<T>);
  sf.gl = gl;
#pragma unroll
  for (int i = 0; i < kNR; ++i) sf.sr[i] = sr[i] * (uint32_t)sizeof(VT<T>);
}

// The mapping after a lane exchange that moved thread bit LB to another
// local position: the lane part and the global index change by that bit's
// images dsw (swizzled LDS index) and dg (global bits); the register
// patterns' offsets are the new mapping's constants.
template <typename T>
__device__ __forceinline__ void stage_swap_remap(uint32_t tid, int lb, uint32_t dsw, uint64_t dg,
                                                 const uint32_t (&sr)[kNR], StageFrame& sf) {
  const bool b = (tid >> lb) & 1u;
  sf.sl ^= b ? dsw * (uint32_t)sizeof(VT<T>) : 0u;
  sf.gl ^= b ? dg : 0ull;
#pragma unroll
  for (int i = 0; i < kNR; ++i) sf.sr[i] = sr[i] * (uint32_t)sizeof(VT<T>);
}

// v_permlane16_swap (LB = 4) / v_permlane32_swap (LB = 5) on the two dwords
// of a double: a's lanes with lane bit LB set trade values with b's lanes
// with it clear (partner lane: lane ^ 2^LB).
template <int LB>
__device__ __forceinline__ void permlane_swap(double& a, double& b) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  u32x2 lo, hi;
  if constexpr (LB == 4) {
    lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  } else {
    lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  }
  a = __longlong_as_double((long long)(((unsigned long long)hi.x << 32) | lo.x));
  b = __longlong_as_double((long long)(((unsigned long long)hi.y << 32) | lo.y));
}

// Trades the qubit of register slot R with the qubit of lane bit LB: for
// every pair (j, j | 2^R) the pattern-j amplitudes of lanes with bit LB set
// swap with the pattern-(j | 2^R) amplitudes of their partner lanes, which
// is the exchange of the two index bits.
template <int R, int LB>
__device__ __forceinline__ void lane_swap(VT<double> (&x)[kNR]) {
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & (1 << R)) continue;
    const int h = j | (1 << R);
    permlane_swap<LB>(x[j].x, x[h].x);
    permlane_swap<LB>(x[j].y, x[h].y);
  }
}

// registers -> LDS (rounded to T), then a barrier
template <typename T>
__device__ __forceinline__ void stage_store(VT<T>* lds, const StageFrame& sf,
                                            const VT<double> (&x)[kNR]) {
#pragma unroll
  for (int i = 0; i < kNR; ++i) lds_at<T>(lds, sf.sl ^ sf.sr[i]) = make_c<T>(x[i].x, x[i].y);
  __syncthreads();
}

}  // namespace smoq
