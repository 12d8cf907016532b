// Device-side helpers shared by the kernel translation units.
#pragma once

// Also compiled by the run-time compiler (as embedded text), which brings its
// own HIP runtime declarations.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include <cstdint>

namespace smoq {

template <typename T>
struct CT;
template <>
struct CT<double> {
  using V = double2;
#ifndef SMOQ_ROWBITS_C128
#define SMOQ_ROWBITS_C128 3
#endif
  static constexpr int kRowBits = SMOQ_ROWBITS_C128;  // amplitudes per swizzle row (log2)
};
template <>
struct CT<float> {
  using V = float2;
#ifndef SMOQ_ROWBITS_C64
#define SMOQ_ROWBITS_C64 4
#endif
  static constexpr int kRowBits = SMOQ_ROWBITS_C64;
};

template <typename T>
using VT_alias = typename CT<T>::V;

template <typename T>
__host__ __device__ inline typename CT<T>::V make_c(double re, double im) {
  typename CT<T>::V v;
  v.x = (T)re;
  v.y = (T)im;
  return v;
}

template <typename V>
__device__ inline double2 to_d2(V a) {
  double2 r;
  r.x = (double)a.x;
  r.y = (double)a.y;
  return r;
}

template <typename V>
__device__ inline V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <typename V>
__device__ inline V cadd(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}

// u0*a + u1*b
template <typename V>
__device__ inline V cfma2(V u0, V a, V u1, V b) {
  return cadd(cmul(u0, a), cmul(u1, b));
}

template <typename T>
struct Mat2 {
  typename CT<T>::V m[4];
};
template <typename T>
struct Mat4 {
  typename CT<T>::V m[16];
};
template <typename T>
struct Diag8 {
  typename CT<T>::V p[8];
};

// index with a zero bit inserted at position q
__device__ inline uint64_t insert_zero(uint64_t t, int q) {
  const uint64_t low = t & ((1ull << q) - 1ull);
  return ((t >> q) << (q + 1)) | low;
}

// scatter the low bits of v into the set bits of mask (ascending)
__device__ inline uint32_t pdep32(uint32_t v, uint32_t mask) {
  uint32_t r = 0;
  while (mask) {
    uint32_t low = mask & (0u - mask);
    if (v & 1u) r |= low;
    v >>= 1;
    mask ^= low;
  }
  return r;
}

__device__ inline uint64_t pdep64(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  while (mask) {
    uint64_t low = mask & (0ull - mask);
    if (v & 1ull) r |= low;
    v >>= 1;
    mask ^= low;
  }
  return r;
}

// XOR swizzle of an LDS element index: the low WB bits (position inside a
// 128 B row) are XORed with a fold of all higher WB-bit groups.  A bijection
// on any power-of-two range; strided gathers spread over all row slots.
template <int WB>
__device__ inline uint32_t swz(uint32_t l) {
  uint32_t f = (l >> WB) ^ (l >> (2 * WB)) ^ (l >> (3 * WB)) ^ (l >> (4 * WB));
  return l ^ (f & ((1u << WB) - 1u));
}

}  // namespace smoq

namespace smoq {
// Sum over the wave by DPP lane moves (quad xor 1, quad xor 2, half-row
// mirror, row mirror: every lane of a 16-lane row then holds the row sum)
// and the four row sums read out in a fixed order; the result is uniform
// and the same on every run.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, 0xf, 0xf,
                                             false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double read_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double wave_total(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// component-wise select: a struct-level `c ? a : b` becomes a pointer select
// plus copy, which keeps register arrays from being promoted out of scratch
template <typename V>
__device__ __forceinline__ V csel(bool c, V a, V b) {
  V r;
  r.x = c ? a.x : b.x;
  r.y = c ? a.y : b.y;
  return r;
}
}  // namespace smoq
