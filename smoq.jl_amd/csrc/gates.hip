// One-gate streaming kernels for gfx950 and the basis-state fill.
//
// Each thread owns one amplitude pair (quad for 2-qubit gates).  Pair index
// t enumerates the indices with the target bit cleared; inserting a zero at
// the target keeps lanes on consecutive addresses for both halves of the
// pair, so every load and store is a coalesced 16 B (c128) / 8 B (c64)
// access.  Batch items are contiguous, so the bits above n carry the item.
// T is the storage type; the arithmetic is complex128 for both dtypes.
#include <cmath>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

template <typename T>
__global__ void __launch_bounds__(256)
    k_u1(VT_alias<T>* __restrict__ psi, uint64_t npairs, int q, uint64_t ctrl, Mat2<double> u) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npairs;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = insert_zero(t, q);
    if ((lo & ctrl) != ctrl) continue;
    const uint64_t hi = lo | (1ull << q);
    const double2 a = to_d2(psi[lo]), b = to_d2(psi[hi]);
    const double2 na = cfma2(u.m[0], a, u.m[1], b);
    const double2 nb = cfma2(u.m[2], a, u.m[3], b);
    psi[lo] = make_c<T>(na.x, na.y);
    psi[hi] = make_c<T>(nb.x, nb.y);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_u2(VT_alias<T>* __restrict__ psi, uint64_t nquads, int qa, int qb, uint64_t ctrl,
         Mat4<double> u) {
  const int lo_b = qa < qb ? qa : qb, hi_b = qa < qb ? qb : qa;
  const uint64_t ma = 1ull << qa, mb = 1ull << qb;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nquads;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i0 = insert_zero(insert_zero(t, lo_b), hi_b);
    if ((i0 & ctrl) != ctrl) continue;
    const uint64_t idx[4] = {i0, i0 | ma, i0 | mb, i0 | ma | mb};
    double2 x[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = to_d2(psi[idx[e]]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double2 acc = cmul(u.m[4 * r], x[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(u.m[4 * r + c], x[c]));
      psi[idx[r]] = make_c<T>(acc.x, acc.y);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_diag(VT_alias<T>* __restrict__ psi, uint64_t total, uint64_t dim_mask, int nd, int q0,
           int q1, int q2, Diag8<double> ph) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = i & dim_mask;
    int e = (int)((g >> q0) & 1);
    if (nd > 1) e |= (int)((g >> q1) & 1) << 1;
    if (nd > 2) e |= (int)((g >> q2) & 1) << 2;
    const double2 v = cmul(ph.p[e], to_d2(psi[i]));
    psi[i] = make_c<T>(v.x, v.y);
  }
}

template <typename T>
static int launch_gate_t(smoq_state* s, const IGate& g) {
  VT_alias<T>* psi = (VT_alias<T>*)s->d;
  hipStream_t st = s->ctx->stream;
  const uint64_t total = s->dim * (uint64_t)s->batch;
  const int B = 256;
  if (g.kind == IK_U1) {
    Mat2<double> u;
    for (int i = 0; i < 4; ++i) u.m[i] = make_c<double>(g.m[2 * i], g.m[2 * i + 1]);
    const uint64_t np = total >> 1;
    hipLaunchKernelGGL(k_u1<T>, dim3(stride_blocks(np, B)), dim3(B), 0, st, psi, np, g.t0,
                       g.ctrl, u);
  } else if (g.kind == IK_U2) {
    Mat4<double> u;
    for (int i = 0; i < 16; ++i) u.m[i] = make_c<double>(g.m[2 * i], g.m[2 * i + 1]);
    const uint64_t nq = total >> 2;
    hipLaunchKernelGGL(k_u2<T>, dim3(stride_blocks(nq, B)), dim3(B), 0, st, psi, nq, g.t0,
                       g.t1, g.ctrl, u);
  } else {
    Diag8<double> ph;
    for (int i = 0; i < (1 << g.nd); ++i) ph.p[i] = make_c<double>(g.m[2 * i], g.m[2 * i + 1]);
    hipLaunchKernelGGL(k_diag<T>, dim3(stride_blocks(total, B)), dim3(B), 0, st, psi, total,
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
    k_set_basis(VT_alias<T>* __restrict__ psi, uint64_t total, uint64_t dim_mask,
                uint64_t index) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    psi[i] = make_c<T>((i & dim_mask) == index ? 1.0 : 0.0, 0.0);
}

// Product state: amplitude i = c_1[b_1] c_2[b_2] ... c_n[b_n] with b_k bit
// k-1 of i, multiplied left to right as tensor_product_ket builds it
// (utils/cpu/cpuQuantumStatePreparation.jl:1108-1123,1182-1191).
struct ProductKets {
  double2 c[kMaxQubits][2];
};

template <typename T>
__global__ void __launch_bounds__(256)
    k_set_product(VT_alias<T>* __restrict__ psi, uint64_t total, int n, ProductKets pk) {
#pragma clang fp contract(off)
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = i & ((1ull << n) - 1ull);
    double2 p = pk.c[0][k & 1];
    for (int q = 1; q < n; ++q) {
      const double2 c = pk.c[q][(k >> q) & 1];
      p = make_double2(p.x * c.x - p.y * c.y, p.x * c.y + p.y * c.x);
    }
    psi[i] = make_c<T>(p.x, p.y);
  }
}

}  // namespace smoq

extern "C" int smoq_state_set_product(smoq_state* s, const double* kets) {
  using namespace smoq;
  int rc = check_state(s);
  if (rc) return rc;
  if (!kets) return set_error(SMOQ_EINVAL, "smoq_state_set_product: null kets");
  if (s->n > kMaxQubits)
    return set_error(SMOQ_ENOTSUP, "smoq_state_set_product: %d qubits (max %d)", s->n,
                     kMaxQubits);
  ProductKets pk;
  std::memset(&pk, 0, sizeof(pk));
  for (int q = 0; q < s->n; ++q)
    for (int b = 0; b < 2; ++b) {
      const double re = kets[4 * q + 2 * b], im = kets[4 * q + 2 * b + 1];
      if (!std::isfinite(re) || !std::isfinite(im))
        return set_error(SMOQ_EINVAL, "smoq_state_set_product: non-finite entry for qubit %d", q);
      pk.c[q][b] = make_double2(re, im);
    }
  const uint64_t total = s->dim * (uint64_t)s->batch;
  const unsigned nb = stride_blocks(total, 256);
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_set_product<double>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (double2*)s->d, total, s->n, pk);
  else
    hipLaunchKernelGGL(k_set_product<float>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (float2*)s->d, total, s->n, pk);
  SMOQ_CHECK_LAUNCH("k_set_product");
  return SMOQ_OK;
}

extern "C" int smoq_state_set_basis(smoq_state* s, uint64_t index) {
  using namespace smoq;
  int rc = check_state(s);
  if (rc) return rc;
  if (index >= s->dim)
    return set_error(SMOQ_EINVAL, "smoq_state_set_basis: index %llu >= 2^%d",
                     (unsigned long long)index, s->n);
  const uint64_t total = s->dim * (uint64_t)s->batch;
  const unsigned nb = stride_blocks(total, 256);
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_set_basis<double>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (double2*)s->d, total, s->dim - 1, index);
  else
    hipLaunchKernelGGL(k_set_basis<float>, dim3(nb), dim3(256), 0, s->ctx->stream,
                       (float2*)s->d, total, s->dim - 1, index);
  SMOQ_CHECK_LAUNCH("k_set_basis");
  return SMOQ_OK;
}
