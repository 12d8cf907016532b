// Density matrices on the device.  rho of N qubits is held as vec(rho), a
// 2N-qubit state: element (i, j) at index i + 2^N j (column-major, as a Julia
// Matrix{ComplexF64} lies in memory).  Gates and dissipators act on it as
// 2N-qubit gates (U on the row qubits, conj(U) on the column qubits; a
// dissipator as a 4x4 map on the pair (k, N + k)), so they run on the fused
// passes.  This file holds what is specific to rho: the trace functional
// Tr(rho P) and the dense propagation rho <- U rho U^H.
#include <rocblas/rocblas.h>

#include <cmath>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kDmBlock = 256;

struct DmTerm {
  uint64_t x, z;
  double fre, fim;  // i^{|x & z|}
};

// Tr(rho P) = sum_j rho[j ^ x, j] i^{|x&z|} (-1)^{|(j ^ x) & z|}: one block
// per (term, item), fixed summation order; out = real part.
template <typename T>
__global__ void __launch_bounds__(kDmBlock)
    k_dm_expect(const VT_alias<T>* __restrict__ rho, int n, uint64_t vdim,
                const DmTerm* __restrict__ terms, int nt, double* __restrict__ out) {
  __shared__ double sh[kDmBlock / 64];
  const int t = blockIdx.x;
  const uint64_t item = blockIdx.y;
  const DmTerm tm = terms[t];
  const VT_alias<T>* r = rho + item * vdim;
  const uint64_t dim = 1ull << n;
  double acc = 0.0;
  for (uint64_t j = threadIdx.x; j < dim; j += kDmBlock) {
    const uint64_t i = j ^ tm.x;
    const VT_alias<T> v = r[i + (j << n)];
    const double s = (__popcll(i & tm.z) & 1) ? -1.0 : 1.0;
    // Re(phase * v)
    acc += s * (tm.fre * (double)v.x - tm.fim * (double)v.y);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < kDmBlock / 64; ++w) v += sh[w];
    out[item * nt + t] = v;
  }
}

int blas_handle(smoq_ctx* ctx, const char* who, void** out) {
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return set_error(SMOQ_EDEVICE, "%s: rocblas_create_handle failed", who);
    ctx->blas = h;
  }
  if (rocblas_set_stream((rocblas_handle)ctx->blas, ctx->stream) != rocblas_status_success)
    return set_error(SMOQ_EDEVICE, "%s: rocblas_set_stream failed", who);
  *out = ctx->blas;
  return SMOQ_OK;
}

void blas_release(smoq_ctx* ctx) {
  if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
  ctx->blas = nullptr;
}

static int dm_qubits(const smoq_state* rho, const char* who, int* N) {
  if (rho->n % 2)
    return set_error(SMOQ_EINVAL, "%s: a density matrix is a state of 2N qubits, got %d", who,
                     rho->n);
  *N = rho->n / 2;
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_dm_expect_pauli(const smoq_state* rho, int n_terms, const uint64_t* xmask,
                                    const uint64_t* zmask, double* out) {
  int rc = check_state(rho);
  if (rc) return rc;
  int N = 0;
  if ((rc = dm_qubits(rho, "smoq_dm_expect_pauli", &N))) return rc;
  if (n_terms < 0 || (n_terms > 0 && (!xmask || !zmask || !out)))
    return set_error(SMOQ_EINVAL, "smoq_dm_expect_pauli: bad term arrays");
  if (n_terms == 0) return SMOQ_OK;
  if (n_terms > 65535) return set_error(SMOQ_EINVAL, "smoq_dm_expect_pauli: too many terms");
  const uint64_t lim = (1ull << N) - 1ull;
  std::vector<DmTerm> tv(n_terms);
  for (int t = 0; t < n_terms; ++t) {
    if ((xmask[t] | zmask[t]) & ~lim)
      return set_error(SMOQ_EINVAL, "smoq_dm_expect_pauli: term %d acts outside the %d qubits", t,
                       N);
    const int c = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
    const double re[4] = {1, 0, -1, 0}, im[4] = {0, 1, 0, -1};
    tv[t] = {xmask[t], zmask[t], re[c], im[c]};
  }
  smoq_ctx* ctx = rho->ctx;
  const size_t tb = ((tv.size() * sizeof(DmTerm)) + 255) & ~(size_t)255;
  const size_t ob = (size_t)rho->batch * n_terms * sizeof(double);
  if ((rc = ensure_scratch(ctx, tb + ob))) return rc;
  DmTerm* dt = (DmTerm*)ctx->scratch;
  double* dout = (double*)((char*)ctx->scratch + tb);
  SMOQ_HIP(hipMemcpyAsync(dt, tv.data(), tv.size() * sizeof(DmTerm), hipMemcpyHostToDevice,
                          ctx->stream));
  const dim3 grid((unsigned)n_terms, (unsigned)rho->batch);
  if (rho->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_dm_expect<double>, grid, dim3(kDmBlock), 0, ctx->stream,
                       (const double2*)rho->d, N, rho->dim, dt, n_terms, dout);
  else
    hipLaunchKernelGGL(k_dm_expect<float>, grid, dim3(kDmBlock), 0, ctx->stream,
                       (const float2*)rho->d, N, rho->dim, dt, n_terms, dout);
  SMOQ_CHECK_LAUNCH("k_dm_expect");
  SMOQ_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

extern "C" int smoq_dm_apply_dense(smoq_state* rho, const smoq_state* u) {
  int rc = check_state(rho);
  if (rc) return rc;
  if ((rc = check_state(u))) return rc;
  int N = 0;
  if ((rc = dm_qubits(rho, "smoq_dm_apply_dense", &N))) return rc;
  if (u->n != rho->n || u->dtype != rho->dtype || u->batch != 1 || u->ctx != rho->ctx ||
      u->d == rho->d)
    return set_error(SMOQ_EINVAL,
                     "smoq_dm_apply_dense: U must be a separate single 2N-qubit state of the "
                     "same dtype and context");
  smoq_ctx* ctx = rho->ctx;
  void* hv = nullptr;
  if ((rc = blas_handle(ctx, "smoq_dm_apply_dense", &hv))) return rc;
  rocblas_handle h = (rocblas_handle)hv;
  const rocblas_int d = (rocblas_int)(1u << N);
  const size_t mb = (size_t)d * d * (rho->dtype == SMOQ_C128 ? 16 : 8);
  void* tmp = nullptr;
  SMOQ_HIP(hipMalloc(&tmp, mb));
  rocblas_status st = rocblas_status_success;
  for (int b = 0; b < rho->batch && st == rocblas_status_success; ++b) {
    char* r = (char*)rho->d + (size_t)b * mb;
    if (rho->dtype == SMOQ_C128) {
      const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
      // tmp = U rho ; rho = tmp U^H
      st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, d, d, d, &one,
                         (const rocblas_double_complex*)u->d, d,
                         (const rocblas_double_complex*)r, d, &zero,
                         (rocblas_double_complex*)tmp, d);
      if (st == rocblas_status_success)
        st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, d, d,
                           d, &one, (const rocblas_double_complex*)tmp, d,
                           (const rocblas_double_complex*)u->d, d, &zero,
                           (rocblas_double_complex*)r, d);
    } else {
      const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
      st = rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, d, d, d, &one,
                         (const rocblas_float_complex*)u->d, d,
                         (const rocblas_float_complex*)r, d, &zero,
                         (rocblas_float_complex*)tmp, d);
      if (st == rocblas_status_success)
        st = rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, d, d,
                           d, &one, (const rocblas_float_complex*)tmp, d,
                           (const rocblas_float_complex*)u->d, d, &zero,
                           (rocblas_float_complex*)r, d);
    }
  }
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  if (st != rocblas_status_success)
    return set_error(SMOQ_EDEVICE, "smoq_dm_apply_dense: rocBLAS gemm failed (%d)", (int)st);
  if (se != hipSuccess) return hip_error(se, "smoq_dm_apply_dense");
  return SMOQ_OK;
}
