// Hamiltonian action (pull form) and Chebyshev-Bessel propagation.
//
// H = sum_t c_t P_t, P_t = i^{|x&z|} X^x Z^z.  Row k of P_t has one entry,
// at column k ^ x: i^{|x&z|} (-1)^{|(k ^ x) & z|}.  Each thread owns output
// amplitudes and writes each of them once, as apply_hamiltonian_bitwise!
// (utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl:55-152) does.
#include <algorithm>
#include <cmath>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

struct TermDev {
  uint64_t x, z;
  double fre, fim;  // c_t * i^{|x&z|}
};

enum HMode : int {
  HM_PLAIN = 0,  // out = H in
  HM_CHEB1 = 1,  // out = (H in - b in)/a ;  acc = J0 in + a1 out
  HM_CHEBK = 2,  // out = 2 (H in - b in)/a - prev ;  acc += ak out   (out may alias prev)
};

struct ChebArgs {
  double a, b;
  double c0;         // J0 (mode 1)
  double cre, cim;   // a1 (mode 1) or ak (mode 2)
};

template <typename T>
__global__ void __launch_bounds__(256)
    k_hpsi(const VT_alias<T>* __restrict__ in, const VT_alias<T>* prev, VT_alias<T>* out,
           VT_alias<T>* __restrict__ acc, uint64_t total, uint64_t dim, int nt,
           const TermDev* __restrict__ terms, int mode, ChebArgs ca) {
#pragma clang fp contract(off)
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
  const uint64_t k = i & (dim - 1);
  const uint64_t base = i - k;
  double hr = 0.0, hi = 0.0;
  // terms come sorted by x: the amplitude at k ^ x is loaded once for all
  // the terms that share x (the diagonal terms share x = 0)
  uint64_t cur = ~0ull;
  double vr = 0.0, vi = 0.0;
  for (int t = 0; t < nt; ++t) {
    const TermDev tm = terms[t];
    const uint64_t j = k ^ tm.x;
    if (tm.x != cur) {
      cur = tm.x;
      const VT_alias<T> v = in[base + j];
      vr = (double)v.x;
      vi = (double)v.y;
    }
    const bool neg = __popcll(j & tm.z) & 1;
    const double sr = neg ? -vr : vr, si = neg ? -vi : vi;
    hr += tm.fre * sr - tm.fim * si;
    hi += tm.fre * si + tm.fim * sr;
  }
  if (mode == HM_PLAIN) {
    out[i] = make_c<T>(hr, hi);
    continue;
  }
  const VT_alias<T> x = in[i];
  const double xr = (double)x.x, xi = (double)x.y;
  // (H x - b x) / a
  double tr = (hr - ca.b * xr) / ca.a, ti = (hi - ca.b * xi) / ca.a;
  if (mode == HM_CHEBK) {
    const VT_alias<T> p = prev[i];
    tr = 2.0 * tr - (double)p.x;
    ti = 2.0 * ti - (double)p.y;
  }
  const VT_alias<T> tv = make_c<T>(tr, ti);
  out[i] = tv;
  // accumulate with the rounded T value, as the reference stores T then adds
  const double ur = (double)tv.x, ui = (double)tv.y;
  const double ar = ca.cre * ur - ca.cim * ui, ai = ca.cre * ui + ca.cim * ur;
  if (mode == HM_CHEB1) {
    acc[i] = make_c<T>(ca.c0 * xr + ar, ca.c0 * xi + ai);
  } else {
    const VT_alias<T> o = acc[i];
    acc[i] = make_c<T>((double)o.x + ar, (double)o.y + ai);
  }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_scale(VT_alias<T>* __restrict__ psi, uint64_t total, double re, double im) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const VT_alias<T> v = psi[i];
    psi[i] = make_c<T>(re * (double)v.x - im * (double)v.y, re * (double)v.y + im * (double)v.x);
  }
}

static int check_terms(const smoq_state* s, int n_terms, const uint64_t* x, const uint64_t* z,
                       const double* c, std::vector<TermDev>& tv) {
  if (n_terms < 0 || (n_terms > 0 && (!x || !z || !c)))
    return set_error(SMOQ_EINVAL, "Pauli sum: bad term arrays");
  const uint64_t lim = s->dim - 1;
  for (int t = 0; t < n_terms; ++t) {
    if ((x[t] | z[t]) & ~lim)
      return set_error(SMOQ_EINVAL, "Pauli sum: term %d acts outside the %d qubits", t, s->n);
    const int cnt = __builtin_popcountll(x[t] & z[t]) & 3;
    // c * i^cnt
    const double re[4] = {1, 0, -1, 0}, im[4] = {0, 1, 0, -1};
    tv.push_back({x[t], z[t], c[t] * re[cnt], c[t] * im[cnt]});
  }
  // terms sharing an x mask next to each other (k_hpsi loads their common
  // amplitude once); the order inside a group is the caller's
  std::stable_sort(tv.begin(), tv.end(),
                   [](const TermDev& a, const TermDev& b) { return a.x < b.x; });
  return SMOQ_OK;
}

template <typename T>
static int hpsi_launch(smoq_state* s, const void* in, const void* prev, void* out, void* acc,
                       const TermDev* dterms, int nt, int mode, ChebArgs ca) {
  const uint64_t total = s->dim * (uint64_t)s->batch;
  hipLaunchKernelGGL(k_hpsi<T>, dim3(stride_blocks(total, 256)), dim3(256), 0,
                     s->ctx->stream, (const VT_alias<T>*)in, (const VT_alias<T>*)prev,
                     (VT_alias<T>*)out, (VT_alias<T>*)acc, total, s->dim, nt, dterms, mode, ca);
  SMOQ_CHECK_LAUNCH("k_hpsi");
  return SMOQ_OK;
}

static int hpsi_any(smoq_state* s, const void* in, const void* prev, void* out, void* acc,
                    const TermDev* dterms, int nt, int mode, ChebArgs ca) {
  return s->dtype == SMOQ_C128 ? hpsi_launch<double>(s, in, prev, out, acc, dterms, nt, mode, ca)
                               : hpsi_launch<float>(s, in, prev, out, acc, dterms, nt, mode, ca);
}

static int upload_terms(smoq_ctx* ctx, const std::vector<TermDev>& tv, TermDev** d) {
  const size_t bytes = tv.size() * sizeof(TermDev) + 64;
  int rc = ensure_scratch(ctx, bytes);
  if (rc) return rc;
  *d = (TermDev*)ctx->scratch;
  if (!tv.empty())
    SMOQ_HIP(hipMemcpyAsync(*d, tv.data(), tv.size() * sizeof(TermDev), hipMemcpyHostToDevice,
                            ctx->stream));
  // tv is host memory owned by the caller's frame: finish the copy now
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

// ---------------------------------------------------------------------------
// Lanczos (estimate_spectral_range_lanczos, …Chebyshev.jl:166-220): w = H v,
// w -= beta_prev v_prev, alpha = Re<v|w>, w -= alpha v, beta = ||w||,
// v_prev = v, v = w / beta.  Block partial sums, added on the host in block
// order (deterministic).
constexpr int kLzBlock = 256;

__device__ inline double lz_block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kLzBlock / 64; ++w) t += sh[w];
  return t;
}

// w -= beta * vp; partial[block] = sum Re(conj(v) w)
template <typename T>
__global__ void __launch_bounds__(kLzBlock)
    k_lz_orth_prev(VT_alias<T>* __restrict__ w, const VT_alias<T>* __restrict__ v,
                   const VT_alias<T>* __restrict__ vp, uint64_t dim, double beta,
                   double* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[kLzBlock / 64];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kLzBlock + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kLzBlock) {
    VT_alias<T> x = w[i];
    double xr = (double)x.x, xi = (double)x.y;
    if (beta != 0.0) {
      const VT_alias<T> p = vp[i];
      x = make_c<T>(xr - beta * (double)p.x, xi - beta * (double)p.y);
      w[i] = x;
      xr = (double)x.x;
      xi = (double)x.y;
    }
    const VT_alias<T> a = v[i];
    acc += (double)a.x * xr + (double)a.y * xi;
  }
  const double t = lz_block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// w -= alpha * v; partial[block] = sum |w|^2
template <typename T>
__global__ void __launch_bounds__(kLzBlock)
    k_lz_orth_cur(VT_alias<T>* __restrict__ w, const VT_alias<T>* __restrict__ v, uint64_t dim,
                  double alpha, double* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[kLzBlock / 64];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kLzBlock + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kLzBlock) {
    const VT_alias<T> x = w[i], a = v[i];
    const VT_alias<T> y = make_c<T>((double)x.x - alpha * (double)a.x,
                                    (double)x.y - alpha * (double)a.y);
    w[i] = y;
    acc += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
  }
  const double t = lz_block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// partial[block] = sum |v|^2
template <typename T>
__global__ void __launch_bounds__(kLzBlock)
    k_lz_norm2(const VT_alias<T>* __restrict__ v, uint64_t dim, double* __restrict__ partial) {
  __shared__ double sh[kLzBlock / 64];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kLzBlock + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kLzBlock) {
    const VT_alias<T> a = v[i];
    acc += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
  }
  const double t = lz_block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// vp = v; v = w * inv
template <typename T>
__global__ void __launch_bounds__(kLzBlock)
    k_lz_next(const VT_alias<T>* __restrict__ w, VT_alias<T>* __restrict__ v,
              VT_alias<T>* __restrict__ vp, uint64_t dim, double inv) {
  for (uint64_t i = (uint64_t)blockIdx.x * kLzBlock + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kLzBlock) {
    vp[i] = v[i];
    const VT_alias<T> x = w[i];
    v[i] = make_c<T>((double)x.x * inv, (double)x.y * inv);
  }
}

// eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm count)
static int sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    const double off = i ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i ? off / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// k-th smallest eigenvalue (0-based) by bisection on the Gershgorin interval
static double tridiag_eig(const std::vector<double>& d, const std::vector<double>& e, int k) {
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < d.size(); ++i) {
    const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < d.size() ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo) + std::fabs(hi));
       ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_count(d, e, mid) > k)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

template <typename T>
static int lanczos_t(const smoq_state* start, const TermDev* dterms, int nt, int k_iter,
                     double* e_min, double* e_max, int* iters) {
  smoq_ctx* ctx = start->ctx;
  const uint64_t dim = start->dim;
  const size_t bytes = dim * sizeof(VT_alias<T>);
  const int nb = (int)std::max<uint64_t>(1, std::min<uint64_t>((dim + kLzBlock - 1) / kLzBlock,
                                                               2048));
  unsigned char* buf = nullptr;
  SMOQ_HIP(hipMalloc(&buf, 3 * bytes + nb * sizeof(double)));
  VT_alias<T>* v = (VT_alias<T>*)buf;
  VT_alias<T>* vp = (VT_alias<T>*)(buf + bytes);
  VT_alias<T>* w = (VT_alias<T>*)(buf + 2 * bytes);
  double* partial = (double*)(buf + 3 * bytes);
  std::vector<double> hp(nb), alphas, betas;
  smoq_state sv = *start;  // view of v for the H·v launches
  sv.d = v;
  int rc = SMOQ_OK;
  auto sum_partial = [&](double& out) -> int {
    SMOQ_HIP(hipMemcpyAsync(hp.data(), partial, nb * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));
    out = 0.0;
    for (int b = 0; b < nb; ++b) out += hp[b];
    return SMOQ_OK;
  };
  do {
    // v = start / ||start||
    SMOQ_HIP(hipMemcpyAsync(v, start->d, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_lz_norm2<T>, dim3(nb), dim3(kLzBlock), 0, ctx->stream, v, dim, partial);
    SMOQ_CHECK_LAUNCH("k_lz_norm2");
    double n2 = 0.0;
    if ((rc = sum_partial(n2))) break;
    if (!(n2 > 0.0) || !std::isfinite(n2)) {
      rc = set_error(SMOQ_EINVAL, "smoq_lanczos_range: start vector has zero or non-finite norm");
      break;
    }
    hipLaunchKernelGGL(k_scale<T>, dim3(stride_blocks(dim, 256)), dim3(256), 0,
                       ctx->stream, v, dim, 1.0 / std::sqrt(n2), 0.0);
    SMOQ_CHECK_LAUNCH("k_scale");
    for (int j = 0; j < k_iter; ++j) {
      if ((rc = hpsi_any(&sv, v, nullptr, w, nullptr, dterms, nt, HM_PLAIN, ChebArgs{}))) break;
      const double bprev = j > 0 ? betas.back() : 0.0;
      hipLaunchKernelGGL(k_lz_orth_prev<T>, dim3(nb), dim3(kLzBlock), 0, ctx->stream, w, v, vp,
                         dim, bprev, partial);
      SMOQ_CHECK_LAUNCH("k_lz_orth_prev");
      double alpha = 0.0;
      if ((rc = sum_partial(alpha))) break;
      alphas.push_back(alpha);
      hipLaunchKernelGGL(k_lz_orth_cur<T>, dim3(nb), dim3(kLzBlock), 0, ctx->stream, w, v, dim,
                         alpha, partial);
      SMOQ_CHECK_LAUNCH("k_lz_orth_cur");
      double b2 = 0.0;
      if ((rc = sum_partial(b2))) break;
      const double beta = std::sqrt(b2);
      if (beta < 1e-12) break;
      betas.push_back(beta);
      hipLaunchKernelGGL(k_lz_next<T>, dim3(nb), dim3(kLzBlock), 0, ctx->stream, w, v, vp, dim,
                         1.0 / beta);
      SMOQ_CHECK_LAUNCH("k_lz_next");
    }
  } while (false);
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  if (rc) return rc;
  if (se != hipSuccess) return hip_error(se, "smoq_lanczos_range");
  // SymTridiagonal(alphas, betas): the off-diagonal is betas[0 .. m-2]
  std::vector<double> e(betas.begin(), betas.begin() + std::max<size_t>(0, alphas.size() - 1));
  *e_min = tridiag_eig(alphas, e, 0);
  *e_max = tridiag_eig(alphas, e, (int)alphas.size() - 1);
  if (iters) *iters = (int)alphas.size();
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_hpsi(const smoq_state* in, smoq_state* out, int n_terms,
                         const uint64_t* xmask, const uint64_t* zmask, const double* coeff) {
  int rc = check_state(in);
  if (rc) return rc;
  if ((rc = check_state(out))) return rc;
  if (in->n != out->n || in->dtype != out->dtype || in->batch != out->batch || in == out ||
      in->d == out->d)
    return set_error(SMOQ_EINVAL, "smoq_hpsi: in/out must be distinct states of one shape");
  std::vector<TermDev> tv;
  if ((rc = check_terms(in, n_terms, xmask, zmask, coeff, tv))) return rc;
  TermDev* d = nullptr;
  if ((rc = upload_terms(out->ctx, tv, &d))) return rc;
  return hpsi_any(out, in->d, nullptr, out->d, nullptr, d, n_terms, HM_PLAIN, ChebArgs{});
}

extern "C" int smoq_chebyshev_evolve(smoq_state* s, int n_terms, const uint64_t* xmask,
                                     const uint64_t* zmask, const double* coeff, double t,
                                     double emin, double emax, double tol, int max_order,
                                     int* orders_used) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!(emax > emin)) return set_error(SMOQ_EINVAL, "smoq_chebyshev_evolve: emax <= emin");
  std::vector<TermDev> tv;
  if ((rc = check_terms(s, n_terms, xmask, zmask, coeff, tv))) return rc;
  smoq_ctx* ctx = s->ctx;
  // get_rescaling_factors (…Chebyshev.jl:232-243), safety margin 0.01
  const double width = emax - emin;
  const double lo = emin - 0.01 * std::fabs(width), hi = emax + 0.01 * std::fabs(width);
  const double a = (hi - lo) / 2.0, b = (hi + lo) / 2.0;
  const double tau = a * t;
  void* A = nullptr;
  void* B = nullptr;
  SMOQ_HIP(hipMalloc(&A, s->bytes));
  hipError_t e = hipMalloc(&B, s->bytes);
  if (e != hipSuccess) {
    (void)hipFree(A);
    return hip_error(e, "hipMalloc(chebyshev)");
  }
  auto done = [&](int code) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(A);
    (void)hipFree(B);
    return code;
  };
  TermDev* dt = nullptr;
  if ((rc = upload_terms(ctx, tv, &dt))) return done(rc);
  // T0 = psi (copy in A); T1 -> B; psi = J0 T0 - 2i J1 T1
  if (hipMemcpyAsync(A, s->d, s->bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
    return done(set_error(SMOQ_EDEVICE, "smoq_chebyshev_evolve: copy failed"));
  ChebArgs ca{a, b, jn(0, tau), 0.0, -2.0 * jn(1, tau)};
  if ((rc = hpsi_any(s, A, nullptr, B, s->d, dt, n_terms, HM_CHEB1, ca))) return done(rc);
  void* prev = A;  // T_{k-2}, overwritten by T_k
  void* cur = B;   // T_{k-1}
  int m_used = 2;
  for (int k = 2; k <= max_order; ++k) {
    const double jk = jn(k, tau);
    // 2 (-i)^k J_k
    const double pre[4] = {1, 0, -1, 0}, pim[4] = {0, -1, 0, 1};
    ChebArgs ck{a, b, 0.0, 2.0 * jk * pre[k & 3], 2.0 * jk * pim[k & 3]};
    if ((rc = hpsi_any(s, cur, prev, prev, s->d, dt, n_terms, HM_CHEBK, ck))) return done(rc);
    m_used = k + 1;
    if (std::fabs(jk) < tol) break;
    std::swap(prev, cur);
  }
  // global phase exp(-i b t)
  const uint64_t total = s->dim * (uint64_t)s->batch;
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_scale<double>, dim3(stride_blocks(total, 256)), dim3(256), 0,
                       ctx->stream, (double2*)s->d, total, std::cos(b * t), -std::sin(b * t));
  else
    hipLaunchKernelGGL(k_scale<float>, dim3(stride_blocks(total, 256)), dim3(256), 0,
                       ctx->stream, (float2*)s->d, total, std::cos(b * t), -std::sin(b * t));
  if (hipGetLastError() != hipSuccess)
    return done(set_error(SMOQ_EDEVICE, "smoq_chebyshev_evolve: k_scale launch failed"));
  if (orders_used) *orders_used = m_used;
  return done(SMOQ_OK);
}

extern "C" int smoq_lanczos_range(const smoq_state* start, int n_terms, const uint64_t* xmask,
                                  const uint64_t* zmask, const double* coeff, int k_iter,
                                  double* e_min, double* e_max, int* iters_used) {
  int rc = check_state(start);
  if (rc) return rc;
  if (start->batch != 1)
    return set_error(SMOQ_EINVAL, "smoq_lanczos_range: one state (batch 1), got %d",
                     start->batch);
  if (k_iter < 1 || !e_min || !e_max)
    return set_error(SMOQ_EINVAL, "smoq_lanczos_range: k_iter %d or null outputs", k_iter);
  std::vector<TermDev> tv;
  if ((rc = check_terms(start, n_terms, xmask, zmask, coeff, tv))) return rc;
  TermDev* d = nullptr;
  if ((rc = upload_terms(start->ctx, tv, &d))) return rc;
  return start->dtype == SMOQ_C128
             ? lanczos_t<double>(start, d, n_terms, k_iter, e_min, e_max, iters_used)
             : lanczos_t<float>(start, d, n_terms, k_iter, e_min, e_max, iters_used);
}
