// Pauli-string expectation values and norms: deterministic two-stage FP64
// reductions (per-block partials, then a fixed-order second stage).
#include <algorithm>
#include <map>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kRedTerms = 16;   // accumulators per launch
constexpr int kRedBlock = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block-wide sum of `nt` accumulators; result valid in thread 0
template <int NT>
__device__ inline void block_sum(double (&acc)[NT], int nt, double* sh /* [NT][4] */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t >= nt) continue;
    double v = wave_sum(acc[t]);
    if (lane == 0) sh[t * 4 + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int t = 0; t < nt; ++t) {
      double v = 0.0;
      for (int w = 0; w < nw; ++w) v += sh[t * 4 + w];
      acc[t] = v;
    }
  }
}

struct TermSet {
  uint64_t z[kRedTerms];
  int kind[kRedTerms];  // x-part: 0: Re(w) *+1, 1: Im(w) *-1, 2: Re(w) *-1, 3: Im(w) *+1
};

// Diagonal terms: <Z^z> = sum_i |psi_i|^2 (-1)^{|i & z|}
template <typename T>
__global__ void __launch_bounds__(kRedBlock)
    k_expect_diag(const typename CT<T>::V* __restrict__ psi, uint64_t dim, int nt, TermSet ts,
                  double* __restrict__ partial) {
  __shared__ double sh[kRedTerms * 4];
  double acc[kRedTerms];
#pragma unroll
  for (int t = 0; t < kRedTerms; ++t) acc[t] = 0.0;
  const uint64_t item = blockIdx.y;
  const typename CT<T>::V* p = psi + item * dim;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    typename CT<T>::V a = p[i];
    double pr = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
#pragma unroll
    for (int t = 0; t < kRedTerms; ++t)
      if (t < nt) acc[t] += (__popcll(i & ts.z[t]) & 1) ? -pr : pr;
  }
  block_sum<kRedTerms>(acc, nt, sh);
  if (threadIdx.x == 0) {
    double* out = partial + ((uint64_t)item * gridDim.x + blockIdx.x) * nt;
    for (int t = 0; t < nt; ++t) out[t] = acc[t];
  }
}

// Terms sharing x != 0: pairs (i, j = i ^ x) with bit h of i cleared.
// With w = conj(psi_j) psi_i and s_i = (-1)^{|i & z|}:
//   |x & z| = c even -> 2 s_i (-1)^{c/2} Re w
//   c odd            -> 2 s_i (-1)^{(c+1)/2} Im w
template <typename T>
__global__ void __launch_bounds__(kRedBlock)
    k_expect_pair(const typename CT<T>::V* __restrict__ psi, uint64_t dim, uint64_t x, int h,
                  int nt, TermSet ts, double* __restrict__ partial, int ref_mode) {
  __shared__ double sh[kRedTerms * 4];
  double acc[kRedTerms];
#pragma unroll
  for (int t = 0; t < kRedTerms; ++t) acc[t] = 0.0;
  const uint64_t item = blockIdx.y;
  const typename CT<T>::V* p = psi + item * dim;
  const uint64_t half = dim >> 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t0 < half; t0 += stride) {
    const uint64_t i = insert_zero(t0, h);
    // reference mode keeps only bras with every X/Y bit cleared
    if (ref_mode && (i & x)) continue;
    const uint64_t j = i ^ x;
    typename CT<T>::V a = p[i], b = p[j];
    // w = conj(b) * a
    const double wr = (double)b.x * (double)a.x + (double)b.y * (double)a.y;
    const double wi = (double)b.x * (double)a.y - (double)b.y * (double)a.x;
#pragma unroll
    for (int t = 0; t < kRedTerms; ++t) {
      if (t >= nt) continue;
      const int k = ts.kind[t];
      double v = (k & 1) ? wi : wr;
      const bool neg = ((k == 1) || (k == 2)) ^ (bool)(__popcll(i & ts.z[t]) & 1);
      acc[t] += neg ? -v : v;
    }
  }
  block_sum<kRedTerms>(acc, nt, sh);
  if (threadIdx.x == 0) {
    double* out = partial + ((uint64_t)item * gridDim.x + blockIdx.x) * nt;
    for (int t = 0; t < nt; ++t) out[t] = acc[t];
  }
}

// Second stage: one thread per (item, term), fixed summation order.
__global__ void k_finish(const double* __restrict__ partial, int nblocks, int nt, double scale,
                         double* __restrict__ out, int out_stride, const int* __restrict__ out_col,
                         int n_items) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n_items * nt) return;
  const int item = id / nt, t = id % nt;
  const double* p = partial + (uint64_t)item * nblocks * nt + t;
  double v = 0.0;
  for (int b = 0; b < nblocks; ++b) v += p[(uint64_t)b * nt];
  out[(uint64_t)item * out_stride + out_col[t]] = scale * v;
}

static int red_blocks(uint64_t work) {
  uint64_t b = (work + kRedBlock * 8 - 1) / (kRedBlock * 8);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, 2048));
}

template <typename T>
static int expect_t(smoq_state* s, int n_terms, const uint64_t* xm, const uint64_t* zm,
                    double* out_host, int ref_mode) {
  smoq_ctx* ctx = s->ctx;
  const int B = s->batch;
  // group term indices by x mask, keeping first-appearance order
  std::vector<uint64_t> xs;
  std::map<uint64_t, std::vector<int>> groups;
  for (int t = 0; t < n_terms; ++t) {
    if (!groups.count(xm[t])) xs.push_back(xm[t]);
    groups[xm[t]].push_back(t);
  }
  const int maxblocks = 2048;
  const size_t part_bytes = (size_t)B * maxblocks * kRedTerms * sizeof(double);
  const size_t out_bytes = (size_t)B * n_terms * sizeof(double);
  const size_t col_bytes = 4096 * sizeof(int);
  int rc = ensure_scratch(ctx, part_bytes + out_bytes + col_bytes + 256);
  if (rc) return rc;
  double* partial = (double*)ctx->scratch;
  double* out_dev = partial + (size_t)B * maxblocks * kRedTerms;
  int* cols = (int*)(out_dev + (size_t)B * n_terms);
  // the column lists are uploaded per chunk; keep a host mirror
  std::vector<int> colh;
  std::vector<std::pair<size_t, size_t>> chunk_cols;
  struct Chunk {
    uint64_t x;
    std::vector<int> idx;
  };
  std::vector<Chunk> chunks;
  for (uint64_t x : xs) {
    const auto& v = groups[x];
    for (size_t a = 0; a < v.size(); a += kRedTerms) {
      Chunk c;
      c.x = x;
      c.idx.assign(v.begin() + a, v.begin() + std::min(v.size(), a + kRedTerms));
      chunks.push_back(c);
    }
  }
  for (auto& c : chunks) {
    chunk_cols.push_back({colh.size(), c.idx.size()});
    colh.insert(colh.end(), c.idx.begin(), c.idx.end());
  }
  if (colh.size() > 4096) return set_error(SMOQ_EINVAL, "smoq_expect_pauli: more than 4096 terms");
  SMOQ_HIP(hipMemcpyAsync(cols, colh.data(), colh.size() * sizeof(int), hipMemcpyHostToDevice,
                          ctx->stream));
  using V = typename CT<T>::V;
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    const Chunk& c = chunks[ci];
    const int nt = (int)c.idx.size();
    TermSet ts;
    std::memset(&ts, 0, sizeof(ts));
    for (int t = 0; t < nt; ++t) {
      const uint64_t z = zm[c.idx[t]];
      ts.z[t] = z;
      const int cnt = __builtin_popcountll(c.x & z) & 3;
      // i^c: c=0 -> +Re, c=1 -> -Im, c=2 -> -Re, c=3 -> +Im
      ts.kind[t] = cnt;
    }
    const uint64_t work = c.x ? (s->dim >> 1) : s->dim;
    const int nb = red_blocks(work);
    dim3 grid(nb, B);
    if (c.x == 0) {
      hipLaunchKernelGGL(k_expect_diag<T>, grid, dim3(kRedBlock), 0, ctx->stream, (const V*)s->d,
                         s->dim, nt, ts, partial);
    } else {
      const int h = 63 - __builtin_clzll(c.x);
      hipLaunchKernelGGL(k_expect_pair<T>, grid, dim3(kRedBlock), 0, ctx->stream, (const V*)s->d,
                         s->dim, c.x, h, nt, ts, partial, ref_mode);
    }
    SMOQ_CHECK_LAUNCH("expectation kernel");
    const int nthreads = B * nt;
    hipLaunchKernelGGL(k_finish, dim3((nthreads + 255) / 256), dim3(256), 0, ctx->stream, partial,
                       nb, nt, (c.x && !ref_mode) ? 2.0 : 1.0, out_dev, n_terms,
                       cols + chunk_cols[ci].first, B);
    SMOQ_CHECK_LAUNCH("k_finish");
  }
  SMOQ_HIP(hipMemcpyAsync(out_host, out_dev, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

static int expect_entry(smoq_state* s, int n_terms, const uint64_t* xmask,
                        const uint64_t* zmask, double* out, int ref_mode) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n_terms < 0 || (n_terms > 0 && (!xmask || !zmask || !out)))
    return set_error(SMOQ_EINVAL, "smoq_expect_pauli: bad term arrays");
  if (n_terms == 0) return SMOQ_OK;
  const uint64_t lim = s->dim - 1;
  for (int t = 0; t < n_terms; ++t)
    if ((xmask[t] | zmask[t]) & ~lim)
      return set_error(SMOQ_EINVAL, "smoq_expect_pauli: term %d acts outside the %d qubits", t,
                       s->n);
  return s->dtype == SMOQ_C128 ? expect_t<double>(s, n_terms, xmask, zmask, out, ref_mode)
                               : expect_t<float>(s, n_terms, xmask, zmask, out, ref_mode);
}

namespace smoq {
int reduce_pauli(smoq_state* s, int n_terms, const uint64_t* xm, const uint64_t* zm,
                 double* out_host, int ref_mode) {
  return expect_entry(s, n_terms, xm, zm, out_host, ref_mode);
}
}  // namespace smoq

extern "C" int smoq_expect_pauli(smoq_state* s, int n_terms, const uint64_t* xmask,
                                 const uint64_t* zmask, double* out) {
  return expect_entry(s, n_terms, xmask, zmask, out, 0);
}

extern "C" int smoq_pauli_string_reference(smoq_state* s, int n_terms, const uint64_t* xmask,
                                           const uint64_t* zmask, double* out) {
  return expect_entry(s, n_terms, xmask, zmask, out, 1);
}

extern "C" int smoq_norm2(smoq_state* s, double* out) {
  uint64_t x = 0, z = 0;
  return smoq_expect_pauli(s, 1, &x, &z, out);
}
