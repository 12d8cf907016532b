// Born-rule sampling with bit-exact reproduction of the reference's prefix
// sums.
//
// SMOQ_SAMPLE_CUMSUM_SEARCH follows run_circuit_sampling
// (SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:424-429):
//   probs = abs2.(ψ); cumprobs = cumsum(probs); idx = searchsortedfirst(cumprobs, r) - 1.
// Julia's cumsum of a Float64 vector is Base.accumulate_pairwise!: c[1] = v[1]
// and the rest is a recursive halving (left n>>1, right n - n>>1) down to
// leaves of fewer than 128 elements; a leaf with carried prefix s writes
// c[i] = s + (v[i1] + ... + v[i]) with the local sum taken left to right,
// a node returns sum_left + sum_right, and the right child is carried
// s + sum_left.  That tree is evaluated here level by level on the device
// with exactly the same floating-point operations, so every c[i] is bit-equal
// to the reference's.  The search is Julia's searchsortedfirst loop.
//
// SMOQ_SAMPLE_SEQUENTIAL follows sample_from_probabilities
// (utils/cpu/cpuQuantumStateMeasurements.jl:203-213): a strictly sequential
// running sum, first index with r <= C_i, else the last index.
#include <algorithm>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr uint32_t kLeaf = 128;  // leaves hold fewer than 128 elements

// |a|^2 as Julia's abs2: re*re + im*im, two roundings, no fused multiply-add.
template <typename T>
__device__ inline double prob_of(typename CT<T>::V a) {
#pragma clang fp contract(off)
  const double x = (double)a.x, y = (double)a.y;
  const double xx = x * x;
  const double yy = y * y;
  return xx + yy;
}

// Level d of the tree has 2^d slots; flag 0 internal, 1 leaf, 2 absent.
struct Level {
  uint64_t* start;
  uint32_t* len;
  uint8_t* flag;
  double* sum;
  double* carry;
};

// One step of each tree phase for slot j of a level; the per-level kernels
// below and the single-workgroup k_cumsum_block run the same steps.
__device__ inline void tree_level_at(Level parent, Level child, uint64_t j) {
  const uint64_t pj = j >> 1;
  const bool right = j & 1;
  const uint8_t pf = parent.flag[pj];
  if (pf != 0) {
    child.flag[j] = 2;
    child.len[j] = 0;
    child.start[j] = 0;
    return;
  }
  const uint64_t ps = parent.start[pj];
  const uint32_t pl = parent.len[pj];
  const uint32_t half = pl >> 1;
  const uint64_t st = right ? ps + half : ps;
  const uint32_t ln = right ? pl - half : half;
  child.start[j] = st;
  child.len[j] = ln;
  child.flag[j] = ln < kLeaf ? 1 : 0;
}

// Leaves are summed left to right; the loads of each group of kBatch
// elements are issued before its additions, so one leaf costs a few memory
// latencies instead of one per element (the sums are the same).
constexpr uint32_t kBatch = 16;

template <typename T>
__device__ inline void leaf_sum_at(const typename CT<T>::V* __restrict__ psi, Level lv,
                                   uint64_t j) {
#pragma clang fp contract(off)
  if (lv.flag[j] != 1) return;
  const uint64_t st = lv.start[j];
  const uint32_t ln = lv.len[j];
  double s = prob_of<T>(psi[st]);
  uint32_t i = 1;
  for (; i + kBatch <= ln; i += kBatch) {
    double p[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) p[u] = prob_of<T>(psi[st + i + u]);
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) s = s + p[u];
  }
  for (; i < ln; ++i) s = s + prob_of<T>(psi[st + i]);
  lv.sum[j] = s;
}

__device__ inline void up_at(Level lv, Level child, uint64_t j) {
#pragma clang fp contract(off)
  if (lv.flag[j] != 0) return;
  lv.sum[j] = child.sum[2 * j] + child.sum[2 * j + 1];
}

__device__ inline void down_at(Level lv, Level child, uint64_t j) {
#pragma clang fp contract(off)
  if (lv.flag[j] != 0) return;
  const double s = lv.carry[j];
  child.carry[2 * j] = s;
  child.carry[2 * j + 1] = s + child.sum[2 * j];
}

template <typename T>
__device__ inline void leaf_write_at(const typename CT<T>::V* __restrict__ psi, Level lv,
                                     uint64_t j, double* __restrict__ c) {
#pragma clang fp contract(off)
  if (lv.flag[j] != 1) return;
  const uint64_t st = lv.start[j];
  const uint32_t ln = lv.len[j];
  const double s = lv.carry[j];
  double loc = prob_of<T>(psi[st]);
  c[st] = s + loc;
  uint32_t i = 1;
  for (; i + kBatch <= ln; i += kBatch) {
    double p[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) p[u] = prob_of<T>(psi[st + i + u]);
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      loc = loc + p[u];
      c[st + i + u] = s + loc;
    }
  }
  for (; i < ln; ++i) {
    loc = loc + prob_of<T>(psi[st + i]);
    c[st + i] = s + loc;
  }
}

// searchsortedfirst(v, x) over 1-based v[1..n]: Julia Base's halving loop
// (lo, len) with lt = isless.  Returns the 1-based index, n+1 if none.
__device__ inline uint64_t search_first(const double* __restrict__ c, uint64_t n, double x) {
  uint64_t lo = 1, len = n;
  while (len != 0) {
    const uint64_t half = len >> 1;
    const uint64_t m = lo + half;
    if (c[m - 1] < x) {
      lo = m + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo - 1;  // 0-based; equals n when r exceeds every prefix
}

__global__ void k_tree_level(Level parent, Level child, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) tree_level_at(parent, child, j);
}

template <typename T>
__global__ void k_leaf_sum(const typename CT<T>::V* __restrict__ psi, Level lv, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) leaf_sum_at<T>(psi, lv, j);
}

__global__ void k_up(Level lv, Level child, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) up_at(lv, child, j);
}

__global__ void k_down(Level lv, Level child, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) down_at(lv, child, j);
}

template <typename T>
__global__ void k_leaf_write(const typename CT<T>::V* __restrict__ psi, Level lv, uint64_t nslots,
                             double* __restrict__ c) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) leaf_write_at<T>(psi, lv, j, c);
}

template <typename T>
__global__ void k_first(const typename CT<T>::V* __restrict__ psi, double* c, Level root) {
  const double p0 = prob_of<T>(psi[0]);
  c[0] = p0;
  root.carry[0] = p0;
}

__global__ void k_search(const double* __restrict__ c, uint64_t n, const double* __restrict__ r,
                         int64_t nd, uint64_t* __restrict__ out) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d < nd) out[d] = search_first(c, n, r[d]);
}

// The whole pairwise cumsum and the draws in one workgroup (short vectors:
// the per-level launches would cost more than the work).  Barriers separate
// the phases; the level arrays live in global memory, which the workgroup's
// barrier makes visible to all its threads.
constexpr int kMaxLevels = 24;
constexpr uint64_t kBlockCumsumMax = 1ull << 16;
struct LevelSet {
  Level lv[kMaxLevels];
};

template <typename T>
__global__ void __launch_bounds__(1024)
    k_cumsum_block(const typename CT<T>::V* __restrict__ psi, uint64_t n, LevelSet ls, int D,
                   double* __restrict__ c, const double* __restrict__ r, int64_t nd,
                   uint64_t* __restrict__ out) {
  const uint64_t t = threadIdx.x, B = blockDim.x;
  if (t == 0) {  // root: elements 2..n (1-based)
    ls.lv[0].start[0] = 1;
    ls.lv[0].len[0] = (uint32_t)(n - 1);
    ls.lv[0].flag[0] = n - 1 < kLeaf ? 1 : 0;
  }
  __syncthreads();
  for (int d = 1; d <= D; ++d) {
    for (uint64_t j = t; j < (1ull << d); j += B) tree_level_at(ls.lv[d - 1], ls.lv[d], j);
    __syncthreads();
  }
  for (int d = 0; d <= D; ++d)
    for (uint64_t j = t; j < (1ull << d); j += B) leaf_sum_at<T>(psi, ls.lv[d], j);
  __syncthreads();
  for (int d = D - 1; d >= 0; --d) {
    for (uint64_t j = t; j < (1ull << d); j += B) up_at(ls.lv[d], ls.lv[d + 1], j);
    __syncthreads();
  }
  if (t == 0) {
    const double p0 = prob_of<T>(psi[0]);
    c[0] = p0;
    ls.lv[0].carry[0] = p0;
  }
  __syncthreads();
  for (int d = 0; d < D; ++d) {
    for (uint64_t j = t; j < (1ull << d); j += B) down_at(ls.lv[d], ls.lv[d + 1], j);
    __syncthreads();
  }
  for (int d = 0; d <= D; ++d)
    for (uint64_t j = t; j < (1ull << d); j += B) leaf_write_at<T>(psi, ls.lv[d], j, c);
  __syncthreads();
  for (int64_t dd = (int64_t)t; dd < nd; dd += (int64_t)B) out[dd] = search_first(c, n, r[dd]);
}

// Sequential running sum, one thread (exact reproduction of the loop).  With
// diag_step > 0 the terms are real(rho_ii) of vec(rho), element i * diag_step.
template <typename T>
__global__ void k_seq_prefix(const typename CT<T>::V* __restrict__ psi, uint64_t n,
                             double* __restrict__ c, uint64_t diag_step) {
#pragma clang fp contract(off)
  double s = 0.0;
  uint64_t i = 0;
  for (; i + kBatch <= n; i += kBatch) {  // loads of a group ahead of its sums
    double p[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u)
      p[u] = diag_step ? (double)psi[(i + u) * diag_step].x : prob_of<T>(psi[i + u]);
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      s = s + p[u];
      c[i + u] = s;
    }
  }
  for (; i < n; ++i) {
    s = s + (diag_step ? (double)psi[i * diag_step].x : prob_of<T>(psi[i]));
    c[i] = s;
  }
}

// first i with r <= c[i] (c is non-decreasing), else n-1
__global__ void k_search_le(const double* __restrict__ c, uint64_t n, const double* __restrict__ r,
                            int64_t nd, uint64_t* __restrict__ out) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const double x = r[d];
  uint64_t lo = 0, hi = n;  // first index with !(c < x) in [lo, hi)
  while (lo < hi) {
    const uint64_t m = lo + ((hi - lo) >> 1);
    if (c[m] < x)
      lo = m + 1;
    else
      hi = m;
  }
  out[d] = lo < n ? lo : n - 1;
}

static unsigned nblk(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename T>
static int sample_t(smoq_state* s, const double* r_host, int64_t nd, int sem, uint64_t* out_host) {
  using V = typename CT<T>::V;
  smoq_ctx* ctx = s->ctx;
  const bool diag = sem == SMOQ_SAMPLE_SEQUENTIAL_DIAG;
  const uint64_t n = diag ? (1ull << (s->n / 2)) : s->dim;
  const uint64_t diag_step = diag ? n + 1 : 0;
  const V* psi = (const V*)s->d;
  // scratch: c[n] | r[nd] | out[nd] | tree
  size_t off_r = n * sizeof(double);
  size_t off_o = off_r + nd * sizeof(double);
  size_t need = off_o + nd * sizeof(uint64_t);
  need = (need + 255) & ~(size_t)255;

  // tree geometry: depth D with every slot at depth D a leaf or absent
  std::vector<uint64_t> lvl_off;
  int D = 0;
  if (sem == SMOQ_SAMPLE_CUMSUM_SEARCH && n > 1) {
    uint64_t hi_len = n - 1;  // the largest node length at a depth
    while (hi_len >= kLeaf) {
      hi_len = hi_len - (hi_len >> 1);
      ++D;
    }
  }
  const size_t per_slot = sizeof(uint64_t) + sizeof(uint32_t) + sizeof(uint8_t) + 2 * sizeof(double);
  size_t tree_bytes = 0;
  if (sem == SMOQ_SAMPLE_CUMSUM_SEARCH && n > 1) {
    for (int d = 0; d <= D; ++d) tree_bytes += ((uint64_t)1 << d) * per_slot + 256 * 5;
  }
  int rc = ensure_scratch(ctx, need + tree_bytes);
  if (rc) return rc;
  unsigned char* base = (unsigned char*)ctx->scratch;
  double* c = (double*)base;
  double* rd = (double*)(base + off_r);
  uint64_t* od = (uint64_t*)(base + off_o);
  SMOQ_HIP(hipMemcpyAsync(rd, r_host, nd * sizeof(double), hipMemcpyHostToDevice, ctx->stream));

  if (sem != SMOQ_SAMPLE_CUMSUM_SEARCH || n == 1) {
    hipLaunchKernelGGL(k_seq_prefix<T>, dim3(1), dim3(1), 0, ctx->stream, psi, n, c, diag_step);
    SMOQ_CHECK_LAUNCH("k_seq_prefix");
    if (sem != SMOQ_SAMPLE_CUMSUM_SEARCH) {
      hipLaunchKernelGGL(k_search_le, dim3(nblk(nd)), dim3(256), 0, ctx->stream, c, n, rd, nd, od);
    } else {
      hipLaunchKernelGGL(k_search, dim3(nblk(nd)), dim3(256), 0, ctx->stream, c, n, rd, nd, od);
    }
    SMOQ_CHECK_LAUNCH("search kernel");
  } else {
    std::vector<Level> L(D + 1);
    unsigned char* t = base + need;
    for (int d = 0; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      auto take = [&](size_t bytes) {
        unsigned char* p = t;
        t += (bytes + 255) & ~(size_t)255;
        return p;
      };
      L[d].start = (uint64_t*)take(ns * sizeof(uint64_t));
      L[d].len = (uint32_t*)take(ns * sizeof(uint32_t));
      L[d].flag = (uint8_t*)take(ns);
      L[d].sum = (double*)take(ns * sizeof(double));
      L[d].carry = (double*)take(ns * sizeof(double));
    }
    if (n <= kBlockCumsumMax && D < kMaxLevels) {
      LevelSet ls;
      for (int d = 0; d <= D; ++d) ls.lv[d] = L[d];
      hipLaunchKernelGGL(k_cumsum_block<T>, dim3(1), dim3(1024), 0, ctx->stream, psi, n, ls, D,
                         c, rd, nd, od);
      SMOQ_CHECK_LAUNCH("k_cumsum_block");
      SMOQ_HIP(hipMemcpyAsync(out_host, od, nd * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
      SMOQ_HIP(hipStreamSynchronize(ctx->stream));
      return SMOQ_OK;
    }
    // root: elements 2..n (1-based) = [1, n) 0-based
    uint64_t rs = 1;
    uint32_t rl = (uint32_t)(n - 1);
    uint8_t rf = rl < kLeaf ? 1 : 0;
    SMOQ_HIP(hipMemcpyAsync(L[0].start, &rs, sizeof(rs), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(L[0].len, &rl, sizeof(rl), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(L[0].flag, &rf, sizeof(rf), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));  // host scalars above live on the stack
    for (int d = 1; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_tree_level, dim3(nblk(ns)), dim3(256), 0, ctx->stream, L[d - 1], L[d], ns);
    }
    for (int d = 0; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_leaf_sum<T>, dim3(nblk(ns, 64)), dim3(64), 0, ctx->stream, psi, L[d], ns);
    }
    for (int d = D - 1; d >= 0; --d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_up, dim3(nblk(ns)), dim3(256), 0, ctx->stream, L[d], L[d + 1], ns);
    }
    hipLaunchKernelGGL(k_first<T>, dim3(1), dim3(1), 0, ctx->stream, psi, c, L[0]);
    for (int d = 0; d < D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_down, dim3(nblk(ns)), dim3(256), 0, ctx->stream, L[d], L[d + 1], ns);
    }
    for (int d = 0; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_leaf_write<T>, dim3(nblk(ns, 64)), dim3(64), 0, ctx->stream, psi, L[d],
                         ns, c);
    }
    SMOQ_CHECK_LAUNCH("pairwise cumsum kernels");
    hipLaunchKernelGGL(k_search, dim3(nblk(nd)), dim3(256), 0, ctx->stream, c, n, rd, nd, od);
    SMOQ_CHECK_LAUNCH("k_search");
  }
  SMOQ_HIP(hipMemcpyAsync(out_host, od, nd * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_sample(smoq_state* s, const double* uniforms, int64_t n_draws, int semantics,
                           uint64_t* out_index) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n_draws < 0 || (n_draws > 0 && (!uniforms || !out_index)))
    return set_error(SMOQ_EINVAL, "smoq_sample: bad draw arrays");
  if (semantics != SMOQ_SAMPLE_CUMSUM_SEARCH && semantics != SMOQ_SAMPLE_SEQUENTIAL &&
      semantics != SMOQ_SAMPLE_SEQUENTIAL_DIAG)
    return set_error(SMOQ_EINVAL, "smoq_sample: unknown semantics %d", semantics);
  if (semantics == SMOQ_SAMPLE_SEQUENTIAL_DIAG && s->n % 2)
    return set_error(SMOQ_EINVAL, "smoq_sample: a density matrix has an even number of vec "
                                  "qubits, got %d", s->n);
  if (n_draws == 0) return SMOQ_OK;
  if (s->dim > (1ull << 32))
    return set_error(SMOQ_EINVAL, "smoq_sample: more than 2^32 amplitudes");
  return s->dtype == SMOQ_C128 ? sample_t<double>(s, uniforms, n_draws, semantics, out_index)
                               : sample_t<float>(s, uniforms, n_draws, semantics, out_index);
}
