// Pauli-string expectation values and norms: deterministic two-stage FP64
// reductions (per-block partials, then a fixed-order second stage).
#include <algorithm>
#include <map>
#include <mutex>

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
// One wave per (item, term): lane i adds blocks i, i+64, ... in order, then
// the fixed-order wave sum.
constexpr int kFinishBlock = 256;
__global__ void __launch_bounds__(kFinishBlock)
    k_finish(const double* __restrict__ partial, int nblocks, int nt, double scale,
             double* __restrict__ out, int out_stride, const int* __restrict__ out_col,
             int n_items) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wid >= (int64_t)n_items * nt) return;  // whole waves
  const int item = (int)(wid / nt), t = (int)(wid % nt);
  const double* p = partial + (uint64_t)item * nblocks * nt + t;
  double v = 0.0;
  for (int b = lane; b < nblocks; b += 64) v += p[(uint64_t)b * nt];
  v = wave_total(v);
  if (lane == 0) out[(uint64_t)item * out_stride + out_col[t]] = scale * v;
}

static unsigned finish_blocks(int n_outputs) {
  return (unsigned)(((int64_t)n_outputs * 64 + kFinishBlock - 1) / kFinishBlock);
}

// ---------------------------------------------------------------------------
// Tiled reductions: many Pauli groups per HBM read.
//
// A tile is the 2^k amplitudes base | dep(l), l < 2^k, where dep scatters
// the local index over k tile bit positions (bits 0..3 always among them,
// so loads run over 256-byte segments) and base enumerates the other bits.
// Every group whose x mask lies inside the tile bits is evaluated from LDS:
// the pair (i, i ^ x) is (dep(l), dep(l ^ lx)) with lx the x mask in local
// bits, and the sign (-1)^{|i & z|} splits into the base's parity and the
// parity of l & lz.  One pass reads the state once for all its groups; a
// chain's local and neighbour terms (X, Y, Z, XX, YY, ZZ) need about
// (N - 12) / 7 + 1 passes instead of one launch per x mask.
constexpr int kObsTileBits = 12;
constexpr int kObsLowBits = 4;
constexpr int kObsMaxTerms = 32;  // terms of one pass: one register accumulator each
constexpr int kObsBlockBits = 9;
constexpr int kObsBlock = 1 << kObsBlockBits;  // 2 workgroups (4 waves per SIMD) per CU
constexpr int kObsPairIters = (1 << (kObsTileBits - 1)) / kObsBlock;  // 4: bits 9, 10 of q
constexpr int kObsJ = (1 << kObsTileBits) / kObsBlock;                // 8: bits 9..11 of l

struct ObsTerm {
  uint64_t z;     // global z mask (base parity)
  uint32_t lz;    // z mask in local tile bits
  int32_t kind;   // -1 diagonal; else |x & z| mod 4 as in k_expect_pair
  uint32_t lx;    // x mask in local tile bits (0: diagonal)
  int32_t hl;     // highest local bit of lx
  // the sign (-1)^{|l & lz|} of a thread's amplitude (pair: of its pair's
  // first index) splits into the thread-index part (-1)^{|tid & la|},
  // applied once after the last tile, and the part of its iteration bits
  // (pair: q bits 9, 10 as lb; diagonal: l bits 9..11 as lb)
  uint32_t la;
  uint32_t lb;
};
// Terms sharing an x mask (lx != 0), or diagonal terms (lx == 0) sharing
// the z bits m above the low kObsBlockBits tile bits (planning only).
struct ObsGroup {
  uint32_t lx;
  int32_t hl;
  int32_t t0;
  int32_t nt;
  int32_t nconst;
  uint32_t m;
};
struct ObsPass {
  int k, lowb, n_cbits;
  int group0, n_groups;
  int term0, n_terms;
  uint8_t tpos[kObsTileBits];  // tile bit positions, ascending
  uint8_t cpos[64];            // the other bit positions, ascending
};

__device__ inline int parity64(uint64_t v) { return __popcll(v) & 1; }

// One pass of the tiled reduction.  Every workgroup walks its tiles of one
// batch item; each thread keeps one FP64 accumulator per term of the pass
// in registers (term s in acc[s]; the tile-invariant signs are factored
// out), so a term costs a few operations per tile and one wave sum per
// workgroup at the end.  Terms of one x mask are consecutive: their pair
// products are computed once per tile.
template <typename T>
__global__ void __launch_bounds__(kObsBlock, 4)
    k_expect_tiles(const typename CT<T>::V* __restrict__ psi, uint64_t dim, ObsPass P,
                   const ObsGroup* __restrict__ groups, const ObsTerm* __restrict__ terms,
                   double* __restrict__ partial, int part_terms) {
  using V = typename CT<T>::V;
  (void)groups;
  __shared__ V tile[1 << kObsTileBits];
  __shared__ uint64_t dep_hi[1 << (kObsTileBits - kObsLowBits)];
  __shared__ double acc_s[kObsBlock / 64][kObsMaxTerms];
  // the pass's term descriptors are read from global memory with uniform
  // addresses (scalar loads)
  const ObsTerm* __restrict__ s_terms = terms + P.term0;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint64_t item = blockIdx.y;
  const V* p = psi + item * dim;
  const int k = P.k, lowb = P.lowb, ntile = 1 << k;
  const int nhi = 1 << (k - lowb);
  for (int h = tid; h < nhi; h += kObsBlock) {
    uint64_t off = 0;
    for (int b = 0; b < k - lowb; ++b)
      if ((h >> b) & 1) off |= 1ull << P.tpos[lowb + b];
    dep_hi[h] = off;
  }
  double acc[kObsMaxTerms];
#pragma unroll
  for (int t = 0; t < kObsMaxTerms; ++t) acc[t] = 0.0;
  __syncthreads();
  const int nterms = P.n_terms;
  const uint64_t n_tiles = dim >> k;
  const int lowmask = (1 << lowb) - 1;
  for (uint64_t tt = blockIdx.x; tt < n_tiles; tt += gridDim.x) {
    uint64_t base = 0;
    for (int b = 0; b < P.n_cbits; ++b)
      if ((tt >> b) & 1) base |= 1ull << P.cpos[b];
    // the thread's loads (in two halves: the registers of all eight would
    // not fit beside the accumulators), the LDS stores and the plain sum of
    // its |psi_l|^2, l = tid + kObsBlock j
    double pr0 = 0.0;
#pragma unroll
    for (int h = 0; h < kObsJ; h += kObsJ / 2) {
      V v[kObsJ / 2];
#pragma unroll
      for (int j = 0; j < kObsJ / 2; ++j) {
        const int l = tid + kObsBlock * (h + j);
        if (l < ntile) v[j] = p[base | (uint64_t)(l & lowmask) | dep_hi[l >> lowb]];
      }
#pragma unroll
      for (int j = 0; j < kObsJ / 2; ++j) {
        const int l = tid + kObsBlock * (h + j);
        if (l < ntile) {
          tile[l] = v[j];
          pr0 += (double)v[j].x * (double)v[j].x + (double)v[j].y * (double)v[j].y;
        }
      }
    }
    __syncthreads();
    // per x mask the plain sums of conj(psi_{l ^ lx}) psi_l over the thread's
    // pairs (l = q with a zero inserted at hl, q = tid + kObsBlock it); terms
    // whose sign also depends on the iteration bits (lb != 0, rare) sum
    // their own
    double sr = 0.0, si = 0.0;
    uint32_t cur = 0;
#pragma unroll
    for (int s = 0; s < kObsMaxTerms; ++s) {
      if (s >= nterms) continue;
      const ObsTerm& tm = s_terms[s];
      const uint32_t lx = tm.lx;
      const uint32_t lb = tm.lb;
      const int kind = tm.kind;
      const uint64_t z = tm.z;
      double u = 0.0;
      if (lx == 0) {
        if (lb == 0) {
          u = pr0;
        } else {
#pragma unroll
          for (int j = 0; j < kObsJ; ++j) {
            const int l = tid + kObsBlock * j;
            if (l < ntile) {
              const V a = tile[l];
              const double pj = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
              u += (__popc((uint32_t)j & lb) & 1) ? -pj : pj;
            }
          }
        }
      } else {
        const int hl = tm.hl;
        if (lx != cur || lb != 0) {
          double ar = 0.0, ai = 0.0;
#pragma unroll
          for (int it = 0; it < kObsPairIters; ++it) {
            const int q = tid + it * kObsBlock;
            if (q < (ntile >> 1)) {
              const uint32_t l =
                  ((uint32_t)q >> hl << (hl + 1)) | ((uint32_t)q & ((1u << hl) - 1u));
              const V a = tile[l], b = tile[l ^ lx];
              double wr = (double)b.x * (double)a.x + (double)b.y * (double)a.y;
              double wi = (double)b.x * (double)a.y - (double)b.y * (double)a.x;
              if (__popc((uint32_t)it & lb) & 1) {
                wr = -wr;
                wi = -wi;
              }
              ar += wr;
              ai += wi;
            }
          }
          if (lb == 0) {
            cur = lx;
            sr = ar;
            si = ai;
          }
          u = (kind & 1) ? ai : ar;
        } else {
          u = (kind & 1) ? si : sr;
        }
      }
      const int sb = parity64(base & z) ^ ((kind == 1 || kind == 2) ? 1 : 0);
      acc[s] += sb ? -u : u;
    }
    __syncthreads();
  }
  // the thread-index signs, then one wave sum per term, the waves in order
#pragma unroll
  for (int t = 0; t < kObsMaxTerms; ++t) {
    if (t < nterms) {
      const uint32_t la = s_terms[t].la;
      double v = (__popc((uint32_t)tid & la) & 1) ? -acc[t] : acc[t];
      v = wave_total(v);
      if (lane == 0) acc_s[wid][t] = v;
    }
  }
  __syncthreads();
  for (int t = tid; t < nterms; t += kObsBlock) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kObsBlock / 64; ++w) v += acc_s[w][t];
    // pair terms count each unordered pair once: 2 Re / 2 Im (exact)
    if (s_terms[t].kind >= 0) v *= 2.0;
    partial[((uint64_t)item * gridDim.x + blockIdx.x) * part_terms + P.term0 + t] = v;
  }
}

struct ObsPlan {
  std::vector<ObsPass> passes;
  std::vector<ObsGroup> groups;
  std::vector<ObsTerm> terms;
  std::vector<int> cols;  // pass-ordered term -> caller's term index
  std::vector<int> rest;  // terms left to the per-mask kernels
};

static uint32_t pext64(uint64_t v, const uint8_t* pos, int k) {
  uint32_t r = 0;
  for (int j = 0; j < k; ++j) r |= (uint32_t)((v >> pos[j]) & 1ull) << j;
  return r;
}

// Greedy cover of the distinct x masks (ordered by highest, then lowest
// bit) by tile bit sets of k bits that always hold bits 0..lowb-1, at most
// kObsMaxTerms terms a pass; the diagonal terms fill the passes' spare
// term slots, then passes of their own.
static void plan_obs_tiles(int n, int n_terms, const uint64_t* xm, const uint64_t* zm,
                           ObsPlan& P) {
  const int k = std::min(n, kObsTileBits), lowb = std::min(k, kObsLowBits);
  const uint64_t lowset = (1ull << lowb) - 1ull;
  std::vector<uint64_t> xs;
  std::map<uint64_t, std::vector<int>> by_x;
  for (int t = 0; t < n_terms; ++t) {
    if (!by_x.count(xm[t])) xs.push_back(xm[t]);
    by_x[xm[t]].push_back(t);
  }
  std::vector<uint64_t> pair_xs;
  for (uint64_t x : xs)
    if (x) pair_xs.push_back(x);
  std::stable_sort(pair_xs.begin(), pair_xs.end(), [](uint64_t a, uint64_t b) {
    const int ha = 63 - __builtin_clzll(a), hb = 63 - __builtin_clzll(b);
    if (ha != hb) return ha < hb;
    return __builtin_ctzll(a) < __builtin_ctzll(b);
  });
  struct Draft {
    uint64_t hi = 0;  // tile bits above the low run
    std::vector<uint64_t> xs;
    std::vector<int> diag;
    int n_terms = 0;
  };
  std::vector<Draft> drafts;
  for (uint64_t x : pair_xs) {
    const uint64_t bits = x & ~lowset;
    const int nt = (int)by_x[x].size();
    if (__builtin_popcountll(bits) > k - lowb || nt > kObsMaxTerms) {
      P.rest.insert(P.rest.end(), by_x[x].begin(), by_x[x].end());
      continue;
    }
    if (!drafts.empty() && __builtin_popcountll(drafts.back().hi | bits) <= k - lowb &&
        drafts.back().n_terms + nt <= kObsMaxTerms) {
      Draft& d = drafts.back();
      d.hi |= bits;
      d.xs.push_back(x);
      d.n_terms += nt;
    } else {
      Draft nd;
      nd.hi = bits;
      nd.xs.push_back(x);
      nd.n_terms = nt;
      drafts.push_back(nd);
    }
  }
  if (by_x.count(0)) {
    const std::vector<int>& diag = by_x[0];
    size_t i = 0;
    for (Draft& d : drafts)
      while (i < diag.size() && d.n_terms < kObsMaxTerms) {
        d.diag.push_back(diag[i++]);
        d.n_terms++;
      }
    while (i < diag.size()) {
      Draft nd;
      while (i < diag.size() && nd.n_terms < kObsMaxTerms) {
        nd.diag.push_back(diag[i++]);
        nd.n_terms++;
      }
      drafts.push_back(nd);
    }
  }
  for (Draft& d : drafts) {
    // fill the tile up to k bits with the lowest free positions
    uint64_t hi = d.hi;
    for (int b = lowb; b < n && __builtin_popcountll(hi) < k - lowb; ++b) hi |= 1ull << b;
    ObsPass ps;
    std::memset(&ps, 0, sizeof(ps));
    ps.k = k;
    ps.lowb = lowb;
    int j = 0, c = 0;
    for (int b = 0; b < n; ++b) {
      if (b < lowb || ((hi >> b) & 1))
        ps.tpos[j++] = (uint8_t)b;
      else
        ps.cpos[c++] = (uint8_t)b;
    }
    ps.n_cbits = c;
    ps.group0 = (int)P.groups.size();
    ps.term0 = (int)P.terms.size();
    auto push_term = [&](uint64_t x, int t, int hl) {
      ObsTerm tm;
      std::memset(&tm, 0, sizeof(tm));
      tm.z = zm[t];
      tm.lz = pext64(zm[t], ps.tpos, k);
      tm.kind = x ? (__builtin_popcountll(x & zm[t]) & 3) : -1;
      tm.lx = x ? pext64(x, ps.tpos, k) : 0u;
      tm.hl = hl;
      if (!x) {
        // l = tid + kObsBlock j: thread bits are l's low bits, j its bits 9..11
        tm.la = tm.lz & ((1u << kObsBlockBits) - 1u);
        tm.lb = tm.lz >> kObsBlockBits;
      } else {
        // l = q with a zero at hl, q = tid + kObsBlock it
        for (int i = 0; i < kObsBlockBits + 2; ++i) {
          const int pos = i < hl ? i : i + 1;
          if (!((tm.lz >> pos) & 1)) continue;
          if (i < kObsBlockBits)
            tm.la |= 1u << i;
          else
            tm.lb |= 1u << (i - kObsBlockBits);
        }
      }
      P.terms.push_back(tm);
      P.cols.push_back(t);
    };
    if (!d.diag.empty()) {
      // one group per value of the z bits above the low kObsBlockBits tile bits
      std::vector<uint32_t> ms;
      std::map<uint32_t, std::vector<int>> by_m;
      for (int t : d.diag) {
        const uint32_t m = pext64(zm[t], ps.tpos, k) >> kObsBlockBits;
        if (!by_m.count(m)) ms.push_back(m);
        by_m[m].push_back(t);
      }
      for (uint32_t m : ms) {
        ObsGroup g{};
        g.m = m;
        g.t0 = (int)P.terms.size();
        g.nt = (int)by_m[m].size();
        for (int t : by_m[m]) push_term(0, t, 0);
        P.groups.push_back(g);
      }
    }
    for (uint64_t x : d.xs) {
      ObsGroup g{};
      g.lx = pext64(x, ps.tpos, k);
      g.hl = 31 - __builtin_clz(g.lx);
      g.t0 = (int)P.terms.size();
      g.nt = (int)by_x[x].size();
      for (int t : by_x[x]) push_term(x, t, g.hl);
      P.groups.push_back(g);
    }
    ps.n_groups = (int)P.groups.size() - ps.group0;
    ps.n_terms = (int)P.terms.size() - ps.term0;
    P.passes.push_back(ps);
  }
}

static int red_blocks(uint64_t work) {
  uint64_t b = (work + kRedBlock * 8 - 1) / (kRedBlock * 8);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, 2048));
}

// Device memory for the partial sums of one reduction: the context's
// grow-only scratch, or a fixed region that is never grown (ptr != null).
struct RedScratch {
  void* ptr = nullptr;
  size_t cap = 0;
};

template <typename T>
static int expect_t(smoq_state* s, int n_terms, const uint64_t* xm, const uint64_t* zm,
                    double* out_host, int ref_mode, RedScratch fixed) {
  smoq_ctx* ctx = s->ctx;
  const int B = s->batch;
  if (n_terms > 4096) return set_error(SMOQ_EINVAL, "smoq_expect_pauli: more than 4096 terms");
  // tiled passes when they read the state fewer times than one launch per
  // x mask would (true values only; the reference-rule path stays per mask)
  ObsPlan plan;
  bool tiled = false;
  // compiled passes for this term list (not in the entry point that
  // allocates nothing: loading a code object allocates device memory)
  std::shared_ptr<ObsJitPlan> jplan;
  if (!ref_mode && ctx->obs_tiles) {
    auto launches = [&](const std::vector<int>& ts) {
      std::map<uint64_t, int> per_x;
      for (int t : ts) per_x[xm[t]]++;
      size_t n = 0;
      for (auto& kv : per_x) n += (kv.second + kRedTerms - 1) / kRedTerms;
      return n;
    };
    std::vector<int> all(n_terms);
    for (int t = 0; t < n_terms; ++t) all[t] = t;
    const bool want_jit =
        !fixed.ptr && s->n >= kOjTileBits &&
        (ctx->obs_jit >= 2 ||
         (ctx->obs_jit == 1 && ((uint64_t)s->dim * (uint64_t)B) >= (1ull << kOjMinLogAmps)));
    if (want_jit) {
      std::string key((const char*)&s->n, sizeof(int));
      key.append((const char*)&s->dtype, sizeof(int));
      key.append((const char*)xm, sizeof(uint64_t) * n_terms);
      key.append((const char*)zm, sizeof(uint64_t) * n_terms);
      auto it = ctx->obs_jit_plans.find(key);
      if (it != ctx->obs_jit_plans.end()) {
        jplan = it->second;
      } else {
        auto jp = std::make_shared<ObsJitPlan>();
        plan_obs_jit(s->n, s->dtype, n_terms, xm, zm, *jp);
        if (!jp->passes.empty()) {
          const int rc = jit_get_source(ctx, jp->src, jp->passes.size(), "smoq_obs_", jp->mod);
          if (rc) return rc;
        }
        jp->src.clear();
        jp->src.shrink_to_fit();
        // a bounded table: many distinct term lists would otherwise keep a
        // code object each (in-flight launches hold no reference: the
        // stream is drained first)
        if (ctx->obs_jit_plans.size() >= kObsJitPlansMax) {
          SMOQ_HIP(hipStreamSynchronize(ctx->stream));
          ctx->obs_jit_plans.clear();
        }
        ctx->obs_jit_plans.emplace(key, jp);
        jplan = jp;
      }
      if (jplan->passes.empty() ||
          jplan->passes.size() + launches(jplan->rest) >= launches(all))
        jplan.reset();
    }
    if (jplan) {
      tiled = true;
    } else {
      plan_obs_tiles(s->n, n_terms, xm, zm, plan);
      tiled = !plan.passes.empty() && plan.passes.size() + launches(plan.rest) < launches(all);
    }
  }
  std::vector<int> plain_terms;
  if (jplan) {
    plain_terms = jplan->rest;
  } else if (tiled) {
    plain_terms = plan.rest;
  } else {
    plain_terms.resize(n_terms);
    for (int t = 0; t < n_terms; ++t) plain_terms[t] = t;
  }
  // group the per-mask terms by x mask, keeping first-appearance order
  std::vector<uint64_t> xs;
  std::map<uint64_t, std::vector<int>> groups;
  for (int t : plain_terms) {
    if (!groups.count(xm[t])) xs.push_back(xm[t]);
    groups[xm[t]].push_back(t);
  }
  const int maxblocks = 2048;
  const int tile_terms = jplan ? (int)jplan->cols.size() : tiled ? (int)plan.terms.size() : 0;
  const uint64_t n_tiles =
      jplan ? (s->dim >> kOjTileBits) : tiled ? (s->dim >> plan.passes[0].k) : 0;
  const int tile_wg = tiled ? (int)std::max<uint64_t>(
                                  1, std::min<uint64_t>(n_tiles, (4096 + B - 1) / B))
                            : 0;
  const size_t part_bytes = (size_t)B * maxblocks * kRedTerms * sizeof(double);
  const size_t tpart_bytes = (size_t)B * tile_wg * tile_terms * sizeof(double);
  const size_t out_bytes = (size_t)B * n_terms * sizeof(double);
  const size_t col_bytes = 2 * 4096 * sizeof(int);
  const size_t grp_bytes = plan.groups.size() * sizeof(ObsGroup);
  const size_t trm_bytes = plan.terms.size() * sizeof(ObsTerm);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t need = up(part_bytes) + up(tpart_bytes) + up(out_bytes) + up(col_bytes) +
                      up(grp_bytes) + up(trm_bytes);
  char* sc = nullptr;
  if (fixed.ptr) {
    if (need > fixed.cap)
      return set_error(SMOQ_ENOMEM,
                       "smoq_expect_pauli_noalloc: %zu bytes of partial sums needed (%d items, "
                       "%d terms), the fixed region holds %zu",
                       need, B, n_terms, fixed.cap);
    sc = (char*)fixed.ptr;
  } else {
    int rc = ensure_scratch(ctx, need);
    if (rc) return rc;
    sc = (char*)ctx->scratch;
  }
  double* partial = (double*)sc;
  double* tpartial = (double*)(sc += up(part_bytes));
  double* out_dev = (double*)(sc += up(tpart_bytes));
  int* cols = (int*)(sc += up(out_bytes));
  int* tcols = cols + 4096;
  ObsGroup* grp_dev = (ObsGroup*)(sc += up(col_bytes));
  ObsTerm* trm_dev = (ObsTerm*)(sc += up(grp_bytes));
  using V = typename CT<T>::V;
  if (jplan) {
    SMOQ_HIP(hipMemcpyAsync(tcols, jplan->cols.data(), jplan->cols.size() * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
    const V* psi = (const V*)s->d;
    unsigned long long dim = s->dim, ntl = n_tiles;
    int pt = tile_terms;
    for (size_t p = 0; p < jplan->passes.size(); ++p) {
      int term0 = jplan->passes[p].term0;
      void* args[] = {&psi, &dim, &ntl, &tpartial, &pt, &term0};
      SMOQ_HIP(hipModuleLaunchKernel(jplan->mod->fn[p], (unsigned)tile_wg, (unsigned)B, 1,
                                     kOjThreads, 1, 1, 0, ctx->stream, args, nullptr));
    }
    hipLaunchKernelGGL(k_finish, dim3(finish_blocks(B * tile_terms)), dim3(kFinishBlock), 0,
                       ctx->stream, tpartial, tile_wg, tile_terms, 1.0, out_dev, n_terms, tcols,
                       B);
    SMOQ_CHECK_LAUNCH("k_finish");
  } else if (tiled) {
    SMOQ_HIP(hipMemcpyAsync(grp_dev, plan.groups.data(), grp_bytes, hipMemcpyHostToDevice,
                            ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(trm_dev, plan.terms.data(), trm_bytes, hipMemcpyHostToDevice,
                            ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(tcols, plan.cols.data(), plan.cols.size() * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
    for (const ObsPass& ps : plan.passes) {
      hipLaunchKernelGGL(k_expect_tiles<T>, dim3(tile_wg, B), dim3(kObsBlock), 0, ctx->stream,
                         (const V*)s->d, s->dim, ps, grp_dev, trm_dev, tpartial, tile_terms);
      SMOQ_CHECK_LAUNCH("k_expect_tiles");
    }
    hipLaunchKernelGGL(k_finish, dim3(finish_blocks(B * tile_terms)), dim3(kFinishBlock), 0,
                       ctx->stream, tpartial, tile_wg, tile_terms, 1.0, out_dev, n_terms, tcols,
                       B);
    SMOQ_CHECK_LAUNCH("k_finish");
  }
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
  if (!colh.empty())
    SMOQ_HIP(hipMemcpyAsync(cols, colh.data(), colh.size() * sizeof(int), hipMemcpyHostToDevice,
                            ctx->stream));
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
    hipLaunchKernelGGL(k_finish, dim3(finish_blocks(B * nt)), dim3(kFinishBlock), 0, ctx->stream,
                       partial, nb, nt, (c.x && !ref_mode) ? 2.0 : 1.0, out_dev, n_terms,
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
                        const uint64_t* zmask, double* out, int ref_mode,
                        RedScratch fixed = RedScratch()) {
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
  return s->dtype == SMOQ_C128 ? expect_t<double>(s, n_terms, xmask, zmask, out, ref_mode, fixed)
                               : expect_t<float>(s, n_terms, xmask, zmask, out, ref_mode, fixed);
}

// ------------------------------------------------- no-allocation variant --
// smoq_expect_pauli_noalloc keeps its partial sums in `scratch`, a fixed
// device array of kNoAllocScratchBytes that every device holds from module
// load to process exit, so the call never allocates device memory.  Host
// code only ever uses the per-device address that hipGetSymbolAddress
// returns for it (taken in smoq_init through noalloc_scratch_init); the
// array's own name is a host-side stand-in, not a device address.  Every
// partial-sum slot a reduction reads is written earlier in the same call,
// so the region needs no clearing between calls; calls that overlap on one
// device share the region and race.
namespace smoq {
constexpr size_t kNoAllocScratchBytes = 16ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices)
    return set_error(SMOQ_ENOTSUP, "device %d beyond the %d the library tracks", device,
                     kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return SMOQ_OK;
  void* ptr = nullptr;
  // the caller has made `device` current
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
  if (!ptr) return set_error(SMOQ_EDEVICE, "hipGetSymbolAddress(scratch) returned null");
  g_noalloc_ptr[device] = ptr;
  return SMOQ_OK;
}

static void* noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc_ptr[device];
}
}  // namespace smoq

extern "C" int smoq_expect_pauli_noalloc(smoq_state* s, int n_terms, const uint64_t* xmask,
                                         const uint64_t* zmask, double* out) {
  int rc = check_state(s);
  if (rc) return rc;
  RedScratch fixed;
  fixed.ptr = noalloc_scratch(s->ctx->device);
  fixed.cap = kNoAllocScratchBytes;
  if (!fixed.ptr)
    return set_error(SMOQ_EDEVICE,
                     "smoq_expect_pauli_noalloc: no fixed region on device %d (smoq_init sets "
                     "it up)",
                     s->ctx->device);
  return expect_entry(s, n_terms, xmask, zmask, out, 0, fixed);
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
