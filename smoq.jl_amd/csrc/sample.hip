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
// Long vectors never store the 2^N prefixes.  A leaf's prefixes rise
// monotonically inside it, so one coalesced pass sums every leaf (staged
// through LDS), the tree sweeps give the carries, and only each leaf's end
// value E = carry + leaf sum (its last prefix) is kept.  A draw searches
// the leaf ends and recomputes the prefixes of the one leaf it lands in.
// That equals Julia's binary search over the full prefix array whenever
// "prefix >= r" is false-then-true along the array.  Across a leaf boundary
// the prefixes can step down by a rounding difference (the carry into a
// subtree is rounded differently from the running chain before it): a draw
// r inside such a step's interval (first prefix of the leaf, last prefix
// before it] could make the reference's search path-dependent, so those
// draws replay Julia's halving search itself with every probed prefix
// recomputed from its leaf.  The intervals are collected and sorted by lower
// end on the device, with the running maximum of their upper ends, so the
// check per draw is one binary search.
//
// SMOQ_SAMPLE_SEQUENTIAL follows sample_from_probabilities
// (utils/cpu/cpuQuantumStateMeasurements.jl:203-213): a strictly sequential
// running sum, first index with r <= C_i, else the last index.
#include <algorithm>
#include <cmath>
#include <vector>

#include "device_common.h"
#include <hipcub/hipcub.hpp>

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

__global__ void k_up(Level lv, Level child, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) up_at(lv, child, j);
}

__global__ void k_down(Level lv, Level child, uint64_t nslots) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nslots) down_at(lv, child, j);
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
constexpr int kMaxLevels = 28;  // 2^32 elements need 27 levels
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

// ------------------------------------------ sequential prefix, in parallel --
// The reference's running sum s_i = fl(s_{i-1} + p_i) is one dependent
// chain.  It is reproduced bit for bit in chunks of kSeqChunk terms:
//
//  1. every chunk is summed from zero, and an exclusive scan of those sums
//     gives each chunk a guess G of its true start S;
//  2. every chunk runs the chain from its guess to its end E and notes
//     whether any step rounded a tie;
//  3. one workgroup walks the chunks in order.  While the chain stays inside
//     one binade [2^e, 2^(e+1)) its grid spacing u is fixed, and adding p
//     to a value and to the same value shifted by a multiple k u of the
//     grid rounds to results shifted by the same k u, except for a tie when
//     k is odd (round-half-to-even flips with the grid index's parity).  So
//     when G, S, E and E + (S - G) share a binade and no tie meets an odd
//     k, the chunk's true end is E + (S - G), exact; a chunk whose terms are
//     all zero ends where it starts; any other chunk is run again from S
//     (the few that start near or cross a power of two);
//  4. every chunk runs the chain from its exact start, writing s_i.
constexpr uint32_t kSeqChunk = 4096;
constexpr uint64_t kSeqParallelMin = 1ull << 18;  // shorter: one thread
constexpr int kSeqFixThreads = 1024;
constexpr int kSeqFixPer = 8;  // consecutive chunks per thread in a round of the walk
constexpr uint64_t kSeqFixRound = (uint64_t)kSeqFixPer * kSeqFixThreads;
constexpr int kSeqSub = 4;  // terms per sub-chunk when a chunk is run again

template <typename T>
__device__ __forceinline__ double seq_term(const typename CT<T>::V* __restrict__ psi, uint64_t i,
                                           uint64_t diag_step) {
  return diag_step ? (double)psi[i * diag_step].x : prob_of<T>(psi[i]);
}

// The chunk passes: one thread per chunk, kSeqStageBlock chunks per
// workgroup.  Each round stages the next kSeqStageTerms terms of all the
// workgroup's chunks through LDS with coalesced loads (16 lanes read one
// chunk's 16 contiguous terms), then every thread adds its chunk's terms in
// order.  The writing pass puts its prefixes back into the same buffer and
// stores them coalesced.
__device__ __forceinline__ uint64_t dbits(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ int dexp(double x) { return (int)((dbits(x) >> 52) & 0x7ff); }

// a + p, and t |= whether the exact sum lay halfway between grid points
// (TwoSum error, compared with the spacing above and below the result)
__device__ __forceinline__ double seq_add_tie(double a, double p, bool& t) {
#pragma clang fp contract(off)
  const double s = a + p;
  const double bp = s - a;
  const double err = (a - (s - bp)) + (p - bp);
  if (err != 0.0) {
    const double e2 = 2.0 * fabs(err);
    const double up = __longlong_as_double((long long)(dbits(s) + 1)) - s;
    t |= e2 == up || 2.0 * e2 == up;
  }
  return s;
}

constexpr int kSeqStageBlock = 256;
constexpr int kSeqStageTerms = 16;
static_assert(kSeqChunk % kSeqStageTerms == 0, "chunk length");

// the next kSeqStageTerms terms of the workgroup's chunks: each thread's
// share of the raw amplitudes (loads in flight until seq_put)
template <typename T>
__device__ __forceinline__ void seq_fetch(const typename CT<T>::V* __restrict__ psi, uint64_t n,
                                          uint64_t diag_step, uint64_t ch0, uint32_t off,
                                          typename CT<T>::V (&v)[kSeqStageTerms]) {
#pragma unroll
  for (int j = 0; j < kSeqStageTerms; ++j) {
    const uint32_t L = (uint32_t)j * kSeqStageBlock + threadIdx.x;
    const uint64_t i = (ch0 + L / kSeqStageTerms) * kSeqChunk + off + L % kSeqStageTerms;
    if (i < n) v[j] = psi[diag_step ? i * diag_step : i];
  }
}
// ... their terms into LDS (a term past the end is zero)
template <typename T>
__device__ __forceinline__ void seq_put(uint64_t n, uint64_t diag_step, uint64_t ch0, uint32_t off,
                                        const typename CT<T>::V (&v)[kSeqStageTerms],
                                        double (*buf)[kSeqStageTerms + 1]) {
#pragma unroll
  for (int j = 0; j < kSeqStageTerms; ++j) {
    const uint32_t L = (uint32_t)j * kSeqStageBlock + threadIdx.x;
    const uint64_t i = (ch0 + L / kSeqStageTerms) * kSeqChunk + off + L % kSeqStageTerms;
    buf[L / kSeqStageTerms][L % kSeqStageTerms] =
        i < n ? (diag_step ? (double)v[j].x : prob_of<T>(v[j])) : 0.0;
  }
}

// MODE 0: the chain of each chunk from zero -> sums[ch]
// MODE 1: from guess[ch] -> ends[ch], tie[ch] = 1 when a step's exact sum lay
//         halfway between grid points (TwoSum error, compared with the
//         spacing above and below the result)
// MODE 2: from start[ch], writing every prefix to c
template <typename T, int MODE>
__global__ void __launch_bounds__(kSeqStageBlock)
    k_seq_chunks(const typename CT<T>::V* __restrict__ psi, uint64_t n, uint64_t diag_step,
                 uint64_t nch, const double* __restrict__ from, double* __restrict__ out,
                 uint8_t* __restrict__ tie) {
#pragma clang fp contract(off)
  __shared__ double buf[kSeqStageBlock][kSeqStageTerms + 1];
  const uint64_t ch0 = (uint64_t)blockIdx.x * kSeqStageBlock;
  const uint64_t ch = ch0 + threadIdx.x;
  const bool mine = ch < nch;
  double a = MODE == 0 ? 0.0 : (mine ? from[ch] : 0.0);
  bool t = false;
  typename CT<T>::V v[kSeqStageTerms];
  seq_fetch<T>(psi, n, diag_step, ch0, 0, v);
  for (uint32_t off = 0; off < kSeqChunk; off += kSeqStageTerms) {
    seq_put<T>(n, diag_step, ch0, off, v, buf);
    __syncthreads();
    // the next round's loads in flight while this round is added
    if (off + kSeqStageTerms < kSeqChunk)
      seq_fetch<T>(psi, n, diag_step, ch0, off + kSeqStageTerms, v);
    if (mine) {
#pragma unroll
      for (int u = 0; u < kSeqStageTerms; ++u) {
        const double p = buf[threadIdx.x][u];
        const double s = MODE == 1 ? seq_add_tie(a, p, t) : a + p;
        a = s;
        if (MODE == 2) buf[threadIdx.x][u] = s;
      }
    }
    __syncthreads();
    if (MODE == 2) {
#pragma unroll
      for (int j = 0; j < kSeqStageTerms; ++j) {
        const uint32_t L = (uint32_t)j * kSeqStageBlock + threadIdx.x;
        const uint64_t c = ch0 + L / kSeqStageTerms;
        const uint64_t i = c * kSeqChunk + off + L % kSeqStageTerms;
        if (c < nch && i < n) out[i] = buf[L / kSeqStageTerms][L % kSeqStageTerms];
      }
      __syncthreads();
    }
  }
  if (mine && MODE != 2) out[ch] = a;
  if (mine && MODE == 1) tie[ch] = t ? 1 : 0;
}

// Block-wide scans of one value per thread (every thread of the workgroup
// takes part; the wave sums pass through `ws`, kSeqFixThreads / 64 slots).
// Returns the exclusive prefix; *total gets the workgroup's sum.  Two
// barriers; `ws` may be reused after the second.
template <typename V>
__device__ __forceinline__ V seq_block_scan(V v, V* ws, V* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  V incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const V y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) ws[wid] = incl;
  __syncthreads();
  V base = incl - v, tot = V(0);
#pragma unroll
  for (int w = 0; w < kSeqFixThreads / 64; ++w) {
    if (w < wid) base += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  *total = tot;
  return base;
}

// exclusive prefix of the chunk sums (a guess only: any order will do), one
// workgroup over rounds of kSeqFixPer consecutive sums per thread
__global__ void __launch_bounds__(kSeqFixThreads)
    k_seq_guess(const double* __restrict__ sums, uint64_t nch, double* __restrict__ guess) {
  __shared__ double ws[kSeqFixThreads / 64];
  double run = 0.0;
  for (uint64_t k0 = 0; k0 < nch; k0 += kSeqFixRound) {
    const uint64_t kb = k0 + (uint64_t)threadIdx.x * kSeqFixPer;
    double v[kSeqFixPer], loc = 0.0;
#pragma unroll
    for (int j = 0; j < kSeqFixPer; ++j) {
      v[j] = kb + j < nch ? sums[kb + j] : 0.0;
      loc += v[j];
    }
    double tot;
    double a = run + seq_block_scan(loc, ws, &tot);
#pragma unroll
    for (int j = 0; j < kSeqFixPer; ++j) {
      if (kb + j < nch) guess[kb + j] = a;
      a += v[j];
    }
    run += tot;
  }
}


// pass 3: the exact start of every chunk (one workgroup).  Inside a binade
// the walk is integer arithmetic on bit patterns: B(S_{k+1}) = B(S_k) +
// (B(E_k) - B(G_k)) for every chunk that meets the conditions above (0 for a
// chunk of zeros), so the starts over a round of chunks are one prefix sum.
// The first chunk in the round that fails them is run again from its exact
// start, and the next round begins after it.  The rerun applies the same
// rule to the chunk's kSeqFixThreads sub-chunks of kSeqSub terms (staged in
// LDS, one per thread, guesses from a scan of their sums), and adds in
// order, on one thread, only the terms of the sub-chunks that fail it.
__device__ __forceinline__ bool seq_shift_ok(double S, double g, double e, bool tie, bool zero) {
  if (zero || S == g) return true;
  const int eg = dexp(g);
  if (!(S > 0.0) || !(g > 0.0) || !(e > 0.0) || eg == 0 || eg == 0x7ff || dexp(S) != eg ||
      dexp(e) != eg)
    return false;
  const long long kd = (long long)(dbits(S) - dbits(g));
  if (tie && (kd & 1)) return false;
  const uint64_t rb = dbits(e) + (uint64_t)kd;
  return (int)((rb >> 52) & 0x7ff) == eg;
}

// One round of the walk: thread t holds items t*P .. t*P+P-1 (in order) with
// their guesses g, chain ends e and flags (1 tie, 2 zero, 4 present); the
// first present item starts exactly at S.  Returns the first present item
// whose shift is not exact (or kNone), after put(j, start) for every item of
// the thread up to it, and leaves in *s_S the exact start of that item, or
// else the end of the round.  Three barriers.
constexpr uint32_t kSeqNone = 0xffffffffu;
template <int P, typename Put>
__device__ __forceinline__ uint32_t seq_walk_round(double S, const double (&g)[P],
                                                   const double (&e)[P], const uint8_t (&fl)[P],
                                                   long long* ws, unsigned* s_bad, double* s_S,
                                                   Put put) {
  long long d[P], loc = 0;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    d[j] = (fl[j] & 6) == 4 ? (long long)(dbits(e[j]) - dbits(g[j])) : 0ll;
    loc += d[j];
  }
  if (threadIdx.x == 0) *s_bad = kSeqNone;
  long long tot;
  const long long base = seq_block_scan(loc, ws, &tot);  // its barriers order s_bad's reset
  uint64_t bS = dbits(S) + (uint64_t)base;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    if ((fl[j] & 4) &&
        !seq_shift_ok(__longlong_as_double((long long)bS), g[j], e[j], fl[j] & 1, fl[j] & 2)) {
      atomicMin(s_bad, (unsigned)(threadIdx.x * P + j));
      break;
    }
    bS += (uint64_t)d[j];
  }
  __syncthreads();
  const uint32_t bad = *s_bad;
  bS = dbits(S) + (uint64_t)base;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const uint32_t r = threadIdx.x * P + j;
    if (r > bad) break;
    if (fl[j] & 4) put(j, __longlong_as_double((long long)bS));
    if (r == bad) *s_S = __longlong_as_double((long long)bS);
    bS += (uint64_t)d[j];
  }
  if (bad == kSeqNone && threadIdx.x == kSeqFixThreads - 1)
    *s_S = __longlong_as_double((long long)(dbits(S) + (uint64_t)tot));
  __syncthreads();
  return bad;
}

template <typename T>
__global__ void __launch_bounds__(kSeqFixThreads)
    k_seq_fix(const typename CT<T>::V* __restrict__ psi, uint64_t n, uint64_t diag_step,
              uint64_t nch, const double* __restrict__ sums, const double* __restrict__ guess,
              const double* __restrict__ ends, const uint8_t* __restrict__ tie,
              double* __restrict__ start, unsigned long long* __restrict__ reruns) {
#pragma clang fp contract(off)
  static_assert(kSeqChunk == kSeqSub * kSeqFixThreads, "one sub-chunk per thread");
  __shared__ double terms[kSeqChunk];
  __shared__ long long ws[kSeqFixThreads / 64];
  __shared__ double wsd[kSeqFixThreads / 64];
  __shared__ unsigned s_bad;
  __shared__ double s_S;
  const int tid = threadIdx.x;
  if (tid == 0) s_S = 0.0;
  __syncthreads();
  unsigned long long nrerun = 0;
  uint64_t k0 = 0;
  while (k0 < nch) {
    const double S0 = s_S;
    const uint64_t kb = k0 + (uint64_t)tid * kSeqFixPer;
    double g[kSeqFixPer], e[kSeqFixPer];
    uint8_t fl[kSeqFixPer];
#pragma unroll
    for (int j = 0; j < kSeqFixPer; ++j) {
      const uint64_t k = kb + j;
      const bool in = k < nch;
      g[j] = in ? guess[k] : 0.0;
      e[j] = in ? ends[k] : 0.0;
      fl[j] = in ? (uint8_t)(4 | (sums[k] == 0.0 ? 2 : 0) | (tie[k] ? 1 : 0)) : 0;
    }
    const uint32_t bad = seq_walk_round<kSeqFixPer>(
        S0, g, e, fl, ws, &s_bad, &s_S, [&](int j, double v) { start[kb + j] = v; });
    if (bad == kSeqNone) {
      k0 += kSeqFixRound;
      continue;
    }
    // chunk k0 + bad again, from its exact start s_S, in sub-chunks
    const uint64_t c = k0 + bad;
    const uint64_t i0 = c * kSeqChunk;
    for (uint32_t L = tid; L < kSeqChunk; L += kSeqFixThreads)
      terms[L] = i0 + L < n ? seq_term<T>(psi, i0 + L, diag_step) : 0.0;
    __syncthreads();
    const double Sc = s_S;
    double u = 0.0;
#pragma unroll
    for (int q = 0; q < kSeqSub; ++q) u = u + terms[tid * kSeqSub + q];
    double utot;
    const double gs = Sc + seq_block_scan(u, wsd, &utot);
    bool t = false;
    double es = gs;
#pragma unroll
    for (int q = 0; q < kSeqSub; ++q) es = seq_add_tie(es, terms[tid * kSeqSub + q], t);
    uint32_t j0 = 0;
    for (;;) {
      const double gj[1] = {gs}, ej[1] = {es};
      const uint8_t fj[1] = {(uint8_t)((uint32_t)tid >= j0 ? 4 | (u == 0.0 ? 2 : 0) | (t ? 1 : 0)
                                                           : 0)};
      const uint32_t sb =
          seq_walk_round<1>(s_S, gj, ej, fj, ws, &s_bad, &s_S, [](int, double) {});
      if (sb == kSeqNone) break;
      if (tid == 0) {
        double S = s_S;
#pragma unroll
        for (int q = 0; q < kSeqSub; ++q) S = S + terms[sb * kSeqSub + q];
        s_S = S;
      }
      __syncthreads();
      j0 = sb + 1;
      if (j0 >= (uint32_t)kSeqFixThreads) break;
    }
    ++nrerun;
    k0 = c + 1;
  }
  if (tid == 0) *reruns = nrerun;
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

// ------------------------------------------------ leaf-end path (long) --
// Leaves of the split tree sit at depths D-1 and D only (lengths at one
// depth differ by at most one, so a depth mixes leaves of 127 elements with
// nodes of 128 that split into two leaves of 64).  Virtual slot v of depth
// D names leaf (D, v), or, when its parent p = v/2 is itself a leaf,
// (D-1, p) for even v and nothing ("duplicate") for odd v.
struct LeafRef {
  int d;
  uint64_t j;
  bool dup;
};

__device__ inline LeafRef vleaf(const Level& lp, int D, uint64_t v) {
  LeafRef r;
  const uint64_t p = v >> 1;
  if (lp.flag[p] == 1) {
    r.d = D - 1;
    r.j = p;
    r.dup = (v & 1) != 0;
  } else {
    r.d = D;
    r.j = v;
    r.dup = false;
  }
  return r;
}

constexpr int kStageBlock = 256;
// depth D-1 nodes per workgroup; such a node holds at most 254 elements
// (two leaves of at most 127)
constexpr int kStageParents = 16;
constexpr int kStageSlots = 2 * kStageParents;
constexpr int kStageMax = kStageParents * 2 * kLeaf;  // elements staged (4096)
constexpr int kStageLoads = kStageMax / kStageBlock;
__device__ inline int stage_pad(int l) { return l + (l >> 6); }  // spreads the leaf starts over banks

// Every leaf's sum (left to right, as the reference's leaf loop) into its
// level slot and its first element's probability into first[v]: a
// workgroup stages the elements of 16 consecutive depth D-1 nodes through
// LDS with coalesced loads, then one lane per virtual slot adds its leaf.
template <typename T>
__global__ void __launch_bounds__(kStageBlock)
    k_leaf_sums_staged(const typename CT<T>::V* __restrict__ psi, LevelSet ls, int D,
                       uint64_t n_parents, double* __restrict__ first) {
#pragma clang fp contract(off)
  __shared__ double p_s[kStageMax + kStageMax / 64];
  const Level lp = ls.lv[D - 1];
  const uint64_t p0 = (uint64_t)blockIdx.x * kStageParents;
  const uint64_t p_last = min(p0 + kStageParents, n_parents) - 1;
  const uint64_t lo = lp.start[p0];
  const int cnt = (int)(lp.start[p_last] + lp.len[p_last] - lo);
  typename CT<T>::V a[kStageLoads];
#pragma unroll
  for (int it = 0; it < kStageLoads; ++it) {
    const int l = threadIdx.x + it * kStageBlock;
    if (l < cnt) a[it] = psi[lo + l];
  }
#pragma unroll
  for (int it = 0; it < kStageLoads; ++it) {
    const int l = threadIdx.x + it * kStageBlock;
    if (l < cnt) p_s[stage_pad(l)] = prob_of<T>(a[it]);
  }
  __syncthreads();
  if (threadIdx.x >= kStageSlots) return;
  const uint64_t v = 2 * p0 + threadIdx.x;
  if ((v >> 1) > p_last) return;
  const LeafRef lf = vleaf(lp, D, v);
  if (lf.dup) return;
  const Level lv = ls.lv[lf.d];
  const int o = (int)(lv.start[lf.j] - lo);
  const int ln = (int)lv.len[lf.j];
  double sum = p_s[stage_pad(o)];
  first[v] = sum;
  for (int i = 1; i < ln; ++i) sum = sum + p_s[stage_pad(o + i)];
  lv.sum[lf.j] = sum;
}

// E[v] = carry + sum of v's leaf (its last prefix; a duplicate slot repeats
// the leaf before it).  A leaf whose first prefix lies below the previous
// leaf's last one appends the interval (first, previous last] to dips.
__global__ void k_leaf_ends(LevelSet ls, int D, uint64_t nv, const double* __restrict__ first,
                            double* __restrict__ E, double* __restrict__ dips,
                            unsigned long long* __restrict__ n_dips, unsigned long long cap) {
#pragma clang fp contract(off)
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const Level lp = ls.lv[D - 1];
  const LeafRef lf = vleaf(lp, D, v);
  const LeafRef own = lf.dup ? vleaf(lp, D, v - 1) : lf;
  const Level lv = ls.lv[own.d];
  const double c = lv.carry[own.j];
  const double e = c + lv.sum[own.j];
  E[v] = e;
  if (lf.dup || v == 0) return;
  LeafRef pr = vleaf(lp, D, v - 1);
  if (pr.dup) pr = vleaf(lp, D, v - 2);
  const Level lq = ls.lv[pr.d];
  const double e_prev = lq.carry[pr.j] + lq.sum[pr.j];
  const double f = c + first[v];
  if (f < e_prev) {
    const unsigned long long k = atomicAdd(n_dips, 1ull);
    if (k < cap) {
      dips[k] = f;
      dips[cap + k] = e_prev;
    }
  }
}

// The up-sweep of levels top-1 .. 0 and the down-sweep of levels 0 ..
// top-1 (every level of at most 2^top slots) in one workgroup, with the
// root's carry c[1] = v[1] set between them; the per-level kernels handle
// the wider levels below.
template <typename T>
__global__ void __launch_bounds__(1024)
    k_sweep_top(const typename CT<T>::V* __restrict__ psi, LevelSet ls, int top, bool up,
                bool down) {
#pragma clang fp contract(off)
  const uint64_t t = threadIdx.x, B = blockDim.x;
  if (up) {
    for (int d = top - 1; d >= 0; --d) {
      for (uint64_t j = t; j < (1ull << d); j += B) up_at(ls.lv[d], ls.lv[d + 1], j);
      __syncthreads();
    }
  }
  if (!down) return;
  if (t == 0) ls.lv[0].carry[0] = prob_of<T>(psi[0]);
  __syncthreads();
  for (int d = 0; d < top; ++d) {
    for (uint64_t j = t; j < (1ull << d); j += B) down_at(ls.lv[d], ls.lv[d + 1], j);
    __syncthreads();
  }
}

// the prefix c[i] (0-based) of the reference's pairwise cumsum, recomputed
// from the leaf holding element i
template <typename T>
__device__ inline double prefix_at(const typename CT<T>::V* __restrict__ psi, const LevelSet& ls,
                                   double c0, uint64_t i) {
#pragma clang fp contract(off)
  if (i == 0) return c0;
  int d = 0;
  uint64_t j = 0;
  while (ls.lv[d].flag[j] == 0) {
    const uint64_t half = ls.lv[d].len[j] >> 1;
    j = 2 * j + (i >= ls.lv[d].start[j] + half ? 1 : 0);
    ++d;
  }
  const uint64_t st = ls.lv[d].start[j];
  double loc = prob_of<T>(psi[st]);
  for (uint64_t k = st + 1; k <= i; ++k) loc = loc + prob_of<T>(psi[k]);
  return ls.lv[d].carry[j] + loc;
}

// One draw per thread: the leaf-end search and the in-leaf recomputation,
// or, for a draw inside a step interval (ulo: lower ends, ascending; uhi:
// running maximum of the upper ends), Julia's searchsortedfirst loop over
// the recomputed prefixes.
template <typename T>
__global__ void k_search_leaves(const typename CT<T>::V* __restrict__ psi, LevelSet ls, int D,
                                uint64_t n, const double* __restrict__ E, uint64_t nv,
                                const double* __restrict__ ulo, const double* __restrict__ uhi,
                                int n_union, const double* __restrict__ r, int64_t nd,
                                uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const double x = r[d];
  const double c0 = ls.lv[0].carry[0];
  if (n_union > 0) {
    // the last interval whose lower end lies below x: x is inside one iff
    // the running maximum of the upper ends there reaches it
    int a = 0, b = n_union;
    while (a < b) {
      const int m = (a + b) >> 1;
      if (ulo[m] < x)
        a = m + 1;
      else
        b = m;
    }
    if (a > 0 && x <= uhi[a - 1]) {
      uint64_t lo = 1, len = n;
      while (len != 0) {
        const uint64_t half = len >> 1;
        const uint64_t m = lo + half;
        if (prefix_at<T>(psi, ls, c0, m - 1) < x) {
          lo = m + 1;
          len -= half + 1;
        } else {
          len = half;
        }
      }
      out[d] = lo - 1;
      return;
    }
  }
  if (!(c0 < x)) {
    out[d] = 0;
    return;
  }
  uint64_t lo = 0, hi = nv;  // first slot with !(E < x)
  while (lo < hi) {
    const uint64_t m = lo + ((hi - lo) >> 1);
    if (E[m] < x)
      lo = m + 1;
    else
      hi = m;
  }
  if (lo == nv) {
    out[d] = n;
    return;
  }
  const LeafRef lf = vleaf(ls.lv[D - 1], D, lo);
  const Level lv = ls.lv[lf.d];
  const uint64_t st = lv.start[lf.j];
  const uint32_t ln = lv.len[lf.j];
  const double c = lv.carry[lf.j];
  double loc = prob_of<T>(psi[st]);
  uint64_t hit = st + ln - 1;
  if (!(c + loc < x)) {
    hit = st;
  } else {
    uint32_t i = 1;
    bool found = false;
    for (; i + kBatch <= ln && !found; i += kBatch) {
      double pb[kBatch];
#pragma unroll
      for (uint32_t u = 0; u < kBatch; ++u) pb[u] = prob_of<T>(psi[st + i + u]);
#pragma unroll
      for (uint32_t u = 0; u < kBatch; ++u) {
        loc = loc + pb[u];
        if (!found && !(c + loc < x)) {
          hit = st + i + u;
          found = true;
        }
      }
    }
    for (; i < ln && !found; ++i) {
      loc = loc + prob_of<T>(psi[st + i]);
      if (!(c + loc < x)) {
        hit = st + i;
        found = true;
      }
    }
  }
  out[d] = hit;
}

static unsigned nblk(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// step intervals (lo, hi] listed at most
constexpr unsigned long long kDipCap = 1ull << 20;

// The draws' uniforms go to the device through the context's pinned
// staging buffer (one host copy, then a DMA at full link speed); the
// indices come back the same way.
static int stage_draws(smoq_ctx* ctx, const double* r_host, int64_t nd, double* rd) {
  const size_t bytes = (size_t)nd * sizeof(uint64_t);
  int rc = ensure_pinned(ctx, bytes);
  if (rc) return rc;
  // the staging buffer may still feed an earlier copy on the stream
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(ctx->host_pinned, r_host, (size_t)nd * sizeof(double));
  SMOQ_HIP(hipMemcpyAsync(rd, ctx->host_pinned, (size_t)nd * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
  return SMOQ_OK;
}

static int finish_draws(smoq_ctx* ctx, const uint64_t* od, int64_t nd, uint64_t* out_host) {
  SMOQ_HIP(hipMemcpyAsync(ctx->host_pinned, od, (size_t)nd * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out_host, ctx->host_pinned, (size_t)nd * sizeof(uint64_t));
  return SMOQ_OK;
}

template <typename T>
static int sample_t(smoq_state* s, const double* r_host, int64_t nd, int sem, uint64_t* out_host) {
  using V = typename CT<T>::V;
  smoq_ctx* ctx = s->ctx;
  const bool diag = sem == SMOQ_SAMPLE_SEQUENTIAL_DIAG;
  const uint64_t n = diag ? (1ull << (s->n / 2)) : s->dim;
  const uint64_t diag_step = diag ? n + 1 : 0;
  const V* psi = (const V*)s->d;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const bool tree = sem == SMOQ_SAMPLE_CUMSUM_SEARCH && n > 1;
  const bool leaf_ends = tree && n > kBlockCumsumMax;

  // tree geometry: depth D with every slot at depth D a leaf or absent
  int D = 0;
  if (tree) {
    uint64_t hi_len = n - 1;  // the largest node length at a depth
    while (hi_len >= kLeaf) {
      hi_len = hi_len - (hi_len >> 1);
      ++D;
    }
  }
  if (tree && D >= kMaxLevels) return set_error(SMOQ_ENOTSUP, "smoq_sample: tree too deep");
  const uint64_t nv = leaf_ends ? (1ull << D) : 0;
  const size_t per_slot =
      sizeof(uint64_t) + sizeof(uint32_t) + sizeof(uint8_t) + 2 * sizeof(double);
  size_t tree_bytes = 0;
  if (tree)
    for (int d = 0; d <= D; ++d) tree_bytes += ((uint64_t)1 << d) * per_slot + 256 * 5;
  // scratch: r[nd] | out[nd] | c (the whole prefix array: short vectors) |
  // tree levels (short vectors; long ones keep theirs in the tree buffer)
  const size_t off_r = 0;
  const size_t off_o = off_r + up(nd * sizeof(double));
  const size_t off_c = off_o + up(nd * sizeof(uint64_t));
  const size_t off_levels = off_c + up((leaf_ends ? 1 : n) * sizeof(double));
  // the parallel sequential prefix keeps per-chunk records after c
  // (probabilities |ψ_i|² only: the chunk rule needs a monotone chain, and
  // the diagonal of a density matrix may hold small negative values)
  const bool seq_par = !tree && n >= kSeqParallelMin && !diag_step;
  const uint64_t nch_seq = seq_par ? (n + kSeqChunk - 1) / kSeqChunk : 0;
  const size_t seq_bytes = nch_seq ? up(nch_seq * (4 * sizeof(double) + 1) + 16) : 0;
  int rc = ensure_scratch(ctx, off_levels + (leaf_ends ? 0 : tree_bytes) + seq_bytes);
  if (rc) return rc;
  unsigned char* base = (unsigned char*)ctx->scratch;
  double* rd = (double*)(base + off_r);
  uint64_t* od = (uint64_t*)(base + off_o);
  double* c = (double*)(base + off_c);
  if ((rc = stage_draws(ctx, r_host, nd, rd))) return rc;

  if (!tree) {
    if (!seq_par) {
      hipLaunchKernelGGL(k_seq_prefix<T>, dim3(1), dim3(1), 0, ctx->stream, psi, n, c, diag_step);
      SMOQ_CHECK_LAUNCH("k_seq_prefix");
    } else {
      const uint64_t nch = (n + kSeqChunk - 1) / kSeqChunk;
      double* sums = (double*)(base + off_levels);
      double* guess = sums + nch;
      double* ends = guess + nch;
      double* start = ends + nch;
      unsigned long long* reruns = (unsigned long long*)(start + nch);
      uint8_t* tie = (uint8_t*)(reruns + 1);
      const unsigned gb = (unsigned)((nch + kSeqStageBlock - 1) / kSeqStageBlock);
      hipLaunchKernelGGL((k_seq_chunks<T, 0>), dim3(gb), dim3(kSeqStageBlock), 0, ctx->stream,
                         psi, n, diag_step, nch, nullptr, sums, nullptr);
      hipLaunchKernelGGL(k_seq_guess, dim3(1), dim3(kSeqFixThreads), 0, ctx->stream, sums, nch,
                         guess);
      hipLaunchKernelGGL((k_seq_chunks<T, 1>), dim3(gb), dim3(kSeqStageBlock), 0, ctx->stream,
                         psi, n, diag_step, nch, guess, ends, tie);
      hipLaunchKernelGGL(k_seq_fix<T>, dim3(1), dim3(kSeqFixThreads), 0, ctx->stream, psi, n,
                         diag_step, nch, sums, guess, ends, tie, start, reruns);
      hipLaunchKernelGGL((k_seq_chunks<T, 2>), dim3(gb), dim3(kSeqStageBlock), 0, ctx->stream,
                         psi, n, diag_step, nch, start, c, nullptr);
      SMOQ_CHECK_LAUNCH("sequential prefix kernels");
    }
    if (sem != SMOQ_SAMPLE_CUMSUM_SEARCH) {
      hipLaunchKernelGGL(k_search_le, dim3(nblk(nd)), dim3(256), 0, ctx->stream, c, n, rd, nd, od);
    } else {
      hipLaunchKernelGGL(k_search, dim3(nblk(nd)), dim3(256), 0, ctx->stream, c, n, rd, nd, od);
    }
    SMOQ_CHECK_LAUNCH("search kernel");
    return finish_draws(ctx, od, nd, out_host);
  }

  LevelSet ls;
  auto lay_levels = [&](unsigned char* t) {
    for (int d = 0; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      auto take = [&](size_t bytes) {
        unsigned char* p = t;
        t += (bytes + 255) & ~(size_t)255;
        return p;
      };
      ls.lv[d].start = (uint64_t*)take(ns * sizeof(uint64_t));
      ls.lv[d].len = (uint32_t*)take(ns * sizeof(uint32_t));
      ls.lv[d].flag = (uint8_t*)take(ns);
      ls.lv[d].sum = (double*)take(ns * sizeof(double));
      ls.lv[d].carry = (double*)take(ns * sizeof(double));
    }
    return t;
  };
  if (!leaf_ends) {
    lay_levels(base + off_levels);
    hipLaunchKernelGGL(k_cumsum_block<T>, dim3(1), dim3(1024), 0, ctx->stream, psi, n, ls, D, c,
                       rd, nd, od);
    SMOQ_CHECK_LAUNCH("k_cumsum_block");
    return finish_draws(ctx, od, nd, out_host);
  }
  // long vectors: the tree and the leaf-end buffers live in the context's
  // sampling buffer, whose node geometry is kept for the next call of the
  // same length
  // the step intervals: lower ends, upper ends, both sorted by lower end,
  // then the device sort's temporary storage
  size_t sort_bytes = 0, scan_bytes = 0;
  SMOQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const double*)nullptr,
                                              (double*)nullptr, (const double*)nullptr,
                                              (double*)nullptr, (int)kDipCap, 0, 64,
                                              ctx->stream));
  SMOQ_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, scan_bytes, (const double*)nullptr,
                                             (double*)nullptr, hipcub::Max(), (int)kDipCap,
                                             ctx->stream));
  const size_t dip_temp = up(std::max(sort_bytes, scan_bytes));
  const size_t leaf_bytes = up(nv * sizeof(double)) * 2 + up(4 * kDipCap * sizeof(double)) +
                            dip_temp + 256;
  if (ctx->tree_dev_bytes < tree_bytes + leaf_bytes) {
    if (ctx->tree_dev) {
      SMOQ_HIP(hipStreamSynchronize(ctx->stream));
      SMOQ_HIP(hipFree(ctx->tree_dev));
      ctx->tree_dev = nullptr;
      ctx->tree_dev_bytes = 0;
    }
    ctx->tree_n = 0;
    const size_t want = tree_bytes + leaf_bytes + 4096;
    hipError_t e = hipMalloc(&ctx->tree_dev, want);
    if (e != hipSuccess) {
      ctx->tree_dev = nullptr;
      return set_error(SMOQ_ENOMEM, "smoq_sample: %zu bytes for the cumsum tree: %s", want,
                       hipGetErrorString(e));
    }
    ctx->tree_dev_bytes = want;
  }
  unsigned char* tb = (unsigned char*)ctx->tree_dev;
  unsigned char* tl = lay_levels(tb);
  double* first = (double*)tl;
  double* E = (double*)(tl + up(nv * sizeof(double)));
  double* dips = (double*)(tl + 2 * up(nv * sizeof(double)));
  double* dips_sorted = dips + 2 * kDipCap;
  void* dip_tmp = (void*)(tl + 2 * up(nv * sizeof(double)) + up(4 * kDipCap * sizeof(double)));
  unsigned long long* ndips = (unsigned long long*)((unsigned char*)dip_tmp + dip_temp);
  if (ctx->tree_n != n) {
    // root: elements 2..n (1-based) = [1, n) 0-based
    const uint64_t rs = 1;
    const uint32_t rl = (uint32_t)(n - 1);
    const uint8_t rf = rl < kLeaf ? 1 : 0;
    SMOQ_HIP(hipMemcpyAsync(ls.lv[0].start, &rs, sizeof(rs), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(ls.lv[0].len, &rl, sizeof(rl), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(ls.lv[0].flag, &rf, sizeof(rf), hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));  // host scalars above live on the stack
    for (int d = 1; d <= D; ++d) {
      const uint64_t ns = (uint64_t)1 << d;
      hipLaunchKernelGGL(k_tree_level, dim3(nblk(ns)), dim3(256), 0, ctx->stream, ls.lv[d - 1],
                         ls.lv[d], ns);
    }
    SMOQ_CHECK_LAUNCH("k_tree_level");
    ctx->tree_n = n;
  }
  const uint64_t n_parents = nv >> 1;
  hipLaunchKernelGGL(k_leaf_sums_staged<T>,
                     dim3((unsigned)((n_parents + kStageParents - 1) / kStageParents)),
                     dim3(kStageBlock), 0, ctx->stream, psi, ls, D, n_parents, first);
  // levels of at most 2^kTopLevels slots sweep in one workgroup
  constexpr int kTopLevels = 12;
  const int top = std::min(D, kTopLevels);
  for (int d = D - 1; d >= top; --d) {
    const uint64_t ns = (uint64_t)1 << d;
    hipLaunchKernelGGL(k_up, dim3(nblk(ns)), dim3(256), 0, ctx->stream, ls.lv[d], ls.lv[d + 1],
                       ns);
  }
  hipLaunchKernelGGL(k_sweep_top<T>, dim3(1), dim3(1024), 0, ctx->stream, psi, ls, top, true,
                     true);
  for (int d = top; d < D; ++d) {
    const uint64_t ns = (uint64_t)1 << d;
    hipLaunchKernelGGL(k_down, dim3(nblk(ns)), dim3(256), 0, ctx->stream, ls.lv[d], ls.lv[d + 1],
                       ns);
  }
  SMOQ_HIP(hipMemsetAsync(ndips, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_leaf_ends, dim3(nblk(nv)), dim3(256), 0, ctx->stream, ls, D, nv, first, E,
                     dips, ndips, kDipCap);
  SMOQ_CHECK_LAUNCH("pairwise cumsum kernels");
  unsigned long long n_dips = 0;
  SMOQ_HIP(hipMemcpyAsync(&n_dips, ndips, sizeof(n_dips), hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  // the step intervals sorted by lower end with the running maximum of the
  // upper ends: a draw x lies in their union iff, for the last interval
  // whose lower end is below x, that maximum reaches x.  More intervals
  // than the list holds: every draw takes the reference's search itself
  // (one interval covering every value).
  int n_union = 0;
  const double* ulo = dips_sorted;
  const double* uhi = dips_sorted + kDipCap;
  if (n_dips > kDipCap) {
    const double all[2] = {-HUGE_VAL, HUGE_VAL};
    SMOQ_HIP(hipMemcpyAsync(dips_sorted, &all[0], sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream));
    SMOQ_HIP(hipMemcpyAsync(dips_sorted + kDipCap, &all[1], sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    SMOQ_HIP(hipStreamSynchronize(ctx->stream));  // the host values live on the stack
    n_union = 1;
  } else if (n_dips > 0) {
    const int nk = (int)n_dips;
    size_t tb_sort = dip_temp, tb_scan = dip_temp;
    SMOQ_HIP(hipcub::DeviceRadixSort::SortPairs(dip_tmp, tb_sort, (const double*)dips,
                                                dips_sorted, (const double*)(dips + kDipCap),
                                                dips, nk, 0, 64, ctx->stream));
    SMOQ_HIP(hipcub::DeviceScan::InclusiveScan(dip_tmp, tb_scan, (const double*)dips,
                                               dips_sorted + kDipCap, hipcub::Max(), nk,
                                               ctx->stream));
    n_union = nk;
  }
  hipLaunchKernelGGL(k_search_leaves<T>, dim3(nblk(nd)), dim3(256), 0, ctx->stream, psi, ls, D, n,
                     E, nv, ulo, uhi, n_union, rd, nd, od);
  SMOQ_CHECK_LAUNCH("k_search_leaves");
  return finish_draws(ctx, od, nd, out_host);
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
