// Reduced density matrices: rho_A = Tr_B |psi><psi| and Tr_B rho
// (utils/cpu/cpuQuantumStatePartialTrace.jl:68-188).
//
// With the kept qubits K (mask keep) and the traced ones T, write
// M[i, t] = psi[pdep(i, keep) | pdep(t, trace)]; then rho_A = M M^H, a
// D x D result (D = 2^|K|) of an inner dimension 2^|T|.
//
//  * D <= 4 (k_ptrace_pairs): each thread loads its columns' D amplitudes and
//    keeps the D^2 products in registers.
//  * D = 8..32 (k_ptrace_tile): each block streams tiles of 2048 amplitudes
//    (all D rows x 2048/D columns of M) through LDS with coalesced loads, and
//    every thread accumulates a register block of rho_A over the tile's
//    columns; block partials are then summed in a fixed order.  Reading psi
//    once bounds it.
//  * D >= 64: M is psi itself when the kept qubits are the lowest ones, else
//    a gathered copy; rocBLAS multiplies M M^H in column chunks (split-K:
//    the inner dimension is long and D x D short), and the chunk products are
//    summed in a fixed order.
//
// The output is vec(rho_A): element (i, j) at i + D j, i and j over the kept
// qubits in increasing order, as the reference's keep_indices.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kPtBlock = 256;
constexpr int kPtTile = 2048;                   // amplitudes per LDS tile (32 KiB)
constexpr int kPtLoads = kPtTile / kPtBlock;   // loads in flight per thread

// Tile geometry: element e (11 bits) of a tile sits at address
// base | off_lo[e & 63] | off_hi[e >> 6] and at LDS index
// lds_lo[e & 63] + lds_hi[e >> 6] = tl * DP + i (DP: the padded column).
struct PtTables {
  uint64_t off_lo[64], off_hi[64];
  uint32_t lds_lo[64], lds_hi[64];
  uint64_t trace_hi;  // trace bits above the tile's column bits
  int tile_elems;     // D * TC
  int tc;             // columns per tile
};

// Register blocks of RB x RB entries of rho_A per thread (rows ti + TI r,
// columns tj + TI c, so a wave's LDS reads are consecutive), G groups of
// threads splitting a tile's columns.
template <int D>
struct PtShape {
  static constexpr int RB = D < 4 ? D : 4;
  static constexpr int TI = D / RB;
  static constexpr int G = kPtBlock / (TI * TI);
  // LDS column stride: a wave holds 64 / TI^2 column groups, each reading TI
  // consecutive rows; the stride spreads the groups over distinct banks (16
  // of 16 B per 256 B) — D = 8 cannot avoid a two-way overlap
  static constexpr int DP = D == 16 ? 20 : (D == 8 ? 9 : D);
};

static int pt_stride(uint64_t D) { return D == 16 ? 20 : (D == 8 ? 9 : (int)D); }

#ifndef SMOQ_PT_GEMM_MIN
#define SMOQ_PT_GEMM_MIN 6  // kept qubits from which the GEMM path runs
#endif

#ifndef SMOQ_PT_MINBLOCKS
#define SMOQ_PT_MINBLOCKS 3  // blocks per CU the tile kernel's registers must allow
#endif

template <typename T, int K>
__global__ void __launch_bounds__(kPtBlock, SMOQ_PT_MINBLOCKS)
    k_ptrace_tile(const VT_alias<T>* __restrict__ psi, int n, PtTables tb, uint64_t tiles_per_blk,
                  double2* __restrict__ part) {
  constexpr int D = 1 << K;
  constexpr int RB = PtShape<D>::RB, TI = PtShape<D>::TI, G = PtShape<D>::G;
  constexpr int DP = PtShape<D>::DP;
  __shared__ double2 L[kPtTile / D * DP];
  // the geometry tables, read per element: from LDS rather than the kernel
  // arguments (per-lane indexed argument reads are global loads that each
  // amplitude load would wait for)
  __shared__ uint64_t s_off[128];
  __shared__ uint32_t s_lds[128];
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const VT_alias<T>* p = psi + (item << n);
  const int tid = threadIdx.x;
  if (tid < 64) {
    s_off[tid] = tb.off_lo[tid];
    s_off[64 + tid] = tb.off_hi[tid];
    s_lds[tid] = tb.lds_lo[tid];
    s_lds[64 + tid] = tb.lds_hi[tid];
  }
  __syncthreads();
  const int g = tid / (TI * TI), ti = (tid / TI) % TI, tj = tid % TI;
  const int tc = tb.tc, E = tb.tile_elems;
  double2 acc[RB][RB];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < RB; ++c) acc[r][c] = make_double2(0.0, 0.0);
  for (uint64_t s = 0; s < tiles_per_blk; ++s) {
    const uint64_t base = pdep64(blk * tiles_per_blk + s, tb.trace_hi);
    // all of the thread's loads in flight before the first LDS write
    VT_alias<T> v[kPtLoads];
#pragma unroll
    for (int r = 0; r < kPtLoads; ++r) {
      const int e = tid + kPtBlock * r;
      if (e < E) v[r] = p[base | s_off[e & 63] | s_off[64 + (e >> 6)]];
    }
    __syncthreads();  // the previous tile's reads are done
#pragma unroll
    for (int r = 0; r < kPtLoads; ++r) {
      const int e = tid + kPtBlock * r;
      if (e < E)
        L[s_lds[e & 63] + s_lds[64 + (e >> 6)]] = make_double2((double)v[r].x, (double)v[r].y);
    }
    __syncthreads();
    for (int tl = g; tl < tc; tl += G) {
      double2 a[RB], b[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) a[r] = L[tl * DP + ti + TI * r];
#pragma unroll
      for (int c = 0; c < RB; ++c) b[c] = L[tl * DP + tj + TI * c];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < RB; ++c) {
          // a * conj(b)
          acc[r][c].x = fma(a[r].x, b[c].x, fma(a[r].y, b[c].y, acc[r][c].x));
          acc[r][c].y = fma(a[r].y, b[c].x, fma(-a[r].x, b[c].y, acc[r][c].y));
        }
    }
  }
  double2* out = part + (item * gridDim.x + blk) * (uint64_t)(D * D);
  if (G == 1) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int c = 0; c < RB; ++c) out[(ti + TI * r) + D * (tj + TI * c)] = acc[r][c];
    return;
  }
  // the column groups' sums, one entry of the register block per round, in
  // group order
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < RB; ++c) {
      __syncthreads();
      L[tid] = acc[r][c];
      __syncthreads();
      if (g == 0) {
        double2 w = L[tid];
        for (int h = 1; h < G; ++h) {
          const double2 x = L[h * TI * TI + tid];
          w.x += x.x;
          w.y += x.y;
        }
        out[(ti + TI * r) + D * (tj + TI * c)] = w;
      }
    }
}

__device__ inline double block_sum_ptrace(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kPtBlock / 64; ++w) t += sh[w];
  return t;
}

// K <= 2 kept qubits (lo < hi; hi unused for K = 1): every thread loads its
// columns' D amplitudes itself (consecutive threads, consecutive columns)
// and keeps the D^2 products in registers; a block sum per entry at the end.
template <typename T, int K>
__global__ void __launch_bounds__(kPtBlock)
    k_ptrace_pairs(const VT_alias<T>* __restrict__ psi, int n, int lo, int hi,
                   double2* __restrict__ part) {
  constexpr int D = 1 << K;
  __shared__ double sh[kPtBlock / 64];
  const uint64_t item = blockIdx.y;
  const VT_alias<T>* p = psi + (item << n);
  const uint64_t cols = 1ull << (n - K);
  double2 acc[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) acc[i][j] = make_double2(0.0, 0.0);
  const uint64_t mlo = 1ull << lo, mhi = K > 1 ? (1ull << hi) : 0;
  for (uint64_t t = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x; t < cols;
       t += (uint64_t)gridDim.x * kPtBlock) {
    const uint64_t a0 = K > 1 ? insert_zero(insert_zero(t, lo), hi) : insert_zero(t, lo);
    double2 x[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const VT_alias<T> v = p[a0 | ((i & 1) ? mlo : 0) | ((i & 2) ? mhi : 0)];
      x[i] = make_double2((double)v.x, (double)v.y);
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        acc[i][j].x = fma(x[i].x, x[j].x, fma(x[i].y, x[j].y, acc[i][j].x));
        acc[i][j].y = fma(x[i].y, x[j].x, fma(-x[i].x, x[j].y, acc[i][j].y));
      }
  }
  double2* out = part + (item * gridDim.x + blockIdx.x) * (uint64_t)(D * D);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const double re = block_sum_ptrace(acc[i][j].x, sh);
      const double im = block_sum_ptrace(acc[i][j].y, sh);
      if (threadIdx.x == 0) out[i + D * j] = make_double2(re, im);
    }
}

// out[item][e] = sum over the nb partials in order, e < D^2
template <typename T, typename P>
__global__ void __launch_bounds__(256)
    k_ptrace_finish(const P* __restrict__ part, int nb, uint64_t dd, uint64_t part_item_stride,
                    VT_alias<T>* __restrict__ out, uint64_t total) {
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t item = o / dd, e = o % dd;
    const P* q = part + item * part_item_stride + e;
    double x = 0.0, y = 0.0;
    for (int b = 0; b < nb; ++b) {
      x += (double)q[(uint64_t)b * dd].x;
      y += (double)q[(uint64_t)b * dd].y;
    }
    out[o] = make_c<T>(x, y);
  }
}

// The same sum with one block per output entry (few entries, many partials):
// each thread sums a strided share in order, then the block sums the threads'
// results in a fixed order.
template <typename T, typename P>
__global__ void __launch_bounds__(256)
    k_ptrace_finish_wide(const P* __restrict__ part, int nb, uint64_t dd,
                         uint64_t part_item_stride, VT_alias<T>* __restrict__ out) {
  __shared__ double sh[kPtBlock / 64];
  const uint64_t o = blockIdx.x, item = o / dd, e = o % dd;
  const P* q = part + item * part_item_stride + e;
  double x = 0.0, y = 0.0;
  for (int b = threadIdx.x; b < nb; b += kPtBlock) {
    x += (double)q[(uint64_t)b * dd].x;
    y += (double)q[(uint64_t)b * dd].y;
  }
  x = block_sum_ptrace(x, sh);
  y = block_sum_ptrace(y, sh);
  if (threadIdx.x == 0) out[o] = make_c<T>(x, y);
}

// partials -> out: a block per entry when entries are few, a thread per
// entry otherwise
template <typename T, typename P>
static int ptrace_finish(hipStream_t st, const P* part, int nb, uint64_t dd, uint64_t stride,
                         VT_alias<T>* out, uint64_t total) {
  if (total <= 4096 && nb >= 64)
    hipLaunchKernelGGL((k_ptrace_finish_wide<T, P>), dim3((unsigned)total), dim3(kPtBlock), 0, st,
                       part, nb, dd, stride, out);
  else
    hipLaunchKernelGGL((k_ptrace_finish<T, P>), dim3(stride_blocks(total, 256)), dim3(256), 0, st,
                       part, nb, dd, stride, out, total);
  SMOQ_CHECK_LAUNCH("k_ptrace_finish");
  return SMOQ_OK;
}

// M[i + D t] = psi[pdep(i, keep) | pdep(t, trace)] for one item; the bit
// deposits come from byte tables in LDS (i: 3 bytes, t: 4 bytes; the state
// has at most 36 qubits and the output at most 2^34 entries, so i < 2^17 and
// t < 2^31)
template <typename T>
__global__ void __launch_bounds__(256)
    k_ptrace_gather(const VT_alias<T>* __restrict__ psi, uint64_t keep, uint64_t trace, int k,
                    uint64_t len, VT_alias<T>* __restrict__ m) {
  __shared__ uint64_t tab[7][256];
  for (int e = threadIdx.x; e < 7 * 256; e += 256) {
    const int b = e >> 8, v = e & 255;
    tab[b][v] = b < 3 ? pdep64((uint64_t)v << (8 * b), keep)
                      : pdep64((uint64_t)v << (8 * (b - 3)), trace);
  }
  __syncthreads();
  const uint64_t dmask = (1ull << k) - 1;
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < len;
       o += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = o & dmask, t = o >> k;
    m[o] = psi[tab[0][i & 255] | tab[1][(i >> 8) & 255] | tab[2][(i >> 16) & 255] |
               tab[3][t & 255] | tab[4][(t >> 8) & 255] | tab[5][(t >> 16) & 255] |
               tab[6][(t >> 24) & 255]];
  }
}

// rho_A[i, j] = sum_t rho[i|t, j|t] over a column chunk of t per blockIdx.y,
// for one item
template <typename T>
__global__ void __launch_bounds__(256)
    k_ptrace_rho(const VT_alias<T>* __restrict__ p, int N, uint64_t keep, uint64_t trace,
                 int k, uint64_t t_per_chunk, double2* __restrict__ part) {
  const uint64_t D = 1ull << k, dd = D * D;
  const uint64_t chunk = blockIdx.y;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < dd;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = pdep64(e & (D - 1), keep), j = pdep64(e >> k, keep);
    double x = 0.0, y = 0.0;
    for (uint64_t t = chunk * t_per_chunk; t < (chunk + 1) * t_per_chunk; ++t) {
      const uint64_t pt = pdep64(t, trace);
      const VT_alias<T> v = p[(i | pt) + ((j | pt) << N)];
      x += (double)v.x;
      y += (double)v.y;
    }
    part[chunk * dd + e] = make_double2(x, y);
  }
}

static uint64_t pdep_host(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  while (mask) {
    const uint64_t low = mask & (0ull - mask);
    if (v & 1ull) r |= low;
    v >>= 1;
    mask ^= low;
  }
  return r;
}

static int ilog2(uint64_t v) { return 63 - __builtin_clzll(v); }

template <typename T>
static int ptrace_tile(const smoq_state* s, uint64_t keep, int k, smoq_state* out) {
  smoq_ctx* ctx = s->ctx;
  const uint64_t full = s->dim - 1, trace = full & ~keep;
  const int tbits = s->n - k;
  const uint64_t D = 1ull << k;
  const uint64_t tc = std::min<uint64_t>(kPtTile / D, 1ull << tbits);
  const int cbits = ilog2(tc);
  // the tile: kept bits and the lowest cbits traced bits
  uint64_t low_trace = 0, rest = trace;
  for (int b = 0; b < cbits; ++b) {
    const uint64_t lb = rest & (0ull - rest);
    low_trace |= lb;
    rest ^= lb;
  }
  const uint64_t tile_mask = keep | low_trace;
  PtTables tb;
  tb.trace_hi = rest;
  tb.tc = (int)tc;
  tb.tile_elems = (int)(D * tc);
  // LDS index contribution of each tile bit, in address order
  std::vector<uint32_t> bit_lds;
  {
    int rk = 0, rt = 0;
    for (int b = 0; b < s->n; ++b) {
      const uint64_t m = 1ull << b;
      if (keep & m)
        bit_lds.push_back(1u << rk++);
      else if (low_trace & m)
        bit_lds.push_back((uint32_t)pt_stride(D) << rt++);
    }
  }
  const int ebits = (int)bit_lds.size();  // k + cbits <= 11
  for (int j = 0; j < 64; ++j) {
    tb.off_lo[j] = pdep_host((uint64_t)j, tile_mask);
    tb.off_hi[j] = pdep_host((uint64_t)j << 6, tile_mask);
    uint32_t lo = 0, hi = 0;
    for (int b = 0; b < 6; ++b) {
      if (((j >> b) & 1) && b < ebits) lo += bit_lds[b];
      if (((j >> b) & 1) && b + 6 < ebits) hi += bit_lds[b + 6];
    }
    tb.lds_lo[j] = lo;
    tb.lds_hi[j] = hi;
  }
  const uint64_t tiles = 1ull << (tbits - cbits);
  const size_t B = (size_t)s->batch;
  uint64_t nb = 1;
  while (nb * 2 <= tiles && nb * 2 * B <= 2048) nb *= 2;
  const uint64_t per = tiles / nb;
  const uint64_t dd = D * D;
  int rc = ensure_scratch(ctx, B * nb * dd * sizeof(double2));
  if (rc) return rc;
  double2* part = (double2*)ctx->scratch;
  const dim3 grid((unsigned)nb, (unsigned)B);
  using V = VT_alias<T>;
  const V* psi = (const V*)s->d;
  switch (k) {
    case 3: hipLaunchKernelGGL((k_ptrace_tile<T, 3>), grid, dim3(kPtBlock), 0, ctx->stream, psi, s->n, tb, per, part); break;
    case 4: hipLaunchKernelGGL((k_ptrace_tile<T, 4>), grid, dim3(kPtBlock), 0, ctx->stream, psi, s->n, tb, per, part); break;
    default: hipLaunchKernelGGL((k_ptrace_tile<T, 5>), grid, dim3(kPtBlock), 0, ctx->stream, psi, s->n, tb, per, part); break;
  }
  SMOQ_CHECK_LAUNCH("k_ptrace_tile");
  const uint64_t total = B * dd;
  if ((rc = ptrace_finish<T, double2>(ctx->stream, part, (int)nb, dd, nb * dd, (V*)out->d, total)))
    return rc;
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

template <typename T>
static int ptrace_pairs(const smoq_state* s, uint64_t keep, int k, smoq_state* out) {
  smoq_ctx* ctx = s->ctx;
  const int lo = __builtin_ctzll(keep), hi = 63 - __builtin_clzll(keep);
  const uint64_t D = 1ull << k, dd = D * D, cols = s->dim >> k;
  const size_t B = (size_t)s->batch;
  // blocks: about 8 columns per thread at least, 2048 blocks at most
  uint64_t nb = std::max<uint64_t>(1, std::min<uint64_t>(cols / (8 * kPtBlock), 2048 / B));
  int rc = ensure_scratch(ctx, B * nb * dd * sizeof(double2));
  if (rc) return rc;
  double2* part = (double2*)ctx->scratch;
  using V = VT_alias<T>;
  const dim3 grid((unsigned)nb, (unsigned)B);
  if (k == 1)
    hipLaunchKernelGGL((k_ptrace_pairs<T, 1>), grid, dim3(kPtBlock), 0, ctx->stream,
                       (const V*)s->d, s->n, lo, lo, part);
  else
    hipLaunchKernelGGL((k_ptrace_pairs<T, 2>), grid, dim3(kPtBlock), 0, ctx->stream,
                       (const V*)s->d, s->n, lo, hi, part);
  SMOQ_CHECK_LAUNCH("k_ptrace_pairs");
  const uint64_t total = B * dd;
  if ((rc = ptrace_finish<T, double2>(ctx->stream, part, (int)nb, dd, nb * dd, (V*)out->d, total)))
    return rc;
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

template <typename T>
static int ptrace_gemm(const smoq_state* s, uint64_t keep, int k, smoq_state* out) {
  using V = VT_alias<T>;
  smoq_ctx* ctx = s->ctx;
  void* hv = nullptr;
  int rc = blas_handle(ctx, "smoq_partial_trace", &hv);
  if (rc) return rc;
  rocblas_handle h = (rocblas_handle)hv;
  const uint64_t full = s->dim - 1, trace = full & ~keep;
  const uint64_t D = 1ull << k, dt = 1ull << (s->n - k), dd = D * D;
  const bool direct = keep == D - 1;  // kept qubits are the lowest: M is psi
  // column chunks: enough products of D x D to fill the chip, 1 GiB of them at most
  const uint64_t tiles = std::max<uint64_t>(1, (D / 64) * (D / 64));
  uint64_t P = 1;
  while (P * 2 <= dt && P * tiles < 1024 && (P * 2) * dd * sizeof(V) <= (1ull << 30)) P *= 2;
  const uint64_t kc = dt / P;
  const size_t m_bytes = direct ? 0 : (size_t)s->dim * sizeof(V);
  const size_t p_bytes = P > 1 ? (size_t)(P * dd * sizeof(V)) : 0;
  if ((rc = ensure_scratch(ctx, ((m_bytes + 255) & ~(size_t)255) + p_bytes))) return rc;
  V* mbuf = (V*)ctx->scratch;
  V* part = (V*)((unsigned char*)ctx->scratch + ((m_bytes + 255) & ~(size_t)255));
  for (int b = 0; b < s->batch; ++b) {
    const V* item = (const V*)s->d + ((uint64_t)b << s->n);
    V* res = (V*)out->d + (uint64_t)b * dd;
    const V* M = item;
    if (!direct) {
      hipLaunchKernelGGL(k_ptrace_gather<T>, dim3(std::min(stride_blocks(s->dim, 256), 4096u)),
                         dim3(256), 0,
                         ctx->stream, item, keep, trace, k, s->dim, mbuf);
      SMOQ_CHECK_LAUNCH("k_ptrace_gather");
      M = mbuf;
    }
    V* C = P > 1 ? part : res;
    rocblas_status st;
    if constexpr (sizeof(V) == 16) {
      const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
      st = rocblas_zgemm_strided_batched(
          h, rocblas_operation_none, rocblas_operation_conjugate_transpose, (rocblas_int)D,
          (rocblas_int)D, (rocblas_int)kc, &one, (const rocblas_double_complex*)M, (rocblas_int)D,
          (rocblas_stride)(D * kc), (const rocblas_double_complex*)M, (rocblas_int)D,
          (rocblas_stride)(D * kc), &zero, (rocblas_double_complex*)C, (rocblas_int)D,
          (rocblas_stride)dd, (rocblas_int)P);
    } else {
      const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
      st = rocblas_cgemm_strided_batched(
          h, rocblas_operation_none, rocblas_operation_conjugate_transpose, (rocblas_int)D,
          (rocblas_int)D, (rocblas_int)kc, &one, (const rocblas_float_complex*)M, (rocblas_int)D,
          (rocblas_stride)(D * kc), (const rocblas_float_complex*)M, (rocblas_int)D,
          (rocblas_stride)(D * kc), &zero, (rocblas_float_complex*)C, (rocblas_int)D,
          (rocblas_stride)dd, (rocblas_int)P);
    }
    if (st != rocblas_status_success)
      return set_error(SMOQ_EDEVICE, "smoq_partial_trace: rocBLAS gemm failed (%d)", (int)st);
    if (P > 1 && (rc = ptrace_finish<T, V>(ctx->stream, part, (int)P, dd, P * dd, res, dd)))
      return rc;
  }
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

template <typename T>
static int ptrace_rho(const smoq_state* rho, uint64_t keep, int k, smoq_state* out) {
  smoq_ctx* ctx = rho->ctx;
  const int N = rho->n / 2;
  const uint64_t trace = ((1ull << N) - 1) & ~keep;
  const uint64_t D = 1ull << k, dd = D * D, dt = 1ull << (N - k);
  const size_t B = (size_t)rho->batch;
  uint64_t P = 1;  // column chunks while few output entries
  while (P * 2 <= dt && dd * P * 2 <= 65536) P *= 2;
  int rc = ensure_scratch(ctx, B * P * dd * sizeof(double2));
  if (rc) return rc;
  double2* part = (double2*)ctx->scratch;
  const unsigned bx = (unsigned)std::min<uint64_t>((dd + 255) / 256, 4096);
  for (size_t b = 0; b < B; ++b) {
    hipLaunchKernelGGL(k_ptrace_rho<T>, dim3(bx, (unsigned)P), dim3(256), 0, ctx->stream,
                       (const VT_alias<T>*)rho->d + (b << (2 * N)), N, keep, trace, k, dt / P,
                       part + b * P * dd);
    SMOQ_CHECK_LAUNCH("k_ptrace_rho");
  }
  const uint64_t total = B * dd;
  if ((rc = ptrace_finish<T, double2>(ctx->stream, part, (int)P, dd, P * dd,
                                      (VT_alias<T>*)out->d, total)))
    return rc;
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

static int ptrace_check(const smoq_state* s, int n_logical, uint64_t keep, const smoq_state* out,
                        const char* who, int* k) {
  const uint64_t full = n_logical >= 64 ? ~0ull : ((1ull << n_logical) - 1);
  if (keep == 0 || (keep & ~full))
    return set_error(SMOQ_EINVAL, "%s: keep mask 0x%llx must name 1..%d of the %d qubits", who,
                     (unsigned long long)keep, n_logical, n_logical);
  *k = __builtin_popcountll(keep);
  if (out->n != 2 * *k || out->dtype != s->dtype || out->batch != s->batch || out->ctx != s->ctx ||
      out->d == s->d)
    return set_error(SMOQ_EINVAL,
                     "%s: out must be a separate state of %d qubits (vec of the %d-qubit reduced "
                     "matrix), batch %d, same dtype and context",
                     who, 2 * *k, *k, s->batch);
  if (s->batch > 65535) return set_error(SMOQ_ENOTSUP, "%s: batch above 65535", who);
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_partial_trace(const smoq_state* psi, uint64_t keep_mask, smoq_state* out) {
  int rc = check_state(psi);
  if (rc) return rc;
  if ((rc = check_state(out))) return rc;
  int k = 0;
  if ((rc = ptrace_check(psi, psi->n, keep_mask, out, "smoq_partial_trace", &k))) return rc;
  if (k <= 2)
    return psi->dtype == SMOQ_C128 ? ptrace_pairs<double>(psi, keep_mask, k, out)
                                   : ptrace_pairs<float>(psi, keep_mask, k, out);
  if (k < SMOQ_PT_GEMM_MIN)
    return psi->dtype == SMOQ_C128 ? ptrace_tile<double>(psi, keep_mask, k, out)
                                   : ptrace_tile<float>(psi, keep_mask, k, out);
  return psi->dtype == SMOQ_C128 ? ptrace_gemm<double>(psi, keep_mask, k, out)
                                 : ptrace_gemm<float>(psi, keep_mask, k, out);
}

extern "C" int smoq_dm_partial_trace(const smoq_state* rho, uint64_t keep_mask, smoq_state* out) {
  int rc = check_state(rho);
  if (rc) return rc;
  if ((rc = check_state(out))) return rc;
  if (rho->n % 2)
    return set_error(SMOQ_EINVAL, "smoq_dm_partial_trace: a density matrix is a state of 2N "
                                  "qubits, got %d", rho->n);
  int k = 0;
  if ((rc = ptrace_check(rho, rho->n / 2, keep_mask, out, "smoq_dm_partial_trace", &k)))
    return rc;
  return rho->dtype == SMOQ_C128 ? ptrace_rho<double>(rho, keep_mask, k, out)
                                 : ptrace_rho<float>(rho, keep_mask, k, out);
}
