// Classical-shadow snapshots on the device
// (utils/cpu/cpuClassicalShadows.jl: classical_shadow_snapshot :1001-1037,
// clifford_shadow_snapshot :580-603, collect_shadows :1098-1106,
// collect_clifford_shadows :636-641).
//
// A snapshot rotates a copy of psi qubit by qubit and draws one outcome with
// sample_from_probabilities' rule.  Snapshots that share a rotation string
// share the rotated state: the host groups them, each unique string becomes a
// batch item (psi broadcast, rotated in registers four consecutive qubits at a
// time, in increasing qubit order); below 2^16 amplitudes each item's
// sequential prefix of |amp|^2 is built by one wave (loads coalesced, the adds
// in the reference's order) and every draw is a binary search of it; from
// 2^16 one wave walks the prefix once per item and answers the item's draws
// on the way.
//
// Pauli strings use the reference's arithmetic (H as s (a + b), s (a - b) with
// s = 1/sqrt(2); S† as (a, b) -> (b, -a)), so the rotated amplitudes, the
// prefix and the outcome are those of the reference bit for bit.  Local
// Clifford strings apply 2x2 matrices from a table the caller passes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kShGroup = 4;  // qubits rotated per pass, in registers

template <typename T>
__global__ void __launch_bounds__(256)
    k_shadow_broadcast(const VT_alias<T>* __restrict__ psi, uint64_t dim, uint64_t total,
                       double2* __restrict__ work) {
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (uint64_t)gridDim.x * blockDim.x) {
    const VT_alias<T> v = psi[o & (dim - 1)];
    work[o] = make_double2((double)v.x, (double)v.y);
  }
}

// Qubits q0 .. q0+nq-1 of every item, in increasing order.  codes[item * n + q]:
// Pauli mode 1 = X (H), 2 = Y (S† then H), 3 = Z (none); Clifford mode the
// 0-based index of a 2x2 matrix in cliff (row-major re, im pairs).
template <int NQ>
__global__ void __launch_bounds__(256)
    k_shadow_rotate(double2* __restrict__ work, int n, int q0, uint64_t total,
                    const uint8_t* __restrict__ codes, int clifford,
                    const double* __restrict__ cliff, double s) {
#pragma clang fp contract(off)
  constexpr int m = 1 << NQ;
  const int rb = n - NQ;  // bits outside the group
  const uint64_t low = (1ull << q0) - 1;
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t item = o >> rb, rest = o & ((1ull << rb) - 1);
    const uint64_t base = (item << n) | ((rest & ~low) << NQ) | (rest & low);
    double2 x[m];
#pragma unroll
    for (int e = 0; e < m; ++e) x[e] = work[base | ((uint64_t)e << q0)];
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int code = codes[item * n + q0 + j];
      const int bit = 1 << j;
      if (!clifford) {
        if (code == 3) continue;
#pragma unroll
        for (int e = 0; e < m; ++e) {
          if (e & bit) continue;
          double2 a = x[e], b = x[e | bit];
          if (code == 2) b = make_double2(b.y, -b.x);  // S†: b * (-i)
          x[e] = make_double2(s * (a.x + b.x), s * (a.y + b.y));
          x[e | bit] = make_double2(s * (a.x - b.x), s * (a.y - b.y));
        }
      } else {
        const double* c = cliff + 8 * code;
#pragma unroll
        for (int e = 0; e < m; ++e) {
          if (e & bit) continue;
          const double2 a = x[e], b = x[e | bit];
          // row r: c[r][0] a + c[r][1] b, complex products as a sum of rows
          x[e] = make_double2(c[0] * a.x - c[1] * a.y + (c[2] * b.x - c[3] * b.y),
                              c[0] * a.y + c[1] * a.x + (c[2] * b.y + c[3] * b.x));
          x[e | bit] = make_double2(c[4] * a.x - c[5] * a.y + (c[6] * b.x - c[7] * b.y),
                                    c[4] * a.y + c[5] * a.x + (c[6] * b.y + c[7] * b.x));
        }
      }
    }
#pragma unroll
    for (int e = 0; e < m; ++e) work[base | ((uint64_t)e << q0)] = x[e];
  }
}

// One wave per item: c[i] = c[i-1] + (re^2 + im^2) in index order, the lanes
// loading 64 amplitudes at a time and the sums taken one lane after another.
__global__ void __launch_bounds__(256)
    k_shadow_prefix(const double2* __restrict__ work, int n, uint64_t items,
                    double* __restrict__ c) {
#pragma clang fp contract(off)
  const uint64_t item = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= items) return;
  const int lane = threadIdx.x & 63;
  const uint64_t dim = 1ull << n;
  const double2* w = work + (item << n);
  double* out = c + (item << n);
  double s = 0.0;
  for (uint64_t b0 = 0; b0 < dim; b0 += 64) {
    double p = 0.0;
    if (b0 + lane < dim) {
      const double2 v = w[b0 + lane];
      const double xx = v.x * v.x;
      const double yy = v.y * v.y;
      p = xx + yy;
    }
    const int cnt = dim - b0 < 64 ? (int)(dim - b0) : 64;
    double mine = 0.0;
    for (int l = 0; l < cnt; ++l) {
      const int lo = __builtin_amdgcn_readlane(__double2loint(p), l);
      const int hi = __builtin_amdgcn_readlane(__double2hiint(p), l);
      s = s + __hiloint2double(hi, lo);
      if (lane == l) mine = s;
    }
    if (lane < cnt) out[b0 + lane] = mine;
  }
}

// first i with r <= c[item][i], else the last index
__global__ void __launch_bounds__(256)
    k_shadow_search(const double* __restrict__ c, int n, const double* __restrict__ r,
                    const uint32_t* __restrict__ draw_item, int64_t nd,
                    uint64_t* __restrict__ out) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const uint64_t dim = 1ull << n;
  const double* ci = c + ((uint64_t)draw_item[d] << n);
  const double x = r[d];
  uint64_t lo = 0, hi = dim;
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (ci[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[d] = lo < dim ? lo : dim - 1;
}

// Large states: one 64-thread block per item walks the sequential prefix and
// answers the item's draws on the way.  The lanes load 64 amplitudes and
// stage their |amp|^2 in LDS; lane 0 adds them in index order and, as the
// sum passes each of the item's uniforms (sorted ascending), records the
// index; the walk stops after the largest uniform.  Uniforms above the last
// sum take the last index.
__global__ void __launch_bounds__(64)
    k_shadow_walk(const double2* __restrict__ work, int n, const uint32_t* __restrict__ off,
                  const double* __restrict__ r_sorted, const uint32_t* __restrict__ pos,
                  uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double buf[64];
  __shared__ int s_done;
  const uint64_t item = blockIdx.x;
  const int lane = threadIdx.x;
  const uint64_t dim = 1ull << n;
  const double2* w = work + (item << n);
  const uint32_t d0 = off[item], d1 = off[item + 1];
  uint32_t k = d0;  // next draw (lane 0)
  double s = 0.0;
  if (lane == 0) s_done = d0 == d1;
  __syncthreads();
  for (uint64_t b0 = 0; b0 < dim && !s_done; b0 += 64) {
    const double2 v = w[b0 + lane];  // dim >= 2^16: whole blocks of 64
    const double xx = v.x * v.x;
    const double yy = v.y * v.y;
    buf[lane] = xx + yy;
    __syncthreads();
    if (lane == 0) {
      double rk = r_sorted[k];
      bool done = false;
      // 16 values at a time: the reads, then the adds, then (only when the
      // sub-block's last sum reaches the next uniform) the draw checks
      for (int sb = 0; sb < 64 && !done; sb += 16) {
        double q[16], c[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) q[j] = buf[sb + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          s = s + q[j];
          c[j] = s;
        }
        if (rk <= s) {
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            while (!done && rk <= c[j]) {
              out[pos[k]] = b0 + sb + j;
              if (++k == d1)
                done = true;
              else
                rk = r_sorted[k];
            }
          }
        }
      }
      s_done = done;
    }
    __syncthreads();
  }
  if (lane == 0)
    for (; k < d1; ++k) out[pos[k]] = dim - 1;
}

// Global Clifford snapshots (global_clifford_shadow_snapshot,
// …ClassicalShadows.jl:884-903): one workgroup per snapshot holds psi in LDS,
// applies the snapshot's gate list — 0 = H on bit a, 1 = S on bit a,
// 2 = CNOT with control bit a and target bit b — and draws with
// sample_from_probabilities' rule (thread 0 walks the prefix in index order).
constexpr int kGcMaxQubits = 12;

template <typename T>
__global__ void __launch_bounds__(256)
    k_global_clifford(const VT_alias<T>* __restrict__ psi, int n, const uint8_t* __restrict__ g,
                      const int64_t* __restrict__ off, const double* __restrict__ r,
                      uint64_t* __restrict__ out, double s) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char gc_smem[];
  double2* x = reinterpret_cast<double2*>(gc_smem);  // 2^n amplitudes
  const uint64_t m = blockIdx.x;
  const int dim = 1 << n, t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < dim; i += nt) {
    const VT_alias<T> v = psi[i];
    x[i] = make_double2((double)v.x, (double)v.y);
  }
  __syncthreads();
  for (int64_t k = off[m]; k < off[m + 1]; ++k) {
    const int type = g[3 * k], a = g[3 * k + 1], b = g[3 * k + 2];
    const int ma = 1 << a, mb = 1 << b;
    for (int p = t; p < dim / 2; p += nt) {
      const int i0 = ((p >> a) << (a + 1)) | (p & (ma - 1)), i1 = i0 | ma;
      if (type == 0) {
        const double2 u = x[i0], w = x[i1];
        x[i0] = make_double2(s * (u.x + w.x), s * (u.y + w.y));
        x[i1] = make_double2(s * (u.x - w.x), s * (u.y - w.y));
      } else if (type == 1) {
        const double2 w = x[i1];
        x[i1] = make_double2(-w.y, w.x);  // times i
      } else if (!(i1 & mb)) {            // control set, target clear: swap with target set
        const double2 u = x[i1];
        x[i1] = x[i1 | mb];
        x[i1 | mb] = u;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    const double rr = r[m];
    double acc = 0.0;
    int idx = dim - 1;
    for (int i = 0; i < dim; ++i) {
      const double xx = x[i].x * x[i].x;
      const double yy = x[i].y * x[i].y;
      acc = acc + (xx + yy);
      if (rr <= acc) {
        idx = i;
        break;
      }
    }
    out[m] = (uint64_t)idx;
  }
}

template <typename T>
static int shadows_t(const smoq_state* s, int64_t n_snap, const uint8_t* codes, int clifford,
                     const double* cliff_host, const double* uniforms, uint64_t* out_index) {
  smoq_ctx* ctx = s->ctx;
  const int n = s->n;
  const uint64_t dim = s->dim;
  // unique rotation strings (2 bits a qubit for Pauli codes, 5 for Clifford
  // indices), in order of first appearance
  const int bits = clifford ? 5 : 2;
  const int per_word = 64 / bits;
  const int words = (n + per_word - 1) / per_word;
  std::unordered_map<std::string, uint32_t> uniq;
  std::vector<uint32_t> item_of(n_snap);
  std::vector<int64_t> first;  // snapshot index of each unique string
  std::string key((size_t)words * 8, '\0');
  for (int64_t m = 0; m < n_snap; ++m) {
    std::fill(key.begin(), key.end(), '\0');
    for (int q = 0; q < n; ++q) {
      uint64_t w;
      std::memcpy(&w, &key[(q / per_word) * 8], 8);
      w |= (uint64_t)codes[m * n + q] << (bits * (q % per_word));
      std::memcpy(&key[(q / per_word) * 8], &w, 8);
    }
    auto it = uniq.find(key);
    if (it == uniq.end()) {
      it = uniq.emplace(key, (uint32_t)first.size()).first;
      first.push_back(m);
    }
    item_of[m] = it->second;
  }
  const uint64_t U = first.size();
  // states of 2^16 amplitudes and more walk their prefix once per item
  // (k_shadow_walk); smaller ones keep it for a binary search per draw
  const bool walk = dim >= (1ull << 16);
  // items per chunk: the rotated states (and prefixes) within 8 GiB
  const uint64_t per_item = dim * (sizeof(double2) + (walk ? 0 : sizeof(double)));
  const uint64_t B = std::max<uint64_t>(1, std::min<uint64_t>(U, (8ull << 30) / per_item));
  // draws of each chunk, grouped
  std::vector<std::vector<int64_t>> draws_of((U + B - 1) / B);
  for (int64_t m = 0; m < n_snap; ++m) draws_of[item_of[m] / B].push_back(m);
  size_t max_draws = 0;
  for (auto& v : draws_of) max_draws = std::max(max_draws, v.size());
  const size_t work_b = B * dim * sizeof(double2), pre_b = walk ? 0 : B * dim * sizeof(double);
  const size_t code_b = B * n, cl_b = 24 * 8 * sizeof(double);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t need = up(work_b) + up(pre_b) + up(code_b) + up(cl_b) + up(max_draws * 8) +
                      up(std::max<size_t>(max_draws, B + 1) * 4) + up(max_draws * 8) +
                      up(max_draws * 4);
  unsigned char* dev = nullptr;
  SMOQ_HIP(hipMalloc(&dev, need));
  double2* work = (double2*)dev;
  double* pre = (double*)(dev + up(work_b));
  uint8_t* dcodes = (uint8_t*)((unsigned char*)pre + up(pre_b));
  double* dcl = (double*)((unsigned char*)dcodes + up(code_b));
  double* dr = (double*)((unsigned char*)dcl + up(cl_b));
  uint32_t* ditem = (uint32_t*)((unsigned char*)dr + up(max_draws * 8));
  uint64_t* dout =
      (uint64_t*)((unsigned char*)ditem + up(std::max<size_t>(max_draws, B + 1) * 4));
  uint32_t* dpos = (uint32_t*)((unsigned char*)dout + up(max_draws * 8));
  const double sq = 1 / std::sqrt(2.0);
  hipError_t e = hipSuccess;
  if (clifford) e = hipMemcpyAsync(dcl, cliff_host, cl_b, hipMemcpyHostToDevice, ctx->stream);
  std::vector<uint8_t> hcodes(code_b);
  std::vector<double> hr(max_draws);
  std::vector<uint32_t> hitem(max_draws);
  std::vector<uint64_t> hout(max_draws);
  std::vector<uint32_t> hoff(B + 1), hpos(max_draws);
  for (uint64_t c0 = 0; c0 < U && e == hipSuccess; c0 += B) {
    const uint64_t nb = std::min<uint64_t>(B, U - c0);
    for (uint64_t b = 0; b < nb; ++b)
      std::memcpy(&hcodes[b * n], codes + first[c0 + b] * n, n);
    const auto& dl = draws_of[c0 / B];
    if (!walk) {
      for (size_t d = 0; d < dl.size(); ++d) {
        hr[d] = uniforms[dl[d]];
        hitem[d] = (uint32_t)(item_of[dl[d]] - c0);
      }
    } else {
      // each item's draws together, uniforms ascending; hpos: the draw's slot
      std::vector<uint32_t> ord(dl.size());
      for (size_t d = 0; d < dl.size(); ++d) ord[d] = (uint32_t)d;
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
        const uint32_t ia = item_of[dl[a]], ib = item_of[dl[b]];
        return ia != ib ? ia < ib : uniforms[dl[a]] < uniforms[dl[b]];
      });
      std::fill(hoff.begin(), hoff.end(), 0u);
      for (size_t j = 0; j < ord.size(); ++j) {
        hr[j] = uniforms[dl[ord[j]]];
        hpos[j] = ord[j];
        ++hoff[item_of[dl[ord[j]]] - c0 + 1];
      }
      for (uint64_t b = 0; b < nb; ++b) hoff[b + 1] += hoff[b];
    }
    e = hipMemcpyAsync(dcodes, hcodes.data(), nb * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dr, hr.data(), dl.size() * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !walk)
      e = hipMemcpyAsync(ditem, hitem.data(), dl.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && walk)
      e = hipMemcpyAsync(ditem, hoff.data(), (nb + 1) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && walk)
      e = hipMemcpyAsync(dpos, hpos.data(), dl.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) break;
    const uint64_t total = nb * dim;
    hipLaunchKernelGGL(k_shadow_broadcast<T>, dim3(stride_blocks(total, 256)), dim3(256), 0,
                       ctx->stream, (const VT_alias<T>*)s->d, dim, total, work);
    for (int q0 = 0; q0 < n; q0 += kShGroup) {
      const int nq = std::min(kShGroup, n - q0);
      const uint64_t t = nb << (n - nq);
      const dim3 g(stride_blocks(t, 256));
      switch (nq) {
        case 1: hipLaunchKernelGGL(k_shadow_rotate<1>, g, dim3(256), 0, ctx->stream, work, n, q0, t, dcodes, clifford, dcl, sq); break;
        case 2: hipLaunchKernelGGL(k_shadow_rotate<2>, g, dim3(256), 0, ctx->stream, work, n, q0, t, dcodes, clifford, dcl, sq); break;
        case 3: hipLaunchKernelGGL(k_shadow_rotate<3>, g, dim3(256), 0, ctx->stream, work, n, q0, t, dcodes, clifford, dcl, sq); break;
        default: hipLaunchKernelGGL(k_shadow_rotate<4>, g, dim3(256), 0, ctx->stream, work, n, q0, t, dcodes, clifford, dcl, sq); break;
      }
    }
    if (walk) {
      hipLaunchKernelGGL(k_shadow_walk, dim3((unsigned)nb), dim3(64), 0, ctx->stream, work, n,
                         ditem, dr, dpos, dout);
    } else {
      hipLaunchKernelGGL(k_shadow_prefix, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0,
                         ctx->stream, work, n, nb, pre);
      if (!dl.empty())
        hipLaunchKernelGGL(k_shadow_search, dim3((unsigned)((dl.size() + 255) / 256)), dim3(256),
                           0, ctx->stream, pre, n, dr, ditem, (int64_t)dl.size(), dout);
    }
    e = hipGetLastError();
    if (e == hipSuccess && !dl.empty())
      e = hipMemcpyAsync(hout.data(), dout, dl.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
      for (size_t d = 0; d < dl.size(); ++d) out_index[dl[d]] = hout[d];
  }
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_error(e, "smoq_shadow_snapshots");
  if (se != hipSuccess) return hip_error(se, "smoq_shadow_snapshots");
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_global_clifford_snapshots(const smoq_state* psi, int64_t n_snap,
                                              const uint8_t* gates, const int64_t* offsets,
                                              const double* uniforms, uint64_t* out_index) {
  int rc = check_state(psi);
  if (rc) return rc;
  if (n_snap < 0 || (n_snap > 0 && (!gates || !offsets || !uniforms || !out_index)))
    return set_error(SMOQ_EINVAL, "smoq_global_clifford_snapshots: bad snapshot arrays");
  if (psi->batch != 1)
    return set_error(SMOQ_EINVAL,
                     "smoq_global_clifford_snapshots: state has %d items; snapshots take one",
                     psi->batch);
  if (psi->n > kGcMaxQubits)
    return set_error(SMOQ_ENOTSUP, "smoq_global_clifford_snapshots: more than %d qubits",
                     kGcMaxQubits);
  if (n_snap == 0) return SMOQ_OK;
  if (n_snap > 0x7fffffff)
    return set_error(SMOQ_ENOTSUP, "smoq_global_clifford_snapshots: more than 2^31 snapshots");
  if (offsets[0] != 0) return set_error(SMOQ_EINVAL, "smoq_global_clifford_snapshots: offsets[0] != 0");
  for (int64_t m = 0; m < n_snap; ++m)
    if (offsets[m + 1] < offsets[m])
      return set_error(SMOQ_EINVAL, "smoq_global_clifford_snapshots: offsets decrease at %lld",
                       (long long)m);
  const int64_t ng = offsets[n_snap];
  for (int64_t k = 0; k < ng; ++k) {
    const int type = gates[3 * k], a = gates[3 * k + 1], b = gates[3 * k + 2];
    if (type > 2 || a >= psi->n || (type == 2 && (b >= psi->n || b == a)))
      return set_error(SMOQ_EINVAL, "smoq_global_clifford_snapshots: bad gate %lld", (long long)k);
  }
  smoq_ctx* ctx = psi->ctx;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t gb = up((size_t)ng * 3 + 1), ob = up((size_t)(n_snap + 1) * 8),
               rb = up((size_t)n_snap * 8);
  unsigned char* dev = nullptr;
  SMOQ_HIP(hipMalloc(&dev, gb + ob + rb + (size_t)n_snap * 8));
  uint8_t* dg = dev;
  int64_t* doff = (int64_t*)(dev + gb);
  double* dr = (double*)(dev + gb + ob);
  uint64_t* dout = (uint64_t*)(dev + gb + ob + rb);
  hipError_t e = hipMemcpyAsync(dg, gates, (size_t)ng * 3, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(doff, offsets, (size_t)(n_snap + 1) * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dr, uniforms, (size_t)n_snap * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    const double sq = 1 / std::sqrt(2.0);
    // one wave for small states (its barriers are cheap), up to 256 threads
    const unsigned nt = (unsigned)std::min<uint64_t>(256, std::max<uint64_t>(64, psi->dim / 2));
    const size_t shm = (size_t)psi->dim * sizeof(double2);
    if (psi->dtype == SMOQ_C128)
      hipLaunchKernelGGL(k_global_clifford<double>, dim3((unsigned)n_snap), dim3(nt), shm,
                         ctx->stream, (const double2*)psi->d, psi->n, dg, doff, dr, dout, sq);
    else
      hipLaunchKernelGGL(k_global_clifford<float>, dim3((unsigned)n_snap), dim3(nt), shm,
                         ctx->stream, (const float2*)psi->d, psi->n, dg, doff, dr, dout, sq);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_index, dout, (size_t)n_snap * 8, hipMemcpyDeviceToHost, ctx->stream);
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_error(e, "smoq_global_clifford_snapshots");
  if (se != hipSuccess) return hip_error(se, "smoq_global_clifford_snapshots");
  return SMOQ_OK;
}

extern "C" int smoq_shadow_snapshots(const smoq_state* psi, int64_t n_snap, int kind,
                                     const uint8_t* codes, const double* clifford_table,
                                     const double* uniforms, uint64_t* out_index) {
  int rc = check_state(psi);
  if (rc) return rc;
  if (n_snap < 0 || (n_snap > 0 && (!codes || !uniforms || !out_index)))
    return set_error(SMOQ_EINVAL, "smoq_shadow_snapshots: bad snapshot arrays");
  if (psi->batch != 1)
    return set_error(SMOQ_EINVAL, "smoq_shadow_snapshots: state has %d items; snapshots take one",
                     psi->batch);
  if (kind != SMOQ_SHADOW_PAULI && kind != SMOQ_SHADOW_LOCAL_CLIFFORD)
    return set_error(SMOQ_EINVAL, "smoq_shadow_snapshots: unknown kind %d", kind);
  const bool cl = kind == SMOQ_SHADOW_LOCAL_CLIFFORD;
  if (cl && !clifford_table)
    return set_error(SMOQ_EINVAL, "smoq_shadow_snapshots: null Clifford table");
  for (int64_t i = 0; i < n_snap * psi->n; ++i)
    if (cl ? codes[i] > 23 : (codes[i] < 1 || codes[i] > 3))
      return set_error(SMOQ_EINVAL, "smoq_shadow_snapshots: code %d at %lld outside %s",
                       (int)codes[i], (long long)i, cl ? "0..23" : "1..3");
  if (n_snap == 0) return SMOQ_OK;
  if (psi->n > 32)
    return set_error(SMOQ_ENOTSUP, "smoq_shadow_snapshots: more than 32 qubits");
  return psi->dtype == SMOQ_C128
             ? shadows_t<double>(psi, n_snap, codes, cl, clifford_table, uniforms, out_index)
             : shadows_t<float>(psi, n_snap, codes, cl, clifford_table, uniforms, out_index);
}
