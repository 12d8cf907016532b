// Batched Monte Carlo wave-function trajectories with σ⁻ jumps in one
// persistent launch: one workgroup per trajectory runs all of its steps.
//
// A trajectory of n <= 12 qubits lives in LDS for the whole run.  Larger
// ones stay in HBM and every step makes two tiled passes: pass A over the
// contiguous 4096-amplitude chunks (local qubits 0..11) applies the pending
// renormalisation, the diagonal bond gates and the site gates of qubits
// 0..11; pass B over tiles made of the low 12-h qubits and the high h = n-12
// qubits applies the remaining site gates, the no-jump decay and the norm
// reduction.  The jump decision is taken by one thread from block sums;
// site populations are reduced only on steps that jump or record <Z>.
// The 1/||ψ|| of step s is folded into the first pass of step s+1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kMcBlock = 256;
constexpr int kMcTileQ = 12;
constexpr int kMcMaxQ = 20;
constexpr int kMcWaves = kMcBlock / 64;

struct McwfGroupDev {
  void* psi;
  int n, n_bonds, uniform_chain;
  int site_rx;  // every site matrix real on the diagonal, imaginary off it
  const int32_t* bond_q;
  const double2* bond_diag;
  const double2* site_u;
  const double* decay_pow;
  const double2* chain_tab;  // open uniform chain: product phase by #unequal neighbours
  const double* uniforms;
  int32_t* jumps;
  double* z_out;
};

struct McwfTrajDev {
  int group, idx;
};

struct McwfArgs {
  const McwfGroupDev* groups;
  const McwfTrajDev* trajs;
  double gamma_dt;
  int n_steps, obs_every;
  // smallest blockIdx.x * n_steps + step whose renormalisation met a zero or
  // non-finite norm; all ones when none did
  unsigned long long* fault;
};

__device__ inline double mc_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block sums of v[0..m), m <= M; every thread reads the results from out[]
template <int M>
__device__ inline void mc_block_sum(const double (&v)[M], int m, double* red, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    if (i < m) {
      const double s = mc_wave_sum(v[i]);
      if (lane == 0) red[i * kMcWaves + wid] = s;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < m) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kMcWaves; ++w) s += red[threadIdx.x * kMcWaves + w];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// global index of tile position l in tile t
struct TileMap {
  int pass_b;  // 0: chunk t of 2^tq contiguous amplitudes; 1: low 12-h bits + high h bits
  int tq, h;
  __device__ inline uint64_t g(uint64_t t, uint32_t l) const {
    if (!pass_b) return (t << tq) | l;
    const int lb = kMcTileQ - h;
    const uint64_t low = l & ((1u << lb) - 1u);
    return low | (t << lb) | ((uint64_t)(l >> lb) << kMcTileQ);
  }
};

// a * p + i b * q for real a, b: (a p.x - b q.y, a p.y + b q.x), each
// product rounded on its own as in the general complex form
__device__ inline double2 rx_pair(double a, double b, double2 p, double2 q) {
#pragma clang fp contract(off)
  double2 r;
  r.x = a * p.x - b * q.y;
  r.y = a * p.y + b * q.x;
  return r;
}

// the step's small tables, copied into LDS by the kernel: the stages read
// them at every tile, and from global memory each read would wait out a
// cache latency
struct McTabs {
  const double2* site_u;     // 4 per site
  const double2* chain_tab;  // uniform chain: n + 1 products
  const double* decay_pow;   // n + 1
};

struct StageOpts {
  double pre_scale;  // applied on load
  bool bonds;        // bond phases before the site gates
  bool decay;        // decay + norm after them
};

// One LDS round trip over the tile: 8 amplitudes per thread spanning three
// tile positions, site gates on the positions with site[s] >= 0.
template <typename T>
__device__ void mc_stage(VT_alias<T>* lds, int tq, const int (&pos)[3], const int (&site)[3],
                         const McwfGroupDev& G, const McTabs& tb, const TileMap& tm, uint64_t t,
                         const StageOpts& o, uint64_t dmask, double& norm_acc) {
  constexpr int WB = CT<T>::kRowBits;
  const uint32_t tmask = (1u << tq) - 1u;
  const uint32_t pm = (1u << pos[0]) | (1u << pos[1]) | (1u << pos[2]);
  const uint32_t groups = 1u << (tq - 3);
  for (uint32_t grp = threadIdx.x; grp < groups; grp += kMcBlock) {
    const uint32_t base = pdep32(grp, tmask & ~pm);
    double2 x[8];
    uint32_t li[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      li[i] = base | ((i & 1) ? (1u << pos[0]) : 0u) | ((i & 2) ? (1u << pos[1]) : 0u) |
              ((i & 4) ? (1u << pos[2]) : 0u);
      x[i] = to_d2(lds[swz<WB>(li[i])]);
      x[i].x *= o.pre_scale;
      x[i].y *= o.pre_scale;
    }
    if (o.bonds) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint64_t gi = tm.g(t, li[i]) & dmask;
        if (G.uniform_chain) {
          const int nd = __popcll((gi ^ (gi >> 1)) & (dmask >> 1));
          x[i] = cmul(tb.chain_tab[nd], x[i]);
        } else {
          for (int b = 0; b < G.n_bonds; ++b) {
            const int e = (int)((gi >> G.bond_q[2 * b]) & 1) |
                          ((int)((gi >> G.bond_q[2 * b + 1]) & 1) << 1);
            x[i] = cmul(G.bond_diag[4 * b + e], x[i]);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      if (site[s] < 0) continue;
      if (G.site_rx) {
        // the products with the matrix's zero parts left out: the same
        // roundings as the general form below
        const double2* u = tb.site_u + 4 * site[s];
        const double a0 = u[0].x, b0 = u[1].y, b1 = u[2].y, a1 = u[3].x;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          if (i & (1 << s)) continue;
          const int j = i | (1 << s);
          const double2 a = x[i], b = x[j];
          x[i] = rx_pair(a0, b0, a, b);
          x[j] = rx_pair(a1, b1, b, a);
        }
        continue;
      }
      const double2* u = tb.site_u + 4 * site[s];
      const double2 u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (i & (1 << s)) continue;
        const int j = i | (1 << s);
        const double2 a = x[i], b = x[j];
        x[i] = cfma2(u0, a, u1, b);
        x[j] = cfma2(u2, a, u3, b);
      }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (o.decay) {
        // the norm is that of the amplitudes as stored
        const int ne = __popcll(tm.g(t, li[i]) & dmask);
        const double d = ne > 0 ? tb.decay_pow[ne] : 1.0;
        const VT_alias<T> r = make_c<T>(x[i].x * d, x[i].y * d);
        lds[swz<WB>(li[i])] = r;
        norm_acc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
      } else {
        lds[swz<WB>(li[i])] = make_c<T>(x[i].x, x[i].y);
      }
    }
  }
  __syncthreads();
}

// site gates site0 + k on tile positions pos0 + k, k < cnt, three per stage;
// the pending scale and the bonds ride on the first stage, decay + norm on
// the last
template <typename T>
__device__ void mc_gates(VT_alias<T>* lds, int tq, int pos0, int site0, int cnt,
                         const McwfGroupDev& G, const McTabs& tb, const TileMap& tm, uint64_t t,
                         StageOpts o, uint64_t dmask, double& norm_acc) {
  const int stages = cnt > 3 ? (cnt + 2) / 3 : 1;
  for (int st = 0; st < stages; ++st) {
    int pos[3], site[3];
    uint32_t used = 0;
    for (int s = 0; s < 3; ++s) {
      const int k = st * 3 + s;
      pos[s] = k < cnt ? pos0 + k : -1;
      site[s] = k < cnt ? site0 + k : -1;
      if (pos[s] >= 0) used |= 1u << pos[s];
    }
    for (int s = 0; s < 3; ++s) {  // pad with free positions that carry no gate
      if (pos[s] >= 0) continue;
      const int b = __ffs(~used & ((1u << tq) - 1u)) - 1;
      pos[s] = b;
      used |= 1u << b;
    }
    StageOpts so{st == 0 ? o.pre_scale : 1.0, o.bonds && st == 0, o.decay && st == stages - 1};
    mc_stage<T>(lds, tq, pos, site, G, tb, tm, t, so, dmask, norm_acc);
  }
}

// A full 2^kMcTileQ tile into LDS with eight of the thread's loads in
// flight at a time (a loop of single loads waits out one memory latency per
// iteration).
template <typename T>
__device__ __forceinline__ void mc_load_full_tile(VT_alias<T>* lds, const VT_alias<T>* psi,
                                                  const TileMap& tm, uint64_t t) {
  constexpr int WB = CT<T>::kRowBits;
  constexpr int kPer = (1 << kMcTileQ) / kMcBlock;
  // plain vector registers (the HIP vector type's copies can go through
  // private memory)
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2* src = reinterpret_cast<const v2*>(psi);
  v2* dst = reinterpret_cast<v2*>(lds);
#pragma unroll
  for (int h = 0; h < kPer; h += 8) {
    v2 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = src[tm.g(t, threadIdx.x + (uint32_t)(h + i) * kMcBlock)];
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[swz<WB>(threadIdx.x + (uint32_t)(h + i) * kMcBlock)] = v[i];
  }
}

// Register stages of 16 amplitudes per thread (kMcBlock threads over one
// 2^kMcTileQ tile), register index i over four tile positions: 0..3, 4..7
// or 8..11, the other eight from the thread index.  The last map is the
// coalesced one (16 contiguous amplitudes per group of 16 lanes).
__device__ __forceinline__ uint32_t mc_l0(uint32_t tx, uint32_t i) { return (tx << 4) | i; }
__device__ __forceinline__ uint32_t mc_l4(uint32_t tx, uint32_t i) {
  return (tx & 15u) | (i << 4) | ((tx >> 4) << 8);
}
__device__ __forceinline__ uint32_t mc_l8(uint32_t tx, uint32_t i) { return tx | (i << 8); }
// LDS row of tile index l in the 16-amplitude stages: the low four bits
// XORed with the two higher nibbles, so that for each of the three maps the
// 16 lanes of a row group read 16 distinct 16-byte bank groups
__device__ __forceinline__ uint32_t mc_swz(uint32_t l) {
  return l ^ (((l >> 4) ^ (l >> 8)) & 15u);
}
// the thread index as a value the compiler cannot hoist out of a loop: the
// 16-amplitude maps are then recomputed per tile (a few integer operations)
// instead of held in registers across it
__device__ __forceinline__ uint32_t mc_tx() {
  uint32_t tx = threadIdx.x;
  asm volatile("" : "+v"(tx));
  return tx;
}

// the site gate of `site` on register bit B of the thread's 16 amplitudes
template <int B>
__device__ __forceinline__ void mc_site16(double2 (&x)[16], int site, const McwfGroupDev& G,
                                          const McTabs& tb) {
  const double2* u = tb.site_u + 4 * site;
  if (G.site_rx) {
    const double a0 = u[0].x, b0 = u[1].y, b1 = u[2].y, a1 = u[3].x;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i & (1 << B)) continue;
      const int j = i | (1 << B);
      const double2 a = x[i], b = x[j];
      x[i] = rx_pair(a0, b0, a, b);
      x[j] = rx_pair(a1, b1, b, a);
    }
  } else {
    const double2 u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i & (1 << B)) continue;
      const int j = i | (1 << B);
      const double2 a = x[i], b = x[j];
      x[i] = cfma2(u0, a, u1, b);
      x[j] = cfma2(u2, a, u3, b);
    }
  }
}

// site gates on register bits lo..hi-1 in ascending order, site site0 + b
// on bit b
__device__ __forceinline__ void mc_sites16(double2 (&x)[16], int lo, int hi, int site0,
                                           const McwfGroupDev& G, const McTabs& tb) {
  // scheduling barriers keep the compiler from interleaving the four gates
  // over all 16 amplitudes (more live values than the register budget)
  if (lo <= 0 && 0 < hi) mc_site16<0>(x, site0, G, tb);
  __builtin_amdgcn_sched_barrier(0);
  if (lo <= 1 && 1 < hi) mc_site16<1>(x, site0 + 1, G, tb);
  __builtin_amdgcn_sched_barrier(0);
  if (lo <= 2 && 2 < hi) mc_site16<2>(x, site0 + 2, G, tb);
  __builtin_amdgcn_sched_barrier(0);
  if (lo <= 3 && 3 < hi) mc_site16<3>(x, site0 + 3, G, tb);
  __builtin_amdgcn_sched_barrier(0);
}

// the bond phases of global index gi
__device__ __forceinline__ double2 mc_bonds(double2 v, uint64_t gi, uint64_t dmask,
                                            const McwfGroupDev& G, const McTabs& tb) {
  gi &= dmask;
  if (G.uniform_chain) {
    const int nd = __popcll((gi ^ (gi >> 1)) & (dmask >> 1));
    return cmul(tb.chain_tab[nd], v);
  }
  for (int b = 0; b < G.n_bonds; ++b) {
    const int e = (int)((gi >> G.bond_q[2 * b]) & 1) | ((int)((gi >> G.bond_q[2 * b + 1]) & 1) << 1);
    v = cmul(G.bond_diag[4 * b + e], v);
  }
  return v;
}

template <typename T>
__global__ void __launch_bounds__(kMcBlock, 2) k_mcwf(McwfArgs A) {
  using V = VT_alias<T>;
  constexpr int WB = CT<T>::kRowBits;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  __shared__ double red[(kMcMaxQ + 1) * kMcWaves];
  __shared__ double sums[kMcMaxQ + 1];
  __shared__ double s_scale;
  __shared__ int s_site;

  const McwfTrajDev tr = A.trajs[blockIdx.x];
  const McwfGroupDev G = A.groups[tr.group];
  const int n = G.n;
  __shared__ double2 tab_u[4 * kMcMaxQ], tab_chain[kMcMaxQ + 1];
  __shared__ double tab_decay[kMcMaxQ + 1];
  for (int i = threadIdx.x; i < 4 * n; i += kMcBlock) tab_u[i] = G.site_u[i];
  for (int i = threadIdx.x; i <= n; i += kMcBlock) {
    tab_decay[i] = G.decay_pow[i];
    tab_chain[i] = G.uniform_chain ? G.chain_tab[i] : make_double2(1.0, 0.0);
  }
  __syncthreads();
  const McTabs tb{tab_u, tab_chain, tab_decay};
  const uint64_t D = 1ull << n, dmask = D - 1;
  // the state through a global-memory pointer (a pointer read from memory is
  // generic, and generic loads also count against the LDS wait counter)
  typedef __attribute__((address_space(1))) V GV;
  V* psi = (V*)((GV*)G.psi + (uint64_t)tr.idx * D);
  const bool resident = n <= kMcTileQ;
  const int tq = resident ? (n < 3 ? 3 : n) : kMcTileQ;  // tiny states pad the tile with zeros
  const int h = resident ? 0 : n - kMcTileQ;
  const uint64_t ntiles = 1ull << h;
  const uint32_t tile = 1u << tq;
  const int rec_n = A.obs_every > 0 ? A.n_steps / A.obs_every : 0;
  const TileMap tmA{0, tq, h}, tmB{1, tq, h};

  const int cntA = n < kMcTileQ ? n : kMcTileQ;

  auto load_tile = [&](const TileMap& tm, uint64_t t) {
    if (!resident) {
      mc_load_full_tile<T>(lds, psi, tm, t);
    } else {
      for (uint32_t l = threadIdx.x; l < tile; l += kMcBlock) {
        const uint64_t gi = tm.g(t, l);
        lds[swz<WB>(l)] = gi < D ? psi[gi] : make_c<T>(0.0, 0.0);
      }
    }
    __syncthreads();
  };
  auto store_tile = [&](const TileMap& tm, uint64_t t) {
    for (uint32_t l = threadIdx.x; l < tile; l += kMcBlock) psi[tm.g(t, l)] = lds[swz<WB>(l)];
    __syncthreads();
  };
  // sums[i] = Σ_{bit i set} |ψ|² for i < n, sums[n] = ||ψ||², over the
  // stored state.  Index k = tid + 256 j: bits 0..7 come from tid, 8.. from j.
  auto pops = [&]() {
    double tot = 0.0, hi[kMcMaxQ - 8];
#pragma unroll
    for (int i = 0; i < kMcMaxQ - 8; ++i) hi[i] = 0.0;
    auto add = [&](uint64_t j, const V& v) {
      const double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      tot += p;
#pragma unroll
      for (int i = 0; i < kMcMaxQ - 8; ++i)
        if ((j >> i) & 1) hi[i] += p;
    };
    if (resident) {
      for (uint64_t j = 0; (j << 8) < D; ++j) {
        const uint64_t k = (j << 8) | threadIdx.x;
        if (k >= D) break;
        add(j, lds[swz<WB>((uint32_t)k)]);
      }
    } else {
      // 2^n >= 2^13 amplitudes: eight loads of the thread in flight at once
      typedef T v2 __attribute__((ext_vector_type(2)));
      const v2* src = reinterpret_cast<const v2*>(psi);
      for (uint64_t j0 = 0; (j0 << 8) < D; j0 += 8) {
        v2 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = src[((j0 + i) << 8) | threadIdx.x];
#pragma unroll
        for (int i = 0; i < 8; ++i) add(j0 + i, make_c<T>(v[i].x, v[i].y));
      }
    }
    double v[kMcMaxQ + 1];
#pragma unroll
    for (int i = 0; i <= kMcMaxQ; ++i) {
      double b = i < 8 ? (((threadIdx.x >> i) & 1) ? tot : 0.0) : hi[i < 8 ? 0 : i - 8];
      if (i == kMcMaxQ) b = 0.0;
      v[i] = (i == n) ? tot : b;  // the total goes to slot n
    }
    mc_block_sum<kMcMaxQ + 1>(v, n + 1, red, sums);
  };

  if (threadIdx.x == 0) s_scale = 1.0;
  if (resident) load_tile(tmA, 0);
  __syncthreads();

  for (int step = 0; step < A.n_steps; ++step) {
    const double sc = s_scale;
    double norm_acc = 0.0;
    if (resident) {
      mc_gates<T>(lds, tq, 0, 0, cntA, G, tb, tmA, 0, StageOpts{sc, true, true}, dmask,
                  norm_acc);
    } else {
      // Two passes per step, 16 amplitudes per thread in registers; every
      // amplitude sees the same operations in the same order as in the
      // gate list (scale, bonds, sites 0..n-1, decay).
      typedef T v2 __attribute__((ext_vector_type(2)));
      const v2* gsrc = reinterpret_cast<const v2*>(psi);
      v2* ldv = reinterpret_cast<v2*>(lds);
      double2 x[16];
      // pass A: tiles of the contiguous qubits 0..11 (sites 0..11)
      for (uint64_t t = 0; t < ntiles; ++t) {
        const uint64_t gb = t << kMcTileQ;
        const uint32_t tx = mc_tx();
        {
          v2 v[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) v[i] = gsrc[gb | mc_l8(tx, i)];
          __syncthreads();  // the previous tile's readers are done with LDS
#pragma unroll
          for (int i = 0; i < 16; ++i) ldv[mc_swz(mc_l8(tx, i))] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const v2 w = ldv[mc_swz(mc_l0(tx, i))];
          x[i] = make_double2((double)w.x * sc, (double)w.y * sc);
          x[i] = mc_bonds(x[i], gb | mc_l0(tx, i), dmask, G, tb);
        }
        mc_sites16(x, 0, 4, 0, G, tb);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const VT_alias<T> r = make_c<T>(x[i].x, x[i].y);
          lds[mc_swz(mc_l0(tx, i))] = r;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = to_d2(lds[mc_swz(mc_l4(tx, i))]);
        mc_sites16(x, 0, 4, 4, G, tb);
#pragma unroll
        for (int i = 0; i < 16; ++i) lds[mc_swz(mc_l4(tx, i))] = make_c<T>(x[i].x, x[i].y);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = to_d2(lds[mc_swz(mc_l8(tx, i))]);
        mc_sites16(x, 0, 4, 8, G, tb);
#pragma unroll
        for (int i = 0; i < 16; ++i) psi[gb | mc_l8(tx, i)] = make_c<T>(x[i].x, x[i].y);
      }
      __syncthreads();  // pass A's stores are seen by every thread
      // pass B: tiles of the low lb = 12 - h qubits and the high h qubits
      // (sites 12..n-1 at tile positions lb..11), decay and norm
      const int lb = kMcTileQ - h;
      const uint32_t lowm = (1u << lb) - 1u;
      auto gB = [&](uint64_t t, uint32_t l) {
        return (uint64_t)(l & lowm) | (t << lb) | ((uint64_t)(l >> lb) << kMcTileQ);
      };
      for (uint64_t t = 0; t < ntiles; ++t) {
        const uint32_t tx = mc_tx();
        if (lb >= 8) {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const v2 w = gsrc[gB(t, mc_l8(tx, i))];
            x[i] = make_double2((double)w.x, (double)w.y);
          }
        } else {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const v2 w = gsrc[gB(t, mc_l4(tx, i))];
            x[i] = make_double2((double)w.x, (double)w.y);
          }
          mc_sites16(x, lb - 4, 4, kMcTileQ - (lb - 4), G, tb);
          __syncthreads();  // the previous tile's readers are done with LDS
#pragma unroll
          for (int i = 0; i < 16; ++i) lds[mc_swz(mc_l4(tx, i))] = make_c<T>(x[i].x, x[i].y);
          __syncthreads();
#pragma unroll
          for (int i = 0; i < 16; ++i) x[i] = to_d2(lds[mc_swz(mc_l8(tx, i))]);
        }
        const int lo8 = lb > 8 ? lb - 8 : 0;
        mc_sites16(x, lo8, 4, kMcTileQ + 8 - lb, G, tb);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const uint64_t gi = gB(t, mc_l8(tx, i));
          const int ne = __popcll(gi & dmask);
          const double d = ne > 0 ? tb.decay_pow[ne] : 1.0;
          const VT_alias<T> r = make_c<T>(x[i].x * d, x[i].y * d);
          psi[gi] = r;
          norm_acc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
        }
      }
    }
    double nv[1] = {norm_acc};
    mc_block_sum<1>(nv, 1, red, sums);
    const double norm_sq = sums[0];
    const double dp = 1.0 - norm_sq;
    const double r = G.uniforms[(uint64_t)tr.idx * A.n_steps + step];
    const bool jump = (r < dp) && (dp > 0.0);
    double new_norm_sq = norm_sq;
    if (jump) {  // uniform over the block: every thread read the same sums[0] and r
      pops();
      if (threadIdx.x == 0) {
        double cum = 0.0;
        int sel = 0;
        for (int k = 0; k < n; ++k) {
          cum += A.gamma_dt * sums[k] / norm_sq;
          if (r < cum) {
            sel = k;
            break;
          }
        }
        s_site = sel;
      }
      __syncthreads();
      const int k = s_site;
      new_norm_sq = sums[k];
      const uint64_t m = 1ull << k;
      if (resident) {
        for (uint64_t p = threadIdx.x; p < (D >> 1); p += kMcBlock) {
          const uint64_t i0 = insert_zero(p, k), i1 = i0 | m;
          lds[swz<WB>((uint32_t)i0)] = lds[swz<WB>((uint32_t)i1)];
          lds[swz<WB>((uint32_t)i1)] = make_c<T>(0.0, 0.0);
        }
      } else {
        // 2^(n-1) >= 2^12 pairs: eight of the thread's loads in flight at once
        typedef T v2 __attribute__((ext_vector_type(2)));
        v2* pv = reinterpret_cast<v2*>(psi);
        const v2 zero = {(T)0, (T)0};
        for (uint64_t p0 = threadIdx.x; p0 < (D >> 1); p0 += 8 * kMcBlock) {
          v2 v[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) v[i] = pv[insert_zero(p0 + (uint64_t)i * kMcBlock, k) | m];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint64_t i0 = insert_zero(p0 + (uint64_t)i * kMcBlock, k);
            pv[i0] = v[i];
            pv[i0 | m] = zero;
          }
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      G.jumps[(uint64_t)tr.idx * A.n_steps + step] = jump ? s_site : -1;
      s_scale = 1.0 / sqrt(new_norm_sq);
      if (!(new_norm_sq > 0.0) || !isfinite(s_scale))
        atomicMin(A.fault, (unsigned long long)blockIdx.x * A.n_steps + step);
    }
    __syncthreads();
    // the trajectory cannot continue; the host reports the first such step
    if (!(new_norm_sq > 0.0) || !isfinite(s_scale)) break;
    if (rec_n > 0 && G.z_out && (step + 1) % A.obs_every == 0) {
      // <Z_i> = (||ψ||² - 2 pop_i) / ||ψ||², independent of the pending scale
      pops();
      const int rix = (step + 1) / A.obs_every - 1;
      if ((int)threadIdx.x < n) {
        const double nn = sums[n];
        G.z_out[((uint64_t)tr.idx * rec_n + rix) * n + threadIdx.x] =
            (nn - 2.0 * sums[threadIdx.x]) / nn;
      }
      __syncthreads();
    }
  }
  // apply the last renormalisation and write the state back
  const double sc = s_scale;
  for (uint64_t k = threadIdx.x; k < D; k += kMcBlock) {
    const V v = resident ? lds[swz<WB>((uint32_t)k)] : psi[k];
    psi[k] = make_c<T>(sc * (double)v.x, sc * (double)v.y);
  }
}

// ------------------------------------------------------------------------
// The dissipative half of one MCWF step for any coherent step (applied
// before, e.g. by a compiled plan), on every trajectory (batch item) at
// once.  Decay (σ⁻: sqrt(1-γdt)^popcount, σ_z: the uniform sqrt(1-nγdt)),
// norm, decision, jump site, jump and renormalisation in two launches over
// chunks of each trajectory: k_jump_sums reads the chunk once for the norm
// of the decayed amplitudes (and, for σ⁻, the site populations);
// k_jump_apply adds the chunks' partial sums in a fixed order, takes the
// decision and rewrites the chunk (decay, jump, 1/norm) in one sweep, with
// the rounding of decaying first and scaling second.
constexpr int kJumpMaxQ = 30;
constexpr int kJumpSlots = kJumpMaxQ + 1;  // norm, then one population per site

int mcwf_jump_blocks(const smoq_state* s) {
  int nb = 1;
  while ((uint64_t)nb * 2 <= (s->dim >> 12) && (uint64_t)nb * 2 * s->batch <= 2048) nb *= 2;
  return nb;
}

size_t mcwf_jump_partial_bytes(const smoq_state* s) {
  return (size_t)s->batch * mcwf_jump_blocks(s) * kJumpSlots * sizeof(double);
}

template <typename T, bool POPS>
__global__ void __launch_bounds__(kMcBlock)
    k_jump_sums(const VT_alias<T>* __restrict__ psi, int n, int lc, int sz,
                const double* __restrict__ decay_pow, double* __restrict__ partial) {
  using V = VT_alias<T>;
  __shared__ double red[kJumpSlots * kMcWaves];
  __shared__ double sums[kJumpSlots];
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const uint64_t chunk = 1ull << lc;
  const uint64_t g0 = blk << lc;
  const V* p = psi + (item << n) + g0;
  const int tid = threadIdx.x;
  double tot = 0.0;
  double hi[POPS ? kJumpMaxQ - 8 : 1];
#pragma unroll
  for (int i = 0; i < (POPS ? kJumpMaxQ - 8 : 1); ++i) hi[i] = 0.0;
  for (uint64_t j = 0; (j << 8) < chunk; ++j) {
    const uint64_t k = (j << 8) | (uint64_t)tid;
    if (k >= chunk) break;
    const double d = decay_pow[sz ? 0 : __popcll(g0 | k)];
    const V v = p[k];
    const V r = make_c<T>((double)v.x * d, (double)v.y * d);
    const double q = (double)r.x * (double)r.x + (double)r.y * (double)r.y;
    tot += q;
    if (POPS) {
#pragma unroll
      for (int i = 0; i < kJumpMaxQ - 8; ++i)
        if (i + 8 < lc && ((j >> i) & 1)) hi[i] += q;
    }
  }
  double v[kJumpSlots];
  v[0] = tot;
#pragma unroll
  for (int b = 0; b < kJumpMaxQ; ++b) {
    double s;
    if (b >= lc)
      s = ((blk >> (b - lc)) & 1) ? tot : 0.0;
    else if (b < 8)
      s = ((tid >> b) & 1) ? tot : 0.0;
    else
      s = POPS ? hi[POPS ? b - 8 : 0] : 0.0;
    v[b + 1] = s;
  }
  const int m = POPS ? n + 1 : 1;
  mc_block_sum<kJumpSlots>(v, m, red, sums);
  if (tid < m) partial[(item * gridDim.x + blk) * kJumpSlots + tid] = sums[tid];
}

template <typename T>
__global__ void __launch_bounds__(kMcBlock)
    k_jump_apply(VT_alias<T>* __restrict__ psi, int n, int lc, int sz, double gamma_dt,
                 const double* __restrict__ decay_pow, const double* __restrict__ partial,
                 const double* __restrict__ uni, int uni_stride, int32_t* __restrict__ jumps,
                 int jump_stride, unsigned long long* fault, int n_steps, int step) {
  using V = VT_alias<T>;
  __shared__ double s_new;
  __shared__ int s_site, s_jump;
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const int nbk = gridDim.x;
  const uint64_t chunk = 1ull << lc;
  V* p = psi + (item << n);
  if (threadIdx.x == 0) {
    const double* pp = partial + item * nbk * kJumpSlots;
    double norm_sq = 0.0;
    for (int b = 0; b < nbk; ++b) norm_sq += pp[b * kJumpSlots];
    const double dp = 1.0 - norm_sq;
    const double r = uni[item * uni_stride];
    const bool jump = (r < dp) && (dp > 0.0);
    double new_norm_sq = norm_sq;
    int site = -1;
    if (jump) {
      if (sz) {
        site = (int)(uni[item * uni_stride + 1] * n);
        if (site > n - 1) site = n - 1;
      } else {
        double cum = 0.0;
        int sel = 0;
        for (int k = 0; k < n; ++k) {
          double pop = 0.0;
          for (int b = 0; b < nbk; ++b) pop += pp[b * kJumpSlots + 1 + k];
          cum += gamma_dt * pop / norm_sq;
          if (r < cum) {
            sel = k;
            break;
          }
        }
        site = sel;
        new_norm_sq = 0.0;
        for (int b = 0; b < nbk; ++b) new_norm_sq += pp[b * kJumpSlots + 1 + sel];
      }
    }
    s_new = new_norm_sq;
    s_site = site;
    s_jump = jump;
    if (blk == 0) {
      jumps[item * jump_stride] = jump ? site : -1;
      if (!(new_norm_sq > 0.0) || !isfinite(1.0 / sqrt(new_norm_sq)))
        atomicMin(fault, (unsigned long long)item * n_steps + step);
    }
  }
  __syncthreads();
  const double sc = 1.0 / sqrt(s_new);
  const int site = s_site;
  const uint64_t g0 = blk << lc;
  if (s_jump && !sz) {
    // σ⁻ at `site`: ψ[i0] = decayed ψ[i0 | m] / norm, ψ[i0 | m] = 0 for
    // every i0 with the site bit clear; a chunk holding the site bit
    // within it handles its own pairs, otherwise the chunk with the bit
    // clear handles the pairs with its partner chunk
    const uint64_t m = 1ull << site;
    if (site < lc) {
      for (uint64_t q = threadIdx.x; q < (chunk >> 1); q += kMcBlock) {
        const uint64_t i0 = g0 | insert_zero(q, site), i1 = i0 | m;
        const V v = p[i1];
        const double d = decay_pow[__popcll(i1)];
        const V r = make_c<T>((double)v.x * d, (double)v.y * d);
        p[i0] = make_c<T>((double)r.x * sc, (double)r.y * sc);
        p[i1] = make_c<T>(0.0, 0.0);
      }
    } else if (!(g0 & m)) {
      for (uint64_t k = threadIdx.x; k < chunk; k += kMcBlock) {
        const uint64_t i0 = g0 | k, i1 = i0 | m;
        const V v = p[i1];
        const double d = decay_pow[__popcll(i1)];
        const V r = make_c<T>((double)v.x * d, (double)v.y * d);
        p[i0] = make_c<T>((double)r.x * sc, (double)r.y * sc);
        p[i1] = make_c<T>(0.0, 0.0);
      }
    }
    return;
  }
  const uint64_t mz = (s_jump && sz) ? (1ull << site) : 0ull;
  for (uint64_t k = threadIdx.x; k < chunk; k += kMcBlock) {
    const uint64_t gk = g0 | k;
    const double d = decay_pow[sz ? 0 : __popcll(gk)];
    const V v = p[gk];
    const V r = make_c<T>((double)v.x * d, (double)v.y * d);
    const double s = (gk & mz) ? -sc : sc;
    p[gk] = make_c<T>((double)r.x * s, (double)r.y * s);
  }
}

int mcwf_jump_launch(smoq_state* s, int kind, double gamma_dt, const double* d_decay,
                     const double* d_uni, int uni_stride, int32_t* d_jumps, int jump_stride,
                     unsigned long long* d_fault, int n_steps, int step, double* d_partial) {
  const int nbk = mcwf_jump_blocks(s);
  const int lc = s->n - __builtin_ctz((unsigned)nbk);
  const int sz = kind == SMOQ_JUMP_SIGMA_Z;
  const dim3 grid((unsigned)nbk, (unsigned)s->batch);
  hipStream_t st = s->ctx->stream;
  if (s->dtype == SMOQ_C128) {
    if (sz)
      hipLaunchKernelGGL((k_jump_sums<double, false>), grid, dim3(kMcBlock), 0, st,
                         (const double2*)s->d, s->n, lc, sz, d_decay, d_partial);
    else
      hipLaunchKernelGGL((k_jump_sums<double, true>), grid, dim3(kMcBlock), 0, st,
                         (const double2*)s->d, s->n, lc, sz, d_decay, d_partial);
    SMOQ_CHECK_LAUNCH("k_jump_sums");
    hipLaunchKernelGGL(k_jump_apply<double>, grid, dim3(kMcBlock), 0, st, (double2*)s->d, s->n,
                       lc, sz, gamma_dt, d_decay, d_partial, d_uni, uni_stride, d_jumps,
                       jump_stride, d_fault, n_steps, step);
  } else {
    if (sz)
      hipLaunchKernelGGL((k_jump_sums<float, false>), grid, dim3(kMcBlock), 0, st,
                         (const float2*)s->d, s->n, lc, sz, d_decay, d_partial);
    else
      hipLaunchKernelGGL((k_jump_sums<float, true>), grid, dim3(kMcBlock), 0, st,
                         (const float2*)s->d, s->n, lc, sz, d_decay, d_partial);
    SMOQ_CHECK_LAUNCH("k_jump_sums");
    hipLaunchKernelGGL(k_jump_apply<float>, grid, dim3(kMcBlock), 0, st, (float2*)s->d, s->n, lc,
                       sz, gamma_dt, d_decay, d_partial, d_uni, uni_stride, d_jumps, jump_stride,
                       d_fault, n_steps, step);
  }
  SMOQ_CHECK_LAUNCH("k_jump_apply");
  return SMOQ_OK;
}

int mcwf_jump_check(const smoq_state* s, int kind, double gamma_dt) {
  int rc = check_state(s);
  if (rc) return rc;
  if (kind != SMOQ_JUMP_SIGMA_MINUS && kind != SMOQ_JUMP_SIGMA_Z)
    return set_error(SMOQ_EINVAL, "mcwf: unknown jump kind %d", kind);
  if (s->n > kJumpMaxQ)
    return set_error(SMOQ_ENOTSUP, "mcwf: %d qubits per trajectory (max %d)", s->n, kJumpMaxQ);
  if (!(gamma_dt >= 0.0 && gamma_dt <= 1.0))
    return set_error(SMOQ_EINVAL, "mcwf: gamma_dt %g outside [0, 1]", gamma_dt);
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_mcwf_evolve(smoq_ctx* ctx, const smoq_mcwf_group* groups, int n_groups,
                                double gamma_dt, int n_steps, int obs_every) {
  if (!ctx || !groups || n_groups < 1)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: null context/groups or n_groups < 1");
  if (n_steps < 0 || obs_every < 0)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: n_steps %d obs_every %d", n_steps,
                     obs_every);
  if (!(gamma_dt >= 0.0 && gamma_dt <= 1.0))
    return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: gamma_dt %g outside [0, 1]", gamma_dt);
  int dtype = -1;
  size_t total_traj = 0;
  for (int g = 0; g < n_groups; ++g) {
    const smoq_mcwf_group& G = groups[g];
    int rc = check_state(G.psi);
    if (rc) return rc;
    if (dtype < 0) dtype = G.psi->dtype;
    if (G.psi->ctx != ctx || G.psi->dtype != dtype)
      return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: groups must share context and dtype");
    const int n = G.psi->n;
    if (n > kMcMaxQ)
      return set_error(SMOQ_ENOTSUP, "smoq_mcwf_evolve: group %d has %d qubits (max %d)", g, n,
                       kMcMaxQ);
    if (G.n_bonds < 0 || (G.n_bonds > 0 && (!G.bond_q || !G.bond_diag)) || !G.site_u ||
        !G.decay_pow || (n_steps > 0 && (!G.uniforms || !G.jumps)))
      return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: group %d has a null array", g);
    for (int b = 0; b < G.n_bonds; ++b) {
      const int a = G.bond_q[2 * b], c = G.bond_q[2 * b + 1];
      if (a < 0 || a >= n || c < 0 || c >= n || a == c)
        return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: group %d bond %d on qubits (%d, %d)", g,
                         b, a, c);
    }
    total_traj += (size_t)G.psi->batch;
  }
  if (n_steps == 0) return SMOQ_OK;
  if (total_traj > 0x7fffffff)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_evolve: %zu trajectories", total_traj);
  const int rec_n = obs_every > 0 ? n_steps / obs_every : 0;

  // one device buffer: tables, descriptors, uniforms, outputs
  std::vector<unsigned char> blob;
  auto reserve = [&](size_t bytes) {
    const size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + bytes);
    return off;
  };
  auto put = [&](const void* p, size_t bytes) {
    const size_t off = reserve(bytes);
    if (bytes) std::memcpy(blob.data() + off, p, bytes);
    return off;
  };
  struct Off {
    size_t bond_q, bond_diag, site_u, decay, chain, uni, jumps, z;
    bool chain_ok;
  };
  std::vector<Off> offs(n_groups);
  size_t device_bytes = 0;  // uniforms and outputs, placed after the blob
  for (int g = 0; g < n_groups; ++g) {
    const smoq_mcwf_group& G = groups[g];
    const int n = G.psi->n;
    Off& o = offs[g];
    o.bond_q = put(G.bond_q, sizeof(int32_t) * 2 * G.n_bonds);
    o.bond_diag = put(G.bond_diag, sizeof(double) * 8 * G.n_bonds);
    o.site_u = put(G.site_u, sizeof(double) * 8 * n);
    o.decay = put(G.decay_pow, sizeof(double) * (n + 1));
    // open chain (q, q+1), q = 0..n-2, every bond diagonal (s, d, d, s) with
    // the same s and d: the product of the bond phases depends only on the
    // number of unequal neighbours
    bool chain = n >= 2 && G.n_bonds == n - 1;
    for (int b = 0; chain && b < G.n_bonds; ++b) {
      chain = G.bond_q[2 * b] == b && G.bond_q[2 * b + 1] == b + 1;
      for (int e = 0; chain && e < 8; ++e) chain = G.bond_diag[8 * b + e] == G.bond_diag[e];
    }
    if (chain)
      chain = G.bond_diag[0] == G.bond_diag[6] && G.bond_diag[1] == G.bond_diag[7] &&
              G.bond_diag[2] == G.bond_diag[4] && G.bond_diag[3] == G.bond_diag[5];
    std::vector<double> tab(2 * (size_t)n + 2, 0.0);
    if (chain) {
      for (int nd = 0; nd < n; ++nd) {
        double re = 1.0, im = 0.0;
        for (int b = 0; b < n - 1; ++b) {
          const double* f = b < n - 1 - nd ? &G.bond_diag[0] : &G.bond_diag[2];
          const double r2 = re * f[0] - im * f[1], i2 = re * f[1] + im * f[0];
          re = r2;
          im = i2;
        }
        tab[2 * nd] = re;
        tab[2 * nd + 1] = im;
      }
    }
    o.chain = put(tab.data(), sizeof(double) * tab.size());
    o.chain_ok = chain;
    const size_t T = (size_t)G.psi->batch;
    o.uni = device_bytes;
    device_bytes += (T * n_steps * sizeof(double) + 255) & ~(size_t)255;
    o.jumps = device_bytes;
    device_bytes += (T * n_steps * sizeof(int32_t) + 255) & ~(size_t)255;
    o.z = device_bytes;
    if (G.z_out && rec_n > 0)
      device_bytes += (T * rec_n * n * sizeof(double) + 255) & ~(size_t)255;
  }
  const size_t grp_off = reserve(sizeof(McwfGroupDev) * n_groups);
  const size_t trj_off = reserve(sizeof(McwfTrajDev) * total_traj);
  const size_t fault_off = reserve(sizeof(unsigned long long));
  const size_t dyn_off = (blob.size() + 255) & ~(size_t)255;

  unsigned char* dev = nullptr;
  SMOQ_HIP(hipMalloc(&dev, dyn_off + device_bytes));
  auto done = [&](int code) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dev);
    return code;
  };

  std::vector<McwfGroupDev> gd(n_groups);
  std::vector<std::pair<int, int>> order;  // (n, group): largest states start first
  for (int g = 0; g < n_groups; ++g) {
    const smoq_mcwf_group& G = groups[g];
    const Off& o = offs[g];
    McwfGroupDev& d = gd[g];
    d.psi = G.psi->d;
    d.n = G.psi->n;
    d.n_bonds = G.n_bonds;
    d.uniform_chain = o.chain_ok ? 1 : 0;
    d.site_rx = 1;
    for (int q = 0; q < d.n && d.site_rx; ++q) {
      const double* u = G.site_u + 8 * q;  // (re, im) of u00, u01, u10, u11
      d.site_rx = u[1] == 0.0 && u[2] == 0.0 && u[4] == 0.0 && u[7] == 0.0;
    }
    d.bond_q = (const int32_t*)(dev + o.bond_q);
    d.bond_diag = (const double2*)(dev + o.bond_diag);
    d.site_u = (const double2*)(dev + o.site_u);
    d.decay_pow = (const double*)(dev + o.decay);
    d.chain_tab = (const double2*)(dev + o.chain);
    d.uniforms = (const double*)(dev + dyn_off + o.uni);
    d.jumps = (int32_t*)(dev + dyn_off + o.jumps);
    d.z_out = (G.z_out && rec_n > 0) ? (double*)(dev + dyn_off + o.z) : nullptr;
    order.push_back({d.n, g});
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                     return a.first > b.first;
                   });
  std::vector<McwfTrajDev> trajs;
  trajs.reserve(total_traj);
  for (const auto& og : order)
    for (int i = 0; i < groups[og.second].psi->batch; ++i) trajs.push_back({og.second, i});
  std::memcpy(blob.data() + grp_off, gd.data(), sizeof(McwfGroupDev) * n_groups);
  std::memcpy(blob.data() + trj_off, trajs.data(), sizeof(McwfTrajDev) * total_traj);
  const unsigned long long no_fault = ~0ull;
  std::memcpy(blob.data() + fault_off, &no_fault, sizeof(no_fault));

  hipError_t e =
      hipMemcpyAsync(dev, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream);
  for (int g = 0; g < n_groups && e == hipSuccess; ++g) {
    const size_t T = (size_t)groups[g].psi->batch;
    e = hipMemcpyAsync(dev + dyn_off + offs[g].uni, groups[g].uniforms,
                       T * n_steps * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  }
  if (e != hipSuccess) return done(hip_error(e, "smoq_mcwf_evolve: upload"));
  McwfArgs args{(const McwfGroupDev*)(dev + grp_off), (const McwfTrajDev*)(dev + trj_off),
                gamma_dt, n_steps, obs_every, (unsigned long long*)(dev + fault_off)};
  const size_t shm = (dtype == SMOQ_C128 ? sizeof(double2) : sizeof(float2)) << kMcTileQ;
  if (dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_mcwf<double>, dim3((unsigned)total_traj), dim3(kMcBlock), shm,
                       ctx->stream, args);
  else
    hipLaunchKernelGGL(k_mcwf<float>, dim3((unsigned)total_traj), dim3(kMcBlock), shm,
                       ctx->stream, args);
  e = hipGetLastError();
  if (e != hipSuccess) return done(hip_error(e, "k_mcwf launch"));
  for (int g = 0; g < n_groups && e == hipSuccess; ++g) {
    const smoq_mcwf_group& G = groups[g];
    const size_t T = (size_t)G.psi->batch;
    e = hipMemcpyAsync(G.jumps, dev + dyn_off + offs[g].jumps, T * n_steps * sizeof(int32_t),
                       hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && gd[g].z_out)
      e = hipMemcpyAsync(G.z_out, gd[g].z_out, T * rec_n * G.psi->n * sizeof(double),
                         hipMemcpyDeviceToHost, ctx->stream);
  }
  unsigned long long fault = no_fault;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&fault, dev + fault_off, sizeof(fault), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e != hipSuccess) return done(hip_error(e, "smoq_mcwf_evolve: download"));
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return done(hip_error(e, "k_mcwf"));
  if (fault != no_fault) {
    const McwfTrajDev& t = trajs[fault / (unsigned long long)n_steps];
    return done(set_error(SMOQ_ENUMERIC,
                          "smoq_mcwf_evolve: zero or non-finite norm at step %d of trajectory %d "
                          "in group %d (%d qubits); that trajectory stopped there",
                          (int)(fault % (unsigned long long)n_steps), t.idx, t.group,
                          groups[t.group].psi->n));
  }
  return done(SMOQ_OK);
}

extern "C" int smoq_mcwf_jump(smoq_state* s, int kind, double gamma_dt, const double* decay_pow,
                              const double* uniforms, int32_t* jumps) {
  int rc = mcwf_jump_check(s, kind, gamma_dt);
  if (rc) return rc;
  if (!decay_pow || !uniforms || !jumps)
    return set_error(SMOQ_EINVAL, "smoq_mcwf_jump: null array");
  smoq_ctx* ctx = s->ctx;
  const size_t nb = (size_t)s->batch;
  const size_t off_u = ((size_t)(s->n + 1) * sizeof(double) + 255) & ~(size_t)255;
  const size_t off_j = off_u + ((2 * nb * sizeof(double) + 255) & ~(size_t)255);
  const size_t off_f = off_j + ((nb * sizeof(int32_t) + 255) & ~(size_t)255);
  const size_t off_p = off_f + 256;
  unsigned char* dev = nullptr;
  SMOQ_HIP(hipMalloc(&dev, off_p + mcwf_jump_partial_bytes(s)));
  const unsigned long long no_fault = ~0ull;
  unsigned long long fault = no_fault;
  hipError_t e = hipMemcpyAsync(dev, decay_pow, (size_t)(s->n + 1) * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev + off_u, uniforms, 2 * nb * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev + off_f, &no_fault, sizeof(no_fault), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    rc = mcwf_jump_launch(s, kind, gamma_dt, (const double*)dev, (const double*)(dev + off_u), 2,
                          (int32_t*)(dev + off_j), 1, (unsigned long long*)(dev + off_f), 1, 0,
                          (double*)(dev + off_p));
  if (e == hipSuccess && rc == SMOQ_OK)
    e = hipMemcpyAsync(jumps, dev + off_j, nb * sizeof(int32_t), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess && rc == SMOQ_OK)
    e = hipMemcpyAsync(&fault, dev + off_f, sizeof(fault), hipMemcpyDeviceToHost, ctx->stream);
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (rc) return rc;
  if (e != hipSuccess) return hip_error(e, "smoq_mcwf_jump: copy");
  if (se != hipSuccess) return hip_error(se, "smoq_mcwf_jump");
  if (fault != no_fault)
    return set_error(SMOQ_ENUMERIC, "smoq_mcwf_jump: zero or non-finite norm in trajectory %llu",
                     fault);
  return SMOQ_OK;
}
