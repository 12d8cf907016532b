// Adjoint gradients of many parameter sets of different qubit counts in one
// persistent launch (smoq_adjoint_gradient_batch).
//
// Every item (qubit count n, its tape, angles and Hamiltonian) runs the same
// steps as the single-state fused adjoint: |0…0>, the forward passes over
// ψ, W = (ψ, Hψ) interleaved, the reverse passes over W with their overlap
// gates, the per-slot sums.  One workgroup owns an item from start to end:
// it walks every tile of every pass itself, with workgroup barriers between
// the steps, so workgroups never wait for one another and the launch needs
// no cross-workgroup signalling.  The host deals the items to the
// workgroups (largest first, each to the least loaded), and each workgroup
// runs its list in order.
//
// Passes run the precompiled pass code (gate and stage records read at run
// time) with overlap gates added, so items of any structure share the one
// kernel; tiles of 2^12 amplitudes by default (2^n for smaller items, whose
// idle threads only join the barriers).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "adjoint_common.h"
#include "fused_ops.h"

namespace smoq {
namespace {

constexpr int kVqeTile = 12;
// device memory for one batch of large items with the same tape (ψ plus the
// two-state W of each item)
constexpr uint64_t kVqeGroupBytes = 8ull << 30;

// Byte offsets from the launch's base pointers (tables; states) rather than
// pointers: addresses formed from a kernel argument are known to be global,
// so the records are read with scalar and global loads, not flat ones.
struct VqeProgDev {
  uint64_t passes, stages, gates, tabs;
};

struct VqeItemDev {
  uint64_t psi;         // state area: 2^n amplitudes
  uint64_t w;           // state area: 2^(n+1), W[2k] = ψ_k, W[2k+1] = λ_k
  VqeProgDev fwd, rev;  // table area
  uint64_t terms;       // table area: AdjTerm[n_terms]
  uint64_t opart;       // table area: overlap partials [slot][reverse tile]
  uint64_t epart;       // table area: energy partials per chunk
  uint64_t sums;        // table area: n_slots slot sums, then the energy
  int32_t n, n_terms, n_slots, kf;
  int32_t n_fwd, n_rev; // passes of the two programs
  uint32_t chunks;      // 2^(n - kf): chunks of W's construction
  uint32_t rtiles;      // tiles of a reverse pass
};

struct VqeProgPtr {
  const PassDev* passes;
  const StageDev* stages;
  const GateDev* gates;
  const double2* tabs;
};

__device__ __forceinline__ VqeProgPtr prog_at(const unsigned char* tb, const VqeProgDev& d) {
  return {(const PassDev*)(tb + d.passes), (const StageDev*)(tb + d.stages),
          (const GateDev*)(tb + d.gates), (const double2*)(tb + d.tabs)};
}

// Σ over the thread's ψ patterns of Im[conj(λ_k) (Gψ)_k] (the per-thread
// part of the compiled passes' overlap gate, with the phase power, z masks
// and controls read at run time)
template <int XR, int SEL>
__device__ __forceinline__ double ovl_sum(const VT<double> (&x)[kNR], const GateDev& g,
                                          uint64_t gl) {
  constexpr int sm = 1 << SEL;
  const int zr = g.ozreg;
  const int zp = __popcll(gl & g.ozglob) & 1;
  const bool ctrl = (g.creg | g.cglob) != 0;
  const bool gon = (gl & g.cglob) == g.cglob;
  const int cnt = g.ocnt;
  double ov = 0.0;
#pragma unroll
  for (int j = 0; j < kNR; ++j) {
    if (j & sm) continue;
    const int jp = j ^ XR;
    const VT<double> l = x[j | sm];
    const VT<double> w = x[jp];
    double ux = (cnt & 1) ? -w.y : w.x;
    double uy = (cnt & 1) ? w.x : w.y;
    if (cnt & 2) {
      ux = -ux;
      uy = -uy;
    }
    double c = l.x * uy - l.y * ux;
    if ((zp ^ __popc(jp & zr)) & 1) c = -c;
    if (ctrl) c = (gon && (j & g.creg) == g.creg) ? c : 0.0;
    ov += c;
  }
  return ov;
}

__device__ __forceinline__ double ovl_any(const VT<double> (&x)[kNR], const GateDev& g,
                                          uint64_t gl) {
  const int xr = (g.a >= 0 ? 1 << g.a : 0) | (g.b >= 0 ? 1 << g.b : 0);
#define SMOQ_OVL_CASE(XR)                                   \
  case XR * 3 + 0: return ovl_sum<XR, 0>(x, g, gl);         \
  case XR * 3 + 1: return ovl_sum<XR, 1>(x, g, gl);         \
  case XR * 3 + 2: return ovl_sum<XR, 2>(x, g, gl);
  switch (xr * 3 + g.osel) {
    SMOQ_OVL_CASE(0)
    SMOQ_OVL_CASE(1)
    SMOQ_OVL_CASE(2)
    SMOQ_OVL_CASE(3)
    SMOQ_OVL_CASE(4)
    SMOQ_OVL_CASE(5)
    SMOQ_OVL_CASE(6)
    SMOQ_OVL_CASE(7)
    default: return 0.0;
  }
#undef SMOQ_OVL_CASE
}

// One tile of one pass: the precompiled pass kernel's body with an explicit
// tile index, threads beyond the tile's 2^(k-R) idle (they join every
// barrier and add zero to the overlaps).  ovl holds one double per (overlap
// gate of the pass, wave).
template <typename T>
__device__ __forceinline__ void vqe_pass_tile(VT<T>* lds, double* ovl, VT<T>* __restrict__ psi,
                              const PassDev* __restrict__ pass, const VqeProgPtr& P,
                              uint64_t tile, double* __restrict__ opart, uint64_t ntiles) {
  constexpr int R = kRegQubits;
  constexpr int WB = CT<T>::kRowBits;
  const int nt = 1 << (pass->k - R);
  const bool act = (int)threadIdx.x < nt;
  const uint32_t tid = act ? threadIdx.x : 0u;
  const uint64_t gbase = pdep64(tile, pass->nl_mask);
  VT<T>* pt = psi + (gbase | pdep64((uint64_t)tid, pass->tid_mask));
  uint64_t off[kNR];
#pragma unroll
  for (int i = 0; i < kNR; ++i) {
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) o |= ((i >> j) & 1) ? pass->hb[j] : 0ull;
    off[i] = o;
  }
  const uint32_t st_lane = swz<WB>(tid);
  if (act) {
    VT<T> v[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) v[i] = pt[off[i]];
#pragma unroll
    for (int i = 0; i < kNR; ++i) lds[st_lane ^ swz<WB>((uint32_t)(i * nt))] = v[i];
  }
  __syncthreads();
  const int nwaves = (int)(blockDim.x >> 6);
  int novl = 0;
  for (int s = pass->s0; s < pass->s1; ++s) {
    const StageDev& st = P.stages[s];
    VT<double> x[kNR];
    StageFrame sf;
    if (act) {
      stage_load<T>(lds, st, gbase, x, sf);
    } else {
#pragma unroll
      for (int i = 0; i < kNR; ++i) x[i] = make_c<double>(0.0, 0.0);
      sf.gl = 0;
    }
    for (int g = st.g0; g < st.g1; ++g) {
      const GateDev& G = P.gates[g];
      if (G.kind == GK_OVL) {
        double v = act ? ovl_any(x, G, sf.gl) : 0.0;
        v = wave_total(v);
        if ((threadIdx.x & 63) == 0) ovl[novl * nwaves + (threadIdx.x >> 6)] = v;
        ++novl;
      } else if (act) {
        apply_gate<double>(x, G, sf.gl, P.tabs);
      }
    }
    if (act) {
#pragma unroll
      for (int i = 0; i < kNR; ++i) lds_at<T>(lds, sf.sl ^ sf.sr[i]) = make_c<T>(x[i].x, x[i].y);
    }
    __syncthreads();
  }
  if (act) {
    VT<T> v[kNR];
#pragma unroll
    for (int i = 0; i < kNR; ++i) v[i] = lds[st_lane ^ swz<WB>((uint32_t)(i * nt))];
#pragma unroll
    for (int i = 0; i < kNR; ++i) pt[off[i]] = v[i];
  }
  // the tile's overlaps, waves added in order, one partial per (slot, tile)
  for (int t = threadIdx.x; t < novl; t += blockDim.x) {
    int c = 0, slot = -1;
    for (int s = pass->s0; s < pass->s1 && slot < 0; ++s)
      for (int g = P.stages[s].g0; g < P.stages[s].g1; ++g)
        if (P.gates[g].kind == GK_OVL) {
          if (c == t) {
            slot = P.gates[g].oslot;
            break;
          }
          ++c;
        }
    double v = 0.0;
    for (int w = 0; w < nwaves; ++w) v += ovl[t * nwaves + w];
    opart[(uint64_t)slot * ntiles + tile] = v;
  }
}

// W for one chunk of ψ indices and the chunk's part of E = Re<ψ|λ> (λ
// rounded to the state's type first, as it is stored); the pull form of
// apply_hamiltonian_bitwise!, as in the single-state sweep
template <typename T>
__device__ __forceinline__ void vqe_init_chunk(const unsigned char* tro, unsigned char* tb,
                                               unsigned char* sb, const VqeItemDev& it,
                                               uint64_t chunk, double* red) {
#pragma clang fp contract(off)
  const VT<T>* p = (const VT<T>*)(sb + it.psi);
  VT<T>* w = (VT<T>*)(sb + it.w);
  const AdjTerm* terms = (const AdjTerm*)(tro + it.terms);
  const uint64_t len = 1ull << it.kf;
  const uint64_t k0 = chunk * len;
  double acc = 0.0;
  for (uint64_t e = threadIdx.x; e < len; e += blockDim.x) {
    const uint64_t k = k0 + e;
    double hr = 0.0, hi = 0.0;
    for (int t = 0; t < it.n_terms; ++t) {
      const AdjTerm tm = terms[t];
      const uint64_t j = k ^ tm.x;
      const VT<T> v = p[j];
      const double vr = (double)v.x, vi = (double)v.y;
      const bool neg = __popcll(j & tm.z) & 1;
      const double sr = neg ? -vr : vr, si = neg ? -vi : vi;
      hr += tm.fre * sr - tm.fim * si;
      hi += tm.fre * si + tm.fim * sr;
    }
    const VT<T> a = p[k];
    const VT<T> l = make_c<T>(hr, hi);
    w[2 * k] = a;
    w[2 * k + 1] = l;
    acc += (double)a.x * (double)l.x + (double)a.y * (double)l.y;
  }
  acc = wave_total(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) v += red[i];
    ((double*)(tb + it.epart))[chunk] = v;
  }
}

// one item from |0…0> to its slot sums and energy, by the whole workgroup
template <typename T>
__device__ __forceinline__ void vqe_item(const unsigned char* tro, unsigned char* tb,
                                         unsigned char* sb, const VqeItemDev& it, VT<T>* lds,
                                         double* ovl) {
  VT<T>* psi = (VT<T>*)(sb + it.psi);
  const uint64_t dim = 1ull << it.n;
  for (uint64_t e = threadIdx.x; e < dim; e += blockDim.x) psi[e] = make_c<T>(e == 0 ? 1.0 : 0.0, 0.0);
  __syncthreads();
  const VqeProgPtr F = prog_at(tro, it.fwd);
  for (int p = 0; p < it.n_fwd; ++p) {
    const PassDev* pass = F.passes + p;
    const uint64_t ntiles = 1ull << pass->n_nonlocal;
    for (uint64_t tile = 0; tile < ntiles; ++tile) {
      vqe_pass_tile<T>(lds, ovl, psi, pass, F, tile, nullptr, 0);
      __syncthreads();
    }
  }
  for (uint32_t c = 0; c < it.chunks; ++c) {
    vqe_init_chunk<T>(tro, tb, sb, it, c, ovl);
    __syncthreads();
  }
  const VqeProgPtr R = prog_at(tro, it.rev);
  VT<T>* w = (VT<T>*)(sb + it.w);
  double* opart = (double*)(tb + it.opart);
  for (int p = 0; p < it.n_rev; ++p) {
    const PassDev* pass = R.passes + p;
    for (uint64_t tile = 0; tile < it.rtiles; ++tile) {
      vqe_pass_tile<T>(lds, ovl, w, pass, R, tile, opart, it.rtiles);
      __syncthreads();
    }
  }
  // slot sums in tile order, the energy from the chunk partials
  const double* epart = (const double*)(tb + it.epart);
  double* sums = (double*)(tb + it.sums);
  for (int r = threadIdx.x; r <= it.n_slots; r += blockDim.x) {
    double v = 0.0;
    if (r < it.n_slots) {
      for (uint32_t j = 0; j < it.rtiles; ++j) v += opart[(uint64_t)r * it.rtiles + j];
    } else {
      for (uint32_t j = 0; j < it.chunks; ++j) v += epart[j];
    }
    sums[r] = v;
  }
  __syncthreads();
}

// workgroup b runs items list[first[b] .. first[b+1]-1] in order
template <typename T>
__global__ void __launch_bounds__(kMaxPassThreads)
    k_vqe_batch(const unsigned char* __restrict__ tro, unsigned char* __restrict__ tb,
                unsigned char* __restrict__ sb, const VqeItemDev* __restrict__ items,
                const int* __restrict__ first, const int* __restrict__ list, int tile_elems) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  VT<T>* lds = reinterpret_cast<VT<T>*>(smem_raw);
  double* ovl = reinterpret_cast<double*>(smem_raw + (size_t)tile_elems * sizeof(VT<T>));
  const int k1 = first[blockIdx.x + 1];
  for (int k = first[blockIdx.x]; k < k1; ++k) vqe_item<T>(tro, tb, sb, items[list[k]], lds, ovl);
}

struct ItemPlan {
  int n = 0, kf = 0, n_slots = 0;
  std::vector<Step> steps;
  std::vector<AdjTerm> terms;
  Program fwd, rev;
  ProgLayout lf, lr;
  int rc = SMOQ_OK;
  std::string err;
};

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Enqueues the one-launch part on stream st: plans, uploads (waiting for the
// upload, so the context's pinned staging buffer is free again on return),
// launches, and queues the copies of every item's sums into the context's
// pinned result buffer; finish_batch waits for them and fills the outputs.
template <typename T>
int run_batch(smoq_ctx* ctx, int dtype, const smoq_vqe_item* items, int n_items, hipStream_t st,
              std::vector<ItemPlan>& plans) {
  using V = VT<T>;
  const int rowbits = CT<T>::kRowBits;
  const int tile_req = ctx->tile_qubits > 0 ? std::min(ctx->tile_qubits, kMaxTileQubits) : kVqeTile;
  // plan every item (host threads over the items)
  plans.assign(n_items, ItemPlan());
  auto plan_one = [&](int i) {
    const smoq_vqe_item& it = items[i];
    ItemPlan& P = plans[i];
    P.n = it.n_qubits;
    int rc = build_steps(it.ops, it.n_ops, it.n_params, P.n, P.steps);
    std::vector<IGate> gs;
    if (!rc) rc = adj_terms(it.n_terms, it.xmask, it.zmask, it.coeff, P.n, P.terms);
    if (!rc) rc = forward_gates(P.steps, it.theta, P.n, gs);
    if (!rc) {
      P.kf = std::min(tile_req, P.n);
      plan_program(gs, P.n, P.kf, ctx->low_qubits, rowbits, P.fwd, ctx->pass_max_gates, false);
      rc = reverse_gates(P.steps, it.theta, P.n, gs, &P.n_slots);
    }
    if (!rc)
      plan_program(gs, P.n + 1, std::min(tile_req, P.n + 1), ctx->low_qubits, rowbits, P.rev,
                   ctx->pass_max_gates, false);
    if (rc) P.err = smoq_last_error();
    P.rc = rc;
    P.lf = program_layout(P.fwd, dtype);
    P.lr = program_layout(P.rev, dtype);
  };
  const int nth = std::max(1, std::min<int>(n_items, (int)std::min(16u,
                                            std::max(1u, std::thread::hardware_concurrency()))));
  if (nth == 1) {
    for (int i = 0; i < n_items; ++i) plan_one(i);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
      pool.emplace_back([&, t] {
        for (int i = t; i < n_items; i += nth) plan_one(i);
      });
    for (auto& th : pool) th.join();
  }
  for (int i = 0; i < n_items; ++i)
    if (plans[i].rc)
      return set_error(plans[i].rc, "smoq_adjoint_gradient_batch: item %d: %s", i,
                       plans[i].err.c_str());
  // device layout: [states] in work slot 4; [items | segments | counters |
  // per item: programs, terms, partials, sums] in work slot 5
  std::vector<size_t> off_psi(n_items), off_w(n_items);
  size_t state_bytes = 0;
  for (int i = 0; i < n_items; ++i) {
    off_psi[i] = state_bytes;
    state_bytes += up256(sizeof(V) << plans[i].n);
    off_w[i] = state_bytes;
    state_bytes += up256(sizeof(V) << (plans[i].n + 1));
  }
  // block size: the largest tile of any pass (at least one wave)
  int block = 64;
  for (const ItemPlan& P : plans) {
    for (const PassDev& pd : P.fwd.passes) block = std::max(block, 1 << (pd.k - kRegQubits));
    for (const PassDev& pd : P.rev.passes) block = std::max(block, 1 << (pd.k - kRegQubits));
  }
  // overlap gates of one pass, for the LDS overlap area
  int max_ovl = 1;
  for (const ItemPlan& P : plans)
    for (const PassDev& pd : P.rev.passes) {
      int c = 0;
      for (int s = pd.s0; s < pd.s1; ++s)
        for (int g = P.rev.stages[s].g0; g < P.rev.stages[s].g1; ++g)
          c += P.rev.gates[g].kind == GK_OVL;
      max_ovl = std::max(max_ovl, c);
    }
  const int tile_elems = 8 * block;
  const size_t shm = (size_t)tile_elems * sizeof(V) + sizeof(double) * (size_t)max_ovl * (block / 64);
  if (shm > 160 * 1024)
    return set_error(SMOQ_ENOTSUP, "smoq_adjoint_gradient_batch: %zu B of LDS for %d overlaps in "
                                   "one pass", shm, max_ovl);
  int per_cu = 0, n_cu = 0;
  SMOQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_vqe_batch<T>, block, shm));
  SMOQ_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(n_items, (int64_t)std::max(1, per_cu) * std::max(1, n_cu)));
  // deal the items: largest cost first, each to the least loaded workgroup
  // (cost ~ amplitudes x passes)
  std::vector<int> order(n_items);
  std::vector<double> cost(n_items);
  for (int i = 0; i < n_items; ++i) {
    order[i] = i;
    cost[i] = std::ldexp(1.0, plans[i].n) *
              (1.0 + plans[i].fwd.passes.size() + 2.0 * (plans[i].rev.passes.size() + 1));
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<std::vector<int>> deal(grid);
  std::vector<double> load(grid, 0.0);
  for (int i : order) {
    const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    deal[w].push_back(i);
    load[w] += cost[i];
  }
  std::vector<int> first(grid + 1, 0), list;
  for (int w = 0; w < grid; ++w) {
    first[w] = (int)list.size();
    list.insert(list.end(), deal[w].begin(), deal[w].end());
  }
  first[grid] = (int)list.size();
  // table area
  size_t tb = 0;
  const size_t off_items = tb;
  tb += up256(sizeof(VqeItemDev) * n_items);
  const size_t off_first = tb;
  tb += up256(sizeof(int) * first.size());
  const size_t off_list = tb;
  tb += up256(sizeof(int) * list.size());
  std::vector<size_t> off_f(n_items), off_r(n_items), off_t(n_items), off_o(n_items),
      off_e(n_items), off_s(n_items);
  for (int i = 0; i < n_items; ++i) {
    const ItemPlan& P = plans[i];
    off_f[i] = tb;
    tb += up256(P.lf.total);
    off_r[i] = tb;
    tb += up256(P.lr.total);
    off_t[i] = tb;
    tb += up256(sizeof(AdjTerm) * (P.terms.size() + 1));
    off_o[i] = tb;
    const uint64_t rtiles = P.rev.passes.empty() ? 1 : (1ull << P.rev.passes[0].n_nonlocal);
    tb += up256(sizeof(double) * (size_t)P.n_slots * rtiles + 8);
    off_e[i] = tb;
    tb += up256(sizeof(double) << (P.n - P.kf));
    off_s[i] = tb;
    tb += up256(sizeof(double) * (P.n_slots + 1));
  }
  void* sbuf = nullptr;
  void* tbuf = nullptr;
  int rc = ensure_work(ctx, 4, state_bytes, &sbuf);
  if (!rc) rc = ensure_work(ctx, 5, tb, &tbuf);
  if (!rc) rc = ensure_pinned(ctx, tb);
  if (rc) return rc;
  // the pinned buffer may feed an earlier copy on either stream
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(st));
  unsigned char* sd = (unsigned char*)sbuf;
  unsigned char* td = (unsigned char*)tbuf;
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  std::vector<VqeItemDev> dev(n_items);
  for (int i = 0; i < n_items; ++i) {
    const ItemPlan& P = plans[i];
    auto put_prog = [&](const Program& prog, const ProgLayout& L, size_t off, VqeProgDev& pd) {
      std::memcpy(h + off, prog.passes.data(), L.pb);
      std::memcpy(h + off + L.pb, prog.stages.data(), L.sb);
      std::memcpy(h + off + L.pb + L.sb, prog.gates.data(), L.gb);
      if (L.ntab) std::memcpy(h + off + L.off_t, prog.tabs.data(), L.ntab * sizeof(double2));
      pd.passes = off;
      pd.stages = off + L.pb;
      pd.gates = off + L.pb + L.sb;
      pd.tabs = off + L.off_t;
    };
    VqeItemDev& d = dev[i];
    put_prog(P.fwd, P.lf, off_f[i], d.fwd);
    put_prog(P.rev, P.lr, off_r[i], d.rev);
    if (!P.terms.empty())
      std::memcpy(h + off_t[i], P.terms.data(), sizeof(AdjTerm) * P.terms.size());
    d.psi = off_psi[i];
    d.w = off_w[i];
    d.terms = off_t[i];
    d.opart = off_o[i];
    d.epart = off_e[i];
    d.sums = off_s[i];
    d.n = P.n;
    d.n_terms = (int)P.terms.size();
    d.n_slots = P.n_slots;
    d.kf = P.kf;
    d.n_fwd = (int)P.fwd.passes.size();
    d.n_rev = (int)P.rev.passes.size();
    d.chunks = (uint32_t)(1u << (P.n - P.kf));
    d.rtiles = P.rev.passes.empty() ? 1u : (uint32_t)(1u << P.rev.passes[0].n_nonlocal);
  }
  std::memcpy(h + off_items, dev.data(), sizeof(VqeItemDev) * n_items);
  std::memcpy(h + off_first, first.data(), sizeof(int) * first.size());
  std::memcpy(h + off_list, list.data(), sizeof(int) * list.size());
  SMOQ_HIP(hipMemcpyAsync(td, h, tb, hipMemcpyHostToDevice, st));
  SMOQ_HIP(hipStreamSynchronize(st));
  // the table area is passed twice: read-only (programs, terms, item
  // records, lists: never written by the kernel, so scalar loads may read
  // them) and writable (the partials and sums, in other byte ranges)
  hipLaunchKernelGGL(k_vqe_batch<T>, dim3((unsigned)grid), dim3(block), shm, st,
                     (const unsigned char*)td, td, sd, (const VqeItemDev*)(td + off_items),
                     (const int*)(td + off_first), (const int*)(td + off_list), tile_elems);
  SMOQ_CHECK_LAUNCH("k_vqe_batch");
  // every item's sums, into the pinned result buffer
  size_t res_n = 0;
  for (const ItemPlan& P : plans) res_n += P.n_slots + 1;
  if (res_n * sizeof(double) > ctx->res_pinned_bytes) {
    if (ctx->res_pinned) SMOQ_HIP(hipHostFree(ctx->res_pinned));
    ctx->res_pinned = nullptr;
    ctx->res_pinned_bytes = 0;
    SMOQ_HIP(hipHostMalloc(&ctx->res_pinned, res_n * sizeof(double), hipHostMallocDefault));
    ctx->res_pinned_bytes = res_n * sizeof(double);
  }
  double* res = (double*)ctx->res_pinned;
  size_t r = 0;
  for (int i = 0; i < n_items; ++i) {
    const size_t cnt = plans[i].n_slots + 1;
    SMOQ_HIP(hipMemcpyAsync(res + r, td + off_s[i], cnt * sizeof(double), hipMemcpyDeviceToHost,
                            st));
    r += cnt;
  }
  return SMOQ_OK;
}

int finish_batch(hipStream_t st, const double* res, const smoq_vqe_item* items, int n_items,
                 const std::vector<ItemPlan>& plans) {
  SMOQ_HIP(hipStreamSynchronize(st));
  size_t r = 0;
  for (int i = 0; i < n_items; ++i) {
    const ItemPlan& P = plans[i];
    grad_from_slots(P.steps, res + r, P.n_slots, items[i].n_params, items[i].grad);
    *items[i].energy = res[r + P.n_slots];
    r += P.n_slots + 1;
  }
  return SMOQ_OK;
}

// items that differ only in their angles (same qubit count, tape and
// Hamiltonian, compared by contents)
static bool same_problem(const smoq_vqe_item& a, const smoq_vqe_item& b) {
  auto eq = [](const void* x, const void* y, size_t bytes) {
    return x == y || (x && y && std::memcmp(x, y, bytes) == 0);
  };
  return a.n_qubits == b.n_qubits && a.n_ops == b.n_ops && a.n_params == b.n_params &&
         a.n_terms == b.n_terms && eq(a.ops, b.ops, sizeof(smoq_op) * a.n_ops) &&
         eq(a.xmask, b.xmask, sizeof(uint64_t) * a.n_terms) &&
         eq(a.zmask, b.zmask, sizeof(uint64_t) * a.n_terms) &&
         eq(a.coeff, b.coeff, sizeof(double) * a.n_terms);
}

// Which items take the one launch and which the whole-device sweep one after
// another: a workgroup moves an item's 2^n amplitudes through its passes
// alone (about 8.6e-10 s per amplitude and tape op, measured at n = 20),
// while the whole-device sweep costs about 1.8e-12 s per amplitude and op
// with a floor of about 6e-7 s per op (launches).  The two run side by side
// on two streams, so the most expensive items move to the sweep, as many as
// make the longer of the two shortest.  SMOQ_VQE_ROUTE=one_launch keeps
// every item in the launch, =sweep sends every item to the sweep.
std::vector<char> route_items(const smoq_ctx* ctx, const smoq_vqe_item* items, int n_items) {
  std::vector<char> big(n_items, 0);
  if (const char* e = std::getenv("SMOQ_VQE_ROUTE")) {
    if (!std::strcmp(e, "one_launch")) return big;
    if (!std::strcmp(e, "sweep")) return std::vector<char>(n_items, 1);
  }
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
      hipSuccess || n_cu <= 0)
    n_cu = 256;
  std::vector<double> tp(n_items), ts(n_items);
  std::vector<int> order(n_items);
  for (int i = 0; i < n_items; ++i) {
    const double amps = std::ldexp(1.0, items[i].n_qubits);
    const double ops = std::max(1, items[i].n_ops);
    tp[i] = 8.6e-10 * amps * ops;
    ts[i] = std::max(6e-7 * ops, 1.8e-12 * amps * ops);
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return tp[a] > tp[b]; });
  double sum_p = 0.0, sum_s = 0.0;
  for (int i = 0; i < n_items; ++i) sum_p += tp[i];
  auto launch_time = [&](size_t from) {
    // the remaining items (order[from..]) on up to n_cu workgroups
    if (from >= order.size()) return 0.0;
    const double per = sum_p / std::min<double>((double)n_cu, (double)(order.size() - from));
    return std::max(tp[order[from]], per);
  };
  // move the k most expensive items, for the k that minimises the longer of
  // the two streams' times
  size_t best_k = 0;
  double best = launch_time(0);
  for (size_t k = 1; k <= order.size(); ++k) {
    sum_p -= tp[order[k - 1]];
    sum_s += ts[order[k - 1]];
    const double t = std::max(launch_time(k), sum_s);
    if (t < best) {
      best = t;
      best_k = k;
    }
  }
  for (size_t k = 0; k < best_k; ++k) big[order[k]] = 1;
  return big;
}

}  // namespace
}  // namespace smoq

using namespace smoq;

extern "C" int smoq_adjoint_gradient_batch(smoq_ctx* ctx, int dtype, const smoq_vqe_item* items,
                                           int n_items) {
  if (!ctx) return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient_batch: null context");
  if (dtype != SMOQ_C128 && dtype != SMOQ_C64)
    return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient_batch: unknown dtype %d", dtype);
  if (n_items < 0 || (n_items > 0 && !items))
    return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient_batch: bad item list");
  if (n_items == 0) return SMOQ_OK;
  for (int i = 0; i < n_items; ++i) {
    const smoq_vqe_item& it = items[i];
    if (it.n_qubits < kRegQubits || it.n_qubits > 30)
      return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient_batch: item %d has %d qubits (3..30)",
                       i, it.n_qubits);
    if (it.n_params < 0 || (it.n_params > 0 && (!it.theta || !it.grad)) || !it.energy)
      return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient_batch: item %d: bad arrays", i);
  }
  const std::vector<char> big = route_items(ctx, items, n_items);
  std::vector<smoq_vqe_item> small;
  for (int i = 0; i < n_items; ++i)
    if (!big[i]) small.push_back(items[i]);
  std::vector<ItemPlan> plans;
  int rc = SMOQ_OK;
  if (!small.empty()) {
    if (!ctx->aux_stream)
      SMOQ_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    rc = dtype == SMOQ_C128
             ? run_batch<double>(ctx, dtype, small.data(), (int)small.size(), ctx->aux_stream, plans)
             : run_batch<float>(ctx, dtype, small.data(), (int)small.size(), ctx->aux_stream, plans);
    if (rc) {
      (void)hipStreamSynchronize(ctx->aux_stream);
      return rc;
    }
  }
  // the large items through the whole-device sweep on the context's stream,
  // while the launch runs: items with the same tape and Hamiltonian (equal
  // contents) as batches of parameter sets (each pass one launch for the
  // batch), up to a memory budget per batch, the others one after another
  std::vector<char> done(n_items, 0);
  for (int i = 0; i < n_items && !rc; ++i) {
    if (!big[i] || done[i]) continue;
    const smoq_vqe_item& it = items[i];
    std::vector<int> group{i};
    const uint64_t state_bytes =
        ((uint64_t)1 << it.n_qubits) * (dtype == SMOQ_C128 ? 16u : 8u);
    // ψ plus W (two states) per item
    const uint64_t max_group = std::max<uint64_t>(1, kVqeGroupBytes / (3 * state_bytes));
    for (int j = i + 1; j < n_items && group.size() < max_group; ++j)
      if (big[j] && !done[j] && same_problem(it, items[j])) group.push_back(j);
    for (int j : group) done[j] = 1;
    const int B = (int)group.size();
    smoq_state* s = nullptr;
    rc = smoq_state_alloc(ctx, it.n_qubits, dtype, B, &s);
    if (!rc) {
      if (B == 1) {
        rc = smoq_adjoint_gradient(s, it.ops, it.n_ops, it.theta, it.n_params, it.n_terms,
                                   it.xmask, it.zmask, it.coeff, it.energy, it.grad);
      } else {
        const size_t P = (size_t)it.n_params;
        std::vector<double> th(P * B), en(B), gr(P * B);
        for (int b = 0; b < B; ++b)
          std::copy(items[group[b]].theta, items[group[b]].theta + P, th.begin() + b * P);
        rc = smoq_adjoint_gradient(s, it.ops, it.n_ops, th.data(), it.n_params, it.n_terms,
                                   it.xmask, it.zmask, it.coeff, en.data(), gr.data());
        for (int b = 0; b < B && !rc; ++b) {
          *items[group[b]].energy = en[b];
          std::copy(gr.begin() + b * P, gr.begin() + (b + 1) * P, items[group[b]].grad);
        }
      }
    }
    if (s) {
      const int frc = smoq_state_free(s);
      if (!rc) rc = frc;
    }
  }
  if (!small.empty()) {
    const int frc = finish_batch(ctx->aux_stream, (const double*)ctx->res_pinned, small.data(),
                                 (int)small.size(), plans);
    if (!rc) rc = frc;
  }
  return rc;
}
