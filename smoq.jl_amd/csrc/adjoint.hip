// Adjoint gradients of parametrised circuits.
//
// A parametrised gate is R(θ) = exp(-iθG/2) with G = Π_C ⊗ P: a Pauli string
// P = i^{|x&z|} X^x Z^z, optionally restricted to indices whose control bits
// C are all set (CRY).  R(θ) = cos(θ/2) I - i sin(θ/2) G on the controlled
// subspace.  The reverse sweep keeps ψ_g (state after gate g) and
// λ_g = U_{>g}† H ψ_N; dE/dθ_g = 2 Re<λ_g|(-i/2) G ψ_g> = Im<λ_g|G|ψ_g>,
// then both vectors are un-rotated by R(-θ_g) in the same pass.
#include <algorithm>
#include <string>
#include <cstring>
#include <thread>
#include <atomic>
#include <cmath>

#include "adjoint_common.h"

namespace smoq {

constexpr int kAdjBlock = 256;

__device__ inline double2 ipow_c(int cnt) {
  double2 r;
  r.x = (cnt == 0) ? 1.0 : (cnt == 2) ? -1.0 : 0.0;
  r.y = (cnt == 1) ? 1.0 : (cnt == 3) ? -1.0 : 0.0;
  return r;
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Rotates the item's vector(s) by R(sign*θ) and, when lam is given,
// accumulates Im<λ|G|ψ> (before the rotation) into partial[item*bpi + block].
template <typename T>
__global__ void __launch_bounds__(kAdjBlock)
    k_prot(VT_alias<T>* __restrict__ psi, VT_alias<T>* __restrict__ lam, int n, PauliRot g,
           const double2* __restrict__ cs, int n_params, int pidx, double sign,
           double* __restrict__ partial) {
  __shared__ double sh[kAdjBlock / 64];
  const uint64_t item = blockIdx.y;
  const uint64_t dim = 1ull << n;
  VT_alias<T>* __restrict__ p = psi + item * dim;
  VT_alias<T>* __restrict__ l = lam ? lam + item * dim : nullptr;
  const double2 ang = cs[item * n_params + pidx];
  const double c = ang.x, s = sign * ang.y;
  const double2 ic = ipow_c(g.cnt);
  double ov = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (g.x) {
    const uint64_t half = dim >> 1;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < half; t += stride) {
      const uint64_t k = insert_zero(t, g.h);
      if ((k & g.ctrl) != g.ctrl) continue;
      const uint64_t j = k ^ g.x;
      // <k|G|j> = i^cnt (-1)^{|j&z|},  <j|G|k> = i^cnt (-1)^{|k&z|}
      double2 fkj = ic, fjk = ic;
      if (__popcll(j & g.z) & 1) fkj = make_c<double>(-ic.x, -ic.y);
      if (__popcll(k & g.z) & 1) fjk = make_c<double>(-ic.x, -ic.y);
      // -i s f
      const double2 mkj = make_c<double>(s * fkj.y, -s * fkj.x);
      const double2 mjk = make_c<double>(s * fjk.y, -s * fjk.x);
      const double2 a = to_d2(p[k]), b = to_d2(p[j]);
      if (l) {
        const double2 la = to_d2(l[k]), lb = to_d2(l[j]);
        // conj(λk) fkj ψj + conj(λj) fjk ψk : imaginary part
        const double2 u = cmul(fkj, b), v = cmul(fjk, a);
        ov += (la.x * u.y - la.y * u.x) + (lb.x * v.y - lb.y * v.x);
        const double2 nla = cadd(make_c<double>(c * la.x, c * la.y), cmul(mkj, lb));
        const double2 nlb = cadd(make_c<double>(c * lb.x, c * lb.y), cmul(mjk, la));
        l[k] = make_c<T>(nla.x, nla.y);
        l[j] = make_c<T>(nlb.x, nlb.y);
      }
      const double2 na = cadd(make_c<double>(c * a.x, c * a.y), cmul(mkj, b));
      const double2 nb = cadd(make_c<double>(c * b.x, c * b.y), cmul(mjk, a));
      p[k] = make_c<T>(na.x, na.y);
      p[j] = make_c<T>(nb.x, nb.y);
    }
  } else {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dim; k += stride) {
      if ((k & g.ctrl) != g.ctrl) continue;
      const double sg = (__popcll(k & g.z) & 1) ? -1.0 : 1.0;
      // (c - i s sg)
      const double2 f = make_c<double>(c, -s * sg);
      const double2 a = to_d2(p[k]);
      if (l) {
        const double2 la = to_d2(l[k]);
        ov += sg * (la.x * a.y - la.y * a.x);
        const double2 nl = cmul(f, la);
        l[k] = make_c<T>(nl.x, nl.y);
      }
      const double2 na = cmul(f, a);
      p[k] = make_c<T>(na.x, na.y);
    }
  }
  if (!partial) return;
  ov = wave_sum_d(ov);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = ov;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += sh[w];
    partial[item * gridDim.x + blockIdx.x] = v;
  }
}

// Re <ψ|λ> per item, block partials
template <typename T>
__global__ void __launch_bounds__(kAdjBlock)
    k_dot_re(const VT_alias<T>* __restrict__ psi, const VT_alias<T>* __restrict__ lam, int n,
             double* __restrict__ partial) {
  __shared__ double sh[kAdjBlock / 64];
  const uint64_t item = blockIdx.y;
  const uint64_t dim = 1ull << n;
  const VT_alias<T>* p = psi + item * dim;
  const VT_alias<T>* l = lam + item * dim;
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dim; k += stride) {
    const double2 a = to_d2(p[k]), b = to_d2(l[k]);
    acc += a.x * b.x + a.y * b.y;
  }
  acc = wave_sum_d(acc);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += sh[w];
    partial[item * gridDim.x + blockIdx.x] = v;
  }
}

// sum of bpi consecutive partials, fixed order: out[r] = Σ_b in[r*bpi + b]
__global__ void k_rowsum(const double* __restrict__ in, int bpi, int rows, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double v = 0.0;
  for (int b = 0; b < bpi; ++b) v += in[(size_t)r * bpi + b];
  out[r] = v;
}

// Start of the fused reverse sweep: W (n+1 qubits per item) interleaves ψ
// and λ = Hψ — W[2k] = ψ_k, W[2k+1] = λ_k, so W's qubit 0 selects the
// vector and its qubit q+1 is ψ's qubit q — and the block partials of
// E = Re<ψ|λ> (λ rounded to the state's type first, as a stored λ is).
// λ is the pull form of apply_hamiltonian_bitwise!, one write per index.
template <typename T>
__global__ void __launch_bounds__(kAdjBlock)
    k_adj_init(const VT_alias<T>* __restrict__ psi, VT_alias<T>* __restrict__ W, int n, int nt,
               const AdjTerm* __restrict__ terms, double* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[kAdjBlock / 64];
  const uint64_t item = blockIdx.y;
  const uint64_t dim = 1ull << n;
  const VT_alias<T>* p = psi + item * dim;
  VT_alias<T>* w = W + 2 * item * dim;
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dim; k += stride) {
    double hr = 0.0, hi = 0.0;
    for (int t = 0; t < nt; ++t) {
      const AdjTerm tm = terms[t];
      const uint64_t j = k ^ tm.x;
      const VT_alias<T> v = p[j];
      const double vr = (double)v.x, vi = (double)v.y;
      const bool neg = __popcll(j & tm.z) & 1;
      const double sr = neg ? -vr : vr, si = neg ? -vi : vi;
      hr += tm.fre * sr - tm.fim * si;
      hi += tm.fre * si + tm.fim * sr;
    }
    const VT_alias<T> a = p[k];
    const VT_alias<T> l = make_c<T>(hr, hi);
    w[2 * k] = a;
    w[2 * k + 1] = l;
    acc += (double)a.x * (double)l.x + (double)a.y * (double)l.y;
  }
  acc = wave_sum_d(acc);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) v += sh[ww];
    partial[item * gridDim.x + blockIdx.x] = v;
  }
}

// out[slot * batch + item] = the sum of the item's tile partials of one
// overlap slot (tiles_per_item consecutive tiles), in a fixed order: strided
// per-thread sums, then a fixed tree over the threads.
constexpr int kOvlRedBlock = 256;
__global__ void __launch_bounds__(kOvlRedBlock)
    k_ovl_reduce(const double* __restrict__ opart, uint64_t ntiles, uint64_t tiles_per_item,
                 int batch, double* __restrict__ out) {
  __shared__ double sh[kOvlRedBlock];
  const uint64_t slot = blockIdx.x / batch, item = blockIdx.x % batch;
  const double* src = opart + slot * ntiles + item * tiles_per_item;
  double v = 0.0;
  for (uint64_t t = threadIdx.x; t < tiles_per_item; t += kOvlRedBlock) v += src[t];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = kOvlRedBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

static IGate inverse_of(const IGate& g) {
  IGate r = g;
  if (g.kind == IK_U1 || g.kind == IK_U2) {
    const int d = g.kind == IK_U1 ? 2 : 4;
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        r.m[2 * (i * d + j)] = g.m[2 * (j * d + i)];
        r.m[2 * (i * d + j) + 1] = -g.m[2 * (j * d + i) + 1];
      }
  } else {
    for (int e = 0; e < (1 << g.nd); ++e) r.m[2 * e + 1] = -g.m[2 * e + 1];
  }
  return r;
}

int build_steps(const smoq_op* ops, int n_ops, int n_params, int n, std::vector<Step>& steps) {
  if (n_ops < 0 || (n_ops > 0 && !ops)) return set_error(SMOQ_EINVAL, "adjoint: bad tape");
  for (int i = 0; i < n_ops; ++i) {
    const smoq_op& op = ops[i];
    const bool rot = op.gate == SMOQ_G_RX || op.gate == SMOQ_G_RY || op.gate == SMOQ_G_RZ ||
                     op.gate == SMOQ_G_RXX || op.gate == SMOQ_G_RYY || op.gate == SMOQ_G_RZZ ||
                     op.gate == SMOQ_G_CRY;
    if (rot && op.param >= 0) {
      if (op.param >= n_params)
        return set_error(SMOQ_EINVAL, "adjoint: op %d uses parameter %d of %d", i, op.param,
                         n_params);
      // validate the qubits through the ordinary lowering
      std::vector<IGate> tmp;
      int rc = lower_op(op, 0.0, n, tmp);
      if (rc) return rc;
      Step st;
      st.param = true;
      st.pidx = op.param;
      st.op = op;
      PauliRot& g = st.rot;
      const uint64_t a = 1ull << op.q[0], b = op.q[1] >= 0 ? 1ull << op.q[1] : 0;
      g.ctrl = 0;
      switch (op.gate) {
        case SMOQ_G_RX: g.x = a; g.z = 0; break;
        case SMOQ_G_RY: g.x = a; g.z = a; break;
        case SMOQ_G_RZ: g.x = 0; g.z = a; break;
        case SMOQ_G_RXX: g.x = a | b; g.z = 0; break;
        case SMOQ_G_RYY: g.x = a | b; g.z = a | b; break;
        case SMOQ_G_RZZ: g.x = 0; g.z = a | b; break;
        default: g.x = b; g.z = b; g.ctrl = a; break;  // CRY: control q0, Y on q1
      }
      g.cnt = __builtin_popcountll(g.x & g.z) & 3;
      g.h = g.x ? 63 - __builtin_clzll(g.x) : 0;
      steps.push_back(st);
    } else {
      std::vector<IGate> gs;
      int rc = lower_op(op, op.theta, n, gs);
      if (rc) return rc;
      for (auto& g : gs) {
        Step st;
        st.fwd = g;
        st.inv = inverse_of(g);
        steps.push_back(st);
      }
    }
  }
  return SMOQ_OK;
}

IGate shifted_up(const IGate& g) {
  IGate r = g;
  if (r.t0 >= 0) r.t0 += 1;
  if (r.t1 >= 0) r.t1 += 1;
  for (int i = 0; i < r.nd; ++i) r.dq[i] += 1;
  r.ctrl <<= 1;
  r.oz <<= 1;
  return r;
}

int adj_terms(int n_terms, const uint64_t* xm, const uint64_t* zm, const double* coeff, int n,
              std::vector<AdjTerm>& out) {
  out.clear();
  if (n_terms < 0 || (n_terms > 0 && (!xm || !zm || !coeff)))
    return set_error(SMOQ_EINVAL, "adjoint: bad term arrays");
  const uint64_t dim = 1ull << n;
  for (int t = 0; t < n_terms; ++t) {
    if ((xm[t] | zm[t]) & ~(dim - 1))
      return set_error(SMOQ_EINVAL, "Pauli sum: term %d acts outside the %d qubits", t, n);
    const int cnt = __builtin_popcountll(xm[t] & zm[t]) & 3;
    const double re[4] = {1, 0, -1, 0}, im[4] = {0, 1, 0, -1};
    out.push_back({xm[t], zm[t], coeff[t] * re[cnt], coeff[t] * im[cnt]});
  }
  return SMOQ_OK;
}

int forward_gates(const std::vector<Step>& steps, const double* theta, int n,
                  std::vector<IGate>& out) {
  out.clear();
  for (const Step& st : steps) {
    if (!st.param) {
      out.push_back(st.fwd);
      continue;
    }
    const int rc = lower_op(st.op, theta[st.pidx], n, out);
    if (rc) return rc;
  }
  return SMOQ_OK;
}

int reverse_gates(const std::vector<Step>& steps, const double* theta, int n,
                  std::vector<IGate>& rg, int* n_slots_out) {
  rg.clear();
  int n_slots = 0;
  for (const Step& st : steps) n_slots += st.param;
  int slot = 0;
  std::vector<IGate> tmp;
  for (int i = (int)steps.size() - 1; i >= 0; --i) {
    const Step& st = steps[i];
    if (!st.param) {
      rg.push_back(shifted_up(st.inv));
      continue;
    }
    IGate o;
    o.kind = IK_OVL;
    const uint64_t x = st.rot.x;
    o.t0 = x ? __builtin_ctzll(x) : -1;
    o.t1 = __builtin_popcountll(x) == 2 ? 63 - __builtin_clzll(x) : -1;
    o.ctrl = st.rot.ctrl;
    o.oz = st.rot.z;
    o.ocnt = st.rot.cnt;
    o.oslot = slot++;
    IGate os = shifted_up(o);
    os.sel = 0;
    rg.push_back(os);
    tmp.clear();
    const int rc = lower_op(st.op, -theta[st.pidx], n, tmp);
    if (rc) return rc;
    for (const IGate& g : tmp) rg.push_back(shifted_up(g));
  }
  *n_slots_out = n_slots;
  return SMOQ_OK;
}

void grad_from_slots(const std::vector<Step>& steps, const double* hs, int n_slots, int n_params,
                     double* grad) {
  for (int p = 0; p < n_params; ++p) grad[p] = 0.0;
  int fwd = 0;
  for (const Step& st : steps) {
    if (!st.param) continue;
    grad[st.pidx] += hs[n_slots - 1 - fwd];
    ++fwd;
  }
}

namespace {

int blocks_per_item(const smoq_state* s) {
  const uint64_t work = s->dim >> 1;
  uint64_t b = (work + kAdjBlock * 4 - 1) / (kAdjBlock * 4);
  const uint64_t cap = std::max<uint64_t>(1, 4096 / (uint64_t)s->batch);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, std::min<uint64_t>(cap, 1024)));
}

template <typename T>
int launch_prot(smoq_state* s, void* lam, const PauliRot& g, const double2* cs, int n_params,
                int pidx, double sign, double* partial, int bpi) {
  hipLaunchKernelGGL(k_prot<T>, dim3(bpi, s->batch), dim3(kAdjBlock), 0, s->ctx->stream,
                     (VT_alias<T>*)s->d, (VT_alias<T>*)lam, s->n, g, cs, n_params, pidx, sign,
                     partial);
  SMOQ_CHECK_LAUNCH("k_prot");
  return SMOQ_OK;
}

int prot(smoq_state* s, void* lam, const PauliRot& g, const double2* cs, int n_params, int pidx,
         double sign, double* partial, int bpi) {
  return s->dtype == SMOQ_C128 ? launch_prot<double>(s, lam, g, cs, n_params, pidx, sign, partial, bpi)
                               : launch_prot<float>(s, lam, g, cs, n_params, pidx, sign, partial, bpi);
}

// Device buffers of one call, all grow-only work buffers of the context
// (kept across calls, so repeated evaluations allocate nothing):
//   slot 0: per-item (cos, sin) of every parameter's half angle
//   slot 1: λ = Hψ
//   slot 2: partial sums and their row sums
//   slot 3: the reverse sweep's working copy of ψ
struct Work {
  void* lam = nullptr;
  double2* cs = nullptr;
  double* partial = nullptr;
  double* sums = nullptr;
};

// forward sweep from |0> on every item; per-item angles from cs
int forward(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta, int n_params,
            const std::vector<Step>& steps, const double2* cs, bool one_gate) {
  int rc = smoq_state_set_basis(s, 0);
  if (rc) return rc;
  if (!one_gate && s->batch == 1 && s->n >= kRegQubits) {
    // one item: angles can be baked into a fused program
    std::vector<IGate>& gs = s->ctx->tape_gates;
    gs.clear();
    if ((rc = lower_tape(ops, n_ops, theta, n_params, s->n, gs))) return rc;
    Program& prog = s->ctx->tape_prog;
    plan_program(gs, s->n, pass_tile(s->ctx, s->n, s->batch, gs.size()), s->ctx->low_qubits,
                 s->dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits, prog,
                 s->ctx->pass_max_gates,
                 s->ctx->lane_swaps && jit_planned(s->ctx, s->n, gs.size()));
    return run_program(s, prog);
  }
  for (const Step& st : steps) {
    if (st.param)
      rc = prot(s, nullptr, st.rot, cs, n_params, st.pidx, 1.0, nullptr, blocks_per_item(s));
    else
      rc = launch_gate(s, st.fwd);
    if (rc) return rc;
  }
  return SMOQ_OK;
}

int setup(smoq_state* s, const double* theta, int n_params, Work& w, std::vector<double2>& hcs) {
  const size_t ncs = (size_t)s->batch * std::max(1, n_params);
  hcs.resize(ncs);
  for (size_t i = 0; i < ncs; ++i) {
    const double th = n_params > 0 ? theta[i] : 0.0;
    hcs[i] = make_c<double>(std::cos(0.5 * th), std::sin(0.5 * th));
  }
  void* p = nullptr;
  int rc = ensure_work(s->ctx, 0, ncs * sizeof(double2), &p);
  if (rc) return rc;
  w.cs = (double2*)p;
  SMOQ_HIP(hipMemcpyAsync(w.cs, hcs.data(), ncs * sizeof(double2), hipMemcpyHostToDevice,
                          s->ctx->stream));
  // hcs is the caller's host vector: the copy must finish before it goes
  SMOQ_HIP(hipStreamSynchronize(s->ctx->stream));
  return SMOQ_OK;
}

// λ = H ψ into w.lam and E_b = Re<ψ|λ>
int energy_of(smoq_state* s, int n_terms, const uint64_t* x, const uint64_t* z, const double* c,
              Work& w, double* energy) {
  int rc = ensure_work(s->ctx, 1, s->bytes, &w.lam);
  if (rc) return rc;
  smoq_state lam = *s;
  lam.d = w.lam;
  rc = smoq_hpsi(s, &lam, n_terms, x, z, c);
  if (rc) return rc;
  const int bpi = blocks_per_item(s);
  void* p = nullptr;
  const size_t part_n = (size_t)bpi * s->batch;
  if ((rc = ensure_work(s->ctx, 2, (part_n + s->batch + 2) * sizeof(double), &p))) return rc;
  double* part = (double*)p;
  double* esum = part + part_n;
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_dot_re<double>, dim3(bpi, s->batch), dim3(kAdjBlock), 0, s->ctx->stream,
                       (const double2*)s->d, (const double2*)w.lam, s->n, part);
  else
    hipLaunchKernelGGL(k_dot_re<float>, dim3(bpi, s->batch), dim3(kAdjBlock), 0, s->ctx->stream,
                       (const float2*)s->d, (const float2*)w.lam, s->n, part);
  SMOQ_CHECK_LAUNCH("k_dot_re");
  hipLaunchKernelGGL(k_rowsum, dim3((s->batch + 255) / 256), dim3(256), 0, s->ctx->stream, part,
                     bpi, s->batch, esum);
  SMOQ_CHECK_LAUNCH("k_rowsum");
  SMOQ_HIP(hipMemcpyAsync(energy, esum, s->batch * sizeof(double), hipMemcpyDeviceToHost,
                          s->ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(s->ctx->stream));
  return SMOQ_OK;
}

// The reverse sweep as fused passes over W (ψ and λ interleaved, one state
// of n+1 qubits): walking the gates backwards, a parametrised rotation
// becomes an overlap gate (Im<λ|G|ψ> into its slot) followed by R(-θ), any
// other gate its inverse; every gate acts on both halves at once.  Batches
// of one parameter set with compiled passes only; *done stays false when
// the caller should take the gate-by-gate sweep instead.
int fused_adjoint(smoq_state* s, const std::vector<Step>& steps, const double* theta,
                  int n_terms, const uint64_t* xm, const uint64_t* zm, const double* coeff,
                  int n_params, double* energy, double* grad, bool* done) {
  *done = false;
  smoq_ctx* ctx = s->ctx;
  const int nW = s->n + 1;
  if (s->batch != 1 || ctx->jit_mode == 0 || nW < 6 || nW > kMaxQubits) return SMOQ_OK;
  std::vector<AdjTerm> tv;
  int rc = adj_terms(n_terms, xm, zm, coeff, s->n, tv);
  if (rc) return rc;
  // the reverse program
  std::vector<IGate>& rg = ctx->tape_gates;
  int n_slots = 0;
  if ((rc = reverse_gates(steps, theta, s->n, rg, &n_slots))) return rc;
  Program& prog = ctx->tape_prog;
  const int rowbits = s->dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits;
  plan_program(rg, nW, pass_tile(ctx, nW, 1, rg.size()), ctx->low_qubits, rowbits, prog,
               ctx->pass_max_gates, ctx->lane_swaps != 0);
  std::shared_ptr<JitModule> jit;
  if (!prog.passes.empty()) {
    const int jrc = jit_get(ctx, prog, s->dtype, jit);
    if (jrc) {
      if (ctx->jit_mode >= 2) return jrc;
      return SMOQ_OK;  // automatic mode: the gate-by-gate sweep
    }
  }
  // buffers: W (2S) | overlap partials, slot sums, energy partials | terms
  void* wbuf = nullptr;
  rc = ensure_work(ctx, 1, 2 * s->bytes, &wbuf);
  if (rc) return rc;
  const uint64_t ntiles = prog.passes.empty() ? 1 : (1ull << prog.passes[0].n_nonlocal);
  const int bpi = blocks_per_item(s);
  const size_t n_op = (size_t)n_slots * ntiles;
  void* pb = nullptr;
  if ((rc = ensure_work(ctx, 2, (n_op + n_slots + bpi + 2) * sizeof(double), &pb))) return rc;
  double* opart = (double*)pb;
  double* sums = opart + n_op;
  double* epart = sums + n_slots;
  double* esum = epart + bpi;
  void* tb = nullptr;
  if ((rc = ensure_work(ctx, 3, (tv.size() + 1) * sizeof(AdjTerm), &tb))) return rc;
  AdjTerm* dterms = (AdjTerm*)tb;
  if (!tv.empty())
    SMOQ_HIP(hipMemcpyAsync(dterms, tv.data(), tv.size() * sizeof(AdjTerm), hipMemcpyHostToDevice,
                            ctx->stream));
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_adj_init<double>, dim3(bpi, 1), dim3(kAdjBlock), 0, ctx->stream,
                       (const double2*)s->d, (double2*)wbuf, s->n, (int)tv.size(), dterms, epart);
  else
    hipLaunchKernelGGL(k_adj_init<float>, dim3(bpi, 1), dim3(kAdjBlock), 0, ctx->stream,
                       (const float2*)s->d, (float2*)wbuf, s->n, (int)tv.size(), dterms, epart);
  SMOQ_CHECK_LAUNCH("k_adj_init");
  hipLaunchKernelGGL(k_rowsum, dim3(1), dim3(256), 0, ctx->stream, epart, bpi, 1, esum);
  SMOQ_CHECK_LAUNCH("k_rowsum");
  if (!prog.passes.empty()) {
    smoq_state ws = *s;
    ws.n = nW;
    ws.dim = 2 * s->dim;
    ws.bytes = 2 * s->bytes;
    ws.d = wbuf;
    const ProgLayout L = program_layout(prog, s->dtype);
    if ((rc = ensure_prog(ctx, L.total))) return rc;
    if ((rc = upload_program(ctx, prog, s->dtype, L, ctx->prog))) return rc;
    if ((rc = launch_program(&ws, prog, L, ctx->prog, jit.get(), opart))) return rc;
    if (n_slots > 0) {
      hipLaunchKernelGGL(k_ovl_reduce, dim3((unsigned)n_slots), dim3(kOvlRedBlock), 0,
                         ctx->stream, opart, ntiles, ntiles, 1, sums);
      SMOQ_CHECK_LAUNCH("k_ovl_reduce");
    }
  }
  std::vector<double> hs(n_slots + 1);
  if (n_slots > 0)
    SMOQ_HIP(hipMemcpyAsync(hs.data(), sums, n_slots * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
  SMOQ_HIP(hipMemcpyAsync(energy, esum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  // each parameter's slots in gate order (slot numbers run backwards)
  grad_from_slots(steps, hs.data(), n_slots, n_params, grad);
  *done = true;
  return SMOQ_OK;
}

// Plans of every item of a batch on host threads: fn(b) for b < B.
template <typename F>
static void for_items(int B, F fn) {
  const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nth = std::min(B, hw);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int b; (b = next.fetch_add(1)) < B;) fn(b);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
}

// every item's program image, item_bytes apart, in one copy through the
// pinned staging buffer
static int upload_items(smoq_ctx* ctx, const std::vector<Program>& progs, const ProgLayout& L,
                        size_t item_bytes, void* d) {
  const size_t total = item_bytes * progs.size();
  int rc = ensure_pinned(ctx, total);
  if (rc) return rc;
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));  // the staging buffer may feed an earlier copy
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  for (size_t b = 0; b < progs.size(); ++b) {
    const Program& pr = progs[b];
    unsigned char* o = h + b * item_bytes;
    std::memcpy(o, pr.passes.data(), L.pb);
    std::memcpy(o + L.pb, pr.stages.data(), L.sb);
    std::memcpy(o + L.pb + L.sb, pr.gates.data(), L.gb);
    if (L.ntab) std::memcpy(o + L.off_t, pr.tabs.data(), L.ntab * sizeof(double2));
  }
  SMOQ_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, ctx->stream));
  return SMOQ_OK;
}

// The batched form of fused_adjoint: B parameter sets of one tape on a
// batch state.  Every item gets its own copy of the forward and reverse
// programs (same structure, its own angles; structures that differ between
// items leave *done false), so each pass is one launch for the whole batch.
int fused_adjoint_batch(smoq_state* s, const smoq_op* ops, int n_ops,
                        const std::vector<Step>& steps, const double* theta, int n_terms,
                        const uint64_t* xm, const uint64_t* zm, const double* coeff,
                        int n_params, double* energy, double* grad, bool* done) {
  *done = false;
  smoq_ctx* ctx = s->ctx;
  const int B = s->batch, n = s->n, nW = n + 1;
  if (B < 2 || ctx->jit_mode == 0 || n < kRegQubits || nW < 6 || nW > kMaxQubits)
    return SMOQ_OK;
  std::vector<AdjTerm> tv;
  int rc = adj_terms(n_terms, xm, zm, coeff, n, tv);
  if (rc) return rc;
  const int rowbits = s->dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits;
  // every item's forward and reverse programs, planned on host threads
  std::vector<Program> fp(B), rp(B);
  std::vector<int> errs(B, 0), slots(B, 0);
  std::vector<std::string> fkey(B), rkey(B);
  size_t n_fwd_gates = 0;
  {
    std::vector<IGate> g0;
    if ((rc = lower_tape(ops, n_ops, theta, n_params, n, g0))) return rc;
    n_fwd_gates = g0.size();
  }
  const int ftile = pass_tile(ctx, n, B, n_fwd_gates);
  const bool fswaps = ctx->lane_swaps != 0 && jit_planned(ctx, n, n_fwd_gates);
  for_items(B, [&](int b) {
    std::vector<IGate> gs, rg;
    const double* th = theta + (size_t)b * n_params;
    int e = lower_tape(ops, n_ops, th, n_params, n, gs);
    if (!e) {
      plan_program(gs, n, ftile, ctx->low_qubits, rowbits, fp[b], ctx->pass_max_gates, fswaps);
      fp[b].per_item = true;
      jit_structure_key(fp[b], s->dtype, fkey[b]);
      e = reverse_gates(steps, th, n, rg, &slots[b]);
    }
    if (!e) {
      plan_program(rg, nW, pass_tile(ctx, nW, B, rg.size()), ctx->low_qubits, rowbits, rp[b],
                   ctx->pass_max_gates, ctx->lane_swaps != 0);
      rp[b].per_item = true;
      jit_structure_key(rp[b], s->dtype, rkey[b]);
    }
    errs[b] = e;
  });
  for (int b = 0; b < B; ++b)
    if (errs[b]) {
      // again on this thread, for its error message (thread-local)
      std::vector<IGate> gs, rg;
      const double* th = theta + (size_t)b * n_params;
      if ((rc = lower_tape(ops, n_ops, th, n_params, n, gs))) return rc;
      int ns = 0;
      if ((rc = reverse_gates(steps, th, n, rg, &ns))) return rc;
      return errs[b];
    }
  for (int b = 1; b < B; ++b)
    if (fkey[b] != fkey[0] || rkey[b] != rkey[0]) return SMOQ_OK;  // item structures differ
  const int n_slots = slots[0];
  if (fp[0].passes.empty() || rp[0].passes.empty()) return SMOQ_OK;
  std::shared_ptr<JitModule> jf, jr;
  if ((rc = jit_get(ctx, fp[0], s->dtype, jf)) || (rc = jit_get(ctx, rp[0], s->dtype, jr)))
    return ctx->jit_mode >= 2 ? rc : SMOQ_OK;
  const ProgLayout Lf = program_layout(fp[0], s->dtype), Lr = program_layout(rp[0], s->dtype);
  const size_t fbytes = (Lf.total + 255) & ~(size_t)255, rbytes = (Lr.total + 255) & ~(size_t)255;
  // buffers: programs (forward, then reverse) | W (2S) | overlap partials,
  // slot sums, energy partials | terms
  void* pbuf = nullptr;
  if ((rc = ensure_work(ctx, 0, std::max(fbytes, rbytes) * B, &pbuf))) return rc;
  void* wbuf = nullptr;
  if ((rc = ensure_work(ctx, 1, 2 * s->bytes, &wbuf))) return rc;
  const uint64_t tiles_item = 1ull << rp[0].passes[0].n_nonlocal;
  const uint64_t ntiles = (uint64_t)B * tiles_item;
  const int bpi = blocks_per_item(s);
  const size_t n_op = (size_t)n_slots * ntiles, n_sums = (size_t)n_slots * B;
  void* pb = nullptr;
  if ((rc = ensure_work(ctx, 2, (n_op + n_sums + (size_t)B * bpi + B + 2) * sizeof(double), &pb)))
    return rc;
  double* opart = (double*)pb;
  double* sums = opart + n_op;
  double* epart = sums + n_sums;
  double* esum = epart + (size_t)B * bpi;
  void* tb = nullptr;
  if ((rc = ensure_work(ctx, 3, (tv.size() + 1) * sizeof(AdjTerm), &tb))) return rc;
  AdjTerm* dterms = (AdjTerm*)tb;
  if (!tv.empty())
    SMOQ_HIP(hipMemcpyAsync(dterms, tv.data(), tv.size() * sizeof(AdjTerm), hipMemcpyHostToDevice,
                            ctx->stream));
  // forward: |0…0> through each item's program
  if ((rc = smoq_state_set_basis(s, 0))) return rc;
  if ((rc = upload_items(ctx, fp, Lf, fbytes, pbuf))) return rc;
  if ((rc = launch_program(s, fp[0], Lf, pbuf, jf.get(), nullptr, -1, fbytes))) return rc;
  // W and the energies
  if (s->dtype == SMOQ_C128)
    hipLaunchKernelGGL(k_adj_init<double>, dim3(bpi, B), dim3(kAdjBlock), 0, ctx->stream,
                       (const double2*)s->d, (double2*)wbuf, n, (int)tv.size(), dterms, epart);
  else
    hipLaunchKernelGGL(k_adj_init<float>, dim3(bpi, B), dim3(kAdjBlock), 0, ctx->stream,
                       (const float2*)s->d, (float2*)wbuf, n, (int)tv.size(), dterms, epart);
  SMOQ_CHECK_LAUNCH("k_adj_init");
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, epart,
                     bpi, B, esum);
  SMOQ_CHECK_LAUNCH("k_rowsum");
  // reverse sweep over every item's W
  smoq_state ws = *s;
  ws.n = nW;
  ws.dim = 2 * s->dim;
  ws.bytes = 2 * s->bytes;
  ws.d = wbuf;
  if ((rc = upload_items(ctx, rp, Lr, rbytes, pbuf))) return rc;
  if ((rc = launch_program(&ws, rp[0], Lr, pbuf, jr.get(), opart, -1, rbytes))) return rc;
  if (n_slots > 0) {
    hipLaunchKernelGGL(k_ovl_reduce, dim3((unsigned)n_sums), dim3(kOvlRedBlock), 0, ctx->stream,
                       opart, ntiles, tiles_item, B, sums);
    SMOQ_CHECK_LAUNCH("k_ovl_reduce");
  }
  std::vector<double> hs(n_sums + 1);
  if (n_sums > 0)
    SMOQ_HIP(hipMemcpyAsync(hs.data(), sums, n_sums * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
  SMOQ_HIP(hipMemcpyAsync(energy, esum, B * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  // each item's parameters from its slots, in gate order
  std::vector<double> hb(n_slots + 1);
  for (int b = 0; b < B; ++b) {
    for (int k = 0; k < n_slots; ++k) hb[k] = hs[(size_t)k * B + b];
    grad_from_slots(steps, hb.data(), n_slots, n_params, grad + (size_t)b * n_params);
  }
  *done = true;
  return SMOQ_OK;
}

}  // namespace
}  // namespace smoq

using namespace smoq;

extern "C" int smoq_tape_energy(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                                int n_params, int n_terms, const uint64_t* xmask,
                                const uint64_t* zmask, const double* coeff, int flags,
                                double* energy) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!energy || n_params < 0 || (n_params > 0 && !theta))
    return set_error(SMOQ_EINVAL, "smoq_tape_energy: bad arguments");
  if (n_terms < 0 || (n_terms > 0 && (!xmask || !zmask || !coeff)))
    return set_error(SMOQ_EINVAL, "smoq_tape_energy: bad term arrays");
  std::vector<Step> steps;
  if ((rc = build_steps(ops, n_ops, n_params, s->n, steps))) return rc;
  const bool one_gate = (flags & SMOQ_ENERGY_ONE_GATE_KERNELS) != 0;
  Work w;
  std::vector<double2> hcs;
  // the angle table feeds the per-gate rotations only (batches, one-gate mode)
  if ((one_gate || s->batch > 1 || s->n < kRegQubits) &&
      (rc = setup(s, theta, n_params, w, hcs)))
    return rc;
  if ((rc = forward(s, ops, n_ops, theta, n_params, steps, w.cs, one_gate))) return rc;
  // E_b = Σ_t c_t <P_t>_b from the tiled Pauli reduction (one read of ψ per
  // tile pass), summed on the host in term order
  std::vector<double> vals((size_t)s->batch * std::max(1, n_terms));
  if (n_terms > 0 && (rc = reduce_pauli(s, n_terms, xmask, zmask, vals.data(), 0))) return rc;
  for (int b = 0; b < s->batch; ++b) {
    double e = 0.0;
    for (int t = 0; t < n_terms; ++t) e += coeff[t] * vals[(size_t)b * n_terms + t];
    energy[b] = e;
  }
  return SMOQ_OK;
}

extern "C" int smoq_adjoint_gradient(smoq_state* s, const smoq_op* ops, int n_ops,
                                     const double* theta, int n_params, int n_terms,
                                     const uint64_t* xmask, const uint64_t* zmask,
                                     const double* coeff, double* energy, double* grad) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!energy || !grad || n_params < 0 || (n_params > 0 && !theta))
    return set_error(SMOQ_EINVAL, "smoq_adjoint_gradient: bad arguments");
  std::vector<Step> steps;
  if ((rc = build_steps(ops, n_ops, n_params, s->n, steps))) return rc;
  Work w;
  std::vector<double2> hcs;
  // one parameter set: fused forward passes, then the fused reverse sweep
  if (s->batch == 1 && s->n >= kRegQubits) {
    if ((rc = forward(s, ops, n_ops, theta, n_params, steps, nullptr, false))) return rc;
    bool done = false;
    if ((rc = fused_adjoint(s, steps, theta, n_terms, xmask, zmask, coeff, n_params, energy, grad,
                            &done)))
      return rc;
    if (done) return SMOQ_OK;
    if ((rc = setup(s, theta, n_params, w, hcs))) return rc;
  } else {
    // a batch of parameter sets: every pass one launch for all items
    bool done = false;
    if ((rc = fused_adjoint_batch(s, ops, n_ops, steps, theta, n_terms, xmask, zmask, coeff,
                                  n_params, energy, grad, &done)))
      return rc;
    if (done) return SMOQ_OK;
    if ((rc = setup(s, theta, n_params, w, hcs))) return rc;
    if ((rc = forward(s, ops, n_ops, theta, n_params, steps, w.cs, false))) return rc;
  }
  if ((rc = energy_of(s, n_terms, xmask, zmask, coeff, w, energy))) return rc;

  // reverse sweep on a working copy of ψ, so that s keeps U(θ)|0>
  void* pcopy = nullptr;
  if ((rc = ensure_work(s->ctx, 3, s->bytes, &pcopy))) return rc;
  SMOQ_HIP(hipMemcpyAsync(pcopy, s->d, s->bytes, hipMemcpyDeviceToDevice, s->ctx->stream));
  smoq_state ps = *s;
  ps.d = pcopy;
  smoq_state ls = *s;
  ls.d = w.lam;
  int n_slots = 0;
  for (const Step& st : steps) n_slots += st.param;
  const int bpi = blocks_per_item(s);
  const size_t rows = (size_t)n_slots * s->batch;
  void* pp = nullptr;
  if ((rc = ensure_work(s->ctx, 2, (rows * bpi + rows + 2) * sizeof(double), &pp))) return rc;
  w.partial = (double*)pp;
  w.sums = w.partial + rows * bpi;
  int slot = n_slots;
  for (int i = (int)steps.size() - 1; i >= 0; --i) {
    const Step& st = steps[i];
    if (st.param) {
      --slot;
      double* part = w.partial + (size_t)slot * s->batch * bpi;
      rc = prot(&ps, w.lam, st.rot, w.cs, n_params, st.pidx, -1.0, part, bpi);
    } else {
      rc = launch_gate(&ps, st.inv);
      if (!rc) rc = launch_gate(&ls, st.inv);
    }
    if (rc) return rc;
  }
  std::vector<double> sums(rows);
  if (rows) {
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0,
                       s->ctx->stream, w.partial, bpi, (int)rows, w.sums);
    SMOQ_CHECK_LAUNCH("k_rowsum");
    SMOQ_HIP(hipMemcpyAsync(sums.data(), w.sums, rows * sizeof(double), hipMemcpyDeviceToHost,
                            s->ctx->stream));
  }
  SMOQ_HIP(hipStreamSynchronize(s->ctx->stream));
  // sum the slots of each parameter in gate order (deterministic)
  for (int b = 0; b < s->batch; ++b)
    for (int p = 0; p < n_params; ++p) grad[(size_t)b * n_params + p] = 0.0;
  slot = 0;
  for (const Step& st : steps) {
    if (!st.param) continue;
    for (int b = 0; b < s->batch; ++b)
      grad[(size_t)b * n_params + st.pidx] += sums[(size_t)slot * s->batch + b];
    ++slot;
  }
  return SMOQ_OK;
}
