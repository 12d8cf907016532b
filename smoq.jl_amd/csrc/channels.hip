// Single-qubit stochastic operations on every batch item, one caller uniform
// per (qubit, item) in place of the reference's rand(): Z measurement with
// collapse, forced reset, and the trajectory forms of the Kraus channels
// (utils/cpu/cpuQuantumChannelKrausOperators.jl:262-385,
// utils/cpu/cpuQuantumStateMeasurements.jl:224-256, 584-635).  The density
// matrix forms (vec(rho) as a 2N-qubit state) of the Z measurement and of the
// forced projection are here too (…Measurements.jl:259-295, 686-708).
//
// Per qubit, all on the context stream with one synchronisation per call:
//   k_qc_sums     P(bit 0), P(bit 1) per chunk (measurement, damping only)
//   k_qc_decide   per item: the chunk sums in a fixed order, the decision
//                 from u and the populations, the record
//   k_qc_apply    the pair rewrite of every chunk, per-chunk norms
//   k_qc_norm     per item: the norm (damping, forced reset)
//   k_qc_scale    division by it
// Items whose decision is "nothing" and need no renormalisation read nothing.
#include <cmath>
#include <vector>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

constexpr int kQcBlock = 256;

enum QcOp : int {
  kOpNone = 0,
  kOpX = 1,
  kOpY = 2,
  kOpZ = 3,
  kOpKeep0 = 4,
  kOpKeep1 = 5,
  kOpJump = 6,     // amplitude damping jump: psi0 += psi1, psi1 = 0
  kOpNoJump = 7,   // psi1 *= sqrt(1 - gamma)
  kOpMove0 = 8,    // forced |0>: psi0 += psi1, psi1 = 0
  kOpMove1 = 9     // forced |1>: psi1 += psi0, psi0 = 0
};

__device__ inline double qc_block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kQcBlock / 64; ++w) t += sh[w];
  return t;
}

__device__ inline double abs2d(double re, double im) { return re * re + im * im; }

// partial[(item * gridDim.x + blk) * 2 + b] = sum of |psi_i|^2 with bit q = b
template <typename T>
__global__ void __launch_bounds__(kQcBlock)
    k_qc_sums(const VT_alias<T>* __restrict__ psi, int n, int lc, int q,
              double* __restrict__ partial) {
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const uint64_t g0 = blk << lc, chunk = 1ull << lc;
  const VT_alias<T>* p = psi + (item << n);
  double s0 = 0.0, s1 = 0.0;
  for (uint64_t k = threadIdx.x; k < chunk; k += kQcBlock) {
    const uint64_t i = g0 | k;
    const VT_alias<T> v = p[i];
    const double a = abs2d((double)v.x, (double)v.y);
    if ((i >> q) & 1)
      s1 += a;
    else
      s0 += a;
  }
  s0 = qc_block_sum(s0, sh);
  s1 = qc_block_sum(s1, sh);
  if (threadIdx.x == 0) {
    partial[(item * gridDim.x + blk) * 2] = s0;
    partial[(item * gridDim.x + blk) * 2 + 1] = s1;
  }
}

// the operation chosen for one item; every block of the item reads it
struct QcDecision {
  int op;
  int record;    // written to out[item]
  double scale;  // multiplies the kept half (measurement); 1 otherwise
};

__device__ inline QcDecision qc_decide(int kind, double p, double u, double p0, double p1) {
  QcDecision d{kOpNone, 0, 1.0};
  switch (kind) {
    case SMOQ_QC_MEASURE_Z: {
      // outcome = rand() < prob_0 ? 0 : 1; psi *= 1 / sqrt(kept norm)
      const int o = u < p0 ? 0 : 1;
      d.op = o ? kOpKeep1 : kOpKeep0;
      d.scale = 1.0 / sqrt(o ? p1 : p0);
      d.record = o;
      break;
    }
    case SMOQ_QC_DEPHASING:
      // Z with probability q, (1 - 2q) = sqrt(1 - p); p == 0 is the identity
      if (p != 0.0 && u < (1.0 - sqrt(1.0 - p)) / 2.0) d.op = kOpZ;
      break;
    case SMOQ_QC_DEPOLARIZING:
      if (u < p / 3)
        d.op = kOpX;
      else if (u < 2 * p / 3)
        d.op = kOpY;
      else if (u < p)
        d.op = kOpZ;
      break;
    case SMOQ_QC_BIT_FLIP:
      if (u < p) d.op = kOpX;
      break;
    case SMOQ_QC_PHASE_FLIP:
      if (u < p) d.op = kOpZ;
      break;
    case SMOQ_QC_AMPLITUDE_DAMPING:
      // jump when rand() < gamma * pop_1 && pop_1 > 1e-12
      d.op = (u < p * p1 && p1 > 1e-12) ? kOpJump : kOpNoJump;
      d.record = d.op == kOpJump ? 4 : 0;
      break;
    case SMOQ_QC_FORCE_ZERO:
      d.op = kOpMove0;
      break;
    default:
      d.op = kOpMove1;
      d.record = 1;
      break;
  }
  if (d.op >= kOpX && d.op <= kOpZ) d.record = d.op;
  return d;
}

__host__ __device__ inline bool qc_renormalises(int kind) {
  return kind == SMOQ_QC_AMPLITUDE_DAMPING || kind == SMOQ_QC_FORCE_ZERO ||
         kind == SMOQ_QC_FORCE_ONE;
}

// Sum of the nb chunk values of one item, nb fixed per call: strided per
// thread, then the block sum, so the order does not depend on scheduling.
__device__ inline double qc_item_sum(const double* __restrict__ v, int nb, int stride,
                                     double* sh) {
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += kQcBlock) a += v[(size_t)b * stride];
  return qc_block_sum(a, sh);
}

// one block per item: populations (when sums != nullptr), then the decision
__global__ void __launch_bounds__(kQcBlock)
    k_qc_decide(int kind, double prm, const double* __restrict__ uni,
                const double* __restrict__ sums, int nb, QcDecision* __restrict__ dec,
                int32_t* __restrict__ out, unsigned long long* __restrict__ fault) {
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.x;
  double p0 = 0.0, p1 = 0.0;
  if (sums) {
    p0 = qc_item_sum(sums + item * nb * 2, nb, 2, sh);
    p1 = qc_item_sum(sums + item * nb * 2 + 1, nb, 2, sh);
  }
  if (threadIdx.x == 0) {
    const QcDecision d = qc_decide(kind, prm, uni[item], p0, p1);
    dec[item] = d;
    out[item] = d.record;
    if (!isfinite(d.scale)) atomicMin(fault, (unsigned long long)item);
  }
}

template <typename T>
__global__ void __launch_bounds__(kQcBlock)
    k_qc_apply(VT_alias<T>* __restrict__ psi, int n, int lc, int q, int kind, double prm,
               const QcDecision* __restrict__ dec, double* __restrict__ normp) {
  using V = VT_alias<T>;
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const int nb = gridDim.x;
  const QcDecision d = dec[item];
  const bool renorm = qc_renormalises(kind);
  if (d.op == kOpNone) return;  // no renormalising kind decides "nothing"
  V* p = psi + (item << n);
  const uint64_t m = 1ull << q;
  const uint64_t g0 = blk << lc, chunk = 1ull << lc;
  const double sq = d.op == kOpNoJump ? sqrt(1.0 - prm) : 1.0;
  // pairs (i0, i0 | m): inside the chunk when q < lc; otherwise the chunk
  // with bit q clear rewrites its partner chunk as well
  const bool own = q < lc;
  double acc = 0.0;
  if (own || !(g0 & m)) {
    const uint64_t npairs = own ? (chunk >> 1) : chunk;
    for (uint64_t t = threadIdx.x; t < npairs; t += kQcBlock) {
      const uint64_t i0 = own ? (g0 | insert_zero(t, q)) : (g0 | t), i1 = i0 | m;
      const V a = p[i0], b = p[i1];
      double ar = a.x, ai = a.y, br = b.x, bi = b.y;
      switch (d.op) {
        case kOpX: {
          const double tr = ar, ti = ai;
          ar = br; ai = bi; br = tr; bi = ti;
          break;
        }
        case kOpY: {  // (psi0, psi1) <- (-i psi1, i psi0)
          const double nar = bi, nai = -br, nbr = -ai, nbi = ar;
          ar = nar; ai = nai; br = nbr; bi = nbi;
          break;
        }
        case kOpZ: br = -br; bi = -bi; break;
        case kOpKeep0: ar *= d.scale; ai *= d.scale; br = 0.0; bi = 0.0; break;
        case kOpKeep1: br *= d.scale; bi *= d.scale; ar = 0.0; ai = 0.0; break;
        case kOpJump:
        case kOpMove0: ar += br; ai += bi; br = 0.0; bi = 0.0; break;
        case kOpNoJump: br *= sq; bi *= sq; break;
        default: br += ar; bi += ai; ar = 0.0; ai = 0.0; break;  // kOpMove1
      }
      const V na = make_c<T>(ar, ai), nbv = make_c<T>(br, bi);
      p[i0] = na;
      p[i1] = nbv;
      if (renorm)
        acc += abs2d((double)na.x, (double)na.y) + abs2d((double)nbv.x, (double)nbv.y);
    }
  }
  if (renorm) {
    acc = qc_block_sum(acc, sh);
    if (threadIdx.x == 0) normp[item * nb + blk] = acc;
  }
}

// one block per item: the divisor of psi ./= norm — amplitude damping divides
// by sqrt(norm2) when that exceeds 1e-12, the forced reset (_normalize!) when
// norm2 exceeds 1e-15; 1 means no division
__global__ void __launch_bounds__(kQcBlock)
    k_qc_norm(int kind, const double* __restrict__ normp, int nb, double* __restrict__ div) {
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.x;
  const double nrm = qc_item_sum(normp + item * nb, nb, 1, sh);
  if (threadIdx.x == 0) {
    const double r = sqrt(nrm);
    const bool go = kind == SMOQ_QC_AMPLITUDE_DAMPING ? r > 1e-12 : nrm > 1e-15;
    div[item] = go ? r : 1.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kQcBlock)
    k_qc_scale(VT_alias<T>* __restrict__ psi, int n, int lc, const double* __restrict__ div) {
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const double dv = div[item];
  if (dv == 1.0) return;
  VT_alias<T>* p = psi + (item << n);
  const uint64_t g0 = blk << lc, chunk = 1ull << lc;
  for (uint64_t k = threadIdx.x; k < chunk; k += kQcBlock) {
    const VT_alias<T> v = p[g0 | k];
    p[g0 | k] = make_c<T>((double)v.x / dv, (double)v.y / dv);
  }
}

// ---- density matrices: vec(rho), element (i, j) at i + 2^N j -------------------
// partial[(item * gridDim.x + blk) * 2 + b] = sum of Re rho_ii with bit q of i = b
template <typename T>
__global__ void __launch_bounds__(kQcBlock)
    k_qc_dm_sums(const VT_alias<T>* __restrict__ rho, int N, int lcd, int q,
                 double* __restrict__ partial) {
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const uint64_t g0 = blk << lcd, chunk = 1ull << lcd;
  const VT_alias<T>* p = rho + (item << (2 * N));
  double s0 = 0.0, s1 = 0.0;
  for (uint64_t k = threadIdx.x; k < chunk; k += kQcBlock) {
    const uint64_t i = g0 | k;
    const double a = (double)p[i | (i << N)].x;
    if ((i >> q) & 1)
      s1 += a;
    else
      s0 += a;
  }
  s0 = qc_block_sum(s0, sh);
  s1 = qc_block_sum(s1, sh);
  if (threadIdx.x == 0) {
    partial[(item * gridDim.x + blk) * 2] = s0;
    partial[(item * gridDim.x + blk) * 2 + 1] = s1;
  }
}

// one block per item: the kept bit b (measurement: b = rand() < P(0) ? 0 : 1,
// prob = P(0) or 1 - P(0); forced: prob = the kept trace) in op, the divisor
// (prob when prob > 1e-15, else 1) in scale
__global__ void __launch_bounds__(kQcBlock)
    k_qc_dm_decide(int kind, const double* __restrict__ uni, const double* __restrict__ sums,
                   int nbd, QcDecision* __restrict__ dec, int32_t* __restrict__ out) {
  __shared__ double sh[kQcBlock / 64];
  const uint64_t item = blockIdx.x;
  const double p0 = qc_item_sum(sums + item * nbd * 2, nbd, 2, sh);
  const double p1 = qc_item_sum(sums + item * nbd * 2 + 1, nbd, 2, sh);
  if (threadIdx.x == 0) {
    int b;
    double prob;
    if (kind == SMOQ_QC_MEASURE_Z) {
      b = uni[item] < p0 ? 0 : 1;
      prob = b ? 1.0 - p0 : p0;
    } else {
      b = kind == SMOQ_QC_FORCE_ONE ? 1 : 0;
      prob = b ? p1 : p0;
    }
    dec[item] = QcDecision{b, b, prob > 1e-15 ? prob : 1.0};
    out[item] = b;
  }
}

// P_b rho P_b / divisor
template <typename T>
__global__ void __launch_bounds__(kQcBlock)
    k_qc_dm_project(VT_alias<T>* __restrict__ rho, int N, int lc, int q,
                    const QcDecision* __restrict__ dec) {
  const uint64_t item = blockIdx.y, blk = blockIdx.x;
  const uint64_t b = (uint64_t)dec[item].op;
  const double dv = dec[item].scale;
  VT_alias<T>* p = rho + (item << (2 * N));
  const uint64_t g0 = blk << lc, chunk = 1ull << lc;
  const uint64_t low = (1ull << N) - 1;
  for (uint64_t k = threadIdx.x; k < chunk; k += kQcBlock) {
    const uint64_t e = g0 | k;
    const bool keep = (((e & low) >> q) & 1) == b && (((e >> N) >> q) & 1) == b;
    if (keep) {
      const VT_alias<T> v = p[e];
      p[e] = make_c<T>((double)v.x / dv, (double)v.y / dv);
    } else {
      p[e] = make_c<T>(0.0, 0.0);
    }
  }
}

// Chunks per item, a power of two: items are cut into as many chunks of at
// least 4096 elements as keep the grid within 2^16 blocks, so that even one
// item's rewrite spreads over the whole chip; blocks of items whose decision
// is "nothing" return at once.
static int qc_chunks(uint64_t len, size_t batch) {
  int nb = 1;
  while ((uint64_t)nb * 2 * 4096 <= len && (uint64_t)nb * 2 * batch <= 65536) nb *= 2;
  return nb;
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T>
static int qc_run(smoq_state* s, int nq, const int* qubits, int kind, double prm,
                  const double* uni_host, int32_t* out_host) {
  using V = VT_alias<T>;
  smoq_ctx* ctx = s->ctx;
  const size_t B = (size_t)s->batch;
  const int nb = qc_chunks(s->dim, B);
  const int lc = s->n - __builtin_ctz((unsigned)nb);
  // scratch: fault | u[nq B] | out[nq B] | dec[B] | div[B] | sums[B nb 2] | norms[B nb]
  const size_t o_u = 256, o_out = o_u + up256(nq * B * 8);
  const size_t o_dec = o_out + up256(nq * B * 4), o_div = o_dec + up256(B * sizeof(QcDecision));
  const size_t o_sums = o_div + up256(B * 8), o_norm = o_sums + up256(B * nb * 16);
  int rc = ensure_scratch(ctx, o_norm + B * nb * 8);
  if (rc) return rc;
  unsigned char* base = (unsigned char*)ctx->scratch;
  unsigned long long* fault = (unsigned long long*)base;
  double* uni = (double*)(base + o_u);
  int32_t* out = (int32_t*)(base + o_out);
  QcDecision* dec = (QcDecision*)(base + o_dec);
  double* div = (double*)(base + o_div);
  double* sums = (double*)(base + o_sums);
  double* normp = (double*)(base + o_norm);
  const unsigned long long no_fault = ~0ull;
  unsigned long long fhost = no_fault;
  SMOQ_HIP(hipMemcpyAsync(fault, &no_fault, 8, hipMemcpyHostToDevice, ctx->stream));
  SMOQ_HIP(hipMemcpyAsync(uni, uni_host, nq * B * 8, hipMemcpyHostToDevice, ctx->stream));
  const dim3 grid((unsigned)nb, (unsigned)B);
  const bool need_sums = kind == SMOQ_QC_MEASURE_Z || kind == SMOQ_QC_AMPLITUDE_DAMPING;
  const bool renorm = qc_renormalises(kind);
  for (int k = 0; k < nq; ++k) {
    const int q = qubits[k];
    if (need_sums)
      hipLaunchKernelGGL(k_qc_sums<T>, grid, dim3(kQcBlock), 0, ctx->stream, (const V*)s->d, s->n,
                         lc, q, sums);
    hipLaunchKernelGGL(k_qc_decide, dim3((unsigned)B), dim3(kQcBlock), 0, ctx->stream, kind, prm,
                       uni + k * B, need_sums ? sums : nullptr, nb, dec, out + k * B, fault);
    hipLaunchKernelGGL(k_qc_apply<T>, grid, dim3(kQcBlock), 0, ctx->stream, (V*)s->d, s->n, lc, q,
                       kind, prm, dec, normp);
    if (renorm) {
      hipLaunchKernelGGL(k_qc_norm, dim3((unsigned)B), dim3(kQcBlock), 0, ctx->stream, kind, normp,
                         nb, div);
      hipLaunchKernelGGL(k_qc_scale<T>, grid, dim3(kQcBlock), 0, ctx->stream, (V*)s->d, s->n, lc,
                         div);
    }
  }
  SMOQ_CHECK_LAUNCH("qubit channel kernels");
  if (out_host)
    SMOQ_HIP(hipMemcpyAsync(out_host, out, nq * B * 4, hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipMemcpyAsync(&fhost, fault, 8, hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  if (fhost != no_fault)
    return set_error(SMOQ_ENUMERIC,
                     "smoq_qubit_channel: a measured outcome has zero probability (item %llu)",
                     fhost);
  return SMOQ_OK;
}

template <typename T>
static int qc_dm_run(smoq_state* rho, int nq, const int* qubits, int kind, const double* uni_host,
                     int32_t* out_host) {
  using V = VT_alias<T>;
  smoq_ctx* ctx = rho->ctx;
  const int N = rho->n / 2;
  const size_t B = (size_t)rho->batch;
  const int nb = qc_chunks(rho->dim, B);
  const int lc = rho->n - __builtin_ctz((unsigned)nb);
  const int nbd = qc_chunks(1ull << N, B);
  const int lcd = N - __builtin_ctz((unsigned)nbd);
  const size_t o_out = up256(nq * B * 8), o_dec = o_out + up256(nq * B * 4);
  const size_t o_sums = o_dec + up256(B * sizeof(QcDecision));
  int rc = ensure_scratch(ctx, o_sums + B * nbd * 16);
  if (rc) return rc;
  unsigned char* base = (unsigned char*)ctx->scratch;
  double* uni = (double*)base;
  int32_t* out = (int32_t*)(base + o_out);
  QcDecision* dec = (QcDecision*)(base + o_dec);
  double* sums = (double*)(base + o_sums);
  SMOQ_HIP(hipMemcpyAsync(uni, uni_host, nq * B * 8, hipMemcpyHostToDevice, ctx->stream));
  for (int k = 0; k < nq; ++k) {
    const int q = qubits[k];
    hipLaunchKernelGGL(k_qc_dm_sums<T>, dim3((unsigned)nbd, (unsigned)B), dim3(kQcBlock), 0,
                       ctx->stream, (const V*)rho->d, N, lcd, q, sums);
    hipLaunchKernelGGL(k_qc_dm_decide, dim3((unsigned)B), dim3(kQcBlock), 0, ctx->stream, kind,
                       uni + k * B, sums, nbd, dec, out + k * B);
    hipLaunchKernelGGL(k_qc_dm_project<T>, dim3((unsigned)nb, (unsigned)B), dim3(kQcBlock), 0,
                       ctx->stream, (V*)rho->d, N, lc, q, dec);
  }
  SMOQ_CHECK_LAUNCH("density-matrix projection kernels");
  if (out_host)
    SMOQ_HIP(hipMemcpyAsync(out_host, out, nq * B * 4, hipMemcpyDeviceToHost, ctx->stream));
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  return SMOQ_OK;
}

static int qc_check(const smoq_state* s, int n_logical, int nq, const int* qubits, const char* fn) {
  if (nq < 0 || (nq > 0 && !qubits)) return set_error(SMOQ_EINVAL, "%s: bad qubit list", fn);
  if (s->batch > 65535)
    return set_error(SMOQ_ENOTSUP, "%s: batch %d above 65535 items", fn, s->batch);
  for (int k = 0; k < nq; ++k)
    if (qubits[k] < 0 || qubits[k] >= n_logical)
      return set_error(SMOQ_EINVAL, "%s: qubit %d outside 0..%d", fn, qubits[k], n_logical - 1);
  return SMOQ_OK;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_qubit_channel(smoq_state* s, int n_qubits, const int* qubits, int kind,
                                  double p, const double* uniforms, int32_t* out) {
  int rc = check_state(s);
  if (rc) return rc;
  if ((rc = qc_check(s, s->n, n_qubits, qubits, "smoq_qubit_channel"))) return rc;
  if (kind < SMOQ_QC_MEASURE_Z || kind > SMOQ_QC_FORCE_ONE)
    return set_error(SMOQ_EINVAL, "smoq_qubit_channel: unknown kind %d", kind);
  const bool needs_u = kind != SMOQ_QC_FORCE_ZERO && kind != SMOQ_QC_FORCE_ONE;
  if (needs_u && n_qubits > 0 && !uniforms)
    return set_error(SMOQ_EINVAL, "smoq_qubit_channel: null uniforms");
  if (kind != SMOQ_QC_MEASURE_Z && needs_u && !(p >= 0.0 && p <= 1.0))
    return set_error(SMOQ_EINVAL, "smoq_qubit_channel: probability %g outside [0, 1]", p);
  if (n_qubits == 0) return SMOQ_OK;
  std::vector<double> zeros;
  if (!needs_u) {
    zeros.assign((size_t)n_qubits * s->batch, 0.0);
    uniforms = zeros.data();
  }
  return s->dtype == SMOQ_C128 ? qc_run<double>(s, n_qubits, qubits, kind, p, uniforms, out)
                               : qc_run<float>(s, n_qubits, qubits, kind, p, uniforms, out);
}

extern "C" int smoq_dm_qubit_channel(smoq_state* rho, int n_qubits, const int* qubits, int kind,
                                     const double* uniforms, int32_t* out) {
  int rc = check_state(rho);
  if (rc) return rc;
  if (rho->n % 2)
    return set_error(SMOQ_EINVAL, "smoq_dm_qubit_channel: a density matrix has an even number "
                                  "of vec qubits, got %d", rho->n);
  if ((rc = qc_check(rho, rho->n / 2, n_qubits, qubits, "smoq_dm_qubit_channel"))) return rc;
  if (kind != SMOQ_QC_MEASURE_Z && kind != SMOQ_QC_FORCE_ZERO && kind != SMOQ_QC_FORCE_ONE)
    return set_error(SMOQ_EINVAL, "smoq_dm_qubit_channel: kind %d is not MEASURE_Z, FORCE_ZERO "
                                  "or FORCE_ONE", kind);
  if (kind == SMOQ_QC_MEASURE_Z && n_qubits > 0 && !uniforms)
    return set_error(SMOQ_EINVAL, "smoq_dm_qubit_channel: null uniforms");
  if (n_qubits == 0) return SMOQ_OK;
  std::vector<double> zeros;
  if (kind != SMOQ_QC_MEASURE_Z) {
    zeros.assign((size_t)n_qubits * rho->batch, 0.0);
    uniforms = zeros.data();
  }
  return rho->dtype == SMOQ_C128 ? qc_dm_run<double>(rho, n_qubits, qubits, kind, uniforms, out)
                                 : qc_dm_run<float>(rho, n_qubits, qubits, kind, uniforms, out);
}
