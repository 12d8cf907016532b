/*
 * smoq_oracle.c — CPU restatement of the SmoQ.jl state-vector loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
 * library and the timed CPU baseline; nothing in the product links it.
 *
 * Each function follows one reference loop (file:line relative to the
 * reference root) with the same arithmetic as Julia's Complex{Float64}
 * methods: z*w = (zr*wr - zi*wi, zr*wi + zi*wr), x*z = (x*zr, x*zi),
 * abs2(z) = zr*zr + zi*zi.  Build with -O3 -ffp-contract=off (Julia does not
 * fuse multiply-adds).  States are interleaved (re, im) double arrays of
 * length 2*2^N; qubits are 1-based like the reference (bit k-1).
 *
 * Functions whose arithmetic lives in a third-party library the reference
 * calls are restated from that library's published algorithm and say so:
 * Julia Base cumsum (accumulate_pairwise!) and searchsortedfirst.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

static inline cx cmul(cx a, cx b) {
  cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cx rmul(double x, cx a) {
  cx r = {x * a.re, x * a.im};
  return r;
}
static inline cx cadd(cx a, cx b) {
  cx r = {a.re + b.re, a.im + b.im};
  return r;
}
static inline cx csub(cx a, cx b) {
  cx r = {a.re - b.re, a.im - b.im};
  return r;
}
static inline double abs2(cx a) { return a.re * a.re + a.im * a.im; }
static inline cx conjc(cx a) {
  cx r = {a.re, -a.im};
  return r;
}
static inline cx mk(double re, double im) {
  cx r = {re, im};
  return r;
}

#define PSI(p) ((cx*)(p))

/* ---------------------------------------------------------------- gates -- */

/* utils/cpu/cpuQuantumChannelGates.jl:142-162 */
void or_apply_rx(double* p, int k, double th, int N) {
  cx* psi = PSI(p);
  double c = cos(th / 2), s = sin(th / 2);
  cx is = mk(0.0, -s);
  int64_t step = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 0) {
      int64_t j = i + step;
      cx a = psi[i], b = psi[j];
      psi[i] = cadd(rmul(c, a), cmul(is, b));
      psi[j] = cadd(cmul(is, a), rmul(c, b));
    }
}

/* cpuQuantumChannelGates.jl:172-191 */
void or_apply_ry(double* p, int k, double th, int N) {
  cx* psi = PSI(p);
  double c = cos(th / 2), s = sin(th / 2);
  int64_t step = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 0) {
      int64_t j = i + step;
      cx a = psi[i], b = psi[j];
      psi[i] = csub(rmul(c, a), rmul(s, b));
      psi[j] = cadd(rmul(s, a), rmul(c, b));
    }
}

/* cpuQuantumChannelGates.jl:201-212 (exp(∓iθ/2) = (cos(θ/2), ∓sin(θ/2))) */
void or_apply_rz(double* p, int k, double th, int N) {
  cx* psi = PSI(p);
  cx ph[2] = {mk(cos(th / 2), -sin(th / 2)), mk(cos(th / 2), sin(th / 2))};
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) psi[i] = cmul(psi[i], ph[(i >> (k - 1)) & 1]);
}

/* cpuQuantumChannelGates.jl:222-238 */
void or_apply_h(double* p, int k, int N) {
  cx* psi = PSI(p);
  double s = 1 / sqrt(2.0);
  int64_t step = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 0) {
      int64_t j = i + step;
      cx a = psi[i], b = psi[j];
      psi[i] = rmul(s, cadd(a, b));
      psi[j] = rmul(s, csub(a, b));
    }
}

/* cpuQuantumChannelGates.jl:245-256 */
void or_apply_x(double* p, int k, int N) {
  cx* psi = PSI(p);
  int64_t step = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 0) {
      cx t = psi[i];
      psi[i] = psi[i + step];
      psi[i + step] = t;
    }
}

/* cpuQuantumChannelGates.jl:265-279: ψ[i] = -im*v1, ψ[j] = im*v0 */
void or_apply_y(double* p, int k, int N) {
  cx* psi = PSI(p);
  int64_t step = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 0) {
      cx v0 = psi[i], v1 = psi[i + step];
      psi[i] = mk(v1.im, -v1.re);
      psi[i + step] = mk(-v0.im, v0.re);
    }
}

/* cpuQuantumChannelGates.jl:286-390: Z, S, S†, T, T† multiply bit-1 amplitudes */
void or_apply_phase1(double* p, int k, double re, double im, int N) {
  cx* psi = PSI(p);
  cx ph = mk(re, im);
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (k - 1)) & 1) == 1) psi[i] = cmul(psi[i], ph);
}

/* cpuQuantumChannelGates.jl:870-882 */
void or_apply_cnot(double* p, int c, int t, int N) {
  cx* psi = PSI(p);
  int64_t step = (int64_t)1 << (t - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (c - 1)) & 1) == 1 && ((i >> (t - 1)) & 1) == 0) {
      cx a = psi[i];
      psi[i] = psi[i + step];
      psi[i + step] = a;
    }
}

/* cpuQuantumChannelGates.jl:889-900 */
void or_apply_cz(double* p, int a, int b, int N) {
  cx* psi = PSI(p);
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) {
    int64_t bi = (i >> (a - 1)) & 1, bj = (i >> (b - 1)) & 1;
    double sign = (double)(1 - 2 * (bi & bj));
    psi[i] = rmul(sign, psi[i]);
  }
}

/* cpuQuantumChannelGates.jl:914-937 */
void or_apply_cry(double* p, int c, int t, double th, int N) {
  cx* psi = PSI(p);
  double cc = cos(th / 2), s = sin(th / 2);
  int64_t step = (int64_t)1 << (t - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (c - 1)) & 1) == 1 && ((i >> (t - 1)) & 1) == 0) {
      cx a = psi[i], b = psi[i + step];
      psi[i] = csub(rmul(cc, a), rmul(s, b));
      psi[i + step] = cadd(rmul(s, a), rmul(cc, b));
    }
}

/* cpuQuantumChannelGates.jl:992-1026 (which = 0) and :1039-1071 (which = 1) */
void or_apply_rxx_ryy(double* p, int qi, int qj, double th, int which, int N) {
  cx* psi = PSI(p);
  double c = cos(th / 2), s = sin(th / 2);
  cx is = mk(0.0, -s), isn = mk(0.0, s);
  cx far = which ? isn : is;
  int64_t si = (int64_t)1 << (qi - 1), sj = (int64_t)1 << (qj - 1), dim = (int64_t)1 << N;
  for (int64_t b = 0; b < dim; ++b) {
    if (((b >> (qi - 1)) & 1) != 0 || ((b >> (qj - 1)) & 1) != 0) continue;
    int64_t i00 = b, i01 = b + sj, i10 = b + si, i11 = b + si + sj;
    cx a00 = psi[i00], a01 = psi[i01], a10 = psi[i10], a11 = psi[i11];
    psi[i00] = cadd(rmul(c, a00), cmul(far, a11));
    psi[i01] = cadd(rmul(c, a01), cmul(is, a10));
    psi[i10] = cadd(cmul(is, a01), rmul(c, a10));
    psi[i11] = cadd(cmul(far, a00), rmul(c, a11));
  }
}

/* cpuQuantumChannelGates.jl:1084-1100 */
void or_apply_rzz(double* p, int qi, int qj, double th, int N) {
  cx* psi = PSI(p);
  cx same = mk(cos(th / 2), -sin(th / 2)), diff = mk(cos(th / 2), sin(th / 2));
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) {
    int64_t bi = (i >> (qi - 1)) & 1, bj = (i >> (qj - 1)) & 1;
    psi[i] = cmul(psi[i], bi == bj ? same : diff);
  }
}

/* cpuQuantumChannelGates.jl:448-461 */
void or_apply_ccz(double* p, int a, int b, int c, int N) {
  cx* psi = PSI(p);
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (a - 1)) & 1) && ((i >> (b - 1)) & 1) && ((i >> (c - 1)) & 1))
      psi[i] = mk(-psi[i].re, -psi[i].im);
}

/* cpuQuantumChannelGates.jl:471-487 */
void or_apply_toffoli(double* p, int c1, int c2, int t, int N) {
  cx* psi = PSI(p);
  int64_t step = (int64_t)1 << (t - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i)
    if (((i >> (c1 - 1)) & 1) && ((i >> (c2 - 1)) & 1) && !((i >> (t - 1)) & 1)) {
      cx a = psi[i];
      psi[i] = psi[i + step];
      psi[i + step] = a;
    }
}

/* cpuQuantumChannelGates.jl:739-771 (plus = 0: σ⁻, plus = 1: σ⁺) */
void or_apply_sigma(double* p, int k, int plus, int N) {
  cx* psi = PSI(p);
  int64_t mask = (int64_t)1 << (k - 1), dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) {
    if (!plus && (i & mask) != 0) {
      int64_t j = i ^ mask;
      psi[j] = psi[i];
      psi[i] = mk(0, 0);
    } else if (plus && (i & mask) == 0) {
      int64_t j = i | mask;
      psi[j] = psi[i];
      psi[i] = mk(0, 0);
    }
  }
}

/* generic 2x2, row-major (re,im): apply_one_qubit_gate_bitwise_cpu!
 * (utils/cpu/cpuQuantumChannelUnitaryEvolutionTrotter.jl:77-99) */
void or_apply_1q(double* p, const double* u, int q, int N) {
  cx* psi = PSI(p);
  cx u00 = mk(u[0], u[1]), u01 = mk(u[2], u[3]), u10 = mk(u[4], u[5]), u11 = mk(u[6], u[7]);
  int64_t step = (int64_t)1 << (q - 1), dim = (int64_t)1 << N;
  for (int64_t b = 0; b < dim; ++b)
    if ((b & step) == 0) {
      cx a0 = psi[b], a1 = psi[b + step];
      psi[b] = cadd(cmul(u00, a0), cmul(u01, a1));
      psi[b + step] = cadd(cmul(u10, a0), cmul(u11, a1));
    }
}

/* apply_two_qubit_gate_bitwise_cpu! (…Trotter.jl:107-142), including the
 * swap of q1/q2 into ascending order without permuting U.  U is 4x4
 * row-major (re,im); basis index 1 = lower qubit set. */
void or_apply_2q_trotter(double* p, const double* u, int q1, int q2, int N) {
  cx* psi = PSI(p);
  if (q1 > q2) { int t = q1; q1 = q2; q2 = t; }
  int64_t s1 = (int64_t)1 << (q1 - 1), s2 = (int64_t)1 << (q2 - 1), dim = (int64_t)1 << N;
  cx U[16];
  for (int i = 0; i < 16; ++i) U[i] = mk(u[2 * i], u[2 * i + 1]);
  for (int64_t b = 0; b < dim; ++b) {
    if ((b & s1) || (b & s2)) continue;
    int64_t id[4] = {b, b + s1, b + s2, b + s1 + s2};
    cx a[4] = {psi[id[0]], psi[id[1]], psi[id[2]], psi[id[3]]};
    for (int r = 0; r < 4; ++r) {
      cx acc = cmul(U[4 * r], a[0]);
      acc = cadd(acc, cmul(U[4 * r + 1], a[1]));
      acc = cadd(acc, cmul(U[4 * r + 2], a[2]));
      acc = cadd(acc, cmul(U[4 * r + 3], a[3]));
      psi[id[r]] = acc;
    }
  }
}

/* apply_fused_gate! block/offset loop
 * (SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:318-339) */
void or_apply_fused_1q(double* p, const double* u, int q, int N) {
  cx* psi = PSI(p);
  int64_t dim = (int64_t)1 << N, step = (int64_t)1 << (q - 1), bs = step << 1, nb = dim >> q;
  cx u11 = mk(u[0], u[1]), u12 = mk(u[2], u[3]), u21 = mk(u[4], u[5]), u22 = mk(u[6], u[7]);
  for (int64_t blk = 0; blk < nb; ++blk) {
    int64_t base = blk * bs;
    for (int64_t off = 0; off < step; ++off) {
      int64_t i = base + off, j = i + step;
      cx a = psi[i], b = psi[j];
      psi[i] = cadd(cmul(u11, a), cmul(u12, b));
      psi[j] = cadd(cmul(u21, a), cmul(u22, b));
    }
  }
}

/* ------------------------------------------------------------ observables -- */

/* _expect_Z / _expect_X / _expect_Y (cpuQuantumStateObservables.jl:201-266) */
double or_expect_local(const double* p, int k, int N, char pauli) {
  const cx* psi = (const cx*)p;
  int64_t dim = (int64_t)1 << N, bp = k - 1, step = (int64_t)1 << bp;
  double r = 0.0;
  if (pauli == 'z') {
    for (int64_t i = 0; i < dim; ++i) r += abs2(psi[i]) * (double)(1 - 2 * ((i >> bp) & 1));
  } else if (pauli == 'x') {
    for (int64_t i = 0; i < dim; ++i)
      if (((i >> bp) & 1) == 0) r += 2 * cmul(conjc(psi[i]), psi[i + step]).re;
  } else {
    for (int64_t i = 0; i < dim; ++i)
      if (((i >> bp) & 1) == 0) r += 2 * cmul(conjc(psi[i]), psi[i + step]).im;
  }
  return r;
}

/* _expect_ZZ / _expect_XX / _expect_YY (cpuQuantumStateObservables.jl:341-445) */
double or_expect_corr(const double* p, int i, int j, int N, char pair) {
  const cx* psi = (const cx*)p;
  int64_t dim = (int64_t)1 << N, bi = i - 1, bj = j - 1;
  int64_t si = (int64_t)1 << bi, sj = (int64_t)1 << bj;
  double r = 0.0;
  if (pair == 'z') {
    for (int64_t s = 0; s < dim; ++s) {
      int64_t x = ((s >> bi) & 1) ^ ((s >> bj) & 1);
      r += abs2(psi[s]) * (double)(1 - 2 * x);
    }
  } else {
    double f = pair == 'x' ? 2.0 : -2.0;
    for (int64_t s = 0; s < dim; ++s)
      if (((s >> bi) & 1) == 0 && ((s >> bj) & 1) == 0) {
        r += f * cmul(conjc(psi[s]), psi[s + si + sj]).re;
        r += 2.0 * cmul(conjc(psi[s + sj]), psi[s + si]).re;
      }
  }
  return r;
}

/* get_expectation_pauli_string (cpuQuantumStateObservables.jl:840-1049).
 * ops[t] in {'X','Y','Z'}, qubits 1-based.  Reproduces the reference value,
 * including the missing factor 2 for strings with X/Y (its :1044). */
double or_pauli_string_ref(const double* p, const char* ops, const int* qubits, int k, int N) {
  const cx* psi = (const cx*)p;
  int64_t dim = (int64_t)1 << N;
  if (k == 0) return 1.0;
  int zq[64], xq[64], nz = 0, nxy = 0;
  char xo[64];
  for (int t = 0; t < k; ++t) {
    if (ops[t] == 'Z' || ops[t] == 'z') zq[nz++] = qubits[t];
    else { xo[nxy] = ops[t]; xq[nxy++] = qubits[t]; }
  }
  double r = 0.0;
  if (nxy == 0) {
    for (int64_t i = 0; i < dim; ++i) {
      int64_t par = 0;
      for (int t = 0; t < nz; ++t) par ^= (i >> (zq[t] - 1)) & 1;
      r += abs2(psi[i]) * (double)(1 - 2 * par);
    }
    return r;
  }
  for (int64_t base = 0; base < dim; ++base) {
    int allz = 1;
    for (int t = 0; t < nxy; ++t)
      if ((base >> (xq[t] - 1)) & 1) { allz = 0; break; }
    if (!allz) continue;
    for (int64_t fp = 0; fp < ((int64_t)1 << nxy); ++fp) {
      int64_t bra = base, ket = base;
      for (int t = 0; t < nxy; ++t)
        if ((fp >> t) & 1) ket += (int64_t)1 << (xq[t] - 1);
      cx coeff = mk(1, 0);
      int zero = 0;
      for (int t = 0; t < nxy; ++t) {
        int bb = (int)((bra >> (xq[t] - 1)) & 1), kb = (int)((ket >> (xq[t] - 1)) & 1);
        if (bb == kb) { zero = 1; break; }
        if (xo[t] == 'Y' || xo[t] == 'y') coeff = cmul(coeff, kb == 0 ? mk(0, 1) : mk(0, -1));
      }
      if (zero) continue;
      int zs = 1;
      for (int t = 0; t < nz; ++t) zs *= 1 - 2 * (int)((ket >> (zq[t] - 1)) & 1);
      cx term = cmul(rmul((double)zs, coeff), cmul(conjc(psi[bra]), psi[ket]));
      r += term.re;
    }
  }
  return r;
}

/* compute_energy (scripts_variational_quantum_circuits/demo_variational_quantum_eigensolver.jl:251-289) */
double or_compute_energy(const double* p, int N, double Jx, double Jy, double Jz, double hx) {
  const cx* psi = (const cx*)p;
  int64_t dim = (int64_t)1 << N;
  double E = 0.0;
  for (int i = 1; i <= N; ++i)
    for (int64_t k = 0; k < dim; ++k) {
      int64_t j = k ^ ((int64_t)1 << (i - 1));
      E += hx * cmul(conjc(psi[k]), psi[j]).re;
    }
  for (int i = 1; i <= N - 1; ++i) {
    int j = i + 1;
    for (int64_t k = 0; k < dim; ++k) {
      int64_t bi = (k >> (i - 1)) & 1, bj = (k >> (j - 1)) & 1;
      double zi = (double)(1 - 2 * bi), zj = (double)(1 - 2 * bj);
      E += Jz * zi * zj * abs2(psi[k]);
      int64_t kxx = (k ^ ((int64_t)1 << (i - 1))) ^ ((int64_t)1 << (j - 1));
      E += Jx * cmul(conjc(psi[k]), psi[kxx]).re;
      cx pi = bi == 0 ? mk(0, 1) : mk(0, -1), pj = bj == 0 ? mk(0, 1) : mk(0, -1);
      cx ph = cmul(pi, pj);
      E += Jy * cmul(cmul(conjc(psi[k]), ph), psi[kxx]).re;
    }
  }
  return E;
}

/* ---------------------------------------------------------------- sampling -- */

/* abs2.(ψ) */
void or_probs(const double* p, int64_t dim, double* out) {
  const cx* psi = (const cx*)p;
  for (int64_t i = 0; i < dim; ++i) out[i] = abs2(psi[i]);
}

/* Julia Base cumsum for a Float64 vector = accumulate_pairwise!(add_sum, c, v)
 * (Julia 1.12.1 base/accumulate.jl; the reference's pinned Julia,
 * Manifest.toml:3).  Leaves hold n < 128 elements. */
static double pw_rec(double* c, const double* v, double s, int64_t i1, int64_t n) {
  double s_;
  if (n < 128) {
    s_ = v[i1];
    c[i1] = s + s_;
    for (int64_t i = i1 + 1; i < i1 + n; ++i) {
      s_ = s_ + v[i];
      c[i] = s + s_;
    }
  } else {
    int64_t n2 = n >> 1;
    s_ = pw_rec(c, v, s, i1, n2);
    s_ = s_ + pw_rec(c, v, s + s_, i1 + n2, n - n2);
  }
  return s_;
}

void or_cumsum_pairwise(const double* v, int64_t n, double* c) {
  if (n == 0) return;
  c[0] = v[0];
  if (n == 1) return;
  pw_rec(c, v, v[0], 1, n - 1);
}

void or_cumsum_sequential(const double* v, int64_t n, double* c) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    s = s + v[i];
    c[i] = s;
  }
}

/* Julia Base searchsortedfirst(v, x) with the (lo, len) halving loop,
 * lt = isless.  Returns a 1-based index (n+1 when x exceeds every entry). */
int64_t or_searchsortedfirst(const double* v, int64_t n, double x) {
  int64_t lo = 1, len = n;
  while (len != 0) {
    int64_t half = len >> 1, m = lo + half;
    if (v[m - 1] < x) {
      lo = m + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

/* run_circuit_sampling's draw (run_benchmark_smoq_Float64.jl:424-429):
 * idx = searchsortedfirst(cumsum(abs2.(ψ)), r) - 1, 0-based. */
void or_sample_cumsum_search(const double* p, int N, const double* r, int64_t nd, int64_t* out) {
  int64_t dim = (int64_t)1 << N;
  double* pr = (double*)malloc(sizeof(double) * dim);
  double* c = (double*)malloc(sizeof(double) * dim);
  or_probs(p, dim, pr);
  or_cumsum_pairwise(pr, dim, c);
  for (int64_t d = 0; d < nd; ++d) out[d] = or_searchsortedfirst(c, dim, r[d]) - 1;
  free(pr);
  free(c);
}

/* sample_from_probabilities (utils/cpu/cpuQuantumStateMeasurements.jl:203-213),
 * one linear scan per draw; 0-based result. */
void or_sample_sequential(const double* p, int N, const double* r, int64_t nd, int64_t* out) {
  int64_t dim = (int64_t)1 << N;
  double* pr = (double*)malloc(sizeof(double) * dim);
  or_probs(p, dim, pr);
  for (int64_t d = 0; d < nd; ++d) {
    double cs = 0.0;
    int64_t hit = dim - 1;
    for (int64_t i = 0; i < dim; ++i) {
      cs += pr[i];
      if (r[d] <= cs) { hit = i; break; }
    }
    out[d] = hit;
  }
  free(pr);
}

/* ---------------------------------------------------------- H·ψ and MCWF -- */

/* apply_hamiltonian_bitwise! pull loop
 * (utils/cpu/cpuQuantumChannelUnitaryEvolutionChebyshev.jl:55-152).
 * bonds: nb x (i, j) 1-based + (Jxx, Jyy, Jzz); fields: N x (hx, hy, hz). */
void or_hpsi(double* out, const double* in, int N, int nb, const int* bi, const int* bj,
             const double* J, const double* h) {
  const cx* x = (const cx*)in;
  cx* y = (cx*)out;
  int64_t dim = (int64_t)1 << N;
  for (int64_t k = 0; k < dim; ++k) {
    cx acc = mk(0, 0);
    for (int b = 0; b < nb; ++b) {
      int ib = bi[b] - 1, jb = bj[b] - 1;
      int64_t mf = ((int64_t)1 << ib) | ((int64_t)1 << jb);
      int64_t vi = (k >> ib) & 1, vj = (k >> jb) & 1;
      double Jxx = J[3 * b], Jyy = J[3 * b + 1], Jzz = J[3 * b + 2];
      if (Jzz != 0.0) acc = cadd(acc, rmul(Jzz * (vi == vj ? 1.0 : -1.0), x[k]));
      int64_t kf = k ^ mf;
      if (Jxx != 0.0) acc = cadd(acc, rmul(Jxx, x[kf]));
      if (Jyy != 0.0) acc = cadd(acc, rmul(Jyy * (vi == vj ? -1.0 : 1.0), x[kf]));
    }
    for (int f = 0; f < N; ++f) {
      int64_t mask = (int64_t)1 << f, v = (k >> f) & 1;
      double hx = h[3 * f], hy = h[3 * f + 1], hz = h[3 * f + 2];
      if (hz != 0.0) acc = cadd(acc, rmul(hz * (v == 0 ? 1.0 : -1.0), x[k]));
      int64_t kf = k ^ mask;
      if (hx != 0.0) acc = cadd(acc, rmul(hx, x[kf]));
      if (hy != 0.0) acc = cadd(acc, cmul(rmul(hy, v == 0 ? mk(0, 1) : mk(0, -1)), x[kf]));
    }
    y[k] = acc;
  }
}

/* population_excited (…MonteCarloWaveFunction.jl:428-439) */
static double pop_excited(const cx* psi, int k, int64_t dim) {
  int64_t mask = (int64_t)1 << (k - 1);
  double r = 0.0;
  for (int64_t i = 0; i < dim; ++i)
    if (i & mask) r += abs2(psi[i]);
  return r;
}

/* One trajectory of lindblad_mcwf_batched_step!
 * (…MonteCarloWaveFunction.jl:340-390) after the Trotter step:
 * decay by decay_pow[popcount(i)], dp = 1 - ||ψ||^2, jump decision with the
 * supplied r (site chosen by the cumulative γdt pop_k / norm² with the same
 * r, fallback site 1), σ⁻, renormalise by the norm.  decay_pow[n] holds
 * Julia's decay_factor^n.  Returns the 1-based jump site, 0 for none.
 * dot/norm are summed sequentially here (BLAS order is unpinned). */
int or_mcwf_decay_jump(double* p, int N, double gamma_dt, const double* decay_pow, double r,
                       double* dp_out) {
  cx* psi = PSI(p);
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) {
    int ne = __builtin_popcountll((uint64_t)i);
    if (ne > 0) psi[i] = rmul(decay_pow[ne], psi[i]);
  }
  double nsq = 0.0;
  for (int64_t i = 0; i < dim; ++i) nsq += abs2(psi[i]);
  double dp = 1.0 - nsq;
  if (dp_out) *dp_out = dp;
  int site = 0;
  if (r < dp && dp > 0) {
    double cum = 0.0;
    int sel = 1;
    for (int k = 1; k <= N; ++k) {
      double pk = gamma_dt * pop_excited(psi, k, dim) / nsq;
      cum += pk;
      if (r < cum) { sel = k; break; }
    }
    or_apply_sigma(p, sel, 0, N);
    site = sel;
  }
  double nn = 0.0;
  for (int64_t i = 0; i < dim; ++i) nn += abs2(psi[i]);
  nn = sqrt(nn);
  for (int64_t i = 0; i < dim; ++i) psi[i] = mk(psi[i].re / nn, psi[i].im / nn);
  return site;
}

/* One step of lindblad_mcwf_step_trotter! / lindblad_mcwf_step!
 * (…MonteCarloWaveFunction.jl:222-306) after the coherent step, σ_z
 * dephasing: ψ *= decay (= sqrt(1 - Nγdt), uniform), dp = 1 - ||ψ||²,
 * jump if r < dp && dp > 0 on site k = floor(u2·N) (clamped to N-1; the
 * reference draws rand(1:N)), σ_z_k flips the sign where bit k is set
 * (apply_sigma_z!, :170-180), renormalise.  Returns the 1-based site or 0. */
int or_mcwf_dephasing(double* p, int N, double decay, double r, double u2, double* dp_out) {
  cx* psi = PSI(p);
  int64_t dim = (int64_t)1 << N;
  for (int64_t i = 0; i < dim; ++i) psi[i] = rmul(decay, psi[i]);
  double nsq = 0.0;
  for (int64_t i = 0; i < dim; ++i) nsq += abs2(psi[i]);
  double dp = 1.0 - nsq;
  if (dp_out) *dp_out = dp;
  int site = 0;
  if (r < dp && dp > 0) {
    int k = (int)(u2 * N);
    if (k > N - 1) k = N - 1;
    int64_t mask = (int64_t)1 << k;
    for (int64_t i = 0; i < dim; ++i)
      if (i & mask) psi[i] = mk(-psi[i].re, -psi[i].im);
    site = k + 1;
  }
  double nn = 0.0;
  for (int64_t i = 0; i < dim; ++i) nn += abs2(psi[i]);
  nn = sqrt(nn);
  for (int64_t i = 0; i < dim; ++i) psi[i] = mk(psi[i].re / nn, psi[i].im / nn);
  return site;
}
