// Pass kernels compiled at run time for one circuit structure.
//
// The precompiled pass kernel (fused.hip) walks the stage and gate lists of
// its program with a per-gate switch; the loop-carried register array then
// costs two register copies per amplitude and gate at the switch joins,
// which outweighs the gate arithmetic.  Here the same kernel is generated
// with the stage and gate lists unrolled: one straight-line function per
// pass that calls the gate helpers with the gate kind and register slots as
// template arguments, compiled with hiprtc for gfx950 and cached per context
// by its source text.  The source holds only the structure (kinds, slots,
// control presence, table widths, stage and gate indices); passes, stages,
// gate matrices and phase tables are still read from the device program, so
// plans with new angles reuse the compiled module.
#include <hip/hiprtc.h>

#include <unistd.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>
#include <thread>
#include <type_traits>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

// the device headers, embedded at build time
static const char kSrcDeviceCommon[] =
#include "device_common.h.inc"
    ;
static const char kSrcFusedTypes[] =
#include "fused_types.h.inc"
    ;
static const char kSrcFusedOps[] =
#include "fused_ops.h.inc"
    ;

static const char* u1_kind_name(int kind) {
  switch (kind) {
    case GK_U1_REAL: return "GK_U1_REAL";
    case GK_U1_RXT: return "GK_U1_RXT";
    default: return "GK_U1_GEN";
  }
}

// The kernel cache key: the generator's output in binary — every literal it
// writes by address, every value by its bytes, every stream manipulator by
// address — so two programs share a key exactly when the generator would
// write the same source for them, without formatting that source.
struct KeyOut {
  std::string k;
  void put(const void* p, size_t n) { k.append(static_cast<const char*>(p), n); }
  KeyOut& operator<<(const char* s) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(s);
    put(&a, sizeof a);
    return *this;
  }
  KeyOut& operator<<(const std::string& s) {
    const uint64_t n = s.size();
    put(&n, sizeof n);
    put(s.data(), s.size());
    return *this;
  }
  template <typename I, typename std::enable_if<std::is_integral<I>::value, int>::type = 0>
  KeyOut& operator<<(I v) {
    const uint64_t w = (uint64_t)v;
    const uint8_t sz = sizeof(I);
    put(&sz, 1);
    put(&w, sizeof w);
    return *this;
  }
  KeyOut& operator<<(std::ios_base& (*f)(std::ios_base&)) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(f);
    put(&a, sizeof a);
    return *this;
  }
};

// "gates[g]" in the source, g in the key
struct GateRef {
  int g;
};
static std::ostream& operator<<(std::ostream& o, GateRef r) { return o << "gates[" << r.g << "]"; }
static KeyOut& operator<<(KeyOut& o, GateRef r) { return o << "gates[" << r.g; }

// Bit-field expression for a linear map of the thread index: thread bit i
// goes to bit dst[i]; runs of consecutive bits move as one field.
template <typename O>
static void bit_map_emit(O& o, const std::vector<int>& dst, bool wide) {
  bool first = true;
  size_t i = 0;
  while (i < dst.size()) {
    size_t j = i + 1;
    while (j < dst.size() && dst[j] == dst[j - 1] + 1) ++j;
    const int len = (int)(j - i);
    if (!first) o << " | ";
    first = false;
    o << "(" << (wide ? "(uint64_t)" : "") << "((tid >> " << i << ") & 0x" << std::hex
      << ((1u << len) - 1u) << std::dec << "u) << " << dst[i] << ")";
    i = j;
  }
  if (first) o << (wide ? "(uint64_t)0" : "0u");
}

// Uncontrolled single-qubit gates with the structure of an Ry or Rx product
// become factored kinds: sigma * N with |t| <= 1 in N, one fused
// multiply-add per output component instead of a multiply and a fused
// multiply-add; the sigmas of a pass multiply into its scale (each
// |sigma| >= 1/sqrt(2)).  The form (unit diagonal or unit anti-diagonal)
// is compiled into the kernel, so the source depends on the angles: only
// circuits whose angles stay fixed across calls are rewritten.
//
// A pass's tile holds the unnormalised amplitudes until its store, which
// grow by 1/|sigma| (up to sqrt 2) per factored gate, and complex64 tiles
// keep them in float.  Factoring therefore stops once the product of a
// pass's sigmas would fall below 2^-40 (complex64) or 2^-500 (complex128);
// later rotations of that pass stay unfactored.
//
// Register diagonals (GK_DIAGREG) of unit-modulus pattern-0 phase become
// GK_DIAGREG1: that phase joins the pass's complex scale and every pattern
// whose phase equals it exactly is skipped, so that a one-qubit phase costs
// four complex products per thread instead of eight.
void factor_rotations(Program& prog, int dtype, std::vector<FactorUndo>* undo) {
  const int max_log2 = dtype == SMOQ_C64 ? 40 : 500;
  for (PassDev& pd : prog.passes) {
    double scale = 1.0;
    int scale_exp = 0;  // scale = frexp mantissa * 2^scale_exp, tracked exactly
    std::complex<double> phase(1.0, 0.0);
    for (int s = pd.s0; s < pd.s1; ++s) {
      for (int g = prog.stages[s].g0; g < prog.stages[s].g1; ++g) {
        GateDev& G = prog.gates[g];
        if (G.kind == GK_DIAGREG) {
          const std::complex<double> p0(G.m[0], G.m[1]);
          if (!(std::fabs(std::abs(p0) - 1.0) <= 1e-6)) continue;
          FactorUndo u{g, G.kind, G.b, G.m[0], {}};
          std::memcpy(u.md, G.m, sizeof u.md);
          uint32_t mask = 0;
          for (int j = 0; j < (1 << kRegQubits); ++j) {
            const std::complex<double> pj(G.m[2 * j], G.m[2 * j + 1]);
            const std::complex<double> q =
                (pj.real() == p0.real() && pj.imag() == p0.imag()) ? 1.0 : pj / p0;
            if (q != 1.0) mask |= 1u << j;
            G.m[2 * j] = q.real();
            G.m[2 * j + 1] = q.imag();
          }
          if (undo) undo->push_back(u);
          G.kind = GK_DIAGREG1;
          G.b = (int32_t)mask;
          phase *= p0;
          continue;
        }
        if (G.kind == GK_U2) {
          // the exactly-zero entries become structure
          uint32_t z = 0;
          for (int e = 0; e < 16; ++e)
            if (G.m[2 * e] == 0.0 && G.m[2 * e + 1] == 0.0) z |= 1u << e;
          if (z) {
            if (undo) undo->push_back({g, G.kind, G.b, G.m[0], {}});
            G.zero = z;
          }
          continue;
        }
        if (G.creg != 0 || G.cglob != 0) continue;
        const double* m = G.m;
        double a, b;
        int kind;
        if (G.kind == GK_U1_REAL && m[0] == m[6] && m[2] == -m[4]) {
          a = m[0];
          b = m[4];
          kind = GK_U1_ROT;
        } else if (G.kind == GK_U1_RXT && m[0] == m[6] && m[3] == m[5]) {
          a = m[0];
          b = m[3];
          kind = GK_U1_RXR;
        } else {
          continue;
        }
        const bool diag = std::fabs(a) >= std::fabs(b);
        const double sigma = diag ? a : b, t = diag ? b / a : a / b;
        if (!(sigma != 0.0) || !std::isfinite(t)) continue;
        int e = 0;
        const double mant = std::frexp(scale * sigma, &e);
        if (-(scale_exp + e) >= max_log2) continue;
        if (undo) undo->push_back({g, G.kind, G.b, G.m[0], {}});
        G.kind = kind;
        G.b = diag ? 0 : 1;
        G.m[0] = t;
        scale = mant;
        scale_exp += e;
      }
    }
    const double mag = std::ldexp(scale, scale_exp);
    pd.scale = mag * phase.real();
    pd.scale_im = mag * phase.imag();
  }
}

void unfactor_rotations(Program& prog, const std::vector<FactorUndo>& undo) {
  for (const FactorUndo& u : undo) {
    GateDev& G = prog.gates[u.g];
    if (G.kind == GK_DIAGREG1) std::memcpy(G.m, u.md, sizeof u.md);
    G.zero = 0;
    G.kind = u.kind;
    G.b = u.b;
    G.m[0] = u.m0;
  }
  for (PassDev& pd : prog.passes) {
    pd.scale = 1.0;
    pd.scale_im = 0.0;
  }
}


template <typename O>
static void jit_emit(const Program& prog, int dtype, O& o) {
  const char* T = dtype == SMOQ_C128 ? "double" : "float";
  o << "#include \"fused_ops.h\"\nusing namespace smoq;\n";
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    const PassDev& pd = prog.passes[p];
    // two 2^k-amplitude tiles per CU (LDS) need four waves per SIMD: keep the
    // registers within that budget
    o << "extern \"C\" __global__ void __launch_bounds__(" << (1 << (pd.k - kRegQubits))
      << ", 4) smoq_pass_" << p << "(VT<" << T
      << ">* __restrict__ psi, int n, const PassDev* __restrict__ pass,"
         " const StageDev* __restrict__ stages, const GateDev* __restrict__ gates,"
         " const double2* __restrict__ tabs, double* __restrict__ opart,"
         " unsigned long long item_bytes) {\n"
      << "  typedef " << T << " T;\n"
      // the tile as a static LDS array: its address is known when the
      // kernel is compiled, so an LDS access is one XOR of a lane part with
      // a uniform part (the pass is launched without dynamic LDS)
      << "  __shared__ __align__(16) VT<T> lds[" << (1 << pd.k) << "];\n"
      << "  PassFrame<T> f;\n";
    // adjoint overlaps of this pass: one LDS slot per (overlap, 16-lane row),
    // or per (overlap, wave) when that many row slots would cost the second
    // workgroup per CU its LDS
    const int ntp = 1 << (pd.k - kRegQubits);
    std::vector<int> oslots;
    for (int s = pd.s0; s < pd.s1; ++s)
      for (int g = prog.stages[s].g0; g < prog.stages[s].g1; ++g)
        if (prog.gates[g].kind == GK_OVL) oslots.push_back(prog.gates[g].oslot);
    const bool ovl_rows = ntp >= 64 && oslots.size() * (size_t)(ntp / 16) * 8 <= 12288;
    const int nwaves = ntp >= 64 ? (ovl_rows ? ntp / 16 : ntp / 64) : 1;
    if (!oslots.empty())
      o << "  __shared__ double ovl_acc[" << (int)oslots.size() * nwaves << "];\n";
    int ovl_i = 0;
    // Passes of few tiles (small states) run one workgroup per CU, and each
    // stage would wait a full memory latency for its gate records on their
    // first scalar load.  Every thread first loads one dword per 128 B line
    // of the pass's gate records, so that the stages' scalar loads find them
    // in L2; the loaded values are consumed after the store.
    const int g0 = prog.stages[pd.s0].g0, g1 = prog.stages[pd.s1 - 1].g1;
    const size_t nw = (size_t)(g1 - g0) * sizeof(GateDev) / 4;
    const size_t nt = (size_t)1 << (pd.k - kRegQubits);
    const bool warm = pd.n_nonlocal <= kWarmMaxNonlocal && nw <= nt * 32 * kWarmLoads;
    if (warm) {
      o << "  uint32_t warm = 0;\n  {\n    const uint32_t* gw = reinterpret_cast<const uint32_t*>(gates + "
        << g0 << ");\n    for (uint32_t w = threadIdx.x * 32u; w < " << nw
        << "u; w += blockDim.x * 32u) warm ^= gw[w];\n  }\n";
    }
    o << "  pass_load<T>(lds, psi, n, pass, f);\n";
    // batches whose items carry their own gate records and tables (same
    // structure, other angles): item_bytes apart.  Only in programs built
    // for it, so that the other passes address their records from the
    // kernel arguments alone.
    if (prog.per_item)
      o << "  {\n"
        << "    const unsigned long long item = f.tile >> pass->n_nonlocal;\n"
        << "    gates = (const GateDev*)((const char*)gates + item * item_bytes);\n"
        << "    tabs = (const double2*)((const char*)tabs + item * item_bytes);\n"
        << "  }\n";
    for (int s = pd.s0; s < pd.s1; ++s) {
      const StageDev& st = prog.stages[s];
      const bool seg_end = s + 1 == pd.s1 || !prog.stages[s + 1].cont;
      // thread bit i -> local position tpos[i] -> its global bit
      std::vector<int> lpos, gpos;
      for (int i = 0; i < st.ntb; ++i) {
        lpos.push_back(st.tpos[i]);
        gpos.push_back(pd.map[st.tpos[i]]);
      }
      if (!st.cont)
        o << "  {\n    VT<double> x[kNR];\n    StageFrame sf;\n"
          << "    const uint32_t tid = threadIdx.x;\n";
      else
        o << "    lane_swap<" << (int)st.swap_slot << ", " << (int)st.swap_bit << ">(x);\n";
      o << "    {\n      const uint32_t sr[kNR] = {";
      for (int i = 0; i < (1 << kRegQubits); ++i) {
        uint32_t r = 0;
        for (int j = 0; j < kRegQubits; ++j)
          if ((i >> j) & 1) r ^= st.sw_reg[j];
        o << (i ? ", " : "") << r << "u";
      }
      if (!st.cont) {
        o << "};\n      stage_load_c<T>(lds, ";
        bit_map_emit(o, lpos, false);
        o << ", f.gbase | ";
        bit_map_emit(o, gpos, true);
        o << ", sr, x, sf);\n    }\n";
      } else {
        // after a lane exchange only thread bit swap_bit maps to another
        // local position: the lane part and the global index change by that
        // bit's images (swz is linear over GF(2)), the register offsets are
        // constants
        const StageDev& pv = prog.stages[s - 1];
        const int lb = st.swap_bit;
        bool one_bit = pv.ntb == st.ntb;
        for (int i = 0; i < st.ntb && one_bit; ++i)
          if (i != lb) one_bit = pv.sw_tid[i] == st.sw_tid[i] && pv.g_tid[i] == st.g_tid[i];
        if (one_bit) {
          o << "};\n      stage_swap_remap<T>(tid, " << lb << ", "
            << (pv.sw_tid[lb] ^ st.sw_tid[lb]) << "u, 0x" << std::hex
            << (pv.g_tid[lb] ^ st.g_tid[lb]) << std::dec << "ull, sr, sf);\n    }\n";
        } else {
          // the swizzled lane part of the new mapping: the XOR of its
          // thread bits' images
          o << "};\n      stage_remap<T>(";
          o << "0u";
          for (int i = 0; i < st.ntb; ++i)
            o << " ^ ((tid >> " << i << ") & 1u ? " << st.sw_tid[i] << "u : 0u)";
          o << ", f.gbase | ";
          bit_map_emit(o, gpos, true);
          o << ", sr, sf);\n    }\n";
        }
      }
      for (int g = st.g0; g < st.g1; ++g) {
        const GateDev& G = prog.gates[g];
        const GateRef gd{g};
        switch (G.kind) {
          case GK_U1_REAL:
          case GK_U1_RXT:
          case GK_U1_GEN:
            if ((G.creg | G.cglob) == 0)
              o << "    reg_u1<double, " << u1_kind_name(G.kind) << ", " << G.a << ", false>(x, "
                << gd << ", true);\n";
            else
              o << "    reg_u1<double, " << u1_kind_name(G.kind) << ", " << G.a << ", true>(x, "
                << gd << ", (sf.gl & " << gd << ".cglob) == " << gd << ".cglob);\n";
            break;
          case GK_U2: {
            // uncontrolled gates (the usual case) get no per-lane selects
            const bool ctrl = (G.creg | G.cglob) != 0;
            if (G.zero)
              o << "    reg_u2z<double, " << G.a << ", " << G.b << ", " << G.zero << "u, "
                << (ctrl ? "true" : "false") << ">(x, " << gd;
            else
              o << "    reg_u2<double, " << G.a << ", " << G.b << ", "
                << (ctrl ? "true" : "false") << ">(x, " << gd;
            if (ctrl)
              o << ", (sf.gl & " << gd << ".cglob) == " << gd << ".cglob);\n";
            else
              o << ", true);\n";
            break;
          }
          case GK_DIAGTAB:
            o << "    reg_diagtab<double, " << G.nu << ">(x, " << gd << ", sf.gl, tabs);\n";
            break;
          case GK_CZSET: {
            // The CZ pattern is structure.  Register pattern j is negated by
            // the parity of the lane-lane pairs (sign word sp), by each slot
            // t set in j whose lane partners have odd parity (sc<t>) and by
            // the register-register pairs (a constant): only the patterns a
            // per-lane term reaches get an XOR of the sign bits, constant
            // negations are left to the compiler.
            o << "    {\n";
            if (G.ncz > 0) {
              o << "      int par = 0;\n";
              for (int i = 0; i < G.ncz; ++i)
                o << "      par ^= __popcll(sf.gl & (sf.gl >> " << G.czs[i] << ") & 0x"
                  << std::hex << G.czm[i] << std::dec << "ull);\n";
              o << "      const uint32_t sp = (uint32_t)(par & 1) << 31;\n";
            }
            for (int t = 0; t < kRegQubits; ++t)
              if (G.czcm[t])
                o << "      const uint32_t sc" << t << " = (uint32_t)(__popcll(sf.gl & 0x"
                  << std::hex << G.czcm[t] << std::dec << "ull) & 1) << 31;\n";
            for (int j = 0; j < (1 << kRegQubits); ++j) {
              std::string terms;
              if (G.ncz > 0) terms = "sp";
              for (int t = 0; t < kRegQubits; ++t)
                if (G.czcm[t] && (j >> t & 1)) terms += (terms.empty() ? "sc" : " ^ sc") + std::to_string(t);
              const bool cneg = (G.czreg >> j) & 1;
              if (terms.empty()) {
                if (cneg) o << "      x[" << j << "].x = -x[" << j << "].x;\n      x[" << j
                            << "].y = -x[" << j << "].y;\n";
              } else {
                o << "      czsign(x[" << j << "], " << terms << (cneg ? " ^ 0x80000000u" : "")
                  << ");\n";
              }
            }
            o << "    }\n";
            break;
          }
          case GK_DIAGREG: o << "    reg_diagreg<double>(x, " << gd << ");\n"; break;
          case GK_DIAGREG1:
            if (G.b) o << "    reg_diagreg_m<double, " << G.b << "u>(x, " << gd << ");\n";
            break;
          case GK_OVL:
            o << "    reg_ovl<" << G.a << ", " << G.b << ", " << G.osel << ", " << G.ocnt << ", "
              << ((G.creg | G.cglob) ? "true" : "false") << ", " << ntp << ", "
              << (ovl_rows ? "true" : "false") << ">(x, " << gd
              << ", sf.gl, ovl_acc + " << ovl_i * nwaves << ");\n";
            ++ovl_i;
            break;
          case GK_U1_ROT:
          case GK_U1_RXR:
            o << "    reg_u1s<double, " << (G.kind == GK_U1_ROT ? "GK_U1_ROT" : "GK_U1_RXR") << ", "
              << G.a << ", " << G.b << ">(x, " << gd << ");\n";
            break;
          default: o << "    reg_diaglane<double, " << G.nu << ">(x, " << gd << ", sf.gl);\n"; break;
        }
      }
      if (seg_end) o << "    stage_store<T>(lds, sf, x);\n  }\n";
    }
    bool scaled = false, phased = false;
    for (int s = pd.s0; s < pd.s1; ++s)
      for (int g = prog.stages[s].g0; g < prog.stages[s].g1; ++g) {
        scaled |= prog.gates[g].kind == GK_U1_ROT || prog.gates[g].kind == GK_U1_RXR;
        phased |= prog.gates[g].kind == GK_DIAGREG1;
      }
    if (phased)
      o << "  pass_store_cscaled<T>(lds, f, pass->scale, pass->scale_im);\n";
    else if (scaled)
      o << "  pass_store_scaled<T>(lds, f, pass->scale);\n";
    else
      o << "  pass_store<T>(lds, f);\n";
    if (!oslots.empty()) {
      // the tile's overlaps: the waves' sums in order, one partial per
      // (overlap slot, tile of the launch)
      o << "  {\n    const int oslot[" << oslots.size() << "] = {";
      for (size_t i = 0; i < oslots.size(); ++i) o << (i ? ", " : "") << oslots[i];
      o << "};\n    const uint64_t ntiles = (uint64_t)gridDim.x * gridDim.y;\n"
        << "    for (int t = threadIdx.x; t < " << oslots.size() << "; t += blockDim.x) {\n"
        << "      double v = 0.0;\n"
        << "      for (int w = 0; w < " << nwaves << "; ++w) v += ovl_acc[t * " << nwaves
        << " + w];\n"
        << "      opart[(uint64_t)oslot[t] * ntiles + f.tile] = v;\n    }\n  }\n";
    }
    if (warm) o << "  asm volatile(\"\" :: \"v\"(warm));\n";
    o << "}\n";
  }
}

std::string jit_source(const Program& prog, int dtype) {
  std::ostringstream o;
  jit_emit(prog, dtype, o);
  return o.str();
}

// the key of prog into buf (whose capacity is reused across calls)
void jit_structure_key(const Program& prog, int dtype, std::string& buf);
static void jit_key(const Program& prog, int dtype, std::string& buf) {
  KeyOut o;
  o.k.swap(buf);
  o.k.clear();
  o.k.reserve(256 * prog.gates.size() + 4096);
  jit_emit(prog, dtype, o);
  o.k.swap(buf);
}

void jit_structure_key(const Program& prog, int dtype, std::string& buf) {
  jit_key(prog, dtype, buf);
}

JitModule::~JitModule() {
  if (mod) (void)hipModuleUnload(mod);
}

// hiprtc options: the flags of the precompiled kernels (no packed FP32, no
// sinking of the per-slot stores)
static std::vector<std::string> rtc_options() {
  return {"--gpu-architecture=gfx950",
          "-O3",
          "-std=c++17",
          "-DSMOQ_REG_QUBITS=" + std::to_string(kRegQubits),
          "-DSMOQ_ROWBITS_C128=" + std::to_string(CT<double>::kRowBits),
          "-DSMOQ_ROWBITS_C64=" + std::to_string(CT<float>::kRowBits),
          "-mllvm",
          "-simplifycfg-sink-common=false",
          "-Xclang",
          "-target-feature",
          "-Xclang",
          "-packed-fp32-ops"};
}

// One module's source to a code object.  Only hiprtc calls (no HIP runtime
// state), so several compile at once on host threads; err receives the
// compiler's log on failure.
static bool rtc_compile(const std::string& src, std::vector<char>& code, std::string& err) {
  const char* headers[3] = {kSrcDeviceCommon, kSrcFusedTypes, kSrcFusedOps};
  const char* names[3] = {"device_common.h", "fused_types.h", "fused_ops.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "smoq_passes.hip", 3, headers, names) !=
      HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const std::vector<std::string> opt = rtc_options();
  std::vector<const char*> opts;
  for (const std::string& o : opt) opts.push_back(o.c_str());
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      (void)hiprtcGetProgramLog(prog, &log[0]);
    }
    (void)hiprtcDestroyProgram(&prog);
    err = std::string(hiprtcGetErrorString(cr)) + ": " + log.substr(0, 400);
    return false;
  }
  size_t code_n = 0;
  code.clear();
  if (hiprtcGetCodeSize(prog, &code_n) == HIPRTC_SUCCESS) {
    code.resize(code_n);
    (void)hiprtcGetCode(prog, code.data());
  }
  (void)hiprtcDestroyProgram(&prog);
  if (code.empty()) {
    err = "compilation produced no code";
    return false;
  }
  return true;
}

static int load_module(const std::vector<char>& code, size_t n_passes,
                       std::shared_ptr<JitModule>& out, const char* prefix = "smoq_pass_") {
  auto m = std::make_shared<JitModule>();
  SMOQ_HIP(hipModuleLoadData(&m->mod, code.data()));
  m->fn.resize(n_passes);
  for (size_t p = 0; p < n_passes; ++p) {
    const std::string name = prefix + std::to_string(p);
    SMOQ_HIP(hipModuleGetFunction(&m->fn[p], m->mod, name.c_str()));
  }
  out = m;
  return SMOQ_OK;
}

// ------------------------------------------------------- on-disk cache --
// With SMOQ_JIT_CACHE_DIR set, code objects are also kept in that
// directory, named by a 128-bit hash of everything that determines them
// (the generated source, the embedded device headers, the compiler options
// and hiprtc's version), so a new process skips the compilation of a
// circuit structure seen before.
static std::string cache_dir() {
  const char* d = std::getenv("SMOQ_JIT_CACHE_DIR");
  return (d && *d) ? std::string(d) : std::string();
}

static void fnv_mix(uint64_t& h, const void* p, size_t n, uint64_t prime) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= prime;
  }
}

static std::string cache_name(const std::string& src) {
  uint64_t h1 = 0xcbf29ce484222325ull, h2 = 0x84222325cbf29ce4ull;
  int maj = 0, mnr = 0;
  (void)hiprtcVersion(&maj, &mnr);
  auto both = [&](const void* p, size_t n) {
    fnv_mix(h1, p, n, 0x100000001b3ull);
    fnv_mix(h2, p, n, 0x00000100000001b3ull ^ 0x9e3779b97f4a7c15ull);
  };
  both(&maj, sizeof maj);
  both(&mnr, sizeof mnr);
  for (const std::string& o : rtc_options()) both(o.data(), o.size() + 1);
  both(kSrcDeviceCommon, sizeof kSrcDeviceCommon);
  both(kSrcFusedTypes, sizeof kSrcFusedTypes);
  both(kSrcFusedOps, sizeof kSrcFusedOps);
  both(src.data(), src.size());
  char name[64];
  std::snprintf(name, sizeof name, "smoq_pass_%016llx%016llx.co", (unsigned long long)h1,
                (unsigned long long)h2);
  return name;
}

static bool disk_load(const std::string& dir, const std::string& name, std::vector<char>& code) {
  std::ifstream f(dir + "/" + name, std::ios::binary);
  if (!f) return false;
  code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !code.empty();
}

static void disk_store(const std::string& dir, const std::string& name,
                       const std::vector<char>& code) {
  // write then rename, so a reader never sees a partial file
  const std::string tmp = dir + "/." + name + "." + std::to_string((long long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(code.data(), (std::streamsize)code.size());
    if (!f) return;
  }
  (void)std::rename(tmp.c_str(), (dir + "/" + name).c_str());
}

int pass_tile(const smoq_ctx* ctx, int n_qubits, int batch, size_t n_gates) {
  if (ctx->tile_qubits > 0) return ctx->tile_qubits;
  const bool jit = jit_planned(ctx, n_qubits, n_gates);
  // the precompiled kernel is fastest on 2^10 tiles; compiled passes go to
  // the largest tile that still leaves >= 2048 workgroups (fewer passes)
  if (!jit) return 10;
  for (int k = kMaxTileQubits; k > 10; --k)
    if (((uint64_t)(batch > 0 ? batch : 1) << (n_qubits > k ? n_qubits - k : 0)) >= 2048) return k;
  return 10;
}

bool jit_planned(const smoq_ctx* ctx, int n_qubits, size_t n_gates) {
  return ctx->jit_mode >= 2 || (ctx->jit_mode == 1 && n_qubits >= kJitMinQubits &&
                                n_gates >= (size_t)kJitMinGates);
}

bool jit_wanted(const smoq_ctx* ctx, const Program& prog, int n_qubits) {
  if (program_has_swaps(prog)) return true;  // only compiled passes exchange lanes
  if (ctx->jit_mode == 0 || prog.passes.empty()) return false;
  if (ctx->jit_mode >= 2) return true;
  // automatic: large states with enough gates to amortise ~0.1 s of
  // compilation per pass on first use
  return n_qubits >= kJitMinQubits && prog.gates.size() >= (size_t)kJitMinGates;
}

int jit_get_many(smoq_ctx* ctx, const std::vector<const Program*>& progs, int dtype,
                 std::vector<std::shared_ptr<JitModule>>& out) {
  const size_t np = progs.size();
  out.assign(np, nullptr);
  std::vector<std::string> keys(np), srcs(np), names(np);
  std::vector<std::vector<char>> codes(np);
  std::vector<size_t> todo;
  const std::string dir = cache_dir();
  for (size_t i = 0; i < np; ++i) {
    jit_key(*progs[i], dtype, keys[i]);
    auto it = ctx->jit_cache.find(keys[i]);
    if (it != ctx->jit_cache.end()) {
      out[i] = it->second;
      continue;
    }
    // the same structure twice in one request: compile it once
    bool dup = false;
    for (size_t j : todo) dup |= keys[j] == keys[i];
    if (dup) continue;
    srcs[i] = jit_source(*progs[i], dtype);
    if (!dir.empty()) {
      names[i] = cache_name(srcs[i]);
      if (disk_load(dir, names[i], codes[i])) {
        std::shared_ptr<JitModule> m;
        if (load_module(codes[i], progs[i]->passes.size(), m) == SMOQ_OK) {
          ctx->jit_cache.emplace(keys[i], m);
          out[i] = m;
          continue;
        }
        (void)hipGetLastError();  // a stale or damaged file: compile again
      }
    }
    todo.push_back(i);
  }
  // compile what is missing, several modules at once
  std::vector<std::string> errs(np);
  std::vector<char> ok(np, 0);
  if (!todo.empty()) {
    const size_t hw = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t t; (t = next.fetch_add(1)) < todo.size();) {
        const size_t i = todo[t];
        ok[i] = rtc_compile(srcs[i], codes[i], errs[i]) ? 1 : 0;
      }
    };
    std::vector<std::thread> th;
    for (size_t w = 1; w < std::min(hw, todo.size()); ++w) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
  }
  for (size_t i : todo) {
    if (!ok[i])
      return set_error(SMOQ_EDEVICE, "pass kernel compilation failed (%s)", errs[i].c_str());
    std::shared_ptr<JitModule> m;
    int rc = load_module(codes[i], progs[i]->passes.size(), m);
    if (rc) return rc;
    if (!dir.empty()) disk_store(dir, names[i], codes[i]);
    ctx->jit_cache.emplace(keys[i], m);
  }
  for (size_t i = 0; i < np; ++i)
    if (!out[i]) out[i] = ctx->jit_cache.at(keys[i]);
  return SMOQ_OK;
}

int jit_get_source(smoq_ctx* ctx, const std::string& src, size_t n_fn, const char* prefix,
                   std::shared_ptr<JitModule>& out) {
  out.reset();
  // the caller keeps the module (not the context's pass-module table)
  const std::string dir = cache_dir();
  const std::string name = dir.empty() ? std::string() : cache_name(src);
  std::vector<char> code;
  if (!dir.empty() && disk_load(dir, name, code)) {
    std::shared_ptr<JitModule> m;
    if (load_module(code, n_fn, m, prefix) == SMOQ_OK) {
      out = m;
      return SMOQ_OK;
    }
    (void)hipGetLastError();  // a stale or damaged file: compile again
  }
  std::string err;
  if (!rtc_compile(src, code, err))
    return set_error(SMOQ_EDEVICE, "kernel compilation failed (%s)", err.c_str());
  std::shared_ptr<JitModule> m;
  const int rc = load_module(code, n_fn, m, prefix);
  if (rc) return rc;
  if (!dir.empty()) disk_store(dir, name, code);
  out = m;
  return SMOQ_OK;
}

int jit_get(smoq_ctx* ctx, const Program& prog, int dtype, std::shared_ptr<JitModule>& out) {
  out.reset();
  std::string& key = ctx->jit_key_buf;
  jit_key(prog, dtype, key);
  auto it = ctx->jit_cache.find(key);
  if (it != ctx->jit_cache.end()) {
    out = it->second;
    return SMOQ_OK;
  }
  std::vector<std::shared_ptr<JitModule>> mods;
  const int rc = jit_get_many(ctx, {&prog}, dtype, mods);
  if (rc) return rc;
  out = mods[0];
  return SMOQ_OK;
}

}  // namespace smoq

extern "C" int smoq_set_jit(smoq_ctx* ctx, int mode) {
  if (!ctx) return smoq::set_error(SMOQ_EINVAL, "smoq_set_jit: null context");
  if (mode < 0 || mode > 2) return smoq::set_error(SMOQ_EINVAL, "smoq_set_jit: mode %d", mode);
  ctx->jit_mode = mode;
  return SMOQ_OK;
}

extern "C" int smoq_set_obs_jit(smoq_ctx* ctx, int mode) {
  if (!ctx) return smoq::set_error(SMOQ_EINVAL, "smoq_set_obs_jit: null context");
  if (mode < 0 || mode > 2)
    return smoq::set_error(SMOQ_EINVAL, "smoq_set_obs_jit: mode %d", mode);
  ctx->obs_jit = mode;
  return SMOQ_OK;
}

extern "C" int smoq_set_lane_swaps(smoq_ctx* ctx, int on) {
  if (!ctx) return smoq::set_error(SMOQ_EINVAL, "smoq_set_lane_swaps: null context");
  ctx->lane_swaps = on != 0;
  return SMOQ_OK;
}
