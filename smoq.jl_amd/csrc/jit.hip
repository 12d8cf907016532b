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
// control presence, stage and gate indices); passes, stages, gate matrices
// and phase tables are still read from the device program, so plans with
// new angles reuse the compiled module.
#include <hip/hiprtc.h>

#include <cstdlib>
#include <sstream>

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

std::string jit_source(const Program& prog, int dtype) {
  std::ostringstream o;
  const char* T = dtype == SMOQ_C128 ? "double" : "float";
  o << "#include \"fused_ops.h\"\nusing namespace smoq;\n";
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    const PassDev& pd = prog.passes[p];
    o << "extern \"C\" __global__ void __launch_bounds__(256) smoq_pass_" << p << "(VT<" << T
      << ">* __restrict__ psi, int n, const PassDev* __restrict__ pass,"
         " const StageDev* __restrict__ stages, const GateDev* __restrict__ gates,"
         " const double2* __restrict__ tabs) {\n"
      << "  typedef " << T << " T;\n"
      << "  extern __shared__ __align__(16) unsigned char smem_raw[];\n"
      << "  VT<T>* lds = reinterpret_cast<VT<T>*>(smem_raw);\n"
      << "  PassFrame<T> f;\n"
      << "  pass_load<T>(lds, psi, n, pass, f);\n";
    for (int s = pd.s0; s < pd.s1; ++s) {
      const StageDev& st = prog.stages[s];
      o << "  {\n    VT<double> x[kNR];\n    StageFrame sf;\n"
        << "    stage_load<T>(lds, stages[" << s << "], f.gbase, x, sf);\n";
      for (int g = st.g0; g < st.g1; ++g) {
        const GateDev& G = prog.gates[g];
        const std::string gd = "gates[" + std::to_string(g) + "]";
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
          case GK_U2:
            o << "    reg_u2<double, " << G.a << ", " << G.b << ">(x, " << gd << ", (sf.gl & " << gd
              << ".cglob) == " << gd << ".cglob);\n";
            break;
          case GK_DIAGTAB: o << "    reg_diagtab<double>(x, " << gd << ", sf.gl, tabs);\n"; break;
          case GK_CZSET: o << "    reg_czset<double>(x, " << gd << ", sf.gl);\n"; break;
          case GK_DIAGREG: o << "    reg_diagreg<double>(x, " << gd << ");\n"; break;
          default: o << "    reg_diaglane<double>(x, " << gd << ", sf.gl);\n"; break;
        }
      }
      o << "    stage_store<T>(lds, sf, x);\n  }\n";
    }
    o << "  pass_store<T>(lds, f);\n}\n";
  }
  return o.str();
}

JitModule::~JitModule() {
  if (mod) (void)hipModuleUnload(mod);
}

static int compile(const std::string& src, size_t n_passes, std::shared_ptr<JitModule>& out) {
  const char* headers[3] = {kSrcDeviceCommon, kSrcFusedTypes, kSrcFusedOps};
  const char* names[3] = {"device_common.h", "fused_types.h", "fused_ops.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "smoq_passes.hip", 3, headers, names) !=
      HIPRTC_SUCCESS)
    return set_error(SMOQ_EDEVICE, "hiprtcCreateProgram failed");
  const std::string reg = "-DSMOQ_REG_QUBITS=" + std::to_string(kRegQubits);
  const std::string rb128 = "-DSMOQ_ROWBITS_C128=" + std::to_string(CT<double>::kRowBits);
  const std::string rb64 = "-DSMOQ_ROWBITS_C64=" + std::to_string(CT<float>::kRowBits);
  // the flags of the precompiled kernels: no packed FP32, no sinking of the
  // per-slot stores
  const char* opts[] = {"--gpu-architecture=gfx950", "-O3", "-std=c++17", reg.c_str(),
                        rb128.c_str(), rb64.c_str(), "-mllvm", "-simplifycfg-sink-common=false",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      (void)hiprtcGetProgramLog(prog, &log[0]);
    }
    (void)hiprtcDestroyProgram(&prog);
    return set_error(SMOQ_EDEVICE, "pass kernel compilation failed (%s): %.400s",
                     hiprtcGetErrorString(cr), log.c_str());
  }
  size_t code_n = 0;
  std::vector<char> code;
  if (hiprtcGetCodeSize(prog, &code_n) == HIPRTC_SUCCESS) {
    code.resize(code_n);
    (void)hiprtcGetCode(prog, code.data());
  }
  (void)hiprtcDestroyProgram(&prog);
  if (code.empty()) return set_error(SMOQ_EDEVICE, "pass kernel compilation produced no code");
  auto m = std::make_shared<JitModule>();
  SMOQ_HIP(hipModuleLoadData(&m->mod, code.data()));
  m->fn.resize(n_passes);
  for (size_t p = 0; p < n_passes; ++p) {
    const std::string name = "smoq_pass_" + std::to_string(p);
    SMOQ_HIP(hipModuleGetFunction(&m->fn[p], m->mod, name.c_str()));
  }
  out = m;
  return SMOQ_OK;
}

bool jit_wanted(const smoq_ctx* ctx, const Program& prog, int n_qubits) {
  if (ctx->jit_mode == 0 || prog.passes.empty()) return false;
  if (ctx->jit_mode >= 2) return true;
  // automatic: large states with enough gates to amortise ~0.1 s of
  // compilation per pass on first use
  return n_qubits >= kJitMinQubits && prog.gates.size() >= (size_t)kJitMinGates;
}

int jit_get(smoq_ctx* ctx, const Program& prog, int dtype, std::shared_ptr<JitModule>& out) {
  out.reset();
  std::string src = jit_source(prog, dtype);
  auto it = ctx->jit_cache.find(src);
  if (it != ctx->jit_cache.end()) {
    out = it->second;
    return SMOQ_OK;
  }
  std::shared_ptr<JitModule> m;
  const int rc = compile(src, prog.passes.size(), m);
  if (rc) return rc;
  ctx->jit_cache.emplace(std::move(src), m);
  out = m;
  return SMOQ_OK;
}

}  // namespace smoq

extern "C" int smoq_set_jit(smoq_ctx* ctx, int mode) {
  if (!ctx) return smoq::set_error(SMOQ_EINVAL, "smoq_set_jit: null context");
  if (mode < 0 || mode > 2) return smoq::set_error(SMOQ_EINVAL, "smoq_set_jit: mode %d", mode);
  ctx->jit_mode = mode;
  return SMOQ_OK;
}
