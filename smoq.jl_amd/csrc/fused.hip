// Fused gate passes: one HBM read and one HBM write of the state per pass,
// with every gate of the pass applied to an LDS-resident tile.
//
// A workgroup owns one tile: the 2^k amplitudes whose non-local bits equal
// the tile id.  The tile is read from HBM into LDS once (coalesced runs of
// 2^m consecutive amplitudes), every register stage then moves 2^R
// amplitudes per thread from LDS into registers, applies its gates there and
// writes them back, and the tile is finally written to HBM once.  LDS rows
// are XOR-swizzled so that the strided per-stage gathers stay free of bank
// conflicts.  Control and diagonal qubits never need to be local: their bits
// come from the thread's base index (global part) or the register slot.
#include "fused_ops.h"
#include "smoq_internal.h"

namespace smoq {

// T is the storage type of the state in HBM and LDS; in registers the
// amplitudes are complex128 whatever T is, so a complex64 state is rounded
// once per register stage instead of after every gate.  This kernel reads
// its stage and gate lists from the program at run time; jit.hip generates
// the same kernel with both lists unrolled for one circuit structure.
template <typename T>
__global__ void __launch_bounds__(kMaxPassThreads)
    k_fused_pass(VT<T>* __restrict__ psi, int n, const PassDev* __restrict__ pass,
                 const StageDev* __restrict__ stages, const GateDev* __restrict__ gates,
                 const double2* __restrict__ tabs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  VT<T>* lds = reinterpret_cast<VT<T>*>(smem_raw);
  // Passes of few tiles (small states) run one workgroup per CU, and each
  // stage would wait a full memory latency for its descriptor and gate
  // records on their first scalar loads: every thread first loads one dword
  // per 128 B line of them (at most four lines a thread), so that those
  // loads find them in L2; the values are consumed after the store.
  uint32_t warm = 0;
  if (pass->n_nonlocal <= kWarmMaxNonlocal) {
    const int g0 = stages[pass->s0].g0, g1 = stages[pass->s1 - 1].g1;
    const uint32_t ns = (uint32_t)(pass->s1 - pass->s0) * (uint32_t)(sizeof(StageDev) / 4);
    const uint32_t ng = (uint32_t)(g1 - g0) * (uint32_t)(sizeof(GateDev) / 4);
    const uint32_t cap = blockDim.x * 32u * kWarmLoads;
    const uint32_t* sw = reinterpret_cast<const uint32_t*>(stages + pass->s0);
    const uint32_t* gw = reinterpret_cast<const uint32_t*>(gates + g0);
    if (ns + ng <= cap) {
      for (uint32_t w = threadIdx.x * 32u; w < ns; w += blockDim.x * 32u) warm ^= sw[w];
      for (uint32_t w = threadIdx.x * 32u; w < ng; w += blockDim.x * 32u) warm ^= gw[w];
    }
  }
  PassFrame<T> f;
  pass_load<T>(lds, psi, n, pass, f);
  for (int s = pass->s0; s < pass->s1; ++s) {
    const StageDev& st = stages[s];
    VT<double> x[kNR];
    StageFrame sf;
    stage_load<T>(lds, st, f.gbase, x, sf);
    for (int g = st.g0; g < st.g1; ++g) apply_gate<double>(x, gates[g], sf.gl, tabs);
    stage_store<T>(lds, sf, x);
  }
  pass_store<T>(lds, f);
  asm volatile("" ::"v"(warm));
}

ProgLayout program_layout(const Program& prog, int dtype) {
  ProgLayout L;
  L.pb = prog.passes.size() * sizeof(PassDev);
  L.sb = prog.stages.size() * sizeof(StageDev);
  L.gb = prog.gates.size() * sizeof(GateDev);
  L.ntab = prog.tabs.size() / 2;
  (void)dtype;  // phase tables are complex128 for every state dtype
  L.off_t = (L.pb + L.sb + L.gb + 255) & ~(size_t)255;
  L.total = L.off_t + (L.ntab + 1) * sizeof(double2);
  return L;
}

// copies the program into device buffer d (L.total bytes) through the
// context's pinned staging buffer, ordered on the context stream
int upload_program(smoq_ctx* ctx, const Program& prog, int dtype, const ProgLayout& L, void* d) {
  int rc = ensure_pinned(ctx, L.total);
  if (rc) return rc;
  // the pinned staging buffer may still feed an earlier copy
  SMOQ_HIP(hipStreamSynchronize(ctx->stream));
  unsigned char* h = (unsigned char*)ctx->host_pinned;
  std::memcpy(h, prog.passes.data(), L.pb);
  std::memcpy(h + L.pb, prog.stages.data(), L.sb);
  std::memcpy(h + L.pb + L.sb, prog.gates.data(), L.gb);
  (void)dtype;
  if (L.ntab) std::memcpy(h + L.off_t, prog.tabs.data(), L.ntab * sizeof(double2));
  SMOQ_HIP(hipMemcpyAsync(d, h, L.total, hipMemcpyHostToDevice, ctx->stream));
  return SMOQ_OK;
}

template <typename T>
static int launch_program_t(smoq_state* s, const Program& prog, const ProgLayout& L,
                            const void* dbuf, const JitModule* jit, double* opart,
                            int only_pass, uint64_t item_bytes) {
  smoq_ctx* ctx = s->ctx;
  int rc = SMOQ_OK;
  const unsigned char* d = (const unsigned char*)dbuf;
  const PassDev* dp = (const PassDev*)d;
  const StageDev* ds = (const StageDev*)(d + L.pb);
  const GateDev* dg = (const GateDev*)(d + L.pb + L.sb);
  const double2* dt = (const double2*)(d + L.off_t);
  if (jit && jit->fn.size() != prog.passes.size())
    return set_error(SMOQ_EINVAL, "fused pass: compiled module has %zu passes, program %zu",
                     jit->fn.size(), prog.passes.size());
  if (!jit)
    for (const GateDev& g : prog.gates)
      if (g.kind == GK_OVL)
        return set_error(SMOQ_EINVAL, "fused pass: overlap gates need compiled passes");
  if (item_bytes && (!jit || !prog.per_item))
    return set_error(SMOQ_EINVAL, "fused pass: per-item programs need passes compiled for them");
  unsigned long long ib = item_bytes;
  VT<T>* psi = (VT<T>*)s->d;
  int n = s->n;
  for (size_t p = 0; p < prog.passes.size(); ++p) {
    if (only_pass >= 0 && (int)p != only_pass) continue;
    const PassDev& pd = prog.passes[p];
    const uint64_t tiles = (uint64_t)s->batch << pd.n_nonlocal;
    const int nt = 1 << (pd.k - kRegQubits);
    if (nt > kMaxPassThreads || pd.k > kMaxTileQubits)
      return set_error(SMOQ_EINVAL, "fused pass: tile of 2^%d needs %d threads (max %d)", pd.k,
                       nt, kMaxPassThreads);
    const size_t shm = sizeof(VT<T>) << pd.k;
    // rows of 2^16 tiles in blockIdx.y once a pass has 16 or more non-tile
    // qubits (the dispatch order is the same as one row's), so that a pass
    // never needs 2^32 work-items in x; one row otherwise
    uint64_t gx = tiles, gy = 1;
    if (pd.n_nonlocal >= 16) {
      gx = 1ull << 16;
      gy = tiles >> 16;
    }
    if (gy > 65535) {
      gx = tiles;
      gy = 1;
    }
    if (gx * (uint64_t)nt >= (1ull << 32))
      return set_error(SMOQ_ENOTSUP, "fused pass: %llu tiles of %d threads exceed the grid",
                       (unsigned long long)tiles, nt);
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
    if (jit) {
      const PassDev* pp = dp + p;
      void* args[] = {&psi, &n, &pp, (void*)&ds, (void*)&dg, (void*)&dt, &opart, &ib};
      SMOQ_HIP(hipModuleLaunchKernel(jit->fn[p], (unsigned)gx, (unsigned)gy, 1, nt, 1, 1,
                                     0u, ctx->stream, args, nullptr));
    } else {
      hipLaunchKernelGGL(k_fused_pass<T>, dim3((unsigned)gx, (unsigned)gy), dim3(nt), shm,
                         ctx->stream, psi, n, dp + p, ds, dg, dt);
      SMOQ_CHECK_LAUNCH("k_fused_pass");
    }
    if (ctx->timing && (rc = time_mark(ctx))) return rc;
  }
  return SMOQ_OK;
}

int launch_program(smoq_state* s, const Program& prog, const ProgLayout& L, const void* dbuf,
                   const JitModule* jit, double* opart, int only_pass, uint64_t item_bytes) {
  return s->dtype == SMOQ_C128
             ? launch_program_t<double>(s, prog, L, dbuf, jit, opart, only_pass, item_bytes)
             : launch_program_t<float>(s, prog, L, dbuf, jit, opart, only_pass, item_bytes);
}

int run_program(smoq_state* s, Program& prog, bool fixed_angles) {
  const ProgLayout L = program_layout(prog, s->dtype);
  int rc = ensure_prog(s->ctx, L.total);
  if (rc) return rc;
  std::shared_ptr<JitModule> jit;
  if (jit_wanted(s->ctx, prog, s->n)) {
    std::vector<FactorUndo> undo;
    if (fixed_angles) factor_rotations(prog, s->dtype, &undo);
    rc = jit_get(s->ctx, prog, s->dtype, jit);
    // automatic mode keeps the precompiled kernel (unfactored gates) when
    // compilation fails, unless the program exchanges lanes
    if (rc && (s->ctx->jit_mode >= 2 || program_has_swaps(prog))) return rc;
    if (rc) unfactor_rotations(prog, undo);
  }
  if ((rc = upload_program(s->ctx, prog, s->dtype, L, s->ctx->prog))) return rc;
  return launch_program(s, prog, L, s->ctx->prog, jit.get());
}

}  // namespace smoq
