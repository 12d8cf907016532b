// Autotuning of the fused-pass geometry.
//
// A circuit structure's fused passes depend on four choices: the tile size
// (2^k amplitudes per LDS tile), the contiguous low qubits every tile keeps
// (its HBM runs are 2^m amplitudes long), an optional cap on the
// non-diagonal gates one pass takes (deep passes are bound by their VALU
// work and LDS round trips, so spreading them can pay for an extra pass)
// and whether stages trade register and lane qubits in registers (fewer LDS
// round trips for more VALU work).  The reference
// fuses by a fixed greedy rule (fuse_gates,
// SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/run_benchmark_smoq_Float64.jl:167-200);
// here smoq_autotune plans the circuit under a set of candidate settings,
// compiles their pass modules on parallel host threads, times every
// candidate on the device and keeps the fastest for the structure.  Later
// applications of a gate list with the same structure (gate kinds and
// qubits; angles may differ), state size, dtype and batch use it.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

// exchanges tried per pass when searching its local qubit set
constexpr int kSearchIters = 300;
// searches (seeds) per candidate setting
constexpr uint64_t kSearchSeeds = 3;

uint64_t gate_structure_hash(const std::vector<IGate>& gs) {
  uint64_t h = 0xcbf29ce484222325ull;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xff;
      h *= 0x100000001b3ull;
    }
  };
  mix(gs.size());
  for (const IGate& g : gs) {
    mix((uint64_t)(uint32_t)g.kind | ((uint64_t)(uint32_t)g.nd << 32));
    mix((uint64_t)(uint32_t)g.t0 | ((uint64_t)(uint32_t)g.t1 << 32));
    mix((uint64_t)(uint32_t)g.dq[0] | ((uint64_t)(uint32_t)g.dq[1] << 32));
    mix((uint64_t)(uint32_t)g.dq[2] | ((uint64_t)(uint32_t)g.sel << 32));
    mix(g.ctrl);
    mix(g.oz);
  }
  return h;
}

static std::string tune_key(int n, int dtype, int batch, uint64_t hash) {
  char k[96];
  std::snprintf(k, sizeof k, "gfx950 %d %d %d %016llx", n, dtype, batch,
                (unsigned long long)hash);
  return k;
}

Fusion fusion_for(const smoq_ctx* ctx, int n, int dtype, int batch, const std::vector<IGate>& gs) {
  Fusion f;
  f.tile = pass_tile(ctx, n, batch, gs.size());
  f.low = ctx->low_qubits;
  f.cap = ctx->pass_max_gates;
  f.swaps = ctx->lane_swaps && jit_planned(ctx, n, gs.size());
  // explicit settings of the context win over tuned ones
  if (!ctx->tune.empty() && ctx->tile_qubits == 0) {
    auto it = ctx->tune.find(tune_key(n, dtype, batch, gate_structure_hash(gs)));
    if (it != ctx->tune.end()) f = it->second;
  }
  return f;
}

}  // namespace smoq

using namespace smoq;

extern "C" int smoq_autotune(smoq_state* s, const smoq_op* ops, int n_ops, const double* theta,
                             int n_theta, smoq_tune_result* out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!out) return set_error(SMOQ_EINVAL, "smoq_autotune: null result");
  std::memset(out, 0, sizeof(*out));
  smoq_ctx* ctx = s->ctx;
  std::vector<IGate> gs;
  if ((rc = lower_tape(ops, n_ops, theta, n_theta, s->n, gs))) return rc;
  bool fixed = true;
  for (int i = 0; i < n_ops; ++i) fixed &= ops[i].param < 0;
  const bool compiled = jit_planned(ctx, s->n, gs.size());
  const Fusion dflt{pass_tile(ctx, s->n, s->batch, gs.size()), ctx->low_qubits,
                    ctx->pass_max_gates, ctx->lane_swaps && compiled};
  out->tile_qubits = dflt.tile;
  out->low_qubits = dflt.low;
  out->pass_max_gates = dflt.cap;
  out->lane_swaps = dflt.swaps;
  if (gs.empty() || s->n < kRegQubits + 2) return SMOQ_OK;  // nothing to choose
  // candidates: the context's setting first, then tiles of 2^12 and 2^11
  // (those leaving >= 256 workgroups per pass) with 4 or 5 low qubits and a
  // few gate caps
  std::vector<Fusion> cand{dflt};
  auto add = [&](int k, int m, int c, int w) {
    if (k > s->n || m > k - 2) return;
    w = w && compiled;  // lane exchanges exist in compiled passes only
    for (const Fusion& f : cand)
      if (f.tile == k && f.low == m && f.cap == c && f.swaps == w && f.locals.empty()) return;
    cand.push_back({k, m, c, w, {}, {}});
  };
  // the same settings with searched local qubit sets (fewer passes where
  // the greedy sets cost some)
  auto add_searched = [&](int k, int m, int c, int w) {
    if (k >= s->n || m > k - 2) return;
    Fusion f{k, m, c, w && compiled, {}, {}};
    // a few seeds of the search; the fewest passes wins
    std::vector<uint64_t> loc;
    for (uint64_t seed = 1; seed <= kSearchSeeds; ++seed) {
      search_local_sets(gs, s->n, k, m, c, kSearchIters, seed, loc);
      if (f.locals.empty() || loc.size() < f.locals.size()) f.locals = loc;
    }
    cand.push_back(f);
  };
  for (int k : {12, 11}) {
    if (k < s->n && ((uint64_t)s->batch << (s->n - k)) < 256) continue;
    if (k == 12) {
      for (int c : {0, 28, 36, 44}) add(12, 4, c, 1);
      for (int c : {0, 36}) add(12, 4, c, 0);
      for (int c : {0, 36}) add(12, 5, c, 1);
      for (int c : {0, 30, 36, 44}) add_searched(12, 4, c, 1);
    } else {
      for (int c : {0, 36}) add(11, 4, c, 1);
      add(11, 5, 0, 1);
    }
  }
  const int rowbits = s->dtype == SMOQ_C128 ? CT<double>::kRowBits : CT<float>::kRowBits;
  std::vector<Program> progs(cand.size());
  std::vector<char> use_jit(cand.size(), 0);
  std::vector<const Program*> to_compile;
  std::vector<size_t> compile_idx;
  for (size_t i = 0; i < cand.size(); ++i) {
    plan_program(gs, s->n, cand[i].tile, cand[i].low, rowbits, progs[i], cand[i].cap,
                 cand[i].swaps != 0, cand[i].locals.empty() ? nullptr : &cand[i].locals);
    if (jit_wanted(ctx, progs[i], s->n)) {
      if (fixed) factor_rotations(progs[i], s->dtype);
      use_jit[i] = 1;
      to_compile.push_back(&progs[i]);
      compile_idx.push_back(i);
    }
  }
  std::vector<std::shared_ptr<JitModule>> mods;
  if (!to_compile.empty() && (rc = jit_get_many(ctx, to_compile, s->dtype, mods))) return rc;
  std::vector<std::shared_ptr<JitModule>> jit(cand.size());
  for (size_t t = 0; t < compile_idx.size(); ++t) jit[compile_idx[t]] = mods[t];
  // time every candidate: one warm application, then the mean of two
  size_t max_total = 0;
  for (const Program& p : progs) max_total = std::max(max_total, program_layout(p, s->dtype).total);
  if ((rc = ensure_prog(ctx, max_total))) return rc;
  const bool timing = ctx->timing;
  ctx->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&](int code) {
    ctx->timing = timing;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return code;
  };
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return cleanup(set_error(SMOQ_EDEVICE, "smoq_autotune: hipEventCreate failed"));
  if ((rc = smoq_state_set_basis(s, 0))) return cleanup(rc);
  std::vector<double> ms(cand.size(), 0.0);
  constexpr int kReps = 2;
  for (size_t i = 0; i < cand.size(); ++i) {
    const ProgLayout L = program_layout(progs[i], s->dtype);
    if ((rc = upload_program(ctx, progs[i], s->dtype, L, ctx->prog))) return cleanup(rc);
    const JitModule* jm = use_jit[i] ? jit[i].get() : nullptr;
    if ((rc = launch_program(s, progs[i], L, ctx->prog, jm))) return cleanup(rc);
    if (hipEventRecord(e0, ctx->stream) != hipSuccess)
      return cleanup(set_error(SMOQ_EDEVICE, "smoq_autotune: event record failed"));
    for (int r = 0; r < kReps; ++r)
      if ((rc = launch_program(s, progs[i], L, ctx->prog, jm))) return cleanup(rc);
    float el = 0.f;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&el, e0, e1) != hipSuccess)
      return cleanup(set_error(SMOQ_EDEVICE, "smoq_autotune: timing failed"));
    ms[i] = el / kReps;
  }
  size_t best = 0;
  for (size_t i = 1; i < cand.size(); ++i)
    if (ms[i] < ms[best]) best = i;
  // SMOQ_TUNE_LOG=1: every candidate's settings and time on stderr
  const char* log = std::getenv("SMOQ_TUNE_LOG");
  if (log && std::atoi(log) != 0)
    for (size_t i = 0; i < cand.size(); ++i)
      std::fprintf(stderr,
                   "smoq_autotune: tile 2^%d low %d cap %d swaps %d %s: %zu passes %.3f ms%s\n",
                   cand[i].tile, cand[i].low, cand[i].cap, cand[i].swaps,
                   cand[i].locals.empty() ? "greedy" : "searched", progs[i].passes.size(), ms[i],
                   i == best ? " (best)" : "");
  // Per pass, the HBM order of its tiles and of the registers that move
  // them (set_pass_io): they change no result, but the time of a pass
  // depends on which address bits its concurrently running tiles span, by
  // up to a quarter on some qubit maps.  Every pass is timed alone under
  // each order and keeps the fastest (the default unless 1% faster).
  Program& bp = progs[best];
  const int nn = bp.passes.empty() ? 0 : bp.passes[0].n_nonlocal;
  if (nn >= 12 && s->batch == 1) {
    const ProgLayout L = program_layout(bp, s->dtype);
    const JitModule* jm = use_jit[best] ? jit[best].get() : nullptr;
    std::vector<int> codes;
    for (int rev : {0, 32})
      for (int r : {0, 3, 6, 9, 12, 15}) codes.push_back(rev | r);
    for (int rev : {0, 32})
      for (int r : {0, 6, 12}) codes.push_back(64 | rev | r);
    std::vector<uint8_t> io(bp.passes.size(), 0);
    const int low = cand[best].low;
    auto time_pass = [&](size_t p, float* out_ms) -> int {
      int r2;
      if ((r2 = upload_program(ctx, bp, s->dtype, L, ctx->prog))) return r2;
      if (hipEventRecord(e0, ctx->stream) != hipSuccess)
        return set_error(SMOQ_EDEVICE, "smoq_autotune: event record failed");
      for (int r = 0; r < kReps; ++r)
        if ((r2 = launch_program(s, bp, L, ctx->prog, jm, nullptr, (int)p))) return r2;
      if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(out_ms, e0, e1) != hipSuccess)
        return set_error(SMOQ_EDEVICE, "smoq_autotune: timing failed");
      return SMOQ_OK;
    };
    for (size_t p = 0; p < bp.passes.size(); ++p) {
      // one untimed launch of the pass first
      if ((rc = upload_program(ctx, bp, s->dtype, L, ctx->prog))) return cleanup(rc);
      if ((rc = launch_program(s, bp, L, ctx->prog, jm, nullptr, (int)p))) return cleanup(rc);
      float base = 0.f;
      if ((rc = time_pass(p, &base))) return cleanup(rc);
      float bt = base;
      int bc = 0;
      for (int c : codes) {
        if (c == 0) continue;
        if ((c & 31) >= bp.passes[p].n_nonlocal) continue;
        set_pass_io(bp.passes[p], c, low);
        float t = 0.f;
        if ((rc = time_pass(p, &t))) return cleanup(rc);
        if (t < bt) {
          bt = t;
          bc = c;
        }
      }
      if (bt > 0.99f * base) bc = 0;
      io[p] = (uint8_t)bc;
      set_pass_io(bp.passes[p], bc, low);
    }
    cand[best].io = io;
    for (uint8_t c : io) out->io_passes += c != 0;
    float total = 0.f;
    for (size_t p = 0; p < bp.passes.size(); ++p) {
      float t = 0.f;
      if ((rc = time_pass(p, &t))) return cleanup(rc);
      total += t / kReps;
    }
    out->io_ms = total;
  }
  ctx->tune[tune_key(s->n, s->dtype, s->batch, gate_structure_hash(gs))] = cand[best];
  out->tile_qubits = cand[best].tile;
  out->low_qubits = cand[best].low;
  out->pass_max_gates = cand[best].cap;
  out->lane_swaps = cand[best].swaps;
  out->searched = !cand[best].locals.empty();
  out->n_passes = (int)progs[best].passes.size();
  out->ms = ms[best];
  out->default_ms = ms[0];
  out->candidates = (int)cand.size();
  return cleanup(SMOQ_OK);
}

extern "C" int smoq_tune_save(smoq_ctx* ctx, const char* path) {
  if (!ctx || !path) return set_error(SMOQ_EINVAL, "smoq_tune_save: null argument");
  std::ofstream f(path);
  if (!f) return set_error(SMOQ_EINVAL, "smoq_tune_save: cannot write %s", path);
  for (const auto& kv : ctx->tune)
  {
    f << kv.first << ' ' << kv.second.tile << ' ' << kv.second.low << ' ' << kv.second.cap
      << ' ' << kv.second.swaps;
    // per-pass HBM orders and searched local sets, when there are any
    if (!kv.second.io.empty()) {
      f << " io " << kv.second.io.size();
      for (uint8_t c : kv.second.io) f << ' ' << (int)c;
    }
    if (!kv.second.locals.empty()) {
      f << " loc " << kv.second.locals.size() << std::hex;
      for (uint64_t l : kv.second.locals) f << ' ' << l;
      f << std::dec;
    }
    f << '\n';
  }
  return f ? SMOQ_OK : set_error(SMOQ_EINVAL, "smoq_tune_save: write to %s failed", path);
}

extern "C" int smoq_tune_load(smoq_ctx* ctx, const char* path, int* n_entries) {
  if (!ctx || !path) return set_error(SMOQ_EINVAL, "smoq_tune_load: null argument");
  std::ifstream f(path);
  if (!f) return set_error(SMOQ_EINVAL, "smoq_tune_load: cannot read %s", path);
  int count = 0;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string arch, hash;
    int n, dtype, batch;
    Fusion fu;
    if (!(is >> arch >> n >> dtype >> batch >> hash >> fu.tile >> fu.low >> fu.cap >> fu.swaps))
      continue;
    if (arch != "gfx950" || fu.tile < 6 || fu.tile > kMaxTileQubits || fu.low < 3 || fu.low > 6 ||
        fu.cap < 0)
      continue;
    bool ok = true;
    std::string tag;
    while (ok && (is >> tag)) {
      size_t cnt = 0;
      ok = (bool)(is >> cnt) && cnt <= 4096 && (tag == "io" || tag == "loc");
      for (size_t p = 0; p < cnt && ok; ++p) {
        if (tag == "io") {
          int c = -1;
          ok = (bool)(is >> c) && c >= 0 && c < 128;
          fu.io.push_back((uint8_t)c);
        } else {
          uint64_t l = 0;
          ok = (bool)(is >> std::hex >> l >> std::dec) && l != 0 && (n >= 64 || (l >> n) == 0);
          fu.locals.push_back(l);
        }
      }
    }
    if (!ok) continue;
    ctx->tune[arch + " " + std::to_string(n) + " " + std::to_string(dtype) + " " +
              std::to_string(batch) + " " + hash] = fu;
    ++count;
  }
  if (n_entries) *n_entries = count;
  return SMOQ_OK;
}
