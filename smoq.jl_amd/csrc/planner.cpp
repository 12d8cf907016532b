// Host-side planner: groups a gate list into HBM passes over LDS tiles and,
// inside each pass, into register stages.
//
// A pass owns a set L of k "local" qubits (always including the low qubits
// 0..m-1 so that every tile is made of contiguous runs of 2^m amplitudes).
// A gate joins the pass when its target qubits fit in L (or L can grow to
// hold them) and it touches no qubit of a gate already deferred to a later
// pass — so it commutes with everything it overtakes.  Diagonal gates and
// control qubits never need to be local: they only phase or select
// amplitudes by their global index.
//
// Inside a stage, runs of consecutive diagonal gates (which commute with one
// another) are merged: CZs into one parity-sign term per qubit distance, the
// others into one phase table indexed by the register slots and at most
// kMaxTabBits other qubits.  2x2 matrices are classified so that real (Ry,
// H, X) and Rx-type matrices use the cheaper arithmetic.
#include <algorithm>
#include <complex>
#include <cstdlib>
#include <iterator>
#include <map>

#include "smoq_internal.h"

namespace smoq {

using cd = std::complex<double>;

// An overlap reads its Pauli string's support; gates on other qubits commute
// with it (both halves change by the same unitary), and so do other
// overlaps, so the selector qubit is not part of what it touches.
static uint64_t ovl_x(const IGate& g) {
  return (g.t0 >= 0 ? 1ull << g.t0 : 0) | (g.t1 >= 0 ? 1ull << g.t1 : 0);
}

static uint64_t touch_mask(const IGate& g) {
  uint64_t m = g.ctrl;
  if (g.kind == IK_U1) m |= 1ull << g.t0;
  if (g.kind == IK_U2) m |= (1ull << g.t0) | (1ull << g.t1);
  if (g.kind == IK_DIAG)
    for (int i = 0; i < g.nd; ++i) m |= 1ull << g.dq[i];
  if (g.kind == IK_OVL) m |= ovl_x(g) | g.oz;
  return m;
}

// qubits that must be local to the pass (and register slots of the stage)
static uint64_t target_mask(const IGate& g) {
  if (g.kind == IK_U1) return 1ull << g.t0;
  if (g.kind == IK_U2) return (1ull << g.t0) | (1ull << g.t1);
  if (g.kind == IK_OVL) return ovl_x(g) | (1ull << g.sel);
  return 0;
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }

static bool is_cz(const IGate& g) {
  return g.kind == IK_DIAG && g.nd == 2 && g.m[0] == 1 && g.m[1] == 0 && g.m[2] == 1 &&
         g.m[3] == 0 && g.m[4] == 1 && g.m[5] == 0 && g.m[6] == -1 && g.m[7] == 0;
}

static int u1_kind(const double* m) {
  if (m[1] == 0 && m[3] == 0 && m[5] == 0 && m[7] == 0) return GK_U1_REAL;
  if (m[1] == 0 && m[7] == 0 && m[2] == 0 && m[4] == 0) return GK_U1_RXT;
  return GK_U1_GEN;
}

namespace {
struct StageCtx {
  const int* g2l;     // global bit -> local position or -1
  const int* l2slot;  // local position -> register slot or -1
  int n;
  int slot_of(int gq) const { return g2l[gq] >= 0 ? l2slot[g2l[gq]] : -1; }
};
}  // namespace

static GateDev blank_gate() {
  GateDev gd;
  std::memset(&gd, 0, sizeof(gd));
  gd.a = gd.b = -1;
  return gd;
}

static void emit_nondiag(const IGate& g, const StageCtx& sc, Program& prog) {
  GateDev gd = blank_gate();
  gd.kind = g.kind == IK_U2 ? GK_U2 : u1_kind(g.m);
  gd.a = sc.slot_of(g.t0);
  gd.b = g.kind == IK_U2 ? sc.slot_of(g.t1) : -1;
  for (int b = 0; b < sc.n; ++b) {
    if (!(g.ctrl >> b & 1)) continue;
    int sl = sc.slot_of(b);
    if (sl >= 0)
      gd.creg |= 1 << sl;
    else
      gd.cglob |= 1ull << b;
  }
  std::memcpy(gd.m, g.m, sizeof(gd.m));
  prog.gates.push_back(gd);
}

static void emit_ovl(const IGate& g, const StageCtx& sc, Program& prog) {
  GateDev gd = blank_gate();
  gd.kind = GK_OVL;
  gd.a = g.t0 >= 0 ? sc.slot_of(g.t0) : -1;
  gd.b = g.t1 >= 0 ? sc.slot_of(g.t1) : -1;
  gd.osel = sc.slot_of(g.sel);
  gd.oslot = g.oslot;
  gd.ocnt = g.ocnt;
  for (int b = 0; b < sc.n; ++b) {
    const int sl = sc.slot_of(b);
    if (g.ctrl >> b & 1) {
      if (sl >= 0)
        gd.creg |= 1 << sl;
      else
        gd.cglob |= 1ull << b;
    }
    if (g.oz >> b & 1) {
      if (sl >= 0)
        gd.ozreg |= 1 << sl;
      else
        gd.ozglob |= 1ull << b;
    }
  }
  prog.gates.push_back(gd);
}

// CZ gates of a diagonal run: lane-lane pairs become one popcount term per
// qubit distance, register-register pairs a fixed 16-bit pattern mask and
// register-lane pairs a per-slot mask of lane qubits.
static void emit_czs(const std::vector<const IGate*>& czs, const StageCtx& sc,
                     Program& prog) {
  std::map<int, uint64_t> by_shift;
  int czreg = 0;
  uint64_t czcm[kRegQubits] = {};
  for (const IGate* g : czs) {
    int a = g->dq[0], b = g->dq[1];
    int sa = sc.slot_of(a), sb = sc.slot_of(b);
    if (sa >= 0 && sb >= 0) {
      for (int j = 0; j < (1 << kRegQubits); ++j)
        if ((j >> sa & 1) && (j >> sb & 1)) czreg ^= 1 << j;
    } else if (sa >= 0) {
      czcm[sa] ^= 1ull << b;
    } else if (sb >= 0) {
      czcm[sb] ^= 1ull << a;
    } else {
      int lo = std::min(a, b), hi = std::max(a, b);
      by_shift[hi - lo] ^= 1ull << lo;  // CZ twice on a pair is the identity
    }
  }
  GateDev gd = blank_gate();
  gd.kind = GK_CZSET;
  gd.czreg = czreg;
  for (int t = 0; t < kRegQubits; ++t) gd.czcm[t] = czcm[t];
  for (auto& kv : by_shift) {
    if (kv.second == 0) continue;
    if (gd.ncz == kMaxCzShifts) {
      prog.gates.push_back(gd);
      gd = blank_gate();
      gd.kind = GK_CZSET;
    }
    gd.czs[gd.ncz] = kv.first;
    gd.czm[gd.ncz] = kv.second;
    gd.ncz++;
  }
  prog.gates.push_back(gd);
}

// diagonal gates on register qubits only: one uniform 16-entry phase vector
static void emit_diag_reg(const std::vector<const IGate*>& ds, const StageCtx& sc,
                          Program& prog) {
  GateDev gd = blank_gate();
  gd.kind = GK_DIAGREG;
  for (int j = 0; j < (1 << kRegQubits); ++j) {
    cd ph(1.0, 0.0);
    for (const IGate* g : ds) {
      int e = 0;
      for (int t = 0; t < g->nd; ++t) e |= ((j >> sc.slot_of(g->dq[t])) & 1) << t;
      ph *= cd(g->m[2 * e], g->m[2 * e + 1]);
    }
    gd.m[2 * j] = ph.real();
    gd.m[2 * j + 1] = ph.imag();
  }
  prog.gates.push_back(gd);
}

// diagonal gates on non-register qubits only: one phase per lane, chosen
// from 2^nu values by nu <= kMaxTabBits index bits
static void emit_diag_lane(const std::vector<const IGate*>& ds, Program& prog) {
  size_t i = 0;
  while (i < ds.size()) {
    std::vector<int> ub;
    size_t j = i;
    for (; j < ds.size(); ++j) {
      std::vector<int> add;
      for (int t = 0; t < ds[j]->nd; ++t) {
        int q = ds[j]->dq[t];
        if (std::find(ub.begin(), ub.end(), q) == ub.end() &&
            std::find(add.begin(), add.end(), q) == add.end())
          add.push_back(q);
      }
      if ((int)(ub.size() + add.size()) > kMaxTabBits) break;
      ub.insert(ub.end(), add.begin(), add.end());
    }
    GateDev gd = blank_gate();
    gd.kind = GK_DIAGLANE;
    gd.nu = (int)ub.size();
    for (int t = 0; t < gd.nu; ++t) gd.ubits[t] = ub[t];
    for (int row = 0; row < (1 << gd.nu); ++row) {
      cd ph(1.0, 0.0);
      for (size_t g = i; g < j; ++g) {
        int e = 0;
        for (int t = 0; t < ds[g]->nd; ++t) {
          int pos = (int)(std::find(ub.begin(), ub.end(), ds[g]->dq[t]) - ub.begin());
          e |= ((row >> pos) & 1) << t;
        }
        ph *= cd(ds[g]->m[2 * e], ds[g]->m[2 * e + 1]);
      }
      gd.m[2 * row] = ph.real();
      gd.m[2 * row + 1] = ph.imag();
    }
    prog.gates.push_back(gd);
    i = j;
  }
}

// other diagonal gates: phase tables over (non-register qubits, slot pattern)
static void emit_tables(const std::vector<const IGate*>& ds, const StageCtx& sc,
                        Program& prog) {
  size_t i = 0;
  while (i < ds.size()) {
    std::vector<int> ub;  // non-register global qubits of this table
    size_t j = i;
    for (; j < ds.size(); ++j) {
      std::vector<int> add;
      for (int t = 0; t < ds[j]->nd; ++t) {
        int q = ds[j]->dq[t];
        if (sc.slot_of(q) < 0 && std::find(ub.begin(), ub.end(), q) == ub.end() &&
            std::find(add.begin(), add.end(), q) == add.end())
          add.push_back(q);
      }
      if ((int)(ub.size() + add.size()) > kMaxTabBits) break;
      ub.insert(ub.end(), add.begin(), add.end());
    }
    GateDev gd = blank_gate();
    gd.kind = GK_DIAGTAB;
    gd.nu = (int)ub.size();
    for (int t = 0; t < gd.nu; ++t) gd.ubits[t] = ub[t];
    gd.tab = (int)(prog.tabs.size() / 2);
    const int rows = 1 << gd.nu;
    for (int row = 0; row < rows; ++row) {
      for (int jj = 0; jj < (1 << kRegQubits); ++jj) {
        cd ph(1.0, 0.0);
        for (size_t g = i; g < j; ++g) {
          int e = 0;
          for (int t = 0; t < ds[g]->nd; ++t) {
            int q = ds[g]->dq[t];
            int sl = sc.slot_of(q);
            int bit = sl >= 0 ? (jj >> sl) & 1
                              : (row >> (int)(std::find(ub.begin(), ub.end(), q) - ub.begin())) & 1;
            e |= bit << t;
          }
          ph *= cd(ds[g]->m[2 * e], ds[g]->m[2 * e + 1]);
        }
        prog.tabs.push_back(ph.real());
        prog.tabs.push_back(ph.imag());
      }
    }
    prog.gates.push_back(gd);
    i = j;
  }
}

// A run of diagonal gates commutes internally, so it is regrouped by where
// its qubits live: CZs, register-only, lane-only and mixed gates.
static void emit_diag_run(const std::vector<const IGate*>& run, const StageCtx& sc,
                          Program& prog) {
  std::vector<const IGate*> czs, reg, lane, mixed;
  for (const IGate* g : run) {
    if (is_cz(*g)) {
      czs.push_back(g);
      continue;
    }
    int nreg = 0;
    for (int t = 0; t < g->nd; ++t) nreg += sc.slot_of(g->dq[t]) >= 0;
    if (nreg == g->nd)
      reg.push_back(g);
    else if (nreg == 0)
      lane.push_back(g);
    else
      mixed.push_back(g);
  }
  if (!czs.empty()) emit_czs(czs, sc, prog);
  if (!reg.empty()) emit_diag_reg(reg, sc, prog);
  if (!lane.empty()) emit_diag_lane(lane, prog);
  if (!mixed.empty()) emit_tables(mixed, sc, prog);
}

// local positions a gate needs in register slots (bit mask)
static uint32_t need_local(const IGate& g, const int* g2l) {
  uint32_t need = 0;
  if (g.kind == IK_U1) need = 1u << g2l[g.t0];
  if (g.kind == IK_U2) need = (1u << g2l[g.t0]) | (1u << g2l[g.t1]);
  if (g.kind == IK_OVL) {
    need = 1u << g2l[g.sel];
    if (g.t0 >= 0) need |= 1u << g2l[g.t0];
    if (g.t1 >= 0) need |= 1u << g2l[g.t1];
  }
  return need;
}

bool program_has_swaps(const Program& prog) {
  for (const StageDev& st : prog.stages)
    if (st.cont) return true;
  return false;
}

namespace {
// Where a stage keeps the tile's local positions: reg[slot], tb[thread bit].
struct Mapping {
  int reg[kRegQubits];
  int tb[kMaxTileQubits - kRegQubits];
  int ntb;
};
}  // namespace

static StageDev stage_of(const Mapping& M, const PassDev& pd, int row_bits) {
  StageDev sd;
  std::memset(&sd, 0, sizeof(sd));
  for (int r = 0; r < kRegQubits; ++r) {
    sd.q[r] = M.reg[r];
    sd.sw_reg[r] = swz_host(1u << M.reg[r], row_bits);
  }
  for (int i = 0; i < M.ntb; ++i) {
    sd.tpos[i] = (int8_t)M.tb[i];
    sd.sw_tid[i] = swz_host(1u << M.tb[i], row_bits);
    sd.g_tid[i] = 1ull << pd.map[M.tb[i]];
  }
  sd.ntb = M.ntb;
  return sd;
}

// Emits the gates of one stage's run under mapping M (slot assignments
// from it), diagonal runs merged as usual.
namespace {
struct GateSink {
  const std::vector<IGate>& gates;
  const int* g2l;
  int n;
  Program& prog;
  std::vector<const IGate*> run;
  int l2slot[kMaxTileQubits];
  void set_mapping(const Mapping& M, int k) {
    for (int b = 0; b < k; ++b) l2slot[b] = -1;
    for (int r = 0; r < kRegQubits; ++r) l2slot[M.reg[r]] = r;
  }
  void flush() {
    if (run.empty()) return;
    StageCtx sc{g2l, l2slot, n};
    emit_diag_run(run, sc, prog);
    run.clear();
  }
  void add(const IGate& g) {
    if (g.kind == IK_DIAG) {
      run.push_back(&g);
      return;
    }
    flush();
    StageCtx sc{g2l, l2slot, n};
    if (g.kind == IK_OVL)
      emit_ovl(g, sc, prog);
    else
      emit_nondiag(g, sc, prog);
  }
};
}  // namespace

// Stages with lane exchanges: a stage (one LDS round trip) takes consecutive
// gates whose targets span at most five local positions, three in register
// slots and two on lane bits 4 and 5.  A gate on a lane-bit qubit first
// swaps it with the register slot whose qubit is needed furthest ahead
// (v_permlane16/32_swap, no LDS); each swap starts a continuation stage with
// the new mapping.  Lane bits 0..row_bits-1 get positions of distinct
// residues mod row_bits where possible, so that a row of the swizzled LDS
// tile is read without bank conflicts.
static void form_stages_swapping(const std::vector<IGate>& gates, const std::vector<int>& taken,
                                 const int* g2l, int n, int k, int row_bits, const PassDev& pd,
                                 Program& prog) {
  const int ntb = k - kRegQubits;
  GateSink sink{gates, g2l, n, prog, {}, {}};
  size_t i = 0;
  while (i < taken.size()) {
    // the segment's positions in order of first need (at most 5)
    std::vector<int> F;
    size_t j = i;
    for (; j < taken.size(); ++j) {
      const uint32_t need = need_local(gates[taken[j]], g2l);
      std::vector<int> add;
      for (int b = 0; b < k; ++b)
        if ((need >> b & 1) && std::find(F.begin(), F.end(), b) == F.end()) add.push_back(b);
      if (F.size() + add.size() > (size_t)kRegQubits + 2) break;
      F.insert(F.end(), add.begin(), add.end());
    }
    Mapping M;
    M.ntb = ntb;
    std::vector<char> used(k, 0);
    int nr = 0;
    for (size_t f = 0; f < F.size() && nr < kRegQubits; ++f) {
      M.reg[nr++] = F[f];
      used[F[f]] = 1;
    }
    for (int b = 0; b < k && nr < kRegQubits; ++b)
      if (!used[b]) {
        M.reg[nr++] = b;
        used[b] = 1;
      }
    for (int t = 0; t < ntb; ++t) M.tb[t] = -1;
    for (int l = 0; l < 2; ++l) {
      const size_t f = kRegQubits + l;
      if (f < F.size()) {
        M.tb[4 + l] = F[f];
        used[F[f]] = 1;
      }
    }
    // lane bits 0..row_bits-1: distinct residues mod row_bits first
    std::vector<char> res_taken(row_bits, 0);
    for (int t = 0; t < row_bits && t < 4; ++t) {
      int pick = -1;
      for (int b = 0; b < k && pick < 0; ++b)
        if (!used[b] && !res_taken[b % row_bits]) pick = b;
      for (int b = 0; b < k && pick < 0; ++b)
        if (!used[b]) pick = b;
      M.tb[t] = pick;
      used[pick] = 1;
      res_taken[pick % row_bits] = 1;
    }
    for (int t = 0; t < ntb; ++t) {
      if (M.tb[t] >= 0) continue;
      for (int b = 0; b < k; ++b)
        if (!used[b]) {
          M.tb[t] = b;
          used[b] = 1;
          break;
        }
    }
    sink.set_mapping(M, k);
    StageDev sd = stage_of(M, pd, row_bits);
    sd.g0 = (int)prog.gates.size();
    for (size_t t = i; t < j; ++t) {
      const IGate& g = gates[taken[t]];
      const uint32_t need = need_local(g, g2l);
      for (int b = 0; b < k; ++b) {
        if (!(need >> b & 1)) continue;
        bool in_reg = false;
        for (int r = 0; r < kRegQubits; ++r) in_reg |= M.reg[r] == b;
        if (in_reg) continue;
        const int lb = M.tb[4] == b ? 4 : 5;
        // victim: the slot not needed now whose next need is furthest
        int victim = -1;
        size_t best = 0;
        for (int r = 0; r < kRegQubits; ++r) {
          if (need >> M.reg[r] & 1) continue;
          size_t nxt = taken.size() + 1;
          for (size_t u = t + 1; u < j; ++u)
            if (need_local(gates[taken[u]], g2l) >> M.reg[r] & 1) {
              nxt = u;
              break;
            }
          if (victim < 0 || nxt > best) {
            victim = r;
            best = nxt;
          }
        }
        // a pending diagonal run may be applied under either mapping: it
        // stays pending, so that it is emitted with the register slots of
        // the gate that ends it
        sd.g1 = (int)prog.gates.size();
        prog.stages.push_back(sd);
        std::swap(M.reg[victim], M.tb[lb]);
        sink.set_mapping(M, k);
        sd = stage_of(M, pd, row_bits);
        sd.cont = 1;
        sd.swap_slot = (int8_t)victim;
        sd.swap_bit = (int8_t)lb;
        sd.g0 = (int)prog.gates.size();
      }
      sink.add(g);
    }
    // a diagonal run still pending at the segment's end moves on to the
    // next segment (the pass's last one emits it)
    if (j == taken.size()) sink.flush();
    sd.g1 = (int)prog.gates.size();
    prog.stages.push_back(sd);
    i = j;
  }
}

// A diagonal gate other than a CZ commutes with every gate that has none of
// its qubits as a target, so inside a pass it moves down to just before the
// first later gate that does (or to the pass's end).  There that gate's
// target is a register slot, so a one-qubit phase becomes a register
// diagonal merged with the others that wait for the same gate, instead of a
// per-lane phase of its own.  CZs stay where they are: the CZs of a layer
// stay together and merge into one sign mask.  The diagonal gates that no
// later gate of the pass stops come back in `tail` (in gate order), at the
// end of `taken` unless tail_out is set.
static void sink_diagonals(const std::vector<IGate>& gates, std::vector<int>& taken,
                           std::vector<int>* tail_out) {
  std::vector<int> order, pending, keep;
  order.reserve(taken.size());
  for (int gi : taken) {
    const IGate& g = gates[gi];
    if (g.kind == IK_DIAG && !is_cz(g)) {
      pending.push_back(gi);
      continue;
    }
    // qubits on which g does not commute with a diagonal gate
    const uint64_t nc = g.kind == IK_DIAG ? 0 : g.kind == IK_OVL ? touch_mask(g) : target_mask(g);
    if (nc && !pending.empty()) {
      keep.clear();
      for (int d : pending) {
        if (touch_mask(gates[d]) & nc)
          order.push_back(d);
        else
          keep.push_back(d);
      }
      pending.swap(keep);
    }
    order.push_back(gi);
  }
  if (tail_out)
    tail_out->swap(pending);
  else
    order.insert(order.end(), pending.begin(), pending.end());
  taken.swap(order);
}

void set_pass_io(PassDev& pd, int io, int low) {
  const int k = pd.k;
  // the three register positions of the HBM <-> LDS moves: the tile's top
  // three (io bit 6 clear) or the three above its contiguous low qubits
  int r0 = k - kRegQubits;
  if ((io & 64) && low + kRegQubits <= k - kRegQubits) r0 = low;
  pd.ltm = 0;
  pd.tid_mask = 0;
  int hj = 0;
  for (int b = 0; b < k; ++b) {
    if (b >= r0 && b < r0 + kRegQubits) {
      pd.hb[hj++] = 1ull << pd.map[b];
    } else {
      pd.ltm |= 1u << b;
      pd.tid_mask |= 1ull << pd.map[b];
    }
  }
  pd.tord = io & 63;
}

void apply_pass_io(Program& prog, const std::vector<uint8_t>& io, int low) {
  if (io.size() != prog.passes.size()) return;
  for (size_t p = 0; p < io.size(); ++p) {
    PassDev& pd = prog.passes[p];
    // the contiguous low qubits as plan_program keeps them
    const int m = pd.n_nonlocal == 0 ? pd.k : std::min(low, pd.k - 2);
    set_pass_io(pd, io[p], m);
  }
}

// One pass's gates in order: a gate joins when none of its qubits is
// touched by a gate already deferred (so it commutes with all it overtakes),
// the pass's cap on non-diagonal gates allows it and its targets fit the
// local set (grown up to k qubits when `grow`, else fixed).  Returns the
// local set.
static uint64_t take_gates(const std::vector<IGate>& gates, const std::vector<int>& remaining,
                           uint64_t L, bool grow, int k, int max_gates, std::vector<int>& taken,
                           std::vector<int>& deferred) {
  uint64_t blocked = 0;
  int nondiag = 0;
  for (int gi : remaining) {
    const IGate& g = gates[gi];
    const uint64_t touch = touch_mask(g);
    if (touch & blocked) {
      deferred.push_back(gi);
      blocked |= touch;
      continue;
    }
    if (max_gates > 0 && g.kind != IK_DIAG && nondiag >= max_gates && !taken.empty()) {
      deferred.push_back(gi);
      blocked |= touch;
      continue;
    }
    const uint64_t nl = L | target_mask(g);
    if (nl == L || (grow && popc(nl) <= k)) {
      L = nl;
      taken.push_back(gi);
      nondiag += g.kind != IK_DIAG;
    } else {
      deferred.push_back(gi);
      blocked |= touch;
    }
  }
  return L;
}

// non-diagonal gates a pass with the fixed local set L would take
static int count_taken(const std::vector<IGate>& gates, const std::vector<int>& remaining,
                       uint64_t L, int max_gates) {
  uint64_t blocked = 0;
  int nondiag = 0, taken = 0;
  for (int gi : remaining) {
    const IGate& g = gates[gi];
    const uint64_t touch = touch_mask(g);
    if (touch & blocked) {
      blocked |= touch;
      continue;
    }
    if (max_gates > 0 && g.kind != IK_DIAG && nondiag >= max_gates && taken > 0) {
      blocked |= touch;
      continue;
    }
    if ((L | target_mask(g)) == L) {
      ++taken;
      nondiag += g.kind != IK_DIAG;
    } else {
      blocked |= touch;
    }
  }
  return nondiag;
}

void search_local_sets(const std::vector<IGate>& gates, int n, int tile_qubits, int low_qubits,
                       int max_gates, int iters, uint64_t seed, std::vector<uint64_t>& locals) {
  locals.clear();
  const int k = std::min(tile_qubits, n);
  if (k >= n) return;  // one tile holds the state: nothing to choose
  const int m = std::min(low_qubits, k - 2);
  const uint64_t low = (1ull << m) - 1;
  const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
  uint64_t rng = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  auto pick = [&](uint64_t mask) {  // a uniformly chosen set bit of mask
    int c = (int)(next() % (uint64_t)popc(mask));
    for (int b = 0; b < 64; ++b)
      if ((mask >> b & 1) && c-- == 0) return b;
    return -1;
  };
  std::vector<int> taken, deferred;
  // one pass's set: the greedy set, then random exchanges of one non-low
  // local qubit with another qubit, kept when the pass takes at least as
  // many non-diagonal gates
  auto one_set = [&](const std::vector<int>& rem) {
    taken.clear();
    deferred.clear();
    uint64_t cur = take_gates(gates, rem, low, true, k, max_gates, taken, deferred);
    for (int b = 0; b < n && popc(cur) < k; ++b) cur |= 1ull << b;
    int best = count_taken(gates, rem, cur, max_gates);
    for (int it = 0; it < iters && (all & ~cur); ++it) {
      const int out = pick(cur & ~low), in = pick(all & ~cur);
      const uint64_t cand = (cur & ~(1ull << out)) | (1ull << in);
      const int sc = count_taken(gates, rem, cand, max_gates);
      if (sc >= best) {
        best = sc;
        cur = cand;
      }
    }
    return cur;
  };
  // Beam search over the passes: each kept plan prefix tries kBeamTries
  // searched sets for its next pass; the kBeamWidth prefixes that leave the
  // fewest gates (distinct remainders) go on.
  constexpr int kBeamWidth = 4, kBeamTries = 6;
  struct Prefix {
    std::vector<int> rem;
    std::vector<uint64_t> sets;
  };
  std::vector<Prefix> beam(1);
  beam[0].rem.resize(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) beam[0].rem[i] = (int)i;
  while (true) {
    for (const Prefix& p : beam)
      if (p.rem.empty()) {
        locals = p.sets;
        return;
      }
    std::vector<Prefix> next_beam;
    for (const Prefix& p : beam) {
      std::vector<uint64_t> tried;
      for (int t = 0; t < kBeamTries; ++t) {
        const uint64_t L = one_set(p.rem);
        if (std::find(tried.begin(), tried.end(), L) != tried.end()) continue;
        tried.push_back(L);
        taken.clear();
        deferred.clear();
        take_gates(gates, p.rem, L, false, k, max_gates, taken, deferred);
        if (taken.empty()) continue;
        Prefix q;
        q.rem = deferred;
        q.sets = p.sets;
        q.sets.push_back(L);
        next_beam.push_back(std::move(q));
      }
    }
    std::stable_sort(next_beam.begin(), next_beam.end(),
                     [](const Prefix& a, const Prefix& b) { return a.rem.size() < b.rem.size(); });
    beam.clear();
    for (Prefix& q : next_beam) {
      bool dup = false;
      for (const Prefix& b : beam) dup |= b.rem == q.rem;
      if (dup) continue;
      beam.push_back(std::move(q));
      if ((int)beam.size() == kBeamWidth) break;
    }
    if (beam.empty()) return;  // no progress (cannot happen: the greedy set takes a gate)
  }
}

void plan_program(const std::vector<IGate>& gates, int n, int tile_qubits, int low_qubits,
                  int row_bits, Program& prog, int max_gates, bool lane_swaps,
                  const std::vector<uint64_t>* locals) {
  prog.passes.clear();
  prog.stages.clear();
  prog.gates.clear();
  prog.tabs.clear();
  prog.per_item = false;
  const int k = std::min(tile_qubits, n);
  // keep room for two non-low targets so that every gate fits an empty pass
  const int m = (k == n) ? k : std::min(low_qubits, k - 2);
  std::vector<int> remaining(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) remaining[i] = (int)i;
  // max_gates: cap on the non-diagonal gates one pass takes (0: none), which
  // spreads deep passes over more passes

  size_t pass_no = 0;
  while (!remaining.empty()) {
    std::vector<int> taken, deferred;
    uint64_t L;
    if (locals && pass_no < locals->size()) {
      // a local set chosen beforehand (search_local_sets): it stays fixed
      L = take_gates(gates, remaining, (*locals)[pass_no], false, k, max_gates, taken, deferred);
      if (taken.empty()) {
        // not this gate list's sets: greedy from here on
        taken.clear();
        deferred.clear();
        locals = nullptr;
      }
    }
    if (!locals || pass_no >= locals->size())
      L = take_gates(gates, remaining, (m >= 64) ? ~0ull : ((1ull << m) - 1), true, k, max_gates,
                     taken, deferred);
    ++pass_no;
    // pad the local set to exactly k qubits with the lowest free bits
    for (int b = 0; b < n && popc(L) < k; ++b) L |= 1ull << b;
    // Diagonal gates that sink to the pass's end move on to the next pass
    // when there is one (some non-diagonal gate was deferred): every gate
    // before them that they do not commute with is in this pass or an
    // earlier one, so the next pass takes them again, and there they sink
    // to the gate that targets their qubit.
    bool more = false;
    for (int gi : deferred) more |= gates[gi].kind != IK_DIAG;
    if (more) {
      std::vector<int> tail;
      sink_diagonals(gates, taken, &tail);
      if (!tail.empty()) {
        std::vector<int> merged;
        merged.reserve(deferred.size() + tail.size());
        std::merge(deferred.begin(), deferred.end(), tail.begin(), tail.end(),
                   std::back_inserter(merged));
        deferred.swap(merged);
      }
    } else {
      sink_diagonals(gates, taken, nullptr);
    }

    PassDev pd;
    std::memset(&pd, 0, sizeof(pd));
    pd.k = k;
    pd.scale = 1.0;
    int li = 0, ni = 0;
    int g2l[kMaxQubits];
    for (int b = 0; b < n; ++b) {
      if (L >> b & 1) {
        g2l[b] = li;
        pd.map[li++] = b;
      } else {
        g2l[b] = -1;
        pd.nl_mask |= 1ull << b;
        ni++;
      }
    }
    pd.n_nonlocal = ni;
    set_pass_io(pd, 0, m);
    pd.s0 = (int)prog.stages.size();

    // register stages with lane exchanges (tiles of at least one wave)
    if (lane_swaps && k - kRegQubits >= 6) {
      form_stages_swapping(gates, taken, g2l, n, k, row_bits, pd, prog);
      pd.s1 = (int)prog.stages.size();
      prog.passes.push_back(pd);
      remaining.swap(deferred);
      continue;
    }

    // register stages: consecutive gates whose targets fit kRegQubits slots
    size_t i = 0;
    while (i < taken.size()) {
      uint32_t S = 0;  // local positions held in registers
      size_t j = i;
      for (; j < taken.size(); ++j) {
        const IGate& g = gates[taken[j]];
        uint32_t need = 0;
        if (g.kind == IK_U1) need = 1u << g2l[g.t0];
        if (g.kind == IK_U2) need = (1u << g2l[g.t0]) | (1u << g2l[g.t1]);
        if (g.kind == IK_OVL) {
          need = 1u << g2l[g.sel];
          if (g.t0 >= 0) need |= 1u << g2l[g.t0];
          if (g.t1 >= 0) need |= 1u << g2l[g.t1];
        }
        if (popc(S | need) > kRegQubits) break;
        S |= need;
      }
      // diagonal gates after the stage's last non-diagonal gate wait for
      // the next stage, where the gate that ends them has its target in a
      // register slot (the pass's last stage keeps them)
      if (j < taken.size()) {
        size_t e = j;
        while (e > i && gates[taken[e - 1]].kind == IK_DIAG) --e;
        if (e > i) j = e;
      }
      for (int b = 0; b < k && popc(S) < kRegQubits; ++b) S |= 1u << b;
      StageDev sd;
      std::memset(&sd, 0, sizeof(sd));
      int slot = 0, ti = 0;
      int l2slot[kMaxTileQubits];
      for (int b = 0; b < k; ++b) {
        l2slot[b] = -1;
        if (S >> b & 1) {
          l2slot[b] = slot;
          sd.sw_reg[slot] = swz_host(1u << b, row_bits);
          sd.q[slot++] = b;
        } else {
          sd.sw_tid[ti] = swz_host(1u << b, row_bits);
          sd.tpos[ti] = (int8_t)b;
          sd.g_tid[ti++] = 1ull << pd.map[b];
        }
      }
      sd.ntb = ti;
      StageCtx sc{g2l, l2slot, n};
      sd.g0 = (int)prog.gates.size();
      std::vector<const IGate*> run;
      for (size_t t = i; t < j; ++t) {
        const IGate& g = gates[taken[t]];
        if (g.kind == IK_DIAG) {
          run.push_back(&g);
          continue;
        }
        if (!run.empty()) {
          emit_diag_run(run, sc, prog);
          run.clear();
        }
        if (g.kind == IK_OVL)
          emit_ovl(g, sc, prog);
        else
          emit_nondiag(g, sc, prog);
      }
      if (!run.empty()) emit_diag_run(run, sc, prog);
      sd.g1 = (int)prog.gates.size();
      prog.stages.push_back(sd);
      i = j;
    }
    pd.s1 = (int)prog.stages.size();
    prog.passes.push_back(pd);
    remaining.swap(deferred);
  }
}

}  // namespace smoq
