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
#include <algorithm>

#include "smoq_internal.h"

namespace smoq {

static uint64_t touch_mask(const IGate& g) {
  uint64_t m = g.ctrl;
  if (g.kind == IK_U1) m |= 1ull << g.t0;
  if (g.kind == IK_U2) m |= (1ull << g.t0) | (1ull << g.t1);
  if (g.kind == IK_DIAG)
    for (int i = 0; i < g.nd; ++i) m |= 1ull << g.dq[i];
  return m;
}

static uint64_t target_mask(const IGate& g) {
  if (g.kind == IK_U1) return 1ull << g.t0;
  if (g.kind == IK_U2) return (1ull << g.t0) | (1ull << g.t1);
  return 0;
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }

void plan_program(const std::vector<IGate>& gates, int n, int tile_qubits, int low_qubits,
                  Program& prog) {
  prog.passes.clear();
  prog.stages.clear();
  prog.gates.clear();
  const int k = std::min(tile_qubits, n);
  // keep room for two non-low targets so that every gate fits an empty pass
  const int m = (k == n) ? k : std::min(low_qubits, k - 2);
  std::vector<int> remaining(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) remaining[i] = (int)i;

  while (!remaining.empty()) {
    uint64_t L = (m >= 64) ? ~0ull : ((1ull << m) - 1);
    uint64_t blocked = 0;
    std::vector<int> taken, deferred;
    for (int gi : remaining) {
      const IGate& g = gates[gi];
      uint64_t touch = touch_mask(g);
      if (touch & blocked) {
        deferred.push_back(gi);
        blocked |= touch;
        continue;
      }
      uint64_t nl = L | target_mask(g);
      if (popc(nl) <= k) {
        L = nl;
        taken.push_back(gi);
      } else {
        deferred.push_back(gi);
        blocked |= touch;
      }
    }
    // pad the local set to exactly k qubits with the lowest free bits
    for (int b = 0; b < n && popc(L) < k; ++b) L |= 1ull << b;

    PassDev pd;
    std::memset(&pd, 0, sizeof(pd));
    pd.k = k;
    int li = 0, ni = 0;
    int g2l[kMaxQubits];
    for (int b = 0; b < n; ++b) {
      if (L >> b & 1) {
        g2l[b] = li;
        pd.map[li++] = b;
      } else {
        g2l[b] = -1;
        pd.nonlocal[ni++] = b;
      }
    }
    pd.n_nonlocal = ni;
    pd.s0 = (int)prog.stages.size();

    // register stages
    size_t i = 0;
    while (i < taken.size()) {
      uint32_t S = 0;  // local positions held in registers
      size_t j = i;
      for (; j < taken.size(); ++j) {
        const IGate& g = gates[taken[j]];
        uint32_t need = 0;
        if (g.kind == IK_U1) need = 1u << g2l[g.t0];
        if (g.kind == IK_U2) need = (1u << g2l[g.t0]) | (1u << g2l[g.t1]);
        if (popc(S | need) > kRegQubits) break;
        S |= need;
      }
      // pad to kRegQubits positions
      for (int b = 0; b < k && popc(S) < kRegQubits; ++b) S |= 1u << b;
      StageDev sd;
      int slot = 0;
      int l2slot[kMaxTileQubits];
      for (int b = 0; b < k; ++b) {
        l2slot[b] = -1;
        if (S >> b & 1) {
          l2slot[b] = slot;
          sd.q[slot++] = b;
        }
      }
      sd.g0 = (int)prog.gates.size();
      for (size_t t = i; t < j; ++t) {
        const IGate& g = gates[taken[t]];
        GateDev gd;
        std::memset(&gd, 0, sizeof(gd));
        gd.kind = g.kind;
        gd.a = g.kind == IK_DIAG ? -1 : l2slot[g2l[g.t0]];
        gd.b = g.kind == IK_U2 ? l2slot[g2l[g.t1]] : -1;
        gd.nd = g.nd;
        for (int d = 0; d < 3; ++d) gd.dq[d] = g.dq[d];
        gd.ctrl = g.ctrl;
        std::memcpy(gd.m, g.m, sizeof(gd.m));
        prog.gates.push_back(gd);
      }
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
