// Compiled tiled Pauli reductions: one kernel per pass, generated for a list
// of Pauli terms (the 3N local and 3(N-1) neighbour observables of
// measure_all_observables_state_vector_cpu!, cpuQuantumStateObservables.jl:
// 1483-1519, or any Hamiltonian's terms) and compiled at run time.
//
// A pass owns 12 tile bits (bits 0..3 always, so loads run over 256-byte
// segments).  A workgroup of 256 threads takes one 2^12-amplitude tile at a
// time, each thread 16 amplitudes in registers.  The tile is read from HBM
// once; the thread's 16 amplitudes then sit at four "register" tile bits,
// and every x mask inside those four bits is evaluated from registers: the
// pairs (l, l ^ x) are (v[j], v[j ^ r]).  Further stages read the tile back
// from LDS with other register bits until every x mask of at most four
// bits is covered; wider masks read their partner amplitude from LDS.  Each
// term keeps one FP64 accumulator per thread across the tiles (terms whose
// per-tile values agree and whose signs differ only by the thread's own
// index bits share one), so a term costs a few FMAs per pair and one wave
// sum at the end.  All indices, signs and register pairings are constants
// of the generated code.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "smoq_internal.h"

namespace smoq {

namespace {

constexpr int kTB = kOjTileBits;  // tile bits
constexpr int kR = kOjRegBits;    // register bits per stage
constexpr int kThr = kTB - kR;    // thread bits (kOjThreads threads)
constexpr int kLowBits = 4;       // tile bits 0..3 are global bits 0..3

uint32_t pext_pos(uint64_t v, const int* pos, int k) {
  uint32_t r = 0;
  for (int j = 0; j < k; ++j) r |= (uint32_t)((v >> pos[j]) & 1ull) << j;
  return r;
}

uint32_t pdep_pos(uint32_t v, const int* pos, int k) {
  uint32_t r = 0;
  for (int j = 0; j < k; ++j) r |= ((v >> j) & 1u) << pos[j];
  return r;
}

// XOR swizzle of an LDS index (W = log2 of the amplitudes per 256-byte row
// slot group): linear over GF(2), so swz(a ^ b) = swz(a) ^ swz(b)
uint32_t swz_w(uint32_t l, int W) {
  const uint32_t f = (l >> W) ^ (l >> (2 * W)) ^ (l >> (3 * W));
  return l ^ (f & ((1u << W) - 1u));
}

// A stage's register bits are usable when every residue class of the tile
// bits mod W keeps a thread bit: lanes 0..2^W-1 then take one bit of each
// class and their LDS rows fall into distinct bank groups.
bool regs_valid(uint32_t R, int W) {
  for (int c = 0; c < W; ++c) {
    bool any = false;
    for (int b = c; b < kTB; b += W) any |= !((R >> b) & 1u);
    if (!any) return false;
  }
  return true;
}

void lane_bits(uint32_t R, int W, int* tbit) {
  uint32_t used = R;
  int i = 0;
  for (int c = 0; c < W; ++c)
    for (int b = c; b < kTB; b += W)
      if (!((used >> b) & 1u)) {
        tbit[i++] = b;
        used |= 1u << b;
        break;
      }
  for (int b = 0; b < kTB; ++b)
    if (!((used >> b) & 1u)) {
      tbit[i++] = b;
      used |= 1u << b;
    }
}

std::vector<uint32_t> all_reg_sets(bool stage0) {
  std::vector<uint32_t> out;
  for (uint32_t m = 0; m < (1u << kTB); ++m)
    if (__builtin_popcount(m) == kR && (!stage0 || (m & ((1u << kLowBits) - 1u)) == 0))
      out.push_back(m);
  return out;
}

struct TermIn {
  uint64_t x, z;
  int col;
};

// Stages, slots and term records of one pass over the tile bits tpos.
void build_pass(int W, const int* tpos, const std::vector<TermIn>& in, OjPass& P) {
  P = OjPass{};
  std::memcpy(P.tpos, tpos, sizeof(P.tpos));
  uint64_t tmask = 0;
  for (int j = 0; j < kTB; ++j) tmask |= 1ull << tpos[j];
  std::vector<uint32_t> masks;
  for (const TermIn& t : in)
    if (t.x) {
      const uint32_t lx = pext_pos(t.x, tpos, kTB);
      if (__builtin_popcount(lx) <= kR && std::find(masks.begin(), masks.end(), lx) == masks.end())
        masks.push_back(lx);
    }
  std::vector<char> done(masks.size(), 0);
  auto pick = [&](const std::vector<uint32_t>& cands) {
    uint32_t best = 0;
    int bc = -1;
    for (uint32_t R : cands) {
      if (!regs_valid(R, W)) continue;
      int c = 0;
      for (size_t i = 0; i < masks.size(); ++i) c += !done[i] && (masks[i] & ~R) == 0;
      if (c > bc) {
        bc = c;
        best = R;
      }
    }
    return std::make_pair(best, bc);
  };
  static const std::vector<uint32_t> sets0 = all_reg_sets(true), sets = all_reg_sets(false);
  auto add_stage = [&](uint32_t R) {
    OjStage st{};
    st.rmask = R;
    int j = 0;
    for (int b = 0; b < kTB; ++b)
      if ((R >> b) & 1u) st.rpos[j++] = b;
    lane_bits(R, W, st.tbit);
    for (size_t i = 0; i < masks.size(); ++i)
      if ((masks[i] & ~R) == 0) done[i] = 1;
    P.st.push_back(st);
  };
  add_stage(pick(sets0).first);
  for (;;) {
    bool left = false;
    for (char d : done) left |= !d;
    if (!left) break;
    const auto b = pick(sets);
    if (b.second <= 0) break;  // the rest read their partners from LDS
    add_stage(b.first);
  }
  std::map<std::tuple<int, int, uint32_t, int, uint32_t, uint64_t>, int> slot_of;
  auto slot = [&](int stage, int kind, uint32_t m, int comp, uint32_t zr, uint64_t zb) {
    const auto key = std::make_tuple(stage, kind, m, comp, zr, zb);
    auto it = slot_of.find(key);
    if (it != slot_of.end()) return it->second;
    OjSlot s{};
    s.stage = stage;
    s.kind = kind;
    s.m = m;
    s.comp = comp;
    s.zr = zr;
    s.zb = zb;
    P.sl.push_back(s);
    const int id = (int)P.sl.size() - 1;
    slot_of.emplace(key, id);
    return id;
  };
  for (const TermIn& t : in) {
    const uint32_t lx = pext_pos(t.x, tpos, kTB), lz = pext_pos(t.z, tpos, kTB);
    const uint64_t zb = t.z & ~tmask;
    const int kind = __builtin_popcountll(t.x & t.z) & 3;
    OjTerm tm{};
    tm.col = t.col;
    if (!t.x) {
      const OjStage& st = P.st[0];
      tm.stage = 0;
      tm.slot = slot(0, 0, 0, 0, pext_pos(lz, st.rpos, kR), zb);
      tm.lzt = lz & ~st.rmask;
      tm.scale = 1;
    } else {
      int s = -1;
      for (size_t i = 0; i < P.st.size() && s < 0; ++i)
        if ((lx & ~P.st[i].rmask) == 0) s = (int)i;
      const bool neg = kind == 1 || kind == 2;
      if (s >= 0) {
        // pairs (j, j ^ r) with j's highest bit of r clear: half the sum,
        // 2 Re or 2 Im of it by the phase i^kind
        const OjStage& st = P.st[s];
        const uint32_t r = pext_pos(lx, st.rpos, kR);
        const uint32_t hr = 1u << (31 - __builtin_clz(r));
        tm.stage = s;
        tm.slot = slot(s, 1, r, kind & 1, pext_pos(lz, st.rpos, kR) & ~hr, zb);
        tm.lzt = lz & ~st.rmask;
        tm.scale = neg ? -2 : 2;
      } else {
        // partner from LDS, every l of the tile: the whole sum
        const OjStage& st = P.st[0];
        tm.stage = 0;
        tm.slot = slot(0, 2, lx, kind & 1, pext_pos(lz, st.rpos, kR), zb);
        tm.lzt = lz & ~st.rmask;
        tm.scale = neg ? -1 : 1;
      }
    }
    P.tm.push_back(tm);
  }
}

// The OR of shifted runs placing bit i of `src` at dst[i].
std::string pdep_expr(const std::string& src, const std::vector<int>& dst, bool wide) {
  std::ostringstream o;
  bool first = true;
  size_t i = 0;
  while (i < dst.size()) {
    size_t j = i + 1;
    while (j < dst.size() && dst[j] == dst[j - 1] + 1) ++j;
    const uint64_t m = ((j - i) >= 64 ? ~0ull : ((1ull << (j - i)) - 1ull)) << i;
    const int sh = dst[i] - (int)i;
    if (!first) o << " | ";
    first = false;
    o << "((" << src << " & 0x" << std::hex << m << std::dec << (wide ? "ull" : "u") << ")";
    if (sh > 0) o << " << " << sh;
    if (sh < 0) o << " >> " << -sh;
    o << ")";
    i = j;
  }
  if (first) o << (wide ? "0ull" : "0u");
  return o.str();
}

std::string hex64(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
  return b;
}

// sum over the listed register indices of sign(j) * NAME_j, in j order
void emit_signed_sum(std::ostringstream& o, const std::string& name, const std::vector<int>& js,
                     uint32_t zr) {
  o << "      const double u = ";
  for (size_t i = 0; i < js.size(); ++i) {
    const bool neg = __builtin_popcount((uint32_t)js[i] & zr) & 1;
    if (i == 0)
      o << (neg ? "-" : "") << name << js[i];
    else
      o << (neg ? " - " : " + ") << name << js[i];
  }
  o << ";\n";
}

void emit_accumulate(std::ostringstream& o, int slot, uint64_t zb) {
  if (zb)
    o << "      a" << slot << " += (__popcll(base & " << hex64(zb) << ") & 1) ? -u : u;\n";
  else
    o << "      a" << slot << " += u;\n";
}

void emit_pass(std::ostringstream& o, int p, const OjPass& P, int n, int dtype) {
  const bool c128 = dtype == SMOQ_C128;
  const int W = c128 ? 4 : 5;
  const int nreg = 1 << kR;
  const int ns = (int)P.st.size(), nslot = (int)P.sl.size(), nterm = (int)P.tm.size();
  uint64_t tmask = 0;
  for (int j = 0; j < kTB; ++j) tmask |= 1ull << P.tpos[j];
  std::vector<int> cpos;
  for (int b = 0; b < n; ++b)
    if (!((tmask >> b) & 1ull)) cpos.push_back(b);
  o << "extern \"C\" __global__ void __launch_bounds__(" << kOjThreads << ", 2) smoq_obs_" << p
    << "(const VT<T>* __restrict__ psi, unsigned long long dim, unsigned long long n_tiles,"
       " double* __restrict__ partial, int part_terms, int term0) {\n"
    << "  __shared__ __align__(16) VT<T> lds[" << (1 << kTB) << "];\n"
    << "  __shared__ double red[" << kOjThreads / 64 << "][" << nterm << "];\n"
    << "  const unsigned tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;\n"
    << "  const VT<T>* __restrict__ pp = psi + (unsigned long long)blockIdx.y * dim;\n";
  // thread parts: stage 0's global load offset, every stage's LDS row
  {
    std::vector<int> g;
    for (int i = 0; i < kThr; ++i) g.push_back(P.tpos[P.st[0].tbit[i]]);
    o << "  const unsigned long long tid64 = tid;\n"
      << "  const unsigned long long g0 = " << pdep_expr("tid64", g, true) << ";\n";
  }
  for (int s = 0; s < ns; ++s) {
    std::vector<int> d(P.st[s].tbit, P.st[s].tbit + kThr);
    o << "  const unsigned lt" << s << " = " << pdep_expr("tid", d, false) << ";\n"
      << "  const unsigned pb" << s << " = lt" << s << " ^ (((lt" << s << " >> " << W << ") ^ (lt"
      << s << " >> " << 2 * W << ") ^ (lt" << s << " >> " << 3 * W << ")) & " << ((1 << W) - 1)
      << "u);\n";
  }
  for (int k = 0; k < nslot; ++k) o << "  double a" << k << " = 0.0;\n";
  o << "  for (unsigned long long tt = blockIdx.x; tt < n_tiles; tt += gridDim.x) {\n"
    << "    const unsigned long long base = " << pdep_expr("tt", cpos, true) << ";\n"
    << "    const VT<T>* __restrict__ q = pp + (base | g0);\n";
  const OjStage& s0 = P.st[0];
  int gr[kR];
  for (int b = 0; b < kR; ++b) gr[b] = P.tpos[s0.rpos[b]];
  for (int j = 0; j < nreg; ++j) {
    uint64_t off = 0;
    for (int b = 0; b < kR; ++b)
      if ((j >> b) & 1) off |= 1ull << gr[b];
    o << "    const VT<T> v" << j << " = ld_stream<T>(q + " << hex64(off) << ");\n";
  }
  for (int j = 0; j < nreg; ++j)
    o << "    lds[pb0 ^ " << swz_w(pdep_pos(j, s0.rpos, kR), W) << "u] = v" << j << ";\n";
  // one stage's slots from registers X_j, Y_j
  auto emit_stage_slots = [&](int s) {
    bool diag = false;
    std::vector<std::pair<uint32_t, int>> prods;  // (r, comp) needed
    for (int k = 0; k < nslot; ++k) {
      const OjSlot& sl = P.sl[k];
      if (sl.stage != s || sl.kind == 2) continue;
      if (sl.kind == 0) diag = true;
      if (sl.kind == 1) {
        const auto rc = std::make_pair(sl.m, sl.comp);
        if (std::find(prods.begin(), prods.end(), rc) == prods.end()) prods.push_back(rc);
      }
    }
    if (diag) {
      for (int j = 0; j < nreg; ++j)
        o << "      const double P" << j << " = X" << j << " * X" << j << " + Y" << j << " * Y" << j
          << ";\n";
      std::vector<int> all;
      for (int j = 0; j < nreg; ++j) all.push_back(j);
      for (int k = 0; k < nslot; ++k) {
        const OjSlot& sl = P.sl[k];
        if (sl.stage != s || sl.kind != 0) continue;
        o << "    {\n";
        emit_signed_sum(o, "P", all, sl.zr);
        emit_accumulate(o, k, sl.zb);
        o << "    }\n";
      }
    }
    for (const auto& rc : prods) {
      const uint32_t r = rc.first;
      const uint32_t hr = 1u << (31 - __builtin_clz(r));
      std::vector<int> js;
      for (int j = 0; j < nreg; ++j)
        if (!(j & hr)) js.push_back(j);
      o << "    {\n";
      const char* nm = rc.second ? "WI" : "WR";
      for (int j : js) {
        const int b = j ^ (int)r;
        if (rc.second)
          o << "      const double WI" << j << " = X" << b << " * Y" << j << " - Y" << b << " * X" << j
            << ";\n";
        else
          o << "      const double WR" << j << " = X" << b << " * X" << j << " + Y" << b << " * Y" << j
            << ";\n";
      }
      for (int k = 0; k < nslot; ++k) {
        const OjSlot& sl = P.sl[k];
        if (sl.stage != s || sl.kind != 1 || sl.m != r || sl.comp != rc.second) continue;
        o << "     {\n";
        emit_signed_sum(o, nm, js, sl.zr);
        emit_accumulate(o, k, sl.zb);
        o << "     }\n";
      }
      o << "    }\n";
    }
  };
  o << "    {\n";
  for (int j = 0; j < nreg; ++j)
    o << "      const double X" << j << " = (double)v" << j << ".x, Y" << j << " = (double)v" << j
      << ".y;\n";
  emit_stage_slots(0);
  o << "    }\n    __syncthreads();\n";
  // masks whose partners are other threads' amplitudes (from LDS)
  std::vector<uint32_t> gms;
  for (const OjSlot& sl : P.sl)
    if (sl.kind == 2 && std::find(gms.begin(), gms.end(), sl.m) == gms.end()) gms.push_back(sl.m);
  for (uint32_t m : gms) {
    o << "    {\n";
    bool need[2] = {false, false};
    for (const OjSlot& sl : P.sl)
      if (sl.kind == 2 && sl.m == m) need[sl.comp] = true;
    std::vector<int> all;
    for (int j = 0; j < nreg; ++j) {
      all.push_back(j);
      o << "      const VT<T> b" << j << " = lds[pb0 ^ "
        << swz_w(pdep_pos(j, s0.rpos, kR) ^ m, W) << "u];\n";
    }
    for (int j = 0; j < nreg; ++j) {
      o << "      const double BX" << j << " = (double)b" << j << ".x, BY" << j << " = (double)b" << j
        << ".y, AX" << j << " = (double)v" << j << ".x, AY" << j << " = (double)v" << j << ".y;\n";
      if (need[0])
        o << "      const double GR" << j << " = BX" << j << " * AX" << j << " + BY" << j << " * AY"
          << j << ";\n";
      if (need[1])
        o << "      const double GI" << j << " = BX" << j << " * AY" << j << " - BY" << j << " * AX"
          << j << ";\n";
    }
    for (int k = 0; k < nslot; ++k) {
      const OjSlot& sl = P.sl[k];
      if (sl.kind != 2 || sl.m != m) continue;
      o << "     {\n";
      emit_signed_sum(o, sl.comp ? "GI" : "GR", all, sl.zr);
      emit_accumulate(o, k, sl.zb);
      o << "     }\n";
    }
    o << "    }\n";
  }
  for (int s = 1; s < ns; ++s) {
    const OjStage& st = P.st[s];
    o << "    {\n";
    for (int j = 0; j < nreg; ++j)
      o << "      const VT<T> w" << j << " = lds[pb" << s << " ^ " << swz_w(pdep_pos(j, st.rpos, kR), W)
        << "u];\n";
    for (int j = 0; j < nreg; ++j)
      o << "      const double X" << j << " = (double)w" << j << ".x, Y" << j << " = (double)w" << j
        << ".y;\n";
    emit_stage_slots(s);
    o << "    }\n";
  }
  o << "    __syncthreads();\n  }\n";
  // the thread-index signs, one wave sum per term, the waves in order
  for (int t = 0; t < nterm; ++t) {
    const OjTerm& tm = P.tm[t];
    o << "  {\n    double v = ";
    if (tm.lzt)
      o << "(__popc(lt" << tm.stage << " & " << tm.lzt << "u) & 1) ? -a" << tm.slot << " : a"
        << tm.slot;
    else
      o << "a" << tm.slot;
    o << ";\n    v = wave_total(v) * " << tm.scale << ".0;\n"
      << "    if (lane == 0) red[wid][" << t << "] = v;\n  }\n";
  }
  o << "  __syncthreads();\n"
    << "  if (tid < " << nterm << "u) {\n"
    << "    double v = red[0][tid];\n";
  for (int w = 1; w < kOjThreads / 64; ++w) o << "    v += red[" << w << "][tid];\n";
  o << "    partial[((unsigned long long)blockIdx.y * gridDim.x + blockIdx.x) * part_terms + term0 + "
       "tid] = v;\n  }\n}\n";
}

}  // namespace

// Tile bit sets as the precompiled reduction plans them (bits 0..3 and up
// to 8 more, masks ordered by highest bit), then each set's terms split
// into passes of at most kOjMaxSlots accumulators; diagonal terms go where
// their z bits are tile bits if there is room.
void plan_obs_jit(int n, int dtype, int n_terms, const uint64_t* xm, const uint64_t* zm,
                  ObsJitPlan& out) {
  out = ObsJitPlan{};
  if (n < kTB) return;
  const int W = dtype == SMOQ_C128 ? 4 : 5;
  const uint64_t lowset = (1ull << kLowBits) - 1ull;
  std::vector<uint64_t> xs;
  std::map<uint64_t, std::vector<int>> by_x;
  for (int t = 0; t < n_terms; ++t) {
    if (!by_x.count(xm[t])) xs.push_back(xm[t]);
    by_x[xm[t]].push_back(t);
  }
  std::vector<uint64_t> pair_xs;
  for (uint64_t x : xs)
    if (x) pair_xs.push_back(x);
  std::stable_sort(pair_xs.begin(), pair_xs.end(), [](uint64_t a, uint64_t b) {
    const int ha = 63 - __builtin_clzll(a), hb = 63 - __builtin_clzll(b);
    if (ha != hb) return ha < hb;
    return __builtin_ctzll(a) < __builtin_ctzll(b);
  });
  struct Draft {
    uint64_t hi = 0;
    std::vector<uint64_t> xs;
    std::vector<int> diag;
    int tpos[kTB];
    uint64_t tmask = 0;
  };
  std::vector<Draft> drafts;
  for (uint64_t x : pair_xs) {
    const uint64_t bits = x & ~lowset;
    if (__builtin_popcountll(bits) > kTB - kLowBits) {
      out.rest.insert(out.rest.end(), by_x[x].begin(), by_x[x].end());
      continue;
    }
    if (!drafts.empty() && __builtin_popcountll(drafts.back().hi | bits) <= kTB - kLowBits) {
      drafts.back().hi |= bits;
      drafts.back().xs.push_back(x);
    } else {
      Draft d;
      d.hi = bits;
      d.xs.push_back(x);
      drafts.push_back(d);
    }
  }
  for (Draft& d : drafts) {
    uint64_t hi = d.hi;
    for (int b = kLowBits; b < n && __builtin_popcountll(hi) < kTB - kLowBits; ++b) hi |= 1ull << b;
    int j = 0;
    for (int b = 0; b < n; ++b)
      if (b < kLowBits || ((hi >> b) & 1ull)) d.tpos[j++] = b;
    d.tmask = lowset | hi;
  }
  auto terms_of = [&](const std::vector<uint64_t>& dxs, const std::vector<int>& dg) {
    std::vector<TermIn> v;
    for (uint64_t x : dxs)
      for (int t : by_x[x]) v.push_back({x, zm[t], t});
    for (int t : dg) v.push_back({0, zm[t], t});
    return v;
  };
  const int max_slots = dtype == SMOQ_C128 ? kOjMaxSlots : kOjMaxSlotsC64;
  // x-mask groups of each tile bit set one after another while the slots fit
  struct Chunk {
    int draft;
    std::vector<uint64_t> xs;
    std::vector<int> diag;
  };
  std::vector<Chunk> chunks;
  OjPass trial;
  for (size_t di = 0; di < drafts.size(); ++di) {
    bool open = false;
    for (uint64_t x : drafts[di].xs) {
      if (open) {
        Chunk c = chunks.back();
        c.xs.push_back(x);
        build_pass(W, drafts[di].tpos, terms_of(c.xs, c.diag), trial);
        if ((int)trial.sl.size() <= max_slots) {
          chunks.back() = c;
          continue;
        }
      }
      chunks.push_back({(int)di, {x}, {}});
      open = true;
    }
  }
  // diagonal terms into a chunk with room, first those whose tile holds
  // their z bits (they share accumulators there), else a chunk of their own
  if (by_x.count(0)) {
    for (int t : by_x[0]) {
      bool placed = false;
      for (int pref = 1; pref >= 0 && !placed; --pref)
        for (Chunk& c : chunks) {
          const bool inside = (zm[t] & ~drafts[c.draft].tmask) == 0;
          if (inside != (pref == 1)) continue;
          std::vector<int> dg = c.diag;
          dg.push_back(t);
          build_pass(W, drafts[c.draft].tpos, terms_of(c.xs, dg), trial);
          if ((int)trial.sl.size() <= max_slots) {
            c.diag.swap(dg);
            placed = true;
            break;
          }
        }
      if (placed) continue;
      if (drafts.empty()) {
        Draft d;
        for (int b = 0; b < kTB; ++b) d.tpos[b] = b;
        d.tmask = (1ull << kTB) - 1ull;
        drafts.push_back(d);
      }
      int di = (int)drafts.size() - 1;
      for (size_t i = 0; i < drafts.size(); ++i)
        if ((zm[t] & ~drafts[i].tmask) == 0) {
          di = (int)i;
          break;
        }
      chunks.push_back({di, {}, {t}});
    }
  }
  for (Chunk& c : chunks) {
    OjPass P;
    build_pass(W, drafts[c.draft].tpos, terms_of(c.xs, c.diag), P);
    P.term0 = (int)out.cols.size();
    for (const OjTerm& tm : P.tm) out.cols.push_back(tm.col);
    out.passes.push_back(P);
  }
  std::ostringstream o;
  o << "#include \"fused_ops.h\"\nusing namespace smoq;\ntypedef "
    << (dtype == SMOQ_C128 ? "double" : "float") << " T;\n";
  for (size_t p = 0; p < out.passes.size(); ++p) emit_pass(o, (int)p, out.passes[p], n, dtype);
  out.src = o.str();
}

}  // namespace smoq
