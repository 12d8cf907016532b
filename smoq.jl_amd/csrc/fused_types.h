// Types of the fused-pass programs: the device image the pass kernels read.
// Included by the host planner, the precompiled kernels and the pass kernels
// compiled at run time (which see this file as embedded text), so it holds
// plain data definitions only.
#pragma once

#ifndef __HIPCC_RTC__
#include <cstdint>
#endif

namespace smoq {

// ------------------------------------------------------- fused programs --
// 8 amplitudes per thread: at 16 per thread the complex128 stage kernel needs
// ~255 VGPRs (one wave per SIMD) and LDS latency between stages goes
// unhidden; at 8 it fits in ~112 VGPRs (four waves per SIMD).
#ifndef SMOQ_REG_QUBITS
#define SMOQ_REG_QUBITS 3
#endif
constexpr int kRegQubits = SMOQ_REG_QUBITS;  // amplitudes per thread = 2^kRegQubits
constexpr int kMaxTileQubits = 12;     // LDS tile of up to 4096 amplitudes
// a tile of 2^k amplitudes runs 2^(k - kRegQubits) threads
constexpr int kMaxPassThreads = 1 << (kMaxTileQubits - kRegQubits);
constexpr int kMaxQubits = 40;
// Passes with at most 2^kWarmMaxNonlocal tiles per item warm their stage and
// gate records into L2 before the stages read them (small states: 1000-gate
// circuits at 12-16 qubits run 25-30% faster), when that takes at most
// kWarmLoads loads per thread (a state that is one tile runs all its gates
// in one pass of 2..64 threads).
constexpr int kWarmMaxNonlocal = 10;
constexpr uint32_t kWarmLoads = 4;

// Gate kinds inside a register stage.
enum GKind : int32_t {
  GK_U1_REAL = 0,  // 2x2 with real entries (Ry, H, X, sigma+-)
  GK_U1_RXT = 1,   // real diagonal, imaginary off-diagonal (Rx, Y)
  GK_U1_GEN = 2,   // general 2x2
  GK_U2 = 3,       // general 4x4
  GK_DIAGTAB = 4,  // merged diagonal gates: phase table[row][slot pattern]
  GK_CZSET = 5,    // merged CZs: per-lane 16-bit negate mask
  GK_DIAGREG = 6,  // diagonal over register qubits only: m = 16 phases
  GK_DIAGLANE = 7, // diagonal over non-register qubits only: m = 2^nu phases
  // Rotations with the real factor taken out (run-time compiled passes of
  // circuits with fixed angles): the gate is sigma * N with N's unit
  // entries on the diagonal (b = 0) or the anti-diagonal (b = 1), |t| <= 1
  // in m[0]; the pass multiplies its amplitudes once by the product of the
  // sigmas.
  GK_U1_ROT = 8,   // [[a, -b], [b, a]] real (Ry and products of Ry)
  GK_U1_RXR = 9,   // [[a, i q], [i q, a]] (Rx and products of Rx)
  // Adjoint-gradient overlap (run-time compiled passes only): the state
  // holds ψ and λ as the two halves of qubit `osel`; the gate leaves the
  // amplitudes alone and adds Im<λ|G|ψ> of its tile, for the Pauli string
  // G = i^ocnt X^x Z^z (x on slots a, b; controls as for the other kinds),
  // to overlap slot `oslot`.
  GK_OVL = 10,
  // GK_DIAGREG with pattern 0's phase taken out into the pass's complex
  // scale (run-time compiled passes of circuits with fixed angles): only the
  // register patterns in the mask b are multiplied (a one-qubit phase on
  // slot t touches the four patterns with bit t set)
  GK_DIAGREG1 = 11,
};
constexpr int kMaxTabBits = 3;   // non-register qubits of one phase table
constexpr int kMaxCzShifts = 4;

// A gate as one register stage sees it.  Control and diagonal qubits are
// split by the planner into register slots (known per stage) and global bit
// positions read from the thread's base index.
struct GateDev {
  int32_t kind;
  int32_t a, b;          // register-qubit slots (U1: a; U2: a low, b high)
  int32_t creg;          // controls held in registers (slot mask)
  uint64_t cglob;        // controls outside the registers (global mask)
  int32_t nu;            // DIAGTAB: number of non-register qubits
  int32_t ubits[kMaxTabBits];
  int32_t tab;           // DIAGTAB: offset (entries) of the 2^nu x 16 table
  // CZSET: a CZ between two non-register qubits adds popc(G & (G >> s) & M)
  // terms evaluated on the lane's base index; a CZ between register slots
  // flips the slot patterns in czreg; a CZ between slot t and a non-register
  // qubit q puts q in czcm[t] (the lane flips pattern j when parity(gl &
  // czcm[t]) and bit t of j are both 1).
  int32_t ncz;
  int32_t czreg;
  int32_t czs[kMaxCzShifts];
  uint64_t czm[kMaxCzShifts];
  uint64_t czcm[kRegQubits];
  // GK_OVL: selector slot, overlap slot, |x & z| mod 4, the z bits held in
  // register slots (slot mask) and outside them (global mask)
  int32_t osel, oslot, ocnt, ozreg;
  uint64_t ozglob;
  // GK_U2 in run-time compiled passes of fixed-angle circuits: bit 4r + c
  // set when entry (r, c) is exactly zero (its product is left out)
  uint32_t zero;
  int32_t pad2;
  double m[32];
};

// Per-stage addressing, precomputed on the host so that a stage costs the
// device one descriptor load: thread-index bit i selects local position
// p_i (the positions not held in registers, ascending).
struct StageDev {
  int32_t q[kRegQubits];             // local bit positions held in registers
  int32_t g0, g1;                    // gate range [g0, g1) in the gate array
  int32_t ntb;                       // thread-index bits (k - kRegQubits)
  int32_t pad;
  uint32_t sw_tid[kMaxTileQubits - kRegQubits];  // swizzled LDS index of bit p_i
  uint32_t sw_reg[kRegQubits];       // swizzled LDS index of slot bit q_j
  uint64_t g_tid[kMaxTileQubits - kRegQubits];   // global index bit of p_i
  // local position p_i of thread-index bit i
  int8_t tpos[kMaxTileQubits - kRegQubits];
  // Lane exchanges (run-time compiled passes): cont = 1 continues the
  // previous stage without an LDS round trip, after register slot
  // swap_slot has traded its qubit with lane bit swap_bit (4 or 5).
  int8_t cont, swap_slot, swap_bit;
};

struct PassDev {
  int32_t k;              // tile qubits
  int32_t n_nonlocal;
  int32_t s0, s1;         // stage range
  uint64_t nl_mask;       // global bits outside the tile (tile id deposits here)
  uint64_t tid_mask;      // global bits of local positions 0..k-R-1 (thread index)
  uint64_t hb[kRegQubits];  // global bits of local positions k-R..k-1
  int32_t map[kMaxTileQubits];       // local bit -> global bit
  // product of the factored-out rotation scalars and diagonal phases
  double scale, scale_im;
  // HBM order of the tiles and of each tile's amplitudes (they move
  // nothing but the time a pass takes): the launch's tile index is rotated
  // left by tord & 31 bits, after a bit reversal when tord & 32, inside its
  // n_nonlocal bits; ltm marks the local positions that the thread index
  // covers when the tile moves between HBM and LDS (tid_mask holds their
  // global bits, hb the other three positions' bits)
  int32_t tord;
  uint32_t ltm;
};

}  // namespace smoq
