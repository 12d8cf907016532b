// Internal declarations shared by the HIP translation units of libsmoq_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "smoq.h"

namespace smoq {

// ---------------------------------------------------------------- errors --
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);

#define SMOQ_HIP(call)                                  \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return smoq::hip_error(_e, #call); \
  } while (0)

#define SMOQ_CHECK_LAUNCH(name)                         \
  do {                                                  \
    hipError_t _e = hipGetLastError();                  \
    if (_e != hipSuccess) return smoq::hip_error(_e, name); \
  } while (0)

// ---------------------------------------------------- internal gate form --
// Every gate of the public API is lowered to one of these before it reaches
// a kernel.  Targets must be local to a fused tile; control and diagonal
// qubits may sit anywhere because they only select or phase amplitudes.
enum IKind : int32_t {
  IK_U1 = 0,    // 2x2 on t0, applied where (index & ctrl) == ctrl
  IK_U2 = 1,    // 4x4 on (t0 low role, t1 high role), same ctrl rule
  IK_DIAG = 2,  // phase table over nd <= 3 global qubits dq[]
};

struct IGate {
  int32_t kind = IK_U1;
  int32_t t0 = -1, t1 = -1;
  int32_t nd = 0;
  int32_t dq[3] = {-1, -1, -1};
  uint64_t ctrl = 0;
  double m[32] = {0};
};

// Lowers one public op (angle already resolved) to internal gates.
int lower_op(const smoq_op& op, double theta, int n_qubits, std::vector<IGate>& out);

// ------------------------------------------------------- fused programs --
// 8 amplitudes per thread: at 16 per thread the complex128 stage kernel needs
// ~255 VGPRs (one wave per SIMD) and LDS latency between stages goes
// unhidden; at 8 it fits in ~112 VGPRs (four waves per SIMD).
#ifndef SMOQ_REG_QUBITS
#define SMOQ_REG_QUBITS 3
#endif
constexpr int kRegQubits = SMOQ_REG_QUBITS;  // amplitudes per thread = 2^kRegQubits
constexpr int kMaxTileQubits = 12;     // LDS tile of up to 4096 amplitudes
constexpr int kMaxQubits = 40;

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
};

struct PassDev {
  int32_t k;              // tile qubits
  int32_t n_nonlocal;
  int32_t s0, s1;         // stage range
  uint64_t nl_mask;       // global bits outside the tile (tile id deposits here)
  uint64_t tid_mask;      // global bits of local positions 0..k-R-1 (thread index)
  uint64_t hb[kRegQubits];  // global bits of local positions k-R..k-1
  int32_t map[kMaxTileQubits];       // local bit -> global bit
};

struct Program {
  std::vector<PassDev> passes;
  std::vector<StageDev> stages;
  std::vector<GateDev> gates;
  std::vector<double> tabs;   // interleaved complex phase tables
};

// Builds the pass/stage program for a gate list on an n-qubit state.
// row_bits: log2 of the amplitudes per LDS swizzle row of the state's dtype.
void plan_program(const std::vector<IGate>& gates, int n_qubits, int tile_qubits,
                  int low_qubits, int row_bits, Program& prog);
// LDS index swizzle (the same function the kernels use)
inline uint32_t swz_host(uint32_t l, int wb) {
  uint32_t f = (l >> wb) ^ (l >> (2 * wb)) ^ (l >> (3 * wb)) ^ (l >> (4 * wb));
  return l ^ (f & ((1u << wb) - 1u));
}

}  // namespace smoq

// ------------------------------------------------------------ handles --
struct smoq_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int tile_qubits = 10;
  int low_qubits = 4;
  // grow-only device scratch, reused across calls on this context
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* prog = nullptr;
  size_t prog_bytes = 0;
  void* host_pinned = nullptr;
  size_t host_pinned_bytes = 0;
  // fused-pass timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;   // begin/end pairs recorded on the stream
  size_t ev_used = 0;
  double timed_ms = 0.0;
  int64_t timed_launches = 0;
  // cached pairwise-cumsum tree (sampling), keyed by vector length
  uint64_t tree_n = 0;
  void* tree_dev = nullptr;
  size_t tree_dev_bytes = 0;
  std::vector<int64_t> tree_level_off;  // node offsets per level
  int64_t tree_nodes = 0, tree_leaves = 0;
};

struct smoq_state_s {
  smoq_ctx* ctx = nullptr;
  int n = 0;
  int dtype = SMOQ_C128;
  int batch = 1;
  uint64_t dim = 0;        // 2^n
  void* d = nullptr;       // device amplitudes
  size_t bytes = 0;
  size_t amp_bytes() const { return dtype == SMOQ_C128 ? 16 : 8; }
};

namespace smoq {
int ensure_scratch(smoq_ctx* ctx, size_t bytes);
int ensure_prog(smoq_ctx* ctx, size_t bytes);
int ensure_pinned(smoq_ctx* ctx, size_t bytes);
int check_state(const smoq_state* s);
int check_qubit(const smoq_state* s, int q, const char* what);
// records the next pooled event on the context stream (timing mode only)
int time_mark(smoq_ctx* ctx);

// launchers implemented in the kernel translation units
int launch_gate(smoq_state* s, const IGate& g);
int run_program(smoq_state* s, const Program& prog);

// device image of a Program: passes | stages | gates | phase tables
struct ProgLayout {
  size_t pb = 0, sb = 0, gb = 0, ntab = 0, off_t = 0, total = 0;
};
ProgLayout program_layout(const Program& prog, int dtype);
int upload_program(smoq_ctx* ctx, const Program& prog, int dtype, const ProgLayout& L, void* d);
int launch_program(smoq_state* s, const Program& prog, const ProgLayout& L, const void* dbuf);
// lowers a tape (angles resolved from theta) to internal gates
int lower_tape(const smoq_op* ops, int n_ops, const double* theta, int n_theta, int n_qubits,
               std::vector<IGate>& out);
int reduce_pauli(smoq_state* s, int n_terms, const uint64_t* xm, const uint64_t* zm,
                 double* out_host, int ref_mode);
}  // namespace smoq
