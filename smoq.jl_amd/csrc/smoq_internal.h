// Internal declarations shared by the HIP translation units of libsmoq_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "smoq.h"
#include "fused_types.h"

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
  // Im<λ|G|ψ> into overlap slot oslot, ψ and λ the halves of qubit sel, G the
  // Pauli string i^ocnt X^x Z^oz with x on t0 (and t1 when >= 0), applied
  // where (index & ctrl) == ctrl; the amplitudes stay unchanged
  IK_OVL = 3,
};

struct IGate {
  int32_t kind = IK_U1;
  int32_t t0 = -1, t1 = -1;
  int32_t nd = 0;
  int32_t dq[3] = {-1, -1, -1};
  uint64_t ctrl = 0;
  // IK_OVL only
  int32_t sel = -1, oslot = -1, ocnt = 0;
  uint64_t oz = 0;
  double m[32] = {0};
};

// Lowers one public op (angle already resolved) to internal gates.
int lower_op(const smoq_op& op, double theta, int n_qubits, std::vector<IGate>& out);



struct Program {
  std::vector<PassDev> passes;
  std::vector<StageDev> stages;
  std::vector<GateDev> gates;
  std::vector<double> tabs;   // interleaved complex phase tables
  // compiled passes read their gate records and tables at an offset per
  // batch item (launched with item_bytes > 0)
  bool per_item = false;
};

// Builds the pass/stage program for a gate list on an n-qubit state.
// row_bits: log2 of the amplitudes per LDS swizzle row of the state's dtype.
// max_gates: at most that many non-diagonal gates per pass (0: no cap).
// lane_swaps: stages may trade register qubits with lane bits 4 and 5 in
// registers instead of an LDS round trip (run-time compiled passes only).
// locals: the local qubit set of each pass, chosen beforehand
// (search_local_sets); passes beyond it, or after a set that takes no gate,
// grow their sets greedily.
void plan_program(const std::vector<IGate>& gates, int n_qubits, int tile_qubits,
                  int low_qubits, int row_bits, Program& prog, int max_gates = 0,
                  bool lane_swaps = false, const std::vector<uint64_t>* locals = nullptr);
// Local qubit sets of the passes by search: a beam of plan prefixes, each
// extended by sets found from the greedy set with `iters` random exchanges
// of one non-low local qubit with another qubit (kept when the pass then
// takes at least as many non-diagonal gates); the prefixes leaving the
// fewest gates go on.  The greedy rule fixes the first qubits it meets,
// which can cost whole passes later.  Deterministic for a seed.
void search_local_sets(const std::vector<IGate>& gates, int n_qubits, int tile_qubits,
                       int low_qubits, int max_gates, int iters, uint64_t seed,
                       std::vector<uint64_t>& locals);
// whether a program has stages that continue through lane exchanges
bool program_has_swaps(const Program& prog);
// a pass's HBM order (io & 63: the tile order of PassDev::tord; io & 64:
// registers on the three positions above the `low` contiguous qubits
// instead of the tile's top three when moving it between HBM and LDS)
void set_pass_io(PassDev& pd, int io, int low);
// set_pass_io of every pass when io has one entry per pass
void apply_pass_io(Program& prog, const std::vector<uint8_t>& io, int low);

// Fusion settings of a gate list: tile qubits, contiguous low qubits and
// the per-pass gate cap.
struct Fusion {
  int tile = 0, low = 0, cap = 0;
  int swaps = 0;  // lane exchanges in compiled passes
  std::vector<uint8_t> io;  // per-pass HBM order (set_pass_io), empty: default
  std::vector<uint64_t> locals;  // per-pass local qubit sets, empty: greedy
};
// 64-bit hash of a lowered gate list's structure (kinds and qubits, not
// the matrices)
uint64_t gate_structure_hash(const std::vector<IGate>& gs);
// LDS index swizzle (the same function the kernels use)
inline uint32_t swz_host(uint32_t l, int wb) {
  uint32_t f = (l >> wb) ^ (l >> (2 * wb)) ^ (l >> (3 * wb)) ^ (l >> (4 * wb));
  return l ^ (f & ((1u << wb) - 1u));
}

}  // namespace smoq

namespace smoq {
// pass kernels compiled at run time for one circuit structure (jit.hip)
struct JitModule {
  hipModule_t mod = nullptr;
  std::vector<hipFunction_t> fn;  // one per pass
  ~JitModule();
};
// automatic mode compiles for states of at least this many qubits and
// programs of at least this many stage gates
constexpr size_t kStatePoolMaxBytes = 64ull << 20;
constexpr size_t kStatePoolMax = 8;
constexpr int kJitMinQubits = 16;
constexpr int kJitMinGates = 16;

// Compiled tiled Pauli reductions (obs_jit.cpp): tiles of 2^kOjTileBits
// amplitudes, kOjThreads threads with 2^kOjRegBits amplitudes each, at
// most kOjMaxSlots FP64 accumulators per thread and pass.
constexpr int kOjTileBits = 12;
constexpr int kOjRegBits = 4;
constexpr int kOjThreads = 1 << (kOjTileBits - kOjRegBits);
constexpr int kOjMaxSlots = 48;
constexpr int kOjMaxSlotsC64 = 40;  // complex64 converts in registers
// automatic mode: at least 2^kOjMinLogAmps amplitudes over the batch
constexpr int kOjMinLogAmps = 22;
constexpr size_t kObsJitPlansMax = 64;
struct OjStage {
  uint32_t rmask;        // register tile bits
  int rpos[kOjRegBits];  // the same, ascending
  int tbit[kOjTileBits - kOjRegBits];  // tile bit of each thread-index bit
};
struct OjSlot {
  int stage;
  int kind;  // 0 diagonal, 1 register pairs, 2 partner from LDS
  uint32_t m;  // kind 1: x mask in register bits; kind 2: in tile bits
  int comp;    // 0 real part, 1 imaginary part
  uint32_t zr;  // z bits among the register bits (sign per register index)
  uint64_t zb;  // z bits outside the tile (sign per tile)
};
struct OjTerm {
  int col;    // the caller's term index
  int slot;
  int stage;
  uint32_t lzt;  // z bits among the stage's thread bits (sign per thread)
  int scale;     // +-1, +-2
};
struct OjPass {
  int tpos[kOjTileBits];  // global bit of each tile bit
  std::vector<OjStage> st;
  std::vector<OjSlot> sl;
  std::vector<OjTerm> tm;
  int term0 = 0;
};
struct ObsJitPlan {
  std::vector<OjPass> passes;
  std::vector<int> cols;  // pass-ordered term -> caller's term index
  std::vector<int> rest;  // terms left to the per-mask kernels
  std::string src;
  std::shared_ptr<JitModule> mod;
};
void plan_obs_jit(int n, int dtype, int n_terms, const uint64_t* xm, const uint64_t* zm,
                  ObsJitPlan& out);
}  // namespace smoq

// ------------------------------------------------------------ handles --
struct smoq_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int tile_qubits = 0;  // 0: chosen per gate list (pass_tile)
  int low_qubits = 4;
  // run-time compiled pass kernels: 0 off, 1 automatic, 2 always
  int jit_mode = 1;
  std::unordered_map<std::string, std::shared_ptr<smoq::JitModule>> jit_cache;
  std::string jit_key_buf;  // the last key (capacity reused across calls)
  // multi-mask Pauli term lists through tiled passes (SMOQ_OBS_TILES=0: off)
  int obs_tiles = 1;
  // ... compiled for the term list: 0 off, 1 automatic (large states), 2
  // always (states of at least 12 qubits); plans and their modules cached
  // by term list, at most kObsJitPlansMax (then the table starts over)
  int obs_jit = 1;
  std::unordered_map<std::string, std::shared_ptr<smoq::ObsJitPlan>> obs_jit_plans;
  // rocBLAS handle for dense density-matrix propagation, created on first use
  void* blas = nullptr;
  // grow-only device scratch, reused across calls on this context
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* prog = nullptr;
  size_t prog_bytes = 0;
  void* host_pinned = nullptr;
  size_t host_pinned_bytes = 0;
  // second stream (default priority) and pinned result buffer of the batched
  // gradients, whose one-launch part runs beside whole-device launches
  hipStream_t aux_stream = nullptr;
  void* res_pinned = nullptr;
  size_t res_pinned_bytes = 0;
  // fused-pass timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;   // begin/end pairs recorded on the stream
  size_t ev_used = 0;
  double timed_ms = 0.0;
  int64_t timed_launches = 0;
  // pairwise-cumsum tree of long-vector sampling: level arrays and leaf-end
  // buffers, grow-only; the node geometry in it is valid for tree_n
  // elements (0: none)
  uint64_t tree_n = 0;
  void* tree_dev = nullptr;
  size_t tree_dev_bytes = 0;
  // host buffers of tape applications (lowered gates, planned program),
  // kept so that their capacity is reused: fresh ones of a 1000-gate tape
  // (~0.6 MB) cost page faults on every call
  std::vector<smoq::IGate> tape_gates;
  smoq::Program tape_prog;
  // freed state buffers of at most kStatePoolMaxBytes (size, pointer), kept
  // for reuse: hipFree synchronises and short calls would pay for it
  std::vector<std::pair<size_t, void*>> state_pool;
  // grow-only device buffers of the gradient calls (angle tables, λ, the
  // reverse sweep's copy of ψ, partial sums), reused across calls
  static constexpr int kWorkSlots = 6;
  void* work[kWorkSlots] = {};
  size_t work_bytes[kWorkSlots] = {};
  // per-pass cap on non-diagonal gates (0: none; SMOQ_PASS_MAX_GATES)
  int pass_max_gates = 0;
  // compiled passes may exchange register and lane qubits instead of LDS
  // round trips (SMOQ_LANE_SWAPS=0: off)
  int lane_swaps = 1;
  // autotuned fusion settings by (qubits, dtype, batch, structure hash)
  std::unordered_map<std::string, smoq::Fusion> tune;
  // states and plans allocated on this context and not yet freed:
  // smoq_fini refuses to release the context while any remain
  int64_t live_objects = 0;
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

// Blocks of a grid-stride launch over n items with `threads` per block: one
// item per thread up to 2^20 blocks (far more than the chip holds at once),
// so the launch stays valid however large batch x 2^n grows.
inline unsigned stride_blocks(uint64_t n, unsigned threads) {
  const uint64_t want = (n + threads - 1) / threads;
  return (unsigned)(want == 0 ? 1 : (want < (1ull << 20) ? want : (1ull << 20)));
}
int ensure_prog(smoq_ctx* ctx, size_t bytes);
// grow-only work buffer `slot` (0 .. kWorkSlots-1) of at least `bytes`
int ensure_work(smoq_ctx* ctx, int slot, size_t bytes, void** out);
// hipFree of every pooled state buffer (after the stream drains)
void release_state_pool(smoq_ctx* ctx);
int ensure_pinned(smoq_ctx* ctx, size_t bytes);
int check_state(const smoq_state* s);
int check_qubit(const smoq_state* s, int q, const char* what);
// records the next pooled event on the context stream (timing mode only)
int time_mark(smoq_ctx* ctx);

// launchers implemented in the kernel translation units
int launch_gate(smoq_state* s, const IGate& g);
// fixed_angles: the gates' angles do not change between calls with this
// structure, so run-time compiled passes may factor the rotations (in place:
// prog may come back factored)
int run_program(smoq_state* s, Program& prog, bool fixed_angles = false);

// device image of a Program: passes | stages | gates | phase tables
struct ProgLayout {
  size_t pb = 0, sb = 0, gb = 0, ntab = 0, off_t = 0, total = 0;
};
ProgLayout program_layout(const Program& prog, int dtype);
int upload_program(smoq_ctx* ctx, const Program& prog, int dtype, const ProgLayout& L, void* d);
// jit: the program's compiled pass kernels, or null for the precompiled one
// opart: per-tile partials of the program's overlap gates (GK_OVL), at
// opart[oslot * tiles + tile]; compiled passes only
// only_pass >= 0: that pass alone
// item_bytes > 0 (compiled passes only): every batch item has its own
// uploaded copy of the program, item_bytes apart, whose gate records and
// tables its tiles read (same structure, other angles)
int launch_program(smoq_state* s, const Program& prog, const ProgLayout& L, const void* dbuf,
                   const JitModule* jit, double* opart = nullptr, int only_pass = -1,
                   uint64_t item_bytes = 0);
// a key of the compiled form of a program (equal keys: same kernels)
void jit_structure_key(const Program& prog, int dtype, std::string& buf);
// run-time compiled pass kernels (jit.hip)
std::string jit_source(const Program& prog, int dtype);
// rewrites rotation gates into the factored kinds of run-time compiled
// passes; with undo, records what unfactor_rotations needs to restore them
struct FactorUndo {
  int g;
  int32_t kind, b;
  double m0;
  double md[16];  // a normalised register diagonal's phases (GK_DIAGREG1)
};
void factor_rotations(Program& prog, int dtype, std::vector<FactorUndo>* undo = nullptr);
void unfactor_rotations(Program& prog, const std::vector<FactorUndo>& undo);
bool jit_wanted(const smoq_ctx* ctx, const Program& prog, int n_qubits);
// whether a gate list of this size will run as compiled passes (decided
// before planning, so that the planner may use lane exchanges)
bool jit_planned(const smoq_ctx* ctx, int n_qubits, size_t n_gates);
// tile qubits of the fused passes for a gate list (automatic unless set)
int pass_tile(const smoq_ctx* ctx, int n_qubits, int batch, size_t n_gates);
// the fusion settings for gs: an autotuned entry of the context when there
// is one, else the context's settings
Fusion fusion_for(const smoq_ctx* ctx, int n_qubits, int dtype, int batch,
                  const std::vector<IGate>& gs);
int jit_get(smoq_ctx* ctx, const Program& prog, int dtype, std::shared_ptr<JitModule>& out);
// a module from generated source with kernels <prefix>0 .. <prefix>(n_fn-1)
// (compiled, or loaded from the on-disk cache); the caller keeps it
int jit_get_source(smoq_ctx* ctx, const std::string& src, size_t n_fn, const char* prefix,
                   std::shared_ptr<JitModule>& out);
// the compiled modules of several programs, the missing ones compiled on
// parallel host threads
int jit_get_many(smoq_ctx* ctx, const std::vector<const Program*>& progs, int dtype,
                 std::vector<std::shared_ptr<JitModule>>& out);
// lowers a tape (angles resolved from theta) to internal gates
int lower_tape(const smoq_op* ops, int n_ops, const double* theta, int n_theta, int n_qubits,
               std::vector<IGate>& out);
int reduce_pauli(smoq_state* s, int n_terms, const uint64_t* xm, const uint64_t* zm,
                 double* out_host, int ref_mode);
// takes the device address of the fixed partial-sum region of
// smoq_expect_pauli_noalloc on `device` (current); once per device
int noalloc_scratch_init(int device);
// the dissipative half of an MCWF step for every batch item (mcwf.hip):
// uniforms of item b at d_uni[b * uni_stride + 0/1], its jump at
// d_jumps[b * jump_stride]; a zero or non-finite norm records
// min(item * n_steps + step) in *d_fault
int mcwf_jump_check(const smoq_state* s, int kind, double gamma_dt);
// d_partial: mcwf_jump_partial_bytes(s) of device memory for the chunk sums
int mcwf_jump_launch(smoq_state* s, int kind, double gamma_dt, const double* d_decay,
                     const double* d_uni, int uni_stride, int32_t* d_jumps, int jump_stride,
                     unsigned long long* d_fault, int n_steps, int step, double* d_partial);
size_t mcwf_jump_partial_bytes(const smoq_state* s);
// releases the context's rocBLAS handle (density.hip)
void blas_release(smoq_ctx* ctx);
// the context's rocBLAS handle (created on first use) bound to its stream
int blas_handle(smoq_ctx* ctx, const char* who, void** out);
}  // namespace smoq
