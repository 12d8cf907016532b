// Pieces of the adjoint gradient shared by the single-state sweep
// (adjoint.hip) and the persistent batch over parameter sets of different
// qubit counts (vqe_batch.hip).
#pragma once

#include <vector>

#include "device_common.h"
#include "smoq_internal.h"

namespace smoq {

// R(θ) = exp(-iθG/2), G = Π_C ⊗ i^{cnt} X^x Z^z on the controls C = ctrl
struct PauliRot {
  uint64_t x, z, ctrl;
  int cnt;  // |x & z| mod 4
  int h;    // highest bit of x (x != 0)
};

// A Pauli term of H with its phase folded in: c * i^{|x&z|}.
struct AdjTerm {
  uint64_t x, z;
  double fre, fim;
};

// one step of a sweep: a fixed gate (and its inverse) or a Pauli rotation
struct Step {
  bool param = false;
  int pidx = -1;
  PauliRot rot{};
  smoq_op op{};  // the tape op of a parametrised rotation
  IGate fwd, inv;
};

int build_steps(const smoq_op* ops, int n_ops, int n_params, int n, std::vector<Step>& steps);
// g with every qubit moved up by one: W's qubit 0 is the ψ/λ selector
IGate shifted_up(const IGate& g);
// H's terms with their phases folded in; checks that they act inside n qubits
int adj_terms(int n_terms, const uint64_t* xm, const uint64_t* zm, const double* coeff, int n,
              std::vector<AdjTerm>& out);
// The forward gate list of the tape at angles theta.
int forward_gates(const std::vector<Step>& steps, const double* theta, int n,
                  std::vector<IGate>& out);
// The reverse sweep over W (n+1 qubits, ψ and λ interleaved): walking the
// steps backwards, a parametrised rotation becomes an overlap gate (slot
// numbers run backwards from the last rotation) followed by R(-θ), any other
// gate its inverse.  *n_slots = number of parametrised steps.
int reverse_gates(const std::vector<Step>& steps, const double* theta, int n,
                  std::vector<IGate>& out, int* n_slots);
// grad[p] = the sum of parameter p's overlap slots in gate order (slot sums
// indexed as reverse_gates numbers them)
void grad_from_slots(const std::vector<Step>& steps, const double* slot_sums, int n_slots,
                     int n_params, double* grad);

}  // namespace smoq
