// Host-side control of one LP's restarted Halpern PDHG loop: the step size,
// the primal weight and its PID state, the Halpern epoch, the restart test,
// the polish schedule and the termination verdicts. The plain solver keeps
// one of these; the batched solver keeps one per case, so every case takes
// exactly the decisions a plain solve of it would.
//
// The decision sequence at a termination check follows pdhg.cpp:926-1202:
// Farkas certificate, non-finite residual, convergence, feasibility polish,
// time limit, cancellation, then the fixed-point restart test. It is split
// in two around the polish so a caller can run the polishing sub-solve (or a
// batch of them) in between.
#pragma once

#include <string>

#include "../../include/anukulan_hip.h"
#include "kernels.hpp"
#include "lp.hpp"

namespace anuk {

struct Control {
  const anuk_pdhg_options* opt = nullptr;
  double eta = 0.0;
  double omega = 1.0;
  bool halpern = false;
  double reflection = 0.0;
  double tol = 0.0;
  double gap_tol = 0.0;

  bool anchor_needs_reset = true;
  Int halpern_k = 0;
  double movement = 0.0;
  Int since_restart = 0;
  double residual_at_restart = kInf;
  double last_residual = kInf;
  double last_candidate_kkt = kInf;
  double kkt_at_restart = 0.0;
  double werr_sum = 0.0;
  double werr_prev = 0.0;
  bool have_prev_restart = false;
  Int next_polish = 0;
  int status = ANUK_PDHG_ITERATION_LIMIT;
  std::string message;
  Int restarts = 0;

  // Step size, primal weight and schedule from the scaled LP (pdhg.cpp:462-477).
  void init(const anuk_pdhg_options& o, double matrix_norm, double c_norm, double q_norm);

  // The scalars a period of fused steps reads, as of the period's start.
  IterParams period_params(Int num_equalities) const;
  // Advances the Halpern counter and epoch length over `steps` steps.
  void advance(int steps);

  bool passes(const Residual& r) const;
  Int polish_budget(Int at_iteration) const;

  enum class Verdict { kContinue, kStop, kPolish, kRestart };

  // First half of a check at iteration index i: Farkas, non-finite,
  // convergence and the polish trigger. kPolish carries the budget in *budget.
  Verdict check_first(Int i, const Residual& r, bool farkas, Int* budget);
  // Second half (after a polish attempt failed or none was due): logging,
  // time limit, cancellation and the restart test.
  Verdict check_second(Int i, const Residual& r, double candidate_kkt, double elapsed);

  // Primal weight update at a restart from how far each side moved since the
  // previous restart (pdhg.cpp:1135-1181), then the epoch reset (1183-1202).
  void restart(double delta_x, double delta_y);
};

}  // namespace anuk
