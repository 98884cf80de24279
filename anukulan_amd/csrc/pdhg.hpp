// The PDHG solve on one device: restarted reflected-Halpern PDHG with the
// cuPDLPx additions, feasibility polishing and a Farkas infeasibility test,
// making the same decisions as solve_pdhg (src/anukulan/pdhg.cpp:377-1285).
//
// What differs from the reference is where the work happens. Between two
// termination checks the loop is a fixed sequence of fused kernels (K^T y with
// the primal step in its epilogue, K x_bar with the dual step in its
// epilogue) whose per-step scalars come from device memory, so the period can
// be replayed as one HIP graph. A check is four more kernels and one 112-byte
// copy: the movement, the Farkas test and all seven convergence terms are
// reduced on the device. The host keeps the control decisions (restart, PID
// primal weight, polish schedule) and computes them from the same numbers.
#pragma once

#include <string>
#include <vector>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"

namespace anuk {

struct Phases {
  double scaling = 0.0, matrix_norm = 0.0, setup = 0.0, iterating = 0.0, checks = 0.0,
         polishing = 0.0;
};

struct SolveResult {
  int status = ANUK_PDHG_NUMERICAL_ERROR;
  std::vector<double> x, y;
  double objective = 0.0;
  Int iterations = 0, restarts = 0, polish_attempts = 0, polish_iterations = 0;
  bool polished = false;
  double solve_seconds = 0.0, matrix_norm_estimate = 0.0, final_step_size = 0.0,
         final_primal_weight = 0.0;
  Phases phases;
  Residual residual;
  std::string message;
  // Host-side bookkeeping for the caller: how many device periods ran as
  // graphs and how many kernels were launched directly.
  long graph_launches = 0;
  long direct_periods = 0;
};

SolveResult solve_pdhg(HipBackend& be, const StandardLp& lp, const anuk_pdhg_options& options);

// Power iteration on K^T K (pdhg.cpp:223-263) on the device.
double estimate_matrix_norm(HipBackend& be, const StandardLp& lp, int iterations,
                            double tolerance);

void default_options(anuk_pdhg_options* o);

}  // namespace anuk
