// Exact finish: from a first-order point to an optimal basis.
//
// The first-order solve gets every case of a stack to moderate accuracy on
// the device; the cases whose answer decides something are then finished
// exactly on the host cores, one case per core. The reference joins the two
// methods the same way: crossover_basis reads a starting basis off the
// PDHG point (src/anukulan/crossover.cpp:35-229) and its simplex starts
// from it instead of from all logicals (app/anukulan_cli.cpp:592-631,
// branch_and_bound.cpp:441-469). Here both halves are this library's own:
// the crossover applies the reference's candidate and pivot rules, and the
// simplex is a bounded primal simplex (composite phase one, Harris ratio
// test, Bland's rule against stalling) over a sparse LU of the basis with
// product-form updates, run on the Ruiz-equilibrated LP.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lp.hpp"

namespace anuk {

// Where a column sits (VarStatus, simplex.hpp:40).
enum class VarState : std::int8_t { kBasic = 0, kLower = 1, kUpper = 2, kFree = 3 };

struct FinishOptions {
  double primal_tolerance = 1e-9;   // bound and row violation allowed, equilibrated units
  double dual_tolerance = 1e-9;     // reduced-cost sign violation allowed
  double pivot_tolerance = 1e-9;    // smallest |alpha| the ratio test divides by
  std::int32_t max_iterations = 200000;
  std::int32_t refactor_interval = 100;
  std::int32_t crossover = 1;       // start from the point's basis; 0: all logicals
  // crossover_basis's own knobs (crossover.hpp:37-83)
  double interior_tolerance = 1e-6;
  double crossover_pivot_tolerance = 1e-7;
  std::int32_t max_candidates = -1;
  std::int32_t crossover_refactor_interval = 50;
  double dual_weight = 1.0;
  std::int32_t ruiz_iterations = 10;  // equilibration sweeps before the simplex
  std::int32_t threads = 0;           // cases finished at once; 0: the host's cores, at most 16
};

enum FinishStatus {
  kFinishOptimal = 0,
  kFinishInfeasible = 1,
  kFinishUnbounded = 2,
  kFinishIterationLimit = 3,
  kFinishNumericalError = 4,
};

struct CrossoverInfo {
  Int candidates = 0;
  Int pushed = 0;
  Int rejected_small_pivot = 0;
  Int logicals_remaining = 0;
  bool stopped_on_refactorization = false;
};

struct FinishResult {
  int status = kFinishNumericalError;
  std::vector<double> x, y;        // unscaled primal (structurals) and row duals
  double objective = 0.0;          // model sense
  Int iterations = 0, phase_one_iterations = 0, refactorizations = 0, bland_iterations = 0;
  bool started_warm = false;
  CrossoverInfo crossover;
  std::vector<Int> basic;          // column of each basis position (n + i: row i's logical)
  std::vector<VarState> state;     // per column of [K | -I]
  double primal_violation = 0.0;   // unscaled: max row and bound violation
  double dual_violation = 0.0;     // unscaled: max reduced-cost sign violation
  double seconds = 0.0;
  std::string message;
};

// A basis to start from: the column of each basis position (n + i: row i's
// logical) and the state of every column of [K | -I], as FinishResult holds
// them.
struct FinishBasis {
  std::vector<Int> basic;
  std::vector<VarState> state;
};

// x (cols) and y (rows) are an unscaled primal-dual point for lp, or null
// for a cold start (y alone may be null: bounds are then chosen by
// proximity, as crossover_basis does without a dual). With `start`, the
// simplex starts from that basis instead (the point is not used): the
// reference's case-stack workflow, where the base plan's final basis
// warm-starts each what-if case (app/anukulan_cli.cpp:1065-1230). A
// nonbasic state the case's bounds cannot hold moves to a bound they can,
// and a singular start is repaired with logicals.
FinishResult exact_finish(const StandardLp& lp, const double* x, const double* y,
                          const FinishOptions& opt, const FinishBasis* start = nullptr);

// Many cases at once, one per host thread; `start`, when given, warm-starts
// every one of them.
std::vector<FinishResult> exact_finish_many(const std::vector<const StandardLp*>& lps,
                                            const std::vector<const double*>& xs,
                                            const std::vector<const double*>& ys,
                                            const FinishOptions& opt,
                                            const FinishBasis* start = nullptr);

}  // namespace anuk
