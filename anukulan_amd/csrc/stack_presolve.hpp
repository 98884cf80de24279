// Presolve for a stack of LPs that share one constraint matrix A.
//
// The reference presolves one model at a time in front of its solvers
// (presolve.hpp:485, `presolve(const Model&, const PresolveOptions&)`, and
// PostsolveStack::apply / apply_dual for the way back). The batched solver
// wants more: after presolve every case must still share one matrix, or the
// stack no longer runs as one batch. So each reduction here is taken only
// when it is valid for every case at once. Structure (which rows and columns
// go, which coefficients remain) is decided for the whole stack; the numbers
// a reduction moves (row bounds, column bounds, costs, the objective offset)
// are moved per case.
//
// Reductions, all of which keep A's remaining coefficients untouched:
//   empty row        no live entries: removed, or the case is infeasible;
//   free row         no finite side in any case: removed, dual 0;
//   singleton row    a*x_j in [lo, hi]: becomes x_j's bounds, row removed;
//   redundant side   a side that the activity bounds of the row already
//                    guarantee in every case is dropped; a row with both
//                    sides dropped is removed (dual 0);
//   fixed column     lower == upper in every case (the value may differ per
//                    case): substituted into its rows and the offset;
//   empty column     in no live row: set to the bound its cost prefers, or
//                    the case is unbounded;
//   slack singleton  a column alone in a row that is an equality in every
//                    case: x_j = (b - rest)/a, its cost moves onto the rest of
//                    the row, and its bounds become the row's bounds.
//
// Every reduction inverts exactly on the dual side, so postsolve returns the
// original LP's row duals, not an estimate: a removed empty, free or
// redundant row cannot bind (dual 0); a singleton row's dual is the reduced
// cost of its column when the column sits on a bound that came from the row;
// a slack singleton's row dual gains c_j / a.
//
// Conventions: minimisation. A row dual is positive when the row holds at its
// lower side and negative at its upper side, and the reduced cost of column j
// is c_j - sum_i a_ij y_i; the library's standard form gives its duals in this
// sense once a model row's standard rows are summed with their signs.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "lp.hpp"

namespace anuk {

// Which reductions may fire. The switches that share a name with the
// reference's PresolveOptions (presolve.hpp:59-70, 157) stand for the same
// reduction.
struct PresolveRules {
  bool empty_rows = true;
  bool singleton_rows = true;
  bool redundant_rows = true;
  bool fixed_columns = true;
  bool empty_columns = true;
  // Off by default: each substitution turns an equality row into a ranged
  // row, which the standard form states as two opposed inequalities, and on
  // the refinery case stack that costs the first-order solves more
  // iterations than the removed columns save (24 periods, 61 cases: 509,680
  // lane iterations to rel KKT 1e-4 unreduced, 683,360 with the 336
  // substitutions). The exact solvers are not affected that way.
  bool slack_column_singletons = false;
  // Feasibility slack, relative to max(1, |bound|), for the infeasibility and
  // redundancy tests.
  double tolerance = 1e-9;
  // A slack singleton's coefficient must be at least this fraction of the
  // largest magnitude in its row, so that dividing by it stays well posed.
  double pivot_ratio = 1e-3;
  int max_rounds = 64;
};

// Case k's vectors start at k * cols (objective, column bounds) or k * rows
// (row bounds); offset has one entry per case. Infinite bounds are +-inf.
struct LpStack {
  Int rows = 0;
  Int cols = 0;
  Int cases = 0;
  Csr a;
  std::vector<double> objective, row_lower, row_upper, col_lower, col_upper, offset;
};

enum class PresolveStatus { kReduced = 0, kInfeasible = 1, kUnbounded = 2 };

struct PresolveCounts {
  Int rows_removed = 0;
  Int cols_removed = 0;
  Int nonzeros_removed = 0;
  Int empty_rows = 0;
  Int free_rows = 0;
  Int singleton_rows = 0;
  Int redundant_rows = 0;
  Int redundant_sides = 0;
  Int fixed_columns = 0;
  Int empty_columns = 0;
  Int slack_column_singletons = 0;
  Int rounds = 0;
};

class StackPresolve {
 public:
  PresolveStatus status = PresolveStatus::kReduced;
  Int failing_case = -1;  // the case that proved infeasibility or unboundedness
  std::string message;
  LpStack reduced;
  std::vector<Int> row_origin;  // reduced row -> original row
  std::vector<Int> col_origin;  // reduced column -> original column
  PresolveCounts counts;

  Int original_rows() const { return rows0_; }
  Int original_cols() const { return cols0_; }

  // Case k's solution of the reduced LP (rx: reduced cols, ry: reduced rows,
  // in the conventions above) mapped to the original LP (x: cols, y: rows).
  void postsolve(Int k, const double* rx, const double* ry, double* x, double* y) const;

 private:
  friend class StackReducer;
  enum class Kind { kFix, kBound, kSlack };
  // One reduction, replayed in reverse by postsolve. `data` indexes the
  // per-case payload (`width` doubles per case, case-major).
  struct Step {
    Kind kind;
    Int row;
    Int col;
    double coef;
    std::size_t terms_begin, terms_end;  // kSlack: the rest of the row
    std::size_t data;
  };
  std::vector<Step> steps_;
  std::vector<Int> term_col_;
  std::vector<double> term_value_;
  std::vector<double> data_;
  Csr at_;  // the original A^T, for reduced costs
  Int rows0_ = 0;
  Int cols0_ = 0;
  Int cases_ = 0;
};

StackPresolve presolve_stack(const LpStack& stack, const PresolveRules& rules);

}  // namespace anuk
