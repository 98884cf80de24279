// Presolve in front of the solver and postsolve behind it, on the natural-form
// model (the one an MPS file states).
//
// The reductions, their order, their tolerances and what each records for
// postsolve follow the reference's presolve (src/anukulan/presolve.hpp:54-489,
// presolve.cpp:290-1317), so a model reduces to the same reduced model, bit
// for bit, and a reduced solution maps back to the same original one:
//   rounds of   row scan    empty, singleton, forcing, redundant rows and
//                           bound tightening by interval arithmetic
//               column scan fixed columns, dual fixing, empty columns, and
//                           column singletons (implied free: row and column
//                           go; otherwise the row takes the column's bounds)
//               duplicate (parallel) rows, in the first two rounds
//   then        coefficient tightening (off by default) and doubleton
//               equations, once.
// A's coefficients are never edited in place: removals are flags and counts,
// and the two reductions that change coefficients (doubletons, coefficient
// tightening) are applied when the reduced matrix is assembled.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mps.hpp"

namespace anuk {

struct PresolveOptions {  // defaults as presolve.hpp:54-330
  bool empty_rows = true;
  bool singleton_rows = true;
  bool redundant_rows = true;
  bool forcing_rows = true;
  bool duplicate_rows = true;
  bool fixed_columns = true;
  bool empty_columns = true;
  bool free_column_singletons = true;
  bool bound_tightening = true;
  bool slack_column_singletons = true;
  bool inequality_column_singletons = true;
  bool dual_fixing = true;
  bool doubleton_equations = true;
  bool coefficient_tightening = false;
  Int max_rounds = 30;
  double feasibility_tolerance = 1e-9;
  double infeasibility_tolerance = 1e-7;
  double bound_relaxation = 0.0;
  double max_new_finite_bound = 1e7;
};

struct PresolveCounts {  // presolve.hpp:418-451
  Int rows_removed = 0, cols_removed = 0, nonzeros_removed = 0;
  Int empty_rows = 0, singleton_rows = 0, redundant_rows = 0, forcing_rows = 0, duplicate_rows = 0;
  Int fixed_columns = 0, empty_columns = 0, free_column_singletons = 0;
  Int slack_column_singletons = 0, inequality_column_singletons = 0, slack_singletons_declined = 0;
  Int doubleton_equations = 0, dual_fixed_columns = 0, coefficients_tightened = 0;
  Int bounds_tightened = 0, bounds_made_finite = 0;
  double largest_new_finite_bound = 0.0;
  Int rounds = 0;
};

enum class PresolveStatus : std::int32_t { kReduced = 0, kInfeasible = 1, kUnbounded = 2 };

// What postsolve replays, in reverse order of recording.
class Postsolve {
 public:
  struct Term {
    Int index;  // a column (primal entries) or a row (doubleton dual entries)
    double coefficient;
  };

  // Primal side: a removed column either took a value, or is read off the one
  // equality row it was substituted from.
  void fixed(Int col, double value);
  void solved_from_row(Int col, double coefficient, double rhs, std::vector<Term> rest);

  // Dual side, per removed (or altered) row.
  void row_zero(Int row);
  void row_free_singleton(Int row, Int col, double coefficient);
  void row_singleton(Int row, Int col, double coefficient);
  void row_doubleton(Int row, Int col, double coefficient, std::vector<Term> col_rows);
  void row_slack_singleton(Int row, Int col, double coefficient);
  void row_unrecoverable(Int row);
  void bounds_tightened() { tightened_ = true; }

  void set_maps(Int original_rows, Int original_cols, std::vector<Int> row_map,
                std::vector<Int> col_map);

  // The original x from the reduced one (reduced_cols entries).
  std::vector<double> primal(const std::vector<double>& reduced_x) const;
  // The original row duals from the reduced model's row duals and reduced
  // costs (either may be shorter, even empty: what depends on a missing entry
  // gets zero) and the original model's objective (presolve.cpp:134-229).
  std::vector<double> dual(const std::vector<double>& reduced_y, const std::vector<double>& reduced_costs,
                           const std::vector<double>& original_cost) const;

  bool dual_exact() const { return unrecoverable_ == 0 && !tightened_; }
  Int unrecoverable_rows() const { return unrecoverable_; }
  Int original_rows() const { return original_rows_; }
  Int original_cols() const { return original_cols_; }
  const std::vector<Int>& row_map() const { return row_map_; }  // reduced row -> original
  const std::vector<Int>& col_map() const { return col_map_; }  // reduced column -> original

 private:
  struct ColStep {
    Int col;
    bool from_row;
    double value;        // fixed value, or the row's right-hand side
    double coefficient;  // the column's coefficient in that row
    std::vector<Term> rest;
  };
  enum class RowRule : std::int8_t { kZero, kFreeSingleton, kSingleton, kDoubleton, kSlackSingleton, kUnrecoverable };
  struct RowStep {
    RowRule rule;
    Int row;
    Int col;
    double coefficient;
    std::vector<Term> col_rows;
  };
  std::vector<ColStep> cols_;
  std::vector<RowStep> rows_;
  Int original_rows_ = 0, original_cols_ = 0, unrecoverable_ = 0;
  bool tightened_ = false;
  std::vector<Int> row_map_, col_map_;
};

struct PresolveOutcome {
  PresolveStatus status = PresolveStatus::kReduced;
  std::string message;
  MpsModel reduced;  // valid when status is kReduced
  Postsolve post;
  PresolveCounts counts;
  Int original_rows = 0, original_cols = 0, original_nnz = 0;
};

// Column-removing reductions are skipped when the model has a Hessian
// (presolve.hpp:481-485).
PresolveOutcome presolve(const MpsModel& model, const PresolveOptions& opt);

}  // namespace anuk
