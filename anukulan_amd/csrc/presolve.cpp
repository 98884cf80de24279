#include "presolve.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

namespace anuk {

// ---- postsolve ------------------------------------------------------------

void Postsolve::fixed(Int col, double value) { cols_.push_back({col, false, value, 1.0, {}}); }

void Postsolve::solved_from_row(Int col, double coefficient, double rhs, std::vector<Term> rest) {
  cols_.push_back({col, true, rhs, coefficient, std::move(rest)});
}

void Postsolve::row_zero(Int row) { rows_.push_back({RowRule::kZero, row, 0, 1.0, {}}); }

void Postsolve::row_free_singleton(Int row, Int col, double coefficient) {
  rows_.push_back({RowRule::kFreeSingleton, row, col, coefficient, {}});
}

void Postsolve::row_singleton(Int row, Int col, double coefficient) {
  rows_.push_back({RowRule::kSingleton, row, col, coefficient, {}});
}

void Postsolve::row_doubleton(Int row, Int col, double coefficient, std::vector<Term> col_rows) {
  rows_.push_back({RowRule::kDoubleton, row, col, coefficient, std::move(col_rows)});
}

void Postsolve::row_slack_singleton(Int row, Int col, double coefficient) {
  rows_.push_back({RowRule::kSlackSingleton, row, col, coefficient, {}});
}

void Postsolve::row_unrecoverable(Int row) {
  rows_.push_back({RowRule::kUnrecoverable, row, 0, 1.0, {}});
  ++unrecoverable_;
}

void Postsolve::set_maps(Int original_rows, Int original_cols, std::vector<Int> row_map,
                         std::vector<Int> col_map) {
  original_rows_ = original_rows;
  original_cols_ = original_cols;
  row_map_ = std::move(row_map);
  col_map_ = std::move(col_map);
}

// Replayed newest first: a column read off a row needs that row's other
// columns, and any of them that were removed were removed later.
std::vector<double> Postsolve::primal(const std::vector<double>& reduced_x) const {
  std::vector<double> x(uz(original_cols_), 0.0);
  const std::size_t n = std::min(reduced_x.size(), col_map_.size());
  for (std::size_t k = 0; k < n; ++k) x[uz(col_map_[k])] = reduced_x[k];
  for (auto it = cols_.rbegin(); it != cols_.rend(); ++it) {
    if (!it->from_row) {
      x[uz(it->col)] = it->value;
      continue;
    }
    double rest = 0.0;
    for (const Term& t : it->rest) rest += t.coefficient * x[uz(t.index)];
    x[uz(it->col)] = (it->value - rest) / it->coefficient;
  }
  return x;
}

std::vector<double> Postsolve::dual(const std::vector<double>& reduced_y,
                                    const std::vector<double>& reduced_costs,
                                    const std::vector<double>& original_cost) const {
  std::vector<double> y(uz(original_rows_), 0.0);
  const std::size_t m = std::min(reduced_y.size(), row_map_.size());
  for (std::size_t k = 0; k < m; ++k) y[uz(row_map_[k])] = reduced_y[k];
  // Reduced costs are by reduced column, the records by original column.
  std::vector<Int> reduced_of(uz(original_cols_), -1);
  for (std::size_t k = 0; k < col_map_.size(); ++k) {
    const Int o = col_map_[k];
    if (o >= 0 && uz(o) < reduced_of.size()) reduced_of[uz(o)] = static_cast<Int>(k);
  }
  auto reduced_cost = [&](Int col) {
    if (col < 0 || uz(col) >= reduced_of.size()) return 0.0;
    const Int k = reduced_of[uz(col)];
    if (k < 0 || uz(k) >= reduced_costs.size()) return 0.0;
    return reduced_costs[uz(k)];
  };
  auto cost = [&](Int col) { return uz(col) < original_cost.size() ? original_cost[uz(col)] : 0.0; };
  auto usable = [](double a) { return std::fabs(a) > 1e-12; };
  for (auto it = rows_.rbegin(); it != rows_.rend(); ++it) {
    double& yr = y[uz(it->row)];
    switch (it->rule) {
      case RowRule::kZero:
      case RowRule::kUnrecoverable:
        yr = 0.0;
        break;
      case RowRule::kFreeSingleton:  // the free column's reduced cost is zero
        yr = usable(it->coefficient) ? cost(it->col) / it->coefficient : 0.0;
        break;
      case RowRule::kDoubleton: {  // the implied-free column's stationarity
        double rest = 0.0;
        for (const Term& t : it->col_rows) rest += t.coefficient * y[uz(t.index)];
        yr = usable(it->coefficient) ? (cost(it->col) - rest) / it->coefficient : 0.0;
        break;
      }
      case RowRule::kSlackSingleton:  // accumulates onto the row's own dual
        if (usable(it->coefficient)) yr += cost(it->col) / it->coefficient;
        break;
      case RowRule::kSingleton:  // the bound's reduced cost moves back to the row
        yr = usable(it->coefficient) ? reduced_cost(it->col) / it->coefficient : 0.0;
        break;
    }
  }
  return y;
}

// ---- reductions -------------------------------------------------------------

namespace {

// max and min as the reference's C library computes them (a NaN yields the
// other operand; of two equal values, zeros of either sign included, the
// second), so that signed zeros come out as they do there: compilers are free
// to inline std::fmax with either zero.
inline double max2(double x, double y) {
  if (std::isnan(x)) return y;
  if (std::isnan(y)) return x;
  return x > y ? x : y;
}
inline double min2(double x, double y) {
  if (std::isnan(x)) return y;
  if (std::isnan(y)) return x;
  return x < y ? x : y;
}

double scale_of(double v) { return max2(1.0, std::fabs(v)); }

struct Activity {
  double min_value = 0.0, max_value = 0.0;
  Int min_infinite = 0, max_infinite = 0;
  double max_abs_coeff = 0.0;
};

// The working state: A (rows) and A^T (columns) stay as given; rows and
// columns die by flag, and live counts are kept per row and column.
class Reducer {
 public:
  Reducer(const MpsModel& m, const PresolveOptions& opt, PresolveOutcome* out)
      : m_(m), a_(m.a), opt_(opt), post_(out->post), n_(out->counts) {}

  void run(PresolveOutcome* out);

 private:
  Int rows() const { return a_.rows; }
  Int cols() const { return a_.cols; }

  void fail(const char* why) {
    if (!infeasible_ && !unbounded_) message_ = why;
    infeasible_ = true;
  }
  void unbounded(const char* why) {
    unbounded_ = true;
    message_ = why;
  }

  void drop_row(Int i);
  void drop_col(Int j);
  void fix(Int j, double value);
  bool raise_lower(Int j, double value);
  bool cut_upper(Int j, double value);
  Activity activity(Int i) const;

  bool pass_rows();
  bool pass_cols(bool allow_removal);
  bool parallel_rows();
  bool tighten_coefficients();
  bool doubletons(bool allow_removal);
  void assemble(PresolveOutcome* out);

  const MpsModel& m_;
  const Csr& a_;
  Csr at_;
  const PresolveOptions& opt_;
  Postsolve& post_;
  PresolveCounts& n_;

  std::vector<char> row_alive_, col_alive_, integral_;
  std::vector<Int> row_count_, col_count_;
  std::vector<double> rlo_, rup_, clo_, cup_, cost_;  // cost in the minimise direction
  double cost_offset_ = 0.0;

  struct Substitution {
    Int col, other;
    double coefficient, other_coefficient, rhs;
  };
  std::vector<Substitution> subs_;
  std::vector<Int> sub_of_col_;
  std::vector<std::pair<Int, double>> overrides_;  // A's value index -> new value

  double tol_ = 1e-9, infeas_ = 1e-7, relax_ = 0.0, max_finite_ = 1e7;
  bool infeasible_ = false, unbounded_ = false;
  std::string message_;
};

void Reducer::drop_row(Int i) {
  if (!row_alive_[uz(i)]) return;
  row_alive_[uz(i)] = 0;
  for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
    const Int j = a_.index[uz(e)];
    if (col_alive_[uz(j)]) {
      --col_count_[uz(j)];
      ++n_.nonzeros_removed;
    }
  }
  row_count_[uz(i)] = 0;
  ++n_.rows_removed;
}

void Reducer::drop_col(Int j) {
  if (!col_alive_[uz(j)]) return;
  col_alive_[uz(j)] = 0;
  for (Int e = at_.start[uz(j)]; e < at_.start[uz(j) + 1]; ++e) {
    const Int i = at_.index[uz(e)];
    if (row_alive_[uz(i)]) {
      --row_count_[uz(i)];
      ++n_.nonzeros_removed;
    }
  }
  col_count_[uz(j)] = 0;
  ++n_.cols_removed;
}

// A known value moves into the bounds of every live row holding the column.
void Reducer::fix(Int j, double value) {
  for (Int e = at_.start[uz(j)]; e < at_.start[uz(j) + 1]; ++e) {
    const Int i = at_.index[uz(e)];
    if (!row_alive_[uz(i)]) continue;
    const double a = at_.value[uz(e)];
    if (!std::isinf(rlo_[uz(i)])) rlo_[uz(i)] -= a * value;
    if (!std::isinf(rup_[uz(i)])) rup_[uz(i)] -= a * value;
  }
  cost_offset_ += cost_[uz(j)] * value;
  post_.fixed(j, value);
  drop_col(j);
}

bool Reducer::raise_lower(Int j, double value) {
  if (!std::isfinite(value)) return false;
  double v = value;
  if (integral_[uz(j)]) {
    v = std::ceil(v - 1e-9);
  } else {
    v -= relax_ * scale_of(v);
  }
  const double now = clo_[uz(j)];
  if (std::isinf(now) && std::fabs(v) > max_finite_) return false;
  const double gate = std::isinf(now) ? 0.0 : 1e-7 * scale_of(now);
  if (!(v > now + gate)) return false;
  if (std::isinf(now)) {
    ++n_.bounds_made_finite;
    n_.largest_new_finite_bound = max2(n_.largest_new_finite_bound, std::fabs(v));
  }
  clo_[uz(j)] = v;
  ++n_.bounds_tightened;
  if (clo_[uz(j)] > cup_[uz(j)] + infeas_ * scale_of(cup_[uz(j)]))
    fail("bound tightening crossed the bounds of a column");
  return true;
}

bool Reducer::cut_upper(Int j, double value) {
  if (!std::isfinite(value)) return false;
  double v = value;
  if (integral_[uz(j)]) {
    v = std::floor(v + 1e-9);
  } else {
    v += relax_ * scale_of(v);
  }
  const double now = cup_[uz(j)];
  if (std::isinf(now) && std::fabs(v) > max_finite_) return false;
  const double gate = std::isinf(now) ? 0.0 : 1e-7 * scale_of(now);
  if (!(v < now - gate)) return false;
  if (std::isinf(now)) {
    ++n_.bounds_made_finite;
    n_.largest_new_finite_bound = max2(n_.largest_new_finite_bound, std::fabs(v));
  }
  cup_[uz(j)] = v;
  ++n_.bounds_tightened;
  if (clo_[uz(j)] > cup_[uz(j)] + infeas_ * scale_of(cup_[uz(j)]))
    fail("bound tightening crossed the bounds of a column");
  return true;
}

// The range of a row's activity over the live columns' boxes, with the
// infinite contributions counted instead of added.
Activity Reducer::activity(Int i) const {
  Activity r;
  for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
    const Int j = a_.index[uz(e)];
    if (!col_alive_[uz(j)]) continue;
    const double a = a_.value[uz(e)];
    r.max_abs_coeff = max2(r.max_abs_coeff, std::fabs(a));
    const double lo = a > 0.0 ? clo_[uz(j)] : cup_[uz(j)];
    const double hi = a > 0.0 ? cup_[uz(j)] : clo_[uz(j)];
    if (std::isinf(lo)) ++r.min_infinite; else r.min_value += a * lo;
    if (std::isinf(hi)) ++r.max_infinite; else r.max_value += a * hi;
  }
  return r;
}

// Rows, in index order (presolve.cpp:409-563): empty, singleton (a bound),
// forcing, redundant, then bound tightening from the row's activity, which is
// computed once per row and not refreshed as bounds move within it.
bool Reducer::pass_rows() {
  bool changed = false;
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)] || infeasible_) continue;
    const double lo = rlo_[uz(i)], hi = rup_[uz(i)];
    if (lo > hi + infeas_ * scale_of(hi)) {
      fail("a row's lower bound rose above its upper bound");
      return changed;
    }
    if (row_count_[uz(i)] == 0) {
      if (!opt_.empty_rows) continue;
      if (lo > infeas_ * scale_of(lo) || hi < -infeas_ * scale_of(hi)) {
        fail("an empty row cannot hold zero");
        return changed;
      }
      ++n_.empty_rows;
      post_.row_zero(i);
      drop_row(i);
      changed = true;
      continue;
    }
    const Activity act = activity(i);
    const double lo_s = std::isinf(lo) ? 1.0 : scale_of(lo);
    const double hi_s = std::isinf(hi) ? 1.0 : scale_of(hi);
    if (act.min_infinite == 0 && !std::isinf(hi) && act.min_value > hi + infeas_ * hi_s) {
      fail("a row's smallest possible activity exceeds its upper bound");
      return changed;
    }
    if (act.max_infinite == 0 && !std::isinf(lo) && act.max_value < lo - infeas_ * lo_s) {
      fail("a row's largest possible activity is below its lower bound");
      return changed;
    }

    if (row_count_[uz(i)] == 1 && opt_.singleton_rows) {
      Int col = -1;
      double a = 0.0;
      for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
        const Int j = a_.index[uz(e)];
        if (!col_alive_[uz(j)]) continue;
        col = j;
        a = a_.value[uz(e)];
        break;
      }
      if (col >= 0 && std::fabs(a) > 1e-30) {
        const double b1 = std::isinf(lo) ? (a > 0.0 ? -kInf : kInf) : lo / a;
        const double b2 = std::isinf(hi) ? (a > 0.0 ? kInf : -kInf) : hi / a;
        raise_lower(col, min2(b1, b2));
        cut_upper(col, max2(b1, b2));
        ++n_.singleton_rows;
        post_.row_singleton(i, col, a);
        drop_row(i);
        changed = true;
        continue;
      }
    }

    // Forcing: met only at one end of the activity range, so every column of
    // the row is pinned to the bound that end needs.
    const bool force_low = act.min_infinite == 0 && !std::isinf(hi) && act.min_value >= hi - tol_ * hi_s;
    const bool force_high = act.max_infinite == 0 && !std::isinf(lo) && act.max_value <= lo + tol_ * lo_s;
    if (opt_.forcing_rows && (force_low || force_high)) {
      for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
        const Int j = a_.index[uz(e)];
        if (!col_alive_[uz(j)]) continue;
        const bool to_lower = (a_.value[uz(e)] > 0.0) == force_low;
        fix(j, to_lower ? clo_[uz(j)] : cup_[uz(j)]);
      }
      ++n_.forcing_rows;
      post_.row_unrecoverable(i);
      drop_row(i);
      changed = true;
      continue;
    }

    if (opt_.redundant_rows) {
      const bool lower_ok = std::isinf(lo) || (act.min_infinite == 0 && act.min_value >= lo - tol_ * lo_s);
      const bool upper_ok = std::isinf(hi) || (act.max_infinite == 0 && act.max_value <= hi + tol_ * hi_s);
      if (lower_ok && upper_ok) {
        ++n_.redundant_rows;
        post_.row_zero(i);
        drop_row(i);
        changed = true;
        continue;
      }
    }

    if (!opt_.bound_tightening) continue;
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      if (!col_alive_[uz(j)]) continue;
      const double a = a_.value[uz(e)];
      if (std::fabs(a) < 1e-12 * max2(1.0, act.max_abs_coeff)) continue;
      const double own_min = a > 0.0 ? clo_[uz(j)] : cup_[uz(j)];
      const double own_max = a > 0.0 ? cup_[uz(j)] : clo_[uz(j)];
      // a x_j <= hi - (least the rest of the row can be)
      if (!std::isinf(hi) && (act.min_infinite == 0 || (act.min_infinite == 1 && std::isinf(own_min)))) {
        const double rest = act.min_infinite == 0 ? act.min_value - a * own_min : act.min_value;
        const double limit = (hi - rest) / a;
        if (a > 0.0 ? cut_upper(j, limit) : raise_lower(j, limit)) {
          changed = true;
          post_.bounds_tightened();
        }
      }
      // a x_j >= lo - (most the rest of the row can be)
      if (!std::isinf(lo) && (act.max_infinite == 0 || (act.max_infinite == 1 && std::isinf(own_max)))) {
        const double rest = act.max_infinite == 0 ? act.max_value - a * own_max : act.max_value;
        const double limit = (lo - rest) / a;
        if (a > 0.0 ? raise_lower(j, limit) : cut_upper(j, limit)) {
          changed = true;
          post_.bounds_tightened();
        }
      }
      if (infeasible_) return changed;
    }
  }
  return changed;
}

// Columns, in index order (presolve.cpp:566-847): fixed, dual fixing, empty,
// then a column alone in one row substituted out of it.
bool Reducer::pass_cols(bool allow_removal) {
  if (!allow_removal) return false;
  bool changed = false;
  for (Int j = 0; j < cols(); ++j) {
    if (!col_alive_[uz(j)] || infeasible_) continue;
    const double cl = clo_[uz(j)], cu = cup_[uz(j)];
    if (cl > cu + infeas_ * scale_of(cu)) {
      fail("a column's bounds crossed");
      return changed;
    }
    if (opt_.fixed_columns && cu - cl <= tol_ * scale_of(cl) && !std::isinf(cl)) {
      ++n_.fixed_columns;
      fix(j, cl);
      changed = true;
      continue;
    }

    // Dual fixing: with no live row stopping a move in the direction the
    // objective prefers (or does not mind), some optimum has the column at
    // that bound.
    if (opt_.dual_fixing && col_count_[uz(j)] > 0) {
      Int down = 0, up = 0;
      for (Int e = at_.start[uz(j)]; e < at_.start[uz(j) + 1]; ++e) {
        const Int i = at_.index[uz(e)];
        if (!row_alive_[uz(i)]) continue;
        const double a = at_.value[uz(e)];
        if (a == 0.0) continue;
        const bool has_lo = !std::isinf(rlo_[uz(i)]), has_hi = !std::isinf(rup_[uz(i)]);
        if (a > 0.0) {
          down += has_lo;
          up += has_hi;
        } else {
          down += has_hi;
          up += has_lo;
        }
        if (down > 0 && up > 0) break;
      }
      const double c = cost_[uz(j)];
      if (down == 0 && c >= 0.0) {
        if (std::isinf(cl)) {
          if (c > 0.0) {
            unbounded("a column with no downward lock has no lower bound");
            return changed;
          }
        } else {
          ++n_.dual_fixed_columns;
          fix(j, cl);
          changed = true;
          continue;
        }
      } else if (up == 0 && c <= 0.0) {
        if (std::isinf(cu)) {
          if (c < 0.0) {
            unbounded("a column with no upward lock has no upper bound");
            return changed;
          }
        } else {
          ++n_.dual_fixed_columns;
          fix(j, cu);
          changed = true;
          continue;
        }
      }
    }

    if (col_count_[uz(j)] == 0) {
      if (!opt_.empty_columns) continue;
      const double c = cost_[uz(j)];
      double value;
      if (c > 0.0) {
        if (std::isinf(cl)) {
          unbounded("a column in no row can decrease the objective forever");
          return changed;
        }
        value = cl;
      } else if (c < 0.0) {
        if (std::isinf(cu)) {
          unbounded("a column in no row can decrease the objective forever");
          return changed;
        }
        value = cu;
      } else {
        value = min2(max2(0.0, cl), cu);
        if (std::isinf(value)) value = 0.0;
      }
      ++n_.empty_columns;
      fix(j, value);
      changed = true;
      continue;
    }

    if (col_count_[uz(j)] != 1 || integral_[uz(j)] ||
        !(opt_.free_column_singletons || opt_.slack_column_singletons))
      continue;
    Int row = -1;
    double a = 0.0;
    for (Int e = at_.start[uz(j)]; e < at_.start[uz(j) + 1]; ++e) {
      const Int i = at_.index[uz(e)];
      if (!row_alive_[uz(i)]) continue;
      row = i;
      a = at_.value[uz(e)];
      break;
    }
    if (row < 0) continue;
    const double lo = rlo_[uz(row)], hi = rup_[uz(row)];
    const Activity act = activity(row);
    if (std::fabs(a) < 1e-7 * act.max_abs_coeff || std::fabs(a) < 1e-10) continue;

    // What the row alone allows x_j, given the rest of the row's range.
    double gives_lo = -kInf, gives_hi = kInf;
    {
      const double own_min = a > 0.0 ? cl : cu;
      const double own_max = a > 0.0 ? cu : cl;
      const bool min_finite = act.min_infinite == 0 || (act.min_infinite == 1 && std::isinf(own_min));
      const bool max_finite = act.max_infinite == 0 || (act.max_infinite == 1 && std::isinf(own_max));
      const double rest_min = act.min_infinite == 0 ? act.min_value - a * own_min : act.min_value;
      const double rest_max = act.max_infinite == 0 ? act.max_value - a * own_max : act.max_value;
      if (std::isfinite(hi) && min_finite) {
        const double t = (hi - rest_min) / a;
        (a > 0.0 ? gives_hi : gives_lo) = t;
      }
      if (std::isfinite(lo) && max_finite) {
        const double t = (lo - rest_max) / a;
        (a > 0.0 ? gives_lo : gives_hi) = t;
      }
    }
    const bool open_above = std::isinf(cu) || gives_hi <= cu + tol_ * scale_of(cu);
    const bool open_below = std::isinf(cl) || gives_lo >= cl - tol_ * scale_of(cl);

    // The right-hand side to stand on: an equality's, or the side of an
    // inequality that the objective's push on x_j makes tight.
    double rhs = 0.0;
    bool have = false, from_inequality = false;
    if (std::isfinite(lo) && std::isfinite(hi) && std::fabs(hi - lo) <= tol_ * scale_of(hi)) {
      rhs = lo;
      have = true;
    } else if (opt_.inequality_column_singletons && cost_[uz(j)] != 0.0) {
      const bool push_up = cost_[uz(j)] < 0.0;
      const double side = (push_up == (a > 0.0)) ? hi : lo;
      if ((push_up ? open_above : open_below) && std::isfinite(side)) {
        rhs = side;
        have = true;
        from_inequality = true;
      }
    }
    if (!have) continue;
    const bool free_col = open_above && open_below;

    // The column's bounds, moved onto the rest of the row.
    double new_lo = -kInf, new_hi = kInf;
    if (!free_col) {
      const double from_upper = std::isinf(cu) ? (a > 0.0 ? -kInf : kInf) : rhs - a * cu;
      const double from_lower = std::isinf(cl) ? (a > 0.0 ? kInf : -kInf) : rhs - a * cl;
      new_lo = min2(from_upper, from_lower);
      new_hi = max2(from_upper, from_lower);
      if ((std::isfinite(new_lo) && std::fabs(new_lo) > max_finite_) ||
          (std::isfinite(new_hi) && std::fabs(new_hi) > max_finite_)) {
        ++n_.slack_singletons_declined;
        continue;
      }
    }
    if (free_col ? !opt_.free_column_singletons : !opt_.slack_column_singletons) continue;

    std::vector<Postsolve::Term> rest;
    for (Int e = a_.start[uz(row)]; e < a_.start[uz(row) + 1]; ++e) {
      const Int k = a_.index[uz(e)];
      if (k == j || !col_alive_[uz(k)]) continue;
      rest.push_back({k, a_.value[uz(e)]});
    }
    // x_j = (rhs - rest) / a moves c_j x_j onto a constant and the rest.
    const double cj = cost_[uz(j)];
    if (cj != 0.0) {
      cost_offset_ += cj * rhs / a;
      for (const Postsolve::Term& t : rest) cost_[uz(t.index)] -= cj * t.coefficient / a;
    }
    post_.solved_from_row(j, a, rhs, std::move(rest));
    if (from_inequality) ++n_.inequality_column_singletons;
    if (free_col) {
      post_.row_free_singleton(row, j, a);
      ++n_.free_column_singletons;
      drop_col(j);
      drop_row(row);
    } else {
      post_.row_slack_singleton(row, j, a);
      ++n_.slack_column_singletons;
      drop_col(j);
      rlo_[uz(row)] = new_lo;
      rup_[uz(row)] = new_hi;
    }
    changed = true;
  }
  return changed;
}

double bound_over(double bound, double r) {
  if (std::isinf(bound)) return r > 0.0 ? bound : -bound;
  return bound / r;
}

// Parallel rows (presolve.cpp:855-944): rows are grouped by a hash of their
// live column pattern, and within a group, in row order, a later row that is
// a multiple of an earlier one hands its bounds over and goes. Groups share
// no rows, so the order in which groups are visited changes nothing.
bool Reducer::parallel_rows() {
  std::vector<std::pair<std::uint64_t, Int>> keyed;
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)] || row_count_[uz(i)] < 2) continue;
    std::uint64_t h = 1469598103934665603ull;
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      if (!col_alive_[uz(j)]) continue;
      h ^= static_cast<std::uint64_t>(static_cast<std::uint32_t>(j));
      h *= 1099511628211ull;
    }
    keyed.push_back({h, i});
  }
  std::stable_sort(keyed.begin(), keyed.end(),
                   [](const auto& x, const auto& y) { return x.first < y.first; });
  bool changed = false;
  std::vector<Int> ca, cb;
  std::vector<double> va, vb;
  auto live_row = [&](Int i, std::vector<Int>& c, std::vector<double>& v) {
    c.clear();
    v.clear();
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      if (!col_alive_[uz(j)]) continue;
      c.push_back(j);
      v.push_back(a_.value[uz(e)]);
    }
  };
  for (std::size_t g = 0; g < keyed.size();) {
    std::size_t end = g + 1;
    while (end < keyed.size() && keyed[end].first == keyed[g].first) ++end;
    const std::size_t limit = std::min<std::size_t>(end - g, 64);  // the quadratic guard
    for (std::size_t x = 0; x < limit; ++x) {
      const Int i1 = keyed[g + x].second;
      if (!row_alive_[uz(i1)]) continue;
      live_row(i1, ca, va);
      if (ca.empty()) continue;
      for (std::size_t y = x + 1; y < limit; ++y) {
        const Int i2 = keyed[g + y].second;
        if (!row_alive_[uz(i2)]) continue;
        live_row(i2, cb, vb);
        if (cb != ca) continue;
        if (std::fabs(va[0]) < 1e-30) continue;
        const double r = vb[0] / va[0];
        if (std::fabs(r) < 1e-12 || !std::isfinite(r)) continue;
        bool parallel = true;
        for (std::size_t k = 1; k < va.size(); ++k) {
          const double want = r * va[k];
          if (std::fabs(vb[k] - want) > 1e-9 * max2(1.0, std::fabs(want))) {
            parallel = false;
            break;
          }
        }
        if (!parallel) continue;
        const double b1 = bound_over(rlo_[uz(i2)], r), b2 = bound_over(rup_[uz(i2)], r);
        rlo_[uz(i1)] = max2(rlo_[uz(i1)], min2(b1, b2));
        rup_[uz(i1)] = min2(rup_[uz(i1)], max2(b1, b2));
        if (rlo_[uz(i1)] > rup_[uz(i1)] + infeas_ * scale_of(rup_[uz(i1)])) {
          fail("two parallel rows contradict each other");
          return changed;
        }
        ++n_.duplicate_rows;
        post_.row_unrecoverable(i2);
        drop_row(i2);
        changed = true;
      }
    }
    g = end;
  }
  return changed;
}

// Coefficient tightening on integer columns (presolve.cpp:947-1011); off by
// default, and a no-op on a pure LP.
bool Reducer::tighten_coefficients() {
  if (!opt_.coefficient_tightening) return false;
  bool changed = false;
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)] || row_count_[uz(i)] < 2) continue;
    const double lo = rlo_[uz(i)], hi = rup_[uz(i)];
    const bool has_lo = !std::isinf(lo), has_hi = !std::isinf(hi);
    if (has_lo == has_hi) continue;  // a range or an equality, or a free row
    const Activity act = activity(i);
    if (act.max_infinite > 0 || act.min_infinite > 0) continue;
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      if (!col_alive_[uz(j)] || !integral_[uz(j)]) continue;
      if (!sub_of_col_.empty() && sub_of_col_[uz(j)] >= 0) continue;
      const double a = a_.value[uz(e)];
      if (a == 0.0) continue;
      const double cl = clo_[uz(j)], cu = cup_[uz(j)];
      if (std::isinf(cl) || std::isinf(cu) || cu - cl < 1.0 - 1e-9) continue;
      const double sign = has_hi ? 1.0 : -1.0;
      const double coeff = sign * a;
      const double bound = has_hi ? hi : -lo;
      const double max_here = coeff > 0.0 ? coeff * cu : coeff * cl;
      const double max_rest = (has_hi ? act.max_value : -act.min_value) - max_here;
      if (coeff <= 0.0) continue;
      const double span = cu - cl;
      const double d = max_rest + coeff * cu - bound;
      if (!(d > tol_ * scale_of(bound))) continue;
      if (max_rest + coeff * cl > bound - tol_ * scale_of(bound)) continue;
      if (d >= coeff * span) continue;
      const double fresh = coeff - d;
      if (!(fresh > 0.0)) continue;
      overrides_.push_back({e, sign * fresh});
      if (has_hi) rup_[uz(i)] = hi - d; else rlo_[uz(i)] = lo + d;
      ++n_.coefficients_tightened;
      changed = true;
      break;  // the activity is stale now
    }
  }
  return changed;
}

// Doubleton equations a*xp + b*xq = c with xp implied free
// (presolve.cpp:1014-1127): xp is written in terms of xq, its other rows'
// bounds absorb the constant, and the coefficient it leaves on xq is emitted
// when the reduced matrix is assembled. Runs once, after the rounds.
bool Reducer::doubletons(bool allow_removal) {
  if (!opt_.doubleton_equations || !allow_removal) return false;
  sub_of_col_.assign(uz(cols()), -1);
  std::vector<char> kept(uz(cols()), 0);  // a substitution's target may not go itself
  bool changed = false;
  Int cand_col[2] = {0, 0};
  double cand_val[2] = {0.0, 0.0};
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)] || row_count_[uz(i)] != 2) continue;
    const double lo = rlo_[uz(i)], hi = rup_[uz(i)];
    if (!std::isfinite(lo) || !std::isfinite(hi)) continue;
    if (std::fabs(hi - lo) > tol_ * scale_of(lo)) continue;
    const double rhs = lo;
    int live = 0;
    bool blocked = false;
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      if (sub_of_col_[uz(j)] >= 0) {  // before the liveness test: substituted is dead too
        blocked = true;
        break;
      }
      if (!col_alive_[uz(j)]) continue;
      if (integral_[uz(j)]) {
        blocked = true;
        break;
      }
      if (live < 2) {
        cand_col[live] = j;
        cand_val[live] = a_.value[uz(e)];
      }
      ++live;
    }
    if (blocked || live != 2) continue;

    int pick = -1;
    for (int t = 0; t < 2; ++t) {
      const Int p = cand_col[t], q = cand_col[1 - t];
      const double ap = cand_val[t], aq = cand_val[1 - t];
      if (kept[uz(p)]) continue;
      if (std::fabs(ap) < 1e-9 * scale_of(aq)) continue;
      auto end_at = [&](double xq) {
        if (std::isinf(xq)) {
          const double slope = -aq / ap;
          if (slope == 0.0) return rhs / ap;
          return ((xq > 0.0) == (slope > 0.0)) ? kInf : -kInf;
        }
        return (rhs - aq * xq) / ap;
      };
      double ilo = end_at(clo_[uz(q)]), ihi = end_at(cup_[uz(q)]);
      if (ilo > ihi) std::swap(ilo, ihi);
      const double give = tol_ * scale_of(rhs);
      if (ilo < clo_[uz(p)] - give || ihi > cup_[uz(p)] + give) continue;
      if (pick < 0 || col_count_[uz(p)] < col_count_[uz(cand_col[pick])]) pick = t;
    }
    if (pick < 0) continue;
    const Int p = cand_col[pick], q = cand_col[1 - pick];
    const double ap = cand_val[pick], aq = cand_val[1 - pick];
    drop_row(i);
    std::vector<Postsolve::Term> col_rows;
    const double constant = rhs / ap;
    for (Int e = at_.start[uz(p)]; e < at_.start[uz(p) + 1]; ++e) {
      const Int k = at_.index[uz(e)];
      if (!row_alive_[uz(k)]) continue;
      const double akp = at_.value[uz(e)];
      if (std::isfinite(rlo_[uz(k)])) rlo_[uz(k)] -= akp * constant;
      if (std::isfinite(rup_[uz(k)])) rup_[uz(k)] -= akp * constant;
      col_rows.push_back({k, akp});
    }
    cost_[uz(q)] -= cost_[uz(p)] * (aq / ap);
    cost_offset_ += cost_[uz(p)] * constant;
    post_.solved_from_row(p, ap, rhs, {{q, aq}});
    post_.row_doubleton(i, p, ap, std::move(col_rows));
    drop_col(p);
    sub_of_col_[uz(p)] = static_cast<Int>(subs_.size());
    subs_.push_back({p, q, ap, aq, rhs});
    kept[uz(q)] = 1;
    ++n_.doubleton_equations;
    changed = true;
  }
  return changed;
}

// The reduced model (presolve.cpp:1130-1211): live rows and columns in their
// original order, substituted columns' entries re-aimed at their targets and
// merged by the triplet build.
void Reducer::assemble(PresolveOutcome* out) {
  std::vector<Int> new_row(uz(rows()), -1), new_col(uz(cols()), -1), row_map, col_map;
  for (Int i = 0; i < rows(); ++i)
    if (row_alive_[uz(i)]) {
      new_row[uz(i)] = static_cast<Int>(row_map.size());
      row_map.push_back(i);
    }
  for (Int j = 0; j < cols(); ++j)
    if (col_alive_[uz(j)]) {
      new_col[uz(j)] = static_cast<Int>(col_map.size());
      col_map.push_back(j);
    }
  const Int r = static_cast<Int>(row_map.size()), c = static_cast<Int>(col_map.size());
  MpsModel& red = out->reduced;
  red.name = m_.name;
  red.maximize = m_.maximize;
  const double s = m_.maximize ? -1.0 : 1.0;
  red.objective_offset = m_.objective_offset + s * cost_offset_;
  std::vector<Entry> entries;
  entries.reserve(a_.value.size());
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)]) continue;
    for (Int e = a_.start[uz(i)]; e < a_.start[uz(i) + 1]; ++e) {
      const Int j = a_.index[uz(e)];
      double v = a_.value[uz(e)];
      for (const auto& ov : overrides_)
        if (ov.first == e) {
          v = ov.second;
          break;
        }
      const Int sub = sub_of_col_.empty() ? -1 : sub_of_col_[uz(j)];
      if (sub >= 0) {
        const Substitution& t = subs_[uz(sub)];
        if (!col_alive_[uz(t.other)]) continue;
        entries.push_back({new_row[uz(i)], new_col[uz(t.other)], -v * t.other_coefficient / t.coefficient});
        continue;
      }
      if (!col_alive_[uz(j)]) continue;
      entries.push_back({new_row[uz(i)], new_col[uz(j)], v});
    }
  }
  red.a = Csr::from_triplets(r, c, std::move(entries));
  for (Int j = 0; j < cols(); ++j) {
    if (!col_alive_[uz(j)]) continue;
    red.objective.push_back(s * cost_[uz(j)]);
    red.col_lower.push_back(clo_[uz(j)]);
    red.col_upper.push_back(cup_[uz(j)]);
    red.col_type.push_back(m_.col_type[uz(j)]);
    if (!m_.col_names.empty()) red.col_names.push_back(m_.col_names[uz(j)]);
  }
  for (Int i = 0; i < rows(); ++i) {
    if (!row_alive_[uz(i)]) continue;
    red.row_lower.push_back(rlo_[uz(i)]);
    red.row_upper.push_back(rup_[uz(i)]);
    if (!m_.row_names.empty()) red.row_names.push_back(m_.row_names[uz(i)]);
  }
  red.hessian = (m_.hessian.nnz() > 0 && c == cols()) ? m_.hessian : Csr::from_triplets(0, 0, {});
  red.fill_coo();
  post_.set_maps(rows(), cols(), std::move(row_map), std::move(col_map));
}

void Reducer::run(PresolveOutcome* out) {
  const Int m = rows(), n = cols();
  at_ = a_.transpose();
  row_alive_.assign(uz(m), 1);
  col_alive_.assign(uz(n), 1);
  row_count_.resize(uz(m));
  col_count_.resize(uz(n));
  for (Int i = 0; i < m; ++i) row_count_[uz(i)] = a_.start[uz(i) + 1] - a_.start[uz(i)];
  for (Int j = 0; j < n; ++j) col_count_[uz(j)] = at_.start[uz(j) + 1] - at_.start[uz(j)];
  rlo_ = m_.row_lower;
  rup_ = m_.row_upper;
  clo_ = m_.col_lower;
  cup_ = m_.col_upper;
  integral_.assign(uz(n), 0);
  for (Int j = 0; j < n; ++j)
    integral_[uz(j)] = (m_.col_type[uz(j)] == kInteger || m_.col_type[uz(j)] == kSemiInteger) ? 1 : 0;

  // Fractional bounds on an integer column are rounded once, up front.
  for (Int j = 0; j < n; ++j) {
    if (!integral_[uz(j)]) continue;
    if (std::isfinite(clo_[uz(j)])) {
      const double v = std::ceil(clo_[uz(j)] - 1e-9);
      if (v > clo_[uz(j)]) {
        clo_[uz(j)] = v;
        ++n_.bounds_tightened;
      }
    }
    if (std::isfinite(cup_[uz(j)])) {
      const double v = std::floor(cup_[uz(j)] + 1e-9);
      if (v < cup_[uz(j)]) {
        cup_[uz(j)] = v;
        ++n_.bounds_tightened;
      }
    }
    if (clo_[uz(j)] > cup_[uz(j)] + opt_.infeasibility_tolerance) {
      out->status = PresolveStatus::kInfeasible;
      out->message = "an integer column has no whole number between its bounds";
      return;
    }
  }
  const double s = m_.maximize ? -1.0 : 1.0;
  cost_.resize(uz(n));
  for (Int j = 0; j < n; ++j) cost_[uz(j)] = s * m_.objective[uz(j)];
  tol_ = opt_.feasibility_tolerance;
  infeas_ = max2(opt_.infeasibility_tolerance, opt_.feasibility_tolerance);
  relax_ = opt_.bound_relaxation;
  max_finite_ = opt_.max_new_finite_bound;
  const bool allow_removal = m_.hessian.nnz() == 0;

  for (Int round = 0; round < opt_.max_rounds; ++round) {
    ++n_.rounds;
    bool changed = pass_rows();
    if (infeasible_ || unbounded_) break;
    changed |= pass_cols(allow_removal);
    if (infeasible_ || unbounded_) break;
    if (opt_.duplicate_rows && round < 2) changed |= parallel_rows();
    if (infeasible_ || unbounded_) break;
    if (!changed) break;
  }
  if (!infeasible_ && !unbounded_) tighten_coefficients();
  if (!infeasible_ && !unbounded_) doubletons(allow_removal);

  if (unbounded_) {
    out->status = PresolveStatus::kUnbounded;
    out->message = message_;
  } else if (infeasible_) {
    out->status = PresolveStatus::kInfeasible;
    out->message = message_;
  } else {
    out->status = PresolveStatus::kReduced;
    assemble(out);
  }
}

}  // namespace

PresolveOutcome presolve(const MpsModel& model, const PresolveOptions& opt) {
  PresolveOutcome out;
  out.original_rows = model.a.rows;
  out.original_cols = model.a.cols;
  out.original_nnz = model.a.nnz();
  std::string why;
  if (!validate_model(model, &why)) {
    out.status = PresolveStatus::kInfeasible;
    out.message = "model did not validate: " + why;
    return out;
  }
  Reducer(model, opt, &out).run(&out);
  return out;
}

}  // namespace anuk
