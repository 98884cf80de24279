// Presolve for a stack of LPs sharing A: the reductions, their undo log and
// the way back. The rules are the textbook ones (Andersen and Andersen, Math.
// Prog. 71, 1995), restated for a stack: a structural decision is taken only
// when every case allows it, and the numbers it moves are moved per case.
#include "stack_presolve.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace anuk {

namespace {

double slack(double bound, double tolerance) {
  return tolerance * std::max(1.0, std::fabs(bound));
}

// The range a*x takes over x in [lower, upper]; infinite ends are counted
// rather than added, so that a finite sum stays usable when one end is not.
struct Range {
  double low = 0.0, high = 0.0;
  Int low_inf = 0, high_inf = 0;

  void add(double a, double lower, double upper) {
    const double lo = a > 0.0 ? lower : upper;
    const double hi = a > 0.0 ? upper : lower;
    if (std::isinf(lo)) ++low_inf; else low += a * lo;
    if (std::isinf(hi)) ++high_inf; else high += a * hi;
  }
};

}  // namespace

// Works on a copy of the stack's vectors; structure lives in the on/off flags
// and the live entry counts of each row and column.
class StackReducer {
 public:
  StackReducer(const LpStack& s, const PresolveRules& r, StackPresolve* out)
      : R_(r), out_(out), m_(s.rows), n_(s.cols), cases_(s.cases), a_(s.a),
        c_(s.objective), rl_(s.row_lower), ru_(s.row_upper), cl_(s.col_lower),
        cu_(s.col_upper), off_(s.offset),
        row_on_(uz(m_), 1), col_on_(uz(n_), 1), row_len_(uz(m_)), col_len_(uz(n_), 0) {
    out_->rows0_ = m_;
    out_->cols0_ = n_;
    out_->cases_ = cases_;
    at_ = a_.transpose();
    for (Int i = 0; i < m_; ++i) row_len_[uz(i)] = a_.start[uz(i) + 1] - a_.start[uz(i)];
    for (Int j = 0; j < n_; ++j) col_len_[uz(j)] = at_.start[uz(j) + 1] - at_.start[uz(j)];
  }

  void run() {
    for (int round = 0; round < R_.max_rounds && !stopped_; ++round) {
      bool changed = false;
      for (Int i = 0; i < m_ && !stopped_; ++i)
        if (row_on_[uz(i)] && row_pass(i)) changed = true;
      for (Int j = 0; j < n_ && !stopped_; ++j)
        if (col_on_[uz(j)] && col_pass(j)) changed = true;
      ++out_->counts.rounds;
      if (!changed) break;
    }
    if (!stopped_) finish();
  }

 private:
  std::size_t at(Int k, Int j) const { return uz(k) * uz(n_) + uz(j); }
  std::size_t rt(Int k, Int i) const { return uz(k) * uz(m_) + uz(i); }

  void stop(PresolveStatus status, Int k, const std::string& why) {
    out_->status = status;
    out_->failing_case = k;
    out_->message = "case " + std::to_string(k) + ": " + why;
    stopped_ = true;
  }

  void remove_row(Int i) {
    row_on_[uz(i)] = 0;
    for (Int p = a_.start[uz(i)]; p < a_.start[uz(i) + 1]; ++p)
      if (col_on_[uz(a_.index[uz(p)])]) --col_len_[uz(a_.index[uz(p)])];
  }

  void remove_col(Int j) {
    col_on_[uz(j)] = 0;
    for (Int p = at_.start[uz(j)]; p < at_.start[uz(j) + 1]; ++p)
      if (row_on_[uz(at_.index[uz(p)])]) --row_len_[uz(at_.index[uz(p)])];
  }

  // Appends a step whose payload holds `width` doubles per case.
  StackPresolve::Step& log(StackPresolve::Kind kind, Int row, Int col, double coef, int width) {
    const std::size_t base = out_->data_.size();
    out_->data_.resize(base + uz(cases_) * static_cast<std::size_t>(width), 0.0);
    const std::size_t t = out_->term_col_.size();
    out_->steps_.push_back({kind, row, col, coef, t, t, base});
    return out_->steps_.back();
  }
  double& payload(const StackPresolve::Step& s, Int k, int width, int field) {
    return out_->data_[s.data + uz(k) * static_cast<std::size_t>(width) + static_cast<std::size_t>(field)];
  }

  bool row_pass(Int i) {
    if (row_len_[uz(i)] == 0) return R_.empty_rows && empty_row(i);
    bool finite = false;
    for (Int k = 0; k < cases_ && !finite; ++k)
      finite = std::isfinite(rl_[rt(k, i)]) || std::isfinite(ru_[rt(k, i)]);
    if (!finite) {
      remove_row(i);
      ++out_->counts.free_rows;
      return true;
    }
    if (row_len_[uz(i)] == 1 && R_.singleton_rows) return singleton_row(i);
    return R_.redundant_rows && redundant_row(i);
  }

  bool empty_row(Int i) {
    for (Int k = 0; k < cases_; ++k) {
      const double lo = rl_[rt(k, i)], hi = ru_[rt(k, i)];
      if (lo > slack(lo, R_.tolerance) || hi < -slack(hi, R_.tolerance)) {
        std::ostringstream s;
        s << "row " << i << " has no entries left but needs [" << lo << ", " << hi << "]";
        stop(PresolveStatus::kInfeasible, k, s.str());
        return false;
      }
    }
    remove_row(i);
    ++out_->counts.empty_rows;
    return true;
  }

  // a*x_j in [lo, hi] becomes x_j's bounds. The payload keeps, per case, the
  // column's cost at this point and which of its bounds now come from the row
  // (bit 0 lower, bit 1 upper): postsolve hands the column's reduced cost to
  // the row exactly when the column rests on such a bound.
  bool singleton_row(Int i) {
    Int j = -1;
    double a = 0.0;
    for (Int p = a_.start[uz(i)]; p < a_.start[uz(i) + 1]; ++p)
      if (col_on_[uz(a_.index[uz(p)])]) {
        j = a_.index[uz(p)];
        a = a_.value[uz(p)];
      }
    StackPresolve::Step& s = log(StackPresolve::Kind::kBound, i, j, a, 2);
    for (Int k = 0; k < cases_; ++k) {
      const double lo = a > 0.0 ? rl_[rt(k, i)] / a : ru_[rt(k, i)] / a;
      const double hi = a > 0.0 ? ru_[rt(k, i)] / a : rl_[rt(k, i)] / a;
      double& l = cl_[at(k, j)];
      double& u = cu_[at(k, j)];
      int from_row = 0;
      if (lo > l) { l = lo; from_row |= 1; }
      if (hi < u) { u = hi; from_row |= 2; }
      if (l > u) {
        if (l - u > slack(std::max(std::fabs(l), std::fabs(u)), R_.tolerance)) {
          std::ostringstream m;
          m << "row " << i << " bounds column " << j << " to [" << l << ", " << u << "]";
          stop(PresolveStatus::kInfeasible, k, m.str());
          return false;
        }
        if (from_row & 1) u = l; else l = u;
      }
      payload(s, k, 2, 0) = c_[at(k, j)];
      payload(s, k, 2, 1) = static_cast<double>(from_row);
    }
    remove_row(i);
    ++out_->counts.singleton_rows;
    return true;
  }

  // Drops every side of row i that the column bounds already guarantee in
  // every case, and the row when both go; proves infeasibility when the
  // activity cannot reach a side.
  bool redundant_row(Int i) {
    bool low_redundant = true, high_redundant = true, low_finite = false, high_finite = false;
    for (Int k = 0; k < cases_; ++k) {
      Range act;
      for (Int p = a_.start[uz(i)]; p < a_.start[uz(i) + 1]; ++p) {
        const Int j = a_.index[uz(p)];
        if (col_on_[uz(j)]) act.add(a_.value[uz(p)], cl_[at(k, j)], cu_[at(k, j)]);
      }
      const double lo = rl_[rt(k, i)], hi = ru_[rt(k, i)];
      if ((act.low_inf == 0 && act.low > hi + slack(hi, R_.tolerance)) ||
          (act.high_inf == 0 && act.high < lo - slack(lo, R_.tolerance))) {
        std::ostringstream m;
        m << "row " << i << " needs [" << lo << ", " << hi << "] but its activity spans ["
          << (act.low_inf ? -kInf : act.low) << ", " << (act.high_inf ? kInf : act.high) << "]";
        stop(PresolveStatus::kInfeasible, k, m.str());
        return false;
      }
      low_finite = low_finite || std::isfinite(lo);
      high_finite = high_finite || std::isfinite(hi);
      low_redundant = low_redundant && (std::isinf(lo) ||
                                        (act.low_inf == 0 && act.low >= lo - slack(lo, R_.tolerance)));
      high_redundant = high_redundant && (std::isinf(hi) ||
                                          (act.high_inf == 0 && act.high <= hi + slack(hi, R_.tolerance)));
    }
    if (low_redundant && high_redundant) {
      remove_row(i);
      ++out_->counts.redundant_rows;
      return true;
    }
    bool changed = false;
    if (low_redundant && low_finite) {
      for (Int k = 0; k < cases_; ++k) rl_[rt(k, i)] = -kInf;
      ++out_->counts.redundant_sides;
      changed = true;
    }
    if (high_redundant && high_finite) {
      for (Int k = 0; k < cases_; ++k) ru_[rt(k, i)] = kInf;
      ++out_->counts.redundant_sides;
      changed = true;
    }
    return changed;
  }

  bool col_pass(Int j) {
    if (col_len_[uz(j)] == 0) return R_.empty_columns && empty_col(j);
    if (R_.fixed_columns) {
      bool fixed = true;
      for (Int k = 0; k < cases_ && fixed; ++k)
        fixed = cl_[at(k, j)] == cu_[at(k, j)] && std::isfinite(cl_[at(k, j)]);
      if (fixed) return fix_col(j);
    }
    return R_.slack_column_singletons && col_len_[uz(j)] == 1 && slack_singleton(j);
  }

  bool empty_col(Int j) {
    StackPresolve::Step& s = log(StackPresolve::Kind::kFix, -1, j, 0.0, 1);
    for (Int k = 0; k < cases_; ++k) {
      const double cost = c_[at(k, j)], l = cl_[at(k, j)], u = cu_[at(k, j)];
      const double v = cost > 0.0 ? l : (cost < 0.0 ? u : std::min(std::max(0.0, l), u));
      if (std::isinf(v)) {
        std::ostringstream m;
        m << "column " << j << " is in no row and its cost " << cost << " drives it to " << v;
        stop(PresolveStatus::kUnbounded, k, m.str());
        return false;
      }
      payload(s, k, 1, 0) = v;
      off_[uz(k)] += cost * v;
    }
    remove_col(j);
    ++out_->counts.empty_columns;
    return true;
  }

  bool fix_col(Int j) {
    StackPresolve::Step& s = log(StackPresolve::Kind::kFix, -1, j, 0.0, 1);
    for (Int k = 0; k < cases_; ++k) {
      const double v = cl_[at(k, j)];
      payload(s, k, 1, 0) = v;
      off_[uz(k)] += c_[at(k, j)] * v;
      for (Int p = at_.start[uz(j)]; p < at_.start[uz(j) + 1]; ++p) {
        const Int i = at_.index[uz(p)];
        if (!row_on_[uz(i)]) continue;
        const double shift = at_.value[uz(p)] * v;
        rl_[rt(k, i)] -= shift;  // an infinite side stays infinite
        ru_[rt(k, i)] -= shift;
      }
    }
    remove_col(j);
    ++out_->counts.fixed_columns;
    return true;
  }

  // Column j alone in row r, which is an equality b in every case:
  // x_j = (b - rest) / a. The rest of the row takes over j's cost
  // (c_t -= c_j a_rt / a, offset += c_j b / a) and j's bounds become the
  // row's: rest in [b - a u, b - a l] for a > 0. Declined when the bounds of
  // x_j are finite in some cases and not in others, or fixed in some and not
  // in others, since the row would then take different forms across cases.
  bool slack_singleton(Int j) {
    Int r = -1;
    double a = 0.0;
    for (Int p = at_.start[uz(j)]; p < at_.start[uz(j) + 1]; ++p)
      if (row_on_[uz(at_.index[uz(p)])]) {
        r = at_.index[uz(p)];
        a = at_.value[uz(p)];
      }
    if (row_len_[uz(r)] < 2) return false;
    double largest = 0.0;
    for (Int p = a_.start[uz(r)]; p < a_.start[uz(r) + 1]; ++p)
      if (col_on_[uz(a_.index[uz(p)])]) largest = std::max(largest, std::fabs(a_.value[uz(p)]));
    if (std::fabs(a) < R_.pivot_ratio * largest) return false;
    const bool lf = std::isfinite(cl_[at(0, j)]), uf = std::isfinite(cu_[at(0, j)]);
    const bool fx = cl_[at(0, j)] == cu_[at(0, j)];
    for (Int k = 0; k < cases_; ++k) {
      const double b = rl_[rt(k, r)];
      if (!(b == ru_[rt(k, r)]) || !std::isfinite(b)) return false;
      if (std::isfinite(cl_[at(k, j)]) != lf || std::isfinite(cu_[at(k, j)]) != uf ||
          (cl_[at(k, j)] == cu_[at(k, j)]) != fx)
        return false;
    }
    StackPresolve::Step& s = log(StackPresolve::Kind::kSlack, r, j, a, 2);
    for (Int p = a_.start[uz(r)]; p < a_.start[uz(r) + 1]; ++p) {
      const Int t = a_.index[uz(p)];
      if (t == j || !col_on_[uz(t)]) continue;
      out_->term_col_.push_back(t);
      out_->term_value_.push_back(a_.value[uz(p)]);
    }
    s.terms_end = out_->term_col_.size();
    for (Int k = 0; k < cases_; ++k) {
      const double b = rl_[rt(k, r)], cj = c_[at(k, j)];
      payload(s, k, 2, 0) = cj;
      payload(s, k, 2, 1) = b;
      if (cj != 0.0) {
        for (std::size_t q = s.terms_begin; q < s.terms_end; ++q)
          c_[at(k, out_->term_col_[q])] -= cj * out_->term_value_[q] / a;
        off_[uz(k)] += cj * b / a;
      }
      const double l = cl_[at(k, j)], u = cu_[at(k, j)];
      rl_[rt(k, r)] = a > 0.0 ? b - a * u : b - a * l;
      ru_[rt(k, r)] = a > 0.0 ? b - a * l : b - a * u;
    }
    remove_col(j);
    ++out_->counts.slack_column_singletons;
    return true;
  }

  void finish() {
    StackPresolve& o = *out_;
    LpStack& red = o.reduced;
    for (Int i = 0; i < m_; ++i)
      if (row_on_[uz(i)]) o.row_origin.push_back(i);
    std::vector<Int> new_col(uz(n_), -1);
    for (Int j = 0; j < n_; ++j)
      if (col_on_[uz(j)]) {
        new_col[uz(j)] = static_cast<Int>(o.col_origin.size());
        o.col_origin.push_back(j);
      }
    red.rows = static_cast<Int>(o.row_origin.size());
    red.cols = static_cast<Int>(o.col_origin.size());
    red.cases = cases_;
    red.a.rows = red.rows;
    red.a.cols = red.cols;
    red.a.start.assign(1, 0);
    for (const Int i : o.row_origin) {
      for (Int p = a_.start[uz(i)]; p < a_.start[uz(i) + 1]; ++p) {
        const Int j = a_.index[uz(p)];
        if (!col_on_[uz(j)]) continue;
        red.a.index.push_back(new_col[uz(j)]);
        red.a.value.push_back(a_.value[uz(p)]);
      }
      red.a.start.push_back(static_cast<Int>(red.a.index.size()));
    }
    red.a.renew_id();
    auto pick = [&](const std::vector<double>& v, const std::vector<Int>& keep, Int width) {
      std::vector<double> r;
      r.reserve(uz(cases_) * keep.size());
      for (Int k = 0; k < cases_; ++k)
        for (const Int e : keep) r.push_back(v[uz(k) * uz(width) + uz(e)]);
      return r;
    };
    red.objective = pick(c_, o.col_origin, n_);
    red.col_lower = pick(cl_, o.col_origin, n_);
    red.col_upper = pick(cu_, o.col_origin, n_);
    red.row_lower = pick(rl_, o.row_origin, m_);
    red.row_upper = pick(ru_, o.row_origin, m_);
    red.offset = off_;
    o.counts.rows_removed = m_ - red.rows;
    o.counts.cols_removed = n_ - red.cols;
    o.counts.nonzeros_removed = a_.nnz() - red.a.nnz();
    o.at_ = std::move(at_);
  }

  const PresolveRules& R_;
  StackPresolve* out_;
  Int m_, n_, cases_;
  const Csr& a_;
  Csr at_;
  std::vector<double> c_, rl_, ru_, cl_, cu_, off_;
  std::vector<char> row_on_, col_on_;
  std::vector<Int> row_len_, col_len_;
  bool stopped_ = false;
};

StackPresolve presolve_stack(const LpStack& s, const PresolveRules& rules) {
  const std::size_t nc = uz(s.cases) * uz(s.cols), nr = uz(s.cases) * uz(s.rows);
  if (s.cases < 1 || s.rows < 0 || s.cols < 0 || s.a.rows != s.rows || s.a.cols != s.cols ||
      s.objective.size() != nc || s.col_lower.size() != nc || s.col_upper.size() != nc ||
      s.row_lower.size() != nr || s.row_upper.size() != nr || s.offset.size() != uz(s.cases))
    throw std::invalid_argument("presolve: the stack's arrays do not match its sizes");
  std::string why;
  if (!s.a.validate(&why)) throw std::invalid_argument("presolve: " + why);
  for (const double v : s.a.value)
    if (v == 0.0 || !std::isfinite(v))
      throw std::invalid_argument("presolve: A holds an explicit zero or a non-finite value");
  for (std::size_t t = 0; t < nc; ++t)
    if (s.col_lower[t] > s.col_upper[t] || std::isnan(s.col_lower[t]) || std::isnan(s.col_upper[t]))
      throw std::invalid_argument("presolve: an empty or NaN column bound interval");
  for (std::size_t t = 0; t < nr; ++t)
    if (s.row_lower[t] > s.row_upper[t] || std::isnan(s.row_lower[t]) || std::isnan(s.row_upper[t]))
      throw std::invalid_argument("presolve: an empty or NaN row bound interval");
  StackPresolve out;
  StackReducer(s, rules, &out).run();
  return out;
}

void StackPresolve::postsolve(Int k, const double* rx, const double* ry, double* x,
                              double* y) const {
  if (status != PresolveStatus::kReduced)
    throw std::invalid_argument("postsolve: the presolve did not produce a reduced stack");
  if (k < 0 || k >= cases_) throw std::invalid_argument("postsolve: no such case");
  std::fill(x, x + cols0_, 0.0);
  std::fill(y, y + rows0_, 0.0);
  for (std::size_t r = 0; r < row_origin.size(); ++r) y[row_origin[r]] = ry[r];
  for (std::size_t c = 0; c < col_origin.size(); ++c) x[col_origin[c]] = rx[c];
  for (auto it = steps_.rbegin(); it != steps_.rend(); ++it) {
    const Step& s = *it;
    switch (s.kind) {
      case Kind::kFix:
        x[s.col] = data_[s.data + uz(k)];
        break;
      case Kind::kSlack: {
        const double cj = data_[s.data + 2 * uz(k)], b = data_[s.data + 2 * uz(k) + 1];
        double rest = 0.0;
        for (std::size_t q = s.terms_begin; q < s.terms_end; ++q) rest += term_value_[q] * x[term_col_[q]];
        x[s.col] = (b - rest) / s.coef;
        y[s.row] += cj / s.coef;
        break;
      }
      case Kind::kBound: {
        const double cj = data_[s.data + 2 * uz(k)];
        const int from_row = static_cast<int>(data_[s.data + 2 * uz(k) + 1]);
        double d = cj;
        for (Int p = at_.start[uz(s.col)]; p < at_.start[uz(s.col) + 1]; ++p)
          d -= at_.value[uz(p)] * y[at_.index[uz(p)]];
        const bool takes = (d > 0.0 && (from_row & 1)) || (d < 0.0 && (from_row & 2));
        y[s.row] = takes ? d / s.coef : 0.0;
        break;
      }
    }
  }
}

}  // namespace anuk
