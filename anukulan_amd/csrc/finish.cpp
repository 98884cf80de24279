#include "finish.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <stdexcept>
#include <thread>
#include <utility>

namespace anuk {

namespace {

// ---- the logical form [K | -I] ---------------------------------------------

// Every row i gets a logical s_i with K_i x - s_i = q_i: fixed at [0, 0] on
// an equality row, [0, inf) on an inequality row (to_logical_form,
// simplex.hpp:13-36). Columns j < n are K's (K^T's rows), column n + i is
// -e_i.
struct Form {
  Int m = 0, n = 0;
  const Csr* kt = nullptr;  // K^T: row j lists column j of K
  std::vector<double> cost, lower, upper, rhs;

  Int columns() const { return n + m; }
  template <class F>
  void for_column(Int j, F f) const {
    if (j >= n) {
      f(j - n, -1.0);
      return;
    }
    for (Int k = kt->start[uz(j)]; k < kt->start[uz(j) + 1]; ++k) f(kt->index[uz(k)], kt->value[uz(k)]);
  }
  Int column_nnz(Int j) const { return j >= n ? 1 : kt->start[uz(j) + 1] - kt->start[uz(j)]; }
};

Form make_form(const StandardLp& lp) {
  Form f;
  f.m = lp.rows();
  f.n = lp.cols();
  f.kt = &lp.kt;
  f.cost.assign(uz(f.n + f.m), 0.0);
  f.lower.assign(uz(f.n + f.m), 0.0);
  f.upper.assign(uz(f.n + f.m), 0.0);
  std::copy(lp.c.begin(), lp.c.end(), f.cost.begin());
  std::copy(lp.lower.begin(), lp.lower.end(), f.lower.begin());
  std::copy(lp.upper.begin(), lp.upper.end(), f.upper.begin());
  for (Int i = lp.num_equalities; i < f.m; ++i) f.upper[uz(f.n + i)] = kInf;
  f.rhs = lp.q;
  return f;
}

// ---- a sparse LU of the basis with product-form updates -------------------

// Left-looking LU (Gilbert-Peierls) with threshold-free partial pivoting on
// the basis columns taken sparsest first, so the logicals pivot on their own
// rows with no fill. P B Q = L U; L is stored by columns in pivot order with
// original row indices, U by columns with pivot-step indices. A pivot after
// the factorisation multiplies B^-1 by an elementary factor (an eta), kept
// until the next refactorisation.
class BasisLu {
 public:
  // Factorises the columns head[0..m). Positions whose column is dependent
  // on the others are returned in *dependent, and the rows left without a
  // pivot in *free_rows (the same number); the caller puts those rows'
  // logicals in those positions and calls factor again.
  bool factor(const Form& f, const std::vector<Int>& head, std::vector<Int>* dependent,
              std::vector<Int>* free_rows);
  // B z = b: b dense in rows (overwritten), z dense in basis positions.
  void ftran(std::vector<double>& b, std::vector<double>& z) const;
  // B^T y = c: c dense in basis positions (overwritten), y dense in rows.
  void btran(std::vector<double>& c, std::vector<double>& y) const;
  // After a pivot at position r with column w = B^-1 a (dense, positions).
  void update(Int r, const std::vector<double>& w);
  Int updates() const { return static_cast<Int>(eta_r_.size()); }
  // Entries in the eta file against those in the factors: once the updates
  // cost more to apply than the factors, refactorising is cheaper.
  bool etas_outweigh_factors() const {
    return eta_pos_.size() > lrow_.size() + ustep_.size() + static_cast<size_t>(m_);
  }

 private:
  Int m_ = 0;
  std::vector<Int> p_, pinv_, q_;                 // pivot row / step of row / position of step
  std::vector<Int> lstart_, lrow_;
  std::vector<double> lval_;
  std::vector<Int> ustart_, ustep_;
  std::vector<double> uval_, udiag_;
  std::vector<Int> eta_r_, eta_start_{0}, eta_pos_;
  std::vector<double> eta_pivot_, eta_val_;
  // scratch
  std::vector<double> x_;
  std::vector<Int> mark_, stack_, child_, topo_, touched_;
  // Pivot-order work vector of ftran/btran (a solve changes no state).
  mutable std::vector<double> yk_;
};

bool BasisLu::factor(const Form& f, const std::vector<Int>& head, std::vector<Int>* dependent,
                     std::vector<Int>* free_rows) {
  const Int m = f.m;
  m_ = m;
  p_.assign(uz(m), -1);
  pinv_.assign(uz(m), -1);
  q_.assign(uz(m), -1);
  lstart_.assign(1, 0);
  lrow_.clear();
  lval_.clear();
  ustart_.assign(1, 0);
  ustep_.clear();
  uval_.clear();
  udiag_.clear();
  eta_r_.clear();
  eta_start_.assign(1, 0);
  eta_pos_.clear();
  eta_pivot_.clear();
  eta_val_.clear();
  x_.assign(uz(m), 0.0);
  mark_.assign(uz(m), 0);
  child_.assign(uz(m), 0);
  dependent->clear();
  free_rows->clear();

  std::vector<Int> order(uz(m));
  for (Int k = 0; k < m; ++k) order[uz(k)] = k;
  std::stable_sort(order.begin(), order.end(), [&](Int a, Int b) {
    return f.column_nnz(head[uz(a)]) < f.column_nnz(head[uz(b)]);
  });

  Int step = 0;
  int stamp = 0;
  for (Int pos : order) {
    const Int col = head[uz(pos)];
    // Reach of the column's rows through L (pivoted rows only), in
    // topological order: a depth-first search from each entry.
    ++stamp;
    topo_.clear();
    touched_.clear();
    f.for_column(col, [&](Int i, double) {
      if (mark_[uz(i)] == stamp) return;
      if (pinv_[uz(i)] < 0) {
        mark_[uz(i)] = stamp;
        touched_.push_back(i);
        return;
      }
      stack_.assign(1, i);
      mark_[uz(i)] = stamp;
      child_[uz(i)] = lstart_[uz(pinv_[uz(i)])];
      while (!stack_.empty()) {
        const Int r = stack_.back();
        const Int s = pinv_[uz(r)];
        bool descended = false;
        for (Int& e = child_[uz(r)]; e < lstart_[uz(s) + 1]; ++e) {
          const Int c = lrow_[uz(e)];
          if (mark_[uz(c)] == stamp) continue;
          mark_[uz(c)] = stamp;
          if (pinv_[uz(c)] < 0) {
            touched_.push_back(c);
            continue;
          }
          child_[uz(c)] = lstart_[uz(pinv_[uz(c)])];
          stack_.push_back(c);
          descended = true;
          ++e;
          break;
        }
        if (!descended) {
          stack_.pop_back();
          topo_.push_back(r);
        }
      }
    });
    // Numerical solve L x = a over the reach (reverse post-order).
    f.for_column(col, [&](Int i, double v) { x_[uz(i)] += v; });
    for (auto it = topo_.rbegin(); it != topo_.rend(); ++it) {
      const Int r = *it;
      const Int s = pinv_[uz(r)];
      const double v = x_[uz(r)];
      if (v == 0.0) continue;
      for (Int e = lstart_[uz(s)]; e < lstart_[uz(s) + 1]; ++e) x_[uz(lrow_[uz(e)])] -= lval_[uz(e)] * v;
    }
    // Pivot: the largest entry among the rows not yet pivoted.
    Int piv = -1;
    double best = 0.0;
    for (Int r : touched_) {
      const double a = std::fabs(x_[uz(r)]);
      if (a > best || (a == best && piv >= 0 && r < piv)) {
        best = a;
        piv = r;
      }
    }
    double scale = 0.0;
    f.for_column(col, [&](Int, double v) { scale = std::max(scale, std::fabs(v)); });
    if (piv < 0 || best <= 1e-11 * std::max(1.0, scale)) {
      dependent->push_back(pos);
      for (Int r : topo_) x_[uz(r)] = 0.0;
      for (Int r : touched_) x_[uz(r)] = 0.0;
      continue;
    }
    // U column: the pivoted rows' values; L column: the rest over the pivot.
    for (Int r : topo_) {
      if (x_[uz(r)] != 0.0) {
        ustep_.push_back(pinv_[uz(r)]);
        uval_.push_back(x_[uz(r)]);
      }
      x_[uz(r)] = 0.0;
    }
    const double d = x_[uz(piv)];
    for (Int r : touched_) {
      if (r != piv && x_[uz(r)] != 0.0) {
        lrow_.push_back(r);
        lval_.push_back(x_[uz(r)] / d);
      }
      x_[uz(r)] = 0.0;
    }
    udiag_.push_back(d);
    ustart_.push_back(static_cast<Int>(ustep_.size()));
    lstart_.push_back(static_cast<Int>(lrow_.size()));
    p_[uz(step)] = piv;
    pinv_[uz(piv)] = step;
    q_[uz(step)] = pos;
    ++step;
  }
  if (!dependent->empty()) {
    for (Int r = 0; r < m; ++r)
      if (pinv_[uz(r)] < 0) free_rows->push_back(r);
    return false;
  }
  yk_.assign(uz(m), 0.0);
  return true;
}

void BasisLu::ftran(std::vector<double>& b, std::vector<double>& z) const {
  const Int m = m_;
  std::vector<double>& y = yk_;
  for (Int k = 0; k < m; ++k) {
    const double v = b[uz(p_[uz(k)])];
    y[uz(k)] = v;
    if (v == 0.0) continue;
    for (Int e = lstart_[uz(k)]; e < lstart_[uz(k) + 1]; ++e) b[uz(lrow_[uz(e)])] -= lval_[uz(e)] * v;
  }
  z.assign(uz(m), 0.0);
  for (Int k = m - 1; k >= 0; --k) {
    const double zk = y[uz(k)] / udiag_[uz(k)];
    z[uz(q_[uz(k)])] = zk;
    if (zk == 0.0) continue;
    for (Int e = ustart_[uz(k)]; e < ustart_[uz(k) + 1]; ++e) y[uz(ustep_[uz(e)])] -= uval_[uz(e)] * zk;
  }
  for (size_t t = 0; t < eta_r_.size(); ++t) {
    const Int r = eta_r_[t];
    const double zr = z[uz(r)] / eta_pivot_[t];
    z[uz(r)] = zr;
    if (zr == 0.0) continue;
    for (Int e = eta_start_[t]; e < eta_start_[t + 1]; ++e) z[uz(eta_pos_[uz(e)])] -= eta_val_[uz(e)] * zr;
  }
}

void BasisLu::btran(std::vector<double>& c, std::vector<double>& y) const {
  const Int m = m_;
  for (size_t t = eta_r_.size(); t-- > 0;) {
    const Int r = eta_r_[t];
    double s = c[uz(r)];
    for (Int e = eta_start_[t]; e < eta_start_[t + 1]; ++e) s -= eta_val_[uz(e)] * c[uz(eta_pos_[uz(e)])];
    c[uz(r)] = s / eta_pivot_[t];
  }
  std::vector<double>& w = yk_;
  for (Int k = 0; k < m; ++k) {
    double s = c[uz(q_[uz(k)])];
    for (Int e = ustart_[uz(k)]; e < ustart_[uz(k) + 1]; ++e) s -= uval_[uz(e)] * w[uz(ustep_[uz(e)])];
    w[uz(k)] = s / udiag_[uz(k)];
  }
  y.assign(uz(m), 0.0);
  for (Int k = m - 1; k >= 0; --k) {
    double s = w[uz(k)];
    for (Int e = lstart_[uz(k)]; e < lstart_[uz(k) + 1]; ++e) s -= lval_[uz(e)] * w[uz(pinv_[uz(lrow_[uz(e)])])];
    w[uz(k)] = s;
    y[uz(p_[uz(k)])] = s;
  }
}

void BasisLu::update(Int r, const std::vector<double>& w) {
  eta_r_.push_back(r);
  eta_pivot_.push_back(w[uz(r)]);
  for (Int i = 0; i < m_; ++i) {
    if (i != r && w[uz(i)] != 0.0) {
      eta_pos_.push_back(i);
      eta_val_.push_back(w[uz(i)]);
    }
  }
  eta_start_.push_back(static_cast<Int>(eta_pos_.size()));
}

// ---- the bounded primal simplex -------------------------------------------

class Simplex {
 public:
  Simplex(const Form& f, const FinishOptions& o) : f_(f), o_(o) {}

  // Installs a basis (head: column per position; state per column) and
  // factorises it, repairing a singular one with logicals.
  void install(std::vector<Int> head, std::vector<VarState> state);
  int run(FinishResult* out);
  // Crossover: pushes candidate columns into positions held by logicals.
  void crossover(const std::vector<double>& x, const std::vector<double>& rc, bool have_dual,
                 const std::vector<double>& row_slack, CrossoverInfo* info);

  const std::vector<Int>& head() const { return head_; }
  const std::vector<VarState>& state() const { return state_; }
  const std::vector<double>& values() const { return x_; }
  const std::vector<double>& duals() const { return y_; }
  Int refactorizations = 0;

 private:
  void refactor();
  void nonbasic_values();
  void basic_values();
  double infeasibility(Int j) const;
  void compute_duals(bool phase_one);
  double reduced_cost(Int j, bool phase_one) const;
  Int price(bool phase_one, bool bland, double* dq);
  void column(Int j, std::vector<double>& w);

  const Form& f_;
  const FinishOptions& o_;
  BasisLu lu_;
  std::vector<Int> head_, pos_;
  std::vector<VarState> state_;
  std::vector<double> x_, y_, work_, w_;
  Int price_start_ = 0;
};

void Simplex::install(std::vector<Int> head, std::vector<VarState> state) {
  head_ = std::move(head);
  state_ = std::move(state);
  x_.assign(uz(f_.columns()), 0.0);
  refactor();
}

void Simplex::refactor() {
  std::vector<Int> dependent, free_rows;
  ++refactorizations;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (lu_.factor(f_, head_, &dependent, &free_rows)) break;
    if (attempt == 2) throw std::runtime_error("exact finish: the basis cannot be repaired");
    // Dependent columns leave for a bound; the unpivoted rows' logicals
    // take their positions.
    for (size_t k = 0; k < dependent.size(); ++k) {
      const Int posn = dependent[k];
      const Int out = head_[uz(posn)];
      const double lo = f_.lower[uz(out)], hi = f_.upper[uz(out)];
      if (lo <= -kInf && hi >= kInf) state_[uz(out)] = VarState::kFree;
      else if (lo <= -kInf) state_[uz(out)] = VarState::kUpper;
      else if (hi >= kInf) state_[uz(out)] = VarState::kLower;
      else state_[uz(out)] = (x_[uz(out)] - lo > hi - x_[uz(out)]) ? VarState::kUpper : VarState::kLower;
      const Int in = f_.n + free_rows[k];
      head_[uz(posn)] = in;
      state_[uz(in)] = VarState::kBasic;
    }
  }
  pos_.assign(uz(f_.columns()), -1);
  for (Int r = 0; r < f_.m; ++r) pos_[uz(head_[uz(r)])] = r;
  nonbasic_values();
  basic_values();
}

void Simplex::nonbasic_values() {
  for (Int j = 0; j < f_.columns(); ++j) {
    switch (state_[uz(j)]) {
      case VarState::kLower: x_[uz(j)] = f_.lower[uz(j)]; break;
      case VarState::kUpper: x_[uz(j)] = f_.upper[uz(j)]; break;
      case VarState::kFree: x_[uz(j)] = 0.0; break;
      case VarState::kBasic: break;
    }
  }
}

// x_B = B^-1 (q - N x_N)
void Simplex::basic_values() {
  work_ = f_.rhs;
  for (Int j = 0; j < f_.columns(); ++j) {
    if (state_[uz(j)] == VarState::kBasic) continue;
    const double v = x_[uz(j)];
    if (v != 0.0) f_.for_column(j, [&](Int i, double a) { work_[uz(i)] -= a * v; });
  }
  lu_.ftran(work_, w_);
  for (Int r = 0; r < f_.m; ++r) x_[uz(head_[uz(r)])] = w_[uz(r)];
}

double Simplex::infeasibility(Int j) const {
  const double v = x_[uz(j)];
  if (v < f_.lower[uz(j)] - o_.primal_tolerance) return f_.lower[uz(j)] - v;
  if (v > f_.upper[uz(j)] + o_.primal_tolerance) return v - f_.upper[uz(j)];
  return 0.0;
}

// y = B^-T c_B, with the phase-one costs (-1 below the lower bound, +1
// above the upper, 0 inside) when phase_one.
void Simplex::compute_duals(bool phase_one) {
  work_.assign(uz(f_.m), 0.0);
  for (Int r = 0; r < f_.m; ++r) {
    const Int j = head_[uz(r)];
    if (phase_one) {
      const double v = x_[uz(j)];
      work_[uz(r)] = v < f_.lower[uz(j)] - o_.primal_tolerance ? -1.0
                     : v > f_.upper[uz(j)] + o_.primal_tolerance ? 1.0 : 0.0;
    } else {
      work_[uz(r)] = f_.cost[uz(j)];
    }
  }
  lu_.btran(work_, y_);
}

// d_j = c_j - a_j' y
double Simplex::reduced_cost(Int j, bool phase_one) const {
  double s = phase_one ? 0.0 : f_.cost[uz(j)];
  f_.for_column(j, [&](Int i, double a) { s -= a * y_[uz(i)]; });
  return s;
}

// Entering column. Dantzig's largest improving reduced cost within a window
// of columns that rotates through the form (partial pricing: the first
// window with an improving column decides), or under Bland's rule the
// lowest-index improving column over all of them. -1: none improves.
Int Simplex::price(bool phase_one, bool bland, double* dq) {
  const Int N = f_.columns();
  const double td = o_.dual_tolerance;
  const Int window = bland ? N : std::max<Int>(2048, N / 4);
  Int q = -1;
  double best = 0.0;
  Int start = bland ? 0 : price_start_;
  for (Int scanned = 0; scanned < N;) {
    const Int stop = std::min(N, start + std::min(window, N - scanned));
    for (Int j = start; j < stop; ++j) {
      const VarState s = state_[uz(j)];
      if (s == VarState::kBasic || f_.lower[uz(j)] == f_.upper[uz(j)]) continue;
      const double dj = reduced_cost(j, phase_one);
      double gain = 0.0;
      if ((s == VarState::kLower || s == VarState::kFree) && dj < -td) gain = -dj;
      if ((s == VarState::kUpper || s == VarState::kFree) && dj > td) gain = dj;
      if (gain > best) {
        best = gain;
        q = j;
        *dq = dj;
        if (bland) return q;
      }
    }
    scanned += stop - start;
    start = stop == N ? 0 : stop;
    if (q >= 0) break;
  }
  price_start_ = start;
  return q;
}

void Simplex::column(Int j, std::vector<double>& w) {
  work_.assign(uz(f_.m), 0.0);
  f_.for_column(j, [&](Int i, double a) { work_[uz(i)] += a; });
  lu_.ftran(work_, w);
}

void Simplex::crossover(const std::vector<double>& x, const std::vector<double>& rc, bool have_dual,
                        const std::vector<double>& row_slack, CrossoverInfo* info) {
  // Candidates: structurals strictly inside their bounds relative to their
  // range, discounted by how far from zero their reduced cost is
  // (crossover.cpp:12-31,79-101).
  const Int n = f_.n, m = f_.m;
  double rc_scale = 0.0;
  if (have_dual)
    for (Int j = 0; j < n; ++j) rc_scale = std::fmax(rc_scale, std::fabs(rc[uz(j)]));
  std::vector<std::pair<double, Int>> cand;
  for (Int j = 0; j < n; ++j) {
    const double lo = f_.lower[uz(j)], hi = f_.upper[uz(j)], v = x[uz(j)];
    const bool has_lo = lo > -kInf, has_hi = hi < kInf;
    double score;
    if (!has_lo && !has_hi) {
      score = kInf;
    } else {
      double slack = kInf;
      if (has_lo) slack = std::fmin(slack, v - lo);
      if (has_hi) slack = std::fmin(slack, hi - v);
      if (slack <= 0.0) {
        score = 0.0;
      } else {
        const double scale = (has_lo && has_hi) ? std::fmax(1.0, hi - lo) : std::fmax(1.0, std::fabs(v));
        score = slack / scale;
      }
    }
    if (!(score > o_.interior_tolerance)) continue;
    if (o_.dual_weight > 0.0 && have_dual && rc_scale > 0.0)
      score /= 1.0 + o_.dual_weight * (std::fabs(rc[uz(j)]) / rc_scale);
    cand.push_back({score, j});
  }
  std::sort(cand.begin(), cand.end(), [](const std::pair<double, Int>& a, const std::pair<double, Int>& b) {
    if (a.first != b.first) return a.first > b.first;
    return a.second < b.second;
  });
  if (o_.max_candidates >= 0 && static_cast<Int>(cand.size()) > o_.max_candidates)
    cand.resize(uz(o_.max_candidates));
  info->candidates = static_cast<Int>(cand.size());

  // Push them in one at a time, each into the position of a logical, the
  // largest pivot first with rows the point has tight preferred
  // (crossover.cpp:119-170); refactorise every few pushes and keep the last
  // basis that factorised.
  std::vector<Int> good_head = head_;
  std::vector<VarState> good_state = state_;
  Int logicals = 0;
  for (Int r = 0; r < m; ++r) logicals += head_[uz(r)] >= n;
  Int since = 0;
  std::vector<double> alpha;
  for (const auto& c : cand) {
    if (logicals == 0) break;
    const Int col = c.second;
    column(col, alpha);
    Int best_row = -1;
    double best_merit = 0.0;
    for (Int r = 0; r < m; ++r) {
      const Int occ = head_[uz(r)];
      if (occ < n) continue;
      const double piv = std::fabs(alpha[uz(r)]);
      if (!(piv > o_.crossover_pivot_tolerance)) continue;
      const double merit = piv / (1.0 + row_slack[uz(occ - n)]);
      if (merit > best_merit) {
        best_merit = merit;
        best_row = r;
      }
    }
    if (best_row < 0) {
      ++info->rejected_small_pivot;
      continue;
    }
    const Int out = head_[uz(best_row)];
    state_[uz(out)] = VarState::kLower;
    head_[uz(best_row)] = col;
    state_[uz(col)] = VarState::kBasic;
    lu_.update(best_row, alpha);
    --logicals;
    if (o_.crossover_refactor_interval > 0 && ++since >= o_.crossover_refactor_interval) {
      since = 0;
      std::vector<Int> dep, fr;
      ++refactorizations;
      if (lu_.factor(f_, head_, &dep, &fr)) {
        good_head = head_;
        good_state = state_;
      } else {
        head_ = good_head;
        state_ = good_state;
        info->stopped_on_refactorization = true;
        break;
      }
    }
  }
  {
    std::vector<Int> dep, fr;
    ++refactorizations;
    if (!lu_.factor(f_, head_, &dep, &fr)) {
      head_ = good_head;
      state_ = good_state;
      info->stopped_on_refactorization = true;
    }
  }
  info->pushed = 0;
  info->logicals_remaining = 0;
  for (Int r = 0; r < m; ++r) (head_[uz(r)] < n ? info->pushed : info->logicals_remaining)++;

  // Nonbasic columns sit where the point says, or where the reduced cost
  // wants them (crossover.cpp:196-222); nonbasic logicals at zero.
  for (Int j = 0; j < n; ++j) {
    if (state_[uz(j)] == VarState::kBasic) continue;
    const double lo = f_.lower[uz(j)], hi = f_.upper[uz(j)];
    if (lo <= -kInf && hi >= kInf) {
      state_[uz(j)] = VarState::kFree;
    } else if (lo <= -kInf) {
      state_[uz(j)] = VarState::kUpper;
    } else if (hi >= kInf) {
      state_[uz(j)] = VarState::kLower;
    } else {
      bool prefer_upper = (x[uz(j)] - lo) > (hi - x[uz(j)]);
      if (have_dual) {
        if (rc[uz(j)] < 0.0) prefer_upper = true;
        if (rc[uz(j)] > 0.0) prefer_upper = false;
      }
      state_[uz(j)] = prefer_upper ? VarState::kUpper : VarState::kLower;
    }
  }
  for (Int i = 0; i < m; ++i)
    if (state_[uz(n + i)] != VarState::kBasic) state_[uz(n + i)] = VarState::kLower;
  refactor();
}

int Simplex::run(FinishResult* out) {
  const double tp = o_.primal_tolerance;
  Int degenerate = 0;
  bool bland = false;
  Int confirmations = 0;
  std::vector<double> w;
  for (;;) {
    if (out->iterations >= o_.max_iterations) return kFinishIterationLimit;
    if (lu_.updates() >= o_.refactor_interval || lu_.etas_outweigh_factors()) refactor();
    bool phase_one = false;
    for (Int r = 0; r < f_.m && !phase_one; ++r) phase_one = infeasibility(head_[uz(r)]) > 0.0;
    compute_duals(phase_one);
    double dq = 0.0;
    const Int q = price(phase_one, bland, &dq);
    if (q < 0) {
      // No improving column: confirm on fresh factors before stopping.
      if (lu_.updates() > 0 && confirmations < 3) {
        ++confirmations;
        refactor();
        continue;
      }
      if (phase_one) return kFinishInfeasible;
      return kFinishOptimal;
    }
    confirmations = 0;
    const double dir = dq < 0.0 ? 1.0 : -1.0;  // x_q increases when its cost is negative
    column(q, w);
    // Harris ratio test over the basics, then the entering column's own range.
    double theta_max = kInf;
    for (Int r = 0; r < f_.m; ++r) {
      const double rate = -dir * w[uz(r)];
      if (std::fabs(w[uz(r)]) <= o_.pivot_tolerance) continue;
      const Int j = head_[uz(r)];
      const double v = x_[uz(j)], lo = f_.lower[uz(j)], hi = f_.upper[uz(j)];
      double bound;
      if (rate < 0.0) {
        if (phase_one && v < lo - tp) continue;
        bound = (phase_one && v > hi + tp) ? hi : lo;
        if (bound <= -kInf) continue;
        theta_max = std::fmin(theta_max, (v - bound + tp) / -rate);
      } else {
        if (phase_one && v > hi + tp) continue;
        bound = (phase_one && v < lo - tp) ? lo : hi;
        if (bound >= kInf) continue;
        theta_max = std::fmin(theta_max, (bound - v + tp) / rate);
      }
    }
    Int leave = -1;
    double leave_bound = 0.0, theta = kInf, leave_piv = 0.0;
    for (Int r = 0; r < f_.m; ++r) {
      const double rate = -dir * w[uz(r)];
      if (std::fabs(w[uz(r)]) <= o_.pivot_tolerance) continue;
      const Int j = head_[uz(r)];
      const double v = x_[uz(j)], lo = f_.lower[uz(j)], hi = f_.upper[uz(j)];
      double bound, ratio;
      if (rate < 0.0) {
        if (phase_one && v < lo - tp) continue;
        bound = (phase_one && v > hi + tp) ? hi : lo;
        if (bound <= -kInf) continue;
        ratio = (v - bound) / -rate;
      } else {
        if (phase_one && v > hi + tp) continue;
        bound = (phase_one && v < lo - tp) ? lo : hi;
        if (bound >= kInf) continue;
        ratio = (bound - v) / rate;
      }
      if (ratio > theta_max) continue;
      const double piv = std::fabs(w[uz(r)]);
      const bool better = leave < 0 || (bland ? head_[uz(r)] < head_[uz(leave)] : piv > leave_piv);
      if (better) {
        leave = r;
        leave_bound = bound;
        leave_piv = piv;
        theta = std::fmax(0.0, ratio);
      }
    }
    const double range = f_.upper[uz(q)] - f_.lower[uz(q)];
    const bool flip = range < kInf && (leave < 0 || range <= theta);
    if (!flip && leave < 0) {
      if (phase_one) return kFinishNumericalError;  // phase one is bounded below
      return kFinishUnbounded;
    }
    if (flip) theta = range;
    ++out->iterations;
    if (phase_one) ++out->phase_one_iterations;
    if (bland) ++out->bland_iterations;
    if (theta <= 1e-12) {
      if (++degenerate > 200) bland = true;
    } else {
      degenerate = 0;
      bland = false;
    }
    // Move.
    x_[uz(q)] += dir * theta;
    if (theta != 0.0)
      for (Int r = 0; r < f_.m; ++r)
        if (w[uz(r)] != 0.0) x_[uz(head_[uz(r)])] -= dir * theta * w[uz(r)];
    if (flip) {
      state_[uz(q)] = dir > 0.0 ? VarState::kUpper : VarState::kLower;
      x_[uz(q)] = dir > 0.0 ? f_.upper[uz(q)] : f_.lower[uz(q)];
      continue;
    }
    const Int out_col = head_[uz(leave)];
    x_[uz(out_col)] = leave_bound;
    state_[uz(out_col)] = leave_bound == f_.lower[uz(out_col)] ? VarState::kLower : VarState::kUpper;
    pos_[uz(out_col)] = -1;
    head_[uz(leave)] = q;
    pos_[uz(q)] = leave;
    state_[uz(q)] = VarState::kBasic;
    lu_.update(leave, w);
  }
}

double seconds_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

FinishResult exact_finish(const StandardLp& lp, const double* x, const double* y,
                          const FinishOptions& opt, const FinishBasis* warm) {
  const auto start = std::chrono::steady_clock::now();
  FinishResult res;
  const Int n = lp.cols(), m = lp.rows();
  // The equilibrated LP: K~ = D1 K D2, so x~ = x / D2 and y~ = y / D1.
  StandardLp s = lp;
  const Scaling sc = scale_lp(&s, std::max(0, opt.ruiz_iterations), false);
  const Form f = make_form(s);

  std::vector<Int> head(uz(m));
  std::vector<VarState> state(uz(n + m), VarState::kLower);
  for (Int j = 0; j < n + m; ++j) {
    if (f.lower[uz(j)] <= -kInf && f.upper[uz(j)] >= kInf) state[uz(j)] = VarState::kFree;
    else if (f.lower[uz(j)] <= -kInf) state[uz(j)] = VarState::kUpper;
  }
  for (Int i = 0; i < m; ++i) {
    head[uz(i)] = n + i;
    state[uz(n + i)] = VarState::kBasic;
  }
  const bool warm_basis = warm != nullptr && warm->basic.size() == uz(m) &&
                         warm->state.size() == uz(n + m);
  if (warm != nullptr && !warm_basis)
    throw std::invalid_argument("finish: a start basis needs one column per row and one state per column");
  if (warm_basis) {
    head = warm->basic;
    state = warm->state;
    std::vector<char> in_head(uz(n + m), 0);
    for (const Int j : head) {
      if (j < 0 || j >= n + m || in_head[uz(j)])
        throw std::invalid_argument("finish: a start basis names a column twice or out of range");
      in_head[uz(j)] = 1;
    }
    // States follow the head, and a nonbasic state sits on a bound this
    // case has.
    for (Int j = 0; j < n + m; ++j) {
      VarState& v = state[uz(j)];
      if (in_head[uz(j)]) {
        v = VarState::kBasic;
        continue;
      }
      const bool lo = f.lower[uz(j)] > -kInf, hi = f.upper[uz(j)] < kInf;
      if (v == VarState::kBasic || (v == VarState::kLower && !lo) || (v == VarState::kUpper && !hi) ||
          (v == VarState::kFree && (lo || hi)))
        v = lo ? VarState::kLower : (hi ? VarState::kUpper : VarState::kFree);
    }
  }
  Simplex sx(f, opt);
  try {
    sx.install(head, state);
    if (warm_basis) {
      res.started_warm = true;
    } else if (opt.crossover && x != nullptr) {
      std::vector<double> xs(uz(n)), rc, slack(uz(m), 0.0);
      for (Int j = 0; j < n; ++j) xs[uz(j)] = x[uz(j)] / sc.col_scale[uz(j)];
      std::vector<double> kx(uz(m), 0.0);
      s.k.multiply(xs.data(), kx.data());
      for (Int i = 0; i < m; ++i) slack[uz(i)] = std::fabs(kx[uz(i)] - s.q[uz(i)]);
      const bool have_dual = y != nullptr;
      if (have_dual) {
        std::vector<double> ys(uz(m));
        for (Int i = 0; i < m; ++i) ys[uz(i)] = y[uz(i)] / sc.row_scale[uz(i)];
        rc.assign(uz(n), 0.0);
        s.kt.multiply(ys.data(), rc.data());
        for (Int j = 0; j < n; ++j) rc[uz(j)] = s.c[uz(j)] - rc[uz(j)];
      }
      sx.crossover(xs, rc, have_dual, slack, &res.crossover);
      res.started_warm = res.crossover.pushed > 0;
    }
    res.status = sx.run(&res);
  } catch (const std::exception& e) {
    res.status = kFinishNumericalError;
    res.message = e.what();
  }
  res.refactorizations = sx.refactorizations;
  if (!sx.values().empty()) {
    std::vector<double> xs(sx.values().begin(), sx.values().begin() + n);
    sc.unscale_primal(xs, &res.x);
    if (!sx.duals().empty()) sc.unscale_dual(sx.duals(), &res.y);
    else res.y.assign(uz(m), 0.0);
    res.basic = sx.head();
    res.state = sx.state();
    res.objective = lp.model_objective(res.x);
    // The answer judged on the LP as given.
    std::vector<double> kx(uz(m), 0.0), kty(uz(n), 0.0);
    lp.k.multiply(res.x.data(), kx.data());
    lp.kt.multiply(res.y.data(), kty.data());
    double pv = 0.0, dv = 0.0;
    for (Int i = 0; i < m; ++i) {
      const double r = kx[uz(i)] - lp.q[uz(i)];
      pv = std::fmax(pv, i < lp.num_equalities ? std::fabs(r) : std::fmax(0.0, -r));
      if (i >= lp.num_equalities) dv = std::fmax(dv, std::fmax(0.0, -res.y[uz(i)]));
    }
    for (Int j = 0; j < n; ++j) {
      pv = std::fmax(pv, std::fmax(lp.lower[uz(j)] - res.x[uz(j)], res.x[uz(j)] - lp.upper[uz(j)]));
      const double dj = lp.c[uz(j)] - kty[uz(j)];
      const VarState st = res.state[uz(j)];
      if (st == VarState::kBasic || st == VarState::kFree) dv = std::fmax(dv, std::fabs(dj));
      else if (st == VarState::kLower) dv = std::fmax(dv, std::fmax(0.0, -dj));
      else dv = std::fmax(dv, std::fmax(0.0, dj));
    }
    res.primal_violation = pv;
    res.dual_violation = dv;
  }
  if (res.message.empty()) {
    static const char* const names[] = {"optimal", "infeasible", "unbounded", "iteration limit",
                                        "numerical error"};
    res.message = names[res.status];
  }
  res.seconds = seconds_since(start);
  return res;
}

std::vector<FinishResult> exact_finish_many(const std::vector<const StandardLp*>& lps,
                                            const std::vector<const double*>& xs,
                                            const std::vector<const double*>& ys,
                                            const FinishOptions& opt, const FinishBasis* start) {
  const size_t count = lps.size();
  if (xs.size() != count || ys.size() != count) throw std::invalid_argument("finish: one point per LP");
  std::vector<FinishResult> out(count);
  const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
  size_t threads = opt.threads > 0 ? static_cast<size_t>(opt.threads) : std::min<size_t>(hw, 16);
  threads = std::max<size_t>(1, std::min(threads, count));
  std::vector<std::thread> pool;
  std::vector<std::exception_ptr> errors(threads);
  for (size_t t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      try {
        for (size_t i = t; i < count; i += threads)
          out[i] = exact_finish(*lps[i], xs[i], ys[i], opt, start);
      } catch (...) {
        errors[t] = std::current_exception();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (auto& e : errors)
    if (e) std::rethrow_exception(e);
  return out;
}

}  // namespace anuk
