#include "lp.hpp"

#include "host_pool.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace anuk {

std::uint64_t fresh_matrix_id() {
  // Identities start high so they cannot collide with small ids a caller of
  // the C ABI hands in for its own matrices.
  static std::atomic<std::uint64_t> next{std::uint64_t{1} << 48};
  return next.fetch_add(1, std::memory_order_relaxed);
}

void Csr::renew_id() { id = fresh_matrix_id(); }

// Row buckets by a counting pass, then each bucket ordered by column and
// merged: equal columns are summed in bucket order and exact zeros dropped.
// The column order within a bucket uses std::sort with a column-only key, the
// same ordering primitive as sparse.cpp:50-52, so duplicate sums associate
// the same way.
Csr Csr::from_triplets(Int rows, Int cols, std::vector<Entry> entries) {
  Csr a;
  a.rows = rows;
  a.cols = cols;
  std::vector<Int> offset(uz(rows) + 1, 0);
  for (const Entry& e : entries) {
    if (e.row < 0 || e.row >= rows || e.col < 0 || e.col >= cols)
      throw std::invalid_argument("triplet index out of range");
    ++offset[uz(e.row) + 1];
  }
  for (std::size_t r = 1; r < offset.size(); ++r) offset[r] += offset[r - 1];
  std::vector<Entry> bucketed(entries.size());
  {
    std::vector<Int> fill(offset.begin(), offset.end() - 1);
    for (const Entry& e : entries) bucketed[uz(fill[uz(e.row)]++)] = e;
  }
  a.start.assign(uz(rows) + 1, 0);
  a.index.reserve(entries.size());
  a.value.reserve(entries.size());
  for (Int r = 0; r < rows; ++r) {
    auto first = bucketed.begin() + offset[uz(r)];
    auto last = bucketed.begin() + offset[uz(r) + 1];
    std::sort(first, last, [](const Entry& p, const Entry& q) { return p.col < q.col; });
    for (auto it = first; it != last;) {
      const Int col = it->col;
      double total = 0.0;
      for (; it != last && it->col == col; ++it) total += it->value;
      if (total != 0.0) {
        a.index.push_back(col);
        a.value.push_back(total);
      }
    }
    a.start[uz(r) + 1] = static_cast<Int>(a.value.size());
  }
  a.renew_id();
  return a;
}

Csr Csr::from_csr(Int rows, Int cols, const Int* start, const Int* index,
                  const double* value) {
  Csr a;
  a.rows = rows;
  a.cols = cols;
  a.start.assign(start, start + rows + 1);
  const Int nnz = a.start.back();
  a.index.assign(index, index + nnz);
  a.value.assign(value, value + nnz);
  std::string err;
  if (!a.validate(&err)) throw std::invalid_argument("invalid CSR: " + err);
  a.renew_id();
  return a;
}

// Counting sort by column; visiting source rows in increasing order leaves
// every transposed row already sorted.
Csr Csr::transpose() const {
  Csr t;
  t.rows = cols;
  t.cols = rows;
  t.start.assign(uz(cols) + 1, 0);
  t.index.resize(value.size());
  t.value.resize(value.size());
  for (const Int c : index) ++t.start[uz(c) + 1];
  for (std::size_t c = 1; c < t.start.size(); ++c) t.start[c] += t.start[c - 1];
  std::vector<Int> next(t.start.begin(), t.start.end() - 1);
  for (Int r = 0; r < rows; ++r) {
    for (Int k = start[uz(r)]; k < start[uz(r) + 1]; ++k) {
      const Int slot = next[uz(index[uz(k)])]++;
      t.index[uz(slot)] = r;
      t.value[uz(slot)] = value[uz(k)];
    }
  }
  t.renew_id();
  return t;
}

// Each row's sum is serial in storage order whichever thread computes it, so
// splitting the rows over threads leaves every bit as it is.
void Csr::multiply(const double* x, double* y) const {
  auto rows_in = [&](Int r0, Int r1) {
    for (Int r = r0; r < r1; ++r) {
      double s = 0.0;
      for (Int k = start[uz(r)]; k < start[uz(r) + 1]; ++k) s += value[uz(k)] * x[index[uz(k)]];
      y[r] = s;
    }
  };
  HostPool& pool = HostPool::shared();
  const Int tasks = static_cast<Int>(4 * pool.threads());
  if (nnz() < (Int{1} << 20) || pool.threads() < 2 || rows < 2 * tasks) {
    rows_in(0, rows);
    return;
  }
  // Row ranges of about equal nonzero counts on the host pool; every row's
  // sum is still serial in storage order.
  std::vector<Int> cut(uz(tasks) + 1, rows);
  cut[0] = 0;
  const double per = static_cast<double>(nnz()) / tasks;
  Int r = 0;
  for (Int t = 1; t < tasks; ++t) {
    const double target = per * t;
    while (r < rows && static_cast<double>(start[uz(r)]) < target) ++r;
    cut[uz(t)] = r;
  }
  pool.run(uz(tasks), [&](std::size_t t) { rows_in(cut[t], cut[t + 1]); });
}

bool Csr::validate(std::string* error) const {
  auto bad = [&](const char* why) {
    if (error) *error = why;
    return false;
  };
  if (rows < 0 || cols < 0) return bad("negative dimensions");
  if (start.size() != uz(rows) + 1) return bad("row start array has the wrong length");
  if (start.front() != 0) return bad("row starts do not begin at 0");
  if (static_cast<std::size_t>(start.back()) != value.size()) return bad("row starts do not end at nnz");
  if (index.size() != value.size()) return bad("index and value lengths differ");
  for (Int r = 0; r < rows; ++r) {
    if (start[uz(r) + 1] < start[uz(r)]) return bad("row starts are not monotone");
    for (Int k = start[uz(r)]; k < start[uz(r) + 1]; ++k) {
      const Int c = index[uz(k)];
      if (c < 0 || c >= cols) return bad("column index out of range");
      if (k > start[uz(r)] && index[uz(k) - 1] >= c) return bad("columns not strictly increasing");
    }
  }
  return true;
}

double StandardLp::standard_objective(const std::vector<double>& x) const {
  double s = 0.0;
  for (std::size_t j = 0; j < c.size(); ++j) s += c[j] * x[j];
  return s;
}

double StandardLp::model_objective(const std::vector<double>& x) const {
  return objective_scale * standard_objective(x) + objective_offset;
}

StandardLp to_standard_form(const Model& model) {
  StandardLp lp;
  const Int n = model.cols;
  const Int m = model.rows;
  auto clip = [](double v) { return v >= kFileInf ? kInf : (v <= -kFileInf ? -kInf : v); };

  lp.objective_scale = model.maximize ? -1.0 : 1.0;
  lp.objective_offset = model.objective_offset;
  lp.c.resize(uz(n));
  lp.lower.resize(uz(n));
  lp.upper.resize(uz(n));
  for (Int j = 0; j < n; ++j) {
    lp.c[uz(j)] = lp.objective_scale * model.objective[uz(j)];
    lp.lower[uz(j)] = clip(model.col_lower[uz(j)]);
    lp.upper[uz(j)] = clip(model.col_upper[uz(j)]);
  }

  const Csr a = Csr::from_triplets(m, n, model.entries);

  // One emitted standard row: which model row, its sign and its rhs.
  struct Out {
    Int model_row;
    double sign;
    double rhs;
  };
  std::vector<Out> eq, ineq;
  for (Int i = 0; i < m; ++i) {
    const double lo = clip(model.row_lower[uz(i)]);
    const double hi = clip(model.row_upper[uz(i)]);
    const bool has_lo = lo > -kInf, has_hi = hi < kInf;
    if (has_lo && has_hi && lo == hi) {
      eq.push_back({i, 1.0, lo});
    } else if (has_lo && has_hi) {
      ineq.push_back({i, 1.0, lo});
      ineq.push_back({i, -1.0, -hi});
    } else if (has_lo) {
      ineq.push_back({i, 1.0, lo});
    } else if (has_hi) {
      ineq.push_back({i, -1.0, -hi});
    }
  }
  lp.num_equalities = static_cast<Int>(eq.size());
  const Int total = static_cast<Int>(eq.size() + ineq.size());
  lp.q.resize(uz(total));
  lp.row_model.resize(uz(total));
  lp.row_sign.resize(uz(total));
  std::vector<Entry> entries;
  entries.reserve(static_cast<std::size_t>(a.nnz()) + 2 * ineq.size());
  Int dest = 0;
  for (const std::vector<Out>* block : {&eq, &ineq}) {
    for (const Out& o : *block) {
      lp.q[uz(dest)] = o.rhs;
      lp.row_model[uz(dest)] = o.model_row;
      lp.row_sign[uz(dest)] = o.sign;
      for (Int k = a.start[uz(o.model_row)]; k < a.start[uz(o.model_row) + 1]; ++k)
        entries.push_back({dest, a.index[uz(k)], o.sign * a.value[uz(k)]});
      ++dest;
    }
  }
  lp.k = Csr::from_triplets(total, n, std::move(entries));
  lp.kt = lp.k.transpose();
  for (Int j = 0; j < n; ++j)
    if (lp.lower[uz(j)] > lp.upper[uz(j)])
      throw std::invalid_argument("empty bound interval at column " + std::to_string(j));
  for (const double v : lp.q)
    if (!std::isfinite(v)) throw std::invalid_argument("right-hand side is not finite");
  return lp;
}

void Scaling::unscale_primal(const std::vector<double>& scaled, std::vector<double>* out) const {
  out->resize(col_scale.size());
  for (std::size_t j = 0; j < col_scale.size(); ++j) (*out)[j] = col_scale[j] * scaled[j];
}

void Scaling::unscale_dual(const std::vector<double>& scaled, std::vector<double>* out) const {
  out->resize(row_scale.size());
  for (std::size_t i = 0; i < row_scale.size(); ++i) (*out)[i] = row_scale[i] * scaled[i];
}

namespace {

// One equilibration sweep of K. Row and column norms are the infinity norm
// (one_norm == false) or the 1-norm, accumulated in storage order; each row
// and column is divided by the square root of its norm, empty ones left alone.
Sweep sweep_matrix(Csr& k, bool one_norm) {
  std::vector<double> rn(uz(k.rows), 0.0), cn(uz(k.cols), 0.0);
  auto fold = [one_norm](double a, double v) { return one_norm ? a + v : std::fmax(a, v); };
  for (Int r = 0; r < k.rows; ++r) {
    double a = 0.0;
    for (Int p = k.start[uz(r)]; p < k.start[uz(r) + 1]; ++p) a = fold(a, std::fabs(k.value[uz(p)]));
    rn[uz(r)] = a;
  }
  for (std::size_t p = 0; p < k.value.size(); ++p) {
    double& a = cn[uz(k.index[p])];
    a = fold(a, std::fabs(k.value[p]));
  }
  for (double& v : rn) v = v > 0.0 ? 1.0 / std::sqrt(v) : 1.0;
  for (double& v : cn) v = v > 0.0 ? 1.0 / std::sqrt(v) : 1.0;
  // (a * d1) * d2, rounded after each product, as two separate passes would.
  for (Int r = 0; r < k.rows; ++r)
    for (Int p = k.start[uz(r)]; p < k.start[uz(r) + 1]; ++p)
      k.value[uz(p)] = k.value[uz(p)] * rn[uz(r)] * cn[uz(k.index[uz(p)])];
  return Sweep{std::move(rn), std::move(cn)};
}

}  // namespace

std::vector<Sweep> scale_matrix(Csr* k, int ruiz_iterations, bool pock_chambolle) {
  std::vector<Sweep> sweeps;
  for (int i = 0; i < ruiz_iterations; ++i) sweeps.push_back(sweep_matrix(*k, false));
  if (pock_chambolle) sweeps.push_back(sweep_matrix(*k, true));
  k->renew_id();
  return sweeps;
}

void apply_sweeps(const std::vector<Sweep>& sweeps, std::vector<double>* q, std::vector<double>* c,
                  std::vector<double>* lower, std::vector<double>* upper, Scaling* acc) {
  acc->row_scale.assign(q->size(), 1.0);
  acc->col_scale.assign(c->size(), 1.0);
  for (const Sweep& s : sweeps) {
    for (std::size_t i = 0; i < s.row.size(); ++i) {
      (*q)[i] *= s.row[i];
      acc->row_scale[i] *= s.row[i];
    }
    // x = D2 x~: a smaller d2 makes the scaled variable larger, so the bounds
    // divide; infinite bounds stay infinite.
    for (std::size_t j = 0; j < s.col.size(); ++j) {
      (*c)[j] *= s.col[j];
      (*lower)[j] /= s.col[j];
      (*upper)[j] /= s.col[j];
      acc->col_scale[j] *= s.col[j];
    }
  }
}

Scaling scale_lp(StandardLp* lp, int ruiz_iterations, bool pock_chambolle) {
  Scaling s;
  const std::vector<Sweep> sweeps = scale_matrix(&lp->k, ruiz_iterations, pock_chambolle);
  apply_sweeps(sweeps, &lp->q, &lp->c, &lp->lower, &lp->upper, &s);
  lp->kt = lp->k.transpose();
  return s;
}

double host_dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double host_two_norm(const std::vector<double>& a) { return std::sqrt(host_dot(a, a)); }
double host_inf_norm(const std::vector<double>& a) {
  double m = 0.0;
  for (const double v : a) m = std::fmax(m, std::fabs(v));
  return m;
}

bool Residual::primal_ok(double tol, double abs_tol, bool inf_norm) const {
  bool ok = relative_primal <= tol;
  if (inf_norm) ok = ok && relative_primal_inf <= tol;
  if (abs_tol > 0.0) ok = ok && primal_residual_inf <= abs_tol;
  return ok;
}

bool Residual::dual_ok(double tol, double abs_tol, bool inf_norm) const {
  bool ok = relative_dual <= tol;
  if (inf_norm) ok = ok && relative_dual_inf <= tol;
  if (abs_tol > 0.0) ok = ok && dual_residual_inf <= abs_tol;
  return ok;
}

bool Residual::converged(double tol, double abs_tol, bool inf_norm, double gap_tol) const {
  return primal_ok(tol, abs_tol, inf_norm) && dual_ok(tol, abs_tol, inf_norm) &&
         gap_ok(gap_tol > 0.0 ? gap_tol : tol);
}

ResidualScales::ResidualScales(const StandardLp& lp)
    : q_two(host_two_norm(lp.q)),
      c_two(host_two_norm(lp.c)),
      q_inf(host_inf_norm(lp.q)),
      c_inf(host_inf_norm(lp.c)) {}

Residual residual_from_terms(const Terms& t, const ResidualScales& s) {
  Residual r;
  r.primal_residual = std::sqrt(t.primal_sq);
  r.primal_residual_inf = t.primal_inf;
  r.dual_residual = std::sqrt(t.dual_sq);
  r.dual_residual_inf = t.dual_inf;
  r.primal_objective = t.c_dot_x;
  r.dual_objective = t.q_dot_y + t.bound_term;
  r.absolute_gap = std::fabs(r.primal_objective - r.dual_objective);
  r.relative_primal = r.primal_residual / (1.0 + s.q_two);
  r.relative_dual = r.dual_residual / (1.0 + s.c_two);
  r.relative_primal_inf = r.primal_residual_inf / (1.0 + s.q_inf);
  r.relative_dual_inf = r.dual_residual_inf / (1.0 + s.c_inf);
  r.relative_gap =
      r.absolute_gap / (1.0 + std::fabs(r.primal_objective) + std::fabs(r.dual_objective));
  return r;
}

// Host residual of an unscaled pair: rows violated (equalities either way,
// >= rows only below), leftover reduced costs per bound class, and the bound
// contribution to the dual objective taken at whichever end the sign favours.
Residual evaluate_residual(const StandardLp& lp, const std::vector<double>& x,
                           const std::vector<double>& y) {
  Residual r;
  std::vector<double> kx(uz(lp.rows()));
  lp.k.multiply(x.data(), kx.data());
  double psq = 0.0, pinf = 0.0;
  for (Int i = 0; i < lp.rows(); ++i) {
    const double slack = kx[uz(i)] - lp.q[uz(i)];
    const double v = i < lp.num_equalities ? slack : std::fmin(slack, 0.0);
    psq += v * v;
    pinf = std::fmax(pinf, std::fabs(v));
  }
  r.primal_residual = std::sqrt(psq);
  r.primal_residual_inf = pinf;

  std::vector<double> kty(uz(lp.cols()));
  lp.kt.multiply(y.data(), kty.data());
  double dsq = 0.0, bound = 0.0;
  for (Int j = 0; j < lp.cols(); ++j) {
    const double lambda = lp.c[uz(j)] - kty[uz(j)];
    const bool lo = lp.lower[uz(j)] > -kInf, hi = lp.upper[uz(j)] < kInf;
    const double left = (lo && hi) ? 0.0 : lo ? std::fmin(lambda, 0.0) : hi ? std::fmax(lambda, 0.0) : lambda;
    dsq += left * left;
    r.dual_residual_inf = std::fmax(r.dual_residual_inf, std::fabs(left));
    const double absorbed = lambda - left;
    if (absorbed > 0.0) {
      bound += absorbed * lp.lower[uz(j)];
    } else if (absorbed < 0.0) {
      bound += absorbed * lp.upper[uz(j)];
    }
  }
  r.dual_residual = std::sqrt(dsq);
  r.primal_objective = lp.standard_objective(x);
  r.dual_objective = host_dot(lp.q, y) + bound;
  r.absolute_gap = std::fabs(r.primal_objective - r.dual_objective);
  r.relative_primal = r.primal_residual / (1.0 + host_two_norm(lp.q));
  r.relative_dual = r.dual_residual / (1.0 + host_two_norm(lp.c));
  r.relative_primal_inf = r.primal_residual_inf / (1.0 + host_inf_norm(lp.q));
  r.relative_dual_inf = r.dual_residual_inf / (1.0 + host_inf_norm(lp.c));
  r.relative_gap =
      r.absolute_gap / (1.0 + std::fabs(r.primal_objective) + std::fabs(r.dual_objective));
  return r;
}

}  // namespace anuk
