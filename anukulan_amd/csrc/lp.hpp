// Host-side LP layout: CSR matrices, the standard form PDHG iterates on, its
// diagonal scaling and the residual it is judged by.
//
// Semantics follow the reference exactly, because the device kernels assume
// the same row order and signs and because parity of the solve depends on
// the scaled problem being the same to the bit:
//   Csr::from_triplets   sparse.cpp:24-70  (sum duplicates, drop exact zeros,
//                                           columns sorted within each row)
//   Csr::transpose       sparse.cpp:93-118 (counting sort, rows ascending)
//   to_standard_form     standard_form.cpp:72-165
//   scale_lp             scaling.cpp:32-100 (Ruiz inf-norm x N, then one
//                                            Pock-Chambolle 1-norm sweep)
//   evaluate_residual    pdhg.cpp:107-174
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace anuk {

using Int = std::int32_t;
inline constexpr double kInf = std::numeric_limits<double>::infinity();
inline constexpr double kFileInf = 1e30;

inline std::size_t uz(Int i) { return static_cast<std::size_t>(i); }

struct Entry {
  Int row;
  Int col;
  double value;
};

// Compressed sparse rows with int32 indices and fp64 values, sorted and unique
// columns within each row. `id` names the contents for the device cache: a new
// value is drawn whenever a matrix is built or its values change.
struct Csr {
  Int rows = 0;
  Int cols = 0;
  std::vector<Int> start{0};
  std::vector<Int> index;
  std::vector<double> value;
  std::uint64_t id = 0;

  Int nnz() const { return static_cast<Int>(value.size()); }

  static Csr from_triplets(Int rows, Int cols, std::vector<Entry> entries);
  static Csr from_csr(Int rows, Int cols, const Int* start, const Int* index,
                      const double* value);
  Csr transpose() const;
  // y = A x, serial row sums in storage order.
  void multiply(const double* x, double* y) const;
  bool validate(std::string* error) const;
  void renew_id();
};

std::uint64_t fresh_matrix_id();

// min c'x  s.t.  K[0:neq) x = q[0:neq),  K[neq:) x >= q[neq:),  lower <= x <= upper
struct StandardLp {
  Csr k;
  Csr kt;
  std::vector<double> c;
  std::vector<double> q;
  std::vector<double> lower;
  std::vector<double> upper;
  Int num_equalities = 0;
  double objective_scale = 1.0;
  double objective_offset = 0.0;
  std::vector<Int> row_model;
  std::vector<double> row_sign;

  Int rows() const { return k.rows; }
  Int cols() const { return k.cols; }
  double standard_objective(const std::vector<double>& x) const;
  double model_objective(const std::vector<double>& x) const;
};

struct Model {
  Int rows = 0;
  Int cols = 0;
  std::vector<Entry> entries;
  std::vector<double> objective;
  std::vector<double> row_lower, row_upper, col_lower, col_upper;
  bool maximize = false;
  double objective_offset = 0.0;
};

StandardLp to_standard_form(const Model& model);

struct Scaling {
  std::vector<double> row_scale;  // D1
  std::vector<double> col_scale;  // D2
  void unscale_primal(const std::vector<double>& scaled, std::vector<double>* out) const;
  void unscale_dual(const std::vector<double>& scaled, std::vector<double>* out) const;
};

// Scales lp in place (K~ = D1 K D2, q~ = D1 q, c~ = D2 c, bounds / D2) and
// rebuilds K^T. Returns the accumulated diagonals.
Scaling scale_lp(StandardLp* lp, int ruiz_iterations, bool pock_chambolle);

// The same in two halves, for LPs that share K: scale_matrix equilibrates K
// in place and returns each sweep's row and column factors; apply_sweeps
// then applies them, sweep by sweep, to one LP's q, c and bounds. Together
// they give exactly scale_lp's bits (the vector updates never feed back into
// the matrix norms).
struct Sweep {
  std::vector<double> row;  // d1 of the sweep
  std::vector<double> col;  // d2 of the sweep
};
std::vector<Sweep> scale_matrix(Csr* k, int ruiz_iterations, bool pock_chambolle);
void apply_sweeps(const std::vector<Sweep>& sweeps, std::vector<double>* q,
                  std::vector<double>* c, std::vector<double>* lower, std::vector<double>* upper,
                  Scaling* acc);

struct Residual {
  double primal_residual = 0.0;
  double dual_residual = 0.0;
  double primal_residual_inf = 0.0;
  double dual_residual_inf = 0.0;
  double primal_objective = 0.0;
  double dual_objective = 0.0;
  double absolute_gap = 0.0;
  double relative_primal = 0.0;
  double relative_dual = 0.0;
  double relative_gap = 0.0;
  double relative_primal_inf = 0.0;
  double relative_dual_inf = 0.0;

  bool primal_ok(double tol, double abs_tol, bool inf_norm) const;
  bool dual_ok(double tol, double abs_tol, bool inf_norm) const;
  bool gap_ok(double tol) const { return relative_gap <= tol; }
  bool converged(double tol, double abs_tol, bool inf_norm, double gap_tol) const;
};

// The seven convergence numbers (backend.hpp:168-176).
struct Terms {
  double primal_sq = 0.0;
  double primal_inf = 0.0;
  double q_dot_y = 0.0;
  double dual_sq = 0.0;
  double dual_inf = 0.0;
  double bound_term = 0.0;
  double c_dot_x = 0.0;
};

// The four data norms every relative measure divides by (pdhg.cpp:55-66).
struct ResidualScales {
  double q_two = 0.0, c_two = 0.0, q_inf = 0.0, c_inf = 0.0;
  ResidualScales() = default;
  explicit ResidualScales(const StandardLp& lp);
};

Residual residual_from_terms(const Terms& t, const ResidualScales& s);
Residual evaluate_residual(const StandardLp& lp, const std::vector<double>& x,
                           const std::vector<double>& y);

double host_dot(const std::vector<double>& a, const std::vector<double>& b);
double host_two_norm(const std::vector<double>& a);
double host_inf_norm(const std::vector<double>& a);

}  // namespace anuk
