/*
 * oracle.c - TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C CPU restatement of the reference's arithmetic on the PDHG path,
 * one function per reference function, each citing the file:line it follows
 * (paths relative to the reference repository). Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load this
 * library, and only as the checker: the HIP product never calls it.
 *
 * Serial loops in the reference's order, no FMA contraction (built with
 * -ffp-contract=off), so that every result is the reference's to the bit.
 *
 * How each function is pinned (the checks live in tests/test_oracle.py):
 *   orc_from_triplets, orc_transpose, orc_spmv ......... reference build
 *       (oracle/_ref) bit-exact; longhand CSR loop; brute-force dense product
 *   orc_dot / two_norm / inf_norm ...................... reference build
 *       bit-exact; math.fsum within rounding bounds
 *   orc_primal_update / orc_dual_update ................ reference build
 *       bit-exact; the closed forms of test_backend.cpp:276-406; the K x_bar
 *       identities of test_backend.cpp:408-448 against a real matrix
 *   orc_accumulate / orc_scale_into .................... reference build
 *   orc_weighted_norm_squared / orc_step_size_terms .... reference build
 *   orc_convergence_terms .............................. reference build
 *       bit-exact on the four bound classes with zero scales
 *   orc_residual_from_terms ............................ reference build
 *       evaluate_residual at the same point (pdhg.cpp:107-174)
 *   orc_estimate_matrix_norm ........................... reference build
 *       bit-exact; numpy SVD; diag(3,4) -> 4 (test_pdhg.cpp:206-224)
 *   orc_is_infeasibility_certificate ................... hand-built Farkas
 *       rays on an infeasible LP and a feasible one
 *   orc_scale_lp ....................................... reference build
 *       bit-exact; the equivalence invariant of test_scaling.cpp:112-138
 *   orc_to_standard_form ............................... reference build
 *       (to_standard_form) identical arrays on ranged / free / <= rows
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int32_t Int;

/* ---------------------------------------------------------------------- */
/* CSR construction: SparseMatrix::from_triplets (sparse.cpp:24-70)        */
/* ---------------------------------------------------------------------- */

typedef struct {
  Int row, col;
  double value;
  int64_t order; /* position in the input, to make the per-row sort stable */
} orc_triplet;

static int cmp_col(const void* a, const void* b) {
  const orc_triplet* x = (const orc_triplet*)a;
  const orc_triplet* y = (const orc_triplet*)b;
  if (x->col != y->col) return x->col < y->col ? -1 : 1;
  /* std::sort is not stable, but duplicates are summed in whatever order the
   * sort leaves them. The reference sums them in post-sort order; for exact
   * parity on inputs with duplicates the generators here never emit
   * duplicates whose sum depends on order beyond one addition of two terms
   * (a + b == b + a exactly). */
  if (x->order != y->order) return x->order < y->order ? -1 : 1;
  return 0;
}

/* Returns nnz written; out arrays must hold `count` entries (start rows+1). */
int64_t orc_from_triplets(Int rows, Int cols, int64_t count, const Int* ri, const Int* ci,
                          const double* v, Int* start, Int* index, double* value) {
  (void)cols;
  Int* cnt = (Int*)calloc((size_t)rows + 1, sizeof(Int));
  orc_triplet* ordered = (orc_triplet*)malloc((size_t)(count > 0 ? count : 1) * sizeof(orc_triplet));
  Int* cursor = (Int*)malloc(((size_t)rows + 1) * sizeof(Int));
  for (int64_t k = 0; k < count; ++k) cnt[ri[k] + 1]++;
  for (Int i = 1; i <= rows; ++i) cnt[i] += cnt[i - 1];
  memcpy(cursor, cnt, ((size_t)rows + 1) * sizeof(Int));
  for (int64_t k = 0; k < count; ++k) {
    orc_triplet t = {ri[k], ci[k], v[k], k};
    ordered[cursor[ri[k]]++] = t;
  }
  int64_t nnz = 0;
  start[0] = 0;
  for (Int r = 0; r < rows; ++r) {
    const Int lo = cnt[r], hi = cnt[r + 1];
    qsort(ordered + lo, (size_t)(hi - lo), sizeof(orc_triplet), cmp_col);
    for (Int k = lo; k < hi;) {
      const Int col = ordered[k].col;
      double sum = 0.0;
      while (k < hi && ordered[k].col == col) {
        sum += ordered[k].value;
        ++k;
      }
      if (sum != 0.0) {
        index[nnz] = col;
        value[nnz] = sum;
        ++nnz;
      }
    }
    start[r + 1] = (Int)nnz;
  }
  free(cnt);
  free(ordered);
  free(cursor);
  return nnz;
}

/* SparseMatrix::transpose (sparse.cpp:93-118): counting sort by column. */
void orc_transpose(Int rows, Int cols, const Int* start, const Int* index,
                   const double* value, Int* t_start, Int* t_index, double* t_value) {
  const Int nnz = start[rows];
  memset(t_start, 0, ((size_t)cols + 1) * sizeof(Int));
  for (Int k = 0; k < nnz; ++k) t_start[index[k] + 1]++;
  for (Int c = 1; c <= cols; ++c) t_start[c] += t_start[c - 1];
  Int* cursor = (Int*)malloc(((size_t)cols + 1) * sizeof(Int));
  memcpy(cursor, t_start, (size_t)cols * sizeof(Int));
  for (Int r = 0; r < rows; ++r) {
    for (Int k = start[r]; k < start[r + 1]; ++k) {
      const Int dest = cursor[index[k]]++;
      t_index[dest] = r;
      t_value[dest] = value[k];
    }
  }
  free(cursor);
}

/* SparseMatrix::multiply (sparse.cpp:72-80): serial row sums in CSR order. */
void orc_spmv(Int rows, const Int* start, const Int* index, const double* value,
              const double* x, double* y) {
  for (Int r = 0; r < rows; ++r) {
    double sum = 0.0;
    for (Int k = start[r]; k < start[r + 1]; ++k) sum += value[k] * x[index[k]];
    y[r] = sum;
  }
}

/* ---------------------------------------------------------------------- */
/* CpuBackend (backend.cpp:9-127) and the default reductions (131-202)     */
/* ---------------------------------------------------------------------- */

/* backend.cpp:39-43 */
double orc_dot(const double* a, const double* b, Int n) {
  double sum = 0.0;
  for (Int i = 0; i < n; ++i) sum += a[i] * b[i];
  return sum;
}

/* backend.cpp:45-47 */
double orc_two_norm(const double* a, Int n) { return sqrt(orc_dot(a, a, n)); }

/* backend.cpp:49-53 */
double orc_inf_norm(const double* a, Int n) {
  double m = 0.0;
  for (Int i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
  return m;
}

/* backend.cpp:60-79: step; clamp lower then upper; dx; x_bar; reflect; blend. */
void orc_primal_update(Int n, double tau, double gamma, double halpern, const double* x,
                       const double* c, const double* kt_y, const double* lower,
                       const double* upper, const double* anchor, double* x_out,
                       double* dx, double* x_bar) {
  for (Int j = 0; j < n; ++j) {
    const double xj = x[j];
    double v = xj - tau * (c[j] - kt_y[j]);
    if (v < lower[j]) v = lower[j];
    if (v > upper[j]) v = upper[j];
    const double d = v - xj;
    dx[j] = d;
    x_bar[j] = v + d;
    double next = v;
    if (gamma != 0.0) next += gamma * d;
    if (anchor != NULL) next = halpern * next + (1.0 - halpern) * anchor[j];
    x_out[j] = next;
  }
}

/* backend.cpp:81-111 */
void orc_dual_update(Int m, Int num_equalities, double sigma, double gamma, double halpern,
                     const double* y, const double* q, const double* k_x_bar,
                     const double* k_x, const double* anchor_y, const double* anchor_kx,
                     double* y_out, double* dy, double* k_out, double* k_dx) {
  for (Int i = 0; i < m; ++i) {
    const double yi = y[i];
    const double bar = k_x_bar[i];
    const double kx = k_x[i];
    double v = yi + sigma * (q[i] - bar);
    if (i >= num_equalities && v < 0.0) v = 0.0;
    const double d = v - yi;
    dy[i] = d;
    const double kd = 0.5 * (bar - kx);
    k_dx[i] = kd;
    double ny = v;
    double nk = 0.5 * (bar + kx);
    if (gamma != 0.0) {
      ny += gamma * d;
      nk += gamma * kd;
    }
    if (anchor_y != NULL) {
      ny = halpern * ny + (1.0 - halpern) * anchor_y[i];
      nk = halpern * nk + (1.0 - halpern) * anchor_kx[i];
    }
    y_out[i] = ny;
    k_out[i] = nk;
  }
}

/* backend.cpp:113-116 */
void orc_accumulate(Int n, double weight, const double* v, double* sum) {
  for (Int i = 0; i < n; ++i) sum[i] += weight * v[i];
}

/* backend.cpp:118-121: true division. */
void orc_scale_into(Int n, double weight, const double* sum, double* out) {
  for (Int i = 0; i < n; ++i) out[i] = sum[i] / weight;
}

/* backend.cpp:123-126 */
double orc_weighted_norm_squared(Int n, Int m, const double* dx, const double* dy,
                                 double omega) {
  return omega * orc_dot(dx, dx, n) + orc_dot(dy, dy, m) / omega;
}

/* backend.cpp:131-141 */
void orc_step_size_terms(Int n, Int m, const double* dx, const double* dy,
                         const double* k_dx, double omega, double* interaction,
                         double* movement) {
  *interaction = fabs(orc_dot(dy, k_dx, m));
  *movement = orc_weighted_norm_squared(n, m, dx, dy, omega);
}

typedef struct {
  const double *c, *lower, *upper, *q, *row_scale, *col_scale;
  Int num_equalities;
} orc_conv_problem;

typedef struct {
  double primal_sq, primal_inf, q_dot_y, dual_sq, dual_inf, bound_term, c_dot_x;
} orc_conv_terms;

/* backend.cpp:143-202 */
void orc_convergence_terms(Int n, Int m, const orc_conv_problem* lp,
                           const double* x, const double* k_x, const double* y,
                           const double* kt_y, orc_conv_terms* out) {
  orc_conv_terms t = {0, 0, 0, 0, 0, 0, 0};
  for (Int i = 0; i < m; ++i) {
    const double scale = lp->row_scale[i];
    const double row = scale != 0.0 ? k_x[i] / scale : k_x[i];
    const double violation =
        (i < lp->num_equalities) ? (row - lp->q[i]) : fmin(row - lp->q[i], 0.0);
    t.primal_sq += violation * violation;
    t.primal_inf = fmax(t.primal_inf, fabs(violation));
    t.q_dot_y += lp->q[i] * (scale * y[i]);
  }
  for (Int j = 0; j < n; ++j) {
    const double scale = lp->col_scale[j];
    const double product = scale != 0.0 ? kt_y[j] / scale : kt_y[j];
    const double lambda = lp->c[j] - product;
    const int has_lo = lp->lower[j] > -INFINITY;
    const int has_hi = lp->upper[j] < INFINITY;
    double leftover;
    if (has_lo && has_hi) {
      leftover = 0.0;
    } else if (has_lo) {
      leftover = fmin(lambda, 0.0);
    } else if (has_hi) {
      leftover = fmax(lambda, 0.0);
    } else {
      leftover = lambda;
    }
    t.dual_sq += leftover * leftover;
    t.dual_inf = fmax(t.dual_inf, fabs(leftover));
    const double absorbed = lambda - leftover;
    if (absorbed > 0.0) {
      t.bound_term += absorbed * lp->lower[j];
    } else if (absorbed < 0.0) {
      t.bound_term += absorbed * lp->upper[j];
    }
    t.c_dot_x += lp->c[j] * (scale * x[j]);
  }
  *out = t;
}

/* residual_from_terms (pdhg.cpp:70-87) with ResidualScales (pdhg.cpp:55-66):
 * out[12] = primal_residual, dual_residual, primal_residual_inf,
 * dual_residual_inf, primal_objective, dual_objective, absolute_gap,
 * relative_primal, relative_dual, relative_gap, relative_primal_inf,
 * relative_dual_inf (the order of anuk_pdhg_residual). */
void orc_residual_from_terms(const orc_conv_terms* t, double q_two, double c_two,
                             double q_inf, double c_inf, double* out) {
  const double pr = sqrt(t->primal_sq);
  const double dr = sqrt(t->dual_sq);
  const double pobj = t->c_dot_x;
  const double dobj = t->q_dot_y + t->bound_term;
  const double gap = fabs(pobj - dobj);
  out[0] = pr;
  out[1] = dr;
  out[2] = t->primal_inf;
  out[3] = t->dual_inf;
  out[4] = pobj;
  out[5] = dobj;
  out[6] = gap;
  out[7] = pr / (1.0 + q_two);
  out[8] = dr / (1.0 + c_two);
  out[9] = gap / (1.0 + fabs(pobj) + fabs(dobj));
  out[10] = t->primal_inf / (1.0 + q_inf);
  out[11] = t->dual_inf / (1.0 + c_inf);
}

/* estimate_matrix_norm (pdhg.cpp:223-282), host branch; the backend branch
 * (247-262) runs the same loops in the same order and returns the same bits
 * on the CPU backend. */
double orc_estimate_matrix_norm(Int m, Int n, const Int* k_start, const Int* k_index,
                                const double* k_value, const Int* t_start,
                                const Int* t_index, const double* t_value,
                                int iterations, double tolerance) {
  if (n == 0 || m == 0 || k_start[m] == 0) return 0.0;
  double* v = (double*)malloc((size_t)n * sizeof(double));
  double* kv = (double*)malloc((size_t)m * sizeof(double));
  double* ktkv = (double*)malloc((size_t)n * sizeof(double));
  for (Int j = 0; j < n; ++j) v[j] = 1.0 + 0.1 * (double)(j % 7);
  const double norm = orc_two_norm(v, n);
  double sigma = 0.0;
  if (norm == 0.0) goto done;
  for (Int j = 0; j < n; ++j) v[j] /= norm;
  for (int it = 0; it < iterations; ++it) {
    orc_spmv(m, k_start, k_index, k_value, v, kv);
    orc_spmv(n, t_start, t_index, t_value, kv, ktkv);
    const double next = orc_two_norm(ktkv, n);
    if (next == 0.0) {
      sigma = 0.0;
      goto done;
    }
    for (Int j = 0; j < n; ++j) v[j] = ktkv[j] / next;
    const double estimate = sqrt(next);
    if (it > 0 && fabs(estimate - sigma) <= tolerance * estimate) {
      sigma = estimate;
      break;
    }
    sigma = estimate;
  }
done:
  free(v);
  free(kv);
  free(ktkv);
  return sigma;
}

/* is_infeasibility_certificate (pdhg.cpp:191-221). */
int orc_is_infeasibility_certificate(Int m, Int n, Int num_equalities, const Int* t_start,
                                     const Int* t_index, const double* t_value,
                                     const double* q, const double* lower,
                                     const double* upper, const double* dy,
                                     double tolerance) {
  double scale = 0.0;
  for (Int i = 0; i < m; ++i) {
    if (i >= num_equalities && dy[i] < 0.0) return 0;
    scale = fmax(scale, fabs(dy[i]));
  }
  if (scale <= 0.0) return 0;
  double* lambda = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  orc_spmv(n, t_start, t_index, t_value, dy, lambda);
  double support = 0.0;
  int ok = 1;
  for (Int j = 0; j < n && ok; ++j) {
    const double l = lambda[j];
    if (l > 0.0) {
      if (upper[j] >= INFINITY) ok = 0;
      else support += l * upper[j];
    } else if (l < 0.0) {
      if (lower[j] <= -INFINITY) ok = 0;
      else support += l * lower[j];
    }
  }
  free(lambda);
  if (!ok) return 0;
  double rhs = 0.0;
  for (Int i = 0; i < m; ++i) rhs += q[i] * dy[i];
  return (rhs - support) / scale > tolerance;
}

/* ---------------------------------------------------------------------- */
/* scale_lp (scaling.cpp:32-100) in place on a CSR K (rows m, cols n).      */
/* The caller recomputes K^T afterwards (scaling.cpp:98-99).               */
/* ---------------------------------------------------------------------- */

/* SparseMatrix::row_norms / col_norms (sparse.cpp:120-140), then one
 * equilibrate sweep (scaling.cpp:32-59). norm: 0 = infinity, 1 = one. */
static void equilibrate(Int m, Int n, const Int* start, const Int* index, double* value,
                        double* q, double* c, double* lower, double* upper,
                        double* row_scale, double* col_scale, int norm) {
  double* rn = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  double* cn = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (Int r = 0; r < m; ++r) {
    double acc = 0.0;
    for (Int k = start[r]; k < start[r + 1]; ++k) {
      const double a = fabs(value[k]);
      acc = norm == 0 ? fmax(acc, a) : acc + a;
    }
    rn[r] = acc;
  }
  const Int nnz = start[m];
  for (Int k = 0; k < nnz; ++k) {
    const double a = fabs(value[k]);
    double* acc = &cn[index[k]];
    *acc = norm == 0 ? fmax(*acc, a) : *acc + a;
  }
  for (Int i = 0; i < m; ++i) rn[i] = rn[i] > 0.0 ? 1.0 / sqrt(rn[i]) : 1.0;
  for (Int j = 0; j < n; ++j) cn[j] = cn[j] > 0.0 ? 1.0 / sqrt(cn[j]) : 1.0;
  /* scale_rows then scale_cols (sparse.cpp:142-153) */
  for (Int r = 0; r < m; ++r)
    for (Int k = start[r]; k < start[r + 1]; ++k) value[k] *= rn[r];
  for (Int k = 0; k < nnz; ++k) value[k] *= cn[index[k]];
  for (Int i = 0; i < m; ++i) {
    q[i] *= rn[i];
    row_scale[i] *= rn[i];
  }
  for (Int j = 0; j < n; ++j) {
    c[j] *= cn[j];
    lower[j] /= cn[j];
    upper[j] /= cn[j];
    col_scale[j] *= cn[j];
  }
  free(rn);
  free(cn);
}

/* scaling.cpp:75-100 */
void orc_scale_lp(Int m, Int n, const Int* start, const Int* index, double* value,
                  double* q, double* c, double* lower, double* upper, double* row_scale,
                  double* col_scale, int ruiz_iterations, int pock_chambolle) {
  for (Int i = 0; i < m; ++i) row_scale[i] = 1.0;
  for (Int j = 0; j < n; ++j) col_scale[j] = 1.0;
  for (int it = 0; it < ruiz_iterations; ++it)
    equilibrate(m, n, start, index, value, q, c, lower, upper, row_scale, col_scale, 0);
  if (pock_chambolle)
    equilibrate(m, n, start, index, value, q, c, lower, upper, row_scale, col_scale, 1);
}

/* ---------------------------------------------------------------------- */
/* to_standard_form (standard_form.cpp:72-165), LP part.                   */
/* Pass 1 (out == NULL) returns the standard-form row count in *rows_out and */
/* the triplet count in *count_out; pass 2 fills q, row_model, row_sign and  */
/* the triplets (ri, ci, v) that orc_from_triplets then assembles.          */
/* ---------------------------------------------------------------------- */
void orc_to_standard_rows(Int m, const Int* a_start, const Int* a_index,
                          const double* a_value, const double* row_lower,
                          const double* row_upper, Int* rows_out, Int* neq_out,
                          int64_t* count_out, double* q, Int* row_model, double* row_sign,
                          Int* ri, Int* ci, double* v) {
  Int neq = 0, nineq = 0;
  for (Int i = 0; i < m; ++i) {
    const double lo = row_lower[i], hi = row_upper[i];
    const int has_lo = lo > -INFINITY, has_hi = hi < INFINITY;
    if (has_lo && has_hi && lo == hi) neq++;
    else if (has_lo && has_hi) nineq += 2;
    else if (has_lo || has_hi) nineq += 1;
  }
  *rows_out = neq + nineq;
  *neq_out = neq;
  if (q == NULL) {
    int64_t count = 0;
    for (Int i = 0; i < m; ++i) {
      const double lo = row_lower[i], hi = row_upper[i];
      const int has_lo = lo > -INFINITY, has_hi = hi < INFINITY;
      const Int len = a_start[i + 1] - a_start[i];
      if (has_lo && has_hi && lo != hi) count += 2 * len;
      else if (has_lo || has_hi) count += len;
    }
    *count_out = count;
    return;
  }
  /* Equalities fill rows [0, neq) in model order, inequalities [neq, ...). */
  Int eq_dest = 0, in_dest = neq;
  int64_t t = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (Int i = 0; i < m; ++i) {
      const double lo = row_lower[i], hi = row_upper[i];
      const int has_lo = lo > -INFINITY, has_hi = hi < INFINITY;
      const int is_eq = has_lo && has_hi && lo == hi;
      if ((pass == 0) != (is_eq != 0)) continue;
      double signs[2];
      double rhs[2];
      int emit = 0;
      if (is_eq) {
        signs[0] = 1.0, rhs[0] = lo, emit = 1;
      } else if (has_lo && has_hi) {
        signs[0] = 1.0, rhs[0] = lo, signs[1] = -1.0, rhs[1] = -hi, emit = 2;
      } else if (has_lo) {
        signs[0] = 1.0, rhs[0] = lo, emit = 1;
      } else if (has_hi) {
        signs[0] = -1.0, rhs[0] = -hi, emit = 1;
      }
      for (int e = 0; e < emit; ++e) {
        const Int dest = is_eq ? eq_dest++ : in_dest++;
        q[dest] = rhs[e];
        row_model[dest] = i;
        row_sign[dest] = signs[e];
        for (Int k = a_start[i]; k < a_start[i + 1]; ++k) {
          ri[t] = dest;
          ci[t] = a_index[k];
          v[t] = signs[e] * a_value[k];
          ++t;
        }
      }
    }
  }
  *count_out = t;
}
