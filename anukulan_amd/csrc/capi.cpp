// extern "C" entry points of libanukulan_hip.so (declared in
// include/anukulan_hip.h). Each one validates its arguments, forwards to the
// C++ backend or solver, and turns any exception into a status code plus a
// message retrievable with anuk_last_error.
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"
#include "mps.hpp"
#include "adjoint.hpp"
#include "batch.hpp"
#include "multi.hpp"
#include "nodes.hpp"
#include "finish.hpp"
#include "scaling.hpp"
#include "pdhg.hpp"
#include "stack_presolve.hpp"

struct anuk_ctx {
  std::unique_ptr<anuk::HipBackend> be;
  std::string error;
};

struct anuk_lp {
  anuk::StandardLp lp;
};

struct anuk_mps {
  anuk::MpsModel m;
};

struct anuk_presolve {
  anuk::StackPresolve p;
};

namespace {

thread_local std::string g_create_error;

int status_of(anuk_ctx* ctx, int code, const char* what) {
  if (ctx != nullptr) {
    ctx->error = what;
  } else {
    g_create_error = what;
  }
  return code;
}

template <class F>
int guarded(anuk_ctx* ctx, F&& f) {
  try {
    f();
    if (ctx != nullptr) ctx->error.clear();
    return ANUK_OK;
  } catch (const std::bad_alloc&) {
    return status_of(ctx, ANUK_ERR_NOMEM, "out of memory");
  } catch (const std::invalid_argument& e) {
    return status_of(ctx, ANUK_ERR_INVALID, e.what());
  } catch (const anuk::MpsError& e) {
    return status_of(ctx, ANUK_ERR_PARSE, e.what());
  } catch (const anuk::FileError& e) {
    return status_of(ctx, ANUK_ERR_IO, e.what());
  } catch (const std::runtime_error& e) {
    return status_of(ctx, ANUK_ERR_DEVICE, e.what());
  } catch (const std::exception& e) {
    return status_of(ctx, ANUK_ERR_INTERNAL, e.what());
  }
}

void need(bool cond, const char* msg) {
  if (!cond) throw std::invalid_argument(msg);
}

anuk::CsrView view(const anuk_csr* a) {
  need(a != nullptr && a->start != nullptr, "matrix is null");
  need(a->rows >= 0 && a->cols >= 0, "matrix has negative dimensions");
  const anuk::Int nnz = a->start[a->rows];
  need(nnz == 0 || (a->index != nullptr && a->value != nullptr), "matrix arrays are null");
  return anuk::CsrView{a->id, a->rows, a->cols, a->start, a->index, a->value};
}

anuk::DevProblem problem_of(const anuk_conv_problem* p, anuk::Int n, anuk::Int m) {
  need(p != nullptr, "convergence problem is null");
  anuk::DevProblem d;
  d.n = n;
  d.m = m;
  d.num_equalities = p->num_equalities;
  d.orig_c = p->c;
  d.orig_lower = p->lower;
  d.orig_upper = p->upper;
  d.orig_q = p->q;
  d.row_scale = p->row_scale;
  d.col_scale = p->col_scale;
  return d;
}

void fill_residual(const anuk::Residual& r, anuk_pdhg_residual* o) {
  o->primal_residual = r.primal_residual;
  o->dual_residual = r.dual_residual;
  o->primal_residual_inf = r.primal_residual_inf;
  o->dual_residual_inf = r.dual_residual_inf;
  o->primal_objective = r.primal_objective;
  o->dual_objective = r.dual_objective;
  o->absolute_gap = r.absolute_gap;
  o->relative_primal = r.relative_primal;
  o->relative_dual = r.relative_dual;
  o->relative_gap = r.relative_gap;
  o->relative_primal_inf = r.relative_primal_inf;
  o->relative_dual_inf = r.relative_dual_inf;
}

anuk::MpsOptions mps_options(const anuk_mps_options* o) {
  anuk::MpsOptions opt;
  if (o == nullptr) return opt;
  need(o->format >= 0 && o->format <= 2, "format must be 0 (auto), 1 (free) or 2 (fixed)");
  need(o->negative_up_bound == 0 || o->negative_up_bound == 1, "negative_up_bound must be 0 or 1");
  need(o->infinity_threshold > 0.0, "infinity_threshold must be positive");
  opt.format = o->format;
  opt.negative_up_minus_infinity = o->negative_up_bound == 1;
  opt.infinity_threshold = o->infinity_threshold;
  return opt;
}

int csr_out(const anuk::Csr& a, const int32_t** start, const int32_t** index, const double** value) {
  if (start) *start = a.start.data();
  if (index) *index = a.index.data();
  if (value) *value = a.value.data();
  return ANUK_OK;
}

}  // namespace

extern "C" {

int anuk_create(int device, anuk_ctx** out) {
  if (out == nullptr) return ANUK_ERR_INVALID;
  *out = nullptr;
  try {
    auto ctx = std::make_unique<anuk_ctx>();
    ctx->be = std::make_unique<anuk::HipBackend>(device);
    *out = ctx.release();
    g_create_error.clear();
    return ANUK_OK;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return ANUK_ERR_DEVICE;
  }
}

void anuk_destroy(anuk_ctx* ctx) { delete ctx; }

const char* anuk_last_error(const anuk_ctx* ctx) {
  return ctx != nullptr ? ctx->error.c_str() : g_create_error.c_str();
}

const char* anuk_name(const anuk_ctx* ctx) {
  (void)ctx;
  return "hip";
}

int anuk_synchronize(anuk_ctx* ctx) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] { ctx->be->synchronize(); });
}

int anuk_allocate(anuk_ctx* ctx, int32_t n, double** out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0, "negative size");
    *out = ctx->be->allocate(n);
  });
}

int anuk_deallocate(anuk_ctx* ctx, double* p) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] { ctx->be->deallocate(p); });
}

int anuk_upload(anuk_ctx* ctx, const double* host, double* target, int32_t n) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (host && target)), "bad upload arguments");
    ctx->be->upload(host, target, n);
  });
}

int anuk_download(anuk_ctx* ctx, const double* source, double* host, int32_t n) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (host && source)), "bad download arguments");
    ctx->be->download(source, host, n);
  });
}

int anuk_fill(anuk_ctx* ctx, double* target, int32_t n, double value) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || target), "bad fill arguments");
    ctx->be->fill(target, n, value);
  });
}

int anuk_copy(anuk_ctx* ctx, const double* source, double* target, int32_t n) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (source && target)), "bad copy arguments");
    ctx->be->copy(source, target, n);
  });
}

int anuk_prepare(anuk_ctx* ctx, const anuk_csr* a) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] { ctx->be->prepare(view(a)); });
}

int anuk_release(anuk_ctx* ctx) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] { ctx->be->release(); });
}

int anuk_multiply(anuk_ctx* ctx, const anuk_csr* a, const double* x, double* y) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const anuk::CsrView v = view(a);
    need(v.rows == 0 || y != nullptr, "output vector is null");
    need(v.cols == 0 || x != nullptr || v.nnz() == 0, "input vector is null");
    ctx->be->multiply(v, x, y);
  });
}

int anuk_multiply_transpose(anuk_ctx* ctx, const anuk_csr* at, const double* y, double* x) {
  return anuk_multiply(ctx, at, y, x);
}

int anuk_dot(anuk_ctx* ctx, const double* a, const double* b, int32_t n, double* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (a && b)), "bad dot arguments");
    *out = ctx->be->dot(a, b, n);
  });
}

int anuk_two_norm(anuk_ctx* ctx, const double* a, int32_t n, double* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || a), "bad two_norm arguments");
    *out = ctx->be->two_norm(a, n);
  });
}

int anuk_inf_norm(anuk_ctx* ctx, const double* a, int32_t n, double* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || a), "bad inf_norm arguments");
    *out = ctx->be->inf_norm(a, n);
  });
}

int anuk_primal_update(anuk_ctx* ctx, int32_t n, double tau, double gamma, double halpern,
                       const double* x, const double* c, const double* kt_y, const double* lower,
                       const double* upper, const double* anchor, double* x_out, double* dx,
                       double* x_bar) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0, "negative size");
    need(n == 0 || (x && c && kt_y && lower && upper && x_out && dx && x_bar), "null vector");
    ctx->be->primal_update(n, tau, gamma, halpern, x, c, kt_y, lower, upper, anchor, x_out, dx, x_bar);
  });
}

int anuk_dual_update(anuk_ctx* ctx, int32_t m, int32_t num_equalities, double sigma, double gamma,
                     double halpern, const double* y, const double* q, const double* k_x_bar,
                     const double* k_x, const double* anchor_y, const double* anchor_kx,
                     double* y_out, double* dy, double* k_out, double* k_dx) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(m >= 0, "negative size");
    need(m == 0 || (y && q && k_x_bar && k_x && y_out && dy && k_out && k_dx), "null vector");
    need(anchor_y == nullptr || anchor_kx != nullptr, "anchor_y given without anchor_kx");
    ctx->be->dual_update(m, num_equalities, sigma, gamma, halpern, y, q, k_x_bar, k_x, anchor_y,
                         anchor_kx, y_out, dy, k_out, k_dx);
  });
}

int anuk_accumulate(anuk_ctx* ctx, int32_t n, double weight, const double* v, double* sum) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (v && sum)), "bad accumulate arguments");
    ctx->be->accumulate(n, weight, v, sum);
  });
}

int anuk_scale_into(anuk_ctx* ctx, int32_t n, double weight, const double* sum, double* out) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && (n == 0 || (sum && out)), "bad scale_into arguments");
    ctx->be->scale_into(n, weight, sum, out);
  });
}

int anuk_weighted_norm_squared(anuk_ctx* ctx, int32_t n, int32_t m, const double* dx,
                               const double* dy, double omega, double* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && m >= 0 && (n == 0 || dx) && (m == 0 || dy), "bad weighted norm arguments");
    *out = ctx->be->weighted_norm_squared(n, m, dx, dy, omega);
  });
}

int anuk_step_size_terms(anuk_ctx* ctx, int32_t n, int32_t m, const double* dx, const double* dy,
                         const double* k_dx, double omega, double* interaction, double* movement) {
  if (ctx == nullptr || interaction == nullptr || movement == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && m >= 0 && (n == 0 || dx) && (m == 0 || (dy && k_dx)), "bad step term arguments");
    ctx->be->step_size_terms(n, m, dx, dy, k_dx, omega, interaction, movement);
  });
}

int anuk_convergence_terms(anuk_ctx* ctx, int32_t n, int32_t m, const anuk_conv_problem* lp,
                           const double* x, const double* k_x, const double* y, const double* kt_y,
                           anuk_conv_terms* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && m >= 0, "negative size");
    need(n == 0 || (x && kt_y && lp && lp->c && lp->lower && lp->upper && lp->col_scale),
         "null column vector");
    need(m == 0 || (k_x && y && lp && lp->q && lp->row_scale), "null row vector");
    const anuk::Terms t = ctx->be->convergence_terms(n, m, problem_of(lp, n, m), x, k_x, y, kt_y);
    out->primal_sq = t.primal_sq;
    out->primal_inf = t.primal_inf;
    out->q_dot_y = t.q_dot_y;
    out->dual_sq = t.dual_sq;
    out->dual_inf = t.dual_inf;
    out->bound_term = t.bound_term;
    out->c_dot_x = t.c_dot_x;
  });
}

int anuk_convergence_terms_noalloc(anuk_ctx* ctx, int32_t n, int32_t m,
                                   const anuk_conv_problem* lp, const double* x,
                                   const double* k_x, const double* y, const double* kt_y,
                                   anuk_conv_terms* out) {
  if (ctx == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(n >= 0 && m >= 0, "negative size");
    need(n == 0 || (x && kt_y && lp && lp->c && lp->lower && lp->upper && lp->col_scale),
         "null column vector");
    need(m == 0 || (k_x && y && lp && lp->q && lp->row_scale), "null row vector");
    const anuk::Terms t =
        ctx->be->convergence_terms_noalloc(n, m, problem_of(lp, n, m), x, k_x, y, kt_y);
    out->primal_sq = t.primal_sq;
    out->primal_inf = t.primal_inf;
    out->q_dot_y = t.q_dot_y;
    out->dual_sq = t.dual_sq;
    out->dual_inf = t.dual_inf;
    out->bound_term = t.bound_term;
    out->c_dot_x = t.c_dot_x;
  });
}

int anuk_set_profiling(anuk_ctx* ctx, int on) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] { ctx->be->set_profiling(on != 0); });
}

int anuk_profile_report(anuk_ctx* ctx, char* buf, size_t cap, size_t* needed) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const std::string r = ctx->be->profile_report();
    if (needed) *needed = r.size() + 1;
    if (buf != nullptr && cap > 0) {
      const size_t k = std::min(cap - 1, r.size());
      std::memcpy(buf, r.data(), k);
      buf[k] = '\0';
    }
  });
}

int anuk_tuning_report(anuk_ctx* ctx, char* buf, size_t cap, size_t* needed) {
  if (ctx == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const std::string r = ctx->be->tuning_report();
    if (needed) *needed = r.size() + 1;
    if (buf != nullptr && cap > 0) {
      const size_t k = std::min(cap - 1, r.size());
      std::memcpy(buf, r.data(), k);
      buf[k] = '\0';
    }
  });
}

// ---- host layout ------------------------------------------------------------

int anuk_lp_from_model(const anuk_model* model, anuk_lp** out) {
  if (model == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  *out = nullptr;
  return guarded(nullptr, [&] {
    need(model->rows >= 0 && model->cols >= 0 && model->nnz >= 0, "negative model size");
    need(model->cols == 0 || (model->objective && model->col_lower && model->col_upper), "null column data");
    need(model->rows == 0 || (model->row_lower && model->row_upper), "null row data");
    need(model->nnz == 0 || (model->row_index && model->col_index && model->value), "null triplets");
    anuk::Model m;
    m.rows = model->rows;
    m.cols = model->cols;
    m.entries.resize(static_cast<size_t>(model->nnz));
    for (int64_t k = 0; k < model->nnz; ++k)
      m.entries[static_cast<size_t>(k)] = {model->row_index[k], model->col_index[k], model->value[k]};
    m.objective.assign(model->objective, model->objective + model->cols);
    m.col_lower.assign(model->col_lower, model->col_lower + model->cols);
    m.col_upper.assign(model->col_upper, model->col_upper + model->cols);
    if (model->rows > 0) {
      m.row_lower.assign(model->row_lower, model->row_lower + model->rows);
      m.row_upper.assign(model->row_upper, model->row_upper + model->rows);
    }
    m.maximize = model->maximize != 0;
    m.objective_offset = model->objective_offset;
    auto h = std::make_unique<anuk_lp>();
    h->lp = anuk::to_standard_form(m);
    *out = h.release();
  });
}

int anuk_lp_from_standard(int32_t rows, int32_t cols, int32_t num_equalities, const int32_t* start,
                          const int32_t* index, const double* value, const double* c,
                          const double* q, const double* lower, const double* upper,
                          double objective_scale, double objective_offset, anuk_lp** out) {
  if (out == nullptr) return ANUK_ERR_INVALID;
  *out = nullptr;
  return guarded(nullptr, [&] {
    need(rows >= 0 && cols >= 0, "negative size");
    need(num_equalities >= 0 && num_equalities <= rows, "num_equalities out of range");
    need(start != nullptr && (cols == 0 || (c && lower && upper)) && (rows == 0 || q), "null array");
    auto h = std::make_unique<anuk_lp>();
    anuk::StandardLp& lp = h->lp;
    lp.k = anuk::Csr::from_csr(rows, cols, start, index, value);
    lp.kt = lp.k.transpose();
    lp.c.assign(c, c + cols);
    lp.q.assign(q, q + rows);
    lp.lower.assign(lower, lower + cols);
    lp.upper.assign(upper, upper + cols);
    lp.num_equalities = num_equalities;
    lp.objective_scale = objective_scale;
    lp.objective_offset = objective_offset;
    lp.row_model.resize(static_cast<size_t>(rows));
    lp.row_sign.assign(static_cast<size_t>(rows), 1.0);
    for (int32_t i = 0; i < rows; ++i) lp.row_model[static_cast<size_t>(i)] = i;
    for (int32_t j = 0; j < cols; ++j)
      need(!(lp.lower[static_cast<size_t>(j)] > lp.upper[static_cast<size_t>(j)]), "empty bound interval");
    *out = h.release();
  });
}

void anuk_lp_free(anuk_lp* lp) { delete lp; }

int anuk_lp_info_get(const anuk_lp* lp, anuk_lp_info* out) {
  if (lp == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  out->rows = lp->lp.rows();
  out->cols = lp->lp.cols();
  out->nnz = lp->lp.k.nnz();
  out->num_equalities = lp->lp.num_equalities;
  out->objective_scale = lp->lp.objective_scale;
  out->objective_offset = lp->lp.objective_offset;
  return ANUK_OK;
}

int anuk_lp_export(const anuk_lp* h, int32_t* start, int32_t* index, double* value, double* c,
                   double* q, double* lower, double* upper, int32_t* row_model, double* row_sign) {
  if (h == nullptr) return ANUK_ERR_INVALID;
  const anuk::StandardLp& lp = h->lp;
  auto put_i = [](int32_t* d, const std::vector<int32_t>& s) {
    if (d && !s.empty()) std::memcpy(d, s.data(), s.size() * sizeof(int32_t));
  };
  auto put_d = [](double* d, const std::vector<double>& s) {
    if (d && !s.empty()) std::memcpy(d, s.data(), s.size() * sizeof(double));
  };
  put_i(start, lp.k.start);
  put_i(index, lp.k.index);
  put_d(value, lp.k.value);
  put_d(c, lp.c);
  put_d(q, lp.q);
  put_d(lower, lp.lower);
  put_d(upper, lp.upper);
  put_i(row_model, lp.row_model);
  put_d(row_sign, lp.row_sign);
  return ANUK_OK;
}

int anuk_lp_scale(const anuk_lp* lp, int32_t ruiz_iterations, int32_t pock_chambolle,
                  anuk_lp** scaled, double* row_scale, double* col_scale) {
  if (lp == nullptr || scaled == nullptr) return ANUK_ERR_INVALID;
  *scaled = nullptr;
  return guarded(nullptr, [&] {
    need(ruiz_iterations >= 0, "negative sweep count");
    auto h = std::make_unique<anuk_lp>();
    h->lp = lp->lp;
    const anuk::Scaling s = anuk::scale_lp(&h->lp, ruiz_iterations, pock_chambolle != 0);
    if (row_scale && !s.row_scale.empty())
      std::memcpy(row_scale, s.row_scale.data(), s.row_scale.size() * sizeof(double));
    if (col_scale && !s.col_scale.empty())
      std::memcpy(col_scale, s.col_scale.data(), s.col_scale.size() * sizeof(double));
    *scaled = h.release();
  });
}

int anuk_lp_scale_device(anuk_ctx* ctx, const anuk_lp* lp, int32_t ruiz_iterations,
                         int32_t pock_chambolle, anuk_lp** scaled, double* row_scale,
                         double* col_scale) {
  if (ctx == nullptr || lp == nullptr || scaled == nullptr) return ANUK_ERR_INVALID;
  *scaled = nullptr;
  return guarded(ctx, [&] {
    need(ruiz_iterations >= 0, "negative sweep count");
    auto h = std::make_unique<anuk_lp>();
    h->lp = lp->lp;
    const anuk::Scaling s =
        anuk::scale_lp_device(*ctx->be, &h->lp, ruiz_iterations, pock_chambolle != 0);
    if (row_scale && !s.row_scale.empty())
      std::memcpy(row_scale, s.row_scale.data(), s.row_scale.size() * sizeof(double));
    if (col_scale && !s.col_scale.empty())
      std::memcpy(col_scale, s.col_scale.data(), s.col_scale.size() * sizeof(double));
    *scaled = h.release();
  });
}

// ---- the solve --------------------------------------------------------------

void anuk_pdhg_default_options(anuk_pdhg_options* options) {
  if (options != nullptr) anuk::default_options(options);
}

namespace {

anuk_pdhg_options checked_options(const anuk_pdhg_options* options) {
  anuk_pdhg_options o;
  if (options != nullptr) {
    o = *options;
  } else {
    anuk::default_options(&o);
  }
  need(o.termination_check_frequency > 0, "termination_check_frequency must be positive");
  need(o.max_iterations >= 0, "max_iterations must be non-negative");
  need(o.polish_first_iteration > 0, "polish_first_iteration must be positive");
  need(o.product_order >= 0 && o.product_order <= 2, "product_order must be 0, 1 or 2");
  need(o.autotune == 0 || o.autotune == 1, "autotune must be 0 or 1");
  need(o.batch_waves == 0 || o.batch_waves == 1 || o.batch_waves == 2 || o.batch_waves == 4 ||
           o.batch_waves == 8,
       "batch_waves must be 0, 1, 2, 4 or 8");
  need(o.batch_rows == 0 || o.batch_rows == 1 || o.batch_rows == 2 || o.batch_rows == 4,
       "batch_rows must be 0, 1, 2 or 4");
  need(o.batch_layout == 0 || o.batch_layout == 1 || o.batch_layout == 2,
       "batch_layout must be 0, 1 or 2");
  need(o.batch_layout != 2 || o.batch_rows == 0 || o.batch_rows == 1,
       "batch_layout 2 (the fixed-width row copy) runs one row per wave");
  need(o.batch_pack == 0 || o.batch_pack == 1, "batch_pack must be 0 or 1");
  need(o.multi_order == 0 || o.multi_order == 1 || o.multi_order == 2,
       "multi_order must be 0, 1 or 2");
  need(o.multi_stream_loads == 0 || o.multi_stream_loads == 1 || o.multi_stream_loads == 2,
       "multi_stream_loads must be 0, 1 or 2");
  need(o.small_lp_engine == 0 || o.small_lp_engine == 1, "small_lp_engine must be 0 or 1");
  return o;
}

void fill_result(const anuk::SolveResult& r, anuk_pdhg_result* result) {
  std::memset(result, 0, sizeof(*result));
  result->status = r.status;
  result->iterations = r.iterations;
  result->restarts = r.restarts;
  result->polish_attempts = r.polish_attempts;
  result->polish_iterations = r.polish_iterations;
  result->polished = r.polished ? 1 : 0;
  result->objective = r.objective;
  result->solve_seconds = r.solve_seconds;
  result->matrix_norm_estimate = r.matrix_norm_estimate;
  result->final_step_size = r.final_step_size;
  result->final_primal_weight = r.final_primal_weight;
  result->phase_scaling = r.phases.scaling;
  result->phase_matrix_norm = r.phases.matrix_norm;
  result->phase_setup = r.phases.setup;
  result->phase_iterating = r.phases.iterating;
  result->phase_checks = r.phases.checks;
  result->phase_polishing = r.phases.polishing;
  fill_residual(r.residual, &result->residual);
  std::strncpy(result->message, r.message.c_str(), sizeof(result->message) - 1);
}

}  // namespace

int anuk_solve_pdhg_batch(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                          const anuk_pdhg_options* options, anuk_pdhg_result* results,
                          double* x_out, double* y_out) {
  if (ctx == nullptr || lps == nullptr || results == nullptr || count < 0) return ANUK_ERR_INVALID;
  const int rc = [&] {
    anuk_pdhg_options o;
    const int g = guarded(ctx, [&] { o = checked_options(options); });
    if (g != ANUK_OK) return g;
    const char* why = nullptr;
    if (!anuk::batch_supports(o, &why)) {
      ctx->error = std::string("batched solve does not run: ") + why;
      return static_cast<int>(ANUK_ERR_UNSUPPORTED);
    }
    return guarded(ctx, [&] {
      std::vector<const anuk::StandardLp*> v;
      for (int32_t i = 0; i < count; ++i) {
        need(lps[i] != nullptr, "null LP in the batch");
        v.push_back(&lps[i]->lp);
      }
      const std::vector<anuk::SolveResult> rs = anuk::solve_pdhg_batch(*ctx->be, v, o);
      for (size_t i = 0; i < rs.size(); ++i) {
        fill_result(rs[i], &results[i]);
        const size_t cols = static_cast<size_t>(v[i]->cols()), rows = static_cast<size_t>(v[i]->rows());
        if (x_out && !rs[i].x.empty()) std::memcpy(x_out + i * cols, rs[i].x.data(), cols * sizeof(double));
        if (y_out && !rs[i].y.empty()) std::memcpy(y_out + i * rows, rs[i].y.data(), rows * sizeof(double));
      }
    });
  }();
  return rc;
}

int anuk_solve_pdhg_multi(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                          const anuk_pdhg_options* options, anuk_pdhg_result* results,
                          double* x_out, double* y_out) {
  if (ctx == nullptr || lps == nullptr || results == nullptr || count < 0) return ANUK_ERR_INVALID;
  anuk_pdhg_options o;
  const int g = guarded(ctx, [&] { o = checked_options(options); });
  if (g != ANUK_OK) return g;
  const char* why = nullptr;
  if (!anuk::batch_supports(o, &why)) {
    ctx->error = std::string("multi-LP solve does not run: ") + why;
    return static_cast<int>(ANUK_ERR_UNSUPPORTED);
  }
  return guarded(ctx, [&] {
    std::vector<const anuk::StandardLp*> v;
    for (int32_t i = 0; i < count; ++i) {
      need(lps[i] != nullptr, "null LP in the stack");
      v.push_back(&lps[i]->lp);
    }
    const std::vector<anuk::SolveResult> rs = anuk::solve_pdhg_multi(*ctx->be, v, o);
    size_t xo = 0, yo = 0;
    for (size_t i = 0; i < rs.size(); ++i) {
      fill_result(rs[i], &results[i]);
      const size_t cols = static_cast<size_t>(v[i]->cols()), rows = static_cast<size_t>(v[i]->rows());
      if (x_out && rs[i].x.size() == cols) std::memcpy(x_out + xo, rs[i].x.data(), cols * sizeof(double));
      if (y_out && rs[i].y.size() == rows) std::memcpy(y_out + yo, rs[i].y.data(), rows * sizeof(double));
      xo += cols;
      yo += rows;
    }
  });
}

int anuk_solve_node_relaxations(anuk_ctx* ctx, const anuk_lp* lp, int32_t count,
                                const double* lower, const double* upper,
                                const anuk_pdhg_options* options, anuk_node_result* results,
                                double* x_out, double* y_out) {
  if (ctx == nullptr || lp == nullptr || count < 0 || (count > 0 && (results == nullptr ||
      lower == nullptr || upper == nullptr)))
    return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const anuk_pdhg_options o = checked_options(options);
    const std::vector<anuk::NodeRelaxation> rs =
        anuk::solve_node_relaxations(*ctx->be, lp->lp, lower, upper, count, o);
    const size_t cols = static_cast<size_t>(lp->lp.cols()), rows = static_cast<size_t>(lp->lp.rows());
    for (size_t i = 0; i < rs.size(); ++i) {
      anuk_node_result& r = results[i];
      std::memset(&r, 0, sizeof(r));
      r.optimal = rs[i].optimal ? 1 : 0;
      r.proved_infeasible = rs[i].proved_infeasible ? 1 : 0;
      r.bound = rs[i].bound;
      fill_result(rs[i].solve, &r.relaxation);
      if (x_out && rs[i].solve.x.size() == cols)
        std::memcpy(x_out + i * cols, rs[i].solve.x.data(), cols * sizeof(double));
      if (y_out && rs[i].solve.y.size() == rows)
        std::memcpy(y_out + i * rows, rs[i].solve.y.data(), rows * sizeof(double));
    }
  });
}

static anuk::FinishOptions finish_options_of(const anuk_finish_options& o) {
  anuk::FinishOptions f;
  f.primal_tolerance = o.primal_tolerance;
  f.dual_tolerance = o.dual_tolerance;
  f.pivot_tolerance = o.pivot_tolerance;
  f.max_iterations = o.max_iterations;
  f.refactor_interval = o.refactor_interval;
  f.crossover = o.crossover;
  f.interior_tolerance = o.interior_tolerance;
  f.crossover_pivot_tolerance = o.crossover_pivot_tolerance;
  f.max_candidates = o.max_candidates;
  f.crossover_refactor_interval = o.crossover_refactor_interval;
  f.dual_weight = o.dual_weight;
  f.ruiz_iterations = o.ruiz_iterations;
  f.threads = o.threads;
  return f;
}

void anuk_finish_default_options(anuk_finish_options* options) {
  if (options == nullptr) return;
  const anuk::FinishOptions f;
  options->primal_tolerance = f.primal_tolerance;
  options->dual_tolerance = f.dual_tolerance;
  options->pivot_tolerance = f.pivot_tolerance;
  options->max_iterations = f.max_iterations;
  options->refactor_interval = f.refactor_interval;
  options->crossover = f.crossover;
  options->interior_tolerance = f.interior_tolerance;
  options->crossover_pivot_tolerance = f.crossover_pivot_tolerance;
  options->max_candidates = f.max_candidates;
  options->crossover_refactor_interval = f.crossover_refactor_interval;
  options->dual_weight = f.dual_weight;
  options->ruiz_iterations = f.ruiz_iterations;
  options->threads = f.threads;
}

namespace {

int finish_impl(const anuk_lp* const* lps, int32_t count, const double* x, const double* y,
                const int32_t* start_basic, const int32_t* start_state,
                const anuk_finish_options* options, anuk_finish_result* results, double* x_out,
                double* y_out, int32_t* basic_out, int32_t* state_out) {
  if (lps == nullptr || count < 0 || (count > 0 && results == nullptr))
    return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  if ((start_basic == nullptr) != (start_state == nullptr))
    return status_of(nullptr, ANUK_ERR_INVALID, "a start basis needs both its columns and its states");
  return guarded(nullptr, [&] {
    anuk_finish_options o;
    anuk_finish_default_options(&o);
    if (options != nullptr) o = *options;
    need(o.primal_tolerance > 0.0 && o.dual_tolerance > 0.0 && o.pivot_tolerance > 0.0,
         "finish tolerances must be positive");
    need(o.max_iterations >= 0 && o.refactor_interval > 0 && o.ruiz_iterations >= 0 && o.threads >= 0,
         "finish limits must be non-negative (refactor_interval positive)");
    std::vector<const anuk::StandardLp*> v;
    std::vector<const double*> xs, ys;
    size_t co = 0, ro = 0;
    for (int32_t i = 0; i < count; ++i) {
      need(lps[i] != nullptr, "null LP among those to finish");
      v.push_back(&lps[i]->lp);
      xs.push_back(x ? x + co : nullptr);
      ys.push_back(y ? y + ro : nullptr);
      co += static_cast<size_t>(lps[i]->lp.cols());
      ro += static_cast<size_t>(lps[i]->lp.rows());
    }
    anuk::FinishBasis start;
    if (start_basic != nullptr && count > 0) {
      const anuk::StandardLp& f = *v[0];
      for (const anuk::StandardLp* lp : v)
        need(lp->rows() == f.rows() && lp->cols() == f.cols(),
             "a shared start basis needs every LP to have the same shape");
      const size_t rows = static_cast<size_t>(f.rows()), cols = static_cast<size_t>(f.cols());
      start.basic.assign(start_basic, start_basic + rows);
      start.state.resize(rows + cols);
      for (size_t j = 0; j < rows + cols; ++j) {
        need(start_state[j] >= 0 && start_state[j] <= 3, "a start state is not 0 to 3");
        start.state[j] = static_cast<anuk::VarState>(start_state[j]);
      }
    }
    const std::vector<anuk::FinishResult> rs = anuk::exact_finish_many(
        v, xs, ys, finish_options_of(o), start_basic != nullptr ? &start : nullptr);
    co = ro = 0;
    for (size_t i = 0; i < rs.size(); ++i) {
      const anuk::FinishResult& r = rs[i];
      anuk_finish_result& o2 = results[i];
      std::memset(&o2, 0, sizeof(o2));
      o2.status = r.status;
      o2.iterations = r.iterations;
      o2.phase_one_iterations = r.phase_one_iterations;
      o2.refactorizations = r.refactorizations;
      o2.bland_iterations = r.bland_iterations;
      o2.started_warm = r.started_warm ? 1 : 0;
      o2.candidates = r.crossover.candidates;
      o2.pushed = r.crossover.pushed;
      o2.rejected_small_pivot = r.crossover.rejected_small_pivot;
      o2.logicals_remaining = r.crossover.logicals_remaining;
      o2.stopped_on_refactorization = r.crossover.stopped_on_refactorization ? 1 : 0;
      o2.objective = r.objective;
      o2.primal_violation = r.primal_violation;
      o2.dual_violation = r.dual_violation;
      o2.seconds = r.seconds;
      std::strncpy(o2.message, r.message.c_str(), sizeof(o2.message) - 1);
      const size_t cols = static_cast<size_t>(v[i]->cols()), rows = static_cast<size_t>(v[i]->rows());
      if (x_out && r.x.size() == cols) std::memcpy(x_out + co, r.x.data(), cols * sizeof(double));
      if (y_out && r.y.size() == rows) std::memcpy(y_out + ro, r.y.data(), rows * sizeof(double));
      if (basic_out && r.basic.size() == rows)
        for (size_t k = 0; k < rows; ++k) basic_out[ro + k] = r.basic[k];
      if (state_out && r.state.size() == rows + cols)
        for (size_t k = 0; k < rows + cols; ++k)
          state_out[co + ro + k] = static_cast<int32_t>(r.state[k]);
      co += cols;
      ro += rows;
    }
  });
}

}  // namespace

int anuk_exact_finish(const anuk_lp* const* lps, int32_t count, const double* x, const double* y,
                      const anuk_finish_options* options, anuk_finish_result* results,
                      double* x_out, double* y_out, int32_t* basic_out) {
  return finish_impl(lps, count, x, y, nullptr, nullptr, options, results, x_out, y_out, basic_out,
                     nullptr);
}

int anuk_exact_finish_basis(const anuk_lp* const* lps, int32_t count, const double* x,
                            const double* y, const int32_t* start_basic,
                            const int32_t* start_state, const anuk_finish_options* options,
                            anuk_finish_result* results, double* x_out, double* y_out,
                            int32_t* basic_out, int32_t* state_out) {
  return finish_impl(lps, count, x, y, start_basic, start_state, options, results, x_out, y_out,
                     basic_out, state_out);
}

void anuk_adjoint_default_options(anuk_adjoint_options* options) {
  if (options != nullptr) anuk::default_adjoint_options(options);
}

int anuk_pdhg_adjoint(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                      const anuk_pdhg_options* scaling, const anuk_adjoint_options* options,
                      const double* x, const double* y, const double* w_x, const double* w_y,
                      double* grad_c, double* grad_q, double* grad_lower, double* grad_upper,
                      anuk_adjoint_info* info) {
  if (ctx == nullptr || lps == nullptr || count < 0 || x == nullptr || y == nullptr)
    return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const anuk_pdhg_options sc = checked_options(scaling);
    anuk_adjoint_options o;
    anuk::default_adjoint_options(&o);
    if (options != nullptr) o = *options;
    std::vector<const anuk::StandardLp*> v;
    std::vector<const double*> xs, ys, wx, wy;
    size_t col_off = 0, row_off = 0;
    for (int32_t i = 0; i < count; ++i) {
      need(lps[i] != nullptr, "null LP in the adjoint batch");
      const anuk::StandardLp& lp = lps[i]->lp;
      v.push_back(&lp);
      xs.push_back(x + col_off);
      ys.push_back(y + row_off);
      wx.push_back(w_x ? w_x + col_off : nullptr);
      wy.push_back(w_y ? w_y + row_off : nullptr);
      col_off += static_cast<size_t>(lp.cols());
      row_off += static_cast<size_t>(lp.rows());
    }
    const std::vector<anuk::AdjointOutcome> out = anuk::pdhg_adjoint(*ctx->be, v, sc, o, xs, ys, wx, wy);
    col_off = row_off = 0;
    for (size_t i = 0; i < out.size(); ++i) {
      const anuk::AdjointOutcome& a = out[i];
      const size_t n = a.grad_c.size(), m = a.grad_q.size();
      if (grad_c) std::memcpy(grad_c + col_off, a.grad_c.data(), n * sizeof(double));
      if (grad_q) std::memcpy(grad_q + row_off, a.grad_q.data(), m * sizeof(double));
      if (grad_lower) std::memcpy(grad_lower + col_off, a.grad_lower.data(), n * sizeof(double));
      if (grad_upper) std::memcpy(grad_upper + col_off, a.grad_upper.data(), n * sizeof(double));
      if (info) {
        anuk_adjoint_info& f = info[i];
        f.iterations = a.iterations;
        f.converged = a.converged ? 1 : 0;
        f.residual_primal = a.residual_primal;
        f.residual_dual = a.residual_dual;
        f.free_columns = a.free_columns;
        f.binding_rows = a.binding_rows;
        f.lower_columns = a.lower_columns;
        f.upper_columns = a.upper_columns;
        f.basis_completed = a.basis_completed ? 1 : 0;
      }
      col_off += n;
      row_off += m;
    }
  });
}

int anuk_solve_pdhg(anuk_ctx* ctx, const anuk_lp* lp, const anuk_pdhg_options* options,
                    anuk_pdhg_result* result, double* x_out, double* y_out) {
  if (ctx == nullptr || lp == nullptr || result == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    const anuk_pdhg_options o = checked_options(options);
    const anuk::SolveResult r = anuk::solve_pdhg(*ctx->be, lp->lp, o);
    fill_result(r, result);
    if (x_out && !r.x.empty()) std::memcpy(x_out, r.x.data(), r.x.size() * sizeof(double));
    if (y_out && !r.y.empty()) std::memcpy(y_out, r.y.data(), r.y.size() * sizeof(double));
  });
}

int anuk_estimate_matrix_norm(anuk_ctx* ctx, const anuk_lp* lp, int32_t iterations,
                              double tolerance, double* out) {
  if (ctx == nullptr || lp == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(ctx, [&] {
    need(iterations >= 0, "negative iteration count");
    *out = anuk::estimate_matrix_norm(*ctx->be, lp->lp, iterations, tolerance);
  });
}

int anuk_evaluate_residual(const anuk_lp* lp, const double* x, const double* y,
                           anuk_pdhg_residual* out) {
  if (lp == nullptr || out == nullptr) return ANUK_ERR_INVALID;
  return guarded(nullptr, [&] {
    need((lp->lp.cols() == 0 || x) && (lp->lp.rows() == 0 || y), "null vector");
    const std::vector<double> vx(x, x + lp->lp.cols());
    const std::vector<double> vy(y, y + lp->lp.rows());
    fill_residual(anuk::evaluate_residual(lp->lp, vx, vy), out);
  });
}

// ---- MPS files --------------------------------------------------------------

void anuk_mps_default_options(anuk_mps_options* out) {
  if (out == nullptr) return;
  const anuk::MpsOptions d;
  out->format = d.format;
  out->negative_up_bound = d.negative_up_minus_infinity ? 1 : 0;
  out->infinity_threshold = d.infinity_threshold;
}

int anuk_mps_read(const char* path, const anuk_mps_options* options, anuk_mps** out) {
  if (path == nullptr || out == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded(nullptr, [&] {
    auto h = std::make_unique<anuk_mps>();
    h->m = anuk::read_mps_file(path, mps_options(options));
    *out = h.release();
  });
}

int anuk_mps_read_text(const char* text, int64_t length, const char* source,
                       const anuk_mps_options* options, anuk_mps** out) {
  if (out == nullptr || length < 0 || (text == nullptr && length > 0))
    return status_of(nullptr, ANUK_ERR_INVALID, "null or negative argument");
  *out = nullptr;
  return guarded(nullptr, [&] {
    auto h = std::make_unique<anuk_mps>();
    h->m = anuk::read_mps_text(std::string_view(text ? text : "", static_cast<size_t>(length)),
                               source ? source : "<text>", mps_options(options));
    *out = h.release();
  });
}

void anuk_mps_free(anuk_mps* mps) { delete mps; }

int anuk_mps_info_get(const anuk_mps* mps, anuk_mps_info* out) {
  if (mps == nullptr || out == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  const anuk::MpsModel& m = mps->m;
  out->rows = m.a.rows;
  out->cols = m.a.cols;
  out->nnz = m.a.nnz();
  out->hessian_nnz = m.hessian.nnz();
  out->maximize = m.maximize ? 1 : 0;
  out->fixed_format = m.fixed_format ? 1 : 0;
  out->objective_offset = m.objective_offset;
  out->free_rows_dropped = m.free_rows_dropped;
  out->warnings = static_cast<int32_t>(m.warnings.size());
  out->objective_nonzeros = m.objective_nonzeros;
  return ANUK_OK;
}

int anuk_mps_model(const anuk_mps* mps, anuk_model* out) {
  if (mps == nullptr || out == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  const anuk::MpsModel& m = mps->m;
  out->rows = m.a.rows;
  out->cols = m.a.cols;
  out->nnz = m.a.nnz();
  out->row_index = m.coo_row.data();
  out->col_index = m.coo_col.data();
  out->value = m.a.value.data();
  out->objective = m.objective.data();
  out->row_lower = m.row_lower.data();
  out->row_upper = m.row_upper.data();
  out->col_lower = m.col_lower.data();
  out->col_upper = m.col_upper.data();
  out->maximize = m.maximize ? 1 : 0;
  out->objective_offset = m.objective_offset;
  return ANUK_OK;
}

int anuk_mps_matrix(const anuk_mps* mps, const int32_t** start, const int32_t** index,
                    const double** value) {
  if (mps == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  return csr_out(mps->m.a, start, index, value);
}

int anuk_mps_hessian(const anuk_mps* mps, const int32_t** start, const int32_t** index,
                     const double** value) {
  if (mps == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  return csr_out(mps->m.hessian, start, index, value);
}

const int32_t* anuk_mps_col_types(const anuk_mps* mps) {
  return mps == nullptr ? nullptr : mps->m.col_type.data();
}

const char* anuk_mps_name(const anuk_mps* mps) { return mps == nullptr ? nullptr : mps->m.name.c_str(); }

const char* anuk_mps_row_name(const anuk_mps* mps, int32_t row) {
  if (mps == nullptr || row < 0 || static_cast<size_t>(row) >= mps->m.row_names.size()) return nullptr;
  return mps->m.row_names[static_cast<size_t>(row)].c_str();
}

const char* anuk_mps_col_name(const anuk_mps* mps, int32_t col) {
  if (mps == nullptr || col < 0 || static_cast<size_t>(col) >= mps->m.col_names.size()) return nullptr;
  return mps->m.col_names[static_cast<size_t>(col)].c_str();
}

const char* anuk_mps_warning(const anuk_mps* mps, int32_t k) {
  if (mps == nullptr || k < 0 || static_cast<size_t>(k) >= mps->m.warnings.size()) return nullptr;
  return mps->m.warnings[static_cast<size_t>(k)].c_str();
}

int anuk_mps_write(const char* path, const anuk_model* model, const char* name,
                   const char* const* row_names, const char* const* col_names,
                   const int32_t* col_types) {
  if (path == nullptr || model == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  return guarded(nullptr, [&] {
    need(model->rows >= 0 && model->cols >= 0 && model->nnz >= 0, "negative model size");
    need(model->cols == 0 || (model->objective && model->col_lower && model->col_upper), "null column data");
    need(model->rows == 0 || (model->row_lower && model->row_upper), "null row data");
    need(model->nnz == 0 || (model->row_index && model->col_index && model->value), "null triplets");
    std::vector<anuk::Entry> e(static_cast<size_t>(model->nnz));
    for (int64_t k = 0; k < model->nnz; ++k) e[static_cast<size_t>(k)] = {model->row_index[k], model->col_index[k], model->value[k]};
    const anuk::Csr a = anuk::Csr::from_triplets(model->rows, model->cols, std::move(e));
    anuk::MpsWriteModel w;
    w.name = name ? name : "";
    w.rows = model->rows;
    w.cols = model->cols;
    w.start = a.start.data();
    w.index = a.index.data();
    w.value = a.value.data();
    w.objective = model->objective;
    w.row_lower = model->row_lower;
    w.row_upper = model->row_upper;
    w.col_lower = model->col_lower;
    w.col_upper = model->col_upper;
    w.maximize = model->maximize != 0;
    w.objective_offset = model->objective_offset;
    w.row_names = row_names;
    w.col_names = col_names;
    w.col_type = col_types;
    if (col_types)
      for (int32_t j = 0; j < model->cols; ++j) need(col_types[j] >= 0 && col_types[j] <= 3, "column type out of range");
    const std::string text = anuk::write_mps_text(w);
    std::FILE* f = std::fopen(path, "wb");
    if (f == nullptr) throw anuk::FileError(std::string("cannot open \"") + path + "\" for writing");
    const size_t wrote = std::fwrite(text.data(), 1, text.size(), f);
    const int closed = std::fclose(f);
    if (wrote != text.size() || closed != 0) throw anuk::FileError(std::string("cannot write \"") + path + "\"");
  });
}


void anuk_presolve_default_rules(anuk_presolve_rules* out) {
  if (out == nullptr) return;
  const anuk::PresolveRules r;
  out->empty_rows = r.empty_rows;
  out->singleton_rows = r.singleton_rows;
  out->redundant_rows = r.redundant_rows;
  out->fixed_columns = r.fixed_columns;
  out->empty_columns = r.empty_columns;
  out->slack_column_singletons = r.slack_column_singletons;
  out->max_rounds = r.max_rounds;
  out->tolerance = r.tolerance;
  out->pivot_ratio = r.pivot_ratio;
}

int anuk_presolve_stack(int32_t rows, int32_t cols, const int32_t* start, const int32_t* index,
                        const double* value, int32_t cases, const double* objective,
                        const double* row_lower, const double* row_upper,
                        const double* col_lower, const double* col_upper, const double* offset,
                        const anuk_presolve_rules* rules, anuk_presolve** out) {
  if (out == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded(nullptr, [&] {
    need(rows >= 0 && cols >= 0 && cases >= 1, "presolve: negative sizes or no cases");
    need(start != nullptr, "presolve: null row starts");
    const size_t nc = static_cast<size_t>(cases) * static_cast<size_t>(cols);
    const size_t nr = static_cast<size_t>(cases) * static_cast<size_t>(rows);
    need(nc == 0 || (objective && col_lower && col_upper), "presolve: null column arrays");
    need(nr == 0 || (row_lower && row_upper), "presolve: null row arrays");
    anuk::PresolveRules r;
    if (rules != nullptr) {
      r.empty_rows = rules->empty_rows != 0;
      r.singleton_rows = rules->singleton_rows != 0;
      r.redundant_rows = rules->redundant_rows != 0;
      r.fixed_columns = rules->fixed_columns != 0;
      r.empty_columns = rules->empty_columns != 0;
      r.slack_column_singletons = rules->slack_column_singletons != 0;
      r.max_rounds = rules->max_rounds;
      r.tolerance = rules->tolerance;
      r.pivot_ratio = rules->pivot_ratio;
      need(r.max_rounds >= 0 && r.tolerance >= 0.0 && r.pivot_ratio >= 0.0 && r.pivot_ratio <= 1.0,
           "presolve: max_rounds, tolerance or pivot_ratio out of range");
    }
    need(start[0] == 0 && start[rows] >= 0, "presolve: row starts must run from 0 to nnz");
    const size_t nnz = static_cast<size_t>(start[rows]);
    need(nnz == 0 || (index && value), "presolve: null entries");
    anuk::LpStack s;
    s.rows = rows;
    s.cols = cols;
    s.cases = cases;
    s.a = anuk::Csr::from_csr(rows, cols, start, index, value);
    s.objective.assign(objective, objective + nc);
    s.col_lower.assign(col_lower, col_lower + nc);
    s.col_upper.assign(col_upper, col_upper + nc);
    s.row_lower.assign(row_lower, row_lower + nr);
    s.row_upper.assign(row_upper, row_upper + nr);
    s.offset.assign(static_cast<size_t>(cases), 0.0);
    if (offset) s.offset.assign(offset, offset + cases);
    auto h = std::make_unique<anuk_presolve>();
    h->p = anuk::presolve_stack(s, r);
    *out = h.release();
  });
}

void anuk_presolve_free(anuk_presolve* p) { delete p; }

int anuk_presolve_get_info(const anuk_presolve* p, anuk_presolve_info* out) {
  if (p == nullptr || out == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  const anuk::StackPresolve& s = p->p;
  const anuk::PresolveCounts& c = s.counts;
  out->status = static_cast<int32_t>(s.status);
  out->failing_case = s.failing_case;
  out->rows = s.reduced.rows;
  out->cols = s.reduced.cols;
  out->nnz = s.reduced.a.nnz();
  out->original_rows = s.original_rows();
  out->original_cols = s.original_cols();
  out->original_nnz = static_cast<int64_t>(s.reduced.a.nnz()) + c.nonzeros_removed;
  out->empty_rows = c.empty_rows;
  out->free_rows = c.free_rows;
  out->singleton_rows = c.singleton_rows;
  out->redundant_rows = c.redundant_rows;
  out->redundant_sides = c.redundant_sides;
  out->fixed_columns = c.fixed_columns;
  out->empty_columns = c.empty_columns;
  out->slack_column_singletons = c.slack_column_singletons;
  out->rounds = c.rounds;
  return ANUK_OK;
}

const char* anuk_presolve_message(const anuk_presolve* p) {
  return p == nullptr ? "" : p->p.message.c_str();
}

int anuk_presolve_reduced(const anuk_presolve* p, int32_t* start, int32_t* index, double* value,
                          double* objective, double* row_lower, double* row_upper,
                          double* col_lower, double* col_upper, double* offset,
                          int32_t* row_origin, int32_t* col_origin) {
  if (p == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  return guarded(nullptr, [&] {
    const anuk::StackPresolve& s = p->p;
    need(s.status == anuk::PresolveStatus::kReduced, "presolve: no reduced stack (see its status)");
    const anuk::LpStack& r = s.reduced;
    auto put = [](auto* dst, const auto& v) {
      if (dst != nullptr && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(start, r.a.start);
    put(index, r.a.index);
    put(value, r.a.value);
    put(objective, r.objective);
    put(row_lower, r.row_lower);
    put(row_upper, r.row_upper);
    put(col_lower, r.col_lower);
    put(col_upper, r.col_upper);
    put(offset, r.offset);
    put(row_origin, s.row_origin);
    put(col_origin, s.col_origin);
  });
}

int anuk_postsolve(const anuk_presolve* p, int32_t case_index, const double* reduced_x,
                   const double* reduced_y, double* x, double* y) {
  if (p == nullptr || x == nullptr || y == nullptr) return status_of(nullptr, ANUK_ERR_INVALID, "null argument");
  return guarded(nullptr, [&] {
    const anuk::StackPresolve& s = p->p;
    need((reduced_x != nullptr || s.reduced.cols == 0) && (reduced_y != nullptr || s.reduced.rows == 0),
         "postsolve: null reduced solution");
    s.postsolve(case_index, reduced_x, reduced_y, x, y);
  });
}

}  // extern "C"
