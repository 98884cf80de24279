// Test infrastructure only: a C entry layer over the reference's own CPU path,
// compiled from the reference sources where they lie (see oracle/Makefile) into
// oracle/_ref/libanukulan_ref.so. Only tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg load it, and only as the checker.
//
// Every function here is a thin marshalling wrapper around one reference
// function; the arithmetic is the reference's.
#include <cmath>
#include <chrono>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>
#include <vector>

#include "anukulan/backend.hpp"
#include "anukulan/crossover.hpp"
#include "anukulan/mps_reader.hpp"
#include "anukulan/pdhg.hpp"
#include "anukulan/presolve.hpp"
#include "anukulan/scaling.hpp"
#include "anukulan/simplex.hpp"
#include "anukulan/standard_form.hpp"
#include "anukulan/threading.hpp"

#include "anukulan_hip.h"

using namespace anukulan;

#ifdef ANUKULAN_REF_HIP
namespace anukulan {
const LinAlgBackend& hip_backend();  // hip_adapter.cpp
}
#endif

namespace {

thread_local std::string g_error;

struct RefLp {
  StandardLp lp;
};

SparseMatrix csr_to_matrix(Int rows, Int cols, const Int* start, const Int* index,
                           const double* value) {
  std::vector<Triplet> t;
  t.reserve(static_cast<std::size_t>(start[rows]));
  for (Int r = 0; r < rows; ++r)
    for (Int k = start[r]; k < start[r + 1]; ++k) t.push_back(Triplet{r, index[k], value[k]});
  return SparseMatrix::from_triplets(rows, cols, std::move(t));
}

PdhgOptions to_options(const anuk_pdhg_options* o) {
  PdhgOptions p;
  if (o == nullptr) return p;
  p.tolerance = o->tolerance;
  p.gap_tolerance = o->gap_tolerance;
  p.max_iterations = o->max_iterations;
  p.time_limit_seconds = o->time_limit_seconds;
  p.termination_check_frequency = o->termination_check_frequency;
  p.absolute_tolerance = o->absolute_tolerance;
  p.require_inf_norm_termination = o->require_inf_norm_termination != 0;
  p.adaptive_step_size = o->adaptive_step_size != 0;
  p.primal_weight_updates = o->primal_weight_updates != 0;
  p.restarts = o->restarts != 0;
  p.halpern = o->halpern != 0;
  p.halpern_restart = o->halpern_restart != 0;
  p.halpern_minimum_epoch = o->halpern_minimum_epoch;
  p.polish_feasibility = o->polish_feasibility != 0;
  p.polish_first_iteration = o->polish_first_iteration;
  p.polish_iteration_factor = o->polish_iteration_factor;
  p.polish_minimum_iterations = o->polish_minimum_iterations;
  p.polish_maximum_iterations = o->polish_maximum_iterations;
  p.polish_on_exit = o->polish_on_exit != 0;
  p.reflection = o->reflection;
  p.constant_step_size = o->constant_step_size != 0;
  p.constant_step_scale = o->constant_step_scale;
  p.restart_on_fixed_point = o->restart_on_fixed_point != 0;
  p.pid_primal_weight = o->pid_primal_weight != 0;
  p.primal_weight_kp = o->primal_weight_kp;
  p.primal_weight_ki = o->primal_weight_ki;
  p.primal_weight_kd = o->primal_weight_kd;
  p.scaling.ruiz_iterations = o->ruiz_iterations;
  p.scaling.pock_chambolle = o->pock_chambolle != 0;
  p.reuse_products = o->reuse_products != 0;
  p.detect_infeasibility = o->detect_infeasibility != 0;
  p.infeasibility_tolerance = o->infeasibility_tolerance;
  switch (o->termination) {
    case ANUK_TERM_PRIMAL_FEASIBILITY:
      p.termination = PdhgOptions::Termination::kPrimalFeasibility;
      break;
    case ANUK_TERM_DUAL_FEASIBILITY:
      p.termination = PdhgOptions::Termination::kDualFeasibility;
      break;
    default:
      p.termination = PdhgOptions::Termination::kFull;
      break;
  }
  p.known_matrix_norm = o->known_matrix_norm;
  p.initial_step_size = o->initial_step_size;
  p.initial_primal_weight = o->initial_primal_weight;
  p.verbose = o->verbose != 0;
  p.log_frequency = o->log_frequency;
  return p;
}

void fill_residual(const PdhgResidual& r, anuk_pdhg_residual* out) {
  out->primal_residual = r.primal_residual;
  out->dual_residual = r.dual_residual;
  out->primal_residual_inf = r.primal_residual_inf;
  out->dual_residual_inf = r.dual_residual_inf;
  out->primal_objective = r.primal_objective;
  out->dual_objective = r.dual_objective;
  out->absolute_gap = r.absolute_gap;
  out->relative_primal = r.relative_primal;
  out->relative_dual = r.relative_dual;
  out->relative_gap = r.relative_gap;
  out->relative_primal_inf = r.relative_primal_inf;
  out->relative_dual_inf = r.relative_dual_inf;
}

int fail(const std::exception& e) {
  g_error = e.what();
  return ANUK_ERR_INTERNAL;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

int ref_default_thread_count() { return default_thread_count(); }

void ref_pdhg_default_options(anuk_pdhg_options* o) {
  const PdhgOptions p;
  std::memset(o, 0, sizeof(*o));
  o->tolerance = p.tolerance;
  o->gap_tolerance = p.gap_tolerance;
  o->max_iterations = p.max_iterations;
  o->time_limit_seconds = p.time_limit_seconds;
  o->termination_check_frequency = p.termination_check_frequency;
  o->absolute_tolerance = p.absolute_tolerance;
  o->require_inf_norm_termination = p.require_inf_norm_termination;
  o->adaptive_step_size = p.adaptive_step_size;
  o->primal_weight_updates = p.primal_weight_updates;
  o->restarts = p.restarts;
  o->halpern = p.halpern;
  o->halpern_restart = p.halpern_restart;
  o->halpern_minimum_epoch = p.halpern_minimum_epoch;
  o->polish_feasibility = p.polish_feasibility;
  o->polish_first_iteration = p.polish_first_iteration;
  o->polish_iteration_factor = p.polish_iteration_factor;
  o->polish_minimum_iterations = p.polish_minimum_iterations;
  o->polish_maximum_iterations = p.polish_maximum_iterations;
  o->polish_on_exit = p.polish_on_exit;
  o->reflection = p.reflection;
  o->constant_step_size = p.constant_step_size;
  o->constant_step_scale = p.constant_step_scale;
  o->restart_on_fixed_point = p.restart_on_fixed_point;
  o->pid_primal_weight = p.pid_primal_weight;
  o->primal_weight_kp = p.primal_weight_kp;
  o->primal_weight_ki = p.primal_weight_ki;
  o->primal_weight_kd = p.primal_weight_kd;
  o->ruiz_iterations = p.scaling.ruiz_iterations;
  o->pock_chambolle = p.scaling.pock_chambolle;
  o->reuse_products = p.reuse_products;
  o->detect_infeasibility = p.detect_infeasibility;
  o->infeasibility_tolerance = p.infeasibility_tolerance;
  o->termination = ANUK_TERM_FULL;
  o->known_matrix_norm = p.known_matrix_norm;
  o->initial_step_size = p.initial_step_size;
  o->initial_primal_weight = p.initial_primal_weight;
  o->verbose = p.verbose;
  o->log_frequency = p.log_frequency;
  o->use_graphs = 1;     // the fields below have no reference counterpart:
  o->product_order = 0;  // the backend's own defaults, so the structs compare
  o->autotune = 1;
  o->batch_waves = 0;
}

// ---- LP construction (standard_form.cpp:72-165, mps_reader.cpp) ----------

void* ref_lp_from_model(const anuk_model* m) {
  try {
    Model model;
    model.sense = m->maximize ? ObjSense::kMaximize : ObjSense::kMinimize;
    model.objective_offset = m->objective_offset;
    model.objective.assign(m->objective, m->objective + m->cols);
    std::vector<Triplet> t(static_cast<std::size_t>(m->nnz));
    for (int64_t k = 0; k < m->nnz; ++k)
      t[static_cast<std::size_t>(k)] = Triplet{m->row_index[k], m->col_index[k], m->value[k]};
    model.constraints = SparseMatrix::from_triplets(m->rows, m->cols, std::move(t));
    auto inf_clip = [](double v) {
      if (v >= kFileInf) return kInf;
      if (v <= -kFileInf) return -kInf;
      return v;
    };
    model.row_lower.resize(static_cast<std::size_t>(m->rows));
    model.row_upper.resize(static_cast<std::size_t>(m->rows));
    for (Int i = 0; i < m->rows; ++i) {
      model.row_lower[sz(i)] = inf_clip(m->row_lower[i]);
      model.row_upper[sz(i)] = inf_clip(m->row_upper[i]);
    }
    model.col_lower.resize(static_cast<std::size_t>(m->cols));
    model.col_upper.resize(static_cast<std::size_t>(m->cols));
    for (Int j = 0; j < m->cols; ++j) {
      model.col_lower[sz(j)] = inf_clip(m->col_lower[j]);
      model.col_upper[sz(j)] = inf_clip(m->col_upper[j]);
    }
    model.col_type.assign(static_cast<std::size_t>(m->cols), VarType::kContinuous);
    StandardFormResult sf = to_standard_form(model);
    if (!sf.ok) {
      g_error = sf.error;
      return nullptr;
    }
    auto* out = new RefLp;
    out->lp = std::move(sf.lp);
    return out;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void* ref_lp_from_mps(const char* text) {
  try {
    std::istringstream in(text);
    const MpsReadResult r = read_mps_stream(in, "<text>");
    if (!r.ok) {
      g_error = r.error;
      return nullptr;
    }
    StandardFormResult sf = to_standard_form(r.model);
    if (!sf.ok) {
      g_error = sf.error;
      return nullptr;
    }
    auto* out = new RefLp;
    out->lp = std::move(sf.lp);
    return out;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// ---- The MPS reader (mps_reader.cpp:711-728), model exported as read --------

struct RefMps {
  MpsReadResult r;
  std::vector<int32_t> col_type;
};

// options: format 0 auto / 1 free / 2 fixed; negative_up 0 HiGHS / 1 CPLEX.
// Returns a handle even when the file is rejected (ref_mps_ok says so).
void* ref_mps_read_text(const char* text, int64_t length, int32_t format, int32_t negative_up,
                        double infinity_threshold) {
  try {
    MpsOptions o;
    o.format = format == 1 ? MpsOptions::Format::kFree
               : format == 2 ? MpsOptions::Format::kFixed
                             : MpsOptions::Format::kAuto;
    o.negative_up_bound = negative_up == 1 ? MpsOptions::NegativeUpBound::kMinusInfinity
                                           : MpsOptions::NegativeUpBound::kKeepLower;
    o.infinity_threshold = infinity_threshold;
    std::istringstream in(std::string(text, static_cast<std::size_t>(length)));
    auto* h = new RefMps;
    h->r = read_mps_stream(in, "<text>", o);
    for (VarType t : h->r.model.col_type) h->col_type.push_back(static_cast<int32_t>(t));
    return h;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_mps_free(void* h) { delete static_cast<RefMps*>(h); }

int ref_mps_ok(const void* h) { return static_cast<const RefMps*>(h)->r.ok ? 1 : 0; }
const char* ref_mps_error(const void* h) { return static_cast<const RefMps*>(h)->r.error.c_str(); }

// sizes[0..7]: rows, cols, nnz, hessian nnz, maximize, free rows dropped,
// objective nonzeros, fixed format; sizes[8]: warning count.
void ref_mps_sizes(const void* h, int64_t* sizes) {
  const MpsReadResult& r = static_cast<const RefMps*>(h)->r;
  sizes[0] = r.model.num_rows();
  sizes[1] = r.model.num_cols();
  sizes[2] = r.model.constraints.nnz();
  sizes[3] = r.model.hessian.nnz();
  sizes[4] = r.model.sense == ObjSense::kMaximize ? 1 : 0;
  sizes[5] = r.free_rows_dropped;
  sizes[6] = r.objective_nonzeros;
  sizes[7] = r.fixed_format ? 1 : 0;
  sizes[8] = static_cast<int64_t>(r.warnings.size());
}

double ref_mps_offset(const void* h) { return static_cast<const RefMps*>(h)->r.model.objective_offset; }
const char* ref_mps_name(const void* h) { return static_cast<const RefMps*>(h)->r.model.name.c_str(); }

// Copies A (and Q when non-empty) as CSR and the vectors; NULL skips one.
void ref_mps_arrays(const void* h, int32_t* start, int32_t* index, double* value,
                    int32_t* q_start, int32_t* q_index, double* q_value, double* objective,
                    double* row_lower, double* row_upper, double* col_lower, double* col_upper,
                    int32_t* col_type) {
  const RefMps* p = static_cast<const RefMps*>(h);
  const Model& m = p->r.model;
  auto put = [](auto* dst, const auto& src) {
    if (dst != nullptr && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
  };
  put(start, m.constraints.start());
  put(index, m.constraints.index());
  put(value, m.constraints.value());
  if (m.hessian.nnz() > 0) {
    put(q_start, m.hessian.start());
    put(q_index, m.hessian.index());
    put(q_value, m.hessian.value());
  }
  put(objective, m.objective);
  put(row_lower, m.row_lower);
  put(row_upper, m.row_upper);
  put(col_lower, m.col_lower);
  put(col_upper, m.col_upper);
  put(col_type, p->col_type);
}

const char* ref_mps_row_name(const void* h, int32_t i) {
  return static_cast<const RefMps*>(h)->r.model.row_names[static_cast<std::size_t>(i)].c_str();
}
const char* ref_mps_col_name(const void* h, int32_t j) {
  return static_cast<const RefMps*>(h)->r.model.col_names[static_cast<std::size_t>(j)].c_str();
}
const char* ref_mps_warning(const void* h, int32_t k) {
  return static_cast<const RefMps*>(h)->r.warnings[static_cast<std::size_t>(k)].c_str();
}

// ---- Presolve (presolve.cpp:1215-1317) ------------------------------------

struct RefPresolve {
  PresolveResult r;
  RefMps reduced;  // the reduced model, through the ref_mps_* accessors
};

// flags: the fourteen PresolveOptions switches in header order, then
// max_rounds; tols: feasibility, infeasibility, bound relaxation, max new
// finite bound. NULL for either: the defaults.
void* ref_presolve_mps(const char* text, int64_t length, const int32_t* flags, const double* tols) {
  try {
    std::istringstream in(std::string(text, static_cast<std::size_t>(length)));
    const MpsReadResult read = read_mps_stream(in, "<text>");
    if (!read.ok) {
      g_error = read.error;
      return nullptr;
    }
    PresolveOptions o;
    if (flags != nullptr) {
      bool* b[] = {&o.empty_rows, &o.singleton_rows, &o.redundant_rows, &o.forcing_rows,
                   &o.duplicate_rows, &o.fixed_columns, &o.empty_columns,
                   &o.free_column_singletons, &o.bound_tightening, &o.slack_column_singletons,
                   &o.inequality_column_singletons, &o.dual_fixing, &o.doubleton_equations,
                   &o.coefficient_tightening};
      for (int k = 0; k < 14; ++k) *b[k] = flags[k] != 0;
      o.max_rounds = flags[14];
    }
    if (tols != nullptr) {
      o.feasibility_tolerance = tols[0];
      o.infeasibility_tolerance = tols[1];
      o.bound_relaxation = tols[2];
      o.max_new_finite_bound = tols[3];
    }
    auto* h = new RefPresolve;
    h->r = presolve(read.model, o);
    h->reduced.r.ok = h->r.status == PresolveStatus::kReduced;
    h->reduced.r.model = h->r.reduced;
    for (VarType t : h->r.reduced.col_type) h->reduced.col_type.push_back(static_cast<int32_t>(t));
    return h;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_presolve_free(void* h) { delete static_cast<RefPresolve*>(h); }
int ref_presolve_status(const void* h) { return static_cast<int>(static_cast<const RefPresolve*>(h)->r.status); }
const char* ref_presolve_message(const void* h) { return static_cast<const RefPresolve*>(h)->r.message.c_str(); }
int ref_presolve_dual_exact(const void* h) {
  return static_cast<const RefPresolve*>(h)->r.postsolve.dual_is_exact() ? 1 : 0;
}
const void* ref_presolve_reduced(const void* h) { return &static_cast<const RefPresolve*>(h)->reduced; }

// counts[0..21]: rows, cols, nonzeros removed; empty, singleton, redundant,
// forcing, duplicate rows; fixed, empty columns; free, slack, inequality
// singletons, declined; doubletons; dual fixed; coefficients tightened;
// bounds tightened, made finite; rounds. counts_f[0]: largest new bound.
void ref_presolve_counts(const void* h, int64_t* counts, double* largest) {
  const PresolveCounts& c = static_cast<const RefPresolve*>(h)->r.counts;
  const Int v[] = {c.rows_removed, c.cols_removed, c.nonzeros_removed, c.empty_rows,
                   c.singleton_rows, c.redundant_rows, c.forcing_rows, c.duplicate_rows,
                   c.fixed_columns, c.empty_columns, c.free_column_singletons,
                   c.slack_column_singletons, c.inequality_column_singletons,
                   c.slack_singletons_declined, c.doubleton_equations, c.dual_fixed_columns,
                   c.coefficients_tightened, c.bounds_tightened, c.bounds_made_finite, c.rounds};
  for (int k = 0; k < 20; ++k) counts[k] = v[k];
  *largest = c.largest_new_finite_bound;
}

void ref_postsolve_primal(const void* h, const double* rx, int64_t n, double* x) {
  const RefPresolve* p = static_cast<const RefPresolve*>(h);
  const std::vector<double> full = p->r.postsolve.apply(std::vector<double>(rx, rx + n));
  std::memcpy(x, full.data(), full.size() * sizeof(double));
}

void ref_postsolve_dual(const void* h, const double* ry, int64_t m, const double* rc, int64_t n,
                        const double* oc, int64_t on, double* y) {
  const RefPresolve* p = static_cast<const RefPresolve*>(h);
  const std::vector<double> full = p->r.postsolve.apply_dual(
      std::vector<double>(ry, ry + m), std::vector<double>(rc, rc + n), std::vector<double>(oc, oc + on));
  std::memcpy(y, full.data(), full.size() * sizeof(double));
}

// The reduced model in standard form (to_standard_form), as a RefLp.
void* ref_presolve_lp(const void* h) {
  try {
    StandardFormResult sf = to_standard_form(static_cast<const RefPresolve*>(h)->r.reduced);
    if (!sf.ok) {
      g_error = sf.error;
      return nullptr;
    }
    auto* out = new RefLp;
    out->lp = std::move(sf.lp);
    return out;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void* ref_lp_from_standard(int32_t rows, int32_t cols, int32_t neq, const int32_t* start,
                           const int32_t* index, const double* value, const double* c,
                           const double* q, const double* lower, const double* upper,
                           double objective_scale, double objective_offset) {
  try {
    auto* out = new RefLp;
    StandardLp& lp = out->lp;
    lp.k = csr_to_matrix(rows, cols, start, index, value);
    lp.kt = lp.k.transpose();
    lp.c.assign(c, c + cols);
    lp.q.assign(q, q + rows);
    lp.lower.assign(lower, lower + cols);
    lp.upper.assign(upper, upper + cols);
    lp.num_equalities = neq;
    lp.objective_scale = objective_scale;
    lp.objective_offset = objective_offset;
    lp.row_origin.resize(static_cast<std::size_t>(rows));
    for (Int i = 0; i < rows; ++i) lp.row_origin[sz(i)] = StandardLp::RowOrigin{i, 1.0};
    return out;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_lp_free(void* h) { delete static_cast<RefLp*>(h); }

int ref_lp_info(const void* h, anuk_lp_info* out) {
  const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
  out->rows = lp.num_rows();
  out->cols = lp.num_cols();
  out->nnz = lp.k.nnz();
  out->num_equalities = lp.num_equalities;
  out->objective_scale = lp.objective_scale;
  out->objective_offset = lp.objective_offset;
  return 0;
}

int ref_lp_export(const void* h, int32_t* start, int32_t* index, double* value, double* c,
                  double* q, double* lower, double* upper, int32_t* row_model,
                  double* row_sign) {
  const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
  auto put_i = [](int32_t* dst, const std::vector<Int>& src) {
    if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(Int));
  };
  auto put_d = [](double* dst, const std::vector<double>& src) {
    if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(double));
  };
  put_i(start, lp.k.start());
  put_i(index, lp.k.index());
  put_d(value, lp.k.value());
  put_d(c, lp.c);
  put_d(q, lp.q);
  put_d(lower, lp.lower);
  put_d(upper, lp.upper);
  for (std::size_t i = 0; i < lp.row_origin.size(); ++i) {
    if (row_model) row_model[i] = lp.row_origin[i].model_row;
    if (row_sign) row_sign[i] = lp.row_origin[i].sign;
  }
  return 0;
}

// scale_lp (scaling.cpp:75-100) on a copy.
void* ref_lp_scale(const void* h, int32_t ruiz, int32_t pc, double* row_scale,
                   double* col_scale) {
  try {
    auto* out = new RefLp;
    out->lp = static_cast<const RefLp*>(h)->lp;
    ScalingOptions so;
    so.ruiz_iterations = ruiz;
    so.pock_chambolle = pc != 0;
    const ScalingReport rep = scale_lp(&out->lp, so);
    std::memcpy(row_scale, rep.scaling.row_scale.data(),
                rep.scaling.row_scale.size() * sizeof(double));
    std::memcpy(col_scale, rep.scaling.col_scale.data(),
                rep.scaling.col_scale.size() * sizeof(double));
    return out;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// ---- The solve (pdhg.cpp:377-1285) ----------------------------------------

int ref_solve_pdhg(const void* h, const anuk_pdhg_options* o, int32_t threads,
                   anuk_pdhg_result* res, double* x, double* y) {
  try {
    const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
    PdhgOptions p = to_options(o);
    std::vector<double> wx, wy;
    if (o != nullptr && o->warm_x != nullptr && o->warm_y != nullptr) {
      wx.assign(o->warm_x, o->warm_x + lp.num_cols());
      wy.assign(o->warm_y, o->warm_y + lp.num_rows());
      p.warm_x = &wx;
      p.warm_y = &wy;
    }
    p.backend = threads > 1 ? &threaded_cpu_backend(threads) : &cpu_backend();
#ifdef ANUKULAN_REF_HIP
    // threads < 0: the HIP library through the LinAlgBackend adapter
    // (hip_adapter.cpp), driven by the reference's own solve_pdhg.
    if (threads < 0) p.backend = &hip_backend();
#endif
    const PdhgResult r = solve_pdhg(lp, p);
    std::memset(res, 0, sizeof(*res));
    res->status = static_cast<int32_t>(r.status);
    res->iterations = r.iterations;
    res->restarts = r.restarts;
    res->polish_attempts = r.polish_attempts;
    res->polish_iterations = r.polish_iterations;
    res->polished = r.polished ? 1 : 0;
    res->objective = r.objective;
    res->solve_seconds = r.solve_seconds;
    res->matrix_norm_estimate = r.matrix_norm_estimate;
    res->final_step_size = r.final_step_size;
    res->final_primal_weight = r.final_primal_weight;
    res->phase_scaling = r.phases.scaling;
    res->phase_matrix_norm = r.phases.matrix_norm;
    res->phase_setup = r.phases.setup;
    res->phase_iterating = r.phases.iterating;
    res->phase_checks = r.phases.checks;
    res->phase_polishing = r.phases.polishing;
    fill_residual(r.residual, &res->residual);
    std::strncpy(res->message, r.message.c_str(), sizeof(res->message) - 1);
    if (x) std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    if (y) std::memcpy(y, r.y.data(), r.y.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

double ref_estimate_matrix_norm(const void* h, int32_t iterations, double tolerance,
                                int32_t use_backend) {
  const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
  return estimate_matrix_norm(lp, iterations, tolerance,
                              use_backend ? &cpu_backend() : nullptr);
}

int ref_evaluate_residual(const void* h, const double* x, const double* y,
                          anuk_pdhg_residual* out) {
  const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
  const std::vector<double> vx(x, x + lp.num_cols());
  const std::vector<double> vy(y, y + lp.num_rows());
  fill_residual(evaluate_residual(lp, vx, vy), out);
  return 0;
}

// The reference simplex (simplex.hpp:441-457): the FP64 optimum and the final
// basis. basic (rows entries) receives final_basic when non-null.
int ref_simplex(const void* h, double* x, double* y, double* objective, int32_t* status,
                int32_t* basic, int32_t* basic_count) {
  try {
    const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
    SimplexOptions so;
    const SimplexResult r = solve_lp(lp, so);
    *status = static_cast<int32_t>(r.status);
    *objective = r.objective;
    if (x) std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    if (y) std::memcpy(y, r.y.data(), r.y.size() * sizeof(double));
    if (basic_count) *basic_count = static_cast<int32_t>(r.final_basic.size());
    if (basic) std::memcpy(basic, r.final_basic.data(), r.final_basic.size() * sizeof(Int));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// The reference simplex with a chosen algorithm (dual: 1) and, when
// start_basic and start_status are given (start_count basic entries,
// start_status one code per variable of the logical form: 0 basic, 1 at
// lower, 2 at upper, 3 free), warm-started from that basis
// (SimplexOptions::start_basic/start_status, simplex.hpp:295-301). Returns
// the final basis and statuses in the same codes, the pivots and the
// seconds the simplex reports.
int ref_simplex_warm(const void* h, int32_t dual, const int32_t* start_basic,
                     int32_t start_count, const int32_t* start_status, int32_t status_count,
                     double* x, double* y, double* objective, int32_t* status, int32_t* basic,
                     int32_t* basic_count, int32_t* var_status, int32_t* var_count,
                     int64_t* pivots, double* seconds) {
  try {
    const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
    SimplexOptions so;
    so.algorithm = dual ? SimplexOptions::Algorithm::kDual : SimplexOptions::Algorithm::kPrimal;
    std::vector<Int> sb;
    std::vector<VarStatus> ss;
    if (start_basic != nullptr && start_status != nullptr) {
      sb.assign(start_basic, start_basic + start_count);
      for (int32_t k = 0; k < status_count; ++k) ss.push_back(static_cast<VarStatus>(start_status[k]));
      so.start_basic = &sb;
      so.start_status = &ss;
    }
    const SimplexResult r = solve_lp(lp, so);
    *status = static_cast<int32_t>(r.status);
    *objective = r.objective;
    if (x) std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    if (y) std::memcpy(y, r.y.data(), r.y.size() * sizeof(double));
    *basic_count = static_cast<int32_t>(r.final_basic.size());
    if (basic) std::memcpy(basic, r.final_basic.data(), r.final_basic.size() * sizeof(Int));
    *var_count = static_cast<int32_t>(r.final_status.size());
    if (var_status)
      for (size_t k = 0; k < r.final_status.size(); ++k)
        var_status[k] = static_cast<int32_t>(r.final_status[k]);
    *pivots = static_cast<int64_t>(r.iterations);
    *seconds = r.solve_seconds;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// The reference crossover (crossover.hpp:114-117) of a point x (cols) and
// dual y (rows, or null), default options: its counts in info[0..5]
// (ok, candidates, pushed, rejected_small_pivot, logicals_remaining,
// stopped_on_refactorization) and its basis (rows entries) in basic.
int ref_crossover(const void* h, const double* x, const double* y, int32_t* info, int32_t* basic) {
  try {
    const StandardLp& lp = static_cast<const RefLp*>(h)->lp;
    std::vector<double> xv(x, x + lp.num_cols());
    std::vector<double> yv;
    if (y) yv.assign(y, y + lp.num_rows());
    const CrossoverResult r = crossover_basis(lp, xv, yv);
    info[0] = r.ok ? 1 : 0;
    info[1] = static_cast<int32_t>(r.candidates);
    info[2] = static_cast<int32_t>(r.pushed);
    info[3] = static_cast<int32_t>(r.rejected_small_pivot);
    info[4] = static_cast<int32_t>(r.logicals_remaining);
    info[5] = r.stopped_on_refactorization ? 1 : 0;
    if (basic && r.ok) std::memcpy(basic, r.basic.data(), r.basic.size() * sizeof(Int));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// ---- The CPU reference backend (backend.cpp:9-202), host pointers ---------

static const LinAlgBackend& pick(int32_t threads) {
  return threads > 1 ? threaded_cpu_backend(threads) : cpu_backend();
}

void ref_op_multiply(int32_t threads, int32_t rows, int32_t cols, const int32_t* start,
                     const int32_t* index, const double* value, const double* x,
                     double* y) {
  const SparseMatrix a = csr_to_matrix(rows, cols, start, index, value);
  pick(threads).multiply(a, x, y);
}

// Seconds per call of the reference backend's multiply (threads as in
// ref_op_multiply) over `reps` calls, the matrix built once beforehand: the
// host SpMV time the CPU baseline reports.
double ref_time_multiply(int32_t threads, int32_t rows, int32_t cols, const int32_t* start,
                         const int32_t* index, const double* value, const double* x, double* y,
                         int32_t reps) {
  const SparseMatrix a = csr_to_matrix(rows, cols, start, index, value);
  const LinAlgBackend& be = pick(threads);
  be.multiply(a, x, y);
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t r = 0; r < reps; ++r) be.multiply(a, x, y);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count() / std::max<int32_t>(reps, 1);
}

void ref_op_transpose(int32_t rows, int32_t cols, const int32_t* start,
                      const int32_t* index, const double* value, int32_t* t_start,
                      int32_t* t_index, double* t_value) {
  const SparseMatrix a = csr_to_matrix(rows, cols, start, index, value);
  const SparseMatrix t = a.transpose();
  std::memcpy(t_start, t.start().data(), t.start().size() * sizeof(Int));
  std::memcpy(t_index, t.index().data(), t.index().size() * sizeof(Int));
  std::memcpy(t_value, t.value().data(), t.value().size() * sizeof(double));
}

double ref_op_dot(const double* a, const double* b, int32_t n) {
  return cpu_backend().dot(a, b, n);
}
double ref_op_two_norm(const double* a, int32_t n) { return cpu_backend().two_norm(a, n); }
double ref_op_inf_norm(const double* a, int32_t n) { return cpu_backend().inf_norm(a, n); }

void ref_op_primal_update(int32_t n, double tau, double gamma, double halpern,
                          const double* x, const double* c, const double* kt_y,
                          const double* lower, const double* upper, const double* anchor,
                          double* x_out, double* dx, double* x_bar) {
  cpu_backend().primal_update(n, tau, gamma, halpern, x, c, kt_y, lower, upper, anchor,
                              x_out, dx, x_bar);
}

void ref_op_dual_update(int32_t m, int32_t neq, double sigma, double gamma, double halpern,
                        const double* y, const double* q, const double* k_x_bar,
                        const double* k_x, const double* anchor_y,
                        const double* anchor_kx, double* y_out, double* dy, double* k_out,
                        double* k_dx) {
  cpu_backend().dual_update(m, neq, sigma, gamma, halpern, y, q, k_x_bar, k_x, anchor_y,
                            anchor_kx, y_out, dy, k_out, k_dx);
}

void ref_op_accumulate(int32_t n, double w, const double* v, double* sum) {
  cpu_backend().accumulate(n, w, v, sum);
}
void ref_op_scale_into(int32_t n, double w, const double* sum, double* out) {
  cpu_backend().scale_into(n, w, sum, out);
}
double ref_op_weighted_norm_squared(int32_t n, int32_t m, const double* dx,
                                    const double* dy, double omega) {
  return cpu_backend().weighted_norm_squared(n, m, dx, dy, omega);
}
void ref_op_step_size_terms(int32_t n, int32_t m, const double* dx, const double* dy,
                            const double* k_dx, double omega, double* interaction,
                            double* movement) {
  cpu_backend().step_size_terms(n, m, dx, dy, k_dx, omega, interaction, movement);
}
void ref_op_convergence_terms(int32_t n, int32_t m, const anuk_conv_problem* p,
                              const double* x, const double* k_x, const double* y,
                              const double* kt_y, anuk_conv_terms* out) {
  LinAlgBackend::ConvergenceProblem lp;
  lp.c = p->c;
  lp.lower = p->lower;
  lp.upper = p->upper;
  lp.q = p->q;
  lp.row_scale = p->row_scale;
  lp.col_scale = p->col_scale;
  lp.num_equalities = p->num_equalities;
  LinAlgBackend::ConvergenceTerms t;
  cpu_backend().convergence_terms(n, m, lp, x, k_x, y, kt_y, &t);
  out->primal_sq = t.primal_sq;
  out->primal_inf = t.primal_inf;
  out->q_dot_y = t.q_dot_y;
  out->dual_sq = t.dual_sq;
  out->dual_inf = t.dual_inf;
  out->bound_term = t.bound_term;
  out->c_dot_x = t.c_dot_x;
}

}  // extern "C"
