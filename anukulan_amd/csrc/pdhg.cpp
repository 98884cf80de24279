#include "pdhg.hpp"

#include "batch.hpp"
#include "host_pool.hpp"
#include "scaling.hpp"

#include <algorithm>
#include <future>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iterator>
#include <map>
#include <stdexcept>
#include <utility>

namespace anuk {

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

// Host weighted KKT error of a scaled pair (pdhg.cpp:295-340). Only the
// averaged-restart and KKT-restart configurations read it.
double weighted_kkt_error(const StandardLp& lp, const std::vector<double>& x,
                          const std::vector<double>& y, double omega) {
  std::vector<double> kx(uz(lp.rows())), kty(uz(lp.cols()));
  lp.k.multiply(x.data(), kx.data());
  lp.kt.multiply(y.data(), kty.data());
  double psq = 0.0;
  for (Int i = 0; i < lp.rows(); ++i) {
    const double s = kx[uz(i)] - lp.q[uz(i)];
    const double v = i < lp.num_equalities ? s : std::fmin(s, 0.0);
    psq += v * v;
  }
  double dsq = 0.0, bound = 0.0;
  for (Int j = 0; j < lp.cols(); ++j) {
    const double lambda = lp.c[uz(j)] - kty[uz(j)];
    const bool lo = lp.lower[uz(j)] > -kInf, hi = lp.upper[uz(j)] < kInf;
    const double left = (lo && hi) ? 0.0 : lo ? std::fmin(lambda, 0.0) : hi ? std::fmax(lambda, 0.0) : lambda;
    dsq += left * left;
    const double absorbed = lambda - left;
    if (absorbed > 0.0) {
      bound += absorbed * lp.lower[uz(j)];
    } else if (absorbed < 0.0) {
      bound += absorbed * lp.upper[uz(j)];
    }
  }
  const double gap = lp.standard_objective(x) - (host_dot(lp.q, y) + bound);
  return std::sqrt(omega * omega * psq + dsq / (omega * omega) + gap * gap);
}

// Farkas test on the host (pdhg.cpp:191-221), for the general loop.
bool host_farkas(const StandardLp& lp, const std::vector<double>& dy, double tol) {
  double scale = 0.0;
  for (Int i = 0; i < lp.rows(); ++i) {
    if (i >= lp.num_equalities && dy[uz(i)] < 0.0) return false;
    scale = std::fmax(scale, std::fabs(dy[uz(i)]));
  }
  if (scale <= 0.0) return false;
  std::vector<double> lambda(uz(lp.cols()));
  lp.kt.multiply(dy.data(), lambda.data());
  double support = 0.0;
  for (Int j = 0; j < lp.cols(); ++j) {
    const double l = lambda[uz(j)];
    if (l > 0.0) {
      if (lp.upper[uz(j)] >= kInf) return false;
      support += l * lp.upper[uz(j)];
    } else if (l < 0.0) {
      if (lp.lower[uz(j)] <= -kInf) return false;
      support += l * lp.lower[uz(j)];
    }
  }
  double rhs = 0.0;
  for (Int i = 0; i < lp.rows(); ++i) rhs += lp.q[uz(i)] * dy[uz(i)];
  return (rhs - support) / scale > tol;
}

// Nesting depth of solves on this thread: polishing sub-solves share the
// enclosing solve's device matrices, so only the outermost frees them.
thread_local int g_depth = 0;
// The backend's pinned slots for a solve's parameter and check staging, one
// per nesting depth from this base.
constexpr int kPinnedSlotBase = 800;

class Solver {
 public:
  Solver(HipBackend& be, const StandardLp& original, const anuk_pdhg_options& opt)
      : be_(be), orig_(original), opt_(opt) {}
  ~Solver();
  SolveResult run();

 private:
  // setup
  void allocate_device();
  void upload_problem(const std::vector<double>& x0, const std::vector<double>& y0);
  // loops
  void loop_fused();
  void loop_general();
  // one fused period of `steps` iterations ending in a check if has_check
  void run_period(int steps, bool has_check);
  void enqueue_steps(int steps, bool has_check, int parity);
  void enqueue_check(int parity);
  hipGraphExec_t graph_for(int parity);
  void bookkeep_step();
  // shared check decisions; returns true when the loop must stop
  bool after_check(Int iteration, const Residual& r, bool farkas, double candidate_kkt,
                   bool use_average);
  void restart(Int iteration);
  void pull_iterate();
  void run_polish(const std::vector<double>& sx, const std::vector<double>& sy, Int budget,
                  std::vector<double>* ox, std::vector<double>* oy, Residual* orr, Int* spent);
  Int polish_budget(Int at) const;
  void finish();
  double phase_end() {
    const auto now = Clock::now();
    const double s = std::chrono::duration<double>(now - phase_mark_).count();
    phase_mark_ = now;
    return s;
  }
  bool passes(const Residual& r) const;

  HipBackend& be_;
  const StandardLp& orig_;
  const anuk_pdhg_options& opt_;
  SolveResult res_;
  // The LP iterated on: the caller's own when no sweep scales it, else the
  // scaled copy.
  StandardLp scaled_;
  const StandardLp* lp_ = nullptr;
  Scaling scaling_;
  Clock::time_point start_, phase_mark_;
  Int n_ = 0, m_ = 0;
  bool outermost_ = false;

  // device state, carved from one slab
  double* slab_ = nullptr;
  std::size_t slab_doubles_ = 0;  // the slab's capacity
  double* xb_[2] = {nullptr, nullptr};
  double* yb_[2] = {nullptr, nullptr};
  double* kxb_[2] = {nullptr, nullptr};
  int cur_ = 0;
  double *x_bar_ = nullptr, *dx_ = nullptr, *dy_ = nullptr, *k_dx_ = nullptr, *k_x_bar_ = nullptr,
         *kt_y_ = nullptr;
  double *anchor_x_ = nullptr, *anchor_y_ = nullptr, *anchor_kx_ = nullptr;
  double *sum_x_ = nullptr, *sum_y_ = nullptr, *avg_x_ = nullptr, *avg_y_ = nullptr;
  double *part_a_ = nullptr, *part_b_ = nullptr, *part_c_ = nullptr, *part_d_ = nullptr;
  double* check_ = nullptr;
  IterParams* d_params_ = nullptr;
  double* h_pinned_ = nullptr;  // kCheckSlots doubles
  DevCsr dk_, dkt_;
  // Fast product order: long rows summed by waves, and K x_bar over K's
  // column panels when x_bar outgrows an XCD's L2 (the panel sums live in
  // panel_sums_, kPanels * m doubles).
  bool fast_ = false;
  bool panels_ = false;
  DevCsr dpk_, dk_fast_, dkt_fast_;
  double* panel_sums_ = nullptr;
  DevProblem prob_;
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs_;

  // algorithm state (pdhg.cpp:586-619, 753-755)
  double eta_ = 0.0, omega_ = 1.0, matrix_norm_ = 0.0;
  bool adaptive_ = false, halpern_ = false;
  double reflection_ = 0.0;
  double tol_ = 0.0, gap_tol_ = 0.0;
  bool anchor_needs_reset_ = true;
  Int halpern_k_ = 0;
  double movement_ = 0.0, epoch_start_movement_ = 0.0;
  double sum_weight_ = 0.0;
  std::vector<double> host_x_, host_y_, host_dy_, host_avg_x_, host_avg_y_;
  bool host_valid_ = false;
  std::vector<double> restart_x_, restart_y_, prev_restart_x_, prev_restart_y_;
  bool have_prev_restart_ = false;
  double kkt_at_restart_ = 0.0, last_candidate_kkt_ = kInf;
  double residual_at_restart_ = kInf, last_residual_ = kInf;
  double werr_sum_ = 0.0, werr_prev_ = 0.0;
  Int iteration_ = 0, since_restart_ = 0;
  int status_ = ANUK_PDHG_ITERATION_LIMIT;
  Int next_polish_ = 0;
  bool have_polished_ = false;
  std::vector<double> polished_x_, polished_y_;
  Residual polished_r_;
  bool feas_built_ = false;
  StandardLp primal_feas_, dual_feas_;
  ResidualScales* scales_ = nullptr;
};

Solver::~Solver() {
  // Teardown statuses are dropped: there is no caller left to report to.
  for (auto& kv : graphs_) {
    (void)hipGraphExecDestroy(kv.second.second);
    (void)hipGraphDestroy(kv.second.first);
  }
  if (slab_ != nullptr) be_.give_slab(slab_, slab_doubles_);
  delete scales_;
  --g_depth;
  if (outermost_ && lp_ != nullptr) {
    be_.release_one(lp_->k.id);
    be_.release_one(lp_->kt.id);
  }
}

bool Solver::passes(const Residual& r) const {
  const bool inf = opt_.require_inf_norm_termination != 0;
  switch (opt_.termination) {
    case ANUK_TERM_PRIMAL_FEASIBILITY:
      return r.primal_ok(tol_, opt_.absolute_tolerance, inf);
    case ANUK_TERM_DUAL_FEASIBILITY:
      return r.dual_ok(tol_, opt_.absolute_tolerance, inf);
    default:
      return r.converged(tol_, opt_.absolute_tolerance, inf, gap_tol_);
  }
}

Int Solver::polish_budget(Int at) const {
  Int b = static_cast<Int>(opt_.polish_iteration_factor * static_cast<double>(at + 1));
  b = std::max(b, static_cast<Int>(opt_.polish_minimum_iterations));
  b = std::min(b, static_cast<Int>(opt_.polish_maximum_iterations));
  return b;
}

void Solver::allocate_device() {
  const size_t N = static_cast<size_t>(n_), M = static_cast<size_t>(m_);
  const size_t kb =
      static_cast<size_t>(std::max({dk_.blocks, dkt_.blocks, panel_combine_blocks(m_)})) * kMaxRed;
  const size_t P = panels_ ? static_cast<size_t>(kPanels) * M : 0;
  // Every vector starts 256-byte aligned.
  auto round = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t layout[] = {
      N, N, N, N, N, N,            // xb0 xb1 x_bar dx anchor_x kt_y
      N, N, N, N, N, N, N, N,      // c lower upper orig_c orig_lower orig_upper col_scale sum_x
      N,                           // avg_x
      M, M, M, M, M, M, M, M, M,   // yb0 yb1 kx0 kx1 dy k_dx k_x_bar anchor_y anchor_kx
      M, M, M, M, M,               // q orig_q row_scale sum_y avg_y
      kb, kb, kb, kb,              // partials
      32, 32,                      // check, params
      P                            // panel sums
  };
  size_t total = 0;
  for (size_t d : layout) total += round(std::max<size_t>(d, 1));
  slab_ = be_.take_slab(total, &slab_doubles_);
  hip_check(hipMemsetAsync(slab_, 0, total * sizeof(double), be_.stream()), "hipMemsetAsync(state)");
  double* p = slab_;
  size_t li = 0;
  auto take = [&]() {
    double* q = p;
    p += round(std::max<size_t>(layout[li++], 1));
    return q;
  };
  xb_[0] = take(); xb_[1] = take(); x_bar_ = take(); dx_ = take(); anchor_x_ = take(); kt_y_ = take();
  double* c = take(); double* lo = take(); double* hi = take();
  double* oc = take(); double* olo = take(); double* ohi = take(); double* cs = take();
  sum_x_ = take(); avg_x_ = take();
  yb_[0] = take(); yb_[1] = take(); kxb_[0] = take(); kxb_[1] = take(); dy_ = take(); k_dx_ = take();
  k_x_bar_ = take(); anchor_y_ = take(); anchor_kx_ = take();
  double* q = take(); double* oq = take(); double* rs = take();
  sum_y_ = take(); avg_y_ = take();
  part_a_ = take(); part_b_ = take(); part_c_ = take(); part_d_ = take();
  check_ = take();
  d_params_ = reinterpret_cast<IterParams*>(take());
  panel_sums_ = take();
  prob_.n = n_;
  prob_.m = m_;
  prob_.num_equalities = lp_->num_equalities;
  prob_.c = c; prob_.lower = lo; prob_.upper = hi; prob_.q = q;
  prob_.orig_c = oc; prob_.orig_lower = olo; prob_.orig_upper = ohi; prob_.orig_q = oq;
  prob_.row_scale = rs; prob_.col_scale = cs;
  // Page-locked staging kept by the backend, one slot per nesting depth (a
  // polishing sub-solve runs inside its parent's solve).
  h_pinned_ = be_.pinned_slot(kPinnedSlotBase + g_depth, 64);
}

void Solver::upload_problem(const std::vector<double>& x0, const std::vector<double>& y0) {
  // One staged transfer for all twelve vectors (two page-locked chunks in
  // turn) rather than twelve synchronous copies from pageable memory.
  std::vector<HipBackend::UploadPiece> pieces;
  auto up = [&](const double* dst, const std::vector<double>& v) {
    pieces.push_back({v.data(), const_cast<double*>(dst), v.size() * sizeof(double)});
  };
  up(prob_.c, lp_->c);
  up(prob_.lower, lp_->lower);
  up(prob_.upper, lp_->upper);
  up(prob_.q, lp_->q);
  up(prob_.orig_c, orig_.c);
  up(prob_.orig_lower, orig_.lower);
  up(prob_.orig_upper, orig_.upper);
  up(prob_.orig_q, orig_.q);
  up(prob_.row_scale, scaling_.row_scale);
  up(prob_.col_scale, scaling_.col_scale);
  up(xb_[0], x0);
  up(yb_[0], y0);
  be_.upload_scatter(pieces);
}

void Solver::pull_iterate() {
  if (host_valid_) return;
  host_x_.resize(uz(n_));
  host_y_.resize(uz(m_));
  be_.download(xb_[cur_], host_x_.data(), n_);
  be_.download(yb_[cur_], host_y_.data(), m_);
  host_valid_ = true;
}

SolveResult Solver::run() {
  start_ = Clock::now();
  phase_mark_ = start_;
  outermost_ = (g_depth == 0);
  ++g_depth;
  n_ = orig_.cols();
  m_ = orig_.rows();
  const bool scales = opt_.ruiz_iterations > 0 || opt_.pock_chambolle != 0;
  if (scales) {
    scaled_ = orig_;
    lp_ = &scaled_;
    scaling_ = scaled_.k.nnz() >= kDeviceScalingNnz
                   ? scale_lp_device(be_, &scaled_, opt_.ruiz_iterations, opt_.pock_chambolle != 0)
                   : scale_lp(&scaled_, opt_.ruiz_iterations, opt_.pock_chambolle != 0);
  } else {
    lp_ = &orig_;
    // No sweep changes nothing: keep K, K^T and their identities, so a
    // polishing sub-solve reuses the enclosing solve's device copies.
    scaling_.row_scale.assign(uz(m_), 1.0);
    scaling_.col_scale.assign(uz(n_), 1.0);
  }
  res_.phases.scaling = phase_end();

  dk_ = be_.prepare(view_of(lp_->k));
  dkt_ = be_.prepare(view_of(lp_->kt));
  // The fast product order by default only where it pays: 4M nonzeros and up.
  fast_ = opt_.product_order == 2 || (opt_.product_order == 0 && lp_->k.nnz() >= (Int{1} << 22));
  panels_ = fast_ && static_cast<double>(n_) * sizeof(double) > kPanelVectorBytes && m_ > 0;
  // The panel stack is built on the host while the device runs the power
  // iteration (which needs only K and K^T), then uploaded.
  std::future<HipBackend::PanelStack> panel_build;
  if (panels_ && !be_.has_panels(view_of(lp_->k))) {
    const CsrView kv = view_of(lp_->k);
    panel_build = std::async(std::launch::async, [kv] { return HipBackend::build_panels(kv); });
  }
  if (fast_) dkt_fast_ = be_.prepare_fast(view_of(lp_->kt));
  if (fast_ && !panels_) dk_fast_ = be_.prepare_fast(view_of(lp_->k));

  matrix_norm_ = opt_.known_matrix_norm > 0.0 ? opt_.known_matrix_norm
                                             : estimate_matrix_norm(be_, *lp_, 200, 1e-8);
  if (panels_) dpk_ = panel_build.valid() ? be_.install_panels(panel_build.get())
                                          : be_.prepare_panels(view_of(lp_->k));
  res_.phases.matrix_norm = phase_end();
  res_.matrix_norm_estimate = matrix_norm_;
  if (matrix_norm_ <= 0.0) {
    // Nothing couples the columns: each sits at the bound its cost prefers
    // (pdhg.cpp:429-458).
    res_.x.assign(uz(n_), 0.0);
    for (Int j = 0; j < n_; ++j) {
      const double lo = orig_.lower[uz(j)], hi = orig_.upper[uz(j)], cj = orig_.c[uz(j)];
      double v = cj > 0.0 ? lo : cj < 0.0 ? hi : (std::isfinite(lo) ? lo : (std::isfinite(hi) ? hi : 0.0));
      res_.x[uz(j)] = std::isfinite(v) ? v : 0.0;
    }
    res_.y.assign(uz(m_), 0.0);
    res_.residual = evaluate_residual(orig_, res_.x, res_.y);
    res_.objective = orig_.model_objective(res_.x);
    res_.status = res_.residual.converged(opt_.tolerance, opt_.absolute_tolerance,
                                          opt_.require_inf_norm_termination != 0, 0.0)
                      ? ANUK_PDHG_OPTIMAL
                      : ANUK_PDHG_NUMERICAL_ERROR;
    res_.message = "no constraint matrix content";
    res_.solve_seconds = seconds_since(start_);
    return std::move(res_);
  }

  eta_ = 0.9 / matrix_norm_;
  adaptive_ = opt_.adaptive_step_size && !opt_.constant_step_size;
  if (opt_.constant_step_size) eta_ = opt_.constant_step_scale / matrix_norm_;
  omega_ = 1.0;
  if (opt_.primal_weight_updates) {
    const double cn = host_two_norm(lp_->c), qn = host_two_norm(lp_->q);
    if (cn > 1e-12 && qn > 1e-12) omega_ = cn / qn;
  }
  if (opt_.initial_step_size > 0.0) eta_ = opt_.initial_step_size;
  if (opt_.initial_primal_weight > 0.0) omega_ = opt_.initial_primal_weight;

  // Starting point: warm start placed in the dual cone and the primal box
  // (pdhg.cpp:479-503).
  std::vector<double> x(uz(n_), 0.0), y(uz(m_), 0.0);
  const bool warm = opt_.warm_x != nullptr && opt_.warm_y != nullptr;
  if (warm) {
    x.assign(opt_.warm_x, opt_.warm_x + n_);
    y.assign(opt_.warm_y, opt_.warm_y + m_);
    for (Int i = lp_->num_equalities; i < m_; ++i)
      if (!(y[uz(i)] > 0.0)) y[uz(i)] = 0.0;
  }
  for (Int j = 0; j < n_; ++j) {
    const double lo = lp_->lower[uz(j)], hi = lp_->upper[uz(j)];
    double v = warm ? x[uz(j)] : 0.0;
    if (!std::isfinite(v)) v = 0.0;
    if (v < lo) v = std::isfinite(lo) ? lo : 0.0;
    if (v > hi) v = std::isfinite(hi) ? hi : 0.0;
    x[uz(j)] = v;
  }

  allocate_device();
  upload_problem(x, y);
  launch_spmv(be_.stream(), dk_, xb_[0], kxb_[0]);
  cur_ = 0;
  scales_ = new ResidualScales(orig_);
  restart_x_ = x;
  restart_y_ = y;
  prev_restart_x_ = x;
  prev_restart_y_ = y;
  host_x_ = x;
  host_y_ = y;
  host_valid_ = false;

  tol_ = opt_.tolerance;
  gap_tol_ = opt_.gap_tolerance > 0.0 ? opt_.gap_tolerance : opt_.tolerance;
  halpern_ = opt_.halpern && opt_.restarts;
  reflection_ = halpern_ ? opt_.reflection : 0.0;
  if (!opt_.restart_on_fixed_point && opt_.restarts)
    kkt_at_restart_ = weighted_kkt_error(*lp_, x, y, omega_);
  next_polish_ = opt_.polish_first_iteration;
  res_.phases.setup = phase_end();

  const bool fused = !adaptive_ && !opt_.halpern_restart && !(opt_.restarts && !halpern_);
  if (fused) {
    loop_fused();
  } else {
    loop_general();
  }
  finish();
  return std::move(res_);
}

// ---- the fused loop ---------------------------------------------------------

void Solver::bookkeep_step() {
  ++since_restart_;
  if (halpern_) {
    if (anchor_needs_reset_) {
      anchor_needs_reset_ = false;
      halpern_k_ = 0;
    } else {
      ++halpern_k_;
    }
  }
}

void Solver::enqueue_steps(int steps, bool has_check, int parity) {
  hipStream_t s = be_.stream();
  int p = parity;
  // Algorithmic bytes per launch: the matrix once (8 B value + 4 B index per
  // nonzero, 4 B per row start), the gathered vector once, and the seven
  // vectors the epilogue reads and writes (eight or nine on the check step).
  const double nnz = static_cast<double>(lp_->k.nnz()), n = n_, m = m_;
  for (int step = 0; step < steps; ++step) {
    const bool diag = has_check && step == steps - 1;
    const unsigned flags = diag ? kStepDiag : 0u;
    PrimalBufs pb{xb_[p], anchor_x_, xb_[1 - p], x_bar_, dx_, anchor_x_};
    const LaunchEvents pe = be_.launch_events(diag ? "primal_step_check" : "primal_step",
                                              12.0 * nnz + 4.0 * (n + 1) + 8.0 * m +
                                                  (diag ? 64.0 : 56.0) * n);
    launch_primal_step(s, fast_ ? dkt_fast_ : dkt_, yb_[p], prob_, d_params_, step, pb, flags,
                       part_a_, &pe, fast_);
    DualBufs db{yb_[p], kxb_[p], anchor_y_, anchor_kx_, yb_[1 - p], kxb_[1 - p],
                dy_, k_dx_, anchor_y_, anchor_kx_};
    const LaunchEvents de = be_.launch_events(diag ? "dual_step_check" : "dual_step",
                                              12.0 * nnz + 4.0 * (m + 1) + 8.0 * n +
                                                  (diag ? 72.0 : 56.0) * m);
    if (panels_) {
      // Two dispatches: the panel products carry the step's bytes, the
      // combine is timed as its own entry (its partial sums are re-read).
      const LaunchEvents ce =
          be_.launch_events(diag ? "dual_step_combine_check" : "dual_step_combine", 0.0);
      launch_dual_step_panels(s, dpk_, panel_sums_, x_bar_, prob_, d_params_, step, db, flags,
                              part_b_, &de, &ce);
    } else {
      launch_dual_step(s, fast_ ? dk_fast_ : dk_, x_bar_, prob_, d_params_, step, db, flags, part_b_,
                       &de, fast_);
    }
    p = 1 - p;
  }
  if (has_check) enqueue_check(p);
}

void Solver::enqueue_check(int p) {
  hipStream_t s = be_.stream();
  HipBackend::Timed t(be_, "check_reductions");
  const DevCsr& kt = fast_ ? dkt_fast_ : dkt_;
  const Int dual_blocks = panels_ ? panel_combine_blocks(m_) : (fast_ ? dk_fast_ : dk_).blocks;
  launch_check_fold_rows(s, kt.blocks, part_a_, dual_blocks, part_b_, check_);
  launch_conv_product(s, kt, yb_[p], xb_[p], prob_, part_c_, fast_);
  const bool farkas = opt_.detect_infeasibility != 0;
  if (farkas) launch_farkas_product(s, kt, dy_, prob_, check_, part_d_, fast_);
  launch_check_fold_cols(s, kt.blocks, part_c_, kt.blocks, farkas ? part_d_ : nullptr, check_);
}

hipGraphExec_t Solver::graph_for(int parity) {
  // The Halpern state at the period's start is read by the kernels from
  // device memory, so one graph per starting buffer parity serves every period.
  const int key = parity;
  auto it = graphs_.find(key);
  if (it != graphs_.end()) return it->second.second;
  hipStream_t s = be_.stream();
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    enqueue_steps(opt_.termination_check_frequency, true, parity);
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  graphs_[key] = {g, e};
  return e;
}

void Solver::run_period(int steps, bool has_check) {
  IterParams P;
  P.tau = eta_ / omega_;
  P.sigma = eta_ * omega_;
  P.gamma = reflection_;
  P.halpern = halpern_ ? 1 : 0;
  P.reset_at_start = anchor_needs_reset_ ? 1 : 0;
  P.k_at_start = halpern_k_;
  P.num_equalities = lp_->num_equalities;
  std::memcpy(h_pinned_ + 32, &P, sizeof(P));
  hip_check(hipMemcpyAsync(d_params_, h_pinned_ + 32, sizeof(P), hipMemcpyHostToDevice, be_.stream()),
            "upload step scalars");
  const bool graph = opt_.use_graphs != 0 && has_check && steps == opt_.termination_check_frequency &&
                     !be_.profiling();
  if (graph) {
    hip_check(hipGraphLaunch(graph_for(cur_), be_.stream()), "hipGraphLaunch");
    ++res_.graph_launches;
  } else {
    enqueue_steps(steps, has_check, cur_);
    ++res_.direct_periods;
  }
  for (int s = 0; s < steps; ++s) bookkeep_step();
  if (steps % 2 == 1) cur_ = 1 - cur_;
  host_valid_ = false;
  if (has_check) {
    hip_check(hipMemcpyAsync(h_pinned_, check_, kCheckSlots * sizeof(double), hipMemcpyDeviceToHost,
                             be_.stream()),
              "fetch check");
  }
  be_.synchronize();
}

void Solver::loop_fused() {
  const Int freq = opt_.termination_check_frequency;
  while (iteration_ < opt_.max_iterations) {
    const Int period_end = std::min<Int>((iteration_ / freq + 1) * freq, opt_.max_iterations);
    const int steps = period_end - iteration_;
    const bool has_check = (period_end % freq == 0);
    run_period(steps, has_check);
    iteration_ = period_end;
    if (!has_check) continue;
    res_.phases.iterating += phase_end();
    const Int i = period_end - 1;  // the check iteration's own index
    const double* h = h_pinned_;
    movement_ = omega_ * h[kSumDx2] + h[kSumDy2] / omega_;
    bool farkas = false;
    if (opt_.detect_infeasibility) {
      farkas = h[kFarkasRowBad] == 0.0 && h[kFarkasScale] > 0.0 && h[kFarkasColBad] == 0.0 &&
               (h[kFarkasRhs] - h[kFarkasSupport]) / h[kFarkasScale] > opt_.infeasibility_tolerance;
    }
    Terms t;
    t.primal_sq = h[kPrimalSq];
    t.primal_inf = h[kPrimalInf];
    t.q_dot_y = h[kQDotY];
    t.dual_sq = h[kDualSq];
    t.dual_inf = h[kDualInf];
    t.bound_term = h[kBoundTerm];
    t.c_dot_x = h[kCDotX];
    Residual r;
    if (opt_.reuse_products) {
      r = residual_from_terms(t, *scales_);
    } else {
      pull_iterate();
      std::vector<double> ux, uy;
      scaling_.unscale_primal(host_x_, &ux);
      scaling_.unscale_dual(host_y_, &uy);
      r = evaluate_residual(orig_, ux, uy);
    }
    double kkt = 0.0;
    const bool need_kkt = (opt_.restarts && !opt_.restart_on_fixed_point) || opt_.verbose;
    if (need_kkt) {
      pull_iterate();
      kkt = weighted_kkt_error(*lp_, host_x_, host_y_, omega_);
    }
    const bool stop = after_check(i, r, farkas, kkt, false);
    res_.phases.checks += phase_end();
    if (stop) break;
  }
}

// ---- the general loop: one backend call per operation, as pdhg.cpp runs it ---

void Solver::loop_general() {
  std::vector<double> host_dy;
  const double* anchor_null = nullptr;
  for (; iteration_ < opt_.max_iterations; ++iteration_) {
    const bool check = ((iteration_ + 1) % opt_.termination_check_frequency == 0);
    const bool blend = halpern_ && !anchor_needs_reset_;
    const double w = blend ? static_cast<double>(halpern_k_ + 1) / static_cast<double>(halpern_k_ + 2) : 1.0;
    const bool need_movement = adaptive_ || opt_.halpern_restart || check;
    const int p = cur_;
    launch_spmv(be_.stream(), dkt_, yb_[p], kt_y_);
    for (int attempt = 0;; ++attempt) {
      const double tau = eta_ / omega_, sigma = eta_ * omega_;
      be_.primal_update(n_, tau, reflection_, w, xb_[p], prob_.c, kt_y_, prob_.lower, prob_.upper,
                        blend ? anchor_x_ : anchor_null, xb_[1 - p], dx_, x_bar_);
      launch_spmv(be_.stream(), dk_, x_bar_, k_x_bar_);
      be_.dual_update(m_, lp_->num_equalities, sigma, reflection_, w, yb_[p], prob_.q, k_x_bar_,
                      kxb_[p], blend ? anchor_y_ : anchor_null, blend ? anchor_kx_ : anchor_null,
                      yb_[1 - p], dy_, kxb_[1 - p], k_dx_);
      if (!adaptive_) {
        if (need_movement) movement_ = be_.weighted_norm_squared(n_, m_, dx_, dy_, omega_);
        break;
      }
      double interaction = 0.0;
      be_.step_size_terms(n_, m_, dx_, dy_, k_dx_, omega_, &interaction, &movement_);
      double eta_bar = kInf;
      if (interaction > 0.0) eta_bar = movement_ / (2.0 * interaction);
      const double k1 = static_cast<double>(iteration_ + 1);
      double eta_next = eta_;
      if (std::isfinite(eta_bar)) {
        eta_next = std::fmin((1.0 - std::pow(k1, -0.3)) * eta_bar, (1.0 + std::pow(k1, -0.6)) * eta_);
      } else {
        eta_next = (1.0 + std::pow(k1, -0.6)) * eta_;
      }
      if (!(eta_next > 0.0) || !std::isfinite(eta_next)) eta_next = eta_;
      const bool accept = (eta_ <= eta_bar);
      eta_ = eta_next;
      if (accept || attempt > 60) break;
    }
    cur_ = 1 - p;
    host_valid_ = false;
    ++since_restart_;
    if (halpern_) {
      if (anchor_needs_reset_) {
        be_.copy(xb_[cur_], anchor_x_, n_);
        be_.copy(yb_[cur_], anchor_y_, m_);
        be_.copy(kxb_[cur_], anchor_kx_, m_);
        anchor_needs_reset_ = false;
        halpern_k_ = 0;
        epoch_start_movement_ = movement_;
      } else {
        ++halpern_k_;
      }
    }
    if (opt_.restarts && !halpern_) {
      be_.accumulate(n_, eta_, xb_[cur_], sum_x_);
      be_.accumulate(m_, eta_, yb_[cur_], sum_y_);
      sum_weight_ += eta_;
    }
    if (halpern_ && opt_.halpern_restart && !anchor_needs_reset_ && epoch_start_movement_ > 0.0 &&
        halpern_k_ >= opt_.halpern_minimum_epoch) {
      if (movement_ <= epoch_start_movement_ / 7.389056098930650) {
        anchor_needs_reset_ = true;
        ++res_.restarts;
      }
    }
    if (!check) continue;
    be_.synchronize();
    res_.phases.iterating += phase_end();

    bool farkas = false;
    if (opt_.detect_infeasibility) {
      host_dy.resize(uz(m_));
      be_.download(dy_, host_dy.data(), m_);
      farkas = host_farkas(*lp_, host_dy, opt_.infeasibility_tolerance);
    }
    const bool need_kkt = (opt_.restarts && !halpern_) ||
                          (opt_.restarts && !opt_.restart_on_fixed_point) || opt_.verbose;
    const bool need_average = opt_.restarts && !halpern_ && sum_weight_ > 0.0;
    if (need_kkt || need_average) pull_iterate();
    const double current_kkt = need_kkt ? weighted_kkt_error(*lp_, host_x_, host_y_, omega_) : 0.0;
    bool use_average = false;
    double candidate_kkt = current_kkt;
    if (need_average) {
      be_.scale_into(n_, sum_weight_, sum_x_, avg_x_);
      be_.scale_into(m_, sum_weight_, sum_y_, avg_y_);
      host_avg_x_.resize(uz(n_));
      host_avg_y_.resize(uz(m_));
      be_.download(avg_x_, host_avg_x_.data(), n_);
      be_.download(avg_y_, host_avg_y_.data(), m_);
      const double avg_kkt = weighted_kkt_error(*lp_, host_avg_x_, host_avg_y_, omega_);
      if (avg_kkt < current_kkt) {
        use_average = true;
        candidate_kkt = avg_kkt;
      }
    }
    Residual r;
    if (opt_.reuse_products && !use_average) {
      launch_spmv(be_.stream(), dkt_, yb_[cur_], kt_y_);
      const Terms t = be_.convergence_terms(n_, m_, prob_, xb_[cur_], kxb_[cur_], yb_[cur_], kt_y_);
      r = residual_from_terms(t, *scales_);
    } else {
      pull_iterate();
      std::vector<double> ux, uy;
      scaling_.unscale_primal(use_average ? host_avg_x_ : host_x_, &ux);
      scaling_.unscale_dual(use_average ? host_avg_y_ : host_y_, &uy);
      r = evaluate_residual(orig_, ux, uy);
    }
    const bool stop = after_check(iteration_, r, farkas, candidate_kkt, use_average);
    res_.phases.checks += phase_end();
    if (stop) {
      ++iteration_;
      break;
    }
  }
}

// ---- decisions at a termination check (pdhg.cpp:926-1202) --------------------
//
// Returns true when the loop stops. In the fused loop `iteration_` already
// points past the check; in the general loop the caller advances it.
bool Solver::after_check(Int i, const Residual& r, bool farkas, double candidate_kkt,
                         bool use_average) {
  auto adopt = [&]() {
    if (!use_average) return;
    be_.copy(avg_x_, xb_[cur_], n_);
    be_.copy(avg_y_, yb_[cur_], m_);
    host_x_ = host_avg_x_;
    host_y_ = host_avg_y_;
    host_valid_ = true;
  };
  if (farkas) {
    status_ = ANUK_PDHG_PRIMAL_INFEASIBLE;
    res_.message = "Farkas certificate found in the dual iterate difference";
    return true;
  }
  if (!std::isfinite(r.primal_residual) || !std::isfinite(r.dual_residual)) {
    status_ = ANUK_PDHG_NUMERICAL_ERROR;
    res_.message = "iterates stopped being finite";
    adopt();
    return true;
  }
  if (passes(r)) {
    status_ = ANUK_PDHG_OPTIMAL;
    adopt();
    return true;
  }
  if (opt_.polish_feasibility && opt_.termination == ANUK_TERM_FULL && i + 1 >= next_polish_) {
    if (next_polish_ < 1) next_polish_ = 1;
    while (next_polish_ <= i + 1) next_polish_ *= 2;
    const Int budget = polish_budget(i);
    if (r.gap_ok(gap_tol_) && budget > 0) {
      pull_iterate();
      std::vector<double> tx, ty;
      Residual tr;
      Int spent = 0;
      const std::vector<double> fx = use_average ? host_avg_x_ : host_x_;
      const std::vector<double> fy = use_average ? host_avg_y_ : host_y_;
      res_.phases.checks += phase_end();
      run_polish(fx, fy, budget, &tx, &ty, &tr, &spent);
      res_.phases.polishing += phase_end();
      ++res_.polish_attempts;
      res_.polish_iterations += spent;
      if (tr.converged(tol_, opt_.absolute_tolerance, opt_.require_inf_norm_termination != 0,
                       gap_tol_)) {
        polished_x_ = std::move(tx);
        polished_y_ = std::move(ty);
        polished_r_ = tr;
        have_polished_ = true;
        res_.polished = true;
        status_ = ANUK_PDHG_OPTIMAL;
        res_.message = "converged after feasibility polishing";
        return true;
      }
    }
  }
  if (opt_.verbose && (i + 1) % opt_.log_frequency == 0) {
    std::printf("  iter %7d  primal %.3e  dual %.3e  gap %.3e  kkt %.3e  eta %.3e  w %.3e\n", i + 1,
                r.relative_primal, r.relative_dual, r.relative_gap, candidate_kkt, eta_, omega_);
  }
  if (seconds_since(start_) > opt_.time_limit_seconds) {
    status_ = ANUK_PDHG_TIME_LIMIT;
    adopt();
    return true;
  }
  if (opt_.stop != nullptr && *opt_.stop != 0) {
    status_ = ANUK_PDHG_TIME_LIMIT;
    adopt();
    return true;
  }
  if (!opt_.restarts) return false;

  const double residual = std::sqrt(std::fmax(movement_, 0.0));
  const bool by_residual = opt_.restart_on_fixed_point != 0;
  if (by_residual && !std::isfinite(residual_at_restart_)) residual_at_restart_ = residual;
  const double measure = by_residual ? residual : candidate_kkt;
  const double at_restart = by_residual ? residual_at_restart_ : kkt_at_restart_;
  const double previous = by_residual ? last_residual_ : last_candidate_kkt_;
  const bool sufficient = measure <= 0.2 * at_restart;
  const bool necessary = measure <= 0.8 * at_restart && measure > previous;
  const bool too_long = static_cast<double>(since_restart_) >= 0.36 * static_cast<double>(i + 1);
  last_candidate_kkt_ = candidate_kkt;
  last_residual_ = residual;
  if (!(sufficient || necessary || too_long)) return false;
  adopt();
  restart(i);
  return false;
}

void Solver::restart(Int) {
  launch_spmv(be_.stream(), dk_, xb_[cur_], kxb_[cur_]);
  pull_iterate();
  if (opt_.primal_weight_updates && have_prev_restart_) {
    const std::vector<double>& bx = opt_.pid_primal_weight ? restart_x_ : prev_restart_x_;
    const std::vector<double>& by = opt_.pid_primal_weight ? restart_y_ : prev_restart_y_;
    double dxs = 0.0, dys = 0.0;
    for (Int j = 0; j < n_; ++j) {
      const double d = host_x_[uz(j)] - bx[uz(j)];
      dxs += d * d;
    }
    for (Int i = 0; i < m_; ++i) {
      const double d = host_y_[uz(i)] - by[uz(i)];
      dys += d * d;
    }
    const double ddx = std::sqrt(dxs), ddy = std::sqrt(dys);
    if (ddx > 1e-12 && ddy > 1e-12) {
      double cand = 0.0;
      if (opt_.pid_primal_weight) {
        // Error in log space, positive when the primal moved further.
        const double e = std::log(omega_ * ddx / ddy);
        werr_sum_ += e;
        const double correction = opt_.primal_weight_kp * e + opt_.primal_weight_ki * werr_sum_ +
                                  opt_.primal_weight_kd * (e - werr_prev_);
        werr_prev_ = e;
        cand = std::exp(std::log(omega_) - correction);
      } else {
        cand = std::exp(0.5 * std::log(ddy / ddx) + 0.5 * std::log(omega_));
      }
      if (std::isfinite(cand) && cand > 1e-12 && cand < 1e12) omega_ = cand;
    }
  }
  prev_restart_x_ = restart_x_;
  prev_restart_y_ = restart_y_;
  have_prev_restart_ = true;
  restart_x_ = host_x_;
  restart_y_ = host_y_;
  if (opt_.restarts && !halpern_) {
    be_.fill(sum_x_, n_, 0.0);
    be_.fill(sum_y_, m_, 0.0);
  }
  sum_weight_ = 0.0;
  anchor_needs_reset_ = true;
  since_restart_ = 0;
  last_candidate_kkt_ = kInf;
  last_residual_ = kInf;
  residual_at_restart_ = std::sqrt(std::fmax(movement_, 0.0));
  if (!opt_.restart_on_fixed_point) kkt_at_restart_ = weighted_kkt_error(*lp_, host_x_, host_y_, omega_);
  ++res_.restarts;
}

// ---- feasibility polishing (pdhg.cpp:680-751) --------------------------------

void Solver::run_polish(const std::vector<double>& sx, const std::vector<double>& sy, Int budget,
                        std::vector<double>* ox, std::vector<double>* oy, Residual* orr,
                        Int* spent) {
  // The two feasibility problems share this solve's K and K^T: when the
  // matrices are this solve's own scaled copy they are lent for each
  // sub-solve's duration (moved in and back, identities with them, so the
  // device copies are reused); when they are the caller's, each problem holds
  // a copy. Primal drops the objective, dual the right-hand side and the
  // finite bound values.
  const bool lend = lp_ == &scaled_;
  if (!feas_built_) {
    // Each vector of both problems is its own host task: at a million
    // columns, first-touch page faults dominate a serial build.
    const StandardLp& s = *lp_;
    auto finite_to_zero = [](std::vector<double>* v, const std::vector<double>& from, double inf) {
      v->resize(from.size());
      for (std::size_t j = 0; j < from.size(); ++j) (*v)[j] = from[j] == inf ? inf : 0.0;
    };
    const std::function<void()> build[] = {
        [&] { primal_feas_.c.assign(s.c.size(), 0.0); },
        [&] { primal_feas_.q = s.q; },
        [&] { primal_feas_.lower = s.lower; },
        [&] { primal_feas_.upper = s.upper; },
        [&] { primal_feas_.row_model = s.row_model; primal_feas_.row_sign = s.row_sign; },
        [&] { dual_feas_.c = s.c; },
        [&] { dual_feas_.q.assign(s.q.size(), 0.0); },
        [&] { finite_to_zero(&dual_feas_.lower, s.lower, -kInf); },
        [&] { finite_to_zero(&dual_feas_.upper, s.upper, kInf); },
        [&] { dual_feas_.row_model = s.row_model; dual_feas_.row_sign = s.row_sign; },
        [&] { if (!lend) { primal_feas_.k = s.k; primal_feas_.kt = s.kt; } },
        [&] { if (!lend) { dual_feas_.k = s.k; dual_feas_.kt = s.kt; } },
    };
    HostPool::shared().run(std::size(build), [&](std::size_t t) { build[t](); });
    for (StandardLp* f : {&primal_feas_, &dual_feas_}) {
      f->num_equalities = s.num_equalities;
      f->objective_scale = s.objective_scale;
      f->objective_offset = 0.0;
    }
    feas_built_ = true;
  }
  struct Lent {  // moves K and K^T into a feasibility problem and back
    StandardLp& from;
    StandardLp& to;
    bool on;
    Lent(StandardLp& a, StandardLp& b, bool lend) : from(a), to(b), on(lend) {
      if (on) {
        std::swap(from.k, to.k);
        std::swap(from.kt, to.kt);
      }
    }
    ~Lent() {
      if (on) {
        std::swap(from.k, to.k);
        std::swap(from.kt, to.kt);
      }
    }
  };
  anuk_pdhg_options sub = opt_;
  sub.polish_feasibility = 0;
  sub.polish_on_exit = 0;
  sub.detect_infeasibility = 0;
  sub.verbose = 0;
  sub.known_matrix_norm = matrix_norm_;
  sub.ruiz_iterations = 0;
  sub.pock_chambolle = 0;
  sub.max_iterations = budget;
  sub.termination_check_frequency =
      std::max<Int>(10, std::min<Int>(opt_.termination_check_frequency, budget / 4));
  sub.time_limit_seconds = std::fmax(0.0, opt_.time_limit_seconds - seconds_since(start_));
  sub.initial_step_size = eta_;
  sub.initial_primal_weight = omega_;
  const std::vector<double> zx(uz(n_), 0.0), zy(uz(m_), 0.0);

  sub.termination = ANUK_TERM_PRIMAL_FEASIBILITY;
  sub.warm_x = sx.data();
  sub.warm_y = zy.data();
  SolveResult pp;
  {
    const Lent lent(scaled_, primal_feas_, lend);
    pp = Solver(be_, primal_feas_, sub).run();
  }

  sub.termination = ANUK_TERM_DUAL_FEASIBILITY;
  sub.warm_x = zx.data();
  sub.warm_y = sy.data();
  SolveResult dp;
  {
    const Lent lent(scaled_, dual_feas_, lend);
    dp = Solver(be_, dual_feas_, sub).run();
  }

  *spent = pp.iterations + dp.iterations;
  scaling_.unscale_primal(pp.x, ox);
  scaling_.unscale_dual(dp.y, oy);
  *orr = evaluate_residual(orig_, *ox, *oy);
  if (opt_.verbose) {
    std::printf("  polish  spent %5d (%d primal, %d dual)  ->  primal %.3e  dual %.3e  gap %.3e\n",
                *spent, pp.iterations, dp.iterations, orr->relative_primal, orr->relative_dual,
                orr->relative_gap);
  }
}

void Solver::finish() {
  if (!have_polished_) {
    pull_iterate();
    scaling_.unscale_primal(host_x_, &res_.x);
    scaling_.unscale_dual(host_y_, &res_.y);
    // A polishing sub-solve's caller reads only its point and iteration
    // count (and a sub-solve never polishes on exit), so the host residual is
    // computed for the outermost solve only.
    if (outermost_) res_.residual = evaluate_residual(orig_, res_.x, res_.y);
    const bool worth = opt_.polish_feasibility && opt_.polish_on_exit &&
                       opt_.termination == ANUK_TERM_FULL && status_ != ANUK_PDHG_OPTIMAL &&
                       status_ != ANUK_PDHG_PRIMAL_INFEASIBLE && status_ != ANUK_PDHG_NUMERICAL_ERROR &&
                       std::isfinite(res_.residual.primal_residual) &&
                       std::isfinite(res_.residual.dual_residual) &&
                       seconds_since(start_) < opt_.time_limit_seconds;
    const Int budget = worth ? polish_budget(iteration_) : 0;
    if (budget > 0) {
      std::vector<double> tx, ty;
      Residual tr;
      Int spent = 0;
      const std::vector<double> sx = host_x_, sy = host_y_;
      run_polish(sx, sy, budget, &tx, &ty, &tr, &spent);
      ++res_.polish_attempts;
      res_.polish_iterations += spent;
      const double ceiling = std::fmax(gap_tol_, res_.residual.relative_gap);
      const bool better = tr.relative_primal <= res_.residual.relative_primal &&
                          tr.relative_dual <= res_.residual.relative_dual &&
                          (tr.relative_primal < res_.residual.relative_primal ||
                           tr.relative_dual < res_.residual.relative_dual);
      if (better && tr.relative_gap <= ceiling) {
        res_.x = std::move(tx);
        res_.y = std::move(ty);
        res_.residual = tr;
        res_.polished = true;
        if (tr.converged(tol_, opt_.absolute_tolerance, opt_.require_inf_norm_termination != 0,
                         gap_tol_)) {
          status_ = ANUK_PDHG_OPTIMAL;
          res_.message = "converged after feasibility polishing";
        } else {
          res_.message = res_.message.empty() ? "feasibility polished"
                                              : res_.message + ", feasibility polished";
        }
      }
    }
  } else {
    res_.x = std::move(polished_x_);
    res_.y = std::move(polished_y_);
    res_.residual = polished_r_;
  }
  res_.objective = orig_.model_objective(res_.x);
  res_.iterations = iteration_;
  res_.status = status_;
  res_.final_step_size = eta_;
  res_.final_primal_weight = omega_;
  res_.solve_seconds = seconds_since(start_);
}

}  // namespace

namespace {
// Small LPs are latency-bound: the plain fused kernels spread one of a few
// thousand rows over a handful of workgroups, while the batched solver's
// launches give each row a wave (several short rows to a wave, a row of up
// to 32 entries summed in one round). As a batch of one it takes the same
// decisions; measured 1.37x faster at 3,586 nonzeros, 1.34x at 10,786, 1.15x
// at 21,586, and 0.81x at 46,786 (refinery plans of 4, 12, 24 and 52
// periods), and slower on LPs with long columns.
constexpr Int kSmallLpNnz = 30000;
constexpr Int kSmallLpRow = 32;

bool small_lp(const StandardLp& lp, const anuk_pdhg_options& o) {
  if (o.small_lp_engine != 0 || !batch_supports(o, nullptr)) return false;
  if (lp.k.nnz() == 0 || lp.k.nnz() > kSmallLpNnz) return false;
  for (const Csr* a : {&lp.k, &lp.kt})
    for (Int r = 0; r < a->rows; ++r)
      if (a->start[uz(r) + 1] - a->start[uz(r)] > kSmallLpRow) return false;
  return true;
}
}  // namespace

SolveResult solve_pdhg(HipBackend& be, const StandardLp& lp, const anuk_pdhg_options& options) {
  if (options.termination_check_frequency <= 0)
    throw std::invalid_argument("termination_check_frequency must be positive");
  if (small_lp(lp, options)) return std::move(solve_pdhg_batch(be, {&lp}, options)[0]);
  return Solver(be, lp, options).run();
}

double estimate_matrix_norm(HipBackend& be, const StandardLp& lp, int iterations, double tolerance) {
  const Int n = lp.cols(), m = lp.rows();
  if (n == 0 || m == 0 || lp.k.nnz() == 0) return 0.0;
  // Deterministic non-uniform start, normalised on the host (pdhg.cpp:231-235).
  std::vector<double> v(uz(n));
  for (Int j = 0; j < n; ++j) v[uz(j)] = 1.0 + 0.1 * static_cast<double>(j % 7);
  const double norm = host_two_norm(v);
  if (norm == 0.0) return 0.0;
  for (double& e : v) e /= norm;
  const DevCsr& k = be.prepare(view_of(lp.k));
  const DevCsr& kt = be.prepare(view_of(lp.kt));
  // The iterations run in rounds of kRound without waiting on the host: each
  // leaves its ||K^T K v||^2 in sq[it] (the reduction two_norm uses) and
  // divides by its square root on the device. The host then replays the
  // test of each iteration of the round in order, so the estimate and the
  // iteration it stops at are those of the loop that waits every time; the
  // iterations a round runs past the stop change nothing that is returned.
  constexpr int kRound = 16;
  double* d_v = be.allocate(n);
  double* d_kv = be.allocate(m);
  double* d_ktkv = be.allocate(n);
  double* d_sq = be.allocate(std::max(iterations, 1));
  double* d_part = be.allocate(1024 * kMaxRed);
  struct Free {
    HipBackend& b;
    double* p[5];
    ~Free() {
      // A teardown error is dropped: a destructor must not throw.
      for (double* q : p) {
        try {
          b.deallocate(q);
        } catch (...) {
        }
      }
    }
  } guard{be, {d_v, d_kv, d_ktkv, d_sq, d_part}};
  be.upload(v.data(), d_v, n);
  hipStream_t s = be.stream();
  std::vector<double> sq(uz(std::max(iterations, 1)));
  double sigma = 0.0;
  for (int it0 = 0; it0 < iterations; it0 += kRound) {
    const int it1 = std::min(iterations, it0 + kRound);
    for (int it = it0; it < it1; ++it) {
      launch_spmv(s, k, d_v, d_kv);
      launch_spmv(s, kt, d_kv, d_ktkv);
      launch_dot(s, d_ktkv, d_ktkv, n, d_part, d_sq + it);
      launch_scale_by_root(s, n, d_sq + it, d_ktkv, d_v);
    }
    be.download(d_sq + it0, sq.data() + it0, it1 - it0);
    for (int it = it0; it < it1; ++it) {
      const double next = std::sqrt(sq[uz(it)]);
      if (next == 0.0) return 0.0;
      const double estimate = std::sqrt(next);
      if (it > 0 && std::fabs(estimate - sigma) <= tolerance * estimate) return estimate;
      sigma = estimate;
    }
  }
  return sigma;
}

void default_options(anuk_pdhg_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->tolerance = 1e-6;
  o->gap_tolerance = 0.0;
  o->max_iterations = 200000;
  o->time_limit_seconds = 300.0;
  o->termination_check_frequency = 40;
  o->absolute_tolerance = 0.0;
  o->require_inf_norm_termination = 1;
  o->adaptive_step_size = 1;
  o->primal_weight_updates = 1;
  o->restarts = 1;
  o->halpern = 1;
  o->halpern_restart = 0;
  o->halpern_minimum_epoch = 16;
  o->polish_feasibility = 1;
  o->polish_first_iteration = 100;
  o->polish_iteration_factor = 0.125;
  o->polish_minimum_iterations = 40;
  o->polish_maximum_iterations = 20000;
  o->polish_on_exit = 1;
  o->reflection = 1.0;
  o->constant_step_size = 1;
  o->constant_step_scale = 0.90;
  o->restart_on_fixed_point = 1;
  o->pid_primal_weight = 1;
  o->primal_weight_kp = 0.5;
  o->primal_weight_ki = 0.0;
  o->primal_weight_kd = 0.3;
  o->ruiz_iterations = 10;
  o->pock_chambolle = 1;
  o->reuse_products = 1;
  o->detect_infeasibility = 1;
  o->infeasibility_tolerance = 1e-8;
  o->termination = ANUK_TERM_FULL;
  o->known_matrix_norm = 0.0;
  o->initial_step_size = 0.0;
  o->initial_primal_weight = 0.0;
  o->verbose = 0;
  o->log_frequency = 5000;
  o->stop = nullptr;
  o->use_graphs = 1;
  o->product_order = 0;
  o->autotune = 1;
  o->batch_waves = 0;
  o->batch_rows = 0;
  o->batch_layout = 0;
  o->batch_pack = 0;
  o->multi_order = 0;
  o->multi_stream_loads = 0;
}

}  // namespace anuk
