// The adjoint solve on the device: masks and seeds are built on the host from
// the final pair, the two linear systems of the transposed fixed point run
// as 64-lane CGLS (or as the literal fixed-point iteration), and the
// gradients are assembled and unscaled on the host. See adjoint.hpp.
#include "adjoint.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "adjoint_kernels.hpp"
#include "batch_kernels.hpp"
#include "finish.hpp"
#include "pdhg.hpp"

namespace anuk {

namespace {

constexpr int kAdjointSlot = 300;  // backend scratch slots 300..
constexpr int kCheckEvery = 32;    // CGLS iterations between host looks

// Interleaved [rows][64] host buffer.
struct Lanes64 {
  Int rows = 0;
  std::vector<double> v;
  explicit Lanes64(Int r) : rows(r), v(static_cast<size_t>(r) * kLanes, 0.0) {}
  double& operator()(Int row, int lane) { return v[static_cast<size_t>(row) * kLanes + lane]; }
  double operator()(Int row, int lane) const { return v[static_cast<size_t>(row) * kLanes + lane]; }
};

struct LaneSets {
  std::vector<char> free_col, at_lower, at_upper, binding;
};

// The active set of one LP from its unscaled pair.
LaneSets classify(const StandardLp& lp, const double* x, double tol) {
  const Int n = lp.cols(), m = lp.rows();
  LaneSets s;
  s.free_col.assign(uz(n), 0);
  s.at_lower.assign(uz(n), 0);
  s.at_upper.assign(uz(n), 0);
  s.binding.assign(uz(m), 0);
  for (Int j = 0; j < n; ++j) {
    const double lo = lp.lower[uz(j)], hi = lp.upper[uz(j)], v = x[j];
    if (std::isfinite(lo) && v - lo <= tol * (1.0 + std::fabs(lo))) {
      s.at_lower[uz(j)] = 1;
    } else if (std::isfinite(hi) && hi - v <= tol * (1.0 + std::fabs(hi))) {
      s.at_upper[uz(j)] = 1;
    } else {
      s.free_col[uz(j)] = 1;
    }
  }
  std::vector<double> kx(uz(m));
  lp.k.multiply(x, kx.data());
  for (Int i = 0; i < m; ++i) {
    const double q = lp.q[uz(i)];
    s.binding[uz(i)] =
        (i < lp.num_equalities || kx[uz(i)] - q <= tol * (1.0 + std::fabs(q))) ? 1 : 0;
  }
  return s;
}

// The active set of an optimal basis (the exact finish's): the basic
// columns are free, the nonbasic columns sit at the bound their state names,
// and a row binds unless its logical is basic.
LaneSets classify_basis(const StandardLp& lp, const FinishResult& r) {
  const Int n = lp.cols(), m = lp.rows();
  LaneSets s;
  s.free_col.assign(uz(n), 0);
  s.at_lower.assign(uz(n), 0);
  s.at_upper.assign(uz(n), 0);
  s.binding.assign(uz(m), 0);
  for (Int j = 0; j < n; ++j) {
    const VarState v = r.state[uz(j)];
    s.free_col[uz(j)] = v == VarState::kBasic;
    s.at_lower[uz(j)] = v == VarState::kLower;
    s.at_upper[uz(j)] = v == VarState::kUpper;
  }
  for (Int i = 0; i < m; ++i) s.binding[uz(i)] = r.state[uz(n + i)] != VarState::kBasic;
  return s;
}

Int count_set(const std::vector<char>& v) {
  Int c = 0;
  for (char e : v) c += e != 0;
  return c;
}

// One lane of an interleaved vector, times an optional 0/1 mask lane.
std::vector<double> lane_of(const Lanes64& v, int lane, const Lanes64* mask = nullptr) {
  std::vector<double> out(uz(v.rows));
  for (Int r = 0; r < v.rows; ++r) out[uz(r)] = v(r, lane) * (mask ? (*mask)(r, lane) : 1.0);
  return out;
}

std::vector<double> product(const Csr& a, const std::vector<double>& v) {
  std::vector<double> out(uz(a.rows));
  a.multiply(v.data(), out.data());
  return out;
}

class AdjointEngine {
 public:
  AdjointEngine(HipBackend& be, const DevCsr& k, const DevCsr& kt, Int n, Int m)
      : be_(be), k_(k), kt_(kt), n_(n), m_(m) {
    const Int big = std::max(n, m);
    const size_t vec = static_cast<size_t>(big) * kLanes;
    int slot = kAdjointSlot;
    for (double** p : {&P_, &R_, &b_, &z_, &r_, &p_, &q_, &s_, &t_, &w_, &pi_, &rho_, &lx_, &ly_})
      *p = be_.scratch(slot++, vec);
    partials_ = be_.scratch(slot++, adjoint_partial_doubles(big));
    norm_ = be_.scratch(slot++, kLanes);
    state_ = reinterpret_cast<CglsLane*>(be_.scratch(slot++, kLanes * sizeof(CglsLane) / 8 + 1));
    fixed_ = reinterpret_cast<FixedPointLane*>(
        be_.scratch(slot++, kLanes * sizeof(FixedPointLane) / 8 + 1));
  }

  void set_masks(const Lanes64& P, const Lanes64& R) {
    be_.upload_staged(P.v.data(), P_, P.v.size());
    be_.upload_staged(R.v.data(), R_, R.v.size());
  }

  // CGLS for min || Aop z - rhs ||, all lanes at once. primal: Aop = B' (z
  // over rows, rhs over columns); otherwise Aop = B. Returns the iterations
  // of the slowest lane; `converged` says whether every lane stopped.
  int cgls(bool primal, const Lanes64& rhs, double tol, int max_iterations, bool* converged,
           Lanes64* solution) {
    hipStream_t s = be_.stream();
    const Int zr = primal ? m_ : n_, rr = primal ? n_ : m_;
    const DevCsr& op = primal ? kt_ : k_;       // Aop applies this CSR...
    const double* op_mask = primal ? P_ : R_;   // ...and masks its output with this
    const DevCsr& adj = primal ? k_ : kt_;
    const double* adj_mask = primal ? R_ : P_;
    std::vector<CglsLane> init(kLanes);
    for (auto& L : init) L = CglsLane{0.0, 0.0, 0.0, 0.0, tol * tol, 0, 0};
    hip_check(hipMemcpyAsync(state_, init.data(), kLanes * sizeof(CglsLane), hipMemcpyHostToDevice, s),
              "adjoint state");
    const size_t zbytes = static_cast<size_t>(zr) * kLanes * sizeof(double);
    hip_check(hipMemsetAsync(z_, 0, zbytes, s), "adjoint z");
    hip_check(hipMemsetAsync(p_, 0, zbytes, s), "adjoint p");
    be_.upload_staged(rhs.v.data(), r_, rhs.v.size());
    launch_lane_masked_product(s, adj, r_, adj_mask, s_, partials_, norm_);
    launch_cgls_scalars(s, state_, norm_, 1, true);
    launch_cgls_direction(s, state_, zr, p_, s_);
    std::vector<CglsLane> now(kLanes);
    int it = 0;
    *converged = false;
    while (it < max_iterations) {
      const int chunk = std::min(kCheckEvery, max_iterations - it);
      for (int c = 0; c < chunk; ++c) {
        launch_lane_masked_product(s, op, p_, op_mask, q_, partials_, norm_);
        launch_cgls_scalars(s, state_, norm_, 0, false);
        launch_cgls_step(s, state_, zr, z_, p_, rr, r_, q_);
        launch_lane_masked_product(s, adj, r_, adj_mask, s_, partials_, norm_);
        launch_cgls_scalars(s, state_, norm_, 1, false);
        launch_cgls_direction(s, state_, zr, p_, s_);
      }
      it += chunk;
      hip_check(hipMemcpyAsync(now.data(), state_, kLanes * sizeof(CglsLane), hipMemcpyDeviceToHost, s),
                "adjoint state");
      be_.synchronize();
      if (std::all_of(now.begin(), now.end(), [](const CglsLane& L) { return L.done != 0; })) {
        *converged = true;
        break;
      }
    }
    int slowest = 0;
    for (const auto& L : now) slowest = std::max(slowest, static_cast<int>(L.iterations));
    hip_check(hipMemcpyAsync(solution->v.data(), z_, zbytes, hipMemcpyDeviceToHost, s), "adjoint z");
    be_.synchronize();
    return slowest;
  }

  // The literal iteration of the transposed map for `iterations` steps.
  // Returns lam_y (rows) and w (columns).
  void fixed_point(const Lanes64& g, const Lanes64& h, const std::vector<FixedPointLane>& steps,
                   int iterations, Lanes64* lam_y, Lanes64* w) {
    hipStream_t s = be_.stream();
    hip_check(hipMemcpyAsync(fixed_, steps.data(), kLanes * sizeof(FixedPointLane),
                             hipMemcpyHostToDevice, s),
              "adjoint steps");
    const size_t nb = static_cast<size_t>(n_) * kLanes * sizeof(double);
    const size_t mb = static_cast<size_t>(m_) * kLanes * sizeof(double);
    hip_check(hipMemsetAsync(lx_, 0, nb, s), "lam_x");
    hip_check(hipMemsetAsync(ly_, 0, mb, s), "lam_y");
    hip_check(hipMemsetAsync(rho_, 0, mb, s), "rho");
    be_.upload_staged(g.v.data(), b_, g.v.size());
    be_.upload_staged(h.v.data(), t_, h.v.size());
    for (int i = 0; i < iterations; ++i)
      launch_fixed_point_step(s, k_, kt_, fixed_, P_, R_, b_, t_, lx_, ly_, w_, pi_, rho_);
    hip_check(hipMemcpyAsync(lam_y->v.data(), ly_, mb, hipMemcpyDeviceToHost, s), "lam_y");
    hip_check(hipMemcpyAsync(w->v.data(), w_, nb, hipMemcpyDeviceToHost, s), "w");
    be_.synchronize();
  }

 private:
  HipBackend& be_;
  DevCsr k_, kt_;
  Int n_, m_;
  double *P_ = nullptr, *R_ = nullptr, *b_ = nullptr, *z_ = nullptr, *r_ = nullptr, *p_ = nullptr,
         *q_ = nullptr, *s_ = nullptr, *t_ = nullptr, *w_ = nullptr, *pi_ = nullptr, *rho_ = nullptr,
         *lx_ = nullptr, *ly_ = nullptr;
  double* partials_ = nullptr;
  double* norm_ = nullptr;
  CglsLane* state_ = nullptr;
  FixedPointLane* fixed_ = nullptr;
};

double norm2(const std::vector<double>& v) {
  double s = 0.0;
  for (double e : v) s += e * e;
  return std::sqrt(s);
}

}  // namespace

void default_adjoint_options(anuk_adjoint_options* o) {
  o->method = ANUK_ADJOINT_CG;
  o->max_iterations = 20000;
  o->tolerance = 1e-11;
  o->active_tolerance = 1e-6;
  o->primal_weight = 1.0;
  o->active_set = ANUK_ACTIVE_AUTO;
}

std::vector<AdjointOutcome> pdhg_adjoint(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                         const anuk_pdhg_options& scaling,
                                         const anuk_adjoint_options& opt,
                                         const std::vector<const double*>& xs,
                                         const std::vector<const double*>& ys,
                                         const std::vector<const double*>& wxs,
                                         const std::vector<const double*>& wys) {
  if (lps.empty()) return {};
  if (opt.method != ANUK_ADJOINT_CG && opt.method != ANUK_ADJOINT_FIXED_POINT)
    throw std::invalid_argument("adjoint: unknown method");
  if (opt.active_set != ANUK_ACTIVE_AUTO && opt.active_set != ANUK_ACTIVE_POINT &&
      opt.active_set != ANUK_ACTIVE_BASIS)
    throw std::invalid_argument("adjoint: unknown active-set rule");
  if (opt.max_iterations < 0 || !(opt.tolerance >= 0.0) || !(opt.active_tolerance >= 0.0) ||
      !(opt.primal_weight > 0.0))
    throw std::invalid_argument("adjoint: invalid options");
  const StandardLp& first = *lps[0];
  for (const StandardLp* lp : lps) {
    if (lp->rows() != first.rows() || lp->cols() != first.cols() ||
        lp->num_equalities != first.num_equalities || lp->k.start != first.k.start ||
        lp->k.index != first.k.index || lp->k.value != first.k.value)
      throw std::invalid_argument("adjoint: every LP must share the same constraint matrix");
  }
  const Int n = first.cols(), m = first.rows();

  // The scaled K of the solve and its diagonals (they depend on K alone).
  Csr k = first.k;
  std::vector<Sweep> sweeps;
  if (scaling.ruiz_iterations > 0 || scaling.pock_chambolle != 0)
    sweeps = scale_matrix(&k, scaling.ruiz_iterations, scaling.pock_chambolle != 0);
  Csr kt = k.transpose();
  Scaling d;
  {
    std::vector<double> q = first.q, c = first.c, lo = first.lower, hi = first.upper;
    apply_sweeps(sweeps, &q, &c, &lo, &hi, &d);
  }
  if (d.row_scale.empty()) d.row_scale.assign(uz(m), 1.0);
  if (d.col_scale.empty()) d.col_scale.assign(uz(n), 1.0);
  const DevCsr& dk = be.prepare(view_of(k));
  const DevCsr& dkt = be.prepare(view_of(kt));
  struct Release {
    HipBackend& b;
    std::uint64_t a, c;
    ~Release() {
      // A teardown error is dropped: the error that unwinds past here is
      // the one the caller gets, and a destructor must not throw.
      try {
        b.release_one(a);
        b.release_one(c);
      } catch (...) {
      }
    }
  } release{be, k.id, kt.id};

  double eta = 0.0;
  if (opt.method == ANUK_ADJOINT_FIXED_POINT) {
    StandardLp shape;
    shape.k = k;
    shape.kt = kt;
    const double norm = estimate_matrix_norm(be, shape, 200, 1e-8);
    eta = norm > 0.0 ? 0.9 / norm : 0.0;
  }

  // The active set of every LP: read off its pair, and completed to an
  // optimal basis by the exact finish (all such LPs at once, one per host
  // thread) where the rule asks for it or the point's system is not square.
  std::vector<LaneSets> all_sets(lps.size());
  std::vector<size_t> complete;
  for (size_t i = 0; i < lps.size(); ++i) {
    all_sets[i] = classify(*lps[i], xs[i], opt.active_tolerance);
    const bool square = count_set(all_sets[i].free_col) == count_set(all_sets[i].binding);
    if (opt.active_set == ANUK_ACTIVE_BASIS || (opt.active_set == ANUK_ACTIVE_AUTO && !square))
      complete.push_back(i);
  }
  std::vector<bool> completed(lps.size(), false);
  if (!complete.empty()) {
    std::vector<const StandardLp*> fl;
    std::vector<const double*> fx, fy;
    for (size_t i : complete) {
      fl.push_back(lps[i]);
      fx.push_back(xs[i]);
      fy.push_back(ys[i]);
    }
    const std::vector<FinishResult> fin = exact_finish_many(fl, fx, fy, FinishOptions{});
    for (size_t k = 0; k < complete.size(); ++k) {
      // A case the finish could not bring to an optimal basis keeps the
      // point's active set (and says so through basis_completed).
      if (fin[k].status != kFinishOptimal) continue;
      all_sets[complete[k]] = classify_basis(*lps[complete[k]], fin[k]);
      completed[complete[k]] = true;
    }
  }

  AdjointEngine engine(be, dk, dkt, n, m);
  std::vector<AdjointOutcome> out(lps.size());
  for (size_t g0 = 0; g0 < lps.size(); g0 += kLanes) {
    const size_t g1 = std::min(lps.size(), g0 + static_cast<size_t>(kLanes));
    Lanes64 P(n), R(m), g(n), h(m), gF(n);
    std::vector<LaneSets> sets(g1 - g0);
    bool any_x = false, any_y = false;
    for (size_t i = g0; i < g1; ++i) {
      const int lane = static_cast<int>(i - g0);
      LaneSets& ls = sets[uz(lane)];
      ls = all_sets[i];
      AdjointOutcome& o = out[i];
      o.basis_completed = completed[i];
      for (Int j = 0; j < n; ++j) {
        P(j, lane) = ls.free_col[uz(j)] ? 1.0 : 0.0;
        const double seed = wxs[i] != nullptr ? wxs[i][j] : 0.0;
        g(j, lane) = d.col_scale[uz(j)] * seed;
        gF(j, lane) = ls.free_col[uz(j)] ? g(j, lane) : 0.0;
        any_x = any_x || seed != 0.0;
        o.free_columns += ls.free_col[uz(j)];
        o.lower_columns += ls.at_lower[uz(j)];
        o.upper_columns += ls.at_upper[uz(j)];
      }
      for (Int r = 0; r < m; ++r) {
        R(r, lane) = ls.binding[uz(r)] ? 1.0 : 0.0;
        const double seed = wys[i] != nullptr ? wys[i][r] : 0.0;
        h(r, lane) = ls.binding[uz(r)] ? d.row_scale[uz(r)] * seed : 0.0;
        any_y = any_y || seed != 0.0;
        o.binding_rows += ls.binding[uz(r)];
      }
    }
    engine.set_masks(P, R);

    Lanes64 mu(m), nu(n);
    int iterations = 0;
    bool converged = true;
    if (opt.method == ANUK_ADJOINT_CG) {
      bool ok = true;
      if (any_x) {
        iterations = std::max(iterations,
                              engine.cgls(true, gF, opt.tolerance, opt.max_iterations, &ok, &mu));
        converged = converged && ok;
      }
      if (any_y) {
        iterations = std::max(iterations,
                              engine.cgls(false, h, opt.tolerance, opt.max_iterations, &ok, &nu));
        converged = converged && ok;
      }
    } else {
      std::vector<FixedPointLane> steps(kLanes, FixedPointLane{eta / opt.primal_weight,
                                                               eta * opt.primal_weight});
      Lanes64 lam_y(m), w(n);
      engine.fixed_point(g, h, steps, opt.max_iterations, &lam_y, &w);
      const double sigma = eta * opt.primal_weight, tau = eta / opt.primal_weight;
      for (size_t i = g0; i < g1; ++i) {
        const int lane = static_cast<int>(i - g0);
        for (Int r = 0; r < m; ++r) mu(r, lane) = sigma * R(r, lane) * lam_y(r, lane);
        for (Int j = 0; j < n; ++j) nu(j, lane) = -tau * P(j, lane) * w(j, lane);
      }
      iterations = opt.max_iterations;
      converged = false;  // judged by the residuals below
    }

    // Gradients, residuals, unscaling.
    for (size_t i = g0; i < g1; ++i) {
      const int lane = static_cast<int>(i - g0);
      const LaneSets& ls = sets[uz(lane)];
      AdjointOutcome& o = out[i];
      o.grad_c.assign(uz(n), 0.0);
      o.grad_q.assign(uz(m), 0.0);
      o.grad_lower.assign(uz(n), 0.0);
      o.grad_upper.assign(uz(n), 0.0);
      const std::vector<double> mu_l = lane_of(mu, lane), nu_l = lane_of(nu, lane);
      const std::vector<double> t = product(kt, mu_l);
      for (Int r = 0; r < m; ++r) o.grad_q[uz(r)] = d.row_scale[uz(r)] * mu_l[uz(r)];
      for (Int j = 0; j < n; ++j) {
        o.grad_c[uz(j)] = d.col_scale[uz(j)] * nu_l[uz(j)];
        const double b = (g(j, lane) - t[uz(j)]) / d.col_scale[uz(j)];
        if (ls.at_lower[uz(j)]) o.grad_lower[uz(j)] = b;
        if (ls.at_upper[uz(j)]) o.grad_upper[uz(j)] = b;
      }
      // Relative normal-equation residuals ||B (B' mu - g_F)|| / ||B g_F||
      // and ||B' (B nu - h_A)|| / ||B' h_A||.
      const std::vector<double> gF_l = lane_of(gF, lane), h_l = lane_of(h, lane);
      std::vector<double> e(uz(n));
      for (Int j = 0; j < n; ++j) e[uz(j)] = P(j, lane) * (t[uz(j)] - gF_l[uz(j)]);
      std::vector<double> top = product(k, e), bottom = product(k, gF_l);
      for (Int r = 0; r < m; ++r) {
        top[uz(r)] *= R(r, lane);
        bottom[uz(r)] *= R(r, lane);
      }
      o.residual_primal = norm2(bottom) > 0.0 ? norm2(top) / norm2(bottom) : 0.0;
      std::vector<double> bn = product(k, nu_l);
      for (Int r = 0; r < m; ++r) bn[uz(r)] = R(r, lane) * (bn[uz(r)] - h_l[uz(r)]);
      top = product(kt, bn);
      bottom = product(kt, h_l);
      for (Int j = 0; j < n; ++j) {
        top[uz(j)] *= P(j, lane);
        bottom[uz(j)] *= P(j, lane);
      }
      o.residual_dual = norm2(bottom) > 0.0 ? norm2(top) / norm2(bottom) : 0.0;
      o.iterations = iterations;
      o.converged = opt.method == ANUK_ADJOINT_CG
                        ? converged
                        : (o.residual_primal <= std::sqrt(opt.tolerance) &&
                           o.residual_dual <= std::sqrt(opt.tolerance));
    }
  }
  return out;
}

}  // namespace anuk
