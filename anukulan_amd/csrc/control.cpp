#include "control.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace anuk {

void Control::init(const anuk_pdhg_options& o, double matrix_norm, double c_norm, double q_norm) {
  opt = &o;
  // Constant step at scale / ||K|| (pdhg.cpp:462-466); the adaptive rule is
  // not run by the fused loops.
  eta = 0.9 / matrix_norm;
  if (o.constant_step_size) eta = o.constant_step_scale / matrix_norm;
  omega = 1.0;
  if (o.primal_weight_updates && c_norm > 1e-12 && q_norm > 1e-12) omega = c_norm / q_norm;
  if (o.initial_step_size > 0.0) eta = o.initial_step_size;
  if (o.initial_primal_weight > 0.0) omega = o.initial_primal_weight;
  tol = o.tolerance;
  gap_tol = o.gap_tolerance > 0.0 ? o.gap_tolerance : o.tolerance;
  halpern = o.halpern && o.restarts;
  reflection = halpern ? o.reflection : 0.0;
  next_polish = o.polish_first_iteration;
}

IterParams Control::period_params(Int num_equalities) const {
  IterParams p;
  p.tau = eta / omega;
  p.sigma = eta * omega;
  p.gamma = reflection;
  p.halpern = halpern ? 1 : 0;
  p.reset_at_start = anchor_needs_reset ? 1 : 0;
  p.k_at_start = halpern_k;
  p.num_equalities = num_equalities;
  return p;
}

void Control::advance(int steps) {
  for (int s = 0; s < steps; ++s) {
    ++since_restart;
    if (halpern) {
      if (anchor_needs_reset) {
        anchor_needs_reset = false;
        halpern_k = 0;
      } else {
        ++halpern_k;
      }
    }
  }
}

bool Control::passes(const Residual& r) const {
  const bool inf = opt->require_inf_norm_termination != 0;
  switch (opt->termination) {
    case ANUK_TERM_PRIMAL_FEASIBILITY:
      return r.primal_ok(tol, opt->absolute_tolerance, inf);
    case ANUK_TERM_DUAL_FEASIBILITY:
      return r.dual_ok(tol, opt->absolute_tolerance, inf);
    default:
      return r.converged(tol, opt->absolute_tolerance, inf, gap_tol);
  }
}

Int Control::polish_budget(Int at) const {
  Int b = static_cast<Int>(opt->polish_iteration_factor * static_cast<double>(at + 1));
  b = std::max(b, static_cast<Int>(opt->polish_minimum_iterations));
  b = std::min(b, static_cast<Int>(opt->polish_maximum_iterations));
  return b;
}

Control::Verdict Control::check_first(Int i, const Residual& r, bool farkas, Int* budget) {
  if (farkas) {
    status = ANUK_PDHG_PRIMAL_INFEASIBLE;
    message = "Farkas certificate found in the dual iterate difference";
    return Verdict::kStop;
  }
  if (!std::isfinite(r.primal_residual) || !std::isfinite(r.dual_residual)) {
    status = ANUK_PDHG_NUMERICAL_ERROR;
    message = "iterates stopped being finite";
    return Verdict::kStop;
  }
  if (passes(r)) {
    status = ANUK_PDHG_OPTIMAL;
    return Verdict::kStop;
  }
  // Polish checkpoints sit at fixed, doubling iterations and advance whether
  // or not this one attempts (pdhg.cpp:1035-1046).
  if (opt->polish_feasibility && opt->termination == ANUK_TERM_FULL && i + 1 >= next_polish) {
    if (next_polish < 1) next_polish = 1;
    while (next_polish <= i + 1) next_polish *= 2;
    const Int b = polish_budget(i);
    if (r.gap_ok(gap_tol) && b > 0) {
      *budget = b;
      return Verdict::kPolish;
    }
  }
  return Verdict::kContinue;
}

Control::Verdict Control::check_second(Int i, const Residual& r, double candidate_kkt,
                                       double elapsed) {
  if (opt->verbose && (i + 1) % opt->log_frequency == 0) {
    std::printf("  iter %7d  primal %.3e  dual %.3e  gap %.3e  kkt %.3e  eta %.3e  w %.3e\n", i + 1,
                r.relative_primal, r.relative_dual, r.relative_gap, candidate_kkt, eta, omega);
  }
  if (elapsed > opt->time_limit_seconds) {
    status = ANUK_PDHG_TIME_LIMIT;
    return Verdict::kStop;
  }
  if (opt->stop != nullptr && *opt->stop != 0) {
    status = ANUK_PDHG_TIME_LIMIT;
    return Verdict::kStop;
  }
  if (!opt->restarts) return Verdict::kContinue;
  // cuPDLP's three restart conditions on the fixed-point residual
  // (pdhg.cpp:1106-1122).
  const double residual = std::sqrt(std::fmax(movement, 0.0));
  const bool by_residual = opt->restart_on_fixed_point != 0;
  if (by_residual && !std::isfinite(residual_at_restart)) residual_at_restart = residual;
  const double measure = by_residual ? residual : candidate_kkt;
  const double at_restart = by_residual ? residual_at_restart : kkt_at_restart;
  const double previous = by_residual ? last_residual : last_candidate_kkt;
  const bool sufficient = measure <= 0.2 * at_restart;
  const bool necessary = measure <= 0.8 * at_restart && measure > previous;
  const bool too_long = static_cast<double>(since_restart) >= 0.36 * static_cast<double>(i + 1);
  last_candidate_kkt = candidate_kkt;
  last_residual = residual;
  return (sufficient || necessary || too_long) ? Verdict::kRestart : Verdict::kContinue;
}

void Control::restart(double delta_x, double delta_y) {
  if (opt->primal_weight_updates && have_prev_restart && delta_x > 1e-12 && delta_y > 1e-12) {
    double cand = 0.0;
    if (opt->pid_primal_weight) {
      // Imbalance in log space, positive when the primal moved further.
      const double e = std::log(omega * delta_x / delta_y);
      werr_sum += e;
      const double correction = opt->primal_weight_kp * e + opt->primal_weight_ki * werr_sum +
                                opt->primal_weight_kd * (e - werr_prev);
      werr_prev = e;
      cand = std::exp(std::log(omega) - correction);
    } else {
      cand = std::exp(0.5 * std::log(delta_y / delta_x) + 0.5 * std::log(omega));
    }
    if (std::isfinite(cand) && cand > 1e-12 && cand < 1e12) omega = cand;
  }
  have_prev_restart = true;
  anchor_needs_reset = true;
  since_restart = 0;
  last_candidate_kkt = kInf;
  last_residual = kInf;
  residual_at_restart = std::sqrt(std::fmax(movement, 0.0));
  ++restarts;
}

}  // namespace anuk
