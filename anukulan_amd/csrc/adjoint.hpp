// Sensitivities of a PDHG solution: dL/d(c, q, lower, upper) for a loss
// L = w_x' x* + w_y' y* of the primal-dual pair a solve returned, by implicit
// differentiation of the PDHG fixed point on its final active set (no tape of
// the iterations). No counterpart in the reference.
//
// The active set is read off the pair: a column is at a bound when it lies
// within active_tolerance * (1 + |bound|) of it, and free otherwise (F); an
// inequality row is binding when its slack is within active_tolerance *
// (1 + |q_i|), and every equality row is binding (A). The PDHG map is
// differentiated in the scaled space the solve iterates in (the same Ruiz and
// Pock-Chambolle scaling), and the gradients are mapped back:
//
//   dL/dq = D1 mu,       B~' mu = D2 w_x  on F        (mu = 0 off A)
//   dL/dc = D2 nu,       B~  nu = D1 w_y  on A        (nu = 0 off F)
//   dL/dbound_j = (D2 w_x - K~' mu)_j / D2_j   for the columns at that bound
//
// with B~ = K~[A, F]. At a nondegenerate vertex B~ is square and these are
// the basis sensitivities B^-1. At a degenerate vertex (columns at a bound
// that a basis keeps, rows whose slack is zero but basic) B~ read off the
// point is not square, and its least-squares, minimum-norm solution pins
// those columns to their bounds, which is not a derivative of the solution.
// There the active set is completed to an optimal basis (active_set AUTO or
// BASIS): the exact finish (finish.hpp) crosses over from the pair and runs
// the simplex, and F becomes the basic columns, A the rows whose logical is
// nonbasic, and the bound sets the nonbasic columns at each bound. B~ is then
// square and nonsingular, and the same solve returns that basis's
// sensitivities.
#pragma once

#include <vector>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"

namespace anuk {

struct AdjointOutcome {
  std::vector<double> grad_c, grad_q, grad_lower, grad_upper;
  int iterations = 0;
  bool converged = false;
  double residual_primal = 0.0;  // relative normal-equation residual of B' mu = g
  double residual_dual = 0.0;    // relative normal-equation residual of B nu = h
  Int free_columns = 0, binding_rows = 0, lower_columns = 0, upper_columns = 0;
  bool basis_completed = false;  // the active set came from an optimal basis
};

// LPs must share K (as for the batched solve). x[i], y[i] are LP i's
// unscaled pair; w_x[i] / w_y[i] may be null (a zero seed).
std::vector<AdjointOutcome> pdhg_adjoint(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                         const anuk_pdhg_options& scaling,
                                         const anuk_adjoint_options& options,
                                         const std::vector<const double*>& x,
                                         const std::vector<const double*>& y,
                                         const std::vector<const double*>& w_x,
                                         const std::vector<const double*>& w_y);

void default_adjoint_options(anuk_adjoint_options* o);

}  // namespace anuk
