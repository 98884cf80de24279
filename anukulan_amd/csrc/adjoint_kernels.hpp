// Kernels of the adjoint solve: up to 64 LPs sharing one constraint matrix,
// one wave64 lane per LP, vectors row-interleaved [row][64] as in the batched
// solver.
//
// On the final active set of a PDHG solution (columns strictly inside their
// bounds: F; binding rows: A) the fixed point of the transposed PDHG map is
// the solution of two linear systems in B = K[A, F]:
//
//     B' mu = g_F        (the x-seed: dL/dq = mu, dL/dbound = g - K' mu)
//     B  nu = h_A        (the y-seed: dL/dc = nu)
//
// Two solvers reach it. The conjugate-gradient one (CGLS on each system) is
// the default; the literal one iterates the transposed map itself,
//
//     v = K'(R lam_y),  w = lam_x - 2 sigma v,
//     lam_x <- P w + sigma v + g,   lam_y <- R lam_y + tau K (P w) + h,
//
// with P, R the 0/1 masks of F and A, and converges at the linearised PDHG
// rate. Both read K and K' once per product for all lanes.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace anuk {

// Per-lane state of one CGLS run, on the device.
struct CglsLane {
  double alpha;   // step of the current iteration
  double beta;    // direction update of the current iteration
  double gamma;   // ||A' r||^2 of the current residual
  double gamma0;  // the same at the start (for the relative stop test)
  double tol2;    // stop when gamma <= tol2 * gamma0
  Int done;       // 1 once converged (or nothing to solve): the lane stops moving
  Int iterations; // iterations taken before done
};

// Per-lane step sizes of the literal fixed-point iteration.
struct FixedPointLane {
  double tau;
  double sigma;
};

// Partials needed by the masked products' norm reductions (doubles).
size_t adjoint_partial_doubles(Int max_rows);

// out = out_mask .* (A v) per lane (out_mask may be null: no mask), and, when
// `norm_out` is set, per-lane ||out||^2 folded in a fixed tree into
// norm_out[64] through `partials`.
void launch_lane_masked_product(hipStream_t s, const DevCsr& a, const double* v,
                                const double* out_mask, double* out, double* partials,
                                double* norm_out);

// CGLS scalar updates after a product. phase 0: delta = ||A p||^2 is in
// `norm`, alpha = gamma / delta. phase 1: gamma_new = ||A' r||^2 is in `norm`,
// beta = gamma_new / gamma; with `init`, gamma0 = gamma = gamma_new.
void launch_cgls_scalars(hipStream_t s, CglsLane* lanes, const double* norm, int phase, bool init);
// z += alpha p (rows_z rows); r -= alpha q (rows_r rows).
void launch_cgls_step(hipStream_t s, const CglsLane* lanes, Int rows_z, double* z, const double* p,
                      Int rows_r, double* r, const double* q);
// p = s + beta p.
void launch_cgls_direction(hipStream_t s, const CglsLane* lanes, Int rows, double* p,
                           const double* sv);

// One step of the transposed PDHG map (see the header comment). `rho` holds
// R .* lam_y and `pi` P .* w; both are rewritten by the step.
void launch_fixed_point_step(hipStream_t s, const DevCsr& k, const DevCsr& kt,
                             const FixedPointLane* lanes, const double* P, const double* R,
                             const double* g, const double* h, double* lam_x, double* lam_y,
                             double* w, double* pi, double* rho);

}  // namespace anuk
