// gfx950 kernels of the adjoint solve (see adjoint_kernels.hpp).
#include "adjoint_kernels.hpp"

#include <stdexcept>
#include <string>

#include "lanes.hpp"

namespace anuk {

namespace {

using lanes::at;
using lanes::kLaneCount;
using lanes::kWavesPerBlock;
using lanes::row_blocks;
using lanes::row_sum;
using lanes::wave_row;

void check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One wave per row; the four waves of a block fold their squares in wave
// order into partials[block][lane].
__global__ void __launch_bounds__(256)
    masked_product_kernel(DevCsr a, const double* __restrict__ v, const double* __restrict__ mask,
                          double* __restrict__ out, double* __restrict__ partials) {
  __shared__ double lds[kWavesPerBlock][kLaneCount];
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const Int r = wave_row();
  double sq = 0.0;
  if (r < a.rows) {
    double s = row_sum(a, r, v, lane);
    if (mask != nullptr) s *= mask[at(r, lane)];
    out[at(r, lane)] = s;
    sq = s * s;
  }
  if (partials == nullptr) return;
  lds[wave][lane] = sq;
  __syncthreads();
  if (wave == 0) {
    double acc = lds[0][lane];
    for (int w = 1; w < kWavesPerBlock; ++w) acc += lds[w][lane];
    // partials are indexed by the logical block, which covers rows in order
    const int blk = lanes::xcd_remap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    partials[static_cast<size_t>(blk) * kLaneCount + lane] = acc;
  }
}

// 16 threads per lane each fold a contiguous run of blocks in order, then the
// 16 runs fold in order: a fixed tree that depends on the block count only.
__global__ void __launch_bounds__(1024)
    lane_fold_kernel(const double* __restrict__ partials, Int blocks, double* __restrict__ out) {
  __shared__ double lds[16][kLaneCount];
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int part = static_cast<int>(threadIdx.x >> 6);
  const Int per = (blocks + 15) / 16;
  const Int b0 = part * per, b1 = min(b0 + per, blocks);
  double acc = 0.0;
  for (Int b = b0; b < b1; ++b) acc += partials[static_cast<size_t>(b) * kLaneCount + lane];
  lds[part][lane] = acc;
  __syncthreads();
  if (part == 0) {
    double t = lds[0][lane];
    for (int k = 1; k < 16; ++k) t += lds[k][lane];
    out[lane] = t;
  }
}

__global__ void cgls_scalars_kernel(CglsLane* lanes, const double* __restrict__ norm, int phase,
                                    int init) {
  const int lane = static_cast<int>(threadIdx.x);
  if (lane >= kLaneCount) return;
  CglsLane L = lanes[lane];
  const double v = norm[lane];
  if (phase == 0) {
    L.alpha = (L.done || !(v > 0.0)) ? 0.0 : L.gamma / v;
  } else if (init) {
    L.gamma0 = v;
    L.gamma = v;
    L.beta = 0.0;
    L.alpha = 0.0;
    L.iterations = 0;
    L.done = !(v > 0.0) ? 1 : 0;
  } else if (L.done) {
    L.alpha = 0.0;
    L.beta = 0.0;
  } else {
    L.beta = L.gamma > 0.0 ? v / L.gamma : 0.0;
    L.gamma = v;
    L.iterations += 1;
    if (!(v > L.tol2 * L.gamma0)) {
      L.done = 1;
      L.beta = 0.0;
    }
  }
  lanes[lane] = L;
}

__global__ void __launch_bounds__(256)
    cgls_step_kernel(const CglsLane* __restrict__ lanes, Int rows_z, double* __restrict__ z,
                     const double* __restrict__ p, Int rows_r, double* __restrict__ r,
                     const double* __restrict__ q) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int row = static_cast<Int>(blockIdx.x) * kWavesPerBlock + static_cast<Int>(threadIdx.x >> 6);
  const double a = lanes[lane].alpha;
  if (row < rows_z) z[at(row, lane)] += a * p[at(row, lane)];
  if (row < rows_r) r[at(row, lane)] -= a * q[at(row, lane)];
}

__global__ void __launch_bounds__(256)
    cgls_direction_kernel(const CglsLane* __restrict__ lanes, Int rows, double* __restrict__ p,
                          const double* __restrict__ sv) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int row = static_cast<Int>(blockIdx.x) * kWavesPerBlock + static_cast<Int>(threadIdx.x >> 6);
  if (row >= rows) return;
  const size_t o = at(row, lane);
  p[o] = sv[o] + lanes[lane].beta * p[o];
}

// Primal half of the transposed map, one wave per column j (row of K'):
// v = K'(R lam_y), w = lam_x - 2 sigma v, lam_x <- P w + sigma v + g, pi = P w.
__global__ void __launch_bounds__(256)
    fixed_point_primal_kernel(DevCsr kt, const FixedPointLane* __restrict__ fl,
                              const double* __restrict__ P, const double* __restrict__ g,
                              const double* __restrict__ rho, double* __restrict__ lam_x,
                              double* __restrict__ w, double* __restrict__ pi) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int j = wave_row();
  if (j >= kt.rows) return;
  const size_t o = at(j, lane);
  const double sigma = fl[lane].sigma;
  const double p = P[o], lx = lam_x[o], gj = g[o];
  const double v = row_sum(kt, j, rho, lane);
  const double wj = lx - 2.0 * sigma * v;
  w[o] = wj;
  pi[o] = p * wj;
  lam_x[o] = p * wj + sigma * v + gj;
}

// Dual half, one wave per row i of K: lam_y <- R lam_y + tau K(P w) + h,
// rho = R lam_y.
__global__ void __launch_bounds__(256)
    fixed_point_dual_kernel(DevCsr k, const FixedPointLane* __restrict__ fl,
                            const double* __restrict__ R, const double* __restrict__ h,
                            const double* __restrict__ pi, double* __restrict__ lam_y,
                            double* __restrict__ rho) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int i = wave_row();
  if (i >= k.rows) return;
  const size_t o = at(i, lane);
  const double tau = fl[lane].tau;
  const double rr = R[o], ly = lam_y[o], hi = h[o];
  const double t = row_sum(k, i, pi, lane);
  const double next = rr * ly + tau * t + hi;
  lam_y[o] = next;
  rho[o] = rr * next;
}

}  // namespace

size_t adjoint_partial_doubles(Int max_rows) {
  return static_cast<size_t>(row_blocks(max_rows)) * kLaneCount + kLaneCount;
}

void launch_lane_masked_product(hipStream_t s, const DevCsr& a, const double* v,
                                const double* out_mask, double* out, double* partials,
                                double* norm_out) {
  if (a.rows <= 0) return;
  const unsigned blocks = row_blocks(a.rows);
  hipLaunchKernelGGL(masked_product_kernel, dim3(blocks), dim3(256), 0, s, a, v, out_mask, out,
                     norm_out != nullptr ? partials : nullptr);
  check_launch("adjoint product");
  if (norm_out != nullptr) {
    hipLaunchKernelGGL(lane_fold_kernel, dim3(1), dim3(1024), 0, s, partials,
                       static_cast<Int>(blocks), norm_out);
    check_launch("adjoint fold");
  }
}

void launch_cgls_scalars(hipStream_t s, CglsLane* lanes, const double* norm, int phase,
                         bool init) {
  hipLaunchKernelGGL(cgls_scalars_kernel, dim3(1), dim3(kLaneCount), 0, s, lanes, norm, phase,
                     init ? 1 : 0);
  check_launch("cgls scalars");
}

void launch_cgls_step(hipStream_t s, const CglsLane* lanes, Int rows_z, double* z, const double* p,
                      Int rows_r, double* r, const double* q) {
  const Int rows = rows_z > rows_r ? rows_z : rows_r;
  if (rows <= 0) return;
  hipLaunchKernelGGL(cgls_step_kernel, dim3(row_blocks(rows)), dim3(256), 0, s, lanes, rows_z, z, p,
                     rows_r, r, q);
  check_launch("cgls step");
}

void launch_cgls_direction(hipStream_t s, const CglsLane* lanes, Int rows, double* p,
                           const double* sv) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(cgls_direction_kernel, dim3(row_blocks(rows)), dim3(256), 0, s, lanes, rows, p,
                     sv);
  check_launch("cgls direction");
}

void launch_fixed_point_step(hipStream_t s, const DevCsr& k, const DevCsr& kt,
                             const FixedPointLane* lanes, const double* P, const double* R,
                             const double* g, const double* h, double* lam_x, double* lam_y,
                             double* w, double* pi, double* rho) {
  if (kt.rows > 0) {
    hipLaunchKernelGGL(fixed_point_primal_kernel, dim3(row_blocks(kt.rows)), dim3(256), 0, s, kt,
                       lanes, P, g, rho, lam_x, w, pi);
    check_launch("fixed-point primal");
  }
  if (k.rows > 0) {
    hipLaunchKernelGGL(fixed_point_dual_kernel, dim3(row_blocks(k.rows)), dim3(256), 0, s, k, lanes,
                       R, h, pi, lam_y, rho);
    check_launch("fixed-point dual");
  }
}

}  // namespace anuk
