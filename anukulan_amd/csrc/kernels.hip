// CDNA4 (gfx950) kernels for the PDHG backend. See kernels.hpp for the
// numerical contract. Compiled with -ffp-contract=off: every multiply and add
// below rounds on its own, as the reference's CPU loops do.
#include "kernels.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace anuk {

namespace {

constexpr int kWaves = kThreads / 64;

// Workgroups are dispatched round-robin over the 8 XCDs; this maps hardware
// workgroup i to a logical block so each XCD gets one contiguous range of
// row blocks, and neighbouring rows share that XCD's L2 for their gathers.
__device__ __forceinline__ int xcd_block(int i, int g) {
  const int x = i & 7, local = i >> 3, per = g >> 3, rem = g & 7;
  return x * per + (x < rem ? x : rem) + local;
}

__device__ __forceinline__ double wave_fold(double v, bool is_max) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = is_max ? fmax(v, o) : v + o;
  }
  return v;
}

// Folds R per-thread values over the workgroup in a fixed tree and has thread
// 0 store them to out[0..R). Every thread of the workgroup must call it.
template <int R>
__device__ void block_fold_store(double (&v)[R], unsigned max_mask, double* out) {
  __shared__ double lds[R][kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = wave_fold(v[r], (max_mask >> r) & 1u);
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) lds[r][wave] = v[r];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double a = lds[r][0];
      const bool mx = (max_mask >> r) & 1u;
      for (int w = 1; w < kWaves; ++w) a = mx ? fmax(a, lds[r][w]) : a + lds[r][w];
      out[r] = a;
    }
  }
}

__device__ double block_sum(double v) {
  __shared__ double lds[kWaves];
  v = wave_fold(v, false);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double a = lds[0];
  for (int w = 1; w < kWaves; ++w) a += lds[w];
  return a;  // every thread gets the same value
}

// ---------------------------------------------------------------------------
// Row-blocked CSR product with an epilogue.
//
// A short-row block stages value*x[index] for all its nonzeros in LDS with
// coalesced loads, then thread t sums row r0+t serially in storage order: the
// CPU's order, so the row sum has the CPU's bits. A long-row block (one row,
// more nonzeros than a tile) strides the row over the workgroup and folds in a
// fixed tree. The epilogue receives (row, sum) from exactly one thread per
// row and may add to up to kR reduction values, folded per block into
// partials[block * kR ...].
// ---------------------------------------------------------------------------
//
// The leading arguments (the CSR's arrays, the block plan, the gathered
// vector, the grid size) are plain scalars so they can be preloaded into
// SGPRs at wave launch (the file is built with kernel-argument preloading),
// and the block's row bounds and nonzero range are loaded before the
// epilogue's skip test, so the block waits on the plan and the row starts
// while its step scalars arrive.
// `mode` holds kModeWide (rows longer than kSerialRowMax summed across a
// wave in a fixed butterfly: the fast order, not the CPU's bits) and
// kModeStream (the index and value streams read non-temporally: the same
// bits).
constexpr int kModeWide = 1;
constexpr int kModeStream = 2;

template <class Epi, int kTile>
__global__ void __launch_bounds__(kThreads)
    spmv_kernel(const Int* __restrict__ a_start, const Int* __restrict__ a_index,
                const double* __restrict__ a_value, const Int* __restrict__ block_row,
                const Int* __restrict__ block_map, const double* __restrict__ x, int blocks,
                int mode, Epi epi, double* __restrict__ partials) {
  __shared__ double tile[kTile];
  constexpr int R = Epi::kR > 0 ? Epi::kR : 1;
  struct {
    const Int* __restrict__ start;
    const Int* __restrict__ index;
    const double* __restrict__ value;
  } a{a_start, a_index, a_value};
  int b = xcd_block(static_cast<int>(blockIdx.x), blocks);
  if (block_map != nullptr) b = block_map[b];
  const Int r0 = block_row[b], r1 = block_row[b + 1];
  const Int k0 = a.start[r0], k1 = a.start[r1];
  const Epi e = epi.resolve(b);
  if (e.skip()) return;  // uniform across the block
  const Int len = k1 - k0;
  double red[R];
#pragma unroll
  for (int r = 0; r < R; ++r) red[r] = 0.0;

  if (len <= kTile) {
    // Every round's row bounds are loaded up front, with the staging loads,
    // so a block waits for memory twice however many rounds it has.
    constexpr int kRounds = (kTile == kFastTileNnz ? kFastTileRows : kTileRows) / kThreads;
    Int lo_r[kRounds], hi_r[kRounds];
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const Int r = r0 + static_cast<Int>(threadIdx.x) + u * kThreads;
      lo_r[u] = 0;
      hi_r[u] = 0;
      if (r < r1) {
        lo_r[u] = a.start[r] - k0;
        hi_r[u] = a.start[r + 1] - k0;
      }
    }
    if (mode & kModeStream) {
      // The index and value streams are read once per product, so they are
      // loaded non-temporally and leave the L2 to the gathered vector's
      // lines.
#pragma unroll
      for (int u = 0; u < kTile / kThreads; ++u) {
        const Int k = static_cast<Int>(threadIdx.x) + u * kThreads;
        if (k < len)
          tile[k] = __builtin_nontemporal_load(a.value + k0 + k) *
                    x[__builtin_nontemporal_load(a.index + k0 + k)];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kTile / kThreads; ++u) {
        const Int k = static_cast<Int>(threadIdx.x) + u * kThreads;
        if (k < len) tile[k] = a.value[k0 + k] * x[a.index[k0 + k]];
      }
    }
    __syncthreads();
    // Rows in rounds of kThreads, one row per thread per round.
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const Int round = r0 + u * kThreads;
      if (round >= r1) break;  // uniform across the block
      const Int r = round + static_cast<Int>(threadIdx.x);
      double sum = 0.0;
      bool wide = false;
      if (r < r1) {
        wide = (mode & kModeWide) != 0 && hi_r[u] - lo_r[u] > kSerialRowMax;
        if (!wide)
          for (Int k = lo_r[u]; k < hi_r[u]; ++k) sum += tile[k];
      }
      if (mode & kModeWide) {
        // The wave's long rows, one at a time: lanes stride the row's staged
        // products, a butterfly adds them (a fixed tree), and the row's own
        // thread keeps the total for its epilogue.
        unsigned long long mask = __ballot(wide);
        const int lane = static_cast<int>(threadIdx.x & 63);
        while (mask != 0ull) {
          const int bit = __ffsll(static_cast<long long>(mask)) - 1;
          mask &= mask - 1ull;
          const Int lo = __shfl(lo_r[u], bit, 64), hi = __shfl(hi_r[u], bit, 64);
          double s = 0.0;
          for (Int k = lo + lane; k < hi; k += 64) s += tile[k];
          s = wave_fold(s, false);
          if (lane == bit) sum = s;
        }
      }
      if (r < r1) e(r, sum, red);
    }
  } else {
    double s = 0.0;
    for (Int k = k0 + static_cast<Int>(threadIdx.x); k < k1; k += kThreads)
      s += a.value[k] * x[a.index[k]];
    s = block_sum(s);
    if (threadIdx.x == 0) e(r0, s, red);
  }
  if constexpr (Epi::kR > 0) {
    block_fold_store<Epi::kR>(red, Epi::kMaxMask, partials + static_cast<size_t>(b) * Epi::kR);
  }
}

struct PlainEpi {
  static constexpr int kR = 0;
  [[maybe_unused]] static constexpr unsigned kMaxMask = 0;
  double* y;
  __device__ PlainEpi resolve(int) const { return *this; }
  __device__ bool skip() const { return false; }
  __device__ void operator()(Int r, double s, double*) const { y[r] = s; }
};

// Where a fused epilogue finds its step scalars: one IterParams for the
// whole matrix (one LP), or, for a block-diagonal stack of LPs whose row
// blocks never straddle two LPs, one per LP picked by the block's LP.
struct OneLp {
  const IterParams* params;
  __device__ int lp(int) const { return 0; }
  __device__ const IterParams* get(int) const { return params; }
};

struct ManyLps {
  const IterParams* params;
  const Int* block_lp;
  __device__ int lp(int b) const { return block_lp[b]; }
  __device__ const IterParams* get(int b) const { return params + block_lp[b]; }
};

// Halpern schedule of one step of a period, from the period's starting state:
// an epoch's first step takes no blend and its result becomes the anchor; the
// k-th step after it blends with weight (k+1)/(k+2).
struct StepScalars {
  double tau, sigma, gamma, w;
  bool blend, open_epoch;
  Int eq_end;
};

__device__ __forceinline__ StepScalars step_scalars(const IterParams* p, int step) {
  StepScalars s;
  s.tau = p->tau;
  s.sigma = p->sigma;
  s.gamma = p->gamma;
  s.eq_end = p->num_equalities;
  if (!p->halpern) {
    s.blend = false;
    s.open_epoch = false;
    s.w = 1.0;
  } else if (p->reset_at_start && step == 0) {
    s.blend = false;
    s.open_epoch = true;
    s.w = 1.0;
  } else {
    const Int k = p->reset_at_start ? step - 1 : p->k_at_start + step;
    s.blend = true;
    s.open_epoch = false;
    s.w = static_cast<double>(k + 1) / static_cast<double>(k + 2);
  }
  return s;
}

// Primal side of one PDHG step on column j, given (K^T y)_j
// (backend.cpp:60-79 order).
template <bool kDiag, class Src = OneLp>
struct PrimalStepEpi {
  static constexpr int kR = kDiag ? 1 : 0;
  static constexpr unsigned kMaxMask = 0;
  DevProblem p;
  PrimalBufs b;
  Src src;
  int step;
  StepScalars sc;
  bool off;
  __device__ PrimalStepEpi resolve(int blk) const {
    PrimalStepEpi e = *this;
    const IterParams* P = src.get(blk);
    e.off = P->active == 0;
    e.sc = step_scalars(P, step);
    return e;
  }
  __device__ bool skip() const { return off; }
  __device__ void operator()(Int j, double kty, double* red) const {
    const double xj = b.x[j];
    double v = xj - sc.tau * (p.c[j] - kty);
    if (v < p.lower[j]) v = p.lower[j];
    if (v > p.upper[j]) v = p.upper[j];
    const double d = v - xj;
    b.x_bar[j] = v + d;
    double next = v;
    if (sc.gamma != 0.0) next += sc.gamma * d;
    if (sc.blend) next = sc.w * next + (1.0 - sc.w) * b.anchor[j];
    b.x_out[j] = next;
    if (sc.open_epoch) b.anchor_out[j] = next;
    if constexpr (kDiag) {
      b.dx[j] = d;
      red[0] += d * d;
    }
  }
};

// Dual side on row i, given (K x_bar)_i (backend.cpp:81-111 order). On a
// diagonal step it also produces the row half of the convergence terms for
// the new point (backend.cpp:153-164, unscaled through D1) and the row half of
// the Farkas test on dy (pdhg.cpp:193-198, 215-216).
template <bool kDiag, class Src = OneLp>
struct DualStepEpi {
  static constexpr int kR = kDiag ? 7 : 0;
  static constexpr unsigned kMaxMask = (1u << 2) | (1u << 4) | (1u << 5);
  DevProblem p;
  DualBufs b;
  Src src;
  int step;
  StepScalars sc;
  bool off;
  __device__ DualStepEpi resolve(int blk) const {
    DualStepEpi e = *this;
    const IterParams* P = src.get(blk);
    e.off = P->active == 0;
    e.sc = step_scalars(P, step);
    return e;
  }
  __device__ bool skip() const { return off; }
  __device__ void operator()(Int i, double kxb, double* red) const {
    const double yi = b.y[i];
    const double kx = b.k_x[i];
    double v = yi + sc.sigma * (p.q[i] - kxb);
    if (i >= sc.eq_end && v < 0.0) v = 0.0;
    const double d = v - yi;
    const double kd = 0.5 * (kxb - kx);
    double ny = v;
    double nk = 0.5 * (kxb + kx);
    if (sc.gamma != 0.0) {
      ny += sc.gamma * d;
      nk += sc.gamma * kd;
    }
    if (sc.blend) {
      ny = sc.w * ny + (1.0 - sc.w) * b.anchor_y[i];
      nk = sc.w * nk + (1.0 - sc.w) * b.anchor_kx[i];
    }
    b.y_out[i] = ny;
    b.k_out[i] = nk;
    if (sc.open_epoch) {
      b.anchor_y_out[i] = ny;
      b.anchor_kx_out[i] = nk;
    }
    if constexpr (kDiag) {
      b.dy[i] = d;
      b.k_dx[i] = kd;
      red[0] += d * d;
      const double scale = p.row_scale[i];
      const double row = scale != 0.0 ? nk / scale : nk;
      const double slack = row - p.orig_q[i];
      const double viol = (i < sc.eq_end) ? slack : fmin(slack, 0.0);
      red[1] += viol * viol;
      red[2] = fmax(red[2], fabs(viol));
      red[3] += p.orig_q[i] * (scale * ny);
      if (i >= sc.eq_end && d < 0.0) red[4] = 1.0;
      red[5] = fmax(red[5], fabs(d));
      red[6] += p.q[i] * d;
    }
  }
};

// Column half of the convergence terms from (K~^T y~)_j (backend.cpp:166-199).
template <class Src = OneLp>
struct ConvProductEpiT {
  static constexpr int kR = 4;
  static constexpr unsigned kMaxMask = 1u << 1;
  DevProblem p;
  const double* x;
  Src src;
  bool off;
  __device__ ConvProductEpiT resolve(int blk) const {
    ConvProductEpiT e = *this;
    e.off = src.params != nullptr && src.get(blk)->active == 0;
    return e;
  }
  __device__ bool skip() const { return off; }
  __device__ void operator()(Int j, double kty, double* red) const {
    const double scale = p.col_scale[j];
    const double product = scale != 0.0 ? kty / scale : kty;
    const double lambda = p.orig_c[j] - product;
    const bool has_lo = p.orig_lower[j] > -__builtin_inf();
    const bool has_hi = p.orig_upper[j] < __builtin_inf();
    double left;
    if (has_lo && has_hi) {
      left = 0.0;
    } else if (has_lo) {
      left = fmin(lambda, 0.0);
    } else if (has_hi) {
      left = fmax(lambda, 0.0);
    } else {
      left = lambda;
    }
    red[0] += left * left;
    red[1] = fmax(red[1], fabs(left));
    const double absorbed = lambda - left;
    if (absorbed > 0.0) {
      red[2] += absorbed * p.orig_lower[j];
    } else if (absorbed < 0.0) {
      red[2] += absorbed * p.orig_upper[j];
    }
    red[3] += p.orig_c[j] * (scale * x[j]);
  }
};

using ConvProductEpi = ConvProductEpiT<OneLp>;

// Column half of the Farkas test from lambda = (K~^T dy)_j (pdhg.cpp:203-213).
// Skipped as a whole (per LP) when the row half already rules a certificate
// out.
template <class Src = OneLp>
struct FarkasProductEpiT {
  static constexpr int kR = 2;
  static constexpr unsigned kMaxMask = 1u << 0;
  DevProblem p;
  const double* check;  // row slots already folded, kCheckSlots per LP
  Src src;
  bool skip_all;
  __device__ FarkasProductEpiT resolve(int blk) const {
    FarkasProductEpiT e = *this;
    const double* c = check + static_cast<size_t>(src.lp(blk)) * kCheckSlots;
    e.skip_all = (src.params != nullptr && src.get(blk)->active == 0) || c[kFarkasRowBad] != 0.0 ||
                 !(c[kFarkasScale] > 0.0);
    return e;
  }
  __device__ bool skip() const { return skip_all; }
  __device__ void operator()(Int j, double lambda, double* red) const {
    if (lambda > 0.0) {
      if (p.upper[j] >= __builtin_inf()) {
        red[0] = 1.0;
      } else {
        red[1] += lambda * p.upper[j];
      }
    } else if (lambda < 0.0) {
      if (p.lower[j] <= -__builtin_inf()) {
        red[0] = 1.0;
      } else {
        red[1] += lambda * p.lower[j];
      }
    }
  }
};

using FarkasProductEpi = FarkasProductEpiT<OneLp>;

// ---------------------------------------------------------------------------
// Folding per-block partials into final values, one workgroup, fixed order.
// ---------------------------------------------------------------------------
struct FoldSeg {
  const double* partials;
  Int blocks;
  int nvals;
  int slot0;
  unsigned max_mask;
};

struct FoldArgs {
  FoldSeg seg[2];
  int nseg;
  double* out;
};

__global__ void __launch_bounds__(kThreads) fold_kernel(FoldArgs a) {
  for (int s = 0; s < a.nseg; ++s) {
    const FoldSeg g = a.seg[s];
    for (int v = 0; v < g.nvals; ++v) {
      const bool mx = (g.max_mask >> v) & 1u;
      double acc = 0.0;
      for (Int b = static_cast<Int>(threadIdx.x); b < g.blocks; b += kThreads) {
        const double t = g.partials[static_cast<size_t>(b) * g.nvals + v];
        acc = mx ? fmax(acc, t) : acc + t;
      }
      double r[1] = {acc};
      block_fold_store<1>(r, mx ? 1u : 0u, a.out + g.slot0 + v);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Grid-stride reductions over vectors (the standalone backend operations).
// ---------------------------------------------------------------------------
template <class Op>
__global__ void __launch_bounds__(kThreads) reduce_kernel(Int len, Op op, double* partials) {
  double red[Op::kR];
#pragma unroll
  for (int r = 0; r < Op::kR; ++r) red[r] = 0.0;
  const Int stride = static_cast<Int>(gridDim.x) * kThreads;
  for (Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x); i < len;
       i += stride)
    op(i, red);
  block_fold_store<Op::kR>(red, Op::kMaxMask, partials + static_cast<size_t>(blockIdx.x) * Op::kR);
}

struct DotOp {
  static constexpr int kR = 1;
  static constexpr unsigned kMaxMask = 0;
  const double* a;
  const double* b;
  __device__ void operator()(Int i, double* red) const { red[0] += a[i] * b[i]; }
};

struct AbsMaxOp {
  static constexpr int kR = 1;
  static constexpr unsigned kMaxMask = 1;
  const double* a;
  __device__ void operator()(Int i, double* red) const { red[0] = fmax(red[0], fabs(a[i])); }
};

struct DotSumSqOp {  // dy'k_dx and dy'dy over the same rows
  static constexpr int kR = 2;
  static constexpr unsigned kMaxMask = 0;
  const double* dy;
  const double* k_dx;
  __device__ void operator()(Int i, double* red) const {
    red[0] += dy[i] * k_dx[i];
    red[1] += dy[i] * dy[i];
  }
};

struct ConvRowOp {  // backend.cpp:153-164
  static constexpr int kR = 3;
  static constexpr unsigned kMaxMask = 1u << 1;
  const double *q, *row_scale, *k_x, *y;
  Int neq;
  __device__ void operator()(Int i, double* red) const {
    const double scale = row_scale[i];
    const double row = scale != 0.0 ? k_x[i] / scale : k_x[i];
    const double slack = row - q[i];
    const double viol = (i < neq) ? slack : fmin(slack, 0.0);
    red[0] += viol * viol;
    red[1] = fmax(red[1], fabs(viol));
    red[2] += q[i] * (scale * y[i]);
  }
};

// backend.cpp:166-199 with a K^T y already in memory.
struct ConvColOp {
  static constexpr int kR = 4;
  static constexpr unsigned kMaxMask = 1u << 1;
  ConvProductEpi inner;
  const double* kt_y;
  __device__ void operator()(Int j, double* red) const { inner(j, kt_y[j], red); }
};

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------
// y = A x on the blocks of active LPs.
struct MultiPlainEpi {
  static constexpr int kR = 0;
  [[maybe_unused]] static constexpr unsigned kMaxMask = 0;
  double* y;
  ManyLps src;
  bool off;
  __device__ MultiPlainEpi resolve(int blk) const {
    MultiPlainEpi e = *this;
    e.off = src.get(blk)->active == 0;
    return e;
  }
  __device__ bool skip() const { return off; }
  __device__ void operator()(Int r, double s, double*) const { y[r] = s; }
};

// One thread per (LP, value): folds the LP's block partials in block order.
__global__ void __launch_bounds__(kThreads)
    multi_fold_kernel(const double* __restrict__ partials, int nvals, int slot0, unsigned max_mask,
                      const Int* __restrict__ lp_blocks, Int lps,
                      const IterParams* __restrict__ params, double* __restrict__ out) {
  const Int t = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  const Int l = t / nvals;
  const int v = static_cast<int>(t % nvals);
  if (l >= lps || params[l].active == 0) return;
  const bool mx = (max_mask >> v) & 1u;
  double acc = 0.0;
  for (Int b = lp_blocks[l]; b < lp_blocks[l + 1]; ++b) {
    const double x = partials[static_cast<size_t>(b) * nvals + v];
    acc = mx ? fmax(acc, x) : acc + x;
  }
  out[static_cast<size_t>(l) * kCheckSlots + slot0 + v] = acc;
}

// One workgroup per chunk (at most kTileNnz elements of one LP).
__global__ void __launch_bounds__(kThreads)
    multi_sumsq_kernel(const Int* __restrict__ chunk_start, const Int* __restrict__ chunk_lp,
                       const double* __restrict__ v, const double* from, double* base,
                       double* prev, int advance, const IterParams* __restrict__ params,
                       double* __restrict__ partials) {
  const Int c = static_cast<Int>(blockIdx.x);
  if (params[chunk_lp[c]].active == 0) return;  // uniform across the block
  const Int i0 = chunk_start[c], i1 = chunk_start[c + 1];
  double acc = 0.0;
  for (Int i = i0 + static_cast<Int>(threadIdx.x); i < i1; i += kThreads) {
    const double cur = v[i];
    const double e = from != nullptr ? cur - from[i] : cur;
    acc += e * e;
    if (advance) {
      if (prev != nullptr) prev[i] = base[i];
      base[i] = cur;
    }
  }
  double r[1] = {acc};
  block_fold_store<1>(r, 0u, partials + c);
}

__global__ void __launch_bounds__(kThreads)
    multi_chunk_fold_kernel(const double* __restrict__ partials, const Int* __restrict__ lp_chunks,
                            Int lps, const IterParams* __restrict__ params, double* __restrict__ out,
                            int stride, int slot) {
  const Int l = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (l >= lps || params[l].active == 0) return;
  double acc = 0.0;
  for (Int c = lp_chunks[l]; c < lp_chunks[l + 1]; ++c) acc += partials[c];
  out[static_cast<size_t>(l) * stride + slot] = acc;
}

__global__ void __launch_bounds__(kThreads)
    multi_divide_kernel(const Int* __restrict__ chunk_start, const Int* __restrict__ chunk_lp,
                        const double* __restrict__ v, const double* __restrict__ divisor,
                        const IterParams* __restrict__ params, double* __restrict__ out) {
  const Int c = static_cast<Int>(blockIdx.x);
  const Int l = chunk_lp[c];
  if (params[l].active == 0) return;
  const double d = divisor[l];
  for (Int i = chunk_start[c] + static_cast<Int>(threadIdx.x); i < chunk_start[c + 1]; i += kThreads)
    out[i] = v[i] / d;
}

__global__ void __launch_bounds__(kThreads)
    row_norms_kernel(const Int* __restrict__ start, Int rows, const double* __restrict__ value,
                     int one_norm, double* __restrict__ out) {
  if (!one_norm) {
    // The largest |a| of a row: kRowLanes lanes take a stride each and the
    // maxima meet by shuffles (a maximum has no order to keep).
    const long t = static_cast<long>(blockIdx.x) * kThreads + static_cast<long>(threadIdx.x);
    const Int r = static_cast<Int>(t / kRowLanes);
    const int sub = static_cast<int>(t % kRowLanes);
    double acc = 0.0;
    if (r < rows)
      for (Int k = start[r] + sub; k < start[r + 1]; k += kRowLanes) acc = fmax(acc, fabs(value[k]));
#pragma unroll
    for (int w = kRowLanes / 2; w > 0; w >>= 1) acc = fmax(acc, __shfl_xor(acc, w, kRowLanes));
    if (r < rows && sub == 0) out[r] = acc;
    return;
  }
  // The one-norm keeps the serial sum in storage order (the CPU's bits).
  const Int r = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (r >= rows) return;
  double acc = 0.0;
  for (Int k = start[r]; k < start[r + 1]; ++k) {
    const double v = fabs(value[k]);
    acc = one_norm ? acc + v : fmax(acc, v);
  }
  out[r] = acc;
}

__global__ void __launch_bounds__(kThreads)
    inv_sqrt_kernel(const double* __restrict__ norm, Int n, double* __restrict__ d) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i >= n) return;
  const double v = norm[i];
  d[i] = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
}

// a[i] += add[l] for the segment l with bounds[l] <= i < bounds[l + 1]
// (bounds: segments + 1 entries, bounds[0] = 0, bounds[segments] = count).
__global__ void __launch_bounds__(kThreads)
    add_segment_offsets_kernel(Int* __restrict__ a, Int count, const Int* __restrict__ bounds,
                               const Int* __restrict__ add, Int segments) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i >= count) return;
  Int lo = 0, hi = segments;
  while (hi - lo > 1) {
    const Int mid = lo + (hi - lo) / 2;
    if (bounds[mid] <= i) lo = mid;
    else hi = mid;
  }
  a[i] += add[lo];
}

__global__ void __launch_bounds__(kThreads)
    scale_values_kernel(const Int* __restrict__ start, const Int* __restrict__ index, Int rows,
                        double* __restrict__ value, const double* __restrict__ first,
                        const double* __restrict__ second, int first_by_row) {
  // kRowLanes lanes per row, striding it: each entry is scaled on its own,
  // so the split changes no bits, and a row's entries load together.
  const long t = static_cast<long>(blockIdx.x) * kThreads + static_cast<long>(threadIdx.x);
  const Int r = static_cast<Int>(t / kRowLanes);
  const Int sub = static_cast<Int>(t % kRowLanes);
  if (r >= rows) return;
  for (Int k = start[r] + sub; k < start[r + 1]; k += kRowLanes) {
    double v = value[k];
    if (first_by_row) {
      v = v * first[r];
      v = v * second[index[k]];
    } else {
      v = v * first[index[k]];
      v = v * second[r];
    }
    value[k] = v;
  }
}

__global__ void __launch_bounds__(kThreads)
    scale_vectors_kernel(Int m, const double* __restrict__ d1, double* __restrict__ q,
                         double* __restrict__ rs, Int n, const double* __restrict__ d2,
                         double* __restrict__ c, double* __restrict__ lo, double* __restrict__ hi,
                         double* __restrict__ cs) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < m) {
    q[i] *= d1[i];
    rs[i] *= d1[i];
  }
  if (i < n) {
    const double d = d2[i];
    c[i] *= d;
    lo[i] /= d;
    hi[i] /= d;
    cs[i] *= d;
  }
}

__global__ void fill_kernel(double* p, Int n, double v) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) p[i] = v;
}

// Bounds of the dual feasibility problem: a finite bound becomes 0, an
// infinite one stays.
__global__ void dual_bounds_kernel(Int n, const double* __restrict__ lo, const double* __restrict__ hi,
                                   double* __restrict__ dlo, double* __restrict__ dhi) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) {
    const double l = lo[i], h = hi[i];
    dlo[i] = l > -__builtin_inf() ? 0.0 : l;
    dhi[i] = h < __builtin_inf() ? 0.0 : h;
  }
}

// The cold primal start (pdhg.cpp:479-503 with no warm start): zero moved
// into the box, onto a finite bound it violates, else left at zero.
__global__ void cold_start_kernel(Int n, const double* __restrict__ lo, const double* __restrict__ hi,
                                  double* __restrict__ x) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) {
    const double l = lo[i], h = hi[i];
    double v = 0.0;
    if (v < l) v = isfinite(l) ? l : 0.0;
    if (v > h) v = isfinite(h) ? h : 0.0;
    x[i] = v;
  }
}

__global__ void accumulate_kernel(Int n, double w, const double* v, double* sum) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) sum[i] += w * v[i];
}

__global__ void scale_into_kernel(Int n, double w, const double* sum, double* out) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) out[i] = sum[i] / w;
}

// out = sum / sqrt(*sq), the divisor read on the device (the device's fp64
// sqrt and division are correctly rounded, so the bits are those of
// scale_into with the host's sqrt of the same value).
__global__ void scale_by_root_kernel(Int n, const double* sq, const double* sum, double* out) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i < n) out[i] = sum[i] / sqrt(*sq);
}

// The standalone fused updates (LinAlgBackend::primal_update/dual_update).
// Every input is read into registers before any output is written, so x_out
// may alias x and y_out / k_out may alias y / k_x.
__global__ void primal_update_kernel(Int n, double tau, double gamma, double h,
                                     const double* x, const double* c, const double* kt_y,
                                     const double* lower, const double* upper,
                                     const double* anchor, double* x_out, double* dx,
                                     double* x_bar) {
  const Int j = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (j >= n) return;
  const double xj = x[j];
  const double a = anchor != nullptr ? anchor[j] : 0.0;
  double v = xj - tau * (c[j] - kt_y[j]);
  if (v < lower[j]) v = lower[j];
  if (v > upper[j]) v = upper[j];
  const double d = v - xj;
  double next = v;
  if (gamma != 0.0) next += gamma * d;
  if (anchor != nullptr) next = h * next + (1.0 - h) * a;
  dx[j] = d;
  x_bar[j] = v + d;
  x_out[j] = next;
}

__global__ void dual_update_kernel(Int m, Int neq, double sigma, double gamma, double h,
                                   const double* y, const double* q, const double* k_x_bar,
                                   const double* k_x, const double* anchor_y,
                                   const double* anchor_kx, double* y_out, double* dy,
                                   double* k_out, double* k_dx) {
  const Int i = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (i >= m) return;
  const double yi = y[i];
  const double bar = k_x_bar[i];
  const double kx = k_x[i];
  const bool blend = anchor_y != nullptr;
  const double ay = blend ? anchor_y[i] : 0.0;
  const double ak = blend ? anchor_kx[i] : 0.0;
  double v = yi + sigma * (q[i] - bar);
  if (i >= neq && v < 0.0) v = 0.0;
  const double d = v - yi;
  const double kd = 0.5 * (bar - kx);
  double ny = v;
  double nk = 0.5 * (bar + kx);
  if (gamma != 0.0) {
    ny += gamma * d;
    nk += gamma * kd;
  }
  if (blend) {
    ny = h * ny + (1.0 - h) * ay;
    nk = h * nk + (1.0 - h) * ak;
  }
  dy[i] = d;
  k_dx[i] = kd;
  y_out[i] = ny;
  k_out[i] = nk;
}

inline unsigned grid_for(Int n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }
// Blocks for kRowLanes lanes per row.
inline unsigned lanes_grid(Int rows) {
  return static_cast<unsigned>((static_cast<long>(rows) * kRowLanes + kThreads - 1) / kThreads);
}

void check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("launch ") + what + ": " + hipGetErrorString(e));
}

void fold(hipStream_t s, std::initializer_list<FoldSeg> segs, double* out) {
  FoldArgs a{};
  a.nseg = 0;
  for (const FoldSeg& g : segs) a.seg[a.nseg++] = g;
  a.out = out;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kThreads), 0, s, a);
  check_launch("fold");
}

template <class Epi>
void spmv_with(hipStream_t s, const DevCsr& a, const double* x, const Epi& e, double* partials,
               const char* what, const LaunchEvents* ev = nullptr, bool fast = false) {
  if (a.blocks <= 0) return;
  if (!fast && a.max_rows > kTileRows)
    throw std::invalid_argument(std::string(what) + ": a fast-order plan in the storage order");
  const dim3 grid(static_cast<unsigned>(a.blocks));
  // The fast order runs on plans with the large tile (prepare_fast), the
  // storage order on plans with the small one.
  validate_launch(what, grid, dim3(kThreads), ev);
  const int blocks = static_cast<int>(a.blocks);
  const int f = (fast ? kModeWide | kModeStream : 0) | (a.stream_loads ? kModeStream : 0);
  auto launch = [&](auto kernel) {
    if (ev != nullptr && ev->start != nullptr) {
      hipExtLaunchKernelGGL(kernel, grid, dim3(kThreads), 0, s, ev->start, ev->stop, 0, a.start,
                            a.index, a.value, a.block_row, a.block_map, x, blocks, f, e,
                            partials);
    } else {
      hipLaunchKernelGGL(kernel, grid, dim3(kThreads), 0, s, a.start, a.index, a.value,
                         a.block_row, a.block_map, x, blocks, f, e, partials);
    }
  };
  if (fast) {
    launch(spmv_kernel<Epi, kFastTileNnz>);
  } else {
    launch(spmv_kernel<Epi, kTileNnz>);
  }
  check_launch(what);
}

// Adds each row's kPanels partial sums in panel order and runs the epilogue.
template <class Epi>
__global__ void __launch_bounds__(kThreads)
    panel_combine_kernel(Int rows, const double* __restrict__ sums, Epi epi,
                         double* __restrict__ partials) {
  constexpr int R = Epi::kR > 0 ? Epi::kR : 1;
  const Epi e = epi.resolve(0);
  if (e.skip()) return;
  double red[R];
#pragma unroll
  for (int i = 0; i < R; ++i) red[i] = 0.0;
  const Int r = static_cast<Int>(blockIdx.x) * kThreads + static_cast<Int>(threadIdx.x);
  if (r < rows) {
    double v = sums[r];
#pragma unroll
    for (int p = 1; p < kPanels; ++p) v += sums[static_cast<size_t>(p) * rows + r];
    e(r, v, red);
  }
  if constexpr (Epi::kR > 0) {
    block_fold_store<Epi::kR>(red, Epi::kMaxMask, partials + static_cast<size_t>(blockIdx.x) * Epi::kR);
  }
}

// The panel products and the combine are two dispatches, each timed by its
// own event pair (ev: the products, ev_combine: the combine).
template <class Epi>
void panel_step_with(hipStream_t s, const DevCsr& panels, double* sums, Int rows,
                     const double* x, const Epi& e, double* partials, const LaunchEvents* ev,
                     const LaunchEvents* ev_combine) {
  if (rows <= 0) return;
  spmv_with(s, panels, x, PlainEpi{sums}, nullptr, "panel products", ev, true);
  const dim3 grid(static_cast<unsigned>((rows + kThreads - 1) / kThreads));
  validate_launch("panel combine", grid, dim3(kThreads), ev_combine);
  if (ev_combine != nullptr && ev_combine->start != nullptr) {
    hipExtLaunchKernelGGL(panel_combine_kernel<Epi>, grid, dim3(kThreads), 0, s, ev_combine->start,
                          ev_combine->stop, 0, rows, sums, e, partials);
  } else {
    hipLaunchKernelGGL(panel_combine_kernel<Epi>, grid, dim3(kThreads), 0, s, rows, sums, e, partials);
  }
  check_launch("panel combine");
}

template <class Op>
void reduce_with(hipStream_t s, Int len, const Op& op, double* partials) {
  hipLaunchKernelGGL(reduce_kernel<Op>, dim3(static_cast<unsigned>(reduce_grid(len))),
                     dim3(kThreads), 0, s, len, op, partials);
  check_launch("reduce");
}

}  // namespace

// Greedy row blocks over the CSR: a block closes when the next row would push
// its nonzeros past a tile or its rows past one workgroup; a row longer than a
// tile is a block of its own.
void plan_blocks(Int rows, const Int* start, Int tile_nnz, int max_rows, std::int32_t** out_rows,
                 Int* out_blocks) {
  std::vector<Int> br;
  br.reserve(static_cast<size_t>(rows) / 16 + 2);
  br.push_back(0);
  Int r = 0;
  while (r < rows) {
    if (start[r + 1] - start[r] > tile_nnz) {
      ++r;
      br.push_back(r);
      continue;
    }
    Int nnz = 0;
    int count = 0;
    while (r < rows && count < max_rows) {
      const Int len = start[r + 1] - start[r];
      if (len > tile_nnz || nnz + len > tile_nnz) break;
      nnz += len;
      ++count;
      ++r;
    }
    br.push_back(r);
  }
  *out_blocks = static_cast<Int>(br.size() - 1);
  auto* mem = static_cast<std::int32_t*>(std::malloc(br.size() * sizeof(std::int32_t)));
  if (mem == nullptr) throw std::bad_alloc();
  std::copy(br.begin(), br.end(), mem);
  *out_rows = mem;
}

int reduce_grid(Int len) {
  const Int per = kThreads * 4;
  Int g = (len + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return static_cast<int>(g);
}

void launch_spmv(hipStream_t s, const DevCsr& a, const double* x, double* y,
                 const LaunchEvents* ev) {
  spmv_with(s, a, x, PlainEpi{y}, nullptr, "spmv", ev);
}

void launch_fill(hipStream_t s, double* p, Int n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
  check_launch("fill");
}

void launch_dual_bounds(hipStream_t s, Int n, const double* lo, const double* hi, double* dlo,
                        double* dhi) {
  if (n <= 0) return;
  hipLaunchKernelGGL(dual_bounds_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, lo, hi, dlo, dhi);
  check_launch("dual bounds");
}

void launch_cold_start(hipStream_t s, Int n, const double* lo, const double* hi, double* x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cold_start_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, lo, hi, x);
  check_launch("cold start");
}

void launch_accumulate(hipStream_t s, Int n, double w, const double* v, double* sum) {
  if (n <= 0) return;
  hipLaunchKernelGGL(accumulate_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, w, v, sum);
  check_launch("accumulate");
}

void launch_scale_into(hipStream_t s, Int n, double w, const double* sum, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_into_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, w, sum, out);
  check_launch("scale_into");
}

void launch_scale_by_root(hipStream_t s, Int n, const double* sq, const double* sum, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_by_root_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, sq, sum, out);
  check_launch("scale_by_root");
}

void launch_primal_update(hipStream_t s, Int n, double tau, double gamma, double h,
                          const double* x, const double* c, const double* kt_y,
                          const double* lower, const double* upper, const double* anchor,
                          double* x_out, double* dx, double* x_bar) {
  if (n <= 0) return;
  hipLaunchKernelGGL(primal_update_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, tau, gamma,
                     h, x, c, kt_y, lower, upper, anchor, x_out, dx, x_bar);
  check_launch("primal_update");
}

void launch_dual_update(hipStream_t s, Int m, Int neq, double sigma, double gamma, double h,
                        const double* y, const double* q, const double* k_x_bar,
                        const double* k_x, const double* anchor_y, const double* anchor_kx,
                        double* y_out, double* dy, double* k_out, double* k_dx) {
  if (m <= 0) return;
  hipLaunchKernelGGL(dual_update_kernel, dim3(grid_for(m)), dim3(kThreads), 0, s, m, neq, sigma,
                     gamma, h, y, q, k_x_bar, k_x, anchor_y, anchor_kx, y_out, dy, k_out, k_dx);
  check_launch("dual_update");
}

void launch_dot(hipStream_t s, const double* a, const double* b, Int n, double* partials,
                double* out) {
  reduce_with(s, n, DotOp{a, b}, partials);
  fold(s, {FoldSeg{partials, reduce_grid(n), 1, 0, 0u}}, out);
}

void launch_sumsq2(hipStream_t s, const double* a, Int n, const double* b, Int m,
                   double* partials_a, double* partials_b, double* out2) {
  reduce_with(s, n, DotOp{a, a}, partials_a);
  reduce_with(s, m, DotOp{b, b}, partials_b);
  fold(s, {FoldSeg{partials_a, reduce_grid(n), 1, 0, 0u}, FoldSeg{partials_b, reduce_grid(m), 1, 1, 0u}},
       out2);
}

void launch_absmax(hipStream_t s, const double* a, Int n, double* partials, double* out) {
  reduce_with(s, n, AbsMaxOp{a}, partials);
  fold(s, {FoldSeg{partials, reduce_grid(n), 1, 0, 1u}}, out);
}

void launch_dot_and_sumsq(hipStream_t s, const double* dy, const double* k_dx, Int m,
                          const double* dx, Int n, double* partials_a, double* partials_b,
                          double* out3) {
  reduce_with(s, m, DotSumSqOp{dy, k_dx}, partials_a);
  reduce_with(s, n, DotOp{dx, dx}, partials_b);
  fold(s, {FoldSeg{partials_a, reduce_grid(m), 2, 0, 0u}, FoldSeg{partials_b, reduce_grid(n), 1, 2, 0u}},
       out3);
}

void launch_convergence_terms(hipStream_t s, Int n, Int m, const DevProblem& p,
                              const double* x, const double* k_x, const double* y,
                              const double* kt_y, double* partials_a, double* partials_b,
                              double* out7) {
  ConvRowOp row{p.orig_q, p.row_scale, k_x, y, p.num_equalities};
  ConvColOp col{ConvProductEpi{p, x, OneLp{nullptr}, false}, kt_y};
  reduce_with(s, m, row, partials_a);
  reduce_with(s, n, col, partials_b);
  fold(s, {FoldSeg{partials_a, reduce_grid(m), 3, 0, ConvRowOp::kMaxMask},
           FoldSeg{partials_b, reduce_grid(n), 4, 3, ConvColOp::kMaxMask}},
       out7);
}

// The allocation-free entry points' scratch: a file-scope device array. In
// host code its name is only a stand-in; the kernels get the address that
// hipGetSymbolAddress returns on each device.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of `scratch` there
}  // namespace

void noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.count(device) != 0) return;
  int current = -1;
  if (hipGetDevice(&current) != hipSuccess || current != device)
    throw std::runtime_error("noalloc scratch: device " + std::to_string(device) + " is not current");
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(ptr);
}

double* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end())
    throw std::runtime_error("noalloc scratch: not set up on device " + std::to_string(device));
  return it->second;
}

void launch_primal_step(hipStream_t s, const DevCsr& kt, const double* y, const DevProblem& p,
                        const IterParams* params, int step, const PrimalBufs& b, unsigned flags,
                        double* partials, const LaunchEvents* ev, bool fast) {
  if (flags & kStepDiag) {
    spmv_with(s, kt, y, PrimalStepEpi<true>{p, b, OneLp{params}, step, {}, false}, partials,
              "primal_step", ev, fast);
  } else {
    spmv_with(s, kt, y, PrimalStepEpi<false>{p, b, OneLp{params}, step, {}, false}, partials,
              "primal_step", ev, fast);
  }
}

void launch_dual_step(hipStream_t s, const DevCsr& k, const double* x_bar, const DevProblem& p,
                      const IterParams* params, int step, const DualBufs& b, unsigned flags,
                      double* partials, const LaunchEvents* ev, bool fast) {
  if (flags & kStepDiag) {
    spmv_with(s, k, x_bar, DualStepEpi<true>{p, b, OneLp{params}, step, {}, false}, partials,
              "dual_step", ev, fast);
  } else {
    spmv_with(s, k, x_bar, DualStepEpi<false>{p, b, OneLp{params}, step, {}, false}, partials,
              "dual_step", ev, fast);
  }
}

Int panel_combine_blocks(Int rows) { return (rows + kThreads - 1) / kThreads; }

void launch_dual_step_panels(hipStream_t s, const DevCsr& panels, double* panel_sums,
                             const double* x_bar, const DevProblem& p, const IterParams* params,
                             int step, const DualBufs& b, unsigned flags, double* partials,
                             const LaunchEvents* ev, const LaunchEvents* ev_combine) {
  const Int m = p.m;
  if (panels.rows != kPanels * m) throw std::invalid_argument("dual step: panel stack does not fit K");
  if (flags & kStepDiag) {
    panel_step_with(s, panels, panel_sums, m, x_bar,
                    DualStepEpi<true>{p, b, OneLp{params}, step, {}, false}, partials, ev,
                    ev_combine);
  } else {
    panel_step_with(s, panels, panel_sums, m, x_bar,
                    DualStepEpi<false>{p, b, OneLp{params}, step, {}, false}, partials, ev,
                    ev_combine);
  }
}

void launch_conv_product(hipStream_t s, const DevCsr& kt, const double* y, const double* x,
                         const DevProblem& p, double* partials, bool fast) {
  spmv_with(s, kt, y, ConvProductEpi{p, x, OneLp{nullptr}, false}, partials, "conv_product",
            nullptr, fast);
}

void launch_farkas_product(hipStream_t s, const DevCsr& kt, const double* dy,
                           const DevProblem& p, const double* check, double* partials, bool fast) {
  spmv_with(s, kt, dy, FarkasProductEpi{p, check, OneLp{nullptr}, false}, partials,
            "farkas_product", nullptr, fast);
}

void launch_multi_primal_step(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                              const double* y, const DevProblem& p, const IterParams* params,
                              int step, const PrimalBufs& b, bool diag, double* partials,
                              const LaunchEvents* ev) {
  const ManyLps src{params, block_lp};
  if (diag) {
    spmv_with(s, kt, y, PrimalStepEpi<true, ManyLps>{p, b, src, step, {}, false}, partials,
              "multi primal_step", ev);
  } else {
    spmv_with(s, kt, y, PrimalStepEpi<false, ManyLps>{p, b, src, step, {}, false}, partials,
              "multi primal_step", ev);
  }
}

void launch_multi_dual_step(hipStream_t s, const DevCsr& k, const Int* block_lp,
                            const double* x_bar, const DevProblem& p, const IterParams* params,
                            int step, const DualBufs& b, bool diag, double* partials,
                            const LaunchEvents* ev) {
  const ManyLps src{params, block_lp};
  if (diag) {
    spmv_with(s, k, x_bar, DualStepEpi<true, ManyLps>{p, b, src, step, {}, false}, partials,
              "multi dual_step", ev);
  } else {
    spmv_with(s, k, x_bar, DualStepEpi<false, ManyLps>{p, b, src, step, {}, false}, partials,
              "multi dual_step", ev);
  }
}

void launch_multi_conv_product(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                               const double* y, const double* x, const DevProblem& p,
                               const IterParams* params, double* partials) {
  spmv_with(s, kt, y, ConvProductEpiT<ManyLps>{p, x, ManyLps{params, block_lp}, false}, partials,
            "multi conv_product");
}

void launch_multi_farkas_product(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                                 const double* dy, const DevProblem& p, const IterParams* params,
                                 const double* check, double* partials) {
  spmv_with(s, kt, dy, FarkasProductEpiT<ManyLps>{p, check, ManyLps{params, block_lp}, false},
            partials, "multi farkas_product");
}

void launch_multi_product(hipStream_t s, const DevCsr& a, const Int* block_lp, const double* x,
                          double* y, const IterParams* params) {
  spmv_with(s, a, x, MultiPlainEpi{y, ManyLps{params, block_lp}, false}, nullptr, "multi product");
}

void launch_multi_fold(hipStream_t s, const double* partials, int nvals, int slot0,
                       unsigned max_mask, const Int* lp_blocks, Int lps, const IterParams* params,
                       double* out) {
  if (lps <= 0) return;
  const Int threads = lps * nvals;
  hipLaunchKernelGGL(multi_fold_kernel, dim3(static_cast<unsigned>((threads + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, partials, nvals, slot0, max_mask, lp_blocks, lps, params,
                     out);
  check_launch("multi fold");
}

void launch_multi_sumsq(hipStream_t s, const ChunkPlan& plan, const double* v, const double* from,
                        double* base, double* prev, bool advance, const IterParams* params,
                        double* partials, const Int* lp_chunks, Int lps, double* out, int stride,
                        int slot) {
  if (plan.chunks <= 0 || lps <= 0) return;
  hipLaunchKernelGGL(multi_sumsq_kernel, dim3(static_cast<unsigned>(plan.chunks)), dim3(kThreads), 0,
                     s, plan.start, plan.lp, v, from, base, prev, advance ? 1 : 0, params, partials);
  check_launch("multi sumsq");
  hipLaunchKernelGGL(multi_chunk_fold_kernel, dim3(static_cast<unsigned>((lps + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, partials, lp_chunks, lps, params, out, stride, slot);
  check_launch("multi chunk fold");
}

void launch_multi_divide(hipStream_t s, const ChunkPlan& plan, const double* v,
                         const double* divisor, const IterParams* params, double* out) {
  if (plan.chunks <= 0) return;
  hipLaunchKernelGGL(multi_divide_kernel, dim3(static_cast<unsigned>(plan.chunks)), dim3(kThreads), 0,
                     s, plan.start, plan.lp, v, divisor, params, out);
  check_launch("multi divide");
}

void launch_row_norms(hipStream_t s, const Int* start, Int rows, const double* value, bool one_norm,
                      double* out) {
  if (rows <= 0) return;
  const unsigned grid = one_norm ? grid_for(rows) : lanes_grid(rows);
  hipLaunchKernelGGL(row_norms_kernel, dim3(grid), dim3(kThreads), 0, s, start, rows, value,
                     one_norm ? 1 : 0, out);
  check_launch("row norms");
}

void launch_inv_sqrt(hipStream_t s, const double* norm, Int n, double* d) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_sqrt_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, norm, n, d);
  check_launch("inverse square roots");
}

void launch_add_segment_offsets(hipStream_t s, Int* a, Int count, const Int* bounds, const Int* add,
                                Int segments) {
  if (count <= 0 || segments <= 0) return;
  hipLaunchKernelGGL(add_segment_offsets_kernel, dim3(grid_for(count)), dim3(kThreads), 0, s, a, count,
                     bounds, add, segments);
  check_launch("segment offsets");
}

void launch_scale_values(hipStream_t s, const Int* start, const Int* index, Int rows,
                         double* value, const double* first, const double* second,
                         bool first_by_row) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(scale_values_kernel, dim3(lanes_grid(rows)), dim3(kThreads), 0, s, start, index,
                     rows, value, first, second, first_by_row ? 1 : 0);
  check_launch("scale values");
}

void launch_scale_vectors(hipStream_t s, Int m, const double* d1, double* q, double* row_scale,
                          Int n, const double* d2, double* c, double* lower, double* upper,
                          double* col_scale) {
  const Int len = m > n ? m : n;
  if (len <= 0) return;
  hipLaunchKernelGGL(scale_vectors_kernel, dim3(grid_for(len)), dim3(kThreads), 0, s, m, d1, q,
                     row_scale, n, d2, c, lower, upper, col_scale);
  check_launch("scale vectors");
}

void launch_check_fold_rows(hipStream_t s, Int primal_blocks, const double* pa, Int dual_blocks,
                            const double* pb, double* out) {
  fold(s, {FoldSeg{pa, primal_blocks, 1, kSumDx2, 0u},
           FoldSeg{pb, dual_blocks, 7, kSumDy2, DualStepEpi<true>::kMaxMask}},
       out);
}

void launch_check_fold_cols(hipStream_t s, Int conv_blocks, const double* pc, Int farkas_blocks,
                            const double* pd, double* out) {
  if (pd != nullptr) {
    fold(s, {FoldSeg{pc, conv_blocks, 4, kDualSq, ConvProductEpi::kMaxMask},
             FoldSeg{pd, farkas_blocks, 2, kFarkasColBad, FarkasProductEpi::kMaxMask}},
         out);
  } else {
    fold(s, {FoldSeg{pc, conv_blocks, 4, kDualSq, ConvProductEpi::kMaxMask}}, out);
  }
}

}  // namespace anuk
