// gfx950 kernels of the batched solver (see batch_kernels.hpp). Compiled with
// -ffp-contract=off like the plain kernels: each lane's arithmetic is the
// reference's, operation for operation.
#include "batch_kernels.hpp"
#include "lanes.hpp"

#include <hip/hip_ext.h>

#include <stdexcept>
#include <string>

namespace anuk {

namespace {

using lanes::at;
using lanes::kWavesPerBlock;
using lanes::row_sum;
using lanes::wave_row;

// The slot span stored after the 64 LaneParams (LaneTable::span).
__device__ __forceinline__ Int span_of(const LaneParams* params) {
  return reinterpret_cast<const LaneTable*>(params)->span;
}

struct LaneScalars {
  double tau, sigma, gamma, w;
  bool blend, open_epoch;
};

// The Halpern schedule of one step, per lane (as step_scalars in kernels.hip).
__device__ __forceinline__ LaneScalars lane_scalars(const LaneParams& p, int step) {
  LaneScalars s;
  s.tau = p.tau;
  s.sigma = p.sigma;
  s.gamma = p.gamma;
  if (!p.halpern) {
    s.blend = false;
    s.open_epoch = false;
    s.w = 1.0;
  } else if (p.reset_at_start && step == 0) {
    s.blend = false;
    s.open_epoch = true;
    s.w = 1.0;
  } else {
    const Int k = p.reset_at_start ? step - 1 : p.k_at_start + step;
    s.blend = true;
    s.open_epoch = false;
    s.w = static_cast<double>(k + 1) / static_cast<double>(k + 2);
  }
  return s;
}

// The step kernels issue every load that does not depend on the row sum
// (the lane's scalars, the epilogue operands) before the gathers, so the
// wave waits for one round of memory latency rather than three. A retired
// lane issues no vector load and stores nothing: with the live lanes kept in
// the lowest slots, a row touches only the cache lines that hold them.
// The leading arguments (the row structure, the gathered vector, the row
// count, the slot span, the grid size, the step and the lane table) are
// plain scalars so they can be preloaded into SGPRs at wave launch
// (compiled with kernel-argument preloading). The epilogue's buffers come as
// structs, loaded while the row sum is in flight.
//
// The row structure is the CSR (start, index, value) when kW is 0, and the
// fixed-width copy (row, index, value of DevCsr::slot_*) of width kW
// otherwise: then the row's first loads depend on the row number alone, and
// a row longer than kW continues in the CSR's index and value arrays (the
// trailing arguments).
// The row-sum halves for either structure: with the CSR, the gather half
// does nothing and the finish half takes the whole row sum.
template <int kW>
struct StepRow {
  lanes::SlotRow<kW> r;
};
template <>
struct StepRow<0> {};

template <int kW>
__device__ __forceinline__ void step_gather(const Int* __restrict__ s0, const Int* __restrict__ s1,
                                            const double* __restrict__ s2, Int row,
                                            const double* __restrict__ v, Int v_rows, int lane,
                                            StepRow<kW>& g) {
  if constexpr (kW != 0)
    lanes::slot_gather<kW>(lanes::SlotArrays{s0, s1, s2}, row, v, v_rows, lane, g.r);
}

template <int kW>
__device__ __forceinline__ double step_finish(const Int* __restrict__ s0,
                                              const Int* __restrict__ s1,
                                              const double* __restrict__ s2,
                                              const Int* __restrict__ index,
                                              const double* __restrict__ value, Int row,
                                              const double* __restrict__ v, int lane,
                                              const StepRow<kW>& g) {
  if constexpr (kW == 0) {
    return lanes::row_sum(lanes::CsrArrays{s0, s1, s2}, row, v, lane);
  } else {
    return lanes::slot_finish<kW>(lanes::SlotArrays{s0, s1, s2},
                                  lanes::CsrArrays{nullptr, index, value}, row, g.r, v, lane);
  }
}

// The primal step's epilogue for one (row, lane): the projected step,
// reflection and Halpern blend of backend.cpp:60-85, in the reference's
// operation order.
template <bool kDiag>
__device__ __forceinline__ void primal_epilogue(const LaneParams& P, int step, size_t o, double xj,
                                                double cj, double lo, double hi, double anc,
                                                double kty, const PrimalBufs& b) {
  const LaneScalars sc = lane_scalars(P, step);
  double v = xj - sc.tau * (cj - kty);
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  const double dd = v - xj;
  stream_out(b.x_bar + o, v + dd);
  double next = v;
  if (sc.gamma != 0.0) next += sc.gamma * dd;
  if (sc.blend) next = sc.w * next + (1.0 - sc.w) * anc;
  stream_out(b.x_out + o, next);
  if (sc.open_epoch) stream_out(b.anchor_out + o, next);
  if constexpr (kDiag) stream_out(b.dx + o, dd);
}

// The dual step's epilogue for one (row, lane) (backend.cpp:87-111).
template <bool kDiag>
__device__ __forceinline__ void dual_epilogue(const LaneParams& P, int step, Int i, Int neq,
                                              size_t o, double yi, double kx, double qi, double ay,
                                              double ak, double kxb, const DualBufs& b) {
  const LaneScalars sc = lane_scalars(P, step);
  double v = yi + sc.sigma * (qi - kxb);
  if (i >= neq && v < 0.0) v = 0.0;
  const double dd = v - yi;
  const double kd = 0.5 * (kxb - kx);
  double ny = v;
  double nk = 0.5 * (kxb + kx);
  if (sc.gamma != 0.0) {
    ny += sc.gamma * dd;
    nk += sc.gamma * kd;
  }
  if (sc.blend) {
    ny = sc.w * ny + (1.0 - sc.w) * ay;
    nk = sc.w * nk + (1.0 - sc.w) * ak;
  }
  stream_out(b.y_out + o, ny);
  stream_out(b.k_out + o, nk);
  if (sc.open_epoch) {
    stream_out(b.anchor_y_out + o, ny);
    stream_out(b.anchor_kx_out + o, nk);
  }
  if constexpr (kDiag) {
    stream_out(b.dy + o, dd);
    stream_out(b.k_dx + o, kd);
  }
}

// The register budget of a step kernel: a wide fixed-width row keeps all its
// gathers and the epilogue's operands in flight, which needs more than the
// 64 VGPRs of eight waves per SIMD; four leave 128.
template <int kW>
inline constexpr int kStepMaxWaves = kW >= 32 ? 4 : 8;

template <bool kDiag, int kWaves, int kW>
__global__ void __launch_bounds__(64 * kWaves)
    __attribute__((amdgpu_waves_per_eu(1, kStepMaxWaves<kW>)))
    bprimal_kernel(const Int* __restrict__ s0, const Int* __restrict__ s1,
                   const double* __restrict__ s2, const double* __restrict__ y, Int rows,
                   Int span, int blocks, int step, Int y_rows, const LaneParams* __restrict__ params,
                   BatchData d, PrimalBufs b, const Int* __restrict__ kt_index,
                   const double* __restrict__ kt_value) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int j = lanes::wave_row_of<kWaves>(blocks);
  if (j >= rows || lane >= span) return;
  const size_t o = at(j, lane);
  StepRow<kW> g;
  step_gather<kW>(s0, s1, s2, j, y, y_rows, lane, g);
  const LaneParams P = params[lane];
  const double xj = b.x[o], cj = d.c[o], lo = d.lower[o], hi = d.upper[o], anc = b.anchor[o];
  const double kty = step_finish<kW>(s0, s1, s2, kt_index, kt_value, j, y, lane, g);
  if (!P.active) return;
  primal_epilogue<kDiag>(P, step, o, xj, cj, lo, hi, anc, kty, b);
}

template <bool kDiag, int kWaves, int kW>
__global__ void __launch_bounds__(64 * kWaves)
    __attribute__((amdgpu_waves_per_eu(1, kStepMaxWaves<kW>)))
    bdual_kernel(const Int* __restrict__ s0, const Int* __restrict__ s1,
                 const double* __restrict__ s2, const double* __restrict__ x_bar, Int rows,
                 Int span, int blocks, int step, Int x_rows, const LaneParams* __restrict__ params,
                 BatchData d, DualBufs b, const Int* __restrict__ k_index,
                 const double* __restrict__ k_value) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int i = lanes::wave_row_of<kWaves>(blocks);
  if (i >= rows || lane >= span) return;
  const size_t o = at(i, lane);
  StepRow<kW> g;
  step_gather<kW>(s0, s1, s2, i, x_bar, x_rows, lane, g);
  const LaneParams P = params[lane];
  const double yi = b.y[o], kx = b.k_x[o], qi = d.q[o], ay = b.anchor_y[o], ak = b.anchor_kx[o];
  const double kxb = step_finish<kW>(s0, s1, s2, k_index, k_value, i, x_bar, lane, g);
  if (!P.active) return;
  dual_epilogue<kDiag>(P, step, i, d.num_equalities, o, yi, kx, qi, ay, ak, kxb, b);
}

// Binned steps for a narrow span: the rows are binned by length
// (DevCsr::bin_plan). The first short_waves waves each take kG rows of at
// most kBinShortMax entries, a lane per (row, slot): slot s of the live
// lanes, kept in the lowest 64 / kG slots by compaction, on row r; every
// longer row then gets a wave of its own, summed in one round
// (lanes::round_sum_loaded: the row's entries in one vector load, all its
// gathers at once). A lane reads its row's {row, start, length} from the
// plan and, in parallel, the row's entries from their fixed-width copy
// (DevCsr::bin_index8 / bin_index32), which is addressed by the plan entry
// alone; a long row past kBinLongWidth entries is summed from the CSR in
// pipelined rounds instead. Same bits as the one-row kernel.
struct BinArgs {
  const int4* __restrict__ plan;
  const Int* __restrict__ index8;
  const double* __restrict__ value8;
  const Int* __restrict__ index32;
  const double* __restrict__ value32;
};
BinArgs bin_args(const DevCsr& a, int p) {
  return {reinterpret_cast<const int4*>(a.bin_plan[p]), a.bin_index8, a.bin_value8, a.bin_index32,
          a.bin_value32};
}

// A short entry's row: its 8 indices and values in 16-byte loads from the
// fixed-width copy (bin8_load, issued before the plan entry is back), then
// the gathers together and the terms in storage order (bin8_sum: a position
// past len adds +0.0; the copy's padding repeats the row's last index, a
// line the row reads anyway).
struct Bin8 {
  int4 i0, i1;
  double2 w0, w1, w2, w3;
};
__device__ __forceinline__ Bin8 bin8_load(const Int* __restrict__ index8,
                                          const double* __restrict__ value8, Int e) {
  const int4* ip = reinterpret_cast<const int4*>(index8 + static_cast<size_t>(e) * kBinShortMax);
  const double2* wp =
      reinterpret_cast<const double2*>(value8 + static_cast<size_t>(e) * kBinShortMax);
  return {ip[0], ip[1], wp[0], wp[1], wp[2], wp[3]};
}
// Marks a short entry's loaded row as used here. The lanes of a padding
// entry or of a retired LP store nothing, so without this the compiler moves
// their test ahead of the row's loads, which then wait for the plan entry
// and the lane's parameters instead of going out beside them.
__device__ __forceinline__ void bin8_keep(const Bin8& r) {
  asm volatile("" ::"v"(r.i0.x), "v"(r.i0.y), "v"(r.i0.z), "v"(r.i0.w), "v"(r.i1.x), "v"(r.i1.y),
               "v"(r.i1.z), "v"(r.i1.w));
  asm volatile("" ::"v"(r.w0.x), "v"(r.w0.y), "v"(r.w1.x), "v"(r.w1.y), "v"(r.w2.x), "v"(r.w2.y),
               "v"(r.w3.x), "v"(r.w3.y));
}
__device__ __forceinline__ double bin8_sum(const Bin8& r, Int len, const double* __restrict__ v,
                                           int slot) {
  const Int ix[8] = {r.i0.x, r.i0.y, r.i0.z, r.i0.w, r.i1.x, r.i1.y, r.i1.z, r.i1.w};
  const double w[8] = {r.w0.x, r.w0.y, r.w1.x, r.w1.y, r.w2.x, r.w2.y, r.w3.x, r.w3.y};
  double x[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) x[u] = v[at(ix[u], slot)];
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const double t = w[u] * x[u];
    s += u < len ? t : 0.0;
  }
  return s;
}

// A long row's sum (every lane of the wave calls this): its first
// kBinLongWidth entries, one per lane, from the fixed-width copy (my_ix,
// my_w: loaded by the caller ahead of the plan entry), summed in one round
// of 16 or 32 gathers; a longer row (or a gathered vector too large for the
// round's 31-bit offsets) in range_sum's rounds from the CSR.
__device__ __forceinline__ double bin32_sum(Int my_ix, double my_w, const Int* __restrict__ index,
                                            const double* __restrict__ value, Int lo, Int len,
                                            const double* __restrict__ v, Int v_rows, int lane,
                                            bool live) {
  if (len > kBinLongWidth || v_rows >= kSlotMaxCols) {
    if (!live) return 0.0;
    return lanes::range_sum(lanes::CsrArrays{nullptr, index, value}, lo, lo + len, v, lane, 0.0);
  }
  if (len <= 16) return lanes::round_sum_loaded<16>(my_ix, my_w, len, v, v_rows, lane, live);
  return lanes::round_sum_loaded<kBinLongWidth>(my_ix, my_w, len, v, v_rows, lane, live);
}

template <bool kDiag, int kWaves, int kG>
__global__ void __launch_bounds__(64 * kWaves)
    bprimal_binned_kernel(const int4* __restrict__ plan, const Int* __restrict__ index8,
                          const double* __restrict__ value8, const Int* __restrict__ index32,
                          const double* __restrict__ value32, Int short_waves, Int units,
                          int blocks, int step, Int span, Int y_rows,
                          const Int* __restrict__ kt_index, const double* __restrict__ kt_value,
                          const double* __restrict__ y, const LaneParams* __restrict__ params,
                          BatchData d, PrimalBufs b) {
  constexpr int kSpan = 64 / kG;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int w = lanes::wave_row_of<kWaves>(blocks);
  if (w >= units) return;  // the last workgroup's spare waves
  if (w < short_waves) {
    const int slot = lane % kSpan;
    const Int en = w * kG + lane / kSpan;
    const Bin8 r = bin8_load(index8, value8, en);
    const int4 e = plan[en];
    const LaneParams P = params[slot];
    bin8_keep(r);
    // A padding entry (row -1) computes on row 0 and stores nothing.
    const bool pad = e.x < 0;
    const Int j = pad ? 0 : e.x;
    const size_t o = at(j, slot);
    const double xj = b.x[o], cj = d.c[o], lo = d.lower[o], hi = d.upper[o], anc = b.anchor[o];
    const double kty = bin8_sum(r, e.z, y, slot);
    if (pad || !P.active) return;
    primal_epilogue<kDiag>(P, step, o, xj, cj, lo, hi, anc, kty, b);
  } else {
    // A long row: every lane takes part in its row sum.
    const Int q = w - short_waves;
    const size_t at32 = static_cast<size_t>(q) * kBinLongWidth + static_cast<size_t>(lane & 31);
    const Int my_ix = index32[at32];
    const double my_w = value32[at32];
    const int4 e = plan[short_waves * kG + q];
    const Int j = e.x;
    const bool live = lane < span;
    const size_t o = at(j, lane);
    const LaneParams P = params[lane];
    double xj = 0.0, cj = 0.0, lo = 0.0, hi = 0.0, anc = 0.0;
    if (live) {
      xj = b.x[o];
      cj = d.c[o];
      lo = d.lower[o];
      hi = d.upper[o];
      anc = b.anchor[o];
    }
    const double kty = bin32_sum(my_ix, my_w, kt_index, kt_value, e.y, e.z, y, y_rows, lane, live);
    if (!live || !P.active) return;
    primal_epilogue<kDiag>(P, step, o, xj, cj, lo, hi, anc, kty, b);
  }
}

template <bool kDiag, int kWaves, int kG>
__global__ void __launch_bounds__(64 * kWaves)
    bdual_binned_kernel(const int4* __restrict__ plan, const Int* __restrict__ index8,
                        const double* __restrict__ value8, const Int* __restrict__ index32,
                        const double* __restrict__ value32, Int short_waves, Int units,
                        int blocks, int step, Int span, Int x_rows,
                        const Int* __restrict__ k_index, const double* __restrict__ k_value,
                        const double* __restrict__ x_bar, const LaneParams* __restrict__ params,
                        BatchData d, DualBufs b) {
  constexpr int kSpan = 64 / kG;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int w = lanes::wave_row_of<kWaves>(blocks);
  if (w >= units) return;  // the last workgroup's spare waves
  if (w < short_waves) {
    const int slot = lane % kSpan;
    const Int en = w * kG + lane / kSpan;
    const Bin8 r = bin8_load(index8, value8, en);
    const int4 e = plan[en];
    const LaneParams P = params[slot];
    bin8_keep(r);
    // A padding entry (row -1) computes on row 0 and stores nothing.
    const bool pad = e.x < 0;
    const Int i = pad ? 0 : e.x;
    const size_t o = at(i, slot);
    const double yi = b.y[o], kx = b.k_x[o], qi = d.q[o], ay = b.anchor_y[o], ak = b.anchor_kx[o];
    const double kxb = bin8_sum(r, e.z, x_bar, slot);
    if (pad || !P.active) return;
    dual_epilogue<kDiag>(P, step, i, d.num_equalities, o, yi, kx, qi, ay, ak, kxb, b);
  } else {
    // A long row: every lane takes part in its row sum.
    const Int q = w - short_waves;
    const size_t at32 = static_cast<size_t>(q) * kBinLongWidth + static_cast<size_t>(lane & 31);
    const Int my_ix = index32[at32];
    const double my_w = value32[at32];
    const int4 e = plan[short_waves * kG + q];
    const Int i = e.x;
    const bool live = lane < span;
    const size_t o = at(i, lane);
    const LaneParams P = params[lane];
    double yi = 0.0, kx = 0.0, qi = 0.0, ay = 0.0, ak = 0.0;
    if (live) {
      yi = b.y[o];
      kx = b.k_x[o];
      qi = d.q[o];
      ay = b.anchor_y[o];
      ak = b.anchor_kx[o];
    }
    const double kxb =
        bin32_sum(my_ix, my_w, k_index, k_value, e.y, e.z, x_bar, x_rows, lane, live);
    if (!live || !P.active) return;
    dual_epilogue<kDiag>(P, step, i, d.num_equalities, o, yi, kx, qi, ay, ak, kxb, b);
  }
}

// The same steps with kRows consecutive rows per wave, their nonzeros taken
// as one segment (lanes::segment_sums): fewer waves, and every load of the
// wave's rows in flight at once. Each lane's arithmetic is the one-row
// kernel's, term for term, so every shape gives the same bits.
constexpr int kSegment = 16;

template <bool kDiag, int kWaves, int kRows>
__global__ void __launch_bounds__(64 * kWaves)
    bprimal_rows_kernel(DevCsr kt, const double* __restrict__ y, BatchData d, PrimalBufs b,
                        const LaneParams* __restrict__ params, int step, Int span) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int j0 = lanes::wave_row_in<kWaves>() * kRows;
  if (j0 >= kt.rows || lane >= span) return;
  const LaneParams P = params[lane];
  double xj[kRows], cj[kRows], lo[kRows], hi[kRows], anc[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (j0 + r < kt.rows) {
      const size_t o = at(j0 + r, lane);
      xj[r] = b.x[o];
      cj[r] = d.c[o];
      lo[r] = d.lower[o];
      hi[r] = d.upper[o];
      anc[r] = b.anchor[o];
    }
  }
  double kty[kRows];
  lanes::segment_sums<kRows, kSegment>(kt, j0, y, lane, kty);
  if (!P.active) return;
  const LaneScalars sc = lane_scalars(P, step);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (j0 + r >= kt.rows) break;
    const size_t o = at(j0 + r, lane);
    double v = xj[r] - sc.tau * (cj[r] - kty[r]);
    if (v < lo[r]) v = lo[r];
    if (v > hi[r]) v = hi[r];
    const double dd = v - xj[r];
    stream_out(b.x_bar + o, v + dd);
    double next = v;
    if (sc.gamma != 0.0) next += sc.gamma * dd;
    if (sc.blend) next = sc.w * next + (1.0 - sc.w) * anc[r];
    stream_out(b.x_out + o, next);
    if (sc.open_epoch) stream_out(b.anchor_out + o, next);
    if constexpr (kDiag) stream_out(b.dx + o, dd);
  }
}

template <bool kDiag, int kWaves, int kRows>
__global__ void __launch_bounds__(64 * kWaves)
    bdual_rows_kernel(DevCsr k, const double* __restrict__ x_bar, BatchData d, DualBufs b,
                      const LaneParams* __restrict__ params, int step, Int span) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int i0 = lanes::wave_row_in<kWaves>() * kRows;
  if (i0 >= k.rows || lane >= span) return;
  const LaneParams P = params[lane];
  double yi[kRows], kx[kRows], qi[kRows], ay[kRows], ak[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (i0 + r < k.rows) {
      const size_t o = at(i0 + r, lane);
      yi[r] = b.y[o];
      kx[r] = b.k_x[o];
      qi[r] = d.q[o];
      ay[r] = b.anchor_y[o];
      ak[r] = b.anchor_kx[o];
    }
  }
  double kxb[kRows];
  lanes::segment_sums<kRows, kSegment>(k, i0, x_bar, lane, kxb);
  if (!P.active) return;
  const LaneScalars sc = lane_scalars(P, step);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const Int i = i0 + r;
    if (i >= k.rows) break;
    const size_t o = at(i, lane);
    double v = yi[r] + sc.sigma * (qi[r] - kxb[r]);
    if (i >= d.num_equalities && v < 0.0) v = 0.0;
    const double dd = v - yi[r];
    const double kd = 0.5 * (kxb[r] - kx[r]);
    double ny = v;
    double nk = 0.5 * (kxb[r] + kx[r]);
    if (sc.gamma != 0.0) {
      ny += sc.gamma * dd;
      nk += sc.gamma * kd;
    }
    if (sc.blend) {
      ny = sc.w * ny + (1.0 - sc.w) * ay[r];
      nk = sc.w * nk + (1.0 - sc.w) * ak[r];
    }
    stream_out(b.y_out + o, ny);
    stream_out(b.k_out + o, nk);
    if (sc.open_epoch) {
      stream_out(b.anchor_y_out + o, ny);
      stream_out(b.anchor_kx_out + o, nk);
    }
    if constexpr (kDiag) {
      stream_out(b.dy + o, dd);
      stream_out(b.k_dx + o, kd);
    }
  }
}

__device__ __forceinline__ bool lane_on(const LaneParams* params, unsigned long long mask,
                                        int lane) {
  return params != nullptr ? params[lane].active != 0 : ((mask >> lane) & 1ull) != 0;
}

template <int kWaves>
__global__ void __launch_bounds__(64 * kWaves)
    bspmm_kernel(DevCsr a, const double* __restrict__ v1, double* __restrict__ out1,
                 const double* __restrict__ v2, double* __restrict__ out2,
                 const LaneParams* __restrict__ params, unsigned long long mask) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const Int r = lanes::wave_row_in<kWaves>();
  if (r >= a.rows) return;  // uniform across the wave
  // The row sums are lane-serial (no cross-lane operation), so lanes beyond
  // the span sit out and the rest store only if live.
  if (params != nullptr && lane >= span_of(params)) return;
  const double s1 = row_sum(a, r, v1, lane);
  const double s2 = v2 != nullptr ? row_sum(a, r, v2, lane) : 0.0;
  if (!lane_on(params, mask, lane)) return;
  out1[at(r, lane)] = s1;
  if (v2 != nullptr) out2[at(r, lane)] = s2;
}

// Check reductions, per lane, in a fixed tree whose shape depends on the row
// counts only: each wave folds a contiguous run of rows in index order, the
// 16 waves of a block fold in wave order into per-block partials
// [block][value][lane], and bfold_kernel folds every 16th block per wave, then
// the waves in order. The per-row terms are those of backend.cpp:143-202
// (convergence terms of the unscaled problem), backend.cpp:123-126 (the
// movement) and pdhg.cpp:191-221 (the Farkas test).
constexpr int kSideVals = 7;
// Row side: sum dy^2, primal_sq, primal_inf (max), q'y, Farkas row flag (max),
// max |dy|, q~'dy.
constexpr unsigned kRowMax = (1u << 2) | (1u << 4) | (1u << 5);
constexpr int kRowSlot[kSideVals] = {kSumDy2, kPrimalSq, kPrimalInf, kQDotY,
                                     kFarkasRowBad, kFarkasScale, kFarkasRhs};
// Column side: sum dx^2, dual_sq, dual_inf (max), bound term, c'x, Farkas
// column flag (max), support.
constexpr unsigned kColMax = (1u << 2) | (1u << 5);
constexpr int kColSlot[kSideVals] = {kSumDx2, kDualSq, kDualInf, kBoundTerm,
                                     kCDotX, kFarkasColBad, kFarkasSupport};

// The per-row terms of the check: movement, the primal residual of the
// unscaled problem, q'y and the row half of the Farkas test.
__device__ __forceinline__ void row_terms(const BatchData& d, Int i, int lane,
                                          const double* __restrict__ y,
                                          const double* __restrict__ k_x,
                                          const double* __restrict__ dy, int farkas,
                                          double (&v)[kSideVals]) {
  const size_t p = at(i, lane);
  const double e = dy[p];
  v[0] += e * e;
  const double scale = d.row_scale[i];
  const double row = scale != 0.0 ? k_x[p] / scale : k_x[p];
  const double slack = row - d.orig_q[p];
  const double viol = (i < d.num_equalities) ? slack : fmin(slack, 0.0);
  v[1] += viol * viol;
  v[2] = fmax(v[2], fabs(viol));
  v[3] += d.orig_q[p] * (scale * y[p]);
  if (farkas) {
    if (i >= d.num_equalities && e < 0.0) v[4] = 1.0;
    v[5] = fmax(v[5], fabs(e));
    v[6] += d.q[p] * e;
  }
}

// The per-column terms: movement, the dual residual, the bound term, c'x and
// the column half of the Farkas test.
__device__ __forceinline__ void col_terms(const BatchData& d, Int j, int lane,
                                          const double* __restrict__ x,
                                          const double* __restrict__ dx,
                                          const double* __restrict__ kt_y,
                                          const double* __restrict__ lambda, int farkas,
                                          double (&v)[kSideVals]) {
  const size_t p = at(j, lane);
  const double e = dx[p];
  v[0] += e * e;
  const double scale = d.col_scale[j];
  const double product = scale != 0.0 ? kt_y[p] / scale : kt_y[p];
  const double c = d.orig_c[p];
  const double lam = c - product;
  const double lo = d.orig_lower[p], hi = d.orig_upper[p];
  const bool has_lo = lo > -__builtin_inf(), has_hi = hi < __builtin_inf();
  double left;
  if (has_lo && has_hi) {
    left = 0.0;
  } else if (has_lo) {
    left = fmin(lam, 0.0);
  } else if (has_hi) {
    left = fmax(lam, 0.0);
  } else {
    left = lam;
  }
  v[1] += left * left;
  v[2] = fmax(v[2], fabs(left));
  const double absorbed = lam - left;
  if (absorbed > 0.0) {
    v[3] += absorbed * lo;
  } else if (absorbed < 0.0) {
    v[3] += absorbed * hi;
  }
  v[4] += c * (scale * x[p]);
  if (farkas) {
    const double l = lambda[p];
    if (l > 0.0) {
      if (d.upper[p] >= __builtin_inf()) {
        v[5] = 1.0;
      } else {
        v[6] += l * d.upper[p];
      }
    } else if (l < 0.0) {
      if (d.lower[p] <= -__builtin_inf()) {
        v[5] = 1.0;
      } else {
        v[6] += l * d.lower[p];
      }
    }
  }
}

// One launch for both sides of the check: blocks [0, row_blocks) take rows of
// K, the rest take columns. Each of a block's 16 waves folds a contiguous run
// of `per_wave` rows in index order; the block folds its waves in wave order
// into partials[block][value][lane].
constexpr int kCheckWaves = 16;

__global__ void __launch_bounds__(1024)
    bcheck_kernel(BatchData d, const double* __restrict__ x, const double* __restrict__ y,
                  const double* __restrict__ k_x, const double* __restrict__ dx,
                  const double* __restrict__ dy, const double* __restrict__ kt_y,
                  const double* __restrict__ lambda, int farkas, Int row_blocks, Int row_per_wave,
                  Int col_per_wave, const LaneParams* __restrict__ params,
                  double* __restrict__ row_partials, double* __restrict__ col_partials) {
  __shared__ double lds[kCheckWaves][kSideVals][kLanes];
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const bool rows = static_cast<Int>(blockIdx.x) < row_blocks;
  const Int block = rows ? static_cast<Int>(blockIdx.x) : static_cast<Int>(blockIdx.x) - row_blocks;
  const Int per = rows ? row_per_wave : col_per_wave;
  const Int total = rows ? d.m : d.n;
  const Int r0 = (block * kCheckWaves + wave) * per;
  const Int r1 = min(r0 + per, total);
  double v[kSideVals] = {0, 0, 0, 0, 0, 0, 0};
  if (params[lane].active != 0) {
    if (rows) {
      for (Int i = r0; i < r1; ++i) row_terms(d, i, lane, y, k_x, dy, farkas, v);
    } else {
      for (Int j = r0; j < r1; ++j) col_terms(d, j, lane, x, dx, kt_y, lambda, farkas, v);
    }
  }
#pragma unroll
  for (int k = 0; k < kSideVals; ++k) lds[wave][k][lane] = v[k];
  __syncthreads();
  if (wave < kSideVals) {
    const int k = wave;  // wave k folds value k over the 16 waves
    const bool mx = (((rows ? kRowMax : kColMax) >> k) & 1u) != 0;
    double a = lds[0][k][lane];
    for (int w = 1; w < kCheckWaves; ++w) a = mx ? fmax(a, lds[w][k][lane]) : a + lds[w][k][lane];
    double* part = rows ? row_partials : col_partials;
    part[(static_cast<size_t>(block) * kSideVals + k) * kLanes + lane] = a;
  }
}

// One block per (side, value): wave w folds blocks w, w + 16, ... in order,
// then wave 0 folds the 16 waves in order into out[lane * kLaneCheckStride +
// slot].
__global__ void __launch_bounds__(1024)
    bfold_kernel(const double* __restrict__ row_partials, Int row_blocks,
                 const double* __restrict__ col_partials, Int col_blocks,
                 const LaneParams* __restrict__ params, double* __restrict__ out) {
  __shared__ double lds[kCheckWaves][kLanes];
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const int side_value = static_cast<int>(blockIdx.x);  // 0..2*kSideVals-1
  const bool rows = side_value < kSideVals;
  const int v = rows ? side_value : side_value - kSideVals;
  const double* part = rows ? row_partials : col_partials;
  const Int blocks = rows ? row_blocks : col_blocks;
  const bool mx = (((rows ? kRowMax : kColMax) >> v) & 1u) != 0;
  double a = 0.0;
  bool first = true;
#pragma unroll 8
  for (Int b = wave; b < blocks; b += kCheckWaves) {
    const double t = part[(static_cast<size_t>(b) * kSideVals + v) * kLanes + lane];
    a = first ? t : (mx ? fmax(a, t) : a + t);
    first = false;
  }
  lds[wave][lane] = a;
  __syncthreads();
  if (wave == 0 && params[lane].active != 0) {
    double r = lds[0][lane];
    for (int w = 1; w < kCheckWaves; ++w) r = mx ? fmax(r, lds[w][lane]) : r + lds[w][lane];
    out[static_cast<size_t>(lane) * kLaneCheckStride + (rows ? kRowSlot[v] : kColSlot[v])] = r;
  }
}

// dst[row][s] = src[row][perm[s]] for the first `count` slots; the other
// slots of dst are left as they are.
__global__ void __launch_bounds__(256)
    bpermute_kernel(Int rows, const double* __restrict__ src, double* __restrict__ dst,
                    SlotMap map) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t r = i / kLanes;
  const int s = static_cast<int>(i % kLanes);
  if (r >= static_cast<size_t>(rows) || s >= map.count) return;
  dst[r * kLanes + static_cast<size_t>(s)] = src[r * kLanes + static_cast<size_t>(map.from[s])];
}

// Distances moved since a restart point, per lane, in the same fixed tree as
// the check: per-wave runs of rows, waves in order, then blocks in order.
// The restart points advance in the same pass when `advance` is set.
__global__ void __launch_bounds__(256)
    brestart_part_kernel(Int rows, Int rows_per_wave, const double* __restrict__ v, double* r,
                         double* pr, int use_previous, int advance, unsigned long long mask,
                         double* __restrict__ partials) {
  __shared__ double lds[4][kLanes];
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const Int r0 = (static_cast<Int>(blockIdx.x) * 4 + wave) * rows_per_wave;
  const Int r1 = min(r0 + rows_per_wave, rows);
  double acc = 0.0;
  if ((mask >> lane) & 1ull) {
    for (Int i = r0; i < r1; ++i) {
      const size_t p = at(i, lane);
      const double cur = v[p];
      const double e = cur - (use_previous ? pr[p] : r[p]);
      acc += e * e;
      if (advance) {
        if (pr != nullptr) pr[p] = r[p];
        r[p] = cur;
      }
    }
  }
  lds[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) {
    double a = lds[0][lane];
    for (int w = 1; w < 4; ++w) a += lds[w][lane];
    partials[static_cast<size_t>(blockIdx.x) * kLanes + lane] = a;
  }
}

__global__ void __launch_bounds__(128)
    brestart_fold_kernel(const double* __restrict__ px, Int bx, const double* __restrict__ py,
                         Int by, unsigned long long mask, double* __restrict__ out) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int side = static_cast<int>(threadIdx.x >> 6);
  if (!((mask >> lane) & 1ull)) return;
  const double* part = side == 0 ? px : py;
  const Int blocks = side == 0 ? bx : by;
  double a = 0.0;
  for (Int b = 0; b < blocks; ++b) a += part[static_cast<size_t>(b) * kLanes + lane];
  out[2 * lane + side] = a;
}

__global__ void bextract_kernel(Int rows, const double* __restrict__ v, unsigned long long mask,
                                double* __restrict__ out) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  if (!((mask >> lane) & 1ull)) return;
  const int slot = __builtin_popcountll(mask & ((1ull << lane) - 1ull));
  const Int stride = static_cast<Int>(gridDim.x * (blockDim.x >> 6));
  for (Int r = static_cast<Int>(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); r < rows;
       r += stride)
    out[static_cast<size_t>(slot) * rows + r] = v[at(r, lane)];
}

// The lane table, word by word (a period's first launch); each thread's
// loads from host memory are all in flight before its stores.
constexpr int kTableWords = static_cast<int>(sizeof(LaneTable) / sizeof(Int));
constexpr int kTableRounds = (kTableWords + 255) / 256;
__global__ void __launch_bounds__(256)
    bparams_kernel(const Int* __restrict__ staged, Int* __restrict__ table) {
  Int w[kTableRounds];
#pragma unroll
  for (int r = 0; r < kTableRounds; ++r) {
    const int i = static_cast<int>(threadIdx.x) + 256 * r;
    w[r] = i < kTableWords ? staged[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < kTableRounds; ++r) {
    const int i = static_cast<int>(threadIdx.x) + 256 * r;
    if (i < kTableWords) table[i] = w[r];
  }
}

// The check values out to the host, then the period's number into *flag
// (a period's last launch): every thread's stores are made visible system
// wide before the workgroup barrier, and the flag is stored after it with
// release semantics at system scope.
constexpr int kPublishRounds = kLanes * kLaneCheckStride / 256;
static_assert(kLanes * kLaneCheckStride % 256 == 0, "the check values split evenly over 256 threads");
__global__ void __launch_bounds__(256)
    bpublish_kernel(const double* __restrict__ check, const LaneTable* __restrict__ table,
                    double* __restrict__ host_check, Int* flag) {
  double v[kPublishRounds];
#pragma unroll
  for (int r = 0; r < kPublishRounds; ++r) v[r] = check[threadIdx.x + 256 * r];
#pragma unroll
  for (int r = 0; r < kPublishRounds; ++r) host_check[threadIdx.x + 256 * r] = v[r];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(flag, table->seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("launch ") + what + ": " + hipGetErrorString(e));
}

// `per_wave` consecutive rows per wave, `waves` waves per workgroup.
template <class K, class... Args>
void launch_rows_per(K kernel, int waves, int per_wave, Int rows, hipStream_t s,
                     const LaunchEvents* ev, const char* what, Args... args) {
  if (rows <= 0) return;
  const Int units = (rows + per_wave - 1) / per_wave;
  const dim3 grid(static_cast<unsigned>((units + waves - 1) / waves));
  const dim3 block(static_cast<unsigned>(64 * waves));
  validate_launch(what, grid, block, ev);
  if (ev != nullptr && ev->start != nullptr) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev->start, ev->stop, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
  }
  check_launch(what);
}

// One wave per row, `waves` rows per workgroup.
template <class K, class... Args>
void launch_rows(K kernel, int waves, Int rows, hipStream_t s, const LaunchEvents* ev,
                 const char* what, Args... args) {
  launch_rows_per(kernel, waves, 1, rows, s, ev, what, args...);
}

// The kernel instance for a shape (row-waves per workgroup).
template <template <int> class Pick>
auto by_waves(int waves) {
  switch (waves) {
    case 1: return Pick<1>::get();
    case 2: return Pick<2>::get();
    case 8: return Pick<8>::get();
    default: return Pick<4>::get();
  }
}
// The one-row step kernel for (waves, row layout): slot width 0 is the CSR.
template <bool kDiag, int kW>
auto primal_by_waves(int waves) {
  switch (waves) {
    case 1: return &bprimal_kernel<kDiag, 1, kW>;
    case 2: return &bprimal_kernel<kDiag, 2, kW>;
    case 8: return &bprimal_kernel<kDiag, 8, kW>;
    default: return &bprimal_kernel<kDiag, 4, kW>;
  }
}
template <bool kDiag>
auto primal_one_row(int waves, int width) {
  switch (width) {
    case 8: return primal_by_waves<kDiag, 8>(waves);
    case 16: return primal_by_waves<kDiag, 16>(waves);
    case 32: return primal_by_waves<kDiag, 32>(waves);
    default: return primal_by_waves<kDiag, 0>(waves);
  }
}
template <bool kDiag, int kW>
auto dual_by_waves(int waves) {
  switch (waves) {
    case 1: return &bdual_kernel<kDiag, 1, kW>;
    case 2: return &bdual_kernel<kDiag, 2, kW>;
    case 8: return &bdual_kernel<kDiag, 8, kW>;
    default: return &bdual_kernel<kDiag, 4, kW>;
  }
}
template <bool kDiag>
auto dual_one_row(int waves, int width) {
  switch (width) {
    case 8: return dual_by_waves<kDiag, 8>(waves);
    case 16: return dual_by_waves<kDiag, 16>(waves);
    case 32: return dual_by_waves<kDiag, 32>(waves);
    default: return dual_by_waves<kDiag, 0>(waves);
  }
}
// The three leading row-structure arguments of a one-row step kernel.
struct RowStructure {
  const Int* s0;
  const Int* s1;
  const double* s2;
  int width;
};
RowStructure row_structure(const DevCsr& a) {
  if (a.slot_width > 0) return {a.slot_row, a.slot_index, a.slot_value, a.slot_width};
  return {a.start, a.index, a.value, 0};
}
template <int W>
struct SpmmAt {
  static auto get() { return &bspmm_kernel<W>; }
};

}  // namespace

namespace {
// The multi-row instance for (waves, rows), rows 2 or 4.
template <bool kDiag, int R>
auto primal_rows_by_waves(int waves) {
  switch (waves) {
    case 1: return &bprimal_rows_kernel<kDiag, 1, R>;
    case 2: return &bprimal_rows_kernel<kDiag, 2, R>;
    case 8: return &bprimal_rows_kernel<kDiag, 8, R>;
    default: return &bprimal_rows_kernel<kDiag, 4, R>;
  }
}
template <bool kDiag, int R>
auto dual_rows_by_waves(int waves) {
  switch (waves) {
    case 1: return &bdual_rows_kernel<kDiag, 1, R>;
    case 2: return &bdual_rows_kernel<kDiag, 2, R>;
    case 8: return &bdual_rows_kernel<kDiag, 8, R>;
    default: return &bdual_rows_kernel<kDiag, 4, R>;
  }
}

// Rows per wave of the binned steps' short rows for a span (0: the span is
// too wide).
int packed_rows(Int span) { return span <= 16 ? 4 : span <= 32 ? 2 : 0; }

// How a narrow span runs over `a`: with the row bins, the short rows two or
// four to a wave from their fixed-width copy and the long ones a wave each
// (packing K of the refinery stacks, whose rows run to 30 entries, whole was
// measured slower: a packed wave waits for its longest row); without them,
// one row per wave.
enum class Narrow { kNone, kBinned };
Narrow narrow_mode(const DevCsr& a, int g) {
  if (g == 0) return Narrow::kNone;
  if (a.bin_short[g == 4 ? 1 : 0] >= 0) return Narrow::kBinned;
  return Narrow::kNone;
}

template <bool kDiag, int kG>
auto primal_binned_by_waves(int waves) {
  switch (waves) {
    case 1: return &bprimal_binned_kernel<kDiag, 1, kG>;
    case 2: return &bprimal_binned_kernel<kDiag, 2, kG>;
    case 8: return &bprimal_binned_kernel<kDiag, 8, kG>;
    default: return &bprimal_binned_kernel<kDiag, 4, kG>;
  }
}
template <bool kDiag, int kG>
auto dual_binned_by_waves(int waves) {
  switch (waves) {
    case 1: return &bdual_binned_kernel<kDiag, 1, kG>;
    case 2: return &bdual_binned_kernel<kDiag, 2, kG>;
    case 8: return &bdual_binned_kernel<kDiag, 8, kG>;
    default: return &bdual_binned_kernel<kDiag, 4, kG>;
  }
}

template <bool kDiag>
void primal_step_shape(hipStream_t s, const DevCsr& kt, const double* y, const BatchData& d,
                       const LaneParams* params, int step, Int span, const PrimalBufs& b,
                       int waves, int rows, bool pack, const LaunchEvents* ev, int narrow_waves) {
  const char* what = "batch_primal_step";
  const int g = pack && rows == 1 && kt.slot_width == 0 ? packed_rows(span) : 0;
  const Narrow mode = narrow_mode(kt, g);
  if (mode != Narrow::kNone && narrow_waves != 0) waves = narrow_waves;
  if (mode == Narrow::kBinned) {
    const int p = g == 4 ? 1 : 0;
    const Int units = kt.bin_short[p] + kt.bin_long;
    const int blocks = static_cast<int>((units + waves - 1) / waves);
    const BinArgs a = bin_args(kt, p);
    launch_rows_per(g == 4 ? primal_binned_by_waves<kDiag, 4>(waves)
                           : primal_binned_by_waves<kDiag, 2>(waves),
                    waves, 1, units, s, ev, what, a.plan, a.index8, a.value8, a.index32, a.value32,
                    kt.bin_short[p], units, blocks, step, span, kt.cols, kt.index, kt.value, y,
                    params, d, b);
  } else if (rows == 4) {
    launch_rows_per(primal_rows_by_waves<kDiag, 4>(waves), waves, 4, kt.rows, s, ev, what, kt, y,
                    d, b, params, step, span);
  } else if (rows == 2) {
    launch_rows_per(primal_rows_by_waves<kDiag, 2>(waves), waves, 2, kt.rows, s, ev, what, kt, y,
                    d, b, params, step, span);
  } else {
    const int blocks = static_cast<int>((kt.rows + waves - 1) / waves);
    const RowStructure t = row_structure(kt);
    launch_rows(primal_one_row<kDiag>(waves, t.width), waves, kt.rows, s, ev, what, t.s0, t.s1,
                t.s2, y, kt.rows, span, blocks, step, kt.cols, params, d, b, kt.index, kt.value);
  }
}

template <bool kDiag>
void dual_step_shape(hipStream_t s, const DevCsr& k, const double* x_bar, const BatchData& d,
                     const LaneParams* params, int step, Int span, const DualBufs& b, int waves,
                     int rows, bool pack, const LaunchEvents* ev, int narrow_waves) {
  const char* what = "batch_dual_step";
  const int g = pack && rows == 1 && k.slot_width == 0 ? packed_rows(span) : 0;
  const Narrow mode = narrow_mode(k, g);
  if (mode != Narrow::kNone && narrow_waves != 0) waves = narrow_waves;
  if (mode == Narrow::kBinned) {
    const int p = g == 4 ? 1 : 0;
    const Int units = k.bin_short[p] + k.bin_long;
    const int blocks = static_cast<int>((units + waves - 1) / waves);
    const BinArgs a = bin_args(k, p);
    launch_rows_per(g == 4 ? dual_binned_by_waves<kDiag, 4>(waves)
                           : dual_binned_by_waves<kDiag, 2>(waves),
                    waves, 1, units, s, ev, what, a.plan, a.index8, a.value8, a.index32, a.value32,
                    k.bin_short[p], units, blocks, step, span, k.cols, k.index, k.value, x_bar,
                    params, d, b);
  } else if (rows == 4) {
    launch_rows_per(dual_rows_by_waves<kDiag, 4>(waves), waves, 4, k.rows, s, ev, what, k, x_bar,
                    d, b, params, step, span);
  } else if (rows == 2) {
    launch_rows_per(dual_rows_by_waves<kDiag, 2>(waves), waves, 2, k.rows, s, ev, what, k, x_bar,
                    d, b, params, step, span);
  } else {
    const int blocks = static_cast<int>((k.rows + waves - 1) / waves);
    const RowStructure t = row_structure(k);
    launch_rows(dual_one_row<kDiag>(waves, t.width), waves, k.rows, s, ev, what, t.s0, t.s1, t.s2,
                x_bar, k.rows, span, blocks, step, k.cols, params, d, b, k.index, k.value);
  }
}
}  // namespace

void launch_batch_primal_step(hipStream_t s, const DevCsr& kt, const double* y,
                              const BatchData& d, const LaneParams* params, int step, Int span,
                              const PrimalBufs& b, bool diag, int waves, int rows, bool pack,
                              const LaunchEvents* ev, int narrow_waves) {
  if (diag) {
    primal_step_shape<true>(s, kt, y, d, params, step, span, b, waves, rows, pack, ev,
                            narrow_waves);
  } else {
    primal_step_shape<false>(s, kt, y, d, params, step, span, b, waves, rows, pack, ev,
                             narrow_waves);
  }
}

void launch_batch_dual_step(hipStream_t s, const DevCsr& k, const double* x_bar,
                            const BatchData& d, const LaneParams* params, int step, Int span,
                            const DualBufs& b, bool diag, int waves, int rows, bool pack,
                            const LaunchEvents* ev, int narrow_waves) {
  if (diag) {
    dual_step_shape<true>(s, k, x_bar, d, params, step, span, b, waves, rows, pack, ev,
                          narrow_waves);
  } else {
    dual_step_shape<false>(s, k, x_bar, d, params, step, span, b, waves, rows, pack, ev,
                           narrow_waves);
  }
}

void launch_batch_params(hipStream_t s, const LaneTable* staged, LaneParams* table) {
  static_assert(sizeof(LaneTable) % sizeof(Int) == 0, "the table copies by words");
  hipLaunchKernelGGL(bparams_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<const Int*>(staged),
                     reinterpret_cast<Int*>(table));
  check_launch("batch_params");
}

void launch_batch_publish(hipStream_t s, const double* check, const LaneParams* table,
                          double* host_check, Int* flag) {
  hipLaunchKernelGGL(bpublish_kernel, dim3(1), dim3(256), 0, s, check,
                     reinterpret_cast<const LaneTable*>(table), host_check, flag);
  check_launch("batch_publish");
}

void launch_batch_spmm(hipStream_t s, const DevCsr& a, const double* v1, double* out1,
                       const double* v2, double* out2, const LaneParams* params,
                       unsigned long long mask, int waves) {
  launch_rows(by_waves<SpmmAt>(waves), waves, a.rows, s, nullptr, "batch_spmm", a, v1, out1, v2,
              out2, params, mask);
}

namespace {
// Rows per wave so a side needs at most kCheckBlocks blocks of four waves
// (the restart distances).
Int check_rows_per_wave(Int rows) {
  const Int waves = kCheckBlocks * 4;
  return rows <= waves ? 1 : (rows + waves - 1) / waves;
}
Int check_blocks(Int rows) {
  const Int per_block = 4 * check_rows_per_wave(rows);
  return rows <= 0 ? 0 : (rows + per_block - 1) / per_block;
}
// The check's shape: one row per wave while a side fits kCheckSideBlocks
// blocks of 16 waves, more rows per wave beyond that.
Int side_rows_per_wave(Int rows) {
  const Int waves = static_cast<Int>(kCheckSideBlocks) * kCheckWaves;
  return rows <= waves ? 1 : (rows + waves - 1) / waves;
}
Int side_blocks(Int rows) {
  const Int per_block = kCheckWaves * side_rows_per_wave(rows);
  return rows <= 0 ? 0 : (rows + per_block - 1) / per_block;
}
}  // namespace

size_t batch_check_scratch_doubles() {
  return static_cast<size_t>(2) * kCheckSideBlocks * kSideVals * kLanes;
}

void launch_batch_check(hipStream_t s, const BatchData& d, const double* x, const double* y,
                        const double* k_x, const double* dx, const double* dy,
                        const double* kt_y, const double* lambda, bool farkas,
                        const LaneParams* params, double* scratch, double* out) {
  double* row_part = scratch;
  double* col_part = scratch + static_cast<size_t>(kCheckSideBlocks) * kSideVals * kLanes;
  const Int rb = side_blocks(d.m), cb = side_blocks(d.n);
  if (rb + cb > 0) {
    hipLaunchKernelGGL(bcheck_kernel, dim3(static_cast<unsigned>(rb + cb)), dim3(1024), 0, s, d, x,
                       y, k_x, dx, dy, kt_y, lambda, farkas ? 1 : 0, rb, side_rows_per_wave(d.m),
                       side_rows_per_wave(d.n), params, row_part, col_part);
    check_launch("batch_check");
  }
  hipLaunchKernelGGL(bfold_kernel, dim3(2 * kSideVals), dim3(1024), 0, s, row_part, rb, col_part,
                     cb, params, out);
  check_launch("batch_check_fold");
}

void launch_batch_restart(hipStream_t s, Int n, Int m, const double* x, const double* y,
                          double* rx, double* ry, double* prx, double* pry, bool use_previous,
                          bool advance, unsigned long long mask, double* scratch, double* out) {
  double* px = scratch;
  double* py = scratch + static_cast<size_t>(kCheckBlocks) * kLanes;
  const Int bx = check_blocks(n), by = check_blocks(m);
  if (bx > 0) {
    hipLaunchKernelGGL(brestart_part_kernel, dim3(static_cast<unsigned>(bx)), dim3(256), 0, s, n,
                       check_rows_per_wave(n), x, rx, prx, use_previous ? 1 : 0, advance ? 1 : 0,
                       mask, px);
    check_launch("batch_restart_x");
  }
  if (by > 0) {
    hipLaunchKernelGGL(brestart_part_kernel, dim3(static_cast<unsigned>(by)), dim3(256), 0, s, m,
                       check_rows_per_wave(m), y, ry, pry, use_previous ? 1 : 0, advance ? 1 : 0,
                       mask, py);
    check_launch("batch_restart_y");
  }
  hipLaunchKernelGGL(brestart_fold_kernel, dim3(1), dim3(128), 0, s, px, bx, py, by, mask, out);
  check_launch("batch_restart_fold");
}

void launch_batch_permute(hipStream_t s, Int rows, double* v, double* tmp, const SlotMap& map) {
  if (rows <= 0 || map.count <= 0) return;
  const size_t n = static_cast<size_t>(rows) * kLanes;
  const hipError_t e = hipMemcpyAsync(tmp, v, n * sizeof(double), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) throw std::runtime_error(std::string("batch permute copy: ") + hipGetErrorString(e));
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(bpermute_kernel, dim3(blocks), dim3(256), 0, s, rows, tmp, v, map);
  check_launch("batch_permute");
}

void launch_batch_extract(hipStream_t s, Int rows, const double* v, unsigned long long mask,
                          double* out) {
  if (rows <= 0 || mask == 0) return;
  const unsigned blocks = static_cast<unsigned>(std::min<Int>((rows + 3) / 4, 1024));
  hipLaunchKernelGGL(bextract_kernel, dim3(blocks), dim3(256), 0, s, rows, v, mask, out);
  check_launch("batch_extract");
}

}  // namespace anuk
