// Kernels of the batched solver: up to 64 LPs that share one constraint
// matrix K, one wave64 lane per LP.
//
// Every per-LP vector is stored row-interleaved, v[row * 64 + lane], so the
// wave that owns a row of K touches one contiguous 512-byte line per vector
// and the matrix itself (value, index, row start) is read once for all the
// LPs with wave-uniform scalar loads. Each lane sums its own row serially in
// storage order and runs the reference's elementwise steps in the
// reference's order, so every product and update of lane i has exactly the
// bits the reference CPU backend computes for LP i. The reductions (check
// terms, restart distances) fold each lane's rows in a fixed tree, which
// agrees with the reference's serial sums to rounding and is the same on
// every run.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace anuk {

inline constexpr int kLanes = 64;

// Per-LP step scalars; the device holds kLanes of them.
struct LaneParams {
  double tau;
  double sigma;
  double gamma;
  Int halpern;
  Int reset_at_start;
  Int k_at_start;
  Int active;  // 0: retired or padding, the lane does nothing
};

// The lane table the kernels read: one LaneParams per slot, then the slot
// span (one past the highest live slot). Loads are issued by the lanes below
// the span, whose bound is wave-uniform and arrives with the row's first
// scalar loads; stores are left to the live lanes.
struct LaneTable {
  LaneParams lane[64];
  Int span;
  Int seq;  // the period's number, which launch_batch_publish hands back
};

// Problem data, per LP row-interleaved except the shared scaling diagonals.
struct BatchData {
  Int n = 0, m = 0, num_equalities = 0;
  const double* c = nullptr;      // [n][64] iteration (scaled) data
  const double* lower = nullptr;
  const double* upper = nullptr;
  const double* q = nullptr;      // [m][64]
  const double* orig_c = nullptr; // [n][64] convergence (unscaled) data
  const double* orig_lower = nullptr;
  const double* orig_upper = nullptr;
  const double* orig_q = nullptr; // [m][64]
  const double* row_scale = nullptr;  // [m] shared
  const double* col_scale = nullptr;  // [n] shared
};

// Slots of the per-lane check vector; the same quantities as CheckSlot.
inline constexpr int kLaneCheckStride = 16;

// `span`: the lanes below it load and compute (the live ones among them
// store); a kernel argument, so the bound costs no dependent load.
// `waves`: waves per workgroup, 1, 2, 4 or 8; `rows`: consecutive rows per
// wave, 1, 2 or 4. With one row per wave the kernels read the matrix's
// fixed-width row copy when it carries one (DevCsr::slot_width > 0), else
// its CSR; from the CSR, with `pack`, a span of 32 slots or fewer and the
// matrix's row bins (DevCsr::bin_*), one wave takes 64 / 32 = 2 or
// 64 / 16 = 4 rows of at most kBinShortMax entries, a lane per (row, slot),
// and a longer row a wave of its own; those narrow launches run
// `narrow_waves` waves per workgroup (0: `waves`).
// Every shape computes the same bits.
void launch_batch_primal_step(hipStream_t s, const DevCsr& kt, const double* y,
                              const BatchData& d, const LaneParams* params, int step, Int span,
                              const PrimalBufs& b, bool diag, int waves = 4, int rows = 1,
                              bool pack = false, const LaunchEvents* ev = nullptr,
                              int narrow_waves = 0);
void launch_batch_dual_step(hipStream_t s, const DevCsr& k, const double* x_bar,
                            const BatchData& d, const LaneParams* params, int step, Int span,
                            const DualBufs& b, bool diag, int waves = 4, int rows = 1,
                            bool pack = false, const LaunchEvents* ev = nullptr,
                            int narrow_waves = 0);
// A period's host traffic without a host wait on the device side: its first
// launch copies the lane table the host staged in page-locked (coherent)
// memory into the device table the step kernels read, and its last copies
// the per-lane check values out to page-locked memory and then sets *flag to
// the table's period number, after the values and with system scope, so the
// host can poll the flag instead of waiting on the stream.
void launch_batch_params(hipStream_t s, const LaneTable* staged, LaneParams* table);
void launch_batch_publish(hipStream_t s, const double* check, const LaneParams* table,
                          double* host_check, Int* flag);
// out1 = A v1 (and out2 = A v2 when v2 is non-null) for the active lanes:
// those whose LaneParams say so when `params` is given (read at run time, so
// a captured graph follows retirements), else those set in `mask`. Other
// lanes are left alone.
void launch_batch_spmm(hipStream_t s, const DevCsr& a, const double* v1, double* out1,
                       const double* v2, double* out2, const LaneParams* params,
                       unsigned long long mask = ~0ull, int waves = 4);
// Per-lane check reductions into out[lane * kLaneCheckStride + slot] (slots
// as CheckSlot) for the active lanes: movement sums, convergence terms of the
// unscaled problem, and both halves of the Farkas test. Deterministic: a
// fixed tree whose shape depends on the row counts only. `scratch` holds
// batch_check_scratch_doubles() doubles.
inline constexpr int kCheckBlocks = 64;       // restart distances: blocks per side
inline constexpr int kCheckSideBlocks = 1024;  // check terms: at most this many blocks per side
size_t batch_check_scratch_doubles();
void launch_batch_check(hipStream_t s, const BatchData& d, const double* x, const double* y,
                        const double* k_x, const double* dx, const double* dy,
                        const double* kt_y, const double* lambda, bool farkas,
                        const LaneParams* params, double* scratch, double* out);
// For masked lanes: out[2 lane] = sum_j (x - base_x)^2 and out[2 lane + 1] =
// sum_i (y - base_y)^2 in a fixed tree, where the base is the last restart
// point (rx, ry) or, with use_previous, the one before it (prx, pry). Then,
// when `advance`, prx <- rx and rx <- x, pry <- ry and ry <- y (prx/pry may
// be null when use_previous is false). `scratch` holds
// batch_check_scratch_doubles() doubles.
void launch_batch_restart(hipStream_t s, Int n, Int m, const double* x, const double* y,
                          double* rx, double* ry, double* prx, double* pry, bool use_previous,
                          bool advance, unsigned long long mask, double* scratch, double* out);
// Moves lanes between slots: slot s of every row receives what slot
// from[s] held, for s < count (passed by value to the kernel).
struct SlotMap {
  std::int8_t from[64];
  int count;
};
// Permutes the slots of v (rows x 64, row-interleaved) by `map`, staging
// through tmp (at least rows x 64 doubles).
void launch_batch_permute(hipStream_t s, Int rows, double* v, double* tmp, const SlotMap& map);
// Copies the columns of the masked lanes, in lane order, into dense
// [count][rows] arrays.
void launch_batch_extract(hipStream_t s, Int rows, const double* v, unsigned long long mask,
                          double* out);

}  // namespace anuk
