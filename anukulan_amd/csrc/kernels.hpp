// Device kernels of the MI355X backend and the host launchers for them.
//
// Everything is fp64. Elementwise arithmetic is written in the reference's
// operation order and compiled without FMA contraction, so a fused step gives
// the same bits as the CPU backend (backend.cpp:60-111). Row sums of the
// sparse products are taken serially in storage order for every row that fits
// one LDS tile, which makes them the same bits as SparseMatrix::multiply
// (sparse.cpp:72-80) too. Reductions are deterministic (fixed grid, fixed tree)
// but not serial, so they agree with the CPU's serial sums to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "launch_check.hpp"

namespace anuk {

using Int = std::int32_t;

inline constexpr int kThreads = 256;      // one workgroup = 4 wave64s
inline constexpr int kRowLanes = 8;      // lanes per row in the scaling sweeps
inline constexpr int kTileNnz = 1024;     // products staged in LDS per row block
inline constexpr int kTileRows = 256;     // most rows per row block
// Plans for the fast product order. Measured on the 1M-column random LP,
// larger tiles (4096 nonzeros, rows in rounds of kThreads) were slower than
// the storage order's tile: the products are bound by the L2 request rate of
// their random gathers, not by per-block latency. The machinery stays
// (kernels are templated on the tile; rounds prefetch their row bounds).
inline constexpr int kFastTileNnz = 1024;
inline constexpr int kFastTileRows = 256;
inline constexpr int kMaxRed = 8;         // most reduction values one kernel emits

// A store of a batched step's result with the nontemporal hint: measured
// 3-5% faster per batched step than plain stores. (In the single-LP
// kernels, whose outputs are megabytes that the next launch gathers from,
// the same hint measured 1.5% slower for the fused steps and 13% slower for
// the panel combine, so they store plainly.)
__device__ __forceinline__ void stream_out(double* p, double v) {
  __builtin_nontemporal_store(v, p);
}

// A CSR matrix on the device plus its row-block plan: block b covers rows
// [block_row[b], block_row[b+1]). A block either holds up to kThreads rows
// whose nonzeros together fit one LDS tile (summed serially per row), or a
// single longer row (summed by the whole workgroup).
struct DevCsr {
  Int rows = 0;
  Int cols = 0;
  Int nnz = 0;
  Int blocks = 0;
  const Int* start = nullptr;
  const Int* index = nullptr;
  const double* value = nullptr;
  const Int* block_row = nullptr;
  Int max_rows = kTileRows;  // most rows in one block of this plan
  Int longest = 0;           // nonzeros in the longest row
  // Optional compaction: launched block i runs planned block block_map[i]
  // (`blocks` then counts the launched ones). Null: every planned block.
  const Int* block_map = nullptr;
  // Read the index and value streams with non-temporal loads, which leaves
  // the L2 to the gathered vector's lines (the same bits either way).
  bool stream_loads = false;
  // Optional fixed-width copy for kernels that give each row one wave
  // (slot_width 0: not built). Row r's first slot_width entries sit at
  // [r * slot_width, (r + 1) * slot_width) of slot_index / slot_value,
  // padded with value 0 at the row's last index (index 0 for an empty row),
  // and slot_row[2r], slot_row[2r + 1] = start[r], the row's length; a
  // longer row continues in the CSR itself.
  Int slot_width = 0;
  const Int* slot_row = nullptr;
  const Int* slot_index = nullptr;
  const double* slot_value = nullptr;
  // Optional row bins for the packed batched steps (bin_short[p] < 0: not
  // built). bin_plan[p] (p = 0 for two rows a wave, 1 for four) holds
  // {row, start, length, 0} entries: bin_short[p] waves of 2 or 4 rows of
  // at most kBinShortMax entries each (row -1 pads the last), then the
  // bin_long longer rows, one wave each.
  const Int* bin_plan[2] = {nullptr, nullptr};
  Int bin_short[2] = {-1, -1};
  Int bin_long = 0;
  // Fixed-width copies of the binned rows, addressed by plan entry so their
  // loads need not wait for the plan's: bin_index8/bin_value8 hold short
  // entry e's entries at [e * 8, e * 8 + 8) (the short entries are the same
  // rows in the same order in both plans), bin_index32/bin_value32 long row
  // q's first kBinLongWidth at [q * 32, q * 32 + 32). Padding repeats the
  // row's last index with value 0 (index 0 for a padding entry).
  const Int* bin_index8 = nullptr;
  const double* bin_value8 = nullptr;
  const Int* bin_index32 = nullptr;
  const double* bin_value32 = nullptr;
};

// The longest row the packed steps take several to a wave: one round of
// their sum.
inline constexpr Int kBinShortMax = 8;
// Entries of a long row in its fixed-width copy (DevCsr::bin_index32).
inline constexpr Int kBinLongWidth = 32;

// Widths of the fixed-width copy: the smallest that holds every row, or the
// widest with longer rows continuing in the CSR.
inline constexpr int kSlotWidths[] = {8, 16, 32};
// Matrices with this many columns or more get no fixed-width copy (its
// kernels address the gathered vector, 512 bytes per row, with 31 bits).
inline constexpr Int kSlotMaxCols = 1 << 22;

// Per-iteration scalars the fused kernels read from device memory, so a
// captured period of iterations can be replayed with new values.
struct IterParams {
  double tau;
  double sigma;
  double gamma;              // reflection
  Int halpern;               // 1 when the Halpern scheme is on
  Int reset_at_start;        // the period's first step begins a new epoch
  Int k_at_start;            // Halpern counter before the period's first step
  Int num_equalities;        // end of the equality rows (a global row index)
  Int active = 1;            // 0: the LP does nothing this period
};

// Which of the per-iteration extras a fused step produces.
enum StepFlags : unsigned {
  kStepDiag = 1u,   // write dx / dy / k_dx and the check reductions
};

// Problem data the fused kernels read. Scaled data drive the iteration and
// the Farkas test; the original data and the two diagonals give the
// convergence measures of the unscaled problem (backend.hpp:150-165).
struct DevProblem {
  Int n = 0, m = 0, num_equalities = 0;
  const double* c = nullptr;       // scaled
  const double* lower = nullptr;
  const double* upper = nullptr;
  const double* q = nullptr;
  const double* orig_c = nullptr;  // unscaled
  const double* orig_lower = nullptr;
  const double* orig_upper = nullptr;
  const double* orig_q = nullptr;
  const double* row_scale = nullptr;
  const double* col_scale = nullptr;
};

// The iterate buffers of one side of a step.
struct PrimalBufs {
  const double* x;      // current
  const double* anchor; // Halpern anchor
  double* x_out;
  double* x_bar;
  double* dx;           // written on diagnostic steps
  double* anchor_out;   // written when the step opens an epoch
};

struct DualBufs {
  const double* y;
  const double* k_x;
  const double* anchor_y;
  const double* anchor_kx;
  double* y_out;
  double* k_out;
  double* dy;
  double* k_dx;
  double* anchor_y_out;
  double* anchor_kx_out;
};

// Slots of the check result vector (all fp64, reduced on the device).
enum CheckSlot : int {
  kSumDx2 = 0,
  kSumDy2,
  kPrimalSq,
  kPrimalInf,
  kQDotY,
  kFarkasRowBad,    // any inequality row with dy < 0 (max of 0/1)
  kFarkasScale,     // max |dy|
  kFarkasRhs,       // q' dy
  kDualSq,
  kDualInf,
  kBoundTerm,
  kCDotX,
  kFarkasColBad,    // lambda pointing at an infinite bound (max of 0/1)
  kFarkasSupport,
  kCheckSlots
};

// An event pair the dispatch itself records at kernel start and end
// (hipExtLaunchKernel), the same interval a kernel-trace profiler reports.
// Null events mean a plain launch.
struct LaunchEvents {
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
};

// Throws std::invalid_argument unless launch_shape_error accepts the launch
// (grid and workgroup shape, and the event pair of `ev`, which may be null).
inline void validate_launch(const char* what, dim3 grid, dim3 block, const LaunchEvents* ev) {
  LaunchShape s;
  s.gx = grid.x;
  s.gy = grid.y;
  s.gz = grid.z;
  s.bx = block.x;
  s.by = block.y;
  s.bz = block.z;
  const char* why = launch_shape_error(s, ev != nullptr ? ev->start : nullptr,
                                       ev != nullptr ? ev->stop : nullptr);
  if (why != nullptr) throw std::invalid_argument(std::string("launch ") + what + ": " + why);
}

// Rows longer than this are summed by a whole wave in the fast product
// order; at most this many are summed serially by one thread.
inline constexpr Int kSerialRowMax = 32;
// Column panels of a K whose gathered vector outgrows one XCD's L2: panel p
// holds the columns of the p-th eighth of K's nonzeros, so the XCD that runs
// it keeps its slice of x in L2. The panels are stacked as one CSR of
// kPanels * rows rows; row p * rows + r holds row r's entries in panel p.
inline constexpr int kPanels = 8;
// A gathered vector larger than this (half an XCD's 4 MB L2) gets panels.
inline constexpr double kPanelVectorBytes = 2.0 * 1024 * 1024;

// ---- launchers (stream-ordered, asynchronous) --------------------------

// Builds the row-block plan of a CSR on the host.
void plan_blocks(Int rows, const Int* start, Int tile_nnz, int max_rows,
                 std::int32_t** out_rows, Int* out_blocks);

void launch_spmv(hipStream_t s, const DevCsr& a, const double* x, double* y,
                 const LaunchEvents* ev = nullptr);

void launch_fill(hipStream_t s, double* p, Int n, double v);
// dlo/dhi = lo/hi with every finite bound replaced by 0.
void launch_dual_bounds(hipStream_t s, Int n, const double* lo, const double* hi, double* dlo,
                        double* dhi);
// x = the cold start in the box [lo, hi]: 0, or the finite bound it violates.
void launch_cold_start(hipStream_t s, Int n, const double* lo, const double* hi, double* x);
void launch_accumulate(hipStream_t s, Int n, double w, const double* v, double* sum);
void launch_scale_into(hipStream_t s, Int n, double w, const double* sum, double* out);
// out = sum / sqrt(*sq) with *sq read on the device.
void launch_scale_by_root(hipStream_t s, Int n, const double* sq, const double* sum, double* out);
void launch_primal_update(hipStream_t s, Int n, double tau, double gamma, double h,
                          const double* x, const double* c, const double* kt_y,
                          const double* lower, const double* upper, const double* anchor,
                          double* x_out, double* dx, double* x_bar);
void launch_dual_update(hipStream_t s, Int m, Int neq, double sigma, double gamma, double h,
                        const double* y, const double* q, const double* k_x_bar,
                        const double* k_x, const double* anchor_y, const double* anchor_kx,
                        double* y_out, double* dy, double* k_out, double* k_dx);

// Reductions. Each writes its results to `out` (device) after a fixed-order
// finish; `partials` must hold reduce_grid(len) * kMaxRed doubles.
int reduce_grid(Int len);
void launch_dot(hipStream_t s, const double* a, const double* b, Int n, double* partials,
                double* out);
void launch_sumsq2(hipStream_t s, const double* a, Int n, const double* b, Int m,
                   double* partials_a, double* partials_b, double* out2);
void launch_absmax(hipStream_t s, const double* a, Int n, double* partials, double* out);
void launch_dot_and_sumsq(hipStream_t s, const double* dy, const double* k_dx, Int m,
                          const double* dx, Int n, double* partials_a, double* partials_b,
                          double* out3);
void launch_convergence_terms(hipStream_t s, Int n, Int m, const DevProblem& p,
                              const double* x, const double* k_x, const double* y,
                              const double* kt_y, double* partials_a, double* partials_b,
                              double* out7);

// The library-wide device scratch array of the allocation-free entry
// points: one copy per device, held for the life of the process, shared by
// every call on that device (calls that overlap on one device race).
// Layout for the convergence terms: the row-side partials at [0,
// 1024 * kMaxRed), the column-side partials after them, the 7 results at
// kNoAllocResultOffset. Every value a call reads is written earlier in the
// same call, so nothing needs clearing between calls.
inline constexpr size_t kNoAllocPartials = static_cast<size_t>(1024) * kMaxRed;
inline constexpr size_t kNoAllocResultOffset = 2 * kNoAllocPartials;
inline constexpr size_t kNoAllocScratchDoubles = kNoAllocResultOffset + 64;
// Looks up the array's device address on `device` (which must be current)
// once and keeps it; throws if the lookup fails. Called from context set-up.
void noalloc_scratch_init(int device);
// The address noalloc_scratch_init found for `device`; throws if it has not
// run on that device.
double* noalloc_scratch(int device);

// The fused PDHG step. `flags` selects the diagnostic variant; `step` is the
// position of this step in its period (for the Halpern weight).
// `fast` selects the fast product order (see anuk_pdhg_options).
void launch_primal_step(hipStream_t s, const DevCsr& kt, const double* y,
                        const DevProblem& p, const IterParams* params, int step,
                        const PrimalBufs& b, unsigned flags, double* partials,
                        const LaunchEvents* ev = nullptr, bool fast = false);
void launch_dual_step(hipStream_t s, const DevCsr& k, const double* x_bar,
                      const DevProblem& p, const IterParams* params, int step,
                      const DualBufs& b, unsigned flags, double* partials,
                      const LaunchEvents* ev = nullptr, bool fast = false);
// The dual step over K's column panels: the panels' partial row sums into
// `panel_sums` (kPanels * m), then one thread per row adds them in panel
// order and runs the dual epilogue. Its check partials come from
// panel_combine_blocks(m) blocks.
void launch_dual_step_panels(hipStream_t s, const DevCsr& panels, double* panel_sums,
                             const double* x_bar, const DevProblem& p, const IterParams* params,
                             int step, const DualBufs& b, unsigned flags, double* partials,
                             const LaunchEvents* ev = nullptr,
                             const LaunchEvents* ev_combine = nullptr);
Int panel_combine_blocks(Int rows);
// Check products: dual-side convergence terms from K^T y, and the Farkas
// support from K^T dy. The Farkas product reads the row slots of `check`
// (folded first by launch_check_fold_rows) and does nothing when the row half
// of the test has already ruled a certificate out.
void launch_conv_product(hipStream_t s, const DevCsr& kt, const double* y,
                         const double* x, const DevProblem& p, double* partials,
                         bool fast = false);
void launch_farkas_product(hipStream_t s, const DevCsr& kt, const double* dy,
                           const DevProblem& p, const double* check, double* partials,
                           bool fast = false);
// Folds the partials of a check into check[kCheckSlots]: first the two step
// kernels' (slots kSumDx2..kFarkasRhs), then the two products' (kDualSq..).
// pd may be null when infeasibility detection is off.
void launch_check_fold_rows(hipStream_t s, Int primal_blocks, const double* pa,
                            Int dual_blocks, const double* pb, double* check);
void launch_check_fold_cols(hipStream_t s, Int conv_blocks, const double* pc,
                            Int farkas_blocks, const double* pd, double* check);

// ---- scaling sweeps (scaling.cpp:24-61) ----------------------------------
// out[r] = infinity norm (max |a|) or one norm (serial sum of |a| in storage
// order) of row r, one thread per row.
void launch_row_norms(hipStream_t s, const Int* start, Int rows, const double* value, bool one_norm,
                      double* out);
// d = 1 / sqrt(norm) where norm > 0, else 1.
void launch_inv_sqrt(hipStream_t s, const double* norm, Int n, double* d);
// a[i] += add[l] where bounds[l] <= i < bounds[l + 1]: shifts each LP's
// slice of a concatenated index or row-start array by that LP's offset.
void launch_add_segment_offsets(hipStream_t s, Int* a, Int count, const Int* bounds, const Int* add,
                                Int segments);
// Every entry a[r][k] becomes (a * first) * second, rounding after each
// product; `first_by_row` says whether `first` is indexed by the row (and
// `second` by the column) or the other way round.
void launch_scale_values(hipStream_t s, const Int* start, const Int* index, Int rows,
                         double* value, const double* first, const double* second,
                         bool first_by_row);
// q *= d1, row_scale *= d1 (m entries); c *= d2, lower /= d2, upper /= d2,
// col_scale *= d2 (n entries).
void launch_scale_vectors(hipStream_t s, Int m, const double* d1, double* q, double* row_scale,
                          Int n, const double* d2, double* c, double* lower, double* upper,
                          double* col_scale);

// ---- block-diagonal stacks of LPs (each with its own K) -----------------
//
// The LPs' matrices are concatenated block-diagonally and their vectors end
// to end. Row blocks are planned per LP, so no block straddles two LPs, and
// `block_lp[b]` names the LP of block b; each LP has its own IterParams
// (num_equalities holds the global end row of its equalities; active = 0
// makes its blocks return at once). Vector segments are cut into chunks of
// at most kTileNnz elements within one LP for the per-LP reductions.
struct ChunkPlan {
  Int chunks = 0;
  const Int* start = nullptr;  // chunks + 1 element offsets
  const Int* lp = nullptr;     // LP of each chunk
};

void launch_multi_primal_step(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                              const double* y, const DevProblem& p, const IterParams* params,
                              int step, const PrimalBufs& b, bool diag, double* partials,
                              const LaunchEvents* ev = nullptr);
void launch_multi_dual_step(hipStream_t s, const DevCsr& k, const Int* block_lp,
                            const double* x_bar, const DevProblem& p, const IterParams* params,
                            int step, const DualBufs& b, bool diag, double* partials,
                            const LaunchEvents* ev = nullptr);
void launch_multi_conv_product(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                               const double* y, const double* x, const DevProblem& p,
                               const IterParams* params, double* partials);
// `check` holds kCheckSlots doubles per LP, row slots already folded.
void launch_multi_farkas_product(hipStream_t s, const DevCsr& kt, const Int* block_lp,
                                 const double* dy, const DevProblem& p, const IterParams* params,
                                 const double* check, double* partials);
// y = A x on the blocks of the LPs whose params are active (others untouched).
void launch_multi_product(hipStream_t s, const DevCsr& a, const Int* block_lp, const double* x,
                          double* y, const IterParams* params);
// Folds nvals per-block partials of each active LP over its blocks
// [lp_blocks[l], lp_blocks[l+1]) in block order into
// out[l * kCheckSlots + slot0 + v].
void launch_multi_fold(hipStream_t s, const double* partials, int nvals, int slot0,
                       unsigned max_mask, const Int* lp_blocks, Int lps, const IterParams* params,
                       double* out);
// Per chunk of each active LP: sum (v - from)^2 (from null: sum v^2) into
// partials[chunk]; with `advance`, prev <- base (when prev is set) and
// base <- v afterwards. Then folded per LP in chunk order into
// out[l * stride + slot].
void launch_multi_sumsq(hipStream_t s, const ChunkPlan& plan, const double* v, const double* from,
                        double* base, double* prev, bool advance, const IterParams* params,
                        double* partials, const Int* lp_chunks, Int lps, double* out, int stride,
                        int slot);
// out = v / divisor[lp] element-wise on the chunks of active LPs.
void launch_multi_divide(hipStream_t s, const ChunkPlan& plan, const double* v,
                         const double* divisor, const IterParams* params, double* out);

}  // namespace anuk
