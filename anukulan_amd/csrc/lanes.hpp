// Device helpers shared by the kernels that run one LP per wave64 lane over
// a shared CSR matrix: vectors are row-interleaved, v[row * 64 + lane], and a
// wave owns one row, so the matrix loads are wave-uniform (scalar) and each
// vector access is one contiguous 512-byte line.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace anuk {
namespace lanes {

inline constexpr int kLaneCount = 64;
inline constexpr int kWavesPerBlock = 4;  // 256 threads, one row per wave

// Workgroups are dispatched round-robin over the 8 XCDs; this gives each XCD
// one contiguous range of logical blocks, so neighbouring rows share its L2.
__device__ __forceinline__ int xcd_remap(int i, int g) {
  const int x = i & 7, local = i >> 3, per = g >> 3, rem = g & 7;
  return x * per + (x < rem ? x : rem) + local;
}

// The row this wave owns in a launch of kWaves row-waves per workgroup,
// made wave-uniform so the matrix loads are scalar.
template <int kWaves>
__device__ __forceinline__ Int wave_row_in() {
  const int blk = xcd_remap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  return __builtin_amdgcn_readfirstlane(blk * kWaves + static_cast<int>(threadIdx.x >> 6));
}

__device__ __forceinline__ Int wave_row() { return wave_row_in<kWavesPerBlock>(); }

// The same with the grid size passed in (a preloaded kernel argument)
// instead of read from the dispatch's argument block.
template <int kWaves>
__device__ __forceinline__ Int wave_row_of(int blocks) {
  const int blk = xcd_remap(static_cast<int>(blockIdx.x), blocks);
  return __builtin_amdgcn_readfirstlane(blk * kWaves + static_cast<int>(threadIdx.x >> 6));
}

__device__ __forceinline__ size_t at(Int row, int lane) {
  return static_cast<size_t>(row) * kLaneCount + static_cast<size_t>(lane);
}

// Lane-serial row sum of A[row, :] v[:, lane] in storage order. The gathers
// of up to eight nonzeros are issued together so a long row is not a chain of
// memory latencies; the adds still happen one by one in storage order.
template <int W>
__device__ __forceinline__ void chunk_sum(const DevCsr& a, Int k, const double* __restrict__ v,
                                          int lane, double& s) {
  double w[W], x[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    w[u] = a.value[k + u];
    x[u] = v[at(a.index[k + u], lane)];
  }
#pragma unroll
  for (int u = 0; u < W; ++u) s += w[u] * x[u];
}

// A CSR's three arrays, passed as leading scalar kernel arguments so the
// hardware can preload them into SGPRs at wave launch: the first load of a
// row sum (its start) then waits on no argument load.
struct CsrArrays {
  const Int* __restrict__ start;
  const Int* __restrict__ index;
  const double* __restrict__ value;
};

template <int W>
__device__ __forceinline__ void chunk_sum(const CsrArrays& a, Int k, const double* __restrict__ v,
                                          int lane, double& s) {
  double w[W], x[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    w[u] = a.value[k + u];
    x[u] = v[at(a.index[k + u], lane)];
  }
#pragma unroll
  for (int u = 0; u < W; ++u) s += w[u] * x[u];
}

// Adds the terms [lo, hi) of a's storage to s in storage order (row_sum
// below, over a range).
__device__ __forceinline__ double range_sum(const CsrArrays& a, Int lo, Int hi,
                                            const double* __restrict__ v, int lane, double s) {
  Int k = lo;
  if (k + 8 <= hi) {
    // Rounds of eight, software-pipelined: the next round's indices and
    // values are requested while this round's gathers are in flight, so a
    // long row waits for one memory latency per round rather than two
    // (1% on the 24-period stack, whose rows of K run to 30 entries).
    Int ix[8];
    double w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      ix[u] = a.index[k + u];
      w[u] = a.value[k + u];
    }
    for (;;) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = v[at(ix[u], lane)];
      const Int next = k + 8;
      const bool more = next + 8 <= hi;
      Int nix[8];
      double nw[8];
      if (more) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          nix[u] = a.index[next + u];
          nw[u] = a.value[next + u];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += w[u] * x[u];
      k = next;
      if (!more) break;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        ix[u] = nix[u];
        w[u] = nw[u];
      }
    }
  }
  if (k + 4 <= hi) {
    chunk_sum<4>(a, k, v, lane, s);
    k += 4;
  }
  const Int rest = hi - k;
  if (rest > 0) {
    double w[3], x[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      if (u < rest) {
        w[u] = a.value[k + u];
        x[u] = v[at(a.index[k + u], lane)];
      }
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (u < rest) s += w[u] * x[u];
  }
  return s;
}

// row_sum below over the arrays alone (the same terms in the same order).
__device__ __forceinline__ double row_sum(const CsrArrays& a, Int row,
                                          const double* __restrict__ v, int lane) {
  return range_sum(a, a.start[row], a.start[row + 1], v, lane, 0.0);
}

// The fixed-width copy of a CSR (DevCsr::slot_*), as leading scalar kernel
// arguments.
struct SlotArrays {
  const Int* __restrict__ row;      // [rows][2]: start, length
  const Int* __restrict__ index;    // [rows][W]
  const double* __restrict__ value; // [rows][W]
};

// A raw buffer resource over `bytes` bytes at p (stride 0, the data-format
// word of the gfx9 family): a load through it at an offset of `bytes` or
// more returns 0 and reads no memory, which turns a gather on or off with
// no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void* p, Int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0), bytes,
                                           0x00020000);
}
// A row sum over the fixed-width copy, in two halves so a kernel can put its
// own loads between them. slot_gather reads the row's start and length and
// its W indices: the addresses depend on the row alone, so the scalar loads
// go out with the wave's first loads, and the row's gathers are then all in
// flight at once. The gathers go through a buffer resource over v
// (`v_rows` rows of 64 lanes); a slot past the row's length gets an offset
// outside it, reads nothing and yields 0. slot_finish adds the row's terms
// in storage order: a slot past the length adds +0.0, which leaves the sum
// unchanged (a sum that starts at +0.0 is never -0.0), so the result has
// the bits of the serial loop. A row longer than W continues from the CSR
// `a`. v_rows * 512 must stay below 2^31 (DevCsr::slot_width is 0 for
// wider matrices).
template <int W>
struct SlotRow {
  Int lo, len;
  double x[W];
};

template <int W>
__device__ __forceinline__ void slot_gather(const SlotArrays& t, Int row,
                                            const double* __restrict__ v, Int v_rows, int lane,
                                            SlotRow<W>& r) {
  r.lo = t.row[2 * row];
  r.len = t.row[2 * row + 1];
  const size_t base = static_cast<size_t>(row) * W;
  const __amdgpu_buffer_rsrc_t rs = buffer_over(v, v_rows * (kLaneCount * 8));
  const unsigned mine = static_cast<unsigned>(lane) * 8u;
  Int ix[W];
#pragma unroll
  for (int u = 0; u < W; ++u) ix[u] = t.index[base + u];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    // Arithmetic rather than a select, so every index load stays
    // unconditional: the top bit is set exactly when u >= len.
    const unsigned past = static_cast<unsigned>(r.len - 1 - u) >> 31;
    const unsigned at_row =
        static_cast<unsigned>(ix[u]) * static_cast<unsigned>(kLaneCount * 8) + (past << 31);
    r.x[u] = __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(rs, static_cast<int>(at_row + mine), 0, 0));
  }
}

// A row of len entries from lo, 1 <= len <= kMax, summed with one load of
// its structure and one round of gathers. Lane u of the wave loads entry
// min(u, len - 1) with a vector load; readlane then hands each entry's index
// and value to every lane, and all kMax gathers go out together. So every
// lane of the wave must execute this call (the broadcasts read the lanes
// that loaded the entries); a lane with `gather` false, and every position
// past the row's length, gathers through the buffer resource at an offset
// outside it, which reads no memory and yields 0. Each lane adds its terms
// in storage order, a position past the length adding +0.0 (which leaves a
// sum that started at +0.0 unchanged), so the sum has the bits of the serial
// loop. v_rows * 512 must stay below 2^31. round_sum_loaded is the part
// after the load, for a caller that loads the entries itself: lane u holds
// entry u (for u < len) in my_ix and my_wd.
template <int kMax>
__device__ __forceinline__ double round_sum_loaded(Int my_ix, double my_wd, Int len,
                                                  const double* __restrict__ v, Int v_rows,
                                                  int lane, bool gather) {
  static_assert(kMax >= 1 && kMax <= kLaneCount, "one entry per lane");
  const unsigned long long my_w = __builtin_bit_cast(unsigned long long, my_wd);
  const int w_lo = static_cast<int>(static_cast<unsigned>(my_w));
  const int w_hi = static_cast<int>(static_cast<unsigned>(my_w >> 32));
  const __amdgpu_buffer_rsrc_t rs = buffer_over(v, v_rows * (kLaneCount * 8));
  const unsigned off = static_cast<unsigned>(lane) * 8u;
  const unsigned dead = gather ? 0u : 1u;
  double x[kMax];
#pragma unroll
  for (int t = 0; t < kMax; ++t) {
    const unsigned ix = static_cast<unsigned>(__builtin_amdgcn_readlane(my_ix, t));
    const unsigned past = static_cast<unsigned>(len - 1 - t) >> 31;
    const unsigned at_row =
        (ix * static_cast<unsigned>(kLaneCount * 8) + off) | ((past | dead) << 31);
    x[t] = __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(rs, static_cast<int>(at_row), 0, 0));
  }
  // Keep every gather ahead of the first add: left to itself, the scheduler
  // interleaves the adds and issues the last gathers only after the first
  // ones have returned.
  __builtin_amdgcn_sched_barrier(0);
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < kMax; ++t) {
    const unsigned lo32 = static_cast<unsigned>(__builtin_amdgcn_readlane(w_lo, t));
    const unsigned hi32 = static_cast<unsigned>(__builtin_amdgcn_readlane(w_hi, t));
    const double w =
        __builtin_bit_cast(double, (static_cast<unsigned long long>(hi32) << 32) | lo32);
    const double p = w * x[t];
    s += t < len ? p : 0.0;
  }
  return s;
}

template <int kMax>
__device__ __forceinline__ double round_sum(const Int* __restrict__ index,
                                           const double* __restrict__ value, Int lo, Int len,
                                           const double* __restrict__ v, Int v_rows, int lane,
                                           bool gather) {
  const Int mine = lo + (lane < len ? lane : len - 1);
  return round_sum_loaded<kMax>(index[mine], value[mine], len, v, v_rows, lane, gather);
}

// The longest row round_row_sum takes in one round.
inline constexpr Int kRoundMax = 32;

// A whole row [lo, lo + len) by round_sum in the narrowest round that holds
// it, and in range_sum's pipelined rounds past kRoundMax (or for a gathered
// vector too large for 31-bit buffer offsets). Every lane of the wave must
// execute this; a lane with `live` false gathers nothing and gets 0. Same
// bits as row_sum.
__device__ __forceinline__ double round_row_sum(const Int* __restrict__ index,
                                               const double* __restrict__ value, Int lo, Int len,
                                               const double* __restrict__ v, Int v_rows, int lane,
                                               bool live) {
  if (len <= 0) return 0.0;
  if (len > kRoundMax || v_rows >= kSlotMaxCols) {
    if (!live) return 0.0;
    return range_sum(CsrArrays{nullptr, index, value}, lo, lo + len, v, lane, 0.0);
  }
  if (len <= 8) return round_sum<8>(index, value, lo, len, v, v_rows, lane, live);
  if (len <= 16) return round_sum<16>(index, value, lo, len, v, v_rows, lane, live);
  return round_sum<kRoundMax>(index, value, lo, len, v, v_rows, lane, live);
}

template <int W>
__device__ __forceinline__ double slot_finish(const SlotArrays& t, const CsrArrays& a, Int row,
                                              const SlotRow<W>& r, const double* __restrict__ v,
                                              int lane) {
  const size_t base = static_cast<size_t>(row) * W;
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < W; ++u) {
    const double p = t.value[base + u] * r.x[u];
    s += u < r.len ? p : 0.0;
  }
  if (r.len > W) s = range_sum(a, r.lo + W, r.lo + r.len, v, lane, s);
  return s;
}

__device__ __forceinline__ double row_sum(const DevCsr& a, Int row, const double* __restrict__ v,
                                          int lane) {
  const Int lo = a.start[row], hi = a.start[row + 1];
  double s = 0.0;
  Int k = lo;
  for (; k + 8 <= hi; k += 8) chunk_sum<8>(a, k, v, lane, s);
  if (k + 4 <= hi) {
    chunk_sum<4>(a, k, v, lane, s);
    k += 4;
  }
  // The last one to three terms: their gathers go out together too (a row
  // of K^T is often this short), and the adds stay in order. (Masking every
  // round to eight instead was slower on the dual side.)
  const Int rest = hi - k;
  if (rest > 0) {
    double w[3], x[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      if (u < rest) {
        w[u] = a.value[k + u];
        x[u] = v[at(a.index[k + u], lane)];
      }
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (u < rest) s += w[u] * x[u];
  }
  return s;
}

// Row sums of the kRows consecutive rows r0, r0 + 1, ... of A times
// v[:, lane] (rows at or past a.rows are empty). Consecutive rows are one
// contiguous run of the CSR, [start[r0], start[r0 + kRows]), so the wave
// takes it in rounds of kSeg nonzeros: a round issues all its index and
// value loads (wave-uniform) and all its gathers together, then adds each
// term to its row's sum in storage order. A term outside a row adds +0.0 to
// that row's sum, which leaves it unchanged (a sum that starts at +0.0
// never becomes -0.0), so every row sum has the bits of the serial
// storage-order loop.
template <int kRows, int kSeg>
__device__ __forceinline__ void segment_sums(const DevCsr& a, Int r0, const double* __restrict__ v,
                                             int lane, double (&s)[kRows]) {
  Int st[kRows + 1];
#pragma unroll
  for (int r = 0; r <= kRows; ++r) st[r] = a.start[min(r0 + r, a.rows)];
#pragma unroll
  for (int r = 0; r < kRows; ++r) s[r] = 0.0;
  for (Int p0 = st[0]; p0 < st[kRows]; p0 += kSeg) {
    double w[kSeg], x[kSeg];
#pragma unroll
    for (int u = 0; u < kSeg; ++u) {
      const Int p = p0 + u;
      if (p < st[kRows]) {
        w[u] = a.value[p];
        x[u] = v[at(a.index[p], lane)];
      } else {
        w[u] = 0.0;
        x[u] = 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
#pragma unroll
      for (int u = 0; u < kSeg; ++u) {
        const Int p = p0 + u;
        const double t = (p >= st[r] && p < st[r + 1]) ? w[u] * x[u] : 0.0;
        s[r] += t;
      }
    }
  }
}

// Blocks needed for one wave per row.
inline unsigned row_blocks(Int rows) {
  return static_cast<unsigned>((rows + kWavesPerBlock - 1) / kWavesPerBlock);
}

}  // namespace lanes
}  // namespace anuk
