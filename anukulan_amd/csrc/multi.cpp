#include "multi.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <stdexcept>
#include <thread>

#include "batch.hpp"
#include "control.hpp"
#include "kernels.hpp"

namespace anuk {

namespace {

using Clock = std::chrono::steady_clock;

// Backend scratch slots of the engines' state slabs, one per nesting depth.
constexpr int kMultiScratchSlot = 500;
constexpr size_t kParamDoubles = (sizeof(IterParams) + 7) / 8;

double seconds_since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

// ---- the stack on the host ------------------------------------------------

struct Stack {
  Int lps = 0;
  Int n = 0, m = 0;
  std::vector<Int> col_off, row_off;  // lps + 1
  std::vector<Int> nz_off;            // lps + 1: where each LP's entries start in K (and K^T)
  Int nnz = 0;
  // The LPs themselves: K and K^T are concatenated on the device only
  // (upload_stack), unscaled, and scaled there.
  std::vector<const StandardLp*> lp;
  std::vector<Int> eq_end;            // global end row of each LP's equalities
  std::vector<double> row_scale, col_scale;
  std::vector<double> c, q, lower, upper;  // scaled
  std::vector<double> oc, oq, olo, ohi;    // as given
  std::vector<Scaling> scaling;            // per LP
  Int cols(Int l) const { return col_off[uz(l) + 1] - col_off[uz(l)]; }
  Int rows(Int l) const { return row_off[uz(l) + 1] - row_off[uz(l)]; }
};

// Runs f(i) for i in [0, count) on up to `threads` host threads.
template <class F>
void parallel_for(size_t count, F f) {
  const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
  const size_t threads = std::min<size_t>(std::min<size_t>(hw, 16), count);
  if (threads <= 1) {
    for (size_t i = 0; i < count; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::exception_ptr> errors(threads);
  for (size_t t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      try {
        for (size_t i = t; i < count; i += threads) f(i);
      } catch (...) {
        errors[t] = std::current_exception();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (auto& e : errors)
    if (e) std::rethrow_exception(e);
}

// Lays the LPs end to end: offsets of each LP's columns, rows and entries,
// and its vectors copied into the stack's (each LP's slice by its own host
// thread). K and K^T are not concatenated here: each LP's arrays go straight
// to their place in the device stack (upload_stack).
Stack concat_stack(const std::vector<const StandardLp*>& lps) {
  Stack st;
  st.lps = static_cast<Int>(lps.size());
  st.lp = lps;
  st.scaling.resize(lps.size());
  st.col_off.assign(lps.size() + 1, 0);
  st.row_off.assign(lps.size() + 1, 0);
  st.nz_off.assign(lps.size() + 1, 0);
  long nz = 0;
  for (size_t i = 0; i < lps.size(); ++i) {
    const long cols = static_cast<long>(st.col_off[i]) + lps[i]->cols();
    const long rows = static_cast<long>(st.row_off[i]) + lps[i]->rows();
    nz += lps[i]->k.nnz();
    if (cols > (1L << 30) || rows > (1L << 30) || nz > (1L << 30))
      throw std::invalid_argument("multi: stack too large");
    st.col_off[i + 1] = static_cast<Int>(cols);
    st.row_off[i + 1] = static_cast<Int>(rows);
    st.nz_off[i + 1] = static_cast<Int>(nz);
  }
  st.n = st.col_off.back();
  st.m = st.row_off.back();
  st.nnz = st.nz_off.back();
  for (std::vector<double>* v : {&st.oc, &st.olo, &st.ohi}) v->resize(uz(st.n));
  st.oq.resize(uz(st.m));
  parallel_for(lps.size(), [&](size_t i) {
    std::copy(lps[i]->c.begin(), lps[i]->c.end(), st.oc.begin() + st.col_off[i]);
    std::copy(lps[i]->lower.begin(), lps[i]->lower.end(), st.olo.begin() + st.col_off[i]);
    std::copy(lps[i]->upper.begin(), lps[i]->upper.end(), st.ohi.begin() + st.col_off[i]);
    std::copy(lps[i]->q.begin(), lps[i]->q.end(), st.oq.begin() + st.row_off[i]);
  });
  for (size_t i = 0; i < lps.size(); ++i) st.eq_end.push_back(st.row_off[i] + lps[i]->num_equalities);
  return st;
}

// ---- the stack on the device ----------------------------------------------

struct DeviceStack {
  HipBackend& be;
  DevCsr k, kt;
  double* k_value = nullptr;   // k.value and kt.value, writable for the scaling
  double* kt_value = nullptr;
  const Int* k_block_lp = nullptr;
  const Int* kt_block_lp = nullptr;
  const Int* k_lp_blocks = nullptr;
  const Int* kt_lp_blocks = nullptr;
  std::vector<Int> k_lp_blocks_host, kt_lp_blocks_host;  // the same, on the host
  // Each planned block's place when the blocks are ordered by work
  // (nonzeros plus rows, heaviest first), and whether the launches bin the
  // blocks by it (launch_blocks) or take the plan's order; chosen by
  // tune_stack, with the index and value streams' load kind (DevCsr::
  // stream_loads). Every choice gives the same bits.
  std::vector<Int> k_rank, kt_rank;
  bool k_binned = false, kt_binned = false;
  ChunkPlan cols, rows;
  const Int* col_lp_chunks = nullptr;
  const Int* row_lp_chunks = nullptr;
  Int col_chunks = 0, row_chunks = 0;
  // The stack's vectors on the device, in the stack's layout (set by
  // scale_stack): scaled, as given, the scales, and the dual problem's
  // bounds. Engines copy their ranges from these on the device.
  const double *c = nullptr, *q = nullptr, *lower = nullptr, *upper = nullptr;
  const double *oc = nullptr, *oq = nullptr, *olo = nullptr, *ohi = nullptr;
  const double *row_scale = nullptr, *col_scale = nullptr;
  const double *dual_lower = nullptr, *dual_upper = nullptr;
  std::vector<void*> owned;

  explicit DeviceStack(HipBackend& b) : be(b) {}
  ~DeviceStack() {
    (void)hipStreamSynchronize(be.stream());
    for (void* p : owned) (void)hipFree(p);
  }
  DeviceStack(const DeviceStack&) = delete;
  DeviceStack& operator=(const DeviceStack&) = delete;

  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc(stack)");
    owned.push_back(p);
    return static_cast<T*>(p);
  }

  template <class T>
  const T* put(const std::vector<T>& h) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(h.size(), 1) * sizeof(T)), "hipMalloc(stack)");
    owned.push_back(p);
    if (!h.empty())
      hip_check(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, be.stream()),
                "upload stack");
    return static_cast<const T*>(p);
  }
};

// Row blocks planned per LP (so none straddles two), with each block's LP
// and each LP's block range.
struct Plan {
  std::vector<Int> block_row, block_lp, lp_blocks;
  std::vector<Int> rank;  // each block's place by work, heaviest first
};

Plan plan_stack(const Stack& st, bool transpose) {
  const std::vector<Int>& off = transpose ? st.col_off : st.row_off;
  Plan p;
  std::vector<long> work;
  p.block_row.push_back(0);
  p.lp_blocks.push_back(0);
  for (size_t l = 0; l + 1 < off.size(); ++l) {
    const Int rows = off[l + 1] - off[l];
    if (rows > 0) {
      const Csr& a = transpose ? st.lp[l]->kt : st.lp[l]->k;
      std::int32_t* br = nullptr;
      Int blocks = 0;
      plan_blocks(rows, a.start.data(), kTileNnz, kTileRows, &br, &blocks);
      for (Int b = 1; b <= blocks; ++b) {
        p.block_row.push_back(off[l] + br[b]);
        p.block_lp.push_back(static_cast<Int>(l));
        work.push_back(static_cast<long>(a.start[uz(br[b])] - a.start[uz(br[b - 1])]) +
                       (br[b] - br[b - 1]));
      }
      std::free(br);
    }
    p.lp_blocks.push_back(static_cast<Int>(p.block_lp.size()));
  }
  std::vector<Int> by_work(work.size());
  for (size_t b = 0; b < by_work.size(); ++b) by_work[b] = static_cast<Int>(b);
  std::stable_sort(by_work.begin(), by_work.end(),
                   [&](Int a, Int b) { return work[uz(a)] > work[uz(b)]; });
  p.rank.assign(work.size(), 0);
  for (size_t i = 0; i < by_work.size(); ++i) p.rank[uz(by_work[i])] = static_cast<Int>(i);
  return p;
}

// The blocks of one side to launch for the LPs marked in `on`, in the plan's
// order, or binned: the kernels give XCD x the x-th of eight contiguous
// ranges of the launch (xcd_block), so each XCD keeps the same LPs (and
// their gathered vectors in its own L2), but within its range the blocks
// run heaviest first, the long single-row blocks ahead of the tiles, so no
// long block starts last. (Dealing the heaviest blocks over all eight
// ranges instead was measured twice as slow on the 512-LP stack: an LP's
// vector then has to be fetched into every XCD's L2.)
std::vector<Int> launch_blocks(const std::vector<Int>& lp_blocks, const std::vector<char>& on,
                               bool binned, const std::vector<Int>& rank) {
  std::vector<Int> bl;
  for (size_t l = 0; l + 1 < lp_blocks.size(); ++l) {
    if (!on[l]) continue;
    for (Int b = lp_blocks[l]; b < lp_blocks[l + 1]; ++b) bl.push_back(b);
  }
  if (!binned) return bl;
  const size_t g = bl.size(), per = g / 8, rem = g % 8;
  size_t lo = 0;
  for (size_t x = 0; x < 8; ++x) {
    const size_t hi = lo + per + (x < rem ? 1 : 0);
    std::sort(bl.begin() + static_cast<long>(lo), bl.begin() + static_cast<long>(hi),
              [&](Int a, Int b) { return rank[uz(a)] < rank[uz(b)]; });
    lo = hi;
  }
  return bl;
}

struct Chunks {
  std::vector<Int> start, lp, lp_chunks;
};

Chunks chunk_stack(const std::vector<Int>& off) {
  Chunks c;
  c.start.push_back(0);
  c.lp_chunks.push_back(0);
  for (size_t l = 0; l + 1 < off.size(); ++l) {
    for (Int i = off[l]; i < off[l + 1]; i += kTileNnz) {
      c.start.push_back(std::min(off[l + 1], i + kTileNnz));
      c.lp.push_back(static_cast<Int>(l));
    }
    c.lp_chunks.push_back(static_cast<Int>(c.lp.size()));
  }
  return c;
}

void upload_stack(const Stack& st, DeviceStack* ds) {
  hipStream_t s = ds->be.stream();
  const Int* d_col_off = ds->put(st.col_off);
  const Int* d_row_off = ds->put(st.row_off);
  const Int* d_nz_off = ds->put(st.nz_off);
  // Each LP's CSR arrays are copied to their place in the stacked CSR as
  // they are (the LPs' arrays end to end, through page-locked chunks), then
  // shifted on the device: row starts by the LP's first entry, column
  // indices by its first column (row, for K^T).
  auto dev_csr = [&](bool transpose, const Plan& p) {
    DevCsr d;
    d.rows = transpose ? st.n : st.m;
    d.cols = transpose ? st.m : st.n;
    d.nnz = st.nnz;
    d.blocks = static_cast<Int>(p.block_lp.size());
    Int* start = ds->alloc<Int>(uz(d.rows) + 1);
    Int* index = ds->alloc<Int>(uz(d.nnz));
    double* value = ds->alloc<double>(uz(d.nnz));
    hip_check(hipMemsetAsync(start, 0, sizeof(Int), s), "stack row starts");
    std::vector<std::pair<const void*, size_t>> starts, indices, values;
    for (Int l = 0; l < st.lps; ++l) {
      const Csr& a = transpose ? st.lp[uz(l)]->kt : st.lp[uz(l)]->k;
      starts.push_back({a.start.data() + 1, uz(a.rows) * sizeof(Int)});
      indices.push_back({a.index.data(), uz(a.nnz()) * sizeof(Int)});
      values.push_back({a.value.data(), uz(a.nnz()) * sizeof(double)});
    }
    ds->be.upload_concat(start + 1, starts);
    ds->be.upload_concat(index, indices);
    ds->be.upload_concat(value, values);
    launch_add_segment_offsets(s, start + 1, d.rows, transpose ? d_col_off : d_row_off, d_nz_off, st.lps);
    launch_add_segment_offsets(s, index, d.nnz, d_nz_off, transpose ? d_row_off : d_col_off, st.lps);
    d.start = start;
    d.index = index;
    d.value = value;
    d.block_row = ds->put(p.block_row);
    return d;
  };
  const Plan pk = plan_stack(st, false);
  const Plan pkt = plan_stack(st, true);
  ds->k = dev_csr(false, pk);
  ds->kt = dev_csr(true, pkt);
  ds->k_value = const_cast<double*>(ds->k.value);
  ds->kt_value = const_cast<double*>(ds->kt.value);
  ds->k_block_lp = ds->put(pk.block_lp);
  ds->kt_block_lp = ds->put(pkt.block_lp);
  ds->k_lp_blocks = ds->put(pk.lp_blocks);
  ds->kt_lp_blocks = ds->put(pkt.lp_blocks);
  ds->k_lp_blocks_host = pk.lp_blocks;
  ds->kt_lp_blocks_host = pkt.lp_blocks;
  ds->k_rank = pk.rank;
  ds->kt_rank = pkt.rank;
  const Chunks cc = chunk_stack(st.col_off), rc = chunk_stack(st.row_off);
  ds->col_chunks = static_cast<Int>(cc.lp.size());
  ds->row_chunks = static_cast<Int>(rc.lp.size());
  ds->cols = ChunkPlan{ds->col_chunks, ds->put(cc.start), ds->put(cc.lp)};
  ds->rows = ChunkPlan{ds->row_chunks, ds->put(rc.start), ds->put(rc.lp)};
  ds->col_lp_chunks = ds->put(cc.lp_chunks);
  ds->row_lp_chunks = ds->put(rc.lp_chunks);
  ds->be.synchronize();
}

// scale_lp (scaling.cpp:32-100) of every LP at once, in place on the device
// stack. A row or column of the block-diagonal stack belongs to one LP, so
// the stack's sweeps are each LP's sweeps, with the same operations in the
// same order: every LP gets exactly its own scale_lp's bits. The scaled
// vectors and the diagonals come back to the host.
void scale_stack(HipBackend& be, const anuk_pdhg_options& opt, Stack* st, DeviceStack* ds) {
  const Int m = st->m, n = st->n;
  const bool sweeps = opt.ruiz_iterations > 0 || opt.pock_chambolle != 0;
  if (sweeps) {
    // Read back from the device below.
    for (std::vector<double>* v : {&st->c, &st->lower, &st->upper, &st->col_scale}) v->resize(uz(n));
    for (std::vector<double>* v : {&st->q, &st->row_scale}) v->resize(uz(m));
  } else {
    st->c = st->oc;
    st->q = st->oq;
    st->lower = st->olo;
    st->upper = st->ohi;
    st->row_scale.assign(uz(m), 1.0);
    st->col_scale.assign(uz(n), 1.0);
  }
  hipStream_t s = be.stream();
  int slot = kMultiScratchSlot + 100;
  auto dev = [&](Int len) { return be.scratch(slot++, uz(std::max<Int>(len, 1))); };
  // Whole-stack vectors cross PCIe through page-locked chunks.
  auto put = [&](const std::vector<double>& h) {
    double* d = dev(static_cast<Int>(h.size()));
    be.upload_concat(d, {{h.data(), h.size() * sizeof(double)}});
    return d;
  };
  auto clone = [&](const double* from, Int len) {
    double* d = dev(len);
    if (len > 0)
      hip_check(hipMemcpyAsync(d, from, uz(len) * sizeof(double), hipMemcpyDeviceToDevice, s),
                "stack vectors");
    return d;
  };
  // The vectors as given cross to the device once; the scaled ones start as
  // copies of them there, and the scales as ones.
  double* oq = put(st->oq);
  double* oc = put(st->oc);
  double* olo = put(st->olo);
  double* ohi = put(st->ohi);
  double* q = clone(oq, m);
  double* c = clone(oc, n);
  double* lo = clone(olo, n);
  double* hi = clone(ohi, n);
  double* rs = dev(m);
  double* cs = dev(n);
  launch_fill(s, rs, m, 1.0);
  launch_fill(s, cs, n, 1.0);
  if (sweeps) {
    int tmp = kMultiScratchSlot + 80;
    double* rn = be.scratch(tmp++, uz(m) + 1);
    double* cn = be.scratch(tmp++, uz(n) + 1);
    double* d1 = be.scratch(tmp++, uz(m) + 1);
    double* d2 = be.scratch(tmp++, uz(n) + 1);
    auto sweep = [&](bool one_norm) {
      launch_row_norms(s, ds->k.start, m, ds->k_value, one_norm, rn);
      launch_row_norms(s, ds->kt.start, n, ds->kt_value, one_norm, cn);
      launch_inv_sqrt(s, rn, m, d1);
      launch_inv_sqrt(s, cn, n, d2);
      launch_scale_values(s, ds->k.start, ds->k.index, m, ds->k_value, d1, d2, true);
      launch_scale_values(s, ds->kt.start, ds->kt.index, n, ds->kt_value, d1, d2, false);
      launch_scale_vectors(s, m, d1, q, rs, n, d2, c, lo, hi, cs);
    };
    for (int i = 0; i < opt.ruiz_iterations; ++i) sweep(false);
    if (opt.pock_chambolle) sweep(true);
    // The host keeps the scaled vectors for the per-LP norms, start points
    // and unscaling.
    auto get = [&](const double* d, std::vector<double>* h) {
      be.download_staged(d, h->data(), h->size() * sizeof(double));
    };
    get(q, &st->q);
    get(c, &st->c);
    get(lo, &st->lower);
    get(hi, &st->upper);
    get(rs, &st->row_scale);
    get(cs, &st->col_scale);
  }
  double* dlo = dev(n);
  double* dhi = dev(n);
  launch_dual_bounds(s, n, lo, hi, dlo, dhi);
  ds->c = c;
  ds->q = q;
  ds->lower = lo;
  ds->upper = hi;
  ds->oc = oc;
  ds->oq = oq;
  ds->olo = olo;
  ds->ohi = ohi;
  ds->row_scale = rs;
  ds->col_scale = cs;
  ds->dual_lower = dlo;
  ds->dual_upper = dhi;
  be.synchronize();
  parallel_for(uz(st->lps), [&](size_t l) {
    Scaling& sc = st->scaling[l];
    sc.row_scale.assign(st->row_scale.begin() + st->row_off[l], st->row_scale.begin() + st->row_off[l + 1]);
    sc.col_scale.assign(st->col_scale.begin() + st->col_off[l], st->col_scale.begin() + st->col_off[l + 1]);
  });
}

// Device-resident per-LP flags as IterParams (only `active` is read).
std::vector<IterParams> flags(const std::vector<char>& on) {
  std::vector<IterParams> p(on.size());
  for (size_t i = 0; i < on.size(); ++i) {
    p[i] = IterParams{};
    p[i].active = on[i] ? 1 : 0;
  }
  return p;
}

// The stack's launch tuner. Per side (K and K^T) it times the plain product
// over all LPs with the blocks launched in the plan's order or binned
// (launch_blocks), each with the index and value streams read plainly or
// non-temporally, and keeps, per stack pattern for the context's life, the
// fastest, if it beats the plan's order with plain loads by kStackMargin.
// Every choice gives the same bits: a block's row sums and epilogue do not
// depend on when it runs, and each LP's reductions fold its partials in
// plan order. A choice is 200 + 10 * binned + non-temporal.
namespace {
constexpr double kStackMargin = 0.03;
constexpr int kStackDefault = 200;
bool stack_binned(int choice) { return (choice / 10) % 10 != 0; }
bool stack_stream(int choice) { return choice % 10 != 0; }

std::uint64_t stack_key(const Stack& st) {
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&](std::uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  for (const std::vector<Int>* v : {&st.col_off, &st.row_off, &st.nz_off})
    for (Int x : *v) mix(static_cast<std::uint64_t>(static_cast<std::uint32_t>(x)));
  return h;
}
}  // namespace

void tune_stack(HipBackend& be, const Stack& st, DeviceStack* ds, const anuk_pdhg_options& opt) {
  auto apply = [&](bool k_side, int choice) {
    (k_side ? ds->k_binned : ds->kt_binned) = stack_binned(choice);
    (k_side ? ds->k : ds->kt).stream_loads = stack_stream(choice);
  };
  if (opt.multi_order != 0 || opt.multi_stream_loads != 0) {
    const int choice = kStackDefault +
                       10 * (opt.multi_order != 0 ? (opt.multi_order == 2 ? 1 : 0)
                                                  : (stack_binned(kStackDefault) ? 1 : 0)) +
                       (opt.multi_stream_loads != 0 ? (opt.multi_stream_loads == 2 ? 1 : 0)
                                                    : (stack_stream(kStackDefault) ? 1 : 0));
    apply(true, choice);
    apply(false, choice);
    return;
  }
  const std::uint64_t key = stack_key(st);
  const HipBackend::Tuning* tk = be.tuned("multi_k_product", key);
  const HipBackend::Tuning* tkt = be.tuned("multi_kt_product", key);
  if (tk != nullptr && tkt != nullptr) {
    apply(true, tk->chosen);
    apply(false, tkt->chosen);
    return;
  }
  if (opt.autotune == 0 || be.profiling() || ds->k.blocks <= 0 || ds->kt.blocks <= 0) return;
  hipStream_t s = be.stream();
  double* x = be.scratch(kMultiScratchSlot + 80, uz(st.n));
  double* y = be.scratch(kMultiScratchSlot + 81, uz(st.m));
  be.fill(x, st.n, 1.0);
  be.fill(y, st.m, 1.0);
  auto* params = reinterpret_cast<IterParams*>(
      be.scratch(kMultiScratchSlot + 82, uz(st.lps) * kParamDoubles + 1));
  const std::vector<char> all(uz(st.lps), 1);
  const std::vector<IterParams> on = flags(all);
  hip_check(hipMemcpyAsync(params, on.data(), on.size() * sizeof(IterParams), hipMemcpyHostToDevice, s),
            "tune flags");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hip_check(hipEventCreate(&e0), "hipEventCreate(tune stack)");
  hip_check(hipEventCreate(&e1), "hipEventCreate(tune stack)");
  constexpr int kReps = 5, kTrials = 3;
  auto side = [&](bool k_side) {
    const DevCsr& a = k_side ? ds->k : ds->kt;
    const Int* block_lp = k_side ? ds->k_block_lp : ds->kt_block_lp;
    const std::vector<Int>& lp_blocks = k_side ? ds->k_lp_blocks_host : ds->kt_lp_blocks_host;
    const std::vector<Int>& rank = k_side ? ds->k_rank : ds->kt_rank;
    HipBackend::Tuning t{k_side ? "multi_k_product" : "multi_kt_product", a.rows, a.nnz, {},
                         kStackDefault};
    // The four candidates, each trial timing all of them in turn (a
    // candidate's time is its best trial).
    std::vector<DevCsr> runs;
    std::vector<int> choices;
    for (int binned : {0, 1}) {
      const std::vector<Int> order = launch_blocks(lp_blocks, all, binned != 0, rank);
      Int* map = reinterpret_cast<Int*>(be.scratch(kMultiScratchSlot + 83 + binned, order.size() / 2 + 1));
      hip_check(hipMemcpyAsync(map, order.data(), order.size() * sizeof(Int), hipMemcpyHostToDevice, s),
                "tune block map");
      be.synchronize();  // the host order goes out of scope
      for (int stream : {0, 1}) {
        DevCsr run = a;
        run.block_map = map;
        run.blocks = static_cast<Int>(order.size());
        run.stream_loads = stream != 0;
        runs.push_back(run);
        choices.push_back(kStackDefault + 10 * binned + stream);
      }
    }
    const double* in = k_side ? x : y;
    double* out = k_side ? y : x;
    std::vector<double> best(runs.size(), 1e300);
    for (const DevCsr& run : runs) launch_multi_product(s, run, block_lp, in, out, params);  // warm
    for (int trial = 0; trial < kTrials; ++trial) {
      for (size_t c = 0; c < runs.size(); ++c) {
        hip_check(hipEventRecord(e0, s), "hipEventRecord(tune stack)");
        for (int r = 0; r < kReps; ++r) launch_multi_product(s, runs[c], block_lp, in, out, params);
        hip_check(hipEventRecord(e1, s), "hipEventRecord(tune stack)");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize(tune stack)");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime(tune stack)");
        best[c] = std::min(best[c], 1e3 * static_cast<double>(ms) / kReps);
      }
    }
    for (size_t c = 0; c < runs.size(); ++c) t.us.push_back({choices[c], best[c]});
    double base = 0.0, fastest = 1e300;
    int pick = kStackDefault;
    for (const auto& c : t.us) {
      if (c.first == kStackDefault) base = c.second;
      if (c.second < fastest) {
        fastest = c.second;
        pick = c.first;
      }
    }
    t.chosen = fastest < (1.0 - kStackMargin) * base ? pick : kStackDefault;
    apply(k_side, t.chosen);
    be.record_tuning(key, std::move(t));
  };
  side(true);
  side(false);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// ||K_l|| for every LP by the power iteration on K_l^T K_l (pdhg.cpp:223-263),
// all LPs in the same launches, each stopping on its own test.
std::vector<double> stack_norms(HipBackend& be, const Stack& st, const DeviceStack& ds,
                                int iterations, double tolerance) {
  const Int L = st.lps;
  std::vector<double> norm(uz(L), 0.0), sigma(uz(L), 0.0);
  std::vector<char> on(uz(L), 0);
  std::vector<double> v(uz(st.n), 0.0);
  for (Int l = 0; l < L; ++l) {
    const Int n = st.cols(l), m = st.rows(l);
    const Int nnz = st.nz_off[uz(l) + 1] - st.nz_off[uz(l)];
    if (n == 0 || m == 0 || nnz == 0) continue;
    std::vector<double> s(uz(n));
    for (Int j = 0; j < n; ++j) s[uz(j)] = 1.0 + 0.1 * static_cast<double>(j % 7);
    const double z = host_two_norm(s);
    for (Int j = 0; j < n; ++j) v[uz(st.col_off[uz(l)] + j)] = s[uz(j)] / z;
    on[uz(l)] = 1;
  }
  const size_t N = uz(st.n), M = uz(st.m);
  double* d_v = be.scratch(kMultiScratchSlot + 90, N);
  double* d_kv = be.scratch(kMultiScratchSlot + 91, M);
  double* d_u = be.scratch(kMultiScratchSlot + 92, N);
  double* d_part = be.scratch(kMultiScratchSlot + 93, uz(ds.col_chunks) + 1);
  double* d_out = be.scratch(kMultiScratchSlot + 94, uz(L) + 1);
  double* d_div = be.scratch(kMultiScratchSlot + 95, uz(L) + 1);
  auto* d_params = reinterpret_cast<IterParams*>(be.scratch(kMultiScratchSlot + 96, uz(L) * kParamDoubles + 1));
  be.upload_staged(v.data(), d_v, N);
  auto put_flags = [&] {
    const std::vector<IterParams> p = flags(on);
    hip_check(hipMemcpyAsync(d_params, p.data(), p.size() * sizeof(IterParams), hipMemcpyHostToDevice,
                             be.stream()),
              "norm flags");
  };
  put_flags();
  // The products run over the blocks of the LPs still iterating (a block
  // map); the map is rebuilt once a quarter of its blocks belong to LPs
  // that have finished.
  DevCsr k_run = ds.k, kt_run = ds.kt;
  std::vector<Int> km, ktm;
  auto compact = [&] {
    // Queued products may still read the previous maps.
    be.synchronize();
    km = launch_blocks(ds.k_lp_blocks_host, on, ds.k_binned, ds.k_rank);
    ktm = launch_blocks(ds.kt_lp_blocks_host, on, ds.kt_binned, ds.kt_rank);
    auto put = [&](int slot, const std::vector<Int>& h) {
      Int* d = reinterpret_cast<Int*>(be.scratch(slot, h.size() / 2 + 1));
      if (!h.empty())
        hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(Int), hipMemcpyHostToDevice, be.stream()),
                  "norm block map");
      return d;
    };
    k_run.block_map = put(kMultiScratchSlot + 97, km);
    kt_run.block_map = put(kMultiScratchSlot + 98, ktm);
    k_run.blocks = static_cast<Int>(km.size());
    kt_run.blocks = static_cast<Int>(ktm.size());
  };
  compact();
  std::vector<double> out(uz(L)), div(uz(L), 1.0);
  hipStream_t s = be.stream();
  for (int it = 0; it < iterations && std::any_of(on.begin(), on.end(), [](char c) { return c; }); ++it) {
    launch_multi_product(s, k_run, ds.k_block_lp, d_v, d_kv, d_params);
    launch_multi_product(s, kt_run, ds.kt_block_lp, d_kv, d_u, d_params);
    launch_multi_sumsq(s, ds.cols, d_u, nullptr, nullptr, nullptr, false, d_params, d_part,
                       ds.col_lp_chunks, L, d_out, 1, 0);
    hip_check(hipMemcpyAsync(out.data(), d_out, uz(L) * sizeof(double), hipMemcpyDeviceToHost, s), "norms");
    be.synchronize();
    bool changed = false;
    for (Int l = 0; l < L; ++l) {
      if (!on[uz(l)]) continue;
      const double next = std::sqrt(out[uz(l)]);
      if (next == 0.0) {
        norm[uz(l)] = 0.0;
        on[uz(l)] = 0;
        changed = true;
        continue;
      }
      div[uz(l)] = next;
      const double estimate = std::sqrt(next);
      if (it > 0 && std::fabs(estimate - sigma[uz(l)]) <= tolerance * estimate) {
        norm[uz(l)] = estimate;
        on[uz(l)] = 0;
        changed = true;
        continue;
      }
      sigma[uz(l)] = estimate;
      norm[uz(l)] = estimate;
    }
    hip_check(hipMemcpyAsync(d_div, div.data(), uz(L) * sizeof(double), hipMemcpyHostToDevice, s), "divisors");
    if (changed) {
      put_flags();
      Int live = 0;
      for (Int l = 0; l < L; ++l)
        if (on[uz(l)]) live += ds.k_lp_blocks_host[uz(l) + 1] - ds.k_lp_blocks_host[uz(l)];
      if (4 * live < 3 * k_run.blocks) compact();
    }
    launch_multi_divide(s, ds.cols, d_u, d_div, d_params, d_v);
  }
  be.synchronize();
  return norm;
}

// ---- one engine: a set of LPs of the stack iterated together --------------

struct LpRun {
  bool participates = true;
  double initial_omega = 0.0;  // 0: from ||c|| / ||q||
  double initial_eta = 0.0;    // 0: 0.9 / ||K||
  std::vector<double> warm_x, warm_y;
  double q_two = 0, c_two = 0, q_inf = 0, c_inf = 0;
};

// One polished pair per seed, unscaled, with the iterations it cost.
struct Polished {
  std::vector<double> x, y;
  Int spent = 0;
};

struct Outcome {
  int status = ANUK_PDHG_ITERATION_LIMIT;
  Int iterations = 0, restarts = 0;
  double eta = 0.0, omega = 0.0, seconds = 0.0;
  std::string message;
  std::vector<double> x, y;  // working space, unless polished
  bool polished = false;
  Residual polished_r;
  Int polish_attempts = 0, polish_iterations = 0;
};

// Iteration and convergence data of an engine, concatenated over the stack.
// One of an engine's problem vectors: where its device copy comes from (a
// device array in the stack's layout, copied range by range, or, when
// `dev` is null, the constant `fill`), and its host copy, read for the
// per-LP norms (objective, right-hand side) and start points (bounds); a
// null host copy reads as zeros.
struct Vec {
  const double* dev = nullptr;
  const std::vector<double>* host = nullptr;
  double fill = 0.0;
};

struct EngineData {
  Vec c, q, lower, upper, oc, oq, olo, ohi, row_scale, col_scale;
};

struct PolishContext {
  const std::vector<const StandardLp*>* originals;
};

std::vector<Outcome> run_engine(HipBackend& be, const Stack& st, const DeviceStack& ds,
                                const std::vector<double>& norms, const anuk_pdhg_options& opt,
                                const EngineData& data, std::vector<LpRun>& runs,
                                Clock::time_point start, const PolishContext* polish,
                                Phases* phases, int depth);

class Engine {
 public:
  Engine(HipBackend& be, const Stack& st, const DeviceStack& ds, const std::vector<double>& norms,
         const anuk_pdhg_options& opt, const EngineData& data, std::vector<LpRun>& runs,
         Clock::time_point start, const PolishContext* polish, Phases* phases, int depth)
      : be_(be), st_(st), ds_(ds), norms_(norms), opt_(opt), data_(data), runs_(runs),
        start_(start), polish_(polish), phases_(phases), depth_(depth) {}
  ~Engine() {
    for (auto& g : graphs_) {
      (void)hipGraphExecDestroy(g.second.second);
      (void)hipGraphDestroy(g.second.first);
    }
    // The pinned staging is the backend's, kept for the next engine.
    if (pinned_) (void)hipStreamSynchronize(be_.stream());
  }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  std::vector<Outcome> run();

 private:
  void allocate();
  void upload();
  void upload_params();
  void enqueue_period(int steps, bool has_check, int parity);
  hipGraphExec_t graph_for(int parity);
  void extract(Int l, Outcome* o);
  void compact();
  void run_restarts(const std::vector<Int>& who);
  // Feasibility polishing runs beside the main periods, as in the batched
  // solver: an attempt takes its seeds at its check, runs its sub-engines
  // through the backend's helper on another host thread, and is applied once
  // it is back. A failed attempt changes nothing; a successful one retires
  // its LP as of that check.
  struct PendingPolish {
    std::vector<Int> lps;
    Int iteration = 0;              // index of the check that started it
    std::vector<Control> at_check;  // each LP's control as that check left it
    std::future<std::vector<Polished>> result;
  };
  void start_polish(const std::vector<Int>& who, Int budget, Int iteration);
  void settle_polish(std::vector<Outcome>& out, bool wait);
  void apply_polish(PendingPolish& p, std::vector<Outcome>& out);

  HipBackend& be_;
  const Stack& st_;
  const DeviceStack& ds_;
  const std::vector<double>& norms_;
  const anuk_pdhg_options& opt_;
  EngineData data_;
  std::vector<LpRun>& runs_;
  Clock::time_point start_;
  const PolishContext* polish_;
  Phases* phases_;
  int depth_;

  Int L_ = 0;
  // The stack's matrices as launched: only the active LPs' row blocks
  // (block_map), rebuilt when a quarter of them have retired.
  DevCsr k_run_, kt_run_;
  Int compacted_blocks_ = 0;
  std::vector<Control> ctl_;
  std::vector<char> active_;
  int cur_ = 0;
  double *X_[2], *Y_[2], *KX_[2];
  double *XBAR_, *DX_, *AX_, *LAM_, *RX_, *PRX_;
  double *DY_, *KDX_, *AY_, *AKX_, *RY_, *PRY_;
  double *CHECK_, *RESTART_, *PA_, *PB_, *PC_, *PD_, *PCH_;
  std::vector<PendingPolish> pending_;
  std::vector<char> polishing_;  // per LP: an attempt in flight
  std::vector<char> settled_;    // per LP: a successful attempt fixed the outcome
  IterParams* d_params_ = nullptr;
  IterParams* d_rparams_ = nullptr;
  DevProblem prob_;
  double* pinned_ = nullptr;  // params staging, then check values
  size_t pinned_doubles_ = 0;
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs_;
};

void Engine::compact() {
  const std::vector<Int> km = launch_blocks(ds_.k_lp_blocks_host, active_, ds_.k_binned, ds_.k_rank);
  const std::vector<Int> ktm =
      launch_blocks(ds_.kt_lp_blocks_host, active_, ds_.kt_binned, ds_.kt_rank);
  auto put = [&](int slot, const std::vector<Int>& h) {
    Int* d = reinterpret_cast<Int*>(be_.scratch(slot, h.size() / 2 + 1));
    if (!h.empty())
      hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(Int), hipMemcpyHostToDevice, be_.stream()),
                "upload block map");
    return d;
  };
  // The maps feed the captured graphs: wait for them before rewriting.
  be_.synchronize();
  for (auto& g : graphs_) {
    (void)hipGraphExecDestroy(g.second.second);
    (void)hipGraphDestroy(g.second.first);
  }
  graphs_.clear();
  k_run_ = ds_.k;
  kt_run_ = ds_.kt;
  k_run_.block_map = put(kMultiScratchSlot + 10 + 2 * depth_, km);
  kt_run_.block_map = put(kMultiScratchSlot + 11 + 2 * depth_, ktm);
  k_run_.blocks = static_cast<Int>(km.size());
  kt_run_.blocks = static_cast<Int>(ktm.size());
  compacted_blocks_ = k_run_.blocks + kt_run_.blocks;
  be_.synchronize();  // the host lists go out of scope
}

void Engine::allocate() {
  const size_t N = uz(st_.n), M = uz(st_.m), L = uz(L_);
  auto round = [](size_t d) { return (std::max<size_t>(d, 1) + 31) / 32 * 32; };
  const size_t pa = uz(ds_.kt.blocks), pb = uz(ds_.k.blocks) * 7, pc = uz(ds_.kt.blocks) * 4,
               pd = uz(ds_.kt.blocks) * 2, pch = uz(std::max(ds_.col_chunks, ds_.row_chunks));
  const size_t sizes[] = {
      N, N, N, N, N, N, N, N,               // X0 X1 XBAR DX AX LAM RX PRX
      N, N, N, N, N, N, N,                  // c lower upper oc olo ohi col_scale
      M, M, M, M, M, M, M, M, M, M,         // Y0 Y1 KX0 KX1 DY KDX AY AKX RY PRY
      M, M, M,                              // q oq row_scale
      L * kCheckSlots, L * 2,               // CHECK RESTART
      L * kParamDoubles, L * kParamDoubles, // params, restart params
      pa, pb, pc, pd, pch};
  size_t total = 0;
  for (size_t s : sizes) total += round(s);
  double* slab = be_.scratch(kMultiScratchSlot + depth_, total);
  hip_check(hipMemsetAsync(slab, 0, total * sizeof(double), be_.stream()), "hipMemsetAsync(multi)");
  double* p = slab;
  size_t i = 0;
  auto take = [&]() {
    double* q = p;
    p += round(sizes[i++]);
    return q;
  };
  X_[0] = take(); X_[1] = take(); XBAR_ = take(); DX_ = take(); AX_ = take(); LAM_ = take();
  RX_ = take(); PRX_ = take();
  double* c = take(); double* lo = take(); double* hi = take();
  double* oc = take(); double* olo = take(); double* ohi = take(); double* cs = take();
  Y_[0] = take(); Y_[1] = take(); KX_[0] = take(); KX_[1] = take(); DY_ = take(); KDX_ = take();
  AY_ = take(); AKX_ = take(); RY_ = take(); PRY_ = take();
  double* q = take(); double* oq = take(); double* rs = take();
  CHECK_ = take(); RESTART_ = take();
  d_params_ = reinterpret_cast<IterParams*>(take());
  d_rparams_ = reinterpret_cast<IterParams*>(take());
  PA_ = take(); PB_ = take(); PC_ = take(); PD_ = take(); PCH_ = take();
  prob_.n = st_.n;
  prob_.m = st_.m;
  prob_.num_equalities = 0;  // per LP, in IterParams
  prob_.c = c; prob_.lower = lo; prob_.upper = hi; prob_.q = q;
  prob_.orig_c = oc; prob_.orig_lower = olo; prob_.orig_upper = ohi; prob_.orig_q = oq;
  prob_.row_scale = rs; prob_.col_scale = cs;
  pinned_doubles_ = std::max(L * kParamDoubles, L * kCheckSlots) + L * 2 + 64;
  pinned_ = be_.pinned_slot(kMultiScratchSlot + depth_, pinned_doubles_);
}

void Engine::upload() {
  // A sub-engine (polishing a few LPs) builds start points for its LPs only.
  Int taking = 0;
  for (const LpRun& r : runs_) taking += r.participates ? 1 : 0;
  const bool whole = taking == L_;
  hipStream_t s = be_.stream();
  // Copied on the device from the stack's arrays, or filled with a
  // constant; the slab is zeroed when allocated, so zeros need no work. A
  // sub-engine takes whole vectors too: one copy costs less than one per
  // LP, and every kernel reads only the ranges of LPs it runs (blocks and
  // chunks are per LP, gated by each LP's flag).
  auto put_range = [&](const Vec& v, const double* to, const std::vector<Int>& off) {
    const Int len = off.back();
    if ((v.dev == nullptr && v.fill == 0.0) || len <= 0) return;
    double* d = const_cast<double*>(to);  // the engine's own slab
    if (v.dev != nullptr)
      hip_check(hipMemcpyAsync(d, v.dev, uz(len) * sizeof(double), hipMemcpyDeviceToDevice, s),
                "engine vectors");
    else
      launch_fill(s, d, len, v.fill);
  };
  put_range(data_.c, prob_.c, st_.col_off);
  put_range(data_.lower, prob_.lower, st_.col_off);
  put_range(data_.upper, prob_.upper, st_.col_off);
  put_range(data_.oc, prob_.orig_c, st_.col_off);
  put_range(data_.olo, prob_.orig_lower, st_.col_off);
  put_range(data_.ohi, prob_.orig_upper, st_.col_off);
  put_range(data_.col_scale, prob_.col_scale, st_.col_off);
  put_range(data_.q, prob_.q, st_.row_off);
  put_range(data_.oq, prob_.orig_q, st_.row_off);
  put_range(data_.row_scale, prob_.row_scale, st_.row_off);
  // Starting points (pdhg.cpp:479-503): a warm start put in the dual cone and
  // the primal box, a cold one the projection of zero. Built per LP, so a
  // sub-engine polishing a few LPs touches only their ranges on the host.
  auto start_point = [&](Int l, double* x, double* y) {
    const LpRun& r = runs_[uz(l)];
    const Int c0 = st_.col_off[uz(l)], r0 = st_.row_off[uz(l)];
    const bool warm = !r.warm_x.empty() && !r.warm_y.empty();
    for (Int i = 0; i < st_.rows(l); ++i) {
      double v = warm ? r.warm_y[uz(i)] : 0.0;
      if (warm && r0 + i >= st_.eq_end[uz(l)] && !(v > 0.0)) v = 0.0;
      y[uz(i)] = v;
    }
    for (Int j = 0; j < st_.cols(l); ++j) {
      const double lo = data_.lower.host != nullptr ? (*data_.lower.host)[uz(c0 + j)] : 0.0;
      const double hi = data_.upper.host != nullptr ? (*data_.upper.host)[uz(c0 + j)] : 0.0;
      double v = warm ? r.warm_x[uz(j)] : 0.0;
      if (!std::isfinite(v)) v = 0.0;
      if (v < lo) v = std::isfinite(lo) ? lo : 0.0;
      if (v > hi) v = std::isfinite(hi) ? hi : 0.0;
      x[uz(j)] = v;
    }
  };
  const bool warm = std::any_of(runs_.begin(), runs_.end(), [](const LpRun& r) {
    return r.participates && !r.warm_x.empty() && !r.warm_y.empty();
  });
  if (whole && !warm) {
    // Every LP cold: y stays zero (the slab), x is zero moved into the box.
    launch_cold_start(s, st_.n, prob_.lower, prob_.upper, X_[0]);
  } else if (whole) {
    std::vector<double> x(uz(st_.n), 0.0), y(uz(st_.m), 0.0);
    for (Int l = 0; l < L_; ++l)
      start_point(l, x.data() + st_.col_off[uz(l)], y.data() + st_.row_off[uz(l)]);
    be_.upload_staged(x.data(), X_[0], x.size());
    be_.upload_staged(y.data(), Y_[0], y.size());
  } else {
    // Each participating LP's start point is written straight into pinned
    // staging and copied asynchronously; one wait covers them all.
    size_t need = 0;
    for (Int l = 0; l < L_; ++l)
      if (runs_[uz(l)].participates) need += uz(st_.cols(l)) + uz(st_.rows(l));
    double* stage = be_.pinned_scratch(std::max<size_t>(need, 1));
    size_t at = 0;
    for (Int l = 0; l < L_; ++l) {
      if (!runs_[uz(l)].participates) continue;
      const size_t nc = uz(st_.cols(l)), nr = uz(st_.rows(l));
      start_point(l, stage + at, stage + at + nc);
      if (nc > 0)
        hip_check(hipMemcpyAsync(X_[0] + st_.col_off[uz(l)], stage + at, nc * sizeof(double),
                                 hipMemcpyHostToDevice, be_.stream()),
                  "upload start points");
      if (nr > 0)
        hip_check(hipMemcpyAsync(Y_[0] + st_.row_off[uz(l)], stage + at + nc, nr * sizeof(double),
                                 hipMemcpyHostToDevice, be_.stream()),
                  "upload start points");
      at += nc + nr;
    }
    be_.synchronize();
  }
  be_.copy(X_[0], RX_, st_.n);
  be_.copy(X_[0], PRX_, st_.n);
  be_.copy(Y_[0], RY_, st_.m);
  be_.copy(Y_[0], PRY_, st_.m);
}

void Engine::upload_params() {
  auto* P = reinterpret_cast<IterParams*>(pinned_);
  for (Int l = 0; l < L_; ++l) {
    IterParams ip{};
    if (active_[uz(l)]) {
      ip = ctl_[uz(l)].period_params(st_.eq_end[uz(l)]);
      ip.active = 1;
    } else {
      ip.active = 0;
      ip.num_equalities = st_.eq_end[uz(l)];
    }
    P[l] = ip;
  }
  hip_check(hipMemcpyAsync(d_params_, P, uz(L_) * sizeof(IterParams), hipMemcpyHostToDevice,
                           be_.stream()),
            "upload multi params");
}

void Engine::enqueue_period(int steps, bool has_check, int parity) {
  hipStream_t s = be_.stream();
  int p = parity;
  for (int step = 0; step < steps; ++step) {
    const bool diag = has_check && step == steps - 1;
    PrimalBufs pb{X_[p], AX_, X_[1 - p], XBAR_, DX_, AX_};
    const LaunchEvents pe = be_.launch_events(diag ? "multi_primal_step_check" : "multi_primal_step");
    launch_multi_primal_step(s, kt_run_, ds_.kt_block_lp, Y_[p], prob_, d_params_, step, pb, diag,
                             PA_, &pe);
    DualBufs db{Y_[p], KX_[p], AY_, AKX_, Y_[1 - p], KX_[1 - p], DY_, KDX_, AY_, AKX_};
    const LaunchEvents de = be_.launch_events(diag ? "multi_dual_step_check" : "multi_dual_step");
    launch_multi_dual_step(s, k_run_, ds_.k_block_lp, XBAR_, prob_, d_params_, step, db, diag, PB_,
                           &de);
    p = 1 - p;
  }
  if (has_check) {
    HipBackend::Timed t(be_, "multi_check");
    launch_multi_fold(s, PA_, 1, kSumDx2, 0u, ds_.kt_lp_blocks, L_, d_params_, CHECK_);
    launch_multi_fold(s, PB_, 7, kSumDy2, (1u << 2) | (1u << 4) | (1u << 5), ds_.k_lp_blocks, L_,
                      d_params_, CHECK_);
    launch_multi_conv_product(s, kt_run_, ds_.kt_block_lp, Y_[p], X_[p], prob_, d_params_, PC_);
    launch_multi_fold(s, PC_, 4, kDualSq, 1u << 1, ds_.kt_lp_blocks, L_, d_params_, CHECK_);
    if (opt_.detect_infeasibility) {
      launch_multi_farkas_product(s, kt_run_, ds_.kt_block_lp, DY_, prob_, d_params_, CHECK_, PD_);
      launch_multi_fold(s, PD_, 2, kFarkasColBad, 1u << 0, ds_.kt_lp_blocks, L_, d_params_, CHECK_);
    }
  }
}

hipGraphExec_t Engine::graph_for(int parity) {
  auto it = graphs_.find(parity);
  if (it != graphs_.end()) return it->second.second;
  hipStream_t s = be_.stream();
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    enqueue_period(opt_.termination_check_frequency, true, parity);
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  graphs_[parity] = {g, e};
  return e;
}

void Engine::extract(Int l, Outcome* o) {
  const Int n = st_.cols(l), m = st_.rows(l);
  o->x.resize(uz(n));
  o->y.resize(uz(m));
  if (n > 0)
    hip_check(hipMemcpyAsync(o->x.data(), X_[cur_] + st_.col_off[uz(l)], uz(n) * sizeof(double),
                             hipMemcpyDeviceToHost, be_.stream()),
              "multi extract");
  if (m > 0)
    hip_check(hipMemcpyAsync(o->y.data(), Y_[cur_] + st_.row_off[uz(l)], uz(m) * sizeof(double),
                             hipMemcpyDeviceToHost, be_.stream()),
              "multi extract");
}

// A restart for the listed LPs (pdhg.cpp:1124-1202): K x recomputed, the
// distances moved since the last restart measured, the restart points
// advanced.
void Engine::run_restarts(const std::vector<Int>& who) {
  if (who.empty()) return;
  std::vector<char> on(uz(L_), 0);
  for (Int l : who) on[uz(l)] = 1;
  const std::vector<IterParams> f = flags(on);
  hipStream_t s = be_.stream();
  hip_check(hipMemcpyAsync(d_rparams_, f.data(), f.size() * sizeof(IterParams), hipMemcpyHostToDevice, s),
            "restart flags");
  launch_multi_product(s, k_run_, ds_.k_block_lp, X_[cur_], KX_[cur_], d_rparams_);
  const bool previous = !opt_.pid_primal_weight;
  launch_multi_sumsq(s, ds_.cols, X_[cur_], previous ? PRX_ : RX_, RX_, PRX_, true, d_rparams_, PCH_,
                     ds_.col_lp_chunks, L_, RESTART_, 2, 0);
  launch_multi_sumsq(s, ds_.rows, Y_[cur_], previous ? PRY_ : RY_, RY_, PRY_, true, d_rparams_, PCH_,
                     ds_.row_lp_chunks, L_, RESTART_, 2, 1);
  double* h = pinned_ + std::max(uz(L_) * kParamDoubles, uz(L_) * kCheckSlots);
  hip_check(hipMemcpyAsync(h, RESTART_, uz(L_) * 2 * sizeof(double), hipMemcpyDeviceToHost, s),
            "restart distances");
  be_.synchronize();
  for (Int l : who) ctl_[uz(l)].restart(std::sqrt(h[2 * l]), std::sqrt(h[2 * l + 1]));
}

// Polished pairs for the LPs `who` (one budget): the primal feasibility
// problem (c = 0) from (x, 0) and the dual one (q = 0, finite bounds zeroed)
// from (0, y), each LP with its step size and primal weight, no scaling (the
// seeds are in the scaled space), as sub-engines over the same stack
// (pdhg.cpp:680-743); then unscaled with each LP's diagonals.
std::vector<Polished> polish_stack(HipBackend& be, const Stack& st, const DeviceStack& ds,
                                   const std::vector<double>& norms, const anuk_pdhg_options& opt,
                                   const std::vector<Int>& who,
                                   const std::vector<std::vector<double>>& seed_x,
                                   const std::vector<std::vector<double>>& seed_y,
                                   const std::vector<double>& omega, const std::vector<double>& eta,
                                   Int budget, Clock::time_point start) {
  anuk_pdhg_options sub = opt;
  sub.polish_feasibility = 0;
  sub.polish_on_exit = 0;
  sub.detect_infeasibility = 0;
  sub.verbose = 0;
  sub.ruiz_iterations = 0;
  sub.pock_chambolle = 0;
  sub.max_iterations = budget;
  sub.termination_check_frequency =
      std::max<Int>(10, std::min<Int>(opt.termination_check_frequency, budget / 4));
  sub.time_limit_seconds = std::fmax(0.0, opt.time_limit_seconds - seconds_since(start));
  sub.warm_x = nullptr;
  sub.warm_y = nullptr;
  auto runs_for = [&](bool primal) {
    std::vector<LpRun> runs(uz(st.lps));
    for (LpRun& r : runs) r.participates = false;
    for (size_t s = 0; s < who.size(); ++s) {
      const Int l = who[s];
      LpRun& r = runs[uz(l)];
      r.participates = true;
      r.initial_omega = omega[s];
      r.initial_eta = eta[s];
      if (primal) {
        r.warm_x = seed_x[s];
        r.warm_y.assign(uz(st.rows(l)), 0.0);
      } else {
        r.warm_x.assign(uz(st.cols(l)), 0.0);
        r.warm_y = seed_y[s];
      }
    }
    return runs;
  };
  std::vector<LpRun> pr = runs_for(true), dr = runs_for(false);
  sub.termination = ANUK_TERM_PRIMAL_FEASIBILITY;
  // The feasibility problems in the scaled space (scales of one): the
  // primal one drops the objective, the dual one the right-hand side and
  // puts the dual problem's bounds on the columns.
  const Vec zero{}, one{nullptr, nullptr, 1.0};
  const Vec q{ds.q, &st.q}, lo{ds.lower, &st.lower}, hi{ds.upper, &st.upper};
  // The dual problem starts every LP at x = 0, inside its bounds (0 or
  // infinite), so its bounds need no host copy for the start points.
  const Vec c{ds.c, &st.c}, dlo{ds.dual_lower}, dhi{ds.dual_upper};
  const EngineData pdata{zero, q, lo, hi, zero, q, lo, hi, one, one};
  const std::vector<Outcome> po = run_engine(be, st, ds, norms, sub, pdata, pr, start, nullptr, nullptr, 1);
  sub.termination = ANUK_TERM_DUAL_FEASIBILITY;
  const EngineData ddata{c, zero, dlo, dhi, c, zero, dlo, dhi, one, one};
  const std::vector<Outcome> dq = run_engine(be, st, ds, norms, sub, ddata, dr, start, nullptr, nullptr, 1);
  std::vector<Polished> out(who.size());
  for (size_t s = 0; s < who.size(); ++s) {
    const Int l = who[s];
    st.scaling[uz(l)].unscale_primal(po[uz(l)].x, &out[s].x);
    st.scaling[uz(l)].unscale_dual(dq[uz(l)].y, &out[s].y);
    out[s].spent = po[uz(l)].iterations + dq[uz(l)].iterations;
  }
  return out;
}

void Engine::start_polish(const std::vector<Int>& who, Int budget, Int iteration) {
  std::vector<std::vector<double>> sx, sy;
  std::vector<double> omega, eta;
  PendingPolish pp;
  pp.lps = who;
  pp.iteration = iteration;
  for (Int l : who) {
    Outcome tmp;
    extract(l, &tmp);
    sx.push_back(std::move(tmp.x));
    sy.push_back(std::move(tmp.y));
    omega.push_back(ctl_[uz(l)].omega);
    eta.push_back(ctl_[uz(l)].eta);
    pp.at_check.push_back(ctl_[uz(l)]);
    polishing_[uz(l)] = 1;
  }
  be_.synchronize();
  HipBackend* helper = &be_.helper();
  const int device = be_.device();
  // The task reads the stack and its device copy (read-only, shared with
  // the main periods) and writes only the helper's memory; the engine waits
  // for every attempt before it returns.
  pp.result = std::async(std::launch::async, [this, helper, device, budget, who,
                                              sx = std::move(sx), sy = std::move(sy),
                                              omega = std::move(omega), eta = std::move(eta)]() {
    hip_check(hipSetDevice(device), "hipSetDevice(polish)");
    return polish_stack(*helper, st_, ds_, norms_, opt_, who, sx, sy, omega, eta, budget, start_);
  });
  pending_.push_back(std::move(pp));
}

void Engine::apply_polish(PendingPolish& pp, std::vector<Outcome>& out) {
  std::vector<Polished> pol = pp.result.get();
  // Each polished point is checked on the original LP, one LP per host thread.
  std::vector<Residual> checked(pp.lps.size());
  parallel_for(pp.lps.size(), [&](size_t s) {
    checked[s] = evaluate_residual(*(*polish_->originals)[uz(pp.lps[s])], pol[s].x, pol[s].y);
  });
  for (size_t s = 0; s < pp.lps.size(); ++s) {
    const Int l = pp.lps[s];
    Outcome& o = out[uz(l)];
    const Residual& tr = checked[s];
    ++o.polish_attempts;
    o.polish_iterations += pol[s].spent;
    polishing_[uz(l)] = 0;
    const Control& c = pp.at_check[s];
    if (!tr.converged(c.tol, opt_.absolute_tolerance, opt_.require_inf_norm_termination != 0,
                      c.gap_tol))
      continue;
    // Converged: the LP's outcome is the polished pair as of the check that
    // started the attempt, whatever the LP did since.
    o.x = std::move(pol[s].x);
    o.y = std::move(pol[s].y);
    o.polished = true;
    o.polished_r = tr;
    o.status = ANUK_PDHG_OPTIMAL;
    o.message = "converged after feasibility polishing";
    o.iterations = pp.iteration + 1;
    o.seconds = seconds_since(start_);
    o.restarts = c.restarts;
    o.eta = c.eta;
    o.omega = c.omega;
    settled_[uz(l)] = 1;
    active_[uz(l)] = 0;
  }
}

void Engine::settle_polish(std::vector<Outcome>& out, bool wait) {
  for (auto it = pending_.begin(); it != pending_.end();) {
    if (wait || it->result.wait_for(std::chrono::seconds(0)) == std::future_status::ready) {
      apply_polish(*it, out);
      it = pending_.erase(it);
    } else {
      ++it;
    }
  }
}

// Norms of a range of a host vector; a null vector is all zeros.
double range_two_norm(const std::vector<double>* v, Int a, Int b) {
  double s = 0.0;
  if (v != nullptr)
    for (Int i = a; i < b; ++i) s += (*v)[uz(i)] * (*v)[uz(i)];
  return std::sqrt(s);
}

double range_inf_norm(const std::vector<double>* v, Int a, Int b) {
  double s = 0.0;
  if (v != nullptr)
    for (Int i = a; i < b; ++i) s = std::fmax(s, std::fabs((*v)[uz(i)]));
  return s;
}

std::vector<Outcome> Engine::run() {
  Clock::time_point mark = Clock::now();
  L_ = st_.lps;
  allocate();
  ctl_.resize(uz(L_));
  active_.assign(uz(L_), 0);
  polishing_.assign(uz(L_), 0);
  settled_.assign(uz(L_), 0);
  for (Int l = 0; l < L_; ++l) {
    LpRun& r = runs_[uz(l)];
    if (!r.participates) continue;
    const Int c0 = st_.col_off[uz(l)], c1 = st_.col_off[uz(l) + 1];
    const Int r0 = st_.row_off[uz(l)], r1 = st_.row_off[uz(l) + 1];
    r.q_two = range_two_norm(data_.oq.host, r0, r1);
    r.c_two = range_two_norm(data_.oc.host, c0, c1);
    r.q_inf = range_inf_norm(data_.oq.host, r0, r1);
    r.c_inf = range_inf_norm(data_.oc.host, c0, c1);
    Control& c = ctl_[uz(l)];
    c.init(opt_, norms_[uz(l)], range_two_norm(data_.c.host, c0, c1), range_two_norm(data_.q.host, r0, r1));
    if (r.initial_eta > 0.0) c.eta = r.initial_eta;
    if (r.initial_omega > 0.0) c.omega = r.initial_omega;
    active_[uz(l)] = norms_[uz(l)] > 0.0 ? 1 : 0;
  }
  upload();
  upload_params();
  compact();
  launch_multi_product(be_.stream(), k_run_, ds_.k_block_lp, X_[0], KX_[0], d_params_);
  cur_ = 0;
  auto lap = [&](double Phases::*field) {
    const Clock::time_point now = Clock::now();
    if (phases_ != nullptr) phases_->*field += std::chrono::duration<double>(now - mark).count();
    mark = now;
  };
  be_.synchronize();
  lap(&Phases::setup);

  std::vector<Outcome> out(uz(L_));
  auto any_active = [&] { return std::any_of(active_.begin(), active_.end(), [](char c) { return c; }); };
  const Int freq = opt_.termination_check_frequency;
  Int iteration = 0;
  while (iteration < opt_.max_iterations && any_active()) {
    const Int period_end = std::min<Int>((iteration / freq + 1) * freq, opt_.max_iterations);
    const int steps = period_end - iteration;
    const bool has_check = (period_end % freq == 0);
    upload_params();
    // Sub-engines (polishing) run a few short periods: capturing and
    // instantiating their graphs costs more than launching directly.
    const bool graph = opt_.use_graphs != 0 && depth_ == 0 && has_check && steps == freq &&
                       !be_.profiling();
    if (graph) {
      hip_check(hipGraphLaunch(graph_for(cur_), be_.stream()), "hipGraphLaunch(multi)");
    } else {
      enqueue_period(steps, has_check, cur_);
    }
    for (Int l = 0; l < L_; ++l)
      if (active_[uz(l)]) ctl_[uz(l)].advance(steps);
    if (steps % 2 == 1) cur_ = 1 - cur_;
    iteration = period_end;
    if (!has_check) {
      be_.synchronize();
      lap(&Phases::iterating);
      continue;
    }
    double* h_check = pinned_;
    hip_check(hipMemcpyAsync(h_check, CHECK_, uz(L_) * kCheckSlots * sizeof(double),
                             hipMemcpyDeviceToHost, be_.stream()),
              "multi check");
    be_.synchronize();
    lap(&Phases::iterating);
    if (!pending_.empty()) settle_polish(out, false);
    const Int i = period_end - 1;
    const double elapsed = seconds_since(start_);
    std::vector<char> stopped(uz(L_), 0);
    std::vector<Residual> res(uz(L_));
    std::map<Int, std::vector<Int>> polish_by_budget;
    for (Int l = 0; l < L_; ++l) {
      if (!active_[uz(l)]) continue;
      Control& c = ctl_[uz(l)];
      const LpRun& r = runs_[uz(l)];
      const double* h = h_check + uz(l) * kCheckSlots;
      c.movement = c.omega * h[kSumDx2] + h[kSumDy2] / c.omega;
      bool farkas = false;
      if (opt_.detect_infeasibility) {
        farkas = h[kFarkasRowBad] == 0.0 && h[kFarkasScale] > 0.0 && h[kFarkasColBad] == 0.0 &&
                 (h[kFarkasRhs] - h[kFarkasSupport]) / h[kFarkasScale] > opt_.infeasibility_tolerance;
      }
      Terms t;
      t.primal_sq = h[kPrimalSq];
      t.primal_inf = h[kPrimalInf];
      t.q_dot_y = h[kQDotY];
      t.dual_sq = h[kDualSq];
      t.dual_inf = h[kDualInf];
      t.bound_term = h[kBoundTerm];
      t.c_dot_x = h[kCDotX];
      ResidualScales rs;
      rs.q_two = r.q_two;
      rs.c_two = r.c_two;
      rs.q_inf = r.q_inf;
      rs.c_inf = r.c_inf;
      res[uz(l)] = residual_from_terms(t, rs);
      Int budget = 0;
      const Control::Verdict v = c.check_first(i, res[uz(l)], farkas, &budget);
      if (v == Control::Verdict::kStop) {
        stopped[uz(l)] = 1;
      } else if (v == Control::Verdict::kPolish) {
        polish_by_budget[budget].push_back(l);
      }
    }
    lap(&Phases::checks);
    if (polish_ != nullptr && !polish_by_budget.empty()) {  // sub-engines never polish
      // An LP due another attempt while one is in flight has that one
      // settled first (it may have succeeded, which retires the LP).
      bool due_in_flight = false;
      for (const auto& kv : polish_by_budget)
        for (Int l : kv.second) due_in_flight = due_in_flight || polishing_[uz(l)] != 0;
      if (due_in_flight) settle_polish(out, true);
      for (const auto& kv : polish_by_budget) {
        std::vector<Int> who;
        for (Int l : kv.second)
          if (active_[uz(l)]) who.push_back(l);
        if (!who.empty()) start_polish(who, kv.first, i);
      }
    }
    lap(&Phases::polishing);
    std::vector<Int> restart;
    for (Int l = 0; l < L_; ++l) {
      if (!active_[uz(l)] || stopped[uz(l)]) continue;
      const Control::Verdict v = ctl_[uz(l)].check_second(i, res[uz(l)], 0.0, elapsed);
      if (v == Control::Verdict::kStop) {
        stopped[uz(l)] = 1;
      } else if (v == Control::Verdict::kRestart) {
        restart.push_back(l);
      }
    }
    for (Int l = 0; l < L_; ++l) {
      if (!stopped[uz(l)] || settled_[uz(l)]) continue;
      Outcome& o = out[uz(l)];
      o.iterations = i + 1;
      o.seconds = seconds_since(start_);
      if (!o.polished) extract(l, &o);
      active_[uz(l)] = 0;
    }
    run_restarts(restart);
    be_.synchronize();
    Int live = 0;
    for (Int l = 0; l < L_; ++l) {
      if (!active_[uz(l)]) continue;
      live += ds_.k_lp_blocks_host[uz(l) + 1] - ds_.k_lp_blocks_host[uz(l)];
      live += ds_.kt_lp_blocks_host[uz(l) + 1] - ds_.kt_lp_blocks_host[uz(l)];
    }
    if (4 * live < 3 * compacted_blocks_) compact();
    lap(&Phases::checks);
  }
  for (Int l = 0; l < L_; ++l) {
    if (!active_[uz(l)]) continue;
    Outcome& o = out[uz(l)];
    o.iterations = iteration;
    o.seconds = seconds_since(start_);
    extract(l, &o);
  }
  be_.synchronize();
  for (Int l = 0; l < L_; ++l) {
    Outcome& o = out[uz(l)];
    const Control& c = ctl_[uz(l)];
    if (!runs_[uz(l)].participates || settled_[uz(l)]) continue;
    if (norms_[uz(l)] <= 0.0) {
      o.status = ANUK_PDHG_NUMERICAL_ERROR;
      o.message = "the constraint matrix is empty";
      continue;
    }
    o.status = c.status;
    o.message = c.message;
    o.restarts = c.restarts;
    o.eta = c.eta;
    o.omega = c.omega;
  }
  // Attempts still in flight decide last: a success stands for its LP
  // whatever the LP did after the attempt's check.
  settle_polish(out, true);
  return out;
}

std::vector<Outcome> run_engine(HipBackend& be, const Stack& st, const DeviceStack& ds,
                                const std::vector<double>& norms, const anuk_pdhg_options& opt,
                                const EngineData& data, std::vector<LpRun>& runs,
                                Clock::time_point start, const PolishContext* polish,
                                Phases* phases, int depth) {
  Engine e(be, st, ds, norms, opt, data, runs, start, polish, phases, depth);
  return e.run();
}

}  // namespace

std::vector<SolveResult> solve_pdhg_multi(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                          const anuk_pdhg_options& opt) {
  const Clock::time_point start = Clock::now();
  const char* why = nullptr;
  if (!batch_supports(opt, &why)) throw std::invalid_argument(std::string("multi: unsupported option: ") + why);
  if (lps.empty()) return {};
  Phases phases;
  Clock::time_point mark = start;
  auto lap = [&](double Phases::*field) {
    const Clock::time_point now = Clock::now();
    phases.*field += std::chrono::duration<double>(now - mark).count();
    mark = now;
  };
  Stack st = concat_stack(lps);
  DeviceStack ds(be);
  upload_stack(st, &ds);
  scale_stack(be, opt, &st, &ds);
  tune_stack(be, st, &ds, opt);
  lap(&Phases::scaling);
  std::vector<double> norms(uz(st.lps), opt.known_matrix_norm);
  if (!(opt.known_matrix_norm > 0.0)) norms = stack_norms(be, st, ds, 200, 1e-8);
  lap(&Phases::matrix_norm);

  std::vector<LpRun> runs(uz(st.lps));
  const EngineData data{{ds.c, &st.c},   {ds.q, &st.q},     {ds.lower, &st.lower}, {ds.upper, &st.upper},
                        {ds.oc, &st.oc}, {ds.oq, &st.oq},   {ds.olo, &st.olo},     {ds.ohi, &st.ohi},
                        {ds.row_scale},  {ds.col_scale}};
  PolishContext pc{&lps};
  std::vector<Outcome> outcomes = run_engine(be, st, ds, norms, opt, data, runs, start,
                                             opt.polish_feasibility ? &pc : nullptr, &phases, 0);

  // Final answers, and one more polish for the LPs that ran out of
  // iterations (pdhg.cpp:1205-1269), grouped by budget.
  std::vector<SolveResult> results(lps.size());
  std::map<Int, std::vector<Int>> exit_polish;
  for (size_t i = 0; i < lps.size(); ++i) {
    Outcome& o = outcomes[i];
    SolveResult& r = results[i];
    r.status = o.status;
    r.message = o.message;
    r.iterations = o.iterations;
    r.restarts = o.restarts;
    r.polish_attempts = o.polish_attempts;
    r.polish_iterations = o.polish_iterations;
    r.matrix_norm_estimate = norms[i];
    r.final_step_size = o.eta;
    r.final_primal_weight = o.omega;
    r.solve_seconds = o.seconds;
    if (o.polished) {
      r.x = std::move(o.x);
      r.y = std::move(o.y);
      r.residual = o.polished_r;
      r.polished = true;
    } else {
      st.scaling[i].unscale_primal(o.x, &r.x);
      st.scaling[i].unscale_dual(o.y, &r.y);
      r.residual = evaluate_residual(*lps[i], r.x, r.y);
      const bool worth = opt.polish_feasibility && opt.polish_on_exit &&
                         opt.termination == ANUK_TERM_FULL && r.status != ANUK_PDHG_OPTIMAL &&
                         r.status != ANUK_PDHG_PRIMAL_INFEASIBLE &&
                         r.status != ANUK_PDHG_NUMERICAL_ERROR &&
                         std::isfinite(r.residual.primal_residual) &&
                         std::isfinite(r.residual.dual_residual) &&
                         seconds_since(start) < opt.time_limit_seconds;
      if (worth) {
        Control c;
        c.opt = &opt;
        const Int budget = c.polish_budget(r.iterations);
        if (budget > 0) exit_polish[budget].push_back(static_cast<Int>(i));
      }
    }
  }
  mark = Clock::now();
  const double gap_tol = opt.gap_tolerance > 0.0 ? opt.gap_tolerance : opt.tolerance;
  for (const auto& kv : exit_polish) {
    std::vector<std::vector<double>> sx, sy;
    std::vector<double> omega, eta;
    for (Int i : kv.second) {
      sx.push_back(outcomes[uz(i)].x);
      sy.push_back(outcomes[uz(i)].y);
      omega.push_back(outcomes[uz(i)].omega);
      eta.push_back(outcomes[uz(i)].eta);
    }
    std::vector<Polished> pol = polish_stack(be, st, ds, norms, opt, kv.second, sx, sy, omega, eta,
                                             kv.first, start);
    for (size_t s = 0; s < kv.second.size(); ++s) {
      SolveResult& r = results[uz(kv.second[s])];
      Polished& p = pol[s];
      const Residual tr = evaluate_residual(*lps[uz(kv.second[s])], p.x, p.y);
      ++r.polish_attempts;
      r.polish_iterations += p.spent;
      const double ceiling = std::fmax(gap_tol, r.residual.relative_gap);
      const bool better = tr.relative_primal <= r.residual.relative_primal &&
                          tr.relative_dual <= r.residual.relative_dual &&
                          (tr.relative_primal < r.residual.relative_primal ||
                           tr.relative_dual < r.residual.relative_dual);
      if (better && tr.relative_gap <= ceiling) {
        r.x = std::move(p.x);
        r.y = std::move(p.y);
        r.residual = tr;
        r.polished = true;
        if (tr.converged(opt.tolerance, opt.absolute_tolerance, opt.require_inf_norm_termination != 0,
                         gap_tol)) {
          r.status = ANUK_PDHG_OPTIMAL;
          r.message = "converged after feasibility polishing";
        } else {
          r.message = r.message.empty() ? "feasibility polished" : r.message + ", feasibility polished";
        }
      }
    }
  }
  lap(&Phases::polishing);
  for (size_t i = 0; i < lps.size(); ++i) {
    results[i].objective = lps[i]->model_objective(results[i].x);
    results[i].phases = phases;
  }
  return results;
}

}  // namespace anuk
