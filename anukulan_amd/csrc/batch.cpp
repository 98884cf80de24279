#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <iterator>
#include <map>
#include <memory>
#include <stdexcept>

#include "batch_kernels.hpp"
#include "control.hpp"
#include "host_pool.hpp"

namespace anuk {

namespace {

using Clock = std::chrono::steady_clock;
using Mask = unsigned long long;

// Backend scratch slots holding the state slabs of batch engines, one per
// nesting depth (the stack, and the polishing sub-batches it runs).
constexpr int kBatchScratchSlot = 100;

double seconds_since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

// One LP of a batch in the solver's working space.
struct Lane {
  std::vector<double> c, q, lower, upper;  // iteration data
  std::vector<double> oc, oq, olo, ohi;    // data the convergence test measures
  std::vector<double> warm_x, warm_y;      // empty: cold start
  double initial_omega = 0.0;              // 0: derive from ||c|| / ||q||
  int termination = -1;                    // -1: the engine's; else this LP's own test
  double q_two = 0, c_two = 0, q_inf = 0, c_inf = 0;  // ResidualScales of (oq, oc)
  void set_scales() {
    q_two = host_two_norm(oq);
    c_two = host_two_norm(oc);
    q_inf = host_inf_norm(oq);
    c_inf = host_inf_norm(oc);
  }
};

struct LaneOutcome {
  int status = ANUK_PDHG_ITERATION_LIMIT;
  Int iterations = 0;
  Int restarts = 0;
  double eta = 0.0, omega = 0.0;
  double seconds = 0.0;
  std::string message;
  std::vector<double> x, y;  // working-space iterate, unless polished
  bool polished = false;     // x, y are the unscaled polished pair
  Residual polished_r;
  Int polish_attempts = 0, polish_iterations = 0;
};

// One polished pair per seed, unscaled, with the iterations it cost.
struct Polished {
  std::vector<double> x, y;
  Int spent = 0;
};

// What a polishing attempt needs from the enclosing batch; absent in
// sub-batches, which never polish.
struct PolishContext {
  const std::vector<const StandardLp*>* originals;  // per lane of this engine
  const Scaling* scaling;
};

class Engine {
 public:
  Engine(HipBackend& be, const Csr& k, const Csr& kt, Int neq, std::vector<double> row_scale,
         std::vector<double> col_scale, double matrix_norm, const anuk_pdhg_options& opt,
         std::vector<Lane>& lanes, Clock::time_point start, const PolishContext* polish,
         Phases* phases, int depth)
      : be_(be), k_(k), kt_(kt), neq_(neq), row_scale_(std::move(row_scale)),
        col_scale_(std::move(col_scale)), norm_(matrix_norm), opt_(opt), lanes_(lanes),
        start_(start), polish_(polish), phases_(phases), depth_(depth) {}
  ~Engine();
  std::vector<LaneOutcome> run();

 private:
  void allocate();
  void upload_lanes();
  // Stages the lane table of the next period in page-locked memory and, unless
  // a captured period copies it itself (`staged_only`), uploads it.
  void upload_params(bool staged_only = false);
  // `host_io`: the period copies the staged lane table in first and
  // publishes its check to the host last (launch_batch_params/_publish).
  void enqueue_period(int steps, bool has_check, int parity, bool host_io = false);
  hipGraphExec_t graph_for(int parity);
  void fetch(const double* dev, double* host, size_t doubles);
  // Waits for the period numbered seq_ to publish its check.
  void wait_published();
  void retire_extract(Mask mask, std::vector<LaneOutcome>& out);
  void run_restarts(Mask mask);
  // Feasibility polishing of the main batch runs beside its periods: an
  // attempt takes its seeds at its check, runs its sub-batches through the
  // backend's helper on another host thread, and is applied once it is back
  // (settle_polish). A failed attempt changes nothing (the lane went on from
  // the check, as it does after a failed attempt); a successful one retires
  // the lane as of that check, and the periods it ran since are discarded.
  struct PendingPolish {
    std::vector<int> lanes;         // in the slot order of the seed extract
    Int iteration = 0;              // index of the check that started it
    std::vector<Control> at_check;  // each lane's control as that check left it
    std::future<std::vector<Polished>> result;
  };
  void start_polish(const std::vector<int>& who, Int budget, Int iteration);
  // Applies the attempts that are back (every one, waiting, with `wait`).
  void settle_polish(std::vector<LaneOutcome>& out, bool wait);
  void apply_polish(PendingPolish& p, std::vector<LaneOutcome>& out);
  void compact();
  void tune();
  // The device slots of a set of lanes.
  Mask slots(Mask lanes) const {
    Mask m = 0;
    for (int l = 0; l < count_; ++l)
      if (((lanes >> l) & 1ull) && slot_of_[l] >= 0) m |= 1ull << slot_of_[l];
    return m;
  }

  HipBackend& be_;
  const Csr& k_;
  const Csr& kt_;
  Int neq_;
  std::vector<double> row_scale_, col_scale_;
  double norm_;
  const anuk_pdhg_options& opt_;
  std::vector<Lane>& lanes_;
  Clock::time_point start_;
  const PolishContext* polish_;
  Phases* phases_;  // where this engine's time goes (null: not recorded)
  int depth_;       // 0: the stack itself, 1: a polishing sub-batch

  Int n_ = 0, m_ = 0;
  int count_ = 0;
  DevCsr dk_, dkt_;    // the CSR
  DevCsr dks_, dkts_;  // the same with their fixed-width row copies
  std::vector<Control> ctl_;
  std::vector<anuk_pdhg_options> lane_opt_;  // each lane's options (its termination test)
  Mask active_ = 0;  // by lane
  int cur_ = 0;
  // Lane l's state sits in device slot slot_of_[l] (-1 once retired and
  // moved out); lane_at_[s] is the lane in slot s (-1: none). Live lanes are
  // kept in the lowest slots (compact).
  int slot_of_[kLanes];
  int lane_at_[kLanes];
  // One past the highest live slot, rounded up to a whole 16-slot line: the
  // step kernels' lane bound, so at most four graph variants per parity.
  Int span_ = kLanes;
  // Waves per workgroup, rows per wave and row layout (the fixed-width copy
  // or the CSR) of the primal (K^T) and dual (K) launches.
  int pw_ = 2, dw_ = 2;
  int pr_ = 1, dr_ = 1;
  bool ps_ = false, ds_ = false;
  // Waves per workgroup of the binned launches of narrow spans.
  int pnw_ = 1, dnw_ = 1;
  // Narrow spans (32 live slots or fewer) run the binned step kernels.
  bool pack_ = true;

  double* slab_ = nullptr;
  double *X_[2], *Y_[2], *KX_[2];
  double *XBAR_, *DX_, *AX_, *KTY_, *LAM_, *RX_, *PRX_;
  double *DY_, *KDX_, *AY_, *AKX_, *RY_, *PRY_;
  double *EXT_, *CHECK_, *RESTART_, *CHECKP_;
  LaneParams* d_params_ = nullptr;
  BatchData data_;
  double* pinned_ = nullptr;  // params + check + restart staging, then the publish flag
  Int* flag_ = nullptr;       // the last period whose check reached pinned_
  Int seq_ = 0;               // the last period staged
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs_;
  std::vector<PendingPolish> pending_;
  Mask polishing_ = 0;  // lanes with an attempt in flight
  Mask settled_ = 0;    // lanes whose outcome a successful attempt fixed
};

Engine::~Engine() {
  for (auto& g : graphs_) {
    (void)hipGraphExecDestroy(g.second.second);
    (void)hipGraphDestroy(g.second.first);
  }
  // The slab and the pinned staging are the backend's, per nesting depth,
  // kept for the next batch of the same depth.
  if (slab_) (void)hipStreamSynchronize(be_.stream());
}

constexpr size_t kPinnedDoubles = 64 * 8 + 64 * kLaneCheckStride + 128 + 8;
static_assert(sizeof(LaneTable) <= 64 * 8 * sizeof(double), "the lane table fits its slab entry");

void Engine::allocate() {
  const size_t N = static_cast<size_t>(n_) * kLanes, M = static_cast<size_t>(m_) * kLanes;
  auto round = [](size_t d) { return (std::max<size_t>(d, 1) + 31) / 32 * 32; };
  const size_t ext = static_cast<size_t>(std::max(n_, m_)) * kLanes;
  const size_t sizes[] = {
      N, N, N, N, N, N, N, N, N, N,            // X0 X1 XBAR DX AX KTY LAM RX PRX c
      N, N, N, N, N,                           // lower upper oc olo ohi
      M, M, M, M, M, M, M, M, M, M, M, M,      // Y0 Y1 KX0 KX1 DY KDX AY AKX RY PRY q oq
      static_cast<size_t>(n_), static_cast<size_t>(m_),  // col_scale row_scale
      ext, 64 * kLaneCheckStride, 128, 64 * 8,           // EXT CHECK RESTART params
      batch_check_scratch_doubles()};                    // check partials
  size_t total = 0;
  for (size_t s : sizes) total += round(s);
  slab_ = be_.scratch(kBatchScratchSlot + depth_, total);
  hip_check(hipMemsetAsync(slab_, 0, total * sizeof(double), be_.stream()), "hipMemsetAsync(batch)");
  double* p = slab_;
  size_t i = 0;
  auto take = [&]() {
    double* q = p;
    p += round(sizes[i++]);
    return q;
  };
  X_[0] = take(); X_[1] = take(); XBAR_ = take(); DX_ = take(); AX_ = take(); KTY_ = take();
  LAM_ = take(); RX_ = take(); PRX_ = take();
  double* c = take(); double* lo = take(); double* hi = take();
  double* oc = take(); double* olo = take(); double* ohi = take();
  Y_[0] = take(); Y_[1] = take(); KX_[0] = take(); KX_[1] = take(); DY_ = take(); KDX_ = take();
  AY_ = take(); AKX_ = take(); RY_ = take(); PRY_ = take();
  double* q = take(); double* oq = take();
  double* cs = take(); double* rs = take();
  EXT_ = take(); CHECK_ = take(); RESTART_ = take();
  d_params_ = reinterpret_cast<LaneParams*>(take());
  CHECKP_ = take();
  data_.n = n_;
  data_.m = m_;
  data_.num_equalities = neq_;
  data_.c = c; data_.lower = lo; data_.upper = hi; data_.q = q;
  data_.orig_c = oc; data_.orig_lower = olo; data_.orig_upper = ohi; data_.orig_q = oq;
  data_.row_scale = rs; data_.col_scale = cs;
  pinned_ = be_.pinned_slot(kBatchScratchSlot + depth_, kPinnedDoubles);
  flag_ = reinterpret_cast<Int*>(pinned_ + kPinnedDoubles - 8);
  __atomic_store_n(flag_, Int{-1}, __ATOMIC_RELEASE);
  seq_ = 0;
}

// Row-interleaves one vector per lane (rows entries each) into `out`, whose
// rows are `width` slots apart: filled row by row (the padding slots up to
// `width` zeroed), so the writes stream instead of striding across the
// buffer.
void interleave(double* out, Int rows, int count, int width, const std::vector<const double*>& src) {
  constexpr Int kRowsPerTask = 1024;
  const Int tasks = (rows + kRowsPerTask - 1) / kRowsPerTask;
  HostPool::shared().run(uz(tasks), [&](size_t t) {
    const Int r1 = std::min<Int>(rows, static_cast<Int>(t + 1) * kRowsPerTask);
    for (Int r = static_cast<Int>(t) * kRowsPerTask; r < r1; ++r) {
      double* line = out + static_cast<size_t>(r) * static_cast<size_t>(width);
      for (int l = 0; l < count; ++l) line[l] = src[static_cast<size_t>(l)][r];
      for (int l = count; l < width; ++l) line[l] = 0.0;
    }
  });
}

void Engine::upload_lanes() {
  // Starting points: a warm start is put in the dual cone and the primal box,
  // a cold one is the projection of zero (pdhg.cpp:479-503).
  std::vector<std::vector<double>> xs(static_cast<size_t>(count_)), ys(static_cast<size_t>(count_));
  HostPool::shared().run(static_cast<size_t>(count_), [&](size_t li) {
    const int l = static_cast<int>(li);
    Lane& ln = lanes_[static_cast<size_t>(l)];
    const bool warm = !ln.warm_x.empty() && !ln.warm_y.empty();
    std::vector<double>& x = xs[static_cast<size_t>(l)];
    std::vector<double>& y = ys[static_cast<size_t>(l)];
    x.assign(uz(n_), 0.0);
    y.assign(uz(m_), 0.0);
    if (warm) {
      x = ln.warm_x;
      y = ln.warm_y;
      for (Int i = neq_; i < m_; ++i)
        if (!(y[uz(i)] > 0.0)) y[uz(i)] = 0.0;
    }
    for (Int j = 0; j < n_; ++j) {
      const double lo = ln.lower[uz(j)], hi = ln.upper[uz(j)];
      double v = warm ? x[uz(j)] : 0.0;
      if (!std::isfinite(v)) v = 0.0;
      if (v < lo) v = std::isfinite(lo) ? lo : 0.0;
      if (v > hi) v = std::isfinite(hi) ? hi : 0.0;
      x[uz(j)] = v;
    }
  });
  // Every lane vector is interleaved straight into one pinned staging area
  // and copied asynchronously; one wait covers them all. Only the slots up to
  // the lane count rounded to 16 are staged and sent, as a strided copy into
  // the rows' 512-byte lines: the slab is zeroed, so the slots above stay
  // the zero padding (a small batch, such as a polishing sub-batch, then
  // moves a fraction of the bytes).
  struct Item {
    double* dst;
    Int rows;
    std::vector<const double*> src;
  };
  auto field = [&](std::vector<double> Lane::*f) {
    std::vector<const double*> v;
    for (Lane& l : lanes_) v.push_back((l.*f).data());
    return v;
  };
  auto own = [&](std::vector<std::vector<double>>& vs) {
    std::vector<const double*> v;
    for (auto& x : vs) v.push_back(x.data());
    return v;
  };
  const Item items[] = {
      {const_cast<double*>(data_.c), n_, field(&Lane::c)},
      {const_cast<double*>(data_.lower), n_, field(&Lane::lower)},
      {const_cast<double*>(data_.upper), n_, field(&Lane::upper)},
      {const_cast<double*>(data_.orig_c), n_, field(&Lane::oc)},
      {const_cast<double*>(data_.orig_lower), n_, field(&Lane::olo)},
      {const_cast<double*>(data_.orig_upper), n_, field(&Lane::ohi)},
      {const_cast<double*>(data_.q), m_, field(&Lane::q)},
      {const_cast<double*>(data_.orig_q), m_, field(&Lane::oq)},
      {X_[0], n_, own(xs)},
      {Y_[0], m_, own(ys)},
  };
  const int width = std::min(kLanes, (count_ + 15) / 16 * 16);
  size_t total = static_cast<size_t>(n_) + static_cast<size_t>(m_);  // the shared diagonals
  for (const Item& it : items) total += static_cast<size_t>(it.rows) * static_cast<size_t>(width);
  double* stage = be_.pinned_scratch(total);
  hipStream_t s = be_.stream();
  size_t off = 0;
  auto send = [&](double* dst, size_t count) {
    if (count > 0)
      hip_check(hipMemcpyAsync(dst, stage + off, count * sizeof(double), hipMemcpyHostToDevice, s),
                "upload lanes");
    off += count;
  };
  for (const Item& it : items) {
    interleave(stage + off, it.rows, count_, width, it.src);
    if (width == kLanes) {
      send(it.dst, static_cast<size_t>(it.rows) * kLanes);
    } else if (it.rows > 0) {
      const size_t line = static_cast<size_t>(width) * sizeof(double);
      hip_check(hipMemcpy2DAsync(it.dst, kLanes * sizeof(double), stage + off, line, line,
                                 static_cast<size_t>(it.rows), hipMemcpyHostToDevice, s),
                "upload lanes");
      off += static_cast<size_t>(it.rows) * static_cast<size_t>(width);
    }
  }
  std::copy(col_scale_.begin(), col_scale_.end(), stage + off);
  send(const_cast<double*>(data_.col_scale), uz(n_));
  std::copy(row_scale_.begin(), row_scale_.end(), stage + off);
  send(const_cast<double*>(data_.row_scale), uz(m_));
  be_.synchronize();
  // The first restart point (and the one before it) is the starting point.
  be_.copy(X_[0], RX_, n_ * kLanes);
  be_.copy(X_[0], PRX_, n_ * kLanes);
  be_.copy(Y_[0], RY_, m_ * kLanes);
  be_.copy(Y_[0], PRY_, m_ * kLanes);
}

void Engine::upload_params(bool staged_only) {
  auto* T = reinterpret_cast<LaneTable*>(pinned_);
  T->seq = ++seq_;
  LaneParams* P = T->lane;
  T->span = 0;
  for (int slot = 0; slot < kLanes; ++slot) {
    LaneParams lp{};
    const int l = lane_at_[slot];
    if (l >= 0 && ((active_ >> l) & 1ull)) {
      T->span = slot + 1;
      const IterParams ip = ctl_[static_cast<size_t>(l)].period_params(neq_);
      lp.tau = ip.tau;
      lp.sigma = ip.sigma;
      lp.gamma = ip.gamma;
      lp.halpern = ip.halpern;
      lp.reset_at_start = ip.reset_at_start;
      lp.k_at_start = ip.k_at_start;
      lp.active = 1;
    }
    P[slot] = lp;
  }
  span_ = std::max<Int>(16, (T->span + 15) / 16 * 16);
  if (staged_only) return;
  hip_check(hipMemcpyAsync(d_params_, T, sizeof(LaneTable), hipMemcpyHostToDevice, be_.stream()),
            "upload lane params");
}

void Engine::enqueue_period(int steps, bool has_check, int parity, bool host_io) {
  hipStream_t s = be_.stream();
  if (host_io) launch_batch_params(s, reinterpret_cast<const LaneTable*>(pinned_), d_params_);
  int p = parity;
  // Algorithmic bytes per launch: the matrix once (8 B value + 4 B index per
  // nonzero, 4 B per row start), the gathered vector once per active LP, and
  // the seven per-LP vectors the epilogue reads and writes (nine or ten on the
  // check step), 8 B each.
  const double lanes = static_cast<double>(__builtin_popcountll(active_));
  const double nnz = static_cast<double>(k_.nnz()), n = n_, m = m_;
  for (int step = 0; step < steps; ++step) {
    const bool diag = has_check && step == steps - 1;
    PrimalBufs pb{X_[p], AX_, X_[1 - p], XBAR_, DX_, AX_};
    const LaunchEvents pe =
        be_.launch_events(diag ? "batch_primal_step_check" : "batch_primal_step",
                          12.0 * nnz + 4.0 * (n + 1) + 8.0 * lanes * (m + (diag ? 8.0 : 7.0) * n));
    launch_batch_primal_step(s, ps_ ? dkts_ : dkt_, Y_[p], data_, d_params_, step, span_, pb, diag,
                             pw_, pr_, pack_, &pe, pnw_);
    DualBufs db{Y_[p], KX_[p], AY_, AKX_, Y_[1 - p], KX_[1 - p], DY_, KDX_, AY_, AKX_};
    const LaunchEvents de =
        be_.launch_events(diag ? "batch_dual_step_check" : "batch_dual_step",
                          12.0 * nnz + 4.0 * (m + 1) + 8.0 * lanes * (n + (diag ? 9.0 : 7.0) * m));
    launch_batch_dual_step(s, ds_ ? dks_ : dk_, XBAR_, data_, d_params_, step, span_, db, diag, dw_,
                           dr_, pack_, &de, dnw_);
    p = 1 - p;
  }
  if (has_check) {
    const bool farkas = opt_.detect_infeasibility != 0;
    HipBackend::Timed t(be_, "batch_check");
    launch_batch_spmm(s, dkt_, Y_[p], KTY_, farkas ? DY_ : nullptr, farkas ? LAM_ : nullptr,
                      d_params_, ~0ull, pw_);
    launch_batch_check(s, data_, X_[p], Y_[p], KX_[p], DX_, DY_, KTY_, LAM_, farkas, d_params_,
                       CHECKP_, CHECK_);
    if (host_io) launch_batch_publish(s, CHECK_, d_params_, pinned_ + 64 * 8, flag_);
  }
}

hipGraphExec_t Engine::graph_for(int parity) {
  // One graph per parity and span (the span is a kernel argument).
  const int key = parity * 8 + span_ / 16;
  auto it = graphs_.find(key);
  if (it != graphs_.end()) return it->second.second;
  hipStream_t s = be_.stream();
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    enqueue_period(opt_.termination_check_frequency, true, parity, true);
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  graphs_[key] = {g, e};
  return e;
}

void Engine::wait_published() {
  // Polling the flag sees the check about as soon as it lands; a period that
  // takes unusually long (or a fault, which leaves the flag unset) falls back
  // to the stream's own wait, which reports a device error.
  const Int want = seq_;
  const Clock::time_point t0 = Clock::now();
  for (unsigned spin = 1;; ++spin) {
    if (__atomic_load_n(flag_, __ATOMIC_ACQUIRE) == want) return;
    if ((spin & 1023u) == 0 && Clock::now() - t0 > std::chrono::milliseconds(50)) break;
  }
  be_.synchronize();
  if (__atomic_load_n(flag_, __ATOMIC_ACQUIRE) != want)
    throw std::runtime_error("batch: a period finished without publishing its check");
}

void Engine::fetch(const double* dev, double* host, size_t doubles) {
  hip_check(hipMemcpyAsync(host, dev, doubles * sizeof(double), hipMemcpyDeviceToHost, be_.stream()),
            "batch fetch");
  be_.synchronize();
}

// Copies the current iterate of the masked lanes out to their outcomes.
void Engine::retire_extract(Mask mask, std::vector<LaneOutcome>& out) {
  if (mask == 0) return;
  const int cnt = __builtin_popcountll(mask);
  const Mask dev = slots(mask);
  std::vector<double> xs(static_cast<size_t>(cnt) * uz(n_)), ys(static_cast<size_t>(cnt) * uz(m_));
  launch_batch_extract(be_.stream(), n_, X_[cur_], dev, EXT_);
  be_.download(EXT_, xs.data(), cnt * n_);
  launch_batch_extract(be_.stream(), m_, Y_[cur_], dev, EXT_);
  be_.download(EXT_, ys.data(), cnt * m_);
  // The extract is in slot order.
  int k = 0;
  for (int slot = 0; slot < kLanes; ++slot) {
    if (!((dev >> slot) & 1ull)) continue;
    LaneOutcome& o = out[static_cast<size_t>(lane_at_[slot])];
    if (!o.polished) {
      o.x.assign(xs.begin() + static_cast<long>(k) * n_, xs.begin() + static_cast<long>(k + 1) * n_);
      o.y.assign(ys.begin() + static_cast<long>(k) * m_, ys.begin() + static_cast<long>(k + 1) * m_);
    }
    ++k;
  }
}

// A restart for the masked lanes (pdhg.cpp:1124-1202): K x recomputed, the
// distances moved since the last restart (for the primal weight) measured
// lane-serially, the restart points advanced.
void Engine::run_restarts(Mask mask) {
  if (mask == 0) return;
  const Mask dev = slots(mask);
  launch_batch_spmm(be_.stream(), dk_, X_[cur_], KX_[cur_], nullptr, nullptr, nullptr, dev, dw_);
  const bool previous = !opt_.pid_primal_weight;
  launch_batch_restart(be_.stream(), n_, m_, X_[cur_], Y_[cur_], RX_, RY_, PRX_, PRY_, previous,
                       true, dev, CHECKP_, RESTART_);
  double* h = pinned_ + 64 * 8 + 64 * kLaneCheckStride;
  fetch(RESTART_, h, 128);
  for (int l = 0; l < count_; ++l) {
    if (!((mask >> l) & 1ull)) continue;
    const int slot = slot_of_[l];
    ctl_[static_cast<size_t>(l)].restart(std::sqrt(h[2 * slot]), std::sqrt(h[2 * slot + 1]));
  }
}

std::vector<LaneOutcome> run_engine(HipBackend& be, const Csr& k, const Csr& kt, Int neq,
                                    std::vector<double> row_scale, std::vector<double> col_scale,
                                    double norm, const anuk_pdhg_options& opt,
                                    std::vector<Lane>& lanes, Clock::time_point start,
                                    const PolishContext* polish, Phases* phases = nullptr,
                                    int depth = 1) {
  Engine e(be, k, kt, neq, std::move(row_scale), std::move(col_scale), norm, opt, lanes, start,
           polish, phases, depth);
  return e.run();
}


// The two feasibility sub-problems of each seed, as one sub-batch each
// (pdhg.cpp:680-743): primal with c = 0 started from (x, 0), dual with q = 0
// and the finite bounds zeroed started from (0, y), both with the caller's
// step size and each LP's primal weight, no scaling (the seeds are already in
// the scaled space), then unscaled with the stack's diagonals.
std::vector<Polished> polish_lanes(HipBackend& be, const Csr& k, const Csr& kt, Int neq,
                                   double norm, const anuk_pdhg_options& opt,
                                   const std::vector<const Lane*>& src,
                                   const std::vector<std::vector<double>>& seed_x,
                                   const std::vector<std::vector<double>>& seed_y,
                                   const std::vector<double>& omega, double eta, Int budget,
                                   const Scaling& scaling, Clock::time_point start) {
  const Int n = k.cols, m = k.rows;
  anuk_pdhg_options sub = opt;
  sub.polish_feasibility = 0;
  sub.polish_on_exit = 0;
  sub.detect_infeasibility = 0;
  sub.verbose = 0;
  sub.known_matrix_norm = norm;
  sub.ruiz_iterations = 0;
  sub.pock_chambolle = 0;
  sub.max_iterations = budget;
  sub.termination_check_frequency =
      std::max<Int>(10, std::min<Int>(opt.termination_check_frequency, budget / 4));
  sub.time_limit_seconds = std::fmax(0.0, opt.time_limit_seconds - seconds_since(start));
  sub.initial_step_size = eta;
  sub.warm_x = nullptr;
  sub.warm_y = nullptr;
  const size_t cnt = src.size();
  std::vector<Lane> primal(cnt), dual(cnt);
  HostPool::shared().run(cnt, [&](size_t s) {
    Lane& p = primal[s];
    p.c.assign(uz(n), 0.0);
    p.q = src[s]->q;
    p.lower = src[s]->lower;
    p.upper = src[s]->upper;
    p.warm_x = seed_x[s];
    p.warm_y.assign(uz(m), 0.0);
    p.initial_omega = omega[s];
    Lane& d = dual[s];
    d.c = src[s]->c;
    d.q.assign(uz(m), 0.0);
    d.lower = src[s]->lower;
    d.upper = src[s]->upper;
    for (Int j = 0; j < n; ++j) {
      if (d.lower[uz(j)] > -kInf) d.lower[uz(j)] = 0.0;
      if (d.upper[uz(j)] < kInf) d.upper[uz(j)] = 0.0;
    }
    d.warm_x.assign(uz(n), 0.0);
    d.warm_y = seed_y[s];
    d.initial_omega = omega[s];
    for (Lane* t : {&p, &d}) {
      t->oc = t->c;
      t->oq = t->q;
      t->olo = t->lower;
      t->ohi = t->upper;
    }
  });
  const std::vector<double> ones_m(uz(m), 1.0), ones_n(uz(n), 1.0);
  std::vector<LaneOutcome> pr, dr;
  if (2 * cnt <= static_cast<size_t>(kLanes)) {
    // Both feasibility problems of every seed share K: one sub-batch, each
    // lane with its own termination test.
    std::vector<Lane> both;
    for (Lane& p : primal) {
      p.termination = ANUK_TERM_PRIMAL_FEASIBILITY;
      both.push_back(std::move(p));
    }
    for (Lane& d : dual) {
      d.termination = ANUK_TERM_DUAL_FEASIBILITY;
      both.push_back(std::move(d));
    }
    std::vector<LaneOutcome> out =
        run_engine(be, k, kt, neq, ones_m, ones_n, norm, sub, both, start, nullptr);
    pr.assign(std::make_move_iterator(out.begin()), std::make_move_iterator(out.begin() + static_cast<long>(cnt)));
    dr.assign(std::make_move_iterator(out.begin() + static_cast<long>(cnt)), std::make_move_iterator(out.end()));
  } else {
    sub.termination = ANUK_TERM_PRIMAL_FEASIBILITY;
    pr = run_engine(be, k, kt, neq, ones_m, ones_n, norm, sub, primal, start, nullptr);
    sub.termination = ANUK_TERM_DUAL_FEASIBILITY;
    dr = run_engine(be, k, kt, neq, ones_m, ones_n, norm, sub, dual, start, nullptr);
  }
  std::vector<Polished> out(cnt);
  HostPool::shared().run(cnt, [&](size_t s) {
    scaling.unscale_primal(pr[s].x, &out[s].x);
    scaling.unscale_dual(dr[s].y, &out[s].y);
    out[s].spent = pr[s].iterations + dr[s].iterations;
  });
  return out;
}

// A polish attempt for the lanes `who` at a check: adopted, and the lane
// stopped as optimal, when the polished pair passes the full test on the
// lane's original LP.
void Engine::start_polish(const std::vector<int>& who, Int budget, Int iteration) {
  // The extract comes out in slot order, so the lanes are taken in that order.
  std::vector<int> order(who);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return slot_of_[a] < slot_of_[b]; });
  Mask mask = 0;
  for (int l : order) mask |= 1ull << l;
  const Mask dev = slots(mask);
  const int cnt = static_cast<int>(order.size());
  std::vector<double> xs(static_cast<size_t>(cnt) * uz(n_)), ys(static_cast<size_t>(cnt) * uz(m_));
  launch_batch_extract(be_.stream(), n_, X_[cur_], dev, EXT_);
  be_.download(EXT_, xs.data(), cnt * n_);
  launch_batch_extract(be_.stream(), m_, Y_[cur_], dev, EXT_);
  be_.download(EXT_, ys.data(), cnt * m_);
  std::vector<const Lane*> src;
  std::vector<std::vector<double>> sx, sy;
  std::vector<double> omega;
  PendingPolish pp;
  pp.lanes = order;
  pp.iteration = iteration;
  for (int s = 0; s < cnt; ++s) {
    const int l = order[static_cast<size_t>(s)];
    src.push_back(&lanes_[static_cast<size_t>(l)]);
    sx.emplace_back(xs.begin() + static_cast<long>(s) * n_, xs.begin() + static_cast<long>(s + 1) * n_);
    sy.emplace_back(ys.begin() + static_cast<long>(s) * m_, ys.begin() + static_cast<long>(s + 1) * m_);
    omega.push_back(ctl_[static_cast<size_t>(l)].omega);
    pp.at_check.push_back(ctl_[static_cast<size_t>(l)]);
    polishing_ |= 1ull << l;
  }
  const double eta = ctl_[static_cast<size_t>(who[0])].eta;
  HipBackend* helper = &be_.helper();
  const int device = be_.device();
  // Everything the task reads is the engine's and outlives it: the engine
  // waits for every attempt before it returns (and its futures wait in its
  // destructor).
  pp.result = std::async(std::launch::async, [this, helper, device, eta, budget,
                                              src = std::move(src), sx = std::move(sx),
                                              sy = std::move(sy), omega = std::move(omega)]() {
    hip_check(hipSetDevice(device), "hipSetDevice(polish)");
    return polish_lanes(*helper, k_, kt_, neq_, norm_, opt_, src, sx, sy, omega, eta, budget,
                        *polish_->scaling, start_);
  });
  pending_.push_back(std::move(pp));
}

void Engine::apply_polish(PendingPolish& pp, std::vector<LaneOutcome>& out) {
  std::vector<Polished> pol = pp.result.get();
  const size_t cnt = pp.lanes.size();
  std::vector<Residual> trs(cnt);
  HostPool::shared().run(cnt, [&](size_t s) {
    const int l = pp.lanes[s];
    trs[s] = evaluate_residual(*(*polish_->originals)[static_cast<size_t>(l)], pol[s].x, pol[s].y);
  });
  for (size_t s = 0; s < cnt; ++s) {
    const int l = pp.lanes[s];
    const Mask bit = 1ull << l;
    LaneOutcome& o = out[static_cast<size_t>(l)];
    ++o.polish_attempts;
    o.polish_iterations += pol[s].spent;
    polishing_ &= ~bit;
    const Control& c = pp.at_check[s];
    const Residual& tr = trs[s];
    if (!tr.converged(c.tol, opt_.absolute_tolerance, opt_.require_inf_norm_termination != 0,
                      c.gap_tol))
      continue;
    // Converged: the lane's outcome is the polished pair as of the check that
    // started the attempt, whatever the lane did since.
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
    settled_ |= bit;
    active_ &= ~bit;
  }
}

void Engine::settle_polish(std::vector<LaneOutcome>& out, bool wait) {
  for (auto it = pending_.begin(); it != pending_.end();) {
    if (wait || it->result.wait_for(std::chrono::seconds(0)) == std::future_status::ready) {
      apply_polish(*it, out);
      it = pending_.erase(it);
    } else {
      ++it;
    }
  }
}

// A key for a matrix's sparsity pattern (the launch shapes depend on nothing
// else): its size and an FNV-1a hash of the row starts and column indices.
std::uint64_t pattern_key(const Csr& a) {
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&](std::uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  mix(static_cast<std::uint64_t>(a.rows));
  mix(static_cast<std::uint64_t>(a.cols));
  for (Int v : a.start) mix(static_cast<std::uint64_t>(static_cast<std::uint32_t>(v)));
  for (Int v : a.index) mix(static_cast<std::uint64_t>(static_cast<std::uint32_t>(v)));
  return h;
}

// The launch-shape tuner: times the primal and dual step kernels in every
// shape (1, 2, 4 or 8 waves per workgroup; 1, 2 or 4 consecutive rows per
// wave from the CSR, or one row per wave from the fixed-width row copy) on
// this stack's K^T and K, and the binned launches of a narrow
// span (16 slots) with 1, 2, 4 or 8 waves per workgroup, and keeps, per
// sparsity pattern for the context's life, the fastest of each, if it beats
// the default by at least kTuneMargin;
// otherwise the default stays, so timing noise between near-equal shapes
// does not flip the choice. A shape's
// time is the best of kTrials means of kReps back-to-back launches. Every
// shape computes the same bits, so the choice changes only the time. The
// trial launches read the current iterate and write only buffers the first
// real step overwrites (the next parity, x-bar), with a step index that
// opens no Halpern epoch.
namespace {
constexpr double kTuneMargin = 0.03;
// A shape is 100 * layout + 10 * waves + rows per wave, layout 1 for the
// fixed-width row copy and 0 for the CSR; a narrow-span shape is 300 + 10 *
// waves + 1. One wave per workgroup measured fastest for the narrow
// launches (a 2% shorter solve of the 24-period stack than two waves; four
// and eight were slower still).
constexpr int kDefaultShape = 21;
constexpr int kDefaultNarrow = 311;
constexpr Int kNarrowSpan = 16;
int shape_waves(int shape) { return (shape % 100) / 10; }
int shape_rows(int shape) { return shape % 10; }
bool shape_slots(int shape) { return shape >= 100; }
}  // namespace

void Engine::tune() {
  pnw_ = dnw_ = shape_waves(kDefaultNarrow);
  if (opt_.batch_waves != 0) pnw_ = dnw_ = opt_.batch_waves;
  if (opt_.batch_waves != 0 || opt_.batch_rows != 0 || opt_.batch_layout != 0) {
    pw_ = dw_ = opt_.batch_waves != 0 ? opt_.batch_waves : shape_waves(kDefaultShape);
    pr_ = dr_ = opt_.batch_rows != 0 ? opt_.batch_rows : shape_rows(kDefaultShape);
    ps_ = ds_ = opt_.batch_layout != 0 ? opt_.batch_layout == 2 : shape_slots(kDefaultShape);
    return;
  }
  pw_ = dw_ = shape_waves(kDefaultShape);
  pr_ = dr_ = shape_rows(kDefaultShape);
  ps_ = ds_ = shape_slots(kDefaultShape);
  const std::uint64_t kp = pattern_key(k_), ktp = pattern_key(kt_);
  const HipBackend::Tuning* tp = be_.tuned("batch_primal_step", ktp);
  const HipBackend::Tuning* td = be_.tuned("batch_dual_step", kp);
  const HipBackend::Tuning* tpn = be_.tuned("batch_primal_step_narrow", ktp);
  const HipBackend::Tuning* tdn = be_.tuned("batch_dual_step_narrow", kp);
  if (tp != nullptr && td != nullptr && tpn != nullptr && tdn != nullptr) {
    pw_ = shape_waves(tp->chosen);
    pr_ = shape_rows(tp->chosen);
    ps_ = shape_slots(tp->chosen);
    dw_ = shape_waves(td->chosen);
    dr_ = shape_rows(td->chosen);
    ds_ = shape_slots(td->chosen);
    pnw_ = shape_waves(tpn->chosen);
    dnw_ = shape_waves(tdn->chosen);
    return;
  }
  if (opt_.autotune == 0 || depth_ != 0 || be_.profiling()) return;
  upload_params();
  hipStream_t s = be_.stream();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hip_check(hipEventCreate(&e0), "hipEventCreate(tune)");
  hip_check(hipEventCreate(&e1), "hipEventCreate(tune)");
  // Each trial times every candidate once (kReps back-to-back launches), so
  // the trials of one shape are spread over the whole tuning run; a shape's
  // time is its best trial. (A shape timed in one burst had read 30% slow in
  // some runs and fast in others.)
  constexpr int kReps = 20, kTrials = 5;
  const PrimalBufs pb{X_[cur_], AX_, X_[1 - cur_], XBAR_, DX_, AX_};
  const DualBufs db{Y_[cur_], KX_[cur_], AY_, AKX_, Y_[1 - cur_], KX_[1 - cur_], DY_, KDX_, AY_, AKX_};
  HipBackend::Tuning p{"batch_primal_step", kt_.rows, kt_.nnz(), {}, kDefaultShape};
  HipBackend::Tuning d{"batch_dual_step", k_.rows, k_.nnz(), {}, kDefaultShape};
  HipBackend::Tuning pn{"batch_primal_step_narrow", kt_.rows, kt_.nnz(), {}, kDefaultNarrow};
  HipBackend::Tuning dn{"batch_dual_step_narrow", k_.rows, k_.nnz(), {}, kDefaultNarrow};
  struct Candidate {
    HipBackend::Tuning* side;
    int shape;
    std::function<void()> launch;
  };
  std::vector<Candidate> cands;
  for (int w : {1, 2, 4, 8}) {
    for (int r : {1, 2, 4}) {
      cands.push_back({&p, w * 10 + r, [&, w, r] {
                         launch_batch_primal_step(s, dkt_, Y_[cur_], data_, d_params_, 1, span_,
                                                  pb, false, w, r);
                       }});
      cands.push_back({&d, w * 10 + r, [&, w, r] {
                         launch_batch_dual_step(s, dk_, XBAR_, data_, d_params_, 1, span_, db,
                                                false, w, r);
                       }});
    }
    cands.push_back({&p, 100 + w * 10 + 1, [&, w] {
                       launch_batch_primal_step(s, dkts_, Y_[cur_], data_, d_params_, 1, span_, pb,
                                                false, w, 1);
                     }});
    cands.push_back({&d, 100 + w * 10 + 1, [&, w] {
                       launch_batch_dual_step(s, dks_, XBAR_, data_, d_params_, 1, span_, db,
                                              false, w, 1);
                     }});
    cands.push_back({&pn, 300 + w * 10 + 1, [&, w] {
                       launch_batch_primal_step(s, dkt_, Y_[cur_], data_, d_params_, 1, kNarrowSpan,
                                                pb, false, w, 1, true, nullptr, w);
                     }});
    cands.push_back({&dn, 300 + w * 10 + 1, [&, w] {
                       launch_batch_dual_step(s, dk_, XBAR_, data_, d_params_, 1, kNarrowSpan, db,
                                              false, w, 1, true, nullptr, w);
                     }});
  }
  std::vector<double> best(cands.size(), 1e300);
  for (int t = 0; t < kTrials; ++t) {
    for (size_t c = 0; c < cands.size(); ++c) {
      hip_check(hipEventRecord(e0, s), "hipEventRecord(tune)");
      for (int r = 0; r < kReps; ++r) cands[c].launch();
      hip_check(hipEventRecord(e1, s), "hipEventRecord(tune)");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize(tune)");
      float ms = 0.0f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime(tune)");
      best[c] = std::min(best[c], 1e3 * static_cast<double>(ms) / kReps);
    }
  }
  for (size_t c = 0; c < cands.size(); ++c) cands[c].side->us.push_back({cands[c].shape, best[c]});
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  auto pick = [](const HipBackend::Tuning& t, int fallback) {
    const auto best = std::min_element(t.us.begin(), t.us.end(), [](const auto& a, const auto& b) {
      return a.second < b.second;
    });
    double base = best->second;
    for (const auto& c : t.us)
      if (c.first == fallback) base = c.second;
    return best->second < (1.0 - kTuneMargin) * base ? best->first : fallback;
  };
  p.chosen = pick(p, kDefaultShape);
  d.chosen = pick(d, kDefaultShape);
  pn.chosen = pick(pn, kDefaultNarrow);
  dn.chosen = pick(dn, kDefaultNarrow);
  pnw_ = shape_waves(pn.chosen);
  dnw_ = shape_waves(dn.chosen);
  pw_ = shape_waves(p.chosen);
  pr_ = shape_rows(p.chosen);
  ps_ = shape_slots(p.chosen);
  dw_ = shape_waves(d.chosen);
  dr_ = shape_rows(d.chosen);
  ds_ = shape_slots(d.chosen);
  be_.record_tuning(ktp, std::move(p));
  be_.record_tuning(kp, std::move(d));
  be_.record_tuning(ktp, std::move(pn));
  be_.record_tuning(kp, std::move(dn));
}

// Moves the live lanes into the lowest slots, in lane order, when that frees
// a whole 16-slot (128-byte) line of every row: the step kernels then touch
// ceil(live / 16) lines per row instead of the span of the scattered slots.
// Only the state that outlives a period moves; a lane's arithmetic does not
// depend on its slot, so the results are unchanged.
void Engine::compact() {
  const int live = __builtin_popcountll(active_);
  if (live == 0) return;
  int top = -1;
  for (int l = 0; l < count_; ++l)
    if ((active_ >> l) & 1ull) top = std::max(top, slot_of_[l]);
  if ((live + 15) / 16 >= (top + 1 + 15) / 16) return;
  SlotMap map{};
  map.count = live;
  int next = 0;
  int lane_at[kLanes];
  std::fill(lane_at, lane_at + kLanes, -1);
  for (int l = 0; l < count_; ++l) {
    if (!((active_ >> l) & 1ull)) continue;
    map.from[next] = static_cast<std::int8_t>(slot_of_[l]);
    lane_at[next] = l;
    ++next;
  }
  hipStream_t st = be_.stream();
  for (double* v : {X_[cur_], AX_, RX_, PRX_, const_cast<double*>(data_.c),
                    const_cast<double*>(data_.lower), const_cast<double*>(data_.upper),
                    const_cast<double*>(data_.orig_c), const_cast<double*>(data_.orig_lower),
                    const_cast<double*>(data_.orig_upper)})
    launch_batch_permute(st, n_, v, EXT_, map);
  for (double* v : {Y_[cur_], KX_[cur_], AY_, AKX_, RY_, PRY_, const_cast<double*>(data_.q),
                    const_cast<double*>(data_.orig_q)})
    launch_batch_permute(st, m_, v, EXT_, map);
  // Retired lanes keep no slot; their outcomes are already extracted.
  for (int l = 0; l < count_; ++l) slot_of_[l] = -1;
  std::copy(lane_at, lane_at + kLanes, lane_at_);
  for (int slot = 0; slot < kLanes; ++slot)
    if (lane_at_[slot] >= 0) slot_of_[lane_at_[slot]] = slot;
}

std::vector<LaneOutcome> Engine::run() {
  Clock::time_point mark = Clock::now();
  count_ = static_cast<int>(lanes_.size());
  if (count_ <= 0 || count_ > kLanes) throw std::invalid_argument("a batch holds 1 to 64 LPs");
  n_ = k_.cols;
  m_ = k_.rows;
  dk_ = be_.prepare_bins(view_of(k_));
  dkt_ = be_.prepare_bins(view_of(kt_));
  dks_ = be_.prepare_slots(view_of(k_));
  dkts_ = be_.prepare_slots(view_of(kt_));
  allocate();
  for (Lane& l : lanes_) l.set_scales();
  ctl_.resize(static_cast<size_t>(count_));
  lane_opt_.assign(static_cast<size_t>(count_), opt_);
  for (int l = 0; l < count_; ++l) {
    const Lane& ln = lanes_[static_cast<size_t>(l)];
    anuk_pdhg_options& o = lane_opt_[static_cast<size_t>(l)];
    if (ln.termination >= 0) o.termination = ln.termination;
    Control& c = ctl_[static_cast<size_t>(l)];
    c.init(o, norm_, host_two_norm(ln.c), host_two_norm(ln.q));
    if (ln.initial_omega > 0.0) c.omega = ln.initial_omega;
  }
  for (int l = 0; l < kLanes; ++l) {
    slot_of_[l] = l;
    lane_at_[l] = l < count_ ? l : -1;
  }
  upload_lanes();
  active_ = (count_ == 64) ? ~0ull : ((1ull << count_) - 1ull);
  pack_ = opt_.batch_pack == 0;
  tune();
  launch_batch_spmm(be_.stream(), dk_, X_[0], KX_[0], nullptr, nullptr, nullptr, active_, dw_);
  cur_ = 0;
  auto lap = [&](double Phases::*field) {
    const Clock::time_point now = Clock::now();
    if (phases_ != nullptr) phases_->*field += std::chrono::duration<double>(now - mark).count();
    mark = now;
  };
  be_.synchronize();
  lap(&Phases::setup);

  std::vector<LaneOutcome> out(static_cast<size_t>(count_));
  const Int freq = opt_.termination_check_frequency;
  Int iteration = 0;
  double* h_check = pinned_ + 64 * 8;
  while (iteration < opt_.max_iterations && active_ != 0) {
    const Int period_end = std::min<Int>((iteration / freq + 1) * freq, opt_.max_iterations);
    const int steps = period_end - iteration;
    const bool has_check = (period_end % freq == 0);
    // Polishing sub-batches run a few periods each; capturing graphs for
    // them costs more than it saves.
    const bool graph =
        opt_.use_graphs != 0 && depth_ == 0 && has_check && steps == freq && !be_.profiling();
    upload_params(graph);
    if (graph) {
      hip_check(hipGraphLaunch(graph_for(cur_), be_.stream()), "hipGraphLaunch(batch)");
    } else {
      enqueue_period(steps, has_check, cur_);
    }
    for (int l = 0; l < count_; ++l)
      if ((active_ >> l) & 1ull) ctl_[static_cast<size_t>(l)].advance(steps);
    if (steps % 2 == 1) cur_ = 1 - cur_;
    iteration = period_end;
    if (!has_check) {
      be_.synchronize();
      lap(&Phases::iterating);
      continue;
    }
    if (graph) {
      wait_published();
    } else {
      fetch(CHECK_, h_check, static_cast<size_t>(64) * kLaneCheckStride);
    }
    lap(&Phases::iterating);
    if (!pending_.empty()) settle_polish(out, false);
    const Int i = period_end - 1;
    const double elapsed = seconds_since(start_);

    std::vector<bool> stopped(static_cast<size_t>(count_), false);
    std::vector<Residual> res(static_cast<size_t>(count_));
    std::map<Int, std::vector<int>> polish_by_budget;
    for (int l = 0; l < count_; ++l) {
      if (!((active_ >> l) & 1ull)) continue;
      Control& c = ctl_[static_cast<size_t>(l)];
      const Lane& ln = lanes_[static_cast<size_t>(l)];
      const double* h = h_check + static_cast<size_t>(slot_of_[l]) * kLaneCheckStride;
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
      rs.q_two = ln.q_two;
      rs.c_two = ln.c_two;
      rs.q_inf = ln.q_inf;
      rs.c_inf = ln.c_inf;
      res[static_cast<size_t>(l)] = residual_from_terms(t, rs);
      Int budget = 0;
      const Control::Verdict v = c.check_first(i, res[static_cast<size_t>(l)], farkas, &budget);
      if (v == Control::Verdict::kStop) {
        stopped[static_cast<size_t>(l)] = true;
      } else if (v == Control::Verdict::kPolish) {
        polish_by_budget[budget].push_back(l);
      }
    }
    lap(&Phases::checks);
    if (polish_ != nullptr && !polish_by_budget.empty()) {  // sub-batches never polish
      // A lane due another attempt while one is in flight has that one
      // settled first (it may have succeeded, which retires the lane).
      Mask due = 0;
      for (const auto& kv : polish_by_budget)
        for (int l : kv.second) due |= 1ull << l;
      if ((due & polishing_) != 0) settle_polish(out, true);
      for (const auto& kv : polish_by_budget) {
        std::vector<int> who;
        for (int l : kv.second)
          if ((active_ >> l) & 1ull) who.push_back(l);
        if (!who.empty()) start_polish(who, kv.first, i);
      }
    }
    lap(&Phases::polishing);
    Mask restart = 0;
    for (int l = 0; l < count_; ++l) {
      if (!((active_ >> l) & 1ull) || stopped[static_cast<size_t>(l)]) continue;
      const Control::Verdict v =
          ctl_[static_cast<size_t>(l)].check_second(i, res[static_cast<size_t>(l)], 0.0, elapsed);
      if (v == Control::Verdict::kStop) {
        stopped[static_cast<size_t>(l)] = true;
      } else if (v == Control::Verdict::kRestart) {
        restart |= 1ull << l;
      }
    }
    Mask retiring = 0;
    for (int l = 0; l < count_; ++l) {
      if (!stopped[static_cast<size_t>(l)] || ((settled_ >> l) & 1ull)) continue;
      retiring |= 1ull << l;
      LaneOutcome& o = out[static_cast<size_t>(l)];
      o.iterations = i + 1;
      o.seconds = seconds_since(start_);
    }
    retire_extract(retiring, out);
    active_ &= ~retiring;
    run_restarts(restart & active_);
    // Lanes waiting on an attempt keep iterating: that costs nothing while
    // the attempt runs, and a failed attempt leaves them that much further.
    compact();
    lap(&Phases::checks);
  }
  // Lanes still running stopped at the iteration limit.
  Mask rest = active_;
  for (int l = 0; l < count_; ++l) {
    if (!((rest >> l) & 1ull)) continue;
    out[static_cast<size_t>(l)].iterations = iteration;
    out[static_cast<size_t>(l)].seconds = seconds_since(start_);
  }
  retire_extract(rest, out);
  for (int l = 0; l < count_; ++l) {
    if ((settled_ >> l) & 1ull) continue;
    LaneOutcome& o = out[static_cast<size_t>(l)];
    const Control& c = ctl_[static_cast<size_t>(l)];
    o.status = c.status;
    o.message = c.message;
    o.restarts = c.restarts;
    o.eta = c.eta;
    o.omega = c.omega;
  }
  // Attempts still in flight decide last: a success stands for its lane
  // whatever the lane did after the attempt's check.
  settle_polish(out, true);
  return out;
}

}  // namespace

bool batch_supports(const anuk_pdhg_options& o, const char** why) {
  auto no = [&](const char* w) {
    if (why) *why = w;
    return false;
  };
  if (o.adaptive_step_size && !o.constant_step_size) return no("adaptive step size");
  if (o.halpern_restart) return no("Halpern restart test");
  if (o.restarts && !o.halpern) return no("averaged restarts");
  if (o.restarts && !o.restart_on_fixed_point) return no("KKT restart test");
  if (!o.reuse_products) return no("host-side residual products");
  if (o.verbose) return no("verbose logging");
  if (o.warm_x != nullptr || o.warm_y != nullptr) return no("a shared warm start");
  if (o.termination_check_frequency <= 0) return no("a non-positive check frequency");
  return true;
}

std::vector<SolveResult> solve_pdhg_batch(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                          const anuk_pdhg_options& opt) {
  const Clock::time_point start = Clock::now();
  const char* why = nullptr;
  if (!batch_supports(opt, &why)) throw std::invalid_argument(std::string("batch: unsupported option: ") + why);
  if (lps.empty()) return {};
  const StandardLp& first = *lps[0];
  for (const StandardLp* lp : lps) {
    if (lp->rows() != first.rows() || lp->cols() != first.cols() ||
        lp->num_equalities != first.num_equalities || lp->k.start != first.k.start ||
        lp->k.index != first.k.index || lp->k.value != first.k.value)
      throw std::invalid_argument("batch: every LP must share the same constraint matrix");
  }

  Phases phases;
  Clock::time_point mark = start;
  auto lap = [&](double Phases::*field) {
    const Clock::time_point now = Clock::now();
    phases.*field += std::chrono::duration<double>(now - mark).count();
    mark = now;
  };

  // Scaling depends on K alone: equilibrate it once and replay the sweeps on
  // each LP's vectors.
  Csr k = first.k;
  Csr kt;
  std::vector<Sweep> sweeps;
  const bool scales = opt.ruiz_iterations > 0 || opt.pock_chambolle != 0;
  if (scales) {
    sweeps = scale_matrix(&k, opt.ruiz_iterations, opt.pock_chambolle != 0);
    kt = k.transpose();
  } else {
    kt = first.kt;
  }
  Scaling scaling;
  std::vector<Lane> lanes(lps.size());
  HostPool::shared().run(lps.size(), [&](size_t i) {
    Lane& ln = lanes[i];
    ln.c = lps[i]->c;
    ln.q = lps[i]->q;
    ln.lower = lps[i]->lower;
    ln.upper = lps[i]->upper;
    Scaling s;
    apply_sweeps(sweeps, &ln.q, &ln.c, &ln.lower, &ln.upper, &s);
    if (i == 0) scaling = s;
    ln.oc = lps[i]->c;
    ln.oq = lps[i]->q;
    ln.olo = lps[i]->lower;
    ln.ohi = lps[i]->upper;
  });
  lap(&Phases::scaling);
  StandardLp shape;
  shape.k = k;
  shape.kt = kt;
  be.prepare(view_of(k));
  be.prepare(view_of(kt));
  struct Release {
    HipBackend& b;
    std::uint64_t a, c;
    ~Release() {
      // A teardown error is dropped: the error that unwinds past here is
      // the one the caller gets, and a destructor must not throw.
      try {
        b.release_one(a);
        b.release_one(c);
      } catch (...) {
      }
    }
  } release{be, k.id, kt.id};
  const double norm = opt.known_matrix_norm > 0.0 ? opt.known_matrix_norm
                                                 : estimate_matrix_norm(be, shape, 200, 1e-8);
  lap(&Phases::matrix_norm);

  std::vector<SolveResult> results(lps.size());
  if (norm <= 0.0) {
    // No coupling: every LP gets its plain solve's componentwise answer
    // (from the plain kernels, which a batch of one would route back here).
    anuk_pdhg_options plain = opt;
    plain.small_lp_engine = 1;
    for (size_t i = 0; i < lps.size(); ++i) results[i] = solve_pdhg(be, *lps[i], plain);
    return results;
  }

  std::vector<LaneOutcome> outcomes;
  outcomes.reserve(lps.size());
  for (size_t g0 = 0; g0 < lps.size(); g0 += kLanes) {
    const size_t g1 = std::min(lps.size(), g0 + kLanes);
    std::vector<Lane> group(lanes.begin() + static_cast<long>(g0), lanes.begin() + static_cast<long>(g1));
    std::vector<const StandardLp*> originals(lps.begin() + static_cast<long>(g0), lps.begin() + static_cast<long>(g1));
    PolishContext pc{&originals, &scaling};
    std::vector<LaneOutcome> part = run_engine(be, k, kt, first.num_equalities, scaling.row_scale,
                                               scaling.col_scale, norm, opt, group, start,
                                               opt.polish_feasibility ? &pc : nullptr, &phases,
                                               0);
    for (auto& o : part) outcomes.push_back(std::move(o));
  }

  // Final answers, and one more polish for the LPs that ran out of
  // iterations (pdhg.cpp:1205-1269), batched by budget.
  std::map<Int, std::vector<size_t>> exit_polish;
  // Each unpolished result unscaled and measured on its own LP, in parallel.
  HostPool::shared().run(lps.size(), [&](size_t i) {
    LaneOutcome& o = outcomes[i];
    if (o.polished) return;
    SolveResult& r = results[i];
    scaling.unscale_primal(o.x, &r.x);
    scaling.unscale_dual(o.y, &r.y);
    r.residual = evaluate_residual(*lps[i], r.x, r.y);
  });
  for (size_t i = 0; i < lps.size(); ++i) {
    LaneOutcome& o = outcomes[i];
    SolveResult& r = results[i];
    r.status = o.status;
    r.message = o.message;
    r.iterations = o.iterations;
    r.restarts = o.restarts;
    r.polish_attempts = o.polish_attempts;
    r.polish_iterations = o.polish_iterations;
    r.matrix_norm_estimate = norm;
    r.final_step_size = o.eta;
    r.final_primal_weight = o.omega;
    r.solve_seconds = o.seconds;
    if (o.polished) {
      r.x = std::move(o.x);
      r.y = std::move(o.y);
      r.residual = o.polished_r;
      r.polished = true;
    } else {
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
        if (budget > 0) exit_polish[budget].push_back(i);
      }
    }
  }
  mark = Clock::now();
  for (const auto& kv : exit_polish) {
    // Seeds are the working-space iterates the loop stopped at; adoption
    // follows pdhg.cpp:1243-1268 (neither feasibility measure worse, one
    // better, the gap within what the caller accepts or already had).
    for (size_t g0 = 0; g0 < kv.second.size(); g0 += kLanes) {
      const size_t g1 = std::min(kv.second.size(), g0 + kLanes);
      std::vector<size_t> idx(kv.second.begin() + static_cast<long>(g0),
                              kv.second.begin() + static_cast<long>(g1));
      // LPs that share a budget reached it at the same iteration and so with
      // the same step size; the primal weight is each LP's own.
      std::vector<const Lane*> src;
      std::vector<std::vector<double>> sx, sy;
      std::vector<double> omega;
      for (size_t i : idx) {
        src.push_back(&lanes[i]);
        sx.push_back(outcomes[i].x);
        sy.push_back(outcomes[i].y);
        omega.push_back(outcomes[i].omega);
      }
      std::vector<Polished> pol = polish_lanes(be, k, kt, first.num_equalities, norm, opt, src, sx,
                                               sy, omega, outcomes[idx[0]].eta, kv.first, scaling,
                                               start);
      const double gap_tol = opt.gap_tolerance > 0.0 ? opt.gap_tolerance : opt.tolerance;
      for (size_t s = 0; s < idx.size(); ++s) {
        SolveResult& r = results[idx[s]];
        Polished& p = pol[s];
        const Residual tr = evaluate_residual(*lps[idx[s]], p.x, p.y);
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
          if (tr.converged(opt.tolerance, opt.absolute_tolerance,
                           opt.require_inf_norm_termination != 0, gap_tol)) {
            r.status = ANUK_PDHG_OPTIMAL;
            r.message = "converged after feasibility polishing";
          } else {
            r.message = r.message.empty() ? "feasibility polished" : r.message + ", feasibility polished";
          }
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
