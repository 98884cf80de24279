#include "backend.hpp"

#include "host_pool.hpp"

#include <algorithm>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace anuk {

namespace {
enum : int { kSlotRedA = 0, kSlotRedB = 1 };
constexpr int kResultDoubles = 64;
}  // namespace

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

HipBackend::HipBackend(int device) : device_(device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw std::runtime_error("hip: no device available");
  if (device < 0 || device >= count) throw std::runtime_error("hip: device index out of range");
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipMalloc(&result_dev_, kResultDoubles * sizeof(double)), "hipMalloc(result)");
  hip_check(hipMemset(result_dev_, 0, kResultDoubles * sizeof(double)), "hipMemset(result)");
  hip_check(hipHostMalloc(&result_host_, kResultDoubles * sizeof(double), hipHostMallocDefault),
            "hipHostMalloc(result)");
  hip_check(hipEventCreate(&stop_event_), "hipEventCreate");
  noalloc_scratch_init(device);
}

HipBackend::~HipBackend() {
  // Teardown: failures here have nowhere to go, so their statuses are dropped.
  helper_.reset();
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  clear_profile();
  for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
  for (auto& kv : matrices_) free_matrix(&kv.second);
  for (auto& kv : scratch_) (void)hipFree(kv.second.first);
  if (result_dev_) (void)hipFree(result_dev_);
  if (result_host_) (void)hipHostFree(result_host_);
  if (pinned_) (void)hipHostFree(pinned_);
  for (auto& kv : pinned_slots_) (void)hipHostFree(kv.second.first);
  if (spare_slab_.first) (void)hipFree(spare_slab_.first);
  if (stop_event_) (void)hipEventDestroy(stop_event_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void HipBackend::synchronize() { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

HipBackend& HipBackend::helper() {
  if (!helper_) helper_ = std::make_unique<HipBackend>(device_);
  return *helper_;
}

hipEvent_t HipBackend::take_event() {
  if (!event_pool_.empty()) {
    hipEvent_t e = event_pool_.back();
    event_pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

// Folds the elapsed times of an entry's recorded pairs into its totals and
// returns the events to the pool. Waits for the stream.
// A pair whose time cannot be read (its launch was skipped, so the events
// were never recorded) is counted as dropped, not timed, and the error the
// query left is cleared so the next launch check does not report it.
void HipBackend::harvest(ProfileEntry& e) const {
  (void)hipStreamSynchronize(stream_);
  for (auto& p : e.pairs) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, p.start, p.stop) == hipSuccess) {
      e.ms += static_cast<double>(t);
      e.bytes += p.bytes;
      ++e.timed;
    } else {
      ++e.dropped;
      (void)hipGetLastError();
    }
    event_pool_.push_back(p.start);
    event_pool_.push_back(p.stop);
  }
  e.pairs.clear();
}

void HipBackend::clear_profile() {
  (void)hipStreamSynchronize(stream_);
  for (auto& kv : profile_) {
    for (auto& p : kv.second.pairs) {
      event_pool_.push_back(p.start);
      event_pool_.push_back(p.stop);
    }
  }
  profile_.clear();
}

LaunchEvents HipBackend::launch_events(const char* name, double bytes) {
  LaunchEvents ev;
  if (!profiling_) return ev;
  auto& e = profile_[name];
  ++e.calls;
  if (static_cast<int>(e.pairs.size()) >= kMaxTimed) harvest(e);
  ev.start = take_event();
  ev.stop = take_event();
  e.pairs.push_back({ev.start, ev.stop, bytes});
  return ev;
}

HipBackend::Timed::Timed(HipBackend& b, const char* name, double bytes)
    : b_(b), name_(name), bytes_(bytes) {
  if (!b_.profiling_) return;
  auto& e = b_.profile_[name_];
  ++e.calls;
  if (static_cast<int>(e.pairs.size()) >= kMaxTimed) b_.harvest(e);
  start_ = b_.take_event();
  hip_check(hipEventRecord(start_, b_.stream_), "hipEventRecord");
  active_ = true;
}

HipBackend::Timed::~Timed() {
  if (!active_) return;
  // A destructor must not throw: without a stop event the interval is
  // dropped and its start event returns to the pool.
  hipEvent_t stop = nullptr;
  try {
    stop = b_.take_event();
  } catch (...) {
    b_.event_pool_.push_back(start_);
    return;
  }
  (void)hipEventRecord(stop, b_.stream_);
  b_.profile_[name_].pairs.push_back({start_, stop, bytes_});
}

// ---- memory ----------------------------------------------------------------

double* HipBackend::allocate(Int n) {
  if (n < 0) throw std::invalid_argument("allocate: negative size");
  if (n == 0) return nullptr;
  double* p = nullptr;
  hip_check(hipMalloc(&p, static_cast<size_t>(n) * sizeof(double)), "hipMalloc");
  // hipMalloc does not zero; the solver accumulates into fresh vectors.
  hip_check(hipMemsetAsync(p, 0, static_cast<size_t>(n) * sizeof(double), stream_), "hipMemsetAsync");
  return p;
}

void HipBackend::deallocate(double* p) {
  if (p == nullptr) return;
  synchronize();
  hip_check(hipFree(p), "hipFree");
}

void HipBackend::upload(const double* host, double* target, Int n) {
  if (n <= 0) return;
  hip_check(hipMemcpyAsync(target, host, static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice,
                           stream_),
            "upload");
  synchronize();
}

void HipBackend::download(const double* source, double* host, Int n) {
  if (n <= 0) return;
  hip_check(hipMemcpyAsync(host, source, static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToHost,
                           stream_),
            "download");
  synchronize();
}

void HipBackend::fill(double* target, Int n, double value) {
  if (n <= 0) return;
  Timed t(*this, "fill");
  launch_fill(stream_, target, n, value);
}

void HipBackend::copy(const double* source, double* target, Int n) {
  if (n <= 0) return;
  Timed t(*this, "copy", 16.0 * n);
  hip_check(hipMemcpyAsync(target, source, static_cast<size_t>(n) * sizeof(double),
                           hipMemcpyDeviceToDevice, stream_),
            "copy");
}

double* HipBackend::scratch(int slot, std::size_t doubles) {
  auto it = scratch_.find(slot);
  if (it != scratch_.end() && it->second.second >= doubles) return it->second.first;
  if (it != scratch_.end()) {
    synchronize();
    (void)hipFree(it->second.first);
    scratch_.erase(it);
  }
  double* p = nullptr;
  const std::size_t size = std::max<std::size_t>(doubles, 1024);
  hip_check(hipMalloc(&p, size * sizeof(double)), "hipMalloc(scratch)");
  scratch_[slot] = {p, size};
  return p;
}

double* HipBackend::pinned_scratch(std::size_t doubles) {
  if (pinned_ != nullptr && pinned_size_ >= doubles) return pinned_;
  if (pinned_ != nullptr) {
    synchronize();
    (void)hipHostFree(pinned_);
    pinned_ = nullptr;
  }
  pinned_size_ = std::max<std::size_t>(doubles, std::size_t{1} << 16);
  hip_check(hipHostMalloc(&pinned_, pinned_size_ * sizeof(double), hipHostMallocDefault),
            "hipHostMalloc(staging)");
  return pinned_;
}

void HipBackend::upload_concat(void* target,
                               const std::vector<std::pair<const void*, std::size_t>>& parts) {
  std::vector<UploadPiece> pieces;
  pieces.reserve(parts.size());
  std::size_t at = 0;
  for (const auto& p : parts) {
    pieces.push_back({p.first, static_cast<char*>(target) + at, p.second});
    at += p.second;
  }
  upload_scatter(pieces);
}

void HipBackend::upload_scatter(const std::vector<UploadPiece>& pieces) {
  std::size_t total = 0;
  for (const auto& p : pieces) total += p.bytes;
  if (total == 0) return;
  constexpr std::size_t kChunk = std::size_t{64} << 20;  // bytes per staging chunk
  constexpr std::size_t kTask = std::size_t{4} << 20;    // bytes per host copy task
  constexpr int kStageSlot = 900;                        // pinned slots 900, 901
  const std::size_t chunk = std::min(kChunk, total);
  char* stage[2];
  for (int b = 0; b < 2; ++b)
    stage[b] = reinterpret_cast<char*>(pinned_slot(kStageSlot + b, (chunk + 7) / 8));
  hipEvent_t done[2] = {take_event(), take_event()};
  bool busy[2] = {false, false};
  // Walks the pieces as one byte stream: piece p from offset `at`.
  std::size_t p = 0, at = 0, out = 0;
  for (int b = 0; out < total; b ^= 1) {
    if (busy[b]) hip_check(hipEventSynchronize(done[b]), "upload_scatter");
    const std::size_t len = std::min(chunk, total - out);
    // The chunk's host copy tasks (offset in the chunk, source, bytes) and
    // its device copies (offset in the chunk, destination, bytes): one per
    // piece fragment that the chunk holds.
    struct Task {
      std::size_t off;
      const char* src;
      std::size_t bytes;
    };
    struct Send {
      std::size_t off;
      char* dst;
      std::size_t bytes;
    };
    std::vector<Task> tasks;
    std::vector<Send> sends;
    for (std::size_t filled = 0; filled < len;) {
      const std::size_t left = std::min(pieces[p].bytes - at, len - filled);
      if (left > 0) {
        sends.push_back({filled, static_cast<char*>(pieces[p].device) + at, left});
        for (std::size_t t = 0; t < left; t += kTask)
          tasks.push_back({filled + t, static_cast<const char*>(pieces[p].host) + at + t,
                           std::min(kTask, left - t)});
      }
      filled += left;
      at += left;
      if (at == pieces[p].bytes) {
        ++p;
        at = 0;
      }
    }
    char* buf = stage[b];
    HostPool::shared().run(tasks.size(), [&](std::size_t i) {
      std::memcpy(buf + tasks[i].off, tasks[i].src, tasks[i].bytes);
    });
    for (const Send& sd : sends)
      hip_check(hipMemcpyAsync(sd.dst, buf + sd.off, sd.bytes, hipMemcpyHostToDevice, stream_),
                "upload_scatter");
    hip_check(hipEventRecord(done[b], stream_), "upload_scatter");
    busy[b] = true;
    out += len;
  }
  synchronize();
  event_pool_.push_back(done[0]);
  event_pool_.push_back(done[1]);
}

void HipBackend::download_staged(const void* source, void* host, std::size_t bytes) {
  if (bytes == 0) return;
  constexpr std::size_t kChunk = std::size_t{64} << 20;
  constexpr std::size_t kTask = std::size_t{4} << 20;
  constexpr int kStageSlot = 900;  // the chunks upload_concat uses
  const std::size_t chunk = std::min(kChunk, bytes);
  char* stage[2];
  for (int b = 0; b < 2; ++b)
    stage[b] = reinterpret_cast<char*>(pinned_slot(kStageSlot + b, (chunk + 7) / 8));
  hipEvent_t done[2] = {take_event(), take_event()};
  const std::size_t count = (bytes + chunk - 1) / chunk;
  auto issue = [&](std::size_t c) {
    const int b = static_cast<int>(c & 1);
    const std::size_t off = c * chunk, len = std::min(chunk, bytes - off);
    hip_check(hipMemcpyAsync(stage[b], static_cast<const char*>(source) + off, len,
                             hipMemcpyDeviceToHost, stream_),
              "download_staged");
    hip_check(hipEventRecord(done[b], stream_), "download_staged");
  };
  issue(0);
  for (std::size_t c = 0; c < count; ++c) {
    // Chunk c + 1 goes out before chunk c is copied out of its buffer.
    if (c + 1 < count) issue(c + 1);
    const int b = static_cast<int>(c & 1);
    hip_check(hipEventSynchronize(done[b]), "download_staged");
    const std::size_t off = c * chunk, len = std::min(chunk, bytes - off);
    const char* buf = stage[b];
    char* dst = static_cast<char*>(host) + off;
    HostPool::shared().run((len + kTask - 1) / kTask, [&](std::size_t i) {
      const std::size_t o = i * kTask;
      std::memcpy(dst + o, buf + o, std::min(kTask, len - o));
    });
  }
  event_pool_.push_back(done[0]);
  event_pool_.push_back(done[1]);
}

double* HipBackend::pinned_slot(int slot, std::size_t doubles) {
  auto it = pinned_slots_.find(slot);
  if (it != pinned_slots_.end() && it->second.second >= doubles) return it->second.first;
  if (it != pinned_slots_.end()) {
    synchronize();
    (void)hipHostFree(it->second.first);
    pinned_slots_.erase(it);
  }
  double* p = nullptr;
  const std::size_t size = std::max<std::size_t>(doubles, 512);
  // Coherent: a kernel may read the host's latest writes here, and the host
  // may poll values a kernel stores.
  hip_check(hipHostMalloc(&p, size * sizeof(double), hipHostMallocCoherent), "hipHostMalloc(slot)");
  pinned_slots_[slot] = {p, size};
  return p;
}

void HipBackend::upload_staged(const double* host, double* target, std::size_t n) {
  if (n == 0) return;
  // Large enough chunks to run at DMA speed, small enough to keep the
  // staging buffer modest; each chunk's copy completes before it is reused.
  const std::size_t chunk = std::size_t{1} << 20;
  double* stage = pinned_scratch(std::min(n, chunk));
  for (std::size_t off = 0; off < n; off += chunk) {
    const std::size_t len = std::min(chunk, n - off);
    std::memcpy(stage, host + off, len * sizeof(double));
    hip_check(hipMemcpyAsync(target + off, stage, len * sizeof(double), hipMemcpyHostToDevice, stream_),
              "upload_staged");
    synchronize();
  }
}

const double* HipBackend::fetch_result(int n) {
  hip_check(hipMemcpyAsync(result_host_, result_dev_, static_cast<size_t>(n) * sizeof(double),
                           hipMemcpyDeviceToHost, stream_),
            "fetch_result");
  synchronize();
  return result_host_;
}

// ---- matrices ----------------------------------------------------------------

void HipBackend::free_matrix(DeviceMatrix* dm) {
  for (void* p : dm->owned) (void)hipFree(p);
  dm->owned.clear();
}

const DevCsr& HipBackend::prepare(const CsrView& a) {
  auto it = matrices_.find(a.id);
  if (it != matrices_.end()) return it->second.view;
  if (a.rows < 0 || a.cols < 0 || a.start == nullptr)
    throw std::invalid_argument("prepare: malformed matrix");
  DeviceMatrix dm;
  const Int nnz = a.nnz();
  // The arrays cross together, staged through the page-locked chunks.
  std::vector<UploadPiece> pieces;
  auto put = [&](const void* src, size_t bytes) -> void* {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(bytes, 8)), "hipMalloc(matrix)");
    dm.owned.push_back(p);
    if (bytes > 0) pieces.push_back({src, p, bytes});
    return p;
  };
  Int* block_rows = nullptr;
  Int blocks = 0;
  plan_blocks(a.rows, a.start, kTileNnz, kTileRows, &block_rows, &blocks);
  dm.view.rows = a.rows;
  dm.view.cols = a.cols;
  dm.view.nnz = nnz;
  dm.view.blocks = blocks;
  for (Int r = 0; r < a.rows; ++r) dm.view.longest = std::max(dm.view.longest, a.start[r + 1] - a.start[r]);
  dm.view.start = static_cast<const Int*>(put(a.start, (static_cast<size_t>(a.rows) + 1) * sizeof(Int)));
  dm.view.index = static_cast<const Int*>(put(a.index, static_cast<size_t>(nnz) * sizeof(Int)));
  dm.view.value = static_cast<const double*>(put(a.value, static_cast<size_t>(nnz) * sizeof(double)));
  dm.view.block_row =
      static_cast<const Int*>(put(block_rows, (static_cast<size_t>(blocks) + 1) * sizeof(Int)));
  try {
    upload_scatter(pieces);
  } catch (...) {
    std::free(block_rows);
    free_matrix(&dm);
    throw;
  }
  std::free(block_rows);
  auto ins = matrices_.emplace(a.id, std::move(dm));
  return ins.first->second.view;
}

namespace {
// Cache keys of a matrix's fast plan and of its panel stack: the matrix's own
// identity with a top bit flipped (identities are drawn from a counter and
// never reach them).
constexpr std::uint64_t kPanelKey = 1ull << 63;
constexpr std::uint64_t kFastKey = 1ull << 62;
constexpr std::uint64_t kSlotKey = 1ull << 61;
constexpr std::uint64_t kBinKey = 1ull << 60;
}  // namespace

void HipBackend::replan_fast(const CsrView& a, DeviceMatrix* dm) {
  Int* block_rows = nullptr;
  Int blocks = 0;
  plan_blocks(a.rows, a.start, kFastTileNnz, kFastTileRows, &block_rows, &blocks);
  void* p = nullptr;
  const size_t bytes = (static_cast<size_t>(blocks) + 1) * sizeof(Int);
  const hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    std::free(block_rows);
    hip_check(e, "hipMalloc(fast plan)");
  }
  dm->owned.push_back(p);
  hip_check(hipMemcpyAsync(p, block_rows, bytes, hipMemcpyHostToDevice, stream_), "upload fast plan");
  synchronize();
  std::free(block_rows);
  dm->view.blocks = blocks;
  dm->view.block_row = static_cast<const Int*>(p);
  dm->view.max_rows = kFastTileRows;
}

const DevCsr& HipBackend::prepare_fast(const CsrView& a) {
  const std::uint64_t key = a.id ^ kFastKey;
  auto it = matrices_.find(key);
  if (it != matrices_.end()) return it->second.view;
  DeviceMatrix dm;
  dm.view = prepare(a);  // the arrays stay owned by the matrix's own entry
  replan_fast(a, &dm);
  auto ins = matrices_.emplace(key, std::move(dm));
  return ins.first->second.view;
}

const DevCsr& HipBackend::prepare_slots(const CsrView& a) {
  const std::uint64_t key = a.id ^ kSlotKey;
  auto it = matrices_.find(key);
  if (it != matrices_.end()) return it->second.view;
  DeviceMatrix dm;
  dm.view = prepare(a);  // the CSR arrays stay owned by the matrix's own entry
  // The kernels gather through a buffer resource with 31-bit offsets: a
  // gathered vector of 64 lanes per row must stay below 2^31 bytes.
  if (a.cols >= kSlotMaxCols) {
    auto ins = matrices_.emplace(key, std::move(dm));
    return ins.first->second.view;
  }
  Int longest = 0;
  for (Int r = 0; r < a.rows; ++r) longest = std::max(longest, a.start[r + 1] - a.start[r]);
  Int width = kSlotWidths[0];
  for (const int w : kSlotWidths) {
    width = w;
    if (longest <= w) break;
  }
  const size_t rows = uz(a.rows), cells = rows * uz(width);
  std::vector<Int> row(2 * rows), index(cells, 0);
  std::vector<double> value(cells, 0.0);
  for (size_t r = 0; r < rows; ++r) {
    const Int lo = a.start[r], len = a.start[r + 1] - lo;
    row[2 * r] = lo;
    row[2 * r + 1] = len;
    for (Int u = 0; u < std::min(len, width); ++u) {
      index[r * uz(width) + uz(u)] = a.index[lo + u];
      value[r * uz(width) + uz(u)] = a.value[lo + u];
    }
    // Padding repeats the row's last index, so its discarded gathers read a
    // line the row reads anyway rather than all waves reading one line.
    for (Int u = std::min(len, width); u < width && len > 0; ++u)
      index[r * uz(width) + uz(u)] = a.index[lo + std::min(len, width) - 1];
  }
  auto put = [&](const void* src, size_t bytes) -> void* {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(bytes, 8)), "hipMalloc(slots)");
    dm.owned.push_back(p);
    if (bytes > 0) hip_check(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream_), "upload slots");
    return p;
  };
  dm.view.slot_width = width;
  dm.view.slot_row = static_cast<const Int*>(put(row.data(), row.size() * sizeof(Int)));
  dm.view.slot_index = static_cast<const Int*>(put(index.data(), index.size() * sizeof(Int)));
  dm.view.slot_value = static_cast<const double*>(put(value.data(), value.size() * sizeof(double)));
  synchronize();
  auto ins = matrices_.emplace(key, std::move(dm));
  return ins.first->second.view;
}

const DevCsr& HipBackend::prepare_bins(const CsrView& a) {
  const std::uint64_t key = a.id ^ kBinKey;
  auto it = matrices_.find(key);
  if (it != matrices_.end()) return it->second.view;
  DeviceMatrix dm;
  dm.view = prepare(a);  // the CSR arrays stay owned by the matrix's own entry
  std::vector<Int> short_rows, long_rows;
  for (Int r = 0; r < a.rows; ++r)
    (a.start[r + 1] - a.start[r] <= kBinShortMax ? short_rows : long_rows).push_back(r);
  auto entry = [&](std::vector<Int>* plan, Int r) {
    if (r < 0) {
      plan->insert(plan->end(), {-1, 0, 0, 0});
    } else {
      plan->insert(plan->end(), {r, a.start[r], a.start[r + 1] - a.start[r], 0});
    }
  };
  for (int p = 0; p < 2; ++p) {
    const Int g = p == 0 ? 2 : 4;
    const Int waves = (static_cast<Int>(short_rows.size()) + g - 1) / g;
    std::vector<Int> plan;
    plan.reserve(4 * (uz(waves * g) + long_rows.size()));
    for (Int e = 0; e < waves * g; ++e)
      entry(&plan, e < static_cast<Int>(short_rows.size()) ? short_rows[uz(e)] : -1);
    for (Int r : long_rows) entry(&plan, r);
    void* d = nullptr;
    hip_check(hipMalloc(&d, std::max<size_t>(plan.size(), 4) * sizeof(Int)), "hipMalloc(bins)");
    dm.owned.push_back(d);
    if (!plan.empty())
      hip_check(hipMemcpyAsync(d, plan.data(), plan.size() * sizeof(Int), hipMemcpyHostToDevice, stream_),
                "upload bins");
    dm.view.bin_plan[p] = static_cast<const Int*>(d);
    dm.view.bin_short[p] = waves;
  }
  dm.view.bin_long = static_cast<Int>(long_rows.size());
  // The fixed-width copies: every short entry of either plan (the four-row
  // plan pads furthest), then every long row's first kBinLongWidth entries.
  auto fill = [&](const std::vector<Int>& rows, size_t count, Int width, std::vector<Int>* ix,
                  std::vector<double>* val) {
    ix->assign(count * uz(width), 0);
    val->assign(count * uz(width), 0.0);
    for (size_t e = 0; e < rows.size(); ++e) {
      const Int lo = a.start[rows[e]], len = std::min(a.start[rows[e] + 1] - lo, width);
      for (Int u = 0; u < width; ++u) {
        const Int at = lo + std::min(u, len - 1);
        (*ix)[e * uz(width) + uz(u)] = len > 0 ? a.index[at] : 0;
        (*val)[e * uz(width) + uz(u)] = u < len ? a.value[at] : 0.0;
      }
    }
  };
  std::vector<Int> ix8, ix32;
  std::vector<double> val8, val32;
  fill(short_rows, (short_rows.size() + 3) / 4 * 4, kBinShortMax, &ix8, &val8);
  fill(long_rows, long_rows.size(), kBinLongWidth, &ix32, &val32);
  auto put = [&](const void* src, size_t bytes) -> void* {
    void* d = nullptr;
    hip_check(hipMalloc(&d, std::max<size_t>(bytes, 16)), "hipMalloc(bin copies)");
    dm.owned.push_back(d);
    if (bytes > 0)
      hip_check(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream_), "upload bin copies");
    return d;
  };
  dm.view.bin_index8 = static_cast<const Int*>(put(ix8.data(), ix8.size() * sizeof(Int)));
  dm.view.bin_value8 = static_cast<const double*>(put(val8.data(), val8.size() * sizeof(double)));
  dm.view.bin_index32 = static_cast<const Int*>(put(ix32.data(), ix32.size() * sizeof(Int)));
  dm.view.bin_value32 =
      static_cast<const double*>(put(val32.data(), val32.size() * sizeof(double)));
  synchronize();
  auto ins = matrices_.emplace(key, std::move(dm));
  return ins.first->second.view;
}

const DevCsr& HipBackend::prepare_panels(const CsrView& a) {
  auto it = matrices_.find(a.id ^ kPanelKey);
  if (it != matrices_.end()) return it->second.view;
  return install_panels(build_panels(a));
}

bool HipBackend::has_panels(const CsrView& a) const {
  return matrices_.count(a.id ^ kPanelKey) != 0;
}

HipBackend::PanelStack HipBackend::build_panels(const CsrView& a) {
  const std::uint64_t key = a.id ^ kPanelKey;
  const Int nnz = a.nnz();
  // Column boundaries that split the nonzeros into kPanels near-equal parts.
  std::vector<long> col_nnz(uz(a.cols) + 1, 0);
  for (Int k = 0; k < nnz; ++k) ++col_nnz[uz(a.index[k]) + 1];
  for (Int j = 0; j < a.cols; ++j) col_nnz[uz(j) + 1] += col_nnz[uz(j)];
  std::vector<Int> bound(kPanels + 1, a.cols);
  bound[0] = 0;
  for (int p = 1; p < kPanels; ++p) {
    const long target = static_cast<long>(nnz) * p / kPanels;
    bound[uz(p)] = static_cast<Int>(std::lower_bound(col_nnz.begin(), col_nnz.end(), target) - col_nnz.begin());
    bound[uz(p)] = std::min(std::max(bound[uz(p)], bound[uz(p) - 1]), a.cols);
  }
  // Row r of panel p: the part of row r with columns in [bound[p], bound[p+1]).
  // The per-row cuts, then each panel's row starts and entries, are
  // independent, so they are built on kPanels threads.
  const Int rows = a.rows;
  std::vector<Int> start(uz(kPanels) * uz(rows) + 1, 0);
  std::vector<Int> cut(uz(rows) * (kPanels + 1));
  auto on_threads = [](int count, const auto& body) {
    std::vector<std::thread> pool;
    for (int t = 1; t < count; ++t) pool.emplace_back(body, t);
    body(0);
    for (std::thread& th : pool) th.join();
  };
  on_threads(kPanels, [&](int t) {
    const Int r0 = static_cast<Int>(static_cast<long>(rows) * t / kPanels);
    const Int r1 = static_cast<Int>(static_cast<long>(rows) * (t + 1) / kPanels);
    for (Int r = r0; r < r1; ++r) {
      const Int* b = a.index + a.start[r];
      const Int* e = a.index + a.start[r + 1];
      for (int p = 0; p <= kPanels; ++p)
        cut[uz(r) * (kPanels + 1) + uz(p)] =
            static_cast<Int>(std::lower_bound(b, e, bound[uz(p)]) - a.index);
      cut[uz(r) * (kPanels + 1) + kPanels] = a.start[r + 1];
    }
  });
  // Where each panel's entries begin: the sum of the earlier panels' sizes.
  std::vector<long> panel_at(kPanels + 1, 0);
  for (int p = 0; p < kPanels; ++p) {
    long n = 0;
    for (Int r = 0; r < rows; ++r)
      n += cut[uz(r) * (kPanels + 1) + uz(p) + 1] - cut[uz(r) * (kPanels + 1) + uz(p)];
    panel_at[uz(p) + 1] = panel_at[uz(p)] + n;
  }
  std::vector<Int> index(uz(nnz));
  std::vector<double> value(uz(nnz));
  on_threads(kPanels, [&](int p) {
    size_t at = static_cast<size_t>(panel_at[uz(p)]);
    for (Int r = 0; r < rows; ++r) {
      const Int k0 = cut[uz(r) * (kPanels + 1) + uz(p)], k1 = cut[uz(r) * (kPanels + 1) + uz(p) + 1];
      for (Int k = k0; k < k1; ++k) {
        index[at] = a.index[k];
        value[at] = a.value[k];
        ++at;
      }
      start[uz(p) * uz(rows) + uz(r) + 1] = static_cast<Int>(at);
    }
  });
  PanelStack out;
  out.key = key;
  out.rows = kPanels * rows;
  out.cols = a.cols;
  out.start = std::move(start);
  out.index = std::move(index);
  out.value = std::move(value);
  return out;
}

const DevCsr& HipBackend::install_panels(PanelStack stack) {
  auto it = matrices_.find(stack.key);
  if (it != matrices_.end()) return it->second.view;
  const CsrView stacked{stack.key, stack.rows, stack.cols, stack.start.data(), stack.index.data(),
                        stack.value.data()};
  prepare(stacked);
  // The panel stack only ever runs in the fast order: give it that plan.
  DeviceMatrix& dm = matrices_.at(stack.key);
  replan_fast(stacked, &dm);
  return dm.view;
}

void HipBackend::release() {
  synchronize();
  for (auto& kv : matrices_) free_matrix(&kv.second);
  matrices_.clear();
  if (spare_slab_.first) hip_check(hipFree(spare_slab_.first), "hipFree(spare slab)");
  spare_slab_ = {nullptr, 0};
}

double* HipBackend::take_slab(std::size_t doubles, std::size_t* capacity) {
  if (spare_slab_.first != nullptr && spare_slab_.second >= doubles) {
    double* p = spare_slab_.first;
    *capacity = spare_slab_.second;
    spare_slab_ = {nullptr, 0};
    return p;
  }
  double* p = nullptr;
  const std::size_t size = std::max<std::size_t>(doubles, 1);
  hip_check(hipMalloc(&p, size * sizeof(double)), "hipMalloc(solve state)");
  *capacity = size;
  return p;
}

void HipBackend::give_slab(double* slab, std::size_t capacity) {
  if (slab == nullptr) return;
  // The slab may still be in use by queued work; the stream orders its next
  // user after that work, and a hipFree waits for the device.
  if (spare_slab_.first == nullptr) {
    spare_slab_ = {slab, capacity};
    return;
  }
  double* drop = slab;
  if (capacity > spare_slab_.second) {
    drop = spare_slab_.first;
    spare_slab_ = {slab, capacity};
  }
  (void)hipStreamSynchronize(stream_);
  (void)hipFree(drop);
}

void HipBackend::release_one(std::uint64_t id) {
  // The fast plan and the slot copy borrow the matrix's arrays; they go first.
  for (const std::uint64_t key : {id ^ kFastKey, id ^ kSlotKey, id ^ kBinKey, id ^ kPanelKey, id}) {
    auto it = matrices_.find(key);
    if (it == matrices_.end()) continue;
    synchronize();
    free_matrix(&it->second);
    matrices_.erase(it);
  }
}

void HipBackend::multiply(const CsrView& a, const double* x, double* y) {
  const DevCsr& d = prepare(a);
  // Algorithmic bytes: value and index once, row starts once, x gathered
  // once per column, y written once.
  const double bytes = 12.0 * d.nnz + 4.0 * (d.rows + 1.0) + 8.0 * d.cols + 8.0 * d.rows;
  const LaunchEvents ev = launch_events("spmv", bytes);
  launch_spmv(stream_, d, x, y, profiling_ ? &ev : nullptr);
}

// ---- reductions ---------------------------------------------------------------

double HipBackend::dot(const double* a, const double* b, Int n) {
  if (n <= 0) return 0.0;
  Timed t(*this, "dot");
  launch_dot(stream_, a, b, n, scratch(kSlotRedA, 1024 * kMaxRed), result_dev_);
  return fetch_result(1)[0];
}

double HipBackend::two_norm(const double* a, Int n) { return std::sqrt(dot(a, a, n)); }

double HipBackend::inf_norm(const double* a, Int n) {
  if (n <= 0) return 0.0;
  Timed t(*this, "inf_norm");
  launch_absmax(stream_, a, n, scratch(kSlotRedA, 1024 * kMaxRed), result_dev_);
  return fetch_result(1)[0];
}

void HipBackend::primal_update(Int n, double tau, double gamma, double halpern, const double* x,
                               const double* c, const double* kt_y, const double* lower,
                               const double* upper, const double* anchor, double* x_out,
                               double* dx, double* x_bar) {
  Timed t(*this, "primal_update");
  launch_primal_update(stream_, n, tau, gamma, halpern, x, c, kt_y, lower, upper, anchor, x_out, dx,
                       x_bar);
}

void HipBackend::dual_update(Int m, Int neq, double sigma, double gamma, double halpern,
                             const double* y, const double* q, const double* k_x_bar,
                             const double* k_x, const double* anchor_y, const double* anchor_kx,
                             double* y_out, double* dy, double* k_out, double* k_dx) {
  Timed t(*this, "dual_update");
  launch_dual_update(stream_, m, neq, sigma, gamma, halpern, y, q, k_x_bar, k_x, anchor_y,
                     anchor_y != nullptr ? anchor_kx : nullptr, y_out, dy, k_out, k_dx);
}

void HipBackend::accumulate(Int n, double weight, const double* v, double* sum) {
  Timed t(*this, "accumulate");
  launch_accumulate(stream_, n, weight, v, sum);
}

void HipBackend::scale_into(Int n, double weight, const double* sum, double* out) {
  Timed t(*this, "scale_into");
  launch_scale_into(stream_, n, weight, sum, out);
}

double HipBackend::weighted_norm_squared(Int n, Int m, const double* dx, const double* dy,
                                         double omega) {
  Timed t(*this, "weighted_norm");
  launch_sumsq2(stream_, dx, n, dy, m, scratch(kSlotRedA, 1024 * kMaxRed),
                scratch(kSlotRedB, 1024 * kMaxRed), result_dev_);
  const double* r = fetch_result(2);
  return omega * r[0] + r[1] / omega;
}

void HipBackend::step_size_terms(Int n, Int m, const double* dx, const double* dy,
                                 const double* k_dx, double omega, double* interaction,
                                 double* movement) {
  Timed t(*this, "step_terms");
  launch_dot_and_sumsq(stream_, dy, k_dx, m, dx, n, scratch(kSlotRedA, 1024 * kMaxRed),
                       scratch(kSlotRedB, 1024 * kMaxRed), result_dev_);
  const double* r = fetch_result(3);
  // The absolute value belongs to the backend (backend.cpp:135-139).
  *interaction = std::fabs(r[0]);
  *movement = omega * r[2] + r[1] / omega;
}

Terms HipBackend::convergence_terms_noalloc(Int n, Int m, const DevProblem& p, const double* x,
                                            const double* k_x, const double* y,
                                            const double* kt_y) {
  if (static_cast<size_t>(reduce_grid(m)) * kMaxRed > kNoAllocPartials ||
      static_cast<size_t>(reduce_grid(n)) * kMaxRed > kNoAllocPartials)
    throw std::invalid_argument("convergence terms: partials exceed the scratch array");
  double* sc = noalloc_scratch(device_);
  launch_convergence_terms(stream_, n, m, p, x, k_x, y, kt_y, sc, sc + kNoAllocPartials,
                           sc + kNoAllocResultOffset);
  hip_check(hipMemcpyAsync(result_host_, sc + kNoAllocResultOffset, 7 * sizeof(double),
                           hipMemcpyDeviceToHost, stream_),
            "convergence terms readback");
  synchronize();
  const double* r = result_host_;
  Terms out;
  out.primal_sq = r[0];
  out.primal_inf = r[1];
  out.q_dot_y = r[2];
  out.dual_sq = r[3];
  out.dual_inf = r[4];
  out.bound_term = r[5];
  out.c_dot_x = r[6];
  return out;
}

Terms HipBackend::convergence_terms(Int n, Int m, const DevProblem& p, const double* x,
                                    const double* k_x, const double* y, const double* kt_y) {
  Timed t(*this, "convergence_terms");
  launch_convergence_terms(stream_, n, m, p, x, k_x, y, kt_y, scratch(kSlotRedA, 1024 * kMaxRed),
                           scratch(kSlotRedB, 1024 * kMaxRed), result_dev_);
  const double* r = fetch_result(7);
  Terms out;
  out.primal_sq = r[0];
  out.primal_inf = r[1];
  out.q_dot_y = r[2];
  out.dual_sq = r[3];
  out.dual_inf = r[4];
  out.bound_term = r[5];
  out.c_dot_x = r[6];
  return out;
}

// ---- tuning ------------------------------------------------------------------

namespace {
// A batched-step shape, 100 * layout + 10 * waves per workgroup + rows per
// wave, as "w<waves>r<rows>" with an "s" for layout 1 (the fixed-width row
// copy); a narrow-span shape, 300 + 10 * waves + 1, as "w<waves>r1n"; a
// stack launch choice, 200 + 10 * binned + non-temporal, as "plan" or
// "binned" with "+nt".
std::string shape_label(int shape) {
  char buf[32];
  if (shape >= 300) {
    std::snprintf(buf, sizeof(buf), "w%dr1n", (shape % 100) / 10);
  } else if (shape >= 200) {
    std::snprintf(buf, sizeof(buf), "%s%s", (shape / 10) % 10 != 0 ? "binned" : "plan",
                  shape % 10 != 0 ? "+nt" : "");
  } else {
    std::snprintf(buf, sizeof(buf), "w%dr%d%s", (shape % 100) / 10, shape % 10,
                  shape >= 100 ? "s" : "");
  }
  return buf;
}
}  // namespace

std::string HipBackend::tuning_report() const {
  std::string out;
  char buf[96];
  for (const auto& kv : tuned_) {
    const Tuning& t = kv.second;
    out += t.kernel;
    std::snprintf(buf, sizeof(buf), " rows=%d nnz=%d", t.rows, t.nnz);
    out += buf;
    for (const auto& c : t.us) {
      std::snprintf(buf, sizeof(buf), " us@%s=%.3f", shape_label(c.first).c_str(), c.second);
      out += buf;
    }
    out += " chosen=" + shape_label(t.chosen) + "\n";
  }
  return out;
}

// ---- profiling -------------------------------------------------------------

void HipBackend::set_profiling(bool on) {
  if (on) clear_profile();
  profiling_ = on;
}

// One line per launch group: calls, timed calls, mean microseconds, the
// total (mean x calls) with its share, GB/s of the timed launches'
// algorithmic bytes, and the pairs whose time could not be read.
std::string HipBackend::profile_report() const {
  if (profile_.empty()) return {};
  (void)hipStreamSynchronize(stream_);
  struct Row {
    std::string name;
    long calls;
    long timed;
    long dropped;
    double mean_us;
    double total_ms;
    double gbs;  // algorithmic bytes of the timed launches over their time
  };
  std::vector<Row> rows;
  double total = 0.0;
  for (auto& kv : profile_) {
    harvest(kv.second);
    const double ms = kv.second.ms;
    const long timed = kv.second.timed;
    const double mean_us = timed > 0 ? 1e3 * ms / static_cast<double>(timed) : 0.0;
    const double est = 1e-3 * mean_us * static_cast<double>(kv.second.calls);
    const double gbs = ms > 0.0 ? kv.second.bytes / (1e-3 * ms) / 1e9 : 0.0;
    rows.push_back({kv.first, kv.second.calls, timed, kv.second.dropped, mean_us, est, gbs});
    total += est;
  }
  std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.total_ms > b.total_ms; });
  std::string out =
      "  kernel                    calls    timed      mean us     total ms   share      GB/s"
      "  dropped\n";
  char line[220];
  for (const auto& r : rows) {
    std::snprintf(line, sizeof(line), "  %-22s %8ld %8ld %12.3f %12.3f %6.1f%% %9.1f %8ld\n",
                  r.name.c_str(), r.calls, r.timed, r.mean_us, r.total_ms,
                  total > 0 ? 100.0 * r.total_ms / total : 0.0, r.gbs, r.dropped);
    out += line;
  }
  return out;
}

}  // namespace anuk
