// HipBackend: the LinAlgBackend operations (src/anukulan/backend.hpp:24-215)
// on one MI355X, one HIP stream, device-resident vectors.
//
// It owns every device allocation it hands out, caches uploaded matrices by
// their 64-bit identity (never by host address, test_backend.cpp:194-225),
// and keeps the small scratch the reductions fold through. Reductions return
// host doubles and therefore synchronise, as the reference interface
// requires; the fused solve in pdhg.cpp avoids them on the hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"
#include "lp.hpp"

namespace anuk {

void hip_check(hipError_t e, const char* what);

// A host CSR as the C ABI passes it: borrowed pointers plus an identity.
struct CsrView {
  std::uint64_t id = 0;
  Int rows = 0;
  Int cols = 0;
  const Int* start = nullptr;
  const Int* index = nullptr;
  const double* value = nullptr;
  Int nnz() const { return start ? start[rows] : 0; }
};

inline CsrView view_of(const Csr& a) {
  return CsrView{a.id, a.rows, a.cols, a.start.data(), a.index.data(), a.value.data()};
}

class HipBackend {
 public:
  explicit HipBackend(int device);
  ~HipBackend();
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  std::string name() const { return "hip"; }
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void synchronize();
  // A second backend on the same device, with its own stream, memory and
  // matrix cache, made on first use and kept for this one's life: work that
  // runs on another host thread beside this backend's (the batched solver's
  // feasibility polishing) goes through it.
  HipBackend& helper();

  // memory (backend.hpp:53-58)
  double* allocate(Int n);
  void deallocate(double* p);
  void upload(const double* host, double* target, Int n);
  void download(const double* source, double* host, Int n);
  void fill(double* target, Int n, double value);
  void copy(const double* source, double* target, Int n);

  // matrices (backend.hpp:35-36, 61-64)
  const DevCsr& prepare(const CsrView& a);
  // `a` with the fast order's row-block plan (large tiles), sharing the
  // arrays of prepare(a); cached with `a`.
  const DevCsr& prepare_fast(const CsrView& a);
  // `a` with its fixed-width copy (DevCsr::slot_*) for the one-wave-per-row
  // batch kernels, sharing the arrays of prepare(a); cached with `a`.
  const DevCsr& prepare_slots(const CsrView& a);
  // `a` with its row bins (DevCsr::bin_*) for the packed batched steps,
  // sharing the arrays of prepare(a); cached with `a`.
  const DevCsr& prepare_bins(const CsrView& a);
  // The column-panel stack of `a` (kPanels panels balanced by nonzeros, see
  // kernels.hpp), built on the host, planned for the fast order and cached
  // with `a`.
  const DevCsr& prepare_panels(const CsrView& a);
  // The same in two halves: the host build (no device call, safe on another
  // thread while this backend runs) and its upload and plan.
  struct PanelStack {
    std::uint64_t key = 0;
    Int rows = 0, cols = 0;
    std::vector<Int> start, index;
    std::vector<double> value;
  };
  bool has_panels(const CsrView& a) const;
  static PanelStack build_panels(const CsrView& a);
  const DevCsr& install_panels(PanelStack stack);
  void release();
  void release_one(std::uint64_t id);
  void multiply(const CsrView& a, const double* x, double* y);

  // reductions and steps (backend.hpp:66-201)
  double dot(const double* a, const double* b, Int n);
  double two_norm(const double* a, Int n);
  double inf_norm(const double* a, Int n);
  void primal_update(Int n, double tau, double gamma, double halpern, const double* x,
                     const double* c, const double* kt_y, const double* lower,
                     const double* upper, const double* anchor, double* x_out, double* dx,
                     double* x_bar);
  void dual_update(Int m, Int neq, double sigma, double gamma, double halpern, const double* y,
                   const double* q, const double* k_x_bar, const double* k_x,
                   const double* anchor_y, const double* anchor_kx, double* y_out, double* dy,
                   double* k_out, double* k_dx);
  void accumulate(Int n, double weight, const double* v, double* sum);
  void scale_into(Int n, double weight, const double* sum, double* out);
  double weighted_norm_squared(Int n, Int m, const double* dx, const double* dy, double omega);
  void step_size_terms(Int n, Int m, const double* dx, const double* dy, const double* k_dx,
                       double omega, double* interaction, double* movement);
  Terms convergence_terms(Int n, Int m, const DevProblem& p, const double* x,
                          const double* k_x, const double* y, const double* kt_y);
  // The same terms with every scratch buffer taken from the library's
  // device array (noalloc_scratch) and the result read back through the
  // pinned buffer made at construction: no device or host allocation, no
  // profiling events. Calls that overlap on one device race.
  Terms convergence_terms_noalloc(Int n, Int m, const DevProblem& p, const double* x,
                                  const double* k_x, const double* y, const double* kt_y);

  // profiling (backend.hpp:203-214)
  void set_profiling(bool on);
  bool profiling() const { return profiling_; }
  std::string profile_report() const;

  // Scratch for the solver: a device buffer of at least `doubles`, one per
  // slot, grown on demand and kept for the context's life.
  double* scratch(int slot, std::size_t doubles);
  // A pinned host staging buffer of at least `doubles`, grown on demand.
  double* pinned_scratch(std::size_t doubles);
  // Page-locked host memory kept per slot for the context's life (engine
  // parameter and readback staging), like scratch(): allocating and freeing
  // it per solve costs more than the solve's own small transfers.
  double* pinned_slot(int slot, std::size_t doubles);
  // Copies n doubles host -> device through the pinned staging buffer and
  // waits (fast DMA for large uploads from pageable memory).
  void upload_staged(const double* host, double* target, std::size_t n);
  // Copies the host ranges `parts` (pointer, bytes), concatenated in order,
  // to `target` on the device, and waits: through two page-locked chunks,
  // one filled by the host pool while the other's copy runs, so a large
  // upload from many pageable pieces moves at DMA speed in few copies.
  void upload_concat(void* target, const std::vector<std::pair<const void*, std::size_t>>& parts);
  // The same for host ranges that go to different places: each piece's
  // bytes to its own device address, staged through the same two chunks.
  struct UploadPiece {
    const void* host;
    void* device;
    std::size_t bytes;
  };
  void upload_scatter(const std::vector<UploadPiece>& pieces);
  // The reverse for one range: `bytes` from `source` on the device to
  // `host`, through the same two page-locked chunks (one chunk's copy runs
  // while the host pool empties the other), and waits.
  void download_staged(const void* source, void* host, std::size_t bytes);

  // A solver's device state slab of at least `doubles`: the spare one when
  // it is large enough, else a new allocation. give_slab hands a slab back;
  // the larger of it and the spare is kept, the other freed. Solves and
  // their polishing sub-solves come and go in quick succession, and a
  // hipMalloc / hipFree pair of a few hundred MB costs milliseconds and a
  // device-wide wait each time.
  // *capacity receives the slab's true size, which give_slab takes back.
  double* take_slab(std::size_t doubles, std::size_t* capacity);
  void give_slab(double* slab, std::size_t capacity);

  // A small device result vector and its pinned host mirror.
  double* result_device() { return result_dev_; }
  double* result_host() { return result_host_; }
  // Copies n doubles of the result vector to the pinned mirror and waits.
  const double* fetch_result(int n);

  // Events for one kernel launch, recorded by the dispatch itself and filed
  // under `name` in the profile, with the launch's algorithmic HBM bytes
  // (its minimum traffic); null events when profiling is off.
  LaunchEvents launch_events(const char* name, double bytes = 0.0);

  // Launch shapes chosen by a tuner, per kernel name and matrix identity,
  // with the timings they were chosen from (for the report).
  struct Tuning {
    std::string kernel;
    Int rows = 0, nnz = 0;
    std::vector<std::pair<int, double>> us;  // (shape, mean microseconds)
    int chosen = 0;
  };
  const Tuning* tuned(const std::string& kernel, std::uint64_t matrix) const {
    auto it = tuned_.find({kernel, matrix});
    return it == tuned_.end() ? nullptr : &it->second;
  }
  void record_tuning(std::uint64_t matrix, Tuning t) {
    tuned_[{t.kernel, matrix}] = std::move(t);
  }
  std::string tuning_report() const;

  // Brackets a launch sequence for the profiler, filed under `name` with the
  // sequence's algorithmic bytes; no-ops when profiling is off.
  class Timed {
   public:
    Timed(HipBackend& b, const char* name, double bytes = 0.0);
    ~Timed();
    Timed(const Timed&) = delete;
    Timed& operator=(const Timed&) = delete;

   private:
    HipBackend& b_;
    const char* name_;
    double bytes_;
    hipEvent_t start_ = nullptr;
    bool active_ = false;
  };

 private:
  struct DeviceMatrix {
    DevCsr view;
    std::vector<void*> owned;
  };
  void free_matrix(DeviceMatrix* dm);
  std::map<std::pair<std::string, std::uint64_t>, Tuning> tuned_;
  // Plans `a` for the fast order and stores the plan in `dm`.
  void replan_fast(const CsrView& a, DeviceMatrix* dm);

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<HipBackend> helper_;
  std::unordered_map<std::uint64_t, DeviceMatrix> matrices_;
  std::map<int, std::pair<double*, std::size_t>> scratch_;
  std::map<int, std::pair<double*, std::size_t>> pinned_slots_;
  std::pair<double*, std::size_t> spare_slab_{nullptr, 0};
  double* pinned_ = nullptr;
  std::size_t pinned_size_ = 0;
  double* result_dev_ = nullptr;
  double* result_host_ = nullptr;
  bool profiling_ = false;
  // Every launch is timed by an event pair recorded without waiting. Once a
  // name holds kMaxTimed outstanding pairs their times are read (one stream
  // wait) and the events reused, which bounds the event count on long solves.
  static constexpr int kMaxTimed = 4096;
  struct PendingPair {
    hipEvent_t start, stop;
    double bytes;  // algorithmic bytes of the launch this pair times
  };
  struct ProfileEntry {
    std::vector<PendingPair> pairs;
    long calls = 0;
    long timed = 0;      // launches whose time is in `ms`
    long dropped = 0;    // pairs whose time could not be read
    double ms = 0.0;
    double bytes = 0.0;  // algorithmic bytes of the timed launches
  };
  void harvest(ProfileEntry& e) const;
  mutable std::map<std::string, ProfileEntry> profile_;
  mutable std::vector<hipEvent_t> event_pool_;
  hipEvent_t take_event();
  void clear_profile();
  hipEvent_t stop_event_ = nullptr;
};

}  // namespace anuk
