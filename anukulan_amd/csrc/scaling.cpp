#include "scaling.hpp"

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace anuk {

namespace {

// Device buffers for one scaling, freed on every path out. Their host
// contents go up together (send) through the page-locked chunks.
struct Buffers {
  std::vector<void*> owned;
  std::vector<HipBackend::UploadPiece> pieces;
  ~Buffers() {
    for (void* p : owned) (void)hipFree(p);
  }
  template <class T>
  T* make(HipBackend&, size_t count, const T* host) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc(scaling)");
    owned.push_back(p);
    if (host != nullptr && count > 0) pieces.push_back({host, p, count * sizeof(T)});
    return static_cast<T*>(p);
  }
  void send(HipBackend& be) {
    be.upload_scatter(pieces);
    pieces.clear();
  }
};

template <class T>
void fetch(HipBackend& be, const T* dev, std::vector<T>* host) {
  if (host->empty()) return;
  be.download_staged(dev, host->data(), host->size() * sizeof(T));
}

}  // namespace

Scaling scale_lp_device(HipBackend& be, StandardLp* lp, int ruiz_iterations, bool pock_chambolle) {
  const Int m = lp->rows(), n = lp->cols();
  hipStream_t s = be.stream();
  Buffers b;
  const Int* ks = b.make(be, uz(m) + 1, lp->k.start.data());
  const Int* ki = b.make(be, lp->k.index.size(), lp->k.index.data());
  double* kv = b.make(be, lp->k.value.size(), lp->k.value.data());
  const Int* ts = b.make(be, uz(n) + 1, lp->kt.start.data());
  const Int* ti = b.make(be, lp->kt.index.size(), lp->kt.index.data());
  double* tv = b.make(be, lp->kt.value.size(), lp->kt.value.data());
  double* q = b.make(be, uz(m), lp->q.data());
  double* c = b.make(be, uz(n), lp->c.data());
  double* lo = b.make(be, uz(n), lp->lower.data());
  double* hi = b.make(be, uz(n), lp->upper.data());
  const std::vector<double> ones_m(uz(m), 1.0), ones_n(uz(n), 1.0);
  double* rs = b.make(be, uz(m), ones_m.data());
  double* cs = b.make(be, uz(n), ones_n.data());
  double* rn = b.make<double>(be, uz(m), nullptr);
  double* cn = b.make<double>(be, uz(n), nullptr);
  double* d1 = b.make<double>(be, uz(m), nullptr);
  double* d2 = b.make<double>(be, uz(n), nullptr);
  b.send(be);
  auto sweep = [&](bool one_norm) {
    // Both norms of the unswept matrix first (scaling.cpp:34-37): row norms
    // over K's rows, column norms over K^T's rows, which hold each column's
    // entries in K's storage order.
    launch_row_norms(s, ks, m, kv, one_norm, rn);
    launch_row_norms(s, ts, n, tv, one_norm, cn);
    launch_inv_sqrt(s, rn, m, d1);
    launch_inv_sqrt(s, cn, n, d2);
    // Rows then columns (scaling.cpp:44-45): (a * d1[i]) * d2[j], in K and
    // in K^T alike.
    launch_scale_values(s, ks, ki, m, kv, d1, d2, true);
    launch_scale_values(s, ts, ti, n, tv, d1, d2, false);
    launch_scale_vectors(s, m, d1, q, rs, n, d2, c, lo, hi, cs);
  };
  for (int i = 0; i < ruiz_iterations; ++i) sweep(false);
  if (pock_chambolle) sweep(true);
  Scaling out;
  out.row_scale.resize(uz(m));
  out.col_scale.resize(uz(n));
  fetch(be, kv, &lp->k.value);
  fetch(be, tv, &lp->kt.value);
  fetch(be, q, &lp->q);
  fetch(be, c, &lp->c);
  fetch(be, lo, &lp->lower);
  fetch(be, hi, &lp->upper);
  fetch(be, rs, &out.row_scale);
  fetch(be, cs, &out.col_scale);
  be.synchronize();
  lp->k.renew_id();
  lp->kt.renew_id();
  return out;
}

}  // namespace anuk
