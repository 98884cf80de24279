// Test infrastructure: the LinAlgBackend adapter a maintainer would add to the
// reference as src/anukulan/hip_backend.cpp, compiled against the reference's
// own headers and linked with libanukulan_hip.so, so that the reference's
// solve_pdhg (pdhg.cpp:377-1285) and its backend contract
// (tests/test_backend.cpp:713-893) run on the HIP library unchanged.
//
// Besides hip_backend(), it binds the reference's GPU backend slot:
// cuda_backend() (backend.hpp:233-236), which the reference's own tests and
// default_backend() call when built with ANUKULAN_WITH_CUDA. Only tests load
// what is built from this file.
#include <stdexcept>
#include <string>

#include "anukulan/backend.hpp"
#include "anukulan_hip.h"

namespace anukulan {
namespace {

void ok(anuk_ctx* ctx, int status, const char* what) {
  if (status != 0)
    throw std::runtime_error(std::string(what) + ": " + anuk_last_error(ctx));
}

anuk_csr view(const SparseMatrix& a) {
  return anuk_csr{a.id(), a.rows(), a.cols(), a.start().data(), a.index().data(),
                  a.value().data()};
}

class HipBackend final : public LinAlgBackend {
 public:
  explicit HipBackend(int device) {
    if (anuk_create(device, &ctx_) != 0)
      throw std::runtime_error(std::string("hip backend: ") + anuk_last_error(nullptr));
  }
  ~HipBackend() override { anuk_destroy(ctx_); }

  std::string name() const override { return anuk_name(ctx_); }

  void prepare(const SparseMatrix& a) const override {
    const anuk_csr v = view(a);
    ok(ctx_, anuk_prepare(ctx_, &v), "prepare");
  }
  void release() const override { ok(ctx_, anuk_release(ctx_), "release"); }

  double* allocate(Int n) const override {
    double* p = nullptr;
    ok(ctx_, anuk_allocate(ctx_, n, &p), "allocate");
    return p;
  }
  void deallocate(double* p) const override { ok(ctx_, anuk_deallocate(ctx_, p), "deallocate"); }
  void upload(const double* h, double* d, Int n) const override {
    ok(ctx_, anuk_upload(ctx_, h, d, n), "upload");
  }
  void download(const double* d, double* h, Int n) const override {
    ok(ctx_, anuk_download(ctx_, d, h, n), "download");
  }
  void fill(double* t, Int n, double v) const override { ok(ctx_, anuk_fill(ctx_, t, n, v), "fill"); }
  void copy(const double* s, double* t, Int n) const override {
    ok(ctx_, anuk_copy(ctx_, s, t, n), "copy");
  }

  void multiply(const SparseMatrix& a, const double* x, double* y) const override {
    const anuk_csr v = view(a);
    ok(ctx_, anuk_multiply(ctx_, &v, x, y), "multiply");
  }
  void multiply_transpose(const SparseMatrix& at, const double* y, double* x) const override {
    const anuk_csr v = view(at);
    ok(ctx_, anuk_multiply_transpose(ctx_, &v, y, x), "multiply_transpose");
  }

  double dot(const double* a, const double* b, Int n) const override {
    double r = 0;
    ok(ctx_, anuk_dot(ctx_, a, b, n, &r), "dot");
    return r;
  }
  double two_norm(const double* a, Int n) const override {
    double r = 0;
    ok(ctx_, anuk_two_norm(ctx_, a, n, &r), "two_norm");
    return r;
  }
  double inf_norm(const double* a, Int n) const override {
    double r = 0;
    ok(ctx_, anuk_inf_norm(ctx_, a, n, &r), "inf_norm");
    return r;
  }

  void primal_update(Int n, double tau, double gamma, double h, const double* x,
                     const double* c, const double* kt_y, const double* lo, const double* up,
                     const double* anchor, double* x_out, double* dx,
                     double* x_bar) const override {
    ok(ctx_, anuk_primal_update(ctx_, n, tau, gamma, h, x, c, kt_y, lo, up, anchor, x_out, dx,
                                x_bar), "primal_update");
  }
  void dual_update(Int m, Int neq, double sigma, double gamma, double h, const double* y,
                   const double* q, const double* k_x_bar, const double* k_x,
                   const double* anchor_y, const double* anchor_kx, double* y_out, double* dy,
                   double* k_out, double* k_dx) const override {
    ok(ctx_, anuk_dual_update(ctx_, m, neq, sigma, gamma, h, y, q, k_x_bar, k_x, anchor_y,
                              anchor_kx, y_out, dy, k_out, k_dx), "dual_update");
  }
  void accumulate(Int n, double w, const double* v, double* sum) const override {
    ok(ctx_, anuk_accumulate(ctx_, n, w, v, sum), "accumulate");
  }
  void scale_into(Int n, double w, const double* sum, double* out) const override {
    ok(ctx_, anuk_scale_into(ctx_, n, w, sum, out), "scale_into");
  }
  double weighted_norm_squared(Int n, Int m, const double* dx, const double* dy,
                               double omega) const override {
    double r = 0;
    ok(ctx_, anuk_weighted_norm_squared(ctx_, n, m, dx, dy, omega, &r), "weighted_norm_squared");
    return r;
  }
  void step_size_terms(Int n, Int m, const double* dx, const double* dy, const double* k_dx,
                       double omega, double* interaction, double* movement) const override {
    ok(ctx_, anuk_step_size_terms(ctx_, n, m, dx, dy, k_dx, omega, interaction, movement),
       "step_size_terms");
  }
  void convergence_terms(Int n, Int m, const ConvergenceProblem& p, const double* x,
                         const double* k_x, const double* y, const double* kt_y,
                         ConvergenceTerms* out) const override {
    const anuk_conv_problem cp{p.c, p.lower, p.upper, p.q, p.row_scale, p.col_scale,
                               p.num_equalities};
    anuk_conv_terms t{};
    ok(ctx_, anuk_convergence_terms(ctx_, n, m, &cp, x, k_x, y, kt_y, &t), "convergence_terms");
    *out = ConvergenceTerms{t.primal_sq, t.primal_inf, t.q_dot_y, t.dual_sq,
                            t.dual_inf,  t.bound_term, t.c_dot_x};
  }

  void set_profiling(bool on) const override { ok(ctx_, anuk_set_profiling(ctx_, on), "profiling"); }
  std::string profile_report() const override {
    size_t need = 0;
    ok(ctx_, anuk_profile_report(ctx_, nullptr, 0, &need), "profile_report");
    std::string s(need, '\0');
    ok(ctx_, anuk_profile_report(ctx_, s.data(), s.size(), &need), "profile_report");
    s.resize(need ? need - 1 : 0);
    return s;
  }

 private:
  anuk_ctx* ctx_ = nullptr;
};

}  // namespace

const LinAlgBackend& hip_backend() {
  static HipBackend backend(0);  // one per process, as cuda_backend() is
  return backend;
}

#ifdef ANUKULAN_WITH_CUDA
// The reference's GPU slot (backend.hpp:233-236), bound to the HIP library.
const LinAlgBackend& cuda_backend() { return hip_backend(); }
#endif

}  // namespace anukulan
