/*
 * anukulan_hip.h - C ABI of the MI355X (gfx950) backend for Anukulan's
 * first-order LP method (restarted reflected-Halpern PDHG).
 *
 * Two layers sit behind this header, both in libanukulan_hip.so:
 *
 *  1. The backend operations. One entry point per virtual of the reference's
 *     LinAlgBackend (src/anukulan/backend.hpp:24-215), with the same argument
 *     lists and the same semantics: device-resident double vectors owned by
 *     the context, CSR matrices cached by a caller-supplied 64-bit identity,
 *     aliasing rules as documented there. A C++ LinAlgBackend subclass that
 *     forwards every virtual here is a complete drop-in for the reference's
 *     CudaBackend (src/anukulan/cuda_backend.cu).
 *
 *  2. Host-side LP layout and a whole device-resident PDHG solve, replacing
 *     solve_pdhg (src/anukulan/pdhg.cpp:377-1285) for callers that want the
 *     fused, launch-light loop rather than one backend call per operation.
 *
 * Conventions: every function returns ANUK_OK (0) or a nonzero status; the
 * message for the last failure on a context is anuk_last_error(ctx). No C++
 * exception crosses this boundary. Sizes are int32 (the reference's Int,
 * src/anukulan/sparse.hpp:11). A context is bound to one device and one HIP
 * stream and is not thread-safe; use one context per host thread.
 */
#ifndef ANUKULAN_HIP_H
#define ANUKULAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum anuk_status {
  ANUK_OK = 0,
  ANUK_ERR_INVALID = 1,     /* bad argument (null, negative size, shape mismatch) */
  ANUK_ERR_DEVICE = 2,      /* HIP runtime error, or no device */
  ANUK_ERR_NOMEM = 3,       /* device or host allocation failed */
  ANUK_ERR_UNSUPPORTED = 4, /* option combination this backend does not run */
  ANUK_ERR_INTERNAL = 5,
  ANUK_ERR_PARSE = 6,       /* an MPS file the reader rejects; the message is "source:line: what" */
  ANUK_ERR_IO = 7           /* a file that cannot be opened, read or written */
};

typedef struct anuk_ctx anuk_ctx;
typedef struct anuk_lp anuk_lp;

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

/* Creates a context on HIP device `device`. Replaces the process-wide
 * cuda_backend() singleton (backend.hpp:233-236, cuda_backend.cu:894-897);
 * fails with ANUK_ERR_DEVICE where the reference constructor throws
 * (cuda_backend.cu:441). */
int anuk_create(int device, anuk_ctx** out);
void anuk_destroy(anuk_ctx* ctx);
/* Message of the last failed call on ctx ("" if none). With ctx == NULL, the
 * last failure on this thread of a call that takes no context (anuk_create,
 * the anuk_lp_* functions, anuk_evaluate_residual). */
const char* anuk_last_error(const anuk_ctx* ctx);
/* LinAlgBackend::name (backend.hpp:28). */
const char* anuk_name(const anuk_ctx* ctx);
/* Blocks until every launch on the context's stream has finished. */
int anuk_synchronize(anuk_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Memory: LinAlgBackend::allocate..copy (backend.hpp:53-58)                 */
/* ------------------------------------------------------------------------ */

/* Device vector of n doubles, zero filled (test_backend.cpp:85-98 relies on
 * this). n == 0 returns NULL and ANUK_OK, as the CPU backend does
 * (backend.cpp:22-24). */
int anuk_allocate(anuk_ctx* ctx, int32_t n, double** out);
int anuk_deallocate(anuk_ctx* ctx, double* p);
int anuk_upload(anuk_ctx* ctx, const double* host, double* target, int32_t n);
int anuk_download(anuk_ctx* ctx, const double* source, double* host, int32_t n);
int anuk_fill(anuk_ctx* ctx, double* target, int32_t n, double value);
int anuk_copy(anuk_ctx* ctx, const double* source, double* target, int32_t n);

/* ------------------------------------------------------------------------ */
/* Matrices                                                                  */
/* ------------------------------------------------------------------------ */

/* A host CSR matrix as the reference stores it (sparse.hpp:37-101): rows+1
 * int32 row starts, int32 column indices sorted and unique per row, fp64
 * values. `id` is SparseMatrix::id() (sparse.hpp:51): the device copy is
 * cached under it, never under the host address. */
typedef struct anuk_csr {
  uint64_t id;
  int32_t rows;
  int32_t cols;
  const int32_t* start;
  const int32_t* index;
  const double* value;
} anuk_csr;

/* LinAlgBackend::prepare (backend.hpp:35). Idempotent per id. */
int anuk_prepare(anuk_ctx* ctx, const anuk_csr* a);
/* LinAlgBackend::release (backend.hpp:36): frees every cached matrix. */
int anuk_release(anuk_ctx* ctx);

/* y = A x (backend.hpp:61). Uploads A on first use if it was not prepared. */
int anuk_multiply(anuk_ctx* ctx, const anuk_csr* a, const double* x, double* y);
/* x = A^T y, given A^T as its own CSR (backend.hpp:62-64). */
int anuk_multiply_transpose(anuk_ctx* ctx, const anuk_csr* at, const double* y,
                            double* x);

/* ------------------------------------------------------------------------ */
/* Reductions and fused steps (backend.hpp:66-201)                           */
/* ------------------------------------------------------------------------ */

int anuk_dot(anuk_ctx* ctx, const double* a, const double* b, int32_t n, double* out);
int anuk_two_norm(anuk_ctx* ctx, const double* a, int32_t n, double* out);
int anuk_inf_norm(anuk_ctx* ctx, const double* a, int32_t n, double* out);

/* LinAlgBackend::primal_update (backend.hpp:70-93). anchor == NULL means no
 * Halpern blend. x_out may alias x. */
int anuk_primal_update(anuk_ctx* ctx, int32_t n, double tau, double gamma,
                       double halpern, const double* x, const double* c,
                       const double* kt_y, const double* lower, const double* upper,
                       const double* anchor, double* x_out, double* dx,
                       double* x_bar);

/* LinAlgBackend::dual_update (backend.hpp:95-119). anchor_y == NULL means no
 * blend (anchor_kx is then ignored). y_out may alias y, k_out may alias k_x. */
int anuk_dual_update(anuk_ctx* ctx, int32_t m, int32_t num_equalities, double sigma,
                     double gamma, double halpern, const double* y, const double* q,
                     const double* k_x_bar, const double* k_x, const double* anchor_y,
                     const double* anchor_kx, double* y_out, double* dy,
                     double* k_out, double* k_dx);

/* sum += weight * v (backend.hpp:121-123). */
int anuk_accumulate(anuk_ctx* ctx, int32_t n, double weight, const double* v,
                    double* sum);
/* out = sum / weight, true division (backend.hpp:124-126, backend.cpp:118-121). */
int anuk_scale_into(anuk_ctx* ctx, int32_t n, double weight, const double* sum,
                    double* out);

/* omega ||dx||^2 + ||dy||^2 / omega (backend.hpp:128-130). */
int anuk_weighted_norm_squared(anuk_ctx* ctx, int32_t n, int32_t m, const double* dx,
                               const double* dy, double omega, double* out);

/* interaction = |dy' k_dx|, movement = weighted norm (backend.hpp:146-148,
 * backend.cpp:131-141), both from one pass. */
int anuk_step_size_terms(anuk_ctx* ctx, int32_t n, int32_t m, const double* dx,
                         const double* dy, const double* k_dx, double omega,
                         double* interaction, double* movement);

/* LinAlgBackend::ConvergenceProblem (backend.hpp:157-165): device pointers. */
typedef struct anuk_conv_problem {
  const double* c;
  const double* lower;
  const double* upper;
  const double* q;
  const double* row_scale;
  const double* col_scale;
  int32_t num_equalities;
} anuk_conv_problem;

/* LinAlgBackend::ConvergenceTerms (backend.hpp:168-176), same field order. */
typedef struct anuk_conv_terms {
  double primal_sq;
  double primal_inf;
  double q_dot_y;
  double dual_sq;
  double dual_inf;
  double bound_term;
  double c_dot_x;
} anuk_conv_terms;

/* LinAlgBackend::convergence_terms (backend.hpp:198-201, backend.cpp:143-202). */
int anuk_convergence_terms(anuk_ctx* ctx, int32_t n, int32_t m,
                           const anuk_conv_problem* lp, const double* x,
                           const double* k_x, const double* y, const double* kt_y,
                           anuk_conv_terms* out);

/* anuk_convergence_terms without allocating: no counterpart in the reference.
 * For callers that cannot allocate device or host memory during a call
 * (code around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, same kernels and the same bits as
 * anuk_convergence_terms; the per-block partials and the 7 results live in
 * one library-wide device array of 2 x 1024 x 8 + 64 doubles (131,584
 * bytes) per device, set up when the context is created and held until the
 * process ends, and the result comes back through the context's pinned
 * buffer. The partial count is at most 1024 blocks per side whatever n and
 * m are; sizes whose partials would not fit are rejected with
 * ANUK_ERR_INVALID. Calls that overlap on one device (several contexts or
 * threads) race on the array: serialise them. */
int anuk_convergence_terms_noalloc(anuk_ctx* ctx, int32_t n, int32_t m,
                                   const anuk_conv_problem* lp, const double* x,
                                   const double* k_x, const double* y, const double* kt_y,
                                   anuk_conv_terms* out);

/* LinAlgBackend::set_profiling / profile_report (backend.hpp:213-214). The
 * report is written NUL-terminated into buf (truncated to cap); *needed gets
 * the full length including the NUL. */
int anuk_set_profiling(anuk_ctx* ctx, int on);
int anuk_profile_report(anuk_ctx* ctx, char* buf, size_t cap, size_t* needed);

/* ------------------------------------------------------------------------ */
/* Host LP layout (standard_form.hpp:23-84, sparse.cpp:24-118)               */
/* ------------------------------------------------------------------------ */

/* A model in its natural form (model.hpp:23-49, LP part only):
 *   optimise  c'x + offset   s.t.  row_lower <= A x <= row_upper,
 *                                  col_lower <= x <= col_upper
 * A is given as triplets; duplicates are summed and exact zeros dropped, as
 * SparseMatrix::from_triplets does (sparse.cpp:24-70). +-INFINITY (or any
 * magnitude >= 1e30) is an absent bound. */
typedef struct anuk_model {
  int32_t rows;
  int32_t cols;
  int64_t nnz;
  const int32_t* row_index;
  const int32_t* col_index;
  const double* value;
  const double* objective;
  const double* row_lower;
  const double* row_upper;
  const double* col_lower;
  const double* col_upper;
  int32_t maximize;
  double objective_offset;
} anuk_model;

/* to_standard_form (standard_form.cpp:72-165): equalities first, <= rows
 * negated, ranged rows split in two, free rows dropped, K^T cached. */
int anuk_lp_from_model(const anuk_model* model, anuk_lp** out);

/* A standard-form LP given directly: K as CSR (rows = num_equalities
 * equality rows first, then >= rows), c, q, bounds, and the map back to the
 * model objective (standard_form.hpp:36-48). The CSR must be valid
 * (sparse.hpp:87-88); it is copied. */
int anuk_lp_from_standard(int32_t rows, int32_t cols, int32_t num_equalities,
                          const int32_t* start, const int32_t* index,
                          const double* value, const double* c, const double* q,
                          const double* lower, const double* upper,
                          double objective_scale, double objective_offset,
                          anuk_lp** out);
void anuk_lp_free(anuk_lp* lp);

typedef struct anuk_lp_info {
  int32_t rows;
  int32_t cols;
  int32_t nnz;
  int32_t num_equalities;
  double objective_scale;
  double objective_offset;
} anuk_lp_info;
int anuk_lp_info_get(const anuk_lp* lp, anuk_lp_info* out);

/* Copies the standard form out. Any pointer may be NULL to skip it. start has
 * rows+1 entries, index/value nnz, c/lower/upper cols, q rows. row_model and
 * row_sign (rows entries) are StandardLp::row_origin (standard_form.hpp:50-55). */
int anuk_lp_export(const anuk_lp* lp, int32_t* start, int32_t* index, double* value,
                   double* c, double* q, double* lower, double* upper,
                   int32_t* row_model, double* row_sign);

/* scale_lp (scaling.cpp:75-100) on a copy of lp: Ruiz infinity-norm sweeps,
 * then one Pock-Chambolle 1-norm sweep. Writes the scaled LP to *scaled and
 * the diagonals D1 (rows) and D2 (cols). */
int anuk_lp_scale(const anuk_lp* lp, int32_t ruiz_iterations, int32_t pock_chambolle,
                  anuk_lp** scaled, double* row_scale, double* col_scale);

/* The same scaling computed on the device of ctx: the sweeps run there with
 * the host's operations in the host's order, so the result is bit for bit
 * anuk_lp_scale's. The solvers use it for matrices of 1M nonzeros and up. */
int anuk_lp_scale_device(anuk_ctx* ctx, const anuk_lp* lp, int32_t ruiz_iterations,
                         int32_t pock_chambolle, anuk_lp** scaled, double* row_scale,
                         double* col_scale);

/* ------------------------------------------------------------------------ */
/* Whole solve (pdhg.hpp:19-543)                                             */
/* ------------------------------------------------------------------------ */

enum anuk_termination {
  ANUK_TERM_FULL = 0,
  ANUK_TERM_PRIMAL_FEASIBILITY = 1,
  ANUK_TERM_DUAL_FEASIBILITY = 2
};

/* PdhgOptions (pdhg.hpp:19-415), every field the plain solve reads. Fill with
 * anuk_pdhg_default_options, which gives the reference defaults. */
typedef struct anuk_pdhg_options {
  double tolerance;                    /* 1e-6 */
  double gap_tolerance;                /* 0: same as tolerance */
  int32_t max_iterations;              /* 200000 */
  double time_limit_seconds;           /* 300 */
  int32_t termination_check_frequency; /* 40 */
  double absolute_tolerance;           /* 0 */
  int32_t require_inf_norm_termination;/* 1 */
  int32_t adaptive_step_size;          /* 1 (inactive while constant_step_size) */
  int32_t primal_weight_updates;       /* 1 */
  int32_t restarts;                    /* 1 */
  int32_t halpern;                     /* 1 */
  int32_t halpern_restart;             /* 0 */
  int32_t halpern_minimum_epoch;       /* 16 */
  int32_t polish_feasibility;          /* 1 */
  int32_t polish_first_iteration;      /* 100 */
  double polish_iteration_factor;      /* 0.125 */
  int32_t polish_minimum_iterations;   /* 40 */
  int32_t polish_maximum_iterations;   /* 20000 */
  int32_t polish_on_exit;              /* 1 */
  double reflection;                   /* 1.0 */
  int32_t constant_step_size;          /* 1 */
  double constant_step_scale;          /* 0.90 */
  int32_t restart_on_fixed_point;      /* 1 */
  int32_t pid_primal_weight;           /* 1 */
  double primal_weight_kp;             /* 0.5 */
  double primal_weight_ki;             /* 0.0 */
  double primal_weight_kd;             /* 0.3 */
  int32_t ruiz_iterations;             /* 10 */
  int32_t pock_chambolle;              /* 1 */
  int32_t reuse_products;              /* 1 */
  int32_t detect_infeasibility;        /* 1 */
  double infeasibility_tolerance;      /* 1e-8 */
  int32_t termination;                 /* ANUK_TERM_FULL */
  const double* warm_x;                /* NULL, or cols entries (see pdhg.hpp:386-391) */
  const double* warm_y;                /* NULL, or rows entries */
  double known_matrix_norm;            /* 0: estimate */
  double initial_step_size;            /* 0: derive */
  double initial_primal_weight;        /* 0: derive */
  int32_t verbose;                     /* 0 */
  int32_t log_frequency;               /* 5000 */
  const volatile int32_t* stop;        /* NULL; nonzero stops at the next check */
  int32_t use_graphs;                  /* 1: replay each check period as a HIP graph */
  int32_t product_order;               /* 0: automatic (storage order below 4M nonzeros,
                                          fastest above), 1: every row sum serially in
                                          CSR storage order (the CPU's bits), 2: fastest
                                          (long rows summed by a wave in a fixed tree,
                                          large K x split into per-XCD column panels);
                                          deterministic either way */
  int32_t autotune;                    /* 1: the batched solver times its launch shapes
                                          (waves per workgroup: 1, 2, 4, 8; rows per
                                          wave: 1, 2, 4; the matrix read from its CSR or
                                          from a fixed-width row copy; and the waves per
                                          workgroup of the binned launches of a narrow
                                          span) on the stack's K
                                          and K^T once per matrix and keeps the fastest
                                          when it beats the default shape by 3%; every
                                          shape computes the same bits. 0: the default
                                          shape */
  int32_t batch_waves;                 /* 0: as autotune decides; 1, 2, 4 or 8: the batched
                                          step kernels use this many waves per workgroup */
  int32_t batch_rows;                  /* 0: as autotune decides; 1, 2 or 4: consecutive
                                          rows per wave of the batched step kernels */
  int32_t batch_layout;                /* 0: as autotune decides; 1: the batched step
                                          kernels read K and K^T from their CSR; 2: from a
                                          fixed-width copy of each row's first entries
                                          (one row per wave only) */
  int32_t batch_pack;                  /* 0: once 32 slots or fewer are live, a batched
                                          step takes its rows of 8 entries or fewer 2 or
                                          4 to a wave (a lane per row and slot) and each
                                          longer row a wave of its own; 1: one row per
                                          wave always. The bits are the same */
  int32_t multi_order;                 /* anuk_solve_pdhg_multi: 0: as autotune decides;
                                          1: row blocks launched in the stack's order; 2:
                                          binned (each XCD's share of the blocks run
                                          heaviest first) */
  int32_t multi_stream_loads;          /* anuk_solve_pdhg_multi: 0: as autotune decides;
                                          1: plain loads of the matrices' index and value
                                          streams; 2: non-temporal loads. With autotune
                                          the solver times the four choices on the
                                          stack's K and K^T once per stack pattern and
                                          keeps the fastest when it beats the stack's
                                          order with plain loads by 3%; every choice
                                          computes the same bits */
  int32_t small_lp_engine;             /* anuk_solve_pdhg: 0: an LP of at most 30,000
                                          nonzeros whose rows of K and K^T hold at most 32
                                          entries, with options the batched solver runs,
                                          is solved as a batch of one (its launches give
                                          each row a wave: 1.15-1.4x faster on such LPs,
                                          with the same decisions); 1: the plain fused
                                          kernels always */
} anuk_pdhg_options;

void anuk_pdhg_default_options(anuk_pdhg_options* options);

enum anuk_pdhg_status {
  ANUK_PDHG_OPTIMAL = 0,
  ANUK_PDHG_PRIMAL_INFEASIBLE = 1,
  ANUK_PDHG_ITERATION_LIMIT = 2,
  ANUK_PDHG_TIME_LIMIT = 3,
  ANUK_PDHG_NUMERICAL_ERROR = 4
};

/* PdhgResidual (pdhg.hpp:429-496), the numbers only. */
typedef struct anuk_pdhg_residual {
  double primal_residual;
  double dual_residual;
  double primal_residual_inf;
  double dual_residual_inf;
  double primal_objective;
  double dual_objective;
  double absolute_gap;
  double relative_primal;
  double relative_dual;
  double relative_gap;
  double relative_primal_inf;
  double relative_dual_inf;
} anuk_pdhg_residual;

/* PdhgResult (pdhg.hpp:498-541) without the vectors, which go to x_out/y_out. */
typedef struct anuk_pdhg_result {
  int32_t status;
  int32_t iterations;
  int32_t restarts;
  int32_t polish_attempts;
  int32_t polish_iterations;
  int32_t polished;
  double objective;
  double solve_seconds;
  double matrix_norm_estimate;
  double final_step_size;
  double final_primal_weight;
  double phase_scaling;
  double phase_matrix_norm;
  double phase_setup;
  double phase_iterating;
  double phase_checks;
  double phase_polishing;
  anuk_pdhg_residual residual;
  char message[128];
} anuk_pdhg_result;

/* solve_pdhg (pdhg.hpp:543, pdhg.cpp:377-1285) on the device of ctx.
 * x_out (cols doubles) and y_out (rows doubles) receive the unscaled
 * primal-dual pair; either may be NULL. */
int anuk_solve_pdhg(anuk_ctx* ctx, const anuk_lp* lp, const anuk_pdhg_options* options,
                    anuk_pdhg_result* result, double* x_out, double* y_out);

/* Batched solve of `count` LPs that share one constraint matrix: the same
 * rows, columns, equality count and CSR arrays (a base plan and its what-if
 * cases), differing in c, q and bounds. No reference counterpart: the
 * reference solves such a stack one LP at a time (solve_pdhg per case). Each
 * LP gets the result its own anuk_solve_pdhg gives: the same steps, bit for
 * bit, and the same decisions and iteration count wherever the check's
 * reductions (folded in a different fixed tree, rounding-level apart) do
 * not tip one; results[i], its primal in x_out + i * cols and its dual in
 * y_out + i * rows (either may be NULL). Runs the options the fused loop runs (constant step size, Halpern
 * or no restarts, fixed-point restart test); others fail with
 * ANUK_ERR_UNSUPPORTED. options->warm_x / warm_y must be NULL. */
int anuk_solve_pdhg_batch(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                          const anuk_pdhg_options* options, anuk_pdhg_result* results,
                          double* x_out, double* y_out);

/* Solves `count` independent LPs, each with its own constraint matrix and
 * of any size, in one set of launches: the scaled matrices are concatenated
 * block-diagonally and every period of fused steps runs over the whole
 * stack, each LP under its own control (as its anuk_solve_pdhg would), an
 * LP that has stopped dropping out of the launches. x_out receives the
 * primals end to end (sum of cols), y_out the duals (sum of rows); either may
 * be NULL. Runs the options anuk_solve_pdhg_batch runs. No reference
 * counterpart: the reference solves LPs one at a time. */
int anuk_solve_pdhg_multi(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                          const anuk_pdhg_options* options, anuk_pdhg_result* results,
                          double* x_out, double* y_out);

/* ------------------------------------------------------------------------ */
/* Branch-and-bound node relaxations                                          */
/* ------------------------------------------------------------------------ */

/* NodeSolution (branch_and_bound.cpp:44-57) of a first-order node solve, the
 * vectors aside: `optimal` when the relaxation converged with its dual
 * feasible to 1e-7 * max(1, |dual objective|), and then `bound` is the dual
 * objective (standard form, a lower bound on the node's subtree); otherwise
 * `bound` is -infinity. `proved_infeasible` when the solve ended with a
 * Farkas certificate. `relaxation` is the node's own solve. */
typedef struct anuk_node_result {
  int32_t optimal;
  int32_t proved_infeasible;
  double bound;
  anuk_pdhg_result relaxation;
} anuk_node_result;

/* Replaces the first-order branch of the reference tree's solve_node
 * (branch_and_bound.cpp:528-566, one solve_pdhg per node) for `count` nodes
 * at once. Node i is `lp` with column bounds lower + i * cols and
 * upper + i * cols (standard-form columns, as Node::lower/upper,
 * branch_and_bound.cpp:17-19); the nodes share lp's matrix, so options the
 * batched loop runs solve them together (anuk_solve_pdhg_batch), and other
 * options solve them one after another with the plain device solver. Every
 * node starts cold, as the reference's does (it passes warm_x without
 * warm_y, which pdhg.cpp:481-484 treats as no warm start): options->warm_x /
 * warm_y must be NULL. x_out (count * cols) and y_out (count * rows)
 * receive each node's unscaled primal and dual; either may be NULL. */
int anuk_solve_node_relaxations(anuk_ctx* ctx, const anuk_lp* lp, int32_t count,
                                const double* lower, const double* upper,
                                const anuk_pdhg_options* options, anuk_node_result* results,
                                double* x_out, double* y_out);

/* ------------------------------------------------------------------------ */
/* Exact finish: crossover and a warm-started simplex on the host            */
/* ------------------------------------------------------------------------ */

/* CrossoverOptions (crossover.hpp:37-83) and the simplex's own settings.
 * Tolerances are in the units of the equilibrated LP the simplex runs on. */
typedef struct anuk_finish_options {
  double primal_tolerance;             /* 1e-9 */
  double dual_tolerance;               /* 1e-9 */
  double pivot_tolerance;              /* 1e-9: smallest |alpha| the ratio test divides by */
  int32_t max_iterations;              /* 200000 */
  int32_t refactor_interval;           /* 100 */
  int32_t crossover;                   /* 1: start from the point's crossover basis; 0: from
                                          all logicals (a cold start) */
  double interior_tolerance;           /* 1e-6 (CrossoverOptions::interior_tolerance) */
  double crossover_pivot_tolerance;    /* 1e-7 (CrossoverOptions::pivot_tolerance) */
  int32_t max_candidates;              /* -1: no cap */
  int32_t crossover_refactor_interval; /* 50 */
  double dual_weight;                  /* 1.0 */
  int32_t ruiz_iterations;             /* 10: equilibration sweeps before the simplex */
  int32_t threads;                     /* 0: one case per host core, at most 16 */
} anuk_finish_options;

void anuk_finish_default_options(anuk_finish_options* options);

enum anuk_finish_status {
  ANUK_FINISH_OPTIMAL = 0,
  ANUK_FINISH_INFEASIBLE = 1,
  ANUK_FINISH_UNBOUNDED = 2,
  ANUK_FINISH_ITERATION_LIMIT = 3,
  ANUK_FINISH_NUMERICAL_ERROR = 4
};

/* SimplexResult (simplex.hpp:385-431) and CrossoverResult
 * (crossover.hpp:86-111), the numbers only. */
typedef struct anuk_finish_result {
  int32_t status;                      /* anuk_finish_status */
  int32_t iterations;                  /* simplex pivots and bound flips */
  int32_t phase_one_iterations;
  int32_t refactorizations;
  int32_t bland_iterations;            /* pivots taken under Bland's rule (stalling) */
  int32_t started_warm;                /* the crossover pushed at least one column */
  int32_t candidates;                  /* crossover: columns the point has interior */
  int32_t pushed;                      /* crossover: structurals basic in the handed-over basis */
  int32_t rejected_small_pivot;
  int32_t logicals_remaining;
  int32_t stopped_on_refactorization;
  double objective;                    /* model sense */
  double primal_violation;             /* on the LP as given: max row / bound violation */
  double dual_violation;               /* max reduced-cost sign violation at the basis */
  double seconds;
  char message[64];
} anuk_finish_result;

/* Finishes `count` LPs exactly: for each, the crossover of its point
 * (crossover_basis, crossover.cpp:35-229, the same candidate, merit and
 * bound rules) gives a starting basis, and a bounded primal simplex of this
 * library takes it to an optimal vertex, as the reference does after its
 * first-order seed (app/anukulan_cli.cpp:592-631). x holds the points end to
 * end (sum of cols; NULL: every LP starts cold from all logicals), y the
 * duals end to end (sum of rows; NULL: nonbasic bounds by proximity).
 * x_out / y_out receive the vertices and row duals end to end, basic_out
 * (sum of rows) each basis position's column (j < cols: structural j;
 * cols + i: row i's logical); any may be NULL. Runs on the host, one LP per
 * thread; it needs no device. */
int anuk_exact_finish(const anuk_lp* const* lps, int32_t count, const double* x, const double* y,
                      const anuk_finish_options* options, anuk_finish_result* results,
                      double* x_out, double* y_out, int32_t* basic_out);

/* The same, with the case-stack workflow of the reference
 * (app/anukulan_cli.cpp:1065-1230, `family --from-first`): given a start
 * basis (start_basic: rows entries as basic_out gives them; start_state:
 * cols + rows entries, 0 basic, 1 at lower, 2 at upper, 3 free, as
 * SimplexOptions::start_basic / start_status, simplex.hpp:295-301), every LP
 * starts its simplex from that basis instead of a crossover of its point,
 * and so all LPs must have the same shape. The reference runs its dual
 * simplex from the base plan's basis; here it is this library's primal
 * simplex, whose composite phase one takes a basis that a case's new bounds
 * or right-hand side made infeasible. A nonbasic state a case's bounds
 * cannot hold moves to a bound they can. start_basic and start_state are
 * both NULL or both given; with NULL this is anuk_exact_finish. state_out
 * (sum of cols + rows, NULL allowed) receives each LP's final states in the
 * same codes, so one LP's result can warm-start the others. */
int anuk_exact_finish_basis(const anuk_lp* const* lps, int32_t count, const double* x,
                            const double* y, const int32_t* start_basic,
                            const int32_t* start_state, const anuk_finish_options* options,
                            anuk_finish_result* results, double* x_out, double* y_out,
                            int32_t* basic_out, int32_t* state_out);

/* ------------------------------------------------------------------------ */
/* Sensitivities of a solution (no reference counterpart)                    */
/* ------------------------------------------------------------------------ */

enum anuk_adjoint_method {
  ANUK_ADJOINT_CG = 0,          /* CGLS on the fixed-point equations */
  ANUK_ADJOINT_FIXED_POINT = 1  /* the transposed PDHG map iterated as is */
};

/* Where the final active set comes from. */
enum anuk_adjoint_active_set {
  /* Read off the point; where that leaves the system non-square (a
   * degenerate vertex: columns at a bound that the basis keeps), complete it
   * to an optimal basis by the exact finish from the point (anuk_exact_finish:
   * crossover, then the simplex) and take the active set from that basis. */
  ANUK_ACTIVE_AUTO = 0,
  ANUK_ACTIVE_POINT = 1,  /* read off the point only (active_tolerance) */
  ANUK_ACTIVE_BASIS = 2   /* always completed to an optimal basis */
};

typedef struct anuk_adjoint_options {
  int32_t method;           /* ANUK_ADJOINT_CG */
  int32_t max_iterations;   /* 20000 */
  double tolerance;         /* 1e-11: CG stops at this relative normal-equation
                               residual; the fixed-point method reports
                               converged when both residuals are below its
                               square root */
  double active_tolerance;  /* 1e-6: a column within this * (1 + |bound|) of a
                               bound is at it; a row within this * (1 + |q|)
                               of its rhs binds */
  double primal_weight;     /* 1.0: omega of the fixed-point method's steps,
                               tau = 0.9 / (||K~|| omega), sigma = 0.9 omega / ||K~|| */
  int32_t active_set;       /* ANUK_ACTIVE_AUTO */
} anuk_adjoint_options;

typedef struct anuk_adjoint_info {
  int32_t iterations;
  int32_t converged;
  double residual_primal;   /* ||B(B'mu - g)|| / ||B g|| on the active set */
  double residual_dual;     /* ||B'(B nu - h)|| / ||B' h|| */
  int32_t free_columns;
  int32_t binding_rows;
  int32_t lower_columns;
  int32_t upper_columns;
  int32_t basis_completed;  /* 1: the active set came from an optimal basis */
} anuk_adjoint_info;

void anuk_adjoint_default_options(anuk_adjoint_options* options);

/* Gradients of L = w_x' x + w_y' y at the PDHG solutions (x, y) of `count`
 * LPs that share K, with respect to each LP's c, q, lower and upper (all in
 * standard form, unscaled), by implicit differentiation of the PDHG fixed
 * point on the final active set: the transposed map's fixed point, reached
 * by the method in options, one wave64 lane per LP. `scaling` gives the
 * solve's ruiz_iterations and pock_chambolle (the space the map lives in).
 * x is count * cols, y count * rows, in LP order; w_x / w_y likewise or NULL
 * for a zero seed; any gradient pointer or info may be NULL. No counterpart
 * in the reference, which has no sensitivities; for callers that want
 * d(x*, y*)/d(c, q, bounds) without re-solving. */
int anuk_pdhg_adjoint(anuk_ctx* ctx, const anuk_lp* const* lps, int32_t count,
                      const anuk_pdhg_options* scaling, const anuk_adjoint_options* options,
                      const double* x, const double* y, const double* w_x, const double* w_y,
                      double* grad_c, double* grad_q, double* grad_lower, double* grad_upper,
                      anuk_adjoint_info* info);

/* The launch shapes the tuners measured and chose on this context, one line
 * per tuned kernel and matrix:
 *   <kernel> rows=<r> nnz=<z> us@<shape>=<t> ... chosen=<shape>
 * For the batched steps (batch_primal_step, batch_dual_step) a shape is
 * w<waves per workgroup>r<rows per wave>, with a trailing "s" for the
 * fixed-width row copy; for the uneven stack's products (multi_k_product,
 * multi_kt_product) it is "plan" or "binned" (the row blocks' launch
 * order), with "+nt" for non-temporal index and value loads.
 * Same buffer protocol as anuk_profile_report. No reference counterpart. */
int anuk_tuning_report(anuk_ctx* ctx, char* buffer, size_t capacity, size_t* needed);

/* estimate_matrix_norm (pdhg.hpp:580-581, pdhg.cpp:223-282) of lp's K as
 * given (no scaling), by power iteration on the device. */
int anuk_estimate_matrix_norm(anuk_ctx* ctx, const anuk_lp* lp, int32_t iterations,
                              double tolerance, double* out);

/* evaluate_residual (pdhg.hpp:548-549, pdhg.cpp:107-174) on the host. */
int anuk_evaluate_residual(const anuk_lp* lp, const double* x, const double* y,
                           anuk_pdhg_residual* out);

/* ------------------------------------------------------------------------ */
/* MPS files: read_mps / read_mps_stream (mps_reader.hpp:13-50,              */
/* mps_reader.cpp:711-728) and a writer the reference does not have         */
/* ------------------------------------------------------------------------ */

typedef struct anuk_mps anuk_mps;

/* MpsOptions (mps_reader.hpp:12-30). format: 0 decides from the ROWS
 * section, 1 free, 2 fixed. negative_up_bound: 0 keeps the lower bound at 0
 * (HiGHS), 1 sets it to -infinity on a continuous column (CPLEX). BOUNDS
 * values at or beyond infinity_threshold in magnitude are infinite. */
typedef struct anuk_mps_options {
  int32_t format;
  int32_t negative_up_bound;
  double infinity_threshold;
} anuk_mps_options;
void anuk_mps_default_options(anuk_mps_options* out);

/* Reads an MPS (or QPS) file into a model owned by *out. A file the reader
 * rejects returns ANUK_ERR_PARSE and the message of MpsReadResult::error
 * ("source:line: what", mps_reader.cpp:164-168), read with
 * anuk_last_error(NULL); a file that cannot be opened returns ANUK_ERR_IO.
 * options may be NULL for the defaults. */
int anuk_mps_read(const char* path, const anuk_mps_options* options, anuk_mps** out);
/* The same from `length` bytes of text in memory (read_mps_stream); `source`
 * names it in messages. */
int anuk_mps_read_text(const char* text, int64_t length, const char* source,
                       const anuk_mps_options* options, anuk_mps** out);
void anuk_mps_free(anuk_mps* mps);

/* Sizes and the reader's counters (MpsReadResult, mps_reader.hpp:32-42). */
typedef struct anuk_mps_info {
  int32_t rows;
  int32_t cols;
  int64_t nnz;
  int64_t hessian_nnz;
  int32_t maximize;
  int32_t fixed_format;
  double objective_offset;
  int32_t free_rows_dropped;
  int32_t warnings;
  int64_t objective_nonzeros;
} anuk_mps_info;
int anuk_mps_info_get(const anuk_mps* mps, anuk_mps_info* out);

/* The model as an anuk_model whose arrays point into `mps` (valid until
 * anuk_mps_free): triplets in CSR order, ready for anuk_lp_from_model.
 * Column types and the Hessian are not part of anuk_model; the LP built from
 * it is the continuous relaxation. */
int anuk_mps_model(const anuk_mps* mps, anuk_model* out);
/* A (rows x cols) and the Hessian Q (cols x cols, full symmetric, empty for
 * an LP) as CSR arrays inside `mps`; any pointer may be NULL. */
int anuk_mps_matrix(const anuk_mps* mps, const int32_t** start, const int32_t** index,
                    const double** value);
int anuk_mps_hessian(const anuk_mps* mps, const int32_t** start, const int32_t** index,
                     const double** value);
/* Column types (VarType order, model.hpp:12: 0 continuous, 1 integer,
 * 2 semi-continuous, 3 semi-integer), cols entries inside `mps`. */
const int32_t* anuk_mps_col_types(const anuk_mps* mps);
/* Names inside `mps`; NULL when out of range. */
const char* anuk_mps_name(const anuk_mps* mps);
const char* anuk_mps_row_name(const anuk_mps* mps, int32_t row);
const char* anuk_mps_col_name(const anuk_mps* mps, int32_t col);
const char* anuk_mps_warning(const anuk_mps* mps, int32_t k);

/* Writes `model` as free-format MPS that anuk_mps_read turns back into the
 * same model, bit for bit (values in their shortest exact form). row_names,
 * col_names and col_types may be NULL (names R<i> / C<j>, all continuous);
 * names may not hold white space. MPS cannot state a free row (written as an
 * N row, which the reader drops), a finite bound at or beyond 1e30, or a
 * ranged row whose bounds no range reproduces exactly. No counterpart in the
 * reference, which has no writer; for callers that store generated stacks. */
int anuk_mps_write(const char* path, const anuk_model* model, const char* name,
                   const char* const* row_names, const char* const* col_names,
                   const int32_t* col_types);

/* ------------------------------------------------------------------------ */
/* Presolve of a case stack in front of the batch: the reference's           */
/* presolve(const Model&, const PresolveOptions&) (presolve.hpp:485) with    */
/* PostsolveStack::apply / apply_dual (presolve.hpp:389-396), taken for a    */
/* stack of cases that share A so that the reduced cases still share one.    */
/* ------------------------------------------------------------------------ */

typedef struct anuk_presolve anuk_presolve;

/* Which reductions may fire; the switches named as in PresolveOptions
 * (presolve.hpp:59-70, 157) stand for the same reduction. All default to on
 * except slack_column_singletons: each such substitution makes its equality
 * row a ranged row, two opposed rows of the standard form, which on the
 * refinery case stack costs the PDHG solves more iterations than the removed
 * columns save (509,680 -> 683,360 lane iterations to 1e-4 over the 61
 * cases of the 24-period stack). tolerance is the
 * feasibility slack relative to max(1, |bound|); a column alone in an
 * equality row is substituted out only when its coefficient is at least
 * pivot_ratio times the largest magnitude in the row. */
typedef struct anuk_presolve_rules {
  int32_t empty_rows;
  int32_t singleton_rows;
  int32_t redundant_rows;
  int32_t fixed_columns;
  int32_t empty_columns;
  int32_t slack_column_singletons;
  int32_t max_rounds;
  double tolerance;
  double pivot_ratio;
} anuk_presolve_rules;
void anuk_presolve_default_rules(anuk_presolve_rules* out);

enum anuk_presolve_status {
  ANUK_PRESOLVE_REDUCED = 0,
  ANUK_PRESOLVE_INFEASIBLE = 1, /* some case proved infeasible: failing_case names it */
  ANUK_PRESOLVE_UNBOUNDED = 2   /* some case has a column in no row that its cost drives to infinity */
};

/* Sizes before and after, and how often each reduction fired
 * (PresolveCounts, presolve.hpp:432-463, for the reductions made here). */
typedef struct anuk_presolve_info {
  int32_t status; /* anuk_presolve_status */
  int32_t failing_case;
  int32_t rows, cols;
  int64_t nnz;
  int32_t original_rows, original_cols;
  int64_t original_nnz;
  int32_t empty_rows, free_rows, singleton_rows, redundant_rows, redundant_sides;
  int32_t fixed_columns, empty_columns, slack_column_singletons;
  int32_t rounds;
} anuk_presolve_info;

/* Presolves `cases` LPs in minimisation form that share A (rows x cols, CSR
 * with strictly increasing columns in each row and no explicit zeros). Case
 * k's objective and column bounds start at k*cols, its row bounds at k*rows;
 * offset holds one objective constant per case (NULL: zeros). Infinite
 * bounds are +-HUGE_VAL. A structural reduction is taken only when every case
 * allows it, so the reduced cases share one matrix again. *out owns the
 * reduced stack and what postsolve needs; an infeasible or unbounded verdict
 * is a status, not an error. rules may be NULL for the defaults. */
int anuk_presolve_stack(int32_t rows, int32_t cols, const int32_t* start, const int32_t* index,
                        const double* value, int32_t cases, const double* objective,
                        const double* row_lower, const double* row_upper,
                        const double* col_lower, const double* col_upper, const double* offset,
                        const anuk_presolve_rules* rules, anuk_presolve** out);
void anuk_presolve_free(anuk_presolve* p);
int anuk_presolve_get_info(const anuk_presolve* p, anuk_presolve_info* out);
/* Why the status is not REDUCED ("case k: ..."), empty otherwise. */
const char* anuk_presolve_message(const anuk_presolve* p);

/* The reduced stack, laid out like the input (start: rows+1, index and value:
 * nnz, objective and column bounds: cases*cols, row bounds: cases*rows,
 * offset: cases) and, for each reduced row and column, the original one it
 * is. Any pointer may be NULL. */
int anuk_presolve_reduced(const anuk_presolve* p, int32_t* start, int32_t* index, double* value,
                          double* objective, double* row_lower, double* row_upper,
                          double* col_lower, double* col_upper, double* offset,
                          int32_t* row_origin, int32_t* col_origin);

/* Case `case_index`'s solution of the reduced LP (reduced_x: reduced cols,
 * reduced_y: one dual per reduced row) mapped back to the original LP
 * (x: original cols, y: original rows). Row duals are positive on a row held
 * at its lower side and negative at its upper side, reduced costs are
 * c - A'y; every reduction made here inverts exactly on the dual side. */
int anuk_postsolve(const anuk_presolve* p, int32_t case_index, const double* reduced_x,
                   const double* reduced_y, double* x, double* y);

#ifdef __cplusplus
}
#endif

#endif /* ANUKULAN_HIP_H */
