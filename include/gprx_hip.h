/*
 * gprx_hip.h -- C ABI of the MI355X (gfx950) backend for gprx's GP hot path.
 *
 * The entry points replace the objective seam of the gprx crate
 * (src/objective.rs:308-340, `Differentiable::value_and_gradient_into`) and
 * the model calls around it.  Plain pointers and sizes only; nothing unwinds
 * across this boundary; every call returns a gprx_status.  A handle owns its
 * device buffers and one HIP stream and is not thread-safe: use one host
 * thread per handle (the reference's models are `&mut self` the same way,
 * src/gpr/objective.rs:14-26).
 *
 * Conventions shared with the reference:
 *   - x is column-major n x d: x[k*n + i] (README.md:7).
 *   - Parameters are log-values, ordered [kernel leaves depth-first |
 *     log sigma^2] (src/kernel/spec.rs:29, src/gpr/objective.rs:10).
 *   - The value is the negative log marginal likelihood including
 *     (n/2) log 2 pi (src/gpr/factor.rs:254-267); the gradient is in log space.
 *   - On failure the model keeps its previous parameters and factor
 *     (src/gpr/exact_fit.rs:196-252).
 */
#ifndef GPRX_HIP_H
#define GPRX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror the GprError variants the path can raise
 * (src/error.rs; docs/design.md:799-827). */
typedef enum gprx_status {
  GPRX_OK = 0,
  GPRX_CHOLESKY_FAILED = 1,          /* GprError::CholeskyFailed */
  GPRX_NON_FINITE_KERNEL_VALUE = 2,  /* GprError::NonFiniteKernelValue */
  GPRX_NON_FINITE_INPUT = 3,         /* GprError::NonFiniteInput */
  GPRX_INVALID_HYPERPARAMETER = 4,   /* GprError::InvalidHyperparameter */
  GPRX_INVALID_NOISE_VARIANCE = 5,   /* GprError::InvalidNoiseVariance */
  GPRX_LENGTH_MISMATCH = 6,          /* GprError::LengthMismatch */
  GPRX_EMPTY_INPUT = 7,              /* GprError::EmptyInput */
  GPRX_UNSUPPORTED = 8,              /* a program or size this backend does not take */
  GPRX_DEVICE_ERROR = 9,             /* a HIP runtime failure */
  GPRX_NOT_CONVERGED = 10,           /* GprError::OptimizationNotConverged */
  GPRX_INVALID_CONFIG = 11,          /* GprError::InvalidConfig */
  GPRX_INVALID_INTERVAL = 12,        /* GprError::InvalidInterval (a start outside its bounds) */
  GPRX_INSUFFICIENT_DATA = 13        /* GprError::InsufficientData (a deletion leaving < 1 point) */
} gprx_status;

/* CholeskyStage of GprError::CholeskyFailed (src/error.rs). */
typedef enum gprx_stage { GPRX_STAGE_FIT = 0, GPRX_STAGE_PREDICT = 1 } gprx_stage;

typedef struct gprx_error {
  int32_t status;
  int32_t stage;
  double jitter;      /* last jitter tried (CholeskyFailed) */
  size_t matrix_size; /* order of the failed factor */
  char message[256];
} gprx_error;

/* Kernel program: a KernelSpec tree in postfix, `n_ops` pairs (op, arg).
 * Leaves push one value; SUM and PROD pop `arg` children and push their sum
 * or product (the flattened associative chains of src/kernel/spec.rs:384-416).
 *   GPRX_OP_CONST     arg unused   theta = [log c]           (src/kernel/constant.rs:12-16)
 *   GPRX_OP_RBF       arg unused   theta = [log l]           (src/kernel/rbf.rs:18-20)
 *   GPRX_OP_RBF_ARD   arg = d      theta = [log l_0..l_d-1]  (src/kernel/rbf_ard.rs:16)
 *   GPRX_OP_PERIODIC  arg unused   theta = [log l, log p]    (src/kernel/periodic.rs:17-23)
 *   GPRX_OP_SUM       arg = k children
 *   GPRX_OP_PROD      arg = k children                                             */
enum {
  GPRX_OP_CONST = 1,
  GPRX_OP_RBF = 2,
  GPRX_OP_RBF_ARD = 3,
  GPRX_OP_PERIODIC = 4,
  GPRX_OP_SUM = 5,
  GPRX_OP_PROD = 6
};

typedef struct gprx_kernel_program {
  const int32_t* ops; /* 2 * n_ops int32: op, arg */
  size_t n_ops;
} gprx_kernel_program;

/* Arithmetic form.  FP64: everything in f64 (the oracle form).  MIXED: all
 * dense trailing and inverse products in f32 on MFMA with f32 accumulation,
 * and the sparse models' kernel fills in f32; f64 for the exact model's
 * kernel fill (rounded to f32 for the factor, kept in f64 for the
 * residuals), the 128 x 128 diagonal-block factorizations and their
 * inverses, the log-determinant, the small m x m factors, and the iterative
 * refinement of alpha with f64 residuals; falls back to FP64 when
 * refinement stalls or when the f32 factor needs more jitter than a set
 * bound (gprx_hip_set_mixed_jitter_bound). */
typedef enum gprx_precision { GPRX_FP64 = 0, GPRX_MIXED = 1 } gprx_precision;

/* Why the last mixed evaluation ran in FP64 (gprx_hip_last_fallback). */
typedef enum gprx_fallback {
  GPRX_FALLBACK_NONE = 0,
  GPRX_FALLBACK_REFINEMENT = 1, /* refinement stalled (src/precision/refine.rs:66-98) */
  GPRX_FALLBACK_JITTER = 2      /* the f32 factor needed jitter above the bound */
} gprx_fallback;

/* JitterPolicy (src/policy.rs:180-396).  The factor is tried without an
 * offset first; on failure with offsets initial, initial*multiplier, ...,
 * at most max_retries of them and none above max_jitter.  fixed(j) is
 * {j, 1, j, 1}; the exact default fixed(0) is {0, 1, 0, 0}. */
typedef struct gprx_jitter_policy {
  double initial;
  double multiplier;
  double max_jitter;
  int32_t max_retries;
} gprx_jitter_policy;

/* VarianceKind (src/prediction.rs). */
typedef enum gprx_variance_kind {
  GPRX_VARIANCE_OBSERVATION = 0, /* latent + sigma^2 (default) */
  GPRX_VARIANCE_LATENT = 1
} gprx_variance_kind;

typedef struct gprx_hip_model gprx_hip_model;

/* Library version and device name; never fails. */
const char* gprx_hip_version(void);
/* Number of visible GPUs (0 without one). */
int gprx_hip_device_count(void);

/* ------------------------------------------------------------------ exact */

/* Replaces Gpr::factor / FittedGpr::prepare (src/gpr/trainer.rs:533-630,
 * src/gpr/fitted.rs:210-270): validates the data, compiles the kernel
 * program, uploads x and y, and sizes the device buffers.  log_theta
 * (p = kernel params + 1) is the starting point; no factor is formed yet. */
int gprx_hip_exact_create(const double* x, size_t n, size_t d, const double* y,
                          const gprx_kernel_program* kernel, const double* log_theta, size_t p,
                          int32_t precision, const gprx_jitter_policy* jitter,
                          gprx_hip_model** out, gprx_error* err);

/* Replaces FittedGpr::value_and_gradient_into (src/gpr/fitted.rs:620) and
 * GprObjective::value_and_gradient_into (src/gpr/objective.rs:125-132):
 * NLML and its gradient at log_theta.  grad may be NULL (value only, as
 * Objective::value, src/objective.rs:193).  On success the model's
 * parameters become log_theta. */
int gprx_hip_exact_value_and_grad(gprx_hip_model* model, const double* log_theta, size_t p,
                                  double* nlml, double* grad, gprx_error* err);

/* Replaces TwiceDifferentiable::value_gradient_hessian_into for the exact
 * model (src/objective.rs; ExactFit::fill_hessian_from_factor,
 * src/gpr/exact_fit.rs:484-596), the Hessian TrustRegion uses: NLML,
 * gradient[p] and the p x p Hessian (row-major, symmetric) with respect to
 * log_theta, always formed in FP64 whatever the model's precision:
 *   H_ij = -1/2 <W, d2A_ij> - 1/2 tr(S_i S_j) + v_i . v_j,
 *   W = alpha alpha^T - A^-1, S_i = L^-1 A_i L^-T, v_i = L^-1 A_i alpha.
 * Needs (p + 2) n_pad^2 doubles of device memory beyond the model's.  On
 * success the model's parameters become log_theta. */
int gprx_hip_exact_hessian(gprx_hip_model* model, const double* log_theta, size_t p,
                           double* nlml, double* grad, double* hess, gprx_error* err);

/* Replaces OnlineGpr::insert (src/gpr/online.rs; the factor update of
 * src/linalg/ldlt.rs) for an FP64 model: appends one point (x_new[d]) and
 * its target, extends the factor and its inverse by one row in O(n^2) on
 * the device, and writes the new NLML (nullable).  When the 128-row padding
 * is full the model is re-padded and factored again.  GPRX_CHOLESKY_FAILED
 * leaves the model as it was. */
int gprx_hip_exact_insert(gprx_hip_model* model, const double* x_new, size_t d, double y_new,
                          double* nlml, gprx_error* err);

/* The factor behind FittedGpr::save_with_factor (src/gpr/fitted.rs:810-830,
 * packed by src/persist/tensors.rs:287-298): forms the model's Cholesky
 * factor L of K + (sigma^2 + jitter) I in FP64 at its current parameters
 * (as L^T = L^-1 K, the device model keeping only L^-1) and writes it
 * column-major into l[n*n] (zeros above the diagonal), with alpha[n] =
 * K^-1 y and the jitter the factor took.  Needs n_pad^2 doubles of device
 * memory beyond the model's.  The model keeps its parameters, and a mixed
 * model is formed again in its own precision afterwards. */
int gprx_hip_exact_export_factor(gprx_hip_model* model, double* l, size_t n, double* alpha,
                                 double* jitter, gprx_error* err);

/* Replaces OnlineGpr::delete (src/gpr/online.rs:564-620, the LDL^T row
 * delete of src/linalg/ldlt.rs:113): drops the point at row `index` (rows
 * keep their order) and writes the new NLML (nullable).  An FP64 model
 * downdates its factor's inverse in O(n^2) on the device (a rank-one update
 * of the trailing factor), keeping the jitter the factor was formed with; a
 * mixed model is factored again at its parameters, O(n^3).  A model of one
 * point refuses (GPRX_INSUFFICIENT_DATA). */
int gprx_hip_exact_delete(gprx_hip_model* model, size_t index, double* nlml, gprx_error* err);

/* Replaces FittedGpr::predict_with (src/gpr/fitted.rs:854-960,
 * src/gpr/shared.rs:754-816) at the model's current parameters.  var may be
 * NULL (mean only). */
int gprx_hip_exact_predict(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                           int32_t variance_kind, double* mean, double* var, gprx_error* err);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, code that must not call hipMalloc).  The same
 * predictions as gprx_hip_exact_predict, computed without any allocation:
 * the work space is a file-scope device array of 32 MiB that each GPU the
 * library runs on holds until the process ends, and queries run in chunks
 * of 128, so any q is taken.  Limits: the model must already be factored
 * (GPRX_INVALID_CONFIG otherwise: evaluate it first) and hold at most 16384
 * points after padding to 128 (GPRX_UNSUPPORTED otherwise).  Every call on
 * a GPU shares that GPU's array and leaves the last call's contents in it,
 * so calls that overlapped on one GPU would race: the library runs them one
 * at a time per device (each returns after its stream work has finished). */
int gprx_hip_exact_predict_noalloc(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                                   int32_t variance_kind, double* mean, double* var,
                                   gprx_error* err);

/* Current parameters (FittedGpr::get_params) and the jitter the last factor
 * took. */
int gprx_hip_get_params(const gprx_hip_model* model, double* log_theta, size_t p);
double gprx_hip_last_jitter(const gprx_hip_model* model);
/* For the mixed form: the gprx_fallback reason if the last evaluation fell
 * back to FP64 (GPRX_FALLBACK_NONE if not), and the number of refinement
 * corrections it took. */
int gprx_hip_last_fallback(const gprx_hip_model* model);
/* Largest jitter the mixed form's f32 factor may add (default 0: it adds
 * none).  When the jitter policy's next offset after a failed f32 factor is
 * above the bound, or the policy has none left, the evaluation forms the
 * FP64 factor at the same jitter and continues the policy in FP64.  No
 * counterpart in the reference (its f32 factor takes the policy's jitter). */
int gprx_hip_set_mixed_jitter_bound(gprx_hip_model* model, double bound, gprx_error* err);
int gprx_hip_last_refine_steps(const gprx_hip_model* model);

/* ------------------------------------------------------------------ sgpr */

/* Replaces Sgpr::factor (src/sgpr/model.rs:448-513): x n x d, z m x d
 * inducing points (column-major), adaptive K_mm jitter
 * (src/policy.rs:222-226). */
int gprx_hip_sgpr_create(const double* x, size_t n, size_t d, const double* y, const double* z,
                         size_t m, const gprx_kernel_program* kernel, const double* log_theta,
                         size_t p, int32_t precision, gprx_hip_model** out, gprx_error* err);

/* Replaces FittedSgpr::value_and_gradient_into (src/sgpr/fitted.rs:306-326):
 * the VFE bound (src/sgpr/factor/vfe.rs:338-369) and its gradient
 * (src/sgpr/factor/derivatives.rs:283-366). */
int gprx_hip_sgpr_value_and_grad(gprx_hip_model* model, const double* log_theta, size_t p,
                                 double* value, double* grad, gprx_error* err);

/* Replaces FittedSgpr::predict_with (src/sgpr/factor/predict.rs:102-260). */
int gprx_hip_sgpr_predict(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                          int32_t variance_kind, double* mean, double* var, gprx_error* err);

/* ------------------------------------------------------------------ svgp */

/* Replaces Svgp::factor's set-up (src/svgp/model.rs): data and inducing
 * points; the variational parameters start at the prior (mean 0, L = I,
 * src/svgp/factor/assemble.rs:144-156). */
int gprx_hip_svgp_create(const double* x, size_t n, size_t d, const double* y, const double* z,
                         size_t m, const gprx_kernel_program* kernel, const double* log_theta,
                         size_t n_theta, int32_t precision, gprx_hip_model** out, gprx_error* err);

/* Total parameter count p = n_theta + m + m(m+1)/2 (src/svgp/factor/assemble.rs:158-164). */
size_t gprx_hip_svgp_num_params(const gprx_hip_model* model);

/* Replaces svgp_value_and_gradient_with (src/svgp/factor/gradient.rs:104-249):
 * -ELBO over the batch (row indices into x) and its gradient at params
 * ([log theta | mu | packed L_q]). */
int gprx_hip_svgp_value_and_grad(gprx_hip_model* model, const double* params, size_t p,
                                 const uint64_t* batch, size_t b, double* value, double* grad,
                                 gprx_error* err);

/* Replaces run_adam_fit (src/svgp/factor/adam.rs:92-141) for `steps` steps
 * (not epochs) of Adam (src/optimizer/adam.rs:179-198) with batches drawn by
 * the reference's Fisher-Yates epoch shuffle (src/rng.rs, seed).  params
 * ([log theta | mu | packed L_q]) is read as the start and overwritten with
 * the result.  values (may be NULL) receives the -ELBO of every step. */
int gprx_hip_svgp_adam(gprx_hip_model* model, double* params, size_t p, size_t batch_size,
                       size_t steps, double lr, double beta1, double beta2, double eps,
                       uint64_t seed, double* values, gprx_error* err);

/* Replaces the SVGP predict (src/svgp/factor/predict.rs) at params. */
int gprx_hip_svgp_predict(gprx_hip_model* model, const double* params, size_t p, const double* xs,
                          size_t q, size_t d, int32_t variance_kind, double* mean, double* var,
                          gprx_error* err);

/* ------------------------------------------------------------------ batch */

/* One problem of a batch: independent exact GPR with its own data and
 * program (the candidates of Lbfgs::with_restarts, src/optimizer/mod.rs:182-227,
 * or separate datasets). */
typedef struct gprx_problem {
  const double* x;
  size_t n, d;
  const double* y;
  const gprx_kernel_program* kernel;
  const double* log_theta;
  size_t p;
} gprx_problem;

/* Creates a handle over k problems that one call evaluates together. */
int gprx_hip_batch_create(const gprx_problem* problems, size_t k, int32_t precision,
                          gprx_hip_model** out, gprx_error* err);

/* NLML and gradient of every problem at its log_theta[i] (length p_i):
 * nlml[k], grad[i] of length p_i; errs[k] receives each problem's status
 * (a failed problem does not fail the others).  Returns GPRX_OK unless the
 * call itself failed. */
int gprx_hip_batch_value_and_grad(gprx_hip_model* model, const double* const* log_theta,
                                  double* nlml, double* const* grad, gprx_error* errs);

/* Launch schedule of the batch's grouped products.  By default the first
 * evaluation times every candidate on the batch itself and keeps the
 * fastest (shared by batches of the same shape in the process).  A policy
 * orders each launch's tiles over the 8 XCDs:
 *   0 build order, one contiguous run of tiles per XCD
 *   1 longest reduction first, each tile to the least-loaded XCD
 *   2 problems as units, heaviest first to the least-loaded XCD
 * and launches with fewer than small_tile_below tiles run 64 x 64 sub-tiles.
 * set_schedule fixes one (policy -1 returns to tuning); schedule reports the
 * one in use (policy -1 before the first evaluation) and the tuned time of
 * the factor + inverse product in ms (0 when fixed by hand).  No counterpart
 * in the reference (it evaluates problems one by one). */
int gprx_hip_batch_set_schedule(gprx_hip_model* model, int32_t policy, int32_t small_tile_below,
                                gprx_error* err);
int gprx_hip_batch_schedule(const gprx_hip_model* model, int32_t* policy,
                            int32_t* small_tile_below, double* tuned_ms);

/* A mixed batch evaluation replays two captured hipGraphs: the first factors
 * and refines alpha twice with f64 residuals on every problem, recording each
 * residual; after one readback the stop rules of
 * src/precision/refine.rs:56-98 decide per problem, the problems still
 * refining continue (at most 10 residual passes), and the second graph forms
 * A^-1 and the gradient.  A problem whose f32 factor fails
 * (GPRX_FALLBACK_JITTER) or whose refinement stalls or does not converge
 * (GPRX_FALLBACK_REFINEMENT) is evaluated again alone in FP64 (as
 * gprx_hip_exact_value_and_grad with the default jitter policy), so its
 * results and status are FP64's.
 * reasons[k] receives each problem's gprx_fallback for the last call.  No
 * counterpart in the reference (it evaluates problems one by one). */
int gprx_hip_batch_last_fallback(const gprx_hip_model* model, int32_t* reasons, size_t k);

/* ------------------------------------------------------------------ fit */

/* Lbfgs settings (src/optimizer/lbfgs.rs:19-75): max_iterations 100,
 * gradient tolerance sqrt(eps) (on the gradient in logit coordinates),
 * history 10; Lbfgs::with_restarts(n, seed) adds n starts drawn
 * log-uniformly inside the bounds from SeededRng(seed)
 * (src/optimizer/mod.rs:182-227, src/optimizer/logit.rs:243-265). */
typedef struct gprx_lbfgs_options {
  uint64_t max_iterations;
  double tolerance;
  int32_t history_size;
  uint32_t restarts;
  uint64_t seed;
} gprx_lbfgs_options;

/* Fills the defaults of Lbfgs::new(), no restarts. */
void gprx_lbfgs_default_options(gprx_lbfgs_options* opt);

/* An objective on the host: value and gradient at x[n]; a nonzero return
 * (or a non-finite value) counts as a failed point. */
typedef int (*gprx_objective_fn)(void* ctx, const double* x, size_t n, double* value,
                                 double* grad);

/* Replaces the unconstrained L-BFGS run (src/optimizer/lbfgs.rs:118-142,
 * argmin 0.11's LBFGS + MoreThuenteLineSearch restated): minimizes fn from
 * x0, writing the best point, its value and the iterations taken.
 * GPRX_NOT_CONVERGED when no point had a finite cost.  Host only. */
int gprx_lbfgs_minimize(gprx_objective_fn fn, void* ctx, const double* x0, size_t n,
                        const gprx_lbfgs_options* opt, double* x_out, double* value,
                        uint64_t* iterations, gprx_error* err);

/* Replaces Gpr::with_optimizer(Lbfgs::..).fit (src/gpr/trainer.rs, the
 * Optimizer impl of src/optimizer/lbfgs.rs:101-116): minimizes the exact
 * NLML over log theta (p = kernel params + 1) through the logit map of
 * `bounds` (2p user-unit values lo, hi per parameter, or NULL for the
 * default (1e-5, 1e5) of every parameter; src/optimizer/logit.rs:75-230)
 * from log_theta and opt->restarts further starts.  All starts run together:
 * every round evaluates each live start's next point in one batched GPU
 * call (gprx_hip_batch_*, factored without jitter); a point whose factor
 * fails there is evaluated again alone under `jitter` (NULL or fixed(0): no
 * retry), as every evaluation of the reference's fit takes the model's
 * policy (src/gpr/exact_fit.rs:280-283); a point that still fails costs 1e300
 * (src/optimizer/adapter.rs:23,116-150).  Writes the best start's best
 * point, its NLML, its iterations and its index (0 = log_theta).  The first
 * start failing fails the fit with the status of its evaluation (e.g.
 * GPRX_CHOLESKY_FAILED); a failed restart is skipped
 * (src/optimizer/mod.rs:210-226). */
int gprx_hip_exact_fit(const double* x, size_t n, size_t d, const double* y,
                       const gprx_kernel_program* kernel, const double* log_theta, size_t p,
                       const double* bounds, int32_t precision, const gprx_jitter_policy* jitter,
                       const gprx_lbfgs_options* opt, double* best_log_theta, double* best_value,
                       uint64_t* iterations, int32_t* best_start, gprx_error* err);

/* TrustRegion settings (src/optimizer/trust_region.rs:19-80): 100
 * iterations, gradient tolerance sqrt(eps), radius 1, largest radius 100;
 * with_restarts(n, seed) as for Lbfgs. */
typedef struct gprx_tr_options {
  uint64_t max_iterations;
  double tolerance;
  double initial_radius;
  double max_radius;
  uint32_t restarts;
  uint64_t seed;
} gprx_tr_options;

/* Fills the defaults of TrustRegion::new(), no restarts. */
void gprx_tr_default_options(gprx_tr_options* opt);

/* An objective with its Hessian on the host: value, gradient[n] and the
 * n x n Hessian (row-major) at x; nonzero (or non-finite) = failed point. */
typedef int (*gprx_objective2_fn)(void* ctx, const double* x, size_t n, double* value,
                                  double* grad, double* hess);

/* Replaces the trust-region run (src/optimizer/trust_region.rs:107-148,
 * argmin 0.11's TrustRegion + Steihaug restated) over a host callback.
 * Host only. */
int gprx_trust_region_minimize(gprx_objective2_fn fn, void* ctx, const double* x0, size_t n,
                               const gprx_tr_options* opt, double* x_out, double* value,
                               uint64_t* iterations, gprx_error* err);

/* Replaces Gpr::with_optimizer(TrustRegion::..).fit for an exact model:
 * trust-region steps on the NLML through the logit map of `bounds` (as
 * gprx_hip_exact_fit), each point's Hessian formed on the GPU
 * (gprx_hip_exact_hessian), from log_theta and then each restart draw in
 * turn.  The model is left at the best point.  A start that cannot be
 * evaluated fails the fit with its evaluation's status
 * (src/optimizer/trust_region.rs:166-168). */
int gprx_hip_exact_fit_trust_region(gprx_hip_model* model, const double* log_theta, size_t p,
                                    const double* bounds, const gprx_tr_options* opt,
                                    double* best_log_theta, double* best_value,
                                    uint64_t* iterations, int32_t* best_start, gprx_error* err);

/* ------------------------------------------------------------------ all */

/* Frees a handle of any kind (NULL is ignored). */
void gprx_hip_destroy(gprx_hip_model* model);

/* Wall time in milliseconds of the last *_value_and_grad call measured with
 * HIP events on the handle's stream, and the time of its dominant dense
 * phase (factor + inverse), for roofline reporting. */
double gprx_hip_last_eval_ms(const gprx_hip_model* model);
double gprx_hip_last_dense_ms(const gprx_hip_model* model);

/* Per-phase device time (ms) of the last exact evaluation, from HIP events
 * on the handle's stream: 0 kernel fill, 1 factor + triangular inverse,
 * 2 alpha (and refinement), 3 inverse product X^T X, 4 gradient contraction.
 * No counterpart in the reference (its timing is wall clock around the whole
 * call, compare/perf/gprx/src/timing.rs:5-22); for roofline reporting. */
double gprx_hip_last_phase_ms(const gprx_hip_model* model, int32_t phase);

#ifdef __cplusplus
}
#endif

#endif /* GPRX_HIP_H */
