/*
 * libtopopt_hip — MI355X (gfx950) SIMP compliance inner loop behind a C ABI.
 *
 * Drop-in for TopOpt.jl v0.14.0's matrix-free displacement solve, compliance
 * sensitivity and density filter.  Every entry point names the reference
 * function it replaces (file:line in the TopOpt.jl source tree).
 *
 * Conventions
 *  - Plain C types only; no exceptions or aborts cross the boundary.
 *  - Return codes: 0 success; > 0 soft failure (TOHIP_NOT_CONVERGED: valid
 *    outputs; TOHIP_BREAKDOWN: NaN displacement, as the reference returns);
 *    < 0 hard error, message via tohip_last_error().
 *  - Host arrays are owned by the caller and only read/written during the
 *    call (Julia: GC.@preserve around the ccall).  Device memory is owned by
 *    the context.  Calls are synchronous; one context per host thread.
 *  - Vectors of DOFs are in the caller's (Ferrite) DOF order when the
 *    context was created with node_dofs, else node-major
 *    (dof = dim*node + d, node = i + (nx+1)*(j + (ny+1)*k), 0-based).
 *    Cell vectors are in cell order i + nx*(j + ny*k) (Ferrite's order for
 *    generate_grid).
 */
#ifndef TOPOPT_HIP_H
#define TOPOPT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tohip_ctx tohip_ctx;

#define TOHIP_OK 0
#define TOHIP_NOT_CONVERGED 1
/* The (P)CG iteration broke down: a NaN or infinite residual, p'Kp <= 0 or
 * r'z <= 0, i.e. K(vars) or its preconditioner is not positive definite (a
 * hard-void design with xmin = 0, test/FEA/test_simulate.jl:17-35).  Detected
 * on the device within the iteration it happens; u is set to NaN, as the
 * reference's direct solver does for a matrix that is not positive definite
 * (src/FEA/solvers_api.jl:152-158), and tohip_last_error says which test
 * failed.  SIMP runs stop at that iteration. */
#define TOHIP_BREAKDOWN 2
#define TOHIP_ERR_ARG (-1)
#define TOHIP_ERR_HIP (-2)
#define TOHIP_ERR_STATE (-3)
#define TOHIP_ERR_DOMAIN (-4) /* EnergyCriteria DomainError: NaN or negative x'Ax */

#define TOHIP_FP64 0  /* FP64 Jacobi PCG throughout */
#define TOHIP_MIXED 1 /* FP32 inner PCG inside FP64 iterative refinement */
#define TOHIP_MIXED_MG 2 /* as TOHIP_MIXED, the FP32 PCG preconditioned by a
                            geometric multigrid V-cycle instead of Jacobi */
#define TOHIP_FP64_MG 3  /* FP64 PCG (operator, vectors, recurrences) preconditioned
                            by the same FP32 multigrid V-cycle: the comparator that
                            separates the preconditioner's gain from FP32's */

/* Library version string. */
const char* tohip_version(void);

/* Number of visible GPUs and the name of device 0 (name may be NULL).
 * Returns TOHIP_ERR_HIP when the HIP runtime finds no device. */
int tohip_device_info(int* count, char* name, size_t name_len);

/* Message of the last hard error on this context (ctx may be NULL: the last
 * error of tohip_create on this thread). */
const char* tohip_last_error(const tohip_ctx* ctx);

/*
 * Create a solver context for a structured Q4 (dim 2) / hex8 (dim 3) grid of
 * nels[0] x nels[1] (x nels[2]) cells of size h[], material E, nu (plane
 * strain in 2D), SIMP interpolation rho = xmin + (1-xmin) x^p
 * (penalty_before_interp = 1) or (xmin + (1-xmin) x)^p (0).
 *   node_dofs : dim x nnodes, column-major, 1-based DOF numbers
 *               (Metadata.node_dofs) or NULL for node-major order.
 *   fixed_mask: ndofs bytes, non-zero = prescribed DOF (homogeneous
 *               Dirichlet), in the caller's DOF order.
 * Replaces: FEASolver(CGMatrixFreeSolver, problem; xmin, penalty, ...)
 *   src/FEA/solvers_api.jl:468-574 (ElementFEAInfo Kes, meandiag,
 *   fixed_dofs) with Ke from src/TopOptProblems/matrices_and_vectors.jl:68-177.
 */
int tohip_create(tohip_ctx** out, int dim, const int64_t* nels, const double* h, double E,
                 double nu, double xmin, double penalty_p, int penalty_before_interp,
                 const int64_t* node_dofs, const uint8_t* fixed_mask);

/* Release every device resource of the context. */
void tohip_destroy(tohip_ctx* ctx);

/* Sizes: number of DOFs and cells. */
int tohip_sizes(const tohip_ctx* ctx, int64_t* ndofs, int64_t* ncells);

/* Device memory: bytes held by the context (every buffer it allocated, plus
 * the attached network's, see tohip_nn_attach) and the device's used / total
 * bytes at the time of the call (any pointer may be NULL).  No reference
 * counterpart (memory report of large runs). */
int tohip_memory(const tohip_ctx* ctx, int64_t* ctx_bytes, int64_t* device_used,
                 int64_t* device_total);

/* setpenalty!(solver, p)  src/FEA/solvers_api.jl:419-429 (keeps the penalty
 * family and projection, changes p). */
int tohip_set_penalty(tohip_ctx* ctx, double penalty_p);

/* Penalty families, src/Utilities/penalties.jl:30-56, and projections applied
 * to x before the penalty (ProjectedPenaltyFun), penalties.jl:63-93. */
#define TOHIP_PEN_POWER 0    /* PowerPenaltyFun(p):    x^p                   */
#define TOHIP_PEN_RATIONAL 1 /* RationalPenaltyFun(p): x / (1 + p (1 - x))   */
#define TOHIP_PEN_SINH 2     /* SinhPenaltyFun(p):     sinh(p x) / sinh(p)   */
#define TOHIP_PROJ_NONE 0
#define TOHIP_PROJ_HEAVISIDE 1 /* HeavisideProjectionFun(b): 1 - e^(-b x) + x e^(-b) */
#define TOHIP_PROJ_SIGMOID 2   /* SigmoidProjectionFun(b): 1 / (1 + e^((b+1)(0.5 - x))) */

/* Select the penalty function of the interpolation (solver.penalty = P with
 * P one of the families above, optionally wrapped in ProjectedPenaltyFun with
 * projection proj of steepness beta): rho = xmin + (1-xmin) P(x) or
 * P(xmin + (1-xmin) x) as chosen at creation, get_rho / get_rho_drho,
 * src/Utilities/penalties.jl:113-130.  p > 0, beta >= 0 (ignored for
 * TOHIP_PROJ_NONE).  Rebuilds the cell moduli and the Jacobi diagonal. */
int tohip_set_penalty_fun(tohip_ctx* ctx, int kind, double penalty_p, int proj, double beta);

/* CG stopping rule, FEASolver's `conv` keyword (src/FEA/solvers_api.jl:482):
 * TOHIP_CONV_DEFAULT is the residual test max(reltol ||r0||, abstol);
 * TOHIP_CONV_ENERGY is EnergyCriteria (src/FEA/convergence_criteria.jl:20-45):
 * before every iteration, with x'Ax = f.x - x.r on the recursive residual,
 * stop when |x'Ax - energy| / x'Ax <= tol and |x.r| / x'Ax <= tol, then
 * energy := x'Ax (the field persists across solves; `energy` sets its value,
 * EnergyCriteria(energy)).  tol is the solve's reltol (tohip_solve) or
 * rel_tol (tohip_simp_run, tohip_beso_run).  A NaN or negative x'Ax ends the
 * solve with TOHIP_ERR_DOMAIN (the reference's DomainError).  EnergyCriteria
 * runs on the FP64 solver only; the mixed paths return TOHIP_ERR_ARG. */
#define TOHIP_CONV_DEFAULT 0
#define TOHIP_CONV_ENERGY 1
int tohip_set_convergence(tohip_ctx* ctx, int kind, double energy);
/* Current EnergyCriteria.energy (x'Ax of the last check). */
int tohip_get_convergence_energy(const tohip_ctx* ctx, double* energy);

/* solver.vars .= x  (src/Functions/compliance.jl:65): the filtered pseudo
 * densities used by the operator; also rebuilds the Jacobi diagonal. */
int tohip_set_densities(tohip_ctx* ctx, const double* vars);

/* y = K(vars) x.  Replaces mul!(y, ::MatrixFreeOperator, x)
 * src/FEA/matrix_free_operator.jl:66-105 (fixed DOFs: y = meandiag * x). */
int tohip_matvec(tohip_ctx* ctx, const double* x, double* y);

/* Solve K(vars) u = f with Jacobi PCG.  Replaces solve_system!(
 * CGMatrixFreeSolver, ...) -> cg_solve! src/FEA/solvers_api.jl:177-220, 253-282.
 * Stops when ||r||_2 <= max(reltol * ||r0||_2, abstol) or after maxiter
 * iterations (then returns TOHIP_NOT_CONVERGED with valid outputs).
 * initially_zero = 1: u starts at 0 (reference default); 0: u is the warm start.
 * precision TOHIP_MIXED: FP32 operator, vectors and Jacobi PCG recurrences
 *   with FP64 dots and alpha/beta; whenever the FP32 residual has dropped 10x
 *   since the last FP64 refresh, the residual is recomputed in FP64
 *   (r = f - K u, u accumulated in FP64: iterative refinement by reliable
 *   updates, the search direction kept).  Same outer stopping rule on the
 *   FP64 residual; iters_out counts FP32 iterations.
 * precision TOHIP_MIXED_MG: TOHIP_MIXED with the FP32 PCG preconditioned by
 *   one V(1,1) cycle of a geometric multigrid (rediscretised coarse levels of
 *   halved cell counts, damped Jacobi smoothing, trilinear transfers) instead
 *   of Jacobi.  No reference counterpart (TopOpt.jl's matrix-free solver is
 *   Jacobi PCG); the outer FP64 stopping rule is the same.  Needs even cell
 *   counts along every axis (TOHIP_ERR_ARG otherwise).
 * Device residency (what ComplianceFun's solve-then-energy sequence needs,
 * src/Functions/compliance.jl:58-67): f = NULL solves for the load of
 * tohip_set_load, already on the device; u = NULL keeps the displacement on
 * the device (no download; tohip_get_u fetches it, tohip_compliance with
 * u = NULL uses it) and, with initially_zero = 0, starts from the
 * displacement the previous solve left there.
 * iters_out / resid_out may be NULL. */
int tohip_solve(tohip_ctx* ctx, const double* f, double* u, double abstol, double reltol,
                int64_t maxiter, int initially_zero, int precision, int64_t* iters_out,
                double* resid_out);

/* z = M r for the multigrid preconditioner of TOHIP_MIXED_MG at the current
 * densities (FP32 V-cycle; r and z in the caller's DOF order, FP64 at the
 * boundary).  No reference counterpart: it exposes the preconditioner for
 * inspection and testing. */
int tohip_mg_apply(tohip_ctx* ctx, const double* r, double* z);

/* Damped-Jacobi weights of the multigrid levels (omega[l]) and the current
 * estimate of lambda_max(D^-1 K) per level (lambda_est[l]; 0 where none is
 * kept: the coarsest level, or TOHIP_MG_WADAPT=0), for the first max_levels
 * levels; *nlevels = the level count (0 before the first multigrid solve).
 * Either array may be NULL.  No reference counterpart: it exposes the
 * preconditioner's state for inspection and testing. */
int tohip_mg_weights(tohip_ctx* ctx, double* omega, double* lambda_est, int max_levels,
                     int* nlevels);

/* Compliance obj = sum_e rho_e u_e^T Ke u_e, cell_comp_e = u_e^T Ke u_e,
 * grad_e = -drho_e cell_comp_e.  Replaces compute_element_energy
 * src/Functions/compute_element_energy.jl:18-38 (ComplianceFun
 * src/Functions/compliance.jl:58-76).  u = NULL: the displacement of the
 * last solve, already on the device (no upload).  cell_comp / grad may be
 * NULL. */
int tohip_compliance(tohip_ctx* ctx, const double* u, double* obj, double* cell_comp,
                     double* grad);

/* Build the density filter of radius rmin.  Replaces DensityFilterFun(solver;
 * rmin) src/CheqFilters/density_filter.jl:25-33, 49-108 (neighbour search
 * src/CheqFilters/CheqFilters.jl:66-118).  TOHIP_ERR_ARG when no node lies
 * within rmin of a cell centre (the reference's ArgumentError). */
int tohip_filter_setup(tohip_ctx* ctx, double rmin);

/* y = J x (transpose = 0; DensityFilterFun call, density_filter.jl:35-40) or
 * y = J^T x (transpose = 1; its rrule pullback, density_filter.jl:41-47). */
int tohip_filter_apply(tohip_ctx* ctx, const double* x, double* y, int transpose);

/* ProjectedDensityFilterFun (src/CheqFilters/density_filter.jl:130-169):
 * y = post(J pre(x)) with pre / post each TOHIP_PROJ_NONE, _HEAVISIDE or
 * _SIGMOID of steepness beta (HeavisideProjectionFun / SigmoidProjectionFun,
 * src/Utilities/penalties.jl:76-95), elementwise.  Needs tohip_filter_setup. */
int tohip_filter_apply_projected(tohip_ctx* ctx, const double* x, double* y, int pre_proj,
                                 double pre_beta, int post_proj, double post_beta);

/* Its pullback at x (what Zygote differentiates through the reference's
 * broadcasts and the DensityFilterFun rrule, density_filter.jl:35-47):
 * dx = pre'(x) .* J^T (post'(J pre(x)) .* dy). */
int tohip_filter_pullback_projected(tohip_ctx* ctx, const double* x, const double* dy,
                                    double* dx, int pre_proj, double pre_beta, int post_proj,
                                    double post_beta);

/* Volume fraction sum_e x_e V_e / V_total.  Replaces VolumeFun
 * src/Functions/volume.jl:59-67 (gradient is the constant V_e / V_total). */
int tohip_volume(tohip_ctx* ctx, const double* x, double* vol);

/* One OC update (build-defined; the reference's SIMP driver is MMA, see
 * docs/tutorials/simp.qmd:103-114).  dc is the sensitivity w.r.t. the design
 * x (after the filter pullback); the volume constraint is on J x.
 * lambda_out may be NULL. */
int tohip_oc_update(tohip_ctx* ctx, const double* x, const double* dc, double volfrac,
                    double move, double tol, double* xnew, double* lambda_out);

/* Load vector (replaces the fixed load of assemble!,
 * src/TopOptProblems/assemble.jl:49-54, 87): used by tohip_simp_run and by
 * tohip_solve with f = NULL.  Prescribed DOFs are zeroed (apply_zero!);
 * a non-finite entry is TOHIP_ERR_ARG. */
int tohip_set_load(tohip_ctx* ctx, const double* f);

/* Options of the device-resident SIMP loop. */
typedef struct tohip_simp_opts {
    int64_t iterations;     /* OC iterations to run */
    double volfrac;         /* volume fraction of J x */
    double rmin;            /* filter radius (filter is built if needed) */
    double move;            /* OC move limit (0.2) */
    double oc_tol;          /* OC bracket tolerance (1e-12) */
    double rel_tol;         /* solve: ||f - K u|| <= rel_tol * ||f|| */
    int64_t maxiter;        /* PCG iteration cap per solve (inner, mixed) */
    int precision;          /* TOHIP_FP64, TOHIP_MIXED or TOHIP_MIXED_MG */
    int warm_start;         /* 0: u0 = 0 (reference default); 1: previous u;
                               2: learned network guess (tohip_nn_*) combined
                               with the previous u, until two iterations in a
                               row leave more than 1/4 of the previous u's
                               starting residual, or a solve from a prediction
                               takes over twice the previous solve's PCG
                               iterations; then 1 for the rest of the run
                               (tohip_nn_calls counts the predictions) */
    double inner_rel_tol;   /* mixed: FP64 residual refresh each time the FP32
                               residual has dropped by this factor (0.1) */
} tohip_simp_opts;

/* Run `iterations` SIMP iterations with every vector resident on the GPU:
 * x~ = J x, solve, compliance + sensitivity, J^T pullback, OC.  x_inout is
 * the design on entry and exit and starts a new run (the first solve has no
 * previous displacement); NULL continues from the design and displacement
 * the previous call left on the device (no host transfer).  Per-iteration outputs (arrays of length
 * iterations, each may be NULL): obj, PCG iterations, OC change.
 * Replaces the optimizer-driven loop of docs/tutorials/simp.qmd:86-114. */
int tohip_simp_run(tohip_ctx* ctx, const tohip_simp_opts* opts, double* x_inout,
                   double* obj_hist, int64_t* iters_hist, double* change_hist);

/* MMA87 (Svanberg 1987; the optimizer of the reference's SIMP tutorials,
 * docs/tutorials/simp.qmd:103-114, provided there by the external
 * NonconvexMMA package) in place of the OC update: min c(x) s.t.
 * V(J x) <= volfrac, 0 <= x <= 1, moving asymptotes L, U (the first two
 * iterations x -+ s_init, then widened by s_incr / narrowed by s_decr as
 * the steps keep / flip sign), the convex separable approximation on
 * [0.9 L + 0.1 x, 0.9 U + 0.1 x], its one-constraint dual solved to
 * opts->oc_tol by bracketing the multiplier.  opts->move is not used.  The
 * asymptote history starts afresh with every new run (x_inout != NULL) and
 * carries on across resident continuation calls (x_inout == NULL), so a run
 * split into several calls follows the same iterates as one long call. */
typedef struct tohip_mma_opts {
    double s_init; /* 0.5 */
    double s_incr; /* 1.2 */
    double s_decr; /* 0.7 */
} tohip_mma_opts;
int tohip_simp_run_mma(tohip_ctx* ctx, const tohip_simp_opts* opts, const tohip_mma_opts* mma,
                       double* x_inout, double* obj_hist, int64_t* iters_hist,
                       double* change_hist);

/* Compliance and sensitivities (tohip_compliance) of device-resident data,
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  u_dev: displacement in the library's
 * node-major order (dof = dim * node + d, node = i + (nx+1)*(j + (ny+1)*k));
 * xt_dev: filtered densities (cells); writes ce_dev, grad_dev (cells) and
 * obj_dev (one double), all device memory, enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) with no allocation, no host copy and
 * no synchronisation.  The per-block partial sums live in a fixed device
 * array of 4096 doubles (32 KiB) per GPU, held until the process ends;
 * grids that would need more slots are rejected (TOHIP_ERR_ARG; the kernel
 * strides over the cells, so no grid does).  Calls that overlap on one GPU
 * share that array and race.  The context's GPU must be current. */
int tohip_compliance_noalloc(tohip_ctx* ctx, const double* u_dev, const double* xt_dev,
                             double* ce_dev, double* grad_dev, double* obj_dev, void* stream);

/* Last displacement of tohip_simp_run (caller order). */
int tohip_get_u(tohip_ctx* ctx, double* u);

/* Filtered densities x~ used by the last solve (cell order). */
int tohip_get_xphys(tohip_ctx* ctx, double* xphys);

/* ------------------------------------------------------------------------
 * Learned warm start (no reference counterpart: TopOpt.jl starts CG from 0
 * or, with initially_zero=false, from the previous solution,
 * src/FEA/solvers_api.jl:193-195).  A 3D convolutional encoder-decoder
 * (channels 16/32/64/128, 3x3x3 kernels, BF16 weights on MFMA with FP32
 * accumulation) maps the filtered densities, the load and the clamped cells
 * to a nodal displacement guess; inside tohip_simp_run (warm_start = 2) PCG
 * starts from the point of span{guess, previous displacement} with the
 * smallest residual ||f - K u0||.  3D grids with cell counts that are
 * multiples of 8.  DOF vectors of this API are node-major
 * (dof = 3*node + d).
 * ---------------------------------------------------------------------- */
typedef struct tohip_nn tohip_nn;

const char* tohip_nn_last_error(const tohip_nn* nn);
/* fixed_mask: ndofs bytes (node-major), the clamped DOFs of the problem
 * family the network sees.  max_samples: capacity of the device dataset. */
int tohip_nn_create(tohip_nn** out, const int64_t* nels, int max_batch, int max_samples,
                    uint64_t seed, const uint8_t* fixed_mask);
void tohip_nn_destroy(tohip_nn* nn);
int64_t tohip_nn_num_params(const tohip_nn* nn);
int tohip_nn_get_params(tohip_nn* nn, float* params);
int tohip_nn_set_params(tohip_nn* nn, const float* params);
/* Training pair: filtered densities (cells), load and displacement (DOFs). */
int tohip_nn_add_sample(tohip_nn* nn, const double* xphys, const double* f, const double* u);
int tohip_nn_clear_samples(tohip_nn* nn);
int tohip_nn_num_samples(const tohip_nn* nn);
/* Adam steps on random mini-batches of dataset samples [first, first +
 * count); loss per step (MSE of u / rms(u)) into loss_hist (may be NULL). */
int tohip_nn_train(tohip_nn* nn, int first, int count, int steps, int batch, double lr,
                   uint64_t seed, double* loss_hist);
/* Mean loss over dataset samples [first, first + count). */
int tohip_nn_eval(tohip_nn* nn, int first, int count, double* loss);
/* Network output for one problem: u_out has ndofs floats (node-major). */
int tohip_nn_predict(tohip_nn* nn, const double* xphys, const double* f, float* u_out);
/* Use nn for warm_start = 2 of tohip_simp_run (NULL detaches). */
int tohip_nn_attach(tohip_ctx* ctx, tohip_nn* nn);
/* Network predictions the current (or last) tohip_simp_run run used. */
int tohip_nn_calls(const tohip_ctx* ctx, int64_t* calls);
/* One convolution layer (stride 1, zero padding, NDHWC) forward and
 * backward on FP32 host data through the BF16 kernels, for gradient checks:
 * x [B][D][H][W][cin], w [cout][k^3][cin], b [cout], dy/y [B][OD][OH][OW]
 * [cout]; outputs y, dx, dw, db.  act: 0 identity, 1 SiLU. */
int tohip_nn_conv_check(int B, const int* dims, int cin, int cout, int k, int pad, int act,
                        const float* x, const float* w, const float* b, const float* dy,
                        float* y, float* dx, float* dw, float* db);

/* Kernel micro-benchmark (no reference counterpart; measurement support):
 * launches one kernel `reps` times back to back on the context's stream
 * between two HIP events and returns the mean duration in microseconds.
 * Operands are the context's current buffers (set them first with
 * tohip_set_densities / tohip_matvec / tohip_filter_setup). */
#define TOHIP_K_MATVEC64 0   /* y = K x, FP64                          */
#define TOHIP_K_MATVEC32 1   /* y = K x, FP32                          */
#define TOHIP_K_PCG64 2      /* fused p = z + beta p, q = K p, p.q FP64 */
#define TOHIP_K_PCG32 3      /* same, FP32 vectors, FP64 reductions     */
#define TOHIP_K_UPDATE64 4   /* x += a p, r -= a q, r.z, r.r  FP64      */
#define TOHIP_K_UPDATE32 5   /* same, FP32 vectors                      */
#define TOHIP_K_FILTER 6     /* x~ = J x                                */
#define TOHIP_K_FILTER_T 7   /* J^T y                                   */
#define TOHIP_K_ENERGY 8     /* compliance + sensitivity                */
#define TOHIP_K_OC 9         /* one full OC update (several launches)   */
#define TOHIP_K_COPY64 10    /* plain copy of one FP64 DOF vector (16 B per DOF moved) */
/* multigrid preconditioner of TOHIP_MIXED_MG (built on first use): */
#define TOHIP_K_MG_PRE 11     /* finest level: x = w D^-1 b, r = b - K x        */
#define TOHIP_K_MG_RESTRICT 12 /* finest -> first coarse level: P^T r          */
#define TOHIP_K_MG_COARSE 13  /* coarsest level: all Jacobi sweeps (1 workgroup) */
#define TOHIP_K_MG_POST 14    /* finest level: x += P xc, z = x + w D^-1 (b - K x) */
#define TOHIP_K_VCYCLE 15     /* one whole V-cycle (every launch of it)         */
#define TOHIP_K_CSR_SPMV 16   /* assembled K (CSR) times x, FP64 (assembles K first) */
int tohip_bench_kernel(tohip_ctx* ctx, int kernel, int reps, double* avg_us);

/* ---- BESO (src/Algorithms/beso.jl) ---------------------------------------
 * Options of BESO(comp, vol, vol_limit, filter; maxiter, tol, p, er, sens_tol,
 * k) (beso.jl:45-87), plus the solve of each compliance evaluation. */
typedef struct tohip_beso_opts {
    int64_t maxiter;      /* 200 */
    double vol_limit;     /* target volume fraction V */
    double rmin;          /* DensityFilterFun radius (the filter is built if needed) */
    double er;            /* evolution rate (0.02) */
    double tol;           /* convergence of the objective trace (1e-4) */
    double sens_tol;      /* threshold bisection tolerance (tol / 100), > 0 */
    int k;                /* length of the objective trace (10) */
    double penalty_p;     /* p: setpenalty!(solver, p) before the loop (3.0) */
    double rel_tol;       /* solve: ||f - K u|| <= rel_tol * ||f|| */
    int64_t pcg_maxiter;  /* PCG iteration cap per solve */
    int precision;        /* TOHIP_FP64, TOHIP_MIXED or TOHIP_MIXED_MG */
} tohip_beso_opts;

/* BESOResult (beso.jl:1-7), plus the final volume fraction. */
typedef struct tohip_beso_result {
    double objval;
    double change;
    int converged;
    int64_t fevals;
    double volume;
} tohip_beso_result;

/* (b::BESO)(x0; black, white) (beso.jl:93-188) on the device: each
 * iteration filters the topology, solves, evaluates the compliance and its
 * pullback (sens = -J^T dc, averaged with the previous iteration's), shrinks
 * the volume target by er down to vol_limit and picks the topology by the
 * reference's threshold bisection on sens.  passive: NULL, or per cell 0
 * free, 1 solid (black), 2 void (white).  topology_out (ncells, may be NULL)
 * receives the 0/1 topology; obj_hist (maxiter, may be NULL) the objective of
 * every iteration.  Uses the load set by tohip_set_load. */
int tohip_beso_run(tohip_ctx* ctx, const tohip_beso_opts* o, const double* x0,
                   const uint8_t* passive, double* topology_out, tohip_beso_result* res,
                   double* obj_hist);

/* ------------------------------------------------------------------------
 * Assembled-matrix comparator: CGAssemblySolver (src/FEA/solvers_api.jl:
 * 223-250).  K(vars) is assembled on the device in CSR (node-major DOF
 * order; every row keeps the full 3x3(x3)-node pattern) from the element
 * matrix as assemble! does (src/TopOptProblems/assemble.jl:28-90), then
 * apply!'s boundary rule: prescribed rows and columns zero, their diagonal
 * the mean |diag| of the unconstrained K.  `safe` also puts that value on a
 * zero free diagonal (solvers_api.jl:232-239).
 * ---------------------------------------------------------------------- */
/* Assemble (or re-assemble at the current densities). */
int tohip_csr_assemble(tohip_ctx* ctx, int safe);
/* Stored entries, device bytes of the CSR arrays, and the prescribed diagonal. */
int tohip_csr_info(const tohip_ctx* ctx, int64_t* nnz, int64_t* bytes, double* meandiag);
/* The CSR arrays as stored (rows and columns in node-major order 3*node + d):
 * rowptr (ndofs + 1), cols and vals (nnz). */
int tohip_csr_download(const tohip_ctx* ctx, int64_t* rowptr, int* cols, double* vals);
/* y = K x (caller DOF order), re-assembled at the current densities. */
int tohip_csr_matvec(tohip_ctx* ctx, const double* x, double* y);
/* solve_system!(CGAssemblySolver, ...): assemble, f = 0 on prescribed DOFs,
 * CG (jacobi = 0: the reference's default identity preconditioner; 1:
 * Jacobi) with IterativeSolvers' order and stopping rule.  0, or
 * TOHIP_NOT_CONVERGED. */
int tohip_csr_solve(tohip_ctx* ctx, const double* f, double* u, double abstol, double reltol,
                    int64_t maxiter, int initially_zero, int jacobi, int safe, int64_t* iters_out,
                    double* resid_out);

/* ------------------------------------------------------------------------
 * Heat conduction (HeatTransfer physics): steady conduction with one
 * temperature DOF per node, conductivity k(rho) = rho(x) k
 * (HeatConductionProblem, src/TopOptProblems/problem_types.jl:783-1030;
 * element conductivity matrix src/TopOptProblems/matrices_and_vectors.jl:
 * 413-496).  Linear cells (Quadrilateral / Hexahedron) on the structured
 * grid; vectors are in the caller's DOF order (node_dofs), cells x-fastest.
 * Solver operator: the element products skip prescribed columns and the
 * prescribed rows are meandiag * T, meandiag = sum of the element traces
 * (src/FEA/matrix_free_operator.jl:66-105).
 * ---------------------------------------------------------------------- */
typedef struct tohip_heat tohip_heat;

const char* tohip_heat_last_error(const tohip_heat* ctx);
/* node_dofs: Metadata.node_dofs of the scalar field (1 x nnodes, 1-based);
 * fixed_mask: ndofs bytes, the ConstraintHandler's prescribed DOFs;
 * fixed_values: ndofs (read at prescribed DOFs; NULL = all zero), the
 * Dirichlet temperatures.  Densities start at 1.  Replaces the host path of
 * FEASolver(Solver, ::HeatConductionProblem) (src/FEA/solvers_api.jl:468-574). */
int tohip_heat_create(tohip_heat** out, int dim, const int64_t* nels, const double* h, double k,
                      double xmin, double penalty_p, int penalty_before_interpolation,
                      const int64_t* node_dofs, const uint8_t* fixed_mask,
                      const double* fixed_values);
void tohip_heat_destroy(tohip_heat* ctx);
/* As tohip_set_penalty_fun (penalty families of src/Utilities/penalties.jl). */
int tohip_heat_set_penalty_fun(tohip_heat* ctx, int kind, double penalty_p, int proj,
                               double beta);
int tohip_heat_set_densities(tohip_heat* ctx, const double* vars);
/* y = K(vars) x with the solver's fixed-DOF rule. */
int tohip_heat_matvec(tohip_heat* ctx, const double* x, double* y);
/* Jacobi PCG (IterativeSolvers' order and rule ||r|| <= max(reltol ||r0||,
 * abstol)).  apply_dirichlet = 1: the forward solve (solver(); assemble! +
 * apply!): T = the prescribed values at fixed DOFs and K_ff T_f = rhs_f -
 * K_fp T_p on the free ones.  apply_dirichlet = 0: homogeneous conditions,
 * the adjoint and pullback solves (apply_zero!).  T is the initial guess when
 * initially_zero == 0.  0, or TOHIP_NOT_CONVERGED. */
int tohip_heat_solve(tohip_heat* ctx, const double* rhs, double* T, double abstol, double reltol,
                     int64_t maxiter, int initially_zero, int apply_dirichlet, int64_t* iters_out,
                     double* resid_out);
/* cell_e = lam_e' Ke T_e with the unconstrained Ke (rawmatrix), grad_e =
 * sign * drho/dx(x_e) * cell_e: the element loop of compute_thermal_compliance!
 * (src/Functions/thermal_compliance.jl, sign +1, lam the adjoint of -Q) and of
 * TemperatureFun's pullback (src/Functions/temperature.jl, sign -1, lam =
 * K^-1 Delta). */
int tohip_heat_cell_products(tohip_heat* ctx, const double* lam, const double* T, double sign,
                             double* cell_out, double* grad_out);
/* Mean duration (us, HIP events on the context's stream) of the solver
 * operator's product on the current temperature vector, over reps launches. */
int tohip_heat_bench_matvec(tohip_heat* ctx, int reps, double* avg_us);
/* The element conductivity matrix (row-major, 4x4 or 8x8, Ferrite corner order). */
int tohip_heat_element_matrix(const tohip_heat* ctx, double* ke_out);

#ifdef __cplusplus
}
#endif
#endif /* TOPOPT_HIP_H */
