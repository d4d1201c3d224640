/*
 * TEST INFRASTRUCTURE ONLY — the oracle (CPU restatement of TopOpt.jl
 * v0.14.0's SIMP compliance path).  Loaded only by tests/, the smoke check and
 * bench.py's cpu_baseline leg, as the checker and the CPU baseline; the
 * product library never links it.  See oracle.h for conventions.
 *
 * Parity status of each function (how it is pinned, besides the GPU-vs-oracle
 * tests):
 *   orc_numbering          pinned: reference force_dof known answers
 *                          (test/topopt_problems/problems.jl:20, :63) and an
 *                          independent numpy restatement (tests/golden).
 *   orc_element_stiffness  pinned: plane-stress Q4 closed form of
 *                          test/topopt_problems/element_stiffness_matrix.jl
 *                          with plane-strain constants; patch test; rigid
 *                          modes; quadrature-order independence; numpy exact
 *                          polynomial integration.
 *   orc_matvec             pinned: dense assembled K (numpy) fixtures,
 *                          symmetry, fixed-row rule.
 *   orc_pcg                pinned: direct solve agreement, u'Ku = f'u.
 *   orc_pcg_energy         pinned: iterates identical to orc_pcg (shared
 *                          recurrence), the stopping test checked against a
 *                          numpy evaluation of both ratios at every iterate
 *                          (stops exactly at the first iterate passing it),
 *                          NaN / negative-energy DomainError cases.
 *   orc_element_energy     pinned: finite differences, obj = f'u.
 *   orc_rho_drho_fun       pinned: the known answers and analytic derivatives
 *                          of test/Utilities/test_penalties.jl (every penalty
 *                          family and projection), central differences.
 *   orc_filter_build       pinned: row-stochastic, closed-form stencil,
 *                          ArgumentError case, numpy restatement.
 *   orc_oc_update          build-defined (the reference has no OC); pinned
 *                          by the closed form of an unclamped step (x_new
 *                          and lambda built backwards from the answer), the
 *                          analytic optimum of sum a/x under a volume bound
 *                          (x* ~ sqrt(a)), and volume / bound invariants.
 *   orc_direct_solve       pinned: dense numpy solve fixtures.
 *   orc_cell_volume        pinned: closed form prod(h) at every quadrature
 *                          order (det J is constant on a box).
 *   orc_volume             pinned: linearity, V(1) = 1 as a fraction,
 *                          V(e_k) = V_k / V_total.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Ferrite reference-cell corner offsets (Quadrilateral / Hexahedron local
 * node order, as generate_grid builds cells; src/TopOptProblems/grids.jl:80-92). */
static const int OFF2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
static const int OFF3[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

static int corner_off(int dim, int a, int d) {
    return dim == 2 ? OFF2[a][d] : OFF3[a][d];
}

int64_t orc_nnodes(int dim, const int64_t* nels) {
    int64_t n = 1;
    for (int d = 0; d < dim; ++d) n *= nels[d] + 1;
    return n;
}

int64_t orc_ncells(int dim, const int64_t* nels) {
    int64_t n = 1;
    for (int d = 0; d < dim; ++d) n *= nels[d];
    return n;
}

/* Node id (0-based, x-fastest) of local corner a of cell c. */
static int64_t cell_node(int dim, const int64_t* nels, int64_t c, int a) {
    int64_t NX = nels[0] + 1, NY = nels[1] + 1;
    int64_t i = c % nels[0];
    int64_t j = (c / nels[0]) % nels[1];
    int64_t k = dim == 3 ? c / (nels[0] * nels[1]) : 0;
    i += corner_off(dim, a, 0);
    j += corner_off(dim, a, 1);
    if (dim == 3) k += corner_off(dim, a, 2);
    return i + NX * (j + NY * k);
}

/* Restates Ferrite's DofHandler close! numbering as used by Metadata
 * (src/TopOptProblems/metadata.jl:40-50, :116-145): cells in order, each new
 * node gets `dim` consecutive DOFs on first touch. */
int64_t orc_numbering(int dim, const int64_t* nels, int64_t* node_dofs,
                      int64_t* cell_dofs, int64_t* cell_nodes) {
    int64_t nn = orc_nnodes(dim, nels), nel = orc_ncells(dim, nels);
    int nnpc = 1 << dim, ndpc = dim * nnpc;
    int64_t* first = (int64_t*)malloc(sizeof(int64_t) * nn);
    if (!first) return -1;
    for (int64_t n = 0; n < nn; ++n) first[n] = -1;
    int64_t next = 0;
    for (int64_t c = 0; c < nel; ++c) {
        for (int a = 0; a < nnpc; ++a) {
            int64_t n = cell_node(dim, nels, c, a);
            if (first[n] < 0) first[n] = next++;
            if (cell_nodes) cell_nodes[c * nnpc + a] = n + 1;
            if (cell_dofs)
                for (int d = 0; d < dim; ++d)
                    cell_dofs[c * ndpc + dim * a + d] = dim * first[n] + d + 1;
        }
    }
    if (node_dofs)
        for (int64_t n = 0; n < nn; ++n)
            for (int d = 0; d < dim; ++d)
                node_dofs[n * dim + d] = dim * first[n] + d + 1;
    free(first);
    return dim * nn;
}

static void gauss_rule(int order, double* pts, double* wts) {
    if (order == 2) {
        double g = 1.0 / sqrt(3.0);
        pts[0] = -g; pts[1] = g; wts[0] = wts[1] = 1.0;
    } else if (order == 3) {
        double g = sqrt(0.6);
        pts[0] = -g; pts[1] = 0.0; pts[2] = g;
        wts[0] = wts[2] = 5.0 / 9.0; wts[1] = 8.0 / 9.0;
    } else {
        double a = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
        double b = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
        double wa = (18.0 + sqrt(30.0)) / 36.0, wb = (18.0 - sqrt(30.0)) / 36.0;
        pts[0] = -b; pts[1] = -a; pts[2] = a; pts[3] = b;
        wts[0] = wb; wts[1] = wa; wts[2] = wa; wts[3] = wb;
    }
}

/* Restates _make_Kes_and_weights (src/TopOptProblems/matrices_and_vectors.jl
 * :126-177) for one isoparametric Q1 cell of size h: C from the 3D Lame
 * constants (:83-87, hence plane strain in 2D), Ke[dim(a-1)+d1, dim(b-1)+d2]
 * += grad(phi_a) . C . grad(phi_b) dOmega. */
int orc_element_stiffness(int dim, double E, double nu, const double* h,
                          int quad_order, double* Ke) {
    if (dim != 2 && dim != 3) return -1;
    if (quad_order < 2 || quad_order > 4) return -1;
    int nnpc = 1 << dim, ndpc = dim * nnpc;
    double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    double mu = E / (2 * (1 + nu));
    double pts[4], wts[4];
    gauss_rule(quad_order, pts, wts);
    memset(Ke, 0, sizeof(double) * ndpc * ndpc);
    double detJ = 1.0;
    for (int d = 0; d < dim; ++d) detJ *= h[d] / 2.0;
    int nq = quad_order;
    int nqt = dim == 2 ? nq * nq : nq * nq * nq;
    for (int q = 0; q < nqt; ++q) {
        double xi[3], w = 1.0;
        int qq = q;
        for (int d = 0; d < dim; ++d) {
            xi[d] = pts[qq % nq];
            w *= wts[qq % nq];
            qq /= nq;
        }
        double dN[8][3];
        for (int a = 0; a < nnpc; ++a) {
            for (int d = 0; d < dim; ++d) {
                double s = 2.0 * corner_off(dim, a, d) - 1.0;
                double v = s / 2.0;
                for (int e = 0; e < dim; ++e)
                    if (e != d) {
                        double se = 2.0 * corner_off(dim, a, e) - 1.0;
                        v *= (1.0 + se * xi[e]) / 2.0;
                    }
                dN[a][d] = v * 2.0 / h[d];
            }
        }
        double dOmega = detJ * w;
        for (int a = 0; a < nnpc; ++a)
            for (int b = 0; b < nnpc; ++b) {
                double gg = 0.0;
                for (int d = 0; d < dim; ++d) gg += dN[a][d] * dN[b][d];
                for (int i = 0; i < dim; ++i)
                    for (int k = 0; k < dim; ++k) {
                        double v = lam * dN[a][i] * dN[b][k] +
                                   mu * ((i == k ? gg : 0.0) + dN[a][k] * dN[b][i]);
                        Ke[(dim * a + i) * ndpc + dim * b + k] += v * dOmega;
                    }
            }
    }
    return 0;
}

/* get_rho_drho (src/Utilities/penalties.jl:121-130) evaluates the penalty on
 * a ForwardDiff dual number; the same here: (value, d/dx) pairs pushed
 * through each elementary operation, so the derivative never comes from a
 * hand-derived closed form. */
typedef struct { double v, d; } dual;
static dual dual_const(double c) { dual r = {c, 0.0}; return r; }
static dual dual_add(dual a, dual b) { dual r = {a.v + b.v, a.d + b.d}; return r; }
static dual dual_sub(dual a, dual b) { dual r = {a.v - b.v, a.d - b.d}; return r; }
static dual dual_mul(dual a, dual b) { dual r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static dual dual_div(dual a, dual b) {
    dual r = {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
    return r;
}
static dual dual_exp(dual a) { double e = exp(a.v); dual r = {e, e * a.d}; return r; }
static dual dual_sinh(dual a) { dual r = {sinh(a.v), cosh(a.v) * a.d}; return r; }
/* x^p for a constant real p (Base.:^(::Dual, ::Real) in ForwardDiff: the
 * derivative p x^(p-1), with p == 1 giving 1 at x == 0). */
static dual dual_pow(dual a, double p) {
    dual r = {pow(a.v, p), (p == 1.0 ? 1.0 : p * pow(a.v, p - 1.0)) * a.d};
    return r;
}

/* Penalty families (src/Utilities/penalties.jl:30-56) after an optional
 * projection (ProjectedPenaltyFun :63-93): kind 0 power, 1 rational, 2 sinh;
 * proj 0 none, 1 Heaviside, 2 sigmoid. */
static dual orc_penalty_dual(dual x, int kind, double p, int proj, double beta) {
    dual y = x;
    const dual one = dual_const(1.0);
    if (proj == 1) {         /* 1 - exp(-b x) + x exp(-b) */
        dual t = dual_exp(dual_mul(dual_const(-beta), x));
        y = dual_add(dual_sub(one, t), dual_mul(x, dual_const(exp(-beta))));
    } else if (proj == 2) {  /* 1 / (1 + exp((b + 1)(-x + 0.5))) */
        dual t = dual_exp(dual_mul(dual_const(beta + 1.0),
                                   dual_add(dual_mul(dual_const(-1.0), x), dual_const(0.5))));
        y = dual_div(one, dual_add(one, t));
    }
    if (kind == 1)           /* y / (1 + p (1 - y)) */
        return dual_div(y, dual_add(one, dual_mul(dual_const(p), dual_sub(one, y))));
    if (kind == 2)           /* sinh(p y) / sinh(p) */
        return dual_div(dual_sinh(dual_mul(dual_const(p), y)), dual_const(sinh(p)));
    return dual_pow(y, p);
}

/* density(x, xmin) = x (1 - xmin) + xmin (src/Utilities/utils.jl:77); pbi
 * selects PENALTY_BEFORE_INTERPOLATION (src/TopOpt.jl:16-23). */
void orc_rho_drho_fun(double x, int kind, double p, int proj, double beta, double xmin,
                      int pbi, double* rho, double* drho) {
    dual d = {x, 1.0}, r;
    const dual a = dual_const(1.0 - xmin), b = dual_const(xmin);
    if (pbi)
        r = dual_add(dual_mul(orc_penalty_dual(d, kind, p, proj, beta), a), b);
    else
        r = orc_penalty_dual(dual_add(dual_mul(d, a), b), kind, p, proj, beta);
    *rho = r.v;
    *drho = r.d;
}

/* HeavisideProjectionFun / SigmoidProjectionFun alone (penalties.jl:76-95):
 * the projection is PowerPenaltyFun(1) after it. */
void orc_project(double x, int proj, double beta, double* v, double* dv) {
    dual d = {x, 1.0};
    dual r = orc_penalty_dual(d, 0, 1.0, proj, beta);
    *v = r.v;
    *dv = r.d;
}

/* PowerPenaltyFun (:30-33) without projection. */
void orc_rho_drho(double x, double p, double xmin, int pbi, double* rho,
                  double* drho) {
    orc_rho_drho_fun(x, 0, p, 0, 0.0, xmin, pbi, rho, drho);
}

int orc_op_init(orc_op* op) {
    int ndpc = op->ndpc;
    int64_t nel = op->nel, ndof = op->ndof;
    op->dc_off = (int64_t*)calloc((size_t)ndof + 1, sizeof(int64_t));
    op->dc_val = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)(nel * ndpc));
    op->xes = (double*)malloc(sizeof(double) * (size_t)(nel * ndpc));
    op->px = (double*)malloc(sizeof(double) * (size_t)nel);
    if (!op->dc_off || !op->dc_val || !op->xes || !op->px) return -1;
    /* dof_cells (src/TopOptProblems/metadata.jl:64-76): per DOF, the
     * (cell, local dof) pairs in ascending cell order. */
    for (int64_t e = 0; e < nel; ++e)
        for (int i = 0; i < ndpc; ++i) op->dc_off[op->cell_dofs[e * ndpc + i]]++;
    for (int64_t d = 0; d < ndof; ++d) op->dc_off[d + 1] += op->dc_off[d];
    int64_t* fill = (int64_t*)calloc((size_t)ndof, sizeof(int64_t));
    if (!fill) return -1;
    for (int64_t e = 0; e < nel; ++e)
        for (int i = 0; i < ndpc; ++i) {
            int64_t d = op->cell_dofs[e * ndpc + i] - 1;
            int64_t pos = op->dc_off[d] + fill[d]++;
            op->dc_val[2 * pos] = e;
            op->dc_val[2 * pos + 1] = i;
        }
    free(fill);
    /* meandiag = sum_e sumdiag(Ke_e) (src/FEA/solvers_api.jl:526-527). */
    double tr = 0.0;
    for (int i = 0; i < ndpc; ++i) tr += op->Ke[i * ndpc + i];
    double md = 0.0;
    for (int64_t e = 0; e < nel; ++e) md += tr;
    op->meandiag = md;
    for (int64_t e = 0; e < nel; ++e) op->px[e] = 1.0;
    return 0;
}

void orc_op_free(orc_op* op) {
    free(op->dc_off); free(op->dc_val); free(op->xes); free(op->px);
    op->dc_off = op->dc_val = NULL; op->xes = op->px = NULL;
}

void orc_op_set_vars(orc_op* op, const double* vars) {
    for (int64_t e = 0; e < op->nel; ++e) {
        double r, dr;
        orc_rho_drho_fun(vars[e], op->kind, op->p, op->proj, op->beta, op->xmin, op->pbi, &r,
                         &dr);
        op->px[e] = r;
    }
}

/* mul!(y, ::MatrixFreeOperator, x) (src/FEA/matrix_free_operator.jl:66-105):
 * per element xes[e] = px_e * bcmatrix(Ke) * x_e, fixed DOFs y = meandiag*x,
 * free DOFs gather xes over dof_cells. */
void orc_matvec(const orc_op* op, const double* x, double* y) {
    int ndpc = op->ndpc;
    const double* Ke = op->Ke;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < op->nel; ++e) {
        double xe[24];
        int mask[24];
        for (int i = 0; i < ndpc; ++i) {
            int64_t d = op->cell_dofs[e * ndpc + i] - 1;
            mask[i] = !op->fixed[d];
            xe[i] = x[d];
        }
        double px = op->px[e];
        double* out = op->xes + e * ndpc;
        for (int i = 0; i < ndpc; ++i) {
            double s = 0.0;
            if (mask[i])
                for (int j = 0; j < ndpc; ++j)
                    if (mask[j]) s += Ke[i * ndpc + j] * xe[j];
            out[i] = px * s;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t d = 0; d < op->ndof; ++d) {
        if (op->fixed[d]) {
            y[d] = op->meandiag * x[d];
        } else {
            double s = 0.0;
            for (int64_t k = op->dc_off[d]; k < op->dc_off[d + 1]; ++k)
                s += op->xes[op->dc_val[2 * k] * ndpc + op->dc_val[2 * k + 1]];
            y[d] = s;
        }
    }
}

/* Jacobi diagonal of K(rho) for the current densities: sum over dof_cells of
 * px_e * Ke[local, local] (the assembled-K diagonal that
 * DiagonalPreconditioner reads, src/FEA/solvers_api.jl:187-192), meandiag on
 * fixed DOFs.  Rebuilt for every call (the reference caches its first). */
void orc_jacobi_diag(const orc_op* op, double* dg) {
    int ndpc = op->ndpc;
#pragma omp parallel for schedule(static)
    for (int64_t d = 0; d < op->ndof; ++d) {
        if (op->fixed[d]) { dg[d] = op->meandiag; continue; }
        double s = 0.0;
        for (int64_t k = op->dc_off[d]; k < op->dc_off[d + 1]; ++k) {
            int64_t e = op->dc_val[2 * k];
            int i = (int)op->dc_val[2 * k + 1];
            s += op->px[e] * op->Ke[i * ndpc + i];
        }
        dg[d] = s;
    }
}

/* Dot product with a fixed summation order for a given thread count: each
 * thread sums one contiguous chunk, the chunk sums are added in thread order
 * (an OpenMP reduction clause combines in an unspecified order, which made
 * repeated solves differ). */
static double vdot(int64_t n, const double* a, const double* b) {
#ifdef _OPENMP
    double part[256];
    int nt = 1;
#pragma omp parallel num_threads(omp_get_max_threads() < 256 ? omp_get_max_threads() : 256)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const int64_t chunk = (n + T - 1) / T;
        const int64_t i0 = t * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
        double s = 0.0;
        for (int64_t i = i0; i < i1; ++i) s += a[i] * b[i];
        part[t] = s;
        if (t == 0) nt = T;
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += part[t];
    return s;
#else
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
#endif
}

/* cg_solve! -> IterativeSolvers.cg! (src/FEA/solvers_api.jl:177-220).  The
 * update order restates the library's preconditioned CG iterable: tolerance
 * max(reltol*||r0||, abstol), c = Pl\r, rho = c.r, u = c + (rho/rho_prev) u,
 * c = A u, alpha = rho/(u.c), x += alpha u, r -= alpha c, stop when
 * ||r|| <= tolerance or maxiter iterations are done. */
int64_t orc_pcg(const orc_op* op, const double* b, double* x, double abstol,
                double reltol, int64_t maxiter, int initially_zero,
                int use_jacobi, double* resid) {
    int64_t n = op->ndof;
    double* r = (double*)malloc(sizeof(double) * n);
    double* u = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)malloc(sizeof(double) * n);
    double* dg = (double*)malloc(sizeof(double) * n);
    if (use_jacobi) orc_jacobi_diag(op, dg);
    if (initially_zero) {
        for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }
    } else {
        orc_matvec(op, x, c);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - c[i];
    }
    double res = sqrt(vdot(n, r, r));
    double tol = reltol * res > abstol ? reltol * res : abstol;
    double rho = 1.0;
    int64_t it = 0;
    while (it < maxiter && !(res <= tol)) {
        for (int64_t i = 0; i < n; ++i) c[i] = use_jacobi ? r[i] / dg[i] : r[i];
        double rho_prev = rho;
        rho = vdot(n, c, r);
        double beta = rho / rho_prev;
        for (int64_t i = 0; i < n; ++i) u[i] = c[i] + beta * u[i];
        orc_matvec(op, u, c);
        double alpha = rho / vdot(n, u, c);
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * u[i];
            r[i] -= alpha * c[i];
        }
        res = sqrt(vdot(n, r, r));
        ++it;
    }
    if (resid) *resid = res;
    free(r); free(u); free(c); free(dg);
    return it;
}

/* PCG stopped by EnergyCriteria (src/FEA/convergence_criteria.jl:26-45) in
 * IterativeSolvers' order: before every iteration, unless the iteration
 * count has reached maxiter (the library tests `iteration >= maxiter` first
 * and short-circuits), evaluate on the current x and recursive residual r
 * x'Ax = b.x - x.r, change = x'Ax - energy; NaN -> status -1, x'Ax < 0 ->
 * status -2 (DomainError); converged when |change| / x'Ax <= tol and
 * |x.r| / x'Ax <= tol (status 1); energy := x'Ax.  *energy carries the
 * criterion's state in and out.  Returns the iteration count. */
int64_t orc_pcg_energy(const orc_op* op, const double* b, double* x, double tol,
                       int64_t maxiter, int initially_zero, int use_jacobi, double* energy,
                       int* status, double* resid) {
    int64_t n = op->ndof;
    double* r = (double*)malloc(sizeof(double) * n);
    double* u = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)malloc(sizeof(double) * n);
    double* dg = (double*)malloc(sizeof(double) * n);
    if (use_jacobi) orc_jacobi_diag(op, dg);
    if (initially_zero) {
        for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }
    } else {
        orc_matvec(op, x, c);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - c[i];
    }
    double rho = 1.0;
    int64_t it = 0;
    *status = 0;
    for (;;) {
        if (it >= maxiter) break;
        double xtr = vdot(n, x, r);
        double xAx = vdot(n, b, x) - xtr;
        double change = xAx - *energy;
        if (isnan(change) || isnan(xAx)) { *status = -1; break; }
        if (xAx < 0) { *status = -2; break; }
        int conv = fabs(change) / xAx <= tol && fabs(xtr) / xAx <= tol;
        *energy = xAx;
        if (conv) { *status = 1; break; }
        for (int64_t i = 0; i < n; ++i) c[i] = use_jacobi ? r[i] / dg[i] : r[i];
        double rho_prev = rho;
        rho = vdot(n, c, r);
        double beta = rho / rho_prev;
        for (int64_t i = 0; i < n; ++i) u[i] = c[i] + beta * u[i];
        orc_matvec(op, u, c);
        double alpha = rho / vdot(n, u, c);
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * u[i];
            r[i] -= alpha * c[i];
        }
        ++it;
    }
    if (resid) *resid = sqrt(vdot(n, r, r));
    free(r); free(u); free(c); free(dg);
    return it;
}

/* compute_element_energy (src/Functions/compute_element_energy.jl:18-38) with
 * the unconstrained Ke (rawmatrix). */
double orc_element_energy(const orc_op* op, const double* u,
                          const double* vars, double* ce, double* grad) {
    int ndpc = op->ndpc;
    double obj = 0.0;
#pragma omp parallel for reduction(+ : obj) schedule(static)
    for (int64_t e = 0; e < op->nel; ++e) {
        double ue[24];
        for (int i = 0; i < ndpc; ++i) ue[i] = u[op->cell_dofs[e * ndpc + i] - 1];
        double s = 0.0;
        for (int w = 0; w < ndpc; ++w)
            for (int v = 0; v < ndpc; ++v) s += ue[v] * op->Ke[v * ndpc + w] * ue[w];
        double r, dr;
        orc_rho_drho_fun(vars[e], op->kind, op->p, op->proj, op->beta, op->xmin, op->pbi, &r,
                         &dr);
        ce[e] = s;
        grad[e] = -dr * s;
        obj += r * s;
    }
    return obj;
}

/* ---------------- density filter ---------------- */

typedef struct { int64_t* v; int64_t n, cap; } ivec;
static void ipush(ivec* a, int64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = (int64_t*)realloc(a->v, sizeof(int64_t) * a->cap);
    }
    a->v[a->n++] = x;
}
typedef struct { double* v; int64_t n, cap; } dvec;
static void dpush(dvec* a, double x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = (double*)realloc(a->v, sizeof(double) * a->cap);
    }
    a->v[a->n++] = x;
}

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Cells containing node n, ascending (Metadata node_cells,
 * src/TopOptProblems/metadata.jl:99-109). */
static int node_cells(int dim, const int64_t* nels, int64_t n, int64_t* out) {
    int64_t NX = nels[0] + 1, NY = nels[1] + 1;
    int64_t i = n % NX, j = (n / NX) % NY, k = dim == 3 ? n / (NX * NY) : 0;
    int cnt = 0;
    int kmax = dim == 3 ? 1 : 0;
    for (int dk = -1; dk <= 0; ++dk) {
        if (dim == 2 && dk != 0) continue;
        int64_t kk = k + dk;
        if (dim == 3 && (kk < 0 || kk >= nels[2])) continue;
        for (int dj = -1; dj <= 0; ++dj) {
            int64_t jj = j + dj;
            if (jj < 0 || jj >= nels[1]) continue;
            for (int di = -1; di <= 0; ++di) {
                int64_t ii = i + di;
                if (ii < 0 || ii >= nels[0]) continue;
                out[cnt++] = ii + nels[0] * (jj + nels[1] * (dim == 3 ? kk : 0));
            }
        }
    }
    (void)kmax;
    return cnt;
}

static void node_coord(int dim, const int64_t* nels, const double* h, int64_t n,
                       double* xc) {
    int64_t NX = nels[0] + 1, NY = nels[1] + 1;
    xc[0] = (double)(n % NX) * h[0];
    xc[1] = (double)((n / NX) % NY) * h[1];
    if (dim == 3) xc[2] = (double)(n / (NX * NY)) * h[2];
}

/* get_neighbour_info (src/CheqFilters/CheqFilters.jl:66-118) for one cell:
 * breadth-first traversal from the cell, pushing every node of every visited
 * cell that lies strictly within rmin of the cell centre (so a node is pushed
 * once per visited cell containing it) with weight max(rmin - d, 0). */
static void bfs_cell(int dim, const int64_t* nels, const double* h, double rmin,
                     int64_t cell, int64_t* stamp, ivec* queue, ivec* nodes,
                     dvec* wts) {
    int nnpc = 1 << dim;
    double center[3] = {0, 0, 0};
    for (int a = 0; a < nnpc; ++a) {
        double xc[3];
        node_coord(dim, nels, h, cell_node(dim, nels, cell, a), xc);
        for (int d = 0; d < dim; ++d) center[d] += xc[d];
    }
    for (int d = 0; d < dim; ++d) center[d] /= nnpc;
    queue->n = 0; nodes->n = 0; wts->n = 0;
    int64_t head = 0;
    ipush(queue, cell);
    stamp[cell] = cell + 1;
    while (head < queue->n) {
        int64_t cid = queue->v[head++];
        for (int a = 0; a < nnpc; ++a) {
            int64_t n = cell_node(dim, nels, cid, a);
            double xc[3], dist = 0.0;
            node_coord(dim, nels, h, n, xc);
            for (int d = 0; d < dim; ++d) dist += (xc[d] - center[d]) * (xc[d] - center[d]);
            dist = sqrt(dist);
            if (dist < rmin) {
                ipush(nodes, n);
                dpush(wts, rmin - dist > 0 ? rmin - dist : 0.0);
                int64_t nc[8];
                int cnt = node_cells(dim, nels, n, nc);
                for (int q = 0; q < cnt; ++q)
                    if (stamp[nc[q]] != cell + 1) {
                        stamp[nc[q]] = cell + 1;
                        ipush(queue, nc[q]);
                    }
            }
        }
    }
}

/* getJacobian (src/CheqFilters/density_filter.jl:49-108): J = W * A with
 * W[i, n] = sum of the pushed weights of node n / sum of all pushed weights of
 * cell i, and A[n, c] = V_c / sum_{c' in n} V_c' (scalecols!, :110-121). */
int64_t orc_filter_build(int dim, const int64_t* nels, const double* h,
                         double rmin, int64_t* rowptr, int64_t* col,
                         double* val) {
    int64_t nel = orc_ncells(dim, nels);
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) vol *= h[d];
    int64_t* rowlen = (int64_t*)calloc((size_t)nel, sizeof(int64_t));
    int empty_all = 1;
    int pass_count = rowptr ? 2 : 1;
    int64_t total = 0;
    for (int pass = 0; pass < pass_count; ++pass) {
        if (pass == 1) {
            rowptr[0] = 0;
            for (int64_t i = 0; i < nel; ++i) rowptr[i + 1] = rowptr[i] + rowlen[i];
        }
#pragma omp parallel
        {
            int64_t* stamp = (int64_t*)calloc((size_t)nel, sizeof(int64_t));
            double* acc = (double*)calloc((size_t)nel, sizeof(double));
            int64_t* seen = (int64_t*)calloc((size_t)nel, sizeof(int64_t));
            ivec queue = {0}, nodes = {0}, touched = {0};
            dvec wts = {0};
            int local_nonempty = 0;
#pragma omp for schedule(dynamic, 64)
            for (int64_t i = 0; i < nel; ++i) {
                bfs_cell(dim, nels, h, rmin, i, stamp, &queue, &nodes, &wts);
                if (nodes.n == 0) { rowlen[i] = 0; continue; }
                local_nonempty = 1;
                double wsum = 0.0;
                for (int64_t q = 0; q < wts.n; ++q) wsum += wts.v[q];
                touched.n = 0;
                for (int64_t q = 0; q < nodes.n; ++q) {
                    int64_t n = nodes.v[q];
                    double wn = wts.v[q] / wsum;
                    int64_t nc[8];
                    int cnt = node_cells(dim, nels, n, nc);
                    double vs = 0.0;
                    for (int t = 0; t < cnt; ++t) vs += vol;
                    for (int t = 0; t < cnt; ++t) {
                        int64_t c = nc[t];
                        if (seen[c] != i + 1) {
                            seen[c] = i + 1;
                            acc[c] = 0.0;
                            ipush(&touched, c);
                        }
                        acc[c] += wn * (vol / vs);
                    }
                }
                qsort(touched.v, (size_t)touched.n, sizeof(int64_t), cmp_i64);
                if (pass == 0 && rowptr == NULL) {
                    rowlen[i] = touched.n;
                } else if (pass == 0) {
                    rowlen[i] = touched.n;
                } else {
                    int64_t base = rowptr[i];
                    for (int64_t q = 0; q < touched.n; ++q) {
                        col[base + q] = touched.v[q];
                        val[base + q] = acc[touched.v[q]];
                    }
                }
            }
            if (local_nonempty) {
#pragma omp atomic write
                empty_all = 0;
            }
            free(stamp); free(acc); free(seen);
            free(queue.v); free(nodes.v); free(touched.v); free(wts.v);
        }
    }
    for (int64_t i = 0; i < nel; ++i) total += rowlen[i];
    free(rowlen);
    if (empty_all) return -1;
    return total;
}

void orc_csr_mv(int64_t nrows, const int64_t* rowptr, const int64_t* col,
                const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; ++i) {
        double s = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

void orc_csr_mtv(int64_t nrows, int64_t ncols, const int64_t* rowptr,
                 const int64_t* col, const double* val, const double* x,
                 double* y) {
    for (int64_t j = 0; j < ncols; ++j) y[j] = 0.0;
    for (int64_t i = 0; i < nrows; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) y[col[k]] += val[k] * x[i];
}

/* Build-defined OC step (the reference has no OC; its SIMP driver is MMA).
 * Optimality-criteria fixed point with a density filter: dv = colsum(J)/nel,
 * x_new = clamp(x * sqrt(max(-dc,0) / (dv * lambda)), [max(0, x-move),
 * min(1, x+move)]), lambda by bisection on [0, 1e9] until
 * (l2 - l1)/(l1 + l2) <= tol so that sum(J x_new) = colsum . x_new meets
 * volfrac * nel.  x_new is the one evaluated at the last midpoint. */
/* OpenMP threads of the oracle's parallel loops (n <= 0 leaves them);
 * returns the count now in effect.  Used to time the CPU baseline. */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* get_cell_volumes (src/TopOptProblems/elementinfo.jl:180-192): the sum
 * over quadrature points of det J dV; on an axis-aligned box det J is the
 * constant prod(h_d / 2) and the weights sum to 2^dim per reference cell. */
double orc_cell_volume(int dim, const double* h, int quad_order) {
    if ((dim != 2 && dim != 3) || quad_order < 2 || quad_order > 4) return -1.0;
    double pts[4], wts[4];
    gauss_rule(quad_order, pts, wts);
    double detJ = 1.0;
    for (int d = 0; d < dim; ++d) detJ *= h[d] / 2.0;
    int nq = quad_order, nqt = dim == 2 ? nq * nq : nq * nq * nq;
    double v = 0.0;
    for (int q = 0; q < nqt; ++q) {
        double w = 1.0;
        int qq = q;
        for (int d = 0; d < dim; ++d) {
            w *= wts[qq % nq];
            qq /= nq;
        }
        v += w * detJ;
    }
    return v;
}

/* VolumeFun call (src/Functions/volume.jl:59-62, compute_volume :77-80):
 * dot(x, cellvolumes), divided by the total volume when fraction. */
double orc_volume(int64_t nel, const double* x, const double* cellvol, int fraction) {
    double v = 0.0, tot = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        v += x[e] * cellvol[e];
        tot += cellvol[e];
    }
    return fraction ? v / tot : v;
}

double orc_oc_update(int64_t nel, const double* x, const double* dc,
                     const double* colsum, double volfrac, double move,
                     double tol, double* xnew) {
    double l1 = 0.0, l2 = 1e9, lmid = 0.5 * (l1 + l2);
    double target = volfrac * (double)nel;
    while ((l2 - l1) / (l1 + l2) > tol) {
        lmid = 0.5 * (l1 + l2);
        double vol = 0.0;
#pragma omp parallel for reduction(+ : vol) schedule(static)
        for (int64_t e = 0; e < nel; ++e) {
            double dv = colsum[e] / (double)nel;
            double g = -dc[e] > 0.0 ? -dc[e] : 0.0;
            double xn = x[e] * sqrt(g / (dv * lmid));
            double lo = x[e] - move > 0.0 ? x[e] - move : 0.0;
            double hi = x[e] + move < 1.0 ? x[e] + move : 1.0;
            xn = xn < hi ? xn : hi;
            xn = xn > lo ? xn : lo;
            xnew[e] = xn;
            vol += colsum[e] * xn;
        }
        if (vol > target) l1 = lmid; else l2 = lmid;
    }
    return lmid;
}

/* DirectSolver (src/FEA/solvers_api.jl:131-168): Cholesky of the assembled
 * K (src/TopOptProblems/assemble.jl:28-90) after apply! — here the reduced
 * free-DOF system in banded form; prescribed DOFs get 0. */
int orc_direct_solve(const orc_op* op, const double* f, double* u) {
    int ndpc = op->ndpc;
    int64_t n = op->ndof;
    int64_t* map = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t nf = 0;
    for (int64_t d = 0; d < n; ++d) map[d] = op->fixed[d] ? -1 : nf++;
    int64_t bw = 0;
    for (int64_t e = 0; e < op->nel; ++e)
        for (int i = 0; i < ndpc; ++i)
            for (int j = 0; j < ndpc; ++j) {
                int64_t a = map[op->cell_dofs[e * ndpc + i] - 1];
                int64_t b = map[op->cell_dofs[e * ndpc + j] - 1];
                if (a >= 0 && b >= 0 && a - b > bw) bw = a - b;
            }
    int64_t W = bw + 1;
    double* L = (double*)calloc((size_t)(nf * W), sizeof(double)); /* L[i*W + (i-j)] */
    if (!L) { free(map); return -1; }
    for (int64_t e = 0; e < op->nel; ++e)
        for (int i = 0; i < ndpc; ++i) {
            int64_t a = map[op->cell_dofs[e * ndpc + i] - 1];
            if (a < 0) continue;
            for (int j = 0; j < ndpc; ++j) {
                int64_t b = map[op->cell_dofs[e * ndpc + j] - 1];
                if (b < 0 || b > a) continue;
                L[a * W + (a - b)] += op->px[e] * op->Ke[i * ndpc + j];
            }
        }
    for (int64_t j = 0; j < nf; ++j) {
        double s = L[j * W];
        int64_t k0 = j - bw > 0 ? j - bw : 0;
        for (int64_t k = k0; k < j; ++k) s -= L[j * W + (j - k)] * L[j * W + (j - k)];
        if (!(s > 0.0)) {
            free(L); free(map);
            for (int64_t d = 0; d < n; ++d) u[d] = NAN;
            return 1;
        }
        double ljj = sqrt(s);
        L[j * W] = ljj;
        int64_t iend = j + bw < nf - 1 ? j + bw : nf - 1;
#pragma omp parallel for schedule(static)
        for (int64_t i = j + 1; i <= iend; ++i) {
            double t = L[i * W + (i - j)];
            int64_t kk0 = i - bw > k0 ? i - bw : k0;
            for (int64_t k = kk0; k < j; ++k) t -= L[i * W + (i - k)] * L[j * W + (j - k)];
            L[i * W + (i - j)] = t / ljj;
        }
    }
    double* y = (double*)malloc(sizeof(double) * nf);
    for (int64_t d = 0; d < n; ++d)
        if (map[d] >= 0) y[map[d]] = f[d];
    for (int64_t i = 0; i < nf; ++i) {
        double s = y[i];
        int64_t k0 = i - bw > 0 ? i - bw : 0;
        for (int64_t k = k0; k < i; ++k) s -= L[i * W + (i - k)] * y[k];
        y[i] = s / L[i * W];
    }
    for (int64_t i = nf - 1; i >= 0; --i) {
        double s = y[i];
        int64_t kend = i + bw < nf - 1 ? i + bw : nf - 1;
        for (int64_t k = i + 1; k <= kend; ++k) s -= L[k * W + (k - i)] * y[k];
        y[i] = s / L[i * W];
    }
    for (int64_t d = 0; d < n; ++d) u[d] = map[d] >= 0 ? y[map[d]] : 0.0;
    free(y); free(L); free(map);
    return 0;
}
