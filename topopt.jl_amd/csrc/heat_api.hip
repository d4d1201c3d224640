// C ABI of the heat-conduction path: context set-up, densities, the
// matrix-free conduction operator, the Jacobi-PCG solve with prescribed
// temperatures, and the element products of the thermal gradients
// (TopOpt.jl HeatConductionProblem, src/TopOptProblems/problem_types.jl:783-1030;
// ThermalComplianceFun, src/Functions/thermal_compliance.jl; TemperatureFun,
// src/Functions/temperature.jl).
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/topopt_hip.h"
#include "heat_internal.h"
#include "k_common.h"

using namespace tohip;

struct tohip_heat {
    HeatGrid g{};
    HeatKe ke{};
    double xmin = 1e-3;
    Penalty pen{TOHIP_PEN_POWER, TOHIP_PROJ_NONE, 1.0, 0.0};
    int pbi = 1;
    double meandiag = 0.0;
    bool inhom = false;              // some prescribed temperature is nonzero
    std::vector<int64_t> dof_node;   // caller DOF (0-based) -> node id
    hipStream_t stream = nullptr;
    uint8_t* fixmask = nullptr;
    double *fixv = nullptr, *x = nullptr, *E = nullptr, *dinv = nullptr;
    double *u = nullptr, *r = nullptr, *p[2] = {nullptr, nullptr}, *q = nullptr, *b = nullptr,
           *t = nullptr, *lam = nullptr, *cell = nullptr, *grad = nullptr;
    double *partA = nullptr, *partB = nullptr;
    PcgScalars *sc = nullptr, *h_sc = nullptr;
    hipGraphExec_t graph = nullptr;
    std::string err;
};

namespace {

constexpr int kHeatGraphIters = 32;
thread_local std::string g_heat_error;

#define HCK(call)                                                  \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) {                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_); \
            return TOHIP_ERR_HIP;                                  \
        }                                                          \
    } while (0)

// Ke[a][b] = k * integral over the box cell of grad phi_a . grad phi_b,
// exactly: the bilinear / trilinear shape functions are products of 1D hat
// functions l(x) with integral l_a' l_b' = +-1/h and integral l_a l_b = h/3 or
// h/6 (same / other end), so each term is a sum over the derivative
// direction of products of 1D integrals.
void heat_ke(int dim, const double* h, double k, HeatKe& ke) {
    const int nc = 1 << dim;
    auto off = [](int b, int d) {
        const int ox = (b & 1) ^ ((b >> 1) & 1), oy = (b >> 1) & 1, oz = b >> 2;
        return d == 0 ? ox : d == 1 ? oy : oz;
    };
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            double s = 0.0;
            for (int d = 0; d < dim; ++d) {
                double term = (off(a, d) == off(b, d) ? 1.0 : -1.0) / h[d];
                for (int e = 0; e < dim; ++e)
                    if (e != d) term *= (off(a, e) == off(b, e) ? h[e] / 3.0 : h[e] / 6.0);
                s += term;
            }
            ke.k[a * nc + b] = k * s;
        }
}

template <typename T>
int halloc(tohip_heat* ctx, T** p, size_t n) {
    HCK(hipMalloc((void**)p, (n ? n : 1) * sizeof(T)));
    return 0;
}

int upload_nodes(tohip_heat* ctx, const double* host, double* dev) {
    const int64_t n = ctx->g.nnodes;
    std::vector<double> tmp(n);
    for (int64_t d = 0; d < n; ++d) tmp[ctx->dof_node[d]] = host[d];
    HCK(hipMemcpyAsync(dev, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int download_nodes(tohip_heat* ctx, const double* dev, double* host) {
    const int64_t n = ctx->g.nnodes;
    std::vector<double> tmp(n);
    HCK(hipMemcpyAsync(tmp.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    for (int64_t d = 0; d < n; ++d) host[d] = tmp[ctx->dof_node[d]];
    return 0;
}

bool finite_all(const double* v, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

int refresh(tohip_heat* ctx) {
    HCK(launch_heat_density(ctx->g.ncells, ctx->x, ctx->xmin, ctx->pen, ctx->pbi, ctx->E,
                            ctx->stream));
    HCK(launch_heat_diag(ctx->g, ctx->ke, ctx->fixmask, ctx->E, ctx->meandiag, ctx->dinv,
                         ctx->stream));
    return 0;
}

int capture(tohip_heat* ctx) {
    if (ctx->graph) return 0;
    hipGraph_t graph;
    HCK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kHeatGraphIters; ++it) {
        int nbA = 0;
        hipError_t e1 = launch_heat_dirmatvec(ctx->g, ctx->ke, ctx->r, ctx->dinv, ctx->p[0],
                                              ctx->p[1], ctx->q, ctx->fixmask, ctx->E,
                                              ctx->meandiag, ctx->sc, ctx->partB, kVecBlocks,
                                              ctx->partA, &nbA, ctx->stream);
        hipError_t e2 = launch_pcg_update<double>(ctx->g.nnodes, ctx->u, ctx->r, ctx->p[0],
                                                  ctx->p[1], ctx->q, ctx->dinv, ctx->sc,
                                                  ctx->partA, nbA, ctx->partB, ctx->stream);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            hipStreamEndCapture(ctx->stream, &graph);
            ctx->err = "heat PCG graph capture failed";
            return TOHIP_ERR_HIP;
        }
    }
    HCK(hipStreamEndCapture(ctx->stream, &graph));
    HCK(hipGraphInstantiate(&ctx->graph, graph, nullptr, nullptr, 0));
    HCK(hipGraphDestroy(graph));
    return 0;
}

// Jacobi PCG on the solver's operator with right-hand side ctx->b, from
// u = 0 or from the current u.  IterativeSolvers' order and stopping rule
// ||r|| <= max(reltol ||r0||, abstol) or maxiter.
int solve(tohip_heat* ctx, bool zero, double abstol, double reltol, int64_t maxiter,
          int64_t* iters, double* resid) {
    const int64_t n = ctx->g.nnodes;
    PcgScalars hs{};
    hs.abstol = abstol;
    hs.reltol = reltol;
    hs.maxiter = maxiter;
    *ctx->h_sc = hs;
    HCK(hipMemcpyAsync(ctx->sc, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice,
                       ctx->stream));
    if (zero) {
        HCK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
        HCK(launch_pcg_start<double>(n, ctx->b, nullptr, ctx->r, ctx->dinv, ctx->p[0], ctx->sc,
                                     ctx->partB, 0, ctx->stream));
    } else {
        HCK(launch_heat_matvec(ctx->g, ctx->ke, ctx->u, ctx->fixmask, ctx->E, ctx->meandiag,
                               ctx->t, 0, ctx->stream));
        HCK(launch_pcg_start<double>(n, ctx->b, ctx->t, ctx->r, ctx->dinv, ctx->p[0], ctx->sc,
                                     ctx->partB, 0, ctx->stream));
    }
    if (int rc = capture(ctx)) return rc;
    for (;;) {
        HCK(hipGraphLaunch(ctx->graph, ctx->stream));
        HCK(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost,
                           ctx->stream));
        HCK(hipStreamSynchronize(ctx->stream));
        if (ctx->h_sc->done) break;
    }
    const PcgScalars& h = *ctx->h_sc;
    if (iters) *iters = h.iter;
    if (resid) *resid = std::sqrt(h.rr);
    if (h.err) {  // NaN temperatures, as the direct solver's convention
        HCK(launch_fill64(n, std::numeric_limits<double>::quiet_NaN(), ctx->u, ctx->stream));
        HCK(hipStreamSynchronize(ctx->stream));
        ctx->err = pcg_err_text(h.err);
        return TOHIP_BREAKDOWN;
    }
    return std::sqrt(h.rr) <= h.tol ? 0 : TOHIP_NOT_CONVERGED;
}

void free_heat(tohip_heat* ctx) {
    if (!ctx) return;
    if (ctx->graph) hipGraphExecDestroy(ctx->graph);
    void* bufs[] = {ctx->fixmask, ctx->fixv, ctx->x,     ctx->E,    ctx->dinv, ctx->u,
                    ctx->r,       ctx->p[0], ctx->p[1],  ctx->q,    ctx->b,    ctx->t,
                    ctx->lam,     ctx->cell, ctx->grad,  ctx->partA, ctx->partB, ctx->sc};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (ctx->h_sc) hipHostFree(ctx->h_sc);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int create_heat(tohip_heat* ctx, int dim, const int64_t* nels, const double* h, double k,
                double xmin, double penalty_p, int pbi, const int64_t* node_dofs,
                const uint8_t* fixed_mask, const double* fixed_values) {
    if (dim != 2 && dim != 3) {
        ctx->err = "dim must be 2 or 3";
        return TOHIP_ERR_ARG;
    }
    if (!nels || !h || !node_dofs || !fixed_mask) {
        ctx->err = "null argument";
        return TOHIP_ERR_ARG;
    }
    for (int d = 0; d < dim; ++d)
        if (nels[d] < 1 || !(h[d] > 0) || !std::isfinite(h[d])) {
            ctx->err = "cell counts must be >= 1 and sizes positive";
            return TOHIP_ERR_ARG;
        }
    if (!(k > 0) || !std::isfinite(k) || !(xmin >= 0) || !(xmin < 1) || !(penalty_p > 0)) {
        ctx->err = "conductivity and penalty must be positive, 0 <= xmin < 1";
        return TOHIP_ERR_ARG;
    }
    HeatGrid& g = ctx->g;
    g.dim = dim;
    g.nx = (int)nels[0];
    g.ny = (int)nels[1];
    g.nz = dim == 3 ? (int)nels[2] : 1;
    g.NX = g.nx + 1;
    g.NY = g.ny + 1;
    g.NZ = dim == 3 ? g.nz + 1 : 1;
    g.nnodes = (int64_t)g.NX * g.NY * g.NZ;
    g.ncells = (int64_t)g.nx * g.ny * g.nz;
    if (g.nnodes >= ((int64_t)1 << 31)) {
        ctx->err = "grid too large";
        return TOHIP_ERR_ARG;
    }
    heat_ke(dim, h, k, ctx->ke);
    double tr = 0.0;
    for (int a = 0; a < (1 << dim); ++a) tr += ctx->ke.k[a * (1 << dim) + a];
    // matrix-free fixed-row rule: the sum of the element traces
    // (src/FEA/solvers_api.jl:526-527)
    ctx->meandiag = tr * (double)g.ncells;
    ctx->xmin = xmin;
    ctx->pbi = pbi ? 1 : 0;
    ctx->pen = Penalty{TOHIP_PEN_POWER, TOHIP_PROJ_NONE, penalty_p, 0.0};
    // node_dofs: the 1 x nnodes Metadata.node_dofs (1-based), a permutation
    const int64_t n = g.nnodes;
    ctx->dof_node.assign(n, -1);
    for (int64_t node = 0; node < n; ++node) {
        const int64_t d = node_dofs[node] - 1;
        if (d < 0 || d >= n || ctx->dof_node[d] >= 0) {
            ctx->err = "node_dofs is not a permutation of 1..nnodes";
            return TOHIP_ERR_ARG;
        }
        ctx->dof_node[d] = node;
    }
    std::vector<uint8_t> mask(n, 0);
    std::vector<double> vals(n, 0.0);
    for (int64_t d = 0; d < n; ++d) {
        const int64_t node = ctx->dof_node[d];
        mask[node] = fixed_mask[d] ? 1 : 0;
        if (mask[node] && fixed_values) {
            if (!std::isfinite(fixed_values[d])) {
                ctx->err = "non-finite prescribed temperature";
                return TOHIP_ERR_ARG;
            }
            vals[node] = fixed_values[d];
            if (vals[node] != 0.0) ctx->inhom = true;
        }
    }
    HCK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int rc;
    if ((rc = halloc(ctx, &ctx->fixmask, n)) || (rc = halloc(ctx, &ctx->fixv, n)) ||
        (rc = halloc(ctx, &ctx->x, g.ncells)) || (rc = halloc(ctx, &ctx->E, g.ncells)) ||
        (rc = halloc(ctx, &ctx->dinv, n)) || (rc = halloc(ctx, &ctx->u, n)) ||
        (rc = halloc(ctx, &ctx->r, n)) || (rc = halloc(ctx, &ctx->p[0], n)) ||
        (rc = halloc(ctx, &ctx->p[1], n)) || (rc = halloc(ctx, &ctx->q, n)) ||
        (rc = halloc(ctx, &ctx->b, n)) || (rc = halloc(ctx, &ctx->t, n)) ||
        (rc = halloc(ctx, &ctx->lam, n)) || (rc = halloc(ctx, &ctx->cell, g.ncells)) ||
        (rc = halloc(ctx, &ctx->grad, g.ncells)) || (rc = halloc(ctx, &ctx->partA, kMaxPartials)) ||
        (rc = halloc(ctx, &ctx->partB, 2 * kMaxPartials)) || (rc = halloc(ctx, &ctx->sc, 1)))
        return rc;
    HCK(hipHostMalloc((void**)&ctx->h_sc, sizeof(PcgScalars), hipHostMallocDefault));
    HCK(hipMemcpyAsync(ctx->fixmask, mask.data(), n, hipMemcpyHostToDevice, ctx->stream));
    HCK(hipMemcpyAsync(ctx->fixv, vals.data(), n * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream));
    std::vector<double> ones(g.ncells, 1.0);
    HCK(hipMemcpyAsync(ctx->x, ones.data(), g.ncells * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream));
    HCK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
    HCK(hipMemsetAsync(ctx->p[0], 0, n * sizeof(double), ctx->stream));
    HCK(hipMemsetAsync(ctx->p[1], 0, n * sizeof(double), ctx->stream));
    if ((rc = refresh(ctx))) return rc;
    HCK(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace

extern "C" {

const char* tohip_heat_last_error(const tohip_heat* ctx) {
    return ctx ? ctx->err.c_str() : g_heat_error.c_str();
}

int tohip_heat_create(tohip_heat** out, int dim, const int64_t* nels, const double* h, double k,
                      double xmin, double penalty_p, int penalty_before_interpolation,
                      const int64_t* node_dofs, const uint8_t* fixed_mask,
                      const double* fixed_values) {
    if (!out) return TOHIP_ERR_ARG;
    *out = nullptr;
    tohip_heat* ctx = new tohip_heat();
    const int rc = create_heat(ctx, dim, nels, h, k, xmin, penalty_p,
                               penalty_before_interpolation, node_dofs, fixed_mask, fixed_values);
    if (rc) {
        g_heat_error = ctx->err;
        free_heat(ctx);
        return rc;
    }
    *out = ctx;
    return 0;
}

void tohip_heat_destroy(tohip_heat* ctx) { free_heat(ctx); }

int tohip_heat_set_penalty_fun(tohip_heat* ctx, int kind, double penalty_p, int proj,
                               double beta) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (kind < TOHIP_PEN_POWER || kind > TOHIP_PEN_SINH || proj < TOHIP_PROJ_NONE ||
        proj > TOHIP_PROJ_SIGMOID || !(penalty_p > 0) || !std::isfinite(penalty_p) ||
        (proj != TOHIP_PROJ_NONE && (!(beta >= 0) || !std::isfinite(beta)))) {
        ctx->err = "tohip_heat_set_penalty_fun: bad penalty kind, projection, p or beta";
        return TOHIP_ERR_ARG;
    }
    ctx->pen = Penalty{kind, proj, penalty_p, proj == TOHIP_PROJ_NONE ? 0.0 : beta};
    if (int rc = refresh(ctx)) return rc;
    HCK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_heat_set_densities(tohip_heat* ctx, const double* vars) {
    if (!ctx || !vars) return TOHIP_ERR_ARG;
    if (!finite_all(vars, ctx->g.ncells)) {
        ctx->err = "tohip_heat_set_densities: non-finite density";
        return TOHIP_ERR_ARG;
    }
    HCK(hipMemcpyAsync(ctx->x, vars, ctx->g.ncells * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream));
    if (int rc = refresh(ctx)) return rc;
    HCK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_heat_matvec(tohip_heat* ctx, const double* x, double* y) {
    if (!ctx || !x || !y) return TOHIP_ERR_ARG;
    if (int rc = upload_nodes(ctx, x, ctx->t)) return rc;
    HCK(launch_heat_matvec(ctx->g, ctx->ke, ctx->t, ctx->fixmask, ctx->E, ctx->meandiag, ctx->q,
                           0, ctx->stream));
    return download_nodes(ctx, ctx->q, y);
}

int tohip_heat_solve(tohip_heat* ctx, const double* rhs, double* T, double abstol, double reltol,
                     int64_t maxiter, int initially_zero, int apply_dirichlet, int64_t* iters_out,
                     double* resid_out) {
    if (!ctx || !rhs || !T) return TOHIP_ERR_ARG;
    if (!(abstol >= 0) || !(reltol >= 0) || maxiter < 0) {
        ctx->err = "tohip_heat_solve: bad tolerance or maxiter";
        return TOHIP_ERR_ARG;
    }
    const int64_t n = ctx->g.nnodes;
    if (!finite_all(rhs, n) || (!initially_zero && !finite_all(T, n))) {
        ctx->err = "tohip_heat_solve: non-finite right-hand side or initial guess";
        return TOHIP_ERR_ARG;
    }
    if (int rc = upload_nodes(ctx, rhs, ctx->q)) return rc;
    const bool lift = apply_dirichlet && ctx->inhom;
    if (lift)
        HCK(launch_heat_matvec(ctx->g, ctx->ke, ctx->fixv, ctx->fixmask, ctx->E, ctx->meandiag,
                               ctx->t, 2, ctx->stream));
    HCK(launch_heat_rhs(n, ctx->q, lift ? ctx->t : nullptr, ctx->fixmask, ctx->fixv,
                        ctx->meandiag, lift ? 1 : 0, ctx->b, ctx->stream));
    if (!initially_zero)
        if (int rc = upload_nodes(ctx, T, ctx->u)) return rc;
    const int rc = solve(ctx, initially_zero != 0, abstol, reltol, maxiter, iters_out, resid_out);
    if (rc < 0) return rc;
    if (int rc2 = download_nodes(ctx, ctx->u, T)) return rc2;
    return rc;
}

int tohip_heat_cell_products(tohip_heat* ctx, const double* lam, const double* T, double sign,
                             double* cell_out, double* grad_out) {
    if (!ctx || !lam || !T || !cell_out || !grad_out) return TOHIP_ERR_ARG;
    if (int rc = upload_nodes(ctx, lam, ctx->lam)) return rc;
    if (int rc = upload_nodes(ctx, T, ctx->t)) return rc;
    HCK(launch_heat_cell_products(ctx->g, ctx->ke, ctx->lam, ctx->t, ctx->x, ctx->xmin, ctx->pen,
                                  ctx->pbi, sign, ctx->cell, ctx->grad, ctx->stream));
    const size_t bytes = ctx->g.ncells * sizeof(double);
    HCK(hipMemcpyAsync(cell_out, ctx->cell, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipMemcpyAsync(grad_out, ctx->grad, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_heat_bench_matvec(tohip_heat* ctx, int reps, double* avg_us) {
    if (!ctx || reps < 1 || !avg_us) return TOHIP_ERR_ARG;
    hipEvent_t e0, e1;
    HCK(hipEventCreate(&e0));
    HCK(hipEventCreate(&e1));
    HCK(launch_heat_matvec(ctx->g, ctx->ke, ctx->u, ctx->fixmask, ctx->E, ctx->meandiag, ctx->q,
                           0, ctx->stream));  // warm-up
    HCK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i)
        HCK(launch_heat_matvec(ctx->g, ctx->ke, ctx->u, ctx->fixmask, ctx->E, ctx->meandiag,
                               ctx->q, 0, ctx->stream));
    HCK(hipEventRecord(e1, ctx->stream));
    HCK(hipEventSynchronize(e1));
    float ms = 0;
    HCK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *avg_us = 1e3 * ms / reps;
    return 0;
}

int tohip_heat_element_matrix(const tohip_heat* ctx, double* ke_out) {
    if (!ctx || !ke_out) return TOHIP_ERR_ARG;
    const int nc = 1 << ctx->g.dim;
    for (int i = 0; i < nc * nc; ++i) ke_out[i] = ctx->ke.k[i];
    return 0;
}

}  // extern "C"
