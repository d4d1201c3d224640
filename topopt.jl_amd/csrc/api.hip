// C-ABI of libtopopt_hip: context, element operator, solves, SIMP loop.
// See include/topopt_hip.h for the contract of every entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/topopt_hip.h"
#include "tohip_internal.h"

using namespace tohip;

namespace {
thread_local std::string g_create_error;
constexpr int kGraphIters = 32;  // PCG iterations per captured graph (even)
}  // namespace

struct tohip_ctx {
    Grid g{};
    double E = 1, nu = 0.3, xmin = 1e-3, p = 1;
    double h[3] = {1, 1, 1};
    int pbi = 1;
    double meandiag = 0;
    ElemOp<double> op64{};
    ElemOp<float> op32{};
    std::vector<int64_t> perm;  // internal DOF -> caller DOF (0-based); empty = identity
    std::vector<uint8_t> fixed_internal;
    hipStream_t stream = nullptr;
    std::string err;

    // device buffers
    uint8_t* fixmask = nullptr;
    double *xt = nullptr, *E64 = nullptr;
    float* E32 = nullptr;
    double *f = nullptr, *u = nullptr, *b64 = nullptr, *r64 = nullptr, *t64 = nullptr,
           *dinv64 = nullptr, *p64[2] = {nullptr, nullptr}, *q64 = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *dinv32 = nullptr,
          *p32[2] = {nullptr, nullptr}, *q32 = nullptr;
    PcgScalars *sc64 = nullptr, *sc32 = nullptr;
    double* partials = nullptr;
    double* dscal = nullptr;           // [0] obj, [1] rr (refinement), [2] volume
    unsigned int* counters = nullptr;  // [0] energy, [1] residual, [2] volume
    // filter / OC
    int* f_offs = nullptr;
    double* f_wts = nullptr;
    int f_count = 0;
    double f_rmin = -1;
    double *f_invden = nullptr, *f_colsum = nullptr;
    double *xd = nullptr, *xnew = nullptr, *dc = nullptr, *ce = nullptr, *grad = nullptr,
           *cellbuf = nullptr;
    OcState* oc = nullptr;
    // pinned host mirrors
    PcgScalars* h_sc = nullptr;
    double* h_scal = nullptr;
    OcState* h_oc = nullptr;
    // captured PCG iteration graphs
    hipGraphExec_t g64 = nullptr, g32 = nullptr;
    bool have_u = false;
};

#define CK(call)                                                                        \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return TOHIP_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)

namespace {

// ---------------------------------------------------------------- Ke ----
// Exact element stiffness of a Q1 cell on an h-box as tensor products of 1D
// integrals (phi0 = 1 - t/h, phi1 = t/h):
//   Mv[a][b] = int phi_a phi_b, Dd[a][b] = int phi_a' phi_b',
//   Gd[a][b] = int phi_a' phi_b.
// Corner c has bit d = 1 when the node sits at the far side of axis d.
// Ke[(c,i),(c',k)] = lam I(i,k) + mu (delta_ik sum_j I(j,j) + I(k,i)) with
// I(i,k)[c][c'] = int d_i N_c d_k N_c'.  Plane strain in 2D (3D Lame lambda),
// as TopOpt.jl's SymmetricTensor{4,dim} built from lambda, mu
// (src/TopOptProblems/matrices_and_vectors.jl:83-87).
void element_stiffness_bits(int dim, double E, double nu, const double* h,
                            std::vector<double>& Ke) {
    const int nc = 1 << dim, nd = dim * nc;
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = E / (2 * (1 + nu));
    Ke.assign(nd * nd, 0.0);
    auto s = [](int a) { return a ? 1.0 : -1.0; };
    auto I = [&](int i, int k, int c, int cp) {
        double v = 1.0;
        for (int d = 0; d < dim; ++d) {
            const int a = (c >> d) & 1, b = (cp >> d) & 1;
            const double hd = h[d];
            double f;
            if (d == i && d == k)
                f = s(a) * s(b) / hd;
            else if (d == i)
                f = s(a) / 2.0;  // int phi_a' phi_b
            else if (d == k)
                f = s(b) / 2.0;  // int phi_a phi_b'
            else
                f = hd / 6.0 * (a == b ? 2.0 : 1.0);
            v *= f;
        }
        return v;
    };
    for (int c = 0; c < nc; ++c)
        for (int cp = 0; cp < nc; ++cp) {
            double tr = 0;
            for (int j = 0; j < dim; ++j) tr += I(j, j, c, cp);
            for (int i = 0; i < dim; ++i)
                for (int k = 0; k < dim; ++k)
                    Ke[(dim * c + i) * nd + dim * cp + k] =
                        lam * I(i, k, c, cp) + mu * ((i == k ? tr : 0.0) + I(k, i, c, cp));
        }
}

// B = H^T Ke H / nc^2 on the label blocks; returns the largest dropped
// (off-block) entry relative to the largest entry.
double build_elem_op(int dim, const std::vector<double>& Ke, ElemOp<double>& op) {
    const int nc = 1 << dim, nd = dim * nc;
    auto H = [&](int c, int m) {
        double v = 1.0;
        for (int a = 0; a < dim; ++a)
            if ((m >> a) & 1) v *= ((c >> a) & 1) ? 1.0 : -1.0;
        return v;
    };
    // M[(m,i),(m',k)] = sum_{c,c'} H[c][m] Ke[(c,i),(c',k)] H[c'][m']
    std::vector<double> M(nd * nd, 0.0);
    for (int m = 0; m < nc; ++m)
        for (int i = 0; i < dim; ++i)
            for (int mp = 0; mp < nc; ++mp)
                for (int k = 0; k < dim; ++k) {
                    double s = 0;
                    for (int c = 0; c < nc; ++c)
                        for (int cp = 0; cp < nc; ++cp)
                            s += H(c, m) * Ke[(dim * c + i) * nd + dim * cp + k] * H(cp, mp);
                    M[(dim * m + i) * nd + dim * mp + k] = s / double(nc * nc);
                }
    std::memset(&op, 0, sizeof(op));
    double maxall = 0, maxoff = 0;
    for (int a = 0; a < nd; ++a)
        for (int b = 0; b < nd; ++b) {
            const int ma = a / dim, ia = a % dim, mb = b / dim, ib = b % dim;
            const double v = M[a * nd + b];
            maxall = std::max(maxall, std::fabs(v));
            if ((ma ^ (1 << ia)) != (mb ^ (1 << ib))) maxoff = std::max(maxoff, std::fabs(v));
        }
    for (int L = 0; L < nc; ++L)
        for (int d = 0; d < dim; ++d)
            for (int d2 = 0; d2 < dim; ++d2) {
                const int ma = L ^ (1 << d), mb = L ^ (1 << d2);
                op.blk[L][d][d2] = M[(dim * ma + d) * nd + dim * mb + d2];
            }
    for (int d = 0; d < dim; ++d) op.kdiag[d] = Ke[d * nd + d];
    return maxall > 0 ? maxoff / maxall : 0.0;
}

template <typename T>
int dalloc(tohip_ctx* ctx, T** p, size_t n) {
    if (n == 0) n = 1;
    CK(hipMalloc((void**)p, n * sizeof(T)));
    return 0;
}

// caller order <-> internal order
void to_internal(const tohip_ctx* ctx, const double* in, std::vector<double>& out) {
    const int64_t n = ctx->g.ndof;
    out.resize(n);
    if (ctx->perm.empty())
        std::memcpy(out.data(), in, n * sizeof(double));
    else
        for (int64_t i = 0; i < n; ++i) out[i] = in[ctx->perm[i]];
}
void to_caller(const tohip_ctx* ctx, const std::vector<double>& in, double* out) {
    const int64_t n = ctx->g.ndof;
    if (ctx->perm.empty())
        std::memcpy(out, in.data(), n * sizeof(double));
    else
        for (int64_t i = 0; i < n; ++i) out[ctx->perm[i]] = in[i];
}

int upload_dofs(tohip_ctx* ctx, const double* host, double* dev) {
    std::vector<double> tmp;
    to_internal(ctx, host, tmp);
    CK(hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}
int download_dofs(tohip_ctx* ctx, const double* dev, double* host) {
    std::vector<double> tmp(ctx->g.ndof);
    CK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    to_caller(ctx, tmp, host);
    return 0;
}
int upload_cells(tohip_ctx* ctx, const double* host, double* dev) {
    CK(hipMemcpyAsync(dev, host, ctx->g.ncells * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    return 0;
}
int download_cells(tohip_ctx* ctx, const double* dev, double* host) {
    CK(hipMemcpyAsync(host, dev, ctx->g.ncells * sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// E and Jacobi diagonals from ctx->xt.
int refresh_operator(tohip_ctx* ctx) {
    CK(launch_density(ctx->g.ncells, ctx->xt, ctx->xmin, ctx->p, ctx->pbi, ctx->E64, ctx->E32,
                      ctx->stream));
    CK(launch_jacobi_inv<double>(ctx->g, ctx->op64, ctx->E64, ctx->fixmask, ctx->meandiag,
                                 ctx->dinv64, ctx->stream));
    CK(launch_jacobi_inv<float>(ctx->g, ctx->op64, ctx->E64, ctx->fixmask, ctx->meandiag,
                                ctx->dinv32, ctx->stream));
    return 0;
}

// Capture kGraphIters PCG iterations (direction+matvec, update) into a graph.
template <typename T>
int capture_pcg_graph(tohip_ctx* ctx, hipGraphExec_t* exec, const ElemOp<T>& op, T* x, T* r,
                      T* dinv, T** p, T* q, const T* E, PcgScalars* sc) {
    if (*exec) return 0;
    hipGraph_t graph;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kGraphIters; ++it) {
        T* pold = p[it & 1];
        T* pnew = p[(it + 1) & 1];
        hipError_t e1 = launch_pcg_dirmatvec<T>(ctx->g, op, r, dinv, pold, pnew, q, ctx->fixmask,
                                                E, ctx->meandiag, sc, ctx->partials, ctx->stream);
        hipError_t e2 = launch_pcg_update<T>(ctx->g.ndof, x, r, pnew, q, dinv, sc, ctx->partials,
                                             ctx->stream);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            hipStreamEndCapture(ctx->stream, &graph);
            ctx->err = "PCG graph capture failed";
            return TOHIP_ERR_HIP;
        }
    }
    CK(hipStreamEndCapture(ctx->stream, &graph));
    CK(hipGraphInstantiate(exec, graph, nullptr, nullptr, 0));
    CK(hipGraphDestroy(graph));
    return 0;
}

// Device-resident Jacobi PCG on K(E) x = b.  x is the initial guess when
// !initially_zero.  Returns 0, or TOHIP_NOT_CONVERGED.
template <typename T>
int pcg_device(tohip_ctx* ctx, const ElemOp<T>& op, const T* E, const T* b, T* x, T* r, T* dinv,
               T** p, T* q, T* tmp, PcgScalars* sc, hipGraphExec_t* exec, double abstol,
               double reltol, int64_t maxiter, bool initially_zero, int64_t* iters,
               double* resid) {
    const int64_t n = ctx->g.ndof;
    int rc = capture_pcg_graph<T>(ctx, exec, op, x, r, dinv, p, q, E, sc);
    if (rc) return rc;
    PcgScalars hs{};
    hs.abstol = abstol;
    hs.reltol = reltol;
    hs.maxiter = maxiter;
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(sc, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice, ctx->stream));
    if (initially_zero) {
        CK(hipMemsetAsync(x, 0, n * sizeof(T), ctx->stream));
        CK(launch_pcg_start<T>(n, b, (const T*)nullptr, r, dinv, p[0], sc, ctx->partials,
                               ctx->stream));
    } else {
        CK(launch_matvec<T>(ctx->g, op, x, ctx->fixmask, E, tmp, ctx->meandiag, ctx->stream));
        CK(launch_pcg_start<T>(n, b, tmp, r, dinv, p[0], sc, ctx->partials, ctx->stream));
    }
    for (;;) {
        CK(hipMemcpyAsync(ctx->h_sc, sc, sizeof(PcgScalars), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        if (ctx->h_sc->done) break;
        CK(hipGraphLaunch(*exec, ctx->stream));
    }
    if (iters) *iters = ctx->h_sc->iter;
    if (resid) *resid = std::sqrt(ctx->h_sc->rr);
    const bool ok = std::sqrt(ctx->h_sc->rr) <= ctx->h_sc->tol;
    return ok ? 0 : TOHIP_NOT_CONVERGED;
}

int solve_fp64(tohip_ctx* ctx, double abstol, double reltol, int64_t maxiter, bool zero,
               int64_t* iters, double* resid) {
    return pcg_device<double>(ctx, ctx->op64, ctx->E64, ctx->b64, ctx->u, ctx->r64, ctx->dinv64,
                              ctx->p64, ctx->q64, ctx->t64, ctx->sc64, &ctx->g64, abstol, reltol,
                              maxiter, zero, iters, resid);
}

// FP64 iterative refinement around an FP32 Jacobi PCG.  Outer rule:
// ||b - K u|| <= max(reltol*||b - K u0||, abstol).  The inner system is
// solved for r/||r|| to relative tolerance inner_rel.
int solve_mixed(tohip_ctx* ctx, double abstol, double reltol, int64_t maxiter, bool zero,
                double inner_rel, int64_t* iters, double* resid) {
    const int64_t n = ctx->g.ndof;
    if (zero) CK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
    int64_t total = 0;
    double tol = -1, res = 0;
    const int max_outer = 60;
    for (int outer = 0; outer <= max_outer; ++outer) {
        CK(launch_matvec<double>(ctx->g, ctx->op64, ctx->u, ctx->fixmask, ctx->E64, ctx->t64,
                                 ctx->meandiag, ctx->stream));
        CK(launch_residual64(n, ctx->b64, ctx->t64, ctx->r64, ctx->partials, ctx->dscal + 1,
                             ctx->counters + 1, ctx->stream));
        CK(hipMemcpyAsync(ctx->h_scal + 1, ctx->dscal + 1, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        res = std::sqrt(ctx->h_scal[1]);
        if (tol < 0) tol = std::max(reltol * res, abstol);
        if (res <= tol || total >= maxiter || outer == max_outer) break;
        CK(launch_scale_to_f32(n, ctx->r64, ctx->dscal + 1, ctx->b32, ctx->stream));
        int64_t it = 0;
        double ires = 0;
        int rc = pcg_device<float>(ctx, ctx->op32, ctx->E32, ctx->b32, ctx->x32, ctx->r32,
                                   ctx->dinv32, ctx->p32, ctx->q32, (float*)nullptr, ctx->sc32,
                                   &ctx->g32, 0.0, inner_rel, maxiter - total, true, &it, &ires);
        if (rc < 0) return rc;
        total += it;
        CK(launch_add_scaled_f32(n, ctx->x32, ctx->dscal + 1, ctx->u, ctx->stream));
    }
    if (iters) *iters = total;
    if (resid) *resid = res;
    return res <= tol ? 0 : TOHIP_NOT_CONVERGED;
}

// Density filter stencil: every node (i+a, j+b, k+c) of the cell (i, j, k)
// neighbourhood with |x_node - x_centre| < rmin, weight rmin - distance.
int build_filter(tohip_ctx* ctx, double rmin) {
    if (!(rmin > 0) || !std::isfinite(rmin)) {
        ctx->err = "DensityFilterFun: rmin must be positive and finite";
        return TOHIP_ERR_ARG;
    }
    if (ctx->f_offs && ctx->f_rmin == rmin) return 0;
    const int dim = ctx->g.dim;
    double hmin = std::min(ctx->h[0], ctx->h[1]);
    if (dim == 3) hmin = std::min(hmin, ctx->h[2]);
    const int R = (int)std::ceil(rmin / hmin) + 1;
    std::vector<int> offs;
    std::vector<double> wts;
    const int cz0 = dim == 3 ? -R : 0, cz1 = dim == 3 ? R + 1 : 0;
    for (int c = cz0; c <= cz1; ++c)
        for (int b = -R; b <= R + 1; ++b)
            for (int a = -R; a <= R + 1; ++a) {
                const double dx = (a - 0.5) * ctx->h[0], dy = (b - 0.5) * ctx->h[1],
                             dz = dim == 3 ? (c - 0.5) * ctx->h[2] : 0.0;
                const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (dist < rmin) {
                    offs.push_back(a);
                    offs.push_back(b);
                    offs.push_back(c);
                    wts.push_back(std::max(rmin - dist, 0.0));
                }
            }
    const int count = (int)wts.size();
    if (count == 0) {
        ctx->err =
            "DensityFilterFun: no neighbouring nodes were found within the filter radius "
            "`rmin` for any element (rmin is smaller than the element size)";
        return TOHIP_ERR_ARG;
    }
    if (count > kMaxStencil) {
        ctx->err = "DensityFilterFun: rmin too large for the stencil table";
        return TOHIP_ERR_ARG;
    }
    if (ctx->f_offs) {
        CK(hipFree(ctx->f_offs));
        CK(hipFree(ctx->f_wts));
        ctx->f_offs = nullptr;
        ctx->f_wts = nullptr;
    }
    if (int rc = dalloc(ctx, &ctx->f_offs, offs.size())) return rc;
    if (int rc = dalloc(ctx, &ctx->f_wts, wts.size())) return rc;
    CK(hipMemcpyAsync(ctx->f_offs, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipMemcpyAsync(ctx->f_wts, wts.data(), wts.size() * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    ctx->f_count = count;
    CK(launch_filter_invden(ctx->g, ctx->f_offs, ctx->f_wts, count, ctx->f_invden, ctx->stream));
    // colsum = J^T 1 (the volume-constraint weights of the OC step)
    std::vector<double> ones(ctx->g.ncells, 1.0);
    CK(hipMemcpyAsync(ctx->cellbuf, ones.data(), ones.size() * sizeof(double),
                      hipMemcpyHostToDevice, ctx->stream));
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, count, ctx->f_invden, ctx->cellbuf,
                     ctx->f_colsum, 1, nullptr, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->f_rmin = rmin;
    return 0;
}

double host_norm(const double* v, int64_t n) {
    long double s = 0;
    for (int64_t i = 0; i < n; ++i) s += (long double)v[i] * v[i];
    return (double)std::sqrt(s);
}

void free_ctx(tohip_ctx* ctx) {
    if (!ctx) return;
    if (ctx->g64) hipGraphExecDestroy(ctx->g64);
    if (ctx->g32) hipGraphExecDestroy(ctx->g32);
    void* bufs[] = {ctx->fixmask, ctx->xt,       ctx->E64,      ctx->E32,    ctx->f,
                    ctx->u,       ctx->b64,      ctx->r64,      ctx->t64,    ctx->dinv64,
                    ctx->p64[0],  ctx->p64[1],   ctx->q64,      ctx->b32,    ctx->x32,
                    ctx->r32,     ctx->dinv32,   ctx->p32[0],   ctx->p32[1], ctx->q32,
                    ctx->sc64,    ctx->sc32,     ctx->partials, ctx->dscal,  ctx->counters,
                    ctx->f_offs,  ctx->f_wts,    ctx->f_invden, ctx->f_colsum, ctx->xd,
                    ctx->xnew,    ctx->dc,       ctx->ce,       ctx->grad,   ctx->cellbuf,
                    ctx->oc};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (ctx->h_sc) hipHostFree(ctx->h_sc);
    if (ctx->h_scal) hipHostFree(ctx->h_scal);
    if (ctx->h_oc) hipHostFree(ctx->h_oc);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int create_impl(tohip_ctx* ctx, int dim, const int64_t* nels, const double* h, double E,
                double nu, double xmin, double pp, int pbi, const int64_t* node_dofs,
                const uint8_t* fixed_mask) {
    if (dim != 2 && dim != 3) {
        ctx->err = "dim must be 2 or 3";
        return TOHIP_ERR_ARG;
    }
    if (!nels || !h || !fixed_mask) {
        ctx->err = "nels, h and fixed_mask are required";
        return TOHIP_ERR_ARG;
    }
    for (int d = 0; d < dim; ++d) {
        if (nels[d] < 1 || nels[d] > (1 << 20)) {
            ctx->err = "cell counts must be in [1, 2^20]";
            return TOHIP_ERR_ARG;
        }
        if (!(h[d] > 0) || !std::isfinite(h[d])) {
            ctx->err = "cell sizes must be positive";
            return TOHIP_ERR_ARG;
        }
    }
    if (!(E > 0) || !(nu > -1 && nu < 0.5) || !(xmin >= 0 && xmin < 1) || !(pp > 0)) {
        ctx->err = "need E > 0, -1 < nu < 0.5, 0 <= xmin < 1, p > 0";
        return TOHIP_ERR_ARG;
    }
    Grid& g = ctx->g;
    g.dim = dim;
    g.nx = (int)nels[0];
    g.ny = (int)nels[1];
    g.nz = dim == 3 ? (int)nels[2] : 1;
    g.NX = g.nx + 1;
    g.NY = g.ny + 1;
    g.NZ = dim == 3 ? g.nz + 1 : 1;
    g.nnodes = (int64_t)g.NX * g.NY * g.NZ;
    g.ncells = (int64_t)g.nx * g.ny * g.nz;
    g.ndof = dim * g.nnodes;
    if (g.ncells > (int64_t)kMaxPartials * 256) {
        ctx->err = "grid too large (more than 16.7M cells)";
        return TOHIP_ERR_ARG;
    }
    ctx->E = E;
    ctx->nu = nu;
    ctx->xmin = xmin;
    ctx->p = pp;
    ctx->pbi = pbi ? 1 : 0;
    for (int d = 0; d < 3; ++d) ctx->h[d] = d < dim ? h[d] : 1.0;

    std::vector<double> Ke;
    element_stiffness_bits(dim, E, nu, ctx->h, Ke);
    const double off = build_elem_op(dim, Ke, ctx->op64);
    if (off > 1e-12) {
        ctx->err = "element operator is not block diagonal in the corner-Hadamard basis";
        return TOHIP_ERR_STATE;
    }
    for (int L = 0; L < 8; ++L)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) ctx->op32.blk[L][a][b] = (float)ctx->op64.blk[L][a][b];
    for (int d = 0; d < 3; ++d) ctx->op32.kdiag[d] = (float)ctx->op64.kdiag[d];
    // meandiag = sum over cells of tr(Ke) (src/FEA/solvers_api.jl:526-527)
    const int nd = dim << dim;
    double tr = 0;
    for (int i = 0; i < nd; ++i) tr += Ke[i * nd + i];
    double md = 0;
    for (int64_t e = 0; e < g.ncells; ++e) md += tr;
    ctx->meandiag = md;

    // DOF permutation from Metadata.node_dofs
    std::vector<uint8_t> nodemask(g.nnodes, 0);
    if (node_dofs) {
        ctx->perm.assign(g.ndof, -1);
        std::vector<uint8_t> seen(g.ndof, 0);
        for (int64_t n = 0; n < g.nnodes; ++n)
            for (int d = 0; d < dim; ++d) {
                const int64_t v = node_dofs[n * dim + d] - 1;
                if (v < 0 || v >= g.ndof || seen[v]) {
                    ctx->err = "node_dofs is not a permutation of 1..ndofs";
                    return TOHIP_ERR_ARG;
                }
                seen[v] = 1;
                ctx->perm[dim * n + d] = v;
            }
    }
    for (int64_t n = 0; n < g.nnodes; ++n)
        for (int d = 0; d < dim; ++d) {
            const int64_t c = node_dofs ? ctx->perm[dim * n + d] : dim * n + d;
            if (fixed_mask[c]) nodemask[n] |= (uint8_t)(1u << d);
        }

    CK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const size_t nD = g.ndof, nC = g.ncells;
    int rc = 0;
    if ((rc = dalloc(ctx, &ctx->fixmask, g.nnodes)) || (rc = dalloc(ctx, &ctx->xt, nC)) ||
        (rc = dalloc(ctx, &ctx->E64, nC)) || (rc = dalloc(ctx, &ctx->E32, nC)) ||
        (rc = dalloc(ctx, &ctx->f, nD)) || (rc = dalloc(ctx, &ctx->u, nD)) ||
        (rc = dalloc(ctx, &ctx->b64, nD)) || (rc = dalloc(ctx, &ctx->r64, nD)) ||
        (rc = dalloc(ctx, &ctx->t64, nD)) || (rc = dalloc(ctx, &ctx->dinv64, nD)) ||
        (rc = dalloc(ctx, &ctx->p64[0], nD)) || (rc = dalloc(ctx, &ctx->p64[1], nD)) ||
        (rc = dalloc(ctx, &ctx->q64, nD)) || (rc = dalloc(ctx, &ctx->b32, nD)) ||
        (rc = dalloc(ctx, &ctx->x32, nD)) || (rc = dalloc(ctx, &ctx->r32, nD)) ||
        (rc = dalloc(ctx, &ctx->dinv32, nD)) || (rc = dalloc(ctx, &ctx->p32[0], nD)) ||
        (rc = dalloc(ctx, &ctx->p32[1], nD)) || (rc = dalloc(ctx, &ctx->q32, nD)) ||
        (rc = dalloc(ctx, &ctx->sc64, 1)) || (rc = dalloc(ctx, &ctx->sc32, 1)) ||
        (rc = dalloc(ctx, &ctx->partials, 2 * (size_t)kMaxPartials)) ||
        (rc = dalloc(ctx, &ctx->dscal, 8)) || (rc = dalloc(ctx, &ctx->counters, 8)) ||
        (rc = dalloc(ctx, &ctx->f_invden, nC)) || (rc = dalloc(ctx, &ctx->f_colsum, nC)) ||
        (rc = dalloc(ctx, &ctx->xd, nC)) || (rc = dalloc(ctx, &ctx->xnew, nC)) ||
        (rc = dalloc(ctx, &ctx->dc, nC)) || (rc = dalloc(ctx, &ctx->ce, nC)) ||
        (rc = dalloc(ctx, &ctx->grad, nC)) || (rc = dalloc(ctx, &ctx->cellbuf, nC)) ||
        (rc = dalloc(ctx, &ctx->oc, 1)))
        return rc;
    CK(hipHostMalloc((void**)&ctx->h_sc, sizeof(PcgScalars), hipHostMallocDefault));
    CK(hipHostMalloc((void**)&ctx->h_scal, 8 * sizeof(double), hipHostMallocDefault));
    CK(hipHostMalloc((void**)&ctx->h_oc, sizeof(OcState), hipHostMallocDefault));
    CK(hipMemcpyAsync(ctx->fixmask, nodemask.data(), g.nnodes, hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipMemsetAsync(ctx->counters, 0, 8 * sizeof(unsigned int), ctx->stream));
    CK(hipMemsetAsync(ctx->sc64, 0, sizeof(PcgScalars), ctx->stream));
    CK(hipMemsetAsync(ctx->sc32, 0, sizeof(PcgScalars), ctx->stream));
    CK(hipMemsetAsync(ctx->oc, 0, sizeof(OcState), ctx->stream));
    CK(hipMemsetAsync(ctx->u, 0, nD * sizeof(double), ctx->stream));
    CK(hipMemsetAsync(ctx->f, 0, nD * sizeof(double), ctx->stream));
    // solver.vars = ones (src/FEA/solvers_api.jl:497-499)
    std::vector<double> ones(nC, 1.0);
    CK(hipMemcpyAsync(ctx->xt, ones.data(), nC * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    if ((rc = refresh_operator(ctx))) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace

extern "C" {

const char* tohip_version(void) { return "libtopopt_hip 0.1 (gfx950)"; }

int tohip_device_info(int* count, char* name, size_t name_len) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        if (count) *count = 0;
        g_create_error = "no HIP device";
        return TOHIP_ERR_HIP;
    }
    if (count) *count = n;
    if (name && name_len > 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return TOHIP_ERR_HIP;
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return 0;
}

const char* tohip_last_error(const tohip_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int tohip_sizes(const tohip_ctx* ctx, int64_t* ndofs, int64_t* ncells) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (ndofs) *ndofs = ctx->g.ndof;
    if (ncells) *ncells = ctx->g.ncells;
    return 0;
}

int tohip_create(tohip_ctx** out, int dim, const int64_t* nels, const double* h, double E,
                 double nu, double xmin, double penalty_p, int penalty_before_interp,
                 const int64_t* node_dofs, const uint8_t* fixed_mask) {
    if (!out) {
        g_create_error = "out is NULL";
        return TOHIP_ERR_ARG;
    }
    *out = nullptr;
    tohip_ctx* ctx = new (std::nothrow) tohip_ctx();
    if (!ctx) {
        g_create_error = "out of host memory";
        return TOHIP_ERR_HIP;
    }
    int rc = create_impl(ctx, dim, nels, h, E, nu, xmin, penalty_p, penalty_before_interp,
                         node_dofs, fixed_mask);
    if (rc != 0) {
        g_create_error = ctx->err;
        free_ctx(ctx);
        return rc;
    }
    *out = ctx;
    return 0;
}

void tohip_destroy(tohip_ctx* ctx) { free_ctx(ctx); }

int tohip_set_penalty(tohip_ctx* ctx, double penalty_p) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (!(penalty_p > 0)) {
        ctx->err = "penalty must be positive";
        return TOHIP_ERR_ARG;
    }
    ctx->p = penalty_p;
    if (int rc = refresh_operator(ctx)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_set_densities(tohip_ctx* ctx, const double* vars) {
    if (!ctx || !vars) return TOHIP_ERR_ARG;
    if (int rc = upload_cells(ctx, vars, ctx->xt)) return rc;
    if (int rc = refresh_operator(ctx)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_matvec(tohip_ctx* ctx, const double* x, double* y) {
    if (!ctx || !x || !y) return TOHIP_ERR_ARG;
    if (int rc = upload_dofs(ctx, x, ctx->r64)) return rc;
    CK(launch_matvec<double>(ctx->g, ctx->op64, ctx->r64, ctx->fixmask, ctx->E64, ctx->t64,
                             ctx->meandiag, ctx->stream));
    return download_dofs(ctx, ctx->t64, y);
}

int tohip_solve(tohip_ctx* ctx, const double* f, double* u, double abstol, double reltol,
                int64_t maxiter, int initially_zero, int precision, int64_t* iters_out,
                double* resid_out) {
    if (!ctx || !f || !u) return TOHIP_ERR_ARG;
    if (!(abstol >= 0) || !(reltol >= 0) || maxiter < 0 ||
        (precision != TOHIP_FP64 && precision != TOHIP_MIXED)) {
        ctx->err = "tohip_solve: bad tolerance, maxiter or precision";
        return TOHIP_ERR_ARG;
    }
    if (int rc = upload_dofs(ctx, f, ctx->b64)) return rc;
    if (!initially_zero)
        if (int rc = upload_dofs(ctx, u, ctx->u)) return rc;
    int rc = precision == TOHIP_FP64
                 ? solve_fp64(ctx, abstol, reltol, maxiter, initially_zero != 0, iters_out,
                              resid_out)
                 : solve_mixed(ctx, abstol, reltol, maxiter, initially_zero != 0, 1e-4,
                               iters_out, resid_out);
    if (rc < 0) return rc;
    ctx->have_u = true;
    if (int rc2 = download_dofs(ctx, ctx->u, u)) return rc2;
    return rc;
}

int tohip_compliance(tohip_ctx* ctx, const double* u, double* obj, double* cell_comp,
                     double* grad) {
    if (!ctx || !u) return TOHIP_ERR_ARG;
    if (int rc = upload_dofs(ctx, u, ctx->t64)) return rc;
    CK(launch_energy(ctx->g, ctx->op64, ctx->t64, ctx->xt, ctx->xmin, ctx->p, ctx->pbi, ctx->ce,
                     ctx->grad, ctx->partials, ctx->dscal, ctx->counters, ctx->stream));
    CK(hipMemcpyAsync(ctx->h_scal, ctx->dscal, sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (obj) *obj = ctx->h_scal[0];
    if (cell_comp)
        if (int rc = download_cells(ctx, ctx->ce, cell_comp)) return rc;
    if (grad)
        if (int rc = download_cells(ctx, ctx->grad, grad)) return rc;
    return 0;
}

int tohip_filter_setup(tohip_ctx* ctx, double rmin) {
    if (!ctx) return TOHIP_ERR_ARG;
    return build_filter(ctx, rmin);
}

int tohip_filter_apply(tohip_ctx* ctx, const double* x, double* y, int transpose) {
    if (!ctx || !x || !y) return TOHIP_ERR_ARG;
    if (!ctx->f_offs) {
        ctx->err = "tohip_filter_apply: call tohip_filter_setup first";
        return TOHIP_ERR_STATE;
    }
    if (int rc = upload_cells(ctx, x, ctx->cellbuf)) return rc;
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_invden, ctx->cellbuf,
                     ctx->xnew, transpose ? 1 : 0, nullptr, ctx->stream));
    return download_cells(ctx, ctx->xnew, y);
}

int tohip_volume(tohip_ctx* ctx, const double* x, double* vol) {
    if (!ctx || !x || !vol) return TOHIP_ERR_ARG;
    if (int rc = upload_cells(ctx, x, ctx->cellbuf)) return rc;
    CK(launch_sum(ctx->g.ncells, ctx->cellbuf, ctx->partials, ctx->dscal + 2, ctx->counters + 2,
                  ctx->stream));
    CK(hipMemcpyAsync(ctx->h_scal + 2, ctx->dscal + 2, sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    // uniform cells: V_e / V_total = 1 / ncells
    *vol = ctx->h_scal[2] / (double)ctx->g.ncells;
    return 0;
}

int tohip_oc_update(tohip_ctx* ctx, const double* x, const double* dc, double volfrac,
                    double move, double tol, double* xnew, double* lambda_out) {
    if (!ctx || !x || !dc || !xnew) return TOHIP_ERR_ARG;
    if (!(volfrac > 0 && volfrac <= 1) || !(move > 0) || !(tol > 0)) {
        ctx->err = "tohip_oc_update: need 0 < volfrac <= 1, move > 0, tol > 0";
        return TOHIP_ERR_ARG;
    }
    if (!ctx->f_offs) {
        ctx->err = "tohip_oc_update: call tohip_filter_setup first (volume is on J x)";
        return TOHIP_ERR_STATE;
    }
    if (int rc = upload_cells(ctx, x, ctx->xd)) return rc;
    if (int rc = upload_cells(ctx, dc, ctx->dc)) return rc;
    CK(launch_oc(ctx->g.ncells, ctx->xd, ctx->dc, ctx->f_colsum, volfrac, move, tol, 16,
                 ctx->xnew, ctx->oc, ctx->partials, ctx->stream));
    CK(hipMemcpyAsync(ctx->h_oc, ctx->oc, sizeof(OcState), hipMemcpyDeviceToHost, ctx->stream));
    if (int rc = download_cells(ctx, ctx->xnew, xnew)) return rc;
    if (lambda_out) *lambda_out = ctx->h_oc->lambda;
    return 0;
}

int tohip_set_load(tohip_ctx* ctx, const double* f) {
    if (!ctx || !f) return TOHIP_ERR_ARG;
    return upload_dofs(ctx, f, ctx->f);
}

int tohip_get_u(tohip_ctx* ctx, double* u) {
    if (!ctx || !u) return TOHIP_ERR_ARG;
    return download_dofs(ctx, ctx->u, u);
}

int tohip_simp_run(tohip_ctx* ctx, const tohip_simp_opts* o, double* x_inout, double* obj_hist,
                   int64_t* iters_hist, double* change_hist) {
    if (!ctx || !o || !x_inout) return TOHIP_ERR_ARG;
    if (o->iterations < 0 || !(o->volfrac > 0 && o->volfrac <= 1) || !(o->move > 0) ||
        !(o->oc_tol > 0) || !(o->rel_tol > 0) || o->maxiter < 1 ||
        (o->precision != TOHIP_FP64 && o->precision != TOHIP_MIXED) || o->warm_start < 0 ||
        o->warm_start > 1 || (o->precision == TOHIP_MIXED && !(o->inner_rel_tol > 0))) {
        ctx->err = "tohip_simp_run: bad options";
        return TOHIP_ERR_ARG;
    }
    if (int rc = build_filter(ctx, o->rmin)) return rc;
    const int64_t nD = ctx->g.ndof, nC = ctx->g.ncells;
    std::vector<double> fh(nD);
    CK(hipMemcpyAsync(fh.data(), ctx->f, nD * sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    const double fnorm = host_norm(fh.data(), nD);
    if (!(fnorm > 0)) {
        ctx->err = "tohip_simp_run: the load vector is zero (tohip_set_load)";
        return TOHIP_ERR_STATE;
    }
    const double abstol = o->rel_tol * fnorm;
    CK(hipMemcpyAsync(ctx->b64, ctx->f, nD * sizeof(double), hipMemcpyDeviceToDevice,
                      ctx->stream));
    if (int rc = upload_cells(ctx, x_inout, ctx->xd)) return rc;
    int final_rc = 0;
    for (int64_t it = 0; it < o->iterations; ++it) {
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_invden, ctx->xd,
                         ctx->xt, 0, nullptr, ctx->stream));
        if (int rc = refresh_operator(ctx)) return rc;
        const bool zero = o->warm_start == 0 || !ctx->have_u;
        int64_t its = 0;
        double res = 0;
        int rc = o->precision == TOHIP_FP64
                     ? solve_fp64(ctx, abstol, 0.0, o->maxiter, zero, &its, &res)
                     : solve_mixed(ctx, abstol, 0.0, o->maxiter, zero, o->inner_rel_tol, &its,
                                   &res);
        if (rc < 0) return rc;
        if (rc > 0) final_rc = rc;
        ctx->have_u = true;
        CK(launch_energy(ctx->g, ctx->op64, ctx->u, ctx->xt, ctx->xmin, ctx->p, ctx->pbi,
                         ctx->ce, ctx->grad, ctx->partials, ctx->dscal, ctx->counters,
                         ctx->stream));
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_invden, ctx->grad,
                         ctx->dc, 1, nullptr, ctx->stream));
        CK(launch_oc(nC, ctx->xd, ctx->dc, ctx->f_colsum, o->volfrac, o->move, o->oc_tol, 16,
                     ctx->xnew, ctx->oc, ctx->partials, ctx->stream));
        std::swap(ctx->xd, ctx->xnew);
        CK(hipMemcpyAsync(ctx->h_scal, ctx->dscal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipMemcpyAsync(ctx->h_oc, ctx->oc, sizeof(OcState), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        if (obj_hist) obj_hist[it] = ctx->h_scal[0];
        if (iters_hist) iters_hist[it] = its;
        if (change_hist) change_hist[it] = ctx->h_oc->change;
    }
    if (int rc = download_cells(ctx, ctx->xd, x_inout)) return rc;
    return final_rc;
}

}  // extern "C"

