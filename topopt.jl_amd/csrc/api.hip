// C-ABI of libtopopt_hip: context, element operator, solves, SIMP loop.
// See include/topopt_hip.h for the contract of every entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/topopt_hip.h"
#include "ctx_internal.h"
#include "tohip_internal.h"

using namespace tohip;

namespace {
thread_local std::string g_create_error;
constexpr int kGraphItersBig = 32;   // PCG iterations per long captured graph
constexpr int kGraphItersSmall = 4;  // ... per short graph (near convergence)
constexpr int kGraphItersMg = 8;     // multigrid-preconditioned PCG: iterations per long graph
constexpr int kGraphItersMgSmall = 2;
// Coarsest multigrid level: Chebyshev iterations and the assumed condition
// number of D^-1 K there (any value keeps the preconditioner SPD).  The V-cycle
// applies the polynomial as a dense matrix, so the degree only prices the
// matrix's rebuild (~7 us per iteration at 15x5x5, once per SIMP iteration).
// On MI355X at 120x40x40 (50 SIMP iterations of the mixed multigrid path):
// degree 16 / kappa 100 took 1756 PCG iterations at 265.7 SIMP it/s, 24 / 300
// 1594 at 281.0, 32 / 1000 1468 at 298.0, 48 / 1000 1396 at 298.1, 64 / 1000
// 1388 at 289.9; kappa 3000 and above took more iterations at every degree.
constexpr int kMgCoarseIters = 32;
constexpr double kMgCoarseKappa = 1000.0;
// Damped Jacobi sweeps of a coarsest level too large for that kernel.
constexpr int kMgFallbackSweeps = 30;
// Mixed precision: the FP64 residual replaces the FP32 recursive one each
// time the latter has dropped by this factor (reliable updates).
constexpr double kDefaultReplaceRatio = 0.1;
// With the multigrid preconditioner a residual drop of 1000x takes ~15
// iterations: fewer FP64 refreshes (each a host round trip) pay off.
constexpr double kDefaultReplaceRatioMg = 1e-3;
// ... and the search direction restarts (p = z) when a replacement finds the
// true residual more than this many times the norm of the recursive one.  On
// 64x32 box cells (1 x 0.5, random design) the first replacement found it
// 3-10 times larger; keeping p took 558 multigrid PCG iterations against 105
// for the FP64 PCG with the same V-cycle (117 with the restart).  Thresholds
// 2 to 5 gave the same counts there, and none of them restarts in the SIMP
// runs measured (120x40x40 and 128x64 cantilevers), where the ratio stays
// below 2 and restarting on any mismatch cost 6-13 % more iterations.
constexpr double kRestartRatio = 3.0;
}  // namespace



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
    // B[(L,d),(L,d')] of label L between components d and d'
    auto Bl = [&](int L, int d, int d2) {
        return M[(dim * (L ^ (1 << d)) + d) * nd + dim * (L ^ (1 << d2)) + d2];
    };
    // entries the compact form keeps (everything else must vanish)
    std::vector<char> kept(nd * nd, 0);
    auto keep = [&](int L, int d, int d2) {
        kept[(dim * (L ^ (1 << d)) + d) * nd + dim * (L ^ (1 << d2)) + d2] = 1;
        kept[(dim * (L ^ (1 << d2)) + d2) * nd + dim * (L ^ (1 << d)) + d] = 1;
        return Bl(L, d, d2);
    };
    if (dim == 3) {
        const int full[2] = {0, 7};
        for (int f = 0; f < 2; ++f) {
            double* k = f == 0 ? op.l0 : op.l7;
            const int L = full[f];
            k[0] = keep(L, 0, 0); k[1] = keep(L, 0, 1); k[2] = keep(L, 0, 2);
            k[3] = keep(L, 1, 1); k[4] = keep(L, 1, 2); k[5] = keep(L, 2, 2);
        }
        // labels 1, 2, 4: rigid member d = log2(L); pair = the other two
        const int lp[3] = {1, 2, 4}, pa[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        for (int t = 0; t < 3; ++t) {
            const int L = lp[t], a = pa[t][0], b = pa[t][1];
            op.pr[t][0] = keep(L, a, a); op.pr[t][1] = keep(L, a, b); op.pr[t][2] = keep(L, b, b);
        }
        // labels 3, 5, 6: pair (a, b) + single c
        const int lq[3] = {3, 5, 6}, qa[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
        for (int t = 0; t < 3; ++t) {
            const int L = lq[t], a = qa[t][0], b = qa[t][1], c = qa[t][2];
            op.qs[t][0] = keep(L, a, a); op.qs[t][1] = keep(L, a, b);
            op.qs[t][2] = keep(L, b, b); op.qs[t][3] = keep(L, c, c);
        }
    } else {
        op.q2[0] = keep(0, 0, 0); op.q2[1] = keep(0, 0, 1); op.q2[2] = keep(0, 1, 1);
        op.q2[3] = keep(3, 0, 0); op.q2[4] = keep(3, 0, 1); op.q2[5] = keep(3, 1, 1);
        op.q2[6] = keep(1, 1, 1);
        op.q2[7] = keep(2, 0, 0);
    }
    double maxall = 0, maxoff = 0;
    for (int a = 0; a < nd * nd; ++a) {
        maxall = std::max(maxall, std::fabs(M[a]));
        if (!kept[a]) maxoff = std::max(maxoff, std::fabs(M[a]));
    }
    for (int d = 0; d < dim; ++d) op.kdiag[d] = Ke[d * nd + d];
    return maxall > 0 ? maxoff / maxall : 0.0;
}

// Cube form (ElemOp::cube) when the compact blocks show the cube symmetry.
void detect_cube(ElemOp<double>& op, int dim) {
    op.is_cube = 0;
    if (dim != 3) return;
    const double A = op.l0[0], B = op.l0[1], C = op.pr[0][0], D = op.pr[0][1];
    const double E = op.qs[0][0], F = op.qs[0][3], G = op.l7[0], H = op.l7[1];
    const double scale = std::max({std::fabs(A), std::fabs(G), std::fabs(C), 1e-300});
    auto eq = [&](double a, double b) { return std::fabs(a - b) <= 1e-13 * scale; };
    bool ok = eq(op.l0[3], A) && eq(op.l0[5], A) && eq(op.l0[2], B) && eq(op.l0[4], B) &&
              eq(op.l7[3], G) && eq(op.l7[5], G) && eq(op.l7[2], H) && eq(op.l7[4], H);
    for (int t = 0; t < 3; ++t) {
        ok = ok && eq(op.pr[t][0], C) && eq(op.pr[t][1], D) && eq(op.pr[t][2], C);
        ok = ok && eq(op.qs[t][0], E) && eq(op.qs[t][1], E) && eq(op.qs[t][2], E) &&
             eq(op.qs[t][3], F);
    }
    if (!ok) return;
    const double c[8] = {A - B, B, C, D, E, F, G - H, H};
    for (int i = 0; i < 8; ++i) op.cube[i] = c[i];
    op.is_cube = 1;
}

template <typename T>
ElemOp<T> cast_elem_op(const ElemOp<double>& a) {
    ElemOp<T> b;
    for (int i = 0; i < 6; ++i) {
        b.l0[i] = (T)a.l0[i];
        b.l7[i] = (T)a.l7[i];
    }
    for (int t = 0; t < 3; ++t) {
        for (int i = 0; i < 3; ++i) b.pr[t][i] = (T)a.pr[t][i];
        for (int i = 0; i < 4; ++i) b.qs[t][i] = (T)a.qs[t][i];
    }
    for (int i = 0; i < 8; ++i) {
        b.q2[i] = (T)a.q2[i];
        b.cube[i] = (T)a.cube[i];
    }
    for (int i = 0; i < 3; ++i) b.kdiag[i] = (T)a.kdiag[i];
    b.is_cube = a.is_cube;
    return b;
}

template <typename T>
int dalloc(tohip_ctx* ctx, T** p, size_t n) {
    if (n == 0) n = 1;
    CK(hipMalloc((void**)p, n * sizeof(T)));
    ctx->dev_bytes += n * sizeof(T);
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

// Coarse-level moduli and Jacobi diagonals of the multigrid hierarchy, from
// the finest level's E64 (after every change of the densities or penalty).
// Smoother weights from lambda_max(D^-1 K) of every smoothed level, on the
// device between solves.  Power steps on I - D^-1 K (whose dominant
// eigenvalue is 1 - lambda_max, below -1 for these operators) from the kept
// iterate: 12 at the first refresh, then 2 per operator refresh, warm
// started (a SIMP iteration changes the design a little).  The Rayleigh
// quotient theta <= lambda_max gives w = 1.9 / (1.2 theta), kept within
// [the design-independent bound's weight, 0.8], once theta has levelled off
// over the last two steps (else the bound's weight): the smoother stays
// convergent (w lambda_max < 2) while theta is within 20 % of lambda_max.
constexpr int kMgPowerFirst = 12, kMgPowerSteps = 2;  // even (two-buffer cycle)
constexpr float kMgWMargin = 1.2f, kMgWMax = 0.8f;
int mg_estimate_weights(tohip_ctx* ctx) {
    std::vector<tohip_mg_level>& M = ctx->mg;
    const int L = (int)M.size() - 1;
    const int steps = ctx->mg_wsteps == 0 ? kMgPowerFirst : kMgPowerSteps;
    // TOHIP_MG_WMARGIN overrides the margin (measurement, read per refresh, >= 1)
    const char* em = getenv("TOHIP_MG_WMARGIN");
    const float margin = em && atof(em) >= 1.0 ? (float)atof(em) : kMgWMargin;
    const float* one = ctx->mg_wd + 4 * (L + 1);
    for (int l = 0; l < L; ++l) {
        tohip_mg_level& V = M[l];
        float* buf[2] = {V.pw, V.rz};
        // partials of the last step into partB, of the one before into partA
        for (int k = 0; k < steps; ++k)
            CK(launch_mg_smooth(5, V.view(), 1.f, nullptr, buf[k % 2], nullptr, nullptr,
                                buf[(k + 1) % 2], nullptr, ctx->meandiag, nullptr,
                                k == steps - 1 ? ctx->partB : ctx->partA, ctx->stream, nullptr,
                                k == 0 ? V.wd + 1 : one));
        CK(launch_mg_omega(ctx->partA, ctx->partB, mg_power_blocks(V.g), V.wd, V.omega, kMgWMax,
                           margin, ctx->stream));
    }
    ctx->mg_wsteps += steps;
    return 0;
}

int mg_refresh(tohip_ctx* ctx) {
    for (size_t l = 1; l < ctx->mg.size(); ++l) {
        tohip_mg_level& F = ctx->mg[l - 1];
        tohip_mg_level& C = ctx->mg[l];
        CK(launch_mg_coarsen(F.g, C.g, F.E64, C.E64, C.E32, ctx->stream));
        CK(launch_jacobi_inv<float>(C.g, C.op64, C.E64, C.fixmask, ctx->meandiag, C.dinv,
                                    ctx->stream));
    }
    if (ctx->mg_M) {
        // the coarsest operator changed: rebuild its dense form on the second
        // stream, after the moduli and diagonals above (TOHIP_MG_OVERLAP=0, read
        // per call: on the solver stream, for comparisons)
        const char* eo = getenv("TOHIP_MG_OVERLAP");
        if (eo && atoi(eo) == 0) {
            CK(launch_mg_coarse_basis(ctx->mg.back().view(), ctx->mg_cheb, ctx->mg_M,
                                      ctx->stream));
        } else {
            if (!ctx->mg_stream) CK(hipStreamCreateWithFlags(&ctx->mg_stream, hipStreamNonBlocking));
            for (hipEvent_t& e : ctx->mg_ev)
                if (!e) CK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            CK(hipEventRecord(ctx->mg_ev[0], ctx->stream));
            CK(hipStreamWaitEvent(ctx->mg_stream, ctx->mg_ev[0], 0));
            CK(launch_mg_coarse_basis(ctx->mg.back().view(), ctx->mg_cheb, ctx->mg_M,
                                      ctx->mg_stream));
            CK(hipEventRecord(ctx->mg_ev[1], ctx->mg_stream));
            ctx->mg_pending = true;
        }
    }
    if (ctx->mg_wadapt && !ctx->mg.empty()) return mg_estimate_weights(ctx);
    return 0;
}

// Called before work that applies a V-cycle (outside graph capture): the
// solver stream waits for a pending rebuild of the coarsest dense operator.
int mg_join(tohip_ctx* ctx) {
    if (!ctx->mg_pending) return 0;
    CK(hipStreamWaitEvent(ctx->stream, ctx->mg_ev[1], 0));
    ctx->mg_pending = false;
    return 0;
}

// E and Jacobi diagonals from ctx->xt.
int refresh_operator(tohip_ctx* ctx) {
    CK(launch_density(ctx->g.ncells, ctx->xt, ctx->xmin, ctx->pen, ctx->pbi, ctx->E64, ctx->E32,
                      ctx->stream));
    CK(launch_jacobi_inv<double>(ctx->g, ctx->op64, ctx->E64, ctx->fixmask, ctx->meandiag,
                                 ctx->dinv64, ctx->stream));
    CK(launch_jacobi_inv<float>(ctx->g, ctx->op64, ctx->E64, ctx->fixmask, ctx->meandiag,
                                ctx->dinv32, ctx->stream));
    return mg_refresh(ctx);
}

// ---- multigrid hierarchy (preconditioner of TOHIP_MIXED_MG) ----
Grid halve_grid(const Grid& f) {
    Grid c = f;
    c.nx = f.nx / 2;
    c.ny = f.ny / 2;
    c.nz = f.dim == 3 ? f.nz / 2 : 1;
    c.NX = c.nx + 1;
    c.NY = c.ny + 1;
    c.NZ = f.dim == 3 ? c.nz + 1 : 1;
    c.nnodes = (int64_t)c.NX * c.NY * c.NZ;
    c.ncells = (int64_t)c.nx * c.ny * c.nz;
    c.ndof = f.dim * c.nnodes;
    return c;
}

// Coarse DOF (N, d) is prescribed when any fine node of its prolongation
// stencil (2I + a, 2J + b, 2K + c), a, b, c in {-1, 0, 1}, has d prescribed.
std::vector<uint8_t> coarse_mask(const Grid& f, const Grid& c, const std::vector<uint8_t>& fm) {
    std::vector<uint8_t> cm(c.nnodes, 0);
    for (int K = 0; K < c.NZ; ++K)
        for (int J = 0; J < c.NY; ++J)
            for (int I = 0; I < c.NX; ++I) {
                uint8_t m = 0;
                for (int dk = (f.dim == 3 ? -1 : 0); dk <= (f.dim == 3 ? 1 : 0); ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                            if (i < 0 || i >= f.NX || j < 0 || j >= f.NY || k < 0 || k >= f.NZ)
                                continue;
                            m |= fm[i + (int64_t)f.NX * (j + (int64_t)f.NY * k)];
                        }
                cm[I + (int64_t)c.NX * (J + (int64_t)c.NY * K)] = m;
            }
    return cm;
}

// Frees the buffers a hierarchy owns (level 0 borrows the context's operator).
void free_mg_levels(std::vector<tohip_mg_level>& lv) {
    for (size_t l = 0; l < lv.size(); ++l) {
        tohip_mg_level& L = lv[l];
        for (void* b : {(void*)L.x, (void*)L.rz, (void*)L.b, (void*)L.pw})
            if (b) hipFree(b);
        if (l > 0)
            for (void* b : {(void*)L.E64, (void*)L.E32, (void*)L.fixmask, (void*)L.dinv})
                if (b) hipFree(b);
    }
    lv.clear();
}

int mg_build_levels(tohip_ctx* ctx, const std::vector<Grid>& gs, std::vector<tohip_mg_level>& lv);

// Largest eigenvalue of diag(Ke)^-1/2 Ke diag(Ke)^-1/2 (power iteration on a
// PSD 24 x 24 / 8 x 8 matrix), plus 1 % margin.  Since
// x^T K x = sum_e E_e x_e^T Ke x_e <= lambda x^T D x for ANY cell moduli, it
// bounds lambda_max(D^-1 K) of the assembled operator, so a Jacobi weight
// below 2 / lambda makes the smoother convergent, and the V-cycle a
// symmetric positive definite preconditioner, whatever the design.
double jacobi_bound(const std::vector<double>& Ke, int nd) {
    std::vector<double> S(nd * nd), v(nd, 1.0), w(nd);
    for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j)
            S[i * nd + j] = Ke[i * nd + j] / std::sqrt(Ke[i * nd + i] * Ke[j * nd + j]);
    for (int i = 0; i < nd; ++i) v[i] = 1.0 + 0.01 * i;  // not orthogonal to the top mode
    double lam = 0;
    for (int it = 0; it < 4000; ++it) {
        double nrm = 0;
        for (int i = 0; i < nd; ++i) {
            double s = 0;
            for (int j = 0; j < nd; ++j) s += S[i * nd + j] * v[j];
            w[i] = s;
            nrm += s * s;
        }
        nrm = std::sqrt(nrm);
        double vw = 0, vv = 0;
        for (int i = 0; i < nd; ++i) {
            vw += v[i] * w[i];
            vv += v[i] * v[i];
        }
        lam = vw / vv;
        for (int i = 0; i < nd; ++i) v[i] = w[i] / nrm;
    }
    return 1.01 * lam;
}

// Builds the hierarchy once: levels are added while every cell count is even,
// until the coarsest fits the one-workgroup coarse kernel.
int mg_setup(tohip_ctx* ctx) {
    if (!ctx->mg.empty()) return 0;
    std::vector<Grid> gs{ctx->g};
    while (gs.size() < 2 || !mg_coarse_fits(gs.back())) {
        const Grid& f = gs.back();
        if (f.nx % 2 || f.ny % 2 || (f.dim == 3 && f.nz % 2)) break;
        gs.push_back(halve_grid(f));
    }
    if (gs.size() < 2) {
        ctx->err = "multigrid: every cell count must be even to build a coarse level";
        return TOHIP_ERR_ARG;
    }
    std::vector<tohip_mg_level> lv(gs.size());
    // coarsest solve: Chebyshev iteration on [lbound / kappa, lbound]
    // (TOHIP_MG_ITERS / TOHIP_MG_KAPPA override the defaults for measurement)
    int iters = kMgCoarseIters;
    double kappa = kMgCoarseKappa;
    const char* ei = getenv("TOHIP_MG_ITERS");
    const char* ek = getenv("TOHIP_MG_KAPPA");
    if (ei && atoi(ei) >= 1 && atoi(ei) <= kMgChebMax) iters = atoi(ei);
    if (ek && atof(ek) > 1.0) kappa = atof(ek);
    if (int rc = mg_build_levels(ctx, gs, lv)) {
        hipStreamSynchronize(ctx->stream);
        free_mg_levels(lv);
        return rc;
    }
    ctx->mg_cheb = mg_chebyshev(lv.back().lbound, kappa, iters);
    ctx->mg_sweeps = kMgFallbackSweeps;
    // the coarsest solve as a dense matrix applied across the chip (rebuilt
    // at every operator refresh); TOHIP_MG_DENSE=0 keeps the one-workgroup
    // iteration (measurement)
    const char* ed = getenv("TOHIP_MG_DENSE");
    if (mg_coarse_fits(lv.back().g) && !(ed && atoi(ed) == 0)) {
        const int64_t n = lv.back().g.ndof;
        int rc;
        if ((rc = dalloc(ctx, &ctx->mg_M, (size_t)(n * n)))) {
            hipStreamSynchronize(ctx->stream);
            free_mg_levels(lv);
            return rc;
        }
    }
    // smoother weights in device memory: the bound's weight until the first
    // estimate; TOHIP_MG_WADAPT=0 (or a forced TOHIP_MG_OMEGA) keeps it
    {
        const int nl = (int)lv.size();
        const char* ea = getenv("TOHIP_MG_WADAPT");
        const char* ew = getenv("TOHIP_MG_OMEGA");
        const double wf = ew ? atof(ew) : 0.0;
        ctx->mg_wadapt = !(ea && atoi(ea) == 0) && !(wf > 0 && wf < 1);
        ctx->mg_wsteps = 0;
        std::vector<float> hw(4 * nl + 1, 0.f);
        for (int l = 0; l < nl; ++l) {
            hw[4 * l] = lv[l].omega;
            hw[4 * l + 1] = 1.f;
        }
        hw[4 * nl] = 1.f;
        // the level count is fixed per context: a set-up retried after a
        // failure reuses the array
        int rc = ctx->mg_wd ? 0 : dalloc(ctx, &ctx->mg_wd, hw.size());
        for (int l = 0; rc == 0 && l < nl; ++l) {
            lv[l].wd = ctx->mg_wd + 4 * l;
            if (l < nl - 1 && ctx->mg_wadapt && (rc = dalloc(ctx, &lv[l].pw, lv[l].g.ndof)) == 0)
                if (launch_mg_fill_start(lv[l].g.ndof, lv[l].pw, ctx->stream) != hipSuccess)
                    rc = TOHIP_ERR_HIP;
        }
        if (rc == 0 && hipMemcpyAsync(ctx->mg_wd, hw.data(), hw.size() * sizeof(float),
                                      hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = TOHIP_ERR_HIP;
        if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = TOHIP_ERR_HIP;
        if (rc) {
            hipStreamSynchronize(ctx->stream);
            free_mg_levels(lv);
            return rc;
        }
    }
    ctx->mg = std::move(lv);
    if (int rc = mg_refresh(ctx)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int mg_build_levels(tohip_ctx* ctx, const std::vector<Grid>& gs, std::vector<tohip_mg_level>& lv) {
    const int dim = ctx->g.dim;
    const int nd = dim << dim;
    // TOHIP_MG_OMEGA overrides the weight of every level (measurement only)
    const char* ew = getenv("TOHIP_MG_OMEGA");
    const double wforce = ew ? atof(ew) : 0.0;
    std::vector<uint8_t> mask = ctx->nodemask;
    for (size_t l = 0; l < gs.size(); ++l) {
        tohip_mg_level& L = lv[l];
        L.g = gs[l];
        double h[3];
        for (int d = 0; d < 3; ++d) h[d] = ctx->h[d] * (double)(1 << l);
        std::vector<double> Ke;
        element_stiffness_bits(dim, ctx->E, ctx->nu, h, Ke);
        L.lbound = jacobi_bound(Ke, nd);
        L.omega = (float)(wforce > 0 && wforce < 1 ? wforce : 1.9 / L.lbound);
        if (l == 0) {
            L.op64 = ctx->op64;
            L.op32 = ctx->op32;
            L.E64 = ctx->E64;
            L.E32 = ctx->E32;
            L.fixmask = ctx->fixmask;
            L.dinv = ctx->dinv32;
        } else {
            if (build_elem_op(dim, Ke, L.op64) > 1e-12) {
                ctx->err = "multigrid: coarse element operator is not block diagonal";
                return TOHIP_ERR_STATE;
            }
            detect_cube(L.op64, dim);
            L.op32 = cast_elem_op<float>(L.op64);
            mask = coarse_mask(gs[l - 1], gs[l], mask);
            int rc = 0;
            if ((rc = dalloc(ctx, &L.E64, L.g.ncells)) || (rc = dalloc(ctx, &L.E32, L.g.ncells)) ||
                (rc = dalloc(ctx, &L.fixmask, L.g.nnodes)) || (rc = dalloc(ctx, &L.dinv, L.g.ndof)) ||
                (rc = dalloc(ctx, &L.b, L.g.ndof)))
                return rc;
            CK(hipMemcpyAsync(L.fixmask, mask.data(), L.g.nnodes, hipMemcpyHostToDevice,
                              ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));  // `mask` is reassigned next level
        }
        int rc = 0;
        if ((rc = dalloc(ctx, &L.x, L.g.ndof)) || (rc = dalloc(ctx, &L.rz, L.g.ndof))) return rc;
    }
    if (!ctx->ones32) {
        std::vector<float> ones(ctx->g.ndof, 1.0f);
        if (int rc = dalloc(ctx, &ctx->ones32, ctx->g.ndof)) return rc;
        CK(hipMemcpyAsync(ctx->ones32, ones.data(), ones.size() * sizeof(float),
                          hipMemcpyHostToDevice, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

// z = M r, one V(1,1) cycle; z is left in mg[0].rz.  With `part`, the finest
// post-smoothing kernel writes the (r.z, r.r) partials the PCG direction
// kernel reduces (*nb_part of them).  With `upd`, the finest pre-smoothing
// first applies the pending PCG update to r and x (the cycle then runs on the
// new residual, which it writes to upd->rnew).
hipError_t launch_vcycle(tohip_ctx* ctx, const float* r, const PcgScalars* sc, double* part,
                         int* nb_part, const MgFusedUpdate* upd = nullptr) {
    std::vector<tohip_mg_level>& M = ctx->mg;
    const int L = (int)M.size() - 1;
    const double md = ctx->meandiag;
    hipStream_t s = ctx->stream;
    hipError_t e;
    for (int l = 0; l < L; ++l) {
        const float* b = l == 0 ? r : M[l].b;
        const MgLevelView F = M[l].view(), C = M[l + 1].view();
        if ((e = launch_mg_smooth(l == 0 && upd ? 3 : 0, F, M[l].omega, b, nullptr, nullptr,
                                  nullptr, M[l].x, M[l].rz, md, sc, nullptr, s,
                                  l == 0 ? upd : nullptr)) != hipSuccess)
            return e;
        if ((e = launch_mg_restrict(F, C, M[l].rz, M[l + 1].b, sc, s)) != hipSuccess) return e;
    }
    // coarsest level: x[L] = sweeps of damped Jacobi
    const float* xc = M[L].x;
    const MgLevelView CL = M[L].view();
    const float w = M[L].omega;
    if (ctx->mg_M) {
        if ((e = launch_mg_dense((int)CL.g.ndof, ctx->mg_M, M[L].b, M[L].x, sc,
                                 s)) != hipSuccess)
            return e;
    } else if (mg_coarse_fits(CL.g)) {
        if ((e = launch_mg_coarse(CL, ctx->mg_cheb, M[L].b, M[L].x, sc, s)) != hipSuccess)
            return e;
    } else {
        if ((e = launch_mg_smooth(0, CL, w, M[L].b, nullptr, nullptr, nullptr, M[L].x, M[L].rz,
                                  md, sc, nullptr, s)) != hipSuccess)
            return e;
        float* cur = M[L].x;
        float* nxt = M[L].rz;
        for (int k = 1; k < ctx->mg_sweeps; ++k) {
            if ((e = launch_mg_smooth(2, CL, w, M[L].b, cur, nullptr, nullptr, nxt, nullptr, md,
                                      sc, nullptr, s)) != hipSuccess)
                return e;
            std::swap(cur, nxt);
        }
        xc = cur;
    }
    for (int l = L - 1; l >= 0; --l) {
        const float* b = l == 0 ? (upd ? upd->rnew : r) : M[l].b;
        const MgLevelView F = M[l].view(), C = M[l + 1].view();
        if ((e = launch_mg_smooth(1, F, M[l].omega, b, M[l].x, &C, xc, M[l].rz, nullptr, md, sc,
                                  l == 0 ? part : nullptr, s)) != hipSuccess)
            return e;
        xc = M[l].rz;
    }
    if (nb_part) *nb_part = mg_blocks(M[0].g);
    return hipSuccess;
}

// One captured PCG graph: `iters` x (direction+matvec, update).  The
// kernels pick their buffers and parity from the device iteration counter,
// so a graph can be replayed from any iteration.
// With `ec` (FP64 only) each iteration ends with the EnergyCriteria check.
template <typename T>
int capture_pcg_graph(tohip_ctx* ctx, hipGraphExec_t* exec, int iters, const ElemOp<T>& op, T* x,
                      T* r, T* dinv, T** p, T* q, const T* E, PcgScalars* sc,
                      EnergyCrit* ec = nullptr) {
    if (*exec) return 0;
    hipGraph_t graph;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < iters; ++it) {
        int nbA = 0;
        hipError_t e1 = launch_pcg_dirmatvec<T>(ctx->g, op, r, dinv, p[0], p[1], q, ctx->fixmask,
                                                E, ctx->meandiag, sc, ctx->partB, kVecBlocks,
                                                ctx->partA, &nbA, ctx->stream);
        hipError_t e2 = launch_pcg_update<T>(ctx->g.ndof, x, r, p[0], p[1], q, dinv, sc,
                                             ctx->partA, nbA, ctx->partB, ctx->stream);
        if constexpr (std::is_same<T, double>::value)
            if (ec && e2 == hipSuccess)
                e2 = launch_energy_check(ctx->g.ndof, x, r, ctx->b64, sc, ec, ctx->partials,
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

// Whether the multigrid PCG graphs fuse each update into the next V-cycle's
// finest pre-smoothing (TOHIP_MG_FUSE=0 keeps the separate update kernel,
// for measurement).
static bool mg_fuse_update() {
    static const bool on = [] {
        const char* e = getenv("TOHIP_MG_FUSE");
        return !(e && atoi(e) == 0);
    }();
    return on;
}

// One captured multigrid-preconditioned FP32 PCG graph: `iters` x (V-cycle
// z = M r with the (r.z, r.r) partials, direction p = z + beta p with
// q = K p, update).  The direction and update kernels see an all-ones
// diagonal, so they compute p = z + beta p and skip the Jacobi scaling.
// Fused form: the update of iteration i runs inside the finest
// pre-smoothing of iteration i + 1, which reads the residual from one buffer
// of the pair (r32, r32b) and writes the new one to the other (neighbouring
// workgroups read the old residual on their halos); the graph's last update
// is a separate kernel that leaves r in r32, so x32 and r32 hold the
// iterate and residual at every graph boundary, as the host expects.
int capture_mg_graph(tohip_ctx* ctx, hipGraphExec_t* exec, int iters) {
    if (*exec) return 0;
    // (the fused update is a mode of the tile smoother: not with z-marching)
    const bool fuse = mg_fuse_update() && !mg_use_zm(ctx->g);
    if (fuse && !ctx->r32b)
        if (int rc = dalloc(ctx, &ctx->r32b, ctx->g.ndof)) return rc;
    float* R[2] = {ctx->r32, ctx->r32b};  // residual of in-graph iteration i: R[i & 1]
    hipGraph_t graph;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int nbA = 0;
    for (int it = 0; it < iters; ++it) {
        int nbB = 0;
        hipError_t e0;
        if (fuse && it > 0) {
            const MgFusedUpdate u{ctx->q32, ctx->p32[0], ctx->p32[1], R[it & 1], ctx->x32,
                                  ctx->partA, nbA, ctx->sc32};
            e0 = launch_vcycle(ctx, R[(it - 1) & 1], ctx->sc32, ctx->partB, &nbB, &u);
        } else {
            e0 = launch_vcycle(ctx, ctx->r32, ctx->sc32, ctx->partB, &nbB);
        }
        hipError_t e1 = e0 == hipSuccess
                            ? launch_pcg_dirmatvec<float>(
                                  ctx->g, ctx->op32, ctx->mg[0].rz, nullptr, ctx->p32[0],
                                  ctx->p32[1], ctx->q32, ctx->fixmask, ctx->E32, ctx->meandiag,
                                  ctx->sc32, ctx->partB, nbB, ctx->partA, &nbA, ctx->stream)
                            : e0;
        hipError_t e2 = hipSuccess;
        if (e1 == hipSuccess && (!fuse || it == iters - 1)) {
            float* rin = fuse ? R[it & 1] : ctx->r32;
            e2 = launch_pcg_update_from<float>(ctx->g.ndof, ctx->x32, rin, ctx->r32, ctx->p32[0],
                                               ctx->p32[1], ctx->q32, nullptr, ctx->sc32,
                                               ctx->partA, nbA, ctx->partB, ctx->stream);
        }
        if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
            hipStreamEndCapture(ctx->stream, &graph);
            ctx->err = "multigrid PCG graph capture failed";
            return TOHIP_ERR_HIP;
        }
    }
    CK(hipStreamEndCapture(ctx->stream, &graph));
    CK(hipGraphInstantiate(exec, graph, nullptr, nullptr, 0));
    CK(hipGraphDestroy(graph));
    return 0;
}

// Replays PCG graphs until the device sets `done`.  A long graph runs while
// the residual history predicts many iterations to go, a short one near the
// end, which keeps the no-op tail after convergence short.
// `nbig` is the iteration count of the long graph.
int replay_pcg_graphs(tohip_ctx* ctx, hipGraphExec_t big, hipGraphExec_t small, int nbig,
                      PcgScalars* sc) {
    double prev_res = -1;
    long long prev_it = -1;
    bool use_small = false;
    int launches = 1;
    for (;;) {
        if (ctx->time_graphs) CK(hipEventRecord(ctx->graph_ev[0], ctx->stream));
        for (int l = 0; l < launches; ++l)
            CK(hipGraphLaunch(use_small ? small : big, ctx->stream));
        if (ctx->time_graphs) CK(hipEventRecord(ctx->graph_ev[1], ctx->stream));
        CK(hipMemcpyAsync(ctx->h_sc, sc, sizeof(PcgScalars), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        ++ctx->host_syncs;
        if (ctx->time_graphs) {
            float ms = 0;
            hipEventElapsedTime(&ms, ctx->graph_ev[0], ctx->graph_ev[1]);
            ctx->graph_ms += ms;
            ctx->graph_launches += launches;
        }
        const PcgScalars& h = *ctx->h_sc;
        if (h.done) break;
        const double res = std::sqrt(h.rr);
        use_small = false;
        launches = 1;
        if (prev_res > 0 && res > 0 && h.tol > 0 && h.iter > prev_it) {
            const double rate = std::log(res / prev_res) / double(h.iter - prev_it);
            if (rate < 0) {
                const double togo = std::log(h.tol / res) / rate;
                use_small = togo < 0.75 * nbig;
                // far from convergence: queue several long graphs per host check
                // (fewer idle gaps; the surplus after `done` is a few no-op launches)
                if (togo > 3.0 * nbig) launches = std::min(4, (int)(togo / (1.5 * nbig)));
            }
        }
        prev_res = res;
        prev_it = h.iter;
    }
    return 0;
}

template <typename T>
int run_pcg_graphs(tohip_ctx* ctx, hipGraphExec_t* big, hipGraphExec_t* small,
                   const ElemOp<T>& op, T* x, T* r, T* dinv, T** p, T* q, const T* E,
                   PcgScalars* sc, EnergyCrit* ec = nullptr) {
    if (int rc = capture_pcg_graph<T>(ctx, big, kGraphItersBig, op, x, r, dinv, p, q, E, sc, ec))
        return rc;
    if (int rc =
            capture_pcg_graph<T>(ctx, small, kGraphItersSmall, op, x, r, dinv, p, q, E, sc, ec))
        return rc;
    return replay_pcg_graphs(ctx, *big, *small, kGraphItersBig, sc);
}

// Iterations per long / short multigrid graph; TOHIP_MG_GRAPH="long,short"
// overrides them (measurement).
std::pair<int, int> mg_graph_iters() {
    static const std::pair<int, int> v = [] {
        std::pair<int, int> r{kGraphItersMg, kGraphItersMgSmall};
        if (const char* e = getenv("TOHIP_MG_GRAPH")) {
            int a = 0, b = 0;
            if (sscanf(e, "%d,%d", &a, &b) == 2 && a >= 1 && a <= 64 && b >= 1 && b <= a) r = {a, b};
        }
        return r;
    }();
    return v;
}

int run_mg_graphs(tohip_ctx* ctx) {
    if (int rc = mg_join(ctx)) return rc;
    const auto [nl, ns] = mg_graph_iters();
    if (int rc = capture_mg_graph(ctx, &ctx->g32mg, nl)) return rc;
    if (int rc = capture_mg_graph(ctx, &ctx->g32mgs, ns)) return rc;
    return replay_pcg_graphs(ctx, ctx->g32mg, ctx->g32mgs, nl, ctx->sc32);
}

// A solve that broke down (PcgScalars::err): u = NaN, the message, and
// TOHIP_BREAKDOWN.
int pcg_breakdown(tohip_ctx* ctx, int err) {
    CK(launch_fill64(ctx->g.ndof, std::numeric_limits<double>::quiet_NaN(), ctx->u, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->err = pcg_err_text(err);
    return TOHIP_BREAKDOWN;
}

// Device-resident Jacobi PCG on K(E) x = b.  x is the initial guess when
// !initially_zero.  Returns 0, or TOHIP_NOT_CONVERGED.  Under EnergyCriteria
// (ctx->conv_kind) the residual test is off and the energy test with
// tolerance `etol` stops the iteration; it is evaluated on the starting
// iterate too, as IterativeSolvers evaluates its stopping test before every
// iteration; a NaN or negative energy returns TOHIP_ERR_DOMAIN.
int solve_fp64(tohip_ctx* ctx, double abstol, double reltol, int64_t maxiter, bool zero,
               int64_t* iters, double* resid, double etol) {
    const int64_t n = ctx->g.ndof;
    const bool energy = ctx->conv_kind == TOHIP_CONV_ENERGY;
    PcgScalars hs{};
    hs.abstol = energy ? 0.0 : abstol;
    hs.reltol = energy ? 0.0 : reltol;
    hs.maxiter = maxiter;
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(ctx->sc64, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice,
                      ctx->stream));
    if (energy) {
        if (!(etol > 0)) {
            ctx->err = "EnergyCriteria needs a positive tolerance";
            return TOHIP_ERR_ARG;
        }
        ctx->h_ecrit->tol = etol;
        ctx->h_ecrit->converged = 0;
        ctx->h_ecrit->err = 0;
        CK(hipMemcpyAsync(ctx->ecrit, ctx->h_ecrit, sizeof(EnergyCrit), hipMemcpyHostToDevice,
                          ctx->stream));
    }
    if (zero) {
        CK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, nullptr, ctx->r64, ctx->dinv64, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    } else {
        CK(launch_matvec<double>(ctx->g, ctx->op64, ctx->u, ctx->fixmask, ctx->E64, ctx->t64,
                                 ctx->meandiag, ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, ctx->t64, ctx->r64, ctx->dinv64, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    }
    if (energy) {
        CK(launch_energy_check(n, ctx->u, ctx->r64, ctx->b64, ctx->sc64, ctx->ecrit,
                               ctx->partials, ctx->stream));
        if (int rc = run_pcg_graphs<double>(ctx, &ctx->g64e, &ctx->g64es, ctx->op64, ctx->u,
                                            ctx->r64, ctx->dinv64, ctx->p64, ctx->q64, ctx->E64,
                                            ctx->sc64, ctx->ecrit))
            return rc;
        CK(hipMemcpyAsync(ctx->h_ecrit, ctx->ecrit, sizeof(EnergyCrit), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
    } else if (int rc = run_pcg_graphs<double>(ctx, &ctx->g64, &ctx->g64s, ctx->op64, ctx->u,
                                               ctx->r64, ctx->dinv64, ctx->p64, ctx->q64,
                                               ctx->E64, ctx->sc64)) {
        return rc;
    }
    const PcgScalars& h = *ctx->h_sc;
    if (iters) *iters = h.iter;
    if (resid) *resid = std::sqrt(h.rr);
    if (h.err) return pcg_breakdown(ctx, h.err);
    if (energy) {
        const EnergyCrit& e = *ctx->h_ecrit;
        if (e.err) {
            ctx->err = e.err == 1 ? "EnergyCriteria: NaN detected in the energy change or x'Ax"
                                  : "EnergyCriteria: expected non-negative x'Ax";
            return TOHIP_ERR_DOMAIN;
        }
        return e.converged ? 0 : TOHIP_NOT_CONVERGED;
    }
    return std::sqrt(h.rr) <= h.tol ? 0 : TOHIP_NOT_CONVERGED;
}

// r64 = b - K64 u (FP64), r32 = r64, partials for the FP32 iteration; returns
// ||r64||.  u += x32 first when `accumulate`.
int refresh_residual(tohip_ctx* ctx, bool accumulate, double* res) {
    const int64_t n = ctx->g.ndof;
    if (accumulate) CK(launch_accumulate_f32(n, ctx->x32, ctx->u, ctx->stream));
    CK(launch_matvec<double>(ctx->g, ctx->op64, ctx->u, ctx->fixmask, ctx->E64, ctx->t64,
                             ctx->meandiag, ctx->stream));
    CK(launch_residual_refresh(n, ctx->b64, ctx->t64, ctx->r64, ctx->r32, ctx->dinv32,
                               ctx->partB, ctx->stream));
    CK(launch_reduce(ctx->partB + 2 * kVecBlocks, kVecBlocks, 1, ctx->dscal + 1, ctx->stream));
    CK(hipMemcpyAsync(ctx->h_scal + 1, ctx->dscal + 1, sizeof(double), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ++ctx->host_syncs;
    *res = std::sqrt(ctx->h_scal[1]);
    return 0;
}

// Mixed precision: Jacobi PCG with FP32 vectors and FP32 matvec, FP64 dots
// and alpha/beta, inside FP64 iterative refinement by residual replacement
// ("reliable updates"): the solution is accumulated in FP64 and, whenever the
// FP32 recursive residual has dropped by `inner_rel` since the last
// replacement, r = b - K u is recomputed in FP64 and the FP32 residual is
// replaced by it while the search direction and rho are kept, so the Krylov
// information is not thrown away as in restarted refinement -- unless the
// true residual comes out more than kRestartRatio times the recursive one:
// then the direction restarts from p = z.  Stops when the
// FP64 residual meets max(reltol * ||r0||, abstol).
int solve_mixed(tohip_ctx* ctx, double abstol, double reltol, int64_t maxiter, bool zero,
                double inner_rel, bool mg, int64_t* iters, double* resid) {
    const int64_t n = ctx->g.ndof;
    if (ctx->conv_kind != TOHIP_CONV_DEFAULT) {
        ctx->err = "EnergyCriteria is implemented for the FP64 solver (precision TOHIP_FP64)";
        return TOHIP_ERR_ARG;
    }
    if (mg)
        if (int rc = mg_setup(ctx)) return rc;
    if (zero) CK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
    CK(hipMemsetAsync(ctx->x32, 0, n * sizeof(float), ctx->stream));
    CK(hipMemsetAsync(ctx->p32[0], 0, n * sizeof(float), ctx->stream));
    double res = 0;
    if (int rc = refresh_residual(ctx, false, &res)) return rc;
    if (!std::isfinite(res)) return pcg_breakdown(ctx, kPcgErrNonFinite);
    const double tol = std::max(reltol * res, abstol);
    PcgScalars hs{};
    hs.maxiter = maxiter;
    hs.iter = 0;
    hs.rho[0] = hs.rho[1] = 1.0;
    hs.tol = std::max(inner_rel * res, tol);
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(ctx->sc32, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice,
                      ctx->stream));
    int64_t total = 0;
    int replacements = 0;
    // TOHIP_RESTART_RATIO overrides kRestartRatio (measurement, read per
    // solve; 0: never restart)
    const char* er = getenv("TOHIP_RESTART_RATIO");
    const double restart = er ? atof(er) : kRestartRatio;
    while (res > tol) {
        if (int rc = mg ? run_mg_graphs(ctx)
                        : run_pcg_graphs<float>(ctx, &ctx->g32, &ctx->g32s, ctx->op32, ctx->x32,
                                                ctx->r32, ctx->dinv32, ctx->p32, ctx->q32,
                                                ctx->E32, ctx->sc32))
            return rc;
        total = ctx->h_sc->iter;
        if (ctx->h_sc->err) {
            ctx->last_replacements = replacements;
            if (iters) *iters = total;
            if (resid) *resid = std::sqrt(ctx->h_sc->rr);
            return pcg_breakdown(ctx, ctx->h_sc->err);
        }
        if (int rc = refresh_residual(ctx, true, &res)) return rc;
        ++replacements;
        if (!std::isfinite(res)) {
            ctx->last_replacements = replacements;
            if (iters) *iters = total;
            if (resid) *resid = res;
            return pcg_breakdown(ctx, kPcgErrNonFinite);
        }
        if (res <= tol || total >= maxiter || replacements > 200) break;
        CK(launch_pcg_rearm(ctx->sc32, ctx->dscal + 1, inner_rel, 0.9 * tol, restart,
                            ctx->stream));
    }
    ctx->last_replacements = replacements;
    if (iters) *iters = total;
    if (resid) *resid = res;
    return res <= tol ? 0 : TOHIP_NOT_CONVERGED;
}

// FP64 PCG preconditioned by the FP32 multigrid V-cycle (TOHIP_FP64_MG): the
// operator, vectors, recurrences and stopping rule are those of solve_fp64;
// z = M r goes through an FP32 copy of r and comes back as FP64.  One captured
// iteration: r32 = r, the V-cycle (with the (r.z, r.r) partials), z64 = z32,
// the FP64 direction + K p kernel on p = z + beta p, the FP64 update.  The
// comparator that separates the preconditioner's gain from FP32 arithmetic's.
int capture_fp64_mg_graph(tohip_ctx* ctx, hipGraphExec_t* exec, int iters) {
    if (*exec) return 0;
    const int64_t n = ctx->g.ndof;
    hipGraph_t graph;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < iters; ++it) {
        int nbB = 0, nbA = 0;
        hipError_t e0 = launch_f64_to_f32(n, ctx->r64, ctx->r32, ctx->stream);
        hipError_t e1 = e0 == hipSuccess ? launch_vcycle(ctx, ctx->r32, ctx->sc64, ctx->partB, &nbB)
                                         : e0;
        hipError_t e2 = e1 == hipSuccess ? launch_f32_to_f64(n, ctx->mg[0].rz, ctx->t64,
                                                             ctx->stream)
                                         : e1;
        hipError_t e3 = e2 == hipSuccess
                            ? launch_pcg_dirmatvec<double>(ctx->g, ctx->op64, ctx->t64,
                                                           nullptr, ctx->p64[0], ctx->p64[1],
                                                           ctx->q64, ctx->fixmask, ctx->E64,
                                                           ctx->meandiag, ctx->sc64, ctx->partB,
                                                           nbB, ctx->partA, &nbA, ctx->stream)
                            : e2;
        hipError_t e4 = e3 == hipSuccess
                            ? launch_pcg_update<double>(n, ctx->u, ctx->r64, ctx->p64[0],
                                                        ctx->p64[1], ctx->q64, nullptr,
                                                        ctx->sc64, ctx->partA, nbA, ctx->partB,
                                                        ctx->stream)
                            : e3;
        if (e4 != hipSuccess) {
            hipStreamEndCapture(ctx->stream, &graph);
            ctx->err = "FP64 multigrid PCG graph capture failed";
            return TOHIP_ERR_HIP;
        }
    }
    CK(hipStreamEndCapture(ctx->stream, &graph));
    CK(hipGraphInstantiate(exec, graph, nullptr, nullptr, 0));
    CK(hipGraphDestroy(graph));
    return 0;
}

int solve_fp64_mg(tohip_ctx* ctx, double abstol, double reltol, int64_t maxiter, bool zero,
                  int64_t* iters, double* resid) {
    const int64_t n = ctx->g.ndof;
    if (ctx->conv_kind != TOHIP_CONV_DEFAULT) {
        ctx->err = "EnergyCriteria is implemented for the FP64 Jacobi solver (TOHIP_FP64)";
        return TOHIP_ERR_ARG;
    }
    if (int rc = mg_setup(ctx)) return rc;
    if (!ctx->ones64) {
        if (int rc = dalloc(ctx, &ctx->ones64, n)) return rc;
        CK(launch_fill64(n, 1.0, ctx->ones64, ctx->stream));
    }
    PcgScalars hs{};
    hs.abstol = abstol;
    hs.reltol = reltol;
    hs.maxiter = maxiter;
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(ctx->sc64, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice,
                      ctx->stream));
    if (zero) {
        CK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, nullptr, ctx->r64, ctx->ones64, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    } else {
        CK(launch_matvec<double>(ctx->g, ctx->op64, ctx->u, ctx->fixmask, ctx->E64, ctx->t64,
                                 ctx->meandiag, ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, ctx->t64, ctx->r64, ctx->ones64, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    }
    if (int rc = mg_join(ctx)) return rc;
    const auto [nl, ns] = mg_graph_iters();
    if (int rc = capture_fp64_mg_graph(ctx, &ctx->g64mg, nl)) return rc;
    if (int rc = capture_fp64_mg_graph(ctx, &ctx->g64mgs, ns)) return rc;
    if (int rc = replay_pcg_graphs(ctx, ctx->g64mg, ctx->g64mgs, nl, ctx->sc64)) return rc;
    const PcgScalars& h = *ctx->h_sc;
    if (iters) *iters = h.iter;
    if (resid) *resid = std::sqrt(h.rr);
    if (h.err) return pcg_breakdown(ctx, h.err);
    return std::sqrt(h.rr) <= h.tol ? 0 : TOHIP_NOT_CONVERGED;
}

// The solve of one precision mode (TOHIP_FP64, _MIXED, _MIXED_MG, _FP64_MG).
// etol: the EnergyCriteria tolerance of the FP64 path; inner: the residual
// replacement ratio of the mixed paths (<= 0: the mode's default).
int solve_mode(tohip_ctx* ctx, int precision, double abstol, double reltol, int64_t maxiter,
               bool zero, double inner, double etol, int64_t* iters, double* resid) {
    switch (precision) {
        case TOHIP_FP64:
            return solve_fp64(ctx, abstol, reltol, maxiter, zero, iters, resid, etol);
        case TOHIP_FP64_MG:
            return solve_fp64_mg(ctx, abstol, reltol, maxiter, zero, iters, resid);
        case TOHIP_MIXED:
            return solve_mixed(ctx, abstol, reltol, maxiter, zero,
                               inner > 0 ? inner : kDefaultReplaceRatio, false, iters, resid);
        case TOHIP_MIXED_MG:
            return solve_mixed(ctx, abstol, reltol, maxiter, zero,
                               inner > 0 ? inner : kDefaultReplaceRatioMg, true, iters, resid);
        default:
            ctx->err = "unknown precision mode";
            return TOHIP_ERR_ARG;
    }
}

bool valid_precision(int p) {
    return p == TOHIP_FP64 || p == TOHIP_MIXED || p == TOHIP_MIXED_MG || p == TOHIP_FP64_MG;
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
    ctx->f_halo = 0;  // largest node offset on any axis (the offsets span [1 - halo, halo])
    for (int v : offs) ctx->f_halo = std::max(ctx->f_halo, v);
    CK(launch_filter_invden(ctx->g, ctx->f_offs, ctx->f_wts, count, ctx->f_invden, ctx->stream));
    // colsum = J^T 1 (the volume-constraint weights of the OC step)
    std::vector<double> ones(ctx->g.ncells, 1.0);
    CK(hipMemcpyAsync(ctx->cellbuf, ones.data(), ones.size() * sizeof(double),
                      hipMemcpyHostToDevice, ctx->stream));
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, count, ctx->f_halo, ctx->f_invden, ctx->cellbuf,
                     ctx->f_colsum, 1, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->f_rmin = rmin;
    return 0;
}

double host_norm(const double* v, int64_t n) {
    long double s = 0;
    for (int64_t i = 0; i < n; ++i) s += (long double)v[i] * v[i];
    return (double)std::sqrt(s);
}

bool all_finite(const double* v, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

void free_ctx(tohip_ctx* ctx) {
    if (!ctx) return;
    if (ctx->nn) nn_detach_ctx(ctx->nn, ctx);
    for (hipGraphExec_t ge : {ctx->g64, ctx->g32, ctx->g64s, ctx->g32s})
        if (ge) hipGraphExecDestroy(ge);
    void* bufs[] = {ctx->fixmask, ctx->xt,       ctx->E64,      ctx->E32,    ctx->f,
                    ctx->u,       ctx->b64,      ctx->r64,      ctx->t64,    ctx->dinv64,
                    ctx->p64[0],  ctx->p64[1],   ctx->q64,      ctx->b32,    ctx->x32,
                    ctx->r32,     ctx->dinv32,   ctx->p32[0],   ctx->p32[1], ctx->q32,
                    ctx->sc64,    ctx->sc32,     ctx->partials, ctx->dscal,  ctx->partA, ctx->partB,
                    ctx->f_offs,  ctx->f_wts,    ctx->f_invden, ctx->f_colsum, ctx->xd,
                    ctx->xnew,    ctx->dc,       ctx->ce,       ctx->grad,   ctx->cellbuf,
                    ctx->oc};
    if (ctx->mg_stream) hipStreamSynchronize(ctx->mg_stream);  // a rebuild may be in flight
    for (void* b : bufs)
        if (b) hipFree(b);
    for (hipGraphExec_t ge : {ctx->g32mg, ctx->g32mgs, ctx->g64e, ctx->g64es, ctx->gcsr,
                              ctx->g64mg, ctx->g64mgs})
        if (ge) hipGraphExecDestroy(ge);
    for (void* b : {(void*)ctx->csr_rowptr, (void*)ctx->csr_cols, (void*)ctx->csr_vals,
                    (void*)ctx->csr_diag, (void*)ctx->csr_dinv, (void*)ctx->csr_ke,
                    (void*)ctx->csr_md, (void*)ctx->mma_L, (void*)ctx->mma_U,
                    (void*)ctx->mma_x1, (void*)ctx->mma_x2, (void*)ctx->mma_st,
                    (void*)ctx->mg_M, (void*)ctx->r32b, (void*)ctx->mg_wd})
        if (b) hipFree(b);
    if (ctx->h_mma) hipHostFree(ctx->h_mma);
    if (ctx->ecrit) hipFree(ctx->ecrit);
    if (ctx->h_ecrit) hipHostFree(ctx->h_ecrit);
    for (hipEvent_t e : ctx->graph_ev)
        if (e) hipEventDestroy(e);
    free_mg_levels(ctx->mg);
    if (ctx->ones32) hipFree(ctx->ones32);
    if (ctx->ones64) hipFree(ctx->ones64);
    if (ctx->h_sc) hipHostFree(ctx->h_sc);
    if (ctx->h_scal) hipHostFree(ctx->h_scal);
    if (ctx->h_oc) hipHostFree(ctx->h_oc);
    for (void* b : {(void*)ctx->beso_sens[0], (void*)ctx->beso_sens[1], (void*)ctx->beso_mask,
                    (void*)ctx->beso_st})
        if (b) hipFree(b);
    if (ctx->h_beso) hipHostFree(ctx->h_beso);
    for (hipEvent_t e : ctx->ws_ev)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : ctx->mg_ev)
        if (e) hipEventDestroy(e);
    if (ctx->mg_stream) hipStreamDestroy(ctx->mg_stream);
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
    // kernels address vectors with 32-bit byte offsets: FP64 vectors < 4 GiB
    if (g.ndof >= (int64_t)1 << 29) {
        ctx->err = "grid too large (2^29 or more DOFs)";
        return TOHIP_ERR_ARG;
    }
    ctx->E = E;
    ctx->nu = nu;
    ctx->xmin = xmin;
    ctx->pen.p = pp;
    ctx->pbi = pbi ? 1 : 0;
    for (int d = 0; d < 3; ++d) ctx->h[d] = d < dim ? h[d] : 1.0;

    std::vector<double> Ke;
    element_stiffness_bits(dim, E, nu, ctx->h, Ke);
    const double off = build_elem_op(dim, Ke, ctx->op64);
    detect_cube(ctx->op64, dim);
    if (off > 1e-12) {
        ctx->err = "element operator is not block diagonal in the corner-Hadamard basis";
        return TOHIP_ERR_STATE;
    }
    ctx->op32 = cast_elem_op<float>(ctx->op64);
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
        (rc = dalloc(ctx, &ctx->partials, kGeneralPartials)) ||
        (rc = dalloc(ctx, &ctx->partA, kMaxPartials)) ||
        (rc = dalloc(ctx, &ctx->partB, 2 * kMaxPartials)) || (rc = dalloc(ctx, &ctx->dscal, 16)) ||
        (rc = dalloc(ctx, &ctx->f_invden, nC)) || (rc = dalloc(ctx, &ctx->f_colsum, nC)) ||
        (rc = dalloc(ctx, &ctx->xd, nC)) || (rc = dalloc(ctx, &ctx->xnew, nC)) ||
        (rc = dalloc(ctx, &ctx->dc, nC)) || (rc = dalloc(ctx, &ctx->ce, nC)) ||
        (rc = dalloc(ctx, &ctx->grad, nC)) || (rc = dalloc(ctx, &ctx->cellbuf, nC)) ||
        (rc = dalloc(ctx, &ctx->oc, 1)))
        return rc;
    CK(hipHostMalloc((void**)&ctx->h_sc, sizeof(PcgScalars), hipHostMallocDefault));
    CK(hipHostMalloc((void**)&ctx->h_scal, 8 * sizeof(double), hipHostMallocDefault));
    CK(hipHostMalloc((void**)&ctx->h_oc, sizeof(OcState), hipHostMallocDefault));
    CK(hipMalloc((void**)&ctx->ecrit, sizeof(EnergyCrit)));
    CK(hipHostMalloc((void**)&ctx->h_ecrit, sizeof(EnergyCrit), hipHostMallocDefault));
    *ctx->h_ecrit = EnergyCrit{};
    CK(hipMemcpyAsync(ctx->fixmask, nodemask.data(), g.nnodes, hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->nodemask = std::move(nodemask);
    CK(hipMemsetAsync(ctx->sc64, 0, sizeof(PcgScalars), ctx->stream));
    CK(hipMemsetAsync(ctx->sc32, 0, sizeof(PcgScalars), ctx->stream));
    CK(hipMemsetAsync(ctx->oc, 0, sizeof(OcState), ctx->stream));
    CK(hipMemsetAsync(ctx->u, 0, nD * sizeof(double), ctx->stream));
    for (double* b : {ctx->p64[0], ctx->p64[1], ctx->q64, ctx->r64, ctx->t64, ctx->b64})
        CK(hipMemsetAsync(b, 0, nD * sizeof(double), ctx->stream));
    for (float* b : {ctx->p32[0], ctx->p32[1], ctx->q32, ctx->r32, ctx->x32, ctx->b32})
        CK(hipMemsetAsync(b, 0, nD * sizeof(float), ctx->stream));
    CK(hipMemsetAsync(ctx->f, 0, nD * sizeof(double), ctx->stream));
    // solver.vars = ones (src/FEA/solvers_api.jl:497-499)
    std::vector<double> ones(nC, 1.0);
    CK(hipMemcpyAsync(ctx->xt, ones.data(), nC * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    if ((rc = refresh_operator(ctx))) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---- assembled K in CSR (CGAssemblySolver comparator) ----
constexpr int kCsrGraphIters = 32;

// Pattern (host row pointers, device columns) on first use, then the values
// of K(current densities) with apply!'s boundary rule and the preconditioner
// diagonal (jacobi: 1 / diag, else identity).
int csr_assemble(tohip_ctx* ctx, int safe, int jacobi) {
    const Grid& g = ctx->g;
    const int dim = g.dim;
    bool pattern = false;
    if (!ctx->csr_rowptr) {
        std::vector<int64_t> rowptr(g.ndof + 1, 0);
        for (int64_t n = 0; n < g.nnodes; ++n) {
            const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
            const int k = dim == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
            const int cx = 1 + (i > 0) + (i < g.NX - 1), cy = 1 + (j > 0) + (j < g.NY - 1);
            const int cz = dim == 3 ? 1 + (k > 0) + (k < g.NZ - 1) : 1;
            for (int d = 0; d < dim; ++d)
                rowptr[dim * n + d + 1] = rowptr[dim * n + d] + (int64_t)dim * cx * cy * cz;
        }
        const int64_t nnz = rowptr[g.ndof];
        if (nnz >= ((int64_t)1 << 31)) {
            ctx->err = "assembled matrix exceeds 2^31 entries";
            return TOHIP_ERR_ARG;
        }
        const int nd = dim << dim;
        std::vector<double> Ke;
        element_stiffness_bits(dim, ctx->E, ctx->nu, ctx->h, Ke);
        int rc;
        if ((rc = dalloc(ctx, &ctx->csr_rowptr, g.ndof + 1)) ||
            (rc = dalloc(ctx, &ctx->csr_cols, nnz)) || (rc = dalloc(ctx, &ctx->csr_vals, nnz)) ||
            (rc = dalloc(ctx, &ctx->csr_diag, g.ndof)) || (rc = dalloc(ctx, &ctx->csr_dinv, g.ndof)) ||
            (rc = dalloc(ctx, &ctx->csr_ke, nd * nd)) || (rc = dalloc(ctx, &ctx->csr_md, 3)))
            return rc;
        CK(hipMemcpyAsync(ctx->csr_rowptr, rowptr.data(), rowptr.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice, ctx->stream));
        CK(hipMemcpyAsync(ctx->csr_ke, Ke.data(), Ke.size() * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));  // host vectors go out of scope
        ctx->csr_nnz = nnz;
        pattern = true;
    }
    CK(launch_csr_build(g, ctx->csr_ke, ctx->E64, ctx->fixmask, ctx->csr_rowptr, ctx->csr_cols,
                        safe, ctx->csr_vals, ctx->csr_diag, ctx->partials, ctx->csr_md, pattern,
                        ctx->stream));
    CK(launch_csr_dinv(g, ctx->csr_diag, ctx->fixmask, ctx->csr_md, jacobi, safe, ctx->csr_dinv,
                       ctx->stream));
    ctx->csr_safe = safe;
    return 0;
}

int capture_csr_graph(tohip_ctx* ctx, hipGraphExec_t* exec, int iters) {
    if (*exec) return 0;
    hipGraph_t graph;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < iters; ++it) {
        hipError_t e = launch_csr_pcg_iter(ctx->g.ndof, ctx->csr_rowptr, ctx->csr_cols,
                                           ctx->csr_vals, ctx->u, ctx->r64, ctx->csr_dinv,
                                           ctx->p64[0], ctx->p64[1], ctx->q64, ctx->sc64,
                                           ctx->partA, ctx->partB, ctx->stream);
        if (e != hipSuccess) {
            hipStreamEndCapture(ctx->stream, &graph);
            ctx->err = "CSR PCG graph capture failed";
            return TOHIP_ERR_HIP;
        }
    }
    CK(hipStreamEndCapture(ctx->stream, &graph));
    CK(hipGraphInstantiate(exec, graph, nullptr, nullptr, 0));
    CK(hipGraphDestroy(graph));
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

int tohip_memory(const tohip_ctx* ctx, int64_t* ctx_bytes, int64_t* device_used,
                 int64_t* device_total) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (ctx_bytes) *ctx_bytes = (int64_t)(ctx->dev_bytes + (ctx->nn ? nn_device_bytes(ctx->nn) : 0));
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return TOHIP_ERR_HIP;
    if (device_used) *device_used = (int64_t)(tot - fr);
    if (device_total) *device_total = (int64_t)tot;
    return 0;
}

}  // extern "C"

// ---- scratch of tohip_compliance_noalloc ----
// The energy kernel's per-block partial sums for the entry point that
// allocates nothing: one slot per workgroup, at most kNoAllocSlots workgroups
// (energy_blocks caps the grid there; the kernel strides over the cells, so
// any grid fits).  Every slot is written before it is read, so nothing needs
// clearing.  Each GPU holds its own copy until the process ends, at its own
// address, resolved once per device with that device current (at context
// creation); every call on a GPU shares it, so calls that overlap on one GPU
// race.
constexpr int kNoAllocSlots = kMaxPartials;
static __device__ __align__(256) double scratch[kNoAllocSlots];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// This device's copy of `scratch` (hipGetSymbolAddress once per device).
static int scratch_init(tohip_ctx* ctx) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        ctx->err = "cannot identify the current GPU";
        return TOHIP_ERR_HIP;
    }
    ctx->device = dev;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return TOHIP_ERR_HIP;
    }
    g_scratch_dev[dev] = static_cast<double*>(p);
    return 0;
}

extern "C" {

int tohip_compliance_noalloc(tohip_ctx* ctx, const double* u_dev, const double* xt_dev,
                             double* ce_dev, double* grad_dev, double* obj_dev, void* stream) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (!u_dev || !xt_dev || !ce_dev || !grad_dev || !obj_dev) {
        ctx->err = "tohip_compliance_noalloc: every device buffer is required";
        return TOHIP_ERR_ARG;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != ctx->device) {
        ctx->err = "tohip_compliance_noalloc: the context's GPU must be the current device";
        return TOHIP_ERR_STATE;
    }
    double* scr = g_scratch_dev[dev];
    if (!scr) {
        ctx->err = "tohip_compliance_noalloc: scratch not resolved on this GPU";
        return TOHIP_ERR_STATE;
    }
    const int nb = energy_blocks(ctx->g);
    if (nb > kNoAllocSlots) {
        ctx->err = "tohip_compliance_noalloc: the grid needs more partial-sum slots than the "
                   "fixed scratch holds (4096)";
        return TOHIP_ERR_ARG;
    }
    CK(launch_energy(ctx->g, ctx->op64, u_dev, xt_dev, ctx->xmin, ctx->pen, ctx->pbi, ce_dev,
                     grad_dev, scr, obj_dev, static_cast<hipStream_t>(stream)));
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
    int rc = scratch_init(ctx);
    if (rc == 0)
        rc = create_impl(ctx, dim, nels, h, E, nu, xmin, penalty_p, penalty_before_interp,
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
    ctx->pen.p = penalty_p;
    if (int rc = refresh_operator(ctx)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_set_penalty_fun(tohip_ctx* ctx, int kind, double penalty_p, int proj, double beta) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (kind < TOHIP_PEN_POWER || kind > TOHIP_PEN_SINH) {
        ctx->err = "tohip_set_penalty_fun: unknown penalty kind";
        return TOHIP_ERR_ARG;
    }
    if (proj < TOHIP_PROJ_NONE || proj > TOHIP_PROJ_SIGMOID) {
        ctx->err = "tohip_set_penalty_fun: unknown projection";
        return TOHIP_ERR_ARG;
    }
    if (!(penalty_p > 0) || !std::isfinite(penalty_p)) {
        ctx->err = "penalty must be positive";
        return TOHIP_ERR_ARG;
    }
    if (proj != TOHIP_PROJ_NONE && (!(beta >= 0) || !std::isfinite(beta))) {
        ctx->err = "tohip_set_penalty_fun: projection steepness must be finite and >= 0";
        return TOHIP_ERR_ARG;
    }
    ctx->pen = Penalty{kind, proj, penalty_p, proj == TOHIP_PROJ_NONE ? 0.0 : beta};
    if (int rc = refresh_operator(ctx)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_set_convergence(tohip_ctx* ctx, int kind, double energy) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (kind != TOHIP_CONV_DEFAULT && kind != TOHIP_CONV_ENERGY) {
        ctx->err = "tohip_set_convergence: unknown criterion";
        return TOHIP_ERR_ARG;
    }
    ctx->conv_kind = kind;
    *ctx->h_ecrit = EnergyCrit{};
    ctx->h_ecrit->energy = energy;
    return 0;
}

int tohip_get_convergence_energy(const tohip_ctx* ctx, double* energy) {
    if (!ctx || !energy) return TOHIP_ERR_ARG;
    *energy = ctx->h_ecrit->energy;
    return 0;
}

int tohip_set_densities(tohip_ctx* ctx, const double* vars) {
    if (!ctx || !vars) return TOHIP_ERR_ARG;
    if (!all_finite(vars, ctx->g.ncells)) {
        ctx->err = "tohip_set_densities: non-finite density";
        return TOHIP_ERR_ARG;
    }
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

int tohip_csr_assemble(tohip_ctx* ctx, int safe) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (int rc = csr_assemble(ctx, safe ? 1 : 0, 1)) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tohip_csr_info(const tohip_ctx* cctx, int64_t* nnz, int64_t* bytes, double* meandiag) {
    if (!cctx) return TOHIP_ERR_ARG;
    tohip_ctx* ctx = const_cast<tohip_ctx*>(cctx);  // for the error message only
    if (!ctx->csr_rowptr) {
        ctx->err = "tohip_csr_info: no assembled matrix";
        return TOHIP_ERR_STATE;
    }
    if (nnz) *nnz = ctx->csr_nnz;
    if (bytes)
        *bytes = ctx->csr_nnz * (int64_t)(sizeof(double) + sizeof(int)) +
                 (ctx->g.ndof + 1) * (int64_t)sizeof(int64_t);
    if (meandiag) {
        double s = 0;
        CK(hipMemcpy(&s, ctx->csr_md, sizeof(double), hipMemcpyDeviceToHost));
        *meandiag = s / (double)ctx->g.ndof;
    }
    return 0;
}

int tohip_csr_download(const tohip_ctx* cctx, int64_t* rowptr, int* cols, double* vals) {
    if (!cctx || !rowptr || !cols || !vals) return TOHIP_ERR_ARG;
    tohip_ctx* ctx = const_cast<tohip_ctx*>(cctx);  // for the error message only
    if (!ctx->csr_rowptr) {
        ctx->err = "tohip_csr_download: no assembled matrix";
        return TOHIP_ERR_STATE;
    }
    CK(hipMemcpy(rowptr, ctx->csr_rowptr, (ctx->g.ndof + 1) * sizeof(int64_t),
                 hipMemcpyDeviceToHost));
    CK(hipMemcpy(cols, ctx->csr_cols, ctx->csr_nnz * sizeof(int), hipMemcpyDeviceToHost));
    CK(hipMemcpy(vals, ctx->csr_vals, ctx->csr_nnz * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int tohip_csr_matvec(tohip_ctx* ctx, const double* x, double* y) {
    if (!ctx || !x || !y) return TOHIP_ERR_ARG;
    if (int rc = csr_assemble(ctx, ctx->csr_safe > 0 ? 1 : 0, 1)) return rc;
    if (int rc = upload_dofs(ctx, x, ctx->r64)) return rc;
    CK(launch_csr_spmv(ctx->g.ndof, ctx->csr_rowptr, ctx->csr_cols, ctx->csr_vals, ctx->r64,
                       ctx->t64, ctx->stream));
    return download_dofs(ctx, ctx->t64, y);
}

int tohip_csr_solve(tohip_ctx* ctx, const double* f, double* u, double abstol, double reltol,
                    int64_t maxiter, int initially_zero, int jacobi, int safe, int64_t* iters_out,
                    double* resid_out) {
    if (!ctx || !f || !u) return TOHIP_ERR_ARG;
    if (!(abstol >= 0) || !(reltol >= 0) || maxiter < 0) {
        ctx->err = "tohip_csr_solve: bad tolerance or maxiter";
        return TOHIP_ERR_ARG;
    }
    const int64_t n = ctx->g.ndof;
    if (!all_finite(f, n) || (!initially_zero && !all_finite(u, n))) {
        ctx->err = "tohip_csr_solve: non-finite load or initial guess";
        return TOHIP_ERR_ARG;
    }
    // apply!(K, f, ch) with zero prescribed values: f = 0 on the prescribed DOFs
    std::vector<double> fz(f, f + n);
    const int dim = ctx->g.dim;
    for (int64_t node = 0; node < ctx->g.nnodes; ++node)
        for (int d = 0; d < dim; ++d)
            if ((ctx->nodemask[node] >> d) & 1)
                fz[ctx->perm.empty() ? dim * node + d : ctx->perm[dim * node + d]] = 0.0;
    if (int rc = upload_dofs(ctx, fz.data(), ctx->b64)) return rc;
    if (int rc = csr_assemble(ctx, safe ? 1 : 0, jacobi ? 1 : 0)) return rc;
    PcgScalars hs{};
    hs.abstol = abstol;
    hs.reltol = reltol;
    hs.maxiter = maxiter;
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(ctx->sc64, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice,
                      ctx->stream));
    if (initially_zero) {
        CK(hipMemsetAsync(ctx->u, 0, n * sizeof(double), ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, nullptr, ctx->r64, ctx->csr_dinv, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    } else {
        if (int rc = upload_dofs(ctx, u, ctx->u)) return rc;
        CK(launch_csr_spmv(n, ctx->csr_rowptr, ctx->csr_cols, ctx->csr_vals, ctx->u, ctx->t64,
                           ctx->stream));
        CK(launch_pcg_start<double>(n, ctx->b64, ctx->t64, ctx->r64, ctx->csr_dinv, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 0, ctx->stream));
    }
    if (int rc = capture_csr_graph(ctx, &ctx->gcsr, kCsrGraphIters)) return rc;
    for (;;) {
        CK(hipGraphLaunch(ctx->gcsr, ctx->stream));
        CK(hipMemcpyAsync(ctx->h_sc, ctx->sc64, sizeof(PcgScalars), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        if (ctx->h_sc->done) break;
    }
    const PcgScalars& h = *ctx->h_sc;
    if (iters_out) *iters_out = h.iter;
    if (resid_out) *resid_out = std::sqrt(h.rr);
    ctx->have_u = true;
    const int bd = h.err ? pcg_breakdown(ctx, h.err) : 0;
    if (int rc = download_dofs(ctx, ctx->u, u)) return rc;
    if (bd) return bd;
    return std::sqrt(h.rr) <= h.tol ? 0 : TOHIP_NOT_CONVERGED;
}

int tohip_mg_apply(tohip_ctx* ctx, const double* r, double* z) {
    if (!ctx || !r || !z) return TOHIP_ERR_ARG;
    if (!all_finite(r, ctx->g.ndof)) {
        ctx->err = "tohip_mg_apply: non-finite input";
        return TOHIP_ERR_ARG;
    }
    if (int rc = mg_setup(ctx)) return rc;
    if (int rc = upload_dofs(ctx, r, ctx->r64)) return rc;
    CK(launch_f64_to_f32(ctx->g.ndof, ctx->r64, ctx->b32, ctx->stream));
    if (int rc = mg_join(ctx)) return rc;
    CK(launch_vcycle(ctx, ctx->b32, nullptr, nullptr, nullptr));
    std::vector<float> zf(ctx->g.ndof);
    CK(hipMemcpyAsync(zf.data(), ctx->mg[0].rz, zf.size() * sizeof(float),
                      hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    std::vector<double> zd(zf.begin(), zf.end());
    to_caller(ctx, zd, z);
    return 0;
}

int tohip_mg_weights(tohip_ctx* ctx, double* omega, double* lambda_est, int max_levels,
                     int* nlevels) {
    if (!ctx || !nlevels || max_levels < 0) return TOHIP_ERR_ARG;
    const int nl = (int)ctx->mg.size();
    *nlevels = nl;
    if (nl == 0) return 0;
    std::vector<float> hw(4 * nl + 1);
    CK(hipMemcpyAsync(hw.data(), ctx->mg_wd, hw.size() * sizeof(float), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    for (int l = 0; l < std::min(nl, max_levels); ++l) {
        if (omega) omega[l] = hw[4 * l];
        if (lambda_est) lambda_est[l] = hw[4 * l + 2];
    }
    return 0;
}

int tohip_solve(tohip_ctx* ctx, const double* f, double* u, double abstol, double reltol,
                int64_t maxiter, int initially_zero, int precision, int64_t* iters_out,
                double* resid_out) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (!(abstol >= 0) || !(reltol >= 0) || maxiter < 0 ||
        !valid_precision(precision)) {
        ctx->err = "tohip_solve: bad tolerance, maxiter or precision";
        return TOHIP_ERR_ARG;
    }
    if ((f && !all_finite(f, ctx->g.ndof)) ||
        (u && !initially_zero && !all_finite(u, ctx->g.ndof))) {
        ctx->err = "tohip_solve: non-finite load or initial guess";
        return TOHIP_ERR_ARG;
    }
    if (!f && !ctx->have_load) {
        ctx->err = "tohip_solve: f = NULL needs a load set by tohip_set_load";
        return TOHIP_ERR_STATE;
    }
    if (!u && !initially_zero && !ctx->have_u) {
        ctx->err = "tohip_solve: u = NULL with initially_zero = 0 needs a previous solve";
        return TOHIP_ERR_STATE;
    }
    if (f) {
        if (int rc = upload_dofs(ctx, f, ctx->b64)) return rc;
    } else {
        CK(hipMemcpyAsync(ctx->b64, ctx->f, ctx->g.ndof * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (!initially_zero && u)
        if (int rc = upload_dofs(ctx, u, ctx->u)) return rc;
    int rc = solve_mode(ctx, precision, abstol, reltol, maxiter, initially_zero != 0, 0.0,
                        reltol, iters_out, resid_out);
    if (rc < 0) return rc;
    ctx->have_u = true;
    if (u)
        if (int rc2 = download_dofs(ctx, ctx->u, u)) return rc2;
    return rc;
}

int tohip_compliance(tohip_ctx* ctx, const double* u, double* obj, double* cell_comp,
                     double* grad) {
    if (!ctx) return TOHIP_ERR_ARG;
    const double* ud = ctx->u;  // u = NULL: the displacement the last solve left on the device
    if (u) {
        if (int rc = upload_dofs(ctx, u, ctx->t64)) return rc;
        ud = ctx->t64;
    } else if (!ctx->have_u) {
        ctx->err = "tohip_compliance: u = NULL needs a previous solve";
        return TOHIP_ERR_STATE;
    }
    CK(launch_energy(ctx->g, ctx->op64, ud, ctx->xt, ctx->xmin, ctx->pen, ctx->pbi, ctx->ce,
                     ctx->grad, ctx->partials, ctx->dscal, ctx->stream));
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
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, ctx->cellbuf,
                     ctx->xnew, transpose ? 1 : 0, ctx->stream));
    return download_cells(ctx, ctx->xnew, y);
}

namespace {
bool valid_projection(int proj, double beta) {
    return proj == TOHIP_PROJ_NONE ||
           ((proj == TOHIP_PROJ_HEAVISIDE || proj == TOHIP_PROJ_SIGMOID) && beta >= 0 &&
            std::isfinite(beta));
}
}  // namespace

int tohip_filter_apply_projected(tohip_ctx* ctx, const double* x, double* y, int pre_proj,
                                 double pre_beta, int post_proj, double post_beta) {
    if (!ctx || !x || !y) return TOHIP_ERR_ARG;
    if (!ctx->f_offs) {
        ctx->err = "tohip_filter_apply_projected: call tohip_filter_setup first";
        return TOHIP_ERR_STATE;
    }
    if (!valid_projection(pre_proj, pre_beta) || !valid_projection(post_proj, post_beta)) {
        ctx->err = "tohip_filter_apply_projected: unknown projection or bad steepness";
        return TOHIP_ERR_ARG;
    }
    const int64_t n = ctx->g.ncells;
    if (int rc = upload_cells(ctx, x, ctx->cellbuf)) return rc;
    const double* t = ctx->cellbuf;
    if (pre_proj != TOHIP_PROJ_NONE) {
        CK(launch_project(n, pre_proj, pre_beta, ctx->cellbuf, nullptr, ctx->ce, ctx->stream));
        t = ctx->ce;
    }
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, t, ctx->grad,
                     0, ctx->stream));
    const double* out = ctx->grad;
    if (post_proj != TOHIP_PROJ_NONE) {
        CK(launch_project(n, post_proj, post_beta, ctx->grad, nullptr, ctx->xnew, ctx->stream));
        out = ctx->xnew;
    }
    return download_cells(ctx, out, y);
}

int tohip_filter_pullback_projected(tohip_ctx* ctx, const double* x, const double* dy,
                                    double* dx, int pre_proj, double pre_beta, int post_proj,
                                    double post_beta) {
    if (!ctx || !x || !dy || !dx) return TOHIP_ERR_ARG;
    if (!ctx->f_offs) {
        ctx->err = "tohip_filter_pullback_projected: call tohip_filter_setup first";
        return TOHIP_ERR_STATE;
    }
    if (!valid_projection(pre_proj, pre_beta) || !valid_projection(post_proj, post_beta)) {
        ctx->err = "tohip_filter_pullback_projected: unknown projection or bad steepness";
        return TOHIP_ERR_ARG;
    }
    const int64_t n = ctx->g.ncells;
    if (int rc = upload_cells(ctx, x, ctx->cellbuf)) return rc;
    if (int rc = upload_cells(ctx, dy, ctx->dc)) return rc;
    const double* g = ctx->dc;  // cotangent of the filter output
    if (post_proj != TOHIP_PROJ_NONE) {
        // forward value at the filter output: post'(J pre(x)) * dy
        const double* t = ctx->cellbuf;
        if (pre_proj != TOHIP_PROJ_NONE) {
            CK(launch_project(n, pre_proj, pre_beta, ctx->cellbuf, nullptr, ctx->ce, ctx->stream));
            t = ctx->ce;
        }
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, t,
                         ctx->grad, 0, ctx->stream));
        CK(launch_project(n, post_proj, post_beta, ctx->grad, ctx->dc, ctx->xnew, ctx->stream));
        g = ctx->xnew;
    }
    CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, g, ctx->grad, 1,
                     ctx->stream));
    const double* out = ctx->grad;
    if (pre_proj != TOHIP_PROJ_NONE) {
        CK(launch_project(n, pre_proj, pre_beta, ctx->cellbuf, ctx->grad, ctx->xnew, ctx->stream));
        out = ctx->xnew;
    }
    return download_cells(ctx, out, dx);
}

int tohip_volume(tohip_ctx* ctx, const double* x, double* vol) {
    if (!ctx || !x || !vol) return TOHIP_ERR_ARG;
    if (int rc = upload_cells(ctx, x, ctx->cellbuf)) return rc;
    CK(launch_sum(ctx->g.ncells, ctx->cellbuf, ctx->partials, ctx->dscal + 2, ctx->stream));
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
    if (!all_finite(f, ctx->g.ndof)) {
        ctx->err = "tohip_set_load: non-finite load";
        return TOHIP_ERR_ARG;
    }
    std::vector<double> tmp;
    to_internal(ctx, f, tmp);
    const int dim = ctx->g.dim;
    for (int64_t n = 0; n < ctx->g.nnodes; ++n)  // apply_zero! on the prescribed DOFs
        for (int d = 0; d < dim; ++d)
            if ((ctx->nodemask[n] >> d) & 1) tmp[dim * n + d] = 0.0;
    CK(hipMemcpyAsync(ctx->f, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->have_load = true;
    return 0;
}

int tohip_get_u(tohip_ctx* ctx, double* u) {
    if (!ctx || !u) return TOHIP_ERR_ARG;
    return download_dofs(ctx, ctx->u, u);
}

int tohip_get_xphys(tohip_ctx* ctx, double* xphys) {
    if (!ctx || !xphys) return TOHIP_ERR_ARG;
    return download_cells(ctx, ctx->xt, xphys);
}

}  // extern "C"

namespace {

// Phase timing of the SIMP loop (TOHIP_SIMP_TIMING=1, measurement only): GPU
// time between events recorded at the phase boundaries, host wall time and
// host round trips, averaged over the run and printed to stderr at its end.
struct SimpTiming {
    bool on = false;
    hipEvent_t ev[4] = {};
    double gpu_ms[3] = {0, 0, 0};  // filter + operator, solve, energy + pullback + update
    double wall_ms = 0;
    int64_t n = 0;
    int64_t syncs0 = 0, graphs0 = 0;
    double graph_ms0 = 0;
    std::chrono::steady_clock::time_point t0;
    explicit SimpTiming(tohip_ctx* ctx) {
        const char* e = getenv("TOHIP_SIMP_TIMING");
        on = e && atoi(e) == 1;
        if (on)
            for (hipEvent_t& v : ev)
                if (hipEventCreate(&v) != hipSuccess) on = false;
        syncs0 = ctx->host_syncs;
        if (on && !ctx->time_graphs) {
            ctx->time_graphs = hipEventCreate(&ctx->graph_ev[0]) == hipSuccess &&
                               hipEventCreate(&ctx->graph_ev[1]) == hipSuccess;
        }
        graph_ms0 = ctx->graph_ms;
        graphs0 = ctx->graph_launches;
    }
    void mark(int i, hipStream_t s) {
        if (!on) return;
        if (i == 0) t0 = std::chrono::steady_clock::now();
        hipEventRecord(ev[i], s);
    }
    void done_iteration() {  // after the iteration's stream synchronisation
        if (!on) return;
        for (int i = 0; i < 3; ++i) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            gpu_ms[i] += ms;
        }
        wall_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                       .count();
        ++n;
    }
    void report(const tohip_ctx* ctx) {
        if (!on || n == 0) return;
        fprintf(stderr,
                "[simp timing] %lld iterations, per iteration: wall %.3f ms; GPU between "
                "events: filter+operator %.3f ms, solve %.3f ms, energy+pullback+update %.3f "
                "ms; host round trips %.1f; PCG graphs %.1f taking %.3f ms on the GPU\n",
                (long long)n, wall_ms / n, gpu_ms[0] / n, gpu_ms[1] / n, gpu_ms[2] / n,
                double(ctx->host_syncs - syncs0) / n, double(ctx->graph_launches - graphs0) / n,
                (ctx->graph_ms - graph_ms0) / n);
    }
    ~SimpTiming() {
        for (hipEvent_t v : ev)
            if (v) hipEventDestroy(v);
    }
};

// The SIMP loop of tohip_simp_run (OC update) and tohip_simp_run_mma (MMA87).
int simp_run_impl(tohip_ctx* ctx, const tohip_simp_opts* o, const tohip_mma_opts* mma,
                  double* x_inout, double* obj_hist, int64_t* iters_hist, double* change_hist) {
    if (!ctx || !o) return TOHIP_ERR_ARG;
    if (o->iterations < 0 || !(o->volfrac > 0 && o->volfrac <= 1) || !(o->move > 0) ||
        !(o->oc_tol > 0) || !(o->rel_tol > 0) || o->maxiter < 1 ||
        !valid_precision(o->precision) || o->warm_start < 0 || o->warm_start > 2 ||
        ((o->precision == TOHIP_MIXED || o->precision == TOHIP_MIXED_MG) &&
         !(o->inner_rel_tol > 0))) {
        ctx->err = "tohip_simp_run: bad options";
        return TOHIP_ERR_ARG;
    }
    if (o->warm_start == 2 && !ctx->nn) {
        ctx->err = "tohip_simp_run: warm_start = 2 needs tohip_nn_attach";
        return TOHIP_ERR_STATE;
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
    if (x_inout) {  // a new run: no previous displacement, the network consulted afresh
        if (int rc = upload_cells(ctx, x_inout, ctx->xd)) return rc;
        ctx->have_u = false;
        ctx->nn_calls = 0;
        ctx->nn_misses = 0;
        ctx->nn_off = false;
    }
    if (mma && !ctx->mma_st) {
        int rc;
        if ((rc = dalloc(ctx, &ctx->mma_L, nC)) || (rc = dalloc(ctx, &ctx->mma_U, nC)) ||
            (rc = dalloc(ctx, &ctx->mma_x1, nC)) || (rc = dalloc(ctx, &ctx->mma_x2, nC)) ||
            (rc = dalloc(ctx, &ctx->mma_st, 1)))
            return rc;
        CK(hipHostMalloc((void**)&ctx->h_mma, sizeof(MmaState), hipHostMallocDefault));
        ctx->mma_iter = 0;
    }
    if (mma && (x_inout || ctx->mma_iter == 0)) {
        // a new MMA run: the asymptote schedule and the design history restart;
        // a resident continuation (x_inout = NULL) keeps both
        ctx->mma_iter = 0;
        CK(hipMemsetAsync(ctx->mma_x1, 0, nC * sizeof(double), ctx->stream));
        CK(hipMemsetAsync(ctx->mma_x2, 0, nC * sizeof(double), ctx->stream));
    }
    const MmaParams mprm = mma ? MmaParams{mma->s_init, mma->s_incr, mma->s_decr}
                               : MmaParams{0.5, 1.2, 0.7};
    int final_rc = 0;
    int64_t prev_its = 0;  // PCG iterations of the previous solve of this call
    SimpTiming timing(ctx);
    for (int64_t it = 0; it < o->iterations; ++it) {
        timing.mark(0, ctx->stream);
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, ctx->xd,
                         ctx->xt, 0, ctx->stream));
        if (int rc = refresh_operator(ctx)) return rc;
        timing.mark(1, ctx->stream);
        bool zero = o->warm_start == 0 || !ctx->have_u;
        // warm_start = 2: the network's prediction combined with the previous
        // displacement, until it stops paying (then the previous one alone)
        const bool use_net = o->warm_start == 2 && !ctx->nn_off;
        const bool rate_net = use_net && ctx->have_u;
        if (use_net) {
            if (int rc = nn_warm_start(ctx, ctx->have_u)) return rc;
            ++ctx->nn_calls;
            zero = false;
        }
        int64_t its = 0;
        double res = 0;
        int rc = solve_mode(ctx, o->precision, abstol, 0.0, o->maxiter, zero, o->inner_rel_tol,
                            o->rel_tol, &its, &res);
        if (rc < 0) return rc;
        if (rc == TOHIP_BREAKDOWN) {  // NaN displacement: stop at this iteration
            if (obj_hist) obj_hist[it] = std::numeric_limits<double>::quiet_NaN();
            if (iters_hist) iters_hist[it] = its;
            final_rc = rc;
            break;
        }
        if (rc > 0) final_rc = rc;
        ctx->have_u = true;
        timing.mark(2, ctx->stream);
        CK(launch_energy(ctx->g, ctx->op64, ctx->u, ctx->xt, ctx->xmin, ctx->pen, ctx->pbi,
                         ctx->ce, ctx->grad, ctx->partials, ctx->dscal, ctx->stream));
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, ctx->grad,
                         ctx->dc, 1, ctx->stream));
        if (mma) {
            CK(launch_mma(nC, ctx->mma_iter++, ctx->xd, ctx->dc, ctx->f_colsum, o->volfrac,
                          o->oc_tol, mprm,
                          ctx->mma_L, ctx->mma_U, ctx->mma_x1, ctx->mma_x2, ctx->xnew,
                          ctx->mma_st, ctx->partials, ctx->stream));
            CK(hipMemcpyAsync(ctx->h_mma, ctx->mma_st, sizeof(MmaState), hipMemcpyDeviceToHost,
                              ctx->stream));
        } else {
            CK(launch_oc(nC, ctx->xd, ctx->dc, ctx->f_colsum, o->volfrac, o->move, o->oc_tol,
                         16, ctx->xnew, ctx->oc, ctx->partials, ctx->stream));
            CK(hipMemcpyAsync(ctx->h_oc, ctx->oc, sizeof(OcState), hipMemcpyDeviceToHost,
                              ctx->stream));
        }
        std::swap(ctx->xd, ctx->xnew);
        timing.mark(3, ctx->stream);
        CK(hipMemcpyAsync(ctx->h_scal, ctx->dscal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
        if (rate_net)
            CK(hipMemcpyAsync(ctx->h_scal + 3, ctx->dscal + kWsRatio, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        timing.done_iteration();
        if (rate_net) {
            const bool blowup = prev_its > 0 && its > kWsBlowup * prev_its;
            ctx->nn_misses = (ctx->h_scal[3] > kWsPayRatio || blowup) ? ctx->nn_misses + 1 : 0;
            if (blowup || ctx->nn_misses >= kWsMaxMisses) ctx->nn_off = true;
        }
        prev_its = its;
        if (obj_hist) obj_hist[it] = ctx->h_scal[0];
        if (iters_hist) iters_hist[it] = its;
        if (change_hist) change_hist[it] = mma ? ctx->h_mma->change : ctx->h_oc->change;
    }
    timing.report(ctx);
    if (x_inout)
        if (int rc = download_cells(ctx, ctx->xd, x_inout)) return rc;
    return final_rc;
}

}  // namespace

extern "C" {

int tohip_simp_run(tohip_ctx* ctx, const tohip_simp_opts* o, double* x_inout, double* obj_hist,
                   int64_t* iters_hist, double* change_hist) {
    return simp_run_impl(ctx, o, nullptr, x_inout, obj_hist, iters_hist, change_hist);
}

int tohip_simp_run_mma(tohip_ctx* ctx, const tohip_simp_opts* o, const tohip_mma_opts* m,
                       double* x_inout, double* obj_hist, int64_t* iters_hist,
                       double* change_hist) {
    if (!ctx || !o || !m) return TOHIP_ERR_ARG;
    if (!(m->s_init > 0 && m->s_init <= 1) || !(m->s_decr > 0 && m->s_decr <= 1) ||
        !(m->s_incr >= 1 && m->s_incr < 10)) {
        ctx->err = "tohip_simp_run_mma: need 0 < s_init <= 1, 0 < s_decr <= 1 <= s_incr < 10";
        return TOHIP_ERR_ARG;
    }
    return simp_run_impl(ctx, o, m, x_inout, obj_hist, iters_hist, change_hist);
}

int tohip_beso_run(tohip_ctx* ctx, const tohip_beso_opts* o, const double* x0,
                   const uint8_t* passive, double* topology_out, tohip_beso_result* res,
                   double* obj_hist) {
    if (!ctx || !o || !x0 || !res) return TOHIP_ERR_ARG;
    // sens_tol = 0 would make the bisection loop of beso.jl:158 run forever
    if (o->maxiter < 0 || !(o->vol_limit >= 0 && o->vol_limit <= 1) || !(o->er >= 0) ||
        !(o->tol >= 0) || !(o->sens_tol > 0) || o->k < 2 || o->k > 4096 ||
        !(o->penalty_p > 0) || !(o->rel_tol > 0) || o->pcg_maxiter < 1 ||
        !valid_precision(o->precision)) {
        ctx->err = "tohip_beso_run: bad options";
        return TOHIP_ERR_ARG;
    }
    const int64_t nD = ctx->g.ndof, nC = ctx->g.ncells;
    if (!all_finite(x0, nC)) {
        ctx->err = "tohip_beso_run: non-finite initial design";
        return TOHIP_ERR_ARG;
    }
    if (int rc = build_filter(ctx, o->rmin)) return rc;
    std::vector<double> fh(nD);
    CK(hipMemcpyAsync(fh.data(), ctx->f, nD * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    const double fnorm = host_norm(fh.data(), nD);
    if (!(fnorm > 0)) {
        ctx->err = "tohip_beso_run: the load vector is zero (tohip_set_load)";
        return TOHIP_ERR_STATE;
    }
    if (!ctx->beso_st) {
        int rc = 0;
        if ((rc = dalloc(ctx, &ctx->beso_sens[0], nC)) || (rc = dalloc(ctx, &ctx->beso_sens[1], nC)) ||
            (rc = dalloc(ctx, &ctx->beso_mask, nC)) || (rc = dalloc(ctx, &ctx->beso_st, 1)))
            return rc;
        CK(hipHostMalloc((void**)&ctx->h_beso, sizeof(BesoState), hipHostMallocDefault));
    }
    // setpenalty!(solver, b.penalty.p) (beso.jl:136-141); the family is kept
    ctx->pen.p = o->penalty_p;
    // topology = round(x0) (ties to even), then passive cells: 1 solid, 2 void
    std::vector<double> topo(nC);
    std::vector<uint8_t> mask(nC, 0);
    double nblack = 0;
    for (int64_t e = 0; e < nC; ++e) {
        topo[e] = std::nearbyint(x0[e]);
        if (passive) {
            if (passive[e] > 2) {
                ctx->err = "tohip_beso_run: passive flags are 0 (free), 1 (solid) or 2 (void)";
                return TOHIP_ERR_ARG;
            }
            mask[e] = passive[e];
            if (passive[e] == 1) topo[e] = 1.0, nblack += 1;
            if (passive[e] == 2) topo[e] = 0.0;
        }
    }
    CK(hipMemcpyAsync(ctx->xd, topo.data(), nC * sizeof(double), hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipMemcpyAsync(ctx->beso_mask, mask.data(), nC, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->b64, ctx->f, nD * sizeof(double), hipMemcpyDeviceToDevice,
                      ctx->stream));
    const double cv = ctx->h[0] * ctx->h[1] * (ctx->g.dim == 3 ? ctx->h[2] : 1.0);
    double total = 0;  // sum(cellvolumes), summed as the reference does
    for (int64_t e = 0; e < nC; ++e) total += cv;
    double dot = 0;
    for (int64_t e = 0; e < nC; ++e) dot += topo[e] * cv;
    double vol = dot / total, true_vol = vol;
    const double V = o->vol_limit;
    const int k = o->k;
    std::vector<double> trace(k, 0.0);
    // passes of 6 bisection steps covering (l2 - l1) / l2 <= 1 down to sens_tol
    const int passes =
        std::max(1, (int)std::ceil((std::log2(1.0 / o->sens_tol) + 2) / kBesoLevels));
    const double abstol = o->rel_tol * fnorm;
    double change = 1.0;
    int64_t iter = 0;
    int which = 0;
    int final_rc = 0;
    ctx->have_u = false;
    while ((change > o->tol || true_vol > V) && iter < o->maxiter) {
        ++iter;
        vol = std::max(vol * (1.0 - o->er), V);
        for (int j = std::max(2, k - (int)std::min<int64_t>(iter, k) + 2); j <= k; ++j)
            trace[j - 2] = trace[j - 1];
        // obj = comp(filter(vars)) and its pullback to vars (x~ = J vars)
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, ctx->xd,
                         ctx->xt, 0, ctx->stream));
        if (int rc = refresh_operator(ctx)) return rc;
        int64_t its = 0;
        double r = 0;
        int rc = solve_mode(ctx, o->precision, abstol, 0.0, o->pcg_maxiter, !ctx->have_u, 0.0,
                            o->rel_tol, &its, &r);
        if (rc < 0 || rc == TOHIP_BREAKDOWN) return rc;
        if (rc > 0) final_rc = rc;
        ctx->have_u = true;
        CK(launch_energy(ctx->g, ctx->op64, ctx->u, ctx->xt, ctx->xmin, ctx->pen, ctx->pbi,
                         ctx->ce, ctx->grad, ctx->partials, ctx->dscal, ctx->stream));
        CK(launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo, ctx->f_invden, ctx->grad,
                         ctx->dc, 1, ctx->stream));
        CK(hipMemcpyAsync(ctx->h_scal, ctx->dscal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
        // sens, threshold bisection, topology
        BesoState hs{};
        hs.sens_tol = o->sens_tol;
        hs.vol = vol;
        hs.cv = cv;
        hs.total = total;
        hs.nblack = nblack;
        *ctx->h_beso = hs;
        CK(hipMemcpyAsync(ctx->beso_st, ctx->h_beso, sizeof(BesoState), hipMemcpyHostToDevice,
                          ctx->stream));
        double* sens = ctx->beso_sens[which];
        const double* prev = ctx->beso_sens[1 - which];
        which = 1 - which;
        CK(launch_beso_update(nC, ctx->dc, prev, iter > 1 ? 1 : 0, sens,
                              passive ? ctx->beso_mask : nullptr, ctx->beso_st, ctx->xd,
                              ctx->partials, passes, ctx->stream));
        for (int extra = 0; extra < 64; ++extra) {  // bracket wider than (l2 - l1) / l2 <= 1
            CK(hipMemcpyAsync(ctx->h_beso, ctx->beso_st, sizeof(BesoState),
                              hipMemcpyDeviceToHost, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
            if (ctx->h_beso->done) break;
            CK(launch_beso_passes(nC, sens, passive ? ctx->beso_mask : nullptr, ctx->beso_st,
                                  ctx->partials, passes, ctx->stream));
        }
        if (!ctx->h_beso->done) {
            ctx->err = "tohip_beso_run: the threshold bisection did not terminate";
            return TOHIP_ERR_STATE;
        }
        CK(launch_beso_apply(nC, sens, passive ? ctx->beso_mask : nullptr, ctx->beso_st, ctx->xd,
                             ctx->partials, ctx->dscal + 2, ctx->stream));
        CK(hipMemcpyAsync(ctx->h_scal + 2, ctx->dscal + 2, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        trace[k - 1] = ctx->h_scal[0];
        if (obj_hist) obj_hist[iter - 1] = ctx->h_scal[0];
        true_vol = ctx->h_scal[2] * cv / total;
        if (iter >= k) {
            double l = 0, h = 0;
            for (int j = 0; j < k / 2; ++j) l += trace[j];
            for (int j = k / 2; j < k; ++j) h += trace[j];
            change = std::fabs(l - h) / h;
        }
    }
    res->objval = trace[k - 1];
    res->change = change;
    res->converged = change <= o->tol ? 1 : 0;
    res->fevals = iter;
    res->volume = true_vol;
    if (topology_out)
        if (int rc = download_cells(ctx, ctx->xd, topology_out)) return rc;
    return final_rc;
}

int tohip_bench_kernel(tohip_ctx* ctx, int kernel, int reps, double* avg_us) {
    if (!ctx || !avg_us || reps < 1 || kernel < 0 || kernel > TOHIP_K_CSR_SPMV)
        return TOHIP_ERR_ARG;
    if (kernel >= TOHIP_K_MG_PRE && kernel <= TOHIP_K_VCYCLE)
        if (int rc = mg_setup(ctx)) return rc;
    if (kernel == TOHIP_K_CSR_SPMV)
        if (int rc = csr_assemble(ctx, 0, 1)) return rc;
    const int64_t nD = ctx->g.ndof;
    hipStream_t s = ctx->stream;
    // FP32 operands from the FP64 ones; PCG scalars in a state that runs the
    // full kernel (done = 0, beta = 0, rho = 1).
    CK(launch_f64_to_f32(nD, ctx->r64, ctx->r32, s));
    CK(launch_f64_to_f32(nD, ctx->t64, ctx->q32, s));
    CK(launch_f64_to_f32(nD, ctx->t64, ctx->x32, s));
    PcgScalars hs{};
    hs.rho[0] = hs.rho[1] = 1.0;
    hs.rho_cur = 1.0;
    hs.tol = 0.0;
    hs.iter = 0;
    hs.k_cur = 1;
    hs.maxiter = 1LL << 60;
    *ctx->h_sc = hs;
    CK(hipMemcpyAsync(ctx->sc64, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(ctx->sc32, ctx->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    CK(hipMemsetAsync(ctx->p64[0], 0, nD * sizeof(double), s));
    CK(hipMemsetAsync(ctx->p32[0], 0, nD * sizeof(float), s));
    // PCG kernels: partials of the current r (start kernel keeping the scalars),
    // then one direction launch so the update kernel has p.q to read.
    if (kernel == TOHIP_K_PCG64 || kernel == TOHIP_K_UPDATE64) {
        CK(launch_pcg_start<double>(nD, ctx->r64, nullptr, ctx->r64, ctx->dinv64, ctx->p64[0],
                                    ctx->sc64, ctx->partB, 1, s));
        CK(launch_pcg_dirmatvec<double>(ctx->g, ctx->op64, ctx->r64, ctx->dinv64, ctx->p64[0],
                                        ctx->p64[1], ctx->q64, ctx->fixmask, ctx->E64,
                                        ctx->meandiag, ctx->sc64, ctx->partB, kVecBlocks,
                                        ctx->partA, nullptr, s));
    }
    if (kernel == TOHIP_K_PCG32 || kernel == TOHIP_K_UPDATE32) {
        CK(launch_pcg_start<float>(nD, ctx->r32, nullptr, ctx->r32, ctx->dinv32, ctx->p32[0],
                                   ctx->sc32, ctx->partB, 1, s));
        CK(launch_pcg_dirmatvec<float>(ctx->g, ctx->op32, ctx->r32, ctx->dinv32, ctx->p32[0],
                                       ctx->p32[1], ctx->q32, ctx->fixmask, ctx->E32,
                                       ctx->meandiag, ctx->sc32, ctx->partB, kVecBlocks,
                                       ctx->partA, nullptr, s));
    }
    if ((kernel == TOHIP_K_FILTER || kernel == TOHIP_K_FILTER_T || kernel == TOHIP_K_OC) &&
        !ctx->f_offs) {
        ctx->err = "tohip_bench_kernel: set up the filter first";
        return TOHIP_ERR_STATE;
    }
    if (kernel >= TOHIP_K_MG_PRE) {
        // operands of the single-kernel timings: one full V-cycle on r32
        // leaves every level's x, rz and b filled
        if (int rc = mg_join(ctx)) return rc;
        CK(launch_vcycle(ctx, ctx->r32, nullptr, nullptr, nullptr));
    }
    if (kernel == TOHIP_K_OC) {
        CK(hipMemcpyAsync(ctx->xd, ctx->xt, ctx->g.ncells * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
        CK(launch_energy(ctx->g, ctx->op64, ctx->r64, ctx->xt, ctx->xmin, ctx->pen, ctx->pbi,
                         ctx->ce, ctx->dc, ctx->partials, ctx->dscal, s));
    }
    auto launch = [&]() -> hipError_t {
        switch (kernel) {
            case TOHIP_K_MATVEC64:
                return launch_matvec<double>(ctx->g, ctx->op64, ctx->r64, ctx->fixmask, ctx->E64,
                                             ctx->t64, ctx->meandiag, s);
            case TOHIP_K_MATVEC32:
                return launch_matvec<float>(ctx->g, ctx->op32, ctx->r32, ctx->fixmask, ctx->E32,
                                            ctx->q32, ctx->meandiag, s);
            case TOHIP_K_PCG64:
                return launch_pcg_dirmatvec<double>(ctx->g, ctx->op64, ctx->r64, ctx->dinv64,
                                                    ctx->p64[0], ctx->p64[1], ctx->q64,
                                                    ctx->fixmask, ctx->E64, ctx->meandiag,
                                                    ctx->sc64, ctx->partB, kVecBlocks,
                                                    ctx->partA, nullptr, s);
            case TOHIP_K_PCG32:
                return launch_pcg_dirmatvec<float>(ctx->g, ctx->op32, ctx->r32, ctx->dinv32,
                                                   ctx->p32[0], ctx->p32[1], ctx->q32,
                                                   ctx->fixmask, ctx->E32, ctx->meandiag,
                                                   ctx->sc32, ctx->partB, kVecBlocks,
                                                   ctx->partA, nullptr, s);
            case TOHIP_K_UPDATE64:
                return launch_pcg_update<double>(nD, ctx->b64, ctx->u, ctx->p64[0], ctx->p64[1],
                                                 ctx->t64, ctx->dinv64, ctx->sc64, ctx->partA,
                                                 pcg_dirmatvec_blocks(ctx->g), ctx->partB, s);
            case TOHIP_K_UPDATE32:
                return launch_pcg_update<float>(nD, ctx->x32, ctx->b32, ctx->p32[0], ctx->p32[1],
                                                ctx->q32, ctx->dinv32, ctx->sc32, ctx->partA,
                                                pcg_dirmatvec_blocks32(ctx->g), ctx->partB, s);
            case TOHIP_K_FILTER:
                return launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo,
                                     ctx->f_invden, ctx->xt, ctx->cellbuf, 0, s);
            case TOHIP_K_FILTER_T:
                return launch_filter(ctx->g, ctx->f_offs, ctx->f_wts, ctx->f_count, ctx->f_halo,
                                     ctx->f_invden, ctx->xt, ctx->cellbuf, 1, s);
            case TOHIP_K_ENERGY:
                return launch_energy(ctx->g, ctx->op64, ctx->r64, ctx->xt, ctx->xmin, ctx->pen,
                                     ctx->pbi, ctx->ce, ctx->grad, ctx->partials, ctx->dscal, s);
            case TOHIP_K_COPY64:
                return launch_copy64(nD, ctx->r64, ctx->t64, s);
            case TOHIP_K_MG_PRE: {
                tohip_mg_level& L0 = ctx->mg[0];
                return launch_mg_smooth(0, L0.view(), L0.omega, ctx->r32, nullptr, nullptr,
                                        nullptr, L0.x, ctx->q32, ctx->meandiag, nullptr, nullptr,
                                        s);
            }
            case TOHIP_K_MG_RESTRICT:
                return launch_mg_restrict(ctx->mg[0].view(), ctx->mg[1].view(), ctx->mg[0].rz,
                                          ctx->mg[1].b, nullptr, s);
            case TOHIP_K_MG_COARSE: {
                tohip_mg_level& C = ctx->mg.back();
                if (!mg_coarse_fits(C.g)) return hipErrorInvalidValue;
                if (ctx->mg_M)
                    return launch_mg_dense((int)C.g.ndof, ctx->mg_M, C.b, C.x,
                                           nullptr, s);
                return launch_mg_coarse(C.view(), ctx->mg_cheb, C.b, C.x, nullptr, s);
            }
            case TOHIP_K_MG_POST: {
                tohip_mg_level& L0 = ctx->mg[0];
                const MgLevelView C = ctx->mg[1].view();
                return launch_mg_smooth(1, L0.view(), L0.omega, ctx->r32, L0.x, &C, ctx->mg[1].rz,
                                        ctx->q32, nullptr, ctx->meandiag, nullptr, ctx->partB, s);
            }
            case TOHIP_K_VCYCLE:
                return launch_vcycle(ctx, ctx->r32, nullptr, nullptr, nullptr);
            case TOHIP_K_CSR_SPMV:
                return launch_csr_spmv(ctx->g.ndof, ctx->csr_rowptr, ctx->csr_cols,
                                       ctx->csr_vals, ctx->r64, ctx->q64, s);
            default:
                return launch_oc(ctx->g.ncells, ctx->xd, ctx->dc, ctx->f_colsum, 0.3, 0.2, 1e-12,
                                 16, ctx->xnew, ctx->oc, ctx->partials, s);
        }
    };
    CK(launch());  // warm-up
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    CK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) CK(launch());
    CK(hipEventRecord(e1, s));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *avg_us = 1000.0 * (double)ms / reps;
    return 0;
}

}  // extern "C"

