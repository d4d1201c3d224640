// Internal declarations shared by the libtopopt_hip translation units.
//
// Device layout (all vectors node-major, components interleaved):
//   node n = i + NX*(j + NY*k), NX = nx+1, ...; DOF index = DIM*n + d.
//   cells  e = i + nx*(j + ny*k)  (same x-fastest order as the caller's).
//   fixmask[n] bit d set  <=>  DOF (n, d) is prescribed (homogeneous Dirichlet).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tohip {

// Structured grid sizes (cells nx, ny, nz; nodes NX = nx+1, ...).
struct Grid {
    int dim;
    int nx, ny, nz;
    int NX, NY, NZ;
    int64_t nnodes, ncells, ndof;
};

// Element operator of one Q1 cell in the corner-Hadamard (modal) basis.
// For corner bit index c = bx + 2*by (+ 4*bz) and modal index m with the same
// bit meaning, H[c][m] = prod_{a in m} (2*bit_a(c) - 1).  The cell stiffness
// satisfies Ke = H B H^T where B couples (d, m) only with (d', m') of equal
// label L = m ^ (1<<d) (B is divided by (number of corners)^2 here).  On any
// axis-aligned box the label blocks have a fixed zero pattern (rigid
// translations are the m = 0 members), so only 33 numbers (3D) are stored:
//   L0, L7         symmetric 3x3 (b00 b01 b02 b11 b12 b22)
//   L1, L2, L4     2x2 on the two non-rigid members (b_aa b_ab b_bb)
//   L3, L5, L6     2x2 pair + one decoupled diagonal (b_aa b_ab b_bb b_cc)
// 2D: L0, L3 full 2x2, L1 and L2 one diagonal entry each.
// On a cube the symmetry group ties these to 8 numbers (cube[]):
//   L0 = (A-B) I + B 11^T, L7 = (G-H) I + H 11^T, L1/2/4 pairs [[C D][D C]],
//   L3/5/6 pairs E 11^T with single F; cube[] = {A-B, B, C, D, E, F, G-H, H}.
template <typename T>
struct ElemOp {
    T l0[6], l7[6];
    T pr[3][3];  // labels 1, 2, 4
    T qs[3][4];  // labels 3, 5, 6
    T q2[8];     // 2D: L0 (b00 b01 b11), L3 (b00 b01 b11), L1 b11, L2 b00
    T cube[8];   // 3D cube form (valid when is_cube)
    T kdiag[3];  // Ke diagonal per component (identical for every corner)
    int is_cube;
};

// SIMP penalty family (TopOpt.jl src/Utilities/penalties.jl:30-101), applied
// to the filtered density x (optionally after a projection y = proj(x)):
//   kind 0 power    y^p                  proj 0 none
//   kind 1 rational y / (1 + p (1 - y))  proj 1 Heaviside 1 - exp(-b x) + x exp(-b)
//   kind 2 sinh     sinh(p y) / sinh(p)  proj 2 sigmoid   1 / (1 + exp((b + 1)(0.5 - x)))
struct Penalty {
    int kind;
    int proj;
    double p;
    double beta;
};

// Density filter stencil table limit (node offsets + weights).
constexpr int kMaxStencil = 1024;

// Device-resident PCG state (FP64 scalars).  Writers: the direction kernel's
// block 0 (rho[], rho_cur, k_cur, rr, rr0, tol, done) and the update kernel's
// block 0 (iter, alpha); every other block only reads fields written by an
// earlier kernel, so there is no intra-kernel race.
struct PcgScalars {
    double rho[2];    // rho_k = z_k . r_k, slot k & 1
    double rho_cur;   // rho of the iteration in flight
    double rr;        // ||r||^2 seen by the last direction kernel
    double rr0;       // ||r0||^2
    double tol;       // stopping threshold on ||r|| (< 0: set by the next direction kernel)
    double abstol, reltol;
    double alpha;
    long long iter;     // completed iterations
    long long k_cur;    // iteration in flight (set by the direction kernel)
    long long maxiter;
    int done;
    int err;  // why the iteration stopped early (kPcgErr*), 0 when it did not
};
// PcgScalars::err codes.  A stop on one of these sets `done` as well, so the
// rest of a replayed graph is a no-op.
constexpr int kPcgErrNonFinite = 1;   // r.r or r.z is NaN or infinite
constexpr int kPcgErrIndefinite = 2;  // p.K p <= 0 or not finite: K not positive definite
constexpr int kPcgErrPrecond = 3;     // r.z <= 0 with r != 0: preconditioner not positive definite

// EnergyCriteria stopping state (TopOpt.jl src/FEA/convergence_criteria.jl:20-45).
// `energy` persists across solves, as the reference's mutable struct does.
struct EnergyCrit {
    double energy;  // x'Ax of the last check
    double tol;     // tolerance of both ratios
    double xtr, xAx;
    int converged;
    int err;  // 1: NaN in the energy or its change, 2: negative x'Ax
    int pad[2];
};

constexpr int kMaxPartials = 4096;  // per-block partial slots of one kernel
constexpr int kVecBlocks = 512;     // grid of the streaming vector kernels
constexpr int kMatvecBlocksCap = 1024;

// ------------------------------------------------------------ launchers --
// Matvec y = K(E) x with the reference boundary rule (fixed DOFs:
// y = meandiag * x; x masked to 0 at fixed DOFs inside the element products).
template <typename T>
hipError_t launch_matvec(const Grid& g, const ElemOp<T>& op, const T* x,
                         const uint8_t* fixmask, const T* E, T* y, double meandiag,
                         hipStream_t s);

// PCG direction + matvec (iteration k = iter + 1): reduces the previous
// update's partials (rz, rr), checks convergence, p_k = dinv.*r + beta p_{k-1}
// (p_k in p[k & 1]), q = K p_k, per-block partials of p.q into partA.
// dinv = nullptr: unit diagonal, not read (r is then the preconditioned
// residual z of the multigrid path); likewise for the update below.
template <typename T>
hipError_t launch_pcg_dirmatvec(const Grid& g, const ElemOp<T>& op, const T* r,
                                const T* dinv, T* p0, T* p1, T* q, const uint8_t* fixmask,
                                const T* E, double meandiag, PcgScalars* sc,
                                const double* partB, int nbB, double* partA, int* nbA_out,
                                hipStream_t s);
int pcg_dirmatvec_blocks(const Grid& g);    // FP64 grid size (partA slots)
int pcg_dirmatvec_blocks32(const Grid& g);  // FP32 grid size

// PCG update: alpha = rho/(p.q) from partA, x += alpha p, r -= alpha q,
// partials of r.z and r.r into partB (kVecBlocks blocks).
template <typename T>
hipError_t launch_pcg_update(int64_t n, T* x, T* r, const T* p0, const T* p1, const T* q,
                             const T* dinv, PcgScalars* sc, const double* partA, int nbA,
                             double* partB, hipStream_t s);
// The same update with the residual read from rin and written to r (rin may
// be r).
template <typename T>
hipError_t launch_pcg_update_from(int64_t n, T* x, const T* rin, T* r, const T* p0, const T* p1,
                                  const T* q, const T* dinv, PcgScalars* sc, const double* partA,
                                  int nbA, double* partB, hipStream_t s);

// PCG start: r = b - Ax (Ax may be null: x0 = 0), p[0] = 0, partials of
// r.z, r.r into partB, scalars reset (iter 0, rho_prev 1, tol pending).
// keep_direction = 1 keeps p and rho (residual replacement in mixed mode).
template <typename T>
hipError_t launch_pcg_start(int64_t n, const T* b, const T* Ax, T* r, const T* dinv, T* p0,
                            PcgScalars* sc, double* partB, int keep_direction, hipStream_t s);

// Jacobi inverse diagonal from E: dinv = 1 / (kdiag[d] * sum_{e in n} E_e),
// fixed DOFs 1/meandiag.
template <typename T>
hipError_t launch_jacobi_inv(const Grid& g, const ElemOp<double>& op, const double* E,
                             const uint8_t* fixmask, double meandiag, T* dinv,
                             hipStream_t s);

// E_e = rho(x~_e) (SIMP power law, either interpolation order).
hipError_t launch_density(int64_t nel, const double* xt, double xmin, const Penalty& pen, int pbi,
                          double* E64, float* E32, hipStream_t s);

// Element energy: ce_e = u_e^T Ke u_e, grad_e = -drho_e ce_e; obj -> *obj_out.
// `partials` holds energy_blocks(g) per-block sums (every slot written before
// it is read).
int energy_blocks(const Grid& g);
hipError_t launch_energy(const Grid& g, const ElemOp<double>& op, const double* u,
                         const double* xt, double xmin, const Penalty& pen, int pbi, double* ce,
                         double* grad, double* partials, double* obj_out, hipStream_t s);

// Filter: forward y = J x (transpose = 0) or y = J^T x (transpose = 1).
// halo = the largest stencil node offset on any axis (the stencil's nodes
// lie in [1 - halo, halo] around the cell's lower corner); 3D stencils with
// halo <= kFilterMaxHalo take the z-marching tile kernels.
constexpr int kFilterMaxHalo = 3;
hipError_t launch_filter(const Grid& g, const int* offs, const double* wts, int count,
                         int halo, const double* invden, const double* x, double* y,
                         int transpose, hipStream_t s);
hipError_t launch_filter_invden(const Grid& g, const int* offs, const double* wts,
                                int count, double* invden, hipStream_t s);
// y = proj(x) (g null) or y = proj'(x) * g; proj as in Penalty::proj.
hipError_t launch_project(int64_t n, int proj, double beta, const double* x, const double* g,
                          double* y, hipStream_t s);

// OC update on device: multi-candidate bracketing of lambda, then x_new.
constexpr int kOcCandidates = 64;  // lambda candidates per bracketing pass
struct OcState {
    double lo, hi, lambda, vol, change;
    int done;
    int passes;
    double isq[kOcCandidates];  // 1/sqrt(candidate) of the next pass
};
hipError_t launch_oc(int64_t nel, const double* x, const double* dc, const double* colsum,
                     double volfrac, double move, double tol, int max_passes, double* xnew,
                     OcState* st, double* partials, hipStream_t s);

// MMA87 update on device (k_mma.hip): asymptotes, the one-constraint dual by
// multi-candidate bracketing of the multiplier, the new design.
struct MmaParams {
    double s_init, s_incr, s_decr;
};
struct MmaState {
    double lo, hi, lambda, target, vol, change;
    int done;
    int passes;
    double cand[kOcCandidates];  // multipliers of the next pass
};
// iter: MMA iterations done so far in this run (asymptotes from s_init while < 2).
// L, U, x1, x2 (ncells) persist across iterations; xnew receives the design.
hipError_t launch_mma(int64_t nel, int64_t iter, const double* x, const double* dc,
                      const double* colsum, double volfrac, double tol, const MmaParams& prm,
                      double* L, double* U, double* x1, double* x2, double* xnew, MmaState* st,
                      double* partials, hipStream_t s);

// BESO update on device (k_beso.hip): bisection of the sensitivity threshold,
// 6 levels of the bisection tree per pass.
constexpr int kBesoLevels = 6;
constexpr int kBesoCand = (1 << kBesoLevels) - 1;
constexpr int kGeneralPartials = 16384;  // size of the context's general partials buffer
struct BesoState {
    double l1, l2;    // bisection bracket of the threshold
    double th;        // last threshold tested
    double sens_tol;  // loop condition (l2 - l1) / l2 > sens_tol
    double vol;       // target volume fraction of this iteration
    double cv, total; // cell volume, total volume
    double nblack;    // passive solid cells (always counted)
    int steps;        // bisection steps taken this iteration
    int done;
    double cand[kBesoCand];  // thresholds of the current tree (heap order)
};
// sens = -dc (averaged with prev when avg), bracket, `passes` tree passes.
hipError_t launch_beso_update(int64_t n, const double* dc, const double* prev, int avg,
                              double* sens, const uint8_t* mask, BesoState* st, double* topo,
                              double* partials, int passes, hipStream_t s);
hipError_t launch_beso_passes(int64_t n, const double* sens, const uint8_t* mask, BesoState* st,
                              double* partials, int passes, hipStream_t s);
// topology of the last threshold; *sum_out = sum of the topology.
hipError_t launch_beso_apply(int64_t n, const double* sens, const uint8_t* mask,
                             const BesoState* st, double* topo, double* partials, double* sum_out,
                             hipStream_t s);

// Assembled K in CSR (k_csr.hip).  Ke: the full element matrix (corner bit
// order, (dim*c+i) rows).  launch_csr_build writes the column pattern when
// `pattern`, the unconstrained diagonal (diag), mdiag_sum[0..2] = (sum |diag|,
// its part over prescribed DOFs, their count) and the values with apply!'s
// rule (prescribed rows / columns 0, diagonal mean |diag| of the unconstrained
// K; `safe`: a zero free diagonal gets mean |diag| of the constrained K).
hipError_t launch_csr_build(const Grid& g, const double* Ke, const double* E,
                            const uint8_t* fixmask, const int64_t* rowptr, int* cols, int safe,
                            double* vals, double* diag, double* partials, double* mdiag_sum,
                            bool pattern, hipStream_t s);
hipError_t launch_csr_dinv(const Grid& g, const double* diag, const uint8_t* fixmask,
                           const double* mdiag_sum, int jacobi, int safe, double* dinv,
                           hipStream_t s);
hipError_t launch_csr_spmv(int64_t nrows, const int64_t* rowptr, const int* cols,
                           const double* vals, const double* x, double* y, hipStream_t s);
// PCG direction (convergence test, p_k = dinv .* r + beta p_{k-1} into
// p[k & 1]) for operators applied by a kernel of their own.
hipError_t launch_pcg_direction(int64_t n, const double* r, const double* dinv, double* p0,
                                double* p1, PcgScalars* sc, const double* partB, int nbB,
                                hipStream_t s);
// One PCG iteration on the CSR matrix: direction, q = K p (+ p.q partials),
// update (k_pcg_update); every kernel is a no-op once `done`.
hipError_t launch_csr_pcg_iter(int64_t n, const int64_t* rowptr, const int* cols,
                               const double* vals, double* x, double* r, const double* dinv,
                               double* p0, double* p1, double* q, PcgScalars* sc, double* partA,
                               double* partB, hipStream_t s);

// EnergyCriteria check on the current iterate (x, r = f - Ax): partial sums
// of x.r and f.x over kVecBlocks blocks into `partials`, then one block sets
// ec and, when converged or on a domain error, sc->done.  No-op once done.
hipError_t launch_energy_check(int64_t n, const double* x, const double* r, const double* f,
                               PcgScalars* sc, EnergyCrit* ec, double* partials, hipStream_t s);

// Sum-reductions finished by a one-block kernel: out[v] = sum_b part[b*nv+v].
hipError_t launch_reduce(const double* partials, int nb, int nv, double* out, hipStream_t s);

// Mixed precision helpers.
// u += x32 (FP64 accumulate), x32 = 0.
hipError_t launch_accumulate_f32(int64_t n, float* x32, double* u, hipStream_t s);
// r64 = b - Ku; r32 = r64; kVecBlocks partials each of (r32.dinv32 r32,
// r32.r32) at partB[2 b], [2 b + 1] and r64.r64 at partB[2 kVecBlocks + b].
// Pass r32 = nullptr for FP64 only.
hipError_t launch_residual_refresh(int64_t n, const double* b, const double* Ku, double* r64,
                                   float* r32, const float* dinv32, double* partB,
                                   hipStream_t s);
// Set PCG scalars for a residual replacement: done = 0, new tolerance, and a
// direction restart when ||r64|| > restart * the recursive ||r|| (restart > 0).
hipError_t launch_pcg_rearm(PcgScalars* sc, const double* rr64, double factor, double floor_tol,
                            double restart, hipStream_t s);
hipError_t launch_f64_to_f32(int64_t n, const double* a, float* b, hipStream_t s);
hipError_t launch_f32_to_f64(int64_t n, const float* a, double* b, hipStream_t s);

// ---- geometric multigrid preconditioner (k_mg.hip) ----
// Coarsest level handled by the one-workgroup LDS kernel (k_mg_coarse).
constexpr int kMgCoarseMaxNodes = 600;
constexpr int kMgCoarseMaxCells = 400;
// One level of the hierarchy as the kernels see it (FP32 operator).
struct MgLevelView {
    Grid g;
    ElemOp<float> op;
    const float* E32;
    const uint8_t* fixmask;
    const float* dinv;
    const float* wdev = nullptr;  // smoother weight in device memory (else the launch's w)
};
int mg_blocks(const Grid& g);        // grid (= partial count) of k_mg_smooth on g
int mg_power_blocks(const Grid& g);  // grid of the power step (mode 5) on g
bool mg_coarse_fits(const Grid& g);
// The pending update of the multigrid PCG, fused into the finest
// pre-smoothing (mode 3): alpha = rho_cur / p.q from the direction kernel's
// partials (breakdown when p.q is not positive and finite, as the update
// kernel does); the right-hand side is r - alpha q, with r = `b` of the launch
// and the new residual written to `rnew` for the nodes the workgroup owns (a
// different buffer: neighbours read r on their halos); x += alpha p.
struct MgFusedUpdate {
    const float* q;
    const float* p0;
    const float* p1;
    float* rnew;
    float* x;
    const double* partA;
    int nbA;
    PcgScalars* sc;
};
// mode 0: x = w dinv b, r = b - K x          (xout = x, rout = r)
// mode 1: x' = xprev + P xc, xout = x' + w dinv (b - K x'); partials
//         (b.xout, b.b) into part[2 * block + {0, 1}] when part != null
// mode 2: xout = xprev + w dinv (b - K xprev)
// mode 3: mode 0 on b - alpha q after the fused update `upd`
// mode 5: xout = u = (I - D^-1 K) v, v = aux[0] xprev, and the partials
//         (v'Du, v'Dv, u'Du) into part[3 * block + {0, 1, 2}] over
//         mg_power_blocks(L.g) blocks (tiles on every level)
hipError_t launch_mg_smooth(int mode, const MgLevelView& L, float w, const float* b,
                            const float* xprev, const MgLevelView* C, const float* xc,
                            float* xout, float* rout, double meandiag, const PcgScalars* sc,
                            double* part, hipStream_t s, const MgFusedUpdate* upd = nullptr,
                            const float* aux = nullptr);
// out = {w, rescale, theta} from the partials of the last two mode-5 launches
// (prev, part) over nb blocks: theta = Rayleigh quotient of D^-1 K at the last,
// w = clamp(1.9 / (margin theta), wsafe, wmax) when theta has levelled off
// (<= 1.01 x the previous step's), else wsafe.
hipError_t launch_mg_omega(const double* prev, const double* part, int nb, float* out,
                          float wsafe, float wmax, float margin, hipStream_t s);
// v[i] uniform in [-0.5, 0.5) from a hash of i (start of the estimate)
hipError_t launch_mg_fill_start(int64_t n, float* v, hipStream_t s);
// Finest-level smoothing by z-marching columns (k_matvec_zm.hip): whether a
// level takes it, and the launch (mode 1: pre on b; mode 2: post on xin = x').
bool mg_use_zm(const Grid& g);
int64_t zm_min_cells();  // smallest grid the z-marching kernels take (cells)
int zm_blocks(const Grid& g);  // grid (= partial count) of the z-marching kernels on g
hipError_t launch_mg_smooth_zm(int mode, const Grid& g, const ElemOp<float>& op, const float* E,
                               const uint8_t* fixmask, const float* dinv, float w,
                               const float* wdev, const float* xin, const float* b, float* xout,
                               float* rout,
                               double meandiag, const PcgScalars* sc, double* part,
                               hipStream_t s, const Grid* gc = nullptr,
                               const float* xc = nullptr);
// bc = P^T r, masked at the coarse level's prescribed DOFs.
hipError_t launch_mg_restrict(const MgLevelView& F, const MgLevelView& C, const float* r,
                              float* bc, const PcgScalars* sc, hipStream_t s);
// Chebyshev iteration on D^-1 K x = D^-1 b from x = 0 (a fixed polynomial):
// x_1 = c0 D^-1 b; x_{k+1} = x_k + a[k] (x_k - x_{k-1}) + g[k] D^-1 (b - K x_k).
constexpr int kMgChebMax = 64;
struct MgCheb {
    int iters;
    float c0;
    float a[kMgChebMax];
    float g[kMgChebMax];
};
// Coefficients for the eigenvalue interval [lmax / kappa, lmax] of D^-1 K.
MgCheb mg_chebyshev(double lmax, double kappa, int iters);
// x = the Chebyshev iteration on the coarsest level (one workgroup).
hipError_t launch_mg_coarse(const MgLevelView& L, const MgCheb& ch, const float* b, float* x,
                            const PcgScalars* sc, hipStream_t s);
// Dense coarsest solve: Mout (ndof x ndof, column j = the Chebyshev operator
// applied to e_j, then symmetrised) once per operator refresh; then x = M b
// per V-cycle, one wave per row.  A no-op once sc->done.
hipError_t launch_mg_coarse_basis(const MgLevelView& L, const MgCheb& ch, float* Mout,
                                  hipStream_t s);
hipError_t launch_mg_dense(int n, const float* M, const float* b, float* x, const PcgScalars* sc,
                           hipStream_t s);
// Coarse cell moduli: mean of the 2^dim children.
hipError_t launch_mg_coarsen(const Grid& gf, const Grid& gc, const double* Ef, double* E64,
                             float* E32, hipStream_t s);
hipError_t launch_copy64(int64_t n, const double* a, double* b, hipStream_t s);
hipError_t launch_fill64(int64_t n, double v, double* x, hipStream_t s);
// Message for a PcgScalars::err code.
inline const char* pcg_err_text(int err) {
    switch (err) {
        case kPcgErrNonFinite:
            return "PCG breakdown: the residual became NaN or infinite (K not positive definite?)";
        case kPcgErrIndefinite:
            return "PCG breakdown: p'Kp <= 0, the operator is not positive definite "
                   "(a hard-void design with xmin = 0?)";
        case kPcgErrPrecond:
            return "PCG breakdown: r'z <= 0, the preconditioner is not positive definite";
        default:
            return "PCG breakdown";
    }
}
hipError_t launch_sum(int64_t n, const double* x, double* partials, double* out,
                      hipStream_t s);

}  // namespace tohip
