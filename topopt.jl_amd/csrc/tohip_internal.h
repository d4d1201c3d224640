// Internal declarations shared by the libtopopt_hip translation units.
//
// Device layout (all vectors node-major, components interleaved):
//   node n = i + NX*(j + NY*k), NX = nx+1, ...; DOF index = DIM*n + d.
//   cells  e = i + nx*(j + ny*k)  (same x-fastest order as the caller's).
//   fixmask[n] bit d set  <=>  DOF (n, d) is prescribed (homogeneous Dirichlet).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

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
// label m ^ (1<<d); blk[L][d][d'] holds those 3x3 (2x2 in 2D) blocks, already
// divided by (number of corners)^2.
template <typename T>
struct ElemOp {
    T blk[8][3][3];
    T kdiag[3];  // Ke diagonal per component (identical for every corner)
};

// Density filter stencil: node offsets relative to the cell's lower corner
// node and weights rmin - |x_node - x_centre| (strictly positive).
constexpr int kMaxStencil = 1024;
struct FilterStencil {
    int count;
    int radius;  // max |offset| in cells, for tiling
};

// Device-resident PCG scalars (always FP64).
struct PcgScalars {
    double rho;       // current r.z
    double beta;      // rho_new / rho_old for the next direction
    double alpha;     // rho / (p.q)
    double pq;
    double rr;        // ||r||^2 after the last update
    double rr0;       // ||r0||^2
    double tol;       // stopping threshold on ||r||
    double abstol, reltol;
    double scale;     // mixed precision: ||r_outer|| used to scale the inner rhs
    long long iter;
    long long maxiter;
    int done;
    int pad;
    unsigned int counter[4];
};

// Reduction workspace size (doubles per partial slot).
constexpr int kMaxPartials = 65536;

// ------------------------------------------------------------ launchers --
// Matvec y = K(E) x with the reference boundary rule (fixed DOFs:
// y = meandiag * x; x masked to 0 at fixed DOFs inside the element products).
template <typename T>
hipError_t launch_matvec(const Grid& g, const ElemOp<T>& op, const T* x,
                         const uint8_t* fixmask, const T* E, T* y, double meandiag,
                         hipStream_t s);

// PCG direction + matvec: p = dinv.*r + beta*p_old (beta from sc), q = K p,
// partial p.q -> last block sets sc->alpha.
template <typename T>
hipError_t launch_pcg_dirmatvec(const Grid& g, const ElemOp<T>& op, const T* r,
                                const T* dinv, const T* p_old, T* p_new, T* q,
                                const uint8_t* fixmask, const T* E, double meandiag,
                                PcgScalars* sc, double* partials, hipStream_t s);

// PCG update: x += alpha p, r -= alpha q, rho = r.(dinv r), rr = r.r; last
// block: beta, convergence.
template <typename T>
hipError_t launch_pcg_update(int64_t n, T* x, T* r, const T* p, const T* q,
                             const T* dinv, PcgScalars* sc, double* partials,
                             hipStream_t s);

// PCG start: r = b - Ax (Ax given, may be null for x0 = 0), rho, rr, tol.
template <typename T>
hipError_t launch_pcg_start(int64_t n, const T* b, const T* Ax, T* r, const T* dinv,
                            T* p_old, PcgScalars* sc, double* partials, hipStream_t s);

// Jacobi inverse diagonal from E: dinv = 1 / (kdiag[d] * sum_{e in n} E_e),
// fixed DOFs 1/meandiag.
template <typename T>
hipError_t launch_jacobi_inv(const Grid& g, const ElemOp<double>& op, const double* E,
                             const uint8_t* fixmask, double meandiag, T* dinv,
                             hipStream_t s);

// E_e = rho(x~_e) (SIMP power law, either interpolation order).
hipError_t launch_density(int64_t nel, const double* xt, double xmin, double p, int pbi,
                          double* E64, float* E32, hipStream_t s);

// Element energy: ce_e = u_e^T Ke u_e, grad_e = -drho_e ce_e, obj partials.
hipError_t launch_energy(const Grid& g, const ElemOp<double>& op, const double* u,
                         const double* xt, double xmin, double p, int pbi, double* ce,
                         double* grad, double* partials, double* obj_out,
                         unsigned int* counter, hipStream_t s);

// Filter: forward y = J x (transpose = 0) or y = J^T x (transpose = 1).
hipError_t launch_filter(const Grid& g, const int* offs, const double* wts, int count,
                         const double* invden, const double* x, double* y, int transpose,
                         double* tmp, hipStream_t s);
hipError_t launch_filter_invden(const Grid& g, const int* offs, const double* wts,
                                int count, double* invden, hipStream_t s);

// OC update on device: multi-candidate bracketing of lambda, then x_new.
struct OcState {
    double lo, hi, lambda, vol, change;
    int done;
    int passes;
    unsigned int counter;
    int pad;
};
hipError_t launch_oc(int64_t nel, const double* x, const double* dc, const double* colsum,
                     double volfrac, double move, double tol, int max_passes, double* xnew,
                     OcState* st, double* partials, hipStream_t s);

// Small vector helpers.
hipError_t launch_axpby(int64_t n, double a, const double* x, double b, double* y,
                        hipStream_t s);
hipError_t launch_residual64(int64_t n, const double* f, const double* Ku, double* r,
                             double* partials, double* rr_out, unsigned int* counter,
                             hipStream_t s);
hipError_t launch_scale_to_f32(int64_t n, const double* r, const double* norm2, float* out,
                               hipStream_t s);
hipError_t launch_add_scaled_f32(int64_t n, const float* d, const double* norm2, double* u,
                                 hipStream_t s);
hipError_t launch_f64_to_f32(int64_t n, const double* a, float* b, hipStream_t s);
hipError_t launch_sum(int64_t n, const double* x, double* partials, double* out,
                      unsigned int* counter, hipStream_t s);

}  // namespace tohip
