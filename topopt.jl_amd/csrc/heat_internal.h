// Internal declarations of the heat-conduction path (scalar temperature, one
// DOF per node): the matrix-free conduction operator, the PCG direction
// kernel built on it, and the element products of the thermal gradients.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tohip_internal.h"

namespace tohip {

// Structured grid of unit-ordered cells: nodes x-fastest, NX = nx + 1.
struct HeatGrid {
    int dim;
    int nx, ny, nz;  // cells (nz = 1 in 2D)
    int NX, NY, NZ;  // nodes (NZ = 1 in 2D)
    int64_t nnodes, ncells;
};

// Element conductivity matrix (Ferrite corner order, row-major, 4x4 or 8x8).
struct HeatKe {
    double k[64];
};

// y = K x over the nodes.  mode 0: the solver's operator (x masked to 0 at
// fixed nodes inside the element products, fixed rows y = meandiag * x);
// mode 2: the full operator (no masking; the Dirichlet lift K v).
hipError_t launch_heat_matvec(const HeatGrid& g, const HeatKe& ke, const double* x,
                              const uint8_t* fixmask, const double* E, double meandiag, double* y,
                              int mode, hipStream_t s);

// PCG direction + matvec (iteration k = iter + 1): launch_pcg_direction
// (convergence test, p_k = dinv .* r + beta p_{k-1} into p[k & 1]), then
// q = K p_k (mode 0) with per-block p.q partials into partA (*nbA blocks).
hipError_t launch_heat_dirmatvec(const HeatGrid& g, const HeatKe& ke, const double* r,
                                 const double* dinv, double* p0, double* p1, double* q,
                                 const uint8_t* fixmask, const double* E, double meandiag,
                                 PcgScalars* sc, const double* partB, int nbB, double* partA,
                                 int* nbA, hipStream_t s);

// E_e = rho(x_e) (get_rho with the penalty), dinv = 1 / diag(K) (free nodes:
// sum of E_c Ke[a][a]; fixed nodes: 1 / meandiag).
hipError_t launch_heat_density(int64_t ncells, const double* x, double xmin, const Penalty& pen,
                               int pbi, double* E, hipStream_t s);
hipError_t launch_heat_diag(const HeatGrid& g, const HeatKe& ke, const uint8_t* fixmask,
                            const double* E, double meandiag, double* dinv, hipStream_t s);

// b = q - lift on free nodes, b = meandiag * v on fixed nodes (v = fixed values).
hipError_t launch_heat_rhs(int64_t n, const double* q, const double* lift,
                           const uint8_t* fixmask, const double* v, double meandiag, int use_v,
                           double* b, hipStream_t s);

// cell_e = lam_e' Ke T_e (unconstrained Ke), grad_e = sign * drho(x_e) * cell_e.
hipError_t launch_heat_cell_products(const HeatGrid& g, const HeatKe& ke, const double* lam,
                                     const double* T, const double* x, double xmin,
                                     const Penalty& pen, int pbi, double sign, double* cell,
                                     double* grad, hipStream_t s);

}  // namespace tohip
