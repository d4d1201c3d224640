// Heat conduction on the device: the scalar counterpart of the elasticity
// operator (TopOpt.jl src/TopOptProblems/matrices_and_vectors.jl:413-496 for
// the element conductivity matrix, src/FEA/matrix_free_operator.jl:66-105
// for the matrix-free product and its fixed-DOF rule).
//
// One DOF per node; K = sum_e rho_e Ke with one shared Ke (k * grad phi_a .
// grad phi_b integrated over the box cell).  A node's row gathers the 4 (2D)
// or 8 (3D) cells around it, each a 4- or 8-term row of Ke, in a fixed order
// (one thread per node: deterministic, no atomics).  The kernels are
// HBM-bound gathers; the 27-point neighbourhood is served from L2.
#include "heat_internal.h"
#include "k_common.h"

namespace tohip {

// Ferrite corner index of the corner at offset (ox, oy, oz) of its cell.
__device__ __forceinline__ int heat_corner(int ox, int oy, int oz) {
    return (oy ? (ox ? 2 : 3) : (ox ? 1 : 0)) + 4 * oz;
}

// MODE 0: y = K x (masked columns, fixed rows md x); MODE 1: PCG, q = K p_k
// with p_k = p[k & 1] written by the direction kernel, per-block p.q partials
// into partA, a no-op once `done`; MODE 2: y = K_full x (no masking).
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_heat_matvec(
    HeatGrid g, HeatKe ke, const double* __restrict__ x0, const double* __restrict__ x1,
    const uint8_t* __restrict__ fixmask, const double* __restrict__ E, double md,
    double* __restrict__ y, const PcgScalars* sc, double* __restrict__ partA) {
    __shared__ double red[kWaves];
    constexpr int NC = DIM == 3 ? 8 : 4;
    const double* x = x0;
    if (MODE == 1) {
        if (sc->done) return;
        x = (sc->k_cur & 1) ? x1 : x0;
    }
    double part = 0.0;
    for (int64_t n = blockIdx.x * (int64_t)kThreads + threadIdx.x; n < g.nnodes;
         n += (int64_t)gridDim.x * kThreads) {
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
        const int k = DIM == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
        const double vn = x[n];
        double yn = 0.0;
        if (MODE != 2 && fixmask[n]) {
            yn = md * vn;
        } else {
            // the 3^DIM neighbourhood once (masked, 0 outside the grid), then
            // the 2^DIM cells around the node, each a row of Ke
            constexpr int NN = DIM == 3 ? 27 : 9;
            double xv[NN];
#pragma unroll
            for (int q = 0; q < NN; ++q) {
                const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = DIM == 3 ? q / 9 - 1 : 0;
                const int mi = i + dx, mj = j + dy, mk = k + dz;
                double v = 0.0;
                if (mi >= 0 && mi < g.NX && mj >= 0 && mj < g.NY &&
                    (DIM == 2 || (mk >= 0 && mk < g.NZ))) {
                    const int64_t m = mi + (int64_t)g.NX * (mj + (int64_t)g.NY * mk);
                    v = (MODE != 2 && fixmask[m]) ? 0.0 : (m == n ? vn : x[m]);
                }
                xv[q] = v;
            }
#pragma unroll
            for (int cq = 0; cq < NC; ++cq) {
                // cell origin offset (ox - 1, oy - 1, oz - 1) relative to the node
                const int ox = cq & 1, oy = (cq >> 1) & 1, oz = cq >> 2;
                const int cx = i - 1 + ox, cy = j - 1 + oy, cz = DIM == 3 ? k - 1 + oz : 0;
                if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny ||
                    (DIM == 3 && (cz < 0 || cz >= g.nz)))
                    continue;
                const int a = heat_corner(1 - ox, 1 - oy, DIM == 3 ? 1 - oz : 0);
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const int bx = (b & 1) ^ ((b >> 1) & 1), by = (b >> 1) & 1, bz = b >> 2;
                    // corner b of the cell is node offset (ox - 1 + bx, ...) from n
                    const int q = (ox + bx) + 3 * ((oy + by) + 3 * (DIM == 3 ? oz + bz : 0));
                    t += ke.k[a * NC + b] * xv[q];
                }
                yn += E[cx + (int64_t)g.nx * (cy + (int64_t)g.ny * cz)] * t;
            }
        }
        y[n] = yn;
        if (MODE == 1) part += vn * yn;
    }
    if (MODE == 1) {
        const double s = block_sum_all(part, red);
        if (threadIdx.x == 0) partA[blockIdx.x] = s;
    }
}

static unsigned heat_blocks(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > kMaxPartials) b = kMaxPartials;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_heat_matvec(const HeatGrid& g, const HeatKe& ke, const double* x,
                              const uint8_t* fixmask, const double* E, double meandiag, double* y,
                              int mode, hipStream_t s) {
    const dim3 grid(heat_blocks(g.nnodes)), block(kThreads);
#define TOHIP_HEAT(D, M)                                                                        \
    hipLaunchKernelGGL((k_heat_matvec<D, M>), grid, block, 0, s, g, ke, x, x, fixmask, E,       \
                       meandiag, y, nullptr, nullptr)
    if (g.dim == 3) {
        if (mode == 2) TOHIP_HEAT(3, 2); else TOHIP_HEAT(3, 0);
    } else {
        if (mode == 2) TOHIP_HEAT(2, 2); else TOHIP_HEAT(2, 0);
    }
#undef TOHIP_HEAT
    return hipGetLastError();
}

hipError_t launch_heat_dirmatvec(const HeatGrid& g, const HeatKe& ke, const double* r,
                                 const double* dinv, double* p0, double* p1, double* q,
                                 const uint8_t* fixmask, const double* E, double meandiag,
                                 PcgScalars* sc, const double* partB, int nbB, double* partA,
                                 int* nbA, hipStream_t s) {
    if (hipError_t e = launch_pcg_direction(g.nnodes, r, dinv, p0, p1, sc, partB, nbB, s))
        return e;
    const unsigned nb = heat_blocks(g.nnodes);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_heat_matvec<3, 1>), dim3(nb), dim3(kThreads), 0, s, g, ke, p0, p1,
                           fixmask, E, meandiag, q, sc, partA);
    else
        hipLaunchKernelGGL((k_heat_matvec<2, 1>), dim3(nb), dim3(kThreads), 0, s, g, ke, p0, p1,
                           fixmask, E, meandiag, q, sc, partA);
    if (nbA) *nbA = (int)nb;
    return hipGetLastError();
}

__global__ void k_heat_density(int64_t ncells, const double* __restrict__ x, double xmin,
                               Penalty pen, int pbi, double* __restrict__ E) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        double r, dr;
        rho_drho(pen, x[c], xmin, pbi, r, dr);
        E[c] = r;
    }
}

hipError_t launch_heat_density(int64_t ncells, const double* x, double xmin, const Penalty& pen,
                               int pbi, double* E, hipStream_t s) {
    hipLaunchKernelGGL(k_heat_density, dim3(heat_blocks(ncells)), dim3(kThreads), 0, s, ncells, x,
                       xmin, pen, pbi, E);
    return hipGetLastError();
}

template <int DIM>
__global__ void k_heat_diag(HeatGrid g, HeatKe ke, const uint8_t* __restrict__ fixmask,
                            const double* __restrict__ E, double md, double* __restrict__ dinv) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        if (fixmask[n]) {
            dinv[n] = 1.0 / md;
            continue;
        }
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
        const int k = DIM == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
        double d = 0.0;
        for (int cz = (DIM == 3 ? k - 1 : 0); cz <= (DIM == 3 ? k : 0); ++cz) {
            if (DIM == 3 && (cz < 0 || cz >= g.nz)) continue;
            for (int cy = j - 1; cy <= j; ++cy) {
                if (cy < 0 || cy >= g.ny) continue;
                for (int cx = i - 1; cx <= i; ++cx) {
                    if (cx < 0 || cx >= g.nx) continue;
                    const int a = heat_corner(i - cx, j - cy, DIM == 3 ? k - cz : 0);
                    d += E[cx + (int64_t)g.nx * (cy + (int64_t)g.ny * cz)] * ke.k[a * NC + a];
                }
            }
        }
        dinv[n] = 1.0 / d;
    }
}

hipError_t launch_heat_diag(const HeatGrid& g, const HeatKe& ke, const uint8_t* fixmask,
                            const double* E, double meandiag, double* dinv, hipStream_t s) {
    const dim3 grid(heat_blocks(g.nnodes)), block(kThreads);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_heat_diag<3>), grid, block, 0, s, g, ke, fixmask, E, meandiag, dinv);
    else
        hipLaunchKernelGGL((k_heat_diag<2>), grid, block, 0, s, g, ke, fixmask, E, meandiag, dinv);
    return hipGetLastError();
}

__global__ void k_heat_rhs(int64_t n, const double* __restrict__ q, const double* __restrict__ lift,
                           const uint8_t* __restrict__ fixmask, const double* __restrict__ v,
                           double md, int use_v, double* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (fixmask[i])
            b[i] = use_v ? md * v[i] : 0.0;
        else
            b[i] = q[i] - (lift ? lift[i] : 0.0);
    }
}

hipError_t launch_heat_rhs(int64_t n, const double* q, const double* lift,
                           const uint8_t* fixmask, const double* v, double meandiag, int use_v,
                           double* b, hipStream_t s) {
    hipLaunchKernelGGL(k_heat_rhs, dim3(heat_blocks(n)), dim3(kThreads), 0, s, n, q, lift, fixmask,
                       v, meandiag, use_v, b);
    return hipGetLastError();
}

template <int DIM>
__global__ void k_heat_cell_products(HeatGrid g, HeatKe ke, const double* __restrict__ lam,
                                     const double* __restrict__ T, const double* __restrict__ x,
                                     double xmin, Penalty pen, int pbi, double sign,
                                     double* __restrict__ cell, double* __restrict__ grad) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int cx = (int)(c % g.nx), cy = (int)((c / g.nx) % g.ny);
        const int cz = DIM == 3 ? (int)(c / ((int64_t)g.nx * g.ny)) : 0;
        double lv[NC], tv[NC];
#pragma unroll
        for (int b = 0; b < NC; ++b) {
            const int ox = (b & 1) ^ ((b >> 1) & 1), oy = (b >> 1) & 1, oz = b >> 2;
            const int64_t m = (cx + ox) + (int64_t)g.NX * ((cy + oy) + (int64_t)g.NY * (cz + oz));
            lv[b] = lam[m];
            tv[b] = T[m];
        }
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < NC; ++b) t += ke.k[a * NC + b] * tv[b];
            s += lv[a] * t;
        }
        double r, dr;
        rho_drho(pen, x[c], xmin, pbi, r, dr);
        cell[c] = s;
        grad[c] = sign * dr * s;
    }
}

hipError_t launch_heat_cell_products(const HeatGrid& g, const HeatKe& ke, const double* lam,
                                     const double* T, const double* x, double xmin,
                                     const Penalty& pen, int pbi, double sign, double* cell,
                                     double* grad, hipStream_t s) {
    const dim3 grid(heat_blocks(g.ncells)), block(kThreads);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_heat_cell_products<3>), grid, block, 0, s, g, ke, lam, T, x, xmin,
                           pen, pbi, sign, cell, grad);
    else
        hipLaunchKernelGGL((k_heat_cell_products<2>), grid, block, 0, s, g, ke, lam, T, x, xmin,
                           pen, pbi, sign, cell, grad);
    return hipGetLastError();
}

}  // namespace tohip
