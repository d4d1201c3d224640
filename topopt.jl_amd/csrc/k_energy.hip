// Compliance and its sensitivity (replaces compute_element_energy, TopOpt.jl
// src/Functions/compute_element_energy.jl:18-38):
//   ce_e = u_e^T Ke u_e (unconstrained Ke), grad_e = -drho_e * ce_e,
//   obj = sum_e rho_e * ce_e (per-block partials, one-block finisher).
// u_e^T Ke u_e is evaluated in the corner-Hadamard basis: (H^T u_e)^T B (H^T u_e).
#include <algorithm>

#include "k_common.h"

namespace tohip {

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_energy(Grid g, ElemOp<double> op,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ xt, double xmin,
                                                     Penalty pen, int pbi, double* __restrict__ ce,
                                                     double* __restrict__ grad,
                                                     double* __restrict__ partials) {
    constexpr int NC = 1 << DIM;
    __shared__ double red[kWaves];
    double part = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < g.ncells;
         e += (int64_t)gridDim.x * kThreads) {
        const int i = e % g.nx, j = (e / g.nx) % g.ny, k = DIM == 3 ? e / ((int64_t)g.nx * g.ny) : 0;
        double ue[DIM][NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int64_t n = (i + (c & 1)) +
                              (int64_t)g.NX * ((j + ((c >> 1) & 1)) +
                                               (int64_t)g.NY * (DIM == 3 ? k + ((c >> 2) & 1) : 0));
#pragma unroll
            for (int d = 0; d < DIM; ++d) ue[d][c] = u[DIM * n + d];
        }
        const double en = op.is_cube ? elem_energy<double, DIM, true>(op, ue)
                                     : elem_energy<double, DIM, false>(op, ue);
        double r, dr;
        rho_drho(pen, xt[e], xmin, pbi, r, dr);
        ce[e] = en;
        grad[e] = -dr * en;
        part += r * en;
    }
    const double s = block_sum_all(part, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

int energy_blocks(const Grid& g) {
    return (int)std::min<int64_t>(kMaxPartials, (g.ncells + kThreads - 1) / kThreads);
}

hipError_t launch_energy(const Grid& g, const ElemOp<double>& op, const double* u,
                         const double* xt, double xmin, const Penalty& pen, int pbi, double* ce,
                         double* grad, double* partials, double* obj_out, hipStream_t s) {
    const int nb = energy_blocks(g);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_energy<3>, dim3(nb), dim3(kThreads), 0, s, g, op, u, xt, xmin, pen,
                           pbi, ce, grad, partials);
    else
        hipLaunchKernelGGL(k_energy<2>, dim3(nb), dim3(kThreads), 0, s, g, op, u, xt, xmin, pen,
                           pbi, ce, grad, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce(partials, nb, 1, obj_out, s);
}

}  // namespace tohip
