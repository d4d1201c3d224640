// Compliance and its sensitivity (replaces compute_element_energy, TopOpt.jl
// src/Functions/compute_element_energy.jl:18-38):
//   ce_e = u_e^T Ke u_e (unconstrained Ke), grad_e = -drho_e * ce_e,
//   obj = sum_e rho_e * ce_e (deterministic last-block reduction).
// u_e^T Ke u_e is evaluated in the corner-Hadamard basis: (H^T u_e)^T B (H^T u_e).
#include "k_common.h"

namespace tohip {

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_energy(Grid g, ElemOp<double> op,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ xt, double xmin,
                                                     double p, int pbi, double* __restrict__ ce,
                                                     double* __restrict__ grad, double* partials,
                                                     double* obj_out, unsigned int* counter) {
    constexpr int NC = 1 << DIM;
    __shared__ double red[kWaves];
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    double part = 0.0;
    if (e < g.ncells) {
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
        const double en = elem_energy<double, DIM>(op, ue);
        const double x = xt[e];
        double r, dr;
        if (pbi) {
            r = pow(x, p) * (1.0 - xmin) + xmin;
            dr = (1.0 - xmin) * p * (p == 1.0 ? 1.0 : pow(x, p - 1.0));
        } else {
            const double dd = x * (1.0 - xmin) + xmin;
            r = pow(dd, p);
            dr = p * (p == 1.0 ? 1.0 : pow(dd, p - 1.0)) * (1.0 - xmin);
        }
        ce[e] = en;
        grad[e] = -dr * en;
        part = r * en;
    }
    double s = block_sum(part, red);
    if (publish_partial(s, partials, blockIdx.x, counter, gridDim.x)) {
        double t = sum_partials(partials, gridDim.x, 1, 0, red);
        if (threadIdx.x == 0) {
            *obj_out = t;
            *counter = 0;
        }
    }
}

hipError_t launch_energy(const Grid& g, const ElemOp<double>& op, const double* u,
                         const double* xt, double xmin, double p, int pbi, double* ce,
                         double* grad, double* partials, double* obj_out, unsigned int* counter,
                         hipStream_t s) {
    const int64_t nb = (g.ncells + kThreads - 1) / kThreads;
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_energy<3>, dim3((unsigned)nb), dim3(kThreads), 0, s, g, op, u, xt,
                           xmin, p, pbi, ce, grad, partials, obj_out, counter);
    else
        hipLaunchKernelGGL(k_energy<2>, dim3((unsigned)nb), dim3(kThreads), 0, s, g, op, u, xt,
                           xmin, p, pbi, ce, grad, partials, obj_out, counter);
    return hipGetLastError();
}

}  // namespace tohip
