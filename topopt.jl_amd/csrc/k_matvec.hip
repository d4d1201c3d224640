// Matrix-free K(E)·x on the structured Q1 grid (replaces mul!(y,
// ::MatrixFreeOperator, x), TopOpt.jl src/FEA/matrix_free_operator.jl:66-105).
//
// One workgroup owns a TX x TY (x TZ) tile of output nodes.  It stages the
// input vector on the tile plus a one-node halo in LDS (structure of arrays,
// fixed DOFs masked to 0 like bcmatrix), computes every cell touching the
// tile (the halo cells are recomputed by the neighbouring tile: no atomics,
// no colouring passes), writes each cell's 8 (4) nodal results into a
// per-corner LDS slot, and finally each output node sums its slots in
// ascending cell order — the same order as the reference's dof_cells gather,
// so results are deterministic run to run.  Fixed DOFs return meandiag * x.
//
// The cell product uses the corner-Hadamard factorisation Ke = H B H^T
// (block-diagonal B, 8 blocks of 3x3): 72 FMA + 2x72 add per cell instead of
// the 576 FMA of a dense 24x24 product.
#include "k_common.h"

namespace tohip {

template <int DIM>
struct Tile;
template <>
struct Tile<3> {
    static constexpr int TX = 8, TY = 8, TZ = 4;
};
template <>
struct Tile<2> {
    static constexpr int TX = 32, TY = 8, TZ = 1;
};

template <int DIM>
static inline int ntiles(const Grid& g) {
    using C = Tile<DIM>;
    int n = ((g.NX + C::TX - 1) / C::TX) * ((g.NY + C::TY - 1) / C::TY);
    if (DIM == 3) n *= (g.NZ + C::TZ - 1) / C::TZ;
    return n;
}

// PCG = false: y = K x.
// PCG = true : p = dinv.*r + beta*p_old on the fly (r passed as xin), p_new
//              written for owned nodes, y = q = K p, block partial of p.q,
//              last block: alpha = rho / (p.q).
template <typename T, int DIM, bool PCG>
__global__ __launch_bounds__(kThreads) void k_matvec(
    Grid g, ElemOp<T> op, const T* __restrict__ xin, const T* __restrict__ dinv,
    const T* __restrict__ p_old, T* __restrict__ p_new, const uint8_t* __restrict__ fixmask,
    const T* __restrict__ E, T* __restrict__ y, double meandiag, PcgScalars* sc,
    double* partials) {
    using C = Tile<DIM>;
    constexpr int TX = C::TX, TY = C::TY, TZ = C::TZ;
    constexpr int NC = 1 << DIM;
    constexpr int IX = TX + 2, IY = TY + 2, IZ = DIM == 3 ? TZ + 2 : 1;
    constexpr int EX = TX + 1, EY = TY + 1, EZ = DIM == 3 ? TZ + 1 : 1;
    constexpr int NIN = IX * IY * IZ, NOUT = TX * TY * TZ, NEL = EX * EY * EZ;

    __shared__ T xs[DIM][NIN];
    __shared__ T slot[NC][DIM][NOUT];
    __shared__ T pown[DIM][NOUT];
    __shared__ double red[kWaves];

    if (PCG) {
        if (__hip_atomic_load(&sc->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    }
    const int tid = threadIdx.x;
    const int ntx = (g.NX + TX - 1) / TX, nty = (g.NY + TY - 1) / TY;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = DIM == 3 ? b / (ntx * nty) : 0;
    const int i0 = tx * TX, j0 = ty * TY, k0 = tz * TZ;
    const T beta = PCG ? (T)sc->beta : (T)0;

    // ---- stage input (tile + halo) ----
    for (int idx = tid; idx < NIN * DIM; idx += kThreads) {
        const int d = idx % DIM, nl = idx / DIM;
        const int li = nl % IX, lj = (nl / IX) % IY, lk = DIM == 3 ? nl / (IX * IY) : 0;
        const int gi = i0 - 1 + li, gj = j0 - 1 + lj, gk = DIM == 3 ? k0 - 1 + lk : 0;
        T v = 0;
        if (gi >= 0 && gi < g.NX && gj >= 0 && gj < g.NY && gk >= 0 && gk < g.NZ) {
            const int64_t n = gi + (int64_t)g.NX * (gj + (int64_t)g.NY * gk);
            const int64_t dof = DIM * n + d;
            T val;
            if (PCG)
                val = dinv[dof] * xin[dof] + beta * p_old[dof];
            else
                val = xin[dof];
            const bool owned = li >= 1 && li <= TX && lj >= 1 && lj <= TY &&
                               (DIM == 2 || (lk >= 1 && lk <= TZ));
            if (owned) {
                const int o = (li - 1) + TX * ((lj - 1) + TY * (DIM == 3 ? lk - 1 : 0));
                pown[d][o] = val;
                if (PCG) p_new[dof] = val;
            }
            v = ((fixmask[n] >> d) & 1) ? (T)0 : val;
        }
        xs[d][nl] = v;
    }
    __syncthreads();

    // ---- cell products into per-corner slots ----
    for (int el = tid; el < NEL; el += kThreads) {
        const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
        const int gi = i0 - 1 + ei, gj = j0 - 1 + ej, gk = DIM == 3 ? k0 - 1 + ek : 0;
        if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || (DIM == 3 && (gk < 0 || gk >= g.nz)))
            continue;
        T xe[DIM][NC], ye[DIM][NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
            const int nl = (ei + bx) + IX * ((ej + by) + IY * (DIM == 3 ? ek + bz : 0));
#pragma unroll
            for (int d = 0; d < DIM; ++d) xe[d][c] = xs[d][nl];
        }
        elem_apply<T, DIM>(op, xe, ye);
        const T Ee = E[gi + (int64_t)g.nx * (gj + (int64_t)g.ny * gk)];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
            const int oi = ei + bx - 1, oj = ej + by - 1, ok = DIM == 3 ? ek + bz - 1 : 0;
            if (oi >= 0 && oi < TX && oj >= 0 && oj < TY && ok >= 0 && ok < TZ) {
                const int o = oi + TX * (oj + TY * ok);
#pragma unroll
                for (int d = 0; d < DIM; ++d) slot[c][d][o] = Ee * ye[d][c];
            }
        }
    }
    __syncthreads();

    // ---- per-node gather in ascending cell order ----
    double part = 0.0;
    for (int o = tid; o < NOUT; o += kThreads) {
        const int oi = o % TX, oj = (o / TX) % TY, ok = DIM == 3 ? o / (TX * TY) : 0;
        const int gi = i0 + oi, gj = j0 + oj, gk = k0 + ok;
        if (gi >= g.NX || gj >= g.NY || gk >= g.NZ) continue;
        const int64_t n = gi + (int64_t)g.NX * (gj + (int64_t)g.NY * gk);
        const uint8_t fm = fixmask[n];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            T s = 0;
            if ((fm >> d) & 1) {
                s = (T)meandiag * pown[d][o];
            } else {
#pragma unroll
                for (int c = NC - 1; c >= 0; --c) {
                    const int ci = gi - (c & 1), cj = gj - ((c >> 1) & 1),
                              ck = DIM == 3 ? gk - ((c >> 2) & 1) : 0;
                    if (ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny &&
                        (DIM == 2 || (ck >= 0 && ck < g.nz)))
                        s += slot[c][d][o];
                }
            }
            y[DIM * n + d] = s;
            if (PCG) part += (double)pown[d][o] * (double)s;
        }
    }

    if (PCG) {
        double bs = block_sum(part, red);
        if (publish_partial(bs, partials, blockIdx.x, &sc->counter[0], gridDim.x)) {
            double pq = sum_partials(partials, gridDim.x, 1, 0, red);
            if (tid == 0) {
                sc->pq = pq;
                sc->alpha = sc->rho / pq;
                sc->counter[0] = 0;
            }
        }
    }
}

template <typename T>
hipError_t launch_matvec(const Grid& g, const ElemOp<T>& op, const T* x, const uint8_t* fixmask,
                         const T* E, T* y, double meandiag, hipStream_t s) {
    if (g.dim == 3) {
        hipLaunchKernelGGL((k_matvec<T, 3, false>), dim3(ntiles<3>(g)), dim3(kThreads), 0, s, g,
                           op, x, (const T*)nullptr, (const T*)nullptr, (T*)nullptr, fixmask, E,
                           y, meandiag, (PcgScalars*)nullptr, (double*)nullptr);
    } else {
        hipLaunchKernelGGL((k_matvec<T, 2, false>), dim3(ntiles<2>(g)), dim3(kThreads), 0, s, g,
                           op, x, (const T*)nullptr, (const T*)nullptr, (T*)nullptr, fixmask, E,
                           y, meandiag, (PcgScalars*)nullptr, (double*)nullptr);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_dirmatvec(const Grid& g, const ElemOp<T>& op, const T* r, const T* dinv,
                                const T* p_old, T* p_new, T* q, const uint8_t* fixmask,
                                const T* E, double meandiag, PcgScalars* sc, double* partials,
                                hipStream_t s) {
    int nb = g.dim == 3 ? ntiles<3>(g) : ntiles<2>(g);
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    if (g.dim == 3) {
        hipLaunchKernelGGL((k_matvec<T, 3, true>), dim3(nb), dim3(kThreads), 0, s, g, op, r,
                           dinv, p_old, p_new, fixmask, E, q, meandiag, sc, partials);
    } else {
        hipLaunchKernelGGL((k_matvec<T, 2, true>), dim3(nb), dim3(kThreads), 0, s, g, op, r,
                           dinv, p_old, p_new, fixmask, E, q, meandiag, sc, partials);
    }
    return hipGetLastError();
}

template hipError_t launch_matvec<double>(const Grid&, const ElemOp<double>&, const double*,
                                          const uint8_t*, const double*, double*, double,
                                          hipStream_t);
template hipError_t launch_matvec<float>(const Grid&, const ElemOp<float>&, const float*,
                                         const uint8_t*, const float*, float*, double,
                                         hipStream_t);
template hipError_t launch_pcg_dirmatvec<double>(const Grid&, const ElemOp<double>&,
                                                 const double*, const double*, const double*,
                                                 double*, double*, const uint8_t*, const double*,
                                                 double, PcgScalars*, double*, hipStream_t);
template hipError_t launch_pcg_dirmatvec<float>(const Grid&, const ElemOp<float>&,
                                                const float*, const float*, const float*, float*,
                                                float*, const uint8_t*, const float*, double,
                                                PcgScalars*, double*, hipStream_t);

// ---- Jacobi inverse diagonal and SIMP densities ----

template <typename T, int DIM>
__global__ __launch_bounds__(kThreads) void k_jacobi_inv(Grid g, ElemOp<double> op,
                                                         const double* __restrict__ E,
                                                         const uint8_t* __restrict__ fixmask,
                                                         double meandiag, T* __restrict__ dinv) {
    const int64_t n = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (n >= g.nnodes) return;
    const int gi = n % g.NX, gj = (n / g.NX) % g.NY, gk = DIM == 3 ? n / ((int64_t)g.NX * g.NY) : 0;
    double se = 0.0;
    for (int c = (1 << DIM) - 1; c >= 0; --c) {
        const int ci = gi - (c & 1), cj = gj - ((c >> 1) & 1), ck = DIM == 3 ? gk - ((c >> 2) & 1) : 0;
        if (ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny && (DIM == 2 || (ck >= 0 && ck < g.nz)))
            se += E[ci + (int64_t)g.nx * (cj + (int64_t)g.ny * ck)];
    }
    const uint8_t fm = fixmask[n];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        double dg = ((fm >> d) & 1) ? meandiag : op.kdiag[d] * se;
        dinv[DIM * n + d] = (T)(1.0 / dg);
    }
}

template <typename T>
hipError_t launch_jacobi_inv(const Grid& g, const ElemOp<double>& op, const double* E,
                             const uint8_t* fixmask, double meandiag, T* dinv, hipStream_t s) {
    dim3 grid((unsigned)((g.nnodes + kThreads - 1) / kThreads));
    if (g.dim == 3)
        hipLaunchKernelGGL((k_jacobi_inv<T, 3>), grid, dim3(kThreads), 0, s, g, op, E, fixmask,
                           meandiag, dinv);
    else
        hipLaunchKernelGGL((k_jacobi_inv<T, 2>), grid, dim3(kThreads), 0, s, g, op, E, fixmask,
                           meandiag, dinv);
    return hipGetLastError();
}
template hipError_t launch_jacobi_inv<double>(const Grid&, const ElemOp<double>&, const double*,
                                              const uint8_t*, double, double*, hipStream_t);
template hipError_t launch_jacobi_inv<float>(const Grid&, const ElemOp<double>&, const double*,
                                             const uint8_t*, double, float*, hipStream_t);

// rho_e = xmin + (1-xmin) x^p  (penalty before interpolation) or
//         (xmin + (1-xmin) x)^p  (TopOpt.jl src/Utilities/penalties.jl:113-119).
__global__ __launch_bounds__(kThreads) void k_density(int64_t nel, const double* __restrict__ xt,
                                                      double xmin, double p, int pbi,
                                                      double* __restrict__ E64,
                                                      float* __restrict__ E32) {
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (e >= nel) return;
    const double x = xt[e];
    double r;
    if (pbi)
        r = pow(x, p) * (1.0 - xmin) + xmin;
    else
        r = pow(x * (1.0 - xmin) + xmin, p);
    E64[e] = r;
    if (E32) E32[e] = (float)r;
}

hipError_t launch_density(int64_t nel, const double* xt, double xmin, double p, int pbi,
                          double* E64, float* E32, hipStream_t s) {
    hipLaunchKernelGGL(k_density, dim3((unsigned)((nel + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, nel, xt, xmin, p, pbi, E64, E32);
    return hipGetLastError();
}

}  // namespace tohip
