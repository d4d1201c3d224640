// Jacobi-PCG vector kernels (replaces the vector part of cg_solve! ->
// IterativeSolvers.cg!, TopOpt.jl src/FEA/solvers_api.jl:177-220).
//
// The iteration follows the library's preconditioned-CG order:
//   z = D^-1 r; rho = z.r; beta = rho/rho_prev; p = z + beta p;
//   q = K p; alpha = rho/(p.q); x += alpha p; r -= alpha q; stop on ||r||.
// Direction update, stopping rule and q = K p are fused in
// k_matvec<PCG=true>; this file holds the start kernel and the update kernel.
// Each kernel leaves one partial per block; the next kernel re-reduces them in
// a fixed order (see k_common.h), so a solve is bitwise reproducible.  Once
// `done` is set every later PCG kernel returns at once, which lets the host
// replay a captured multi-iteration graph without checking in between.
#include "k_common.h"

namespace tohip {

template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_start(int64_t n, const T* __restrict__ b,
                                                        const T* __restrict__ Ax,
                                                        T* __restrict__ r,
                                                        const T* __restrict__ dinv,
                                                        T* __restrict__ p0, PcgScalars* sc,
                                                        double* __restrict__ partB, int keep) {
    __shared__ double red[kWaves];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const T ri = Ax ? (T)(b[i] - Ax[i]) : b[i];
        r[i] = ri;
        if (!keep) p0[i] = (T)0;
        const double rd = (double)ri;
        rz += (double)(dinv[i] * ri) * rd;
        rr += rd * rd;
    }
    const double s1 = block_sum_all(rz, red);
    const double s2 = block_sum_all(rr, red);
    if (threadIdx.x == 0) {
        partB[2 * (size_t)blockIdx.x] = s1;
        partB[2 * (size_t)blockIdx.x + 1] = s2;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->done = 0;
        sc->err = 0;
        if (!keep) {
            sc->iter = 0;
            sc->k_cur = 0;
            sc->rho[0] = 1.0;  // IterativeSolvers starts from rho_prev = 1 (multiplies p = 0)
            sc->rho[1] = 1.0;
            sc->tol = -1.0;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_update(int64_t n, T* __restrict__ x,
                                                         const T* rin, T* r,
                                                         const T* __restrict__ p0,
                                                         const T* __restrict__ p1,
                                                         const T* __restrict__ q,
                                                         const T* __restrict__ dinv,
                                                         PcgScalars* sc,
                                                         const double* __restrict__ partA,
                                                         int nbA, double* __restrict__ partB) {
    __shared__ double red[kWaves];
    if (sc->done) return;
    const long long k = sc->k_cur;
    const T* __restrict__ p = (k & 1) ? p1 : p0;
    // The first kChunk elements of every thread are loaded before alpha is
    // known, overlapping their latency with the reduction of the partials.
    constexpr int kChunk = 4;
    // 32-bit indices (vectors hold fewer than 2^29 entries, see tohip_create)
    const uint32_t n32 = (uint32_t)n;
    const uint32_t i0 = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
    const uint32_t stride = gridDim.x * (uint32_t)kThreads;
    // dinv == nullptr: unit diagonal (the multigrid path, whose preconditioner
    // is the V-cycle), not read
    const bool unit = dinv == nullptr;
    T cp[kChunk], cq[kChunk], cr[kChunk], cd[kChunk], cx[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
        const uint32_t i = i0 + c * stride;
        if (i < n32) {
            cp[c] = at(p, i);
            cq[c] = at(q, i);
            cr[c] = at(rin, i);
            cd[c] = unit ? (T)1 : at(dinv, i);
            cx[c] = at((const T*)x, i);
        }
    }
    const double pq = reduce_partials(partA, nbA, 1, 0, red);
    // Breakdown: p.Kp must be positive and finite for a positive definite K
    // (it is not for, e.g., a hard-void design with xmin = 0).  Every block
    // sees the same pq and leaves x and r as they are; block 0 records it.
    if (!(pq > 0.0 && pq < INFINITY)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->err = kPcgErrIndefinite;
            sc->done = 1;
        }
        return;
    }
    const double alpha = sc->rho_cur / pq;
    const T a = (T)alpha;
    double rz = 0.0, rr = 0.0;
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
        const uint32_t i = i0 + c * stride;
        if (i < n32) {
            at(x, i) = cx[c] + a * cp[c];
            const T ri = cr[c] - a * cq[c];
            at(r, i) = ri;
            const double rd = (double)ri;
            rz += (double)(cd[c] * ri) * rd;
            rr += rd * rd;
        }
    }
#pragma unroll 4
    for (uint32_t i = i0 + kChunk * stride; i < n32; i += stride) {
        const T pi = at(p, i), qi = at(q, i), ri0 = at(rin, i),
                di = unit ? (T)1 : at(dinv, i), xi = at((const T*)x, i);
        at(x, i) = xi + a * pi;
        const T ri = ri0 - a * qi;
        at(r, i) = ri;
        const double rd = (double)ri;
        rz += (double)(di * ri) * rd;
        rr += rd * rd;
    }
    const double s1 = block_sum_all(rz, red);
    const double s2 = block_sum_all(rr, red);
    if (threadIdx.x == 0) {
        partB[2 * (size_t)blockIdx.x] = s1;
        partB[2 * (size_t)blockIdx.x + 1] = s2;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->iter = k;
        sc->alpha = alpha;
    }
}

template <typename T>
hipError_t launch_pcg_start(int64_t n, const T* b, const T* Ax, T* r, const T* dinv, T* p0,
                            PcgScalars* sc, double* partB, int keep, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_start<T>), dim3(kVecBlocks), dim3(kThreads), 0, s, n, b, Ax, r,
                       dinv, p0, sc, partB, keep);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update(int64_t n, T* x, T* r, const T* p0, const T* p1, const T* q,
                             const T* dinv, PcgScalars* sc, const double* partA, int nbA,
                             double* partB, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_update<T>), dim3(kVecBlocks), dim3(kThreads), 0, s, n, x, r, r, p0,
                       p1, q, dinv, sc, partA, nbA, partB);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update_from(int64_t n, T* x, const T* rin, T* r, const T* p0, const T* p1,
                                  const T* q, const T* dinv, PcgScalars* sc, const double* partA,
                                  int nbA, double* partB, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_update<T>), dim3(kVecBlocks), dim3(kThreads), 0, s, n, x, rin, r,
                       p0, p1, q, dinv, sc, partA, nbA, partB);
    return hipGetLastError();
}

template hipError_t launch_pcg_start<double>(int64_t, const double*, const double*, double*,
                                             const double*, double*, PcgScalars*, double*, int,
                                             hipStream_t);
template hipError_t launch_pcg_start<float>(int64_t, const float*, const float*, float*,
                                            const float*, float*, PcgScalars*, double*, int,
                                            hipStream_t);
template hipError_t launch_pcg_update<double>(int64_t, double*, double*, const double*,
                                              const double*, const double*, const double*,
                                              PcgScalars*, const double*, int, double*,
                                              hipStream_t);
template hipError_t launch_pcg_update<float>(int64_t, float*, float*, const float*, const float*,
                                             const float*, const float*, PcgScalars*,
                                             const double*, int, double*, hipStream_t);
template hipError_t launch_pcg_update_from<float>(int64_t, float*, const float*, float*,
                                                  const float*, const float*, const float*,
                                                  const float*, PcgScalars*, const double*, int,
                                                  double*, hipStream_t);

// ---- EnergyCriteria (TopOpt.jl src/FEA/convergence_criteria.jl:26-45) ----
// IterativeSolvers evaluates the criterion before every iteration on the
// current x and recursive residual r: x'Ax = f.x - x.r; converged when
// |x'Ax - energy| / x'Ax <= tol and |x.r| / x'Ax <= tol; energy := x'Ax.
// Its stopping test is `iteration >= maxiter || converged`, so the criterion
// is not evaluated once maxiter iterations are done.
__global__ __launch_bounds__(kThreads) void k_ecrit_partials(int64_t n,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ f,
                                                             const PcgScalars* sc,
                                                             double* __restrict__ partials) {
    __shared__ double red[kWaves];
    if (sc->done || sc->iter >= sc->maxiter) return;
    double xr = 0.0, fx = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double xi = x[i];
        xr += xi * r[i];
        fx += f[i] * xi;
    }
    const double s1 = block_sum_all(xr, red);
    const double s2 = block_sum_all(fx, red);
    if (threadIdx.x == 0) {
        partials[2 * (size_t)blockIdx.x] = s1;
        partials[2 * (size_t)blockIdx.x + 1] = s2;
    }
}

__global__ __launch_bounds__(kThreads) void k_ecrit_decide(int nb,
                                                           const double* __restrict__ partials,
                                                           PcgScalars* sc, EnergyCrit* ec) {
    __shared__ double red[kWaves];
    if (sc->done || sc->iter >= sc->maxiter) return;
    const double xtr = reduce_partials(partials, nb, 2, 0, red);
    const double fx = reduce_partials(partials, nb, 2, 1, red);
    if (threadIdx.x != 0) return;
    const double xAx = fx - xtr;
    const double change = xAx - ec->energy;
    ec->xtr = xtr;
    ec->xAx = xAx;
    if (isnan(change) || isnan(xAx)) {
        ec->err = 1;
        sc->done = 1;
        return;
    }
    if (xAx < 0.0) {
        ec->err = 2;
        sc->done = 1;
        return;
    }
    // x'Ax = 0 gives NaN or Inf ratios: not converged, as in the reference
    const bool conv = fabs(change) / xAx <= ec->tol && fabs(xtr) / xAx <= ec->tol;
    ec->energy = xAx;
    if (conv) {
        ec->converged = 1;
        sc->done = 1;
    }
}

hipError_t launch_energy_check(int64_t n, const double* x, const double* r, const double* f,
                               PcgScalars* sc, EnergyCrit* ec, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_ecrit_partials, dim3(kVecBlocks), dim3(kThreads), 0, s, n, x, r, f, sc,
                       partials);
    hipLaunchKernelGGL(k_ecrit_decide, dim3(1), dim3(kThreads), 0, s, kVecBlocks, partials, sc,
                       ec);
    return hipGetLastError();
}

// ---- one-block finisher of a partial-sum array ----
__global__ __launch_bounds__(1024) void k_reduce(const double* __restrict__ partials, int nb,
                                                 int nv, double* __restrict__ out) {
    __shared__ double red[16];
    for (int v = 0; v < nv; ++v) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partials[(size_t)i * nv + v];
        const double t = block_sum_all(s, red);
        if (threadIdx.x == 0) out[v] = t;
    }
}

hipError_t launch_reduce(const double* partials, int nb, int nv, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, partials, nb, nv, out);
    return hipGetLastError();
}

// ---- mixed-precision residual replacement ----

// u += x32, x32 = 0.
__global__ __launch_bounds__(kThreads) void k_accumulate_f32(int64_t n, float* __restrict__ x32,
                                                             double* __restrict__ u) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        u[i] += (double)x32[i];
        x32[i] = 0.0f;
    }
}

hipError_t launch_accumulate_f32(int64_t n, float* x32, double* u, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate_f32, dim3(kVecBlocks), dim3(kThreads), 0, s, n, x32, u);
    return hipGetLastError();
}

// r64 = b - Ku; r32 = (float) r64 with partials (r32.dinv r32, r32.r32) in
// partB (stride 2, the layout the direction kernel reads) and r64.r64 in
// partB[2*kVecBlocks + block].
__global__ __launch_bounds__(kThreads) void k_residual_refresh(
    int64_t n, const double* __restrict__ b, const double* __restrict__ Ku,
    double* __restrict__ r64, float* __restrict__ r32, const float* __restrict__ dinv32,
    double* __restrict__ partB) {
    __shared__ double red[kWaves];
    double rz = 0.0, rr = 0.0, r2 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double ri = b[i] - Ku[i];
        r64[i] = ri;
        r2 += ri * ri;
        if (r32) {
            const float rf = (float)ri;
            r32[i] = rf;
            const double rd = (double)rf;
            rz += (double)(dinv32[i] * rf) * rd;
            rr += rd * rd;
        }
    }
    const double s1 = block_sum_all(rz, red);
    const double s2 = block_sum_all(rr, red);
    const double s3 = block_sum_all(r2, red);
    if (threadIdx.x == 0) {
        partB[2 * (size_t)blockIdx.x] = s1;
        partB[2 * (size_t)blockIdx.x + 1] = s2;
        partB[2 * (size_t)kVecBlocks + blockIdx.x] = s3;
    }
}

hipError_t launch_residual_refresh(int64_t n, const double* b, const double* Ku, double* r64,
                                   float* r32, const float* dinv32, double* partB,
                                   hipStream_t s) {
    hipLaunchKernelGGL(k_residual_refresh, dim3(kVecBlocks), dim3(kThreads), 0, s, n, b, Ku, r64,
                       r32, dinv32, partB);
    return hipGetLastError();
}

// After a residual replacement: continue the FP32 iteration until its
// recursive residual drops to max(factor * ||r64||, floor_tol).  The search
// direction and rho are kept (reliable updates), unless the true residual
// has come out more than `restart` times the norm of the recursive one
// (sc->rr, the last direction kernel's r.r): the FP32 iteration had run past
// what FP32 can resolve, and p, scaled for the recursive residual, would
// outweigh the new z through beta = rho / rho_prev for many iterations.  Then
// the direction restarts: rho_prev = +inf makes the next beta exactly 0, so
// p = z.  (restart <= 0: never.)
__global__ void k_pcg_rearm(PcgScalars* sc, const double* __restrict__ rr64, double factor,
                            double floor_tol, double restart) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->done = 0;
        sc->tol = fmax(factor * sqrt(*rr64), floor_tol);
        if (restart > 0 && *rr64 > restart * restart * sc->rr) sc->rho[sc->iter & 1] = INFINITY;
    }
}

hipError_t launch_pcg_rearm(PcgScalars* sc, const double* rr64, double factor, double floor_tol,
                            double restart, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_rearm, dim3(1), dim3(64), 0, s, sc, rr64, factor, floor_tol,
                       restart);
    return hipGetLastError();
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        b[i] = (float)a[i];
}

hipError_t launch_f64_to_f32(int64_t n, const double* a, float* b, hipStream_t s) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(kVecBlocks), dim3(kThreads), 0, s, n, a, b);
    return hipGetLastError();
}

__global__ void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        b[i] = (double)a[i];
}

hipError_t launch_f32_to_f64(int64_t n, const float* a, double* b, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(kVecBlocks), dim3(kThreads), 0, s, n, a, b);
    return hipGetLastError();
}

// x = v on n entries (the NaN result of a broken-down solve).
__global__ __launch_bounds__(kThreads) void k_fill64(int64_t n, double v, double* __restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        x[i] = v;
}

hipError_t launch_fill64(int64_t n, double v, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_fill64, dim3(kVecBlocks), dim3(kThreads), 0, s, n, v, x);
    return hipGetLastError();
}

// Plain streaming copy b = a (16 bytes moved per element): the reference
// kernel of known traffic against which hardware byte counters are scaled.
__global__ __launch_bounds__(kThreads) void k_copy64(int64_t n, const double* __restrict__ a,
                                                     double* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        b[i] = a[i];
}

hipError_t launch_copy64(int64_t n, const double* a, double* b, hipStream_t s) {
    hipLaunchKernelGGL(k_copy64, dim3(4 * kVecBlocks), dim3(kThreads), 0, s, n, a, b);
    return hipGetLastError();
}

// partials of sum_i x_i (kVecBlocks blocks) then the one-block finisher.
__global__ __launch_bounds__(kThreads) void k_sum_part(int64_t n, const double* __restrict__ x,
                                                       double* __restrict__ partials) {
    __shared__ double red[kWaves];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        s += x[i];
    const double b = block_sum_all(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

hipError_t launch_sum(int64_t n, const double* x, double* partials, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_part, dim3(kVecBlocks), dim3(kThreads), 0, s, n, x, partials);
    return launch_reduce(partials, kVecBlocks, 1, out, s);
}

}  // namespace tohip
