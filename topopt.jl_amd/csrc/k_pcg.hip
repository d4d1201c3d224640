// Jacobi-PCG vector kernels (replaces the vector part of cg_solve! ->
// IterativeSolvers.cg!, TopOpt.jl src/FEA/solvers_api.jl:177-220).
//
// The iteration follows the library's preconditioned-CG order:
//   z = D^-1 r; rho = z.r; beta = rho/rho_prev; p = z + beta p;
//   q = K p; alpha = rho/(p.q); x += alpha p; r -= alpha q; stop on ||r||.
// The direction update and q = K p are fused in k_matvec<PCG=true>; this file
// holds the start kernel (r, first rho, tolerance) and the update kernel
// (x, r, next rho, ||r||^2, beta, convergence flag).  Scalars live on the
// device in PcgScalars and are finalised by the last block to arrive, with a
// fixed summation order, so a solve is bitwise reproducible.  Once `done` is
// set, every later PCG kernel returns immediately, which lets the host
// replay a captured multi-iteration graph without checking in between.
#include "k_common.h"

namespace tohip {

static inline unsigned vec_blocks(int64_t n) {
    int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (unsigned)b;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_start(int64_t n, const T* __restrict__ b,
                                                        const T* __restrict__ Ax,
                                                        T* __restrict__ r,
                                                        const T* __restrict__ dinv,
                                                        T* __restrict__ p_old, PcgScalars* sc,
                                                        double* partials) {
    __shared__ double red[kWaves];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        T ri = Ax ? (T)(b[i] - Ax[i]) : b[i];
        r[i] = ri;
        p_old[i] = (T)0;
        const double rd = (double)ri;
        rz += (double)(dinv[i] * ri) * rd;
        rr += rd * rd;
    }
    double s1 = block_sum(rz, red);
    double s2 = block_sum(rr, red);
    if (threadIdx.x == 0) partials[2 * (size_t)blockIdx.x + 1] = s2;
    if (publish_partial(s1, partials, 2 * blockIdx.x, &sc->counter[1], gridDim.x)) {
        double tz = sum_partials(partials, gridDim.x, 2, 0, red);
        double t2 = sum_partials(partials, gridDim.x, 2, 1, red);
        if (threadIdx.x == 0) {
            sc->rho = tz;
            sc->beta = tz;  // rho / rho_prev with rho_prev = 1; multiplies p_old = 0
            sc->rr = t2;
            sc->rr0 = t2;
            const double res = sqrt(t2);
            const double tol = fmax(sc->reltol * res, sc->abstol);
            sc->tol = tol;
            sc->iter = 0;
            sc->done = (res <= tol || sc->maxiter <= 0) ? 1 : 0;
            sc->counter[1] = 0;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_update(int64_t n, T* __restrict__ x,
                                                         T* __restrict__ r,
                                                         const T* __restrict__ p,
                                                         const T* __restrict__ q,
                                                         const T* __restrict__ dinv,
                                                         PcgScalars* sc, double* partials) {
    __shared__ double red[kWaves];
    if (__hip_atomic_load(&sc->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const double alpha = sc->alpha;
    const T a = (T)alpha;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        const double rd = (double)ri;
        rz += (double)(dinv[i] * ri) * rd;
        rr += rd * rd;
    }
    double s1 = block_sum(rz, red);
    double s2 = block_sum(rr, red);
    if (threadIdx.x == 0) partials[2 * (size_t)blockIdx.x + 1] = s2;
    if (publish_partial(s1, partials, 2 * blockIdx.x, &sc->counter[1], gridDim.x)) {
        double tz = sum_partials(partials, gridDim.x, 2, 0, red);
        double t2 = sum_partials(partials, gridDim.x, 2, 1, red);
        if (threadIdx.x == 0) {
            sc->beta = tz / sc->rho;
            sc->rho = tz;
            sc->rr = t2;
            sc->iter += 1;
            const double res = sqrt(t2);
            sc->done = (res <= sc->tol || sc->iter >= sc->maxiter) ? 1 : 0;
            sc->counter[1] = 0;
        }
    }
}

template <typename T>
hipError_t launch_pcg_start(int64_t n, const T* b, const T* Ax, T* r, const T* dinv, T* p_old,
                            PcgScalars* sc, double* partials, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_start<T>), dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, b, Ax,
                       r, dinv, p_old, sc, partials);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update(int64_t n, T* x, T* r, const T* p, const T* q, const T* dinv,
                             PcgScalars* sc, double* partials, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_update<T>), dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, x, r,
                       p, q, dinv, sc, partials);
    return hipGetLastError();
}

template hipError_t launch_pcg_start<double>(int64_t, const double*, const double*, double*,
                                             const double*, double*, PcgScalars*, double*,
                                             hipStream_t);
template hipError_t launch_pcg_start<float>(int64_t, const float*, const float*, float*,
                                            const float*, float*, PcgScalars*, double*,
                                            hipStream_t);
template hipError_t launch_pcg_update<double>(int64_t, double*, double*, const double*,
                                              const double*, const double*, PcgScalars*,
                                              double*, hipStream_t);
template hipError_t launch_pcg_update<float>(int64_t, float*, float*, const float*,
                                             const float*, const float*, PcgScalars*, double*,
                                             hipStream_t);

// ---- mixed-precision refinement helpers ----

// r = f - Ku (FP64), ||r||^2 -> *rr_out (last block).
__global__ __launch_bounds__(kThreads) void k_residual64(int64_t n, const double* __restrict__ f,
                                                         const double* __restrict__ Ku,
                                                         double* __restrict__ r,
                                                         double* partials, double* rr_out,
                                                         unsigned int* counter) {
    __shared__ double red[kWaves];
    double rr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double ri = f[i] - Ku[i];
        r[i] = ri;
        rr += ri * ri;
    }
    double s = block_sum(rr, red);
    if (publish_partial(s, partials, blockIdx.x, counter, gridDim.x)) {
        double t = sum_partials(partials, gridDim.x, 1, 0, red);
        if (threadIdx.x == 0) {
            *rr_out = t;
            *counter = 0;
        }
    }
}

hipError_t launch_residual64(int64_t n, const double* f, const double* Ku, double* r,
                             double* partials, double* rr_out, unsigned int* counter,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_residual64, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, f, Ku, r,
                       partials, rr_out, counter);
    return hipGetLastError();
}

// out = r / ||r|| in FP32 (scaling keeps the inner system O(1) in FP32).
__global__ void k_scale_to_f32(int64_t n, const double* __restrict__ r,
                               const double* __restrict__ norm2, float* __restrict__ out) {
    const double s = *norm2 > 0.0 ? 1.0 / sqrt(*norm2) : 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        out[i] = (float)(r[i] * s);
}

hipError_t launch_scale_to_f32(int64_t n, const double* r, const double* norm2, float* out,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_scale_to_f32, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, r, norm2,
                       out);
    return hipGetLastError();
}

// u += ||r|| * d (FP64 accumulation of the FP32 correction).
__global__ void k_add_scaled_f32(int64_t n, const float* __restrict__ d,
                                 const double* __restrict__ norm2, double* __restrict__ u) {
    const double s = sqrt(*norm2);
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        u[i] += s * (double)d[i];
}

hipError_t launch_add_scaled_f32(int64_t n, const float* d, const double* norm2, double* u,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_add_scaled_f32, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, d, norm2,
                       u);
    return hipGetLastError();
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        b[i] = (float)a[i];
}

hipError_t launch_f64_to_f32(int64_t n, const double* a, float* b, hipStream_t s) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, a, b);
    return hipGetLastError();
}

__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b,
                        double* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        y[i] = a * x[i] + b * y[i];
}

hipError_t launch_axpby(int64_t n, double a, const double* x, double b, double* y,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, a, x, b, y);
    return hipGetLastError();
}

}  // namespace tohip

namespace tohip {

// s = sum_i x_i (deterministic), written to *out by the last block.
__global__ __launch_bounds__(kThreads) void k_sum(int64_t n, const double* __restrict__ x,
                                                  double* partials, double* out,
                                                  unsigned int* counter) {
    __shared__ double red[kWaves];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        s += x[i];
    double b = block_sum(s, red);
    if (publish_partial(b, partials, blockIdx.x, counter, gridDim.x)) {
        double t = sum_partials(partials, gridDim.x, 1, 0, red);
        if (threadIdx.x == 0) {
            *out = t;
            *counter = 0;
        }
    }
}

hipError_t launch_sum(int64_t n, const double* x, double* partials, double* out,
                      unsigned int* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_sum, dim3(vec_blocks(n)), dim3(kThreads), 0, s, n, x, partials, out,
                       counter);
    return hipGetLastError();
}

}  // namespace tohip
