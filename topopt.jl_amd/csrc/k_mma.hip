// MMA87 design update on the device (Svanberg 1987, "The method of moving
// asymptotes", the optimizer TopOpt.jl's SIMP tutorials drive through
// NonconvexMMA: docs/tutorials/simp.qmd:103-114) for min c(x) subject to
// the one volume constraint g(x) = V(J x) - volfrac <= 0, 0 <= x <= 1.
//
// At iterate x with asymptotes L < x < U the convex approximation is
//   c~(y) = sum_j p0_j / (U_j - y_j) + q0_j / (y_j - L_j) + const,
//   g~(y) = g(x) + sum_j p1_j (1 / (U_j - y_j) - 1 / (U_j - x_j)),
//   p0 = (U - x)^2 max(dc, 0), q0 = (x - L)^2 max(-dc, 0),
//   p1 = (U - x)^2 dg (dg = colsum(J) / nel >= 0, so q1 = 0),
// on the box [alpha, beta] = [max(0, 0.9 L + 0.1 x), min(1, 0.9 U + 0.1 x)].
// With one constraint the dual is one-dimensional: for a multiplier lam the
// Lagrangian's minimiser is, per cell, y = (L sqrt(P) + U sqrt(Q)) /
// (sqrt(P) + sqrt(Q)) clamped to the box, P = p0 + lam p1, Q = q0, and
// F(lam) = sum_j p1_j / (U_j - y_j(lam)) is non-increasing in lam; the dual
// optimum is where g~ = 0, i.e. F(lam) = T = sum_j p1_j / (U_j - x_j) - g(x)
// (lam -> 0 when the constraint is inactive there).  As the OC update does,
// each pass evaluates F at kOcCand geometrically spaced multipliers of the
// current bracket in one sweep (per-block partials) and a one-block kernel
// keeps the sub-bracket holding the crossing.
//
// Asymptotes (1987 rule): the first two iterations L, U = x -+ s_init; then
// L = x - s (x_prev - L_prev), U = x + s (U_prev - x_prev) with s = s_decr
// when the last two steps oscillate, s_incr when they keep their sign.
#include <algorithm>

#include "k_common.h"

namespace tohip {

constexpr int kMmaCand = kOcCandidates;
constexpr int kMmaBlocks = 256;
constexpr int kMmaSelThreads = 1024;
constexpr int kMmaGroups = kMmaSelThreads / kMmaCand;
constexpr double kMmaLamLo = 1e-40, kMmaLamHi = 1e15;

__device__ __forceinline__ double mma_cand(double lo, double hi, int j) {
    return lo * exp(log(hi / lo) * (double)(j + 1) / (double)(kMmaCand + 1));
}

// Per-cell subproblem data from the iterate, its sensitivities and the asymptotes.
struct MmaCell {
    double L, U, a, b, p0, q0, p1;
};

__device__ __forceinline__ MmaCell mma_cell(double x, double dc, double dg, double L, double U) {
    MmaCell c;
    c.L = L;
    c.U = U;
    c.a = fmax(0.0, 0.9 * L + 0.1 * x);
    c.b = fmin(1.0, 0.9 * U + 0.1 * x);
    const double ux = U - x, xl = x - L;
    c.p0 = ux * ux * fmax(dc, 0.0);
    c.q0 = xl * xl * fmax(-dc, 0.0);
    c.p1 = ux * ux * dg;
    return c;
}

__device__ __forceinline__ double mma_y(const MmaCell& c, double lam, double x) {
    const double P = c.p0 + lam * c.p1, Q = c.q0;
    const double sp = sqrt(P), sq = sqrt(Q);
    const double y = (sp + sq) > 0.0 ? (c.L * sp + c.U * sq) / (sp + sq) : x;
    return fmax(c.a, fmin(c.b, y));
}

// Asymptote update and the partial sums of T: per block (sum colsum (U - x),
// sum colsum x) / nel.
__global__ __launch_bounds__(kThreads) void k_mma_prepare(
    int64_t nel, int64_t iter, const double* __restrict__ x, const double* __restrict__ x1,
    const double* __restrict__ x2, double* __restrict__ L, double* __restrict__ U,
    const double* __restrict__ colsum, double s_init, double s_incr, double s_decr,
    double* __restrict__ partials) {
    __shared__ double red[kWaves];
    const double inel = 1.0 / (double)nel;
    double s1 = 0.0, s2 = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * kThreads) {
        const double xe = x[e];
        double l, u;
        if (iter < 2) {
            l = xe - s_init;
            u = xe + s_init;
        } else {
            const double osc = (xe - x1[e]) * (x1[e] - x2[e]);
            const double s = osc < 0.0 ? s_decr : (osc > 0.0 ? s_incr : 1.0);
            l = xe - s * (x1[e] - L[e]);
            u = xe + s * (U[e] - x1[e]);
        }
        L[e] = l;
        U[e] = u;
        const double cs = colsum[e] * inel;
        s1 += cs * (u - xe);
        s2 += cs * xe;
    }
    const double b1 = block_sum_all(s1, red);
    const double b2 = block_sum_all(s2, red);
    if (threadIdx.x == 0) {
        partials[2 * (size_t)blockIdx.x] = b1;
        partials[2 * (size_t)blockIdx.x + 1] = b2;
    }
}

// T = sum p1 / (U - x) - g(x) = S1 - (S2 - volfrac); the first candidate table.
__global__ void k_mma_init(int nb, const double* __restrict__ partials, double volfrac,
                           MmaState* st) {
    __shared__ double tot[2];
    if (threadIdx.x < 2) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partials[2 * (size_t)b + threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < kMmaCand) st->cand[threadIdx.x] = mma_cand(kMmaLamLo, kMmaLamHi, threadIdx.x);
    if (threadIdx.x == 0) {
        st->lo = kMmaLamLo;
        st->hi = kMmaLamHi;
        st->lambda = 0.5 * (kMmaLamLo + kMmaLamHi);
        st->target = tot[0] - (tot[1] - volfrac);
        st->done = 0;
        st->passes = 0;
    }
}

__global__ __launch_bounds__(kThreads) void k_mma_pass(int64_t nel, const double* __restrict__ x,
                                                       const double* __restrict__ dc,
                                                       const double* __restrict__ colsum,
                                                       const double* __restrict__ L,
                                                       const double* __restrict__ U,
                                                       const MmaState* st,
                                                       double* __restrict__ partials) {
    __shared__ double red[kMmaCand][kWaves];
    if (st->done) return;
    double lam[kMmaCand];
#pragma unroll
    for (int j = 0; j < kMmaCand; ++j) lam[j] = st->cand[j];
    double acc[kMmaCand];
#pragma unroll
    for (int j = 0; j < kMmaCand; ++j) acc[j] = 0.0;
    const double inel = 1.0 / (double)nel;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * kThreads) {
        const double xe = x[e];
        const MmaCell c = mma_cell(xe, dc[e], colsum[e] * inel, L[e], U[e]);
        const double sq = sqrt(c.q0);
#pragma unroll
        for (int j = 0; j < kMmaCand; ++j) {
            const double sp = sqrt(c.p0 + lam[j] * c.p1);
            double y = (sp + sq) > 0.0 ? (c.L * sp + c.U * sq) / (sp + sq) : xe;
            y = fmax(c.a, fmin(c.b, y));
            acc[j] += c.p1 / (c.U - y);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kMmaCand; ++j) {
        double v = acc[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[j][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < kMmaCand) {
        double s = 0.0;
        for (int w = 0; w < kWaves; ++w) s += red[threadIdx.x][w];
        partials[(size_t)blockIdx.x * kMmaCand + threadIdx.x] = s;
    }
}

// Totals per candidate (fixed order), the sub-bracket holding F = T, the next table.
__global__ __launch_bounds__(kMmaSelThreads) void k_mma_select(int nb, double tol, MmaState* st,
                                                               const double* __restrict__ partials) {
    __shared__ double red[kMmaCand][kMmaGroups];
    __shared__ double tot[kMmaCand];
    __shared__ double bracket[2];
    if (st->done) return;
    const double lo = st->lo, hi = st->hi;
    {
        const int j = threadIdx.x % kMmaCand, q = threadIdx.x / kMmaCand;
        double s = 0.0;
        for (int b = q; b < nb; b += kMmaGroups) s += partials[(size_t)b * kMmaCand + j];
        red[j][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < kMmaCand) {
        double s = 0.0;
        for (int q = 0; q < kMmaGroups; ++q) s += red[threadIdx.x][q];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double target = st->target;
        int first = kMmaCand;
        for (int j = 0; j < kMmaCand; ++j)
            if (!(tot[j] > target)) {
                first = j;
                break;
            }
        double nlo = lo, nhi = hi;
        if (first == kMmaCand) {
            nlo = mma_cand(lo, hi, kMmaCand - 1);
        } else {
            nhi = mma_cand(lo, hi, first);
            if (first > 0) nlo = mma_cand(lo, hi, first - 1);
        }
        bracket[0] = nlo;
        bracket[1] = nhi;
        st->lo = nlo;
        st->hi = nhi;
        st->lambda = 0.5 * (nlo + nhi);
        st->passes += 1;
        if ((nhi - nlo) / (nhi + nlo) <= tol) st->done = 1;
    }
    __syncthreads();
    if (threadIdx.x < kMmaCand) st->cand[threadIdx.x] = mma_cand(bracket[0], bracket[1], threadIdx.x);
}

// y = x(lambda): the new design; history rotates (x2 <- x1 <- x); partials
// of colsum . y and max |y - x|.
__global__ __launch_bounds__(kThreads) void k_mma_final(
    int64_t nel, const double* __restrict__ x, const double* __restrict__ dc,
    const double* __restrict__ colsum, const double* __restrict__ L,
    const double* __restrict__ U, double* __restrict__ x1, double* __restrict__ x2,
    double* __restrict__ xnew, const MmaState* st, double* __restrict__ partials) {
    __shared__ double red[kWaves];
    __shared__ double redm[kWaves];
    const double lam = st->lambda;
    const double inel = 1.0 / (double)nel;
    double vol = 0.0, chg = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * kThreads) {
        const double xe = x[e];
        const MmaCell c = mma_cell(xe, dc[e], colsum[e] * inel, L[e], U[e]);
        const double y = mma_y(c, lam, xe);
        xnew[e] = y;
        x2[e] = x1[e];
        x1[e] = xe;
        vol += colsum[e] * y;
        chg = fmax(chg, fabs(y - xe));
    }
    const double v = block_sum_all(vol, red);
    for (int off = 32; off > 0; off >>= 1) chg = fmax(chg, __shfl_down(chg, off, 64));
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = chg;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < kWaves; ++w) m = fmax(m, redm[w]);
        partials[2 * (size_t)blockIdx.x] = v;
        partials[2 * (size_t)blockIdx.x + 1] = m;
    }
}

__global__ __launch_bounds__(kThreads) void k_mma_finish(int64_t nel, int nb, MmaState* st,
                                                         const double* __restrict__ partials) {
    __shared__ double red[kWaves];
    __shared__ double redm[kWaves];
    const double t = reduce_partials(partials, nb, 2, 0, red);
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) m = fmax(m, partials[2 * (size_t)b + 1]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int w = 0; w < kWaves; ++w) mm = fmax(mm, redm[w]);
        st->vol = t / (double)nel;
        st->change = mm;
    }
}

hipError_t launch_mma(int64_t nel, int64_t iter, const double* x, const double* dc,
                      const double* colsum, double volfrac, double tol, const MmaParams& prm,
                      double* L, double* U, double* x1, double* x2, double* xnew, MmaState* st,
                      double* partials, hipStream_t s) {
    int nb = (int)std::min<int64_t>(kMmaBlocks, (nel + kThreads - 1) / kThreads);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_mma_prepare, dim3(nb), dim3(kThreads), 0, s, nel, iter, x, x1, x2, L, U,
                       colsum, prm.s_init, prm.s_incr, prm.s_decr, partials);
    hipLaunchKernelGGL(k_mma_init, dim3(1), dim3(64), 0, s, nb, partials, volfrac, st);
    const double w0 = std::log(kMmaLamHi / kMmaLamLo);
    const int passes = std::max(
        1, (int)std::ceil(std::log(w0 / (2.0 * std::max(tol, 1e-300))) /
                          std::log(kMmaCand + 1.0)) + 1);
    for (int i = 0; i < std::min(passes, 24); ++i) {
        hipLaunchKernelGGL(k_mma_pass, dim3(nb), dim3(kThreads), 0, s, nel, x, dc, colsum, L, U,
                           st, partials);
        hipLaunchKernelGGL(k_mma_select, dim3(1), dim3(kMmaSelThreads), 0, s, nb, tol, st,
                           partials);
    }
    hipLaunchKernelGGL(k_mma_final, dim3(nb), dim3(kThreads), 0, s, nel, x, dc, colsum, L, U, x1,
                       x2, xnew, st, partials);
    hipLaunchKernelGGL(k_mma_finish, dim3(1), dim3(kThreads), 0, s, nel, nb, st, partials);
    return hipGetLastError();
}

}  // namespace tohip
