// Optimality-criteria update on the device.  TopOpt.jl has no OC step (its
// SIMP driver is MMA); this build defines one (density-filter variant):
//   dv_e = colsum(J)_e / nel,
//   x_new(lambda) = clamp(x * sqrt(max(-dc,0) / (dv * lambda)),
//                         [max(0, x-move), min(1, x+move)]),
//   lambda chosen so that vol(J x_new) = colsum . x_new / nel meets volfrac
//   (the largest lambda with colsum . x_new > volfrac * nel).
// vol(lambda) is monotone, so instead of one bisection step per launch each
// pass evaluates kOcCand geometrically spaced candidates inside the current
// bracket in one sweep over the cells (k_oc_pass, per-block partials) and a
// one-block kernel keeps the sub-interval that holds the crossing
// (k_oc_select); ~8 passes reach (hi-lo)/(hi+lo) <= 1e-12 from [1e-40, 1e9].
#include <algorithm>

#include "k_common.h"

namespace tohip {

constexpr int kOcCand = kOcCandidates;
constexpr int kOcBlocks = 256;  // one wave per SIMD at 256 VGPRs: 256 x 4 waves fill 1024 SIMDs
constexpr int kOcSelThreads = 1024;
constexpr int kOcGroups = kOcSelThreads / kOcCand;  // partial-sum groups per candidate

__device__ __forceinline__ double oc_cand(double lo, double hi, int j) {
    return lo * exp(log(hi / lo) * (double)(j + 1) / (double)(kOcCand + 1));
}

// 1/sqrt of the candidates of the bracket [lo, hi] (one thread per candidate).
__device__ __forceinline__ void oc_table(OcState* st, double lo, double hi) {
    if (threadIdx.x < kOcCand) st->isq[threadIdx.x] = 1.0 / sqrt(oc_cand(lo, hi, threadIdx.x));
}

// One halving step of the wave's reduce-scatter over the candidates: the
// lane keeps the half of its N running sums selected by lane bit N/2 and adds
// the partner lane's copy of that half.  After the steps N = 64 .. 2, lane l
// holds acc[0] = the wave's sum for candidate l.
template <int N>
__device__ __forceinline__ void oc_reduce_scatter_step(double (&acc)[kOcCand], int lane) {
    constexpr int H = N / 2;
    const bool up = (lane & H) != 0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const double lo = acc[j], hi = acc[j + H];
        const double other = __shfl_xor(up ? lo : hi, H, 64);
        acc[j] = (up ? hi : lo) + other;
    }
}

__global__ void k_oc_init(OcState* st) {
    if (blockIdx.x == 0) {
        oc_table(st, 1e-40, 1e9);
        if (threadIdx.x == 0) {
            st->lo = 1e-40;
            st->hi = 1e9;
            st->lambda = 0.5 * (st->lo + st->hi);
            st->done = 0;
            st->passes = 0;
            st->vol = 0.0;
            st->change = 0.0;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_oc_pass(int64_t nel, const double* __restrict__ x,
                                                      const double* __restrict__ dc,
                                                      const double* __restrict__ colsum,
                                                      double move, const OcState* st,
                                                      double* __restrict__ partials) {
    __shared__ double red[kOcCand][kWaves];
    if (st->done) return;
    double isq[kOcCand];
#pragma unroll
    for (int j = 0; j < kOcCand; ++j) isq[j] = st->isq[j];
    double acc[kOcCand];
#pragma unroll
    for (int j = 0; j < kOcCand; ++j) acc[j] = 0.0;
    const double inel = 1.0 / (double)nel;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * kThreads) {
        const double xe = x[e], cs = colsum[e];
        const double gdc = fmax(-dc[e], 0.0);
        const double base = xe * sqrt(gdc / (cs * inel));
        const double lb = fmax(0.0, xe - move), ub = fmin(1.0, xe + move);
#pragma unroll
        for (int j = 0; j < kOcCand; ++j) acc[j] += cs * fmax(lb, fmin(ub, base * isq[j]));
    }
    // 63 shuffles instead of a 6-level tree per candidate (384): with one
    // wave per SIMD nothing hides the shuffles' latency
    static_assert(kOcCand == 64, "one candidate per lane after the reduce-scatter");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    oc_reduce_scatter_step<64>(acc, lane);
    oc_reduce_scatter_step<32>(acc, lane);
    oc_reduce_scatter_step<16>(acc, lane);
    oc_reduce_scatter_step<8>(acc, lane);
    oc_reduce_scatter_step<4>(acc, lane);
    oc_reduce_scatter_step<2>(acc, lane);
    red[lane][wave] = acc[0];
    __syncthreads();
    if (threadIdx.x < kOcCand) {
        double s = 0.0;
        for (int w = 0; w < kWaves; ++w) s += red[threadIdx.x][w];
        partials[(size_t)blockIdx.x * kOcCand + threadIdx.x] = s;
    }
}

// One block: totals per candidate (kOcGroups threads per candidate, fixed
// split and combine order), the new bracket and the next candidate table.
__global__ __launch_bounds__(kOcSelThreads) void k_oc_select(int64_t nel, int nb, double volfrac,
                                                             double tol, OcState* st,
                                                             const double* __restrict__ partials) {
    __shared__ double red[kOcCand][kOcGroups];
    __shared__ double tot[kOcCand];
    __shared__ double bracket[2];
    if (st->done) return;
    const double lo = st->lo, hi = st->hi;
    {
        const int j = threadIdx.x % kOcCand, q = threadIdx.x / kOcCand;
        double s = 0.0;
#pragma unroll 8
        for (int b = q; b < nb; b += kOcGroups) s += partials[(size_t)b * kOcCand + j];
        red[j][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < kOcCand) {
        double s = 0.0;
        for (int q = 0; q < kOcGroups; ++q) s += red[threadIdx.x][q];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // the first candidate whose volume is not above the target (tot falls
        // as lambda grows): lowest set bit of the wave's ballot
        const double target = volfrac * (double)nel;
        const unsigned long long below = __ballot(!(tot[threadIdx.x] > target));
        if (threadIdx.x == 0) {
            const int first = below ? __ffsll((unsigned long long)below) - 1 : kOcCand;
            double nlo = lo, nhi = hi;
            if (first == kOcCand) {
                nlo = oc_cand(lo, hi, kOcCand - 1);
            } else {
                nhi = oc_cand(lo, hi, first);
                if (first > 0) nlo = oc_cand(lo, hi, first - 1);
            }
            bracket[0] = nlo;
            bracket[1] = nhi;
            st->lo = nlo;
            st->hi = nhi;
            st->lambda = 0.5 * (nlo + nhi);
            st->passes += 1;
            if ((nhi - nlo) / (nhi + nlo) <= tol) st->done = 1;
        }
    }
    __syncthreads();
    oc_table(st, bracket[0], bracket[1]);
}

// x_new at the final lambda; partials of colsum . x_new and max |x_new - x|.
__global__ __launch_bounds__(kThreads) void k_oc_final(int64_t nel, const double* __restrict__ x,
                                                       const double* __restrict__ dc,
                                                       const double* __restrict__ colsum,
                                                       double move, double* __restrict__ xnew,
                                                       const OcState* st,
                                                       double* __restrict__ partials) {
    __shared__ double red[kWaves];
    __shared__ double redm[kWaves];
    const double lam = st->lambda;
    const double inel = 1.0 / (double)nel;
    double vol = 0.0, chg = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * kThreads) {
        const double xe = x[e], cs = colsum[e];
        const double gdc = fmax(-dc[e], 0.0);
        double xn = xe * sqrt(gdc / (cs * inel * lam));
        const double lb = fmax(0.0, xe - move), ub = fmin(1.0, xe + move);
        xn = fmax(lb, fmin(ub, xn));
        xnew[e] = xn;
        vol += cs * xn;
        chg = fmax(chg, fabs(xn - xe));
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

__global__ __launch_bounds__(kThreads) void k_oc_finish(int64_t nel, int nb, OcState* st,
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

hipError_t launch_oc(int64_t nel, const double* x, const double* dc, const double* colsum,
                     double volfrac, double move, double tol, int max_passes, double* xnew,
                     OcState* st, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, s, st);
    int nb = (int)std::min<int64_t>(kOcBlocks, (nel + kThreads - 1) / kThreads);
    if (nb < 1) nb = 1;
    // Each pass shrinks the log-width of the bracket by kOcCand + 1; the
    // stopping test (hi-lo)/(hi+lo) <= tol holds once the log-width is below
    // 2 tol.  Launch only the passes that can still narrow it (plus one).
    const double w0 = std::log(1e9 / 1e-40);
    const int needed =
        (int)std::ceil(std::log(w0 / (2.0 * std::max(tol, 1e-300))) / std::log(kOcCand + 1.0)) + 1;
    const int passes = std::max(1, std::min(max_passes, needed));
    for (int i = 0; i < passes; ++i) {
        hipLaunchKernelGGL(k_oc_pass, dim3(nb), dim3(kThreads), 0, s, nel, x, dc, colsum, move,
                           st, partials);
        hipLaunchKernelGGL(k_oc_select, dim3(1), dim3(kOcSelThreads), 0, s, nel, nb, volfrac, tol,
                           st, partials);
    }
    hipLaunchKernelGGL(k_oc_final, dim3(nb), dim3(kThreads), 0, s, nel, x, dc, colsum, move,
                       xnew, st, partials);
    hipLaunchKernelGGL(k_oc_finish, dim3(1), dim3(kThreads), 0, s, nel, nb, st, partials);
    return hipGetLastError();
}

}  // namespace tohip
