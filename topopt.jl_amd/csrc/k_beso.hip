// BESO topology update on the device (TopOpt.jl src/Algorithms/beso.jl:93-188).
//
// Per BESO iteration, after the compliance sensitivity has been pulled back
// through the filter (dc = J^T grad):
//   sens = -dc, averaged with the previous iteration's sens after the first;
//   l1, l2 = min, max of sens;
//   while (l2 - l1) / l2 > sens_tol:  th = (l1 + l2) / 2;
//       topology = sens > th (passive cells keep 1 / 0);
//       if volume(topology) > vol * total: l1 = th else l2 = th.
// The bisection is sequential, but each pass here evaluates a whole 6-level
// bisection tree at once: the 63 thresholds (l1 + l2) / 2 of every bracket
// the next 6 steps can reach, computed exactly as the sequential loop would.
// One sweep over the cells counts, for every tree node, the free cells above
// its threshold (k_beso_count, per-block partials); a one-block kernel then
// walks the tree along the path the sequential bisection takes, stopping
// where its loop condition fails (k_beso_select).  The topology of the last
// threshold tested is written once at the end (k_beso_apply), as the
// reference's last loop iteration leaves it.
#include "k_common.h"

namespace tohip {

constexpr int kBesoBlocks = 256;
constexpr int kBesoSelThreads = 64;
static_assert(kBesoBlocks * kBesoCand <= kGeneralPartials, "partial buffer");

// The 63 thresholds of the tree rooted at the bracket [l1, l2] (heap order:
// node i has children 2i + 1 (l2 = th) and 2i + 2 (l1 = th)).
__device__ void beso_tree(BesoState* st) {
    double a[kBesoCand], b[kBesoCand];
    a[0] = st->l1;
    b[0] = st->l2;
    for (int i = 0; i < kBesoCand; ++i) {
        const double m = (a[i] + b[i]) / 2;
        st->cand[i] = m;
        if (2 * i + 2 < kBesoCand) {
            a[2 * i + 1] = a[i];
            b[2 * i + 1] = m;
            a[2 * i + 2] = m;
            b[2 * i + 2] = b[i];
        }
    }
}

// sens = -dc (averaged with prev when avg), block min / max partials.
__global__ __launch_bounds__(kThreads) void k_beso_sens(int64_t n, const double* __restrict__ dc,
                                                        const double* __restrict__ prev, int avg,
                                                        double* __restrict__ sens,
                                                        double* __restrict__ partials) {
    __shared__ double rmin[kThreads], rmax[kThreads];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        double s = -dc[i];
        if (avg) s = (s + prev[i]) / 2;
        sens[i] = s;
        lo = fmin(lo, s);
        hi = fmax(hi, s);
    }
    rmin[threadIdx.x] = lo;
    rmax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + w]);
            rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * (size_t)blockIdx.x] = rmin[0];
        partials[2 * (size_t)blockIdx.x + 1] = rmax[0];
    }
}

// Bracket from the min / max partials; the first tree.
__global__ void k_beso_init(int nb, const double* __restrict__ partials, BesoState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nb; ++b) {
        lo = fmin(lo, partials[2 * b]);
        hi = fmax(hi, partials[2 * b + 1]);
    }
    st->l1 = lo;
    st->l2 = hi;
    st->steps = 0;
    st->done = !((hi - lo) / hi > st->sens_tol);
    beso_tree(st);
}

// For every tree node j: number of free cells (mask 0) with sens > cand_j.
__global__ __launch_bounds__(kThreads) void k_beso_count(int64_t n, const double* __restrict__ sens,
                                                         const uint8_t* __restrict__ mask,
                                                         const BesoState* st,
                                                         double* __restrict__ partials) {
    __shared__ double cand[kBesoCand];
    __shared__ unsigned red[kBesoCand][kWaves];
    if (st->done) return;
    if (threadIdx.x < kBesoCand) cand[threadIdx.x] = st->cand[threadIdx.x];
    __syncthreads();
    unsigned cnt[kBesoCand];
#pragma unroll
    for (int j = 0; j < kBesoCand; ++j) cnt[j] = 0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        if (mask && mask[i]) continue;
        const double s = sens[i];
#pragma unroll
        for (int j = 0; j < kBesoCand; ++j) cnt[j] += s > cand[j] ? 1u : 0u;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kBesoCand; ++j) {
        unsigned v = cnt[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[j][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < kBesoCand) {
        unsigned t = 0;
        for (int w = 0; w < kWaves; ++w) t += red[threadIdx.x][w];
        partials[(size_t)blockIdx.x * kBesoCand + threadIdx.x] = (double)t;
    }
}

// Walk the tree along the sequential bisection's path.
__global__ __launch_bounds__(kBesoSelThreads) void k_beso_select(int nb,
                                                                 const double* __restrict__ partials,
                                                                 BesoState* st) {
    __shared__ double count[kBesoCand];
    if (st->done) return;
    if (threadIdx.x < kBesoCand) {
        double t = 0;
        for (int b = 0; b < nb; ++b) t += partials[(size_t)b * kBesoCand + threadIdx.x];
        count[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double l1 = st->l1, l2 = st->l2;
    int node = 0;
    for (int level = 0; level < kBesoLevels; ++level) {
        if (!((l2 - l1) / l2 > st->sens_tol)) {
            st->done = 1;
            break;
        }
        const double th = st->cand[node];
        st->th = th;
        st->steps += 1;
        const double vol_now = (count[node] + st->nblack) * st->cv;
        if (vol_now - st->vol * st->total > 0) {
            l1 = th;
            node = 2 * node + 2;
        } else {
            l2 = th;
            node = 2 * node + 1;
        }
    }
    st->l1 = l1;
    st->l2 = l2;
    if (!st->done) {
        if (!((l2 - l1) / l2 > st->sens_tol))
            st->done = 1;
        else
            beso_tree(st);
    }
}

// topology = sens > th for free cells (when the bisection ran), block sums.
__global__ __launch_bounds__(kThreads) void k_beso_apply(int64_t n, const double* __restrict__ sens,
                                                         const uint8_t* __restrict__ mask,
                                                         const BesoState* st,
                                                         double* __restrict__ topo,
                                                         double* __restrict__ partials) {
    __shared__ double red[kWaves];
    const bool run = st->steps > 0;
    const double th = st->th;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        double t = topo[i];
        if (run && !(mask && mask[i])) {
            t = sens[i] > th ? 1.0 : 0.0;
            topo[i] = t;
        }
        s += t;
    }
    const double b = block_sum_all(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

hipError_t launch_beso_update(int64_t n, const double* dc, const double* prev, int avg,
                              double* sens, const uint8_t* mask, BesoState* st, double* topo,
                              double* partials, int passes, hipStream_t s) {
    hipLaunchKernelGGL(k_beso_sens, dim3(kBesoBlocks), dim3(kThreads), 0, s, n, dc, prev, avg, sens,
                       partials);
    hipLaunchKernelGGL(k_beso_init, dim3(1), dim3(64), 0, s, kBesoBlocks, partials, st);
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_beso_count, dim3(kBesoBlocks), dim3(kThreads), 0, s, n, sens, mask,
                           st, partials);
        hipLaunchKernelGGL(k_beso_select, dim3(1), dim3(kBesoSelThreads), 0, s, kBesoBlocks,
                           partials, st);
    }
    return hipGetLastError();
}

hipError_t launch_beso_passes(int64_t n, const double* sens, const uint8_t* mask, BesoState* st,
                              double* partials, int passes, hipStream_t s) {
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_beso_count, dim3(kBesoBlocks), dim3(kThreads), 0, s, n, sens, mask,
                           st, partials);
        hipLaunchKernelGGL(k_beso_select, dim3(1), dim3(kBesoSelThreads), 0, s, kBesoBlocks,
                           partials, st);
    }
    return hipGetLastError();
}

hipError_t launch_beso_apply(int64_t n, const double* sens, const uint8_t* mask,
                             const BesoState* st, double* topo, double* partials, double* sum_out,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_beso_apply, dim3(kBesoBlocks), dim3(kThreads), 0, s, n, sens, mask, st,
                       topo, partials);
    return launch_reduce(partials, kBesoBlocks, 1, sum_out, s);
}

}  // namespace tohip
