// Device helpers shared by the kernels: deterministic block reductions,
// last-block-done finalisation, XCD-aware block remapping, and the
// corner-Hadamard element operator.
#pragma once
#include "tohip_internal.h"

namespace tohip {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Sum of v over the block; result valid in thread 0.  Fixed shuffle tree and
// fixed wave order, so the result does not depend on scheduling.
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kWaves; ++w) t += red[w];
    __syncthreads();
    return t;
}

// Publishes this block's partial (thread 0 holds it) and reports whether this
// block is the last one to arrive.  The last block then sees every partial.
__device__ __forceinline__ bool publish_partial(double part, double* partials, int slot,
                                                unsigned int* counter, unsigned int nblocks) {
    __shared__ int am_last;
    if (threadIdx.x == 0) {
        partials[slot] = part;
        __threadfence();
        unsigned int prev = atomicAdd(counter, 1u);
        am_last = (prev == nblocks - 1);
    }
    __syncthreads();
    if (am_last) __threadfence();
    return am_last != 0;
}

// Deterministic sum of partials[0..n) by one block (strided per thread, then
// block tree); result valid in thread 0.
__device__ __forceinline__ double sum_partials(const double* partials, int n, int stride,
                                               int offset, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kThreads)
        s += __hip_atomic_load(&partials[(size_t)i * stride + offset], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    return block_sum(s, red);
}

// Blocks are dispatched round-robin over the 8 XCDs; map them so that each
// XCD works on a contiguous range of tiles (neighbouring tiles share halo
// data in that XCD's L2).  Bijective on [0, nb).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int xcd = b & 7, idx = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// ---- corner-Hadamard element operator ----
template <typename T, int DIM>
__device__ __forceinline__ void hadamard_fwd(T (&v)[1 << DIM]) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const int bit = 1 << a;
#pragma unroll
        for (int c = 0; c < (1 << DIM); ++c)
            if (!(c & bit)) {
                T v0 = v[c], v1 = v[c | bit];
                v[c] = v0 + v1;
                v[c | bit] = v1 - v0;
            }
    }
}

template <typename T, int DIM>
__device__ __forceinline__ void hadamard_inv(T (&v)[1 << DIM]) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const int bit = 1 << a;
#pragma unroll
        for (int c = 0; c < (1 << DIM); ++c)
            if (!(c & bit)) {
                T e = v[c], o = v[c | bit];
                v[c] = e - o;
                v[c | bit] = e + o;
            }
    }
}

// z = B xm on the label blocks (xm, z in the modal basis, per component).
template <typename T, int DIM>
__device__ __forceinline__ void modal_blocks(const ElemOp<T>& op, const T (&xm)[DIM][1 << DIM],
                                             T (&z)[DIM][1 << DIM]) {
#pragma unroll
    for (int L = 0; L < (1 << DIM); ++L) {
        T in[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) in[d] = xm[d][L ^ (1 << d)];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            T s = op.blk[L][d][0] * in[0];
#pragma unroll
            for (int d2 = 1; d2 < DIM; ++d2) s += op.blk[L][d][d2] * in[d2];
            z[d][L ^ (1 << d)] = s;
        }
    }
}

// ye = Ke xe for one cell (xe, ye indexed [component][corner bit index]).
template <typename T, int DIM>
__device__ __forceinline__ void elem_apply(const ElemOp<T>& op, T (&xe)[DIM][1 << DIM],
                                           T (&ye)[DIM][1 << DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_fwd<T, DIM>(xe[d]);
    modal_blocks<T, DIM>(op, xe, ye);
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_inv<T, DIM>(ye[d]);
}

// xe^T Ke xe (destroys xe).
template <typename T, int DIM>
__device__ __forceinline__ T elem_energy(const ElemOp<T>& op, T (&xe)[DIM][1 << DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_fwd<T, DIM>(xe[d]);
    T z[DIM][1 << DIM];
    modal_blocks<T, DIM>(op, xe, z);
    T s = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int c = 0; c < (1 << DIM); ++c) s += xe[d][c] * z[d][c];
    return s;
}

}  // namespace tohip
