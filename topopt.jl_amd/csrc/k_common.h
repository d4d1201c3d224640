// Device helpers shared by the kernels: deterministic block reductions,
// consumer-side reduction of a previous kernel's per-block partials,
// XCD-aware tile mapping, and the corner-Hadamard element operator.
//
// Grid-wide sums never use a last-block-done handshake: on this part the
// eight XCD L2 caches are not coherent with each other, so a device-scope
// fence per block costs an L2 write-back + invalidate.  Instead every kernel
// writes one partial per block, and whichever kernel consumes the sum (the
// next one in stream order, after the implicit kernel-boundary coherence)
// re-reduces the partials in a fixed order.  Each consuming block computes
// the identical value, so results stay bitwise reproducible.
#pragma once
#include "tohip_internal.h"

namespace tohip {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Sum of v over the block; the result is returned to every thread.  Fixed
// shuffle tree and fixed wave order.
__device__ __forceinline__ double block_sum_all(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;
}

// Sum over b < n of partials[b*stride + offset], identical in every block and
// every thread (strided per-thread sums, then the fixed block tree).
__device__ __forceinline__ double reduce_partials(const double* __restrict__ partials, int n,
                                                  int stride, int offset, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[(size_t)i * stride + offset];
    return block_sum_all(s, red);
}

// Stopping rule of the PCG direction step for iteration k = it + 1, given
// the reduced r.z and r.r of the current residual: converged when
// ||r|| <= tol, out of iterations when it >= maxiter, or broken down
// (err = kPcgErrNonFinite / kPcgErrPrecond; 0 otherwise).  Every block
// evaluates it on the same reduced values, so every block agrees.
__device__ __forceinline__ bool pcg_stop_test(const PcgScalars* sc, double rz, double rr,
                                              long long it, double tol, int& err) {
    err = 0;
    if (!(isfinite(rr) && isfinite(rz)))
        err = kPcgErrNonFinite;
    else if (!(rz > 0.0) && rr > 0.0 && !(sqrt(rr) <= tol))
        err = kPcgErrPrecond;
    return err != 0 || sqrt(rr) <= tol || it >= sc->maxiter;
}

// Element i of a device array addressed with a 32-bit byte offset, so the
// compiler can use the scalar-base + 32-bit vector-offset form of global
// loads and stores (no 64-bit address arithmetic per access).  Valid for
// arrays below 4 GiB; tohip_create limits grids to 2^29 DOFs.
template <typename T>
__device__ __forceinline__ T& at(T* base, uint32_t i) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) +
                                 (size_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T& at(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       (size_t)(i * (uint32_t)sizeof(T)));
}

// The D values of node n of a node-major vector, addressed as `at` does but
// with one 32-bit byte offset per node: the D accesses differ only by
// immediate offsets, so the compiler issues one wide load or store
// (global_load_dwordx3 for three floats) instead of D narrow ones.
template <int D, typename T>
__device__ __forceinline__ void ld_node(const T* base, uint32_t n, T (&v)[D]) {
    const char* p =
        reinterpret_cast<const char*>(base) + (size_t)(n * (uint32_t)(D * sizeof(T)));
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = *reinterpret_cast<const T*>(p + d * sizeof(T));
}
template <int D, typename T>
__device__ __forceinline__ void st_node(T* base, uint32_t n, const T (&v)[D]) {
    char* p = reinterpret_cast<char*>(base) + (size_t)(n * (uint32_t)(D * sizeof(T)));
#pragma unroll
    for (int d = 0; d < D; ++d) *reinterpret_cast<T*>(p + d * sizeof(T)) = v[d];
}

// Projection value and derivative at x (TopOpt.jl src/Utilities/penalties.jl
// :76-95): proj 0 identity, 1 Heaviside 1 - exp(-b x) + x exp(-b),
// 2 sigmoid 1 / (1 + exp((b + 1)(0.5 - x))).
__device__ __forceinline__ void projection_eval(int proj, double beta, double x, double& y,
                                                double& dy) {
    y = x;
    dy = 1.0;
    if (proj == 1) {
        const double e = exp(-beta * x), eb = exp(-beta);
        y = 1.0 - e + x * eb;
        dy = beta * e + eb;
    } else if (proj == 2) {
        const double e = exp((beta + 1.0) * (0.5 - x));
        y = 1.0 / (1.0 + e);
        dy = (beta + 1.0) * e / ((1.0 + e) * (1.0 + e));
    }
}

// Penalty value and derivative at x (projection first, then the penalty).
__device__ __forceinline__ void penalty_eval(const Penalty& P, double x, double& v, double& dv) {
    double y, dy;
    projection_eval(P.proj, P.beta, x, y, dy);
    double f, df;
    if (P.kind == 1) {
        const double d = 1.0 + P.p * (1.0 - y);
        f = y / d;
        df = (1.0 + P.p) / (d * d);
    } else if (P.kind == 2) {
        const double sp = sinh(P.p);
        f = sinh(P.p * y) / sp;
        df = P.p * cosh(P.p * y) / sp;
    } else {
        f = pow(y, P.p);
        df = P.p * (P.p == 1.0 ? 1.0 : pow(y, P.p - 1.0));
    }
    v = f;
    dv = df * dy;
}

// Penalised density rho(x) and d rho / d x in either interpolation order
// (get_rho / get_rho_drho, src/Utilities/penalties.jl:113-130):
// pbi: xmin + (1 - xmin) pen(x); else pen(xmin + (1 - xmin) x).
__device__ __forceinline__ void rho_drho(const Penalty& P, double x, double xmin, int pbi,
                                         double& r, double& dr) {
    double v, dv;
    if (pbi) {
        penalty_eval(P, x, v, dv);
        r = v * (1.0 - xmin) + xmin;
        dr = (1.0 - xmin) * dv;
    } else {
        penalty_eval(P, x * (1.0 - xmin) + xmin, v, dv);
        r = v;
        dr = dv * (1.0 - xmin);
    }
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

// Symmetric 3x3 block (b00 b01 b02 b11 b12 b22) times (a, b, c).
template <typename T>
__device__ __forceinline__ void sym3(const T* k, T a, T b, T c, T& za, T& zb, T& zc) {
    za = k[0] * a + k[1] * b + k[2] * c;
    zb = k[1] * a + k[3] * b + k[4] * c;
    zc = k[2] * a + k[4] * b + k[5] * c;
}
// Symmetric 2x2 block (b00 b01 b11) times (a, b).
template <typename T>
__device__ __forceinline__ void sym2(const T* k, T a, T b, T& za, T& zb) {
    za = k[0] * a + k[1] * b;
    zb = k[1] * a + k[2] * b;
}

// z = B xm on the label blocks (xm, z in the modal basis, [component][mode]).
// Member of label L for component d: mode L ^ (1 << d).  CUBE uses the
// 8-number cube form (see ElemOp).
// The 3D cube form on explicit coefficients cube[8] (see ElemOp).
template <typename T>
__device__ __forceinline__ void modal_blocks_cube(const T (&cube)[8], const T (&xm)[3][8],
                                                  T (&z)[3][8]) {
    {
        const T AmB = cube[0], B = cube[1], C = cube[2], D = cube[3];
        const T E = cube[4], F = cube[5], GmH = cube[6], H = cube[7];
        // L0: (A-B) x + B (x0 + x1 + x2)
        const T s0 = B * (xm[0][1] + xm[1][2] + xm[2][4]);
        z[0][1] = AmB * xm[0][1] + s0;
        z[1][2] = AmB * xm[1][2] + s0;
        z[2][4] = AmB * xm[2][4] + s0;
        // L7
        const T s7 = H * (xm[0][6] + xm[1][5] + xm[2][3]);
        z[0][6] = GmH * xm[0][6] + s7;
        z[1][5] = GmH * xm[1][5] + s7;
        z[2][3] = GmH * xm[2][3] + s7;
        z[0][0] = z[1][0] = z[2][0] = 0;
        // L1, L2, L4: [[C D][D C]]
        z[1][3] = C * xm[1][3] + D * xm[2][5];
        z[2][5] = D * xm[1][3] + C * xm[2][5];
        z[0][3] = C * xm[0][3] + D * xm[2][6];
        z[2][6] = D * xm[0][3] + C * xm[2][6];
        z[0][5] = C * xm[0][5] + D * xm[1][6];
        z[1][6] = D * xm[0][5] + C * xm[1][6];
        // L3, L5, L6: E (a + b) on the pair, F on the single
        const T e3 = E * (xm[0][2] + xm[1][1]);
        z[0][2] = e3;
        z[1][1] = e3;
        z[2][7] = F * xm[2][7];
        const T e5 = E * (xm[0][4] + xm[2][1]);
        z[0][4] = e5;
        z[2][1] = e5;
        z[1][7] = F * xm[1][7];
        const T e6 = E * (xm[1][4] + xm[2][2]);
        z[1][4] = e6;
        z[2][2] = e6;
        z[0][7] = F * xm[0][7];
    }
}

template <typename T, int DIM, bool CUBE = false>
__device__ __forceinline__ void modal_blocks(const ElemOp<T>& op, const T (&xm)[DIM][1 << DIM],
                                             T (&z)[DIM][1 << DIM]) {
    if constexpr (DIM == 3 && CUBE) {
        modal_blocks_cube<T>(op.cube, xm, z);
    } else if constexpr (DIM == 3) {
        sym3(op.l0, xm[0][1], xm[1][2], xm[2][4], z[0][1], z[1][2], z[2][4]);  // L0
        sym3(op.l7, xm[0][6], xm[1][5], xm[2][3], z[0][6], z[1][5], z[2][3]);  // L7
        z[0][0] = z[1][0] = z[2][0] = 0;  // rigid translations (L1 d0, L2 d1, L4 d2)
        sym2(op.pr[0], xm[1][3], xm[2][5], z[1][3], z[2][5]);  // L1
        sym2(op.pr[1], xm[0][3], xm[2][6], z[0][3], z[2][6]);  // L2
        sym2(op.pr[2], xm[0][5], xm[1][6], z[0][5], z[1][6]);  // L4
        sym2(op.qs[0], xm[0][2], xm[1][1], z[0][2], z[1][1]);  // L3 pair
        z[2][7] = op.qs[0][3] * xm[2][7];                      // L3 single
        sym2(op.qs[1], xm[0][4], xm[2][1], z[0][4], z[2][1]);  // L5 pair
        z[1][7] = op.qs[1][3] * xm[1][7];                      // L5 single
        sym2(op.qs[2], xm[1][4], xm[2][2], z[1][4], z[2][2]);  // L6 pair
        z[0][7] = op.qs[2][3] * xm[0][7];                      // L6 single
    } else {
        sym2(op.q2 + 0, xm[0][1], xm[1][2], z[0][1], z[1][2]);  // L0
        sym2(op.q2 + 3, xm[0][2], xm[1][1], z[0][2], z[1][1]);  // L3
        z[0][0] = z[1][0] = 0;                                  // rigid (L1 d0, L2 d1)
        z[1][3] = op.q2[6] * xm[1][3];                          // L1
        z[0][3] = op.q2[7] * xm[0][3];                          // L2
    }
}

// ye = Ke xe for one cell (xe, ye indexed [component][corner bit index]).
template <typename T, int DIM, bool CUBE = false>
__device__ __forceinline__ void elem_apply(const ElemOp<T>& op, T (&xe)[DIM][1 << DIM],
                                           T (&ye)[DIM][1 << DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_fwd<T, DIM>(xe[d]);
    modal_blocks<T, DIM, CUBE>(op, xe, ye);
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_inv<T, DIM>(ye[d]);
}

// ye = s Ke xe for one cube cell, the scale folded into the 8 coefficients
// (8 products instead of 24 on the outputs).
template <typename T>
__device__ __forceinline__ void elem_apply_cube_scaled(const ElemOp<T>& op, T scale,
                                                       T (&xe)[3][8], T (&ye)[3][8]) {
    T c[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) c[i] = op.cube[i] * scale;
#pragma unroll
    for (int d = 0; d < 3; ++d) hadamard_fwd<T, 3>(xe[d]);
    modal_blocks_cube<T>(c, xe, ye);
#pragma unroll
    for (int d = 0; d < 3; ++d) hadamard_inv<T, 3>(ye[d]);
}

// xe^T Ke xe (destroys xe).
template <typename T, int DIM, bool CUBE = false>
__device__ __forceinline__ T elem_energy(const ElemOp<T>& op, T (&xe)[DIM][1 << DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) hadamard_fwd<T, DIM>(xe[d]);
    T z[DIM][1 << DIM];
    modal_blocks<T, DIM, CUBE>(op, xe, z);
    T s = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int c = 0; c < (1 << DIM); ++c) s += xe[d][c] * z[d][c];
    return s;
}

}  // namespace tohip
