// Assembled stiffness matrix in CSR on the device: the comparator of the
// matrix-free operator, TopOpt.jl's CGAssemblySolver
// (src/FEA/solvers_api.jl:223-250; assemble!, src/TopOptProblems/
// assemble.jl:28-90, then Ferrite's apply!).
//
// Row 3n+d (node-major DOFs) holds, for every node m of the 3x3x3 block
// around n inside the grid (ascending node id) and every component e, the
// entry K[3n+d][3m+e] = sum over the cells containing n and m of
// rho_c Ke[(3a+d)][(3b+e)] (a, b the corner bit indices of n, m in c); the
// structural pattern is kept for prescribed rows and columns, whose entries
// are zero except the diagonal = mean |diag K| of the unconstrained matrix
// (apply!).  Each (node, neighbour) pair is assembled by one thread as a
// dim x dim block: no atomics, deterministic.
//
// SpMV: 16 lanes per row (a row has at most 81 entries), lane-strided over
// the row's entries, a width-16 shuffle tree; HBM-bound on the 12 bytes per
// stored entry (value + column).
#include "k_common.h"

namespace tohip {

// Valid neighbour offsets of node (i, j, k) in ascending node order; returns
// the count and, when o >= 0, the slot of offset o in *slot (-1 if invalid).
template <int DIM>
__device__ __forceinline__ int csr_neighbours(const Grid& g, int i, int j, int k, int o,
                                              int* slot) {
    int cnt = 0;
    if (slot) *slot = -1;
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int oo = (dx + 1) + 3 * ((dy + 1) + 3 * (DIM == 3 ? dz + 1 : 0));
                const int mi = i + dx, mj = j + dy, mk = k + dz;
                const bool ok = mi >= 0 && mi < g.NX && mj >= 0 && mj < g.NY &&
                                (DIM == 2 || (mk >= 0 && mk < g.NZ));
                if (!ok) continue;
                if (slot && oo == o) *slot = cnt;
                ++cnt;
            }
    return cnt;
}

template <int DIM>
__global__ void k_csr_cols(Grid g, const int64_t* __restrict__ rowptr, int* __restrict__ cols) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
        const int k = DIM == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
        int s = 0;
        for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int mi = i + dx, mj = j + dy, mk = k + dz;
                    if (mi < 0 || mi >= g.NX || mj < 0 || mj >= g.NY ||
                        (DIM == 3 && (mk < 0 || mk >= g.NZ)))
                        continue;
                    const int64_t m = mi + (int64_t)g.NX * (mj + (int64_t)g.NY * mk);
                    for (int d = 0; d < DIM; ++d)
                        for (int e = 0; e < DIM; ++e)
                            cols[rowptr[DIM * n + d] + DIM * s + e] = (int)(DIM * m + e);
                    ++s;
                }
    }
}

// Unconstrained diagonal of K (per DOF) and per-block partials of
// (sum |diag|, sum |diag| over prescribed DOFs, number of prescribed DOFs).
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_csr_diag(Grid g, const double* __restrict__ Ke,
                                                       const double* __restrict__ E,
                                                       const uint8_t* __restrict__ fixmask,
                                                       double* __restrict__ diag,
                                                       double* __restrict__ partials) {
    __shared__ double red[kWaves];
    constexpr int ND = DIM << DIM;
    double acc = 0.0, accf = 0.0, nf = 0.0;
    for (int64_t n = blockIdx.x * (int64_t)kThreads + threadIdx.x; n < g.nnodes;
         n += (int64_t)gridDim.x * kThreads) {
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
        const int k = DIM == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
        double dv[DIM];
        for (int d = 0; d < DIM; ++d) dv[d] = 0.0;
        for (int cz = (DIM == 3 ? k - 1 : 0); cz <= (DIM == 3 ? k : 0); ++cz) {
            if (DIM == 3 && (cz < 0 || cz >= g.nz)) continue;
            for (int cy = j - 1; cy <= j; ++cy) {
                if (cy < 0 || cy >= g.ny) continue;
                for (int cx = i - 1; cx <= i; ++cx) {
                    if (cx < 0 || cx >= g.nx) continue;
                    const double ec = E[cx + (int64_t)g.nx * (cy + (int64_t)g.ny * cz)];
                    const int a = (i - cx) + 2 * (j - cy) + (DIM == 3 ? 4 * (k - cz) : 0);
                    for (int d = 0; d < DIM; ++d) dv[d] += ec * Ke[(DIM * a + d) * ND + DIM * a + d];
                }
            }
        }
        const uint8_t fm = fixmask[n];
        for (int d = 0; d < DIM; ++d) {
            diag[DIM * n + d] = dv[d];
            acc += fabs(dv[d]);
            if ((fm >> d) & 1) {
                accf += fabs(dv[d]);
                nf += 1.0;
            }
        }
    }
    const double s0 = block_sum_all(acc, red);
    const double s1 = block_sum_all(accf, red);
    const double s2 = block_sum_all(nf, red);
    if (threadIdx.x == 0) {
        partials[3 * (size_t)blockIdx.x] = s0;
        partials[3 * (size_t)blockIdx.x + 1] = s1;
        partials[3 * (size_t)blockIdx.x + 2] = s2;
    }
}

// apply!'s prescribed diagonal: mean |diag| of the unconstrained K; the safe
// value: mean |diag| of K after apply! (solvers_api.jl:232-239 runs after it).
__device__ __forceinline__ void csr_diag_values(const double* md, int64_t ndof, double& m_apply,
                                                double& m_safe) {
    m_apply = md[0] / (double)ndof;
    m_safe = (md[0] - md[1] + md[2] * m_apply) / (double)ndof;
}

// One thread per (node n, neighbour offset o): the dim x dim block K[n][m].
// mdiag_sum: the three sums of k_csr_diag (see csr_diag_values).
template <int DIM>
__global__ void k_csr_values(Grid g, const double* __restrict__ Ke, const double* __restrict__ E,
                             const uint8_t* __restrict__ fixmask,
                             const int64_t* __restrict__ rowptr,
                             const double* __restrict__ mdiag_sum, int safe,
                             double* __restrict__ vals) {
    constexpr int NO = DIM == 3 ? 27 : 9;
    constexpr int ND = DIM << DIM;
    double md, msafe;
    csr_diag_values(mdiag_sum, g.ndof, md, msafe);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.nnodes * NO;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = t / NO;
        const int o = (int)(t % NO);
        const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = DIM == 3 ? o / 9 - 1 : 0;
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY);
        const int k = DIM == 3 ? (int)(n / ((int64_t)g.NX * g.NY)) : 0;
        const int mi = i + dx, mj = j + dy, mk = k + dz;
        if (mi < 0 || mi >= g.NX || mj < 0 || mj >= g.NY || (DIM == 3 && (mk < 0 || mk >= g.NZ)))
            continue;
        int slot;
        csr_neighbours<DIM>(g, i, j, k, o, &slot);
        const int64_t m = mi + (int64_t)g.NX * (mj + (int64_t)g.NY * mk);
        double blk[DIM][DIM];
        for (int d = 0; d < DIM; ++d)
            for (int e = 0; e < DIM; ++e) blk[d][e] = 0.0;
        for (int cz = (DIM == 3 ? k - 1 : 0); cz <= (DIM == 3 ? k : 0); ++cz) {
            if (DIM == 3 && (cz < 0 || cz >= g.nz)) continue;
            for (int cy = j - 1; cy <= j; ++cy) {
                if (cy < 0 || cy >= g.ny) continue;
                for (int cx = i - 1; cx <= i; ++cx) {
                    if (cx < 0 || cx >= g.nx) continue;
                    const int bx = mi - cx, by = mj - cy, bz = DIM == 3 ? mk - cz : 0;
                    if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                    const double ec = E[cx + (int64_t)g.nx * (cy + (int64_t)g.ny * cz)];
                    const int a = (i - cx) + 2 * (j - cy) + (DIM == 3 ? 4 * (k - cz) : 0);
                    const int b = bx + 2 * by + 4 * bz;
                    for (int d = 0; d < DIM; ++d)
                        for (int e = 0; e < DIM; ++e)
                            blk[d][e] += ec * Ke[(DIM * a + d) * ND + DIM * b + e];
                }
            }
        }
        const uint8_t fn = fixmask[n], fm = fixmask[m];
        for (int d = 0; d < DIM; ++d)
            for (int e = 0; e < DIM; ++e) {
                double v = blk[d][e];
                const bool diag = m == n && d == e;
                if (((fn >> d) & 1) || ((fm >> e) & 1)) v = diag ? md : 0.0;
                else if (safe && diag && v == 0.0) v = msafe;
                vals[rowptr[DIM * n + d] + DIM * slot + e] = v;
            }
    }
}

// dinv = 1 / diag(K after apply!) (jacobi), or 1 (identity preconditioner).
template <int DIM>
__global__ void k_csr_dinv(int64_t ndof, const double* __restrict__ diag,
                           const uint8_t* __restrict__ fixmask,
                           const double* __restrict__ mdiag_sum, int jacobi, int safe,
                           double* __restrict__ dinv) {
    double md, msafe;
    csr_diag_values(mdiag_sum, ndof, md, msafe);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ndof;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!jacobi) {
            dinv[i] = 1.0;
            continue;
        }
        const bool fixed = (fixmask[i / DIM] >> (i % DIM)) & 1;
        double d = fixed ? md : diag[i];
        if (safe && !fixed && d == 0.0) d = msafe;
        dinv[i] = 1.0 / d;
    }
}

// y = K x, 16 lanes per row.  PCG: x is p_k (chosen by the iteration counter),
// per-block partials of p.q into part; a no-op once `done`.
template <bool PCG>
__global__ __launch_bounds__(kThreads) void k_csr_spmv(int64_t nrows,
                                                       const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ cols,
                                                       const double* __restrict__ vals,
                                                       const double* x0, const double* x1,
                                                       double* __restrict__ y, PcgScalars* sc,
                                                       double* __restrict__ part) {
    __shared__ double red[kWaves];
    const double* x = x0;
    if (PCG) {
        if (sc->done) return;
        x = (sc->k_cur & 1) ? x1 : x0;
    }
    const int lane = threadIdx.x & 15;
    const int64_t groups = (int64_t)gridDim.x * (kThreads / 16);
    double acc = 0.0;
    for (int64_t row = blockIdx.x * (int64_t)(kThreads / 16) + (threadIdx.x >> 4); row < nrows;
         row += groups) {
        const int64_t b = rowptr[row], e = rowptr[row + 1];
        double s = 0.0;
        for (int64_t kk = b + lane; kk < e; kk += 16) s += vals[kk] * x[cols[kk]];
        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
        if (lane == 0) {
            y[row] = s;
            if (PCG) acc += x[row] * s;
        }
    }
    if (PCG) {
        const double t = block_sum_all(acc, red);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// PCG direction for operators that are applied in a separate kernel (the
// assembled matrix, the heat-conduction operator): reduces the update's
// (r.z, r.r) partials, tests convergence (the scalar bookkeeping of the fused
// matrix-free kernel) and writes p_k = dinv .* r + beta p_{k-1} into p[k & 1].
__global__ __launch_bounds__(kThreads) void k_pcg_direction(int64_t n, const double* __restrict__ r,
                                                      const double* __restrict__ dinv,
                                                      const double* p0c, const double* p1c,
                                                      double* p0, double* p1, PcgScalars* sc,
                                                      const double* __restrict__ partB, int nbB) {
    __shared__ double red[kWaves];
    if (sc->done) return;
    const long long it = sc->iter;
    const long long k = it + 1;
    const double* pold = ((k - 1) & 1) ? p1c : p0c;
    double* pnew = (k & 1) ? p1 : p0;
    const double rz = reduce_partials(partB, nbB, 2, 0, red);
    const double rr = reduce_partials(partB, nbB, 2, 1, red);
    const double tol_in = sc->tol;
    const double tol = tol_in < 0.0 ? fmax(sc->reltol * sqrt(rr), sc->abstol) : tol_in;
    int err;
    const bool stop = pcg_stop_test(sc, rz, rr, it, tol, err);
    const double rho_prev = sc->rho[(k - 1) & 1];
    // Block 0's writes below cannot change what another block computes: rho
    // goes to the other slot, a block that sees the new tol derives the same
    // value, and one that sees done = 1 would stop anyway.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rr = rr;
        if (tol_in < 0.0) {
            sc->tol = tol;
            sc->rr0 = rr;
        }
        if (stop) {
            sc->done = 1;
            if (err) sc->err = err;
        } else {
            sc->rho[k & 1] = rz;
            sc->rho_cur = rz;
            sc->k_cur = k;
        }
    }
    if (stop) return;
    const double beta = rz / rho_prev;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        pnew[i] = dinv[i] * r[i] + beta * pold[i];
}

static unsigned csr_blocks(int64_t work, int per_block) {
    int64_t b = (work + per_block - 1) / per_block;
    if (b > kMaxPartials) b = kMaxPartials;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_csr_build(const Grid& g, const double* Ke, const double* E,
                            const uint8_t* fixmask, const int64_t* rowptr, int* cols, int safe,
                            double* vals, double* diag, double* partials, double* mdiag_sum,
                            bool pattern, hipStream_t s) {
    const unsigned nbn = csr_blocks(g.nnodes, kThreads);
    if (pattern) {
        if (g.dim == 3)
            hipLaunchKernelGGL((k_csr_cols<3>), dim3(nbn), dim3(kThreads), 0, s, g, rowptr, cols);
        else
            hipLaunchKernelGGL((k_csr_cols<2>), dim3(nbn), dim3(kThreads), 0, s, g, rowptr, cols);
    }
    if (g.dim == 3)
        hipLaunchKernelGGL((k_csr_diag<3>), dim3(nbn), dim3(kThreads), 0, s, g, Ke, E, fixmask,
                           diag, partials);
    else
        hipLaunchKernelGGL((k_csr_diag<2>), dim3(nbn), dim3(kThreads), 0, s, g, Ke, E, fixmask,
                           diag, partials);
    hipError_t e = launch_reduce(partials, (int)nbn, 3, mdiag_sum, s);
    if (e != hipSuccess) return e;
    const int no = g.dim == 3 ? 27 : 9;
    const unsigned nbv = csr_blocks(g.nnodes * no, kThreads);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_csr_values<3>), dim3(nbv), dim3(kThreads), 0, s, g, Ke, E, fixmask,
                           rowptr, mdiag_sum, safe, vals);
    else
        hipLaunchKernelGGL((k_csr_values<2>), dim3(nbv), dim3(kThreads), 0, s, g, Ke, E, fixmask,
                           rowptr, mdiag_sum, safe, vals);
    return hipGetLastError();
}

hipError_t launch_csr_dinv(const Grid& g, const double* diag, const uint8_t* fixmask,
                           const double* mdiag_sum, int jacobi, int safe, double* dinv,
                           hipStream_t s) {
    const unsigned nb = csr_blocks(g.ndof, kThreads);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_csr_dinv<3>), dim3(nb), dim3(kThreads), 0, s, g.ndof, diag, fixmask,
                           mdiag_sum, jacobi, safe, dinv);
    else
        hipLaunchKernelGGL((k_csr_dinv<2>), dim3(nb), dim3(kThreads), 0, s, g.ndof, diag, fixmask,
                           mdiag_sum, jacobi, safe, dinv);
    return hipGetLastError();
}

hipError_t launch_csr_spmv(int64_t nrows, const int64_t* rowptr, const int* cols,
                           const double* vals, const double* x, double* y, hipStream_t s) {
    hipLaunchKernelGGL((k_csr_spmv<false>), dim3(csr_blocks(nrows, kThreads / 16)),
                       dim3(kThreads), 0, s, nrows, rowptr, cols, vals, x, x, y, nullptr,
                       nullptr);
    return hipGetLastError();
}

hipError_t launch_pcg_direction(int64_t n, const double* r, const double* dinv, double* p0,
                                double* p1, PcgScalars* sc, const double* partB, int nbB,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_direction, dim3(kVecBlocks), dim3(kThreads), 0, s, n, r, dinv, p0,
                       p1, p0, p1, sc, partB, nbB);
    return hipGetLastError();
}

hipError_t launch_csr_pcg_iter(int64_t n, const int64_t* rowptr, const int* cols,
                               const double* vals, double* x, double* r, const double* dinv,
                               double* p0, double* p1, double* q, PcgScalars* sc, double* partA,
                               double* partB, hipStream_t s) {
    if (hipError_t e = launch_pcg_direction(n, r, dinv, p0, p1, sc, partB, kVecBlocks, s))
        return e;
    const unsigned nbA = csr_blocks(n, kThreads / 16);
    hipLaunchKernelGGL((k_csr_spmv<true>), dim3(nbA), dim3(kThreads), 0, s, n, rowptr, cols,
                       vals, p0, p1, q, sc, partA);
    return launch_pcg_update<double>(n, x, r, p0, p1, q, dinv, sc, partA, (int)nbA, partB, s);
}

}  // namespace tohip
