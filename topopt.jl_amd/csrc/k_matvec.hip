// Matrix-free K(E)·x on the structured Q1 grid (replaces mul!(y,
// ::MatrixFreeOperator, x), TopOpt.jl src/FEA/matrix_free_operator.jl:66-105).
//
// One workgroup owns a TX x TY (x TZ) tile of output nodes (persistent loop
// over tiles).  It stages the input vector on the tile plus a one-node halo in
// LDS (structure of arrays; prescribed DOFs masked to 0 like bcmatrix), lets
// one thread per cell of the tile's neighbourhood compute that cell's 24 (8)
// nodal results into an LDS slot array (cells outside the mesh write zeros,
// halo cells are recomputed by the neighbouring tile: no atomics, no
// colouring passes, no branches per corner), and finally each output node
// sums its 8 (4) cell slots in ascending cell order — the order of the
// reference's dof_cells gather — so results are deterministic run to run.
// Prescribed DOFs return meandiag * x.
//
// The cell product uses the corner-Hadamard factorisation Ke = H B H^T with
// the compact block form of ElemOp: 2 x 72 add/sub + 45 mul/FMA per cell
// instead of the 576 FMA of a dense 24x24 product.
#include <cstdlib>

#include "k_common.h"

namespace tohip {

// Tile shapes (output nodes per workgroup) and workgroup sizes; 3D variants
// are selectable at run time (TOHIP_TILE32 / TOHIP_TILE64) for measurement.
struct TileShape {
    int tx, ty, tz, threads;
};
constexpr TileShape kTiles3[] = {
    {8, 8, 2, 256}, {8, 4, 2, 128}, {8, 8, 4, 256}, {16, 4, 2, 256}, {4, 4, 4, 128}};
constexpr int kNumTiles3 = 5;
constexpr TileShape kTile2 = {32, 8, 1, 256};

// Measured on MI355X (PCG direction + update pair, 120x40x40 and 256x128x128
// cantilevers, grid capped at 1024 workgroups): FP32 is fastest with 16x4x2
// tiles of 256 threads at both sizes; FP64 with 4x4x4 tiles of 128 threads
// below ~1M cells (more workgroups in flight at 3 waves/SIMD) and 16x4x2 above
// (less halo recomputation per output node).
template <typename T>
static int tile_variant(const Grid& g) {
    static int forced = [] {
        const char* e = getenv(sizeof(T) == 4 ? "TOHIP_TILE32" : "TOHIP_TILE64");
        int t = e ? atoi(e) : -1;
        return (t >= 0 && t < kNumTiles3) ? t : -1;
    }();
    if (forced >= 0) return forced;
    if (sizeof(T) == 4) return 3;
    return g.ncells < (int64_t)1 << 20 ? 4 : 3;
}

template <typename T>
static inline TileShape tile_of(const Grid& g) {
    return g.dim == 3 ? kTiles3[tile_variant<T>(g)] : kTile2;
}

template <typename T>
static inline int ntiles(const Grid& g) {
    const TileShape t = tile_of<T>(g);
    int n = ((g.NX + t.tx - 1) / t.tx) * ((g.NY + t.ty - 1) / t.ty);
    if (g.dim == 3) n *= (g.NZ + t.tz - 1) / t.tz;
    return n;
}

// Persistent grid: at most kMatvecBlocksCap workgroups (TOHIP_MVCAP overrides
// it for measurement, within [256, kMaxPartials]).
static inline int grid_blocks(int nt) {
    static int cap = [] {
        const char* e = getenv("TOHIP_MVCAP");
        const int c = e ? atoi(e) : kMatvecBlocksCap;
        return (c >= 256 && c <= kMaxPartials) ? c & ~7 : kMatvecBlocksCap;
    }();
    return nt <= cap ? nt : cap;  // multiple of 8: keeps the XCD tile ranges contiguous
}

// 3D grids of at least 2^20 cells take the z-marching kernel
// (k_matvec_zm.hip); smaller grids and 2D the tile kernel below.  Measured on
// MI355X: at 256x128x128 z-marching runs the FP64 matvec in 114 us (tiles:
// 232 us) and the FP32 PCG direction + K p in 128 us (164 us); at 120x40x40,
// where the whole working set is cached and ~240 marching workgroups leave
// the chip half idle, tiles are faster (FP32 PCG 14.4 us, z-marching 19.3).
// TOHIP_MV_ZM = 0 / 1 / 2: never / default threshold / every 3D grid
// (measurement and testing).
bool zm_fits(const Grid& g);
int zm_blocks(const Grid& g);
template <typename T, bool PCG>
void launch_matvec_zm(const Grid& g, const ElemOp<T>& op, const T* xin, const T* dinv, T* p0,
                      T* p1, const uint8_t* fixmask, const T* E, T* y, double meandiag,
                      PcgScalars* sc, const double* partB, int nbB, double* partA, bool pre,
                      hipStream_t s);
static bool use_zm(const Grid& g) {
    static int mode = [] {
        const char* e = getenv("TOHIP_MV_ZM");
        return e ? atoi(e) : 1;
    }();
    if (mode == 0 || !zm_fits(g)) return false;
    return mode == 2 || g.ncells >= zm_min_cells();
}

template <typename T>
int pcg_blocks(const Grid& g) {
    return use_zm(g) ? zm_blocks(g) : grid_blocks(ntiles<T>(g));
}
int pcg_dirmatvec_blocks(const Grid& g) { return pcg_blocks<double>(g); }
int pcg_dirmatvec_blocks32(const Grid& g) { return pcg_blocks<float>(g); }

// PCG = false: y = K x.
// PCG = true : iteration k = iter + 1 of Jacobi PCG.  Every block reduces the
//   previous update's partials (r.z, r.r), applies the stopping rule, then
//   p_k = dinv.*r + beta p_{k-1} on the fly (r passed as xin), p_k written for
//   owned nodes, y = q = K p_k, block partial of p.q into partA.
template <typename T, int DIM, bool PCG, int TX, int TY, int TZ, int NT, bool CUBE, bool PRE>
__global__ __launch_bounds__(NT) void k_matvec(
    Grid g, ElemOp<T> op, const T* __restrict__ xin, const T* __restrict__ dinv,
    T* __restrict__ p0, T* __restrict__ p1, const uint8_t* __restrict__ fixmask,
    const T* __restrict__ E, T* __restrict__ y, double meandiag, PcgScalars* sc,
    const double* __restrict__ partB, int nbB, double* __restrict__ partA) {
    constexpr int NC = 1 << DIM;
    constexpr int IX = TX + 2, IY = TY + 2, IZ = DIM == 3 ? TZ + 2 : 1;
    constexpr int EX = TX + 1, EY = TY + 1, EZ = DIM == 3 ? TZ + 1 : 1;
    constexpr int NIN = IX * IY * IZ, NOUT = TX * TY * TZ, NEL = EX * EY * EZ;
    constexpr int NLN = (NIN + NT - 1) / NT;   // staged nodes per thread
    constexpr int NEP = (NEL + NT - 1) / NT;   // cells per thread
    constexpr int NOP = (NOUT + NT - 1) / NT;  // output nodes per thread

    __shared__ T xs[DIM][NIN];
    __shared__ T res[DIM * NC][NEL];
    __shared__ T pown[DIM][NOUT];
    __shared__ uint8_t smask[NOUT];
    __shared__ double red[kWaves];

    const int tid = threadIdx.x;
    T beta = 0;
    T* pold = nullptr;
    T* pnew = nullptr;
    long long it = 0;
    if (PCG) {
        if (sc->done) return;
        it = sc->iter;
        const long long k = it + 1;
        pold = ((k - 1) & 1) ? p1 : p0;
        pnew = (k & 1) ? p1 : p0;
    }

    const int ntx = (g.NX + TX - 1) / TX, nty = (g.NY + TY - 1) / TY;
    const int ntz = DIM == 3 ? (g.NZ + TZ - 1) / TZ : 1;
    const int nt = ntx * nty * ntz;
    double part = 0.0;

    // Tile-relative coordinates are recomputed where needed (cheap constant
    // divisions) rather than held in registers across the tile loop.  Global
    // indices are 32-bit: tohip_create rejects grids with 2^29 DOFs or more.
    T ev[NEP];
    T va[NLN][DIM], vb[NLN][DIM], vc[NLN][DIM];
    uint8_t vm[NLN];
    bool vin[NLN];
    // Issue every global load of one tile's stage: cell moduli, then nodes.
    // Halo positions outside the mesh load a clamped (valid) address and are
    // zeroed afterwards, so the loads need no branches.
    auto load_tile = [&](int i0, int j0, int k0) {
#pragma unroll
        for (int q = 0; q < NEP; ++q) {
            const int el = tid + q * NT;
            const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
            const int gi = i0 - 1 + ei, gj = j0 - 1 + ej, gk = DIM == 3 ? k0 - 1 + ek : 0;
            const int ci = min(max(gi, 0), g.nx - 1), cj = min(max(gj, 0), g.ny - 1),
                      ck = min(max(gk, 0), g.nz - 1);
            const T e = at(E, (uint32_t)ci + (uint32_t)g.nx * ((uint32_t)cj + (uint32_t)g.ny * ck));
            ev[q] = (el < NEL && gi == ci && gj == cj && gk == ck) ? e : (T)0;
        }
#pragma unroll
        for (int q = 0; q < NLN; ++q) {
            const int nl = tid + q * NT;
            const int li = nl % IX, lj = (nl / IX) % IY, lk = DIM == 3 ? nl / (IX * IY) : 0;
            const int gi = i0 - 1 + li, gj = j0 - 1 + lj, gk = DIM == 3 ? k0 - 1 + lk : 0;
            const int ci = min(max(gi, 0), g.NX - 1), cj = min(max(gj, 0), g.NY - 1),
                      ck = min(max(gk, 0), g.NZ - 1);
            const uint32_t n = (uint32_t)ci + (uint32_t)g.NX * ((uint32_t)cj + (uint32_t)g.NY * ck);
            vin[q] = gi == ci && gj == cj && gk == ck;
            if (nl < NIN) {
                const uint8_t m = fixmask[n];
                vm[q] = vin[q] ? m : (uint8_t)0xff;  // outside the mesh: all components 0
                ld_node<DIM>(xin, n, va[q]);
                if (PCG) {
                    // dinv == nullptr: unit diagonal (multigrid path), not read
                    if (dinv) {
                        ld_node<DIM>(dinv, n, vb[q]);
                    } else {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) vb[q][d] = (T)1;
                    }
                    ld_node<DIM>((const T*)pold, n, vc[q]);
                }
            }
        }
    };
    auto tile_origin = [&](int tb, int& i0, int& j0, int& k0) {
        const int b = xcd_remap(tb, nt);
        const int tx = b % ntx, ty = (b / ntx) % nty, tz = DIM == 3 ? b / (ntx * nty) : 0;
        i0 = tx * TX;
        j0 = ty * TY;
        k0 = tz * TZ;
    };

    // The first tile's loads do not depend on beta: they go out before the
    // previous kernel's partials are reduced, so the two latencies overlap.
    if (PRE && (int)blockIdx.x < nt) {
        int i0, j0, k0;
        tile_origin(blockIdx.x, i0, j0, k0);
        load_tile(i0, j0, k0);
    }
    if (PCG) {
        const double rz = reduce_partials(partB, nbB, 2, 0, red);
        const double rr = reduce_partials(partB, nbB, 2, 1, red);
        const long long k = it + 1;
        const double tol_in = sc->tol;
        // tol < 0: first iteration after a start; every block derives the same value.
        const double tol = tol_in < 0.0 ? fmax(sc->reltol * sqrt(rr), sc->abstol) : tol_in;
        int err;
        const bool stop = pcg_stop_test(sc, rz, rr, it, tol, err);
        const double rho_prev = sc->rho[(k - 1) & 1];
        if (blockIdx.x == 0 && tid == 0) {
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
        beta = (T)(rz / rho_prev);
    }

    for (int tb = blockIdx.x; tb < nt; tb += gridDim.x) {
        int i0, j0, k0;
        tile_origin(tb, i0, j0, k0);
        if (!PRE || tb != (int)blockIdx.x) load_tile(i0, j0, k0);
        // ---- stage into LDS (and write p_k of the owned nodes) ----
#pragma unroll
        for (int q = 0; q < NLN; ++q) {
            const int nl = tid + q * NT;
            if (nl < NIN) {
                const int li = nl % IX, lj = (nl / IX) % IY, lk = DIM == 3 ? nl / (IX * IY) : 0;
                const bool owned = vin[q] && li >= 1 && li <= TX && lj >= 1 && lj <= TY &&
                                   (DIM == 2 || (lk >= 1 && lk <= TZ));
                const int o = (li - 1) + TX * ((lj - 1) + TY * (DIM == 3 ? lk - 1 : 0));
                const uint32_t n = (uint32_t)(i0 - 1 + li) +
                                   (uint32_t)g.NX * ((uint32_t)(j0 - 1 + lj) +
                                                     (uint32_t)g.NY * (DIM == 3 ? k0 - 1 + lk : 0));
                T val[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    val[d] = PCG ? vb[q][d] * va[q][d] + beta * vc[q][d] : va[q][d];
                    xs[d][nl] = ((vm[q] >> d) & 1) ? (T)0 : val[d];
                    if (owned) pown[d][o] = val[d];
                }
                if (PCG && owned) st_node<DIM>(pnew, n, val);
                if (owned) smask[o] = vm[q];
            }
        }
        __syncthreads();

        // ---- one thread per cell: Ke x_e scaled by E_e into the slots ----
#pragma unroll
        for (int q = 0; q < NEP; ++q) {
            const int el = tid + q * NT;
            if (el >= NEL) continue;
            const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
            const T Ee = ev[q];
            if (Ee != (T)0) {
                T xe[DIM][NC], ye[DIM][NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const int nl = (ei + (c & 1)) +
                                   IX * ((ej + ((c >> 1) & 1)) + IY * (ek + ((c >> 2) & 1)));
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xe[d][c] = xs[d][nl];
                }
                elem_apply<T, DIM, CUBE>(op, xe, ye);
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) res[c * DIM + d][el] = Ee * ye[d][c];
            } else {
#pragma unroll
                for (int j = 0; j < DIM * NC; ++j) res[j][el] = 0;
            }
        }
        __syncthreads();

        // ---- per-node gather in ascending cell order ----
#pragma unroll
        for (int q = 0; q < NOP; ++q) {
            const int o = tid + q * NT;
            if (o >= NOUT) continue;
            const int oi = o % TX, oj = (o / TX) % TY, ok = DIM == 3 ? o / (TX * TY) : 0;
            const int gi = i0 + oi, gj = j0 + oj, gk = k0 + ok;
            if (gi >= g.NX || gj >= g.NY || gk >= g.NZ) continue;
            const uint32_t n = (uint32_t)gi + (uint32_t)g.NX * ((uint32_t)gj + (uint32_t)g.NY * gk);
            const uint8_t fm = smask[o];
            T s[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) s[d] = 0;
#pragma unroll
            for (int c = NC - 1; c >= 0; --c) {
                const int e = (oi + 1 - (c & 1)) +
                              EX * ((oj + 1 - ((c >> 1) & 1)) +
                                    EY * (DIM == 3 ? ok + 1 - ((c >> 2) & 1) : 0));
#pragma unroll
                for (int d = 0; d < DIM; ++d) s[d] += res[c * DIM + d][e];
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                if ((fm >> d) & 1) s[d] = (T)meandiag * pown[d][o];
                if (PCG) part += (double)pown[d][o] * (double)s[d];
            }
            st_node<DIM>(y, n, s);
        }
        __syncthreads();  // LDS is reused by the next tile
    }

    if (PCG) {
        const double bs = block_sum_all(part, red);
        if (tid == 0) partA[blockIdx.x] = bs;
    }
}

// PCG variant: issue the first tile's loads before reducing the previous
// kernel's partials (overlaps the two latencies, costs registers).
// TOHIP_MVPRE = 0 / 1 overrides the per-precision default for measurement.
template <typename T>
static bool preload_first_tile() {
    static int v = [] {
        const char* e = getenv("TOHIP_MVPRE");
        return e ? (atoi(e) != 0) : (sizeof(T) == 8);
    }();
    return v != 0;
}

// Launches k_matvec with the selected tile shape.
template <typename T, bool PCG>
static void dispatch_matvec(const Grid& g, int nb, hipStream_t s, const ElemOp<T>& op,
                            const T* xin, const T* dinv, T* p0, T* p1, const uint8_t* fixmask,
                            const T* E, T* y, double meandiag, PcgScalars* sc,
                            const double* partB, int nbB, double* partA) {
    const bool pre = PCG && preload_first_tile<T>();
    if (use_zm(g)) {
        launch_matvec_zm<T, PCG>(g, op, xin, dinv, p0, p1, fixmask, E, y, meandiag, sc, partB, nbB,
                                 partA, pre, s);
        return;
    }
#define TOHIP_MV1(D, X, Y, Z, N, CB, PR)                                                       \
    hipLaunchKernelGGL((k_matvec<T, D, PCG, X, Y, Z, N, CB, PR>), dim3(nb), dim3(N), 0, s, g, \
                       op, xin, dinv, p0, p1, fixmask, E, y, meandiag, sc, partB, nbB, partA)
#define TOHIP_MV(D, X, Y, Z, N, CB)           \
    if (pre)                                  \
        TOHIP_MV1(D, X, Y, Z, N, CB, PCG);    \
    else                                      \
        TOHIP_MV1(D, X, Y, Z, N, CB, false)
#define TOHIP_MV3(X, Y, Z, N)        \
    if (op.is_cube)                  \
        TOHIP_MV(3, X, Y, Z, N, true); \
    else                             \
        TOHIP_MV(3, X, Y, Z, N, false)
    if (g.dim == 2) {
        TOHIP_MV(2, 32, 8, 1, 256, false);
        return;
    }
    switch (tile_variant<T>(g)) {
        case 0: TOHIP_MV3(8, 8, 2, 256); break;
        case 1: TOHIP_MV3(8, 4, 2, 128); break;
        case 2: TOHIP_MV3(8, 8, 4, 256); break;
        case 3: TOHIP_MV3(16, 4, 2, 256); break;
        default: TOHIP_MV3(4, 4, 4, 128); break;
    }
#undef TOHIP_MV3
#undef TOHIP_MV
#undef TOHIP_MV1
}

template <typename T>
hipError_t launch_matvec(const Grid& g, const ElemOp<T>& op, const T* x, const uint8_t* fixmask,
                         const T* E, T* y, double meandiag, hipStream_t s) {
    dispatch_matvec<T, false>(g, pcg_blocks<T>(g), s, op, x, (const T*)nullptr,
                              (T*)nullptr, (T*)nullptr, fixmask, E, y, meandiag,
                              (PcgScalars*)nullptr, (const double*)nullptr, 0, (double*)nullptr);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_dirmatvec(const Grid& g, const ElemOp<T>& op, const T* r, const T* dinv,
                                T* p0, T* p1, T* q, const uint8_t* fixmask, const T* E,
                                double meandiag, PcgScalars* sc, const double* partB, int nbB,
                                double* partA, int* nbA_out, hipStream_t s) {
    const int nb = pcg_blocks<T>(g);
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    if (nbA_out) *nbA_out = nb;
    dispatch_matvec<T, true>(g, nb, s, op, r, dinv, p0, p1, fixmask, E, q, meandiag, sc, partB,
                             nbB, partA);
    return hipGetLastError();
}

template hipError_t launch_matvec<double>(const Grid&, const ElemOp<double>&, const double*,
                                          const uint8_t*, const double*, double*, double,
                                          hipStream_t);
template hipError_t launch_matvec<float>(const Grid&, const ElemOp<float>&, const float*,
                                         const uint8_t*, const float*, float*, double,
                                         hipStream_t);
template hipError_t launch_pcg_dirmatvec<double>(const Grid&, const ElemOp<double>&,
                                                 const double*, const double*, double*, double*,
                                                 double*, const uint8_t*, const double*, double,
                                                 PcgScalars*, const double*, int, double*, int*,
                                                 hipStream_t);
template hipError_t launch_pcg_dirmatvec<float>(const Grid&, const ElemOp<float>&, const float*,
                                                const float*, float*, float*, float*,
                                                const uint8_t*, const float*, double,
                                                PcgScalars*, const double*, int, double*, int*,
                                                hipStream_t);

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

// rho_e = xmin + (1-xmin) P(x)  (penalty before interpolation) or
//         P(xmin + (1-xmin) x), P the selected penalty family with its optional
//         projection (get_rho, TopOpt.jl src/Utilities/penalties.jl:113-130).
__global__ __launch_bounds__(kThreads) void k_density(int64_t nel, const double* __restrict__ xt,
                                                      double xmin, Penalty pen, int pbi,
                                                      double* __restrict__ E64,
                                                      float* __restrict__ E32) {
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (e >= nel) return;
    double r, dr;
    rho_drho(pen, xt[e], xmin, pbi, r, dr);
    E64[e] = r;
    if (E32) E32[e] = (float)r;
}

hipError_t launch_density(int64_t nel, const double* xt, double xmin, const Penalty& pen, int pbi,
                          double* E64, float* E32, hipStream_t s) {
    hipLaunchKernelGGL(k_density, dim3((unsigned)((nel + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, nel, xt, xmin, pen, pbi, E64, E32);
    return hipGetLastError();
}

}  // namespace tohip
