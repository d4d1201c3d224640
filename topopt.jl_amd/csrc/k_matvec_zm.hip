// Matrix-free K(E)·x on 3D grids by z-marching columns (replaces mul!(y,
// ::MatrixFreeOperator, x), TopOpt.jl src/FEA/matrix_free_operator.jl:66-105,
// and the fused direction + matvec step of its Jacobi PCG).
//
// One 256-thread workgroup owns a column of 15 x 15 output nodes and a chunk
// of node planes [kb, ke).  It walks up the chunk one cell layer at a time:
//   * the next node plane (17 x 17 nodes: the column plus a one-node halo) is
//     staged in a two-plane LDS ring with prescribed DOFs zeroed (the cells'
//     gathers need no test), its own 15 x 15 nodes also unmasked for the
//     prescribed-DOF rule; the plane after it is already being loaded into
//     registers;
//   * thread (ci, cj) of the 16 x 16 cell layer applies E_e Ke to its cell
//     (corner-Hadamard factorisation, prescribed DOFs masked to 0 as bcmatrix
//     does);
//   * the corner sums of the nodes are formed without an LDS slot array: the
//     left neighbour's x = 1 corners come through a DPP row shift (a wave row
//     of 16 lanes is one row of cells), the row below's partial through a
//     small LDS exchange, and the sum over the layer below is carried in
//     registers from one step to the next.
// Every input node and cell modulus is read from memory once per chunk (plus
// the one-node xy halo, shared with the neighbouring columns through L2),
// every output once.  The sum at a node adds the layer below, then the layer
// above; within a layer the lower row, then the upper one; within a row the
// lower x cell first: ascending cell order in pairs.  Deterministic run to
// run.  Prescribed DOFs return meandiag * x.
#include <algorithm>
#include <cstdlib>

#include "k_common.h"

namespace tohip {

namespace {

constexpr int kZC = 16;                  // cells per row and rows per layer
constexpr int kZT = kZC * kZC;           // threads (one cell each)
constexpr int kZO = kZC - 1;             // output nodes per row (15)
constexpr int kZS = kZC + 1;             // staged nodes per row (17)
constexpr int kZSP = kZS * kZS;          // staged nodes per plane (289)
constexpr int kZL = (kZSP + kZT - 1) / kZT;  // staged nodes per thread (2)

// v of lane l - 1 within each 16-lane row (0 in the row's first lane).
__device__ __forceinline__ float row_shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}
__device__ __forceinline__ double row_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x111, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x111, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace

// Chunking: columns of 15 x 15 nodes, node planes split so the launch has
// about `target` workgroups (TOHIP_ZM_WGS overrides it for measurement),
// at least 4 planes per chunk, at most kMaxPartials workgroups in all.
struct ZmShape {
    int ncx, ncy, zc, nch;
    int blocks() const { return ncx * ncy * nch; }
};
static ZmShape zm_shape(const Grid& g) {
    static int target = [] {
        const char* e = getenv("TOHIP_ZM_WGS");
        const int t = e ? atoi(e) : 1536;
        return (t >= 64 && t <= kMaxPartials) ? t : 1536;
    }();
    // TOHIP_ZM_MINP: fewest node planes per chunk (measurement; default 4)
    static int minp = [] {
        const char* e = getenv("TOHIP_ZM_MINP");
        const int v = e ? atoi(e) : 4;
        return (v >= 1 && v <= 64) ? v : 4;
    }();
    ZmShape s;
    s.ncx = (g.NX + kZO - 1) / kZO;
    s.ncy = (g.NY + kZO - 1) / kZO;
    const int ncol = s.ncx * s.ncy;
    int nch = (target + ncol - 1) / ncol;
    nch = std::max(1, std::min(nch, g.NZ / minp));
    nch = std::min(nch, std::max(1, kMaxPartials / ncol));
    s.zc = (g.NZ + nch - 1) / nch;
    s.nch = (g.NZ + s.zc - 1) / s.zc;
    return s;
}

bool zm_fits(const Grid& g) { return g.dim == 3 && zm_shape(g).blocks() <= kMaxPartials; }
int zm_blocks(const Grid& g) { return zm_shape(g).blocks(); }

// Multigrid smoothing on the finest level (MG = 1, 2), FP32: the weight, the
// right-hand side b, the outputs and the partials of the post-smoothing.
struct ZmMgArgs {
    float w;
    const float* wdev;  // the weight in device memory when given
    const float* b;
    float* xout;
    float* rout;
    double* part;
    const PcgScalars* sc;
    const float* xc;  // MG = 2: the coarser level's correction (x' = xin + P xc), or null
    Grid gc;          // and its grid
};

// PCG = false: y = K x.
// PCG = true : iteration k = iter + 1 of Jacobi PCG, as k_matvec<PCG>: every
//   block reduces the previous update's partials (r.z, r.r), applies the
//   stopping rule, stages p_k = dinv .* r + beta p_{k-1} (writing p_k of the
//   nodes it owns), y = q = K p_k and a block partial of p.q into partA.
// MG = 1: the multigrid pre-smoothing (xin = b): stages x = w dinv b and
//   writes xout = x, rout = b - K x.
// MG = 2: the multigrid post-smoothing on x' = xin + P xc: writes xout =
//   z = x' + w dinv (b - K x') and, with mg.part, the block partials (b.z,
//   b.b).  With mg.xc the prolongation is formed while staging: the column's
//   coarse nodes (10 x 10 per coarse plane) go through a two-plane LDS ring,
//   coarse plane C written at step 2C - 3 after the first barrier (its slot
//   last held plane C - 2, last read by step 2C - 4's staging) and first read
//   by step 2C - 2's staging; the sum follows k_mg_prolong_add's order, so x'
//   has the same bits.  Without mg.xc, xin is x' already.
//   Prescribed DOFs: x = 0, r = 0, z = 0.
template <typename T, bool PCG, bool CUBE, bool PRE, int PD, int MG = 0>
__global__ __launch_bounds__(kZT, MG == 2 ? 5 : 1) void k_matvec_zm(
    Grid g, ElemOp<T> op, int zc, const T* __restrict__ xin, const T* __restrict__ dinv,
    T* __restrict__ p0, T* __restrict__ p1, const uint8_t* __restrict__ fixmask,
    const T* __restrict__ E, T* __restrict__ y, double meandiag, PcgScalars* sc,
    const double* __restrict__ partB, int nbB, double* __restrict__ partA, ZmMgArgs mg) {
    static_assert(MG == 0 || (!PCG && sizeof(T) == 4), "smoothing: FP32, no PCG");
    constexpr int NO = kZO * kZO;
    __shared__ T xs[2][3][kZSP];       // staged node planes, prescribed DOFs and outside: 0
    // the column's own nodes, unmasked (epilogue); smoothing reads xs there
    // (x is 0 at prescribed DOFs, so the masked value is the one it needs)
    __shared__ T xo[MG ? 1 : 2][3][MG ? 1 : NO];
    constexpr int CNN = 10, CNP = CNN * CNN;  // coarse nodes of a column's tile, per row / plane
    __shared__ T xcs[MG == 2 ? 2 : 1][3][MG == 2 ? CNP : 1];  // post: coarse plane ring
    __shared__ uint8_t ms[2][kZSP];    // prescribed-DOF bits (0xff outside the mesh)
    __shared__ T xch[6][kZT];          // row partials sent to the row above
    __shared__ double red[kZT / 64];
    __shared__ T bo_s[MG ? 2 : 1][3][MG ? NO : 1];       // smoothing: b of the own nodes
    __shared__ T do_s[MG == 2 ? 2 : 1][3][MG == 2 ? NO : 1];  // post: dinv of the own nodes
    if (MG && mg.sc && mg.sc->done) return;
    if (MG && mg.wdev) mg.w = *mg.wdev;
    const bool prol = MG == 2 && mg.xc != nullptr;

    const int tid = threadIdx.x;
    const int ci = tid % kZC, cj = tid / kZC;
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
    const int ncx = (g.NX + kZO - 1) / kZO, ncy = (g.NY + kZO - 1) / kZO;
    const int nch = (g.NZ + zc - 1) / zc;
    const int b = xcd_remap(blockIdx.x, ncx * ncy * nch);
    const int i0 = (b % ncx) * kZO, j0 = ((b / ncx) % ncy) * kZO;
    const int kb = (b / (ncx * ncy)) * zc, ke = min(kb + zc, g.NZ);

    // ---- register stage of one node plane and one cell layer ----
    struct Stage {
        T va[kZL][3], vb[kZL][3], vc[kZL][3];
        uint8_t vm[kZL];
        T ev;
    };
    // node plane q (staged nodes i0 - 1 .. i0 + 15, j0 - 1 .. j0 + 15); clamped
    // addresses, values outside the mesh flagged 0xff and zeroed at use
    auto load_plane = [&](Stage& R, int q) {
#pragma unroll
        for (int l = 0; l < kZL; ++l) {
            const int s = min(tid + l * kZT, kZSP - 1);
            const int gi = i0 - 1 + s % kZS, gj = j0 - 1 + s / kZS;
            const int ci_ = min(max(gi, 0), g.NX - 1), cj_ = min(max(gj, 0), g.NY - 1),
                      ck_ = min(max(q, 0), g.NZ - 1);
            const uint32_t n = (uint32_t)ci_ + (uint32_t)g.NX * ((uint32_t)cj_ + (uint32_t)g.NY * ck_);
            const bool in = gi == ci_ && gj == cj_ && q == ck_;
            const uint8_t m = fixmask[n];
            R.vm[l] = in ? m : (uint8_t)0xff;
            ld_node<3>(xin, n, R.va[l]);
            if (PCG) {
                // dinv == nullptr: unit diagonal (multigrid path), not read
                if (dinv) {
                    ld_node<3>(dinv, n, R.vb[l]);
                } else {
#pragma unroll
                    for (int d = 0; d < 3; ++d) R.vb[l][d] = (T)1;
                }
                ld_node<3>((const T*)pold, n, R.vc[l]);
            }
            if (MG == 1) ld_node<3>(dinv, n, R.vb[l]);
            if (MG == 2) {
                ld_node<3>((const T*)mg.b, n, R.vb[l]);
                ld_node<3>(dinv, n, R.vc[l]);
            }
        }
    };
    // post-smoothing with the prolongation: coarse node (CI0 + t % 10,
    // CJ0 + t / 10) of a coarse plane for thread t < 100 (clamped; the
    // column's fine nodes in the mesh only reach coarse nodes in the grid)
    const int CI0 = (i0 - 1) >> 1, CJ0 = (j0 - 1) >> 1;
    T cr[3] = {0, 0, 0};
    auto load_coarse = [&](int K, T(&v)[3]) {
        if (MG == 2 && tid < CNP) {
            const int I = min(max(CI0 + tid % CNN, 0), mg.gc.NX - 1);
            const int J = min(max(CJ0 + tid / CNN, 0), mg.gc.NY - 1);
            const int Kc = min(max(K, 0), mg.gc.NZ - 1);
            ld_node<3>((const T*)mg.xc,
                       (uint32_t)I + (uint32_t)mg.gc.NX * ((uint32_t)J + (uint32_t)mg.gc.NY * Kc), v);
        }
    };
    auto write_coarse = [&](int K, const T(&v)[3]) {
        if (MG == 2 && tid < CNP)
#pragma unroll
            for (int d = 0; d < 3; ++d) xcs[MG == 2 ? (K & 1) : 0][d][MG == 2 ? tid : 0] = v[d];
    };
    // cell layer q: thread (ci, cj) <-> cell (i0 - 1 + ci, j0 - 1 + cj, q)
    auto load_cell = [&](Stage& R, int q) {
        const int gi = i0 - 1 + ci, gj = j0 - 1 + cj;
        const int ci_ = min(max(gi, 0), g.nx - 1), cj_ = min(max(gj, 0), g.ny - 1),
                  ck_ = min(max(q, 0), g.nz - 1);
        const T e = at(E, (uint32_t)ci_ + (uint32_t)g.nx * ((uint32_t)cj_ + (uint32_t)g.ny * ck_));
        R.ev = (gi == ci_ && gj == cj_ && q == ck_) ? e : (T)0;
    };
    // registers -> LDS ring slot of plane q (PCG: p_k, written back for owned nodes)
    auto stage_plane = [&](const Stage& R, int q, bool owned_plane) {
        const int slot = q & 1;
#pragma unroll
        for (int l = 0; l < kZL; ++l) {
            const int s = tid + l * kZT;
            if (s < kZSP) {
                const int li = s % kZS, lj = s / kZS;
                const uint8_t m = R.vm[l];
                // owned: the column's own 15 x 15 nodes (li = 16 is the next column's)
                const bool owned = owned_plane && m != 0xff && li >= 1 && li <= kZO && lj >= 1 &&
                                   lj <= kZO;
                const uint32_t n = (uint32_t)(i0 - 1 + li) +
                                   (uint32_t)g.NX * ((uint32_t)(j0 - 1 + lj) + (uint32_t)g.NY * q);
                // the column's own nodes keep their unmasked value for the
                // prescribed-DOF rule of the epilogue (also on non-owned planes:
                // only planes of the chunk are ever read back)
                const bool own = li >= 1 && li <= kZO && lj >= 1 && lj <= kZO;
                const int o = (li - 1) + kZO * (lj - 1);
                T xp[3] = {R.va[l][0], R.va[l][1], R.va[l][2]};
                if (MG == 2 && prol) {
                    // x' = x + P xc: coarse node I = gi / 2 (weight 1) or the two
                    // neighbours (gi -+ 1) / 2 (1/2 each), per axis, ascending corner
                    const int gi = i0 - 1 + li, gj = j0 - 1 + lj;
                    const int I0 = gi >> 1, J0 = gj >> 1, K0 = q >> 1;
                    const int oi = gi & 1, oj = gj & 1, ok = q & 1;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int a = c & 1, bb = (c >> 1) & 1, cc = (c >> 2) & 1;
                        if ((a && !oi) || (bb && !oj) || (cc && !ok)) continue;
                        const T wt = (oi ? (T)0.5 : (T)1) * (oj ? (T)0.5 : (T)1) *
                                     (ok ? (T)0.5 : (T)1);
                        const int p = (I0 + a - CI0) + CNN * (J0 + bb - CJ0);
                        const int cs = (K0 + cc) & 1;
#pragma unroll
                        for (int d = 0; d < 3; ++d) xp[d] += wt * xcs[MG == 2 ? cs : 0][d][MG == 2 ? p : 0];
                    }
                }
                T pv[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T v = PCG ? R.vb[l][d] * R.va[l][d] + beta * R.vc[l][d] : xp[d];
                    if (MG == 1) v = (T)mg.w * R.vb[l][d] * R.va[l][d];  // w dinv b
                    if (m == 0xff) v = 0;
                    pv[d] = v;
                    const T vm = ((m >> d) & 1) ? (T)0 : v;
                    xs[slot][d][s] = vm;
                    if (!MG && own) xo[MG ? 0 : slot][d][MG ? 0 : o] = v;
                    if (MG == 1 && own) bo_s[MG ? slot : 0][d][MG ? o : 0] = R.va[l][d];
                    if (MG == 2 && own) {
                        bo_s[MG ? slot : 0][d][MG ? o : 0] = R.vb[l][d];
                        do_s[MG == 2 ? slot : 0][d][MG == 2 ? o : 0] = R.vc[l][d];
                    }
                }
                if (PCG && owned) st_node<3>(pnew, n, pv);
                ms[slot][s] = m;
            }
        }
    };

    Stage A, B;  // B is used with a prefetch depth of 2
    // The first plane's loads do not depend on beta: issued before the
    // previous kernel's partials are reduced, so the two latencies overlap.
    if (PRE) load_plane(A, kb - 1);
    double part = 0.0, pz = 0.0, pb = 0.0;
    if (PCG) {
        const double rz = reduce_partials(partB, nbB, 2, 0, red);
        const double rr = reduce_partials(partB, nbB, 2, 1, red);
        const long long k = it + 1;
        const double tol_in = sc->tol;
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
    if (!PRE) load_plane(A, kb - 1);
    if (MG == 2 && prol) {
        // the coarse planes of fine planes kb - 1 and kb, before their staging;
        // then the plane that the first odd step writes
        T c0[3], c1[3];
        load_coarse((kb - 1) >> 1, c0);
        load_coarse((kb + 1) >> 1, c1);
        write_coarse((kb - 1) >> 1, c0);
        write_coarse((kb + 1) >> 1, c1);
        __syncthreads();
        load_coarse((((kb - 1) | 1) + 3) >> 1, cr);
    }
    stage_plane(A, kb - 1, false);  // the plane below the chunk: halo only
    // stage S holds node plane k + 1 and cell layer k when step k begins
    load_plane(A, kb);
    load_cell(A, kb - 1);
    if (PD == 2) {
        load_plane(B, kb + 1);
        load_cell(B, kb);
    }

    const bool node_thread = ci >= 1 && cj >= 1;
    const int oi = ci - 1, oj = cj - 1;  // output node of this thread
    const int gi = i0 + oi, gj = j0 + oj;
    const bool node_in = node_thread && gi < g.NX && gj < g.NY;
    T carry[3] = {0, 0, 0};

    auto step = [&](int k, Stage& S) {
        // plane k + 1 into LDS (owned when inside the chunk); S then takes the
        // plane and cell layer PD steps ahead
        stage_plane(S, k + 1, k + 1 < ke);
        const T Ee = S.ev;
        if (k + 1 + PD <= ke) load_plane(S, k + 1 + PD);
        if (k + PD < ke) load_cell(S, k + PD);
        __syncthreads();
        if (MG == 2 && prol && (k & 1)) {
            // odd step: coarse plane (k + 3) / 2 into the ring (read from step
            // k + 1's staging on), then the registers take the next odd step's
            write_coarse((k + 3) >> 1, cr);
            load_coarse((k + 5) >> 1, cr);
        }

        // ---- this thread's cell of layer k: ye = E_e Ke x_e ----
        const int s0 = ci + kZS * cj;  // staged index of the cell's (0, 0) corner
        T ye[3][8];
        if (Ee != (T)0) {
            T xe[3][8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int s = s0 + (c & 1) + kZS * ((c >> 1) & 1);
                const int slot = (k + (c >> 2)) & 1;
#pragma unroll
                for (int d = 0; d < 3; ++d) xe[d][c] = xs[slot][d][s];
            }
            if constexpr (CUBE) {
                elem_apply_cube_scaled<T>(op, Ee, xe, ye);
            } else {
                elem_apply<T, 3, false>(op, xe, ye);
#pragma unroll
                for (int c = 0; c < 8; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) ye[d][c] *= Ee;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) ye[d][c] = 0;
        }
        // ---- pairs in x: left cell's x = 1 corner + own x = 0 corner ----
        // P[d][by][bz] belongs to node (ci - 1, cj - 1 + by) at plane k + bz
        T P[3][2][2];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int by = 0; by < 2; ++by)
#pragma unroll
                for (int bz = 0; bz < 2; ++bz) {
                    const int c = 2 * by + 4 * bz;
                    P[d][by][bz] = row_shr1(ye[d][c + 1]) + ye[d][c];
                }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xch[2 * d][tid] = P[d][1][0];
            xch[2 * d + 1][tid] = P[d][1][1];
        }
        // the epilogue's values of plane k, read before the barrier: the next
        // step stages plane k + 2 into this ring slot without waiting for
        // the waves still in this step's epilogue
        uint8_t fm = 0;
        T xv[3] = {0, 0, 0}, bv[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
        if (node_thread) {
            const int slot = k & 1;
            fm = ms[slot][ci + kZS * cj];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xv[d] = MG ? xs[slot][d][ci + kZS * cj] : xo[MG ? 0 : slot][d][MG ? 0 : oi + kZO * oj];
                if (MG) bv[d] = bo_s[MG ? slot : 0][d][MG ? oi + kZO * oj : 0];
                if (MG == 2) dv[d] = do_s[MG == 2 ? slot : 0][d][MG == 2 ? oi + kZO * oj : 0];
            }
        }
        __syncthreads();

        // ---- node (oi, oj): row below's pair, then this row's pair ----
        if (node_thread) {
            const int tb = tid - kZC;  // the cell below: (ci, cj - 1)
            T z0[3], z1[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                z0[d] = xch[2 * d][tb] + P[d][0][0];
                z1[d] = xch[2 * d + 1][tb] + P[d][0][1];
            }
            if (k >= kb && node_in) {
                const uint32_t n = (uint32_t)gi + (uint32_t)g.NX * ((uint32_t)gj + (uint32_t)g.NY * k);
                T o1[3], o2[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T v = carry[d] + z0[d];
                    const bool fixd = (fm >> d) & 1;
                    if (MG == 0) {
                        if (fixd) v = (T)meandiag * xv[d];
                        o1[d] = v;
                        if (PCG) part += (double)xv[d] * (double)v;
                    } else if (MG == 1) {
                        o1[d] = xv[d];
                        o2[d] = fixd ? (T)0 : bv[d] - v;
                    } else {
                        const T z = fixd ? (T)0 : xv[d] + (T)mg.w * dv[d] * (bv[d] - v);
                        o1[d] = z;
                        pz += (double)bv[d] * (double)z;
                        pb += (double)bv[d] * (double)bv[d];
                    }
                }
                if (MG == 0) st_node<3>(y, n, o1);
                if (MG != 0) st_node<3>((T*)mg.xout, n, o1);
                if (MG == 1) st_node<3>((T*)mg.rout, n, o2);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) carry[d] = z1[d];
        }
    };
    if (PD == 2) {
        for (int k = kb - 1; k < ke; k += 2) {
            step(k, A);
            if (k + 1 < ke) step(k + 1, B);
        }
    } else {
        for (int k = kb - 1; k < ke; ++k) step(k, A);
    }
    if (PCG) {
        const double bs = block_sum_all(part, red);
        if (tid == 0) partA[blockIdx.x] = bs;
    }
    if (MG == 2 && mg.part) {
        const double s1 = block_sum_all(pz, red);
        const double s2 = block_sum_all(pb, red);
        if (tid == 0) {
            mg.part[2 * (size_t)blockIdx.x] = s1;
            mg.part[2 * (size_t)blockIdx.x + 1] = s2;
        }
    }
}

// Prefetch depth of the node planes: 1 (TOHIP_ZM_PD = 2 selects two
// alternating register stages, for measurement).  Measured on MI355X at
// 256x128x128, depth 2 is no faster for the plain products (FP64 115 vs 116
// us) and slower for the PCG kernels (FP64 252 vs 224 us: 184 VGPRs halve
// the occupancy), at every workgroup target from 1024 to 2560.
template <typename T, bool PCG>
static int zm_depth() {
    static int d = [] {
        const char* e = getenv("TOHIP_ZM_PD");
        const int v = e ? atoi(e) : 1;
        return v == 2 ? 2 : 1;
    }();
    return d;
}

template <typename T, bool PCG>
void launch_matvec_zm(const Grid& g, const ElemOp<T>& op, const T* xin, const T* dinv, T* p0,
                      T* p1, const uint8_t* fixmask, const T* E, T* y, double meandiag,
                      PcgScalars* sc, const double* partB, int nbB, double* partA, bool pre,
                      hipStream_t s) {
    const ZmShape sh = zm_shape(g);
    const int pd = zm_depth<T, PCG>();
#define TOHIP_ZM(CB, PR, PDV)                                                                    \
    hipLaunchKernelGGL((k_matvec_zm<T, PCG, CB, PR, PDV>), dim3(sh.blocks()), dim3(kZT), 0, s, g, \
                       op, sh.zc, xin, dinv, p0, p1, fixmask, E, y, meandiag, sc, partB, nbB,   \
                       partA, ZmMgArgs{})
#define TOHIP_ZM2(CB, PR) \
    if (pd == 2)          \
        TOHIP_ZM(CB, PR, 2); \
    else                  \
        TOHIP_ZM(CB, PR, 1)
    if (op.is_cube) {
        if (pre) TOHIP_ZM2(true, PCG); else TOHIP_ZM2(true, false);
    } else {
        if (pre) TOHIP_ZM2(false, PCG); else TOHIP_ZM2(false, false);
    }
#undef TOHIP_ZM2
#undef TOHIP_ZM
}

// Smallest grid (cells) that the z-marching kernels take over from the tiles:
// 2^20 (TOHIP_ZM_MINCELLS overrides it for measurement).
int64_t zm_min_cells() {
    static int64_t v = [] {
        const char* e = getenv("TOHIP_ZM_MINCELLS");
        const long long c = e ? atoll(e) : 0;
        return c > 0 ? (int64_t)c : ((int64_t)1 << 20);
    }();
    return v;
}

// Multigrid smoothing of a 3D FP32 level by z-marching columns (k_mg.hip
// routes the finest level here from zm_min_cells() up; TOHIP_MG_ZM = 0 / 2:
// never / every 3D level, for measurement and testing).
bool mg_use_zm(const Grid& g) {
    static int mode = [] {
        const char* e = getenv("TOHIP_MG_ZM");
        return e ? atoi(e) : 1;
    }();
    if (mode == 0 || g.dim != 3 || !zm_fits(g)) return false;
    return mode == 2 || g.ncells >= zm_min_cells();
}

// mode 1: pre-smoothing on b; mode 2: post-smoothing on xin = x' (the
// prolongation already added), with partials into `part` when it is given.
hipError_t launch_mg_smooth_zm(int mode, const Grid& g, const ElemOp<float>& op, const float* E,
                               const uint8_t* fixmask, const float* dinv, float w,
                               const float* wdev, const float* xin, const float* b, float* xout,
                               float* rout, double meandiag, const PcgScalars* sc, double* part,
                               hipStream_t s, const Grid* gc, const float* xc) {
    const ZmShape sh = zm_shape(g);
    ZmMgArgs a{w, wdev, b, xout, rout, part, sc, nullptr, Grid{}};
    if (mode == 2 && xc && gc) {
        a.xc = xc;
        a.gc = *gc;
    }
    const float* in = mode == 1 ? b : xin;
#define TOHIP_ZMG(CB, M)                                                                      \
    hipLaunchKernelGGL((k_matvec_zm<float, false, CB, false, 1, M>), dim3(sh.blocks()),       \
                       dim3(kZT), 0, s, g, op, sh.zc, in, dinv, nullptr, nullptr, fixmask, E, \
                       nullptr, meandiag, nullptr, nullptr, 0, nullptr, a)
    if (op.is_cube) {
        if (mode == 1) TOHIP_ZMG(true, 1); else TOHIP_ZMG(true, 2);
    } else {
        if (mode == 1) TOHIP_ZMG(false, 1); else TOHIP_ZMG(false, 2);
    }
#undef TOHIP_ZMG
    return hipGetLastError();
}

#define TOHIP_ZM_INST(T, PCG)                                                                    \
    template void launch_matvec_zm<T, PCG>(const Grid&, const ElemOp<T>&, const T*, const T*,   \
                                           T*, T*, const uint8_t*, const T*, T*, double,        \
                                           PcgScalars*, const double*, int, double*, bool,      \
                                           hipStream_t);
TOHIP_ZM_INST(double, false)
TOHIP_ZM_INST(double, true)
TOHIP_ZM_INST(float, false)
TOHIP_ZM_INST(float, true)
#undef TOHIP_ZM_INST

}  // namespace tohip
