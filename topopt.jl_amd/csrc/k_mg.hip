// Geometric multigrid V-cycle in FP32, the preconditioner of the
// mixed-precision PCG when it is asked for (TOHIP_MIXED_MG).  TopOpt.jl has no
// multigrid: its solver is Jacobi PCG (src/FEA/solvers_api.jl:177-220).  The
// outer FP64 residual test is unchanged, so only the iteration count changes.
//
// Hierarchy: level l + 1 halves every cell count of level l (cell size
// doubles).  Its operator is rediscretised, not Galerkin: the same Q1 element
// operator on the 2h box, scaled per coarse cell by the mean of the moduli of
// its 2^dim child cells.  A coarse DOF is prescribed when any fine node of its
// prolongation stencil has that DOF prescribed.  Transfers are the trilinear
// (bilinear in 2D) prolongation P and restriction P^T.
//
// V(1,1) cycle, damped Jacobi (weight w) both sides, zero initial guess:
//   x_l = w D^-1 b_l;  r_l = b_l - K_l x_l;  b_{l+1} = P^T r_l
//   ... coarsest: a fixed number of damped Jacobi sweeps from zero ...
//   x'_l = x_l + P x_{l+1};  z_l = x'_l + w D^-1 (b_l - K_l x'_l)
// Pre- and post-smoothing are adjoint and the coarsest sweeps are a fixed
// polynomial in D^-1 K, so the preconditioner is a fixed symmetric operator
// (plain PCG stays valid).
//
// Kernels:
//   k_mg_smooth<MODE>: the tile scheme of k_matvec (LDS-staged nodes with a
//     one-node halo, one thread per cell, per-node gather in ascending cell
//     order) with the smoothing arithmetic fused into staging and epilogue:
//     MODE 0 (pre)   stage x = w dinv b, write x and r = b - K x;
//     MODE 1 (post)  stage x' = x + P xc, write z = x' + w dinv (b - K x'),
//                    and at the finest level the partials (b.z, b.b) that the
//                    PCG direction kernel reads;
//     MODE 2 (sweep) stage x, write x + w dinv (b - K x) (coarsest fallback);
//     MODE 5 (power step) stage v = s xprev (s = aux[0]), write
//                    u = (I - D^-1 K) v and the partials (v'Du, v'Dv, u'Du)
//                    of the lambda_max(D^-1 K) estimate below.
//   The weight w is read from device memory (wdev) when it is given, so that
//   the captured iteration graphs follow the weight the estimate sets.
//   k_mg_restrict: b_{l+1} = P^T r_l, one thread per coarse node.
//   k_mg_coarse:   all coarsest sweeps in one workgroup, everything in LDS.
//   k_mg_coarsen:  coarse moduli.
// Every kernel returns at once when the PCG state says the solve is done, so
// the iteration graphs replay as no-ops after convergence.
#include <algorithm>
#include <cstdlib>

#include "k_common.h"

namespace tohip {

namespace {

constexpr int kMgBlocksCap = 1024;  // persistent grid cap (partials of the finest post kernel)

// Tile shapes of k_mg_smooth (output nodes per workgroup, threads); 3D
// variants selectable with TOHIP_MGTILE for measurement.  16x4x2 with 256
// threads was the fastest on MI355X at 120x40x40 (finest-level pre 10.4 us,
// post 15.2 us vs 11.9 / 17.7 us for 8x4x4 with 128 threads).
struct MgTileShape {
    int tx, ty, tz, nt;
};
constexpr MgTileShape kMgTiles3[] = {{8, 4, 4, 128}, {16, 4, 2, 256}, {8, 8, 2, 128}};
constexpr MgTileShape kMgTile2 = {16, 8, 1, 128};
static int mg_tile_variant() {
    static int v = [] {
        const char* e = getenv("TOHIP_MGTILE");
        const int t = e ? atoi(e) : 1;
        return (t >= 0 && t < 3) ? t : 1;
    }();
    return v;
}
static MgTileShape mg_tile(int dim) { return dim == 3 ? kMgTiles3[mg_tile_variant()] : kMgTile2; }

__device__ __forceinline__ bool mg_done(const PcgScalars* sc) { return sc && sc->done; }

}  // namespace

template <int DIM, int MODE, bool CUBE, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT, MODE == 3 ? 4 : 1) void k_mg_smooth(
    Grid g, ElemOp<float> op, const float* __restrict__ E, const uint8_t* __restrict__ fixmask,
    const float* __restrict__ dinv, float w, const float* __restrict__ b,
    const float* __restrict__ xprev, Grid gc, const float* __restrict__ xc,
    float* __restrict__ xout, float* __restrict__ rout, double meandiag, const PcgScalars* sc,
    double* __restrict__ part, MgFusedUpdate upd, const float* __restrict__ wdev,
    const float* __restrict__ aux) {
    constexpr int NC = 1 << DIM;
    constexpr bool UPD = MODE == 3;  // MODE 0 after the fused PCG update
    constexpr bool PW = MODE == 5;   // power step of the weight estimate
    constexpr int IX = TX + 2, IY = TY + 2, IZ = DIM == 3 ? TZ + 2 : 1;
    constexpr int EX = TX + 1, EY = TY + 1, EZ = DIM == 3 ? TZ + 1 : 1;
    constexpr int NIN = IX * IY * IZ, NOUT = TX * TY * TZ, NEL = EX * EY * EZ;
    // coarse nodes whose prolongation reaches the tile + halo (tile origins are
    // even, so they start one coarse node before origin / 2)
    constexpr int CX = TX / 2 + 3, CY = TY / 2 + 3, CZ = DIM == 3 ? TZ / 2 + 3 : 1;
    constexpr int CN = MODE == 1 ? CX * CY * CZ : 1;
    static_assert(TX % 2 == 0 && TY % 2 == 0 && (DIM == 2 || TZ % 2 == 0), "even tiles");

    __shared__ float xs[DIM][NIN];
    __shared__ float res[DIM * NC][NEL];
    __shared__ float pown[DIM][NOUT];
    __shared__ float eloc[NEL];
    __shared__ float xcs[DIM][CN];
    __shared__ double red[NT / 64];

    if (mg_done(sc)) return;
    if (wdev) w = *wdev;
    const float ps = PW ? aux[0] : 1.f;
    const int tid = threadIdx.x;
    const int ntx = (g.NX + TX - 1) / TX, nty = (g.NY + TY - 1) / TY;
    const int ntz = DIM == 3 ? (g.NZ + TZ - 1) / TZ : 1;
    const int nt = ntx * nty * ntz;
    double pz = 0.0, pb = 0.0, pu = 0.0;
    // fused update: the direction p of the iteration in flight, and alpha
    // (reduced once, after the first tile's loads are issued)
    const long long uk = UPD ? upd.sc->k_cur : 0;
    const float* __restrict__ up = UPD ? ((uk & 1) ? upd.p1 : upd.p0) : nullptr;
    double ualpha = 0.0;
    float ua = 0.f;
    bool ufirst = true;

    for (int tb = blockIdx.x; tb < nt; tb += gridDim.x) {
        const int t = xcd_remap(tb, nt);
        const int i0 = (t % ntx) * TX, j0 = ((t / ntx) % nty) * TY;
        const int k0 = DIM == 3 ? (t / (ntx * nty)) * TZ : 0;
        // ---- every global load of the tile first: cell moduli, coarse
        //      correction, iterate / rhs on tile + halo and the epilogue
        //      operands, at clamped (in-bounds) indices; values outside the
        //      mesh are selected away when LDS is written.  Surplus slots of
        //      the last pass reload the tile's own last entry, so they add no
        //      traffic beyond the tile.  The few-workgroup
        //      coarse levels are latency-bound, so one round trip per tile
        //      instead of one per staging phase. ----
        constexpr int SE = (NEL + NT - 1) / NT, SI = (NIN + NT - 1) / NT;
        constexpr int SC = MODE == 1 ? (CN + NT - 1) / NT : 1, SO = (NOUT + NT - 1) / NT;
        const int cI0 = i0 / 2 - 1, cJ0 = j0 / 2 - 1, cK0 = DIM == 3 ? k0 / 2 - 1 : 0;
        float er[SE];
#pragma unroll
        for (int q = 0; q < SE; ++q) {
            const int el = min(tid + q * NT, NEL - 1);
            const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
            const int gi = min(max(i0 - 1 + ei, 0), g.nx - 1), gj = min(max(j0 - 1 + ej, 0), g.ny - 1);
            const int gk = DIM == 3 ? min(max(k0 - 1 + ek, 0), g.nz - 1) : 0;
            er[q] = E[gi + g.nx * (gj + g.ny * gk)];
        }
        float cr[SC][DIM];
        if (MODE == 1) {
#pragma unroll
            for (int q = 0; q < SC; ++q) {
                const int p = min(tid + q * NT, CN - 1);
                const int I = min(max(cI0 + p % CX, 0), gc.NX - 1);
                const int J = min(max(cJ0 + (p / CX) % CY, 0), gc.NY - 1);
                const int K = DIM == 3 ? min(max(cK0 + p / (CX * CY), 0), gc.NZ - 1) : 0;
                const int N = I + gc.NX * (J + gc.NY * K);
#pragma unroll
                for (int d = 0; d < DIM; ++d) cr[q][d] = xc[DIM * N + d];
            }
        }
        uint8_t mr[SI];
        // MODE 0 / 3: dinv, b (and q); else: xprev (ra)
        float ra[SI][DIM], rb[SI][DIM], rq[UPD ? SI : 1][DIM];
#pragma unroll
        for (int q = 0; q < SI; ++q) {
            const int nl = min(tid + q * NT, NIN - 1);
            const int li = nl % IX, lj = (nl / IX) % IY, lk = DIM == 3 ? nl / (IX * IY) : 0;
            const int gi = min(max(i0 - 1 + li, 0), g.NX - 1), gj = min(max(j0 - 1 + lj, 0), g.NY - 1);
            const int gk = DIM == 3 ? min(max(k0 - 1 + lk, 0), g.NZ - 1) : 0;
            const int n = gi + g.NX * (gj + g.NY * gk);
            mr[q] = fixmask[n];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                if (MODE == 0 || UPD) {
                    ra[q][d] = dinv[DIM * n + d];
                    rb[q][d] = b[DIM * n + d];
                    if (UPD) rq[UPD ? q : 0][d] = upd.q[DIM * n + d];
                } else {
                    ra[q][d] = xprev[DIM * n + d];
                }
            }
        }
        uint8_t fo[SO];
        float bo[SO][DIM], dvo[SO][DIM];
        constexpr int SU = UPD ? SO : 1;
        float qo[SU][DIM], xo[SU][DIM], po[SU][DIM];  // fused update: q, x, p
#pragma unroll
        for (int q = 0; q < SO; ++q) {
            const int o = min(tid + q * NT, NOUT - 1);
            const int oi = o % TX, oj = (o / TX) % TY, ok = DIM == 3 ? o / (TX * TY) : 0;
            const int gi = min(i0 + oi, g.NX - 1), gj = min(j0 + oj, g.NY - 1);
            const int gk = DIM == 3 ? min(k0 + ok, g.NZ - 1) : 0;
            const int n = gi + g.NX * (gj + g.NY * gk);
            fo[q] = fixmask[n];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                bo[q][d] = PW ? 0.f : b[DIM * n + d];
                dvo[q][d] = (MODE == 0 || UPD) ? 0.f : dinv[DIM * n + d];
                if (UPD) {
                    qo[UPD ? q : 0][d] = upd.q[DIM * n + d];
                    xo[UPD ? q : 0][d] = upd.x[DIM * n + d];
                    po[UPD ? q : 0][d] = up[DIM * n + d];
                }
            }
        }
        // alpha = rho / p.q once per workgroup, while the loads above are in
        // flight; a p.q that is not positive and finite stops the solve before
        // x or r change (every workgroup reduces the same partials)
        if (UPD && ufirst) {
            ufirst = false;
            const double pq = reduce_partials(upd.partA, upd.nbA, 1, 0, red);
            if (!(pq > 0.0 && pq < INFINITY)) {
                if (blockIdx.x == 0 && tid == 0) {
                    upd.sc->err = kPcgErrIndefinite;
                    upd.sc->done = 1;
                }
                return;
            }
            ualpha = upd.sc->rho_cur / pq;
            ua = (float)ualpha;
        }
        if (UPD) {
            // the updated residual, the same FMA on every node whichever
            // workgroup forms it (owner or neighbour's halo), and x += alpha p
            // on the owned nodes, done here so that q, x and p leave the
            // registers before the cell products
#pragma unroll
            for (int q = 0; q < SI; ++q)
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                    rb[q][d] = __fmaf_rn(-ua, rq[UPD ? q : 0][d], rb[q][d]);
#pragma unroll
            for (int q = 0; q < SO; ++q) {
                const int o = tid + q * NT;
                const int oi = o % TX, oj = (o / TX) % TY, ok = DIM == 3 ? o / (TX * TY) : 0;
                const int gi = i0 + oi, gj = j0 + oj, gk = k0 + ok;
                const bool in = o < NOUT && gi < g.NX && gj < g.NY && gk < g.NZ;
                const int n = gi + g.NX * (gj + g.NY * gk);
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    bo[q][d] = __fmaf_rn(-ua, qo[UPD ? q : 0][d], bo[q][d]);
                    if (in) upd.x[DIM * n + d] = __fmaf_rn(ua, po[UPD ? q : 0][d], xo[UPD ? q : 0][d]);
                }
            }
        }
        // ---- cell moduli of the neighbourhood (0 outside the mesh) ----
#pragma unroll
        for (int q = 0; q < SE; ++q) {
            const int el = tid + q * NT;
            if (el >= NEL) break;
            const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
            const int gi = i0 - 1 + ei, gj = j0 - 1 + ej, gk = DIM == 3 ? k0 - 1 + ek : 0;
            const bool in = gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz;
            eloc[el] = in ? er[q] : 0.f;
        }
        // ---- coarse correction of the neighbourhood into LDS (post-smoothing) ----
        if (MODE == 1) {
#pragma unroll
            for (int q = 0; q < SC; ++q) {
                const int p = tid + q * NT;
                if (p >= CN) break;
                const int I = cI0 + p % CX, J = cJ0 + (p / CX) % CY,
                          K = DIM == 3 ? cK0 + p / (CX * CY) : 0;
                const bool inr = I >= 0 && I < gc.NX && J >= 0 && J < gc.NY && K >= 0 && K < gc.NZ;
#pragma unroll
                for (int d = 0; d < DIM; ++d) xcs[d][p] = inr ? cr[q][d] : 0.f;
            }
            __syncthreads();
        }
        // ---- stage the iterate on tile + halo (prescribed DOFs and outside: 0) ----
#pragma unroll
        for (int q = 0; q < SI; ++q) {
            const int nl = tid + q * NT;
            if (nl >= NIN) break;
            const int li = nl % IX, lj = (nl / IX) % IY, lk = DIM == 3 ? nl / (IX * IY) : 0;
            const int gi = i0 - 1 + li, gj = j0 - 1 + lj, gk = DIM == 3 ? k0 - 1 + lk : 0;
            const bool in = gi >= 0 && gi < g.NX && gj >= 0 && gj < g.NY && gk >= 0 && gk < g.NZ;
            float v[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) v[d] = 0.f;
            if (in) {
                const uint8_t m = mr[q];
                if (MODE == 0) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) v[d] = w * ra[q][d] * rb[q][d];
                } else if (UPD) {  // rb holds r - alpha q
#pragma unroll
                    for (int d = 0; d < DIM; ++d) v[d] = w * ra[q][d] * rb[q][d];
                } else {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) v[d] = PW ? ps * ra[q][d] : ra[q][d];
                    if (MODE == 1) {
                        // + P xc: coarse node I = gi/2 (weight 1) or the two
                        // neighbours (gi -+ 1)/2 (weight 1/2 each), per axis
                        const int I0 = gi >> 1, J0 = gj >> 1, K0 = gk >> 1;
                        const int oi = gi & 1, oj = gj & 1, ok = DIM == 3 ? (gk & 1) : 0;
                        // every lane reads all 2^DIM corners (uniform, no divergent
                        // branch) with weight 0 on those off its stencil: the same
                        // terms in the same ascending order, plus zeros
                        const float wx[2] = {oi ? 0.5f : 1.f, oi ? 0.5f : 0.f};
                        const float wy[2] = {oj ? 0.5f : 1.f, oj ? 0.5f : 0.f};
                        const float wz[2] = {ok ? 0.5f : 1.f, ok ? 0.5f : 0.f};
#pragma unroll
                        for (int c = 0; c < NC; ++c) {
                            const int a = c & 1, bb = (c >> 1) & 1, cc = (c >> 2) & 1;
                            const float wt = wx[a] * wy[bb] * (DIM == 3 ? wz[cc] : 1.f);
                            const int p = (I0 + a - cI0) + CX * ((J0 + bb - cJ0) + CY * (K0 + cc - cK0));
#pragma unroll
                            for (int d = 0; d < DIM; ++d) v[d] += wt * xcs[d][p];
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                    if ((m >> d) & 1) v[d] = 0.f;
            }
            const bool owned = in && li >= 1 && li <= TX && lj >= 1 && lj <= TY &&
                               (DIM == 2 || (lk >= 1 && lk <= TZ));
            const int o = (li - 1) + TX * ((lj - 1) + TY * (DIM == 3 ? lk - 1 : 0));
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xs[d][nl] = v[d];
                if (owned) pown[d][o] = v[d];
            }
        }
        __syncthreads();
        // ---- one thread per cell: E_e Ke x_e into the slots ----
        for (int el = tid; el < NEL; el += NT) {
            const int ei = el % EX, ej = (el / EX) % EY, ek = DIM == 3 ? el / (EX * EY) : 0;
            const float Ee = eloc[el];
            if (Ee != 0.f) {
                float xe[DIM][NC], ye[DIM][NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const int nl = (ei + (c & 1)) +
                                   IX * ((ej + ((c >> 1) & 1)) + IY * (ek + ((c >> 2) & 1)));
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xe[d][c] = xs[d][nl];
                }
                elem_apply<float, DIM, CUBE>(op, xe, ye);
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) res[c * DIM + d][el] = Ee * ye[d][c];
            } else {
#pragma unroll
                for (int j = 0; j < DIM * NC; ++j) res[j][el] = 0.f;
            }
        }
        __syncthreads();
        // ---- per owned node: K x in ascending cell order, then the epilogue ----
#pragma unroll
        for (int q = 0; q < SO; ++q) {
            const int o = tid + q * NT;
            if (o >= NOUT) break;
            const int oi = o % TX, oj = (o / TX) % TY, ok = DIM == 3 ? o / (TX * TY) : 0;
            const int gi = i0 + oi, gj = j0 + oj, gk = k0 + ok;
            if (gi >= g.NX || gj >= g.NY || gk >= g.NZ) continue;
            const int n = gi + g.NX * (gj + g.NY * gk);
            const uint8_t fm = fo[q];
            float s[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) s[d] = 0.f;
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
                const int i = DIM * n + d;
                const bool fixd = (fm >> d) & 1;
                const float x = pown[d][o];
                const float bi = bo[q][d];  // fused update: r - alpha q
                const float y = fixd ? (float)meandiag * x : s[d];
                if (MODE == 0 || UPD) {
                    xout[i] = x;
                    rout[i] = fixd ? 0.f : bi - y;
                    if (UPD) upd.rnew[i] = bi;
                } else if (PW) {
                    // u = v - D^-1 K v; D-weighted partials (prescribed: v = u = 0)
                    const float u = fixd ? 0.f : x - dvo[q][d] * y;
                    xout[i] = u;
                    if (!fixd) {
                        const double dd = 1.0 / (double)dvo[q][d];
                        pz += dd * (double)x * (double)u;
                        pb += dd * (double)x * (double)x;
                        pu += dd * (double)u * (double)u;
                    }
                } else {
                    const float z = fixd ? 0.f : x + w * dvo[q][d] * (bi - y);
                    xout[i] = z;
                    if (MODE == 1 && part) {
                        pz += (double)bi * (double)z;
                        pb += (double)bi * (double)bi;
                    }
                }
            }
        }
        __syncthreads();  // LDS is reused by the next tile
    }
    if (UPD && blockIdx.x == 0 && tid == 0 && !ufirst) {
        upd.sc->iter = uk;
        upd.sc->alpha = ualpha;
    }
    if (MODE == 1 && part) {
        const double s1 = block_sum_all(pz, red);
        const double s2 = block_sum_all(pb, red);
        if (tid == 0) {
            part[2 * (size_t)blockIdx.x] = s1;
            part[2 * (size_t)blockIdx.x + 1] = s2;
        }
    }
    if (PW) {
        const double s1 = block_sum_all(pz, red);
        const double s2 = block_sum_all(pb, red);
        const double s3 = block_sum_all(pu, red);
        if (tid == 0) {
            part[3 * (size_t)blockIdx.x] = s1;
            part[3 * (size_t)blockIdx.x + 1] = s2;
            part[3 * (size_t)blockIdx.x + 2] = s3;
        }
    }
}

// b_c = P^T r (masked at prescribed coarse DOFs): coarse node (I, J, K) sums
// the fine nodes (2I + a, 2J + b, 2K + c), a, b, c in {-1, 0, 1}, with weight
// prod (1 - |a|/2), in ascending (c, b, a) order.  The stencil is fully
// unrolled with clamped indices and out-of-grid terms selected to zero, so all
// 3^DIM * DIM loads issue before the first add (the coarse levels are a few
// workgroups, latency-bound: a guarded loop serialises one load per term).
template <int DIM>
__global__ __launch_bounds__(256) void k_mg_restrict(Grid g, Grid gc,
                                                     const uint8_t* __restrict__ cmask,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ bc,
                                                     const PcgScalars* sc) {
    if (mg_done(sc)) return;
    const int N = blockIdx.x * 256 + threadIdx.x;
    if (N >= gc.nnodes) return;
    const int I = N % gc.NX, J = (N / gc.NX) % gc.NY, K = DIM == 3 ? N / (gc.NX * gc.NY) : 0;
    constexpr int NC = DIM == 3 ? 3 : 1;
    float v[NC][3][3][DIM];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int k = 2 * K + c - (DIM == 3 ? 1 : 0);
        const bool kin = k >= 0 && k < g.NZ;
        const int kk = min(max(k, 0), g.NZ - 1);
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            const int j = 2 * J + bb - 1;
            const bool jin = kin && j >= 0 && j < g.NY;
            const int jj = min(max(j, 0), g.NY - 1);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int i = 2 * I + a - 1;
                const bool in = jin && i >= 0 && i < g.NX;
                const int n = min(max(i, 0), g.NX - 1) + g.NX * (jj + g.NY * kk);
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const float x = r[DIM * n + d];
                    v[c][bb][a][d] = in ? x : 0.f;
                }
            }
        }
    }
    float s[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) s[d] = 0.f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const float wc = (DIM == 3 && c != 1) ? 0.5f : 1.f;
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            const float wb = wc * (bb != 1 ? 0.5f : 1.f);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float wt = wb * (a != 1 ? 0.5f : 1.f);
#pragma unroll
                for (int d = 0; d < DIM; ++d) s[d] += wt * v[c][bb][a][d];
            }
        }
    }
    const uint8_t m = cmask[N];
#pragma unroll
    for (int d = 0; d < DIM; ++d) bc[DIM * N + d] = ((m >> d) & 1) ? 0.f : s[d];
}

// Coarsest level: a Chebyshev iteration on D^-1 K from zero, in one workgroup
// with the whole level in LDS (sizes checked by the host against the limits).
// With fixed coefficients the result is a fixed polynomial in D^-1 K applied
// to D^-1 b, positive on (0, lmax]: the coarse solve stays a symmetric
// positive definite operator, so the V-cycle remains a valid PCG
// preconditioner.  Thread t owns cell t and node t (both counts fit the
// workgroup), so the cell's corner base index, the node's adjacent-cell mask,
// its b, D^-1 and previous iterate stay in registers across the iterations.
// BASIS: workgroup j applies the same operator to the unit vector e_j and
// writes column j of its matrix (xout + j * ndof), for the dense form below.
constexpr int kMgCoarseThreads = 1024;
static_assert(kMgCoarseMaxNodes <= kMgCoarseThreads && kMgCoarseMaxCells <= kMgCoarseThreads,
              "one node and one cell per thread");
// Launched with just enough whole waves for one node and one cell per thread
// (576 at 15x5x5): idle waves of a 1024-thread block would hold wave slots
// and limit the basis build's workgroups per CU.
static unsigned mg_coarse_threads(const Grid& g) {
    const int64_t n = g.nnodes > g.ncells ? g.nnodes : g.ncells;
    return (unsigned)((n + 63) / 64 * 64);
}
template <int DIM, bool CUBE, bool BASIS = false>
__global__ __launch_bounds__(kMgCoarseThreads) void k_mg_coarse(
    Grid g, ElemOp<float> op, const float* __restrict__ E, const uint8_t* __restrict__ fixmask,
    const float* __restrict__ dinv, MgCheb ch, const float* __restrict__ b,
    float* __restrict__ xout, const PcgScalars* sc) {
    constexpr int NC = 1 << DIM;
    __shared__ float xs[DIM][kMgCoarseMaxNodes];
    __shared__ float res[DIM * NC][kMgCoarseMaxCells];
    if (!BASIS && mg_done(sc)) return;
    if (BASIS) xout += (size_t)blockIdx.x * (size_t)g.ndof;
    const int t = threadIdx.x;
    if (BASIS && ((fixmask[blockIdx.x / DIM] >> (blockIdx.x % DIM)) & 1)) {
        // a prescribed DOF's right-hand side is 0, so its column is 0 (the
        // same +0 bits the iteration would leave)
        for (int i = t; i < (int)g.ndof; i += blockDim.x) xout[i] = 0.f;
        return;
    }
    const int nn = (int)g.nnodes, ne = (int)g.ncells;
    const bool has_node = t < nn, has_cell = t < ne;
    // node t: right-hand side, inverse diagonal (0 when prescribed), previous
    // iterate, and which of its 2^DIM adjacent cells exist
    float bn[DIM], dn[DIM], xo[DIM];
    int ecell0 = 0;
    unsigned cmask = 0;
    if (has_node) {
        const int gi = t % g.NX, gj = (t / g.NX) % g.NY, gk = DIM == 3 ? t / (g.NX * g.NY) : 0;
        const uint8_t m = fixmask[t];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const bool fx = (m >> d) & 1;
            if (BASIS)
                bn[d] = (!fx && DIM * t + d == (int)blockIdx.x) ? 1.f : 0.f;
            else
                bn[d] = fx ? 0.f : b[DIM * t + d];
            dn[d] = fx ? 0.f : dinv[DIM * t + d];
            xo[d] = 0.f;
            xs[d][t] = ch.c0 * dn[d] * bn[d];
        }
        ecell0 = gi + g.nx * (gj + g.ny * gk);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int ci = gi - (c & 1), cj = gj - ((c >> 1) & 1),
                      ck = DIM == 3 ? gk - ((c >> 2) & 1) : 0;
            if (ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny && ck >= 0 && ck < g.nz)
                cmask |= 1u << c;
        }
    }
    // cell t: modulus and the node index of its corner 0
    float Ee = 0.f;
    int ncorner0 = 0;
    if (has_cell) {
        const int ei = t % g.nx, ej = (t / g.nx) % g.ny, ek = DIM == 3 ? t / (g.nx * g.ny) : 0;
        Ee = E[t];
        ncorner0 = ei + g.NX * (ej + g.NY * ek);
    }
    const int sx = 1, sy = g.NX, sz = g.NX * g.NY;  // corner offsets
    const int cy = g.nx, cz = g.nx * g.ny;          // adjacent-cell offsets
    __syncthreads();
    for (int k = 1; k < ch.iters; ++k) {
        if (has_cell) {
            float xe[DIM][NC], ye[DIM][NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int n = ncorner0 + (c & 1) * sx + ((c >> 1) & 1) * sy + ((c >> 2) & 1) * sz;
#pragma unroll
                for (int d = 0; d < DIM; ++d) xe[d][c] = xs[d][n];
            }
            elem_apply<float, DIM, CUBE>(op, xe, ye);
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < DIM; ++d) res[c * DIM + d][t] = Ee * ye[d][c];
        }
        __syncthreads();
        if (has_node) {
            float s[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) s[d] = 0.f;
#pragma unroll
            for (int c = NC - 1; c >= 0; --c) {
                if (!((cmask >> c) & 1)) continue;
                const int e = ecell0 - (c & 1) - ((c >> 1) & 1) * cy - ((c >> 2) & 1) * cz;
#pragma unroll
                for (int d = 0; d < DIM; ++d) s[d] += res[c * DIM + d][e];
            }
            const float ak = ch.a[k], gk = ch.g[k];
            // prescribed DOFs have dn = 0 and stay 0
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const float x = xs[d][t];
                xs[d][t] = x + ak * (x - xo[d]) + gk * dn[d] * (bn[d] - s[d]);
                xo[d] = x;
            }
        }
        __syncthreads();
    }
    if (has_node)
#pragma unroll
        for (int d = 0; d < DIM; ++d) xout[DIM * t + d] = xs[d][t];
}

// Chebyshev coefficients (three-term recurrence) for [lmax / kappa, lmax].
MgCheb mg_chebyshev(double lmax, double kappa, int iters) {
    MgCheb ch{};
    ch.iters = iters < 1 ? 1 : (iters > kMgChebMax ? kMgChebMax : iters);
    const double lo = lmax / kappa;
    const double theta = 0.5 * (lmax + lo), delta = 0.5 * (lmax - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    ch.c0 = (float)(1.0 / theta);
    for (int k = 1; k < ch.iters; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        ch.a[k] = (float)(rn * rho);
        ch.g[k] = (float)(2.0 * rn / delta);
        rho = rn;
    }
    return ch;
}

// Coarse moduli: mean of the 2^dim child cells.
template <int DIM>
__global__ __launch_bounds__(256) void k_mg_coarsen(Grid gf, Grid gc, const double* __restrict__ Ef,
                                                    double* __restrict__ E64,
                                                    float* __restrict__ E32) {
    const int C = blockIdx.x * 256 + threadIdx.x;
    if (C >= gc.ncells) return;
    const int I = C % gc.nx, J = (C / gc.nx) % gc.ny, K = DIM == 3 ? C / (gc.nx * gc.ny) : 0;
    double s = 0.0;
    for (int c = 0; c < (1 << DIM); ++c) {
        const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = DIM == 3 ? 2 * K + (c >> 2) : 0;
        s += Ef[i + gf.nx * (j + gf.ny * k)];
    }
    s *= 1.0 / (1 << DIM);
    E64[C] = s;
    E32[C] = (float)s;
}

// x' = x + P xc in place, 0 at the fine prescribed DOFs: the first half of the
// post-smoothing when the finest level smooths by z-marching columns.  Same
// terms and order as the tile kernel's staging (coarse corners ascending).
template <int DIM>
__global__ __launch_bounds__(256) void k_mg_prolong_add(Grid g, Grid gc,
                                                        const uint8_t* __restrict__ fixmask,
                                                        const float* __restrict__ xc,
                                                        float* __restrict__ x,
                                                        const PcgScalars* sc) {
    if (mg_done(sc)) return;
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nnodes) return;
    const int gi = n % g.NX, gj = (n / g.NX) % g.NY, gk = DIM == 3 ? n / (g.NX * g.NY) : 0;
    constexpr int NC = 1 << DIM;
    const int I0 = gi >> 1, J0 = gj >> 1, K0 = gk >> 1;
    const int oi = gi & 1, oj = gj & 1, ok = DIM == 3 ? (gk & 1) : 0;
    float v[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[d] = x[DIM * n + d];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int a = c & 1, bb = (c >> 1) & 1, cc = (c >> 2) & 1;
        if ((a && !oi) || (bb && !oj) || (cc && !ok)) continue;
        const float wt = (oi ? 0.5f : 1.f) * (oj ? 0.5f : 1.f) * (DIM == 3 && ok ? 0.5f : 1.f);
        const int N = (I0 + a) + gc.NX * ((J0 + bb) + gc.NY * (K0 + cc));
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] += wt * xc[DIM * N + d];
    }
    const uint8_t m = fixmask[n];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[DIM * n + d] = ((m >> d) & 1) ? 0.f : v[d];
}

// Smoother weight from the partials of the last two power steps, one
// workgroup.  The Rayleigh quotient theta = v'Kv / v'Dv = 1 - v'Du / v'Dv of
// D^-1 K at a step's input v is a lower bound of lambda_max(D^-1 K).  When
// the last step's theta exceeds the previous step's by at most 1 % (the
// iteration has levelled off), the weight is 1.9 / (margin theta) within
// [wsafe, wmax]; otherwise (a fresh start, or a design that changed a lot)
// it is wsafe, the design-independent bound's weight, until a later refresh
// levels off.  out = {w, 1 / sqrt(u'Du) (rescales the kept vector at the next
// refresh), theta}.
__global__ __launch_bounds__(256) void k_mg_omega(const double* __restrict__ prev,
                                                  const double* __restrict__ part, int nb,
                                                  float* __restrict__ out, float wsafe,
                                                  float wmax, float margin) {
    __shared__ double red[4];
    const double a0 = reduce_partials(prev, nb, 3, 0, red);
    const double b0 = reduce_partials(prev, nb, 3, 1, red);
    const double a = reduce_partials(part, nb, 3, 0, red);
    const double b = reduce_partials(part, nb, 3, 1, red);
    const double c = reduce_partials(part, nb, 3, 2, red);
    if (threadIdx.x != 0) return;
    const double theta0 = 1.0 - a0 / b0, theta = 1.0 - a / b;
    float w = wsafe;
    if (theta > 0.0 && theta < INFINITY && theta <= 1.01 * theta0)
        w = (float)fmin(fmax(1.9 / ((double)margin * theta), (double)wsafe), (double)wmax);
    out[0] = w;
    out[1] = (c > 0.0 && c < INFINITY) ? (float)(1.0 / sqrt(c)) : 1.f;
    out[2] = (float)theta;
}

// ------------------------------------------------------------ launchers --

static int mg_tile_blocks(const Grid& g) {
    const MgTileShape t = mg_tile(g.dim);
    int n = ((g.NX + t.tx - 1) / t.tx) * ((g.NY + t.ty - 1) / t.ty);
    if (g.dim == 3) n *= (g.NZ + t.tz - 1) / t.tz;
    // multiple of 8 when capped: keeps the XCD tile ranges contiguous
    // (TOHIP_MGCAP overrides the cap for measurement, within [256, kMaxPartials])
    static int cap = [] {
        const char* e = getenv("TOHIP_MGCAP");
        const int c = e ? atoi(e) : kMgBlocksCap;
        return (c >= 256 && c <= kMaxPartials) ? c & ~7 : kMgBlocksCap;
    }();
    return n <= cap ? n : cap;
}

int mg_blocks(const Grid& g) {
    if (mg_use_zm(g)) return zm_blocks(g);  // the z-marching smoother's grid
    return mg_tile_blocks(g);
}

// three partials per block in a buffer of kMaxPartials
int mg_power_blocks(const Grid& g) { return std::min(mg_tile_blocks(g), kMaxPartials / 3); }

hipError_t launch_mg_smooth(int mode, const MgLevelView& L, float w, const float* b,
                            const float* xprev, const MgLevelView* C, const float* xc,
                            float* xout, float* rout, double meandiag, const PcgScalars* sc,
                            double* part, hipStream_t s, const MgFusedUpdate* upd,
                            const float* aux) {
    if (mode == 3 && (!upd || !upd->sc || upd->sc != sc)) return hipErrorInvalidValue;
    if (mode == 5 && (!aux || !xprev || !part)) return hipErrorInvalidValue;
    if ((mode == 0 || mode == 1) && mg_use_zm(L.g)) {
        // z-marching columns; the post-smoothing forms x' = xprev + P xc while
        // staging (TOHIP_ZM_PROL=0, read on every call for comparisons: in a
        // separate pass first, in place in xprev, which the cycle does not
        // read again)
        float* xp = const_cast<float*>(xprev);
        if (mode == 1) {
            if (!C || !xc || !xp || C->g.nx * 2 != L.g.nx || C->g.ny * 2 != L.g.ny ||
                C->g.nz * 2 != L.g.nz)
                return hipErrorInvalidValue;
            const char* e = getenv("TOHIP_ZM_PROL");
            if (!(e && atoi(e) == 0))
                return launch_mg_smooth_zm(2, L.g, L.op, L.E32, L.fixmask, L.dinv, w, L.wdev, xp,
                                           b, xout, rout, meandiag, sc, part, s, &C->g, xc);
            hipLaunchKernelGGL(k_mg_prolong_add<3>, dim3((unsigned)((L.g.nnodes + 255) / 256)),
                               dim3(256), 0, s, L.g, C->g, L.fixmask, xc, xp, sc);
        }
        return launch_mg_smooth_zm(mode == 0 ? 1 : 2, L.g, L.op, L.E32, L.fixmask, L.dinv, w,
                                   L.wdev, xp, b, xout, rout, meandiag, sc, part, s);
    }
    const int nb = mode == 5 ? mg_power_blocks(L.g) : mg_tile_blocks(L.g);
    const Grid gc = C ? C->g : L.g;
    const MgFusedUpdate u = upd ? *upd : MgFusedUpdate{};
#define TOHIP_MGS(D, M, CB, X, Y, Z, N)                                                       \
    hipLaunchKernelGGL((k_mg_smooth<D, M, CB, X, Y, Z, N>), dim3(nb), dim3(N), 0, s, L.g, L.op, \
                       L.E32, L.fixmask, L.dinv, w, b, xprev, gc, xc, xout, rout, meandiag, sc, \
                       part, u, L.wdev, aux)
#define TOHIP_MGM(D, CB, X, Y, Z, N)             \
    if (mode == 0)                               \
        TOHIP_MGS(D, 0, CB, X, Y, Z, N);         \
    else if (mode == 1)                          \
        TOHIP_MGS(D, 1, CB, X, Y, Z, N);         \
    else if (mode == 3)                          \
        TOHIP_MGS(D, 3, CB, X, Y, Z, N);         \
    else if (mode == 5)                          \
        TOHIP_MGS(D, 5, CB, X, Y, Z, N);         \
    else                                         \
        TOHIP_MGS(D, 2, CB, X, Y, Z, N)
#define TOHIP_MG3(X, Y, Z, N)              \
    if (L.op.is_cube)                      \
        TOHIP_MGM(3, true, X, Y, Z, N);    \
    else                                   \
        TOHIP_MGM(3, false, X, Y, Z, N)
    if (L.g.dim == 3) {
        switch (mg_tile_variant()) {
            case 1: TOHIP_MG3(16, 4, 2, 256); break;
            case 2: TOHIP_MG3(8, 8, 2, 128); break;
            default: TOHIP_MG3(8, 4, 4, 128); break;
        }
    } else {
        TOHIP_MGM(2, false, 16, 8, 1, 128);
    }
#undef TOHIP_MG3
#undef TOHIP_MGM
#undef TOHIP_MGS
    return hipGetLastError();
}

// A fixed start vector for the lambda_max estimate: uniform in [-0.5, 0.5)
// from an integer hash of the index (prescribed DOFs are zeroed by the
// power step itself).
__global__ __launch_bounds__(256) void k_mg_fill_start(int64_t n, float* __restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    v[i] = (float)((h >> 40) & 0xFFFFFF) * (1.0f / 16777216.0f) - 0.5f;
}

hipError_t launch_mg_fill_start(int64_t n, float* v, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_fill_start, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v);
    return hipGetLastError();
}

hipError_t launch_mg_omega(const double* prev, const double* part, int nb, float* out,
                          float wsafe, float wmax, float margin, hipStream_t s) {
    if (nb < 1 || 3 * nb > kMaxPartials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mg_omega, dim3(1), dim3(256), 0, s, prev, part, nb, out, wsafe, wmax,
                       margin);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(const MgLevelView& F, const MgLevelView& C, const float* r,
                              float* bc, const PcgScalars* sc, hipStream_t s) {
    const dim3 grid((unsigned)((C.g.nnodes + 255) / 256));
    if (F.g.dim == 3)
        hipLaunchKernelGGL((k_mg_restrict<3>), grid, dim3(256), 0, s, F.g, C.g, C.fixmask, r, bc,
                           sc);
    else
        hipLaunchKernelGGL((k_mg_restrict<2>), grid, dim3(256), 0, s, F.g, C.g, C.fixmask, r, bc,
                           sc);
    return hipGetLastError();
}

bool mg_coarse_fits(const Grid& g) {
    return g.nnodes <= kMgCoarseMaxNodes && g.ncells <= kMgCoarseMaxCells;
}

hipError_t launch_mg_coarse(const MgLevelView& L, const MgCheb& ch, const float* b, float* x,
                            const PcgScalars* sc, hipStream_t s) {
    if (!mg_coarse_fits(L.g) || ch.iters < 1 || ch.iters > kMgChebMax) return hipErrorInvalidValue;
    const dim3 kMgCoarseThreads(mg_coarse_threads(L.g));
    if (L.g.dim == 3) {
        if (L.op.is_cube)
            hipLaunchKernelGGL((k_mg_coarse<3, true>), dim3(1), dim3(kMgCoarseThreads), 0, s, L.g,
                               L.op, L.E32, L.fixmask, L.dinv, ch, b, x, sc);
        else
            hipLaunchKernelGGL((k_mg_coarse<3, false>), dim3(1), dim3(kMgCoarseThreads), 0, s,
                               L.g, L.op, L.E32, L.fixmask, L.dinv, ch, b, x, sc);
    } else {
        hipLaunchKernelGGL((k_mg_coarse<2, false>), dim3(1), dim3(kMgCoarseThreads), 0, s, L.g,
                           L.op, L.E32, L.fixmask, L.dinv, ch, b, x, sc);
    }
    return hipGetLastError();
}

// Dense form of the coarsest solve: x = M b with M the matrix of the fixed
// Chebyshev operator above, built by k_mg_coarse<BASIS> once per operator
// refresh (column j = M e_j) and then symmetrised, M <- (M + M^T) / 2: the
// operator is symmetric in exact arithmetic (a polynomial in D^-1 K times
// D^-1), so this only removes FP32 rounding, and it lets row i be read as
// the contiguous column i.  Per V-cycle one wave per row: lane l sums
// M[i][j] b[j] over j = l, l + 64, ... in ascending order, then a fixed
// shuffle tree adds the 64 lane sums (deterministic).  The same operator as
// the iteration, spread over the chip instead of a chain of barriers in one
// workgroup, in one launch.
constexpr int kMgDenseMaxSteps = (3 * kMgCoarseMaxNodes + 63) / 64;

__global__ __launch_bounds__(256) void k_mg_dense_sym(int n, float* __restrict__ M) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    if (i >= j) return;  // each unordered pair {i, j} once
    const float s = 0.5f * (M[(size_t)j * n + i] + M[(size_t)i * n + j]);
    M[(size_t)j * n + i] = s;
    M[(size_t)i * n + j] = s;
}

__global__ __launch_bounds__(256) void k_mg_dense_rows(int n, const float* __restrict__ M,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ x,
                                                       const PcgScalars* sc) {
    if (mg_done(sc)) return;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;  // whole waves
    const float* row = M + (size_t)i * n;
    float v[kMgDenseMaxSteps], w[kMgDenseMaxSteps];
#pragma unroll
    for (int q = 0; q < kMgDenseMaxSteps; ++q) {  // every load issued before the sum
        const int j = min(lane + 64 * q, n - 1);
        v[q] = row[j];
        w[q] = b[j];
    }
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < kMgDenseMaxSteps; ++q)
        if (lane + 64 * q < n) acc += v[q] * w[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) x[i] = acc;
}

hipError_t launch_mg_coarse_basis(const MgLevelView& L, const MgCheb& ch, float* Mout,
                                  hipStream_t s) {
    if (!mg_coarse_fits(L.g) || ch.iters < 1 || ch.iters > kMgChebMax) return hipErrorInvalidValue;
    const dim3 grid((unsigned)L.g.ndof);
    const dim3 kMgCoarseThreads(mg_coarse_threads(L.g));
    if (L.g.dim == 3) {
        if (L.op.is_cube)
            hipLaunchKernelGGL((k_mg_coarse<3, true, true>), grid, dim3(kMgCoarseThreads), 0, s,
                               L.g, L.op, L.E32, L.fixmask, L.dinv, ch, nullptr, Mout, nullptr);
        else
            hipLaunchKernelGGL((k_mg_coarse<3, false, true>), grid, dim3(kMgCoarseThreads), 0, s,
                               L.g, L.op, L.E32, L.fixmask, L.dinv, ch, nullptr, Mout, nullptr);
    } else {
        hipLaunchKernelGGL((k_mg_coarse<2, false, true>), grid, dim3(kMgCoarseThreads), 0, s,
                           L.g, L.op, L.E32, L.fixmask, L.dinv, ch, nullptr, Mout, nullptr);
    }
    const int n = (int)L.g.ndof;
    const unsigned nb = (unsigned)(((int64_t)n * n + 255) / 256);
    hipLaunchKernelGGL(k_mg_dense_sym, dim3(nb), dim3(256), 0, s, n, Mout);
    return hipGetLastError();
}

hipError_t launch_mg_dense(int n, const float* M, const float* b, float* x, const PcgScalars* sc,
                           hipStream_t s) {
    if (n < 1 || n > 3 * kMgCoarseMaxNodes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mg_dense_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, M, b, x,
                       sc);
    return hipGetLastError();
}

hipError_t launch_mg_coarsen(const Grid& gf, const Grid& gc, const double* Ef, double* E64,
                             float* E32, hipStream_t s) {
    const dim3 grid((unsigned)((gc.ncells + 255) / 256));
    if (gf.dim == 3)
        hipLaunchKernelGGL((k_mg_coarsen<3>), grid, dim3(256), 0, s, gf, gc, Ef, E64, E32);
    else
        hipLaunchKernelGGL((k_mg_coarsen<2>), grid, dim3(256), 0, s, gf, gc, Ef, E64, E32);
    return hipGetLastError();
}

}  // namespace tohip
