// DensityFilterFun as a stencil (replaces the sparse J and J^T products of
// TopOpt.jl src/CheqFilters/density_filter.jl:35-47, J built at :49-108 from
// the breadth-first neighbour search of src/CheqFilters/CheqFilters.jl:66-118).
//
// That construction reduces to a closed form on a structured grid: the search
// pushes every node n within rmin of the cell centre once per adjacent cell,
// so with w_n = rmin - |x_n - x_i| and valence v_n (cells touching n),
//   (J x)_i = sum_n w_n * sum_{c in n} x_c  /  sum_n v_n w_n .
// The host lists the stencil nodes once (offsets from the cell's lower corner
// node, weights); `invden` holds 1 / sum_n v_n w_n per cell (boundary aware).
//   forward :  y_i = invden_i * sum_s w_s * S(n_s),   S(n) = sum_{c in n} x_c
//   adjoint :  z_c = sum_{n in c} sum_s w_s * invden_{n-s} * y_{n-s}
#include <algorithm>
#include <cstdlib>

#include "k_common.h"

namespace tohip {

template <int DIM>
__device__ __forceinline__ bool cell_ok(const Grid& g, int i, int j, int k) {
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && (DIM == 2 || (k >= 0 && k < g.nz));
}
template <int DIM>
__device__ __forceinline__ bool node_ok(const Grid& g, int i, int j, int k) {
    return i >= 0 && i < g.NX && j >= 0 && j < g.NY && (DIM == 2 || (k >= 0 && k < g.NZ));
}
template <int DIM>
__device__ __forceinline__ int64_t cid(const Grid& g, int i, int j, int k) {
    return i + (int64_t)g.nx * (j + (int64_t)g.ny * (DIM == 3 ? k : 0));
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_filter_invden(Grid g, const int* __restrict__ offs,
                                                            const double* __restrict__ wts,
                                                            int count, double* __restrict__ invden) {
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (e >= g.ncells) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = DIM == 3 ? e / ((int64_t)g.nx * g.ny) : 0;
    double den = 0.0;
    for (int s = 0; s < count; ++s) {
        const int ni = i + offs[3 * s], nj = j + offs[3 * s + 1], nk = k + offs[3 * s + 2];
        if (!node_ok<DIM>(g, ni, nj, nk)) continue;
        int val = 0;
        for (int c = 0; c < (1 << DIM); ++c)
            val += cell_ok<DIM>(g, ni - (c & 1), nj - ((c >> 1) & 1), nk - ((c >> 2) & 1));
        den += val * wts[s];
    }
    invden[e] = den > 0.0 ? 1.0 / den : 0.0;
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_filter_fwd(Grid g, const int* __restrict__ offs,
                                                         const double* __restrict__ wts, int count,
                                                         const double* __restrict__ invden,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y) {
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (e >= g.ncells) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = DIM == 3 ? e / ((int64_t)g.nx * g.ny) : 0;
    double num = 0.0;
    for (int s = 0; s < count; ++s) {
        const int ni = i + offs[3 * s], nj = j + offs[3 * s + 1], nk = k + offs[3 * s + 2];
        if (!node_ok<DIM>(g, ni, nj, nk)) continue;
        double sn = 0.0;
        for (int c = (1 << DIM) - 1; c >= 0; --c) {
            const int ci = ni - (c & 1), cj = nj - ((c >> 1) & 1), ck = nk - ((c >> 2) & 1);
            if (cell_ok<DIM>(g, ci, cj, ck)) sn += x[cid<DIM>(g, ci, cj, ck)];
        }
        num += wts[s] * sn;
    }
    y[e] = num * invden[e];
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_filter_adj(Grid g, const int* __restrict__ offs,
                                                         const double* __restrict__ wts, int count,
                                                         const double* __restrict__ invden,
                                                         const double* __restrict__ y,
                                                         double* __restrict__ z) {
    const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (e >= g.ncells) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = DIM == 3 ? e / ((int64_t)g.nx * g.ny) : 0;
    double acc = 0.0;
    for (int c = 0; c < (1 << DIM); ++c) {
        const int ni = i + (c & 1), nj = j + ((c >> 1) & 1), nk = DIM == 3 ? k + ((c >> 2) & 1) : 0;
        double tn = 0.0;
        for (int s = 0; s < count; ++s) {
            const int ci = ni - offs[3 * s], cj = nj - offs[3 * s + 1], ck = nk - offs[3 * s + 2];
            if (!cell_ok<DIM>(g, ci, cj, ck)) continue;
            const int64_t q = cid<DIM>(g, ci, cj, ck);
            tn += wts[s] * (y[q] * invden[q]);
        }
        acc += tn;
    }
    z[e] = acc;
}

// ---- z-marching tile kernels (3D, stencil halo H <= kFilterMaxHalo) ----
//
// One workgroup owns a TX x TY column of output cells and a chunk of z
// planes; it walks the chunk plane by plane with rolling LDS buffers of cell
// planes (with an H-cell halo in x and y) and node planes, so every input
// cell is read from memory once per chunk (plus the xy halo, shared with the
// neighbouring columns through L2) instead of once per stencil tap.
//   forward : cell plane q -> node sums S on node plane q (cells q-1, q)
//             -> output cell plane k = q - H:  y = invden * sum_s w_s S(n_s)
//   adjoint : v = invden * y on cell plane q -> T on node plane n = q - H + 1
//             (T(n) = sum_s w_s v(n - off_s)) -> output cell plane n - 1:
//             z = sum of T over the cell's corners
// Stencil node offsets lie in [1 - H, H] on every axis (the nodes within
// rmin of the cell centre, symmetric about it).  Summation orders are those
// of the per-cell kernels above.
constexpr int kFilterTX = 32, kFilterTY = 8;  // one output cell per thread per plane
constexpr int kFilterMaxTaps = 256;

template <bool ADJ, int H>
__global__ __launch_bounds__(kFilterTX * kFilterTY) void k_filter_zm(
    Grid g, const int* __restrict__ offs, const double* __restrict__ wts, int count,
    const double* __restrict__ invden, const double* __restrict__ x, double* __restrict__ y,
    int zchunk) {
    constexpr int TX = kFilterTX, TY = kFilterTY, NT = TX * TY;
    constexpr int RA = H - 1;
    constexpr int CX = TX + 2 * H, CY = TY + 2 * H, CP = CX * CY;  // cell plane + halo
    constexpr int NX = ADJ ? TX + 1 : TX + 2 * RA + 1;               // node plane
    constexpr int NY = ADJ ? TY + 1 : TY + 2 * RA + 1, NP = NX * NY;
    constexpr int CR = ADJ ? 2 * H : 2;  // cell planes kept
    constexpr int NR = ADJ ? 2 : 2 * H;  // node planes kept
    constexpr int NLC = (CP + NT - 1) / NT, NLN = (NP + NT - 1) / NT;
    __shared__ double cs[CR][CP];
    __shared__ double ns[NR][NP];
    __shared__ int s_off[kFilterMaxTaps];
    __shared__ int s_dz[kFilterMaxTaps];
    __shared__ double s_w[kFilterMaxTaps];

    const int tid = threadIdx.x;
    const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY;
    const int nch = (g.nz + zchunk - 1) / zchunk;
    const int b = xcd_remap(blockIdx.x, ntx * nty * nch);
    const int i0 = (b % ntx) * TX, j0 = ((b / ntx) % nty) * TY;
    const int kb = (b / (ntx * nty)) * zchunk, ke = min(kb + zchunk, g.nz);

    // stencil taps as (in-plane LDS offset, plane offset)
    for (int t = tid; t < count; t += NT) {
        const int a = offs[3 * t], bb = offs[3 * t + 1], c = offs[3 * t + 2];
        s_off[t] = ADJ ? -(a + CX * bb) : a + NX * bb;
        s_dz[t] = ADJ ? -c : c;
        s_w[t] = wts[t];
    }

    // cell plane q of the input (adjoint: invden * y), 0 outside the mesh,
    // loaded into a register stage; two stages alternate, so each plane's
    // loads are issued two steps before the plane is stored to LDS
    struct Stage {
        double v[NLC];
    };
    auto load_plane = [&](Stage& R, int q) {
#pragma unroll
        for (int l = 0; l < NLC; ++l) {
            const int e = tid + l * NT;
            const int gi = i0 - H + e % CX, gj = j0 - H + e / CX;
            const bool in = e < CP && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && q >= 0 &&
                            q < g.nz;
            double v = 0.0;
            if (in) {
                const uint32_t c = (uint32_t)gi + (uint32_t)g.nx * ((uint32_t)gj + (uint32_t)g.ny * q);
                v = ADJ ? at(x, c) * at(invden, c) : at(x, c);
            }
            R.v[l] = v;
        }
    };
    auto ring = [](int q, int r) { return (q + 4 * H) % r; };  // q >= -2H throughout

    auto step = [&](int q, Stage& R) {
        double* cq = cs[ring(q, CR)];
#pragma unroll
        for (int l = 0; l < NLC; ++l)
            if (tid + l * NT < CP) cq[tid + l * NT] = R.v[l];
        if (q + 2 < ke + H) load_plane(R, q + 2);  // this stage's next plane
        __syncthreads();
        if (!ADJ) {
            // node plane q: S(n) = sum of the (up to) 8 cells around n, from the
            // lowest cell up; node (ni, nj) of the plane is global (i0 - RA + ni, ...)
            if (q >= kb - RA) {
                const double* c0 = cs[ring(q - 1, CR)];
                const double* c1 = cq;
                double* nq = ns[ring(q, NR)];
#pragma unroll
                for (int l = 0; l < NLN; ++l) {
                    const int e = tid + l * NT;
                    if (e < NP) {
                        const int ni = e % NX, nj = e / NX;
                        const int b00 = ni + CX * nj;  // lowest cell (ni - RA - 1 + H = ni)
                        double sn = 0.0;
                        sn += c0[b00];
                        sn += c0[b00 + 1];
                        sn += c0[b00 + CX];
                        sn += c0[b00 + CX + 1];
                        sn += c1[b00];
                        sn += c1[b00 + 1];
                        sn += c1[b00 + CX];
                        sn += c1[b00 + CX + 1];
                        nq[e] = sn;
                    }
                }
            }
            __syncthreads();
            const int k = q - H;  // output plane: needs node planes k - RA .. k + RA + 1 = q
            if (k >= kb) {
                const int ci = tid % TX, cj = tid / TX;
                const int gi = i0 + ci, gj = j0 + cj;
                if (gi < g.nx && gj < g.ny) {
                    const int base = (ci + RA) + NX * (cj + RA);
                    double num = 0.0;
                    for (int t = 0; t < count; ++t)
                        num += s_w[t] * ns[ring(k + s_dz[t], NR)][base + s_off[t]];
                    const uint32_t c = (uint32_t)gi + (uint32_t)g.nx * ((uint32_t)gj + (uint32_t)g.ny * k);
                    at(y, c) = num * at(invden, c);
                }
            }
        } else {
            // node plane n = q - RA: T(n) = sum_s w_s v(n - off_s), cell planes
            // n - c for c in [-RA, RA + 1] = q - 2H + 1 .. q
            const int n = q - RA;
            if (n >= kb && n <= ke) {
                double* nq = ns[ring(n, NR)];
#pragma unroll
                for (int l = 0; l < NLN; ++l) {
                    const int e = tid + l * NT;
                    if (e < NP) {
                        const int ni = e % NX, nj = e / NX;
                        const int base = (ni + H) + CX * (nj + H);
                        double tn = 0.0;
                        for (int t = 0; t < count; ++t)
                            tn += s_w[t] * cs[ring(n + s_dz[t], CR)][base + s_off[t]];
                        nq[e] = tn;
                    }
                }
            }
            __syncthreads();
            const int k = n - 1;  // output plane: corners on node planes k, k + 1
            if (k >= kb) {
                const int ci = tid % TX, cj = tid / TX;
                const int gi = i0 + ci, gj = j0 + cj;
                if (gi < g.nx && gj < g.ny) {
                    const double* t0 = ns[ring(k, NR)];
                    const double* t1 = ns[ring(k + 1, NR)];
                    const int b00 = ci + NX * cj;
                    double acc = 0.0;  // corners in bit order (x, then y, then z)
                    acc += t0[b00];
                    acc += t0[b00 + 1];
                    acc += t0[b00 + NX];
                    acc += t0[b00 + NX + 1];
                    acc += t1[b00];
                    acc += t1[b00 + 1];
                    acc += t1[b00 + NX];
                    acc += t1[b00 + NX + 1];
                    const uint32_t c = (uint32_t)gi + (uint32_t)g.nx * ((uint32_t)gj + (uint32_t)g.ny * k);
                    at(y, c) = acc;
                }
            }
        }
        // the next plane's stores reuse the ring slots read above
        __syncthreads();
    };

    Stage A, B;
    load_plane(A, kb - H);
    load_plane(B, kb - H + 1);
    for (int q = kb - H; q < ke + H; q += 2) {
        step(q, A);
        if (q + 1 < ke + H) step(q + 1, B);
    }
}

// z chunk of the tile kernels: enough workgroups to fill the chip (about
// 2048 when the grid allows), at least 2H planes per chunk.
// TOHIP_FILTER_WGS overrides the workgroup target (measurement).
static int filter_zchunk(const Grid& g, int halo) {
    static int target = [] {
        const char* e = getenv("TOHIP_FILTER_WGS");
        const int t = e ? atoi(e) : 2048;
        return (t >= 64 && t <= 65536) ? t : 2048;
    }();
    const int ncol = ((g.nx + kFilterTX - 1) / kFilterTX) * ((g.ny + kFilterTY - 1) / kFilterTY);
    const int want = (target + ncol - 1) / ncol;  // chunks per column
    int zc = (g.nz + want - 1) / want;
    return std::max(zc, std::min(2 * halo, g.nz));
}

static bool filter_tiled(const Grid& g, int halo, int count) {
    return g.dim == 3 && halo >= 1 && halo <= kFilterMaxHalo && count <= kFilterMaxTaps;
}

hipError_t launch_filter_invden(const Grid& g, const int* offs, const double* wts, int count,
                                double* invden, hipStream_t s) {
    dim3 grid((unsigned)((g.ncells + kThreads - 1) / kThreads));
    if (g.dim == 3)
        hipLaunchKernelGGL(k_filter_invden<3>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                           invden);
    else
        hipLaunchKernelGGL(k_filter_invden<2>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                           invden);
    return hipGetLastError();
}

hipError_t launch_filter(const Grid& g, const int* offs, const double* wts, int count,
                         int halo, const double* invden, const double* x, double* y,
                         int transpose, hipStream_t s) {
    if (filter_tiled(g, halo, count)) {
        const int zc = filter_zchunk(g, halo);
        const int nb = ((g.nx + kFilterTX - 1) / kFilterTX) * ((g.ny + kFilterTY - 1) / kFilterTY) *
                       ((g.nz + zc - 1) / zc);
        const dim3 grid((unsigned)nb), blk(kFilterTX * kFilterTY);
#define TOHIP_FZM(ADJ, H)                                                                   \
    hipLaunchKernelGGL((k_filter_zm<ADJ, H>), grid, blk, 0, s, g, offs, wts, count, invden, x, \
                       y, zc)
        switch (halo * 2 + (transpose ? 1 : 0)) {
            case 2: TOHIP_FZM(false, 1); break;
            case 3: TOHIP_FZM(true, 1); break;
            case 4: TOHIP_FZM(false, 2); break;
            case 5: TOHIP_FZM(true, 2); break;
            case 6: TOHIP_FZM(false, 3); break;
            default: TOHIP_FZM(true, 3); break;
        }
#undef TOHIP_FZM
        return hipGetLastError();
    }
    dim3 grid((unsigned)((g.ncells + kThreads - 1) / kThreads));
    if (g.dim == 3) {
        if (transpose)
            hipLaunchKernelGGL(k_filter_adj<3>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                               invden, x, y);
        else
            hipLaunchKernelGGL(k_filter_fwd<3>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                               invden, x, y);
    } else {
        if (transpose)
            hipLaunchKernelGGL(k_filter_adj<2>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                               invden, x, y);
        else
            hipLaunchKernelGGL(k_filter_fwd<2>, grid, dim3(kThreads), 0, s, g, offs, wts, count,
                               invden, x, y);
    }
    return hipGetLastError();
}

// Elementwise projection (ProjectedDensityFilterFun's pre/post projections,
// src/CheqFilters/density_filter.jl:138-150): y = proj(x), or with `g` the
// pullback y = proj'(x) * g.
__global__ __launch_bounds__(kThreads) void k_project(int64_t n, int proj, double beta,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ g,
                                                      double* __restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (i >= n) return;
    double v, dv;
    projection_eval(proj, beta, x[i], v, dv);
    y[i] = g ? dv * g[i] : v;
}

hipError_t launch_project(int64_t n, int proj, double beta, const double* x, const double* g,
                          double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_project, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, s, n, proj, beta, x, g, y);
    return hipGetLastError();
}

}  // namespace tohip
