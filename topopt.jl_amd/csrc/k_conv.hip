// BF16 implicit-GEMM 3D convolution on MFMA (FP32 accumulate) and the
// element-wise kernels of the learned warm-start network.
//
// Layout: activations are NDHWC (channel fastest), BF16, with a multiple of
// 8 channels.  Weights of a conv with T = kd*kh*kw taps are [Cout][T][Cin32]
// BF16 with Cin32 = Cin rounded up to 32 (zero padded).
//
// Forward (stride 1, zero padding): Y[p][co] = act(b[co] + sum_{tap, ci}
//   W[co][tap][ci] X[p + tap - pad][ci]) as a GEMM with M = output positions,
//   N = Cout, K = taps x channels.  One wave owns 16 positions x 16*NSUB
//   channels and issues v_mfma_f32_16x16x32_bf16; operand lanes follow the
//   instruction's layout: A[m = lane%16][k = 8*(lane/16) + 0..7],
//   B[k = 8*(lane/16) + 0..7][n = lane%16], D[m = 4*(lane/16) + 0..3][n = lane%16].
// Backward-data is the same kernel on dY with the tap-flipped, channel-
//   transposed weights Wt[ci][T-1-tap][co] and padding k-1-pad.
// Backward-weight: dW[co][tap][ci] = sum_p dY[p][co] X[p + tap - pad][ci], a
//   GEMM with K = positions, split over position ranges into FP32 partials
//   that a second kernel sums in a fixed order (deterministic).
#include <stdint.h>

#include <cstdlib>

#include "k_common.h"
#include "nn_internal.h"

namespace tohip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// Round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

// ---------------------------------------------------------------- forward --
// The K dimension is walked in groups of 8 channels of one tap: group g is
// (tap = g / gpt, channels 8 (g % gpt) .. +7) with gpt = Cin / 8 groups per
// tap.  One 32-wide MFMA K-step takes 4 consecutive groups (one per 16-lane
// quarter of the wave), which may belong to different taps: a narrow input
// (8, 16, 48 channels) is packed across taps instead of padded to 32
// channels with zeros (the weights' zero padding past Cin is never read).
//
// Block = 4 waves, each wave 16 output positions x 16*NSUB output channels.
// KS = 1: the waves own 4 consecutive position tiles and walk all K-steps.
// KS = 4: the waves share one position tile and take every 4th K-step; the
//   4 partial accumulators are summed through LDS in wave order.  For the
//   coarse layers, where P / 64 blocks cannot fill the chip, this gives 4x
//   the blocks and a quarter of the dependent load -> MFMA chain per wave.
constexpr int kConvMaxTaps = 32;

template <int NSUB, int KS>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvDesc c, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ W,
                                                  const float* __restrict__ bias,
                                                  uint16_t* __restrict__ Y,
                                                  uint16_t* __restrict__ Z,
                                                  float* __restrict__ Yf) {
    __shared__ int4 taps[kConvMaxTaps];  // dz, dy, dx, element offset of the tap
    __shared__ f4 red[KS > 1 ? KS - 1 : 1][NSUB][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int T = c.kd * c.kh * c.kw;
    if ((int)threadIdx.x < T) {
        const int t = threadIdx.x;
        const int dz = t / (c.kh * c.kw), dy = (t / c.kw) % c.kh, dx = t % c.kw;
        taps[t] = make_int4(dz, dy, dx, ((dz * c.H + dy) * c.W + dx) * c.Cin);
    }
    __syncthreads();
    const int64_t P = (int64_t)c.B * c.OD * c.OH * c.OW;
    const int64_t p0 = (KS == 1 ? (int64_t)blockIdx.x * 4 + wave : (int64_t)blockIdx.x) * 16;
    const int co0 = blockIdx.y * 16 * NSUB;

    // this lane's A row: output position p0 + row
    const int64_t p = p0 + row;
    const bool pvalid = p < P;
    int ox = 0, oy = 0, oz = 0, bt = 0;
    if (pvalid) {
        int64_t t = p;
        ox = (int)(t % c.OW); t /= c.OW;
        oy = (int)(t % c.OH); t /= c.OH;
        oz = (int)(t % c.OD); bt = (int)(t / c.OD);
    }
    // element index of input voxel (oz - pd, oy - ph, ox - pw); a tap adds taps[].w
    const int64_t xbase =
        ((((int64_t)bt * c.D + oz - c.pd) * c.H + oy - c.ph) * c.W + ox - c.pw) * c.Cin;
    const int gpt = c.Cin / 8;
    const float inv_gpt = 1.0f / (float)gpt;
    const int S = (T * gpt + 3) / 4;  // K-steps
    int64_t wrow[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
        const int co = co0 + 16 * j + row;
        wrow[j] = co < c.Cout ? (int64_t)co * T * c.Cin32 : -1;
    }

    f4 acc[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero = {};
    // U K-steps per iteration: all their loads are issued before the MFMAs,
    // so a wave keeps U gathers in flight (steps past S load zeros).
    constexpr int U = NSUB >= 4 ? 2 : 4;
    for (int s0 = (KS == 1 ? 0 : wave); s0 < S; s0 += U * KS) {
        bf16x8 a[U], bv[U][NSUB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int g = 4 * (s0 + u * KS) + kq;
            // g / gpt, exact: (g + 0.5) / gpt is at least 0.5 / gpt from an integer
            const int tap = (int)(((float)g + 0.5f) * inv_gpt);
            const int ci = 8 * (g - tap * gpt);
            a[u] = zero;
            const bool tok = tap < T;
            if (tok) {
                const int4 td = taps[tap];
                const int iz = oz + td.x - c.pd, iy = oy + td.y - c.ph, ix = ox + td.z - c.pw;
                if (pvalid && iz >= 0 && iz < c.D && iy >= 0 && iy < c.H &&
                    ix >= 0 && ix < c.W)
                    a[u] = load8(X + xbase + td.w + ci);
            }
#pragma unroll
            for (int j = 0; j < NSUB; ++j)
                bv[u][j] = (tok && wrow[j] >= 0)
                               ? load8(W + wrow[j] + (int64_t)tap * c.Cin32 + ci)
                               : zero;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < NSUB; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], bv[u][j], acc[j], 0, 0, 0);
    }
    if (KS > 1) {
        if (wave > 0)
#pragma unroll
            for (int j = 0; j < NSUB; ++j) red[wave - 1][j][lane] = acc[j];
        __syncthreads();
        if (wave > 0) return;
#pragma unroll
        for (int w = 0; w < KS - 1; ++w)
#pragma unroll
            for (int j = 0; j < NSUB; ++j) acc[j] += red[w][j][lane];
    }
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
        const int co = co0 + 16 * j + row;
        if (co >= c.Cout) continue;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t pos = p0 + 4 * kq + i;
            if (pos >= P) continue;
            const float z = acc[j][i] + bb;
            const float y = c.act == 1 ? silu(z) : z;
            if (Yf) {
                if (co < c.out_cvalid) Yf[pos * c.out_cvalid + co] = y;
            } else {
                Y[pos * c.out_cstride + c.out_coff + co] = f2bf(y);
            }
            if (Z) Z[pos * c.Cout + co] = f2bf(z);
        }
    }
}

// Launch shape: KS = 1 with the widest channel tile while P / 64 blocks fill
// the chip; otherwise KS = 4 and the narrowest tile that still gives about
// 4 blocks per CU.  TOHIP_CONV_KS = 1 / 4 forces the split (measurement).
hipError_t launch_conv_fwd(const ConvDesc& c, const uint16_t* X, const uint16_t* W,
                           const float* bias, uint16_t* Y, uint16_t* Z, float* Yf,
                           hipStream_t s) {
    static const int forced = [] {
        const char* e = getenv("TOHIP_CONV_KS");
        return e ? atoi(e) : 0;
    }();
    const int T = c.kd * c.kh * c.kw;
    if (T > kConvMaxTaps || c.Cin % 8 != 0 || c.Cin32 % 32 != 0 || c.Cin > c.Cin32)
        return hipErrorInvalidValue;
    const int64_t P = (int64_t)c.B * c.OD * c.OH * c.OW;
    const int nsub_wide = c.Cout >= 64 ? 4 : c.Cout >= 32 ? 2 : 1;
    const int64_t wide_blocks = (P + 63) / 64 * ((c.Cout + 16 * nsub_wide - 1) / (16 * nsub_wide));
    const int ks = (forced == 1 || forced == 4) ? forced : wide_blocks >= 1024 ? 1 : 4;
    int nsub = nsub_wide;
    if (ks == 4)
        while (nsub > 1 && (P + 15) / 16 * ((c.Cout + 16 * nsub - 1) / (16 * nsub)) < 1024)
            nsub /= 2;
    const unsigned gx = (unsigned)(ks == 1 ? (P + 63) / 64 : (P + 15) / 16);
    const unsigned gy = (unsigned)((c.Cout + 16 * nsub - 1) / (16 * nsub));
#define TOHIP_CONV(NS, K)                                                                     \
    hipLaunchKernelGGL((k_conv_fwd<NS, K>), dim3(gx, gy), dim3(256), 0, s, c, X, W, bias, Y, Z, \
                       Yf)
    if (ks == 1) {
        if (nsub == 4) TOHIP_CONV(4, 1);
        else if (nsub == 2) TOHIP_CONV(2, 1);
        else TOHIP_CONV(1, 1);
    } else {
        if (nsub == 4) TOHIP_CONV(4, 4);
        else if (nsub == 2) TOHIP_CONV(2, 4);
        else TOHIP_CONV(1, 4);
    }
#undef TOHIP_CONV
    return hipGetLastError();
}

// ---------------------------------------------------------- weight grad --
// Block = one wave: (tap, 16 input channels, 16 output channels, position
// split).  A[m = co][k = p] = dY[p][co], B[k = p][n = ci] = X[p + tap - pad][ci].
__global__ __launch_bounds__(64) void k_conv_bwd_w(ConvDesc c, const uint16_t* __restrict__ X,
                                                   const uint16_t* __restrict__ dY,
                                                   float* __restrict__ part, int nsplit) {
    const int lane = threadIdx.x, row = lane & 15, kq = lane >> 4;
    const int T = c.kd * c.kh * c.kw;
    const int nci = c.Cin32 / 16;
    const int tap = blockIdx.x / nci, ci0 = (blockIdx.x % nci) * 16;
    const int co0 = blockIdx.y * 16;
    const int split = blockIdx.z;
    const int64_t P = (int64_t)c.B * c.OD * c.OH * c.OW;
    const int64_t chunk = ((P + nsplit - 1) / nsplit + 31) / 32 * 32;
    const int64_t pb = split * chunk, pe = pb + chunk < P ? pb + chunk : P;
    const int dz = tap / (c.kh * c.kw), dy = (tap / c.kw) % c.kh, dx = tap % c.kw;
    const int co = co0 + row, ci = ci0 + row;
    const bool co_ok = co < c.Cout, ci_ok = ci < c.Cin;
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    // Output position of this lane's first element, advanced incrementally
    // (no div/mod chain per element): p = ((b*OD + oz)*OH + oy)*OW + ox.
    int ox, oy, oz, bb;
    {
        int64_t t = pb + 8 * kq;
        ox = (int)(t % c.OW);
        t /= c.OW;
        oy = (int)(t % c.OH);
        t /= c.OH;
        oz = (int)(t % c.OD);
        bb = (int)(t / c.OD);
    }
    for (int64_t base = pb; base < pe; base += 32) {
        bf16x8 a, bv;
        int x0 = ox, y0 = oy, z0 = oz, b0 = bb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t p = base + 8 * kq + i;
            uint16_t av = 0, xv = 0;
            if (p < pe) {
                if (co_ok) av = dY[p * c.Cout + co];
                const int iz = z0 + dz - c.pd, iy = y0 + dy - c.ph, ix = x0 + dx - c.pw;
                if (ci_ok && iz >= 0 && iz < c.D && iy >= 0 && iy < c.H && ix >= 0 && ix < c.W)
                    xv = X[((((int64_t)b0 * c.D + iz) * c.H + iy) * c.W + ix) * c.Cin + ci];
            }
            a[i] = __builtin_bit_cast(__bf16, av);
            bv[i] = __builtin_bit_cast(__bf16, xv);
            // next position
            if (++x0 == c.OW) {
                x0 = 0;
                if (++y0 == c.OH) {
                    y0 = 0;
                    if (++z0 == c.OD) {
                        z0 = 0;
                        ++b0;
                    }
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
        // advance this lane's start by 32 positions
        int step = 32;
        while (step > 0) {
            const int room = c.OW - ox;
            if (step < room) {
                ox += step;
                step = 0;
            } else {
                step -= room;
                ox = 0;
                if (++oy == c.OH) {
                    oy = 0;
                    if (++oz == c.OD) {
                        oz = 0;
                        ++bb;
                    }
                }
            }
        }
    }
    // D[m = co0 + 4kq + i][n = ci0 + row]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int com = co0 + 4 * kq + i;
        if (com < c.Cout)
            part[(((int64_t)split * c.Cout + com) * T + tap) * c.Cin32 + ci0 + row] = acc[i];
    }
}

// The same tile through LDS: per step of 32 positions each lane loads one
// position's 8 output-gradient and 8 input channels with 16-byte loads (lane
// = 2 * position + half), writes them transposed into LDS ([channel][k]), and
// the MFMA operands come back as one 16-byte LDS read per lane, instead of 16
// scattered 2-byte global loads per lane per MFMA.  Needs Cin and Cout
// multiples of 8 and 16-byte aligned tensors (checked at launch).
__global__ __launch_bounds__(64) void k_conv_bwd_w_lds(ConvDesc c, const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ dY,
                                                       float* __restrict__ part, int nsplit) {
    constexpr int KP = 32 + 8;  // LDS row: 32 positions + 8 of padding (16-byte aligned rows)
    __shared__ __align__(16) uint16_t sA[16][KP];  // [co - co0][k]
    __shared__ __align__(16) uint16_t sB[16][KP];  // [ci - ci0][k]
    const int lane = threadIdx.x, row = lane & 15, kq = lane >> 4;
    const int lp = lane >> 1, lh = lane & 1;  // loading role: position, channel half
    const int T = c.kd * c.kh * c.kw;
    const int nci = c.Cin32 / 16;
    const int tap = blockIdx.x / nci, ci0 = (blockIdx.x % nci) * 16;
    const int co0 = blockIdx.y * 16;
    const int split = blockIdx.z;
    const int64_t P = (int64_t)c.B * c.OD * c.OH * c.OW;
    const int64_t chunk = ((P + nsplit - 1) / nsplit + 31) / 32 * 32;
    const int64_t pb = split * chunk, pe = pb + chunk < P ? pb + chunk : P;
    const int dz = tap / (c.kh * c.kw), dy = (tap / c.kw) % c.kh, dx = tap % c.kw;
    const int col = co0 + 8 * lh, cil = ci0 + 8 * lh;  // this lane's 8 channels
    const bool co_ok = col < c.Cout, ci_ok = cil < c.Cin;
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    // output position p = base + lp, advanced by 32 per step
    int ox, oy, oz, bb;
    {
        int64_t t = pb + lp;
        ox = (int)(t % c.OW);
        t /= c.OW;
        oy = (int)(t % c.OH);
        t /= c.OH;
        oz = (int)(t % c.OD);
        bb = (int)(t / c.OD);
    }
    for (int64_t base = pb; base < pe; base += 32) {
        const int64_t p = base + lp;
        uint4 av = make_uint4(0, 0, 0, 0), xv = make_uint4(0, 0, 0, 0);
        if (p < pe) {
            if (co_ok) av = *reinterpret_cast<const uint4*>(dY + p * c.Cout + col);
            const int iz = oz + dz - c.pd, iy = oy + dy - c.ph, ix = ox + dx - c.pw;
            if (ci_ok && iz >= 0 && iz < c.D && iy >= 0 && iy < c.H && ix >= 0 && ix < c.W)
                xv = *reinterpret_cast<const uint4*>(
                    X + ((((int64_t)bb * c.D + iz) * c.H + iy) * c.W + ix) * c.Cin + cil);
        }
        const uint32_t aw[4] = {av.x, av.y, av.z, av.w}, xw[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sA[8 * lh + 2 * j][lp] = (uint16_t)(aw[j] & 0xffffu);
            sA[8 * lh + 2 * j + 1][lp] = (uint16_t)(aw[j] >> 16);
            sB[8 * lh + 2 * j][lp] = (uint16_t)(xw[j] & 0xffffu);
            sB[8 * lh + 2 * j + 1][lp] = (uint16_t)(xw[j] >> 16);
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[row][8 * kq]);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sB[row][8 * kq]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
        __syncthreads();  // the next step's stores reuse the tiles
        // advance this lane's position by 32
        int step = 32;
        while (step > 0) {
            const int room = c.OW - ox;
            if (step < room) {
                ox += step;
                step = 0;
            } else {
                step -= room;
                ox = 0;
                if (++oy == c.OH) {
                    oy = 0;
                    if (++oz == c.OD) {
                        oz = 0;
                        ++bb;
                    }
                }
            }
        }
    }
    // D[m = co0 + 4kq + i][n = ci0 + row]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int com = co0 + 4 * kq + i;
        if (com < c.Cout)
            part[(((int64_t)split * c.Cout + com) * T + tap) * c.Cin32 + ci0 + row] = acc[i];
    }
}

// dW = sum over splits (fixed order); db[co] = sum_p dY[p][co] (one block per co).
__global__ __launch_bounds__(256) void k_conv_bwd_w_reduce(int n, int nsplit,
                                                           const float* __restrict__ part,
                                                           float* __restrict__ dW) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * n + i];
    dW[i] = s;
}

// db[co] = sum_p dY[p][co]: kBiasBlocks blocks sweep contiguous rows of dY
// (coalesced), one partial per (block, co); a second kernel sums the
// partials of each channel in block order.
constexpr int kBiasBlocks = 256;
__global__ __launch_bounds__(256) void k_bias_grad(int64_t P, int Cout,
                                                   const uint16_t* __restrict__ dY,
                                                   float* __restrict__ part) {
    __shared__ float red[256];
    const int rows = 256 / Cout;  // Cout divides 256 (8..128)
    const int co = threadIdx.x % Cout, r = threadIdx.x / Cout;
    const int64_t chunk = (P + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = blockIdx.x * chunk, p1 = p0 + chunk < P ? p0 + chunk : P;
    float s = 0.f;
    if (r < rows)
        for (int64_t p = p0 + r; p < p1; p += rows) s += bf2f(dY[p * Cout + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < Cout) {
        float t = 0.f;
        for (int q = 0; q < rows; ++q) t += red[q * Cout + threadIdx.x];
        part[(size_t)blockIdx.x * Cout + threadIdx.x] = t;
    }
}

// One block per channel: thread b loads the partial of block b, then a fixed
// reduction tree (deterministic).
__global__ __launch_bounds__(kBiasBlocks) void k_bias_grad_finish(int Cout,
                                                                  const float* __restrict__ part,
                                                                  float* __restrict__ db) {
    __shared__ float red[kBiasBlocks];
    const int co = blockIdx.x;
    red[threadIdx.x] = part[(size_t)threadIdx.x * Cout + co];
    __syncthreads();
    for (int w = kBiasBlocks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = red[0];
}

hipError_t launch_conv_bwd_w(const ConvDesc& c, const uint16_t* X, const uint16_t* dY,
                             float* part, int nsplit, float* dW, float* db, hipStream_t s) {
    const int T = c.kd * c.kh * c.kw;
    dim3 grid(T * (c.Cin32 / 16), (c.Cout + 15) / 16, nsplit);
    // TOHIP_CONV_BWDW=0 keeps the direct-gather kernel (measurement)
    static const bool lds = [] {
        const char* e = getenv("TOHIP_CONV_BWDW");
        return !(e && atoi(e) == 0);
    }();
    const bool vec = lds && c.Cout % 8 == 0 && c.Cin % 8 == 0 &&
                     reinterpret_cast<uintptr_t>(X) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(dY) % 16 == 0;
    if (vec)
        hipLaunchKernelGGL(k_conv_bwd_w_lds, grid, dim3(64), 0, s, c, X, dY, part, nsplit);
    else
        hipLaunchKernelGGL(k_conv_bwd_w, grid, dim3(64), 0, s, c, X, dY, part, nsplit);
    const int n = c.Cout * T * c.Cin32;
    hipLaunchKernelGGL(k_conv_bwd_w_reduce, dim3((n + 255) / 256), dim3(256), 0, s, n, nsplit,
                       part, dW);
    if (db) {
        // the bias partials reuse the split partial buffer after the reduce above
        if (256 % c.Cout != 0 || (int64_t)kBiasBlocks * c.Cout > (int64_t)nsplit * n)
            return hipErrorInvalidValue;
        const int64_t P = (int64_t)c.B * c.OD * c.OH * c.OW;
        hipLaunchKernelGGL(k_bias_grad, dim3(kBiasBlocks), dim3(256), 0, s, P, c.Cout, dY, part);
        hipLaunchKernelGGL(k_bias_grad_finish, dim3(c.Cout), dim3(kBiasBlocks), 0, s, c.Cout, part,
                           db);
    }
    return hipGetLastError();
}

// --------------------------------------------------- weight conversions --
// Master FP32 [Cout][T][Cin] -> BF16 forward [Cout][T][Cin32] and
// backward-data Wt[Cin][T-1-tap][Cout32].
__global__ void k_weights_to_bf16(int Cout, int T, int Cin, int Cin32, int Cout32,
                                  const float* __restrict__ w, uint16_t* __restrict__ wf,
                                  uint16_t* __restrict__ wt) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)Cout * T * Cin32;
    if (i < n) {
        const int ci = (int)(i % Cin32);
        const int tap = (int)((i / Cin32) % T);
        const int co = (int)(i / ((int64_t)Cin32 * T));
        const float v = ci < Cin ? w[((int64_t)co * T + tap) * Cin + ci] : 0.f;
        wf[i] = f2bf(v);
    }
    const int64_t nt = (int64_t)Cin * T * Cout32;
    if (wt && i < nt) {
        const int co = (int)(i % Cout32);
        const int tapf = (int)((i / Cout32) % T);
        const int ci = (int)(i / ((int64_t)Cout32 * T));
        const int tap = T - 1 - tapf;
        const float v = co < Cout ? w[((int64_t)co * T + tap) * Cin + ci] : 0.f;
        wt[i] = f2bf(v);
    }
}

hipError_t launch_weights_to_bf16(int Cout, int T, int Cin, int Cin32, int Cout32, const float* w,
                                  uint16_t* wf, uint16_t* wt, hipStream_t s) {
    int64_t n = (int64_t)Cout * T * Cin32;
    const int64_t nt = (int64_t)Cin * T * Cout32;
    if (wt && nt > n) n = nt;
    hipLaunchKernelGGL(k_weights_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       Cout, T, Cin, Cin32, Cout32, w, wf, wt);
    return hipGetLastError();
}

// -------------------------------------------------------- element-wise --
// dZ = dA * silu'(z)
__global__ void k_silu_bwd(int64_t n, const uint16_t* __restrict__ dA,
                           const uint16_t* __restrict__ Z, uint16_t* __restrict__ dZ) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float z = bf2f(Z[i]);
        const float sg = 1.0f / (1.0f + __expf(-z));
        dZ[i] = f2bf(bf2f(dA[i]) * sg * (1.0f + z * (1.0f - sg)));
    }
}

// Strided channel view helpers: tensor [P][cstride], view [coff, coff + C).
// 2x2x2 average pool: out (B, D/2, H/2, W/2, C) from in (B, D, H, W) view.
__global__ void k_avgpool_fwd(int B, int D, int H, int W, int C, const uint16_t* __restrict__ in,
                              int in_cs, int in_co, uint16_t* __restrict__ out) {
    const int64_t n = (int64_t)B * (D / 2) * (H / 2) * (W / 2) * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % (W / 2)); t /= (W / 2);
        const int y = (int)(t % (H / 2)); t /= (H / 2);
        const int z = (int)(t % (D / 2));
        const int b = (int)(t / (D / 2));
        float s = 0.f;
        for (int c8 = 0; c8 < 8; ++c8) {
            const int zz = 2 * z + ((c8 >> 2) & 1), yy = 2 * y + ((c8 >> 1) & 1),
                      xx = 2 * x + (c8 & 1);
            s += bf2f(in[((((int64_t)b * D + zz) * H + yy) * W + xx) * in_cs + in_co + ch]);
        }
        out[i] = f2bf(0.125f * s);
    }
}

// d(in view) += broadcast(d out) / 8  (accumulate = 1 adds to the view)
__global__ void k_avgpool_bwd(int B, int D, int H, int W, int C, const uint16_t* __restrict__ dout,
                              uint16_t* __restrict__ din, int in_cs, int in_co, int accumulate) {
    const int64_t n = (int64_t)B * D * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % W); t /= W;
        const int y = (int)(t % H); t /= H;
        const int z = (int)(t % D);
        const int b = (int)(t / D);
        const float g =
            0.125f * bf2f(dout[((((int64_t)b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) +
                               x / 2) * C + ch]);
        const int64_t o = (((int64_t)b * D + z) * H + y) * W + x;
        uint16_t* q = din + o * in_cs + in_co + ch;
        *q = f2bf(accumulate ? bf2f(*q) + g : g);
    }
}

// nearest 2x upsample of in (B, D, H, W, C) into the view [coff, coff+C) of
// out (B, 2D, 2H, 2W, out_cs)
__global__ void k_upsample_fwd(int B, int D, int H, int W, int C, const uint16_t* __restrict__ in,
                               uint16_t* __restrict__ out, int out_cs, int out_co) {
    const int64_t n = (int64_t)B * 8 * D * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % (2 * W)); t /= (2 * W);
        const int y = (int)(t % (2 * H)); t /= (2 * H);
        const int z = (int)(t % (2 * D));
        const int b = (int)(t / (2 * D));
        out[((((int64_t)b * 2 * D + z) * 2 * H + y) * 2 * W + x) * out_cs + out_co + ch] =
            in[((((int64_t)b * D + z / 2) * H + y / 2) * W + x / 2) * C + ch];
    }
}

// The same with 8 channels (16 bytes) per thread: C, out_cs and out_co are
// multiples of 8 and the tensors 16-byte aligned.
__global__ void k_upsample_fwd8(int B, int D, int H, int W, int C, const uint16_t* __restrict__ in,
                                uint16_t* __restrict__ out, int out_cs, int out_co) {
    const int C8 = C / 8;
    const int64_t n = (int64_t)B * 8 * D * H * W * C8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = 8 * (int)(i % C8);
        int64_t t = i / C8;
        const int x = (int)(t % (2 * W)); t /= (2 * W);
        const int y = (int)(t % (2 * H)); t /= (2 * H);
        const int z = (int)(t % (2 * D));
        const int b = (int)(t / (2 * D));
        *reinterpret_cast<uint4*>(
            out + ((((int64_t)b * 2 * D + z) * 2 * H + y) * 2 * W + x) * out_cs + out_co + ch) =
            *reinterpret_cast<const uint4*>(
                in + ((((int64_t)b * D + z / 2) * H + y / 2) * W + x / 2) * C + ch);
    }
}

// d in (B, D, H, W, C) = sum of the 8 children in the view of d out
__global__ void k_upsample_bwd(int B, int D, int H, int W, int C,
                               const uint16_t* __restrict__ dout, int out_cs, int out_co,
                               uint16_t* __restrict__ din) {
    const int64_t n = (int64_t)B * D * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % W); t /= W;
        const int y = (int)(t % H); t /= H;
        const int z = (int)(t % D);
        const int b = (int)(t / D);
        float s = 0.f;
        for (int c8 = 0; c8 < 8; ++c8) {
            const int zz = 2 * z + ((c8 >> 2) & 1), yy = 2 * y + ((c8 >> 1) & 1),
                      xx = 2 * x + (c8 & 1);
            s += bf2f(dout[((((int64_t)b * 2 * D + zz) * 2 * H + yy) * 2 * W + xx) * out_cs +
                           out_co + ch]);
        }
        din[i] = f2bf(s);
    }
}

// dst[p][c] (+)= src[p][coff + c] for c < C (dst has C channels)
__global__ void k_slice_copy(int64_t P, int C, const uint16_t* __restrict__ src, int src_cs,
                             int src_co, uint16_t* __restrict__ dst, int accumulate) {
    const int64_t n = P * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % C);
        const int64_t p = i / C;
        const float v = bf2f(src[p * src_cs + src_co + ch]);
        dst[i] = f2bf(accumulate ? bf2f(dst[i]) + v : v);
    }
}

static inline unsigned ew_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_silu_bwd(int64_t n, const uint16_t* dA, const uint16_t* Z, uint16_t* dZ,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_silu_bwd, dim3(ew_blocks(n)), dim3(256), 0, s, n, dA, Z, dZ);
    return hipGetLastError();
}
hipError_t launch_avgpool_fwd(int B, int D, int H, int W, int C, const uint16_t* in, int in_cs,
                              int in_co, uint16_t* out, hipStream_t s) {
    const int64_t n = (int64_t)B * (D / 2) * (H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(k_avgpool_fwd, dim3(ew_blocks(n)), dim3(256), 0, s, B, D, H, W, C, in, in_cs,
                       in_co, out);
    return hipGetLastError();
}
hipError_t launch_avgpool_bwd(int B, int D, int H, int W, int C, const uint16_t* dout,
                              uint16_t* din, int in_cs, int in_co, int accumulate, hipStream_t s) {
    const int64_t n = (int64_t)B * D * H * W * C;
    hipLaunchKernelGGL(k_avgpool_bwd, dim3(ew_blocks(n)), dim3(256), 0, s, B, D, H, W, C, dout,
                       din, in_cs, in_co, accumulate);
    return hipGetLastError();
}
hipError_t launch_upsample_fwd(int B, int D, int H, int W, int C, const uint16_t* in,
                               uint16_t* out, int out_cs, int out_co, hipStream_t s) {
    const int64_t n = (int64_t)B * 8 * D * H * W * C;
    if (C % 8 == 0 && out_cs % 8 == 0 && out_co % 8 == 0 && ((uintptr_t)in & 15) == 0 &&
        ((uintptr_t)out & 15) == 0)
        hipLaunchKernelGGL(k_upsample_fwd8, dim3(ew_blocks(n / 8)), dim3(256), 0, s, B, D, H, W, C,
                           in, out, out_cs, out_co);
    else
        hipLaunchKernelGGL(k_upsample_fwd, dim3(ew_blocks(n)), dim3(256), 0, s, B, D, H, W, C, in,
                           out, out_cs, out_co);
    return hipGetLastError();
}
hipError_t launch_upsample_bwd(int B, int D, int H, int W, int C, const uint16_t* dout,
                               int out_cs, int out_co, uint16_t* din, hipStream_t s) {
    const int64_t n = (int64_t)B * D * H * W * C;
    hipLaunchKernelGGL(k_upsample_bwd, dim3(ew_blocks(n)), dim3(256), 0, s, B, D, H, W, C, dout,
                       out_cs, out_co, din);
    return hipGetLastError();
}
hipError_t launch_slice_copy(int64_t P, int C, const uint16_t* src, int src_cs, int src_co,
                             uint16_t* dst, int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_slice_copy, dim3(ew_blocks(P * C)), dim3(256), 0, s, P, C, src, src_cs,
                       src_co, dst, accumulate);
    return hipGetLastError();
}

}  // namespace tohip
