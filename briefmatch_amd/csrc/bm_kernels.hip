// HIP kernels of the BriefMatch flow path for gfx950 (CDNA4, wave64).
//
// Every stage of the reference's device code (kernels/briefmatch_cuda.cu,
// kernels/util_cuda.cu) has a kernel here, written for this hardware:
//   * one work-group per tile, one-dimensional grids whose work-group order is
//     remapped so each of the 8 XCDs (own L2) walks a contiguous band of rows;
//   * BRIEF stages the (up-sampled) image tile plus a halo in LDS, and the
//     fused variant computes the bilinear up-sampling while filling LDS so the
//     up-sampled frame never reaches HBM;
//   * jump flooding is level-synchronous (one launch per level, ping-pong
//     buffers), with all four candidate gathers issued before any cost is
//     compared, and the random search fused into the last level (or, after a
//     zero initialisation, into the initialisation);
//   * the proposal-seeded initialisation of the learned path gathers its
//     candidates in groups and carries the random search that follows it;
//   * the down-sampling chain is one kernel that evaluates the separable
//     Gaussian only at the nearest-neighbour sample points, from an LDS copy
//     of each output tile's field window;
//   * the trilateral filter streams bands of pre-scaled (u, v, I) through LDS
//     and evaluates one exp2 per tap; a variant accumulates its three weighted
//     sums with 4x4x1 FP32 MFMA blocks.
#include <algorithm>

#include "bm_internal.h"
#include "bm_device.h"
#include "bm_median_net.h"

namespace bm {
namespace {

__device__ __align__(16) const int32_t d_ndtri_q16[4096] = {
#include "bm_ndtri_q16.inc"
};

constexpr int kThreads = 256;

__device__ __forceinline__ void philox4x32_10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                              uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = c0 * 0xD2511F53u, hi0 = __umulhi(c0, 0xD2511F53u);
        const uint32_t lo1 = c2 * 0xCD9E8D57u, hi1 = __umulhi(c2, 0xCD9E8D57u);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// round-half-away-from-zero of a Q32 product
__device__ __forceinline__ int round_q32(int64_t v)
{
    const int64_t a = v < 0 ? -v : v;
    const int64_t r = (a + (int64_t(1) << 31)) >> 32;
    return (int)(v < 0 ? -r : r);
}

// Fixed-point bilinear sample (Q16 gray levels) at the packed (x0 << 8 | frac)
// coordinates of one up-sampled column / row.
__device__ __forceinline__ uint32_t bilinear_q16(const uint8_t *__restrict__ g8, int W, int H, uint32_t ux,
                                                 uint32_t uy)
{
    const int x0 = (int)(ux >> 8), y0 = (int)(uy >> 8);
    const uint32_t ax = ux & 255u, ay = uy & 255u;
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, H - 1);
    const uint8_t *r0 = g8 + (size_t)y0 * W;
    const uint8_t *r1 = g8 + (size_t)y1 * W;
    const uint32_t top = r0[x0] * (256u - ax) + r0[x1] * ax;
    const uint32_t bot = r1[x0] * (256u - ax) + r1[x1] * ax;
    return top * (256u - ay) + bot * ay;
}

// ---------------------------------------------------------------------------
// gray: integer BT.601 of interleaved RGB

__device__ __forceinline__ void gray_tile(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ g8,
                                          float *__restrict__ g32, int n, int t)
{
    const int i = t * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rgb[3 * (size_t)i], g = rgb[3 * (size_t)i + 1], b = rgb[3 * (size_t)i + 2];
    const uint32_t y = (77u * r + 150u * g + 29u * b + 128u) >> 8;
    if (g8) g8[i] = (uint8_t)y;
    if (g32) g32[i] = (float)y;
}

__global__ void __launch_bounds__(kThreads) k_gray(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ g8,
                                                   float *__restrict__ g32, int n)
{
    gray_tile(rgb, g8, g32, n, blockIdx.x);
}

// Both frames of a pair in one launch (the first also as FP32): work-groups
// [0, nb) take the first frame, [nb, 2 nb) the second.
__global__ void __launch_bounds__(kThreads) k_gray_pair(const uint8_t *__restrict__ rgb1, uint8_t *__restrict__ g8a,
                                                        float *__restrict__ g32a, const uint8_t *__restrict__ rgb2,
                                                        uint8_t *__restrict__ g8b, int n, int nb)
{
    if ((int)blockIdx.x < nb)
        gray_tile(rgb1, g8a, g32a, n, blockIdx.x);
    else
        gray_tile(rgb2, g8b, nullptr, n, blockIdx.x - nb);
}

// ---------------------------------------------------------------------------
// bilinear up-sampling (stand-alone; the matching path uses the fused BRIEF)

__global__ void __launch_bounds__(kThreads) k_upsample(const uint8_t *__restrict__ g8,
                                                       const uint32_t *__restrict__ up_x,
                                                       const uint32_t *__restrict__ up_y, int W, int H,
                                                       int sx, int sy, float *__restrict__ L)
{
    const int gx = (sx + kThreads - 1) / kThreads;
    const int b = xcd_remap(blockIdx.x, gx * sy);
    const int x = (b % gx) * kThreads + threadIdx.x, y = b / gx;
    if (x >= sx) return;
    L[(size_t)y * sx + x] = (float)bilinear_q16(g8, W, H, up_x[x], up_y[y]) * (1.0f / 65536.0f);
}

// ---------------------------------------------------------------------------
// Lanczos-3 up-sampling (the reference's resizeLanczos, util_cuda.cu:126-158).
// Its 6 x 6 weights are separable: per output column x the first source column
// lx0[x] and six normalised weights lwx[6x..6x+5], likewise per row; the value
// is sum_j wy_j * (sum_i wx_i * g[clamp(y0 + j)][clamp(x0 + i)]) with fused
// multiply-adds in increasing i, j (the host oracle uses the same tables and
// order, so the result is bit-identical).

__global__ void __launch_bounds__(kThreads) k_upsample_lanczos(const uint8_t *__restrict__ g8,
                                                               const int *__restrict__ lx0,
                                                               const float *__restrict__ lwx,
                                                               const int *__restrict__ ly0,
                                                               const float *__restrict__ lwy, int W, int H,
                                                               int sx, int sy, float *__restrict__ L)
{
    const int gx = (sx + kThreads - 1) / kThreads;
    const int b = xcd_remap(blockIdx.x, gx * sy);
    const int x = (b % gx) * kThreads + threadIdx.x, y = b / gx;
    if (x >= sx) return;
    const int x0 = lx0[x], y0 = ly0[y];
    float wx[6];
    int cx[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        wx[i] = lwx[6 * x + i];
        cx[i] = clampi(x0 + i, 0, W - 1);
    }
    float v = 0.0f;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const uint8_t *row = g8 + (size_t)clampi(y0 + j, 0, H - 1) * W;
        float h = 0.0f;
#pragma unroll
        for (int i = 0; i < 6; ++i) h = __fmaf_rn(wx[i], (float)row[cx[i]], h);
        v = __fmaf_rn(lwy[6 * y + j], h, v);
    }
    L[(size_t)y * sx + x] = v;
}

// ---------------------------------------------------------------------------
// BRIEF.  A 256-thread work-group describes a 64 x 16 tile; thread (lx, ly)
// owns pixels (lx, ly + 4k), k = 0..3, so each wave covers one image row per k.
// The tile plus an R-pixel halo sits in LDS (clamped to the image, which is
// exactly the reference's per-axis offset clamp).  Tests whose four offsets all
// lie within R read LDS; the others (a wave-uniform branch, because the pattern
// is shared by all pixels) read the image in global memory.

constexpr int kBriefTW = 64, kBriefTH = 16, kBriefPY = 4;
constexpr int kBriefLdsBytes = 64 * 1024;

template <int NSP, bool FUSED>
__global__ void __launch_bounds__(kThreads) k_brief(const float *__restrict__ L, const uint8_t *__restrict__ g8,
                                                    const uint32_t *__restrict__ up_x,
                                                    const uint32_t *__restrict__ up_y,
                                                    const int16_t *__restrict__ pattern, int N, int W, int H,
                                                    int sx, int sy, int R, uint32_t *__restrict__ F)
{
    extern __shared__ float tile[];
    const int pitch = kBriefTW + 2 * R;
    const int rows = kBriefTH + 2 * R;
    const int gxb = (sx + kBriefTW - 1) / kBriefTW, gyb = (sy + kBriefTH - 1) / kBriefTH;
    const int b = xcd_remap(blockIdx.x, gxb * gyb);
    const int tx0 = (b % gxb) * kBriefTW, ty0 = (b / gxb) * kBriefTH;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;

    for (int r = ly; r < rows; r += 4) {
        const int gy = clampi(ty0 - R + r, 0, sy - 1);
        for (int c = lx; c < pitch; c += 64) {
            const int gx = clampi(tx0 - R + c, 0, sx - 1);
            float v;
            if constexpr (FUSED)
                v = (float)bilinear_q16(g8, W, H, up_x[gx], up_y[gy]) * (1.0f / 65536.0f);
            else
                v = L[(size_t)gy * sx + gx];
            tile[r * pitch + c] = v;
        }
    }
    __syncthreads();

    const int x = tx0 + lx;
    const int base = (ly + R) * pitch + lx + R;
    const uint2 *pat = reinterpret_cast<const uint2 *>(pattern);
    uint32_t words[NSP][kBriefPY];
#pragma unroll
    for (int w = 0; w < NSP; ++w)
#pragma unroll
        for (int k = 0; k < kBriefPY; ++k) words[w][k] = 0u;

#pragma unroll
    for (int w = 0; w < NSP; ++w) {
        const int nb = min(32, N - 32 * w);
        for (int bit = nb - 1; bit >= 0; --bit) {
            const uint2 pv = pat[32 * w + bit];
            const int ox1 = (int16_t)(pv.x & 0xffffu), oy1 = (int16_t)(pv.x >> 16);
            const int ox2 = (int16_t)(pv.y & 0xffffu), oy2 = (int16_t)(pv.y >> 16);
            const bool inside = max(abs(ox1), abs(ox2)) <= R && max(abs(oy1), abs(oy2)) <= R;
            if (inside) {
                const int o1 = base + oy1 * pitch + ox1, o2 = base + oy2 * pitch + ox2;
#pragma unroll
                for (int k = 0; k < kBriefPY; ++k) {
                    const float a = tile[o1 + 4 * k * pitch], c = tile[o2 + 4 * k * pitch];
                    words[w][k] = (words[w][k] << 1) | (a > c ? 1u : 0u);
                }
            } else {
#pragma unroll
                for (int k = 0; k < kBriefPY; ++k) {
                    const int y = ty0 + ly + 4 * k;
                    const int xa = clampi(x + ox1, 0, sx - 1), ya = clampi(y + oy1, 0, sy - 1);
                    const int xb = clampi(x + ox2, 0, sx - 1), yb = clampi(y + oy2, 0, sy - 1);
                    float a, c;
                    if constexpr (FUSED) {
                        a = (float)bilinear_q16(g8, W, H, up_x[xa], up_y[ya]);
                        c = (float)bilinear_q16(g8, W, H, up_x[xb], up_y[yb]);
                    } else {
                        a = L[(size_t)ya * sx + xa];
                        c = L[(size_t)yb * sx + xb];
                    }
                    words[w][k] = (words[w][k] << 1) | (a > c ? 1u : 0u);
                }
            }
        }
    }

    if (x < sx) {
#pragma unroll
        for (int k = 0; k < kBriefPY; ++k) {
            const int y = ty0 + ly + 4 * k;
            if (y < sy) {
                uint32_t out[NSP];
#pragma unroll
                for (int w = 0; w < NSP; ++w) out[w] = words[w][k];
                store_desc<NSP>(F, (uint32_t)(y * sx + x), out);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// BRIEF with the up-sampling fused, everything in LDS.  Per work-group (64 x TH
// matched pixels, TH = 4 PY, thread (lx, ly) owns rows ly + 4k):
//   gt  the input-resolution gray tile covering the tile plus the FAR halo Rf
//       (the largest pattern offset) as horizontal pairs: 16 bits per source
//       pixel x, the gray level at x (low byte) and at its clamped right
//       neighbour min(x + 1, W - 1) (high byte);
//   xt  per matched column of that halo: the bilinear source column (relative
//       to gt) and the 8-bit weight a;
//   yt  per matched row: the byte offsets in gt of the source row and of its
//       clamped lower neighbour (16 bits each), and the weight pair
//       (256 - a, a);
//   lt  the up-sampled tile plus a NEAR halo RN <= Rf, FP32, computed from gt.
// Tests whose offsets lie within RN read lt (two LDS loads per pixel, the
// per-pixel offsets are immediates because the pitch is a template constant);
// the rest evaluate the bilinear samples from gt.  No test reads global memory.

typedef unsigned short bm_u16x2 __attribute__((ext_vector_type(2)));

// Weight pair (256 - a, a) as two 16-bit lanes.
__device__ __forceinline__ uint32_t weight_pair(uint32_t a) { return (256u - a) | (a << 16); }

// u16 lanes (lo, hi) . (wlo, whi) + 0 on the 16-bit dot-product unit.
__device__ __forceinline__ uint32_t dot_u16x2(uint32_t v, uint32_t w)
{
    return __builtin_amdgcn_udot2(__builtin_bit_cast(bm_u16x2, v), __builtin_bit_cast(bm_u16x2, w), 0u, false);
}

// Bilinear sample in 16.16 fixed point (exact: sum of g_ij wx_j wy_i, below
// 2^24) from the pair tile at column x0 with weight pair wx, source rows
// {upper, lower} offsets and weight pair yv: two 16-bit loads, the byte pairs
// spread to 16-bit lanes and weighted horizontally, then the two row values
// (each below 2^16) weighted vertically, three dot products in all.
__device__ __forceinline__ uint32_t bil_lds(const uint16_t *gt, int x0, uint32_t wx, uint2 yv)
{
    const unsigned char *p = reinterpret_cast<const unsigned char *>(gt + x0);
    const uint32_t t = *reinterpret_cast<const uint16_t *>(p + (yv.x & 0xffffu));
    const uint32_t b = *reinterpret_cast<const uint16_t *>(p + (yv.x >> 16));
    const uint32_t top = dot_u16x2(__builtin_amdgcn_perm(0u, t, 0x0c010c00u), wx);
    const uint32_t bot = dot_u16x2(__builtin_amdgcn_perm(0u, b, 0x0c010c00u), wx);
    return dot_u16x2(top | (bot << 16), yv.y);
}

// A raw buffer resource over n bytes from p (dword 3 as CDNA takes it: 32-bit
// data format, no swizzle, no stride): a load through it takes the base from
// SGPRs and an unsigned 32-bit offset from one VGPR, and reads 0 at or past n.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(const float *p, uint32_t n)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), 0, (int)n, 0x00020000);
}

// Bytes of an M-float frame from float offset o on (o may be negative: a test
// offset up and left), saturated to the 32-bit range of a buffer resource;
// the largest byte offset a load takes, 4 (M - 1), stays below it.
__device__ __forceinline__ uint32_t frame_bytes_from(int M, int o)
{
    const uint64_t n = ((uint64_t)(int64_t)M - (uint64_t)(int64_t)o) * 4u;
    return n > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)n;
}

// One 64 x TH tile b (row-major tile index) of the fused BRIEF; smem is the
// work-group's dynamic LDS.  GL: the up-sampled frame is in global memory
// (a.L, written by the up-sampler before): the tile and its near halo are
// copied from it and the far tests read it directly, so the work-group needs
// neither the pair tile nor the coordinate tables and no test interpolates.

template <int NSP, int RN, int PY, bool GL, int G = 4>
__device__ __forceinline__ void brief_tile(const BriefJob &a, int b, unsigned char *smem)
{
    // G waves per work-group, wave ly taking rows ly + G k of the 64 x TH tile
    constexpr int TH = G * PY, PITCH = kBriefTW + 2 * RN, ROWS = TH + 2 * RN, NT = 64 * G;
    float *lt = reinterpret_cast<float *>(smem);
    const int sx = a.sx, sy = a.sy;
    const int gxb = (sx + kBriefTW - 1) / kBriefTW;
    const int tx0 = (b % gxb) * kBriefTW, ty0 = (b / gxb) * TH;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const uint32_t *xrow = nullptr;
    const uint2 *ycol = nullptr;
    const uint16_t *gt = nullptr;
    if constexpr (GL) {
        // the tile and its halo, clamped to the frame per axis (as the reference's
        // per-axis offset clamp): element i of a thread is row ly + G (i / NC),
        // column lx + 64 (i % NC); FB loads in flight before their stores (the
        // copy is latency-bound)
        constexpr int NC = (PITCH + 63) / 64, NR = (ROWS + G - 1) / G, TOT = NC * NR, FB = 16;
        for (int i0 = 0; i0 < TOT; i0 += FB) {
            float v[FB];
#pragma unroll
            for (int q = 0; q < FB; ++q) {
                const int i = min(i0 + q, TOT - 1), r = ly + G * (i / NC), c = lx + 64 * (i % NC);
                v[q] = a.L[(size_t)clampi(ty0 - RN + r, 0, sy - 1) * sx + clampi(tx0 - RN + c, 0, sx - 1)];
            }
#pragma unroll
            for (int q = 0; q < FB; ++q) {
                const int i = i0 + q, r = ly + G * (i / NC), c = lx + 64 * (i % NC);
                if (i < TOT && r < ROWS && c < PITCH) lt[r * PITCH + c] = v[q];
            }
        }
        __syncthreads();
    } else {
        const int Rf = a.Rf, XT = kBriefTW + 2 * Rf, YT = TH + 2 * Rf;
        uint32_t *xt = reinterpret_cast<uint32_t *>(lt + PITCH * ROWS);
        uint2 *yt = reinterpret_cast<uint2 *>(xt + XT + (XT & 1));
        uint16_t *gtw = reinterpret_cast<uint16_t *>(yt + YT);
        const int W = a.W, H = a.H;

        const int gx0 = (int)(a.up_x[clampi(tx0 - Rf, 0, sx - 1)] >> 8);
        const int gy0 = (int)(a.up_y[clampi(ty0 - Rf, 0, sy - 1)] >> 8);
        const int gx1 = min((int)(a.up_x[clampi(tx0 + kBriefTW - 1 + Rf, 0, sx - 1)] >> 8) + 1, W - 1);
        const int gy1 = min((int)(a.up_y[clampi(ty0 + TH - 1 + Rf, 0, sy - 1)] >> 8) + 1, H - 1);
        const int gw = gx1 - gx0 + 1, gh = gy1 - gy0 + 1;

        for (int i = threadIdx.x; i < XT; i += NT) {
            const uint32_t u = a.up_x[clampi(tx0 - Rf + i, 0, sx - 1)];
            xt[i] = (uint32_t)((int)(u >> 8) - gx0) | ((u & 255u) << 16);
        }
        for (int i = threadIdx.x; i < YT; i += NT) {
            const uint32_t u = a.up_y[clampi(ty0 - Rf + i, 0, sy - 1)];
            const int y0 = (int)(u >> 8), y1 = min(y0 + 1, H - 1);
            yt[i] = make_uint2((uint32_t)((y0 - gy0) * gw * 2) | ((uint32_t)((y1 - gy0) * gw * 2) << 16),
                               weight_pair(u & 255u));
        }
        // The pair tile: each gray byte is read from memory once, FB rows at a time
        // (their loads in flight together: the fill is latency-bound, 8 rows
        // measured best against 2, 4, 12 and 16); the right neighbour comes from
        // the next lane, except in lane 63 and past the tile's last column.
        constexpr int FB = 8;
        for (int r0 = ly; r0 < gh; r0 += G * FB) {
            for (int c0 = 0; c0 < gw; c0 += 64) {
                const int c = c0 + lx, x = gx0 + min(c, gw - 1);
                uint32_t v[FB], n[FB];
    #pragma unroll
                for (int q = 0; q < FB; ++q) v[q] = a.g8[(size_t)(gy0 + min(r0 + G * q, gh - 1)) * W + x];
    #pragma unroll
                for (int q = 0; q < FB; ++q) n[q] = (uint32_t)__shfl_down((int)v[q], 1, 64);
                if (lx == 63 || c + 1 >= gw) {
    #pragma unroll
                    for (int q = 0; q < FB; ++q)
                        n[q] = a.g8[(size_t)(gy0 + min(r0 + G * q, gh - 1)) * W + min(x + 1, W - 1)];
                }
    #pragma unroll
                for (int q = 0; q < FB; ++q)
                    if (c < gw && r0 + G * q < gh) gtw[(r0 + G * q) * gw + c] = (uint16_t)(v[q] | (n[q] << 8));
            }
        }
        __syncthreads();
        for (int c = lx; c < PITCH; c += 64) {
            const uint32_t xv = xt[c - RN + Rf];
            const int x0 = (int)(xv & 0xffffu);
            const uint32_t wx = weight_pair(xv >> 16);
            for (int r = ly; r < ROWS; r += G)
                lt[r * PITCH + c] = (float)bil_lds(gtw, x0, wx, yt[r - RN + Rf]) * (1.0f / 65536.0f);
        }
        __syncthreads();
        xrow = xt + lx + Rf;
        ycol = yt + ly + Rf;
        gt = gtw;
    }

    const float *lbase = lt + ly * PITCH + lx;  // the thread's cell in the halo-padded tile, less (RN, RN)
    // GL: a tile whose far halo (a.Rf, the largest pattern offset) lies inside
    // the frame, and the byte offset of the pixel of each of the thread's rows
    // (below 2^32: sx, sy <= 32767), an unsigned 32-bit offset from a scalar
    // base as the global loads take it
    bool interior = false;
    uint32_t pix[PY];
    if constexpr (GL) {
        interior = tx0 - a.Rf >= 0 && tx0 + kBriefTW - 1 + a.Rf <= sx - 1 && ty0 - a.Rf >= 0 &&
                   ty0 + TH - 1 + a.Rf <= sy - 1;
#pragma unroll
        for (int k = 0; k < PY; ++k) pix[k] = (uint32_t)((ty0 + ly + G * k) * sx + tx0 + lx) * 4u;
    }
    uint32_t words[NSP][PY];
#pragma unroll
    for (int w = 0; w < NSP; ++w)
#pragma unroll
        for (int k = 0; k < PY; ++k) words[w][k] = 0u;

    // The plan holds, per word, 32 near records {o1, o2} (int2) for bits 31 ..
    // 0: the LDS offsets ((RN + oy) * PITCH + RN + ox) of the two samples from
    // lbase (so a sample's address is lbase plus one scalar, and the rows k
    // are immediate offsets of 4 k PITCH floats), both 0 for a
    // bit beyond the descriptor length or a far test; after all words, up to
    // 32 far records per word {packed int16 offsets, bit} (uint4).  A near bit
    // is the sign of b - a on the FP32 bit patterns (samples are non-negative:
    // gray levels below 256, such floats order like their bit patterns),
    // shifted in at the bottom: word = word << 1 | sign = alignbit(word, b - a,
    // 31); after the 32nd record every bit sits at its position, the (0, 0)
    // records having shifted in zeros.  The far records then OR their bits into
    // place.  The near records are taken STEP at a time (one scalar load), so
    // each pass of the loop has 2 * STEP * PY samples in flight from LDS:
    // 4 records per pass for 16-row tiles, 2 for 32-row tiles (measured best).
    constexpr int STEP = PY >= 8 ? 2 : 4;
    const int2 *near_rec = reinterpret_cast<const int2 *>(a.plan);
    const uint4 *far_rec = reinterpret_cast<const uint4 *>(a.plan + 64 * NSP);
#pragma unroll
    for (int w = 0; w < NSP; ++w) {
#pragma unroll 1
        for (int i = 0; i < 32; i += STEP) {
            int2 r[STEP];
#pragma unroll
            for (int j = 0; j < STEP; ++j) r[j] = near_rec[32 * w + i + j];
#pragma unroll
            for (int j = 0; j < STEP; j += 2) {
                const uint32_t *p1 = reinterpret_cast<const uint32_t *>(lbase + r[j].x);
                const uint32_t *p2 = reinterpret_cast<const uint32_t *>(lbase + r[j].y);
                const uint32_t *p3 = reinterpret_cast<const uint32_t *>(lbase + r[j + 1].x);
                const uint32_t *p4 = reinterpret_cast<const uint32_t *>(lbase + r[j + 1].y);
#pragma unroll
                for (int k = 0; k < PY; ++k) {
                    const uint32_t d0 = p2[G * k * PITCH] - p1[G * k * PITCH];
                    const uint32_t d1 = p4[G * k * PITCH] - p3[G * k * PITCH];
                    words[w][k] = __builtin_amdgcn_alignbit(__builtin_amdgcn_alignbit(words[w][k], d0, 31), d1, 31);
                }
            }
        }
        const int nfar = a.plan_count[2 * w + 1];
        if constexpr (GL) {
            // a far sample is one load of the up-sampled frame; its FP32 bit
            // pattern orders like the value (non-negative samples).  A tile
            // whose far halo lies inside the frame needs no clamp: the sample
            // of row k is the pixel's own index plus the test's offset
            // oy * sx + ox, a work-group-uniform base, so the loads take their
            // address from a buffer resource over the frame from that offset
            // on (SGPRs) and the row's byte offset (no VALU)
            const int x = tx0 + lx;
            if (interior) {
                for (int f = 0; f < nfar; ++f) {
                    const uint4 r0 = far_rec[32 * w + f];
                    const int ox1 = (int16_t)(r0.x & 0xffffu), oy1 = (int16_t)(r0.x >> 16);
                    const int ox2 = (int16_t)(r0.y & 0xffffu), oy2 = (int16_t)(r0.y >> 16);
                    const int M = sx * sy, o1 = oy1 * sx + ox1, o2 = oy2 * sx + ox2;
                    const __amdgpu_buffer_rsrc_t r1 = frame_rsrc(a.L + o1, frame_bytes_from(M, o1));
                    const __amdgpu_buffer_rsrc_t r2 = frame_rsrc(a.L + o2, frame_bytes_from(M, o2));
                    uint32_t va[PY], vb[PY];
#pragma unroll
                    for (int k = 0; k < PY; ++k) {
                        va[k] = __builtin_amdgcn_raw_buffer_load_b32(r1, (int)pix[k], 0, 0);
                        vb[k] = __builtin_amdgcn_raw_buffer_load_b32(r2, (int)pix[k], 0, 0);
                    }
#pragma unroll
                    for (int k = 0; k < PY; ++k) words[w][k] |= ((vb[k] - va[k]) >> 31) << r0.z;
                }
                continue;
            }
            for (int f = 0; f < nfar; ++f) {
                const uint4 r0 = far_rec[32 * w + f];
                const int ox1 = (int16_t)(r0.x & 0xffffu), oy1 = (int16_t)(r0.x >> 16);
                const int ox2 = (int16_t)(r0.y & 0xffffu), oy2 = (int16_t)(r0.y >> 16);
                const int xa = clampi(x + ox1, 0, sx - 1), xb = clampi(x + ox2, 0, sx - 1);
                uint32_t va[PY], vb[PY];
#pragma unroll
                for (int k = 0; k < PY; ++k) {
                    const int y = ty0 + ly + G * k;
                    va[k] = __float_as_uint(a.L[(size_t)clampi(y + oy1, 0, sy - 1) * sx + xa]);
                    vb[k] = __float_as_uint(a.L[(size_t)clampi(y + oy2, 0, sy - 1) * sx + xb]);
                }
#pragma unroll
                for (int k = 0; k < PY; ++k) words[w][k] |= ((vb[k] - va[k]) >> 31) << r0.z;
            }
            continue;
        }
        for (int f = 0; f < nfar; ++f) {
            const uint4 r0 = far_rec[32 * w + f];
            const int ox1 = (int16_t)(r0.x & 0xffffu), oy1 = (int16_t)(r0.x >> 16);
            const int ox2 = (int16_t)(r0.y & 0xffffu), oy2 = (int16_t)(r0.y >> 16);
            const uint32_t xa = xrow[ox1], xb = xrow[ox2];
            const int xa0 = (int)(xa & 0xffffu), xb0 = (int)(xb & 0xffffu);
            const uint32_t wxa = weight_pair(xa >> 16), wxb = weight_pair(xb >> 16);
#pragma unroll
            for (int k = 0; k < PY; ++k) {
                const uint32_t va = bil_lds(gt, xa0, wxa, ycol[G * k + oy1]);
                const uint32_t vb = bil_lds(gt, xb0, wxb, ycol[G * k + oy2]);
                words[w][k] |= ((vb - va) >> 31) << r0.z;
            }
        }
    }
    const int x = tx0 + lx;
    if (x < sx) {
#pragma unroll
        for (int k = 0; k < PY; ++k) {
            const int y = ty0 + ly + G * k;
            if (y < sy) {
                uint32_t out[NSP];
#pragma unroll
                for (int w = 0; w < NSP; ++w) out[w] = words[w][k];
                store_desc<NSP>(a.F, (uint32_t)(y * sx + x), out);
            }
        }
    }
}

template <int NSP, int RN, int PY, bool GL, int G = 4>
__global__ void __launch_bounds__(64 * G) k_brief_lds(BriefJob a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TH = G * PY;
    const int gxb = (a.sx + kBriefTW - 1) / kBriefTW, gyb = (a.sy + TH - 1) / TH;
    brief_tile<NSP, RN, PY, GL, G>(a, xcd_remap(blockIdx.x, gxb * gyb), smem);
}

// The fused kernel over both frames of a pair in one launch (their jobs differ
// only in the gray frame and the descriptors): for frames whose tiles alone
// leave most of the device idle.
template <int NSP, int RN, int PY, int G>
__global__ void __launch_bounds__(64 * G) k_brief_lds_pair(BriefJob a, const uint8_t *__restrict__ g8b,
                                                           uint32_t *__restrict__ Fb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TH = G * PY;
    const int gxb = (a.sx + kBriefTW - 1) / kBriefTW, gyb = (a.sy + TH - 1) / TH, tiles = gxb * gyb;
    int b = xcd_remap(blockIdx.x, 2 * tiles);
    if (b >= tiles) {
        a.g8 = g8b;
        a.F = Fb;
        b -= tiles;
    }
    brief_tile<NSP, RN, PY, false, G>(a, b, smem);
}

// ---------------------------------------------------------------------------
// field initialisation

struct RSArgs {
    int enabled;
    uint32_t iteration, pair, key0, key1;
    int64_t sigma_q16;
    const int32_t *ndtri;  // the inverse-normal table as the kernel reads it (set on the device)
};

// The inverse-normal table copied into the work-group's LDS: a random-search
// candidate looks it up at two random indices per pixel, gathers that would
// otherwise take the address units these kernels are bound by.  Every thread
// of the work-group calls it before any returns.
__device__ __forceinline__ const int32_t *ndtri_to_lds(int32_t *t)
{
    const int4 *src = reinterpret_cast<const int4 *>(d_ndtri_q16);
    int4 *dst = reinterpret_cast<int4 *>(t);
    for (int i = threadIdx.x; i < 4096 / 4; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    return t;
}

template <int NSP>
__device__ __forceinline__ void random_candidate(const uint32_t (&f1)[NSP], const uint32_t *__restrict__ F2,
                                                 int idx, int x, int y, int sx, int sy, const RSArgs &rs,
                                                 uint32_t &best, int &bcost)
{
    uint32_t c0 = (uint32_t)idx, c1 = rs.iteration, c2 = rs.pair, c3 = 0u;
    philox4x32_10(c0, c1, c2, c3, rs.key0, rs.key1);
    int nx = round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c0 >> 20]);
    int ny = round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c1 >> 20]);
    nx = clampi(nx, -x, sx - 1 - x);
    ny = clampi(ny, -y, sy - 1 - y);
    const uint32_t cand = pack_off(nx, ny);
    // insert() skips a candidate equal to the current offset (briefmatch_cuda.cu:131-132)
    if (cand != best) {
        uint32_t f2[NSP];
        load_desc<NSP>(F2, idx + ny * sx + nx, f2);
        const int c = hamming<NSP>(f1, f2);
        if (c < bcost) {
            bcost = c;
            best = cand;
        }
    }
}


// RS: the random search of rs.iteration is inserted after the initial
// candidate.  After a zero initialisation that is the whole first iteration:
// every propagation candidate of its levels equals the pixel's own (0, 0) and
// is skipped, so only the random search of its last level changes the field.
// lb: the row-segment block in the strip-major order (before strip_block).
template <int NSP, bool RS>
__device__ __forceinline__ void init_tile(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2, int first,
                                          const uint32_t *__restrict__ prev, uint32_t *__restrict__ off,
                                          uint16_t *__restrict__ cost, int sx, int sy, int strip_rows,
                                          const RSArgs &rs, int lb)
{
    const int gx = (sx + kThreads - 1) / kThreads;
    int bx, y;
    strip_block(lb, gx, sy, strip_rows, bx, y);
    const int x = bx * kThreads + threadIdx.x;
    if (x >= sx) return;
    const int idx = y * sx + x;
    int nx = 0, ny = 0;
    if (!first) {
        const uint32_t o0 = prev[idx];
        const int dx0 = clampi(off_dx(o0), -x, sx - 1 - x), dy0 = clampi(off_dy(o0), -y, sy - 1 - y);
        const uint32_t o1 = prev[idx + dy0 * sx + dx0];
        nx = clampi(off_dx(o1), -x, sx - 1 - x);
        ny = clampi(off_dy(o1), -y, sy - 1 - y);
    }
    uint32_t f1[NSP], f2[NSP];
    load_desc<NSP>(F1, idx, f1);
    load_desc<NSP>(F2, idx + ny * sx + nx, f2);
    uint32_t best = pack_off(nx, ny);
    int bcost = hamming<NSP>(f1, f2);
    if constexpr (RS) random_candidate<NSP>(f1, F2, idx, x, y, sx, sy, rs, best, bcost);
    off[idx] = best;
    cost[idx] = (uint16_t)bcost;
}

// init_tile with L lanes per pixel (as k_level_lanes): lane part reads words
// [NW part, NW part + NW) of the three descriptors (own, initial candidate,
// random candidate), all loads issued first, partial counts summed over the
// pixel's lanes.
template <int NSP, bool RS, int L>
__device__ __forceinline__ void init_tile_lanes(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2,
                                                int first, const uint32_t *__restrict__ prev,
                                                uint32_t *__restrict__ off, uint16_t *__restrict__ cost, int sx,
                                                int sy, int strip_rows, const RSArgs &rs, int lb)
{
    constexpr int NW = NSP / L, NV = NW / 4, span = kThreads / L;
    static_assert(NV >= 1 && NW % 4 == 0, "at least 16 bytes per lane");
    const int gx = (sx + span - 1) / span;
    int bx, y;
    strip_block(lb, gx, sy, strip_rows, bx, y);
    const int part = (int)threadIdx.x % L;
    const int xq = bx * span + (int)threadIdx.x / L;
    const bool live = xq < sx;
    const int x = min(xq, sx - 1), idx = y * sx + x;
    int nx = 0, ny = 0;
    if (!first) {
        const uint32_t o0 = prev[idx];
        const int dx0 = clampi(off_dx(o0), -x, sx - 1 - x), dy0 = clampi(off_dy(o0), -y, sy - 1 - y);
        const uint32_t o1 = prev[idx + dy0 * sx + dx0];
        nx = clampi(off_dx(o1), -x, sx - 1 - x);
        ny = clampi(off_dy(o1), -y, sy - 1 - y);
    }
    const uint32_t init = pack_off(nx, ny);
    uint32_t rcand = init;
    int rtgt = idx;
    if constexpr (RS) {
        uint32_t c0 = (uint32_t)idx, c1 = rs.iteration, c2 = rs.pair, c3 = 0u;
        philox4x32_10(c0, c1, c2, c3, rs.key0, rs.key1);
        const int rx = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c0 >> 20]), -x, sx - 1 - x);
        const int ry = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c1 >> 20]), -y, sy - 1 - y);
        rcand = pack_off(rx, ry);
        rtgt = idx + ry * sx + rx;
    }
    // insert() skips a random candidate equal to the current offset
    const bool rs_need = RS && rcand != init;
    uint4 f1[NV], f2[NV], fr[NV];
    {
        const uint4 *a = reinterpret_cast<const uint4 *>(F1 + (size_t)idx * NSP + NW * part);
        const uint4 *b = reinterpret_cast<const uint4 *>(F2 + (size_t)(idx + ny * sx + nx) * NSP + NW * part);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            f1[v] = a[v];
            f2[v] = b[v];
            fr[v] = f1[v];
        }
        if (rs_need) {
            const uint4 *r = reinterpret_cast<const uint4 *>(F2 + (size_t)rtgt * NSP + NW * part);
#pragma unroll
            for (int v = 0; v < NV; ++v) fr[v] = r[v];
        }
    }
    int d0 = 0, d1 = 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        d0 += __popc(f1[v].x ^ f2[v].x) + __popc(f1[v].y ^ f2[v].y) + __popc(f1[v].z ^ f2[v].z) +
              __popc(f1[v].w ^ f2[v].w);
        d1 += __popc(f1[v].x ^ fr[v].x) + __popc(f1[v].y ^ fr[v].y) + __popc(f1[v].z ^ fr[v].z) +
              __popc(f1[v].w ^ fr[v].w);
    }
#pragma unroll
    for (int m = 1; m < L; m <<= 1) {  // every lane: summed over the pixel's group
        d0 += __shfl_xor(d0, m);
        d1 += __shfl_xor(d1, m);
    }
    uint32_t best = init;
    int bcost = d0;
    if (rs_need && d1 < bcost) {
        bcost = d1;
        best = rcand;
    }
    if (live && part == 0) {
        off[idx] = best;
        cost[idx] = (uint16_t)bcost;
    }
}

template <int NSP, bool RS, int L>
__global__ void __launch_bounds__(kThreads) k_init_lanes(const uint32_t *__restrict__ F1,
                                                         const uint32_t *__restrict__ F2, int first,
                                                         const uint32_t *__restrict__ prev,
                                                         uint32_t *__restrict__ off, uint16_t *__restrict__ cost,
                                                         int sx, int sy, int strip_rows, RSArgs rs)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    if constexpr (RS) rs.ndtri = ndtri_to_lds(nd);
    constexpr int span = kThreads / L;
    const int gx = (sx + span - 1) / span;
    init_tile_lanes<NSP, RS, L>(F1, F2, first, prev, off, cost, sx, sy, strip_rows, rs,
                                xcd_remap(blockIdx.x, gx * sy));
}

template <int NSP, bool RS>
__global__ void __launch_bounds__(kThreads) k_init(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2,
                                                   int first, const uint32_t *__restrict__ prev,
                                                   uint32_t *__restrict__ off, uint16_t *__restrict__ cost,
                                                   int sx, int sy, int strip_rows, RSArgs rs)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    if constexpr (RS) rs.ndtri = ndtri_to_lds(nd);
    const int gx = (sx + kThreads - 1) / kThreads;
    init_tile<NSP, RS>(F1, F2, first, prev, off, cost, sx, sy, strip_rows, rs, xcd_remap(blockIdx.x, gx * sy));
}

// ---------------------------------------------------------------------------
// one jump-flooding level (+ optionally the random search of the iteration)

// PX pixels per thread (256 apart, so every access of a wave stays one
// coalesced row segment); a tuning choice, results do not depend on it.
template <int NSP, bool RS, int PX>
__device__ __forceinline__ void level_tile(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2,
                                           const uint32_t *__restrict__ off_in, const uint16_t *__restrict__ cost_in,
                                           uint32_t *__restrict__ off_out, uint16_t *__restrict__ cost_out, int sx,
                                           int sy, int p, int pd, int double_prop, int strip_rows, const RSArgs &rs,
                                           int lb)
{
    constexpr int span = kThreads * PX;
    const int gx = (sx + span - 1) / span;
    int bx, y;
    strip_block(lb, gx, sy, strip_rows, bx, y);
    const int M = sx * sy;
    const int sv = pd * p * sx, sh = pd * p;

    int xq[PX], idx[PX];
    bool live[PX];
    uint32_t own[PX];
    int own_cost[PX];
    uint32_t o[PX][4];
    bool valid[PX][4];
    // round 1: own entry and the four neighbours' offsets of the snapshot
#pragma unroll
    for (int q = 0; q < PX; ++q) {
        xq[q] = bx * span + q * kThreads + threadIdx.x;
        live[q] = xq[q] < sx;
        idx[q] = y * sx + min(xq[q], sx - 1);
        own[q] = live[q] ? off_in[idx[q]] : 0u;
        own_cost[q] = live[q] ? (int)cost_in[idx[q]] : 0;
        // candidate order of the reference: d = -1 (vertical, horizontal), d = +1 (...)
        const int nidx[4] = {idx[q] - sv, idx[q] - sh, idx[q] + sv, idx[q] + sh};
#pragma unroll
        for (int c = 0; c < 4; ++c)
            valid[q][c] = live[q] && (c >= 2 || double_prop) && (unsigned)nidx[c] < (unsigned)M;
        // a horizontal neighbour less than a wave away in the same row is another
        // lane's own entry: taken from that lane (through LDS, idle here) instead
        // of a load, which relieves the address units that bound this kernel
        bool lane_h[4] = {false, false, false, false};
        if (p < 64) {  // uniform
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int c = 1; c < 4; c += 2) {
                const int dl = c == 1 ? -sh : sh, src = lane + dl, xs = xq[q] + dl;
                const uint32_t v = __shfl(own[q], src & 63);
                lane_h[c] = valid[q][c] && src >= 0 && src < 64 && xs >= 0 && xs < sx;
                o[q][c] = v;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (!lane_h[c]) o[q][c] = valid[q][c] ? off_in[nidx[c]] : 0u;
    }
    // the random-search candidate depends only on (pixel, iteration, pair)
    uint32_t rcand[PX];
    int rtgt[PX];
    if constexpr (RS) {
#pragma unroll
        for (int q = 0; q < PX; ++q) {
            uint32_t c0 = (uint32_t)idx[q], c1 = rs.iteration, c2 = rs.pair, c3 = 0u;
            philox4x32_10(c0, c1, c2, c3, rs.key0, rs.key1);
            const int x = min(xq[q], sx - 1);
            const int nx = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c0 >> 20]), -x, sx - 1 - x);
            const int ny = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c1 >> 20]), -y, sy - 1 - y);
            rcand[q] = pack_off(nx, ny);
            rtgt[q] = idx[q] + ny * sx + nx;
        }
    }
    uint32_t cand[PX][4];
    int tgt[PX][4];
    bool need[PX][4], any[PX];
#pragma unroll
    for (int q = 0; q < PX; ++q) {
        const int x = min(xq[q], sx - 1);
        any[q] = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int nx = clampi(off_dx(o[q][c]), -x, sx - 1 - x);
            const int ny = clampi(off_dy(o[q][c]), -y, sy - 1 - y);
            cand[q][c] = pack_off(nx, ny);
            tgt[q][c] = idx[q] + ny * sx + nx;
            // a candidate equal to the pixel's own offset or to an earlier candidate
            // has a known cost: no gather (insert() likewise skips equal offsets)
            bool fresh = valid[q][c] && cand[q][c] != own[q];
#pragma unroll
            for (int e = 0; e < c; ++e) fresh = fresh && !(valid[q][e] && cand[q][e] == cand[q][c]);
            need[q][c] = fresh;
            any[q] = any[q] || fresh;
        }
    }
    // round 2: own descriptor, candidate gathers and the random-search gather,
    // all issued before any cost is formed
    uint32_t f1[PX][NSP], f2[PX][4][NSP], fr[PX][NSP];
    bool rs_need[PX];
#pragma unroll
    for (int q = 0; q < PX; ++q) {
        rs_need[q] = RS && live[q] && rcand[q] != own[q];
        if (any[q] || rs_need[q]) load_desc<NSP>(F1, (uint32_t)idx[q], f1[q]);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (need[q][c]) load_desc<NSP>(F2, (uint32_t)tgt[q][c], f2[q][c]);
        if (rs_need[q]) load_desc<NSP>(F2, (uint32_t)rtgt[q], fr[q]);
    }
#pragma unroll
    for (int q = 0; q < PX; ++q) {
        int cst[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            cst[c] = 0x7fffffff;
            if (need[q][c]) {
                cst[c] = hamming<NSP>(f1[q], f2[q][c]);
            } else if (valid[q][c]) {
                cst[c] = own_cost[q];
#pragma unroll
                for (int e = c - 1; e >= 0; --e)
                    if (valid[q][e] && cand[q][e] == cand[q][c]) cst[c] = cst[e];
            }
        }
        const int rcost = rs_need[q] ? hamming<NSP>(f1[q], fr[q]) : 0x7fffffff;
        // sequential insert() semantics: skip a candidate equal to the running
        // best, accept a strictly lower cost; the random candidate comes last
        uint32_t best = own[q];
        int bcost = own_cost[q];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (valid[q][c] && cand[q][c] != best && cst[c] < bcost) {
                bcost = cst[c];
                best = cand[q][c];
            }
        // equal to own: skipped when own is still the best, not cheaper otherwise
        if (rs_need[q] && rcand[q] != best && rcost < bcost) {
            bcost = rcost;
            best = rcand[q];
        }
        if (live[q]) {
            off_out[idx[q]] = best;
            cost_out[idx[q]] = (uint16_t)bcost;
        }
    }
}

// The same level with L lanes per pixel for wide descriptors (at least 16
// bytes per lane): lane part of a pixel's group reads words
// [NW part, NW part + NW) of each descriptor (NW = NSP / L), so one load
// instruction reads 64 / L descriptors in whole 16-byte pieces, and the
// Hamming distance is the sum of the group's partial counts.  The pixel's
// selection logic runs on each lane of its group (identical inputs, identical
// results); lane part 0 writes.
// The value of the other lane of an even / odd lane pair (DPP quad_perm
// [1, 0, 3, 2]: one VALU move, no LDS); every lane of the wave is active
// where it is used.
__device__ __forceinline__ uint32_t lane_pair_swap(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

template <int NSP, bool RS, int L>
__device__ __forceinline__ void level_tile_lanes(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2,
                                                 const uint32_t *__restrict__ off_in,
                                                 const uint16_t *__restrict__ cost_in, uint32_t *__restrict__ off_out,
                                                 uint16_t *__restrict__ cost_out, int sx, int sy, int p, int pd,
                                                 int double_prop, int strip_rows, const RSArgs &rs, int lb)
{
    constexpr int NW = NSP / L, NV = NW / 4, span = kThreads / L, PW = 64 / L;  // per lane: words, 16-byte loads
    static_assert(NV >= 1 && NW % 4 == 0, "at least 16 bytes per lane");
    const int gx = (sx + span - 1) / span;
    int bx, y;
    strip_block(lb, gx, sy, strip_rows, bx, y);
    const int M = sx * sy;
    const int sv = pd * p * sx, sh = pd * p;
    const int part = (int)threadIdx.x % L;
    const int xq = bx * span + (int)threadIdx.x / L;
    const bool live = xq < sx;
    const int idx = y * sx + min(xq, sx - 1);
    const int nidx[4] = {idx - sv, idx - sh, idx + sv, idx + sh};
    bool valid[4];
    uint32_t o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) valid[c] = live && (c >= 2 || double_prop) && (unsigned)nidx[c] < (unsigned)M;
    uint32_t own;
    int own_cost;
    if constexpr (L == 2) {
        // the pixel's field words are read once per pixel, not once per lane:
        // lane part 0 reads its own offset and cost, part 1 the upper
        // neighbour's offset, and the pair swap (one load instruction for two
        // words; the loads bound this kernel's address units)
        const uint32_t v = part == 0 ? (live ? off_in[idx] : 0u) : (valid[0] ? off_in[nidx[0]] : 0u);
        const int cst = part == 0 && live ? (int)cost_in[idx] : 0;
        const uint32_t w = lane_pair_swap(v);
        own = part == 0 ? v : w;
        o[0] = part == 0 ? w : v;
        own_cost = part == 0 ? cst : (int)lane_pair_swap((uint32_t)cst);
    } else {
        own = live ? off_in[idx] : 0u;
        own_cost = live ? (int)cost_in[idx] : 0;
    }
    bool lane_h[4] = {false, false, false, false};
    if (p < PW) {  // uniform: a horizontal neighbour in the same wave is another group's own entry
        const int pl = ((int)threadIdx.x & 63) / L;
#pragma unroll
        for (int c = 1; c < 4; c += 2) {
            const int dl = c == 1 ? -sh : sh, src = pl + dl, xs = xq + dl;
            const uint32_t v = __shfl(own, ((src & (PW - 1)) * L + part) & 63);
            lane_h[c] = valid[c] && src >= 0 && src < PW && xs >= 0 && xs < sx;
            o[c] = v;
        }
    }
    if constexpr (L == 2) {
        // the other three neighbours the same way: part 0 the left and right,
        // part 1 the lower one (both lanes of a pixel agree on lane_h)
        const bool n1 = !lane_h[1] && valid[1], n3 = !lane_h[3] && valid[3];
        const uint32_t a = part == 0 ? (n1 ? off_in[nidx[1]] : 0u) : (valid[2] ? off_in[nidx[2]] : 0u);
        const uint32_t b = part == 0 && n3 ? off_in[nidx[3]] : 0u;
        const uint32_t a2 = lane_pair_swap(a), b2 = lane_pair_swap(b);
        if (!lane_h[1]) o[1] = part == 0 ? a : a2;
        o[2] = part == 0 ? a2 : a;
        if (!lane_h[3]) o[3] = part == 0 ? b : b2;
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (!lane_h[c]) o[c] = valid[c] ? off_in[nidx[c]] : 0u;
    }
    uint32_t rcand = 0u;
    int rtgt = 0;
    if constexpr (RS) {
        uint32_t c0 = (uint32_t)idx, c1 = rs.iteration, c2 = rs.pair, c3 = 0u;
        philox4x32_10(c0, c1, c2, c3, rs.key0, rs.key1);
        const int x = min(xq, sx - 1);
        const int nx = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c0 >> 20]), -x, sx - 1 - x);
        const int ny = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c1 >> 20]), -y, sy - 1 - y);
        rcand = pack_off(nx, ny);
        rtgt = idx + ny * sx + nx;
    }
    uint32_t cand[4];
    int tgt[4];
    bool need[4], any = false;
    {
        const int x = min(xq, sx - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int nx = clampi(off_dx(o[c]), -x, sx - 1 - x);
            const int ny = clampi(off_dy(o[c]), -y, sy - 1 - y);
            cand[c] = pack_off(nx, ny);
            tgt[c] = idx + ny * sx + nx;
            bool fresh = valid[c] && cand[c] != own;
#pragma unroll
            for (int e = 0; e < c; ++e) fresh = fresh && !(valid[e] && cand[e] == cand[c]);
            need[c] = fresh;
            any = any || fresh;
        }
    }
    const bool rs_need = RS && live && rcand != own;
    // round 2: every load issued before any count; a slot not loaded holds the
    // own descriptor's part, which counts 0
    uint4 f1[NV], f2[4][NV], fr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) f1[v] = make_uint4(0u, 0u, 0u, 0u);
    if (any || rs_need) {
        const uint4 *src = reinterpret_cast<const uint4 *>(F1 + (size_t)idx * NSP + NW * part);
#pragma unroll
        for (int v = 0; v < NV; ++v) f1[v] = src[v];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int v = 0; v < NV; ++v) f2[c][v] = f1[v];
        if (need[c]) {
            const uint4 *src = reinterpret_cast<const uint4 *>(F2 + (size_t)tgt[c] * NSP + NW * part);
#pragma unroll
            for (int v = 0; v < NV; ++v) f2[c][v] = src[v];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) fr[v] = f1[v];
    if (rs_need) {
        const uint4 *src = reinterpret_cast<const uint4 *>(F2 + (size_t)rtgt * NSP + NW * part);
#pragma unroll
        for (int v = 0; v < NV; ++v) fr[v] = src[v];
    }
    auto partial = [&](const uint4 (&b)[NV]) {
        int d = 0;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            d += __popc(f1[v].x ^ b[v].x) + __popc(f1[v].y ^ b[v].y) + __popc(f1[v].z ^ b[v].z) +
                 __popc(f1[v].w ^ b[v].w);
#pragma unroll
        for (int m = 1; m < L; m <<= 1) d += __shfl_xor(d, m);  // every lane, summed over the group
        return d;
    };
    int cst[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int d = partial(f2[c]);
        cst[c] = 0x7fffffff;
        if (need[c]) {
            cst[c] = d;
        } else if (valid[c]) {
            cst[c] = own_cost;
#pragma unroll
            for (int e = c - 1; e >= 0; --e)
                if (valid[e] && cand[e] == cand[c]) cst[c] = cst[e];
        }
    }
    const int dr = partial(fr);
    const int rcost = rs_need ? dr : 0x7fffffff;
    uint32_t best = own;
    int bcost = own_cost;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (valid[c] && cand[c] != best && cst[c] < bcost) {
            bcost = cst[c];
            best = cand[c];
        }
    if (rs_need && rcand != best && rcost < bcost) {
        bcost = rcost;
        best = rcand;
    }
    if (live && part == 0) {
        off_out[idx] = best;
        cost_out[idx] = (uint16_t)bcost;
    }
}

template <int NSP, bool RS, int L>
__global__ void __launch_bounds__(kThreads) k_level_lanes(const uint32_t *__restrict__ F1,
                                                          const uint32_t *__restrict__ F2,
                                                          const uint32_t *__restrict__ off_in,
                                                          const uint16_t *__restrict__ cost_in,
                                                          uint32_t *__restrict__ off_out,
                                                          uint16_t *__restrict__ cost_out, int sx, int sy, int p,
                                                          int pd, int double_prop, int strip_rows, RSArgs rs)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    if constexpr (RS) rs.ndtri = ndtri_to_lds(nd);
    constexpr int span = kThreads / L;
    const int gx = (sx + span - 1) / span;
    level_tile_lanes<NSP, RS, L>(F1, F2, off_in, cost_in, off_out, cost_out, sx, sy, p, pd, double_prop, strip_rows,
                                 rs, xcd_remap(blockIdx.x, gx * sy));
}

template <int NSP, bool RS, int PX>
__global__ void __launch_bounds__(kThreads) k_level(const uint32_t *__restrict__ F1, const uint32_t *__restrict__ F2,
                                                    const uint32_t *__restrict__ off_in,
                                                    const uint16_t *__restrict__ cost_in,
                                                    uint32_t *__restrict__ off_out,
                                                    uint16_t *__restrict__ cost_out, int sx, int sy, int p,
                                                    int pd, int double_prop, int strip_rows, RSArgs rs)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    if constexpr (RS) rs.ndtri = ndtri_to_lds(nd);
    const int gx = (sx + kThreads * PX - 1) / (kThreads * PX);
    level_tile<NSP, RS, PX>(F1, F2, off_in, cost_in, off_out, cost_out, sx, sy, p, pd, double_prop, strip_rows, rs,
                            xcd_remap(blockIdx.x, gx * sy));
}

template <int NSP>
__global__ void __launch_bounds__(kThreads) k_random_search(const uint32_t *__restrict__ F1,
                                                            const uint32_t *__restrict__ F2,
                                                            uint32_t *__restrict__ off,
                                                            uint16_t *__restrict__ cost, int sx, int sy,
                                                            int strip_rows, RSArgs rs)
{
    __shared__ __align__(16) int32_t nd[4096];
    rs.ndtri = ndtri_to_lds(nd);
    const int gx = (sx + kThreads - 1) / kThreads;
    int bx, y;
    strip_block(xcd_remap(blockIdx.x, gx * sy), gx, sy, strip_rows, bx, y);
    const int x = bx * kThreads + threadIdx.x;
    if (x >= sx) return;
    const int idx = y * sx + x;
    uint32_t f1[NSP];
    load_desc<NSP>(F1, idx, f1);
    uint32_t best = off[idx];
    int bcost = cost[idx];
    random_candidate<NSP>(f1, F2, idx, x, y, sx, sy, rs, best, bcost);
    off[idx] = best;
    cost[idx] = (uint16_t)bcost;
}

// Proposal-seeded initialisation.  The plain candidate (zero offset, or the
// temporal chain through prev) is evaluated first; then, for the network's
// proposal and then the cost-volume argmin, the bilinear interpolation of the
// four surrounding coarse cells' offsets and the four cells' own offsets.  A
// candidate is skipped when it points outside the frame or equals the running
// best, and taken only when strictly cheaper.  coff [cell][4] holds the network
// (u, v) and argmin (u, v) offsets in Q8 matched pixels.  With RS the random-
// search candidate of rs.iteration is inserted last, as k_random_search would.
template <int NSP, bool RS>
__global__ void __launch_bounds__(kThreads) k_init_proposal(const uint32_t *__restrict__ F1,
                                                            const uint32_t *__restrict__ F2, int first,
                                                            const uint32_t *__restrict__ prev,
                                                            const int32_t *__restrict__ coff, int cw, int ch, int W,
                                                            int H, int sx, int sy, int strip_rows, RSArgs rs,
                                                            uint32_t *__restrict__ off, uint16_t *__restrict__ cost)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    if constexpr (RS) rs.ndtri = ndtri_to_lds(nd);
    // row segments of 256 pixels, strip-major and XCD-contiguous, so the
    // candidate gathers of the segments sharing an L2 overlap
    const int gx = (sx + kThreads - 1) / kThreads;
    int bx, y;
    strip_block(xcd_remap(blockIdx.x, gx * sy), gx, sy, strip_rows, bx, y);
    const int x = bx * kThreads + threadIdx.x;
    if (x >= sx) return;
    const int i = y * sx + x;
    int nx = 0, ny = 0;
    if (!first) {
        const uint32_t o0 = prev[i];
        const int dx0 = clampi(off_dx(o0), -x, sx - 1 - x), dy0 = clampi(off_dy(o0), -y, sy - 1 - y);
        const uint32_t o1 = prev[i + dy0 * sx + dx0];
        nx = clampi(off_dx(o1), -x, sx - 1 - x);
        ny = clampi(off_dy(o1), -y, sy - 1 - y);
    }
    const int plain = i + ny * sx + nx;
    uint32_t f1[NSP], f0[NSP], fr[NSP];
    load_desc<NSP>(F1, (uint32_t)i, f1);
    load_desc<NSP>(F2, (uint32_t)plain, f0);
    // the random-search candidate depends only on (pixel, iteration, pair)
    uint32_t rcand = 0u;
    if constexpr (RS) {
        uint32_t c0 = (uint32_t)i, c1 = rs.iteration, c2 = rs.pair, c3 = 0u;
        philox4x32_10(c0, c1, c2, c3, rs.key0, rs.key1);
        const int rx = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c0 >> 20]), -x, sx - 1 - x);
        const int ry = clampi(round_q32(rs.sigma_q16 * (int64_t)rs.ndtri[c1 >> 20]), -y, sy - 1 - y);
        rcand = pack_off(rx, ry);
        load_desc<NSP>(F2, (uint32_t)(i + ry * sx + rx), fr);
    }

    // coarse coordinate in Q8: cell c's centre is input pixel 8c + 3.5
    const int64_t qx = min(max((int64_t)x * W * 32 / sx - 112, (int64_t)0), (int64_t)(cw - 1) * 256);
    const int64_t qy = min(max((int64_t)y * H * 32 / sy - 112, (int64_t)0), (int64_t)(ch - 1) * 256);
    const int x0 = (int)(qx >> 8), y0 = (int)(qy >> 8), x1 = min(x0 + 1, cw - 1), y1 = min(y0 + 1, ch - 1);
    const int64_t ax = qx & 255, ay = qy & 255;
    const int c00 = (y0 * cw + x0) * 4, c01 = (y0 * cw + x1) * 4, c10 = (y1 * cw + x0) * 4, c11 = (y1 * cw + x1) * 4;
    int cand[10][2];
#pragma unroll
    for (int src = 0; src < 2; ++src)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int e = 2 * src + d;
            const int64_t o00 = coff[c00 + e], o01 = coff[c01 + e], o10 = coff[c10 + e], o11 = coff[c11 + e];
            const int64_t v = (o00 * (256 - ax) + o01 * ax) * (256 - ay) + (o10 * (256 - ax) + o11 * ax) * ay;
            cand[5 * src][d] = (int)((v + (1 << 23)) >> 24);
            cand[5 * src + 1][d] = (int)((o00 + 128) >> 8);
            cand[5 * src + 2][d] = (int)((o01 + 128) >> 8);
            cand[5 * src + 3][d] = (int)((o10 + 128) >> 8);
            cand[5 * src + 4][d] = (int)((o11 + 128) >> 8);
        }
    // a candidate equal to the plain one or to an earlier candidate has the same
    // cost and cannot be strictly cheaper than the running best: its cost is not
    // needed.  Every candidate's gather is still issued (a skipped one re-reads
    // the plain candidate's descriptor, always in the frame), so that the loads
    // of a group are in flight together instead of one round trip each.
    uint32_t bo = pack_off(nx, ny);
    uint32_t seen[10];
    bool ok[10];
    int tgt[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const int cx = cand[k][0], cy = cand[k][1];
        seen[k] = pack_off(cx, cy);
        bool fresh = x + cx >= 0 && x + cx < sx && y + cy >= 0 && y + cy < sy && seen[k] != bo;
#pragma unroll
        for (int e = 0; e < k; ++e) fresh = fresh && !(ok[e] && seen[e] == seen[k]);
        ok[k] = fresh;
        tgt[k] = fresh ? i + cy * sx + cx : plain;
    }
    int best = hamming<NSP>(f1, f0);
    // groups of G gathers (10 for 128-bit descriptors: 40 registers)
    constexpr int G = NSP <= 4 ? 10 : (NSP <= 8 ? 5 : 2);
#pragma unroll
    for (int k0 = 0; k0 < 10; k0 += G) {
        uint32_t f2[G][NSP];
#pragma unroll
        for (int g = 0; g < G; ++g) load_desc<NSP>(F2, (uint32_t)tgt[k0 + g], f2[g]);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = hamming<NSP>(f1, f2[g]);
            if (ok[k0 + g] && c < best) {
                best = c;
                bo = seen[k0 + g];
            }
        }
    }
    if constexpr (RS) {
        const int c = hamming<NSP>(f1, fr);
        if (rcand != bo && c < best) {
            best = c;
            bo = rcand;
        }
    }
    off[i] = bo;
    cost[i] = (uint16_t)best;
}

// Local search around each pixel's current offset (the learned path's
// refinement of its seeds; the reference's random search samples around zero
// instead).  Sample s = 0 .. samples - 1 draws a displacement uniformly from
// the square [-r, r]^2, r = max(radius >> s, 1): Philox4x32-10(ctr = {pixel,
// iteration, pair, 1 + s}, key = seed) words 0 and 1 give d = floor(word *
// (2r + 1) / 2^32) - r (counter word 3 is 0 in the random search, so the two
// streams never share a counter).  The displacement is added to the current
// offset, clamped to the frame like a random-search candidate and inserted
// (skip a repeat of the current offset, accept a strictly lower cost), so
// each sample searches around the best offset so far.  A pixel reads and
// writes only its own entry: the field is updated in place.
template <int NSP>
__global__ void __launch_bounds__(kThreads) k_local_search(const uint32_t *__restrict__ F1,
                                                           const uint32_t *__restrict__ F2, int sx, int sy,
                                                           int strip_rows, int radius, int samples, uint32_t iteration,
                                                           uint32_t pair, uint32_t key0, uint32_t key1,
                                                           uint32_t *__restrict__ off, uint16_t *__restrict__ cost)
{
    const int gx = (sx + kThreads - 1) / kThreads;
    int bx, y;
    strip_block(xcd_remap(blockIdx.x, gx * sy), gx, sy, strip_rows, bx, y);
    const int x = bx * kThreads + threadIdx.x;
    if (x >= sx) return;
    const int i = y * sx + x;
    uint32_t best = off[i];
    int bcost = cost[i];
    uint32_t f1[NSP];
    load_desc<NSP>(F1, (uint32_t)i, f1);
    for (int s = 0; s < samples; ++s) {
        const int r = max(radius >> s, 1);
        uint32_t c0 = (uint32_t)i, c1 = iteration, c2 = pair, c3 = 1u + (uint32_t)s;
        philox4x32_10(c0, c1, c2, c3, key0, key1);
        const uint32_t n = 2u * (uint32_t)r + 1u;
        const int dx = (int)__umulhi(c0, n) - r, dy = (int)__umulhi(c1, n) - r;
        const int nx = clampi(off_dx(best) + dx, -x, sx - 1 - x);
        const int ny = clampi(off_dy(best) + dy, -y, sy - 1 - y);
        const uint32_t cand = pack_off(nx, ny);
        if (cand != best) {
            uint32_t f2[NSP];
            load_desc<NSP>(F2, (uint32_t)(i + ny * sx + nx), f2);
            const int c = hamming<NSP>(f1, f2);
            if (c < bcost) {
                bcost = c;
                best = cand;
            }
        }
    }
    off[i] = best;
    cost[i] = (uint16_t)bcost;
}

// ---------------------------------------------------------------------------
// down-sampling: Gaussian (horizontal then vertical, as filterConst twice)
// evaluated only at the nearest-neighbour sample point, then scaled.  Every
// product and sum is rounded on its own (contraction into fused multiply-adds
// is off in these functions), in the
// filter's order (rows j, then taps i within a row), so the result is the
// same float as the separable filter's: the trilateral filter's weights
// depend on flow differences, and a 1-ulp change of a 60-pixel flow moves an
// output whose weight sum sits at the 1e-5 floor by up to 4e-3 pixels.

__device__ __forceinline__ void ds_direct_tile(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nn_x,
                                               const uint32_t *__restrict__ nn_y, const float *__restrict__ gauss,
                                               int P, int sx, int sy, int W, int H, float scx, float scy,
                                               int identity, float *__restrict__ U, float *__restrict__ V, int b)
{
#pragma clang fp contract(off)
    const int gx = (W + kThreads - 1) / kThreads;
    const int x = (b % gx) * kThreads + threadIdx.x, y = b / gx;
    if (x >= W) return;
    const size_t o = (size_t)y * W + x;
    if (identity) {
        const uint32_t v = off[(size_t)y * sx + x];
        U[o] = (float)off_dx(v);
        V[o] = (float)off_dy(v);
        return;
    }
    const int x1 = (int)nn_x[x], y1 = (int)nn_y[y], P2 = P / 2;
    float gu = 0.0f, gv = 0.0f;
    for (int j = 0; j < P; ++j) {
        const uint32_t *row = off + (size_t)clampi(y1 + j - P2, 0, sy - 1) * sx;
        float hu = 0.0f, hv = 0.0f;
        for (int i = 0; i < P; ++i) {
            const uint32_t v = row[clampi(x1 + i - P2, 0, sx - 1)];
            hu += gauss[i] * (float)off_dx(v);
            hv += gauss[i] * (float)off_dy(v);
        }
        gu += gauss[j] * hu;
        gv += gauss[j] * hv;
    }
    U[o] = gu * scx;
    V[o] = gv * scy;
}

__global__ void __launch_bounds__(kThreads) k_downsample(const uint32_t *__restrict__ off,
                                                         const uint32_t *__restrict__ nn_x,
                                                         const uint32_t *__restrict__ nn_y,
                                                         const float *__restrict__ gauss, int P, int sx, int sy,
                                                         int W, int H, float scx, float scy, int identity,
                                                         float *__restrict__ U, float *__restrict__ V)
{
    const int gx = (W + kThreads - 1) / kThreads;
    ds_direct_tile(off, nn_x, nn_y, gauss, P, sx, sy, W, H, scx, scy, identity, U, V, xcd_remap(blockIdx.x, gx * H));
}

// The same down-sampling per output tile of kDsTileW x kDsTileH input pixels:
// the field window the tile's NN samples need (clamped to the frame, as the
// taps are) is staged in LDS once, so the P x P taps of neighbouring outputs,
// whose windows overlap, read LDS instead of memory.  Same products in the
// same order as k_downsample.
__device__ __forceinline__ void ds_tile(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nn_x,
                                        const uint32_t *__restrict__ nn_y, const float *__restrict__ gauss, int P,
                                        int sx, int sy, int W, int H, float scx, float scy, float *__restrict__ U,
                                        float *__restrict__ V, int blk, uint32_t *win)
{
#pragma clang fp contract(off)
    const int gxb = (W + kDsTileW - 1) / kDsTileW;
    const int X0 = (blk % gxb) * kDsTileW, Y0 = (blk / gxb) * kDsTileH;
    const int P2 = P / 2;
    const int cx0 = (int)nn_x[X0] - P2, ry0 = (int)nn_y[Y0] - P2;
    const int cols = (int)nn_x[min(X0 + kDsTileW, W) - 1] + P2 - cx0 + 1;
    const int rows = (int)nn_y[min(Y0 + kDsTileH, H) - 1] + P2 - ry0 + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int r = wave; r < rows; r += kThreads / 64) {
        const uint32_t *row = off + (size_t)clampi(ry0 + r, 0, sy - 1) * sx;
        for (int c = lane; c < cols; c += 64) win[r * cols + c] = row[clampi(cx0 + c, 0, sx - 1)];
    }
    __syncthreads();
    const int x = X0 + (threadIdx.x % kDsTileW), y = Y0 + (threadIdx.x / kDsTileW);
    if (x >= W || y >= H) return;
    const int wx = (int)nn_x[x] - P2 - cx0, wy = (int)nn_y[y] - P2 - ry0;
    float gu = 0.0f, gv = 0.0f;
    for (int j = 0; j < P; ++j) {
        const uint32_t *row = win + (wy + j) * cols + wx;
        float hu = 0.0f, hv = 0.0f;
        for (int i = 0; i < P; ++i) {
            const uint32_t v = row[i];
            hu += gauss[i] * (float)off_dx(v);
            hv += gauss[i] * (float)off_dy(v);
        }
        gu += gauss[j] * hu;
        gv += gauss[j] * hv;
    }
    const size_t o = (size_t)y * W + x;
    U[o] = gu * scx;
    V[o] = gv * scy;
}

__global__ void __launch_bounds__(kThreads) k_downsample_tile(const uint32_t *__restrict__ off,
                                                              const uint32_t *__restrict__ nn_x,
                                                              const uint32_t *__restrict__ nn_y,
                                                              const float *__restrict__ gauss, int P, int sx,
                                                              int sy, int W, int H, float scx, float scy,
                                                              float *__restrict__ U, float *__restrict__ V)
{
    extern __shared__ uint32_t win[];
    const int gxb = (W + kDsTileW - 1) / kDsTileW, gyb = (H + kDsTileH - 1) / kDsTileH;
    ds_tile(off, nn_x, nn_y, gauss, P, sx, sy, W, H, scx, scy, U, V, xcd_remap(blockIdx.x, gxb * gyb), win);
}

// ---------------------------------------------------------------------------
// 1-D median: the fsize window (clamped) in registers, reduced to its middle
// element by a pruned selection network (bm_median_net.h): 186 min/max
// instructions at fsize 23 where a partial selection sort takes 396.

template <int FS, int I>
__device__ __forceinline__ void median_ops(float *v)
{
    if constexpr (I < MedianNet<FS>::net.count) {
        constexpr MedOp o = MedianNet<FS>::net.op[I];
        if constexpr (o.kind == 0) {
            v[o.a] = v[o.b];
        } else {
            const float lo = fminf(v[o.a], v[o.b]), hi = fmaxf(v[o.a], v[o.b]);
            if constexpr ((o.kind & 1) != 0) v[o.a] = lo;
            if constexpr ((o.kind & 2) != 0) v[o.b] = hi;
        }
        median_ops<FS, I + 1>(v);
    }
}

template <int FS>
__device__ __forceinline__ void median_tile(const float *__restrict__ in, float *__restrict__ out, int W, int H,
                                            int dir, int b)
{
    constexpr int KS = FS / 2;
    const int gx = (W + kThreads - 1) / kThreads;
    const int x = (b % gx) * kThreads + threadIdx.x, y = b / gx;
    if (x >= W) return;
    float v[MedianNet<FS>::N];
    if (dir) {
        const float *row = in + (size_t)y * W;
#pragma unroll
        for (int i = 0; i < FS; ++i) v[i] = row[clampi(x - KS + i, 0, W - 1)];
    } else {
#pragma unroll
        for (int i = 0; i < FS; ++i) v[i] = in[(size_t)clampi(y - KS + i, 0, H - 1) * W + x];
    }
    median_ops<FS, 0>(v);
    out[(size_t)y * W + x] = v[KS];
}

template <int FS>
__global__ void __launch_bounds__(kThreads) k_median(const float *__restrict__ in0, const float *__restrict__ in1,
                                                     float *__restrict__ out0, float *__restrict__ out1, int W,
                                                     int H, int dir)
{
    const int gx = (W + kThreads - 1) / kThreads;
    median_tile<FS>(blockIdx.y ? in1 : in0, blockIdx.y ? out1 : out0, W, H, dir, xcd_remap(blockIdx.x, gx * H));
}

// ---------------------------------------------------------------------------
// Cross-trilateral filter.  Work-group: 64 x 16 output tile, thread (lx, ly)
// owns the four consecutive rows 4 ly + k.  Source rows stream through LDS in
// bands of BR rows as float4 (u, v, b I, 0), where b^2 = log2(e) / (2 sigmaI
// 255^2), and one tap is
//   t = a2 ((du)^2 + (dv)^2) + (dI')^2 + log2(e) i^2 / (2 sigmaS),  w = exp2(-t),
// a2 = log2(e) / (2 sigmaEPE): the product of the reference's three
// exponentials.  The flow differences are taken before any scaling, so their
// rounding is relative to the difference, not to the flow: scaled values
// (a u - a um) carry an error of about |a u| 2^-24 into every exponent, which
// the normalisation cancels except where the weight sum falls below its 1e-5
// floor, and there reaches 1e-3 px at 36 px of flow.  The j (row) part of the
// spatial term is applied once per source row.

constexpr int kTriTW = 64, kTriTH = 16, kTriPY = 4;
constexpr int kTriLdsBytes = 64 * 1024;

// Split of a tile's source rows over K work-groups (K = 1: the whole range):
// part p takes rows [src0 + p n / K, src0 + (p + 1) n / K) of the n rows, and
// writes its sums of w, w u and w v per output pixel to partial (three planes
// of W H floats per part, in that order) for k_trilateral_finish.
__device__ __forceinline__ void tri_part_rows(int part, int K, int &src0, int &src1)
{
    if (K > 1) {
        const int n = src1 - src0, a = src0;
        src0 = a + part * n / K;
        src1 = a + (part + 1) * n / K;
    }
}

__device__ __forceinline__ void tri_store_partial(float *__restrict__ partial, int part, int W, int H, size_t o,
                                                  float sw, float su, float sv)
{
    const size_t WH = (size_t)W * H;
    float *q = partial + (size_t)3 * part * WH + o;
    q[0] = sw;
    q[WH] = su;
    q[2 * WH] = sv;
}

__device__ __forceinline__ void tri_tile(const float *__restrict__ Ui, const float *__restrict__ Vi,
                                         const float *__restrict__ Um, const float *__restrict__ Vm,
                                         const float *__restrict__ I, float *__restrict__ Uo, float *__restrict__ Vo,
                                         int W, int H, int s, int BR, float a2, float bI,
                                         const float *__restrict__ spt, int blk, float4 *band, int part = 0,
                                         int K = 1, float *__restrict__ partial = nullptr)
{
    const int pitch = kTriTW + 2 * s;
    // the column part of the spatial weight (2s entries) after the band
    float *spl = reinterpret_cast<float *>(band + (size_t)BR * pitch);
    for (int i = threadIdx.x; i < 2 * s; i += kThreads) spl[i] = spt[i];
    const int gxb = (W + kTriTW - 1) / kTriTW;
    const int tx0 = (blk % gxb) * kTriTW, ty0 = (blk / gxb) * kTriTH;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int x = tx0 + lx, y0 = ty0 + 4 * ly;
    const int xc = min(x, W - 1);

    float um[kTriPY], vm[kTriPY], g0[kTriPY], au[kTriPY], av[kTriPY], aw[kTriPY];
#pragma unroll
    for (int k = 0; k < kTriPY; ++k) {
        const size_t o = (size_t)min(y0 + k, H - 1) * W + xc;
        um[k] = Um[o];
        vm[k] = Vm[o];
        g0[k] = I[o] * bI;
        au[k] = av[k] = aw[k] = 0.0f;
    }

    int src0 = ty0 - s, src1 = ty0 + kTriTH + s;        // source rows of the tile
    tri_part_rows(part, K, src0, src1);                 // this work-group's share of them
    const int my0 = y0 - s, my1 = y0 + kTriPY - 1 + s;  // source rows this wave needs
    for (int band0 = src0; band0 < src1; band0 += BR) {
        __syncthreads();
        for (int r = ly; r < BR; r += 4) {
            const int gy = clampi(band0 + r, 0, H - 1);
            const float *ru = Ui + (size_t)gy * W, *rv = Vi + (size_t)gy * W, *ri = I + (size_t)gy * W;
            for (int c = lx; c < pitch; c += 64) {
                const int gx = clampi(tx0 - s + c, 0, W - 1);
                band[r * pitch + c] = make_float4(ru[gx], rv[gx], ri[gx] * bI, 0.0f);
            }
        }
        __syncthreads();
        const int rend = min(BR, src1 - band0);
        for (int r = 0; r < rend; ++r) {
            const int ys = band0 + r;
            if (ys < my0 || ys >= my1) continue;  // wave-uniform
            float ru[kTriPY], rv[kTriPY], rw[kTriPY];
#pragma unroll
            for (int k = 0; k < kTriPY; ++k) ru[k] = rv[k] = rw[k] = 0.0f;
            // two taps per step: their sources and spatial terms are read from
            // LDS together (2s is even), one wait per pair of taps
            const float4 *src = band + r * pitch + lx;
            for (int i = 0; i < 2 * s; i += 2) {
                const float4 qa = src[i], qb = src[i + 1];
                const float2 sp = *reinterpret_cast<const float2 *>(spl + i);
#pragma unroll
                for (int k = 0; k < kTriPY; ++k) {
                    float du = qa.x - um[k], dv = qa.y - vm[k], dg = qa.z - g0[k];
                    float t = fmaf(dg, dg, sp.x);
                    float e = fmaf(dv, dv, du * du);
                    t = fmaf(e, a2, t);
                    float w = __builtin_amdgcn_exp2f(-t);
                    rw[k] += w;
                    ru[k] = fmaf(w, qa.x, ru[k]);
                    rv[k] = fmaf(w, qa.y, rv[k]);
                    du = qb.x - um[k];
                    dv = qb.y - vm[k];
                    dg = qb.z - g0[k];
                    t = fmaf(dg, dg, sp.y);
                    e = fmaf(dv, dv, du * du);
                    t = fmaf(e, a2, t);
                    w = __builtin_amdgcn_exp2f(-t);
                    rw[k] += w;
                    ru[k] = fmaf(w, qb.x, ru[k]);
                    rv[k] = fmaf(w, qb.y, rv[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < kTriPY; ++k) {
                const int j = ys - (y0 + k);
                const float ej = __builtin_amdgcn_exp2f(-spt[clampi(j + s, 0, 2 * s - 1)]);
                const float e = (j >= -s && j < s) ? ej : 0.0f;
                aw[k] = fmaf(e, rw[k], aw[k]);
                au[k] = fmaf(e, ru[k], au[k]);
                av[k] = fmaf(e, rv[k], av[k]);
            }
        }
    }
    if (x < W) {
#pragma unroll
        for (int k = 0; k < kTriPY; ++k) {
            const int y = y0 + k;
            if (y < H) {
                if (K > 1) {
                    tri_store_partial(partial, part, W, H, (size_t)y * W + x, aw[k], au[k], av[k]);
                } else {
                    const float den = fmaxf(1e-5f, aw[k]);
                    Uo[(size_t)y * W + x] = au[k] / den;
                    Vo[(size_t)y * W + x] = av[k] / den;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_trilateral(const float *__restrict__ Ui, const float *__restrict__ Vi,
                                                         const float *__restrict__ Um, const float *__restrict__ Vm,
                                                         const float *__restrict__ I, float *__restrict__ Uo,
                                                         float *__restrict__ Vo, int W, int H, int s, int BR,
                                                         float a2, float bI, const float *__restrict__ spt, int K,
                                                         float *__restrict__ partial)
{
    extern __shared__ float4 band[];
    const int gxb = (W + kTriTW - 1) / kTriTW, gyb = (H + kTriTH - 1) / kTriTH;
    // K work-groups per tile (the parts of its source rows), adjacent in the remapped order
    const int lb = xcd_remap(blockIdx.x, gxb * gyb * K);
    tri_tile(Ui, Vi, Um, Vm, I, Uo, Vo, W, H, s, BR, a2, bI, spt, lb / K, band, lb % K, K, partial);
}

// The split trilateral's last step: per pixel the parts' sums, added in part
// order, and the normalisation of the unsplit kernels.
__global__ void __launch_bounds__(kThreads) k_trilateral_finish(const float *__restrict__ partial, int K, int W,
                                                                int H, float *__restrict__ Uo,
                                                                float *__restrict__ Vo)
{
    const size_t WH = (size_t)W * H, i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= WH) return;
    float sw = 0.0f, su = 0.0f, sv = 0.0f;
    for (int p = 0; p < K; ++p) {
        const float *q = partial + (size_t)3 * p * WH + i;
        sw += q[0];
        su += q[WH];
        sv += q[2 * WH];
    }
    const float den = fmaxf(1e-5f, sw);
    Uo[i] = su / den;
    Vo[i] = sv / den;
}

// Cross-trilateral filter with the three weighted sums on the matrix cores.
// v_mfma_f32_4x4x1f32 runs 16 independent 4x4 outer products per wave; block b
// is lanes 4b..4b+3 and is one output column.  Lane 4b+k computes the weight w
// of output row k for the current source q (the A operand, w of the four rows
// of a column share q); lane 4b+j supplies component j of (1, u, v, I) of q
// times the column factor exp2(-spatial_i) (the B operand); D[i][j], held by
// lane 4b+j in register i, accumulates row i's sum over q of w * B_j.  The row
// part of the spatial term is the first addend of the exponent (+inf outside
// the window), so a tap costs 3 sub, 1 mul, 3 FMA and one exp2 per pixel on the VALU;
// the per-source B scaling is shared by the R groups of four rows a wave owns.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int R, int MASK>
__device__ __forceinline__ void tri_taps(const float4 *__restrict__ src, const float *__restrict__ srcf,
                                         const float *__restrict__ es, int n, float a2, const float (&um)[R],
                                         const float (&vm)[R], const float (&g0)[R], const float (&ck)[R],
                                         v4f (&acc)[R])
{
#pragma unroll 2
    for (int i = 0; i < n; ++i) {
        const float4 q = src[i];
        const float bq = srcf[4 * i] * es[i];
#pragma unroll
        for (int g = 0; g < R; ++g) {
            if (!((MASK >> g) & 1)) continue;
            const float du = q.x - um[g], dv = q.y - vm[g], dg = q.z - g0[g];
            float t = fmaf(dg, dg, ck[g]);
            t = fmaf(fmaf(dv, dv, du * du), a2, t);
            const float w = __builtin_amdgcn_exp2f(-t);
            acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(w, bq, acc[g], 0, 0, 0);
        }
    }
}

// groups whose rows see source row ys form a contiguous range; one loop per mask
template <int R, int M>
__device__ __forceinline__ void tri_taps_mask(int mask, const float4 *__restrict__ src,
                                              const float *__restrict__ srcf, const float *__restrict__ es, int n,
                                              float a2, const float (&um)[R], const float (&vm)[R],
                                              const float (&g0)[R], const float (&ck)[R], v4f (&acc)[R])
{
    if constexpr (M > 0) {
        if (mask == M)
            tri_taps<R, M>(src, srcf, es, n, a2, um, vm, g0, ck, acc);
        else
            tri_taps_mask<R, M - 1>(mask, src, srcf, es, n, a2, um, vm, g0, ck, acc);
    }
}

template <int R>
__device__ __forceinline__ void tri_mfma_tile(const float *__restrict__ Ui, const float *__restrict__ Vi,
                                              const float *__restrict__ Um, const float *__restrict__ Vm,
                                              const float *__restrict__ I, float *__restrict__ Uo,
                                              float *__restrict__ Vo, int W, int H, int s, int BR, float a2, float bI,
                                              const float *__restrict__ spt, int blk, float4 *band, int part = 0,
                                              int K = 1, float *__restrict__ partial = nullptr)
{
    constexpr int TH = 4 * R;
    const int pitch = kTriTW + 2 * s;
    // exp2 of minus the column part of the spatial term (2s entries) after the band
    float *es = reinterpret_cast<float *>(band + (size_t)BR * pitch);
    for (int i = threadIdx.x; i < 2 * s; i += kThreads) es[i] = __builtin_amdgcn_exp2f(-spt[i]);
    const int gxb = (W + kTriTW - 1) / kTriTW;
    const int tx0 = (blk % gxb) * kTriTW, ty0 = (blk / gxb) * TH;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = wave * 16 + (lane >> 2);  // output column within the tile
    const int k = lane & 3;                   // row of the group (A), component (B)
    const int comp = (k + 3) & 3;             // B: 1 (the .w), u, v, I
    const int x = tx0 + col, xc = min(x, W - 1);

    float um[R], vm[R], g0[R];
    v4f acc[R];
#pragma unroll
    for (int g = 0; g < R; ++g) {
        const size_t o = (size_t)min(ty0 + 4 * g + k, H - 1) * W + xc;
        um[g] = Um[o];
        vm[g] = Vm[o];
        g0[g] = I[o] * bI;
        acc[g] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }

    int src0 = ty0 - s, src1 = ty0 + TH + s;  // source rows of the tile
    tri_part_rows(part, K, src0, src1);       // this work-group's share of them
    for (int band0 = src0; band0 < src1; band0 += BR) {
        __syncthreads();
        for (int r = wave; r < BR; r += 4) {
            const int gy = clampi(band0 + r, 0, H - 1);
            const float *ru = Ui + (size_t)gy * W, *rv = Vi + (size_t)gy * W, *ri = I + (size_t)gy * W;
            for (int c = lane; c < pitch; c += 64) {
                const int gx = clampi(tx0 - s + c, 0, W - 1);
                band[r * pitch + c] = make_float4(ru[gx], rv[gx], ri[gx] * bI, 1.0f);
            }
        }
        __syncthreads();
        const int rend = min(BR, src1 - band0);
        for (int r = 0; r < rend; ++r) {
            const int ys = band0 + r;
            float ck[R];
            int mask = 0;  // wave-uniform: groups with a row inside the window of ys
#pragma unroll
            for (int g = 0; g < R; ++g) {
                const int j = ys - (ty0 + 4 * g + k);
                ck[g] = (j >= -s && j < s) ? spt[j + s] : __builtin_inff();
                const int y0g = ty0 + 4 * g;
                if (ys >= y0g - s && ys < y0g + 3 + s) mask |= 1 << g;
            }
            const float4 *src = band + r * pitch + col;
            const float *srcf = reinterpret_cast<const float *>(src) + comp;
            tri_taps_mask<R, (1 << R) - 1>(mask, src, srcf, es, 2 * s, a2, um, vm, g0, ck, acc);
        }
    }
    // lane 4b+j holds, in register i, row i's sum of w * (1, u, v, I)_j * es
    if (K > 1) {  // lanes 4b+0..2 hold the sums of w, w u, w v: one plane each
        if (x < W && k < 3) {
            const size_t WH = (size_t)W * H;
            float *q = partial + ((size_t)3 * part + k) * WH;
#pragma unroll
            for (int g = 0; g < R; ++g)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int y = ty0 + 4 * g + i;
                    if (y < H) q[(size_t)y * W + x] = acc[g][i];
                }
        }
        return;
    }
#pragma unroll
    for (int g = 0; g < R; ++g) {
        float den[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) den[i] = fmaxf(1e-5f, __shfl(acc[g][i], lane & ~3));
        if (x < W && (k == 1 || k == 2)) {
            float *out = k == 1 ? Uo : Vo;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y = ty0 + 4 * g + i;
                if (y < H) out[(size_t)y * W + x] = acc[g][i] / den[i];
            }
        }
    }
}

template <int R>
__global__ void __launch_bounds__(kThreads) k_trilateral_mfma(const float *__restrict__ Ui,
                                                              const float *__restrict__ Vi,
                                                              const float *__restrict__ Um,
                                                              const float *__restrict__ Vm,
                                                              const float *__restrict__ I, float *__restrict__ Uo,
                                                              float *__restrict__ Vo, int W, int H, int s, int BR,
                                                              float a2, float bI, const float *__restrict__ spt, int K,
                                                              float *__restrict__ partial)
{
    extern __shared__ float4 band[];
    const int gxb = (W + kTriTW - 1) / kTriTW, gyb = (H + 4 * R - 1) / (4 * R);
    const int lb = xcd_remap(blockIdx.x, gxb * gyb * K);
    tri_mfma_tile<R>(Ui, Vi, Um, Vm, I, Uo, Vo, W, H, s, BR, a2, bI, spt, lb / K, band, lb % K, K, partial);
}

// plain 16-byte-per-lane copy, grid-stride
__global__ void __launch_bounds__(kThreads) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Multi-pair kernels: each runs the single-pair tile body for the tiles of all
// its jobs (see tile_of / job_of); a work-group may take several tiles, and
// kernels with LDS tiles wait for the previous tile's readers before reusing it.

// `slots` is start[njobs]; it_ counts the tiles this work-group has taken
#define BM_MULTI_TILES(...)                                                              \
    for (int slot_ = (int)blockIdx.x, it_ = 0; slot_ < slots; slot_ += (int)gridDim.x) { \
        int lt_ = 0;                                                                     \
        const int j_ = job_of(start, njobs, slot_, &lt_);                               \
        if (j_ < 0) continue;                                                            \
        __VA_ARGS__                                                                      \
        ++it_;                                                                           \
    }

__global__ void __launch_bounds__(kThreads) k_gray_multi(const GrayJob *__restrict__ jobs,
                                                         const int *__restrict__ start, int njobs, int slots)
{
    BM_MULTI_TILES({
        const GrayJob &J = jobs[j_];
        gray_tile(J.rgb, J.g8, J.g32, J.n, lt_);
    })
}

template <int NSP, int RN, int PY>
__global__ void __launch_bounds__(kThreads) k_brief_multi(const BriefJob *__restrict__ jobs,
                                                          const int *__restrict__ start, int njobs, int slots)
{
    extern __shared__ __align__(16) unsigned char smem[];
    BM_MULTI_TILES({
        if (it_) __syncthreads();
        brief_tile<NSP, RN, PY, false>(jobs[j_], lt_, smem);
    })
}

__device__ __forceinline__ RSArgs job_rs(const MatchJob &J, int rs_iteration, const int32_t *ndtri)
{
    RSArgs rs;
    rs.enabled = rs_iteration >= 0;
    rs.iteration = (uint32_t)(rs_iteration < 0 ? 0 : rs_iteration);
    rs.pair = J.pair;
    rs.key0 = J.key0;
    rs.key1 = J.key1;
    rs.sigma_q16 = J.sigma_q16;
    rs.ndtri = ndtri;
    return rs;
}

// zero initialisation into off[0] / cost[0]
template <int NSP, bool RS>
__global__ void __launch_bounds__(kThreads) k_init_multi(const MatchJob *__restrict__ jobs,
                                                         const int *__restrict__ start, int njobs, int slots,
                                                         int rs_iteration)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    const int32_t *ndtri = RS ? ndtri_to_lds(nd) : d_ndtri_q16;
    BM_MULTI_TILES({
        const MatchJob &J = jobs[j_];
        const RSArgs rs = job_rs(J, rs_iteration, ndtri);
        init_tile<NSP, RS>(J.F1, J.F2, 1, nullptr, J.off[0], J.cost[0], J.sx, J.sy, J.strip_rows, rs, lt_);
    })
}

// one level from off[src] / cost[src] into the other buffer
template <int NSP, bool RS, int PX>
__global__ void __launch_bounds__(kThreads) k_level_multi(const MatchJob *__restrict__ jobs,
                                                          const int *__restrict__ start, int njobs, int slots, int p,
                                                          int pd, int src, int rs_iteration)
{
    __shared__ __align__(16) int32_t nd[RS ? 4096 : 4];
    const int32_t *ndtri = RS ? ndtri_to_lds(nd) : d_ndtri_q16;
    BM_MULTI_TILES({
        const MatchJob &J = jobs[j_];
        const RSArgs rs = job_rs(J, rs_iteration, ndtri);
        level_tile<NSP, RS, PX>(J.F1, J.F2, J.off[src], J.cost[src], J.off[1 - src], J.cost[1 - src], J.sx, J.sy,
                                p, pd, J.double_prop, J.strip_rows, rs, lt_);
    })
}

__global__ void __launch_bounds__(kThreads) k_downsample_multi(const DsJob *__restrict__ jobs,
                                                               const int *__restrict__ start, int njobs, int slots,
                                                               int tiled)
{
    extern __shared__ uint32_t win[];
    BM_MULTI_TILES({
        const DsJob &J = jobs[j_];
        if (tiled) {
            if (it_) __syncthreads();
            ds_tile(J.off, J.nn_x, J.nn_y, J.gauss, J.P, J.sx, J.sy, J.W, J.H, J.scx, J.scy, J.U, J.V, lt_, win);
        } else {
            ds_direct_tile(J.off, J.nn_x, J.nn_y, J.gauss, J.P, J.sx, J.sy, J.W, J.H, J.scx, J.scy, J.identity, J.U,
                           J.V, lt_);
        }
    })
}

// a job's tiles are its first component's row blocks, then its second's
template <int FS>
__global__ void __launch_bounds__(kThreads) k_median_multi(const MedJob *__restrict__ jobs,
                                                           const int *__restrict__ start, int njobs, int slots,
                                                           int dir)
{
    BM_MULTI_TILES({
        const MedJob &J = jobs[j_];
        const int per = ((J.W + kThreads - 1) / kThreads) * J.H;
        const int c = lt_ >= per;
        median_tile<FS>(c ? J.in1 : J.in0, c ? J.out1 : J.out0, J.W, J.H, dir, lt_ - c * per);
    })
}

// groups = 0: VALU sums; 2, 4: MFMA sums with that many 4-row groups
template <int GROUPS>
__global__ void __launch_bounds__(kThreads) k_trilateral_multi(const TriJob *__restrict__ jobs,
                                                               const int *__restrict__ start, int njobs, int slots)
{
    extern __shared__ float4 band[];
    BM_MULTI_TILES({
        const TriJob &J = jobs[j_];
        if (it_) __syncthreads();
        if constexpr (GROUPS == 0)
            tri_tile(J.Ui, J.Vi, J.Um, J.Vm, J.I, J.Uo, J.Vo, J.W, J.H, J.s, J.BR, J.a2, J.bI, J.spt, lt_, band);
        else
            tri_mfma_tile<GROUPS>(J.Ui, J.Vi, J.Um, J.Vm, J.I, J.Uo, J.Vo, J.W, J.H, J.s, J.BR, J.a2, J.bI, J.spt,
                                  lt_, band);
    })
}

#undef BM_MULTI_TILES

inline unsigned blocks_rows(int width, int rows) { return (unsigned)(((width + kThreads - 1) / kThreads) * rows); }

}  // namespace

int brief_max_radius()
{
    int R = 0;
    while ((kBriefTW + 2 * (R + 1)) * (kBriefTH + 2 * (R + 1)) * 4 <= kBriefLdsBytes) ++R;
    return R;
}

size_t brief_lds_bytes(const Geometry &g, int RN, int TH)
{
    const size_t lt = (size_t)(kBriefTW + 2 * RN) * (TH + 2 * RN) * sizeof(float);
    const size_t xt = (size_t)(kBriefTW + 2 * g.brief_far + 1) / 2 * 2;  // padded to 8 bytes
    const size_t tabs = xt * sizeof(uint32_t) + (size_t)(TH + 2 * g.brief_far) * sizeof(uint2);
    const int which = TH == 32 ? 1 : 0;
    return lt + tabs + (size_t)g.brief_gw[which] * g.brief_gh[which] * sizeof(uint16_t);
}

int brief_tile_rows(int NSP) { return NSP <= 4 ? 32 : 16; }

int tri_max_band(int s)
{
    const int BR = (kTriLdsBytes - 2 * s * (int)sizeof(float)) / ((kTriTW + 2 * s) * (int)sizeof(float4));
    return BR < 16 ? BR : 16;
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 256 * 32);
    hipLaunchKernelGGL(k_copy, dim3(std::max(blocks, 1u)), dim3(kThreads), 0, st, (const uint4 *)src, (uint4 *)dst, n);
    return hipGetLastError();
}

hipError_t launch_gray(const Geometry &g, const uint8_t *rgb, uint8_t *g8, float *g32, hipStream_t st)
{
    const int n = g.W * g.H;
    hipLaunchKernelGGL(k_gray, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, rgb, g8, g32, n);
    return hipGetLastError();
}

hipError_t launch_gray_pair(const Geometry &g, const uint8_t *rgb1, uint8_t *g8a, float *g32a, const uint8_t *rgb2,
                            uint8_t *g8b, hipStream_t st)
{
    const int n = g.W * g.H, nb = (n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_gray_pair, dim3(2 * nb), dim3(kThreads), 0, st, rgb1, g8a, g32a, rgb2, g8b, n, nb);
    return hipGetLastError();
}

hipError_t launch_upsample_lanczos(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st)
{
    if (!g.lz_x0 || !g.lz_wx || !g.lz_y0 || !g.lz_wy) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_upsample_lanczos, dim3(blocks_rows(g.sx, g.sy)), dim3(kThreads), 0, st, g8, g.lz_x0, g.lz_wx,
                       g.lz_y0, g.lz_wy, g.W, g.H, g.sx, g.sy, L);
    return hipGetLastError();
}

hipError_t launch_upsample(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st)
{
    hipLaunchKernelGGL(k_upsample, dim3(blocks_rows(g.sx, g.sy)), dim3(kThreads), 0, st, g8, g.up_x, g.up_y, g.W,
                       g.H, g.sx, g.sy, L);
    return hipGetLastError();
}

template <int NSP, bool FUSED>
static hipError_t launch_brief_t(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F, hipStream_t st)
{
    const int R = g.brief_radius;
    const size_t lds = (size_t)(kBriefTW + 2 * R) * (kBriefTH + 2 * R) * sizeof(float);
    const int gxb = (g.sx + kBriefTW - 1) / kBriefTW, gyb = (g.sy + kBriefTH - 1) / kBriefTH;
    hipLaunchKernelGGL((k_brief<NSP, FUSED>), dim3(gxb * gyb), dim3(kThreads), lds, st, L, g8, g.up_x, g.up_y,
                       g.pattern, g.N, g.W, g.H, g.sx, g.sy, R, F);
    return hipGetLastError();
}

// L != nullptr: the up-sampled frame is in memory (the GL kernel, LDS for
// the near tile only); else the fused kernel interpolates from g8.
// The frame kernel's work-group shape: G waves over a 64 x (G PY) tile, PY
// rows per thread.  Its LDS tile limits the work-groups per CU, so more waves
// per tile hide more of the LDS latency; fewer rows per thread also free
// registers.  Measured (BRIEF + up-sampling per pair, best halo of each, two
// interleaved rounds on one box): 128 bits 4 waves x 8 rows 0.745 ms, 8 x 4
// 0.650, 16 x 2 0.657, 16 x 4 0.670; 512 bits 4 x 4 3.79 ms, 8 x 2 3.52,
// 16 x 1 3.78, 16 x 2 2.92 (32-row tiles: the 48-px halo fits one
// work-group of 16 waves per CU and leaves 31 far tests of 512), 16 x 4 3.54.
template <int NSP> struct BriefFrameShape {
    static constexpr int PY = NSP <= 4 ? 4 : 2, G = NSP <= 4 ? 8 : 16;
};

// The fused kernel's: 8 waves over the same tiles as its 4-wave form (64 x 32
// up to 128 bits, 64 x 16 above), whose LDS plan (pair tile, tables) is sized
// by the tile height; the multi-pair kernel keeps 4 waves.
template <int NSP> struct BriefFusedShape {
    static constexpr int G = 8, PY = (NSP <= 4 ? 32 : 16) / G;
};

template <int NSP, int RN>
static hipError_t launch_brief_lds_t(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F,
                                     hipStream_t st)
{
    constexpr int PY = NSP <= 4 ? 8 : 4;
    constexpr int TH = 4 * PY;
    BriefJob a{g8,  g.up_x, g.up_y, g.brief_plan, g.brief_plan_count, g.N, g.W, g.H, g.sx, g.sy, g.brief_far,
               g.brief_gw[PY == 8 ? 1 : 0], g.brief_gh[PY == 8 ? 1 : 0], F};
    a.L = L;
    const int gxb = (g.sx + kBriefTW - 1) / kBriefTW;
    if (L) {
        constexpr int PYF = BriefFrameShape<NSP>::PY, GF = BriefFrameShape<NSP>::G, THF = PYF * GF;
        const int gyb = (g.sy + THF - 1) / THF;
        const size_t lds = (size_t)(kBriefTW + 2 * RN) * (THF + 2 * RN) * sizeof(float);
        hipLaunchKernelGGL((k_brief_lds<NSP, RN, PYF, true, GF>), dim3(gxb * gyb), dim3(64 * GF), lds, st, a);
    } else {
        constexpr int PYU = BriefFusedShape<NSP>::PY, GU = BriefFusedShape<NSP>::G;
        static_assert(PYU * GU == TH, "the fused kernel's tile height is its LDS plan's");
        const int gyb = (g.sy + TH - 1) / TH;
        const size_t lds = brief_lds_bytes(g, RN, TH);
        hipLaunchKernelGGL((k_brief_lds<NSP, RN, PYU, false, GU>), dim3(gxb * gyb), dim3(64 * GU), lds, st, a);
    }
    return hipGetLastError();
}

template <int NSP, int RN>
static hipError_t launch_brief_pair_t(const Geometry &g, const uint8_t *g8a, uint32_t *Fa, const uint8_t *g8b,
                                      uint32_t *Fb, hipStream_t st)
{
    constexpr int PY = NSP <= 4 ? 8 : 4;
    constexpr int TH = 4 * PY;
    constexpr int PYU = BriefFusedShape<NSP>::PY, GU = BriefFusedShape<NSP>::G;
    static_assert(PYU * GU == TH, "the fused kernel's tile height is its LDS plan's");
    const BriefJob a{g8a, g.up_x, g.up_y, g.brief_plan, g.brief_plan_count, g.N, g.W, g.H, g.sx, g.sy, g.brief_far,
                     g.brief_gw[PY == 8 ? 1 : 0], g.brief_gh[PY == 8 ? 1 : 0], Fa};
    const int gxb = (g.sx + kBriefTW - 1) / kBriefTW, gyb = (g.sy + TH - 1) / TH;
    const size_t lds = brief_lds_bytes(g, RN, TH);
    hipLaunchKernelGGL((k_brief_lds_pair<NSP, RN, PYU, GU>), dim3(2 * gxb * gyb), dim3(64 * GU), lds, st, a, g8b,
                       Fb);
    return hipGetLastError();
}

template <int NSP>
static hipError_t launch_brief_lds_n(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F,
                                     hipStream_t st)
{
    switch (g.brief_near) {
    case 8: return launch_brief_lds_t<NSP, 8>(g, L, g8, F, st);
    case 16: return launch_brief_lds_t<NSP, 16>(g, L, g8, F, st);
    case 24: return launch_brief_lds_t<NSP, 24>(g, L, g8, F, st);
    case 32: return launch_brief_lds_t<NSP, 32>(g, L, g8, F, st);
    case 40: return launch_brief_lds_t<NSP, 40>(g, L, g8, F, st);
    case 48: return launch_brief_lds_t<NSP, 48>(g, L, g8, F, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_brief_pair(const Geometry &g, const uint8_t *g8a, uint32_t *Fa, const uint8_t *g8b, uint32_t *Fb,
                             hipStream_t st)
{
    if (g.brief_near <= 0) return hipErrorInvalidValue;
#define BM_PAIR_RN(NSPV)                                                                                         \
    switch (g.brief_near) {                                                                                      \
    case 8: return launch_brief_pair_t<NSPV, 8>(g, g8a, Fa, g8b, Fb, st);                                        \
    case 16: return launch_brief_pair_t<NSPV, 16>(g, g8a, Fa, g8b, Fb, st);                                      \
    case 24: return launch_brief_pair_t<NSPV, 24>(g, g8a, Fa, g8b, Fb, st);                                      \
    case 32: return launch_brief_pair_t<NSPV, 32>(g, g8a, Fa, g8b, Fb, st);                                      \
    case 40: return launch_brief_pair_t<NSPV, 40>(g, g8a, Fa, g8b, Fb, st);                                      \
    case 48: return launch_brief_pair_t<NSPV, 48>(g, g8a, Fa, g8b, Fb, st);                                      \
    default: return hipErrorInvalidValue;                                                                        \
    }
    switch (g.NSP) {
    case 1: BM_PAIR_RN(1)
    case 2: BM_PAIR_RN(2)
    case 4: BM_PAIR_RN(4)
    case 8: BM_PAIR_RN(8)
    case 16: BM_PAIR_RN(16)
    default: return hipErrorInvalidValue;
    }
#undef BM_PAIR_RN
}

hipError_t launch_brief(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F, hipStream_t st)
{
    const bool fused = (L == nullptr);
    if (g.brief_near > 0) {
        switch (g.NSP) {
        case 1: return launch_brief_lds_n<1>(g, L, g8, F, st);
        case 2: return launch_brief_lds_n<2>(g, L, g8, F, st);
        case 4: return launch_brief_lds_n<4>(g, L, g8, F, st);
        case 8: return launch_brief_lds_n<8>(g, L, g8, F, st);
        case 16: return launch_brief_lds_n<16>(g, L, g8, F, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (g.NSP) {
    case 1: return fused ? launch_brief_t<1, true>(g, L, g8, F, st) : launch_brief_t<1, false>(g, L, g8, F, st);
    case 2: return fused ? launch_brief_t<2, true>(g, L, g8, F, st) : launch_brief_t<2, false>(g, L, g8, F, st);
    case 4: return fused ? launch_brief_t<4, true>(g, L, g8, F, st) : launch_brief_t<4, false>(g, L, g8, F, st);
    case 8: return fused ? launch_brief_t<8, true>(g, L, g8, F, st) : launch_brief_t<8, false>(g, L, g8, F, st);
    case 16: return fused ? launch_brief_t<16, true>(g, L, g8, F, st) : launch_brief_t<16, false>(g, L, g8, F, st);
    default: return hipErrorInvalidValue;
    }
}

#define BM_NSP_DISPATCH(NSPV, EXPR)                                                                             \
    switch (NSPV) {                                                                                             \
    case 1: { constexpr int NSP = 1; EXPR; } break;                                                             \
    case 2: { constexpr int NSP = 2; EXPR; } break;                                                             \
    case 4: { constexpr int NSP = 4; EXPR; } break;                                                             \
    case 8: { constexpr int NSP = 8; EXPR; } break;                                                             \
    case 16: { constexpr int NSP = 16; EXPR; } break;                                                           \
    default: return hipErrorInvalidValue;                                                                       \
    }

static RSArgs make_rs(const Geometry &g, int enabled, uint32_t iteration, uint32_t pair)
{
    RSArgs rs;
    rs.enabled = enabled;
    rs.iteration = iteration;
    rs.pair = pair;
    rs.key0 = (uint32_t)g.seed;
    rs.key1 = (uint32_t)(g.seed >> 32);
    rs.sigma_q16 = g.rs_sigma_q16;
    rs.ndtri = nullptr;  // each kernel points it at its LDS copy
    return rs;
}

hipError_t launch_init(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int first, const uint32_t *prev,
                       uint32_t *off, uint16_t *cost, int rs_iteration, uint32_t pair, hipStream_t st)
{
    const dim3 grid(blocks_rows(g.sx, g.sy));
    const RSArgs rs = make_rs(g, rs_iteration >= 0, (uint32_t)(rs_iteration < 0 ? 0 : rs_iteration), pair);
    if (g.NSP == 16) {  // two lanes per pixel, as the levels (1280x720 x4.9: 0.827 -> 0.783 ms)
        constexpr int L = 2, span = kThreads / L;
        const dim3 grid2((unsigned)(((g.sx + span - 1) / span) * g.sy));
        if (rs.enabled)
            hipLaunchKernelGGL((k_init_lanes<16, true, L>), grid2, dim3(kThreads), 0, st, F1, F2, first, prev, off,
                               cost, g.sx, g.sy, g.strip_rows, rs);
        else
            hipLaunchKernelGGL((k_init_lanes<16, false, L>), grid2, dim3(kThreads), 0, st, F1, F2, first, prev, off,
                               cost, g.sx, g.sy, g.strip_rows, rs);
        return hipGetLastError();
    }
    if (rs.enabled) {
        BM_NSP_DISPATCH(g.NSP, hipLaunchKernelGGL((k_init<NSP, true>), grid, dim3(kThreads), 0, st, F1, F2, first,
                                                  prev, off, cost, g.sx, g.sy, g.strip_rows, rs));
    } else {
        BM_NSP_DISPATCH(g.NSP, hipLaunchKernelGGL((k_init<NSP, false>), grid, dim3(kThreads), 0, st, F1, F2, first,
                                                  prev, off, cost, g.sx, g.sy, g.strip_rows, rs));
    }
    return hipGetLastError();
}

template <int NSP, bool RS, int PX>
static void launch_level_t(const Geometry &g, const uint32_t *F1, const uint32_t *F2, const uint32_t *off_in,
                           const uint16_t *cost_in, uint32_t *off_out, uint16_t *cost_out, int p, int pd,
                           const RSArgs &rs, hipStream_t st)
{
    const int span = kThreads * PX;
    const dim3 grid((unsigned)(((g.sx + span - 1) / span) * g.sy));
    hipLaunchKernelGGL((k_level<NSP, RS, PX>), grid, dim3(kThreads), 0, st, F1, F2, off_in, cost_in, off_out,
                       cost_out, g.sx, g.sy, p, pd, g.doubleProp, g.strip_rows, rs);
}

// pixels per thread: 4 only for descriptors of at most 128 bits (registers)
template <int NSP, bool RSV>
static void level_px_t(const Geometry &g, const uint32_t *F1, const uint32_t *F2, const uint32_t *off_in,
                       const uint16_t *cost_in, uint32_t *off_out, uint16_t *cost_out, int p, int pd, const RSArgs &rs,
                       hipStream_t st)
{
    // 512-bit descriptors: two lanes per pixel, 32 bytes each (1280x720 x4.9
    // pair: 9.26 -> 8.42 ms of levels against one lane per pixel, 8.80 with
    // four lanes of 16 bytes)
    if constexpr (NSP == 16) {
        constexpr int L = 2, span = kThreads / L;
        const dim3 grid((unsigned)(((g.sx + span - 1) / span) * g.sy));
        hipLaunchKernelGGL((k_level_lanes<NSP, RSV, L>), grid, dim3(kThreads), 0, st, F1, F2, off_in, cost_in,
                           off_out, cost_out, g.sx, g.sy, p, pd, g.doubleProp, g.strip_rows, rs);
        return;
    }
    if constexpr (NSP <= 4) {
        if (g.level_px == 4) {
            launch_level_t<NSP, RSV, 4>(g, F1, F2, off_in, cost_in, off_out, cost_out, p, pd, rs, st);
            return;
        }
    }
    if (g.level_px == 2)
        launch_level_t<NSP, RSV, 2>(g, F1, F2, off_in, cost_in, off_out, cost_out, p, pd, rs, st);
    else
        launch_level_t<NSP, RSV, 1>(g, F1, F2, off_in, cost_in, off_out, cost_out, p, pd, rs, st);
}

hipError_t launch_level(const Geometry &g, const uint32_t *F1, const uint32_t *F2, const uint32_t *off_in,
                        const uint16_t *cost_in, uint32_t *off_out, uint16_t *cost_out, int p, int pd,
                        int rs_iteration, uint32_t pair, hipStream_t st)
{
    const RSArgs rs = make_rs(g, rs_iteration >= 0, (uint32_t)(rs_iteration < 0 ? 0 : rs_iteration), pair);
#define BM_LEVEL_CALL(RSV) level_px_t<NSP, RSV>(g, F1, F2, off_in, cost_in, off_out, cost_out, p, pd, rs, st)
    if (rs.enabled) {
        BM_NSP_DISPATCH(g.NSP, BM_LEVEL_CALL(true));
    } else {
        BM_NSP_DISPATCH(g.NSP, BM_LEVEL_CALL(false));
    }
#undef BM_LEVEL_CALL
    return hipGetLastError();
}

hipError_t launch_random_search(const Geometry &g, const uint32_t *F1, const uint32_t *F2, uint32_t *off,
                                uint16_t *cost, uint32_t iteration, uint32_t pair, hipStream_t st)
{
    const dim3 grid(blocks_rows(g.sx, g.sy));
    const RSArgs rs = make_rs(g, 1, iteration, pair);
    BM_NSP_DISPATCH(g.NSP, hipLaunchKernelGGL((k_random_search<NSP>), grid, dim3(kThreads), 0, st, F1, F2, off,
                                              cost, g.sx, g.sy, g.strip_rows, rs));
    return hipGetLastError();
}

hipError_t launch_init_proposal(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int first,
                                const uint32_t *prev, const int32_t *coff, int cw, int ch, int rs_iteration,
                                uint32_t pair, uint32_t *off, uint16_t *cost, hipStream_t st)
{
    const dim3 grid(blocks_rows(g.sx, g.sy));
    const RSArgs rs = make_rs(g, rs_iteration >= 0, (uint32_t)(rs_iteration < 0 ? 0 : rs_iteration), pair);
#define BM_INIT_PROPOSAL(RSV)                                                                                       \
    hipLaunchKernelGGL((k_init_proposal<NSP, RSV>), grid, dim3(kThreads), 0, st, F1, F2, first, prev, coff, cw, ch, \
                       g.W, g.H, g.sx, g.sy, g.strip_rows, rs, off, cost)
    if (rs.enabled) {
        BM_NSP_DISPATCH(g.NSP, BM_INIT_PROPOSAL(true));
    } else {
        BM_NSP_DISPATCH(g.NSP, BM_INIT_PROPOSAL(false));
    }
#undef BM_INIT_PROPOSAL
    return hipGetLastError();
}

hipError_t launch_local_search(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int radius, int samples,
                               uint32_t iteration, uint32_t pair, uint32_t *off, uint16_t *cost, hipStream_t st)
{
    if (samples <= 0) return hipSuccess;
    if (radius < 1) return hipErrorInvalidValue;
    const dim3 grid(blocks_rows(g.sx, g.sy));
    BM_NSP_DISPATCH(g.NSP, hipLaunchKernelGGL((k_local_search<NSP>), grid, dim3(kThreads), 0, st, F1, F2, g.sx, g.sy,
                                              g.strip_rows, radius, samples, iteration, pair, (uint32_t)g.seed,
                                              (uint32_t)(g.seed >> 32), off, cost));
    return hipGetLastError();
}

hipError_t launch_downsample(const Geometry &g, const uint32_t *off, float *U, float *V, hipStream_t st)
{
    const int identity = (g.sx == g.W && g.sy == g.H) ? 1 : 0;
    const float scx = (float)g.W / (float)g.sx, scy = (float)g.H / (float)g.sy;
    static_assert(kDsTileW * kDsTileH == kThreads, "one output pixel per thread");
    if (!identity && g.ds_lds > 0) {
        const int tiles = ((g.W + kDsTileW - 1) / kDsTileW) * ((g.H + kDsTileH - 1) / kDsTileH);
        hipLaunchKernelGGL(k_downsample_tile, dim3(tiles), dim3(kThreads), (size_t)g.ds_lds, st, off, g.nn_x, g.nn_y,
                           g.gauss, g.P, g.sx, g.sy, g.W, g.H, scx, scy, U, V);
    } else {
        hipLaunchKernelGGL(k_downsample, dim3(blocks_rows(g.W, g.H)), dim3(kThreads), 0, st, off, g.nn_x, g.nn_y,
                           g.gauss, g.P, g.sx, g.sy, g.W, g.H, scx, scy, identity, U, V);
    }
    return hipGetLastError();
}

template <int FS>
static hipError_t launch_median_t(const Geometry &g, const float *in0, const float *in1, float *out0, float *out1,
                                  int dir, hipStream_t st)
{
    const dim3 grid(blocks_rows(g.W, g.H), in1 ? 2 : 1);
    hipLaunchKernelGGL((k_median<FS>), grid, dim3(kThreads), 0, st, in0, in1, out0, out1, g.W, g.H, dir);
    return hipGetLastError();
}

hipError_t launch_median(const Geometry &g, const float *in0, const float *in1, float *out0, float *out1, int dir,
                         hipStream_t st)
{
    switch (g.fsize) {
#define BM_MED(FSV) case FSV: return launch_median_t<FSV>(g, in0, in1, out0, out1, dir, st);
    BM_MED(1) BM_MED(3) BM_MED(5) BM_MED(7) BM_MED(9) BM_MED(11) BM_MED(13) BM_MED(15) BM_MED(17) BM_MED(19)
    BM_MED(21) BM_MED(23) BM_MED(25) BM_MED(27) BM_MED(29) BM_MED(31) BM_MED(33) BM_MED(35) BM_MED(37)
    BM_MED(39) BM_MED(41) BM_MED(43) BM_MED(45) BM_MED(47)
#undef BM_MED
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trilateral(const Geometry &g, const float *Ui, const float *Vi, const float *Um, const float *Vm,
                             const float *I, float *Uo, float *Vo, hipStream_t st)
{
    const int s = g.tri_s;
    const int pitch = kTriTW + 2 * s;
    int BR = (kTriLdsBytes - 2 * s * (int)sizeof(float)) / (pitch * (int)sizeof(float4));
    if (BR < 1) return hipErrorInvalidValue;
    if (BR > g.tri_band) BR = g.tri_band;
    const size_t lds = (size_t)BR * pitch * sizeof(float4) + (size_t)2 * s * sizeof(float);
    const int gxb = (g.W + kTriTW - 1) / kTriTW;
    // tri_split > 1: that many work-groups per tile, each over a share of its
    // source rows, then the finishing kernel (partial sums in g.tri_part)
    const int K = g.tri_split > 1 ? g.tri_split : 1;
    if (K > 1 && !g.tri_part) return hipErrorInvalidValue;
    float *part = K > 1 ? g.tri_part : nullptr;
    switch (g.tri_groups) {
    case 0:
        hipLaunchKernelGGL(k_trilateral, dim3(gxb * ((g.H + kTriTH - 1) / kTriTH) * K), dim3(kThreads), lds, st, Ui,
                           Vi, Um, Vm, I, Uo, Vo, g.W, g.H, s, BR, g.tri_a2, g.tri_b, g.tri_spatial, K, part);
        break;
    case 2:
        hipLaunchKernelGGL(k_trilateral_mfma<2>, dim3(gxb * ((g.H + 7) / 8) * K), dim3(kThreads), lds, st, Ui, Vi, Um,
                           Vm, I, Uo, Vo, g.W, g.H, s, BR, g.tri_a2, g.tri_b, g.tri_spatial, K, part);
        break;
    case 4:
        hipLaunchKernelGGL(k_trilateral_mfma<4>, dim3(gxb * ((g.H + 15) / 16) * K), dim3(kThreads), lds, st, Ui, Vi,
                           Um, Vm, I, Uo, Vo, g.W, g.H, s, BR, g.tri_a2, g.tri_b, g.tri_spatial, K, part);
        break;
    default: return hipErrorInvalidValue;
    }
    if (K > 1) {
        const size_t WH = (size_t)g.W * g.H;
        hipLaunchKernelGGL(k_trilateral_finish, dim3((unsigned)((WH + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                           part, K, g.W, g.H, Uo, Vo);
    }
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// multi-pair launches: tile counts per job (the single-pair grids) and launchers

int gray_tiles(int n) { return (n + kThreads - 1) / kThreads; }

int brief_tiles(const Geometry &g)
{
    const int TH = brief_tile_rows(g.NSP);
    return ((g.sx + kBriefTW - 1) / kBriefTW) * ((g.sy + TH - 1) / TH);
}

int match_tiles(const Geometry &g, int px)
{
    const int span = kThreads * (px > 0 ? px : 1);
    return ((g.sx + span - 1) / span) * g.sy;
}

int downsample_tiles(const Geometry &g, bool tiled)
{
    if (tiled) return ((g.W + kDsTileW - 1) / kDsTileW) * ((g.H + kDsTileH - 1) / kDsTileH);
    return (int)blocks_rows(g.W, g.H);
}

int median_tiles(const Geometry &g) { return (int)blocks_rows(g.W, g.H); }

int trilateral_tiles(const Geometry &g, int groups)
{
    const int TH = groups == 0 ? kTriTH : 4 * groups;
    return ((g.W + kTriTW - 1) / kTriTW) * ((g.H + TH - 1) / TH);
}

size_t brief_lds(const Geometry &g)
{
    const int TH = brief_tile_rows(g.NSP);
    return brief_lds_bytes(g, g.brief_near, TH);
}

size_t trilateral_lds(const Geometry &g, int *band)
{
    const int s = g.tri_s, pitch = kTriTW + 2 * s;
    int BR = (kTriLdsBytes - 2 * s * (int)sizeof(float)) / (pitch * (int)sizeof(float4));
    if (BR > g.tri_band) BR = g.tri_band;
    if (band) *band = BR;
    if (BR < 1) return 0;
    return (size_t)BR * pitch * sizeof(float4) + (size_t)2 * s * sizeof(float);
}

static dim3 multi_grid(int tiles, int grid) { return dim3((unsigned)std::max(1, std::min(grid, tiles))); }

hipError_t launch_multi_gray(const GrayJob *jobs, const int *start, int njobs, int tiles, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_gray_multi, multi_grid(tiles, grid), dim3(kThreads), 0, st, jobs, start, njobs, tiles);
    return hipGetLastError();
}

template <int NSP, int RN>
static void brief_multi_t(const BriefJob *jobs, const int *start, int njobs, int tiles, int grid, size_t lds,
                          hipStream_t st)
{
    constexpr int PY = NSP <= 4 ? 8 : 4;
    hipLaunchKernelGGL((k_brief_multi<NSP, RN, PY>), multi_grid(tiles, grid), dim3(kThreads), lds, st, jobs, start,
                       njobs, tiles);
}

template <int NSP>
static hipError_t brief_multi_n(int RN, const BriefJob *jobs, const int *start, int njobs, int tiles, int grid,
                                size_t lds, hipStream_t st)
{
    switch (RN) {
    case 8: brief_multi_t<NSP, 8>(jobs, start, njobs, tiles, grid, lds, st); break;
    case 16: brief_multi_t<NSP, 16>(jobs, start, njobs, tiles, grid, lds, st); break;
    case 24: brief_multi_t<NSP, 24>(jobs, start, njobs, tiles, grid, lds, st); break;
    case 32: brief_multi_t<NSP, 32>(jobs, start, njobs, tiles, grid, lds, st); break;
    case 40: brief_multi_t<NSP, 40>(jobs, start, njobs, tiles, grid, lds, st); break;
    case 48: brief_multi_t<NSP, 48>(jobs, start, njobs, tiles, grid, lds, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_multi_brief(int NSP, int RN, const BriefJob *jobs, const int *start, int njobs, int tiles, int grid,
                              size_t lds, hipStream_t st)
{
    switch (NSP) {
    case 1: return brief_multi_n<1>(RN, jobs, start, njobs, tiles, grid, lds, st);
    case 2: return brief_multi_n<2>(RN, jobs, start, njobs, tiles, grid, lds, st);
    case 4: return brief_multi_n<4>(RN, jobs, start, njobs, tiles, grid, lds, st);
    case 8: return brief_multi_n<8>(RN, jobs, start, njobs, tiles, grid, lds, st);
    case 16: return brief_multi_n<16>(RN, jobs, start, njobs, tiles, grid, lds, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_multi_init(int NSP, const MatchJob *jobs, const int *start, int njobs, int tiles, int grid,
                             int rs_iteration, hipStream_t st)
{
    const dim3 gd = multi_grid(tiles, grid);
    if (rs_iteration >= 0) {
        BM_NSP_DISPATCH(NSP, hipLaunchKernelGGL((k_init_multi<NSP, true>), gd, dim3(kThreads), 0, st, jobs, start,
                                                njobs, tiles, rs_iteration));
    } else {
        BM_NSP_DISPATCH(NSP, hipLaunchKernelGGL((k_init_multi<NSP, false>), gd, dim3(kThreads), 0, st, jobs, start,
                                                njobs, tiles, rs_iteration));
    }
    return hipGetLastError();
}

template <int NSP, bool RS>
static hipError_t level_multi_px(int PX, const MatchJob *jobs, const int *start, int njobs, int tiles, int grid,
                                 int p, int pd, int src, int rs_iteration, hipStream_t st)
{
    const dim3 gd = multi_grid(tiles, grid);
    if constexpr (NSP <= 4) {
        if (PX == 4) {
            hipLaunchKernelGGL((k_level_multi<NSP, RS, 4>), gd, dim3(kThreads), 0, st, jobs, start, njobs, tiles, p,
                               pd, src, rs_iteration);
            return hipGetLastError();
        }
    }
    if (PX == 2 && NSP <= 8) {
        hipLaunchKernelGGL((k_level_multi<NSP, RS, (NSP <= 8 ? 2 : 1)>), gd, dim3(kThreads), 0, st, jobs, start,
                           njobs, tiles, p, pd, src, rs_iteration);
    } else if (PX == 1) {
        hipLaunchKernelGGL((k_level_multi<NSP, RS, 1>), gd, dim3(kThreads), 0, st, jobs, start, njobs, tiles, p, pd,
                           src, rs_iteration);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_multi_level(int NSP, int PX, const MatchJob *jobs, const int *start, int njobs, int tiles, int grid,
                              int p, int pd, int src, int rs_iteration, hipStream_t st)
{
    if (rs_iteration >= 0) {
        BM_NSP_DISPATCH(NSP, return (level_multi_px<NSP, true>(PX, jobs, start, njobs, tiles, grid, p, pd, src,
                                                               rs_iteration, st)));
    } else {
        BM_NSP_DISPATCH(NSP, return (level_multi_px<NSP, false>(PX, jobs, start, njobs, tiles, grid, p, pd, src,
                                                                rs_iteration, st)));
    }
    return hipErrorInvalidValue;
}

hipError_t launch_multi_downsample(bool tiled, const DsJob *jobs, const int *start, int njobs, int tiles, int grid,
                                   size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL(k_downsample_multi, multi_grid(tiles, grid), dim3(kThreads), tiled ? lds : 0, st, jobs, start,
                       njobs, tiles, tiled ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_multi_median(int FS, int dir, const MedJob *jobs, const int *start, int njobs, int tiles, int grid,
                               hipStream_t st)
{
    const dim3 gd = multi_grid(tiles, grid);
    switch (FS) {
#define BM_MED_MULTI(FSV)                                                                                      \
    case FSV:                                                                                                  \
        hipLaunchKernelGGL(k_median_multi<FSV>, gd, dim3(kThreads), 0, st, jobs, start, njobs, tiles, dir); \
        break;
    BM_MED_MULTI(1) BM_MED_MULTI(3) BM_MED_MULTI(5) BM_MED_MULTI(7) BM_MED_MULTI(9) BM_MED_MULTI(11)
    BM_MED_MULTI(13) BM_MED_MULTI(15) BM_MED_MULTI(17) BM_MED_MULTI(19) BM_MED_MULTI(21) BM_MED_MULTI(23)
    BM_MED_MULTI(25) BM_MED_MULTI(27) BM_MED_MULTI(29) BM_MED_MULTI(31) BM_MED_MULTI(33) BM_MED_MULTI(35)
    BM_MED_MULTI(37) BM_MED_MULTI(39) BM_MED_MULTI(41) BM_MED_MULTI(43) BM_MED_MULTI(45) BM_MED_MULTI(47)
#undef BM_MED_MULTI
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_multi_trilateral(int groups, const TriJob *jobs, const int *start, int njobs, int tiles, int grid,
                                   size_t lds, hipStream_t st)
{
    const dim3 gd = multi_grid(tiles, grid);
    switch (groups) {
    case 0: hipLaunchKernelGGL(k_trilateral_multi<0>, gd, dim3(kThreads), lds, st, jobs, start, njobs, tiles); break;
    case 2: hipLaunchKernelGGL(k_trilateral_multi<2>, gd, dim3(kThreads), lds, st, jobs, start, njobs, tiles); break;
    case 4: hipLaunchKernelGGL(k_trilateral_multi<4>, gd, dim3(kThreads), lds, st, jobs, start, njobs, tiles); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace bm
