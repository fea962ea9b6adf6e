// Learned PatchMatch proposal: a small convolutional network at 1/8 of the
// input resolution, BF16 weights and activations on MFMA with FP32
// accumulation, trained by the kernels in this file.
//
// Input features per coarse pixel (NHWC, 128 BF16 channels): a cost volume of
// 121 displacements of -5..5 coarse pixels in x and y (sum of absolute
// differences of 8x8 block sums over a 3x3 block window, clamped and scaled to
// [-1, 1]), the block means of both gray frames, the lowest-cost displacement
// (dx / 8, dy / 8, exact in BF16) and its cost, and zero padding.  Layers
// (3x3, zero padding):
//   128 -> 64 ReLU, 64 -> 64 ReLU dilation 2, 64 -> 64 ReLU dilation 4,
//   64 -> 64 ReLU, 64 -> 16 linear; channels 0, 1 plus the lowest-cost
//   displacement (a fixed skip connection) are the coarse flow in coarse
//   pixels, i.e. units of 8 input pixels.
// 193,808 parameters; the 14 unused output channels (8,078 parameters) get no
// gradient and stay at zero.
//
// Every conv is an implicit GEMM on v_mfma_f32_16x16x16_bf16: a wave owns
// 16 * PT pixels x all output channels; K runs over the 9 taps x input
// channels.  The backward pass reuses the same kernel with flipped, transposed
// weights (data gradient, masked by the ReLU of the layer below) and an
// LDS-staged MFMA GEMM over pixels for the weight gradient, reduced over
// fixed pixel chunks in a fixed order (bit-reproducible training).
#include <hip/hip_runtime.h>

#include "bm_net.h"

namespace bm {
namespace net {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (values are finite)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// num / den rounded half away from zero (den > 0)
__device__ __forceinline__ int rdiv_sym(int64_t num, int64_t den)
{
    return num >= 0 ? (int)((2 * num + den) / (2 * den)) : -(int)((-2 * num + den) / (2 * den));
}

// ---------------------------------------------------------------------------
// features

// 8x8 block sums of both gray frames: S[f][cy][cx], f = 0, 1
__global__ void __launch_bounds__(256) k_block_sums(const uint8_t *__restrict__ g1, const uint8_t *__restrict__ g2,
                                                    int W, int cw, int ch, int32_t *__restrict__ S)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 2 * cw * ch) return;
    const int f = i / (cw * ch), cp = i % (cw * ch), cx = cp % cw, cy = cp / cw;
    const uint8_t *g = f ? g2 : g1;
    int sum = 0;
    for (int yy = 0; yy < kCell; ++yy)
        for (int xx = 0; xx < kCell; ++xx) sum += g[(size_t)(kCell * cy + yy) * W + kCell * cx + xx];
    S[i] = sum;
}

// Feature map [cw * ch][128] BF16, one 128-thread work-group per coarse pixel;
// thread c computes channel c.  The argmin over the cost volume is a min over
// the key cost * 128 + c, so ties go to the lowest channel.
__global__ void __launch_bounds__(128) k_features(const int32_t *__restrict__ S, int cw, int ch,
                                                  uint16_t *__restrict__ feat)
{
    __shared__ int red[2];
    const int cp = blockIdx.x, c = threadIdx.x;
    const int cx = cp % cw, cy = cp / cw;
    const int32_t *S1 = S, *S2 = S + cw * ch;
    float v = 0.0f;
    int key = 0x7fffffff;
    if (c < kCost) {
        const int dx = c % (2 * kRadius + 1) - kRadius, dy = c / (2 * kRadius + 1) - kRadius;
        int cost = 0;
#pragma unroll
        for (int j = -1; j <= 1; ++j)
#pragma unroll
            for (int i = -1; i <= 1; ++i) {
                const int ay = min(max(cy + j, 0), ch - 1), ax = min(max(cx + i, 0), cw - 1);
                const int by = min(max(cy + j + dy, 0), ch - 1), bx = min(max(cx + i + dx, 0), cw - 1);
                cost += abs(S1[ay * cw + ax] - S2[by * cw + bx]);
            }
        v = __fdiv_rn((float)(min(cost, kSadClamp) - kSadClamp / 2), (float)(kSadClamp / 2));
        key = cost * 128 + c;
    } else if (c == kCost || c == kCost + 1) {
        v = __fdiv_rn((float)((c == kCost ? S1 : S2)[cp] - 8160), 16320.0f);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) key = min(key, __shfl_xor(key, o));
    if ((c & 63) == 0) red[c >> 6] = key;
    __syncthreads();
    const int best = min(red[0], red[1]);
    const int bc = best & 127, bcost = best >> 7;
    if (c == kCost + 2) v = (float)(bc % (2 * kRadius + 1) - kRadius) * 0.125f;
    if (c == kCost + 3) v = (float)(bc / (2 * kRadius + 1) - kRadius) * 0.125f;
    if (c == kCost + 4) v = __fdiv_rn((float)(min(bcost, kSadClamp) - kSadClamp / 2), (float)(kSadClamp / 2));
    feat[(size_t)cp * kFeatC + c] = f2bf(v);
}

// Far block match, one 128-thread work-group per 16x16 block (X, Y) of a
// cfw x cfh grid (cfw = cw / 2): the 16x16 block sums of the 3x3 window of
// frame 1 and of the 13x13 window of frame 2 around the block (clamped to
// the grid) are summed from four 8x8 block sums each into LDS; thread c
// scores displacement c = (dy + 5) * 11 + dx + 5.  The minimum of the key
// (SAD * 6 + ring) * 121 + c picks the lowest SAD, then the lowest ring
// max(|dx|, |dy|), then the lowest c (SAD <= 9 * 65280 keeps the key below
// 2^29, and key / 726 recovers the SAD).  The best displacement over all rings
// is reported when its SAD is below 3/4 of the best over rings < kFarMin,
// else the latter.
__global__ void __launch_bounds__(128) k_far_search(const int32_t *__restrict__ S, int cw, int ch,
                                                    int32_t *__restrict__ far)
{
    constexpr int R = kFarRadius, P = 2 * (R + 1) + 1;  // frame-2 window side
    constexpr int KD = (R + 1) * kFarN;                  // key divisor of the SAD
    __shared__ int s1[9], s2[P * P], red[2][2];
    const int cfw = cw >> 1, cfh = ch >> 1;
    const int b = blockIdx.x, X = b % cfw, Y = b / cfw, t = threadIdx.x;
    const int32_t *S2 = S + cw * ch;
    for (int k = t; k < 9 + P * P; k += 128) {
        const bool f1 = k < 9;
        const int q = f1 ? k : k - 9, side = f1 ? 3 : P, half = f1 ? 1 : R + 1;
        const int by = min(max(Y + q / side - half, 0), cfh - 1), bx = min(max(X + q % side - half, 0), cfw - 1);
        const int32_t *s = (f1 ? S : S2) + 2 * by * cw + 2 * bx;
        const int v = s[0] + s[1] + s[cw] + s[cw + 1];
        if (f1) s1[q] = v;
        else s2[q] = v;
    }
    __syncthreads();
    int key = 0x7fffffff, near = 0x7fffffff;
    if (t < kFarN) {
        const int dx = t % (2 * R + 1) - R, dy = t / (2 * R + 1) - R, ring = max(abs(dx), abs(dy));
        int cost = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) cost += abs(s1[3 * j + i] - s2[(j + dy + R) * P + i + dx + R]);
        key = (cost * (R + 1) + ring) * kFarN + t;
        if (ring < kFarMin) near = key;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        key = min(key, __shfl_xor(key, o));
        near = min(near, __shfl_xor(near, o));
    }
    if ((t & 63) == 0) {
        red[t >> 6][0] = key;
        red[t >> 6][1] = near;
    }
    __syncthreads();
    if (t == 0) {
        const int ka = min(red[0][0], red[1][0]), kn = min(red[0][1], red[1][1]);
        const int c = (4 * (ka / KD) < 3 * (kn / KD) ? ka : kn) % kFarN;
        far[2 * b] = c % (2 * R + 1) - R;
        far[2 * b + 1] = c / (2 * R + 1) - R;
    }
}

// ---------------------------------------------------------------------------
// conv (forward, and data gradient with flipped/transposed weights)
//
// MFMA 16x16x16 BF16 operand layout (wave64): lane l supplies A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16], j = 0..3, and receives D[4 (l / 16) + j][l % 16].
// A = activations (pixel x channel), B = weights (channel x output channel).

// SPLIT: the four waves of a work-group share 16 * PT pixels and each takes a
// quarter of the output-channel tiles (four times the waves for small maps);
// otherwise each wave owns its own pixels and all output channels.
template <int COUT, int PT, bool SPLIT>
__global__ void __launch_bounds__(256) k_conv(const uint16_t *__restrict__ in, int cin,
                                              const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                              int npix, int H, int W, int dil, int act,
                                              const uint16_t *__restrict__ ref, const uint16_t *__restrict__ skip,
                                              uint16_t *__restrict__ out_bf, float *__restrict__ out_f)
{
    static_assert(!SPLIT || COUT % 64 == 0, "split needs four output-channel tiles per wave group");
    constexpr int CT = SPLIT ? COUT / 64 : COUT / 16;  // output-channel tiles per wave
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int p0 = (SPLIT ? blockIdx.x : blockIdx.x * 4 + wave) * 16 * PT;
    const int ct0 = SPLIT ? wave * CT : 0;  // first output-channel tile of this wave
    const int hw = H * W;
    int py[PT], px[PT], pb[PT];
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int p = p0 + 16 * t + m;
        const int n = p / hw, rem = p - n * hw;
        py[t] = p < npix ? rem / W : -(1 << 20);  // out-of-range pixels never find a tap in the image
        px[t] = rem % W;
        pb[t] = n * hw;
    }
    v4f acc[PT][CT];
#pragma unroll
    for (int t = 0; t < PT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u) acc[t][u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 9; ++tap) {
        const int oy = (tap / 3 - 1) * dil, ox = (tap % 3 - 1) * dil;
        const uint16_t *src[PT];
        bool inb[PT];
#pragma unroll
        for (int t = 0; t < PT; ++t) {
            const int ty = py[t] + oy, tx = px[t] + ox;
            inb[t] = ty >= 0 && ty < H && tx >= 0 && tx < W;
            src[t] = in + (size_t)(inb[t] ? pb[t] + ty * W + tx : 0) * cin + 4 * kq;
        }
        const uint16_t *wt = w + (size_t)tap * COUT * cin + (size_t)(16 * ct0 + m) * cin + 4 * kq;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            v4s a[PT];
#pragma unroll
            for (int t = 0; t < PT; ++t) {
                a[t] = v4s{0, 0, 0, 0};
                if (inb[t]) a[t] = *reinterpret_cast<const v4s *>(src[t] + c0);
            }
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                const v4s b = *reinterpret_cast<const v4s *>(wt + (size_t)16 * u * cin + c0);
#pragma unroll
                for (int t = 0; t < PT; ++t)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], b, acc[t][u], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < PT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int q = p0 + 16 * t + 4 * kq + j, co = 16 * (ct0 + u) + m;
                if (q >= npix) continue;
                float v = acc[t][u][j] + (bias ? bias[co] : 0.0f);
                if (skip && co < 2) v += 8.0f * bf2f(skip[(size_t)q * kFeatC + kSkipChannel + co]);
                if (act == ACT_RELU) v = fmaxf(v, 0.0f);
                if (act == ACT_MASK && !(bf2f(ref[(size_t)q * COUT + co]) > 0.0f)) v = 0.0f;
                if (out_f) out_f[(size_t)q * COUT + co] = v;
                if (out_bf) out_bf[(size_t)q * COUT + co] = f2bf(v);
            }
}

// ---------------------------------------------------------------------------
// weight gradient: part[chunk][tap][co][ci] = sum over the chunk's pixels p of
// dz[p][co] * x[p + tap offset][ci] (zero outside the image); bias partials
// part_bias[chunk][co] = sum of dz[p][co].  Operands are staged transposed in
// LDS ([channel][pixel]) so each lane reads its four K values in one load.

constexpr int kWgPix = 64;      // pixels per LDS stage
constexpr int kWgChunk = 2048;  // pixels per work-group
constexpr int kWgKP = kWgPix + 4;

template <int COUT, int CIN>
__global__ void __launch_bounds__(256) k_wgrad(const uint16_t *__restrict__ dz, const uint16_t *__restrict__ x,
                                               int npix, int H, int W, int dil, float *__restrict__ part,
                                               float *__restrict__ part_bias)
{
    __shared__ __attribute__((aligned(16))) uint16_t sdz[COUT * kWgKP];
    __shared__ __attribute__((aligned(16))) uint16_t sxx[CIN * kWgKP];
    constexpr int TILES = (COUT / 16) * (CIN / 16);
    constexpr int TPW = TILES / 4;  // tiles per wave
    static_assert(TILES % 4 == 0 && (CIN / 16) % TPW == 0, "each wave's tiles lie in one output-channel block");
    const int chunk = blockIdx.x, tap = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int hw = H * W;
    const int oy = (tap / 3 - 1) * dil, ox = (tap % 3 - 1) * dil;
    v4f acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const int pbeg = chunk * kWgChunk, pend = min(npix, pbeg + kWgChunk);
    for (int s0 = pbeg; s0 < pend; s0 += kWgPix) {
        __syncthreads();
        // four channels (8 bytes) per load, transposed into LDS
#pragma unroll 2
        for (int i = threadIdx.x; i < kWgPix * COUT / 4; i += 256) {
            const int k = i / (COUT / 4), c4 = 4 * (i % (COUT / 4)), pp = s0 + k;
            v4s v = {0, 0, 0, 0};
            if (pp < pend) v = *reinterpret_cast<const v4s *>(dz + (size_t)pp * COUT + c4);
#pragma unroll
            for (int j = 0; j < 4; ++j) sdz[(c4 + j) * kWgKP + k] = (uint16_t)v[j];
        }
#pragma unroll 2
        for (int i = threadIdx.x; i < kWgPix * CIN / 4; i += 256) {
            const int k = i / (CIN / 4), c4 = 4 * (i % (CIN / 4)), pp = s0 + k;
            v4s v = {0, 0, 0, 0};
            if (pp < pend) {
                const int n = pp / hw, rem = pp - n * hw, yy = rem / W + oy, xx = rem % W + ox;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    v = *reinterpret_cast<const v4s *>(x + ((size_t)n * hw + yy * W + xx) * CIN + c4);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) sxx[(c4 + j) * kWgKP + k] = (uint16_t)v[j];
        }
        __syncthreads();
        if (tap == 0 && threadIdx.x < COUT)
            for (int k = 0; k < kWgPix; ++k) bsum += bf2f(sdz[threadIdx.x * kWgKP + k]);
        // a wave's tiles share one output-channel block: one A fragment per K step
        const int ct = wave * TPW / (CIN / 16);
#pragma unroll 1
        for (int ks = 0; ks < kWgPix; ks += 16) {
            const v4s a = *reinterpret_cast<const v4s *>(&sdz[(16 * ct + m) * kWgKP + ks + 4 * kq]);
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int it = (wave * TPW + i) % (CIN / 16);
                const v4s b = *reinterpret_cast<const v4s *>(&sxx[(16 * it + m) * kWgKP + ks + 4 * kq]);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i], 0, 0, 0);
            }
        }
    }
    float *dst = part + ((size_t)chunk * 9 + tap) * COUT * CIN;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int t = wave * TPW + i;
        const int ct = t / (CIN / 16), it = t % (CIN / 16);
#pragma unroll
        for (int j = 0; j < 4; ++j) dst[(size_t)(16 * ct + 4 * kq + j) * CIN + 16 * it + m] = acc[i][j];
    }
    if (tap == 0 && threadIdx.x < COUT) part_bias[(size_t)chunk * COUT + threadIdx.x] = bsum;
}

// out[i] = sum over c of part[c][i], c in increasing order
__global__ void k_reduce(const float *__restrict__ part, int chunks, int n, float *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * n + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// loss, optimiser, weight packing

// Mean absolute error over both flow channels (mean over 2 x pixels); labels
// that are not finite contribute nothing.  pred [npix][16] FP32, label [npix][2].
// dz receives the gradient (BF16), loss_part one partial sum per block.
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ pred, const float *__restrict__ label,
                                              int npix, uint16_t *__restrict__ dz, float *__restrict__ loss_part)
{
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < npix) {
        const float inv = 1.0f / (2.0f * (float)npix);
#pragma unroll
        for (int c = 0; c < kOutC; ++c) {
            float g = 0.0f;
            if (c < 2) {
                const float t = label[(size_t)p * 2 + c];
                if (isfinite(t)) {
                    const float d = pred[(size_t)p * kOutC + c] - t;
                    l += fabsf(d);
                    g = d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f);
                }
            }
            dz[(size_t)p * kOutC + c] = f2bf(g);
        }
        l *= inv;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// out[0] = sum of x[0..n) by one work-group in a fixed order
__global__ void __launch_bounds__(256) k_sum(const float *__restrict__ x, int n, float *__restrict__ out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ mom, float *__restrict__ vel,
                       const float *__restrict__ g, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * mom[i] + (1.0f - b1) * gi;
    const float vi = b2 * vel[i] + (1.0f - b2) * gi * gi;
    mom[i] = mi;
    vel[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// W [9][cout][cin] FP32 -> wb [9][cout][cin] BF16 and wt [9][cin][cout] BF16
// with the taps reversed (the weights of the data-gradient conv).
__global__ void k_pack(const float *__restrict__ w, int cout, int cin, uint16_t *__restrict__ wb,
                       uint16_t *__restrict__ wt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 9 * cout * cin) return;
    const int tap = i / (cout * cin), r = i % (cout * cin), co = r / cin, ci = r % cin;
    const uint16_t v = f2bf(w[i]);
    wb[i] = v;
    wt[((size_t)(8 - tap) * cin + ci) * cout + co] = v;
}

// ---------------------------------------------------------------------------
// labels and the proposal

// 8x8 block means of a flow field in coarse pixels (sum / 64 / 8): out [ch][cw][2]
__global__ void k_coarse_flow(const float *__restrict__ U, const float *__restrict__ V, int W, int cw, int ch,
                              float *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cw * ch) return;
    const int cx = i % cw, cy = i / cw;
    float su = 0.0f, sv = 0.0f;
    for (int yy = 0; yy < kCell; ++yy)
        for (int xx = 0; xx < kCell; ++xx) {
            const size_t o = (size_t)(kCell * cy + yy) * W + kCell * cx + xx;
            su += U[o];
            sv += V[o];
        }
    out[2 * i] = su * (1.0f / 512.0f);
    out[2 * i + 1] = sv * (1.0f / 512.0f);
}

// Per coarse cell, two proposals as matched-pixel offsets in Q8: the network's
// coarse flow (output channels 0, 1) and the lowest-cost displacement of the
// feature map (channels 123, 124 times 8), each rint(clamp(u * scale)),
// scale = 256 * 8 * sx / W; NaN maps to -lim.  coff [n][4] = (net u, net v,
// argmin u, argmin v).  With a far match, a cell whose 16-pixel block (its
// coordinates halved, clamped to the far grid) matched at ring >= kFarMin
// takes (2 dx, 2 dy) coarse pixels in place of the network's flow.
__global__ void k_proposal_offsets(const float *__restrict__ out, const uint16_t *__restrict__ feat,
                                   const int32_t *__restrict__ far, int cw, int ch, float scale_x, float scale_y,
                                   int lim_x, int lim_y, int32_t *__restrict__ coff)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cw * ch) return;
    float src[4] = {out[(size_t)i * kOutC], out[(size_t)i * kOutC + 1],
                    8.0f * bf2f(feat[(size_t)i * kFeatC + kSkipChannel]),
                    8.0f * bf2f(feat[(size_t)i * kFeatC + kSkipChannel + 1])};
    if (far) {
        const int cfw = cw >> 1, cfh = ch >> 1;
        const int X = min((i % cw) >> 1, cfw - 1), Y = min((i / cw) >> 1, cfh - 1);
        const int dx = far[2 * (Y * cfw + X)], dy = far[2 * (Y * cfw + X) + 1];
        if (max(abs(dx), abs(dy)) >= kFarMin) {
            src[0] = (float)(2 * dx);
            src[1] = (float)(2 * dy);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float sc = (k & 1) ? scale_y : scale_x, lim = (float)((k & 1) ? lim_y : lim_x);
        coff[4 * i + k] = (int32_t)rintf(fminf(fmaxf(src[k] * sc, -lim), lim));
    }
}

// dst[r] = src[rl.idx[r]] for rows of row_words 32-bit words
__global__ void k_gather_rows(const uint32_t *__restrict__ src, const RowList rl, size_t row_words, int rows,
                              uint32_t *__restrict__ dst)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= row_words * rows) return;
    const size_t r = i / row_words, c = i % row_words;
    dst[i] = src[(size_t)rl.idx[r] * row_words + c];
}

// ---------------------------------------------------------------------------
// launchers

hipError_t launch_features(const uint8_t *g1, const uint8_t *g2, int W, int cw, int ch, int32_t *S, uint16_t *feat,
                           hipStream_t st)
{
    static_assert(kFeatC == 128 && kCost + 5 <= kFeatC, "one thread per channel, two waves");
    hipLaunchKernelGGL(k_block_sums, dim3((2 * cw * ch + 255) / 256), dim3(256), 0, st, g1, g2, W, cw, ch, S);
    hipLaunchKernelGGL(k_features, dim3((unsigned)cw * ch), dim3(kFeatC), 0, st, S, cw, ch, feat);
    return hipGetLastError();
}

hipError_t launch_block_sums(const uint8_t *g1, const uint8_t *g2, int W, int cw, int ch, int32_t *S, hipStream_t st)
{
    hipLaunchKernelGGL(k_block_sums, dim3((2 * cw * ch + 255) / 256), dim3(256), 0, st, g1, g2, W, cw, ch, S);
    return hipGetLastError();
}

hipError_t launch_far_search(const int32_t *S, int cw, int ch, int32_t *far, hipStream_t st)
{
    static_assert(kFarN <= 128, "one thread per displacement");
    if (cw < 2 || ch < 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_far_search, dim3((unsigned)(cw >> 1) * (ch >> 1)), dim3(128), 0, st, S, cw, ch, far);
    return hipGetLastError();
}

template <int COUT, int PT, bool SPLIT>
static void conv_t(const uint16_t *in, int cin, const uint16_t *w, const float *bias, int npix, int H, int W, int dil,
                   int act, const uint16_t *ref, const uint16_t *skip, uint16_t *out_bf, float *out_f, hipStream_t st)
{
    const int per_block = (SPLIT ? 1 : 4) * 16 * PT;
    hipLaunchKernelGGL((k_conv<COUT, PT, SPLIT>), dim3((npix + per_block - 1) / per_block), dim3(256), 0, st, in, cin,
                       w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f);
}

hipError_t launch_conv(int cout, const uint16_t *in, int cin, const uint16_t *w, const float *bias, int npix, int H,
                       int W, int dil, int act, const uint16_t *ref, const uint16_t *skip, uint16_t *out_bf,
                       float *out_f, hipStream_t st)
{
    if (cin % 16 || npix <= 0) return hipErrorInvalidValue;
    // enough waves to cover the chip before blocking pixels per wave; below
    // 16384 pixels the four waves of a group split the output channels
    // (measured: 4,800 pixels 31 -> 23 us, 32,400 pixels 55 -> 83 us)
    const int pt = npix >= 4 * 64 * 1024 ? 4 : (npix >= 2 * 64 * 512 ? 2 : 1);
    if (cout == 64) {
        if (pt == 4) conv_t<64, 4, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
        else if (pt == 2) conv_t<64, 2, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
        else if (npix >= 16384) conv_t<64, 1, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
        else conv_t<64, 1, true>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
    } else if (cout == 16) {
        if (pt == 4) conv_t<16, 4, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
        else if (pt == 2) conv_t<16, 2, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
        else conv_t<16, 1, false>(in, cin, w, bias, npix, H, W, dil, act, ref, skip, out_bf, out_f, st);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int wgrad_chunks(int npix) { return (npix + kWgChunk - 1) / kWgChunk; }

hipError_t launch_wgrad(int cout, int cin, const uint16_t *dz, const uint16_t *x, int npix, int H, int W, int dil,
                        float *part, float *part_bias, float *grad_w, float *grad_b, hipStream_t st)
{
    const int chunks = wgrad_chunks(npix);
    const dim3 grid(chunks, 9);
    if (cout == 64 && cin == 128)
        hipLaunchKernelGGL((k_wgrad<64, 128>), grid, dim3(256), 0, st, dz, x, npix, H, W, dil, part, part_bias);
    else if (cout == 64 && cin == 64)
        hipLaunchKernelGGL((k_wgrad<64, 64>), grid, dim3(256), 0, st, dz, x, npix, H, W, dil, part, part_bias);
    else if (cout == 16 && cin == 64)
        hipLaunchKernelGGL((k_wgrad<16, 64>), grid, dim3(256), 0, st, dz, x, npix, H, W, dil, part, part_bias);
    else
        return hipErrorInvalidValue;
    const int nw = 9 * cout * cin;
    hipLaunchKernelGGL(k_reduce, dim3((nw + 255) / 256), dim3(256), 0, st, part, chunks, nw, grad_w);
    hipLaunchKernelGGL(k_reduce, dim3((cout + 255) / 256), dim3(256), 0, st, part_bias, chunks, cout, grad_b);
    return hipGetLastError();
}

int loss_blocks(int npix) { return (npix + 255) / 256; }

hipError_t launch_loss(const float *pred, const float *label, int npix, uint16_t *dz, float *loss_part, hipStream_t st)
{
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks(npix)), dim3(256), 0, st, pred, label, npix, dz, loss_part);
    return hipGetLastError();
}

hipError_t launch_sum(const float *x, int n, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, x, n, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, w, m, v, g, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_pack(const float *w, int cout, int cin, uint16_t *wb, uint16_t *wt, hipStream_t st)
{
    const int n = 9 * cout * cin;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, st, w, cout, cin, wb, wt);
    return hipGetLastError();
}

hipError_t launch_coarse_flow(const float *U, const float *V, int W, int cw, int ch, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_coarse_flow, dim3((cw * ch + 255) / 256), dim3(256), 0, st, U, V, W, cw, ch, out);
    return hipGetLastError();
}

hipError_t launch_proposal_offsets(const float *net_out, const uint16_t *feat, const int32_t *far, int cw, int ch,
                                   float scale_x, float scale_y, int lim_x, int lim_y, int32_t *coff, hipStream_t st)
{
    if (far && (cw < 2 || ch < 2)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_proposal_offsets, dim3((cw * ch + 255) / 256), dim3(256), 0, st, net_out, feat, far, cw, ch,
                       scale_x, scale_y, lim_x, lim_y, coff);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const void *src, const RowList &rl, size_t row_bytes, int rows, void *dst,
                              hipStream_t st)
{
    if (row_bytes % 4 || rows <= 0 || rows > kMaxBatch) return hipErrorInvalidValue;
    const size_t words = row_bytes / 4, n = words * rows;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       (const uint32_t *)src, rl, words, rows, (uint32_t *)dst);
    return hipGetLastError();
}

hipError_t forward(const Packed &pk, const uint16_t *feat, int n, int ch, int cw, uint16_t *act0, uint16_t *act1,
                   float *out, hipStream_t st)
{
    const int npix = n * ch * cw;
    const uint16_t *x = feat;
    uint16_t *bufs[2] = {act0, act1};
    for (int l = 0; l < kLayers; ++l) {
        const LayerSpec &s = kSpec[l];
        const bool last = l == kLayers - 1;
        uint16_t *y = last ? nullptr : bufs[l & 1];
        hipError_t e = launch_conv(s.cout, x, s.cin, pk.w[l], pk.b[l], npix, ch, cw, s.dil,
                                   s.relu ? ACT_RELU : ACT_NONE, nullptr, last ? feat : nullptr, y,
                                   last ? out : nullptr, st);
        if (e != hipSuccess) return e;
        x = y;
    }
    return hipSuccess;
}

}  // namespace net
}  // namespace bm
