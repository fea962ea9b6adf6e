/*
 * CPU oracle for the BriefMatch flow path -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar C++ with OpenMP over rows (static schedule), no hand SIMD.  Built with
 * -ffp-contract=off so every float operation rounds on its own.  Each function
 * names the reference code it restates.  Deterministic rules that replace the
 * reference's racy or unseeded behaviour are stated in the comment of the
 * function that applies them:
 *   - gray conversion: integer BT.601, (77R + 150G + 29B + 128) >> 8;
 *   - up-sampling: bilinear in 8-bit fixed point (the reference uses Lanczos-3);
 *   - BRIEF pattern: Philox4x32-10 + a fixed-point inverse-normal table instead
 *     of a time-seeded Boost mt19937;
 *   - propagation: level-synchronous (each jump-flooding level reads a snapshot);
 *   - random search: stateless Philox4x32-10 instead of per-pixel cuRAND state;
 *   - median: clamp-to-edge borders (the reference reads out of bounds).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark load
 * this library, and only as the checker.  The product never links it.
 */
#include "oracle_bm.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int32_t kNdtriQ16[4096] = {
#include "../briefmatch_amd/csrc/bm_ndtri_q16.inc"
};

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

inline uint32_t pack_off(int dx, int dy)
{
    return (uint32_t)(uint16_t)(int16_t)dx | ((uint32_t)(uint16_t)(int16_t)dy << 16);
}
inline int off_dx(uint32_t o) { return (int16_t)(o & 0xffffu); }
inline int off_dy(uint32_t o) { return (int16_t)(o >> 16); }

/* Hamming distance between descriptor words: the integer form of featDist
 * (kernels/briefmatch_cuda.cu:60-80).  The reference returns dist/N as float and
 * exits early once dist exceeds N*dmax; since fl(a/N) is strictly monotone in the
 * integer a for a <= 2048, "accept iff dist/N < D" is "accept iff dist < cost". */
inline int hamming(const uint32_t *a, const uint32_t *b, int NS)
{
    int d = 0;
    for (int w = 0; w < NS; ++w) d += __builtin_popcount(a[w] ^ b[w]);
    return d;
}

/* Round-half-away-from-zero of a Q32 fixed-point product. */
inline int round_q32(int64_t v)
{
    int64_t a = v < 0 ? -v : v;
    int64_t r = (a + (int64_t(1) << 31)) >> 32;
    return (int)(v < 0 ? -r : r);
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the parallel regions that follow (timing at a chosen core count). */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11):
 * ten rounds of two 32x32->64 multiplies with the published multipliers and
 * Weyl key increments. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* BMParams defaults (briefmatch.h:67-84) followed by the CLI derivation of
 * of_pipeline.cpp:178-190 with its exact mixed float/double expression types,
 * and the Gaussian tap count of briefmatch.cpp:153-154. */
int orc_derive_params(unsigned W, unsigned H, float ux, float uy, float patchRad,
                      float sigmaRad, float fsizeRad, float sSRad, orc_params *p)
{
    p->iterations = 2;
    p->jfMax = 256;
    p->doubleProp = 1;
    p->filter = 1;
    p->tInit = 1;
    p->N = 128;
    p->sEPE = 0.4f;
    p->sI = 0.1f;
    p->maxmot = 5.0f;
    if (ux < 0.1f) ux = 2.740f;
    if (uy < 0.1f) uy = 3.052f;
    unsigned sx = (unsigned)(ux * (float)W);
    unsigned sy = (unsigned)(uy * (float)H);
    p->sx = sx;
    p->sy = sy;
    p->sx_in = W;
    p->sy_in = H;
    double diag = std::sqrt((double)(sx * sx + sy * sy));
    double diag_in = std::sqrt((double)(W * W + H * H));
    p->patchArea = (float)((double)(0.01f * patchRad) * diag);
    p->sigma = (float)((double)(0.01f * sigmaRad) * diag);
    p->fsize = (unsigned)(2.0 * std::round((double)(0.005f * fsizeRad) * diag_in) + 1.0);
    p->sS = (float)((double)(0.01f * sSRad) * diag_in);
    p->P = (unsigned)(std::ceil(p->sigma * 3.0f) * 2.0f + 1.0f);
    return 0;
}

/* BRIEF test pattern, the law of briefmatch.cpp:124-148 with a seeded stream:
 * for test s, Philox4x32-10(ctr = {s, 0, 0, 1}, key = seed) gives four words;
 * word c gives v = 0.2 * Q[word >> 20] / 65536 (a N(0, 0.2) draw), then
 * val = sign(v) v^2 clamped to [-1, 1], offset = trunc(patchArea * val) -- the
 * truncation the kernel applies at kernels/briefmatch_cuda.cu:110-113.
 * Layout: pattern[4s + {0,1,2,3}] = (ox1, oy1, ox2, oy2). */
void orc_make_pattern(uint64_t seed, unsigned N, float patchArea, int16_t *pattern)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (unsigned s = 0; s < N; ++s) {
        const uint32_t ctr[4] = {s, 0u, 0u, 1u};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        for (int c = 0; c < 4; ++c) {
            float v = 0.2f * ((float)kNdtriQ16[r[c] >> 20] * (1.0f / 65536.0f));
            float sq = v * v;
            float val = v > 0.0f ? sq : -sq;
            val = std::max(-1.0f, std::min(1.0f, val));
            float o = patchArea * val;
            pattern[4 * s + c] = (int16_t)(int)o;
        }
    }
}

/* Normalised Gaussian down-sampling kernel, briefmatch.cpp:153-172, with the
 * size limit of briefmatch.cpp:90-96 (CONST_KERNEL_SIZE = 100). */
int orc_gauss_kernel(float sigma, float *kernel, unsigned *P)
{
    unsigned n = (unsigned)(std::ceil(sigma * 3.0f) * 2.0f + 1.0f);
    if (n > 100) return -1;
    int P2 = (int)n / 2;
    float w = 0.0f;
    for (int i = -P2; i <= P2; ++i) {
        float e = -(float)(i * i) / (2.0f * sigma * sigma);
        kernel[i + P2] = std::exp(e);
        w += kernel[i + P2];
    }
    for (unsigned i = 0; i < n; ++i) kernel[i] /= w;
    *P = n;
    return 0;
}

/* Random-search standard deviation, briefmatch.cpp:271. */
float orc_rs_sigma(unsigned sx) { return std::sqrt(std::ceil((float)sx * 0.5f / 3.0f)); }

int64_t orc_rs_sigma_q16(unsigned sx)
{
    return (int64_t)std::llround((double)orc_rs_sigma(sx) * 65536.0);
}

/* Gray conversion of interleaved RGB: integer BT.601 weights (the weights of the
 * reference's grayscale ingest, of_pipeline.cpp:161,269). */
void orc_gray(const uint8_t *rgb, int W, int H, uint8_t *gray)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const uint8_t *px = rgb + 3 * ((size_t)y * W + x);
            gray[(size_t)y * W + x] = (uint8_t)((77u * px[0] + 150u * px[1] + 29u * px[2] + 128u) >> 8);
        }
}

/* Up-sampling W x H -> sx x sy with the reference's coordinate map
 * xp = x * W / sx (kernels/util_cuda.cu:134-135), bilinear in fixed point:
 * X = floor(x W 256 / sx), integer part X >> 8, 8-bit weight X & 255, right /
 * lower neighbour clamped to the image.  The weighted sum is an integer below
 * 2^24; L = sum / 65536 is exact in FP32. */
void orc_upsample(const uint8_t *gray, int W, int H, int sx, int sy, float *L)
{
    std::vector<int> x0(sx), x1(sx), ax(sx);
    for (int x = 0; x < sx; ++x) {
        uint64_t X = (uint64_t)x * (uint64_t)W * 256u / (uint64_t)sx;
        x0[x] = (int)(X >> 8);
        ax[x] = (int)(X & 255u);
        x1[x] = std::min(x0[x] + 1, W - 1);
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y) {
        uint64_t Y = (uint64_t)y * (uint64_t)H * 256u / (uint64_t)sy;
        int y0 = (int)(Y >> 8), ay = (int)(Y & 255u), y1 = std::min(y0 + 1, H - 1);
        const uint8_t *r0 = gray + (size_t)y0 * W, *r1 = gray + (size_t)y1 * W;
        for (int x = 0; x < sx; ++x) {
            uint32_t top = r0[x0[x]] * (256u - ax[x]) + r0[x1[x]] * (uint32_t)ax[x];
            uint32_t bot = r1[x0[x]] * (256u - ax[x]) + r1[x1[x]] * (uint32_t)ax[x];
            uint32_t v = top * (256u - ay) + bot * (uint32_t)ay;
            L[(size_t)y * sx + x] = (float)v * (1.0f / 65536.0f);
        }
    }
}

/* Lanczos-3 up-sampling, resizeLanczos (kernels/util_cuda.cu:126-158): output
 * pixel x samples source coordinate xp = x * float(W) / sx (float, as :133-134),
 * x1 = floor(xp), taps i = x1 - 2 .. x1 + 3 with weight 3 sin(t) sin(t / 3) / t^2,
 * t = pi (xp - i) (1 at t = 0), sources clamped to the frame (:153), the sum
 * divided by the sum of the weights (:157).  The weights are separable, so each
 * axis gets a table of six weights (computed in double, normalised, then
 * rounded to float) and the 36-tap sum is taken as
 * sum_j wy_j * (sum_i wx_i * g[j][i]) with fused multiply-adds in increasing i
 * and j; the device kernel uses the same tables and order. */
void orc_lanczos_table(int in, int out, int *first, float *w)
{
    const double pi = 3.14159265358979323846;
    for (int x = 0; x < out; ++x) {
        const float xp = (float)x * (float)in / (float)out;
        const int x1 = (int)std::floor(xp);
        double ws[6], sum = 0.0;
        for (int k = 0; k < 6; ++k) {
            const double t = pi * ((double)xp - (double)(x1 - 2 + k));
            ws[k] = t == 0.0 ? 1.0 : 3.0 * std::sin(t) * std::sin(t / 3.0) / (t * t);
            sum += ws[k];
        }
        first[x] = x1 - 2;
        for (int k = 0; k < 6; ++k) w[6 * x + k] = (float)(ws[k] / sum);
    }
}

void orc_upsample_lanczos(const uint8_t *gray, int W, int H, int sx, int sy, float *L)
{
    std::vector<int> fx(sx), fy(sy);
    std::vector<float> wx(6 * (size_t)sx), wy(6 * (size_t)sy);
    orc_lanczos_table(W, sx, fx.data(), wx.data());
    orc_lanczos_table(H, sy, fy.data(), wy.data());
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y) {
        for (int x = 0; x < sx; ++x) {
            float v = 0.0f;
            for (int j = 0; j < 6; ++j) {
                const uint8_t *row = gray + (size_t)std::min(std::max(fy[y] + j, 0), H - 1) * W;
                float h = 0.0f;
                for (int i = 0; i < 6; ++i)
                    h = std::fmaf(wx[6 * (size_t)x + i], (float)row[std::min(std::max(fx[x] + i, 0), W - 1)], h);
                v = std::fmaf(wy[6 * (size_t)y + j], h, v);
            }
            L[(size_t)y * sx + x] = v;
        }
    }
}

/* BRIEF descriptors, binaryFeatures (kernels/briefmatch_cuda.cu:82-124): bit
 * s % 32 of word s / 32 is I[p + o1] > I[p + o2] (strict), each offset clamped
 * per axis so the sample stays inside the image (:110-113).  Bits >= N are 0.
 * Layout F[(y * sx + x) * NS + w], NS = ceil(N / 32) (:91-92). */
void orc_brief(const float *L, int sx, int sy, const int16_t *pattern, int N, uint32_t *F)
{
    const int NS = (N + 31) / 32;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x) {
            uint32_t *f = F + ((size_t)y * sx + x) * NS;
            for (int w = 0; w < NS; ++w) f[w] = 0u;
            for (int s = 0; s < N; ++s) {
                int ox1 = clampi(pattern[4 * s + 0], -x, sx - x - 1);
                int oy1 = clampi(pattern[4 * s + 1], -y, sy - y - 1);
                int ox2 = clampi(pattern[4 * s + 2], -x, sx - x - 1);
                int oy2 = clampi(pattern[4 * s + 3], -y, sy - y - 1);
                float i1 = L[(size_t)(y + oy1) * sx + (x + ox1)];
                float i2 = L[(size_t)(y + oy2) * sx + (x + ox2)];
                if (i1 > i2) f[s / 32] |= 1u << (s % 32);
            }
        }
}

/* Field initialisation, init (kernels/briefmatch_cuda.cu:149-184).  first: offset
 * (0,0) with its cost.  Otherwise (temporal): follow the previous field once,
 * n = prev[p + prev[p]], clamped so the target stays in the image, read from the
 * previous field as a snapshot (the reference reads it while overwriting it).
 * The first hop's offset is clamped to the image too, which is a no-op for any
 * field this library produces. */
void orc_init(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS, int first,
              const uint32_t *prev_off, uint32_t *off, uint16_t *cost)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x) {
            size_t idx = (size_t)y * sx + x;
            int nx = 0, ny = 0;
            if (!first) {
                uint32_t o0 = prev_off[idx];
                int dx0 = clampi(off_dx(o0), -x, sx - x - 1);
                int dy0 = clampi(off_dy(o0), -y, sy - y - 1);
                uint32_t o1 = prev_off[idx + (std::ptrdiff_t)dy0 * sx + dx0];
                nx = clampi(off_dx(o1), -x, sx - x - 1);
                ny = clampi(off_dy(o1), -y, sy - y - 1);
            }
            size_t t = idx + (std::ptrdiff_t)ny * sx + nx;
            off[idx] = pack_off(nx, ny);
            cost[idx] = (uint16_t)hamming(F1 + idx * NS, F2 + t * NS, NS);
        }
}

/* One jump-flooding level of propagate (kernels/briefmatch_cuda.cu:186-220),
 * level-synchronous: every pixel reads candidates from the level-start snapshot
 * (off_in/cost_in) and writes off_out/cost_out.  Candidate order is d in
 * {-1, +1} ({+1} without doubleProp), then xy in {0 (vertical), 1 (horizontal)};
 * the neighbour is the linear index idx + pd*p*d*(xy ? 1 : sx) (:212), skipped
 * when outside [0, sx*sy) (in-range row wrap kept, as in the reference).  The
 * neighbour's offset is clamped to keep the target inside the image (:213-214)
 * and accepted iff its cost is strictly lower than the pixel's running best
 * (insert, :126-147). */
void orc_propagate_level(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                         int p, int pd, int doubleProp,
                         const uint32_t *off_in, const uint16_t *cost_in,
                         uint32_t *off_out, uint16_t *cost_out)
{
    const int64_t M = (int64_t)sx * sy;
    const int dstart = doubleProp ? -1 : 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x) {
            int64_t idx = (int64_t)y * sx + x;
            uint32_t best = off_in[idx];
            int bcost = cost_in[idx];
            for (int d = dstart; d <= 1; d += 2)
                for (int xy = 0; xy <= 1; ++xy) {
                    int64_t doff = (int64_t)pd * p * d * (xy ? 1 : sx);
                    int64_t j = idx + doff;
                    if (j < 0 || j >= M) continue;
                    uint32_t o = off_in[j];
                    int nx = clampi(off_dx(o), -x, sx - x - 1);
                    int ny = clampi(off_dy(o), -y, sy - y - 1);
                    uint32_t cand = pack_off(nx, ny);
                    if (cand == best) continue;
                    int64_t t = idx + (int64_t)ny * sx + nx;
                    int c = hamming(F1 + idx * NS, F2 + t * NS, NS);
                    if (c < bcost) {
                        bcost = c;
                        best = cand;
                    }
                }
            off_out[idx] = best;
            cost_out[idx] = (uint16_t)bcost;
        }
}

/* All levels p = jfMax, jfMax/2, ..., 1 of one propagate call (:206). */
void orc_propagate(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                   int pd, int jfMax, int doubleProp, uint32_t *off, uint16_t *cost)
{
    const size_t M = (size_t)sx * sy;
    std::vector<uint32_t> o2(M);
    std::vector<uint16_t> c2(M);
    for (int p = jfMax; p > 0; p /= 2) {
        orc_propagate_level(F1, F2, sx, sy, NS, p, pd, doubleProp, off, cost, o2.data(), c2.data());
        std::memcpy(off, o2.data(), M * sizeof(uint32_t));
        std::memcpy(cost, c2.data(), M * sizeof(uint16_t));
    }
}

/* Random search, randomSearch (kernels/briefmatch_cuda.cu:222-246), stateless:
 * r = Philox4x32-10(ctr = {pixel index, iteration, pair, 0}, key = seed);
 * offset = round_half_away(sigma_q16 * Q[r >> 20] / 2^32) per axis from r[0]
 * and r[1] (a N(0, sigma^2) draw, sigma = sqrt(ceil(sx / 6))), clamped to keep
 * the target inside the image, accepted iff strictly cheaper. */
void orc_random_search(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                       int64_t sigma_q16, uint32_t iteration, uint32_t pair, uint64_t seed,
                       uint32_t *off, uint16_t *cost)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x) {
            size_t idx = (size_t)y * sx + x;
            const uint32_t ctr[4] = {(uint32_t)idx, iteration, pair, 0u};
            uint32_t r[4];
            orc_philox4x32_10(ctr, key, r);
            int nx = round_q32(sigma_q16 * (int64_t)kNdtriQ16[r[0] >> 20]);
            int ny = round_q32(sigma_q16 * (int64_t)kNdtriQ16[r[1] >> 20]);
            nx = clampi(nx, -x, sx - x - 1);
            ny = clampi(ny, -y, sy - y - 1);
            uint32_t cand = pack_off(nx, ny);
            if (cand == off[idx]) continue;
            size_t t = idx + (std::ptrdiff_t)ny * sx + nx;
            int c = hamming(F1 + idx * NS, F2 + t * NS, NS);
            if (c < cost[idx]) {
                cost[idx] = (uint16_t)c;
                off[idx] = cand;
            }
        }
}

/* BriefMatch::match (briefmatch.cpp:248-293) after the descriptors: init, then
 * `iterations` times {propagate(pd); randomSearch; pd = -pd} with pd = -1 first. */
void orc_match(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
               int iterations, int jfMax, int doubleProp, int first, const uint32_t *prev_off,
               uint32_t pair, uint64_t seed, uint32_t *off, uint16_t *cost)
{
    const int64_t sq16 = orc_rs_sigma_q16((unsigned)sx);
    orc_init(F1, F2, sx, sy, NS, first, prev_off, off, cost);
    int pd = -1;
    for (int i = 0; i < iterations; ++i) {
        orc_propagate(F1, F2, sx, sy, NS, pd, jfMax, doubleProp, off, cost);
        orc_random_search(F1, F2, sx, sy, NS, sq16, (uint32_t)i, pair, seed, off, cost);
        pd = -pd;
    }
}

/* BriefMatch::downsample (briefmatch.cpp:296-324): int16 -> FP32 (convert,
 * util_cuda.cu:174-183); if up-sampled, separable Gaussian horizontal then
 * vertical with clamped borders (filterConst, :48-73), nearest-neighbour
 * decimation (resizeNN, :160-172) and scaling by W/sx, H/sy (mult, :113-124).
 * The NN index round(x * sx / W) is computed exactly in integers and clamped to
 * the matched grid. */
void orc_downsample(const uint32_t *off, int sx, int sy, int W, int H,
                    const float *kernel, int P, float *U, float *V)
{
    const size_t M = (size_t)sx * sy;
    if (sx == W && sy == H) {
        for (size_t i = 0; i < M; ++i) {
            U[i] = (float)off_dx(off[i]);
            V[i] = (float)off_dy(off[i]);
        }
        return;
    }
    std::vector<float> Uf(M), Vf(M), tmp(M);
    for (size_t i = 0; i < M; ++i) {
        Uf[i] = (float)off_dx(off[i]);
        Vf[i] = (float)off_dy(off[i]);
    }
    const int P2 = P / 2;
    for (float *c : {Uf.data(), Vf.data()}) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < sy; ++y)
            for (int x = 0; x < sx; ++x) {
                float s = 0.0f;
                for (int i = 0; i < P; ++i)
                    s += kernel[i] * c[(size_t)y * sx + clampi(x + i - P2, 0, sx - 1)];
                tmp[(size_t)y * sx + x] = s;
            }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < sy; ++y)
            for (int x = 0; x < sx; ++x) {
                float s = 0.0f;
                for (int i = 0; i < P; ++i)
                    s += kernel[i] * tmp[(size_t)clampi(y + i - P2, 0, sy - 1) * sx + x];
                c[(size_t)y * sx + x] = s;
            }
    }
    const float scx = (float)W / (float)sx, scy = (float)H / (float)sy;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        int y1 = (int)std::min<int64_t>(((int64_t)2 * y * sy + H) / (2 * (int64_t)H), sy - 1);
        for (int x = 0; x < W; ++x) {
            int x1 = (int)std::min<int64_t>(((int64_t)2 * x * sx + W) / (2 * (int64_t)W), sx - 1);
            U[(size_t)y * W + x] = Uf[(size_t)y1 * sx + x1] * scx;
            V[(size_t)y * W + x] = Vf[(size_t)y1 * sx + x1] * scy;
        }
    }
}

/* 1-D median, medianFilter1D (kernels/util_cuda.cu:75-111): dir = 1 horizontal,
 * 0 vertical; the window is the fsize samples starting at -fsize/2, clamped to
 * the image (the reference reads outside it); the result is the (fsize/2)-th
 * smallest (0-based), which is what the reference's partial bubble sort leaves
 * in v[ks]. */
void orc_median1d(const float *in, float *out, int W, int H, int dir, int fsize)
{
    const int ks = fsize / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        std::vector<float> vbuf((size_t)fsize);
        float *v = vbuf.data();
        for (int x = 0; x < W; ++x) {
            for (int i = 0; i < fsize; ++i) {
                int xx = dir ? clampi(x - ks + i, 0, W - 1) : x;
                int yy = dir ? y : clampi(y - ks + i, 0, H - 1);
                v[i] = in[(size_t)yy * W + xx];
            }
            std::nth_element(v, v + ks, v + fsize);
            out[(size_t)y * W + x] = v[ks];
        }
    }
}

/* Cross-trilateral refinement, cbFilter (kernels/briefmatch_cuda.cu:248-297),
 * literally: window i, j in [-s, s) with s = ceil(2.5 sqrt(sigmaS)), clamped
 * indexing, w = exp(-dEPE / 2 sigmaEPE) exp(-dI^2 / 2 sigmaI) exp(-(i^2+j^2) /
 * 2 sigmaS) with dI on I/255, output sum(w u) / max(1e-5, sum w).  sigma*
 * arrive squared, as BriefMatch::process passes them (briefmatch.cpp:340-345).
 * Rows [row0, row1) only, so a test can check a band of a large frame. */
void orc_trilateral(const float *Ui, const float *Vi, const float *Um, const float *Vm,
                    const float *I, int W, int H, float sigmaEPE, float sigmaI, float sigmaS,
                    int row0, int row1, float *Uo, float *Vo)
{
    const int s = (int)std::ceil(2.5 * (double)std::sqrt(sigmaS));
    row0 = std::max(row0, 0);
    row1 = std::min(row1, H);
#pragma omp parallel for schedule(static)
    for (int y = row0; y < row1; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t o = (size_t)y * W + x;
            const float um0 = Um[o], vm0 = Vm[o], I0 = I[o] / 255.0f;
            float uf = 0.0f, vf = 0.0f, Wsum = 0.0f;
            for (int i = -s; i < s; ++i)
                for (int j = -s; j < s; ++j) {
                    const size_t o1 = (size_t)clampi(y + j, 0, H - 1) * W + clampi(x + i, 0, W - 1);
                    const float u1 = Ui[o1], v1 = Vi[o1], I1 = I[o1] / 255.0f;
                    const float du = u1 - um0, dv = v1 - vm0;
                    const float dEPE = du * du + dv * dv;
                    const float dI = I1 - I0;
                    const float w = std::exp(-dEPE / (2.0f * sigmaEPE)) *
                                    std::exp(-(dI * dI) / (2.0f * sigmaI)) *
                                    std::exp((float)(-(i * i + j * j)) / (2.0f * sigmaS));
                    uf += w * u1;
                    vf += w * v1;
                    Wsum += w;
                }
            const double den = std::max(1e-5, (double)Wsum);
            Uo[o] = (float)((double)uf / den);
            Vo[o] = (float)((double)vf / den);
        }
}

/* The whole flow path for one isolated pair, in the order of OFPipeline::
 * processFrame (of_pipeline.cpp:297-325): gray, up-sample, describe both frames,
 * match (first-frame init), down-sample, then median + trilateral guided by the
 * first frame at input resolution (process(dev_Lin_prev), :324). */
int orc_flow_pair(const orc_params *p, uint64_t seed, uint32_t pair,
                  const uint8_t *rgb1, const uint8_t *rgb2,
                  uint32_t *off_out, uint16_t *cost_out, float *U, float *V)
{
    return orc_flow_pair_up(p, seed, pair, rgb1, rgb2, 0, off_out, cost_out, U, V);
}

int orc_flow_pair_up(const orc_params *p, uint64_t seed, uint32_t pair,
                     const uint8_t *rgb1, const uint8_t *rgb2, int lanczos,
                     uint32_t *off_out, uint16_t *cost_out, float *U, float *V)
{
    const int W = (int)p->sx_in, H = (int)p->sy_in, sx = (int)p->sx, sy = (int)p->sy;
    const int N = (int)p->N, NS = (N + 31) / 32;
    const size_t Win = (size_t)W * H, M = (size_t)sx * sy;
    if (N < 2 || N > 512 || p->fsize > 47 || p->fsize == 0 || (p->fsize % 2) == 0) return -1;
    float kernel[128];
    unsigned P = 0;
    if (orc_gauss_kernel(p->sigma, kernel, &P) != 0) return -1;

    std::vector<int16_t> pattern(4 * (size_t)N);
    orc_make_pattern(seed, (unsigned)N, p->patchArea, pattern.data());
    std::vector<uint8_t> g1(Win), g2(Win);
    orc_gray(rgb1, W, H, g1.data());
    orc_gray(rgb2, W, H, g2.data());
    std::vector<float> L(M);
    std::vector<uint32_t> F1(M * NS), F2(M * NS);
    (lanczos ? orc_upsample_lanczos : orc_upsample)(g1.data(), W, H, sx, sy, L.data());
    orc_brief(L.data(), sx, sy, pattern.data(), N, F1.data());
    (lanczos ? orc_upsample_lanczos : orc_upsample)(g2.data(), W, H, sx, sy, L.data());
    orc_brief(L.data(), sx, sy, pattern.data(), N, F2.data());
    std::vector<float>().swap(L);

    std::vector<uint32_t> off(M);
    std::vector<uint16_t> cost(M);
    orc_match(F1.data(), F2.data(), sx, sy, NS, (int)p->iterations, (int)p->jfMax,
              p->doubleProp, 1, nullptr, pair, seed, off.data(), cost.data());
    if (off_out) std::memcpy(off_out, off.data(), M * sizeof(uint32_t));
    if (cost_out) std::memcpy(cost_out, cost.data(), M * sizeof(uint16_t));

    std::vector<float> U0(Win), V0(Win);
    orc_downsample(off.data(), sx, sy, W, H, kernel, (int)P, U0.data(), V0.data());
    if (!p->filter) {
        std::memcpy(U, U0.data(), Win * sizeof(float));
        std::memcpy(V, V0.data(), Win * sizeof(float));
        return 0;
    }
    std::vector<float> tmp(Win), Um(Win), Vm(Win), guide(Win);
    orc_median1d(U0.data(), tmp.data(), W, H, 1, (int)p->fsize);
    orc_median1d(tmp.data(), Um.data(), W, H, 0, (int)p->fsize);
    orc_median1d(V0.data(), tmp.data(), W, H, 1, (int)p->fsize);
    orc_median1d(tmp.data(), Vm.data(), W, H, 0, (int)p->fsize);
    for (size_t i = 0; i < Win; ++i) guide[i] = (float)g1[i];
    orc_trilateral(U0.data(), V0.data(), Um.data(), Vm.data(), guide.data(), W, H,
                   p->sEPE * p->sEPE, p->sI * p->sI, p->sS * p->sS, 0, H, U, V);
    return 0;
}

/* End-point error, BriefMatch::error (briefmatch.cpp:468-514): mean Euclidean
 * error over pixels whose ground truth is known (unknown_flow, flowIO.cpp:35-39:
 * |u| or |v| > 1e9 or NaN), accumulated in double; epe_noc also drops pixels
 * flagged occluded. */
void orc_epe(const float *U, const float *V, const float *gtU, const float *gtV,
             const uint8_t *occluded, int64_t n, double *epe_all, double *epe_noc)
{
    double sa = 0.0, sn = 0.0;
    int64_t na = 0, nn = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float gu = gtU[i], gv = gtV[i];
        if (std::fabs(gu) > 1e9f || std::fabs(gv) > 1e9f || std::isnan(gu) || std::isnan(gv)) continue;
        const double du = (double)gu - U[i], dv = (double)gv - V[i];
        const double e = std::sqrt(du * du + dv * dv);
        sa += e;
        ++na;
        if (!occluded || !occluded[i]) {
            sn += e;
            ++nn;
        }
    }
    *epe_all = na ? sa / (double)na : 0.0;
    *epe_noc = nn ? sn / (double)nn : 0.0;
}

}  // extern "C"
