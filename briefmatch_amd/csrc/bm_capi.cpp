// C-ABI of the BriefMatch flow engine: parameters, context (tables, device
// buffers, stream), stage entry points, the whole-pair path and sequence mode,
// and the pair path without device allocation (its one device array is the
// only device code here; the kernels are in bm_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "bm_internal.h"
#include "bm_net.h"

namespace {

const int32_t kNdtriQ16[4096] = {
#include "bm_ndtri_q16.inc"
};

thread_local std::string g_thread_error;

void philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

}  // namespace

struct bm_ctx {
    bm_params p{};
    bm::Geometry g{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<int16_t> pattern;
    std::vector<float> gauss;
    // tables
    int16_t *d_pattern = nullptr;
    uint32_t *d_plan = nullptr;  // fused BRIEF test records
    int *d_plan_count = nullptr;
    std::vector<uint32_t> plan;
    std::vector<int> plan_count;
    uint32_t *d_up_x = nullptr, *d_up_y = nullptr, *d_nn_x = nullptr, *d_nn_y = nullptr;
    float *d_gauss = nullptr, *d_spt = nullptr;
    // work buffers
    uint8_t *d_g8[2] = {nullptr, nullptr};
    float *d_g32[2] = {nullptr, nullptr};
    uint32_t *d_F[2] = {nullptr, nullptr};
    uint32_t *d_off[2] = {nullptr, nullptr};
    uint16_t *d_cost[2] = {nullptr, nullptr};
    float *d_U0 = nullptr, *d_V0 = nullptr, *d_tU = nullptr, *d_tV = nullptr, *d_Um = nullptr, *d_Vm = nullptr;
    int field = 0;  // d_off[field] holds the last match field
    // configuration generation: bumped by every change that alters launch
    // arguments (pattern and BRIEF plan, tuning, up-sampler), so captured
    // batch graphs know when to re-capture
    uint64_t gen = 0;
    // a batch's origin stream when the context belongs to a batch: its graph
    // may still read the context's tables when the configuration changes
    hipStream_t batch_stream = nullptr;
    // up-sampler of the whole-path calls (Lanczos tables and frame allocated on first use)
    int upsampler = BM_UPSAMPLE_BILINEAR;
    int *d_lz_x0 = nullptr, *d_lz_y0 = nullptr;
    float *d_lz_wx = nullptr, *d_lz_wy = nullptr;
    float *d_L = nullptr;  // up-sampled frame of the Lanczos mode
    float *d_tri_part = nullptr;  // partial sums of the split trilateral filter
    size_t tri_part_cap = 0;      // its size in floats
    // learned proposal (allocated on first use)
    uint16_t *d_feat = nullptr;            // [ch][cw][128] BF16
    uint16_t *d_nact[2] = {nullptr, nullptr};  // [ch][cw][64] BF16
    float *d_nout = nullptr;               // [ch][cw][16]
    int32_t *d_coff = nullptr;             // [ch][cw][4] Q8
    int32_t *d_bsum = nullptr;             // [2][ch][cw] block sums
    int32_t *d_far = nullptr;              // [ch / 2][cw / 2][2] far block match (frames of 16x16 and up)
    int ls_radius = 0, ls_samples = 0;     // bm_set_local_search (0 samples: off)
    // second stream (own hardware queue): the proposal network beside the
    // descriptor kernels, and the refinement tail of a pipelined sequence
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, tail_done = nullptr;
    int pipelined = 0;     // bm_set_pipelining
    int tail_pending = 0;  // a pipelined tail may still read d_U0, d_V0, d_guide
    float *d_guide = nullptr;  // guide frame of the pipelined tail
    // sequence mode
    int have_prev = 0, cur = 0, field_valid = 0;
    uint32_t frame_counter = 0;
    std::vector<void *> allocs;
    // per-stage event timing
    struct ProfEntry {
        int stage;
        hipEvent_t a, b;
    };
    bool prof = false;
    std::vector<ProfEntry> pending;
    std::vector<hipEvent_t> pool;
    double prof_ms[BM_STAGE_COUNT] = {};
    unsigned prof_n[BM_STAGE_COUNT] = {};
};

namespace {

int fail(bm_ctx *ctx, int code, const std::string &msg)
{
    if (ctx)
        ctx->err = msg;
    else
        g_thread_error = msg;
    return code;
}

int hip_fail(bm_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, BM_E_LAUNCH, std::string(what) + ": " + hipGetErrorString(e));
}

#define BM_HIP(ctx, expr)                                        \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
    } while (0)

hipEvent_t take_event(bm_ctx *ctx)
{
    if (!ctx->pool.empty()) {
        hipEvent_t e = ctx->pool.back();
        ctx->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Enqueue one kernel; while profiling, bracket it with events on the stream.
// a launch on stream st, bracketed by timing events when profiling is on
#define BM_LAUNCH_ON(ctx, stage, st, call)                                           \
    do {                                                                             \
        hipEvent_t a_ = nullptr;                                                     \
        if ((ctx)->prof && (a_ = take_event(ctx)) != nullptr)                        \
            (void)hipEventRecord(a_, (st));                                          \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);                     \
        if (a_) {                                                                    \
            hipEvent_t b_ = take_event(ctx);                                         \
            if (b_) {                                                                \
                (void)hipEventRecord(b_, (st));                                      \
                (ctx)->pending.push_back({(stage), a_, b_});                         \
            } else {                                                                 \
                (ctx)->pool.push_back(a_);                                           \
            }                                                                        \
        }                                                                            \
    } while (0)
#define BM_LAUNCH(ctx, stage, call) BM_LAUNCH_ON(ctx, stage, (ctx)->stream, call)

template <class T>
int dev_alloc(bm_ctx *ctx, T **ptr, size_t count)
{
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, BM_E_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    ctx->allocs.push_back(p);
    *ptr = static_cast<T *>(p);
    return BM_OK;
}

template <class T>
int upload(bm_ctx *ctx, T *dst, const std::vector<T> &src)
{
    BM_HIP(ctx, hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return BM_OK;
}

unsigned next_pow2(unsigned v)
{
    unsigned p = 1;
    while (p < v) p <<= 1;
    return p;
}

int make_gauss(float sigma, std::vector<float> &k)
{
    const unsigned n = (unsigned)(std::ceil(sigma * 3.0f) * 2.0f + 1.0f);
    if (n > BM_MAX_GAUSS_TAPS) return -1;
    const int P2 = (int)n / 2;
    k.assign(n, 0.0f);
    float w = 0.0f;
    for (int i = -P2; i <= P2; ++i) {
        k[i + P2] = std::exp(-(float)(i * i) / (2.0f * sigma * sigma));
        w += k[i + P2];
    }
    for (auto &v : k) v /= w;
    return 0;
}

void make_pattern(uint64_t seed, unsigned N, float patchArea, std::vector<int16_t> &pat)
{
    pat.assign(4 * (size_t)N, 0);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (unsigned s = 0; s < N; ++s) {
        const uint32_t ctr[4] = {s, 0u, 0u, 1u};
        uint32_t r[4];
        philox_host(ctr, key, r);
        for (int c = 0; c < 4; ++c) {
            const float v = 0.2f * ((float)kNdtriQ16[r[c] >> 20] * (1.0f / 65536.0f));
            const float sq = v * v;
            float val = v > 0.0f ? sq : -sq;
            val = std::max(-1.0f, std::min(1.0f, val));
            pat[4 * s + c] = (int16_t)(int)(patchArea * val);
        }
    }
}

// Fused BRIEF tiling: the far halo is the largest pattern offset, the near halo
// the smallest of 8/16/24/32 covering min(far, 32); the gray tile bound is the
// widest input-resolution extent of any tile.  The kernel is used when one
// work-group's LDS fits in 64 KiB.
// LDS bytes one work-group may use on the current device (160 KiB on gfx950),
// queried once; 64 KiB if the query fails.
size_t device_lds_limit()
{
    static size_t limit = 0;
    if (limit == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            limit = (size_t)v;
        else
            limit = 64 * 1024;
    }
    return limit;
}

bool plan_brief(bm_ctx *ctx, int rmax, int rn_req)
{
    bm::Geometry &g = ctx->g;
    const int want = std::min(rmax, 32);
    const int rn = rn_req > 0 ? rn_req : (want <= 8 ? 8 : want <= 16 ? 16 : want <= 24 ? 24 : 32);
    rmax = std::max(rmax, rn);  // the coordinate tables also cover the near tile
    g.brief_far = rmax;
    g.brief_near = 0;
    std::vector<uint32_t> upx(g.sx), upy(g.sy);
    for (int x = 0; x < g.sx; ++x) upx[x] = (uint32_t)((uint64_t)x * g.W * 256u / g.sx);
    for (int y = 0; y < g.sy; ++y) upy[y] = (uint32_t)((uint64_t)y * g.H * 256u / g.sy);
    auto extent = [&](const std::vector<uint32_t> &up, int n, int in, int t0, int tw) {
        const int c0 = std::max(0, std::min(n - 1, t0 - rmax));
        const int c1 = std::max(0, std::min(n - 1, t0 + tw - 1 + rmax));
        const int a = (int)(up[c0] >> 8), b = std::min((int)(up[c1] >> 8) + 1, in - 1);
        return b - a + 1;
    };
    int gw = 0;
    for (int t0 = 0; t0 < g.sx; t0 += 64) gw = std::max(gw, extent(upx, g.sx, g.W, t0, 64));
    const int TH = bm::brief_tile_rows(g.NSP), which = TH == 32 ? 1 : 0;
    int gh = 0;
    for (int t0 = 0; t0 < g.sy; t0 += TH) gh = std::max(gh, extent(upy, g.sy, g.H, t0, TH));
    g.brief_gw[which] = gw;
    g.brief_gh[which] = gh;
    g.brief_gw[1 - which] = g.brief_gh[1 - which] = 0;
    if (rn_req == 0) return true;  // stand-alone kernel requested
    if (gw >= 4096 || (size_t)gh * gw >= 32768) return false;
    if (bm::brief_lds_bytes(g, rn, TH) > device_lds_limit()) return false;
    g.brief_near = rn;
    return true;
}

// Test records of the fused BRIEF.  Near part: per descriptor word 32 records
// {o1, o2} (two int32) for bits 31 .. 0, the LDS offsets of both samples for a
// near test (all offsets within the near halo) from the thread's top-left
// halo cell, (RN + oy) * pitch + RN + ox: the kernel's lane base then needs no
// column constant that the row-pair LDS loads cannot take as an immediate;
// both 0 for a far test or a bit beyond the descriptor length (the kernel
// then shifts in a 0).  Far part,
// after the NSP * 64 near words: per descriptor word up to 32 records {packed
// int16 offsets of both samples, bit position, 0}.  Counts per word: bits,
// far tests.  NSP * 192 words in all, within the NSP * 256 of the buffers.
void build_brief_plan(bm_ctx *ctx)
{
    const bm::Geometry &g = ctx->g;
    const int NSP = g.NSP, RN = g.brief_near, pitch = 64 + 2 * RN;
    ctx->plan.assign((size_t)NSP * 64 * 4, 0u);
    ctx->plan_count.assign((size_t)NSP * 2, 0);
    uint32_t *near_rec = ctx->plan.data();
    uint32_t *far_rec = ctx->plan.data() + (size_t)NSP * 64;
    for (int w = 0; w < g.NS; ++w) {
        const int nbits = std::min(32, g.N - 32 * w);
        int nfar = 0;
        for (int i = 0; i < 32; ++i) {
            const int b = 31 - i;
            if (b >= nbits) continue;  // shifts in a 0 above the last bit
            const int16_t *o = &ctx->pattern[4 * (32 * w + b)];
            const int m = std::max(std::max(std::abs((int)o[0]), std::abs((int)o[1])),
                                   std::max(std::abs((int)o[2]), std::abs((int)o[3])));
            if (m <= RN) {
                near_rec[(size_t)64 * w + 2 * i] = (uint32_t)((RN + o[1]) * pitch + RN + o[0]);
                near_rec[(size_t)64 * w + 2 * i + 1] = (uint32_t)((RN + o[3]) * pitch + RN + o[2]);
            } else {
                uint32_t *fr = far_rec + ((size_t)32 * w + nfar++) * 4;
                fr[0] = (uint32_t)(uint16_t)o[0] | ((uint32_t)(uint16_t)o[1] << 16);
                fr[1] = (uint32_t)(uint16_t)o[2] | ((uint32_t)(uint16_t)o[3] << 16);
                fr[2] = (uint32_t)b;
            }
        }
        ctx->plan_count[2 * w] = nbits;
        ctx->plan_count[2 * w + 1] = nfar;
    }
}

int upload_pattern(bm_ctx *ctx)
{
    int r = upload(ctx, ctx->d_pattern, ctx->pattern);
    if (r == BM_OK) r = upload(ctx, ctx->d_plan, ctx->plan);
    if (r == BM_OK) r = upload(ctx, ctx->d_plan_count, ctx->plan_count);
    return r;
}

int pattern_rmax(const bm_ctx *ctx)
{
    int rmax = 0;
    for (int16_t v : ctx->pattern) rmax = std::max(rmax, std::abs((int)v));
    return rmax;
}

// rn_req: -1 default near halo (the stand-alone kernel when the fused one's
// tiles do not fit), 0 stand-alone kernel, else the requested halo, which must
// fit: then nothing changes on failure.  The plan buffers on the device hold
// the largest plan (NSP * 64 records, NSP * 2 counts) from creation on, so a
// pattern set later never outgrows them.
int set_pattern_internal(bm_ctx *ctx, const std::vector<int16_t> &pat, int rn_req = -1)
{
    const std::vector<int16_t> old_pat = ctx->pattern;
    const bm::Geometry old_g = ctx->g;
    ctx->pattern = pat;
    const int rmax = pattern_rmax(ctx);
    ctx->g.brief_radius = std::min(rmax, bm::brief_max_radius());
    if (!plan_brief(ctx, rmax, rn_req)) {
        if (rn_req > 0) {
            ctx->pattern = old_pat;
            ctx->g = old_g;
            return fail(ctx, BM_E_PARAM, "BRIEF near halo does not fit in LDS");
        }
        plan_brief(ctx, rmax, 0);  // stand-alone kernel
    }
    if (ctx->g.brief_near > 0) {
        build_brief_plan(ctx);
    } else {
        ctx->plan.assign((size_t)ctx->g.NSP * 64 * 4, 0u);
        ctx->plan_count.assign((size_t)ctx->g.NSP * 2, 0);
        ctx->g.brief_frame = 0;  // the frame kernel needs the near plan
    }
    ctx->gen++;
    if (ctx->d_pattern) return upload_pattern(ctx);
    return BM_OK;
}

// Waits for every stream that may still read the context's tables: its own
// two and, for a batch member, the batch's origin stream (a captured batch
// graph runs there).  Called before a configuration change re-uploads tables.
int quiesce(bm_ctx *ctx)
{
    BM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) BM_HIP(ctx, hipStreamSynchronize(ctx->side));
    if (ctx->batch_stream) BM_HIP(ctx, hipStreamSynchronize(ctx->batch_stream));
    return BM_OK;
}

int check(bm_ctx *ctx, const void *p, const char *name)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    if (!p) return fail(ctx, BM_E_PARAM, std::string("null pointer: ") + name);
    return BM_OK;
}

// Lanczos-3 weight table of one axis: first source index and six normalised
// weights per output index (resizeLanczos, util_cuda.cu:133-157: xp = x *
// float(in) / out, taps floor(xp) - 2 .. floor(xp) + 3, 3 sin(t) sin(t / 3) / t^2).
void lanczos_table(int in, int out, std::vector<int> &first, std::vector<float> &w)
{
    const double pi = 3.14159265358979323846;
    first.resize(out);
    w.resize(6 * (size_t)out);
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

int lanczos_ready(bm_ctx *ctx, bool frame)
{
    if (!ctx->d_lz_x0) {
        std::vector<int> fx, fy;
        std::vector<float> wx, wy;
        lanczos_table(ctx->g.W, ctx->g.sx, fx, wx);
        lanczos_table(ctx->g.H, ctx->g.sy, fy, wy);
        int *x0 = nullptr, *y0 = nullptr;
        float *dwx = nullptr, *dwy = nullptr;
        int r = dev_alloc(ctx, &x0, fx.size());
        if (!r) r = dev_alloc(ctx, &y0, fy.size());
        if (!r) r = dev_alloc(ctx, &dwx, wx.size());
        if (!r) r = dev_alloc(ctx, &dwy, wy.size());
        if (!r) r = upload(ctx, x0, fx);
        if (!r) r = upload(ctx, y0, fy);
        if (!r) r = upload(ctx, dwx, wx);
        if (!r) r = upload(ctx, dwy, wy);
        if (r) return r;
        ctx->g.lz_x0 = ctx->d_lz_x0 = x0;
        ctx->g.lz_y0 = ctx->d_lz_y0 = y0;
        ctx->g.lz_wx = ctx->d_lz_wx = dwx;
        ctx->g.lz_wy = ctx->d_lz_wy = dwy;
    }
    if (frame && !ctx->d_L) return dev_alloc(ctx, &ctx->d_L, (size_t)ctx->g.sx * ctx->g.sy);
    return BM_OK;
}

// Descriptors of one gray frame with the context's up-sampler: bilinear fused
// into BRIEF, bilinear into the context frame followed by BRIEF on it
// (brief_frame), or Lanczos-3 into the context frame followed by BRIEF on it.
int describe(bm_ctx *ctx, const uint8_t *g8, uint32_t *F)
{
    if (ctx->upsampler == BM_UPSAMPLE_LANCZOS3) {
        BM_LAUNCH(ctx, BM_STAGE_UPSAMPLE, bm::launch_upsample_lanczos(ctx->g, g8, ctx->d_L, ctx->stream));
        BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief(ctx->g, ctx->d_L, nullptr, F, ctx->stream));
    } else if (ctx->g.brief_frame) {
        BM_LAUNCH(ctx, BM_STAGE_UPSAMPLE, bm::launch_upsample(ctx->g, g8, ctx->d_L, ctx->stream));
        BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief(ctx->g, ctx->d_L, nullptr, F, ctx->stream));
    } else {
        BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief(ctx->g, nullptr, g8, F, ctx->stream));
    }
    return BM_OK;
}

// Descriptors of both frames of a pair: in the fused bilinear mode one launch
// over both frames' tiles (small frames leave most of the device idle per
// frame), else describe() per frame.
int describe_pair(bm_ctx *ctx, const uint8_t *g8a, uint32_t *Fa, const uint8_t *g8b, uint32_t *Fb)
{
    if (ctx->upsampler == BM_UPSAMPLE_BILINEAR && !ctx->g.brief_frame && ctx->g.brief_near > 0) {
        BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief_pair(ctx->g, g8a, Fa, g8b, Fb, ctx->stream));
        return BM_OK;
    }
    const int r = describe(ctx, g8a, Fa);
    return r ? r : describe(ctx, g8b, Fb);
}

// Ping-pong propagation of one iteration over the context's field buffers,
// starting from ctx->field; the random search of `iteration` is fused into the
// last level when fuse_rs.
int run_iteration(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, int pd, int iteration, uint32_t pair,
                  bool fuse_rs)
{
    const bm::Geometry &g = ctx->g;
    for (int p = g.jfMax; p > 0; p /= 2) {
        const int rs = (fuse_rs && p == 1) ? iteration : -1;
        const int a = ctx->field, b = 1 - a;
        BM_LAUNCH(ctx, (rs >= 0 ? BM_STAGE_LEVEL_RS : BM_STAGE_LEVEL), bm::launch_level(g, F1, F2, ctx->d_off[a], ctx->d_cost[a], ctx->d_off[b], ctx->d_cost[b], p, pd,
                                     rs, pair, ctx->stream));
        ctx->field = b;
    }
    if (g.jfMax == 0 && fuse_rs) {
        BM_LAUNCH(ctx, BM_STAGE_RANDOM_SEARCH, bm::launch_random_search(g, F1, F2, ctx->d_off[ctx->field], ctx->d_cost[ctx->field],
                                             (uint32_t)iteration, pair, ctx->stream));
    }
    return BM_OK;
}

// init + iterations into the context field buffers.
int run_match(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, int first, const uint32_t *prev, uint32_t pair)
{
    const bm::Geometry &g = ctx->g;
    int dst = ctx->field;
    if (!first && prev == ctx->d_off[dst]) dst = 1 - dst;
    // After a zero initialisation the first iteration is the identity except for
    // its random search (every propagation candidate equals the pixel's own
    // (0, 0)): the initialisation kernel inserts that random search itself.
    const bool fused = first && g.iterations > 0;
    BM_LAUNCH(ctx, BM_STAGE_INIT, bm::launch_init(g, F1, F2, first, prev, ctx->d_off[dst], ctx->d_cost[dst],
                                                  fused ? 0 : -1, pair, ctx->stream));
    ctx->field = dst;
    int pd = -1;
    for (int i = 0; i < g.iterations; ++i) {
        if (!(i == 0 && fused)) {
            const int r = run_iteration(ctx, F1, F2, pd, i, pair, true);
            if (r != BM_OK) return r;
        }
        pd = -pd;
    }
    return BM_OK;
}

// The context stream waits for a pipelined refinement tail still in flight
// (it reads d_U0, d_V0 and d_guide and writes the median buffers).
int settle_tail(bm_ctx *ctx)
{
    if (!ctx->tail_pending) return BM_OK;
    BM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->tail_done, 0));
    ctx->tail_pending = 0;
    return BM_OK;
}

// medians x then y, then the trilateral filter, on stream st
int run_refine_on(bm_ctx *ctx, const float *U, const float *V, const float *I, float *Uo, float *Vo, hipStream_t st)
{
    const bm::Geometry &g = ctx->g;
    BM_LAUNCH_ON(ctx, BM_STAGE_MEDIAN, st, bm::launch_median(g, U, V, ctx->d_tU, ctx->d_tV, 1, st));
    BM_LAUNCH_ON(ctx, BM_STAGE_MEDIAN, st, bm::launch_median(g, ctx->d_tU, ctx->d_tV, ctx->d_Um, ctx->d_Vm, 0, st));
    BM_LAUNCH_ON(ctx, BM_STAGE_TRILATERAL, st, bm::launch_trilateral(g, U, V, ctx->d_Um, ctx->d_Vm, I, Uo, Vo, st));
    return BM_OK;
}

int run_refine(bm_ctx *ctx, const float *U, const float *V, const float *I, float *Uo, float *Vo)
{
    const int r = settle_tail(ctx);
    if (r) return r;
    return run_refine_on(ctx, U, V, I, Uo, Vo, ctx->stream);
}

int run_flow_tail(bm_ctx *ctx, const float *guide, float *U, float *V)
{
    const bm::Geometry &g = ctx->g;
    const int r = settle_tail(ctx);
    if (r) return r;
    if (!ctx->p.filter) {
        BM_LAUNCH(ctx, BM_STAGE_DOWNSAMPLE, bm::launch_downsample(g, ctx->d_off[ctx->field], U, V, ctx->stream));
        return BM_OK;
    }
    BM_LAUNCH(ctx, BM_STAGE_DOWNSAMPLE, bm::launch_downsample(g, ctx->d_off[ctx->field], ctx->d_U0, ctx->d_V0, ctx->stream));
    return run_refine(ctx, ctx->d_U0, ctx->d_V0, guide, U, V);
}

// Host-side coordinate and filter tables of a configuration (uploaded to the
// device by the context or the no-allocation entry point).
struct HostTables {
    std::vector<uint32_t> upx, upy, nnx, nny;
    std::vector<float> spt;
};

// The limits of BriefMatch::setup (briefmatch.cpp:78-96) and of this build's
// 32-bit indexing and int16 offsets.
int validate_params(const bm_params &p)
{
    char m[512];
    if (p.N < 2 || p.N > BM_MAX_FEATURE_BITS) {
        std::snprintf(m, sizeof m, "Feature length exceeds maximum value. Current size is %u and maximum is %d.", p.N,
                      BM_MAX_FEATURE_BITS);
        return fail(nullptr, BM_E_PARAM, m);
    }
    if (p.fsize > BM_MAX_MEDIAN) {
        std::snprintf(m, sizeof m, "Median filter size exceeds maximum value. Current size is %u and maximum is %d.",
                      p.fsize, BM_MAX_MEDIAN);
        return fail(nullptr, BM_E_PARAM, m);
    }
    if (p.filter && (p.fsize % 2) == 0)
        return fail(nullptr, BM_E_PARAM, "Median filter size must be odd.");
    if (std::ceil(p.sigma * 3.0f) * 2.0f + 1.0f > BM_MAX_GAUSS_TAPS) {
        std::snprintf(m, sizeof m,
                      "Maximum size of Gaussian down-sampling filter exceeded. Currently using sigma = %0.4f and "
                      "maximum is %0.4f.",
                      p.sigma, std::floor((BM_MAX_GAUSS_TAPS - 1.0f) / 2.0f) / 3.0f);
        return fail(nullptr, BM_E_PARAM, m);
    }
    if (p.sx_in == 0 || p.sy_in == 0 || p.sx == 0 || p.sy == 0)
        return fail(nullptr, BM_E_PARAM, "empty resolution");
    if (p.sx > 32767 || p.sy > 32767)
        return fail(nullptr, BM_E_PARAM, "matched resolution above 32767 (offsets are int16)");
    if (p.sigma <= 0.0f) return fail(nullptr, BM_E_PARAM, "down-sampling sigma must be positive");
    if (p.filter && (p.sEPE <= 0.0f || p.sI <= 0.0f || p.sS <= 0.0f))
        return fail(nullptr, BM_E_PARAM, "trilateral sigmas must be positive");
    const size_t M = (size_t)p.sx * p.sy;
    const unsigned NS = (p.N + 31) / 32, NSP = next_pow2(NS);
    if ((double)M * NSP >= 4294967296.0 || (double)p.jfMax * p.sx + (double)M >= 2147483647.0)
        return fail(nullptr, BM_E_PARAM, "matched resolution too large for 32-bit indexing");

    return BM_OK;
}

// Everything of a context that lives on the host: parameters, geometry,
// tuning defaults, Gaussian, pattern and fused-BRIEF plan, and the tables in t.
int prepare_host(bm_ctx *ctx, const bm_params &p, uint64_t seed, HostTables &t)
{
    const unsigned NS = (p.N + 31) / 32, NSP = next_pow2(NS);
    ctx->p = p;
    bm::Geometry &g = ctx->g;
    g.W = (int)p.sx_in;
    g.H = (int)p.sy_in;
    g.sx = (int)p.sx;
    g.sy = (int)p.sy;
    g.N = (int)p.N;
    g.NS = (int)NS;
    g.NSP = (int)NSP;
    g.fsize = (int)p.fsize;
    g.jfMax = (int)p.jfMax;
    g.iterations = (int)p.iterations;
    g.doubleProp = p.doubleProp ? 1 : 0;
    g.seed = seed;
    g.level_px = 1;  // the autotuner's choice at 320x240, 1080p and 720p x4.9
    g.strip_rows = 256;
    g.brief_frame = 0;
    g.tri_split = 1;
    g.tri_part = nullptr;
    const float rs_sigma = std::sqrt(std::ceil((float)p.sx * 0.5f / 3.0f));
    g.rs_sigma_q16 = (int64_t)std::llround((double)rs_sigma * 65536.0);

    make_gauss(p.sigma, ctx->gauss);
    g.P = (int)ctx->gauss.size();
    ctx->p.P = (unsigned)g.P;
    std::vector<int16_t> pat;
    make_pattern(seed, p.N, p.patchArea, pat);
    set_pattern_internal(ctx, pat);

    // trilateral constants (sigmas squared as BriefMatch::process passes them)
    const float sE = p.sEPE * p.sEPE, sI = p.sI * p.sI, sS = p.sS * p.sS;
    g.tri_sigma_epe = sE;
    g.tri_sigma_i = sI;
    g.tri_sigma_s = sS;
    const double log2e = 1.4426950408889634;
    g.tri_s = p.filter ? (int)std::ceil(2.5 * (double)std::sqrt(sS)) : 0;
    g.tri_a2 = p.filter ? (float)(log2e / (2.0 * (double)sE)) : 1.0f;
    g.tri_b = p.filter ? (float)(std::sqrt(log2e / (2.0 * (double)sI)) / 255.0) : 1.0f;
    if (p.filter && (g.tri_s < 1 || g.tri_s > 2000))
        return fail(nullptr, BM_E_PARAM, "trilateral window outside [1, 2000]");
    g.tri_band = p.filter ? bm::tri_max_band(g.tri_s) : 1;
    // the MFMA kernel's 64 x 8 tiles fill the device where the VALU kernel's
    // 64 x 16 tiles cannot (measured: 2x faster at 320x240, 75 tiles; 3 % at
    // 1280x720, 900 tiles; 2 % slower at 1920x1080, 2040 tiles)
    g.tri_groups = ((p.sx_in + 63) / 64) * ((p.sy_in + 15) / 16) <= 1024 ? 2 : 0;
    std::vector<float> &spt = t.spt;
    spt.assign(std::max(2 * g.tri_s, 1), 0.0f);
    for (int m_ = 0; m_ < 2 * g.tri_s; ++m_) {
        const double i = m_ - g.tri_s;
        spt[m_] = (float)(log2e * i * i / (2.0 * (double)sS));
    }

    // coordinate tables
    std::vector<uint32_t> &upx = t.upx, &upy = t.upy, &nnx = t.nnx, &nny = t.nny;
    upx.assign(p.sx, 0u);
    upy.assign(p.sy, 0u);
    nnx.assign(p.sx_in, 0u);
    nny.assign(p.sy_in, 0u);
    for (unsigned x = 0; x < p.sx; ++x) upx[x] = (uint32_t)((uint64_t)x * p.sx_in * 256u / p.sx);
    for (unsigned y = 0; y < p.sy; ++y) upy[y] = (uint32_t)((uint64_t)y * p.sy_in * 256u / p.sy);
    for (unsigned x = 0; x < p.sx_in; ++x)
        nnx[x] = (uint32_t)std::min<uint64_t>((2ull * x * p.sx + p.sx_in) / (2ull * p.sx_in), p.sx - 1);
    for (unsigned y = 0; y < p.sy_in; ++y)
        nny[y] = (uint32_t)std::min<uint64_t>((2ull * y * p.sy + p.sy_in) / (2ull * p.sy_in), p.sy - 1);
    // largest field window of a down-sampling output tile (NN samples +- P/2)
    {
        unsigned cols = 0, rows = 0;
        for (unsigned x0 = 0; x0 < p.sx_in; x0 += bm::kDsTileW)
            cols = std::max(cols, nnx[std::min(x0 + bm::kDsTileW, p.sx_in) - 1] - nnx[x0] + 1);
        for (unsigned y0 = 0; y0 < p.sy_in; y0 += bm::kDsTileH)
            rows = std::max(rows, nny[std::min(y0 + bm::kDsTileH, p.sy_in) - 1] - nny[y0] + 1);
        const size_t bytes = (size_t)(cols + g.P - 1) * (rows + g.P - 1) * sizeof(uint32_t);
        g.ds_lds = bytes <= (size_t)bm::kDsLdsMax ? (int)bytes : 0;
    }

    return BM_OK;
}

// Device table pointers of the geometry, from the context's buffers.
void bind_tables(bm_ctx *ctx)
{
    bm::Geometry &g = ctx->g;
    g.pattern = ctx->d_pattern;
    g.brief_plan = ctx->d_plan;
    g.brief_plan_count = ctx->d_plan_count;
    g.up_x = ctx->d_up_x;
    g.up_y = ctx->d_up_y;
    g.nn_x = ctx->d_nn_x;
    g.nn_y = ctx->d_nn_y;
    g.gauss = ctx->d_gauss;
    g.tri_spatial = ctx->d_spt;
}


// One isolated pair on the context's buffers and stream (bm_flow_pair).
int run_pair(bm_ctx *ctx, const uint8_t *rgb1, const uint8_t *rgb2, uint32_t pair, float *U, float *V)
{
    const bm::Geometry &g = ctx->g;
    BM_LAUNCH(ctx, BM_STAGE_GRAY,
              bm::launch_gray_pair(g, rgb1, ctx->d_g8[0], ctx->d_g32[0], rgb2, ctx->d_g8[1], ctx->stream));
    int r = describe_pair(ctx, ctx->d_g8[0], ctx->d_F[0], ctx->d_g8[1], ctx->d_F[1]);
    if (r) return r;
    r = run_match(ctx, ctx->d_F[0], ctx->d_F[1], 1, nullptr, pair);
    if (r) return r;
    // an isolated pair breaks any running sequence
    ctx->have_prev = 0;
    ctx->field_valid = 0;
    return run_flow_tail(ctx, ctx->d_g32[0], U, V);
}

}  // namespace

extern "C" {

void bm_params_default(bm_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->sx = p->sy = p->sx_in = p->sy_in = 64;
    p->iterations = 2;
    p->jfMax = 256;
    p->doubleProp = 1;
    p->filter = 1;
    p->tInit = 1;
    p->sigma = 0.5f;
    p->P = 11;
    p->fsize = 13;
    p->N = 128;
    p->patchArea = 75.0f;
    p->sEPE = 0.4f;
    p->sI = 0.1f;
    p->sS = 6.0f;
    p->maxmot = 5.0f;
}

int bm_params_derive(unsigned W, unsigned H, float ux, float uy, float patchRad, float sigmaRad, float fsizeRad,
                     float sSRad, bm_params *p)
{
    if (!p) return fail(nullptr, BM_E_PARAM, "null params");
    if (W == 0 || H == 0) return fail(nullptr, BM_E_PARAM, "empty input resolution");
    bm_params_default(p);
    if (ux < 0.1f) ux = 2.740f;
    if (uy < 0.1f) uy = 3.052f;
    const unsigned sx = (unsigned)(ux * (float)W), sy = (unsigned)(uy * (float)H);
    p->sx = sx;
    p->sy = sy;
    p->sx_in = W;
    p->sy_in = H;
    const double diag = std::sqrt((double)(sx * sx + sy * sy));
    const double diag_in = std::sqrt((double)(W * W + H * H));
    p->patchArea = (float)((double)(0.01f * patchRad) * diag);
    p->sigma = (float)((double)(0.01f * sigmaRad) * diag);
    p->fsize = (unsigned)(2.0 * std::round((double)(0.005f * fsizeRad) * diag_in) + 1.0);
    p->sS = (float)((double)(0.01f * sSRad) * diag_in);
    p->P = (unsigned)(std::ceil(p->sigma * 3.0f) * 2.0f + 1.0f);
    return BM_OK;
}

const char *bm_last_error(const bm_ctx *ctx) { return ctx ? ctx->err.c_str() : g_thread_error.c_str(); }

int bm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    if (!ctr || !key || !out) return fail(nullptr, BM_E_PARAM, "null pointer");
    philox_host(ctr, key, out);
    return BM_OK;
}

// Matched pixels from which a new context up-samples into a frame for BRIEF
// (brief_frame 1) by default.
constexpr size_t kBriefFrameMinPixels = (size_t)1 << 21;

namespace {
int apply_tuning(bm_ctx *ctx, const bm_tuning &t);
bm_tuning current_tuning(const bm_ctx *ctx);
}  // namespace

int bm_ctx_create(const bm_params *params, uint64_t seed, bm_ctx **out)
{
    if (!params || !out) return fail(nullptr, BM_E_PARAM, "null argument");
    *out = nullptr;
    int r = validate_params(*params);
    if (r) return r;
    bm_ctx *ctx = new (std::nothrow) bm_ctx;
    if (!ctx) return fail(nullptr, BM_E_ALLOC, "out of host memory");
    HostTables t;
    r = prepare_host(ctx, *params, seed, t);
    if (r) {
        delete ctx;
        return r;
    }
    const bm_params &p = ctx->p;
    const bm::Geometry &g = ctx->g;

    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e != hipSuccess) {
        std::string msg = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
        bm_ctx_destroy(ctx);
        return fail(nullptr, BM_E_LAUNCH, msg);
    }

    const size_t Win = (size_t)p.sx_in * p.sy_in, M = (size_t)p.sx * p.sy, NSP = (size_t)g.NSP;
#define BM_ALLOC(ptr, n) \
    if (r == BM_OK) r = dev_alloc(ctx, &(ptr), (n))
    BM_ALLOC(ctx->d_pattern, ctx->pattern.size());
    BM_ALLOC(ctx->d_plan, NSP * 64 * 4);
    BM_ALLOC(ctx->d_plan_count, NSP * 2);
    BM_ALLOC(ctx->d_up_x, t.upx.size());
    BM_ALLOC(ctx->d_up_y, t.upy.size());
    BM_ALLOC(ctx->d_nn_x, t.nnx.size());
    BM_ALLOC(ctx->d_nn_y, t.nny.size());
    BM_ALLOC(ctx->d_gauss, ctx->gauss.size());
    BM_ALLOC(ctx->d_spt, t.spt.size());
    for (int k = 0; k < 2; ++k) {
        BM_ALLOC(ctx->d_g8[k], Win);
        BM_ALLOC(ctx->d_g32[k], Win);
        BM_ALLOC(ctx->d_F[k], M * NSP);
        BM_ALLOC(ctx->d_off[k], M);
        BM_ALLOC(ctx->d_cost[k], M);
    }
    BM_ALLOC(ctx->d_U0, Win);
    BM_ALLOC(ctx->d_V0, Win);
    BM_ALLOC(ctx->d_tU, Win);
    BM_ALLOC(ctx->d_tV, Win);
    BM_ALLOC(ctx->d_Um, Win);
    BM_ALLOC(ctx->d_Vm, Win);
    // large matched frames default to the up-sampled frame in memory for BRIEF
    // (measured: 1080p x2.9 and 720p x4.9 faster, 320x240 x1.5 not)
    if (g.brief_near > 0 && M >= kBriefFrameMinPixels) {
        BM_ALLOC(ctx->d_L, M);
        if (r == BM_OK) ctx->g.brief_frame = 1;
    }
#undef BM_ALLOC
    if (r == BM_OK) r = upload_pattern(ctx);
    if (r == BM_OK) r = upload(ctx, ctx->d_up_x, t.upx);
    if (r == BM_OK) r = upload(ctx, ctx->d_up_y, t.upy);
    if (r == BM_OK) r = upload(ctx, ctx->d_nn_x, t.nnx);
    if (r == BM_OK) r = upload(ctx, ctx->d_nn_y, t.nny);
    if (r == BM_OK) r = upload(ctx, ctx->d_gauss, ctx->gauss);
    if (r == BM_OK) r = upload(ctx, ctx->d_spt, t.spt);
    if (r != BM_OK) {
        std::string msg = ctx->err;
        bm_ctx_destroy(ctx);
        return fail(nullptr, r, msg);
    }
    bind_tables(ctx);
    // 512-bit frame mode: the 48-px near halo is the faster default (its
    // 16-wave work-groups hold the larger tile; 720p x4.9: 2.92 against 3.71 ms
    // of BRIEF at 32 px), where the pattern reaches beyond 32 px and the plan fits
    if (ctx->g.brief_frame && ctx->g.NSP >= 16 && ctx->g.brief_near == 32 && pattern_rmax(ctx) > 32) {
        const bool fits = plan_brief(ctx, pattern_rmax(ctx), 48);
        plan_brief(ctx, pattern_rmax(ctx), 32);
        bm_tuning t = current_tuning(ctx);
        t.brief_near = 48;
        const int rt = fits ? apply_tuning(ctx, t) : BM_OK;
        if (rt != BM_OK) {
            std::string msg = ctx->err;
            bm_ctx_destroy(ctx);
            return fail(nullptr, rt, msg);
        }
    }
    *out = ctx;
    return BM_OK;
}

void bm_ctx_destroy(bm_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (void *p : ctx->allocs) (void)hipFree(p);
    for (auto &pe : ctx->pending) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->tail_done) (void)hipEventDestroy(ctx->tail_done);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int bm_ctx_get_params(const bm_ctx *ctx, bm_params *out)
{
    if (!ctx || !out) return fail(nullptr, BM_E_PARAM, "null argument");
    *out = ctx->p;
    return BM_OK;
}

void *bm_ctx_stream(bm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int bm_sync(bm_ctx *ctx)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    BM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) BM_HIP(ctx, hipStreamSynchronize(ctx->side));
    return BM_OK;
}

unsigned bm_descriptor_stride(const bm_ctx *ctx) { return ctx ? (unsigned)ctx->g.NSP : 0u; }

int bm_get_pattern(const bm_ctx *ctx, int16_t *out)
{
    if (!ctx || !out) return fail(nullptr, BM_E_PARAM, "null argument");
    std::memcpy(out, ctx->pattern.data(), ctx->pattern.size() * sizeof(int16_t));
    return BM_OK;
}

int bm_set_pattern(bm_ctx *ctx, const int16_t *pat)
{
    if (!ctx || !pat) return fail(nullptr, BM_E_PARAM, "null argument");
    std::vector<int16_t> v(pat, pat + 4 * (size_t)ctx->p.N);
    int q = quiesce(ctx);
    if (q) return q;
    const int keep = ctx->g.brief_near > 0 ? ctx->g.brief_near : -1;
    int r = set_pattern_internal(ctx, v, keep);
    if (r != BM_OK && keep > 0) r = set_pattern_internal(ctx, v, -1);
    return r;
}

int bm_get_gauss_kernel(const bm_ctx *ctx, float *out, unsigned *P)
{
    if (!ctx || !out || !P) return fail(nullptr, BM_E_PARAM, "null argument");
    std::memcpy(out, ctx->gauss.data(), ctx->gauss.size() * sizeof(float));
    *P = (unsigned)ctx->gauss.size();
    return BM_OK;
}

int bm_ctx_field(bm_ctx *ctx, uint32_t **off, uint16_t **cost)
{
    if (!ctx || !off || !cost) return fail(nullptr, BM_E_PARAM, "null argument");
    *off = ctx->d_off[ctx->field];
    *cost = ctx->d_cost[ctx->field];
    return BM_OK;
}

int bm_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(nullptr, BM_E_PARAM, "null argument");
    hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(nullptr, BM_E_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return BM_OK;
}

int bm_free(void *ptr)
{
    if (!ptr) return BM_OK;
    // the buffer may still be read by queued work on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFree(ptr);
    if (e != hipSuccess) return fail(nullptr, BM_E_LAUNCH, std::string("hipFree failed: ") + hipGetErrorString(e));
    return BM_OK;
}

int bm_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    if ((!dst || !src) && bytes) return fail(nullptr, BM_E_PARAM, "null pointer");
    hipMemcpyKind k;
    switch (kind) {
    case BM_HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
    case BM_DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
    case BM_DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
    case BM_HOST_TO_HOST: k = hipMemcpyHostToHost; break;
    default: return fail(nullptr, BM_E_PARAM, "unknown copy kind");
    }
    hipError_t e = hipMemcpy(dst, src, bytes, k);
    if (e != hipSuccess) return fail(nullptr, BM_E_LAUNCH, std::string("hipMemcpy failed: ") + hipGetErrorString(e));
    return BM_OK;
}

int bm_memset(void *ptr, int value, size_t bytes)
{
    if (!ptr && bytes) return fail(nullptr, BM_E_PARAM, "null pointer");
    hipError_t e = hipMemset(ptr, value, bytes);
    if (e != hipSuccess) return fail(nullptr, BM_E_LAUNCH, std::string("hipMemset failed: ") + hipGetErrorString(e));
    return BM_OK;
}

int bm_gray(bm_ctx *ctx, const uint8_t *rgb, uint8_t *g8, float *g32)
{
    int r = check(ctx, rgb, "rgb");
    if (r) return r;
    if (!g8 && !g32) return fail(ctx, BM_E_PARAM, "no output");
    BM_LAUNCH(ctx, BM_STAGE_GRAY, bm::launch_gray(ctx->g, rgb, g8, g32, ctx->stream));
    return BM_OK;
}

int bm_upsample(bm_ctx *ctx, const uint8_t *g8, float *L)
{
    int r = check(ctx, g8, "gray8");
    if (!r) r = check(ctx, L, "L");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_UPSAMPLE, bm::launch_upsample(ctx->g, g8, L, ctx->stream));
    return BM_OK;
}

int bm_upsample_lanczos(bm_ctx *ctx, const uint8_t *g8, float *L)
{
    int r = check(ctx, g8, "gray8");
    if (!r) r = check(ctx, L, "L");
    if (!r) r = lanczos_ready(ctx, false);
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_UPSAMPLE, bm::launch_upsample_lanczos(ctx->g, g8, L, ctx->stream));
    return BM_OK;
}

int bm_set_upsampler(bm_ctx *ctx, int mode)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    if (mode != BM_UPSAMPLE_BILINEAR && mode != BM_UPSAMPLE_LANCZOS3)
        return fail(ctx, BM_E_PARAM, "unknown up-sampler " + std::to_string(mode));
    if (mode == ctx->upsampler) return BM_OK;
    int r = quiesce(ctx);
    if (!r && mode == BM_UPSAMPLE_LANCZOS3) r = lanczos_ready(ctx, true);
    if (r) return r;
    ctx->upsampler = mode;
    ctx->gen++;
    return BM_OK;
}

int bm_get_upsampler(const bm_ctx *ctx, int *mode)
{
    if (!ctx || !mode) return fail(nullptr, BM_E_PARAM, "null argument");
    *mode = ctx->upsampler;
    return BM_OK;
}

int bm_brief(bm_ctx *ctx, const float *L, uint32_t *F)
{
    int r = check(ctx, L, "L");
    if (!r) r = check(ctx, F, "F");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief(ctx->g, L, nullptr, F, ctx->stream));
    return BM_OK;
}

int bm_brief_gray(bm_ctx *ctx, const uint8_t *g8, uint32_t *F)
{
    int r = check(ctx, g8, "gray8");
    if (!r) r = check(ctx, F, "F");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_BRIEF, bm::launch_brief(ctx->g, nullptr, g8, F, ctx->stream));
    return BM_OK;
}

int bm_init(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, int first, const uint32_t *prev, uint32_t *off,
            uint16_t *cost)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, off, "off");
    if (!r) r = check(ctx, cost, "cost");
    if (r) return r;
    if (!first && (!prev || prev == off)) return fail(ctx, BM_E_PARAM, "temporal init needs a separate prev_off");
    BM_LAUNCH(ctx, BM_STAGE_INIT, bm::launch_init(ctx->g, F1, F2, first, prev, off, cost, -1, 0, ctx->stream));
    return BM_OK;
}

int bm_propagate_level(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, const uint32_t *oi, const uint16_t *ci,
                       uint32_t *oo, uint16_t *co, int p, int pd)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, oi, "off_in");
    if (!r) r = check(ctx, ci, "cost_in");
    if (!r) r = check(ctx, oo, "off_out");
    if (!r) r = check(ctx, co, "cost_out");
    if (r) return r;
    if ((const void *)oi == (const void *)oo || (const void *)ci == (const void *)co)
        return fail(ctx, BM_E_PARAM, "level input and output must differ");
    if (p < 1 || (pd != 1 && pd != -1)) return fail(ctx, BM_E_PARAM, "bad level step or direction");
    BM_LAUNCH(ctx, BM_STAGE_LEVEL, bm::launch_level(ctx->g, F1, F2, oi, ci, oo, co, p, pd, -1, 0, ctx->stream));
    return BM_OK;
}

int bm_propagate(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, uint32_t *off, uint16_t *cost, int pd)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, off, "off");
    if (!r) r = check(ctx, cost, "cost");
    if (r) return r;
    if (pd != 1 && pd != -1) return fail(ctx, BM_E_PARAM, "pd must be +1 or -1");
    const size_t M = (size_t)ctx->g.sx * ctx->g.sy;
    // the caller's field is the first snapshot; levels ping-pong through the context
    uint32_t *o[2] = {off, ctx->d_off[1 - ctx->field]};
    uint16_t *c[2] = {cost, ctx->d_cost[1 - ctx->field]};
    int a = 0;
    for (int p = ctx->g.jfMax; p > 0; p /= 2) {
        BM_LAUNCH(ctx, BM_STAGE_LEVEL, bm::launch_level(ctx->g, F1, F2, o[a], c[a], o[1 - a], c[1 - a], p, pd, -1, 0, ctx->stream));
        a = 1 - a;
    }
    if (a != 0) {
        BM_HIP(ctx, hipMemcpyAsync(off, o[a], M * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
        BM_HIP(ctx, hipMemcpyAsync(cost, c[a], M * sizeof(uint16_t), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return BM_OK;
}

int bm_random_search(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, uint32_t *off, uint16_t *cost,
                     uint32_t iteration, uint32_t pair)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, off, "off");
    if (!r) r = check(ctx, cost, "cost");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_RANDOM_SEARCH, bm::launch_random_search(ctx->g, F1, F2, off, cost, iteration, pair, ctx->stream));
    return BM_OK;
}

int bm_match(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, int first, const uint32_t *prev, uint32_t pair,
             uint32_t *off, uint16_t *cost)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (r) return r;
    if (!first && !prev) return fail(ctx, BM_E_PARAM, "temporal init needs prev_off");
    r = run_match(ctx, F1, F2, first, prev, pair);
    if (r) return r;
    const size_t M = (size_t)ctx->g.sx * ctx->g.sy;
    if (off && off != ctx->d_off[ctx->field])
        BM_HIP(ctx, hipMemcpyAsync(off, ctx->d_off[ctx->field], M * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    if (cost && cost != ctx->d_cost[ctx->field])
        BM_HIP(ctx, hipMemcpyAsync(cost, ctx->d_cost[ctx->field], M * sizeof(uint16_t), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    return BM_OK;
}

int bm_downsample(bm_ctx *ctx, const uint32_t *off, float *U, float *V)
{
    int r = check(ctx, off, "off");
    if (!r) r = check(ctx, U, "U");
    if (!r) r = check(ctx, V, "V");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_DOWNSAMPLE, bm::launch_downsample(ctx->g, off, U, V, ctx->stream));
    return BM_OK;
}

int bm_median1d(bm_ctx *ctx, const float *in, float *out, int dir)
{
    int r = check(ctx, in, "in");
    if (!r) r = check(ctx, out, "out");
    if (r) return r;
    if (in == out) return fail(ctx, BM_E_PARAM, "median input and output must differ");
    BM_LAUNCH(ctx, BM_STAGE_MEDIAN, bm::launch_median(ctx->g, in, nullptr, out, nullptr, dir ? 1 : 0, ctx->stream));
    return BM_OK;
}

int bm_trilateral(bm_ctx *ctx, const float *Ui, const float *Vi, const float *Um, const float *Vm, const float *I,
                  float *Uo, float *Vo)
{
    int r = check(ctx, Ui, "Ui");
    if (!r) r = check(ctx, Vi, "Vi");
    if (!r) r = check(ctx, Um, "Um");
    if (!r) r = check(ctx, Vm, "Vm");
    if (!r) r = check(ctx, I, "I");
    if (!r) r = check(ctx, Uo, "Uo");
    if (!r) r = check(ctx, Vo, "Vo");
    if (r) return r;
    if (!ctx->p.filter) return fail(ctx, BM_E_STATE, "context created with filter = 0");
    BM_LAUNCH(ctx, BM_STAGE_TRILATERAL, bm::launch_trilateral(ctx->g, Ui, Vi, Um, Vm, I, Uo, Vo, ctx->stream));
    return BM_OK;
}

int bm_refine(bm_ctx *ctx, const float *U, const float *V, const float *I, float *Uo, float *Vo)
{
    int r = check(ctx, U, "U");
    if (!r) r = check(ctx, V, "V");
    if (!r) r = check(ctx, I, "I");
    if (!r) r = check(ctx, Uo, "Uo");
    if (!r) r = check(ctx, Vo, "Vo");
    if (r) return r;
    if (!ctx->p.filter) return fail(ctx, BM_E_STATE, "context created with filter = 0");
    return run_refine(ctx, U, V, I, Uo, Vo);
}

int bm_flow_pair(bm_ctx *ctx, const uint8_t *rgb1, const uint8_t *rgb2, uint32_t pair, float *U, float *V)
{
    int r = check(ctx, rgb1, "rgb1");
    if (!r) r = check(ctx, rgb2, "rgb2");
    if (!r) r = check(ctx, U, "U");
    if (!r) r = check(ctx, V, "V");
    if (r) return r;
    return run_pair(ctx, rgb1, rgb2, pair, U, V);
}

int bm_proposal_dims(const bm_ctx *ctx, unsigned *cw, unsigned *ch)
{
    if (!ctx || !cw || !ch) return fail(nullptr, BM_E_PARAM, "null argument");
    *cw = (unsigned)(ctx->g.W / bm::net::kCell);
    *ch = (unsigned)(ctx->g.H / bm::net::kCell);
    return BM_OK;
}

int bm_ctx_frames(bm_ctx *ctx, uint8_t **g8, uint32_t **F)
{
    if (!ctx || !g8 || !F) return fail(nullptr, BM_E_PARAM, "null argument");
    for (int i = 0; i < 2; ++i) {
        g8[i] = ctx->d_g8[i];
        F[i] = ctx->d_F[i];
    }
    return BM_OK;
}

namespace {

// The second stream and its events, created on first use.  A stream with a CU
// mask (here every CU) gets a hardware queue of its own; plain streams beyond
// the runtime's queue count share queues, and work on a shared queue runs in
// submission order, never side by side.
int side_ready(bm_ctx *ctx)
{
    if (ctx->side) return BM_OK;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<uint32_t> mask((size_t)std::max(1, (cus + 31) / 32), 0xffffffffu);
    hipStream_t s = nullptr;
    if (e == hipSuccess) e = hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data());
    if (e == hipSuccess && !ctx->fork) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e == hipSuccess && !ctx->join) e = hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming);
    if (e == hipSuccess && !ctx->tail_done) e = hipEventCreateWithFlags(&ctx->tail_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (s) (void)hipStreamDestroy(s);
        return hip_fail(ctx, e, "second stream");
    }
    ctx->side = s;
    return BM_OK;
}

int proposal_ready(bm_ctx *ctx)
{
    const int cw = ctx->g.W / bm::net::kCell, ch = ctx->g.H / bm::net::kCell;
    if (cw < 1 || ch < 1) return fail(ctx, BM_E_PARAM, "the learned proposal needs frames of at least 8x8 pixels");
    if (ctx->d_feat) return BM_OK;
    const size_t cells = (size_t)cw * ch;
    int r = dev_alloc(ctx, &ctx->d_feat, cells * bm::net::kFeatC);
    if (!r) r = dev_alloc(ctx, &ctx->d_nact[0], cells * 64);
    if (!r) r = dev_alloc(ctx, &ctx->d_nact[1], cells * 64);
    if (!r) r = dev_alloc(ctx, &ctx->d_nout, cells * bm::net::kOutC);
    if (!r) r = dev_alloc(ctx, &ctx->d_coff, cells * 4);
    if (!r) r = dev_alloc(ctx, &ctx->d_bsum, cells * 2);
    if (!r && cw >= 2 && ch >= 2) r = dev_alloc(ctx, &ctx->d_far, (size_t)(cw / 2) * (ch / 2) * 2);
    if (!r) r = side_ready(ctx);
    if (r) ctx->d_feat = nullptr;
    return r;
}

hipError_t launch_offsets(bm_ctx *ctx, const float *nout, const uint16_t *feat, const int32_t *far, int32_t *coff,
                          hipStream_t st)
{
    const bm::Geometry &g = ctx->g;
    const int cw = g.W / bm::net::kCell, ch = g.H / bm::net::kCell;
    const float scx = (float)(256 * bm::net::kCell * g.sx) / (float)g.W;
    const float scy = (float)(256 * bm::net::kCell * g.sy) / (float)g.H;
    return bm::net::launch_proposal_offsets(nout, feat, far, cw, ch, scx, scy, 256 * g.sx, 256 * g.sy, coff, st);
}

}  // namespace

int bm_proposal_features(bm_ctx *ctx, const uint8_t *g1, const uint8_t *g2, uint16_t *feat)
{
    int r = check(ctx, g1, "g1");
    if (!r) r = check(ctx, g2, "g2");
    if (!r) r = check(ctx, feat, "feat");
    if (!r) r = proposal_ready(ctx);
    if (r) return r;
    const bm::Geometry &g = ctx->g;
    const int cw = g.W / bm::net::kCell, ch = g.H / bm::net::kCell;
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL, bm::net::launch_features(g1, g2, g.W, cw, ch, ctx->d_bsum, feat, ctx->stream));
    return BM_OK;
}

int bm_proposal_far(bm_ctx *ctx, const uint8_t *g1, const uint8_t *g2, int32_t *far)
{
    int r = check(ctx, g1, "g1");
    if (!r) r = check(ctx, g2, "g2");
    if (!r) r = check(ctx, far, "far");
    if (!r && (ctx->g.W < bm::net::kFarCell || ctx->g.H < bm::net::kFarCell))
        r = fail(ctx, BM_E_PARAM, "the far block match needs frames of at least 16x16 pixels");
    if (!r) r = proposal_ready(ctx);
    if (r) return r;
    const bm::Geometry &g = ctx->g;
    const int cw = g.W / bm::net::kCell, ch = g.H / bm::net::kCell;
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL, bm::net::launch_block_sums(g1, g2, g.W, cw, ch, ctx->d_bsum, ctx->stream));
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL, bm::net::launch_far_search(ctx->d_bsum, cw, ch, far, ctx->stream));
    return BM_OK;
}

int bm_coarse_flow(bm_ctx *ctx, const float *U, const float *V, float *label)
{
    int r = check(ctx, U, "U");
    if (!r) r = check(ctx, V, "V");
    if (!r) r = check(ctx, label, "label");
    if (r) return r;
    const int cw = ctx->g.W / bm::net::kCell, ch = ctx->g.H / bm::net::kCell;
    if (cw < 1 || ch < 1) return fail(ctx, BM_E_PARAM, "the learned proposal needs frames of at least 8x8 pixels");
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL, bm::net::launch_coarse_flow(U, V, ctx->g.W, cw, ch, label, ctx->stream));
    return BM_OK;
}

int bm_proposal_offsets(bm_ctx *ctx, const float *nout, const uint16_t *feat, const int32_t *far, int32_t *coff)
{
    int r = check(ctx, nout, "net_out");
    if (!r) r = check(ctx, feat, "feat");
    if (!r) r = check(ctx, coff, "coff");
    if (r) return r;
    if (ctx->g.W < bm::net::kCell || ctx->g.H < bm::net::kCell)
        return fail(ctx, BM_E_PARAM, "the learned proposal needs frames of at least 8x8 pixels");
    if (far && (ctx->g.W < bm::net::kFarCell || ctx->g.H < bm::net::kFarCell))
        return fail(ctx, BM_E_PARAM, "the far block match needs frames of at least 16x16 pixels");
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL, launch_offsets(ctx, nout, feat, far, coff, ctx->stream));
    return BM_OK;
}

int bm_init_proposal(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, int first, const uint32_t *prev,
                     const int32_t *coff, uint32_t *off, uint16_t *cost)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, coff, "coff");
    if (!r) r = check(ctx, off, "off");
    if (!r) r = check(ctx, cost, "cost");
    if (r) return r;
    if (!first && !prev) return fail(ctx, BM_E_PARAM, "temporal init needs prev_off");
    if (!first && prev == off) return fail(ctx, BM_E_PARAM, "prev_off and off must differ");
    const bm::Geometry &g = ctx->g;
    const int cw = g.W / bm::net::kCell, ch = g.H / bm::net::kCell;
    if (cw < 1 || ch < 1) return fail(ctx, BM_E_PARAM, "the learned proposal needs frames of at least 8x8 pixels");
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL,
              bm::launch_init_proposal(g, F1, F2, first, prev, coff, cw, ch, -1, 0, off, cost, ctx->stream));
    return BM_OK;
}

int bm_set_local_search(bm_ctx *ctx, int radius, int samples)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    if (samples < 0 || samples > BM_LOCAL_SEARCH_MAX_SAMPLES)
        return fail(ctx, BM_E_PARAM, "local search samples must be in [0, " +
                                         std::to_string(BM_LOCAL_SEARCH_MAX_SAMPLES) + "]");
    if (samples > 0 && (radius < 1 || radius > BM_LOCAL_SEARCH_MAX_RADIUS))
        return fail(ctx, BM_E_PARAM, "local search radius must be in [1, " +
                                         std::to_string(BM_LOCAL_SEARCH_MAX_RADIUS) + "]");
    ctx->ls_radius = samples > 0 ? radius : 0;
    ctx->ls_samples = samples;
    return BM_OK;
}

int bm_get_local_search(const bm_ctx *ctx, int *radius, int *samples)
{
    if (!ctx || !radius || !samples) return fail(nullptr, BM_E_PARAM, "null argument");
    *radius = ctx->ls_radius;
    *samples = ctx->ls_samples;
    return BM_OK;
}

int bm_local_search(bm_ctx *ctx, const uint32_t *F1, const uint32_t *F2, uint32_t *off, uint16_t *cost,
                    uint32_t iteration, uint32_t pair)
{
    int r = check(ctx, F1, "F1");
    if (!r) r = check(ctx, F2, "F2");
    if (!r) r = check(ctx, off, "off");
    if (!r) r = check(ctx, cost, "cost");
    if (r) return r;
    BM_LAUNCH(ctx, BM_STAGE_LOCAL_SEARCH,
              bm::launch_local_search(ctx->g, F1, F2, ctx->ls_radius, ctx->ls_samples, iteration, pair, off, cost,
                                      ctx->stream));
    return BM_OK;
}

int bm_flow_pair_learned(bm_ctx *ctx, bm_net *net, const uint8_t *rgb1, const uint8_t *rgb2, uint32_t pair, float *U,
                         float *V)
{
    int r = check(ctx, net, "net");
    if (!r) r = check(ctx, rgb1, "rgb1");
    if (!r) r = check(ctx, rgb2, "rgb2");
    if (!r) r = check(ctx, U, "U");
    if (!r) r = check(ctx, V, "V");
    if (!r) r = proposal_ready(ctx);
    if (r) return r;
    const bm::Geometry &g = ctx->g;
    const int cw = g.W / bm::net::kCell, ch = g.H / bm::net::kCell;
    BM_LAUNCH(ctx, BM_STAGE_GRAY,
              bm::launch_gray_pair(g, rgb1, ctx->d_g8[0], ctx->d_g32[0], rgb2, ctx->d_g8[1], ctx->stream));
    // the proposal (features, network, offsets) needs only the gray frames: it
    // runs on the side stream while the descriptor kernels run on the main one
    hipStream_t side = ctx->side;
    BM_HIP(ctx, hipEventRecord(ctx->fork, ctx->stream));
    BM_HIP(ctx, hipStreamWaitEvent(side, ctx->fork, 0));
    BM_LAUNCH_ON(ctx, BM_STAGE_PROPOSAL, side,
                 bm::net::launch_features(ctx->d_g8[0], ctx->d_g8[1], g.W, cw, ch, ctx->d_bsum, ctx->d_feat, side));
    if (ctx->d_far)
        BM_LAUNCH_ON(ctx, BM_STAGE_PROPOSAL, side, bm::net::launch_far_search(ctx->d_bsum, cw, ch, ctx->d_far, side));
    // the packed weights are written on the net's stream
    BM_HIP(ctx, hipStreamWaitEvent(side, net->ready, 0));
    BM_LAUNCH_ON(ctx, BM_STAGE_PROPOSAL, side,
                 bm::net::forward(net->pk, ctx->d_feat, 1, ch, cw, ctx->d_nact[0], ctx->d_nact[1], ctx->d_nout, side));
    BM_LAUNCH_ON(ctx, BM_STAGE_PROPOSAL, side,
                 launch_offsets(ctx, ctx->d_nout, ctx->d_feat, ctx->d_far, ctx->d_coff, side));
    BM_HIP(ctx, hipEventRecord(ctx->join, side));
    r = describe_pair(ctx, ctx->d_g8[0], ctx->d_F[0], ctx->d_g8[1], ctx->d_F[1]);
    if (r) return r;
    BM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->join, 0));
    const int dst = ctx->field;
    // the seeded field, then one random search around it on its own stream
    // index (the propagation iterations use 0 .. iterations - 1), in one kernel
    BM_LAUNCH(ctx, BM_STAGE_PROPOSAL,
              bm::launch_init_proposal(g, ctx->d_F[0], ctx->d_F[1], 1, nullptr, ctx->d_coff, cw, ch, g.iterations,
                                       pair, ctx->d_off[dst], ctx->d_cost[dst], ctx->stream));
    if (ctx->ls_samples > 0)
        BM_LAUNCH(ctx, BM_STAGE_LOCAL_SEARCH,
                  bm::launch_local_search(g, ctx->d_F[0], ctx->d_F[1], ctx->ls_radius, ctx->ls_samples,
                                          (uint32_t)g.iterations, pair, ctx->d_off[dst], ctx->d_cost[dst],
                                          ctx->stream));
    int pd = -1;
    for (int i = 0; i < g.iterations; ++i) {
        r = run_iteration(ctx, ctx->d_F[0], ctx->d_F[1], pd, i, pair, true);
        if (r) return r;
        pd = -pd;
    }
    ctx->have_prev = 0;
    ctx->field_valid = 0;
    return run_flow_tail(ctx, ctx->d_g32[0], U, V);
}

int bm_set_pipelining(bm_ctx *ctx, int on)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    if (!on) {
        const int r = settle_tail(ctx);
        if (r) return r;
        ctx->pipelined = 0;
        return BM_OK;
    }
    int r = side_ready(ctx);
    if (!r && !ctx->d_guide) r = dev_alloc(ctx, &ctx->d_guide, (size_t)ctx->g.W * ctx->g.H);
    if (r) return r;
    ctx->pipelined = 1;
    return BM_OK;
}

int bm_stream_reset(bm_ctx *ctx)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    ctx->have_prev = 0;
    ctx->field_valid = 0;
    ctx->frame_counter = 0;
    return BM_OK;
}

int bm_stream_frame(bm_ctx *ctx, const uint8_t *rgb, float *U, float *V, int *has_flow)
{
    int r = check(ctx, rgb, "rgb");
    if (!r) r = check(ctx, has_flow, "has_flow");
    if (r) return r;
    const bm::Geometry &g = ctx->g;
    const int cur = ctx->cur, prev = 1 - cur;
    BM_LAUNCH(ctx, BM_STAGE_GRAY, bm::launch_gray(g, rgb, ctx->d_g8[cur], ctx->d_g32[cur], ctx->stream));
    r = describe(ctx, ctx->d_g8[cur], ctx->d_F[cur]);
    if (r) return r;
    *has_flow = 0;
    if (ctx->have_prev) {
        if (!U || !V) return fail(ctx, BM_E_PARAM, "null flow output");
        const int first = (ctx->p.tInit && ctx->field_valid) ? 0 : 1;
        r = run_match(ctx, ctx->d_F[prev], ctx->d_F[cur], first, first ? nullptr : ctx->d_off[ctx->field],
                      ctx->frame_counter);
        if (r) return r;
        if (ctx->pipelined && ctx->p.filter) {
            // down-sample here, then medians and trilateral on the second stream:
            // the next frame's descriptors and matching run beside them and wait
            // for them only before its own down-sampling (settle_tail)
            r = settle_tail(ctx);
            if (r) return r;
            BM_LAUNCH(ctx, BM_STAGE_DOWNSAMPLE,
                      bm::launch_downsample(g, ctx->d_off[ctx->field], ctx->d_U0, ctx->d_V0, ctx->stream));
            // the guide is the previous frame, whose gray buffer the next call overwrites
            BM_HIP(ctx, hipMemcpyAsync(ctx->d_guide, ctx->d_g32[prev], (size_t)g.W * g.H * sizeof(float),
                                       hipMemcpyDeviceToDevice, ctx->stream));
            BM_HIP(ctx, hipEventRecord(ctx->fork, ctx->stream));
            BM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork, 0));
            r = run_refine_on(ctx, ctx->d_U0, ctx->d_V0, ctx->d_guide, U, V, ctx->side);
            if (r) return r;
            BM_HIP(ctx, hipEventRecord(ctx->tail_done, ctx->side));
            ctx->tail_pending = 1;
        } else {
            r = run_flow_tail(ctx, ctx->d_g32[prev], U, V);
            if (r) return r;
        }
        ctx->field_valid = 1;
        ctx->frame_counter++;
        *has_flow = 1;
    }
    ctx->have_prev = 1;
    ctx->cur = prev;
    return BM_OK;
}

int bm_timer_start(bm_ctx *ctx)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    BM_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return BM_OK;
}

int bm_timer_stop(bm_ctx *ctx, float *ms)
{
    if (!ctx || !ms) return fail(nullptr, BM_E_PARAM, "null argument");
    const int r = settle_tail(ctx);
    if (r) return r;
    BM_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    BM_HIP(ctx, hipEventSynchronize(ctx->ev1));
    BM_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return BM_OK;
}

int bm_profile_enable(bm_ctx *ctx, int on)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    ctx->prof = on != 0;
    return BM_OK;
}

int bm_profile_read(bm_ctx *ctx, double *ms, unsigned *launches)
{
    if (!ctx) return fail(nullptr, BM_E_PARAM, "null context");
    BM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) BM_HIP(ctx, hipStreamSynchronize(ctx->side));
    for (auto &pe : ctx->pending) {
        float t = 0.0f;
        BM_HIP(ctx, hipEventElapsedTime(&t, pe.a, pe.b));
        ctx->prof_ms[pe.stage] += t;
        ctx->prof_n[pe.stage] += 1;
        ctx->pool.push_back(pe.a);
        ctx->pool.push_back(pe.b);
    }
    ctx->pending.clear();
    for (int i = 0; i < BM_STAGE_COUNT; ++i) {
        if (ms) ms[i] = ctx->prof_ms[i];
        if (launches) launches[i] = ctx->prof_n[i];
        ctx->prof_ms[i] = 0.0;
        ctx->prof_n[i] = 0;
    }
    return BM_OK;
}

int bm_device_copy(bm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    int r = check(ctx, dst, "dst");
    if (!r) r = check(ctx, src, "src");
    if (r) return r;
    if (bytes % 16) return fail(ctx, BM_E_PARAM, "copy size must be a multiple of 16 bytes");
    BM_LAUNCH(ctx, BM_STAGE_COPY, bm::launch_copy(src, dst, bytes, ctx->stream));
    return BM_OK;
}

int bm_epe(const float *U, const float *V, const float *gu, const float *gv, const uint8_t *occ, size_t n,
           double *epe_all, double *epe_noc)
{
    if (!U || !V || !gu || !gv || !epe_all || !epe_noc) return fail(nullptr, BM_E_PARAM, "null argument");
    double sa = 0.0, sn = 0.0;
    size_t na = 0, nn = 0;
    for (size_t i = 0; i < n; ++i) {
        const float a = gu[i], b = gv[i];
        if (std::fabs(a) > 1e9f || std::fabs(b) > 1e9f || std::isnan(a) || std::isnan(b)) continue;
        const double du = (double)a - U[i], dv = (double)b - V[i];
        const double e = std::sqrt(du * du + dv * dv);
        sa += e;
        ++na;
        if (!occ || !occ[i]) {
            sn += e;
            ++nn;
        }
    }
    *epe_all = na ? sa / (double)na : 0.0;
    *epe_noc = nn ? sn / (double)nn : 0.0;
    return BM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// tuning, autotuning, batches

// One multi-pair launch of a batch's plan.
struct MultiOp {
    enum Kind { GRAY, BRIEF, INIT, LEVEL, DOWNSAMPLE, MEDIAN, TRILATERAL } kind;
    int a = 0, b = 0;                     // template keys: (NSP, RN), (NSP, PX), FS, trilateral groups, tiled
    int p = 0, pd = 0, src = 0, rs = -1;  // level / init arguments
    int dir = 0;                          // median direction
    size_t jobs = 0, start = 0;           // byte offsets of the job and tile-start arrays in the plan buffer
    int njobs = 0, tiles = 0;
    size_t lds = 0;
};

struct bm_batch {
    std::vector<bm_ctx *> ctx;
    std::vector<hipStream_t> streams;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> joins;
    bm_batch_schedule sched{BM_BATCH_GRAPH, 0, 0, 0};
    int cus = 256;
    // captured graph of the graph mode / of the multi-pair mode, with the key
    // (buffers, configuration generations, schedule) each was captured for
    hipGraph_t graph = nullptr, mgraph = nullptr;
    hipGraphExec_t exec = nullptr, mexec = nullptr;
    std::vector<uintptr_t> key, mkey;
    // multi-pair plan: launches, their job arrays on the device, the field
    // buffer each pair ends in, and the pairs that take the per-pair path
    std::vector<std::vector<MultiOp>> lanes;  // launches of each lane (a subset of the pairs)
    void *plan_dev = nullptr;
    std::vector<int> final_field, fallback, lane_of;
    std::vector<uintptr_t> pkey;
    // the field buffer each pair's result lands in when the captured graph of
    // the graph mode / the multi-pair mode replays (host state the replay
    // itself does not update)
    std::vector<int> gfield, mfield;
    std::string err;
};

namespace {

int tuning_valid(const bm_ctx *ctx, const bm_tuning &t, std::string &why)
{
    if (t.level_px != 1 && t.level_px != 2 && t.level_px != 4) return why = "level_px must be 1, 2 or 4", 0;
    if (t.level_px == 2 && ctx->g.NSP > 8) return why = "level_px 2 needs at most 256 feature bits", 0;
    if (t.level_px == 4 && ctx->g.NSP > 4) return why = "level_px 4 needs at most 128 feature bits", 0;
    if (t.strip_rows < 16 || t.strip_rows > 4096) return why = "strip_rows outside [16, 4096]", 0;
    if (!(t.brief_near == 0 || (t.brief_near % 8 == 0 && t.brief_near >= 8 && t.brief_near <= 48)))
        return why = "brief_near must be 0, 8, 16, 24, 32, 40 or 48", 0;
    const int maxb = ctx->p.filter ? bm::tri_max_band(ctx->g.tri_s) : 16;
    if (t.tri_band < 1 || t.tri_band > std::max(1, maxb)) return why = "tri_band outside the LDS limit", 0;
    if (t.tri_groups != 0 && t.tri_groups != 2 && t.tri_groups != 4) return why = "tri_groups must be 0, 2 or 4", 0;
    if (t.brief_frame != 0 && t.brief_frame != 1) return why = "brief_frame must be 0 or 1", 0;
    if (t.brief_frame && t.brief_near == 0) return why = "brief_frame 1 needs brief_near > 0", 0;
    if (t.tri_split != 1 && t.tri_split != 2 && t.tri_split != 4 && t.tri_split != 8)
        return why = "tri_split must be 1, 2, 4 or 8", 0;
    return 1;
}

int apply_tuning(bm_ctx *ctx, const bm_tuning &t)
{
    std::string why;
    if (!tuning_valid(ctx, t, why)) return fail(ctx, BM_E_PARAM, why);
    const int q = quiesce(ctx);
    if (q) return q;
    ctx->gen++;
    if (t.brief_near != ctx->g.brief_near) {
        const std::vector<int16_t> pat = ctx->pattern;
        const int r = set_pattern_internal(ctx, pat, t.brief_near);
        if (r != BM_OK) return r;
    }
    if (t.brief_frame && !ctx->d_L) {
        const int r = dev_alloc(ctx, &ctx->d_L, (size_t)ctx->g.sx * ctx->g.sy);
        if (r != BM_OK) return r;
    }
    if (t.tri_split > 1) {  // three planes of W H partial sums per part
        const size_t need = (size_t)3 * t.tri_split * ctx->g.W * ctx->g.H;
        if (need > ctx->tri_part_cap) {
            // the smaller buffer goes (the context's streams are idle, and a
            // captured batch graph that holds it is re-captured: generation)
            if (ctx->d_tri_part) {
                const auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), (void *)ctx->d_tri_part);
                if (it != ctx->allocs.end()) ctx->allocs.erase(it);
                (void)hipFree(ctx->d_tri_part);
                ctx->d_tri_part = nullptr;
                ctx->g.tri_part = nullptr;
                ctx->g.tri_split = 1;
                ctx->tri_part_cap = 0;
            }
            const int r = dev_alloc(ctx, &ctx->d_tri_part, need);
            if (r != BM_OK) return r;
            ctx->tri_part_cap = need;
        }
    }
    ctx->g.level_px = t.level_px;
    ctx->g.strip_rows = t.strip_rows;
    ctx->g.tri_band = t.tri_band;
    ctx->g.tri_groups = t.tri_groups;
    ctx->g.brief_frame = t.brief_frame;
    ctx->g.tri_split = t.tri_split;
    ctx->g.tri_part = ctx->d_tri_part;
    return BM_OK;
}

bm_tuning current_tuning(const bm_ctx *ctx)
{
    bm_tuning t;
    t.level_px = ctx->g.level_px;
    t.strip_rows = ctx->g.strip_rows;
    t.brief_near = ctx->g.brief_near;
    t.tri_band = ctx->g.tri_band;
    t.tri_groups = ctx->g.tri_groups;
    t.brief_frame = ctx->g.brief_frame;
    t.tri_split = ctx->g.tri_split;
    return t;
}

std::string tuning_key(const bm_ctx *ctx)
{
    hipDeviceProp_t prop;
    int dev = 0;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
    }
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    std::ostringstream k;
    k << arch << "/" << cus << "cu/" << ctx->g.W << "x" << ctx->g.H << "/" << ctx->g.sx << "x" << ctx->g.sy << "/n"
      << ctx->g.N << "/m" << ctx->g.fsize << "/s" << ctx->g.tri_s << "/r" << pattern_rmax(ctx) << "/f"
      << ctx->p.filter;
    return k.str();
}

}  // namespace

extern "C" {

int bm_get_tuning(const bm_ctx *ctx, bm_tuning *out)
{
    if (!ctx || !out) return fail(nullptr, BM_E_PARAM, "null argument");
    *out = current_tuning(ctx);
    return BM_OK;
}

int bm_set_tuning(bm_ctx *ctx, const bm_tuning *t)
{
    if (!ctx || !t) return fail(nullptr, BM_E_PARAM, "null argument");
    return apply_tuning(ctx, *t);
}

int bm_autotune(bm_ctx *ctx, const uint8_t *rgb1, const uint8_t *rgb2, const char *cache_path, bm_tuning *out)
{
    int r = check(ctx, rgb1, "rgb1");
    if (!r) r = check(ctx, rgb2, "rgb2");
    if (r) return r;
    const std::string key = tuning_key(ctx);
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            bm_tuning t;
            if (ls >> k >> t.level_px >> t.strip_rows >> t.brief_near >> t.tri_band >> t.tri_groups >> t.brief_frame >>
                    t.tri_split &&
                k == key) {
                std::string why;
                if (tuning_valid(ctx, t, why) && apply_tuning(ctx, t) == BM_OK) {
                    if (out) *out = t;
                    return BM_OK;
                }
            }
        }
    }
    const size_t Win = (size_t)ctx->g.W * ctx->g.H;
    float *U = nullptr, *V = nullptr;
    BM_HIP(ctx, hipMalloc(&U, Win * sizeof(float)));
    if (hipMalloc(&V, Win * sizeof(float)) != hipSuccess) {
        (void)hipFree(U);
        return fail(ctx, BM_E_ALLOC, "autotune scratch allocation failed");
    }
    const bool prof = ctx->prof;
    ctx->prof = true;
    // device time of the stages a knob affects: the least of three runs after
    // a warm-up run (one run's time of a sub-millisecond stage varies by a few
    // per cent, as much as the candidates differ)
    auto measure = [&](const bm_tuning &t, std::initializer_list<int> stages, double &ms) -> int {
        int e = apply_tuning(ctx, t);
        if (e != BM_OK) return e;
        double tot[BM_STAGE_COUNT];
        unsigned n[BM_STAGE_COUNT];
        ms = -1.0;
        for (int rep = 0; rep < 4; ++rep) {
            e = bm_profile_read(ctx, nullptr, nullptr);
            if (e == BM_OK) e = bm_flow_pair(ctx, rgb1, rgb2, 0, U, V);
            if (e == BM_OK) e = bm_profile_read(ctx, tot, n);
            if (e != BM_OK) return e;
            if (rep == 0) continue;
            double sum = 0.0;
            for (int st : stages) sum += tot[st];
            if (ms < 0.0 || sum < ms) ms = sum;
        }
        return BM_OK;
    };
    bm_tuning best = current_tuning(ctx);
    struct Knob {
        int bm_tuning::*field;
        std::vector<int> values;
        std::initializer_list<int> stages;
    };
    const std::initializer_list<int> lv = {BM_STAGE_INIT, BM_STAGE_LEVEL, BM_STAGE_LEVEL_RS};
    const std::initializer_list<int> br = {BM_STAGE_BRIEF, BM_STAGE_UPSAMPLE};
    const std::initializer_list<int> tr = {BM_STAGE_TRILATERAL};
    std::vector<int> bands;
    const int maxb = ctx->p.filter ? bm::tri_max_band(ctx->g.tri_s) : 0;
    for (int v : {4, 8, 12, 16})
        if (v <= maxb) bands.push_back(v);
    // the BRIEF knobs interact (with the frame in memory the far tests are
    // cheap and a smaller near halo, with more work-groups per CU, wins), so
    // they are timed jointly first
    int e = BM_OK;
    {
        double best_ms = -1.0;
        bm_tuning winner = best;
        for (int frame : {0, 1})
            for (int near : {16, 24, 32, 40, 48}) {
                bm_tuning t = best;
                t.brief_frame = frame;
                t.brief_near = near;
                std::string why;
                if (!tuning_valid(ctx, t, why)) continue;
                const int save = ctx->g.brief_near;  // check that the near halo fits before timing it
                const bool fits = plan_brief(ctx, pattern_rmax(ctx), near);
                plan_brief(ctx, pattern_rmax(ctx), save);
                if (!fits) continue;
                double ms = 0.0;
                e = measure(t, br, ms);
                if (e != BM_OK) break;
                if (best_ms < 0 || ms < best_ms) {
                    best_ms = ms;
                    winner = t;
                }
            }
        best = winner;
    }
    // the trilateral kernel and its split interact too (the split gives the
    // device work where tiles are few, which is also where the MFMA kernel's
    // smaller tiles help): timed jointly, unless the pair is not filtered
    if (ctx->p.filter) {
        double best_ms = -1.0;
        bm_tuning winner = best;
        for (int groups : {0, 2, 4})
            for (int split : {1, 2, 4, 8}) {
                if (e != BM_OK) break;
                bm_tuning t = best;
                t.tri_groups = groups;
                t.tri_split = split;
                std::string why;
                if (!tuning_valid(ctx, t, why)) continue;
                double ms = 0.0;
                e = measure(t, tr, ms);
                if (e != BM_OK) break;
                if (best_ms < 0 || ms < best_ms) {
                    best_ms = ms;
                    winner = t;
                }
            }
        best = winner;
    }
    const Knob knobs[] = {{&bm_tuning::level_px, {1, 2, 4}, lv},
                          {&bm_tuning::strip_rows, {64, 128, 256, 512, 1024}, lv},
                          {&bm_tuning::tri_band, bands, tr}};
    for (const Knob &kn : knobs) {
        if (e != BM_OK) break;
        double best_ms = -1.0;
        bm_tuning winner = best;
        for (int v : kn.values) {
            bm_tuning t = best;
            t.*(kn.field) = v;
            std::string why;
            if (!tuning_valid(ctx, t, why)) continue;
            if (t.brief_near > 0) {  // check that the near halo fits before timing it
                const int save = ctx->g.brief_near;
                const bool fits = plan_brief(ctx, pattern_rmax(ctx), t.brief_near);
                plan_brief(ctx, pattern_rmax(ctx), save);
                if (!fits) continue;
            }
            double ms = 0.0;
            e = measure(t, kn.stages, ms);
            if (e != BM_OK) break;
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                winner = t;
            }
        }
        if (e != BM_OK) break;
        best = winner;
    }
    ctx->prof = prof;
    (void)hipFree(U);
    (void)hipFree(V);
    if (e != BM_OK) return e;
    e = apply_tuning(ctx, best);
    if (e != BM_OK) return e;
    if (out) *out = best;
    if (cache_path) {
        std::ofstream o(cache_path, std::ios::app);
        o << key << " " << best.level_px << " " << best.strip_rows << " " << best.brief_near << " " << best.tri_band
          << " " << best.tri_groups << " " << best.brief_frame << " " << best.tri_split << "\n";
    }
    return BM_OK;
}

const char *bm_batch_last_error(const bm_batch *b) { return b ? b->err.c_str() : g_thread_error.c_str(); }

void bm_batch_destroy(bm_batch *b)
{
    if (!b) return;
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->mexec) (void)hipGraphExecDestroy(b->mexec);
    if (b->mgraph) (void)hipGraphDestroy(b->mgraph);
    if (b->plan_dev) (void)hipFree(b->plan_dev);
    for (bm_ctx *c : b->ctx) bm_ctx_destroy(c);
    for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->fork, b->ev0, b->ev1})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

int bm_batch_create(int n, const bm_params *params, uint64_t seed, int streams, bm_batch **out)
{
    if (!params || !out || n <= 0) return fail(nullptr, BM_E_PARAM, "bad batch arguments");
    *out = nullptr;
    bm_batch *b = new (std::nothrow) bm_batch;
    if (!b) return fail(nullptr, BM_E_ALLOC, "out of host memory");
    for (int i = 0; i < n; ++i) {
        bm_ctx *c = nullptr;
        const int r = bm_ctx_create(&params[i], seed, &c);
        if (r != BM_OK) {
            std::string m = "pair " + std::to_string(i) + ": " + g_thread_error;
            bm_batch_destroy(b);
            return fail(nullptr, r, m);
        }
        b->ctx.push_back(c);
    }
    streams = std::max(1, std::min(streams, n));
    hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
    for (int k = 0; k < streams && e == hipSuccess; ++k) {
        hipStream_t s = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) b->streams.push_back(s);
        hipEvent_t j = nullptr;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(j);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    int dev = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
        std::string m = std::string("HIP stream/event creation failed: ") + hipGetErrorString(e);
        bm_batch_destroy(b);
        return fail(nullptr, BM_E_LAUNCH, m);
    }
    for (bm_ctx *c : b->ctx) c->batch_stream = b->origin;
    b->sched.branches = (int)b->streams.size();
    *out = b;
    return BM_OK;
}

int bm_batch_ctx(bm_batch *b, int i, bm_ctx **ctx)
{
    if (!b || !ctx || i < 0 || i >= (int)b->ctx.size()) return fail(nullptr, BM_E_PARAM, "bad batch index");
    *ctx = b->ctx[i];
    return BM_OK;
}

}  // extern "C"

namespace {

int batch_fail(bm_batch *b, int code, const std::string &m)
{
    b->err = m;
    return code;
}

// Work estimate of a pair for ordering: matched pixels (matching stages) plus
// trilateral taps per input pixel over 64 (refinement).
double pair_work(const bm_ctx *c)
{
    const double M = (double)c->g.sx * c->g.sy, Win = (double)c->g.W * c->g.H;
    const double taps = c->p.filter ? 4.0 * c->g.tri_s * c->g.tri_s : 0.0;
    return M * (1.0 + 0.25 * c->g.NSP) + Win * taps / 64.0;
}

// Pair indices in the schedule's order: index order, or largest first.
std::vector<int> pair_order(const bm_batch *b)
{
    std::vector<int> idx(b->ctx.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    if (b->sched.order == 1)
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int x, int y) { return pair_work(b->ctx[x]) > pair_work(b->ctx[y]); });
    return idx;
}

// Per-pair launches: pair i on branch branch[i] (all on the origin stream when one_stream).
int batch_enqueue(bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2, float *const *U, float *const *V,
                  bool one_stream)
{
    const int n = (int)b->ctx.size();
    const int S = std::max(1, std::min(b->sched.branches, (int)b->streams.size()));
    std::vector<int> branch(n, 0);
    if (b->sched.order == 1) {  // longest first onto the least loaded branch
        std::vector<double> load(S, 0.0);
        for (int i : pair_order(b)) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            branch[i] = k;
            load[k] += pair_work(b->ctx[i]);
        }
    } else {
        for (int i = 0; i < n; ++i) branch[i] = i % S;
    }
    for (int i = 0; i < n; ++i) {
        bm_ctx *c = b->ctx[i];
        hipStream_t own = c->stream;
        const bool prof = c->prof;
        c->stream = one_stream ? b->origin : b->streams[branch[i]];
        c->prof = false;
        const int r = bm_flow_pair(c, r1[i], r2[i], (uint32_t)i, U[i], V[i]);
        c->stream = own;
        c->prof = prof;
        if (r != BM_OK) return batch_fail(b, r, "pair " + std::to_string(i) + ": " + c->err);
    }
    return BM_OK;
}

// What a captured graph depends on: every pair's buffers and configuration
// generation, and the schedule.
std::vector<uintptr_t> batch_key(const bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2,
                                 float *const *U, float *const *V)
{
    std::vector<uintptr_t> key;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        key.push_back((uintptr_t)r1[i]);
        key.push_back((uintptr_t)r2[i]);
        key.push_back((uintptr_t)U[i]);
        key.push_back((uintptr_t)V[i]);
        key.push_back((uintptr_t)b->ctx[i]->gen);
    }
    key.push_back((uintptr_t)b->sched.branches);
    key.push_back((uintptr_t)b->sched.order);
    key.push_back((uintptr_t)b->sched.persist);
    return key;
}

// Host image of the plan buffer: job arrays and tile-start arrays, 16-byte aligned.
struct PlanBytes {
    std::vector<unsigned char> buf;
    template <class T>
    size_t put(const std::vector<T> &v)
    {
        const size_t at = (buf.size() + 15) & ~(size_t)15;
        buf.resize(at + std::max<size_t>(v.size() * sizeof(T), 1));
        if (!v.empty()) std::memcpy(buf.data() + at, v.data(), v.size() * sizeof(T));
        return at;
    }
};

// One launch over a group's jobs (job j has tiles_per_job[j] tiles); the job
// array is stored at `jobs_at`, or appended when jobs_at is npos.
template <class Job>
MultiOp group_op(PlanBytes &pb, MultiOp::Kind kind, int a, int bk, const std::vector<Job> &jobs,
                 const std::vector<int> &tiles_per_job, size_t lds, size_t jobs_at = std::string::npos)
{
    MultiOp op;
    op.kind = kind;
    op.a = a;
    op.b = bk;
    // slot starts padded to multiples of 8 (each job spread over all XCDs as
    // in its own launch), then the jobs' tile counts
    const size_t nj = tiles_per_job.size();
    std::vector<int> start(2 * nj + 1, 0);
    for (size_t j = 0; j < nj; ++j) {
        start[j + 1] = start[j] + (tiles_per_job[j] + 7) / 8 * 8;
        start[nj + 1 + j] = tiles_per_job[j];
    }
    op.jobs = jobs_at == std::string::npos ? pb.put(jobs) : jobs_at;
    op.start = pb.put(start);
    op.njobs = (int)nj;
    op.tiles = start[nj];
    op.lds = lds;
    return op;
}

// A pair runs in the multi-pair launches when its context uses the fused
// bilinear BRIEF and the jump-flooding levels; otherwise it takes the
// per-pair path (bm_flow_pair) inside the same sequence.
bool multi_eligible(const bm_ctx *c)
{
    // a split trilateral filter sums in another order than the multi-pair
    // kernel: such a pair keeps its own launches
    return c->upsampler == BM_UPSAMPLE_BILINEAR && c->g.brief_near > 0 && c->g.jfMax >= 1 && c->g.tri_split <= 1;
}

// The plan of the multi-pair mode for these buffers: every stage of every
// eligible pair as one launch per kernel variant, the job arrays in one
// device buffer (re-made only when a buffer, a configuration or the schedule changes).
// The launches of one lane: every stage of its pairs as one launch per kernel variant.
int plan_lane(bm_batch *b, PlanBytes &pb, const std::vector<int> &pairs, const uint8_t *const *r1,
              const uint8_t *const *r2, float *const *U, float *const *V, std::vector<MultiOp> &ops)
{
    // gray: both frames of every pair
    {
        std::vector<bm::GrayJob> jobs;
        std::vector<int> tiles;
        for (int i : pairs) {
            bm_ctx *c = b->ctx[i];
            const int nW = c->g.W * c->g.H;
            jobs.push_back({r1[i], c->d_g8[0], c->d_g32[0], nW});
            jobs.push_back({r2[i], c->d_g8[1], nullptr, nW});
            tiles.push_back(bm::gray_tiles(nW));
            tiles.push_back(bm::gray_tiles(nW));
        }
        if (!jobs.empty()) ops.push_back(group_op(pb, MultiOp::GRAY, 0, 0, jobs, tiles, 0));
    }
    // BRIEF: per (descriptor stride, near halo), both frames
    std::vector<std::pair<int, int>> bkeys;
    for (int i : pairs) {
        const std::pair<int, int> k(b->ctx[i]->g.NSP, b->ctx[i]->g.brief_near);
        if (std::find(bkeys.begin(), bkeys.end(), k) == bkeys.end()) bkeys.push_back(k);
    }
    for (const auto &k : bkeys) {
        std::vector<bm::BriefJob> jobs;
        std::vector<int> tiles;
        size_t lds = 0;
        for (int i : pairs) {
            bm_ctx *c = b->ctx[i];
            const bm::Geometry &g = c->g;
            if (g.NSP != k.first || g.brief_near != k.second) continue;
            const int w = bm::brief_tile_rows(g.NSP) == 32 ? 1 : 0;
            for (int f = 0; f < 2; ++f) {
                jobs.push_back({c->d_g8[f], g.up_x, g.up_y, g.brief_plan, g.brief_plan_count, g.N, g.W, g.H, g.sx,
                                g.sy, g.brief_far, g.brief_gw[w], g.brief_gh[w], c->d_F[f]});
                tiles.push_back(bm::brief_tiles(g));
            }
            lds = std::max(lds, bm::brief_lds(g));
        }
        ops.push_back(group_op(pb, MultiOp::BRIEF, k.first, k.second, jobs, tiles, lds));
    }
    // matching: per (stride, pixels per thread, jump range, iterations)
    std::vector<std::array<int, 4>> mkeys;
    for (int i : pairs) {
        const bm::Geometry &g = b->ctx[i]->g;
        const std::array<int, 4> k{g.NSP, g.level_px, g.jfMax, g.iterations};
        if (std::find(mkeys.begin(), mkeys.end(), k) == mkeys.end()) mkeys.push_back(k);
    }
    for (const auto &k : mkeys) {
        std::vector<bm::MatchJob> jobs;
        std::vector<int> tinit, tlev, members;
        for (int i : pairs) {
            bm_ctx *c = b->ctx[i];
            const bm::Geometry &g = c->g;
            if (g.NSP != k[0] || g.level_px != k[1] || g.jfMax != k[2] || g.iterations != k[3]) continue;
            bm::MatchJob J{};
            J.F1 = c->d_F[0];
            J.F2 = c->d_F[1];
            J.off[0] = c->d_off[0];
            J.off[1] = c->d_off[1];
            J.cost[0] = c->d_cost[0];
            J.cost[1] = c->d_cost[1];
            J.sx = g.sx;
            J.sy = g.sy;
            J.double_prop = g.doubleProp;
            J.strip_rows = g.strip_rows;
            J.pair = (uint32_t)i;
            J.key0 = (uint32_t)g.seed;
            J.key1 = (uint32_t)(g.seed >> 32);
            J.sigma_q16 = g.rs_sigma_q16;
            jobs.push_back(J);
            tinit.push_back(bm::match_tiles(g, 0));
            tlev.push_back(bm::match_tiles(g, g.level_px));
            members.push_back(i);
        }
        // the initialisation carries the first iteration's random search (the
        // levels of that iteration are the identity after a zero field)
        MultiOp init = group_op(pb, MultiOp::INIT, k[0], 0, jobs, tinit, 0);
        init.rs = k[3] > 0 ? 0 : -1;
        ops.push_back(init);
        // the same job array, with the level kernels' tiles
        const MultiOp lev = group_op(pb, MultiOp::LEVEL, k[0], k[1], jobs, tlev, 0, init.jobs);
        int src = 0, pd = -1;
        for (int it = 0; it < k[3]; ++it) {
            if (it > 0) {
                for (int p = k[2]; p > 0; p /= 2) {
                    MultiOp o = lev;
                    o.p = p;
                    o.pd = pd;
                    o.src = src;
                    o.rs = p == 1 ? it : -1;
                    ops.push_back(o);
                    src = 1 - src;
                }
            }
            pd = -pd;
        }
        for (int i : members) b->final_field[i] = src;
    }
    // down-sampling: tiled (LDS window) or direct
    for (int tiled = 1; tiled >= 0; --tiled) {
        std::vector<bm::DsJob> jobs;
        std::vector<int> tiles;
        size_t lds = 0;
        for (int i : pairs) {
            bm_ctx *c = b->ctx[i];
            const bm::Geometry &g = c->g;
            const int identity = (g.sx == g.W && g.sy == g.H) ? 1 : 0;
            const int t = (!identity && g.ds_lds > 0) ? 1 : 0;
            if (t != tiled) continue;
            float *u = c->p.filter ? c->d_U0 : U[i], *v = c->p.filter ? c->d_V0 : V[i];
            jobs.push_back({c->d_off[b->final_field[i]], g.nn_x, g.nn_y, g.gauss, g.P, g.sx, g.sy, g.W, g.H,
                            identity, (float)g.W / (float)g.sx, (float)g.H / (float)g.sy, u, v});
            tiles.push_back(bm::downsample_tiles(g, t != 0));
            lds = std::max(lds, (size_t)g.ds_lds);
        }
        if (!jobs.empty()) ops.push_back(group_op(pb, MultiOp::DOWNSAMPLE, tiled, 0, jobs, tiles, tiled ? lds : 0));
    }
    // medians (x then y) per window size, then the trilateral filter per kernel
    std::vector<int> fss, tgs;
    for (int i : pairs) {
        const bm_ctx *c = b->ctx[i];
        if (!c->p.filter) continue;
        if (std::find(fss.begin(), fss.end(), c->g.fsize) == fss.end()) fss.push_back(c->g.fsize);
        if (std::find(tgs.begin(), tgs.end(), c->g.tri_groups) == tgs.end()) tgs.push_back(c->g.tri_groups);
    }
    for (int dir = 1; dir >= 0; --dir) {
        for (int fs : fss) {
            std::vector<bm::MedJob> jobs;
            std::vector<int> tiles;
            for (int i : pairs) {
                bm_ctx *c = b->ctx[i];
                if (!c->p.filter || c->g.fsize != fs) continue;
                if (dir)
                    jobs.push_back({c->d_U0, c->d_V0, c->d_tU, c->d_tV, c->g.W, c->g.H});
                else
                    jobs.push_back({c->d_tU, c->d_tV, c->d_Um, c->d_Vm, c->g.W, c->g.H});
                tiles.push_back(2 * bm::median_tiles(c->g));
            }
            MultiOp o = group_op(pb, MultiOp::MEDIAN, fs, 0, jobs, tiles, 0);
            o.dir = dir;
            ops.push_back(o);
        }
    }
    for (int tg : tgs) {
        std::vector<bm::TriJob> jobs;
        std::vector<int> tiles;
        size_t lds = 0;
        for (int i : pairs) {
            bm_ctx *c = b->ctx[i];
            const bm::Geometry &g = c->g;
            if (!c->p.filter || g.tri_groups != tg) continue;
            int BR = 0;
            const size_t l = bm::trilateral_lds(g, &BR);
            if (BR < 1) return batch_fail(b, BM_E_PARAM, "pair " + std::to_string(i) + ": trilateral band too small");
            jobs.push_back({c->d_U0, c->d_V0, c->d_Um, c->d_Vm, c->d_g32[0], U[i], V[i], g.tri_spatial, g.W, g.H,
                            g.tri_s, BR, g.tri_a2, g.tri_b});
            tiles.push_back(bm::trilateral_tiles(g, tg));
            lds = std::max(lds, l);
        }
        ops.push_back(group_op(pb, MultiOp::TRILATERAL, tg, 0, jobs, tiles, lds));
    }
    return BM_OK;
}

// The plan of the multi-pair mode for these buffers: the eligible pairs split
// into `branches` lanes (largest first onto the least-loaded lane, or
// round-robin in index order), each lane's stages as multi-pair launches on its
// own branch, so the lanes' different stages overlap; the job arrays in one
// device buffer (re-made only when a buffer, a configuration or the schedule
// changes).
int build_plan(bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2, float *const *U, float *const *V)
{
    const int n = (int)b->ctx.size();
    const int L = std::max(1, std::min(b->sched.branches, (int)b->streams.size()));
    b->fallback.clear();
    b->final_field.assign(n, 0);
    b->lane_of.assign(n, 0);
    std::vector<std::vector<int>> members(L);
    std::vector<double> load(L, 0.0);
    int rr = 0;
    for (int i : pair_order(b)) {
        const int l = b->sched.order == 1 ? (int)(std::min_element(load.begin(), load.end()) - load.begin())
                                           : rr++ % L;
        load[l] += pair_work(b->ctx[i]);
        b->lane_of[i] = l;
        if (multi_eligible(b->ctx[i]))
            members[l].push_back(i);
        else
            b->fallback.push_back(i);
    }
    PlanBytes pb;
    std::vector<std::vector<MultiOp>> lanes(L);
    for (int l = 0; l < L; ++l) {
        const int r = plan_lane(b, pb, members[l], r1, r2, U, V, lanes[l]);
        if (r != BM_OK) return r;
    }
    // the job arrays on the device (the previous plan's buffer may still be read)
    hipError_t e = hipStreamSynchronize(b->origin);
    if (e == hipSuccess && b->plan_dev) {
        e = hipFree(b->plan_dev);
        b->plan_dev = nullptr;
    }
    if (e == hipSuccess) e = hipMalloc(&b->plan_dev, pb.buf.size());
    if (e == hipSuccess) e = hipMemcpy(b->plan_dev, pb.buf.data(), pb.buf.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        b->plan_dev = nullptr;
        (void)hipGetLastError();
        return batch_fail(b, BM_E_ALLOC, std::string("batch plan: ") + hipGetErrorString(e));
    }
    b->lanes.swap(lanes);
    return BM_OK;
}

int multi_grid(const bm_batch *b, int tiles)
{
    if (b->sched.persist <= 0) return tiles;
    const int g = b->sched.persist * b->cus;
    return g >= tiles ? tiles : std::max(8, g / 8 * 8);
}

// The plan's launches on the origin stream, and the per-pair path of the
// pairs the plan does not cover.
int run_plan(bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2, float *const *U, float *const *V)
{
    const unsigned char *base = static_cast<const unsigned char *>(b->plan_dev);
    const bool branched = b->lanes.size() > 1;
    for (size_t l = 0; l < b->lanes.size(); ++l) {
    const hipStream_t st = branched ? b->streams[l] : b->origin;
    for (const MultiOp &o : b->lanes[l]) {
        const int *start = reinterpret_cast<const int *>(base + o.start);
        const void *jobs = base + o.jobs;
        const int grid = multi_grid(b, o.tiles);
        hipError_t e = hipSuccess;
        switch (o.kind) {
        case MultiOp::GRAY:
            e = bm::launch_multi_gray(static_cast<const bm::GrayJob *>(jobs), start, o.njobs, o.tiles, grid, st);
            break;
        case MultiOp::BRIEF:
            e = bm::launch_multi_brief(o.a, o.b, static_cast<const bm::BriefJob *>(jobs), start, o.njobs, o.tiles,
                                       grid, o.lds, st);
            break;
        case MultiOp::INIT:
            e = bm::launch_multi_init(o.a, static_cast<const bm::MatchJob *>(jobs), start, o.njobs, o.tiles, grid,
                                      o.rs, st);
            break;
        case MultiOp::LEVEL:
            e = bm::launch_multi_level(o.a, o.b, static_cast<const bm::MatchJob *>(jobs), start, o.njobs, o.tiles,
                                       grid, o.p, o.pd, o.src, o.rs, st);
            break;
        case MultiOp::DOWNSAMPLE:
            e = bm::launch_multi_downsample(o.a != 0, static_cast<const bm::DsJob *>(jobs), start, o.njobs, o.tiles,
                                            grid, o.lds, st);
            break;
        case MultiOp::MEDIAN:
            e = bm::launch_multi_median(o.a, o.dir, static_cast<const bm::MedJob *>(jobs), start, o.njobs, o.tiles,
                                        grid, st);
            break;
        case MultiOp::TRILATERAL:
            e = bm::launch_multi_trilateral(o.a, static_cast<const bm::TriJob *>(jobs), start, o.njobs, o.tiles,
                                            grid, o.lds, st);
            break;
        }
        if (e != hipSuccess) return batch_fail(b, BM_E_LAUNCH, std::string("multi-pair launch: ") + hipGetErrorString(e));
    }
    }
    for (int i : b->fallback) {
        bm_ctx *c = b->ctx[i];
        hipStream_t own = c->stream;
        const bool prof = c->prof;
        c->stream = branched ? b->streams[b->lane_of[i]] : b->origin;
        c->prof = false;
        const int r = bm_flow_pair(c, r1[i], r2[i], (uint32_t)i, U[i], V[i]);
        c->stream = own;
        c->prof = prof;
        if (r != BM_OK) return batch_fail(b, r, "pair " + std::to_string(i) + ": " + c->err);
    }
    return BM_OK;
}

// Captures `enqueue` on the origin stream (forked to the branch streams when
// `branches`) into a graph, instantiated into *exec.
template <class F>
int capture(bm_batch *b, bool branches, hipGraph_t *graph, hipGraphExec_t *exec, F enqueue)
{
    if (*exec) (void)hipGraphExecDestroy(*exec);
    if (*graph) (void)hipGraphDestroy(*graph);
    *exec = nullptr;
    *graph = nullptr;
    hipError_t e = hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal);
    if (branches) {
        if (e == hipSuccess) e = hipEventRecord(b->fork, b->origin);
        for (hipStream_t s : b->streams)
            if (e == hipSuccess) e = hipStreamWaitEvent(s, b->fork, 0);
    }
    int r = BM_OK;
    if (e == hipSuccess) r = enqueue();
    if (branches) {
        for (size_t k = 0; k < b->streams.size() && e == hipSuccess; ++k) {
            e = hipEventRecord(b->joins[k], b->streams[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->joins[k], 0);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->origin, &g);
    if (e == hipSuccess) e = e2;
    if (r != BM_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return r != BM_OK ? r : batch_fail(b, BM_E_LAUNCH, std::string("graph capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        *exec = nullptr;
        return batch_fail(b, BM_E_LAUNCH, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    *graph = g;
    return BM_OK;
}

int schedule_valid(const bm_batch *b, const bm_batch_schedule &s, std::string &why)
{
    if (s.mode != BM_BATCH_PER_PAIR && s.mode != BM_BATCH_GRAPH && s.mode != BM_BATCH_MULTI)
        return why = "schedule mode must be BM_BATCH_PER_PAIR, BM_BATCH_GRAPH or BM_BATCH_MULTI", 0;
    if (s.branches < 1 || s.branches > (int)b->streams.size())
        return why = "branches outside [1, streams of the batch]", 0;
    if (s.order != 0 && s.order != 1) return why = "order must be 0 or 1", 0;
    if (s.persist < 0 || s.persist > 64) return why = "persist outside [0, 64]", 0;
    return 1;
}

}  // namespace

extern "C" {

int bm_batch_flow(bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2, float *const *U, float *const *V,
                  int mode)
{
    if (!b || !r1 || !r2 || !U || !V) return fail(nullptr, BM_E_PARAM, "null argument");
    const int n = (int)b->ctx.size();
    for (int i = 0; i < n; ++i)
        if (!r1[i] || !r2[i] || !U[i] || !V[i]) return batch_fail(b, BM_E_PARAM, "null pointer in batch");
    if (mode == BM_BATCH_AUTO) mode = b->sched.mode;
    if (mode != BM_BATCH_PER_PAIR && mode != BM_BATCH_GRAPH && mode != BM_BATCH_MULTI)
        return batch_fail(b, BM_E_PARAM, "unknown batch mode");
    for (bm_ctx *c : b->ctx)  // work already queued on a pair's own stream comes first
        if (hipStreamSynchronize(c->stream) != hipSuccess) return batch_fail(b, BM_E_LAUNCH, "stream sync failed");
    if (mode == BM_BATCH_PER_PAIR) return batch_enqueue(b, r1, r2, U, V, true);
    const std::vector<uintptr_t> key = batch_key(b, r1, r2, U, V);
    // after a replay every context's field index is the one its captured
    // launches end in, and no sequence continues from it
    auto settle = [&](const std::vector<int> &field) {
        for (int i = 0; i < n; ++i) {
            b->ctx[i]->field = field[i];
            b->ctx[i]->have_prev = 0;
            b->ctx[i]->field_valid = 0;
        }
    };
    if (mode == BM_BATCH_GRAPH) {
        if (!b->exec || key != b->key) {
            b->key.clear();
            const int r = capture(b, true, &b->graph, &b->exec, [&] { return batch_enqueue(b, r1, r2, U, V, false); });
            if (r != BM_OK) return r;
            b->gfield.resize(n);
            for (int i = 0; i < n; ++i) b->gfield[i] = b->ctx[i]->field;
            b->key = key;
        }
        const hipError_t e = hipGraphLaunch(b->exec, b->origin);
        if (e != hipSuccess) return batch_fail(b, BM_E_LAUNCH, std::string("graph launch: ") + hipGetErrorString(e));
        settle(b->gfield);
        return BM_OK;
    }
    // multi-pair launches, captured once per key and replayed
    if (!b->mexec || key != b->mkey) {
        b->mkey.clear();
        int r = BM_OK;
        if (b->pkey != key) {
            b->pkey.clear();
            r = build_plan(b, r1, r2, U, V);
            if (r != BM_OK) return r;
            b->pkey = key;
        }
        r = capture(b, b->lanes.size() > 1, &b->mgraph, &b->mexec, [&] { return run_plan(b, r1, r2, U, V); });
        if (r != BM_OK) return r;
        // the plan's pairs end in final_field; the per-pair path of the
        // others set theirs while it was captured
        b->mfield.resize(n);
        for (int i = 0; i < n; ++i)
            b->mfield[i] = std::find(b->fallback.begin(), b->fallback.end(), i) != b->fallback.end()
                               ? b->ctx[i]->field
                               : b->final_field[i];
        b->mkey = key;
    }
    const hipError_t e = hipGraphLaunch(b->mexec, b->origin);
    if (e != hipSuccess) return batch_fail(b, BM_E_LAUNCH, std::string("graph launch: ") + hipGetErrorString(e));
    settle(b->mfield);
    return BM_OK;
}

int bm_batch_get_schedule(const bm_batch *b, bm_batch_schedule *out)
{
    if (!b || !out) return fail(nullptr, BM_E_PARAM, "null argument");
    *out = b->sched;
    return BM_OK;
}

int bm_batch_set_schedule(bm_batch *b, const bm_batch_schedule *s)
{
    if (!b || !s) return fail(nullptr, BM_E_PARAM, "null argument");
    std::string why;
    if (!schedule_valid(b, *s, why)) return batch_fail(b, BM_E_PARAM, why);
    b->sched = *s;
    return BM_OK;
}

int bm_batch_autotune(bm_batch *b, const uint8_t *const *r1, const uint8_t *const *r2, float *const *U,
                      float *const *V, bm_batch_schedule *out)
{
    if (!b || !r1 || !r2 || !U || !V) return fail(nullptr, BM_E_PARAM, "null argument");
    const bm_batch_schedule keep = b->sched;
    std::vector<bm_batch_schedule> cands;
    const int S = (int)b->streams.size();
    std::vector<int> branch_counts{S};
    if (S / 2 >= 1 && S / 2 != S) branch_counts.push_back(S / 2);
    for (int order = 0; order <= 1; ++order) {
        for (int br : branch_counts) cands.push_back({BM_BATCH_GRAPH, br, order, 0});
        for (int lanes : {1, 2, 4})
            if (lanes <= S)
                for (int persist : {0, 4}) cands.push_back({BM_BATCH_MULTI, lanes, order, persist});
    }
    float best_ms = -1.0f;
    bm_batch_schedule best = keep;
    for (const bm_batch_schedule &c : cands) {
        b->sched = c;
        float ms = 0.0f;
        int r = BM_OK;
        for (int rep = 0; rep < 3 && r == BM_OK; ++rep) {  // capture, then two timed replays
            if (rep == 2 && (r = bm_batch_timer_start(b)) != BM_OK) break;
            r = bm_batch_flow(b, r1, r2, U, V, c.mode);
            if (r == BM_OK && rep < 2) r = bm_batch_sync(b);
        }
        if (r == BM_OK) r = bm_batch_timer_stop(b, &ms);
        if (r != BM_OK) {
            b->sched = keep;
            return r;
        }
        if (best_ms < 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    b->sched = best;
    if (out) *out = best;
    return BM_OK;
}

int bm_batch_sync(bm_batch *b)
{
    if (!b) return fail(nullptr, BM_E_PARAM, "null batch");
    const hipError_t e = hipStreamSynchronize(b->origin);
    if (e != hipSuccess) return batch_fail(b, BM_E_LAUNCH, std::string("sync: ") + hipGetErrorString(e));
    return BM_OK;
}

int bm_batch_timer_start(bm_batch *b)
{
    if (!b) return fail(nullptr, BM_E_PARAM, "null batch");
    if (hipEventRecord(b->ev0, b->origin) != hipSuccess) return batch_fail(b, BM_E_LAUNCH, "event record failed");
    return BM_OK;
}

int bm_batch_timer_stop(bm_batch *b, float *ms)
{
    if (!b || !ms) return fail(nullptr, BM_E_PARAM, "null argument");
    if (hipEventRecord(b->ev1, b->origin) != hipSuccess || hipEventSynchronize(b->ev1) != hipSuccess ||
        hipEventElapsedTime(ms, b->ev0, b->ev1) != hipSuccess)
        return batch_fail(b, BM_E_LAUNCH, "batch timer failed");
    return BM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// bm_flow_pair_static: the pair path without device allocation.  Its tables
// and buffers are carved out of one file-scope device array; each GPU holds
// its own copy of the array (from the first call on that GPU until the process
// ends), found with hipGetSymbolAddress once per device.  In host code the
// array's name is only a stand-in for that address: every use goes through the
// per-device pointer.

static __device__ __align__(256) unsigned char scratch[BM_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kStaticMaxDevices = 64;

struct StaticDevice {
    unsigned char *base = nullptr;  // device address of `scratch` on this GPU, null until set up
    bm_ctx ctx;                     // geometry and buffer pointers carved from base
    HostTables tabs;                // host tables of the configuration on the device
    bm_params key_p{};              // that configuration
    uint64_t key_seed = 0;
    bool have_key = false;
};

std::mutex g_static_mu;
StaticDevice g_static[kStaticMaxDevices];
// the last configuration whose scratch size was checked
bm_params g_checked_p{};
uint64_t g_checked_seed = 0;
bool g_checked = false;

template <class T>
void carve_one(T **ptr, size_t count, unsigned char *base, size_t &at)
{
    at = (at + 255) & ~(size_t)255;
    *ptr = base ? reinterpret_cast<T *>(base + at) : nullptr;
    at += std::max<size_t>(count * sizeof(T), 1);
}

// Lays the context's tables and buffers out from `base` (each 256-byte
// aligned); with base == nullptr it only measures.  Returns the bytes used.
size_t carve_static(bm_ctx *ctx, const HostTables &t, unsigned char *base)
{
    size_t at = 0;
    const bm::Geometry &g = ctx->g;
    const size_t Win = (size_t)g.W * g.H, M = (size_t)g.sx * g.sy, NSP = (size_t)g.NSP;
    carve_one(&ctx->d_pattern, ctx->pattern.size(), base, at);
    carve_one(&ctx->d_plan, NSP * 64 * 4, base, at);
    carve_one(&ctx->d_plan_count, NSP * 2, base, at);
    carve_one(&ctx->d_up_x, t.upx.size(), base, at);
    carve_one(&ctx->d_up_y, t.upy.size(), base, at);
    carve_one(&ctx->d_nn_x, t.nnx.size(), base, at);
    carve_one(&ctx->d_nn_y, t.nny.size(), base, at);
    carve_one(&ctx->d_gauss, ctx->gauss.size(), base, at);
    carve_one(&ctx->d_spt, t.spt.size(), base, at);
    for (int k = 0; k < 2; ++k) {
        carve_one(&ctx->d_g8[k], Win, base, at);
        carve_one(&ctx->d_F[k], M * NSP, base, at);
        carve_one(&ctx->d_off[k], M, base, at);
        carve_one(&ctx->d_cost[k], M, base, at);
    }
    carve_one(&ctx->d_g32[0], Win, base, at);  // the guide; frame 2 is never needed in FP32
    ctx->d_g32[1] = nullptr;
    for (float **b : {&ctx->d_U0, &ctx->d_V0, &ctx->d_tU, &ctx->d_tV, &ctx->d_Um, &ctx->d_Vm})
        carve_one(b, Win, base, at);
    return at;
}

bool same_config(const bm_params &a, uint64_t sa, const bm_params &b, uint64_t sb)
{
    return sa == sb && std::memcmp(&a, &b, sizeof a) == 0;
}

}  // namespace

extern "C" int bm_flow_pair_static(const bm_params *params, uint64_t seed, const uint8_t *rgb1, const uint8_t *rgb2,
                                   uint32_t pair, float *U, float *V, uint32_t *off_out, uint16_t *cost_out,
                                   void *stream)
{
    if (!params || !rgb1 || !rgb2 || !U || !V) return fail(nullptr, BM_E_PARAM, "null argument");
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_checked || !same_config(g_checked_p, g_checked_seed, *params, seed)) {
        // limits and scratch size on the host, before any device work
        g_checked = false;
        int r = validate_params(*params);
        if (r) return r;
        bm_ctx probe;
        HostTables tabs;
        r = prepare_host(&probe, *params, seed, tabs);
        if (r) return r;
        const size_t need = carve_static(&probe, tabs, nullptr);
        if (need > (size_t)BM_STATIC_SCRATCH_BYTES) {
            char m[320];
            std::snprintf(m, sizeof m,
                          "bm_flow_pair_static: %ux%u -> %ux%u with %u-bit descriptors needs %zu bytes of scratch, "
                          "more than the static maximum of %u bytes (BM_STATIC_SCRATCH_BYTES)",
                          params->sx_in, params->sy_in, params->sx, params->sy, params->N, need,
                          (unsigned)BM_STATIC_SCRATCH_BYTES);
            return fail(nullptr, BM_E_PARAM, m);
        }
        g_checked_p = *params;
        g_checked_seed = seed;
        g_checked = true;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, BM_E_LAUNCH, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kStaticMaxDevices) return fail(nullptr, BM_E_PARAM, "device index out of range");
    StaticDevice &sd = g_static[dev];
    if (!sd.base) {
        // once per device, with that device current: the address of its copy
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, BM_E_LAUNCH, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        sd.base = static_cast<unsigned char *>(ptr);
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    bm_ctx &ctx = sd.ctx;
    if (!sd.have_key || !same_config(sd.key_p, sd.key_seed, *params, seed)) {
        // a new configuration on this GPU: geometry and tables into its array,
        // copied from host tables that stay alive until the copies are done
        sd.have_key = false;
        ctx = bm_ctx();
        int r = prepare_host(&ctx, *params, seed, sd.tabs);
        if (r) return r;
        carve_static(&ctx, sd.tabs, sd.base);
        bind_tables(&ctx);
        e = hipSuccess;
        auto put = [&](void *dst, const void *src, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        };
        put(ctx.d_pattern, ctx.pattern.data(), ctx.pattern.size() * sizeof(int16_t));
        put(ctx.d_plan, ctx.plan.data(), ctx.plan.size() * sizeof(uint32_t));
        put(ctx.d_plan_count, ctx.plan_count.data(), ctx.plan_count.size() * sizeof(int));
        put(ctx.d_up_x, sd.tabs.upx.data(), sd.tabs.upx.size() * sizeof(uint32_t));
        put(ctx.d_up_y, sd.tabs.upy.data(), sd.tabs.upy.size() * sizeof(uint32_t));
        put(ctx.d_nn_x, sd.tabs.nnx.data(), sd.tabs.nnx.size() * sizeof(uint32_t));
        put(ctx.d_nn_y, sd.tabs.nny.data(), sd.tabs.nny.size() * sizeof(uint32_t));
        put(ctx.d_gauss, ctx.gauss.data(), ctx.gauss.size() * sizeof(float));
        put(ctx.d_spt, sd.tabs.spt.data(), sd.tabs.spt.size() * sizeof(float));
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(nullptr, BM_E_LAUNCH, std::string("table upload: ") + hipGetErrorString(e));
        sd.key_p = *params;
        sd.key_seed = seed;
        sd.have_key = true;
    }
    ctx.stream = st;
    ctx.field = 0;
    ctx.err.clear();
    int r = run_pair(&ctx, rgb1, rgb2, pair, U, V);
    const size_t M = (size_t)ctx.g.sx * ctx.g.sy;
    if (!r && off_out) {
        e = hipMemcpyAsync(off_out, ctx.d_off[ctx.field], M * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) r = hip_fail(&ctx, e, "field copy");
    }
    if (!r && cost_out) {
        e = hipMemcpyAsync(cost_out, ctx.d_cost[ctx.field], M * sizeof(uint16_t), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) r = hip_fail(&ctx, e, "cost copy");
    }
    if (r) return fail(nullptr, r, ctx.err);
    return BM_OK;
}
