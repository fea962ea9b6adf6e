/*
 * briefmatch_amd.h -- C ABI of the MI355X (gfx950) BriefMatch flow engine.
 *
 * The drop-in boundary for the reference's extern "C" launch layer
 * (kernels/briefmatch_cuda.h:64-93, kernels/util_cuda.h:114-129).  Differences
 * from that layer, all deliberate:
 *   - every entry point returns a status (BM_OK = 0) instead of throwing a C++
 *     exception across the C boundary (util_cuda.h:102-112); bm_last_error()
 *     gives the message;
 *   - sizes and configuration live in an opaque bm_ctx (one per BMParams, like
 *     the reference's BriefMatch object) instead of module-global __constant__
 *     tables, so several configurations can coexist in one process;
 *   - sizes are size_t; device work is asynchronous on the context's HIP stream
 *     and bm_sync() waits for it (the reference synchronises after every launch);
 *   - the match field is (int16 dx, int16 dy) packed in one uint32 per pixel
 *     plus a uint16 Hamming cost, instead of float D / sint NX / sint NY.
 * All pointers named *_dev are device pointers; *_host are host pointers.
 */
#ifndef BRIEFMATCH_AMD_H
#define BRIEFMATCH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define BM_OK 0
#define BM_E_PARAM 1   /* parameter outside the supported range */
#define BM_E_ALLOC 2   /* device or host allocation failed */
#define BM_E_LAUNCH 3  /* HIP runtime / kernel launch error */
#define BM_E_IO 4      /* file read / write error */
#define BM_E_STATE 5   /* call out of order (e.g. no previous frame) */

/* Copy directions for bm_memcpy (util_cuda.h:74-77). */
#define BM_HOST_TO_DEVICE 0
#define BM_DEVICE_TO_HOST 1
#define BM_DEVICE_TO_DEVICE 2
#define BM_HOST_TO_HOST 3

/* Limits of the reference: PATCH_MAX/4 feature bits (briefmatch_cuda.h:62),
 * MAX_MEDIAN (util_cuda.h:56), CONST_KERNEL_SIZE (util_cuda.h:58). */
#define BM_MAX_FEATURE_BITS 512
#define BM_MAX_MEDIAN 47
#define BM_MAX_GAUSS_TAPS 100

/* Matching parameters: the fields and meaning of BMParams (briefmatch.h:65-89). */
typedef struct bm_params {
    unsigned sx, sy;          /* matched (up-sampled) resolution */
    unsigned sx_in, sy_in;    /* input resolution */
    unsigned N;               /* BRIEF length in bits, 2..512 */
    unsigned iterations;      /* propagate + random-search rounds */
    unsigned jfMax;           /* largest jump-flooding step */
    unsigned P;               /* Gaussian taps (derived from sigma) */
    unsigned fsize;           /* median window, odd, <= 47 */
    float patchArea;          /* BRIEF patch radius in matched pixels */
    float sigma;              /* Gaussian down-sampling sigma (matched pixels) */
    float sEPE, sI, sS;       /* trilateral sigmas: flow, image (I/255), spatial */
    float maxmot;             /* colour-coding max motion */
    int doubleProp;           /* 4 propagation candidates (else 2) */
    int filter;               /* median + trilateral refinement */
    int tInit;                /* temporal initialisation in sequences */
} bm_params;

typedef struct bm_ctx bm_ctx;

/* ---- parameters ------------------------------------------------------- */

/* BMParams::BMParams() defaults (briefmatch.h:67-84). */
void bm_params_default(bm_params *out);

/* The CLI derivation of OFPipeline::setParams (of_pipeline.cpp:178-190) and the
 * Gaussian tap count of BriefMatch::setup (briefmatch.cpp:153-154): up-sampling
 * factors ux/uy (0 selects the reference defaults 2.740/3.052), patch radius,
 * down-sampling sigma, median size and spatial sigma in percent of the diagonal. */
int bm_params_derive(unsigned W, unsigned H, float ux, float uy, float patch_radius,
                     float sigma_ds, float median_size, float range_spatial, bm_params *out);

/* ---- context (BriefMatch::setup / clean, briefmatch.cpp:72-238) -------- */

/* Validates the limits of briefmatch.cpp:78-96, builds the BRIEF pattern from
 * `seed` (replacing the time-seeded Boost pattern, briefmatch.cpp:124-148), the
 * Gaussian kernel (:153-172) and the device buffers, on a new HIP stream. */
int bm_ctx_create(const bm_params *params, uint64_t seed, bm_ctx **out);
void bm_ctx_destroy(bm_ctx *ctx);
int bm_ctx_get_params(const bm_ctx *ctx, bm_params *out);
/* The context's hipStream_t, as void*. */
void *bm_ctx_stream(bm_ctx *ctx);
/* Last error message of this context (ctx == NULL: of the calling thread). */
const char *bm_last_error(const bm_ctx *ctx);
int bm_sync(bm_ctx *ctx);
/* Host Philox4x32-10 (the counter-based stream of the BRIEF pattern and of the
 * random search; replaces cuRAND, briefmatch_cuda.cu:57,237-245). */
int bm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Descriptor stride in uint32 words per pixel: ceil(N/32) rounded up to a power
 * of two (words beyond ceil(N/32) are zero). */
unsigned bm_descriptor_stride(const bm_ctx *ctx);
/* The BRIEF pattern: 4N int16 (ox1, oy1, ox2, oy2) per test. */
int bm_get_pattern(const bm_ctx *ctx, int16_t *pattern_host);
/* Replace the pattern (CUDA_copyPatchToConst, briefmatch_cuda.cu:376-379). */
int bm_set_pattern(bm_ctx *ctx, const int16_t *pattern_host);
/* The normalised Gaussian kernel (CUDA_copyToConst, util_cuda.cu:277-280). */
int bm_get_gauss_kernel(const bm_ctx *ctx, float *kernel_host, unsigned *P);
/* Device pointers of the context's last match field (sx*sy entries). */
int bm_ctx_field(bm_ctx *ctx, uint32_t **off_dev, uint16_t **cost_dev);

/* ---- device memory (CUDA_constructor / destructor / copy, util_cuda.cu:206-239) */

int bm_malloc(void **ptr_dev, size_t bytes);
/* Waits for all queued device work before releasing the buffer. */
int bm_free(void *ptr_dev);
int bm_memcpy(void *dst, const void *src, size_t bytes, int kind);
int bm_memset(void *ptr_dev, int value, size_t bytes);

/* ---- stages (asynchronous on the context stream) ------------------------ */

/* Interleaved RGB uint8 (W*H*3) -> gray uint8 and/or FP32 (either may be NULL). */
int bm_gray(bm_ctx *ctx, const uint8_t *rgb_dev, uint8_t *gray8_dev, float *gray32_dev);
/* Up-sampling sx_in x sy_in gray8 -> sx x sy FP32, bilinear in fixed point:
 * the default up-sampler of the whole-path calls (no counterpart in the
 * reference, whose CUDA_resizeLanczos, util_cuda.h:124, maps to
 * bm_upsample_lanczos below). */
int bm_upsample(bm_ctx *ctx, const uint8_t *gray8_dev, float *L_dev);
/* Lanczos-3 up-sampling sx_in x sy_in gray8 -> sx x sy FP32: CUDA_resizeLanczos
 * (util_cuda.h:124, util_cuda.cu:250-257), the reference's resizeLanczos
 * (util_cuda.cu:126-158) with its 6 x 6 weights factored into
 * per-column and per-row tables (normalised, FP32) and fused multiply-adds. */
int bm_upsample_lanczos(bm_ctx *ctx, const uint8_t *gray8_dev, float *L_dev);
/* Up-sampler of the whole-path calls (bm_flow_pair, bm_stream_frame,
 * bm_flow_pair_learned, batches): bilinear fused into BRIEF (default) or
 * Lanczos-3 into a context frame followed by BRIEF on it. */
#define BM_UPSAMPLE_BILINEAR 0
#define BM_UPSAMPLE_LANCZOS3 1
int bm_set_upsampler(bm_ctx *ctx, int mode);
int bm_get_upsampler(const bm_ctx *ctx, int *mode);
/* BRIEF descriptors of an up-sampled FP32 frame (CUDA_binaryFeatures,
 * briefmatch_cuda.cu:311-319).  F has sx*sy*bm_descriptor_stride() words. */
int bm_brief(bm_ctx *ctx, const float *L_dev, uint32_t *F_dev);
/* Up-sampling fused into BRIEF: the same descriptors as bm_upsample followed by
 * bm_brief, without writing the up-sampled frame. */
int bm_brief_gray(bm_ctx *ctx, const uint8_t *gray8_dev, uint32_t *F_dev);
/* Field initialisation (CUDA_initialize, briefmatch_cuda.cu:321-332): first != 0
 * gives offset (0,0); otherwise temporal, following prev_off (a different buffer
 * from off) once. */
int bm_init(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev, int first,
            const uint32_t *prev_off_dev, uint32_t *off_dev, uint16_t *cost_dev);
/* One level of jump flooding from a snapshot (off_in/cost_in) into a different
 * buffer (off_out/cost_out).  cost_in[i] must be the Hamming cost of off_in[i],
 * as in every field this library produces (and in the reference's D/NX/NY):
 * candidates equal to a known offset reuse its cost instead of a new gather. */
int bm_propagate_level(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev,
                       const uint32_t *off_in_dev, const uint16_t *cost_in_dev,
                       uint32_t *off_out_dev, uint16_t *cost_out_dev, int p, int pd);
/* All levels jfMax..1 of one propagation (CUDA_propagate, briefmatch_cuda.cu:334-347),
 * updating off/cost; pd = -1 for even iterations, +1 for odd (briefmatch.cpp:270,289). */
int bm_propagate(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev,
                 uint32_t *off_dev, uint16_t *cost_dev, int pd);
/* Random search (CUDA_randomSearch, briefmatch_cuda.cu:349-360) with the stateless
 * Philox stream of (pixel, iteration, pair, seed). */
int bm_random_search(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev,
                     uint32_t *off_dev, uint16_t *cost_dev, uint32_t iteration, uint32_t pair);
/* BriefMatch::match after the descriptors (briefmatch.cpp:265-290): init, then
 * `iterations` x {propagate, random search}.  Random search runs fused into the
 * last propagation level; results equal the separate calls. */
int bm_match(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev, int first,
             const uint32_t *prev_off_dev, uint32_t pair, uint32_t *off_dev, uint16_t *cost_dev);
/* BriefMatch::downsample (briefmatch.cpp:296-324): field -> FP32 flow at input
 * resolution (Gaussian, nearest-neighbour decimation, scaling), fused. */
int bm_downsample(bm_ctx *ctx, const uint32_t *off_dev, float *U_dev, float *V_dev);
/* One 1-D median pass (CUDA_medianFilter1D, util_cuda.cu:197-204), dir 1 = x. */
int bm_median1d(bm_ctx *ctx, const float *in_dev, float *out_dev, int dir);
/* Cross-trilateral filter (CUDA_cbFilter, briefmatch_cuda.cu:362-374); sigmas are
 * taken from the context (sEPE^2, sI^2, sS^2 as BriefMatch::process passes them). */
int bm_trilateral(bm_ctx *ctx, const float *Ui_dev, const float *Vi_dev, const float *Um_dev,
                  const float *Vm_dev, const float *I_dev, float *Uo_dev, float *Vo_dev);
/* BriefMatch::process (briefmatch.cpp:327-352): medians x then y of U and V, then
 * the trilateral filter guided by I (the first frame at input resolution, FP32). */
int bm_refine(bm_ctx *ctx, const float *U_dev, const float *V_dev, const float *I_dev,
              float *Uo_dev, float *Vo_dev);

/* ---- whole path ------------------------------------------------------------ */

/* One isolated pair: device-resident interleaved RGB uint8 frames (sx_in x sy_in)
 * -> FP32 flow U, V (frame 1 -> frame 2, on the frame-1 grid).  `pair` selects the
 * random-search stream.  The match field stays in the context (bm_ctx_field). */
int bm_flow_pair(bm_ctx *ctx, const uint8_t *rgb1_dev, const uint8_t *rgb2_dev, uint32_t pair,
                 float *U_dev, float *V_dev);
/* Sequence mode (OFPipeline::processFrame, of_pipeline.cpp:241-387): describe one
 * new frame, match it against the previous one with temporal initialisation,
 * refine, and keep its descriptors for the next call.  *has_flow is 0 for the
 * first frame of a sequence (no previous frame). */
int bm_stream_frame(bm_ctx *ctx, const uint8_t *rgb_dev, float *U_dev, float *V_dev, int *has_flow);
/* Forget the previous frame (the next bm_stream_frame starts a new sequence). */
int bm_stream_reset(bm_ctx *ctx);
/* Pipelined sequence mode (off by default; no counterpart in the reference,
 * whose OFPipeline runs each frame to completion).  When on, bm_stream_frame
 * runs the medians and the trilateral filter of its frame on a second stream of
 * the context, so that the next call's descriptors and matching run beside
 * them; results are identical.  U_dev, V_dev of a frame are complete after
 * bm_sync(ctx) or bm_timer_stop(ctx), not in order with the stream returned by
 * bm_ctx_stream; calls on the context that use its refinement buffers wait for
 * them on the device. */
int bm_set_pipelining(bm_ctx *ctx, int on);

/* ---- the pair path without device allocation ----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  bm_flow_pair for one isolated pair without a context:
 * every table and intermediate buffer (pattern, BRIEF plan, coordinate tables,
 * gray frames, descriptors, both match fields, refinement planes) is carved
 * from one static device array of BM_STATIC_SCRATCH_BYTES that each GPU holds
 * from the first call on it until the process ends; the call allocates no
 * device memory.  Configurations whose buffers need more are rejected with
 * BM_E_PARAM (640x480 at x2 with 128-bit descriptors fits; 1920x1080 at x2.9
 * does not).  Bilinear up-sampling and default launch shapes; results are
 * those of bm_flow_pair on a context created with the same parameters and
 * seed.  Work is queued on `stream` (a hipStream_t, NULL = the null stream);
 * off_dev / cost_dev (sx*sy entries each, either may be NULL) receive the
 * match field.  A call with new parameters or seed uploads their tables and
 * waits for `stream`.  All calls on one GPU share its array, which keeps the
 * last call's contents (every buffer is written in full before it is read,
 * so none needs clearing): calls that overlap on one GPU, on different streams
 * or from several threads, race; consecutive calls on one stream are ordered. */
#define BM_STATIC_SCRATCH_BYTES (64u * 1024u * 1024u)
int bm_flow_pair_static(const bm_params *params, uint64_t seed, const uint8_t *rgb1_dev, const uint8_t *rgb2_dev,
                        uint32_t pair, float *U_dev, float *V_dev, uint32_t *off_dev, uint16_t *cost_dev,
                        void *stream);

/* ---- tuning ------------------------------------------------------------------ */

/* Launch-shape choices of one context.  Every valid tuning computes identical
 * descriptors and match fields; the flow is identical too except that the
 * trilateral filter's kernel choices (tri_groups, tri_split) add its weighted
 * sums in different orders (agreeing to far within the 1e-3 px tolerance).
 * Only the speed changes. */
typedef struct bm_tuning {
    int level_px;     /* pixels per thread of the jump-flooding kernels: 1, 2 (N <= 256) or 4 (N <= 128);
                         above 256 bits they run two threads per pixel whatever the value */
    int strip_rows;   /* rows per strip of the matching kernels' work-group order: 16..4096 */
    int brief_near;   /* LDS near halo of the fused BRIEF: 8 .. 48 in steps of 8 (40, 48 need > 64 KiB of LDS
                         per work-group, as on gfx950); 0 = kernel with global far taps */
    int tri_band;     /* source rows per LDS band of the trilateral filter: 1..(LDS limit, <= 16) */
    int tri_groups;   /* trilateral kernel: 0 = weighted sums on the VALU, 4 output rows per thread;
                         2 or 4 = weighted sums on MFMA (4x4x1 f32), that many 4-row groups per wave */
    int brief_frame;  /* bilinear up-sampling of the whole-path calls: 0 = fused into the BRIEF kernel
                         (interpolated from the gray frame in LDS); 1 = written to a context frame first,
                         the BRIEF kernel then copies its tile from that frame and reads its far tests
                         from it (needs brief_near > 0); a new context takes 1 when its matched frame
                         has at least 2^21 pixels, with brief_near 48 at 512 bits where the plan fits */
    int tri_split;    /* work-groups per trilateral tile: 1, 2, 4 or 8; above 1 each sums a share of the
                         tile's source rows into a context buffer (3 tri_split W H floats) and a
                         finishing kernel adds the shares, which fills the device on small frames */
} bm_tuning;

int bm_get_tuning(const bm_ctx *ctx, bm_tuning *out);
int bm_set_tuning(bm_ctx *ctx, const bm_tuning *t);
/* Times the pair (device RGB frames) under candidate tunings, one knob at a
 * time, keeps the fastest and returns it in *out (may be NULL).  With a
 * cache_path, a tuning stored for the same device and geometry is reused and a
 * new one is appended. */
int bm_autotune(bm_ctx *ctx, const uint8_t *rgb1_dev, const uint8_t *rgb2_dev, const char *cache_path,
                bm_tuning *out);

/* ---- batches of uneven pairs ------------------------------------------------ */

/* n pairs, each with its own parameters (resolution, up-sampling, feature
 * length may all differ) and context; `streams` parallel branches for the graph
 * mode.  Pair i uses random-search stream index i, so a batched result equals
 * bm_flow_pair(ctx_i, ..., pair = i). */
typedef struct bm_batch bm_batch;
int bm_batch_create(int n, const bm_params *params, uint64_t seed, int streams, bm_batch **out);
void bm_batch_destroy(bm_batch *b);
int bm_batch_ctx(bm_batch *b, int i, bm_ctx **ctx);
#define BM_BATCH_PER_PAIR 0  /* the pairs' launches one pair after the other on one stream */
#define BM_BATCH_GRAPH 1     /* every pair's own launches on parallel branches, one hipGraph launch */
#define BM_BATCH_MULTI 2     /* one launch per stage and kernel variant for all pairs together (per-pair
                                parameter blocks and a tile index across pairs), captured as one hipGraph */
#define BM_BATCH_AUTO 3      /* the mode of the batch's schedule (bm_batch_autotune) */
/* The buffers (and every pair's configuration) of a graph are captured on first
 * use and replayed while they stay the same.  In the multi mode, pairs whose
 * context cannot use it (Lanczos up-sampler, a BRIEF pattern too wide for the
 * fused kernel, jfMax 0) run their own launches inside the same graph. */
int bm_batch_flow(bm_batch *b, const uint8_t *const *rgb1_dev, const uint8_t *const *rgb2_dev, float *const *U_dev,
                  float *const *V_dev, int mode);
/* Result-invariant schedule of a batch: which mode BM_BATCH_AUTO runs, how
 * pairs are spread over parallel branches, their order, and the multi mode's
 * resident work-groups per CU. */
typedef struct bm_batch_schedule {
    int mode;      /* BM_BATCH_PER_PAIR, BM_BATCH_GRAPH or BM_BATCH_MULTI */
    int branches;  /* parallel branches, 1 .. the streams of bm_batch_create: the graph mode's per-pair
                      branches; the multi mode's lanes, each a subset of the pairs with its own
                      multi-pair launches, so different stages of different lanes overlap */
    int order;     /* 0: pairs in index order, round-robin over branches;
                      1: largest first, each onto the least-loaded branch */
    int persist;   /* multi mode: resident work-groups per CU walking the tiles; 0 = one per tile */
} bm_batch_schedule;
int bm_batch_get_schedule(const bm_batch *b, bm_batch_schedule *out);
int bm_batch_set_schedule(bm_batch *b, const bm_batch_schedule *s);
/* Times the batch on these buffers under candidate schedules (both orders;
 * the graph mode with half and all branches; the multi mode with 1, 2 and 4
 * lanes, one work-group per tile or 4 resident per CU), keeps the fastest and
 * returns it in *out (may be NULL).  Results do not depend on the schedule. */
int bm_batch_autotune(bm_batch *b, const uint8_t *const *rgb1_dev, const uint8_t *const *rgb2_dev,
                      float *const *U_dev, float *const *V_dev, bm_batch_schedule *out);
int bm_batch_sync(bm_batch *b);
int bm_batch_timer_start(bm_batch *b);
int bm_batch_timer_stop(bm_batch *b, float *ms);
const char *bm_batch_last_error(const bm_batch *b);

/* ---- learned proposal --------------------------------------------------------
 * No counterpart in the reference: its only initialisations are the zero and
 * the temporal field (briefmatch_cuda.cu:149-184, the hook these calls extend).
 *
 * A convolutional network at 1/8 of the input resolution predicts a coarse
 * flow from a block-matching cost volume; its offsets seed PatchMatch next to the plain
 * initialisation, and each pixel keeps the strictly cheaper candidate, so the
 * plain initialisation stays the fallback.  BF16 weights and activations on
 * MFMA, FP32 accumulation; trained by bm_net_train_step.
 *
 * Feature maps are [n][ch][cw][BM_NET_FEATURES] BF16 (uint16) with
 * cw = W / 8, ch = H / 8 (bm_proposal_dims); network outputs are
 * [n][ch][cw][BM_NET_OUTPUTS] FP32, channels 0 and 1 = coarse flow (u, v) in
 * coarse pixels (8 input pixels): the network's residual plus 8 x feature
 * channels 123, 124 (a fixed skip connection); labels are [n][ch][cw][2] FP32
 * in the same unit (non-finite = ignored).
 * The parameter vector is, per layer, W[9][cout][cin] then b[cout]
 * (layers 128->64, 64->64 dil 2, 64->64 dil 4, 64->64, 64->16). */
#define BM_NET_FEATURES 128
#define BM_NET_OUTPUTS 16
#define BM_NET_PARAMS 193808
typedef struct bm_net bm_net;
int bm_net_create(bm_net **out);            /* all weights zero */
void bm_net_destroy(bm_net *net);           /* waits for all device work (contexts may read its weights) */
const char *bm_net_last_error(const bm_net *net);
int bm_net_set_weights(bm_net *net, const float *w_host, size_t count);  /* count = BM_NET_PARAMS; resets Adam */
int bm_net_get_weights(bm_net *net, float *w_host, size_t count);
int bm_net_load(bm_net *net, const char *path);  /* "BMNET001", uint32 count, count FP32 (little endian) */
int bm_net_save(bm_net *net, const char *path);
/* Adam moments decay rates and epsilon (defaults 0.9, 0.999, 1e-8). */
int bm_net_set_optimizer(bm_net *net, float beta1, float beta2, float eps);
/* Inference on the net's stream (synchronous on return). */
int bm_net_forward(bm_net *net, const uint16_t *feat_dev, unsigned n, unsigned ch, unsigned cw, float *out_dev);
/* Mean absolute error of channels 0, 1 against the labels and its gradient
 * with respect to every parameter (host, BM_NET_PARAMS floats); no update. */
int bm_net_gradients(bm_net *net, const uint16_t *feat_dev, const float *label_dev, unsigned n, unsigned ch,
                     unsigned cw, float *loss, float *grad_host);
/* One Adam step on the mini-batch rows idx_host[0..batch) (batch <= 256) of a
 * dataset of `rows` feature maps and labels (device, row = one pair).  *loss
 * (if not NULL) receives the batch loss before the update and makes the call
 * wait.  Contexts must not run bm_flow_pair_learned with a net in training. */
int bm_net_train_step(bm_net *net, const uint16_t *feat_set_dev, const float *label_set_dev, unsigned rows,
                      const int *idx_host, unsigned batch, unsigned ch, unsigned cw, float lr, float *loss);
int bm_net_sync(bm_net *net);

/* Coarse grid of a context: cw = W / 8, ch = H / 8 (both >= 1). */
int bm_proposal_dims(const bm_ctx *ctx, unsigned *cw, unsigned *ch);
/* The context's own gray frames and descriptors of the last bm_flow_pair /
 * bm_flow_pair_learned (g8[i], F[i] for frame i + 1). */
int bm_ctx_frames(bm_ctx *ctx, uint8_t **g8_dev, uint32_t **F_dev);
/* Feature map of a pair from its gray frames: channels 0..120 the cost volume
 * over displacements (dx, dy) in -5..5 coarse pixels, channel (dy + 5) * 11 +
 * dx + 5 (sum of absolute differences of 8x8 block sums over a 3x3 block
 * window, clamp-to-edge, min(SAD, 36864) / 18432 - 1); 121, 122 the block
 * means of both frames / 255 - 1/2; 123, 124 the lowest-cost displacement
 * (lowest channel on ties) / 8; 125 its scaled cost; 126, 127 zero. */
int bm_proposal_features(bm_ctx *ctx, const uint8_t *g1_dev, const uint8_t *g2_dev, uint16_t *feat_dev);
/* Far block match of a pair from its gray frames, [H / 16][W / 16][2] int32
 * (frames of at least 16x16): per 16x16 block the displacement (dx, dy) in
 * -5..5 blocks (+-80 pixels, twice the cost volume's reach) with the lowest
 * sum of absolute differences of 16x16 block sums over a 3x3 block window
 * (clamp-to-edge); ties go to the lower ring max(|dx|, |dy|), then to the
 * lower (dy + 5) * 11 + dx + 5.  A displacement of ring >= 3 is reported
 * only when its sum is below 3/4 of the best sum of rings 0..2; otherwise the
 * best of rings 0..2 is. */
int bm_proposal_far(bm_ctx *ctx, const uint8_t *g1_dev, const uint8_t *g2_dev, int32_t *far_dev);
/* 8x8 block means of a flow field divided by 8 (training labels), [ch][cw][2]. */
int bm_coarse_flow(bm_ctx *ctx, const float *U_dev, const float *V_dev, float *label_dev);
/* Per-cell proposals as matched-pixel offsets in Q8, [ch][cw][4] int32: the
 * network's coarse flow (u, v) and the feature map's lowest-cost displacement
 * (8 x channels 123, 124), each rint(clamp(u * (float)(2048 sx) / (float)W,
 * +-256 sx)) (v likewise with sy, H).  far_dev (bm_proposal_far, or NULL):
 * a cell whose block (cx / 2, cy / 2, clamped to the far grid) reports ring
 * max(|dx|, |dy|) >= 3 (48 pixels and more, beyond the cost volume) takes
 * (2 dx, 2 dy) coarse pixels in place of the network's flow. */
int bm_proposal_offsets(bm_ctx *ctx, const float *net_out_dev, const uint16_t *feat_dev, const int32_t *far_dev,
                        int32_t *coff_dev);
/* Proposal-seeded initialisation: the plain bm_init candidate first, then for
 * the network's and then the argmin proposal the bilinear interpolation of
 * the four surrounding cells and the four cells' own offsets; out-of-frame
 * candidates and repeats of the running best are skipped, strictly cheaper
 * wins. */
int bm_init_proposal(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev, int first,
                     const uint32_t *prev_off_dev, const int32_t *coff_dev, uint32_t *off_dev, uint16_t *cost_dev);
/* bm_flow_pair with the proposal-seeded initialisation followed by one random
 * search (random-search iteration index `iterations`), then the context's
 * `iterations` propagation iterations.  The offsets take the far block match
 * when the frames are at least 16x16 pixels. */
int bm_flow_pair_learned(bm_ctx *ctx, bm_net *net, const uint8_t *rgb1_dev, const uint8_t *rgb2_dev, uint32_t pair,
                         float *U_dev, float *V_dev);
/* Local search around the seeds (no counterpart in the reference, whose random
 * search samples around offset zero; PatchMatch's search around the current
 * offset with a shrinking radius).  With samples > 0, bm_flow_pair_learned
 * runs it after the seeded initialisation's random search, before the
 * propagation iterations: per matched pixel, sample s = 0 .. samples - 1 is
 * the running best offset plus a displacement uniform in [-r, r]^2, r =
 * max(radius >> s, 1) matched pixels, from Philox4x32-10(ctr = {pixel,
 * iteration, pair, 1 + s}, key = seed) (d = floor(word * (2r + 1) / 2^32) -
 * r, words 0 and 1), clamped to the frame, inserted like a random-search
 * candidate.  bm_flow_pair_learned passes iteration = `iterations`, the seeded
 * initialisation's random-search index.  0 samples (the default) turns it
 * off. */
#define BM_LOCAL_SEARCH_MAX_SAMPLES 16
#define BM_LOCAL_SEARCH_MAX_RADIUS 4096
int bm_set_local_search(bm_ctx *ctx, int radius, int samples);
int bm_get_local_search(const bm_ctx *ctx, int *radius, int *samples);
/* The local search alone on a field (off/cost updated in place) with the
 * context's radius and samples. */
int bm_local_search(bm_ctx *ctx, const uint32_t *F1_dev, const uint32_t *F2_dev, uint32_t *off_dev,
                    uint16_t *cost_dev, uint32_t iteration, uint32_t pair);

/* ---- timing (CudaTimer, util_cuda.h:131-141) ----------------------------- */

int bm_timer_start(bm_ctx *ctx);
/* Milliseconds since bm_timer_start on the context stream (waits for it). */
int bm_timer_stop(bm_ctx *ctx, float *ms);

/* Per-stage device time.  While enabled, every kernel the context launches is
 * bracketed by HIP events on its stream (the reference times whole phases with
 * CudaTimer, of_pipeline.cpp:267-325).  bm_profile_read waits for the stream and
 * returns, per stage id, the summed milliseconds and the launch count since the
 * last read, then resets them. */
#define BM_STAGE_GRAY 0
#define BM_STAGE_UPSAMPLE 1
#define BM_STAGE_BRIEF 2
#define BM_STAGE_INIT 3
#define BM_STAGE_LEVEL 4          /* jump-flooding level */
#define BM_STAGE_LEVEL_RS 5       /* last level with the fused random search */
#define BM_STAGE_RANDOM_SEARCH 6
#define BM_STAGE_DOWNSAMPLE 7
#define BM_STAGE_MEDIAN 8
#define BM_STAGE_TRILATERAL 9
#define BM_STAGE_COPY 10
#define BM_STAGE_PROPOSAL 11      /* learned proposal: features, network, seeded init */
#define BM_STAGE_LOCAL_SEARCH 12  /* learned path: local search around the seeds */
#define BM_STAGE_COUNT 13
int bm_profile_enable(bm_ctx *ctx, int on);
int bm_profile_read(bm_ctx *ctx, double *ms, unsigned *launches);

/* Plain device-to-device copy kernel on the context stream (bytes % 16 == 0);
 * the reference point for HBM traffic counters. */
int bm_device_copy(bm_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* ---- metrics ------------------------------------------------------------- */

/* End-point error against ground truth (BriefMatch::error, briefmatch.cpp:468-514),
 * host arrays; occluded may be NULL.  epe_noc skips occluded pixels. */
int bm_epe(const float *U_host, const float *V_host, const float *gtU_host, const float *gtV_host,
           const uint8_t *occluded_host, size_t n, double *epe_all, double *epe_noc);

#ifdef __cplusplus
}
#endif
#endif
