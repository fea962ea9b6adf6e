// Internal interface between the C-ABI layer (bm_capi.cpp) and the HIP kernels
// (bm_kernels.hip).  Launchers enqueue on the given stream and return the
// hipError_t of the launch; they never synchronise.
#ifndef BM_INTERNAL_H
#define BM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/briefmatch_amd.h"

namespace bm {

// Geometry and tables of one configuration, passed by value to the launchers.
struct Geometry {
    int W, H;        // input resolution
    int sx, sy;      // matched resolution
    int N;           // feature bits
    int NS;          // ceil(N / 32)
    int NSP;         // descriptor stride: NS rounded up to a power of two
    int P;           // Gaussian taps
    int fsize;       // median window
    int jfMax, iterations, doubleProp;
    int brief_radius;             // LDS halo of the stand-alone BRIEF tile (pixels)
    int brief_far;                // largest |pattern offset| (fused BRIEF far halo)
    int brief_near;               // fused BRIEF near halo: 8, 16, 24 or 32; 0 = kernel unavailable
    int brief_gw[2], brief_gh[2]; // largest gray tile of the fused BRIEF, [0]: 16-row tiles, [1]: 32-row tiles
    const uint32_t *brief_plan;   // fused BRIEF test records (4 words each, 64 per descriptor word)
    const int *brief_plan_count;  // per descriptor word: bit records, far records
    uint64_t seed;                // pattern + random-search key
    int64_t rs_sigma_q16;         // random-search sigma in Q16
    float tri_a2, tri_b;          // trilateral range scales: of the squared flow difference, of the gray level
    float tri_sigma_epe, tri_sigma_i, tri_sigma_s;  // squared sigmas
    int tri_s;                    // trilateral half window
    // device tables
    const int16_t *pattern;       // 4N
    const uint32_t *up_x, *up_y;  // bilinear (x0 << 8 | frac) per matched column / row
    const uint32_t *nn_x, *nn_y;  // nearest-neighbour source column / row per input pixel
    int ds_lds;                   // LDS bytes of the tiled down-sampling (largest tile window), 0 = direct kernel
    const float *gauss;           // P taps
    const float *tri_spatial;     // 2s entries: log2(e) i^2 / (2 sigmaS^2), i = -s..s-1
    const int *lz_x0, *lz_y0;     // Lanczos-3: first source column / row per output column / row
    const float *lz_wx, *lz_wy;   // Lanczos-3: six normalised weights per output column / row
    // tuning (results do not depend on it)
    int level_px;                 // pixels per thread of the level kernels: 1 or 2
    int strip_rows;               // strip height of the matching kernels' block order
    int tri_band;                 // max source rows per LDS band of the trilateral filter
    int tri_groups;               // trilateral kernel: 0 VALU sums, 2 / 4 MFMA sums with that many 4-row groups
    int brief_frame;              // 1: the whole-path calls up-sample into a frame and BRIEF reads it
    int tri_split;                // work-groups per trilateral tile (shares of its source rows): 1, 2, 4 or 8
    float *tri_part;              // their partial sums, 3 tri_split W H floats (tri_split > 1)
};

// Largest trilateral band (rows) that fits the LDS budget for half window s.
int tri_max_band(int s);

hipError_t launch_gray(const Geometry &g, const uint8_t *rgb, uint8_t *g8, float *g32, hipStream_t st);
// both frames of a pair in one launch (FP32 copy of the first only)
hipError_t launch_gray_pair(const Geometry &g, const uint8_t *rgb1, uint8_t *g8a, float *g32a, const uint8_t *rgb2,
                            uint8_t *g8b, hipStream_t st);
hipError_t launch_upsample(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st);
hipError_t launch_upsample_lanczos(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st);
hipError_t launch_brief(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F, hipStream_t st);
// the fused BRIEF (bilinear from the gray frame, brief_near > 0) of both frames in one launch
hipError_t launch_brief_pair(const Geometry &g, const uint8_t *g8a, uint32_t *Fa, const uint8_t *g8b, uint32_t *Fb,
                             hipStream_t st);
// rs_iteration >= 0 inserts that iteration's random search after the initial candidate.
hipError_t launch_init(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int first,
                       const uint32_t *prev_off, uint32_t *off, uint16_t *cost, int rs_iteration, uint32_t pair,
                       hipStream_t st);
// One propagation level; rs_iteration >= 0 fuses the random search of that
// iteration after the level.
hipError_t launch_level(const Geometry &g, const uint32_t *F1, const uint32_t *F2,
                        const uint32_t *off_in, const uint16_t *cost_in, uint32_t *off_out,
                        uint16_t *cost_out, int p, int pd, int rs_iteration, uint32_t pair,
                        hipStream_t st);
hipError_t launch_random_search(const Geometry &g, const uint32_t *F1, const uint32_t *F2,
                                uint32_t *off, uint16_t *cost, uint32_t iteration, uint32_t pair,
                                hipStream_t st);
// Proposal-seeded initialisation from coff [ch * cw][4] (network and argmin
// offsets, Q8 matched pixels); rs_iteration >= 0 fuses that random search.
hipError_t launch_init_proposal(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int first,
                                const uint32_t *prev, const int32_t *coff, int cw, int ch, int rs_iteration,
                                uint32_t pair, uint32_t *off, uint16_t *cost, hipStream_t st);
// Local search around the current offsets (in place): `samples` candidates,
// sample s uniform in [-r, r]^2 around the running best, r = max(radius >> s, 1).
hipError_t launch_local_search(const Geometry &g, const uint32_t *F1, const uint32_t *F2, int radius, int samples,
                               uint32_t iteration, uint32_t pair, uint32_t *off, uint16_t *cost, hipStream_t st);
hipError_t launch_downsample(const Geometry &g, const uint32_t *off, float *U, float *V, hipStream_t st);
// Output tile of the LDS-tiled down-sampling (columns x rows of input pixels).
constexpr int kDsTileW = 64, kDsTileH = 4, kDsLdsMax = 64 * 1024;
hipError_t launch_median(const Geometry &g, const float *in0, const float *in1, float *out0, float *out1,
                         int dir, hipStream_t st);
hipError_t launch_trilateral(const Geometry &g, const float *Ui, const float *Vi, const float *Um,
                             const float *Vm, const float *I, float *Uo, float *Vo, hipStream_t st);

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);

// ---------------------------------------------------------------------------
// Multi-pair launches (batches of uneven pairs).  A job is one pair's (or one
// frame's) arguments of a stage, kept in device memory; one launch covers the
// tiles of all its jobs with the single-pair kernels' tile shapes, so results
// are identical.  `start` holds njobs + 1 work-group slot offsets, each job's
// padded to a multiple of 8, then the njobs tile counts; `tiles` is the slot
// total.  Each job's tiles are spread over the 8 XCDs as in its own launch
// (job_of in bm_device.h).  grid == slots gives one work-group per slot; a
// smaller grid (a multiple of 8) keeps its work-groups resident, each walking
// the slots of its XCD.

struct GrayJob {
    const uint8_t *rgb;
    uint8_t *g8;
    float *g32;  // may be null
    int n;       // pixels
};

// Fused up-sampling + BRIEF of one frame.
struct BriefJob {
    const uint8_t *g8;
    const uint32_t *up_x, *up_y;
    const uint32_t *plan;   // 4 words per record, 64 records per descriptor word
    const int *plan_count;  // per word: bit records, far records
    int N, W, H, sx, sy, Rf, gw, gh;
    uint32_t *F;
    const float *L = nullptr;  // up-sampled frame: tile and far tests read it (else interpolated from g8)
};

// Initialisation and propagation levels of one pair: the field ping-pongs
// between off[0] / cost[0] and off[1] / cost[1].
struct MatchJob {
    const uint32_t *F1, *F2;
    uint32_t *off[2];
    uint16_t *cost[2];
    int sx, sy, double_prop, strip_rows;
    uint32_t pair, key0, key1;
    int64_t sigma_q16;
};

struct DsJob {
    const uint32_t *off, *nn_x, *nn_y;
    const float *gauss;
    int P, sx, sy, W, H;
    int identity;  // matched = input resolution: the field itself (direct kernel only)
    float scx, scy;
    float *U, *V;
};

struct MedJob {
    const float *in0, *in1;
    float *out0, *out1;
    int W, H;
};

struct TriJob {
    const float *Ui, *Vi, *Um, *Vm, *I;
    float *Uo, *Vo;
    const float *spt;
    int W, H, s, BR;
    float a2, bI;
};

// Tiles of one job in each stage (the single-pair launch's grid).
int gray_tiles(int n);
int brief_tiles(const Geometry &g);
int match_tiles(const Geometry &g, int px);  // px = 0: init (one pixel per thread)
int downsample_tiles(const Geometry &g, bool tiled);
int median_tiles(const Geometry &g);         // per flow component
int trilateral_tiles(const Geometry &g, int groups);
// LDS bytes of the single-pair launch of a stage for this geometry.
size_t brief_lds(const Geometry &g);
size_t trilateral_lds(const Geometry &g, int *band);

hipError_t launch_multi_gray(const GrayJob *jobs, const int *start, int njobs, int tiles, int grid, hipStream_t st);
hipError_t launch_multi_brief(int NSP, int RN, const BriefJob *jobs, const int *start, int njobs, int tiles, int grid,
                              size_t lds, hipStream_t st);
// rs_iteration >= 0 inserts that random search after the zero / level candidate.
hipError_t launch_multi_init(int NSP, const MatchJob *jobs, const int *start, int njobs, int tiles, int grid,
                             int rs_iteration, hipStream_t st);
hipError_t launch_multi_level(int NSP, int PX, const MatchJob *jobs, const int *start, int njobs, int tiles, int grid,
                              int p, int pd, int src, int rs_iteration, hipStream_t st);
hipError_t launch_multi_downsample(bool tiled, const DsJob *jobs, const int *start, int njobs, int tiles, int grid,
                                   size_t lds, hipStream_t st);
// both flow components: tiles = 2 x the per-component tiles of each job
hipError_t launch_multi_median(int FS, int dir, const MedJob *jobs, const int *start, int njobs, int tiles, int grid,
                               hipStream_t st);
hipError_t launch_multi_trilateral(int groups, const TriJob *jobs, const int *start, int njobs, int tiles, int grid,
                                   size_t lds, hipStream_t st);

// LDS budget of the BRIEF tile: the largest halo radius that fits.
int brief_max_radius();
// Fused BRIEF: LDS bytes of one work-group for near halo RN and tile height TH,
// and the tile height used for a descriptor stride.
size_t brief_lds_bytes(const Geometry &g, int RN, int TH);
int brief_tile_rows(int NSP);

}  // namespace bm

#endif
