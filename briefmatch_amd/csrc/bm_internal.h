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
};

// Largest trilateral band (rows) that fits the LDS budget for half window s.
int tri_max_band(int s);

hipError_t launch_gray(const Geometry &g, const uint8_t *rgb, uint8_t *g8, float *g32, hipStream_t st);
hipError_t launch_upsample(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st);
hipError_t launch_upsample_lanczos(const Geometry &g, const uint8_t *g8, float *L, hipStream_t st);
hipError_t launch_brief(const Geometry &g, const float *L, const uint8_t *g8, uint32_t *F, hipStream_t st);
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
hipError_t launch_downsample(const Geometry &g, const uint32_t *off, float *U, float *V, hipStream_t st);
// Output tile of the LDS-tiled down-sampling (columns x rows of input pixels).
constexpr int kDsTileW = 64, kDsTileH = 4, kDsLdsMax = 64 * 1024;
hipError_t launch_median(const Geometry &g, const float *in0, const float *in1, float *out0, float *out1,
                         int dir, hipStream_t st);
hipError_t launch_trilateral(const Geometry &g, const float *Ui, const float *Vi, const float *Um,
                             const float *Vm, const float *I, float *Uo, float *Vo, hipStream_t st);

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);

// LDS budget of the BRIEF tile: the largest halo radius that fits.
int brief_max_radius();
// Fused BRIEF: LDS bytes of one work-group for near halo RN and tile height TH,
// and the tile height used for a descriptor stride.
size_t brief_lds_bytes(const Geometry &g, int RN, int TH);
int brief_tile_rows(int NSP);

}  // namespace bm

#endif
