// Internal interface of the learned proposal: network layout, the bm_net
// object and the launchers of bm_net.hip.
#ifndef BM_NET_H
#define BM_NET_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/briefmatch_amd.h"

namespace bm {
namespace net {

constexpr int kRadius = 5;                                        // cost-volume radius, coarse pixels
constexpr int kCost = (2 * kRadius + 1) * (2 * kRadius + 1);      // 121 matching-score channels
constexpr int kFeatC = BM_NET_FEATURES;                           // 128 input channels
constexpr int kOutC = BM_NET_OUTPUTS;                             // 16 output channels
constexpr int kCell = 8;                                          // input pixels per coarse pixel
constexpr int kLayers = 5;
// cost-volume scaling: SAD of 9 block sums; 9 * 64 * 64 (a mean absolute
// difference of 64 gray levels) maps to 1, zero to -1
constexpr int kSadClamp = 9 * 64 * 64;
// feature channels of the lowest-cost displacement (dx / 8, dy / 8), added
// times 8 to output channels 0, 1
constexpr int kSkipChannel = kCost + 2;
// far block match: 16x16 blocks (two coarse pixels), displacements -5..5
// blocks (+-80 input pixels, twice the cost volume's reach); a match at ring
// max(|dx|, |dy|) >= kFarMin whose SAD is below 3/4 of the best nearer one
// replaces the network's proposal of the cells under the block
constexpr int kFarCell = 2 * kCell;
constexpr int kFarRadius = 5;
constexpr int kFarN = (2 * kFarRadius + 1) * (2 * kFarRadius + 1);
constexpr int kFarMin = 3;

struct LayerSpec {
    int cin, cout, dil, relu;
};
constexpr LayerSpec kSpec[kLayers] = {{128, 64, 1, 1}, {64, 64, 2, 1}, {64, 64, 4, 1}, {64, 64, 1, 1}, {64, 16, 1, 0}};

// Offset of layer l's weights [9][cout][cin] and bias [cout] in the flat parameter vector.
constexpr size_t layer_offset(int l)
{
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)9 * kSpec[i].cout * kSpec[i].cin + kSpec[i].cout;
    return o;
}
constexpr size_t kParams = layer_offset(kLayers);
static_assert(kParams == BM_NET_PARAMS, "parameter count");

// Packed BF16 copies of the weights: forward [9][cout][cin] and data-gradient
// [9][cin][cout] with the taps reversed.
struct Packed {
    uint16_t *w[kLayers];
    uint16_t *wt[kLayers];
    const float *b[kLayers];
};

enum { ACT_NONE = 0, ACT_RELU = 1, ACT_MASK = 2 };

// S: scratch of 2 * cw * ch int32 (block sums)
hipError_t launch_features(const uint8_t *g1, const uint8_t *g2, int W, int cw, int ch, int32_t *S, uint16_t *feat,
                           hipStream_t st);
hipError_t launch_block_sums(const uint8_t *g1, const uint8_t *g2, int W, int cw, int ch, int32_t *S, hipStream_t st);
// Far block match of the 8x8 block sums S [2][ch][cw] (as launch_features
// leaves them): far [ch / 2][cw / 2][2] int32, (dx, dy) in 16-pixel blocks.
hipError_t launch_far_search(const int32_t *S, int cw, int ch, int32_t *far, hipStream_t st);
// skip (optional, the feature map): output channels 0, 1 += 8 * skip[p][kSkipChannel + c]
hipError_t launch_conv(int cout, const uint16_t *in, int cin, const uint16_t *w, const float *bias, int npix, int H,
                       int W, int dil, int act, const uint16_t *ref, const uint16_t *skip, uint16_t *out_bf,
                       float *out_f, hipStream_t st);
int wgrad_chunks(int npix);
hipError_t launch_wgrad(int cout, int cin, const uint16_t *dz, const uint16_t *x, int npix, int H, int W, int dil,
                        float *part, float *part_bias, float *grad_w, float *grad_b, hipStream_t st);
int loss_blocks(int npix);
hipError_t launch_loss(const float *pred, const float *label, int npix, uint16_t *dz, float *loss_part, hipStream_t st);
hipError_t launch_sum(const float *x, int n, float *out, hipStream_t st);
hipError_t launch_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st);
hipError_t launch_pack(const float *w, int cout, int cin, uint16_t *wb, uint16_t *wt, hipStream_t st);
hipError_t launch_coarse_flow(const float *U, const float *V, int W, int cw, int ch, float *out, hipStream_t st);
// coff [ch][cw][4]: network (u, v) and cost-volume argmin (u, v), Q8 matched
// pixels; far (optional, [ch / 2][cw / 2][2]) overrides the network's (u, v)
// where it reaches ring kFarMin
hipError_t launch_proposal_offsets(const float *net_out, const uint16_t *feat, const int32_t *far, int cw, int ch,
                                   float scale_x, float scale_y, int lim_x, int lim_y, int32_t *coff, hipStream_t st);
// Row indices of a mini-batch, passed by value as a kernel argument.
constexpr int kMaxBatch = 256;
struct RowList {
    int idx[kMaxBatch];
};
hipError_t launch_gather_rows(const void *src, const RowList &rows_idx, size_t row_bytes, int rows, void *dst,
                              hipStream_t st);

// Inference: the forward pass of a batch of n feature maps [n][ch][cw][128]
// into out [n][ch][cw][16] (FP32), with two activation buffers of
// n * ch * cw * 64 BF16 each.
hipError_t forward(const Packed &pk, const uint16_t *feat, int n, int ch, int cw, uint16_t *act0, uint16_t *act1,
                   float *out, hipStream_t st);

}  // namespace net
}  // namespace bm

// The network object: FP32 master weights, Adam moments, packed BF16 weights
// and the training buffers (grown on demand), all on the net's own stream.
struct bm_net {
    hipStream_t stream = nullptr;
    std::string err;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // kParams each
    bm::net::Packed pk{};
    uint16_t *packed = nullptr;                                      // storage of pk.w / pk.wt
    int64_t step = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    // training buffers, sized for cap_pix pixels
    int cap_pix = 0, cap_rows = 0;
    uint16_t *x0 = nullptr;            // gathered features [P][128]
    float *label = nullptr;            // gathered labels [P][2]
    uint16_t *act[4] = {};             // ReLU activations [P][64]
    float *out = nullptr;              // [P][16]
    uint16_t *dz4 = nullptr;           // [P][16]
    uint16_t *dz[2] = {};              // [P][64]
    float *part = nullptr, *part_bias = nullptr, *loss_part = nullptr, *loss = nullptr;
    size_t part_cap = 0;
    hipEvent_t ready = nullptr;        // recorded after every change of the packed weights
    std::vector<void *> bufs;          // weights, moments, packed weights
    std::vector<void *> tbufs;         // training buffers
};

#endif
