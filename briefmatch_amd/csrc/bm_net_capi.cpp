// C-ABI of the learned proposal network: weights, Adam training on device
// mini-batches, gradients and inference.  Host code only; the kernels are in
// bm_net.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bm_net.h"

using namespace bm::net;

namespace {

thread_local std::string g_net_error;

int nfail(bm_net *net, int code, const std::string &msg)
{
    if (net)
        net->err = msg;
    else
        g_net_error = msg;
    return code;
}

#define NET_HIP(net, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return nfail((net), BM_E_LAUNCH, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

template <class T>
int nalloc(bm_net *net, std::vector<void *> &list, T **p, size_t count)
{
    void *q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return nfail(net, BM_E_ALLOC, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    list.push_back(q);
    *p = static_cast<T *>(q);
    return BM_OK;
}

size_t weight_count(int l) { return (size_t)9 * kSpec[l].cout * kSpec[l].cin; }

int repack(bm_net *net)
{
    for (int l = 0; l < kLayers; ++l)
        NET_HIP(net, launch_pack(net->w + layer_offset(l), kSpec[l].cout, kSpec[l].cin, net->pk.w[l], net->pk.wt[l],
                                 net->stream));
    NET_HIP(net, hipEventRecord(net->ready, net->stream));
    return BM_OK;
}

void free_training(bm_net *net)
{
    if (net->tbufs.empty()) return;
    (void)hipStreamSynchronize(net->stream);
    for (void *p : net->tbufs) (void)hipFree(p);
    net->tbufs.clear();
    net->cap_pix = 0;
    net->part_cap = 0;
}

// Training / inference buffers for npix pixels.
int ensure_buffers(bm_net *net, int npix)
{
    if (npix <= net->cap_pix) return BM_OK;
    free_training(net);
    const size_t P = (size_t)npix;
    int r = nalloc(net, net->tbufs, &net->x0, P * kFeatC);
    if (!r) r = nalloc(net, net->tbufs, &net->label, P * 2);
    for (int i = 0; i < 4 && !r; ++i) r = nalloc(net, net->tbufs, &net->act[i], P * 64);
    if (!r) r = nalloc(net, net->tbufs, &net->out, P * kOutC);
    if (!r) r = nalloc(net, net->tbufs, &net->dz4, P * kOutC);
    for (int i = 0; i < 2 && !r; ++i) r = nalloc(net, net->tbufs, &net->dz[i], P * 64);
    const size_t chunks = (size_t)wgrad_chunks(npix);
    if (!r) r = nalloc(net, net->tbufs, &net->part, chunks * 9 * 64 * 128);
    if (!r) r = nalloc(net, net->tbufs, &net->part_bias, chunks * 64);
    if (!r) r = nalloc(net, net->tbufs, &net->loss_part, (size_t)loss_blocks(npix));
    if (!r) r = nalloc(net, net->tbufs, &net->loss, 1);
    if (r) {
        free_training(net);
        return r;
    }
    net->cap_pix = npix;
    return BM_OK;
}

// Forward with all activations kept, loss, and the gradient of every
// parameter into net->g (the loss into net->loss).
int forward_backward(bm_net *net, const uint16_t *x0, const float *label, int n, int ch, int cw)
{
    const int npix = n * ch * cw;
    hipStream_t st = net->stream;
    const uint16_t *x = x0;
    for (int l = 0; l < kLayers - 1; ++l) {
        NET_HIP(net, launch_conv(kSpec[l].cout, x, kSpec[l].cin, net->pk.w[l], net->pk.b[l], npix, ch, cw,
                                 kSpec[l].dil, ACT_RELU, nullptr, nullptr, net->act[l], nullptr, st));
        x = net->act[l];
    }
    const int L = kLayers - 1;
    NET_HIP(net, launch_conv(kSpec[L].cout, x, kSpec[L].cin, net->pk.w[L], net->pk.b[L], npix, ch, cw, kSpec[L].dil,
                             ACT_NONE, nullptr, x0, nullptr, net->out, st));
    NET_HIP(net, launch_loss(net->out, label, npix, net->dz4, net->loss_part, st));
    NET_HIP(net, launch_sum(net->loss_part, loss_blocks(npix), net->loss, st));
    const uint16_t *dzl = net->dz4;
    for (int l = kLayers - 1; l >= 0; --l) {
        const LayerSpec &s = kSpec[l];
        const uint16_t *xin = l ? net->act[l - 1] : x0;
        float *gw = net->g + layer_offset(l);
        NET_HIP(net, launch_wgrad(s.cout, s.cin, dzl, xin, npix, ch, cw, s.dil, net->part, net->part_bias, gw,
                                  gw + weight_count(l), st));
        if (l > 0) {
            uint16_t *dnext = net->dz[l & 1];
            NET_HIP(net, launch_conv(s.cin, dzl, s.cout, net->pk.wt[l], nullptr, npix, ch, cw, s.dil, ACT_MASK,
                                     net->act[l - 1], nullptr, dnext, nullptr, st));
            dzl = dnext;
        }
    }
    return BM_OK;
}

int check_dims(bm_net *net, unsigned n, unsigned ch, unsigned cw)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!n || !ch || !cw) return nfail(net, BM_E_PARAM, "empty feature map");
    if ((uint64_t)n * ch * cw > (1u << 26)) return nfail(net, BM_E_PARAM, "feature map larger than 2^26 pixels");
    return BM_OK;
}

}  // namespace

extern "C" {

const char *bm_net_last_error(const bm_net *net) { return net ? net->err.c_str() : g_net_error.c_str(); }

int bm_net_create(bm_net **out)
{
    if (!out) return nfail(nullptr, BM_E_PARAM, "null output pointer");
    *out = nullptr;
    bm_net *net = new (std::nothrow) bm_net();
    if (!net) return nfail(nullptr, BM_E_ALLOC, "host allocation failed");
    int r = BM_OK;
    if (hipStreamCreate(&net->stream) != hipSuccess || hipEventCreateWithFlags(&net->ready, hipEventDisableTiming) != hipSuccess)
        r = nfail(nullptr, BM_E_LAUNCH, "stream / event creation failed");
    if (!r) r = nalloc(net, net->bufs, &net->w, kParams);
    if (!r) r = nalloc(net, net->bufs, &net->m, kParams);
    if (!r) r = nalloc(net, net->bufs, &net->v, kParams);
    if (!r) r = nalloc(net, net->bufs, &net->g, kParams);
    size_t packed = 0;
    for (int l = 0; l < kLayers; ++l) packed += 2 * weight_count(l);
    if (!r) r = nalloc(net, net->bufs, &net->packed, packed);
    if (r) {
        if (!net->err.empty()) g_net_error = net->err;
        bm_net_destroy(net);
        return r;
    }
    uint16_t *q = net->packed;
    for (int l = 0; l < kLayers; ++l) {
        net->pk.w[l] = q;
        q += weight_count(l);
        net->pk.wt[l] = q;
        q += weight_count(l);
        net->pk.b[l] = net->w + layer_offset(l) + weight_count(l);
    }
    if (hipMemsetAsync(net->w, 0, kParams * sizeof(float), net->stream) != hipSuccess ||
        hipMemsetAsync(net->m, 0, kParams * sizeof(float), net->stream) != hipSuccess ||
        hipMemsetAsync(net->v, 0, kParams * sizeof(float), net->stream) != hipSuccess ||
        hipMemsetAsync(net->g, 0, kParams * sizeof(float), net->stream) != hipSuccess || repack(net) != BM_OK ||
        hipStreamSynchronize(net->stream) != hipSuccess) {
        g_net_error = "net initialisation failed";
        bm_net_destroy(net);
        return BM_E_LAUNCH;
    }
    *out = net;
    return BM_OK;
}

void bm_net_destroy(bm_net *net)
{
    if (!net) return;
    // contexts may still have queued learned-path work reading the packed weights
    (void)hipDeviceSynchronize();
    free_training(net);
    for (void *p : net->bufs) (void)hipFree(p);
    if (net->ready) (void)hipEventDestroy(net->ready);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
}

int bm_net_set_weights(bm_net *net, const float *w, size_t count)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!w) return nfail(net, BM_E_PARAM, "null weights");
    if (count != kParams) return nfail(net, BM_E_PARAM, "weight count must be " + std::to_string(kParams));
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(w[i])) return nfail(net, BM_E_PARAM, "weights must be finite");
    NET_HIP(net, hipStreamSynchronize(net->stream));
    NET_HIP(net, hipMemcpy(net->w, w, kParams * sizeof(float), hipMemcpyHostToDevice));
    NET_HIP(net, hipMemsetAsync(net->m, 0, kParams * sizeof(float), net->stream));
    NET_HIP(net, hipMemsetAsync(net->v, 0, kParams * sizeof(float), net->stream));
    net->step = 0;
    int r = repack(net);
    if (r) return r;
    NET_HIP(net, hipStreamSynchronize(net->stream));
    return BM_OK;
}

int bm_net_get_weights(bm_net *net, float *w, size_t count)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!w) return nfail(net, BM_E_PARAM, "null weights");
    if (count != kParams) return nfail(net, BM_E_PARAM, "weight count must be " + std::to_string(kParams));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    NET_HIP(net, hipMemcpy(w, net->w, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return BM_OK;
}

int bm_net_save(bm_net *net, const char *path)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!path) return nfail(net, BM_E_PARAM, "null path");
    std::vector<float> w(kParams);
    int r = bm_net_get_weights(net, w.data(), kParams);
    if (r) return r;
    FILE *f = std::fopen(path, "wb");
    if (!f) return nfail(net, BM_E_IO, std::string("cannot open ") + path);
    const uint32_t count = (uint32_t)kParams;
    const bool ok = std::fwrite("BMNET001", 1, 8, f) == 8 && std::fwrite(&count, 4, 1, f) == 1 &&
                    std::fwrite(w.data(), sizeof(float), kParams, f) == kParams;
    if (std::fclose(f) != 0 || !ok) return nfail(net, BM_E_IO, std::string("write failed: ") + path);
    return BM_OK;
}

int bm_net_load(bm_net *net, const char *path)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!path) return nfail(net, BM_E_PARAM, "null path");
    FILE *f = std::fopen(path, "rb");
    if (!f) return nfail(net, BM_E_IO, std::string("cannot open ") + path);
    char magic[8];
    uint32_t count = 0;
    std::vector<float> w(kParams);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "BMNET001", 8) == 0 &&
              std::fread(&count, 4, 1, f) == 1 && count == kParams &&
              std::fread(w.data(), sizeof(float), kParams, f) == kParams;
    char extra;
    ok = ok && std::fread(&extra, 1, 1, f) == 0;
    std::fclose(f);
    if (!ok) return nfail(net, BM_E_IO, std::string("not a ") + std::to_string(kParams) + "-parameter net file: " + path);
    return bm_net_set_weights(net, w.data(), kParams);
}

int bm_net_set_optimizer(bm_net *net, float beta1, float beta2, float eps)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    if (!(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f))
        return nfail(net, BM_E_PARAM, "Adam needs 0 <= beta < 1 and eps > 0");
    net->beta1 = beta1;
    net->beta2 = beta2;
    net->eps = eps;
    return BM_OK;
}

int bm_net_sync(bm_net *net)
{
    if (!net) return nfail(nullptr, BM_E_PARAM, "null net");
    NET_HIP(net, hipStreamSynchronize(net->stream));
    return BM_OK;
}

int bm_net_forward(bm_net *net, const uint16_t *feat, unsigned n, unsigned ch, unsigned cw, float *out)
{
    int r = check_dims(net, n, ch, cw);
    if (r) return r;
    if (!feat || !out) return nfail(net, BM_E_PARAM, "null feature / output pointer");
    r = ensure_buffers(net, (int)(n * ch * cw));
    if (r) return r;
    NET_HIP(net, forward(net->pk, feat, (int)n, (int)ch, (int)cw, net->act[0], net->act[1], out, net->stream));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    return BM_OK;
}

int bm_net_gradients(bm_net *net, const uint16_t *feat, const float *label, unsigned n, unsigned ch, unsigned cw,
                     float *loss, float *grad)
{
    int r = check_dims(net, n, ch, cw);
    if (r) return r;
    if (!feat || !label) return nfail(net, BM_E_PARAM, "null feature / label pointer");
    r = ensure_buffers(net, (int)(n * ch * cw));
    if (!r) r = forward_backward(net, feat, label, (int)n, (int)ch, (int)cw);
    if (r) return r;
    NET_HIP(net, hipStreamSynchronize(net->stream));
    if (loss) NET_HIP(net, hipMemcpy(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (grad) NET_HIP(net, hipMemcpy(grad, net->g, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return BM_OK;
}

int bm_net_train_step(bm_net *net, const uint16_t *feat_set, const float *label_set, unsigned rows, const int *idx,
                      unsigned batch, unsigned ch, unsigned cw, float lr, float *loss)
{
    int r = check_dims(net, batch, ch, cw);
    if (r) return r;
    if (!feat_set || !label_set || !idx) return nfail(net, BM_E_PARAM, "null dataset / index pointer");
    if (batch > (unsigned)kMaxBatch) return nfail(net, BM_E_PARAM, "batch larger than " + std::to_string(kMaxBatch));
    if (!(lr >= 0.0f && std::isfinite(lr))) return nfail(net, BM_E_PARAM, "learning rate must be finite and >= 0");
    RowList rl{};
    for (unsigned i = 0; i < batch; ++i) {
        if (idx[i] < 0 || (unsigned)idx[i] >= rows) return nfail(net, BM_E_PARAM, "row index outside the dataset");
        rl.idx[i] = idx[i];
    }
    const int npix = (int)(batch * ch * cw);
    r = ensure_buffers(net, npix);
    if (r) return r;
    const size_t cells = (size_t)ch * cw;
    NET_HIP(net, launch_gather_rows(feat_set, rl, cells * kFeatC * sizeof(uint16_t), (int)batch, net->x0, net->stream));
    NET_HIP(net, launch_gather_rows(label_set, rl, cells * 2 * sizeof(float), (int)batch, net->label, net->stream));
    r = forward_backward(net, net->x0, net->label, (int)batch, (int)ch, (int)cw);
    if (r) return r;
    ++net->step;
    const float c1 = (float)(1.0 - std::pow((double)net->beta1, (double)net->step));
    const float c2 = (float)(1.0 - std::pow((double)net->beta2, (double)net->step));
    NET_HIP(net, launch_adam(net->w, net->m, net->v, net->g, (int)kParams, lr, net->beta1, net->beta2, net->eps, c1, c2,
                             net->stream));
    r = repack(net);
    if (r) return r;
    if (loss) {
        NET_HIP(net, hipStreamSynchronize(net->stream));
        NET_HIP(net, hipMemcpy(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost));
    }
    return BM_OK;
}

}  // extern "C"
