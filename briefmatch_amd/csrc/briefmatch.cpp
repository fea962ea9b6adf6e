// BriefMatch (the reference's algorithm object, briefmatch.cpp) over the C ABI.
#include "../../include/briefmatch.h"

#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/bm_io.h"

namespace {

bm_params to_c(const BMParams &p)
{
    bm_params c;
    bm_params_default(&c);
    c.sx = p.sx;
    c.sy = p.sy;
    c.sx_in = p.sx_in;
    c.sy_in = p.sy_in;
    c.N = p.N;
    c.iterations = p.iterations;
    c.jfMax = p.jfMax;
    c.P = p.P;
    c.fsize = p.fsize;
    c.patchArea = p.patchArea;
    c.sigma = p.sigma;
    c.sEPE = p.sEPE;
    c.sI = p.sI;
    c.sS = p.sS;
    c.maxmot = p.maxmot;
    c.doubleProp = p.doubleProp ? 1 : 0;
    c.filter = p.filter ? 1 : 0;
    c.tInit = p.tInit ? 1 : 0;
    return c;
}

}  // namespace

BriefMatch::BriefMatch()
    : m_seed(0), m_setup(false), m_verbose(false), m_firstFrame(true), m_haveF1(false), m_pairIndex(0),
      m_ctx(nullptr), m_f1(0), m_field(0), m_U0(nullptr), m_V0(nullptr), m_Uf(nullptr), m_Vf(nullptr),
      m_outU(nullptr), m_outV(nullptr)
{
    m_F[0] = m_F[1] = nullptr;
    m_off[0] = m_off[1] = nullptr;
    m_cost[0] = m_cost[1] = nullptr;
}

BriefMatch::~BriefMatch() { clean(); }

void BriefMatch::check(int status)
{
    if (status == BM_OK) return;
    const char *m = bm_last_error(m_ctx);
    if (!m || !*m) m = bm_last_error(nullptr);
    throw BriefMatchException(std::string(m ? m : "unknown error"));
}

// Limits, pattern, Gaussian kernel and buffers (briefmatch.cpp:72-214).
void BriefMatch::setup()
{
    if (m_setup) return;
    bm_params cp = to_c(m_params);
    bm_ctx *ctx = nullptr;
    if (bm_ctx_create(&cp, m_seed, &ctx) != BM_OK) throw BriefMatchException(std::string(bm_last_error(nullptr)));
    m_ctx = ctx;
    bm_ctx_get_params(m_ctx, &cp);
    m_params.P = cp.P;

    const float h = std::sqrt((float)(m_params.sx * m_params.sx + m_params.sy * m_params.sy));
    const float h_in = std::sqrt((float)(m_params.sx_in * m_params.sx_in + m_params.sy_in * m_params.sy_in));
    std::fprintf(stderr, "BriefMatch (MI355X) settings:\n");
    std::fprintf(stderr, "  BRIEF: %u bits, patch radius %.1f px (%.4f%% of the up-sampled diagonal)\n", m_params.N,
                 m_params.patchArea, 100 * m_params.patchArea / h);
    std::fprintf(stderr, "  up-sampling: %ux%u -> %ux%u (x%.4f, x%.4f), down-sampling sigma %.4f px\n",
                 m_params.sx_in, m_params.sy_in, m_params.sx, m_params.sy, (float)m_params.sx / m_params.sx_in,
                 (float)m_params.sy / m_params.sy_in, m_params.sigma);
    std::fprintf(stderr, "  search: %u iterations, jump flooding from %u, temporal init %s, pattern seed %llu\n",
                 m_params.iterations, m_params.jfMax, m_params.tInit ? "yes" : "no", (unsigned long long)m_seed);
    std::fprintf(stderr, "  refinement: median %u px (%.4f%%), sigma EPE %.4f, image %.4f, spatial %.4f px\n",
                 m_params.fsize, 100 * m_params.fsize / h_in, m_params.sEPE, m_params.sI, m_params.sS);

    const size_t M = (size_t)m_params.sx * m_params.sy, Win = (size_t)m_params.sx_in * m_params.sy_in;
    const size_t stride = bm_descriptor_stride(m_ctx);
    void *p = nullptr;
    for (int k = 0; k < 2; ++k) {
        check(bm_malloc(&p, M * stride * 4));
        m_F[k] = (uint32_t *)p;
        check(bm_malloc(&p, M * 4));
        m_off[k] = (uint32_t *)p;
        check(bm_malloc(&p, M * 2));
        m_cost[k] = (uint16_t *)p;
    }
    float **bufs[4] = {&m_U0, &m_V0, &m_Uf, &m_Vf};
    for (float **b : bufs) {
        check(bm_malloc(&p, Win * 4));
        *b = (float *)p;
    }
    m_outU = m_U0;
    m_outV = m_V0;
    m_U.assign(Win, 0.0f);
    m_V.assign(Win, 0.0f);
    m_firstFrame = true;
    m_haveF1 = false;
    m_pairIndex = 0;
    m_setup = true;
}

void BriefMatch::clean()
{
    void *bufs[] = {m_F[0], m_F[1], m_off[0], m_off[1], m_cost[0], m_cost[1], m_U0, m_V0, m_Uf, m_Vf};
    for (void *b : bufs)
        if (b) bm_free(b);
    m_F[0] = m_F[1] = nullptr;
    m_off[0] = m_off[1] = nullptr;
    m_cost[0] = m_cost[1] = nullptr;
    m_U0 = m_V0 = m_Uf = m_Vf = m_outU = m_outV = nullptr;
    if (m_ctx) bm_ctx_destroy(m_ctx);
    m_ctx = nullptr;
    m_setup = false;
}

// The current frame's descriptors become the previous ones (briefmatch.cpp:240-245).
void BriefMatch::swap() { m_f1 = 1 - m_f1; }

// briefmatch.cpp:248-293: describe im2, initialise (zero or temporal), then
// iterations x {propagate, random search}.
void BriefMatch::match(float *im1, float *im2)
{
    if (!m_setup) throw BriefMatchException("BriefMatch is not initialized before use.");
    if (!im2 || (!m_haveF1 && !im1)) throw BriefMatchException("null device image passed to BriefMatch::match");
    const int f2 = 1 - m_f1;
    check(bm_brief(m_ctx, im2, m_F[f2]));
    if (!m_haveF1) {
        check(bm_brief(m_ctx, im1, m_F[m_f1]));
        m_haveF1 = true;
    }
    const bool first = m_firstFrame || !m_params.tInit;
    const int dst = 1 - m_field;
    check(bm_match(m_ctx, m_F[m_f1], m_F[f2], first ? 1 : 0, first ? nullptr : m_off[m_field], m_pairIndex,
                   m_off[dst], m_cost[dst]));
    m_field = dst;
    m_firstFrame = false;
    ++m_pairIndex;
}

void BriefMatch::matchLearned(bm_net *net, const uint8_t *rgb1, const uint8_t *rgb2)
{
    if (!m_setup) throw BriefMatchException("BriefMatch is not initialized before use.");
    if (!net || !rgb1 || !rgb2) throw BriefMatchException("null argument passed to BriefMatch::matchLearned");
    check(bm_flow_pair_learned(m_ctx, net, rgb1, rgb2, m_pairIndex, m_Uf, m_Vf));
    m_outU = m_Uf;
    m_outV = m_Vf;
    m_haveF1 = false;  // the descriptors of the stage path no longer describe the previous frame
    m_firstFrame = true;
    ++m_pairIndex;
}

// briefmatch.cpp:296-324
void BriefMatch::downsample()
{
    if (!m_setup) throw BriefMatchException("BriefMatch is not initialized before use.");
    check(bm_downsample(m_ctx, m_off[m_field], m_U0, m_V0));
    m_outU = m_U0;
    m_outV = m_V0;
}

// briefmatch.cpp:327-352
void BriefMatch::process(float *im1)
{
    if (!m_setup) throw BriefMatchException("BriefMatch is not initialized before use.");
    if (!m_params.filter) {
        m_outU = m_U0;
        m_outV = m_V0;
        return;
    }
    if (!im1) throw BriefMatchException("BriefMatch::process needs the previous frame (device, input resolution)");
    check(bm_refine(m_ctx, m_U0, m_V0, im1, m_Uf, m_Vf));
    m_outU = m_Uf;
    m_outV = m_Vf;
}

// Download (briefmatch.cpp:355-370).
void BriefMatch::convert()
{
    if (!m_setup) throw BriefMatchException("BriefMatch is not initialized before use.");
    check(bm_sync(m_ctx));
    const size_t Win = (size_t)m_params.sx_in * m_params.sy_in;
    check(bm_memcpy(m_U.data(), m_outU, Win * 4, BM_DEVICE_TO_HOST));
    check(bm_memcpy(m_V.data(), m_outV, Win * 4, BM_DEVICE_TO_HOST));
    m_vis.clear();
}

void BriefMatch::computeColors()
{
    m_vis.resize((size_t)m_params.sx_in * m_params.sy_in * 3);
    bmio::flow_to_rgb((int)m_params.sx_in, (int)m_params.sy_in, m_U.data(), m_V.data(), m_params.maxmot,
                      m_vis.data());
}

void BriefMatch::printVis(std::string file)
{
    try {
        if (m_vis.empty()) computeColors();
        bmio::write_png_rgb(file, (int)m_params.sx_in, (int)m_params.sy_in, m_vis.data());
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(std::string("Cannot write to image file.\n\t") + e.what());
    }
}

void BriefMatch::printRaw(std::string file)
{
    try {
        bmio::write_raw(file, (int)m_params.sx_in, (int)m_params.sy_in, m_U.data(), m_V.data());
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(e.what());
    }
}

void BriefMatch::print(std::string file)
{
    try {
        bmio::write_flo(file, (int)m_params.sx_in, (int)m_params.sy_in, m_U.data(), m_V.data());
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(std::string("Cannot write .flo file.\n\t") + e.what());
    }
}

// Mean EPE against a .flo ground truth (briefmatch.cpp:468-514).
float BriefMatch::error(std::string file)
{
    try {
        int W = 0, H = 0;
        std::vector<float> gu, gv;
        bmio::read_flo(file, W, H, gu, gv);
        if ((unsigned)W != m_params.sx_in || (unsigned)H != m_params.sy_in)
            throw bmio::IOError("ground truth is " + std::to_string(W) + "x" + std::to_string(H) +
                                ", flow is " + std::to_string(m_params.sx_in) + "x" + std::to_string(m_params.sy_in));
        double all = 0, noc = 0;
        bmio::epe(m_U.data(), m_V.data(), gu.data(), gv.data(), nullptr, gu.size(), all, noc);
        return (float)all;
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(std::string("Failure in flow error computation:\n\t") + e.what());
    }
}
