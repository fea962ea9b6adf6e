// OFPipeline (the reference's frame loop, of_pipeline.cpp) over the C ABI.
#include "../../include/of_pipeline.h"

#include <strings.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <utility>

#include "../../include/bm_io.h"

namespace {

struct Option {
    const char *name, *shortName, *info;
    enum Kind { STR, UINT, FLOAT, FLAG } kind;
    void *var;
    double lo, hi;
};

bool has_extension(const std::string &f, const char *ext)
{
    const size_t n = std::strlen(ext);
    return f.size() > n && strcasecmp(f.c_str() + f.size() - n, ext) == 0;
}

}  // namespace

OFPipeline::OFPipeline()
    : m_doWriteMOT(false), m_doWriteFLO(false), m_doWriteLDR(false), m_verbose(false), m_doPrintError(false),
      m_havePrev(false), m_endFrame(0), m_frame(0), m_stepFrame(1), m_startFrame(0), m_sx(0), m_sy(0), m_sx_in(0),
      m_sy_in(0), m_c(0), m_seed(0), m_upsampler("bilinear"), m_lsRadius(0), m_lsSamples(0), m_net(nullptr), m_bm(nullptr), m_bmP(nullptr), m_timerCtx(nullptr),
      dev_rgb(nullptr), dev_g8(nullptr), dev_rgbPrev(nullptr), m_cur(0)
{
    dev_Lin[0] = dev_Lin[1] = dev_L[0] = dev_L[1] = nullptr;
}

OFPipeline::~OFPipeline() { clear(); }

void OFPipeline::clear()
{
    void *bufs[] = {dev_rgb, dev_g8, dev_rgbPrev, dev_Lin[0], dev_Lin[1], dev_L[0], dev_L[1]};
    for (void *b : bufs)
        if (b) bm_free(b);
    dev_rgb = dev_g8 = dev_rgbPrev = nullptr;
    dev_Lin[0] = dev_Lin[1] = dev_L[0] = dev_L[1] = nullptr;
    delete m_bm;
    delete m_bmP;
    if (m_net) bm_net_destroy(m_net);
    m_net = nullptr;
    m_bm = nullptr;
    m_bmP = nullptr;
    m_timerCtx = nullptr;
}

namespace {

// A frame pattern is a printf format applied to the frame number alone: "%%"
// escapes and at most one integer conversion (flags, width, precision and an
// h or hh length allowed).  The frame number is passed as an unsigned int, so
// an l, ll, j, z or t length (or any other conversion) would make snprintf read
// an argument of a type that was never passed.
bool valid_frame_pattern(const std::string &s)
{
    int conversions = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] != '%') continue;
        if (++i < s.size() && s[i] == '%') continue;
        while (i < s.size() && std::strchr("-+ 0#", s[i])) ++i;
        while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        if (i < s.size() && s[i] == '.') {
            ++i;
            while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        }
        for (int h = 0; h < 2 && i < s.size() && s[i] == 'h'; ++h) ++i;
        if (i >= s.size() || !std::strchr("diuxXo", s[i])) return false;
        ++conversions;
    }
    return conversions <= 1;
}

}  // namespace

std::string OFPipeline::framePath(const std::string &pattern, unsigned int frame)
{
    char buf[4096];
    std::snprintf(buf, sizeof buf, pattern.c_str(), frame);
    return buf;
}

// start:step:end or start:end (of_pipeline.cpp:408-434)
bool OFPipeline::getFrameRange(std::string &frames, unsigned int &start, unsigned int &step, unsigned int &end)
{
    std::vector<long> v;
    const char *p = frames.c_str();
    while (true) {
        char *e = nullptr;
        errno = 0;
        const long x = std::strtol(p, &e, 10);
        if (e == p || errno || x < 0) return false;
        v.push_back(x);
        if (*e == '\0') break;
        if (*e != ':') return false;
        p = e + 1;
    }
    if (v.size() == 2) {
        start = (unsigned)v[0];
        end = (unsigned)v[1];
    } else if (v.size() == 3) {
        start = (unsigned)v[0];
        step = (unsigned)v[1];
        end = (unsigned)v[2];
    } else {
        return false;
    }
    return step > 0;
}

// Options and derivations of of_pipeline.cpp:90-195.
bool OFPipeline::setParams(int argc, char *argv[])
{
    m_bmP = new BMParams;
    std::string frames;
    float ux = 0, uy = 0, patchRad = 3.0f, sigmaRad = 0.025f, fsizeRad = 1.0f, sSRad = 0.85f;
    unsigned long long seed = 0;
    const Option opts[] = {
        {"--input", "-i", "Input sequence of frames (printf pattern with the frame number)", Option::STR, &m_input, 0,
         0},
        {"--output", "-o", "Output location of the estimated optical flow (.flo, .bin or .png)", Option::STR,
         &m_output, 0, 0},
        {"--frames", "-f", "Input frames, formatted as startframe:step:endframe", Option::STR, &frames, 0, 0},
        {"--vis-max-motion", "-mm", "Max-motion clamping in the colour coding of .png output", Option::FLOAT,
         &m_bmP->maxmot, 0, 1e10},
        {"--verbose", "-v", "Verbose mode", Option::FLAG, &m_verbose, 0, 0},
        {"--feature-length", "-fl", "Length of BRIEF binary feature vectors", Option::UINT, &m_bmP->N, 2, 512},
        {"--feature_length", "-fl", "(alias of --feature-length)", Option::UINT, &m_bmP->N, 2, 512},
        {"--patch-radius", "-pr", "BRIEF patch radius (percent of the up-sampled diagonal)", Option::FLOAT,
         &patchRad, 0, 100},
        {"--up-sampling-x", "-ux", "Up-sampling factor for image width", Option::FLOAT, &ux, 1, 100},
        {"--up-sampling-y", "-uy", "Up-sampling factor for image height", Option::FLOAT, &uy, 1, 100},
        {"--sigma-downsampling", "-ss", "Down-sampling filter size (percent of the diagonal)", Option::FLOAT,
         &sigmaRad, 0, 100},
        {"--iterations", "-it", "Number of iterations in neighbor search", Option::UINT, &m_bmP->iterations, 0,
         100},
        {"--jf-max", "-jm", "Max distance of the jump flooding scheme", Option::UINT, &m_bmP->jfMax, 2, 1e4},
        {"--median-size", "-ms", "Median filter size (percent of the diagonal)", Option::FLOAT, &fsizeRad, 0, 100},
        {"--range-epe", "-re", "Sigma of the flow (EPE) term of the trilateral filter", Option::FLOAT,
         &m_bmP->sEPE, 0, 1e10},
        {"--range-i", "-ri", "Sigma of the image term of the trilateral filter", Option::FLOAT, &m_bmP->sI, 0,
         1e10},
        {"--range-spatial", "-rs", "Spatial sigma (percent of the diagonal)", Option::FLOAT, &sSRad, 0, 100},
        {"--flow-gt", "-gt", "Ground truth optical flow (.flo) for an EPE print", Option::STR, &m_flowGT, 0, 0},
        {"--seed", "-sd", "Seed of the BRIEF pattern and the random search", Option::UINT, nullptr, 0, 0},
        {"--upsampler", "-us", "Up-sampling filter: bilinear (default) or lanczos3 (the reference's)", Option::STR,
         &m_upsampler, 0, 0},
        {"--proposal", "-pp", "Learned-proposal weights (.bnet): every pair through the proposal-seeded path",
         Option::STR, &m_proposal, 0, 0},
        {"--local-search", "-ls",
         "With --proposal: RADIUS:SAMPLES local search around the seeds (radius in input pixels)", Option::STR,
         &m_localSearch, 0, 0},
    };
    auto usage = [&]() {
        std::printf("briefmatch -- computes the optical flow of a sequence of frames on an AMD MI355X\n\n"
                    "Usage: briefmatch --input <frames> --frames <start:step:end> --output <output>\n\n");
        for (const Option &o : opts) std::printf("  %-22s %-4s %s\n", o.name, o.shortName, o.info);
        std::printf("\nExample: briefmatch -i data/RubberWhale/frame%%02d.png -f 7:14 -o output/mot_%%02d.flo\n");
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-help" || a == "--h" || a == "-h") {
            usage();
            return false;
        }
        const Option *o = nullptr;
        for (const Option &c : opts)
            if (a == c.name || a == c.shortName) {
                o = &c;
                break;
            }
        if (!o) throw BriefMatchException("'" + a + "' is not a valid input option. Use --help for the options.");
        if (o->kind == Option::FLAG) {
            *(bool *)o->var = true;
            continue;
        }
        if (++i >= argc) throw BriefMatchException("No value provided for input option '" + a + "'");
        const std::string v = argv[i];
        if (o->kind == Option::STR) {
            *(std::string *)o->var = v;
            continue;
        }
        char *e = nullptr;
        const double x = std::strtod(v.c_str(), &e);
        if (e == v.c_str() || *e) throw BriefMatchException("Argument '" + a + "' expects a number, got '" + v + "'");
        if (!o->var) {
            seed = (unsigned long long)std::strtoull(v.c_str(), nullptr, 10);
            continue;
        }
        if (x < o->lo || x > o->hi) {
            char m[256];
            std::snprintf(m, sizeof m, "Argument '%s' with value '%s' is out of range. Valid range is [%g, %g]",
                          a.c_str(), v.c_str(), o->lo, o->hi);
            throw BriefMatchException(m);
        }
        if (o->kind == Option::UINT)
            *(unsigned int *)o->var = (unsigned int)x;
        else
            *(float *)o->var = (float)x;
    }
    if (m_input.empty()) throw BriefMatchException("Required input option '--input' (-i) is missing.");
    for (const std::string *pat : {&m_input, &m_output})
        if (!valid_frame_pattern(*pat))
            throw BriefMatchException("Frame pattern '" + *pat +
                                      "' may hold one integer conversion for the frame number (e.g. frame%02d.png) "
                                      "and %% escapes only.");
    if (m_upsampler != "bilinear" && m_upsampler != "lanczos3")
        throw BriefMatchException("Argument '--upsampler' expects bilinear or lanczos3, got '" + m_upsampler + "'");
    if (!m_localSearch.empty()) {
        char *e = nullptr;
        const char *c = m_localSearch.c_str();
        const double r = std::strtod(c, &e);
        const bool sep = e != c && *e == ':';
        char *e2 = nullptr;
        const long n = sep ? std::strtol(e + 1, &e2, 10) : -1;
        if (!sep || e2 == e + 1 || *e2 || !(r > 0 && r <= 1024) || n < 1 || n > BM_LOCAL_SEARCH_MAX_SAMPLES)
            throw BriefMatchException("Argument '--local-search' expects RADIUS:SAMPLES with a radius in (0, 1024] "
                                      "input pixels and 1 to " + std::to_string(BM_LOCAL_SEARCH_MAX_SAMPLES) +
                                      " samples, got '" + m_localSearch + "'");
        if (m_proposal.empty()) throw BriefMatchException("Argument '--local-search' needs '--proposal'");
        m_lsRadius = (float)r;
        m_lsSamples = (int)n;
    }
    m_seed = seed;
    if (!m_output.empty()) {
        if (has_extension(m_output, ".flo"))
            m_doWriteFLO = true;
        else if (has_extension(m_output, ".bin"))
            m_doWriteMOT = true;
        else if (has_extension(m_output, ".png"))
            m_doWriteLDR = true;
        else
            throw BriefMatchException(
                "Unsupported output format. Supported formats are binary (.bin), Middlebury flow vector format "
                "(.flo) and color encoded images (.png).");
    }
    if (!frames.empty() && !getFrameRange(frames, m_frame, m_stepFrame, m_endFrame))
        throw BriefMatchException("Unable to parse frame range from '" + frames +
                                  "'. Valid format is startframe:step:endframe");
    if (m_endFrame < m_frame)
        throw BriefMatchException("Invalid frame range '" + frames + "'. End frame should be >= start frame");
    m_startFrame = m_frame;
    m_c = 0;

    int W = 0, H = 0;
    std::vector<uint8_t> rgb;
    try {
        bmio::read_image_rgb(framePath(m_input, m_frame), W, H, rgb);
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(std::string("Cannot read image file.\n\t") + e.what());
    }
    m_sx_in = (unsigned)W;
    m_sy_in = (unsigned)H;
    bm_params cp;
    if (bm_params_derive(m_sx_in, m_sy_in, ux, uy, patchRad, sigmaRad, fsizeRad, sSRad, &cp) != BM_OK)
        throw BriefMatchException(bm_last_error(nullptr));
    m_sx = cp.sx;
    m_sy = cp.sy;
    m_bmP->sx = cp.sx;
    m_bmP->sy = cp.sy;
    m_bmP->sx_in = cp.sx_in;
    m_bmP->sy_in = cp.sy_in;
    m_bmP->patchArea = cp.patchArea;
    m_bmP->sigma = cp.sigma;
    m_bmP->fsize = cp.fsize;
    m_bmP->sS = cp.sS;
    m_bmP->P = cp.P;
    m_doPrintError = m_flowGT.size() > 1;
    return true;
}

void OFPipeline::setup()
{
    std::fprintf(stderr, "Input: %s, %ux%u, frames %u:%u:%u, output %s\n", m_input.c_str(), m_sx_in, m_sy_in,
                 m_frame, m_stepFrame, m_endFrame, m_output.empty() ? "--" : m_output.c_str());
    m_bm = new BriefMatch;
    m_bm->setVerbose(m_verbose);
    m_bm->setParams(*m_bmP);
    m_bm->setSeed(m_seed);
    m_bm->setup();
    m_timerCtx = m_bm->context();
    if (!m_proposal.empty()) {
        if (bm_net_create(&m_net) != BM_OK) throw BriefMatchException(bm_last_error(nullptr));
        if (bm_net_load(m_net, m_proposal.c_str()) != BM_OK)
            throw BriefMatchException(std::string("Cannot load the proposal weights.\n\t") + bm_net_last_error(m_net));
        if (m_upsampler == "lanczos3" && bm_set_upsampler(m_timerCtx, BM_UPSAMPLE_LANCZOS3) != BM_OK)
            throw BriefMatchException(bm_last_error(m_timerCtx));
        if (m_lsSamples > 0) {
            // the radius in matched pixels: input pixels times the larger up-sampling factor
            const double f = std::max((double)m_sx / m_sx_in, (double)m_sy / m_sy_in);
            const int r = std::max(1, std::min(BM_LOCAL_SEARCH_MAX_RADIUS, (int)std::lround(m_lsRadius * f)));
            if (bm_set_local_search(m_timerCtx, r, m_lsSamples) != BM_OK)
                throw BriefMatchException(bm_last_error(m_timerCtx));
        }
    }
    const size_t Win = (size_t)m_sx_in * m_sy_in, M = (size_t)m_sx * m_sy;
    void *p = nullptr;
    auto alloc = [&](size_t bytes) {
        if (bm_malloc(&p, bytes) != BM_OK) throw BriefMatchException(bm_last_error(nullptr));
        return p;
    };
    dev_rgb = (uint8_t *)alloc(Win * 3);
    if (m_net) dev_rgbPrev = (uint8_t *)alloc(Win * 3);
    dev_g8 = (uint8_t *)alloc(Win);
    for (int k = 0; k < 2; ++k) {
        dev_Lin[k] = (float *)alloc(Win * 4);
        dev_L[k] = (float *)alloc(M * 4);
    }
    m_havePrev = false;
    m_cur = 0;
}

// Download the last pair's flow and write the requested outputs (of_pipeline.cpp:341-373).
void OFPipeline::writeOutputs()
{
    if (m_doWriteMOT || m_doWriteFLO || m_doWriteLDR || m_doPrintError) m_bm->convert();
    const std::string out = m_output.empty() ? std::string() : framePath(m_output, m_frame);
    if (m_doWriteMOT) m_bm->printRaw(out);
    if (m_doWriteFLO) m_bm->print(out);
    if (m_doWriteLDR) {
        m_bm->computeColors();
        m_bm->printVis(out);
    }
    if (m_doPrintError) std::printf("Error (EPE) = %f\n\n", m_bm->error(m_flowGT));
}

// One frame (of_pipeline.cpp:241-387): read, upload, gray + up-sample, match
// against the previous frame, down-sample, refine, swap, write.
bool OFPipeline::processFrame()
{
    if (m_frame > m_endFrame) return false;
    const std::string path = framePath(m_input, m_frame);
    if (!m_verbose) {
        const unsigned kv = m_endFrame > m_startFrame ? 100 * (m_frame - m_startFrame) / (m_endFrame - m_startFrame)
                                                      : 100;
        if (kv >= m_c) {
            m_c += 5;
            std::printf("\t%3u%% (frame %3u in range [%u,%u])\n", kv, m_frame, m_startFrame, m_endFrame);
        }
    } else {
        std::printf("Processing frame '%s'\n", path.c_str());
    }
    int W = 0, H = 0;
    std::vector<uint8_t> rgb;
    try {
        bmio::read_image_rgb(path, W, H, rgb);
    } catch (const bmio::IOError &e) {
        throw BriefMatchException(std::string("Cannot read image file.\n\t") + e.what());
    }
    if ((unsigned)W != m_sx_in || (unsigned)H != m_sy_in)
        throw BriefMatchException("Frame '" + path + "' has a different size than the first frame.");
    bm_ctx *ctx = m_timerCtx;
    auto ok = [&](int s) {
        if (s != BM_OK) {
            const char *m = bm_last_error(ctx);
            throw BriefMatchException(std::string(m && *m ? m : bm_last_error(nullptr)));
        }
    };
    ok(bm_sync(ctx));  // the upload buffer may still be read by the previous frame
    if (m_net) std::swap(dev_rgb, dev_rgbPrev);  // the last frame's RGB stays on the device
    ok(bm_memcpy(dev_rgb, rgb.data(), rgb.size(), BM_HOST_TO_DEVICE));
    ok(bm_timer_start(ctx));
    const int cur = m_cur, prev = 1 - m_cur;
    if (m_net) {
        // the learned path takes both RGB frames and does gray, up-sampling,
        // matching and refinement itself
        if (m_havePrev) {
            m_bm->matchLearned(m_net, dev_rgbPrev, dev_rgb);
            float ms = 0;
            ok(bm_timer_stop(ctx, &ms));
            if (m_verbose) std::printf("\tdevice time (learned path, whole pair): %.2f ms\n", ms);
            writeOutputs();
        } else {
            float ms = 0;
            ok(bm_timer_stop(ctx, &ms));
        }
        m_havePrev = true;
        m_frame += m_stepFrame;
        return true;
    }
    ok(bm_gray(ctx, dev_rgb, dev_g8, dev_Lin[cur]));
    if (m_upsampler == "lanczos3")
        ok(bm_upsample_lanczos(ctx, dev_g8, dev_L[cur]));
    else
        ok(bm_upsample(ctx, dev_g8, dev_L[cur]));
    bool wrote = false;
    if (m_havePrev) {
        m_bm->match(dev_L[prev], dev_L[cur]);
        m_bm->downsample();
        m_bm->process(dev_Lin[prev]);
        m_bm->swap();
        float ms = 0;
        ok(bm_timer_stop(ctx, &ms));
        if (m_verbose) std::printf("\tdevice time (gray, up-sample, match, refine): %.2f ms\n", ms);
        writeOutputs();
        wrote = true;
    } else {
        // the first frame has no predecessor: describe it on the next call
        ok(bm_sync(ctx));
    }
    (void)wrote;
    m_havePrev = true;
    m_cur = prev;
    m_frame += m_stepFrame;
    return true;
}
