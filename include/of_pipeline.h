/*
 * of_pipeline.h -- the reference's sequence driver (of_pipeline.h:66-97) on the
 * MI355X engine: command-line parameters, frame loop, up-sampling, matching,
 * refinement and output, with the same public methods.
 *
 * Command-line options are those of OFPipeline::setParams (of_pipeline.cpp:
 * 106-126), all always available (the reference hides some behind a build
 * flag), plus --feature_length (the README's spelling), --seed,
 * --upsampler (bilinear by default; lanczos3 is the reference's filter) and
 * --proposal (network weights: every pair through the learned proposal).
 * Frames are PNG/PGM/PPM; the first frame of a range produces no flow file
 * (the reference wrote flow against uninitialised memory for it).
 */
#ifndef OF_PIPELINE_H
#define OF_PIPELINE_H

#include <stdint.h>

#include <string>
#include <vector>

#include "briefmatch.h"

class OFPipeline {
public:
    OFPipeline();
    ~OFPipeline();
    void clear();
    bool setParams(int argc, char *argv[]);
    void setup();
    bool processFrame();

private:
    bool getFrameRange(std::string &frames, unsigned int &start, unsigned int &step, unsigned int &end);
    std::string framePath(const std::string &pattern, unsigned int frame);
    void writeOutputs();

    bool m_doWriteMOT, m_doWriteFLO, m_doWriteLDR, m_verbose, m_doPrintError, m_havePrev;
    unsigned int m_endFrame, m_frame, m_stepFrame, m_startFrame, m_sx, m_sy, m_sx_in, m_sy_in, m_c;
    uint64_t m_seed;
    std::string m_upsampler;  // "bilinear" or "lanczos3"
    std::string m_proposal;   // weights file of the learned proposal, or empty
    std::string m_localSearch;  // --local-search RADIUS:SAMPLES (learned path)
    float m_lsRadius;         // its radius in input pixels
    int m_lsSamples;          // its samples (0: off)
    bm_net *m_net;
    std::string m_input, m_flowGT, m_output;

    BriefMatch *m_bm;
    BMParams *m_bmP;
    bm_ctx *m_timerCtx;

    // device buffers: RGB upload, gray8 scratch, FP32 gray (input res) and
    // up-sampled frame, current / previous
    uint8_t *dev_rgb, *dev_g8;
    uint8_t *dev_rgbPrev;  // with --proposal: the previous frame's RGB
    float *dev_Lin[2], *dev_L[2];
    int m_cur;
};

#endif
