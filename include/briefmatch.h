/*
 * briefmatch.h -- the C++ interface of the reference (briefmatch.h:65-155) on
 * top of the MI355X C ABI (briefmatch_amd.h).
 *
 * Same class, method names, arguments and exception type as the reference so
 * that a caller of BriefMatch switches by relinking:
 *   - match(im1, im2) takes DEVICE pointers to the up-sampled FP32 previous and
 *     current frames (sx x sy); only im2 is described, the previous frame's
 *     descriptors come from swap() -- except on the first call, where im1 is
 *     described too (the reference matched against uninitialised memory);
 *   - process(im1) takes the DEVICE pointer of the previous frame at input
 *     resolution (FP32, 0..255), the guide of the trilateral filter;
 *   - failures throw BriefMatchException, including every C-ABI error.
 * Additions: setSeed() (the reference seeded its pattern with time(NULL)),
 * flowU()/flowV() host access, and device flow pointers for zero-copy use.
 */
#ifndef BRIEF_MATCH_H
#define BRIEF_MATCH_H

#include <stdint.h>

#include <exception>
#include <string>
#include <vector>

#include "briefmatch_amd.h"

// Matching parameters (briefmatch.h:65-89), same fields and defaults.
struct BMParams {
    BMParams()
    {
        sx = sy = sx_in = sy_in = 64;
        iterations = 2;
        jfMax = 256;
        doubleProp = true;
        filter = true;
        tInit = true;
        sigma = 0.5f;
        P = 11;
        fsize = 13;
        N = 128;
        patchArea = 75;
        sEPE = 0.4f;
        sI = 0.1f;
        sS = 6.0f;
        maxmot = 5.0f;
    }

    unsigned int sx, sy, sx_in, sy_in, N, iterations, jfMax, P, fsize;
    float patchArea, sigma, sEPE, sI, sS, maxmot;
    bool doubleProp, filter, tInit;
};

class BriefMatchException : public std::exception {
    std::string msg;

public:
    BriefMatchException(const std::string &message) : msg(message) {}
    BriefMatchException(const char *message) : msg(message) {}
    ~BriefMatchException() throw() {}
    virtual const char *what() const throw() { return msg.c_str(); }
};

class BriefMatch {
public:
    BriefMatch();
    ~BriefMatch();

    void setup();
    void clean();

    void setParams(BMParams params) { m_params = params; }
    BMParams *getParams() { return &m_params; }

    void swap();
    void match(float *im1, float *im2);
    // No counterpart in the reference: the whole pair through the learned
    // proposal (bm_flow_pair_learned) from two device RGB frames at input
    // resolution, flow into the object's output buffers (no temporal
    // initialisation); convert() and the writers then work as after process().
    void matchLearned(bm_net *net, const uint8_t *rgb1, const uint8_t *rgb2);
    void downsample();
    void process(float *im1 = NULL);

    void convert();
    void print(std::string file);
    void computeColors();
    void printVis(std::string file);
    void printRaw(std::string file);
    float error(std::string filename);

    bool isSetup() { return m_setup; }
    void setVerbose(bool v) { m_verbose = v; }

    // additions
    void setSeed(uint64_t seed) { m_seed = seed; }
    const float *flowU() const { return m_U.data(); }
    const float *flowV() const { return m_V.data(); }
    float *deviceFlowU() { return m_outU; }
    float *deviceFlowV() { return m_outV; }
    bm_ctx *context() { return m_ctx; }

private:
    void check(int status);

    BMParams m_params;
    uint64_t m_seed;
    bool m_setup, m_verbose, m_firstFrame, m_haveF1;
    uint32_t m_pairIndex;
    bm_ctx *m_ctx;
    uint32_t *m_F[2];
    int m_f1;  // index of the previous frame's descriptors
    uint32_t *m_off[2];
    uint16_t *m_cost[2];
    int m_field;  // index of the latest field
    float *m_U0, *m_V0, *m_Uf, *m_Vf, *m_outU, *m_outV;
    std::vector<float> m_U, m_V;
    std::vector<uint8_t> m_vis;
};

#endif
