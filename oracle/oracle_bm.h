/*
 * CPU oracle for the BriefMatch flow path -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C++ (+ OpenMP over rows) restatement of the reference kernels
 * kernels/briefmatch_cuda.cu:44-297 and kernels/util_cuda.cu:48-183, of the
 * host-side set-up briefmatch.cpp:72-214 / :248-352 and of the parameter
 * derivation of_pipeline.cpp:178-190.  Where the reference is racy or
 * non-deterministic (time-seeded pattern, cuRAND stream, in-place propagation,
 * out-of-bounds median reads) it follows the deterministic rules written next
 * to each function below; the GPU library implements the same rules, so the
 * integer match field must agree bit for bit and the FP32 flow to 1e-3 px.
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this library, and only as the checker.  The product never links it.
 */
#ifndef ORACLE_BM_H
#define ORACLE_BM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields and meaning as BMParams (briefmatch.h:65-89). */
typedef struct {
    unsigned sx, sy, sx_in, sy_in, N, iterations, jfMax, P, fsize;
    float patchArea, sigma, sEPE, sI, sS, maxmot;
    int doubleProp, filter, tInit;
} orc_params;

int  orc_version(void);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int  orc_derive_params(unsigned W, unsigned H, float ux, float uy, float patch_radius,
                       float sigma_ds, float median_size, float range_spatial,
                       orc_params *out);
void orc_make_pattern(uint64_t seed, unsigned N, float patchArea, int16_t *pattern);
int  orc_gauss_kernel(float sigma, float *kernel, unsigned *P);
float orc_rs_sigma(unsigned sx);
int64_t orc_rs_sigma_q16(unsigned sx);

void orc_gray(const uint8_t *rgb, int W, int H, uint8_t *gray);
void orc_upsample(const uint8_t *gray, int W, int H, int sx, int sy, float *L);
/* Lanczos-3 (util_cuda.cu:126-158) with separable normalised float tables:
 * first[x] = floor(x * float(in) / out) - 2, w[6x + k] the weight of tap k. */
void orc_lanczos_table(int in, int out, int *first, float *w);
void orc_upsample_lanczos(const uint8_t *gray, int W, int H, int sx, int sy, float *L);
void orc_brief(const float *L, int sx, int sy, const int16_t *pattern, int N, uint32_t *F);
void orc_init(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS, int first,
              const uint32_t *prev_off, uint32_t *off, uint16_t *cost);
void orc_propagate_level(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                         int p, int pd, int doubleProp,
                         const uint32_t *off_in, const uint16_t *cost_in,
                         uint32_t *off_out, uint16_t *cost_out);
void orc_propagate(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                   int pd, int jfMax, int doubleProp, uint32_t *off, uint16_t *cost);
void orc_random_search(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
                       int64_t sigma_q16, uint32_t iteration, uint32_t pair, uint64_t seed,
                       uint32_t *off, uint16_t *cost);
void orc_match(const uint32_t *F1, const uint32_t *F2, int sx, int sy, int NS,
               int iterations, int jfMax, int doubleProp, int first, const uint32_t *prev_off,
               uint32_t pair, uint64_t seed, uint32_t *off, uint16_t *cost);
void orc_downsample(const uint32_t *off, int sx, int sy, int W, int H,
                    const float *kernel, int P, float *U, float *V);
void orc_median1d(const float *in, float *out, int W, int H, int dir, int fsize);
void orc_trilateral(const float *Ui, const float *Vi, const float *Um, const float *Vm,
                    const float *I, int W, int H, float sigmaEPE, float sigmaI, float sigmaS,
                    int row0, int row1, float *Uo, float *Vo);
int  orc_flow_pair(const orc_params *p, uint64_t seed, uint32_t pair,
                   const uint8_t *rgb1, const uint8_t *rgb2,
                   uint32_t *off_out, uint16_t *cost_out, float *U, float *V);
/* as orc_flow_pair; lanczos != 0 up-samples with orc_upsample_lanczos */
int  orc_flow_pair_up(const orc_params *p, uint64_t seed, uint32_t pair,
                      const uint8_t *rgb1, const uint8_t *rgb2, int lanczos,
                      uint32_t *off_out, uint16_t *cost_out, float *U, float *V);
void orc_epe(const float *U, const float *V, const float *gtU, const float *gtV,
             const uint8_t *occluded, int64_t n, double *epe_all, double *epe_noc);
int  orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
