/*
 * bm_io.h -- host-side file formats of the BriefMatch tools.
 *
 * Replaces the reference's util/flowIO (Middlebury .flo, flowIO.cpp:35-133),
 * util/colorcode (Middlebury colour wheel, colorcode.cpp:26-74), the raw .bin
 * writer (BriefMatch::printRaw, briefmatch.cpp:429-440) and the image input
 * that the reference takes from OpenCV / imageLib (of_pipeline.cpp:157-176,
 * 266-292).  Errors throw bmio::IOError.
 */
#ifndef BM_IO_H
#define BM_IO_H

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace bmio {

struct IOError : public std::runtime_error {
    explicit IOError(const std::string &m) : std::runtime_error(m) {}
};

/* |u| or |v| above 1e9, or NaN: ground truth unknown (flowIO.h:5, flowIO.cpp:35-39). */
bool unknown_flow(float u, float v);

/* Middlebury .flo: "PIEH", int32 width, int32 height, then interleaved (u, v)
 * float32 rows (flowIO.cpp:5-20). */
void write_flo(const std::string &path, int W, int H, const float *U, const float *V);
void read_flo(const std::string &path, int &W, int &H, std::vector<float> &U, std::vector<float> &V);

/* Raw float32: all W*H u values, then all v values, no header. */
void write_raw(const std::string &path, int W, int H, const float *U, const float *V);

/* Middlebury colour coding: hue = direction, saturation = magnitude / maxrad;
 * maxmot > 0 fixes maxrad (BriefMatch::computeColors, briefmatch.cpp:372-418).
 * rgb receives W*H*3 bytes (R, G, B). */
void flow_to_rgb(int W, int H, const float *U, const float *V, float maxmot, uint8_t *rgb);

/* 8-bit RGB image: PNG (8/16-bit gray, gray+alpha, RGB, RGBA, palette;
 * non-interlaced) or binary PGM/PPM.  Gray inputs are replicated to R = G = B. */
void read_image_rgb(const std::string &path, int &W, int &H, std::vector<uint8_t> &rgb);
void write_png_rgb(const std::string &path, int W, int H, const uint8_t *rgb);

/* Mean end-point error over pixels with known ground truth (and, for noc,
 * not flagged in occluded), accumulated in double. */
void epe(const float *U, const float *V, const float *gtU, const float *gtV, const uint8_t *occluded, size_t n,
         double &epe_all, double &epe_noc);

}  // namespace bmio

/* C entry points of the same functions (libbriefmatch_cpp.so), status as in
 * briefmatch_amd.h; bm_io_last_error() gives the message of the last failure. */
extern "C" {
const char *bm_io_last_error(void);
int bm_write_flo(const char *path, int W, int H, const float *U, const float *V);
/* Call with U = V = NULL to query W and H, then again with buffers of W*H floats. */
int bm_read_flo(const char *path, int *W, int *H, float *U, float *V);
int bm_write_raw(const char *path, int W, int H, const float *U, const float *V);
int bm_flow_to_rgb(int W, int H, const float *U, const float *V, float maxmot, uint8_t *rgb);
/* Call with rgb = NULL to query W and H, then again with a W*H*3 buffer. */
int bm_read_image_rgb(const char *path, int *W, int *H, uint8_t *rgb);
int bm_write_png_rgb(const char *path, int W, int H, const uint8_t *rgb);
}

#endif
