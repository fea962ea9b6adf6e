// Device helpers shared by the flow and learned-proposal kernels: offsets,
// descriptors and work-group orderings.
#ifndef BM_DEVICE_H
#define BM_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// field offsets: dx in the low, dy in the high 16 bits (two's complement)
__device__ __forceinline__ uint32_t pack_off(int dx, int dy)
{
    return (uint32_t)(uint16_t)dx | ((uint32_t)(uint16_t)dy << 16);
}
__device__ __forceinline__ int off_dx(uint32_t o) { return (int)(int16_t)(o & 0xffffu); }
__device__ __forceinline__ int off_dy(uint32_t o) { return (int)(int16_t)(o >> 16); }

// descriptor idx of NSP 32-bit words (rows of NSP words, 16-byte aligned for NSP >= 4)
template <int NSP>
__device__ __forceinline__ void load_desc(const uint32_t *__restrict__ F, uint32_t idx, uint32_t (&w)[NSP])
{
    const uint32_t *p = F + (size_t)idx * NSP;
    if constexpr (NSP == 1) {
        w[0] = p[0];
    } else if constexpr (NSP == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        w[0] = v.x;
        w[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < NSP / 4; ++k) {
            const uint4 v = reinterpret_cast<const uint4 *>(p)[k];
            w[4 * k + 0] = v.x;
            w[4 * k + 1] = v.y;
            w[4 * k + 2] = v.z;
            w[4 * k + 3] = v.w;
        }
    }
}

template <int NSP>
__device__ __forceinline__ void store_desc(uint32_t *__restrict__ F, uint32_t idx, const uint32_t (&w)[NSP])
{
    uint32_t *p = F + (size_t)idx * NSP;
    if constexpr (NSP == 1) {
        p[0] = w[0];
    } else if constexpr (NSP == 2) {
        *reinterpret_cast<uint2 *>(p) = make_uint2(w[0], w[1]);
    } else {
#pragma unroll
        for (int k = 0; k < NSP / 4; ++k)
            reinterpret_cast<uint4 *>(p)[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
    }
}

template <int NSP>
__device__ __forceinline__ int hamming(const uint32_t (&a)[NSP], const uint32_t (&b)[NSP])
{
    int d = 0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) d += __popc(a[k] ^ b[k]);
    return d;
}

// Work-groups are handed to the 8 XCDs round-robin in dispatch order.  Give
// XCD k the k-th contiguous slice of the tile order instead, so that the tiles
// sharing one L2 are neighbours in the image.
__device__ __forceinline__ int xcd_remap(int b, int B)
{
    const int q = B >> 3, r = B & 7, k = b & 7;
    return k * q + min(k, r) + (b >> 3);
}

// Row-segment blocks (kThreads pixels of one row) visited strip by strip: within
// a strip of strip_rows rows the order runs down a column of blocks first, so the
// blocks resident on one XCD at any time cover a compact 2-D patch and the
// data-dependent descriptor gathers around it stay in that XCD's L2.
__device__ __forceinline__ void strip_block(int b, int gx, int rows_total, int strip_rows, int &bx, int &by)
{
    const int per = gx * strip_rows;
    const int strip = b / per;
    const int rows = min(strip_rows, rows_total - strip * strip_rows);
    const int rem = b - strip * per;
    bx = rem / rows;
    by = strip * strip_rows + (rem - bx * rows);
}

// Multi-pair launches: the tile a work-group takes in iteration `it` of a
// launch over T tiles with G work-groups, or -1 when it has none left.  With
// G >= T (one work-group per tile) work-group b takes tile xcd_remap(b, T);
// with fewer (G a multiple of 8, the work-groups stay resident) the
// work-groups of XCD k walk the k-th contiguous eighth of the tiles.
__device__ __forceinline__ int tile_of(int b, int G, int T, int it)
{
    if (G >= T) return (it == 0 && b < T) ? xcd_remap(b, T) : -1;
    const int k = b & 7, r = b >> 3, per = G >> 3;
    const int q = T >> 3, rem = T & 7;
    const int lo = k * q + min(k, rem), hi = lo + q + (k < rem ? 1 : 0);
    const int t = lo + r + it * per;
    return t < hi ? t : -1;
}

// The job owning tile t: start[j] <= t < start[j + 1] (start[0] = 0).
__device__ __forceinline__ int job_of(const int *__restrict__ start, int njobs, int t)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace bm

#endif
