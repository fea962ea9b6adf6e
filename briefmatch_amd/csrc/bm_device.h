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

// Multi-pair launches.  Job j owns the work-group slots [start[j], start[j + 1])
// of the launch, a multiple of 8 slots (start[] is padded so), of which the
// first count[j] = start[njobs + 1 + j] are its tiles; slot s of the job is
// taken by a work-group on XCD s mod 8, and takes the job's tile
// xcd_remap(s, count[j]), the tile the same work-group would take in the
// job's own launch: every pair is spread over all XCDs, each XCD holding a
// contiguous part of it.  With fewer work-groups than slots (a multiple of 8,
// resident), work-group b walks the slots b, b + G, b + 2G, ... (all on its XCD).
// Returns the job and sets *tile, or -1 for an idle slot.
__device__ __forceinline__ int job_of(const int *__restrict__ start, int njobs, int slot, int *tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= slot)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int s = slot - start[lo], count = start[njobs + 1 + lo];
    if (s >= count) return -1;
    *tile = xcd_remap(s, count);
    return lo;
}

}  // namespace bm

#endif
