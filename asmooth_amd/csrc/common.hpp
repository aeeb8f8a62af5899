// Shared definitions for the gfx950 smoothing kernels and their host launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace asm_hip {

constexpr int kWave = 64;       // CDNA wavefront width
constexpr int kSlots = 33;      // neighbour slots incl. the particle itself (k_neighbors.h:47)

// One tree node as read by a traversing wavefront: 32 bytes, so a node is one
// scalar dwordx8 load and the two children of node i (2i, 2i+1) are 64
// contiguous bytes.  `lo_w` / `hi_w` carry per-node payload: for leaves the
// first sorted particle index and the count.
struct __align__(16) Box {
    float lo[3];
    int lo_w;
    float hi[3];
    int hi_w;
};

// Particles per group box: the tree also keeps the bounding box of every 8
// consecutive sorted particles, so a scan of an opened leaf can skip groups
// that no lane's current bound reaches.
constexpr int kGroup = 8;

// Key ordering neighbour candidates canonically: squared distance first (the
// bit pattern of a non-negative float orders like the value), then the
// particle's original index.
__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx)
{
    return (static_cast<uint64_t>(__float_as_uint(d2)) << 32) | idx;
}
__device__ __forceinline__ float key_d2(uint64_t k)
{
    return __uint_as_float(static_cast<uint32_t>(k >> 32));
}
__device__ __forceinline__ uint32_t key_idx(uint64_t k)
{
    return static_cast<uint32_t>(k & 0xffffffffu);
}
// +inf distance, largest index: sorts after every real candidate.
constexpr uint64_t kEmptyKey = (static_cast<uint64_t>(0x7f800000u) << 32) | 0xffffffffu;

// Squared distance exactly as nbr_search.cpp:19-25 evaluates it in FP32:
// dx = x - y per axis, then (dx*dx + dy*dy) + dz*dz with every operation
// rounded (the library is compiled with -ffp-contract=off).
__device__ __forceinline__ float norm2(float x0, float x1, float x2, float y0, float y1, float y2)
{
    const float dx = x0 - y0;
    const float dy = x1 - y1;
    const float dz = x2 - y2;
    float s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// Conservative lower bound of norm2(q, p) over all p inside the box, using the
// same operation order: every rounding step is monotone, so the bound never
// exceeds the FP32 distance of any point in the box.
__device__ __forceinline__ float box_d2(const Box& b, float q0, float q1, float q2)
{
    const float g0 = fmaxf(fmaxf(b.lo[0] - q0, q0 - b.hi[0]), 0.0f);
    const float g1 = fmaxf(fmaxf(b.lo[1] - q1, q1 - b.hi[1]), 0.0f);
    const float g2 = fmaxf(fmaxf(b.lo[2] - q2, q2 - b.hi[2]), 0.0f);
    float s = g0 * g0;
    s = s + g1 * g1;
    s = s + g2 * g2;
    return s;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): workgroup w runs on XCD w % 8.  Returns the logical block for w such
// that each XCD takes one contiguous range of logical blocks, so blocks that
// are neighbours in Morton (or tile) order share an L2 instead of fetching
// the same particles once per XCD.  A bijection of [0, nblocks).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t w, uint32_t nblocks)
{
    const uint32_t per = nblocks / kXcds, rem = nblocks % kXcds;
    const uint32_t x = w % kXcds, k = w / kXcds;
    return x * per + (x < rem ? x : rem) + k;
}

// Wavefront-wide helpers.
__device__ __forceinline__ uint64_t wave_ballot(bool p)
{
    return __builtin_amdgcn_ballot_w64(p);
}
__device__ __forceinline__ bool wave_any(bool p)
{
    return __builtin_amdgcn_ballot_w64(p) != 0ull;
}
__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

}  // namespace asm_hip

#define ASM_HIP_CHECK(expr)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return _e;                      \
    } while (0)
