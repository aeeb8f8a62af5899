// GPU tree build.
//
// Replaces KDTreeSimple::construct (kdtree_balanced.cpp:212-296).  Particles
// are put in Morton order (9 bits per axis over the root cube the reference
// passes, clumping_tree_main.cpp:392-394) by a stable radix sort, then cut into
// leaves of `leaf_size` consecutive particles.  Over the leaves sits an
// implicit complete binary tree (node i has children 2i and 2i+1, leaves at
// [nleaf_pad, 2 nleaf_pad)), every split being the median of its range in
// Morton order, so the tree is perfectly balanced and needs no pointers.
// Node boxes are the exact bounding boxes of the particles below them.
// Exact kNN makes the neighbour result independent of the tree shape.
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

// Spread the low 10 bits of v so that bit i moves to bit 3i (keys use the
// low kMortonAxisBits).
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void morton_keys_kernel(const ParticleRec* __restrict__ p, int64_t n, float left,
                                   float scale, uint32_t* __restrict__ keys,
                                   uint32_t* __restrict__ vals)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float f = (p[i].x[a] - left) * scale;
        f = fminf(fmaxf(f, 0.0f), static_cast<float>((1 << kMortonAxisBits) - 1));
        c[a] = static_cast<uint32_t>(f);
    }
    keys[i] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
    vals[i] = static_cast<uint32_t>(i);
}

__global__ void gather_sorted_kernel(const ParticleRec* __restrict__ p,
                                     const uint32_t* __restrict__ perm, int64_t n,
                                     float4* __restrict__ spos, float4* __restrict__ sattr,
                                     int32_t* __restrict__ inv_perm)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = perm[i];
    const ParticleRec r = p[o];
    spos[i] = make_float4(r.x[0], r.x[1], r.x[2], __uint_as_float(o));
    sattr[i] = make_float4(r.v[0], r.v[1], r.v[2], __int_as_float(r.igrp));
    inv_perm[o] = static_cast<int32_t>(i);
}

__global__ void tree_leaves_kernel(const float4* __restrict__ spos, int64_t n, int leaf_size,
                                   int64_t nleaf, int64_t nleaf_pad, const Box* __restrict__ groups,
                                   Box* __restrict__ nodes)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nleaf_pad) return;
    Box b;
    if (j < nleaf) {
        const int64_t start = j * leaf_size;
        const int64_t end = (start + leaf_size < n) ? start + leaf_size : n;
        float lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY;
        float hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
        if (groups) {
            // The leaf is whole groups (leaf_size is a multiple of kGroup):
            // the union of their boxes, read as consecutive 32-B records.
            for (int64_t g = start / kGroup; g <= (end - 1) / kGroup; ++g) {
                const Box q = groups[g];
                lo0 = fminf(lo0, q.lo[0]); hi0 = fmaxf(hi0, q.hi[0]);
                lo1 = fminf(lo1, q.lo[1]); hi1 = fmaxf(hi1, q.hi[1]);
                lo2 = fminf(lo2, q.lo[2]); hi2 = fmaxf(hi2, q.hi[2]);
            }
        } else {
            for (int64_t i = start; i < end; ++i) {
                const float4 q = spos[i];
                lo0 = fminf(lo0, q.x); hi0 = fmaxf(hi0, q.x);
                lo1 = fminf(lo1, q.y); hi1 = fmaxf(hi1, q.y);
                lo2 = fminf(lo2, q.z); hi2 = fmaxf(hi2, q.z);
            }
        }
        b.lo[0] = lo0; b.lo[1] = lo1; b.lo[2] = lo2;
        b.hi[0] = hi0; b.hi[1] = hi1; b.hi[2] = hi2;
        b.lo_w = static_cast<int>(start);
        b.hi_w = static_cast<int>(end - start);
    } else {
        b.lo[0] = b.lo[1] = b.lo[2] = INFINITY;
        b.hi[0] = b.hi[1] = b.hi[2] = -INFINITY;
        b.lo_w = static_cast<int>(n);
        b.hi_w = 0;
    }
    nodes[nleaf_pad + j] = b;
}

// Box of each kGroup consecutive sorted particles: one thread per particle,
// reduced over the kGroup lanes of a group with cross-lane exchanges.
__global__ void tree_groups_kernel(const float4* __restrict__ spos, int64_t n, Box* __restrict__ groups)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ngroups = (n + kGroup - 1) / kGroup;
    if (i >= ngroups * kGroup) return;   // whole groups only: every lane of a group is present
    const bool have = i < n;
    const float4 q = spos[have ? i : n - 1];
    float lo0 = have ? q.x : INFINITY, lo1 = have ? q.y : INFINITY, lo2 = have ? q.z : INFINITY;
    float hi0 = have ? q.x : -INFINITY, hi1 = have ? q.y : -INFINITY, hi2 = have ? q.z : -INFINITY;
#pragma unroll
    for (int o = 1; o < kGroup; o <<= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, o)); hi0 = fmaxf(hi0, __shfl_xor(hi0, o));
        lo1 = fminf(lo1, __shfl_xor(lo1, o)); hi1 = fmaxf(hi1, __shfl_xor(hi1, o));
        lo2 = fminf(lo2, __shfl_xor(lo2, o)); hi2 = fmaxf(hi2, __shfl_xor(hi2, o));
    }
    if ((i & (kGroup - 1)) == 0) {
        Box b;
        b.lo[0] = lo0; b.lo[1] = lo1; b.lo[2] = lo2;
        b.hi[0] = hi0; b.hi[1] = hi1; b.hi[2] = hi2;
        b.lo_w = static_cast<int>(i);
        b.hi_w = static_cast<int>((n - i < kGroup) ? n - i : kGroup);
        groups[i / kGroup] = b;
    }
}

__global__ void tree_level_kernel(Box* __restrict__ nodes, int64_t first, int64_t count)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t i = first + t;
    const Box a = nodes[2 * i];
    const Box c = nodes[2 * i + 1];
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = fminf(a.lo[k], c.lo[k]);
        b.hi[k] = fmaxf(a.hi[k], c.hi[k]);
    }
    b.lo_w = a.lo_w;
    b.hi_w = a.hi_w + c.hi_w;
    nodes[i] = b;
}

}  // namespace

hipError_t launch_morton_keys(const ParticleRec* p, int64_t n, float left, float right,
                              uint32_t* keys, uint32_t* vals, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const float scale = static_cast<float>(1 << kMortonAxisBits) / (right - left);
    const int th = 256;
    morton_keys_kernel<<<static_cast<unsigned>((n + th - 1) / th), th, 0, s>>>(p, n, left, scale, keys, vals);
    return hipGetLastError();
}

hipError_t launch_gather_sorted(const ParticleRec* p, const uint32_t* perm, int64_t n,
                                float4* spos, float4* sattr, int32_t* inv_perm, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int th = 256;
    gather_sorted_kernel<<<static_cast<unsigned>((n + th - 1) / th), th, 0, s>>>(p, perm, n, spos, sattr, inv_perm);
    return hipGetLastError();
}

hipError_t launch_build_tree(const float4* spos, int64_t n, int leaf_size, int64_t nleaf,
                             int64_t nleaf_pad, Box* nodes, Box* groups, hipStream_t s)
{
    const int th = 256;
    if (groups && n > 0) {
        const int64_t m = (n + kGroup - 1) / kGroup * kGroup;
        tree_groups_kernel<<<static_cast<unsigned>((m + th - 1) / th), th, 0, s>>>(spos, n, groups);
        ASM_HIP_CHECK(hipGetLastError());
    }
    // Group boxes (computed above) stand in for the particles of whole-group leaves.
    const Box* const leaf_groups = (groups && n > 0 && leaf_size % kGroup == 0) ? groups : nullptr;
    tree_leaves_kernel<<<static_cast<unsigned>((nleaf_pad + th - 1) / th), th, 0, s>>>(
        spos, n, leaf_size, nleaf, nleaf_pad, leaf_groups, nodes);
    ASM_HIP_CHECK(hipGetLastError());
    for (int64_t first = nleaf_pad / 2; first >= 1; first /= 2) {
        tree_level_kernel<<<static_cast<unsigned>((first + th - 1) / th), th, 0, s>>>(nodes, first, first);
        ASM_HIP_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

}  // namespace asm_hip
