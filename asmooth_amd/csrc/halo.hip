// Halo excision: mark particles inside spherical-overdensity halos.
//
// Replaces mark_halo_particles (nbr_search.cpp:140-165) and the loop over
// halos that calls it (clumping_tree_main.cpp:465-485).  One wavefront per
// halo walks the tree with the conservative box test; every lane tests one
// particle of a visited leaf with the reference's FP32 distance and the
// strict test d2 < r*r (nbr_search.cpp:150-153), and sets igrp = -1.  The
// count adds one per (halo, particle) pair, as the reference's sum does.
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

constexpr int kHaloThreads = 256;
constexpr int kHaloWaves = kHaloThreads / kWave;

__global__ __launch_bounds__(kHaloThreads) void mark_halo_kernel(
    const float4* __restrict__ spos, float4* __restrict__ sattr, int64_t n,
    const Box* __restrict__ nodes, int64_t nleaf_pad, const HaloRec* __restrict__ halos, int64_t nh,
    float left, float right, unsigned long long* __restrict__ count)
{
    const int lane = threadIdx.x % kWave;
    const int64_t h = static_cast<int64_t>(blockIdx.x) * kHaloWaves + threadIdx.x / kWave;
    if (h >= nh) return;
    const HaloRec H = halos[h];
    // Only halos whose centre lies inside the buffered cube
    // (clumping_tree_main.cpp:478-480).
    if (!(left < H.x[0] && H.x[0] < right && left < H.x[1] && H.x[1] < right && left < H.x[2] &&
          H.x[2] < right))
        return;
    const float r2 = H.r * H.r;
    unsigned long long marked = 0;
    int stack = (lane == 0) ? 1 : 0;
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int node = __builtin_amdgcn_readlane(stack, sp);
        const Box b = nodes[node];
        // Every particle of the box has d2 >= box_d2, so none passes d2 < r2.
        if (!(box_d2(b, H.x[0], H.x[1], H.x[2]) < r2)) continue;
        if (node >= nleaf_pad) {
            const int start = b.lo_w;
            const int cnt = b.hi_w;
            for (int j = lane; j < cnt; j += kWave) {
                const float4 p = spos[start + j];
                if (norm2(H.x[0], H.x[1], H.x[2], p.x, p.y, p.z) < r2) {
                    reinterpret_cast<int*>(sattr)[4 * (static_cast<int64_t>(start) + j) + 3] = -1;
                    ++marked;
                }
            }
        } else {
            stack = (lane == sp) ? 2 * node + 1 : stack;
            ++sp;
            stack = (lane == sp) ? 2 * node : stack;
            ++sp;
        }
    }
    for (int o = 32; o > 0; o >>= 1) marked += __shfl_xor(marked, o);
    if (lane == 0 && marked) atomicAdd(count, marked);
}

}  // namespace

hipError_t launch_mark_halo(const float4* spos, float4* sattr, int64_t n, const Box* nodes,
                            int64_t nleaf_pad, const HaloRec* halos, int64_t nh, float left,
                            float right, unsigned long long* count, hipStream_t s)
{
    if (nh <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((nh + kHaloWaves - 1) / kHaloWaves);
    mark_halo_kernel<<<blocks, kHaloThreads, 0, s>>>(spos, sattr, n, nodes, nleaf_pad, halos, nh, left,
                                                     right, count);
    return hipGetLastError();
}

}  // namespace asm_hip
