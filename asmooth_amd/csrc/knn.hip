// k-nearest-neighbour search, smoothing length and density scatter.
//
// Replaces neighbor_search / find_nbr_recursive (nbr_search.cpp:61-127) with
// the KNeighbors list (k_neighbors.h:9-51) and the OpenMP loop that calls them
// (clumping_tree_main.cpp:418-430).
//
// One query per lane.  A wavefront takes 64 consecutive particles in Morton
// order, so its queries are spatial neighbours, and walks the tree as a
// packet: a node is opened when any lane's conservative box distance is
// within that lane's current k-th distance; a leaf's particles are staged in
// LDS and every lane tests every staged particle.  The traversal stack lives
// in one VGPR (lane i holds entry i), so control flow stays wave-uniform and
// node records come through the scalar cache.
//
// Each lane keeps its 33 smallest candidates as an ascending array of 64-bit
// keys (FP32 squared distance bits, original index) in registers.  That is
// the canonical (d2, index) order: ties are broken by index, and distances
// use exactly the reference's FP32 arithmetic, so h and the neighbour sets
// are reproducible bit for bit.  h = sqrtf(d2 of slot 32) (nbr_search.cpp:73),
// the density kernel goes to slots 0..31 (nbr_search.cpp:78-84: the particle
// itself and its 31 nearest others), accumulated in FP64 with native atomics.
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

constexpr int kKnnThreads = 256;
constexpr int kWavesPerBlock = kKnnThreads / kWave;

// Insert ck (known to be below key[K-1]) into the ascending array, dropping
// the largest entry.  Slot i takes key[i-1] if ck sorts before it, else ck if
// ck sorts before key[i], else keeps key[i].
template <int K>
__device__ __forceinline__ void sorted_insert(uint64_t (&key)[K], uint64_t ck)
{
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
        const uint64_t prev = key[i - 1];
        const uint64_t cur = key[i];
        key[i] = (ck < prev) ? prev : ((ck < cur) ? ck : cur);
    }
    key[0] = (ck < key[0]) ? ck : key[0];
}

// nbr_search.cpp:27-37 with x = the particle, y = its neighbour.
__device__ __forceinline__ float tent_weight(float x0, float x1, float x2, float y0, float y1,
                                             float y2, float hinv)
{
    float w = hinv * hinv;
    w = w * hinv;
    w = w * (1.0f - fabsf(x0 - y0) * hinv);
    w = w * (1.0f - fabsf(x1 - y1) * hinv);
    w = w * (1.0f - fabsf(x2 - y2) * hinv);
    return w;
}

__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Every lane tests particles [start, start + cnt) of the sorted array,
// staged through LDS 64 at a time.
template <int K>
__device__ __forceinline__ void scan_range(const float4* __restrict__ spos, int64_t start, int64_t cnt,
                                           float4* stage, int lane, float qx, float qy, float qz,
                                           uint64_t (&key)[K], float& kth, unsigned long long& n_ins)
{
    for (int64_t base = 0; base < cnt; base += kWave) {
        const int m = static_cast<int>(cnt - base < kWave ? cnt - base : kWave);
        if (lane < m) stage[lane] = spos[start + base + lane];
        wave_lds_fence();
        for (int j = 0; j < m; ++j) {
            const float4 c = stage[j];
            const float d2 = norm2(qx, qy, qz, c.x, c.y, c.z);
            if (d2 <= kth) {
                const uint64_t ck = make_key(d2, __float_as_uint(c.w));
                if (ck < key[K - 1]) {
                    sorted_insert<K>(key, ck);
                    kth = key_d2(key[K - 1]);
                    ++n_ins;
                }
            }
        }
        wave_lds_fence();
    }
}

template <int K>
__global__ __launch_bounds__(kKnnThreads) void knn_dens_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad,
    int64_t nleaf, int leaf_size, int init_leaves,
    const int32_t* __restrict__ inv_perm, float* __restrict__ rk_sorted, double* __restrict__ dens64,
    int32_t* __restrict__ nbr_out, KnnStats* __restrict__ stats)
{
    __shared__ float4 stage[kWavesPerBlock][kWave];
    const int lane = threadIdx.x % kWave;
    const int wid = threadIdx.x / kWave;
    const int64_t wave0 = (static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid) * kWave;
    if (wave0 >= n) return;  // whole wavefront past the end
    const int64_t q = wave0 + lane;
    const bool valid = q < n;
    const float4 me = spos[valid ? q : n - 1];
    const float qx = me.x, qy = me.y, qz = me.z;

    uint64_t key[K];
#pragma unroll
    for (int i = 0; i < K; ++i) key[i] = kEmptyKey;
    float kth = INFINITY;  // d2 of key[K-1]

    unsigned long long n_node = 0, n_leaf = 0, n_dist = 0, n_ins = 0;
    const unsigned long long n_valid_lanes = __popcll(__ballot(valid));

    // Seed every lane's list from the wavefront's own leaves and init_leaves
    // leaves on either side in Morton order: a compact neighbourhood, so the
    // k-th distances are close to final before the tree walk starts, and the
    // walk then opens few nodes.  These leaves are skipped by the walk.
    const int64_t last = (n < wave0 + kWave ? n : wave0 + kWave) - 1;
    int64_t l0 = wave0 / leaf_size - init_leaves;
    int64_t l1 = last / leaf_size + init_leaves;
    if (l0 < 0) l0 = 0;
    if (l1 > nleaf - 1) l1 = nleaf - 1;
    {
        const int64_t p0 = l0 * leaf_size;
        const int64_t p1 = ((l1 + 1) * leaf_size < n) ? (l1 + 1) * leaf_size : n;
        scan_range<K>(spos, p0, p1 - p0, stage[wid], lane, qx, qy, qz, key, kth, n_ins);
        n_leaf += static_cast<unsigned long long>(l1 - l0 + 1);
        n_dist += static_cast<unsigned long long>(p1 - p0);
    }

    int stack = 0;
    int sp = 1;
    stack = (lane == 0) ? 1 : stack;
    while (sp > 0) {
        --sp;
        const int node = __builtin_amdgcn_readlane(stack, sp);
        const Box b = nodes[node];
        ++n_node;
        if (!wave_any(valid && box_d2(b, qx, qy, qz) <= kth)) continue;
        if (node >= nleaf_pad) {
            const int64_t leaf = node - nleaf_pad;
            if (leaf >= l0 && leaf <= l1) continue;  // scanned while seeding
            ++n_leaf;
            n_dist += static_cast<unsigned long long>(b.hi_w);
            scan_range<K>(spos, b.lo_w, b.hi_w, stage[wid], lane, qx, qy, qz, key, kth, n_ins);
        } else {
            const Box c0 = nodes[2 * node];
            const Box c1 = nodes[2 * node + 1];
            const float d0 = box_d2(c0, qx, qy, qz);
            const float d1 = box_d2(c1, qx, qy, qz);
            const bool need0 = wave_any(valid && d0 <= kth);
            const bool need1 = wave_any(valid && d1 <= kth);
            const int votes0 = __popcll(__ballot(valid && d0 <= d1));
            const bool near0 = 2 * votes0 >= static_cast<int>(n_valid_lanes);
            const int first = near0 ? 2 * node + 1 : 2 * node;   // pushed first = visited last
            const int second = near0 ? 2 * node : 2 * node + 1;
            const bool need_first = near0 ? need1 : need0;
            const bool need_second = near0 ? need0 : need1;
            if (need_first) {
                stack = (lane == sp) ? first : stack;
                ++sp;
            }
            if (need_second) {
                stack = (lane == sp) ? second : stack;
                ++sp;
            }
        }
    }

    if (valid) {
        const float kd = key_d2(key[K - 1]);
        const float h = sqrtf(kd);
        rk_sorted[q] = h;
        const float hinv = 1.0f / h;
#pragma unroll
        for (int s = 0; s < K - 1; ++s) {
            const uint32_t o = key_idx(key[s]);
            const int32_t sq = inv_perm[o];
            const float4 c = spos[sq];
            const float w = tent_weight(qx, qy, qz, c.x, c.y, c.z, hinv);
            unsafeAtomicAdd(&dens64[sq], static_cast<double>(w));
        }
        if (nbr_out) {
            const int64_t oq = static_cast<int64_t>(__float_as_uint(me.w));
#pragma unroll
            for (int s = 0; s < K; ++s) nbr_out[oq * K + s] = static_cast<int32_t>(key_idx(key[s]));
        }
    }
    if (stats && lane == 0) {
        atomicAdd(&stats->node_tests, n_node);
        atomicAdd(&stats->leaf_scans, n_leaf);
        atomicAdd(&stats->dist_evals, n_dist * n_valid_lanes);
    }
    if (stats) {
        // per-lane insert counts summed over the wave
        unsigned long long v = n_ins;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) atomicAdd(&stats->inserts, v);
    }
}

__global__ void dens_finalize_kernel(const double* __restrict__ d64, int64_t n, float* __restrict__ d)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = static_cast<float>(d64[i]);
}

}  // namespace

hipError_t launch_knn_dens(const float4* spos, int64_t n, const Box* nodes, int64_t nleaf_pad,
                           int64_t nleaf, int leaf_size, const int32_t* inv_perm, float* rk_sorted, double* dens64,
                           int32_t* nbr_out_orig, KnnStats* stats, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t waves = (n + kWave - 1) / kWave;
    const unsigned blocks = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    const int init_leaves = leaf_size >= kWave ? 1 : kWave / leaf_size;
    knn_dens_kernel<kSlots><<<blocks, kKnnThreads, 0, s>>>(spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves,
                                                          inv_perm, rk_sorted, dens64, nbr_out_orig, stats);
    return hipGetLastError();
}

hipError_t launch_dens_finalize(const double* dens64, int64_t n, float* dens, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int th = 256;
    dens_finalize_kernel<<<static_cast<unsigned>((n + th - 1) / th), th, 0, s>>>(dens64, n, dens);
    return hipGetLastError();
}

}  // namespace asm_hip
