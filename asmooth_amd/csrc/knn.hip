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
// are reproducible bit for bit.  h = sqrtf(d2 of slot 32) (nbr_search.cpp:73).
//
// The density scatter of nbr_search.cpp:78-84 (particle p adds its tent
// kernel to its 32 nearest, itself included) runs at the end of the search
// with native FP64 atomics (the FP32 result is the rounding of an FP64 sum,
// so the summation order does not show).  A bitwise-reproducible mode
// computes it instead as a gather: q receives W(p, q, h_p) from every p whose
// 33rd key is above (d2(p,q), q); the kNN kernel stores each 33rd key and a
// second walk, pruned by the largest 33rd distance below each node, sums the
// terms in FP64 registers in a fixed order.
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
// staged through LDS 64 at a time.  When any lane accepts a candidate, the
// whole wavefront runs the insertion with uniform control flow; lanes that
// did not accept insert an empty key, which leaves their array unchanged.
// (Divergent insertion makes the compiler keep two copies of the 33 keys.)
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
            const bool acc = d2 <= kth;
            if (__ballot(acc)) {
                const uint64_t ck = acc ? make_key(d2, __float_as_uint(c.w)) : kEmptyKey;
                sorted_insert<K>(key, ck);
                kth = key_d2(key[K - 1]);
                n_ins += acc ? 1u : 0u;
            }
        }
        wave_lds_fence();
    }
}

template <int K>
__global__ __launch_bounds__(kKnnThreads) void knn_dens_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad,
    int64_t nleaf, int leaf_size, int init_leaves,
    float* __restrict__ rk_sorted, uint64_t* __restrict__ kkey_sorted, const int32_t* __restrict__ inv_perm,
    double* __restrict__ dens64, int32_t* __restrict__ nbr_out, KnnStats* __restrict__ stats)
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
        kkey_sorted[q] = key[K - 1];
        if (dens64) {
            // nbr_search.cpp:75-84: slots 0..K-2 (the particle itself and its
            // 31 nearest others) receive W(p, q, 1/h_p); FP64 atomic sums.
            const float hinv = 1.0f / h;
#pragma unroll
            for (int s = 0; s < K - 1; ++s) {
                const int32_t sq = inv_perm[key_idx(key[s])];
                const float4 c = spos[sq];
                unsafeAtomicAdd(&dens64[sq], static_cast<double>(tent_weight(qx, qy, qz, c.x, c.y, c.z, hinv)));
            }
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


// Largest 33rd squared distance over the particles below each node.
__global__ void kmax_leaves_kernel(const uint64_t* __restrict__ kkey, int64_t n, int leaf_size, int64_t nleaf,
                                   int64_t nleaf_pad, float* __restrict__ kmax)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nleaf_pad) return;
    float m = -INFINITY;
    if (j < nleaf) {
        const int64_t start = j * leaf_size;
        const int64_t end = (start + leaf_size < n) ? start + leaf_size : n;
        for (int64_t i = start; i < end; ++i) m = fmaxf(m, key_d2(kkey[i]));
    }
    kmax[nleaf_pad + j] = m;
}

__global__ void kmax_level_kernel(float* __restrict__ kmax, int64_t first, int64_t count)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t i = first + t;
    kmax[i] = fmaxf(kmax[2 * i], kmax[2 * i + 1]);
}

// dens[q] = sum over p with (d2(p,q), q) < kkey[p] of W(p, q, 1/h_p), in FP64.
// d2 is symmetric bit for bit (fl(a-b) = -fl(b-a)), so the test reproduces
// "q is in p's 32 nearest" exactly; W uses p as x and q as y as
// nbr_search.cpp:81 does.  One query per lane, packet walk as the search.
__global__ __launch_bounds__(kKnnThreads) void dens_gather_kernel(
    const float4* __restrict__ spos, const uint64_t* __restrict__ kkey, const float* __restrict__ rk, int64_t n,
    const Box* __restrict__ nodes, const float* __restrict__ kmax, int64_t nleaf_pad, float* __restrict__ dens)
{
    __shared__ float4 stage_p[kWavesPerBlock][kWave];
    __shared__ float4 stage_k[kWavesPerBlock][kWave];   // (kkey lo, kkey hi, hinv, -)
    const int lane = threadIdx.x % kWave;
    const int wid = threadIdx.x / kWave;
    const int64_t wave0 = (static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid) * kWave;
    if (wave0 >= n) return;
    const int64_t q = wave0 + lane;
    const bool valid = q < n;
    const float4 me = spos[valid ? q : n - 1];
    const float qx = me.x, qy = me.y, qz = me.z;
    const uint32_t qidx = __float_as_uint(me.w);
    double acc = 0.0;

    int stack = (lane == 0) ? 1 : 0;
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int node = __builtin_amdgcn_readlane(stack, sp);
        const Box b = nodes[node];
        const float km = kmax[node];
        if (!wave_any(valid && box_d2(b, qx, qy, qz) <= km)) continue;
        if (node >= nleaf_pad) {
            const int start = b.lo_w;
            const int cnt = b.hi_w;
            if (lane < cnt) {
                const int64_t i = start + lane;
                stage_p[wid][lane] = spos[i];
                const uint64_t kk = kkey[i];
                stage_k[wid][lane] = make_float4(__uint_as_float(static_cast<uint32_t>(kk)),
                                                 __uint_as_float(static_cast<uint32_t>(kk >> 32)),
                                                 1.0f / rk[i], 0.0f);
            }
            wave_lds_fence();
            for (int j = 0; j < cnt; ++j) {
                const float4 p = stage_p[wid][j];
                const float4 pk = stage_k[wid][j];
                const float d2 = norm2(p.x, p.y, p.z, qx, qy, qz);
                const uint64_t kp = (static_cast<uint64_t>(__float_as_uint(pk.y)) << 32) | __float_as_uint(pk.x);
                if (make_key(d2, qidx) < kp)
                    acc += static_cast<double>(tent_weight(p.x, p.y, p.z, qx, qy, qz, pk.z));
            }
            wave_lds_fence();
        } else {
            stack = (lane == sp) ? 2 * node + 1 : stack;
            ++sp;
            stack = (lane == sp) ? 2 * node : stack;
            ++sp;
        }
    }
    if (valid) dens[q] = static_cast<float>(acc);
}

__global__ void dens_finalize_kernel(const double* __restrict__ d64, int64_t n, float* __restrict__ d)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = static_cast<float>(d64[i]);
}

}  // namespace

hipError_t launch_dens_finalize(const double* dens64, int64_t n, float* dens, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dens_finalize_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(dens64, n, dens);
    return hipGetLastError();
}

hipError_t launch_knn(const float4* spos, int64_t n, const Box* nodes, int64_t nleaf_pad, int64_t nleaf,
                      int leaf_size, float* rk_sorted, uint64_t* kkey_sorted, const int32_t* inv_perm, double* dens64,
                      int32_t* nbr_out_orig, KnnStats* stats, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t waves = (n + kWave - 1) / kWave;
    const unsigned blocks = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    const int init_leaves = leaf_size >= kWave ? 1 : kWave / leaf_size;
    knn_dens_kernel<kSlots><<<blocks, kKnnThreads, 0, s>>>(spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves,
                                                          rk_sorted, kkey_sorted, inv_perm, dens64, nbr_out_orig,
                                                          stats);
    return hipGetLastError();
}

hipError_t launch_kmax(const uint64_t* kkey, int64_t n, int leaf_size, int64_t nleaf, int64_t nleaf_pad, float* kmax,
                       hipStream_t s)
{
    const int th = 256;
    kmax_leaves_kernel<<<static_cast<unsigned>((nleaf_pad + th - 1) / th), th, 0, s>>>(kkey, n, leaf_size, nleaf,
                                                                                      nleaf_pad, kmax);
    ASM_HIP_CHECK(hipGetLastError());
    for (int64_t first = nleaf_pad / 2; first >= 1; first /= 2) {
        kmax_level_kernel<<<static_cast<unsigned>((first + th - 1) / th), th, 0, s>>>(kmax, first, first);
        ASM_HIP_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t launch_dens_gather(const float4* spos, const uint64_t* kkey, const float* rk, int64_t n, const Box* nodes,
                              const float* kmax, int64_t nleaf_pad, float* dens, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t waves = (n + kWave - 1) / kWave;
    const unsigned blocks = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    dens_gather_kernel<<<blocks, kKnnThreads, 0, s>>>(spos, kkey, rk, n, nodes, kmax, nleaf_pad, dens);
    return hipGetLastError();
}

}  // namespace asm_hip
