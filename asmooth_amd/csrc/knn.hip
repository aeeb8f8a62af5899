// k-nearest-neighbour search, smoothing length and density scatter.
//
// Replaces neighbor_search / find_nbr_recursive (nbr_search.cpp:61-127) with
// the KNeighbors list (k_neighbors.h:9-51) and the OpenMP loop that calls them
// (clumping_tree_main.cpp:418-430).
//
// One query per lane.  A wavefront takes 64 consecutive particles in Morton
// order, so its queries are spatial neighbours, and walks the tree as a
// packet: a node is opened when any lane's conservative box distance is
// within that lane's current bound; a leaf's particles are staged in LDS and
// every lane tests every staged particle.  The traversal stack lives in one
// VGPR (lane i holds entry i), so control flow stays wave-uniform and node
// records come through the scalar cache.
//
// The neighbour set is defined canonically: the 33 smallest keys (FP32
// squared distance bits, original index), distances in exactly the
// reference's FP32 arithmetic, ties broken by index; h = sqrtf(d2 of the
// 33rd) (nbr_search.cpp:73).  It is found in two walks:
//
//  1. bound: each lane keeps only the 33 smallest squared distances, as an
//     ascending float array in registers.  Inserting d into it is one
//     v_med3_f32 per slot (slot i becomes med3(t[i-1], d, t[i])), with no
//     compares or selects, and a lane whose d is not below t[32] is left
//     unchanged by the same instructions.  At the end t[32] = tau, the 33rd
//     smallest squared distance over all particles.
//     By default walk 1 is instead a selection (knn_select_kernel): after a
//     per-lane bound from the seed range and the union of the leaves within
//     it, the wavefront settles its 64 queries one at a time, all lanes on
//     one query, and takes tau as the 33rd of the query's candidates within
//     its bound, sorted across the wavefront.  Same tau; fewer instructions.
//  2. collect: a second walk with tau fixed opens exactly the leaves whose
//     box distance is within tau and appends every particle with d2 <= tau to
//     a per-lane list in LDS: normally 33 entries, more only when distances
//     tie at tau.  Entries below tau are the first slots in any order; ties at
//     tau are ranked by index.
//
// The density scatter of nbr_search.cpp:78-84 (particle p adds its tent
// kernel to its 32 nearest, itself included) runs at the end of the collect
// walk with native FP64 atomics (the FP32 result is the rounding of an FP64
// sum, so the summation order does not show).  A wavefront whose tie list
// overflows the LDS list is redone by the insertion kernel below, which keeps
// the 33 (d2, index) keys sorted in registers (also selectable for every
// wavefront).  A bitwise-reproducible mode computes the density instead as a
// gather: q receives W(p, q, h_p) from every p whose 33rd key is above
// (d2(p,q), q); the search stores each 33rd key and a second walk, pruned by
// the largest 33rd distance below each node, sums the terms in FP64 registers
// in a fixed order.
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

constexpr int kKnnThreads = 256;
// Flags the collect walk leaves for the insertion kernel, per wavefront.
constexpr uint32_t kRedoAll = 1u;    // nothing done: redo the search and the whole scatter
constexpr uint32_t kRedoTies = 2u;   // terms below tau scattered: redo the search, scatter ties only
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

// The density kernels: the reference's separable tent (ker,
// nbr_search.cpp:27-37, the one its neighbour loop uses) and its cubic
// spline (spline_ker, nbr_search.cpp:39-53, present there but not called),
// selected per context (kKernTent / kKernSpline).  The spline is radial: it
// takes the squared distance d2 = norm2(x, y), which is the same bit for bit
// with x and y swapped.  Its arithmetic follows spline_ker: the factor
// 8/pi hinv^3 evaluated in FP64 (M_PI is a double) and rounded once to FP32,
// r = sqrtf(d2) * hinv, then the two polynomial pieces in FP32, left to right.
__device__ __forceinline__ float spline_weight(float d2, float hinv)
{
    const double h = static_cast<double>(hinv);
    const float fac = static_cast<float>(((8.0 / M_PI) * h) * h * h);
    const float r = sqrtf(d2) * hinv;
    if (r < 0.5f) {
        const float r6 = 6.0f * r;
        return fac * ((1.0f - r6 * r) + (r6 * r) * r);
    }
    const float r1 = 1.0f - r;
    return fac * (((2.0f * r1) * r1) * r1);
}

__device__ __forceinline__ float dens_weight(int kern, float x0, float x1, float x2, float y0, float y1, float y2,
                                             float hinv, float d2)
{
    return kern == kKernSpline ? spline_weight(d2, hinv) : tent_weight(x0, x1, x2, y0, y1, y2, hinv);
}

// The wavefront's index in its block, as a scalar: the compiler cannot tell
// that threadIdx.x / 64 is uniform, and per-wavefront loop bounds derived
// from it would otherwise live in VGPRs with divergent loop exits.
__device__ __forceinline__ int wave_in_block()
{
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
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
    double* __restrict__ dens64, int32_t* __restrict__ nbr_out, KnnStats* __restrict__ stats,
    const uint32_t* __restrict__ only_flagged, int kern)
{
    __shared__ float4 stage[kWavesPerBlock][kWave];
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock + wid) * kWave;
    if (wave0 >= n) return;  // whole wavefront past the end
    const uint32_t redo = only_flagged ? only_flagged[wave0 / kWave] : kRedoAll;
    if (redo == 0u) return;  // done by the collect walk
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
    // One call site of the scan (the seed range first, then each leaf the
    // walk opens), so the compiler keeps a single copy of the key array.
    int64_t r_start = l0 * leaf_size;
    int64_t r_cnt = (((l1 + 1) * leaf_size < n) ? (l1 + 1) * leaf_size : n) - r_start;
    n_leaf += static_cast<unsigned long long>(l1 - l0 + 1);
    int stack = (lane == 0) ? 1 : 0;
    int sp = 1;
    for (;;) {
        n_dist += static_cast<unsigned long long>(r_cnt);
        scan_range<K>(spos, r_start, r_cnt, stage[wid], lane, qx, qy, qz, key, kth, n_ins);
        bool found = false;
        while (sp > 0 && !found) {
            --sp;
            const int node = __builtin_amdgcn_readlane(stack, sp);
            const Box b = nodes[node];
            ++n_node;
            if (!wave_any(valid && box_d2(b, qx, qy, qz) <= kth)) continue;
            if (node >= nleaf_pad) {
                const int64_t leaf = node - nleaf_pad;
                if (leaf >= l0 && leaf <= l1) continue;  // scanned while seeding
                ++n_leaf;
                r_start = b.lo_w;
                r_cnt = b.hi_w;
                found = true;
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
        if (!found) break;
    }

    if (valid) {
        const float kd = key_d2(key[K - 1]);
        const float h = sqrtf(kd);
        rk_sorted[q] = h;
        kkey_sorted[q] = key[K - 1];
        if (dens64) {
            // nbr_search.cpp:75-84: slots 0..K-2 (the particle itself and its
            // 31 nearest others) receive W(p, q, 1/h_p); FP64 atomic sums.
            // After a collect walk that overflowed its lists, the terms
            // below kd are already scattered; only ties at kd are left.
            const float hinv = 1.0f / h;
            const bool ties_only = redo == kRedoTies;
#pragma unroll
            for (int s = 0; s < K - 1; ++s) {
                if (ties_only && key_d2(key[s]) < kd) continue;
                const int32_t sq = inv_perm[key_idx(key[s])];
                const float4 c = spos[sq];
                unsafeAtomicAdd(&dens64[sq], static_cast<double>(dens_weight(kern, qx, qy, qz, c.x, c.y, c.z, hinv,
                                                                             key_d2(key[s]))));
            }
        }
        if (nbr_out) {
            const int64_t oq = static_cast<int64_t>(__float_as_uint(me.w));
#pragma unroll
            for (int s = 0; s < K; ++s) nbr_out[oq * K + s] = static_cast<int32_t>(key_idx(key[s]));
        }
    }
    if (stats) {
        // per-lane insert counts summed over the wave
        unsigned long long v = n_ins;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        KnnStats* const st = stats + ((wave0 / kWave) & (kStatSlots - 1));
        if (lane == 0) {
            atomicAdd(&st->node_tests, n_node);
            atomicAdd(&st->leaf_scans, n_leaf);
            atomicAdd(&st->dist_evals, n_dist * n_valid_lanes);
            atomicAdd(&st->inserts, v);
        }
    }
}


// ---------------------------------------------------------------------------
// Bound walk and collect walk.

constexpr int kBoundThreads = 64;
constexpr int kBoundWaves = kBoundThreads / kWave;
constexpr int kCollectThreads = 64;
constexpr int kCollectWaves = kCollectThreads / kWave;
constexpr int kCollectSlots = 40;   // per-lane list of all entries: 33 plus up to 7 ties at tau
constexpr int kTieSlots = 8;        // per-lane list of the entries at tau

// Insert d into the ascending array t, dropping the largest entry: slot i
// becomes the median of (t[i-1], d, t[i]), slot 0 the median of (0, d, t[0])
// (all values are >= 0).  A d that is not below t[K-1] leaves every slot
// unchanged.
template <int K>
__device__ __forceinline__ void med3_insert(float (&t)[K], float d)
{
#pragma unroll
    for (int i = K - 1; i > 0; --i) t[i] = __builtin_amdgcn_fmed3f(t[i - 1], d, t[i]);
    t[0] = __builtin_amdgcn_fmed3f(0.0f, d, t[0]);
}

// Stage up to 64 particles of [first, first + left) in LDS; slots past the
// end hold a point at infinity (squared distance +inf: never accepted), so
// the scan loops can take whole groups of candidates.  With group boxes
// (first is then a multiple of kGroup), the boxes of the 8 groups are staged
// too: a Box is kGroup dwords, so group first / kGroup starts at dword
// `first` and lane l loads dword first + l (clamped to the array; only groups
// holding staged particles are ever read).  Returns the number of slots to
// scan, rounded up to whole groups.
__device__ __forceinline__ int stage_batch(const float4* __restrict__ spos, int first, int left,
                                           float4* stage, const uint32_t* __restrict__ gwords, int gmax,
                                           uint32_t* gstage, int lane)
{
    static_assert(sizeof(Box) == kGroup * sizeof(uint32_t), "a group box is kGroup dwords");
    const int m = left < kWave ? left : kWave;
    // Addresses as a uniform base plus a 32-bit lane offset (no 64-bit
    // per-lane pointers held across the walk).
    const uint32_t idx = static_cast<uint32_t>(first + lane);
    const float4 c = (lane < m) ? *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(spos) + idx * 16u)
                                : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    uint32_t gw = 0u;
    if (gwords) {
        const uint32_t gi = idx < static_cast<uint32_t>(gmax) ? idx : static_cast<uint32_t>(gmax);
        gw = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(gwords) + gi * 4u);
    }
    stage[lane] = c;
    gstage[lane] = gw;
    wave_lds_fence();
    return (m + kGroup - 1) & ~(kGroup - 1);
}

template <int K>
__device__ __forceinline__ void bound_scan(const float4* __restrict__ spos, int start, int cnt,
                                           float4* stage, const uint32_t* __restrict__ gwords, int gmax,
                                           uint32_t* gstage, int lane, float qx, float qy, float qz, float (&t)[K],
                                           uint32_t& n_acc, uint32_t& n_dist)
{
    for (int base = 0; base < cnt; base += kWave) {
        const int m8 = stage_batch(spos, start + base, cnt - base, stage, gwords, gmax, gstage, lane);
        for (int g = 0; g < m8; g += kGroup) {
            // Skip a group whose box no lane's current bound reaches: every
            // candidate in it is at least the box distance away.
            if (gwords && !wave_any(box_d2(*reinterpret_cast<const Box*>(gstage + g), qx, qy, qz) <= t[K - 1]))
                continue;
            n_dist += kGroup;
#pragma unroll
            for (int j = g; j < g + kGroup; j += 4) {
                float d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float4 c = stage[j + u];
                    d[u] = norm2(qx, qy, qz, c.x, c.y, c.z);
                }
                // Sort the four ascending per lane (a five-exchange network on
                // the bit patterns: non-negative floats order like unsigned
                // integers), then insert them in order while some lane still
                // accepts one.  Insertion order does not change the final
                // set, and the chain runs as many times as the busiest lane
                // needs rather than once per candidate that any lane accepts.
                uint32_t a0 = __float_as_uint(d[0]), a1 = __float_as_uint(d[1]);
                uint32_t a2 = __float_as_uint(d[2]), a3 = __float_as_uint(d[3]);
                uint32_t lo = min(a0, a1), hi = max(a0, a1);
                a0 = lo; a1 = hi;
                lo = min(a2, a3); hi = max(a2, a3);
                a2 = lo; a3 = hi;
                lo = min(a0, a2); hi = max(a0, a2);
                a0 = lo; a2 = hi;
                lo = min(a1, a3); hi = max(a1, a3);
                a1 = lo; a3 = hi;
                lo = min(a1, a2); hi = max(a1, a2);
                a1 = lo; a2 = hi;
                const float sv[4] = {__uint_as_float(a0), __uint_as_float(a1), __uint_as_float(a2),
                                     __uint_as_float(a3)};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool acc = sv[u] < t[K - 1];
                    if (!wave_any(acc)) break;
                    med3_insert<K>(t, sv[u]);
                    n_acc += acc ? 1u : 0u;
                }
            }
        }
        wave_lds_fence();
    }
}

// Walk 1: tau[q] = 33rd smallest squared distance from q (itself included).
// Seeded from the wavefront's own leaves and init_leaves leaves on either
// side in Morton order, which are then skipped by the walk.  A node is opened
// when some lane's box distance is within its current t[K-1] (>= tau), so
// the leaves opened here include every leaf that walk 2 needs; their ids are
// recorded per wavefront (up to kLeafCap; beyond that the count alone is
// kept and walk 2 hands the wavefront to the insertion kernel).
//
// With a learned bound tau_init (per sorted particle), t starts as 33 copies
// of it: real distances below it displace the copies, so t[K-1] < tau_init at
// the end means the 33 smallest were all found below the bound and the
// result is exact; otherwise the lane's wavefront is flagged in seed_redo
// and searched again with only_flagged = seed_redo and no bound.
template <int K>
__global__ __launch_bounds__(kBoundThreads, 8) void knn_bound_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad, int64_t nleaf,
    int leaf_size, int init_leaves, float* __restrict__ tau_out, uint32_t* __restrict__ wleaf,
    uint32_t* __restrict__ wleaf_n, KnnStats* __restrict__ stats, const float* __restrict__ tau_init,
    const uint32_t* __restrict__ only_flagged, uint32_t* __restrict__ seed_redo,
    unsigned long long* __restrict__ seed_fallbacks, const Box* __restrict__ groups)
{
    __shared__ float4 stage[kBoundWaves][kWave];
    __shared__ __align__(16) uint32_t gstage[kBoundWaves][kWave];
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kBoundWaves + wid) * kWave;
    if (wave0 >= n) return;
    const int64_t wave = wave0 / kWave;
    if (only_flagged && only_flagged[wave] == 0u) return;
    const int64_t q = wave0 + lane;
    const bool valid = q < n;
    const float4 me = spos[valid ? q : n - 1];
    const float qx = me.x, qy = me.y, qz = me.z;
    // Lanes past the end repeat the last particle (position and bound), so
    // they never need a node or a group that its lane does not need: the
    // wavefront-wide tests below need no validity mask.
    const float t0 = tau_init ? tau_init[valid ? q : n - 1] : INFINITY;

    float t[K];
#pragma unroll
    for (int i = 0; i < K; ++i) t[i] = t0;

    // Particle, node and leaf indices fit in 32 bits (the launcher checks
    // n); the walk keeps them in 32-bit scalars.
    const int ni = static_cast<int>(n);
    const int w0 = static_cast<int>(wave0);
    const int nlp = static_cast<int>(nleaf_pad);
    const uint32_t* const gwords = reinterpret_cast<const uint32_t*>(groups);
    const int gmax = (ni + kGroup - 1) / kGroup * kGroup - 1;   // last dword of the group boxes

    uint32_t n_node = 0, n_dist = 0, n_acc = 0;
    const unsigned long long n_valid_lanes = __popcll(wave_ballot(valid));

    const int last = (ni < w0 + kWave ? ni : w0 + kWave) - 1;
    int l0 = w0 / leaf_size - init_leaves;
    int l1 = last / leaf_size + init_leaves;
    if (l0 < 0) l0 = 0;
    if (l1 > static_cast<int>(nleaf) - 1) l1 = static_cast<int>(nleaf) - 1;
    uint32_t n_leaf = static_cast<uint32_t>(l1 - l0 + 1);

    // Recorded leaf ids: 64 at a time in one VGPR (lane i holds entry i),
    // flushed with one coalesced store.
    uint32_t rec = 0;
    int n_rec = 0;
    uint32_t* const my_leaves = wleaf + wave * kLeafCap;

    // One call site of the scan (the seed range first, then each leaf the
    // walk opens), so the compiler keeps a single copy of t.
    int r_start = l0 * leaf_size;
    int r_cnt = (((l1 + 1) * leaf_size < ni) ? (l1 + 1) * leaf_size : ni) - r_start;
    // Traversal stack in one VGPR: lane i holds entry i (a depth-first walk
    // of the binary tree keeps at most one entry per level plus one, well
    // below 64).
    int stack = (lane == 0) ? 1 : 0;
    int sp = 1;
    const auto record = [&](int leaf) {
        if (n_rec < kLeafCap) {
            rec = (lane == (n_rec & (kWave - 1))) ? static_cast<uint32_t>(leaf) : rec;
            if ((n_rec & (kWave - 1)) == kWave - 1) my_leaves[(n_rec & ~(kWave - 1)) + lane] = rec;
        }
        ++n_rec;
    };
    for (;;) {
        bound_scan<K>(spos, r_start, r_cnt, stage[wid], gwords, gmax, gstage[wid], lane, qx, qy, qz, t, n_acc,
                      n_dist);
        // The stack top was tested against the current bound when it was
        // pushed, until the next scan lowers the bound.
        bool fresh = false;
        bool found = false;
        while (sp > 0 && !found) {
            --sp;
            const int node = __builtin_amdgcn_readlane(stack, sp);
            ++n_node;
            const float kth = t[K - 1];
            if (node >= nlp) {
                // A leaf is only ever popped as the root of a one-leaf tree.
                const Box b = nodes[node];
                if (!wave_any(box_d2(b, qx, qy, qz) <= kth)) continue;
                const int leaf = node - nlp;
                if (leaf >= l0 && leaf <= l1) continue;  // scanned while seeding
                ++n_leaf;
                r_start = b.lo_w;
                r_cnt = b.hi_w;
                found = true;
                record(leaf);
                continue;
            }
            if (!fresh && !wave_any(box_d2(nodes[node], qx, qy, qz) <= kth)) continue;
            fresh = false;
            if (4 * node >= nlp && 2 * node < nlp) {
                // The grandchildren are four leaves, consecutive in particle
                // order: test their boxes directly and scan the span from the
                // first needed one to the last (the group boxes weed out the
                // rest of the span); record only the needed ones.  A span
                // holding a leaf of the seed range (already scanned) takes
                // the two-level path below instead.
                const int g0 = 4 * node;
                const int la = g0 - nlp;
                uint32_t mask = 0, seeded = 0;
                Box lb[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    lb[j] = nodes[g0 + j];
                    const bool sj = la + j >= l0 && la + j <= l1;
                    seeded |= sj ? (1u << j) : 0u;
                    if (!sj && wave_any(box_d2(lb[j], qx, qy, qz) <= kth)) mask |= 1u << j;
                }
                if (mask == 0u) continue;
                const int jlo = __builtin_ctz(mask), jhi = 31 - __builtin_clz(mask);
                const uint32_t span = (2u << jhi) - (1u << jlo);
                if ((span & seeded) == 0u) {
                    int lo_w = lb[0].lo_w, hi_end = lb[0].lo_w + lb[0].hi_w;
#pragma unroll
                    for (int j = 1; j < 4; ++j) {
                        lo_w = (j == jlo) ? lb[j].lo_w : lo_w;
                        hi_end = (j == jhi) ? lb[j].lo_w + lb[j].hi_w : hi_end;
                    }
                    r_start = lo_w;
                    r_cnt = hi_end - lo_w;
                    n_leaf += static_cast<uint32_t>(__popc(mask));
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (mask & (1u << j)) record(la + j);
                    found = true;
                    continue;
                }
            }
            const int c0 = 2 * node;
            const Box b0 = nodes[c0];
            const Box b1 = nodes[c0 + 1];
            const float d0 = box_d2(b0, qx, qy, qz);
            const float d1 = box_d2(b1, qx, qy, qz);
            const bool need0 = wave_any(d0 <= kth);
            const bool need1 = wave_any(d1 <= kth);
            if (c0 >= nlp) {
                // Both children are leaves, adjacent in particle order: the
                // needed ones (outside the seed range) are scanned together.
                const int la = c0 - nlp;
                const bool sa = need0 && !(la >= l0 && la <= l1);
                const bool sb = need1 && !(la + 1 >= l0 && la + 1 <= l1);
                if (!sa && !sb) continue;
                r_start = sa ? b0.lo_w : b1.lo_w;
                r_cnt = (sa ? b0.hi_w : 0) + (sb ? b1.hi_w : 0);
                n_leaf += static_cast<uint32_t>(sa) + static_cast<uint32_t>(sb);
                if (sa) record(la);
                if (sb) record(la + 1);
                found = true;
            } else {
                // Near child (by lane vote) on top, visited first.  Slot sp
                // takes the far child when it is needed (else the near one),
                // slot sp + 1 the near one; sp advances by the number needed,
                // so an unused slot lies above the stack top.
                const bool near0 = 2 * __popcll(wave_ballot(d0 <= d1)) >= kWave;
                const int nearc = near0 ? c0 : c0 + 1;
                const bool need_near = near0 ? need0 : need1;
                const bool need_far = near0 ? need1 : need0;
                stack = (lane == sp) ? (need_far ? (nearc ^ 1) : nearc) : stack;
                stack = (lane == sp + 1) ? nearc : stack;
                sp += static_cast<int>(need_far) + static_cast<int>(need_near);
                fresh = need_far || need_near;
            }
        }
        if (!found) break;
    }
    if (n_rec <= kLeafCap && (n_rec & (kWave - 1)) != 0 && lane < (n_rec & (kWave - 1)))
        my_leaves[(n_rec & ~(kWave - 1)) + lane] = rec;
    if (lane == 0) wleaf_n[wave] = static_cast<uint32_t>(n_rec);
    if (valid) tau_out[q] = t[K - 1];
    if (tau_init) {
        const uint64_t fb = wave_ballot(valid && !(t[K - 1] < t0));
        if (lane == 0) {
            seed_redo[wave] = fb ? 1u : 0u;
            if (fb) atomicAdd(&seed_fallbacks[wave & (kStatSlots - 1)], static_cast<unsigned long long>(__popcll(fb)));
        }
    }

    if (stats) {
        unsigned long long v = n_acc;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        KnnStats* const st = stats + (wave & (kStatSlots - 1));
        if (lane == 0) {
            atomicAdd(&st->node_tests, static_cast<unsigned long long>(n_node));
            atomicAdd(&st->leaf_scans, static_cast<unsigned long long>(n_leaf));
            atomicAdd(&st->dist_evals, static_cast<unsigned long long>(n_dist) * n_valid_lanes);
            atomicAdd(&st->inserts, v);
        }
    }
}

// ---------------------------------------------------------------------------
// Walk 1 by selection (the default).  The packet walk above makes every lane
// evaluate every candidate that some lane of its wavefront needs: about 53
// times the 33 slots kept.  Here a wavefront still takes 64 consecutive
// particles, but after two cheap wavefront-wide phases it settles them one
// query at a time, all 64 lanes on one query:
//
//  A. Bound: every lane counts the squared distances from its query to the
//     particles of its seed (the wavefront's particles and 95 on either side
//     in Morton order) in a byte histogram in LDS (64 bins of an eighth of
//     an octave of d2, below the seed's own box diagonal).
//     The upper edge of the bin holding the 33rd count is a bound rA2 >= tau:
//     at least 33 particles lie within it.
//  B. Union: a walk of the tree with the bounds fixed lists every leaf whose
//     box lies within rA2 of some lane, with the mask of the lanes it does
//     (the ballot of the box test).
//  C. Per query: the listed leaves with the query's bit set give their group
//     boxes, those within the query's bound their particles, and the squared
//     distances within the bound go to a list in LDS.  The list is cut back
//     whenever it passes 64 entries: the 33rd smallest of the first 64 is
//     found across the wavefront (two 32-lane bitonic sorts on the bit
//     patterns, which order non-negative floats, then a half-cleaner and a
//     minimum), the entries up to it are kept (33 of them with ties) and the
//     bound drops to it.  At the end tau is the list's 33rd smallest, found
//     the same way.  Every particle with d2 <= tau lies within every bound the query
//     had, so it was listed: tau is the 33rd smallest squared distance over
//     all particles, ties included, in the reference's FP32 arithmetic.
//
// The listed leaves outside the seed range are recorded for the collect walk
// (they include every leaf within tau of some lane).  A wavefront whose
// union would exceed kLeafCap leaves, or whose seed holds fewer than 33
// particles, is flagged in `redo` and left to the packet walk.
constexpr int kSelThreads = 64;
constexpr int kSelWaves = kSelThreads / kWave;
constexpr int kSelBins = 64;                     // byte counters per lane
constexpr int kSelHistWords = kSelBins / 4;
constexpr int kSelSeedHalo = 95;                 // seed particles on either side of the wavefront's 64
static_assert(kWave + 2 * kSelSeedHalo <= 255, "a byte bin never overflows");
constexpr int kSelBinShift = 20;                 // d2 bit pattern >> 20: exponent and 3 mantissa bits

struct SelectLds {
    union {
        struct {
            uint32_t hist[kSelHistWords][kWave];  // lane l's bins 4w..4w+3 in word [w][l]
            float4 stage[kWave];
        } a;
        struct {
            uint64_t umask[kLeafCap];  // lanes whose bound reaches union leaf e
            uint32_t uleaf[kLeafCap];
            union {
                struct {
                    uint32_t front0[kLeafCap];  // a frontier
                    uint32_t front1[kLeafCap];  // a frontier, or the kept leaves of a batch
                };                              // B
                struct {
                    uint32_t slist[2 * kWave];  // union leaves one query needs, 128 union entries at a time
                    uint32_t glist[2 * kWave];  // their groups within its bound, one round
                    uint32_t cand[2 * kWave];   // its squared distances within the bound ...
                    uint32_t cpos[2 * kWave];   // ... and their sorted positions
                } c;                            // C
            };
        } b;
    };
};

// Partner value of lane ^ X.  Patterns within a row of 16 lanes are DPP
// moves, the others within 32 lanes ds_swizzle (no memory access), and the
// mirror across the whole wavefront a ds_bpermute.
template <int X>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int lane)
{
    const int s = static_cast<int>(v);
    if constexpr (X == 1) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, s, 0xB1, 0xF, 0xF, true));
    else if constexpr (X == 2) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, s, 0x4E, 0xF, 0xF, true));
    else if constexpr (X == 3) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, s, 0x1B, 0xF, 0xF, true));
    else if constexpr (X == 7) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, s, 0x141, 0xF, 0xF, true));
    else if constexpr (X == 15) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, s, 0x140, 0xF, 0xF, true));
    else if constexpr (X == 63) return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute((63 - lane) << 2, s));
    else {
        static_assert(X == 4 || X == 8 || X == 16 || X == 31, "xor pattern");
        return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(s, 0x1f | (X << 10)));
    }
}

// One compare-exchange stage: lanes with bit M clear keep the smaller value.
template <int X, int M>
__device__ __forceinline__ uint32_t cx_stage(uint32_t v, int lane)
{
    const uint32_t p = xor_lane<X>(v, lane);
    return (lane & M) ? max(v, p) : min(v, p);
}

// The 33rd smallest of the 64 lanes' values.  Each half of 32 lanes is
// sorted ascending by a bitonic network whose first stage of each merge of
// size k compares mirrored lanes (lane ^ (k - 1)), the rest lane ^ j, the
// smaller value going to the lower lane.  With A and B the sorted halves,
// the 32 largest values of the wavefront are max(A[l], B[31 - l]) for
// l = 0..31 (the upper half of the bitonic half-cleaner on A followed by B
// reversed), so the 33rd smallest is their minimum.  Lane l and lane 63 - l
// form the same pair, so every lane holds one of the 32 maxima and a minimum
// within each half of 32 lanes covers them all.  Must be called with all 64
// lanes active; returns the value in every lane.
__device__ __forceinline__ uint32_t wave_33rd(uint32_t v, int lane)
{
    v = cx_stage<1, 1>(v, lane);
    v = cx_stage<3, 2>(v, lane);
    v = cx_stage<1, 1>(v, lane);
    v = cx_stage<7, 4>(v, lane);
    v = cx_stage<2, 2>(v, lane);
    v = cx_stage<1, 1>(v, lane);
    v = cx_stage<15, 8>(v, lane);
    v = cx_stage<4, 4>(v, lane);
    v = cx_stage<2, 2>(v, lane);
    v = cx_stage<1, 1>(v, lane);
    v = cx_stage<31, 16>(v, lane);
    v = cx_stage<8, 8>(v, lane);
    v = cx_stage<4, 4>(v, lane);
    v = cx_stage<2, 2>(v, lane);
    v = cx_stage<1, 1>(v, lane);
    v = max(v, xor_lane<63>(v, lane));
    // minimum over each half: within quads, 8-lane and 16-lane rows, then 32
    v = min(v, xor_lane<1>(v, lane));
    v = min(v, xor_lane<2>(v, lane));
    v = min(v, xor_lane<7>(v, lane));
    v = min(v, xor_lane<15>(v, lane));
    v = min(v, xor_lane<31>(v, lane));
    return v;
}

// Entry r of the list of sorted particle q: one row of kListCap per
// particle, so that a query's list is one contiguous store.
__device__ __forceinline__ int64_t list_slot(int64_t q, int r)
{
    return q * kListCap + r;
}

__device__ __forceinline__ int lane_rank(uint64_t mask)
{
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
}

template <int K>
__global__ __launch_bounds__(kSelThreads) __attribute__((amdgpu_waves_per_eu(8))) void knn_select_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad, int64_t nleaf,
    int leaf_size, int init_leaves, float* __restrict__ tau_out, uint32_t* __restrict__ wleaf,
    uint32_t* __restrict__ wleaf_n, uint32_t* __restrict__ redo, KnnStats* __restrict__ stats,
    const Box* __restrict__ groups, uint32_t* __restrict__ list, uint32_t* __restrict__ list_n,
    float* __restrict__ rk_sorted, uint64_t* __restrict__ kkey_sorted, const float* __restrict__ tau_init,
    const uint32_t* __restrict__ only_flagged, uint32_t* __restrict__ seed_redo,
    unsigned long long* __restrict__ seed_fallbacks)
{
    static_assert(K == 33, "the selection takes the 33rd smallest of 64 (wave_33rd)");
    __shared__ SelectLds lds[kSelWaves];
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    SelectLds& S = lds[wid];
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kSelWaves + wid) * kWave;
    if (wave0 >= n) return;
    const int64_t wave = wave0 / kWave;
    if (only_flagged && only_flagged[wave] == 0u) return;
    // Particle, leaf and node indices fit in 32 bits (the launcher checks n).
    const int ni = static_cast<int>(n);
    const int w0 = static_cast<int>(wave0);
    const int nlp = static_cast<int>(nleaf_pad);
    const int q = w0 + lane;
    const bool valid = q < ni;
    // Lanes past the end repeat the last particle, so wavefront-wide box
    // tests need no validity mask.
    const float4 me = spos[valid ? q : ni - 1];
    const int nvalid = ni - w0 < kWave ? ni - w0 : kWave;

    const int last = (ni < w0 + kWave ? ni : w0 + kWave) - 1;
    int l0 = w0 / leaf_size - init_leaves;
    int l1 = last / leaf_size + init_leaves;
    if (l0 < 0) l0 = 0;
    if (l1 > static_cast<int>(nleaf) - 1) l1 = static_cast<int>(nleaf) - 1;
    // With learned bounds (tau_init, an upper estimate of each particle's
    // 33rd squared distance) the selection runs in two passes.  The first
    // (only_flagged null) skips phase A and bounds every query by its learned
    // bound; a query whose bound holds fewer than 33 particles is left
    // unsettled (tau_out -1) and flags its wavefront in seed_redo.  The second
    // (only_flagged = seed_redo) settles those queries with the seed-range
    // bound; the wavefront's other queries keep their results and bound the
    // union by the smaller of their two bounds.
    const bool learned = tau_init && !only_flagged;
    const bool redo_pass = tau_init && only_flagged;
    const bool need = valid && (!redo_pass || tau_out[q] < 0.0f);   // this pass settles the query
    float rA2;
    int n_seed = 0;   // particles of the seed range, whose distances phase A evaluated
    if (learned) {
        rA2 = tau_init[valid ? q : ni - 1];
        if (lane == 0) seed_redo[wave] = 0u;
    } else {
        // Seed of the bound: the wavefront's particles and kSelSeedHalo on
        // either side in Morton order (at most 254 particles).
        const int s0 = w0 - kSelSeedHalo > 0 ? w0 - kSelSeedHalo : 0;
        const int s1 = w0 + kWave + kSelSeedHalo < ni ? w0 + kWave + kSelSeedHalo : ni;
        if (s1 - s0 < K) {
            if (lane == 0) redo[wave] = 1u;
            return;
        }
        n_seed = s1 - s0;

        // ---- A: per-lane bound from the seed range --------------------------
        float blo0 = INFINITY, blo1 = INFINITY, blo2 = INFINITY;
        float bhi0 = -INFINITY, bhi1 = -INFINITY, bhi2 = -INFINITY;
        for (int l = s0 / leaf_size; l <= (s1 - 1) / leaf_size; ++l) {
            const Box b = nodes[nlp + l];
            blo0 = fminf(blo0, b.lo[0]);
            blo1 = fminf(blo1, b.lo[1]);
            blo2 = fminf(blo2, b.lo[2]);
            bhi0 = fmaxf(bhi0, b.hi[0]);
            bhi1 = fmaxf(bhi1, b.hi[1]);
            bhi2 = fmaxf(bhi2, b.hi[2]);
        }
        // Every squared distance within the seed is at most the box
        // diagonal's (plus rounding, under one bin step): bins end at its bin + 1.
        const float diag2 = norm2(bhi0, bhi1, bhi2, blo0, blo1, blo2);
        const int base = static_cast<int>(__float_as_uint(diag2) >> kSelBinShift) + 1 - (kSelBins - 1);
    #pragma unroll
        for (int w = 0; w < kSelHistWords; ++w) S.a.hist[w][lane] = 0u;
        float dmax = 0.0f;
        for (int b0 = s0; b0 < s1; b0 += kWave) {
            const int m = s1 - b0 < kWave ? s1 - b0 : kWave;
            S.a.stage[lane] = lane < m ? spos[b0 + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            wave_lds_fence();
            for (int j = 0; j < m; ++j) {
                const float4 c = S.a.stage[j];
                const float d2 = norm2(me.x, me.y, me.z, c.x, c.y, c.z);
                dmax = fmaxf(dmax, d2);
                const int k = static_cast<int>(__float_as_uint(d2) >> kSelBinShift) - base;
                const int bin = k < 0 ? 0 : (k > kSelBins - 1 ? kSelBins - 1 : k);
                atomicAdd(&S.a.hist[bin >> 2][lane], 1u << ((bin & 3) * 8));
            }
            wave_lds_fence();
        }
        int cum = 0, bsel = kSelBins - 1;
        bool found = false;
    #pragma unroll
        for (int w = 0; w < kSelHistWords; ++w) {
            const uint32_t h = S.a.hist[w][lane];
    #pragma unroll
            for (int t = 0; t < 4; ++t) {
                cum += static_cast<int>((h >> (8 * t)) & 0xffu);
                if (!found && cum >= K) {
                    found = true;
                    bsel = 4 * w + t;
                }
            }
        }
        // The top bin also holds anything clamped into it: its bound is the
        // largest squared distance seen.
        rA2 = bsel == kSelBins - 1 ? dmax : __uint_as_float(static_cast<uint32_t>(base + bsel + 1) << kSelBinShift);
        if (redo_pass && !need) rA2 = fminf(rA2, tau_init[valid ? q : ni - 1]);
    }
    wave_lds_fence();   // the histogram's LDS is reused below

    // ---- B: union of the leaves within some lane's bound ----------------
    // Level by level from the root, one frontier node per lane.  A child is
    // kept when its box is within the largest bound of the box of the
    // wavefront's own leaves (which holds every query); at the leaf level
    // each lane takes one candidate leaf and tests it against every query
    // of the wavefront, giving the leaf's mask of lanes.  Every node of a
    // level lies at the same depth of the complete tree.
    float qlo0 = INFINITY, qlo1 = INFINITY, qlo2 = INFINITY;
    float qhi0 = -INFINITY, qhi1 = -INFINITY, qhi2 = -INFINITY;
    for (int l = w0 / leaf_size; l <= last / leaf_size; ++l) {
        const Box b = nodes[nlp + l];
        qlo0 = fminf(qlo0, b.lo[0]);
        qlo1 = fminf(qlo1, b.lo[1]);
        qlo2 = fminf(qlo2, b.lo[2]);
        qhi0 = fmaxf(qhi0, b.hi[0]);
        qhi1 = fmaxf(qhi1, b.hi[1]);
        qhi2 = fmaxf(qhi2, b.hi[2]);
    }
    const float rmax2 = wave_max(rA2);
    // Lower bound of norm2 between any point of b and any query (monotone
    // rounding, as box_d2).
    const auto near_queries = [&](const Box& b) {
        const float g0 = fmaxf(fmaxf(b.lo[0] - qhi0, qlo0 - b.hi[0]), 0.0f);
        const float g1 = fmaxf(fmaxf(b.lo[1] - qhi1, qlo1 - b.hi[1]), 0.0f);
        const float g2 = fmaxf(fmaxf(b.lo[2] - qhi2, qlo2 - b.hi[2]), 0.0f);
        float d = g0 * g0;
        d = d + g1 * g1;
        d = d + g2 * g2;
        return d <= rmax2;
    };
    // Each lane's mask of the queries whose bound reaches box b.
    const auto query_mask = [&](const Box& b) {
        uint32_t mlo = 0u, mhi = 0u;
        for (int j = 0; j < 32; ++j) {
            const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.x), j));
            const float y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.y), j));
            const float z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.z), j));
            const float rb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rA2), j));
            mlo |= (box_d2(b, x, y, z) <= rb) ? (1u << j) : 0u;
        }
        for (int j = 0; j < 32; ++j) {
            const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.x), 32 + j));
            const float y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.y), 32 + j));
            const float z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.z), 32 + j));
            const float rb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rA2), 32 + j));
            mhi |= (box_d2(b, x, y, z) <= rb) ? (1u << j) : 0u;
        }
        return (static_cast<uint64_t>(mhi) << 32) | mlo;
    };
    int nu = 0;
    uint32_t n_node = 0;
    bool over = false;
    // Masks of `cnt` kept leaves listed in `kept`: lane e takes leaf e and
    // tests it against every query; leaves some query needs join the union.
    const auto add_leaves = [&](const uint32_t* kept, int cnt) {
        for (int cb = 0; cb < cnt; cb += kWave) {
            const int e = cb + lane;
            const bool act = e < cnt;
            const int leaf = act ? static_cast<int>(kept[e]) : 0;
            const uint64_t mk = query_mask(nodes[nlp + leaf]);
            const bool add = act && mk != 0ull;
            const uint64_t ab = wave_ballot(add);
            const int slot = nu + lane_rank(ab);
            if (add && slot < kLeafCap) {
                S.b.uleaf[slot] = static_cast<uint32_t>(leaf);
                S.b.umask[slot] = mk;
            }
            nu += __popcll(ab);
        }
        if (nu > kLeafCap) over = true;
    };
    if (nlp == 1) {
        // one-leaf tree: the root is the leaf
        if (lane == 0) S.b.front1[0] = 0u;
        wave_lds_fence();
        add_leaves(S.b.front1, 1);
    } else {
        uint32_t* front = S.b.front0;
        uint32_t* next = S.b.front1;
        if (lane == 0) front[0] = 1u;
        int nf = 1;
        int level_first = 1;   // index of the first node of the frontier's level
        wave_lds_fence();
        while (nf > 0 && !over) {
            const bool leaves = 2 * level_first >= nlp;
            int nn = 0;
            for (int fb = 0; fb < nf && !over; fb += kWave) {
                const int j = fb + lane;
                const bool act = j < nf;
                const int node = act ? static_cast<int>(front[j]) : 1;
                n_node += static_cast<uint32_t>(__popcll(wave_ballot(act)));
                const int c0 = 2 * node;
                const Box b0 = nodes[c0];
                const Box b1 = nodes[c0 + 1];
                const bool k0 = act && near_queries(b0);
                const bool k1 = act && near_queries(b1);
                const uint64_t m0 = wave_ballot(k0);
                const uint64_t m1 = wave_ballot(k1);
                if (leaves) {
                    // the batch's kept leaves (at most 128), then their masks
                    const int s0r = lane_rank(m0);
                    const int s1r = __popcll(m0) + lane_rank(m1);
                    if (k0) next[s0r] = static_cast<uint32_t>(c0 - nlp);
                    if (k1) next[s1r] = static_cast<uint32_t>(c0 + 1 - nlp);
                    wave_lds_fence();
                    add_leaves(next, __popcll(m0) + __popcll(m1));
                    wave_lds_fence();
                } else {
                    const int s0r = nn + lane_rank(m0);
                    const int s1r = nn + __popcll(m0) + lane_rank(m1);
                    if (k0 && s0r < kLeafCap) next[s0r] = static_cast<uint32_t>(c0);
                    if (k1 && s1r < kLeafCap) next[s1r] = static_cast<uint32_t>(c0 + 1);
                    nn += __popcll(m0) + __popcll(m1);
                }
            }
            if (!leaves && nn > kLeafCap) {
                // The box test against all queries at once kept too many:
                // the level again, keeping only nodes some query's own
                // bound reaches.
                nn = 0;
                for (int fb = 0; fb < nf; fb += kWave) {
                    const int j = fb + lane;
                    const bool act = j < nf;
                    const int node = act ? static_cast<int>(front[j]) : 1;
                    const int c0 = 2 * node;
                    const Box b0 = nodes[c0];
                    const Box b1 = nodes[c0 + 1];
                    const bool k0 = act && near_queries(b0) && query_mask(b0) != 0ull;
                    const bool k1 = act && near_queries(b1) && query_mask(b1) != 0ull;
                    const uint64_t m0 = wave_ballot(k0);
                    const uint64_t m1 = wave_ballot(k1);
                    const int s0r = nn + lane_rank(m0);
                    const int s1r = nn + __popcll(m0) + lane_rank(m1);
                    if (k0 && s0r < kLeafCap) next[s0r] = static_cast<uint32_t>(c0);
                    if (k1 && s1r < kLeafCap) next[s1r] = static_cast<uint32_t>(c0 + 1);
                    nn += __popcll(m0) + __popcll(m1);
                }
            }
            if (nn > kLeafCap) over = true;
            wave_lds_fence();
            if (leaves) break;
            uint32_t* const tmp = front;
            front = next;
            next = tmp;
            nf = nn;
            level_first *= 2;
        }
    }
    if (over) {
        if (lane == 0) redo[wave] = 1u;
        return;
    }
    wave_lds_fence();

    // ---- C: one query at a time -----------------------------------------
    // Three stages per query, each issuing its loads together: the union
    // leaves the query needs (from the masks), the boxes of their groups (up
    // to 128 per round, two loads per lane), and the particles of the groups
    // within the bound (up to 32 groups per round, four loads per lane),
    // whose squared distances within the bound go to a list.  Whenever the
    // list passes 64 entries the bound drops to the 33rd smallest of its
    // first 64 (wave_33rd) and only those at or below it stay, 33 of them
    // with ties; at the end tau is the 33rd smallest of the list.
    const int gpl = leaf_size / kGroup;      // groups per leaf (leaf_size % kGroup == 0)
    const int lpr = 2 * kWave / gpl;         // leaves per round of group boxes
    const uint32_t ngroups = static_cast<uint32_t>((ni + kGroup - 1) / kGroup);
    int pr_leaf[2];
    uint32_t pr_grp[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int pr = k * kWave + lane;     // (leaf, group) pair of the round
        const int l = pr / gpl;
        pr_grp[k] = static_cast<uint32_t>(pr - l * gpl);
        // a pair past the round's leaves (gpl not a power of 2) never has one
        pr_leaf[k] = l < lpr ? l : (1 << 24);
    }
    const int g_row = lane / kGroup;            // group of a batch of 64 particles ...
    const uint32_t g_off = static_cast<uint32_t>(lane % kGroup);   // ... and particle in it
    uint32_t* const slist = S.b.c.slist;
    uint32_t* const glist = S.b.c.glist;
    uint32_t* const cand = S.b.c.cand;
    uint32_t* const cpos = S.b.c.cpos;
    // The union in registers: lane l holds the masks and the first group
    // of entries l + 64 j (mask zero past nu), so a query's test is a shift.
    static_assert(kLeafCap == 4 * kWave, "two chunks of 128 union entries");
    uint64_t um[4];
    uint32_t ug[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int e = j * kWave + lane;
        um[j] = e < nu ? S.b.umask[e] : 0ull;
        ug[j] = e < nu ? S.b.uleaf[e] * static_cast<uint32_t>(gpl) : 0u;
    }
    uint32_t n_eval = 0, n_acc = 0;
    float tau = 0.0f;
    uint32_t my_n = 0;          // with lists: lane i's list length ...
    uint32_t my_p33 = 0;        // ... and the sorted position of its 33rd key's particle
    const uint64_t need_mask = wave_ballot(need);
    uint64_t fail_mask = 0ull;   // first pass: queries whose learned bound held fewer than 33
    for (int i = 0; i < nvalid; ++i) {
        if (redo_pass && ((need_mask >> i) & 1ull) == 0ull) continue;
        float4 qq;
        qq.x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.x), i));
        qq.y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.y), i));
        qq.z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(me.z), i));
        float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rA2), i));
        int cc = 0;
        for (int uc = 0; uc < nu; uc += 2 * kWave) {
            // slist: the first groups of the union leaves the query needs
            int ns = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint64_t m = uc == 0 ? um[h] : um[2 + h];
                const bool has = ((m >> i) & 1ull) != 0ull;
                const uint64_t sb = wave_ballot(has);
                if (has) slist[ns + lane_rank(sb)] = uc == 0 ? ug[h] : ug[2 + h];
                ns += __popcll(sb);
            }
            wave_lds_fence();
            for (int t = 0; t < ns; t += lpr) {
                // Loads are unconditional (clamped indices); lanes without a
                // group fail the test through gv.
                Box gbx[2];
                bool gv[2];
                uint64_t gvm[2];   // ballots of gv, each from compares
                uint32_t gidx[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int r = t + pr_leaf[k];
                    const bool rv = r < ns;
                    const uint32_t g = slist[rv ? r : 0] + pr_grp[k];
                    const bool g_ok = g < ngroups;
                    gv[k] = rv && g_ok;
                    gvm[k] = wave_ballot(rv) & wave_ballot(g_ok);
                    gidx[k] = g;
                    gbx[k] = groups[gv[k] ? g : 0u];
                }
                int ng = 0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const float bd = box_d2(gbx[k], qq.x, qq.y, qq.z);
                    const uint64_t gb = wave_ballot(bd <= r2) & gvm[k];
                    if (gv[k] && bd <= r2)
                        glist[__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(gb >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(gb),
                                                                                  static_cast<uint32_t>(ng)))] = gidx[k];
                    ng += __popcll(gb);
                }
                wave_lds_fence();
                // Rounds of up to 32 groups in four batches of 64 particles
                // (8 groups of kGroup).  A full round gives lane row r the
                // groups u + 4 r .. u + 4 r + 3 in one 16-byte load, one per
                // batch; a partial round gives batch k the groups u + 8 k + r,
                // so that its batches stay full.  Either way batch k holds
                // groups of row r exactly when 8 k + r < ng - u.
                constexpr int kGpb = kWave / kGroup;   // groups per batch
                static_assert(kGpb == 8 && (2 * kWave) % (4 * kGpb) == 0, "rounds of 32 groups");
                for (int u = 0; u < ng; u += 4 * kGpb) {
                    const int left = ng - u;
                    uint32_t g4[4];
                    if (left >= 4 * kGpb) {
                        const uint4 v = *reinterpret_cast<const uint4*>(glist + u + 4 * g_row);
                        g4[0] = v.x;
                        g4[1] = v.y;
                        g4[2] = v.z;
                        g4[3] = v.w;
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; ++k) g4[k] = glist[u + k * kGpb + g_row];
                    }
                    float4 cp[4];
                    bool pv[4];
                    uint64_t pvm[4];   // ballots of pv, each from compares
                    uint32_t p_of[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const uint32_t p = g4[k] * static_cast<uint32_t>(kGroup) + g_off;
                        const bool row_ok = k * kGpb + g_row < left;
                        const bool p_ok = p < static_cast<uint32_t>(ni);
                        pv[k] = row_ok && p_ok;
                        pvm[k] = wave_ballot(row_ok) & wave_ballot(p_ok);
                        p_of[k] = p;
                        cp[k] = spos[pv[k] ? p : static_cast<uint32_t>(w0)];
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (k * kGpb >= left) break;
                        const float d2 = norm2(qq.x, qq.y, qq.z, cp[k].x, cp[k].y, cp[k].z);
                        const bool acc = pv[k] && d2 <= r2;
                        const uint64_t ab = wave_ballot(d2 <= r2) & pvm[k];
                        if (stats) n_eval += static_cast<uint32_t>(__popcll(pvm[k]));
                        if (acc) {
                            const int slot = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                static_cast<uint32_t>(ab >> 32),
                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ab), static_cast<uint32_t>(cc))));
                            cand[slot] = __float_as_uint(d2);
                            cpos[slot] = p_of[k];
                        }
                        cc += __popcll(ab);
                        n_acc += static_cast<uint32_t>(__popcll(ab));
                        wave_lds_fence();
                        while (cc > kWave) {
                            // Bound to the 33rd smallest of the first 64.  Keep
                            // those below it and, at it, all of them when lists
                            // are written (the list of a query holds every
                            // particle within its tau), else enough to make 33;
                            // then the rest (at most 64) at or below it; again
                            // while more than 64 remain.  A pass that keeps
                            // everything (ties at the bound filling the list)
                            // sends the wavefront to the packet walk.
                            const uint32_t x = cand[lane];
                            const uint32_t xp = cpos[lane];
                            const bool more = kWave + lane < cc;
                            const uint32_t xe = more ? cand[kWave + lane] : 0xffffffffu;
                            const uint32_t xep = more ? cpos[kWave + lane] : 0u;
                            const uint32_t kb = static_cast<uint32_t>(
                                __builtin_amdgcn_readfirstlane(static_cast<int>(wave_33rd(x, lane))));
                            r2 = __uint_as_float(kb);
                            const int n_lt = __popcll(wave_ballot(x < kb));
                            const uint64_t eq = wave_ballot(x == kb);
                            const bool keep_x = x < kb || (x == kb && (list || lane_rank(eq) < K - n_lt));
                            const uint64_t kx = wave_ballot(keep_x);
                            const bool keep_e = more && xe <= kb;
                            const uint64_t ke = wave_ballot(keep_e);
                            const int nk = __popcll(kx) + __popcll(ke);
                            if (nk >= cc) {
                                if (lane == 0) redo[wave] = 1u;
                                return;
                            }
                            if (keep_x) {
                                cand[lane_rank(kx)] = x;
                                cpos[lane_rank(kx)] = xp;
                            }
                            if (keep_e) {
                                cand[__popcll(kx) + lane_rank(ke)] = xe;
                                cpos[__popcll(kx) + lane_rank(ke)] = xep;
                            }
                            cc = nk;
                            wave_lds_fence();
                        }
                    }
                }
            }
        }
        if (learned && cc < K) {
            fail_mask |= 1ull << i;   // settled by the second pass
            continue;
        }
        const uint32_t x = lane < cc ? cand[lane] : 0xffffffffu;
        const uint32_t tb = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(wave_33rd(x, lane))));
        const float ti = __uint_as_float(tb);
        tau = (lane == i) ? ti : tau;
        if (list) {
            // The query's list: every candidate at or below tau (the
            // collect walk's list), in any order; its 33rd key: tau and the
            // (33 - n_lt)-th smallest original index among the entries at tau.
            const bool in = lane < cc && x <= tb;
            const uint64_t im = wave_ballot(in);
            const int n_in = __popcll(im);
            if (n_in > kListCap) {
                if (lane == 0) redo[wave] = 1u;
                return;
            }
            const uint32_t xp = cpos[lane];
            if (in) {
                const int r = lane_rank(im);
                list[list_slot(static_cast<int64_t>(w0) + i, r)] = xp;
            }
            const bool at = lane < cc && x == tb;
            const uint64_t am = wave_ballot(at);
            uint32_t p33;   // sorted position of the 33rd key's particle
            if ((am & (am - 1)) == 0ull) {
                // one entry at tau: it is the 33rd
                p33 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(xp), __builtin_ctzll(am)));
            } else {
                // ties at tau: the (33 - n_lt)-th smallest original index
                const int n_lt = __popcll(wave_ballot(lane < cc && x < tb));
                const uint32_t orig = at ? __float_as_uint(spos[xp].w) : 0xffffffffu;
                int rank = 0;
                for (uint64_t m = am; m; m &= m - 1) {
                    const uint32_t o =
                        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(orig), __builtin_ctzll(m)));
                    rank += o < orig ? 1 : 0;
                }
                const uint64_t sel = wave_ballot(at && rank == K - 1 - n_lt);
                p33 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(xp), __builtin_ctzll(sel)));
            }
            my_n = (lane == i) ? static_cast<uint32_t>(n_in) : my_n;
            my_p33 = (lane == i) ? p33 : my_p33;
        }
        wave_lds_fence();
    }
    const bool failed = ((fail_mask >> lane) & 1ull) != 0ull;
    if (need && !failed) {
        tau_out[q] = tau;
        if (list) {
            list_n[q] = my_n;
            rk_sorted[q] = sqrtf(tau);
            kkey_sorted[q] = make_key(tau, __float_as_uint(spos[my_p33].w));
        }
    }
    if (learned) {
        if (failed) tau_out[q] = -1.0f;
        if (lane == 0 && fail_mask != 0ull) {
            seed_redo[wave] = 1u;
            atomicAdd(&seed_fallbacks[wave & (kStatSlots - 1)], static_cast<unsigned long long>(__popcll(fail_mask)));
        }
    }

    // Leaves for the collect walk: the union outside the seed range.
    uint32_t* const my_leaves = wleaf + wave * kLeafCap;
    int n_rec = 0;
    for (int ub = 0; ub < nu; ub += kWave) {
        const int e = ub + lane;
        const int leaf = e < nu ? static_cast<int>(S.b.uleaf[e]) : 0;
        const bool rec = e < nu && (leaf < l0 || leaf > l1);
        const uint64_t rb = wave_ballot(rec);
        if (rec) my_leaves[n_rec + lane_rank(rb)] = static_cast<uint32_t>(leaf);
        n_rec += __popcll(rb);
    }
    if (lane == 0) {
        wleaf_n[wave] = static_cast<uint32_t>(n_rec);
        redo[wave] = 0u;
    }
    if (stats && lane == 0) {
        KnnStats* const st = stats + (wave & (kStatSlots - 1));
        atomicAdd(&st->node_tests, static_cast<unsigned long long>(n_node));
        atomicAdd(&st->leaf_scans, static_cast<unsigned long long>(nu));
        atomicAdd(&st->dist_evals, static_cast<unsigned long long>(n_seed) * static_cast<unsigned long long>(nvalid) +
                                       static_cast<unsigned long long>(n_eval));
        atomicAdd(&st->inserts, static_cast<unsigned long long>(n_acc));
    }
}

// LDS of one wavefront in the collect walk.
template <int M>
struct CollectLds {
    uint32_t list[kWave][M + 1];  // per-lane lists; row stride M + 1 keeps rows in distinct banks
    double pacc[kWave];           // density terms pulled by each lane
    float4 qtab[kWave];           // participating lanes of a leaf: (x, y, z, tau) ...
    float htab[kWave];            // ... their 1/h ...
    uint32_t otab[kWave];         // ... and their lane ids
    uint32_t lcnt[kWave];         // entries within tau
    uint32_t ecnt[kWave];         // entries at exactly tau
    uint32_t tidx[kWave];         // original index of one entry at tau
};

// One leaf of the collect walk.  Only the lanes in `mask` (box distance
// within their tau) take part, and the (lane, particle) pairs are packed
// onto the wavefront: with the leaf padded to P = 64 / G particles, lane l
// tests particle l % P for the participating lanes of rank l / P, l / P + G,
// ...  An accepted pair (d2 <= tau_o for owner o and particle c) is counted
// and listed for o, and its density terms are settled:
//   * c's term on o (o is in c's first 32 slots when d2 < tau_c) is pulled
//     by o into its LDS accumulator;
//   * o's term on c (c in o's first 32 slots when d2 < tau_o) is pulled by
//     c in its own walk when c sees o (d2 <= tau_c), so it is scattered with
//     a global atomic only when d2 > tau_c;
//   * terms at exactly tau_o depend on index order and are left to the end.
template <int M, int P, bool kAll, int KERN>
__device__ __forceinline__ void collect_leaf(CollectLds<M>& L, int start, int count, uint64_t mask, float4 c,
                                             float tc, int lane, float qx, float qy, float qz, float tau, float hinv,
                                             double* __restrict__ dens64, unsigned long long& n_pairs)
{
    constexpr int G = kWave / P;
    const int j = lane % P;
    const int grp = lane / P;
    const int a = __popcll(mask);
    n_pairs += static_cast<unsigned long long>(a) * static_cast<unsigned long long>(count);
    if ((mask >> lane) & 1u) {
        const int k = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
        L.qtab[k] = make_float4(qx, qy, qz, tau);
        L.htab[k] = hinv;
        L.otab[k] = static_cast<uint32_t>(lane);
    }
    wave_lds_fence();
    const float hc = 1.0f / sqrtf(tc);   // c's 1/h, as c's own lane computes it
    for (int k = grp; k < a; k += G) {
        const float4 o = L.qtab[k];
        const float d2 = norm2(o.x, o.y, o.z, c.x, c.y, c.z);
        if (d2 <= o.w) {
            const uint32_t own = L.otab[k];
            const uint32_t slot = atomicAdd(&L.lcnt[own], 1u);
            if (kAll && slot < static_cast<uint32_t>(M)) L.list[own][slot] = static_cast<uint32_t>(start + j);
            if (d2 < o.w) {
                if (dens64 && d2 > tc)
                    unsafeAtomicAdd(&dens64[start + j], static_cast<double>(dens_weight(
                                                            KERN, o.x, o.y, o.z, c.x, c.y, c.z, L.htab[k], d2)));
            } else {
                const uint32_t e = atomicAdd(&L.ecnt[own], 1u);
                L.tidx[own] = __float_as_uint(c.w);
                if (!kAll && e < static_cast<uint32_t>(M)) L.list[own][e] = static_cast<uint32_t>(start + j);
            }
            if (dens64 && d2 < tc)
                atomicAdd(&L.pacc[own], static_cast<double>(dens_weight(KERN, c.x, c.y, c.z, o.x, o.y, o.z, hc, d2)));
        }
    }
    wave_lds_fence();
}

// Walk 2: with tau fixed (from walk 1), visit every leaf whose box distance
// is within some lane's tau, count and list each lane's entries with
// d2 <= tau, settle the density terms (collect_leaf), then derive h, the
// 33rd key, the terms at tau and (optionally) the ordered neighbour list.
// Lists hold every entry (kAll, when the ordered lists are wanted) or only
// the entries at tau.
//
// The leaves are the seed range and those recorded by walk 1, taken 64 at a
// time: lane i loads the box of item i, each box is broadcast to test every
// lane, and the particles of the next participating leaf are loaded while
// the current one is processed.  A wavefront whose walk 1 opened more leaves
// than were recorded walks the tree again instead.  A lane with more
// entries than its list holds sends the wavefront to the insertion kernel,
// which redoes the search and scatters the terms at tau only (everything
// else is settled here).
template <int K, int M, int P, bool kAll, int KERN>
__global__ __launch_bounds__(kCollectThreads) void knn_collect_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad, int64_t nleaf,
    int leaf_size, int init_leaves, const uint32_t* __restrict__ wleaf, const uint32_t* __restrict__ wleaf_n,
    const float* __restrict__ tau_sorted, float* __restrict__ rk_sorted, uint64_t* __restrict__ kkey_sorted,
    double* __restrict__ dens64, double* __restrict__ dens_pull, int32_t* __restrict__ nbr_out,
    uint32_t* __restrict__ wave_flags, KnnStats* __restrict__ stats, bool only_flagged)
{
    __shared__ CollectLds<M> lds[kCollectWaves];
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    CollectLds<M>& L = lds[wid];
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kCollectWaves + wid) * kWave;
    if (wave0 >= n) return;
    const int64_t wave = wave0 / kWave;
    // With only_flagged, the wavefronts whose flag is clear were settled
    // from the selection walk's lists.
    if (only_flagged && wave_flags[wave] == 0u) return;
    const uint32_t n_rec = wleaf_n[wave];
    const int64_t q = wave0 + lane;
    const bool valid = q < n;
    const float4 me = spos[valid ? q : n - 1];
    const float qx = me.x, qy = me.y, qz = me.z;
    // Squared distances and box distances are >= 0, so a past-the-end lane
    // with tau = -1 opens and accepts nothing.
    const float tau = valid ? tau_sorted[q] : -1.0f;
    const float h = sqrtf(tau);
    const float hinv = 1.0f / h;
    L.lcnt[lane] = 0u;
    L.ecnt[lane] = 0u;
    L.pacc[lane] = 0.0;
    wave_lds_fence();

    const int j = lane % P;
    unsigned long long n_leaf = 0, n_pairs = 0;
    if (n_rec <= static_cast<uint32_t>(kLeafCap)) {
        const int64_t last = (n < wave0 + kWave ? n : wave0 + kWave) - 1;
        int64_t l0 = wave0 / leaf_size - init_leaves;
        int64_t l1 = last / leaf_size + init_leaves;
        if (l0 < 0) l0 = 0;
        if (l1 > nleaf - 1) l1 = nleaf - 1;
        const int64_t n_seed = l1 - l0 + 1;
        const uint32_t* const my_leaves = wleaf + wave * kLeafCap;
        const int64_t n_items = n_seed + static_cast<int64_t>(n_rec);
        for (int64_t base = 0; base < n_items; base += kWave) {
            const int nb = static_cast<int>(n_items - base < kWave ? n_items - base : kWave);
            Box mb;
            {
                const int64_t my_it = base + (lane < nb ? lane : 0);
                const int64_t leaf = my_it < n_seed ? l0 + my_it : static_cast<int64_t>(my_leaves[my_it - n_seed]);
                mb = nodes[nleaf_pad + leaf];
            }
            uint32_t m_lo = 0, m_hi = 0;   // lane i: participation mask of item base + i
            for (int i = 0; i < nb; ++i) {
                Box bb;
                bb.lo[0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.lo[0]), i));
                bb.lo[1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.lo[1]), i));
                bb.lo[2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.lo[2]), i));
                bb.hi[0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.hi[0]), i));
                bb.hi[1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.hi[1]), i));
                bb.hi[2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mb.hi[2]), i));
                const uint64_t mk = __ballot(box_d2(bb, qx, qy, qz) <= tau);
                m_lo = (lane == i) ? static_cast<uint32_t>(mk) : m_lo;
                m_hi = (lane == i) ? static_cast<uint32_t>(mk >> 32) : m_hi;
            }
            uint64_t part = __ballot((m_lo | m_hi) != 0u);
            if (part == 0) continue;
            int nxt = __builtin_ctzll(part);
            int nstart = __builtin_amdgcn_readlane(mb.lo_w, nxt);
            int ncount = __builtin_amdgcn_readlane(mb.hi_w, nxt);
            float4 cn = (j < ncount) ? spos[nstart + j] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            float tn = (j < ncount) ? tau_sorted[nstart + j] : 0.0f;
            while (part != 0) {
                const int cur = nxt;
                const int start = nstart;
                const int count = ncount;
                const float4 c = cn;
                const float tc = tn;
                part &= part - 1;
                if (part != 0) {
                    nxt = __builtin_ctzll(part);
                    nstart = __builtin_amdgcn_readlane(mb.lo_w, nxt);
                    ncount = __builtin_amdgcn_readlane(mb.hi_w, nxt);
                    cn = (j < ncount) ? spos[nstart + j] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
                    tn = (j < ncount) ? tau_sorted[nstart + j] : 0.0f;
                }
                const uint64_t mask = (static_cast<uint64_t>(__builtin_amdgcn_readlane(m_hi, cur)) << 32) |
                                      static_cast<uint32_t>(__builtin_amdgcn_readlane(m_lo, cur));
                ++n_leaf;
                collect_leaf<M, P, kAll, KERN>(L, start, count, mask, c, tc, lane, qx, qy, qz, tau, hinv, dens64,
                                               n_pairs);
            }
        }
    } else {
        // More leaves than walk 1 recorded: walk the tree with tau.
        int stack = (lane == 0) ? 1 : 0;
        int sp = 1;
        while (sp > 0) {
            --sp;
            const int node = __builtin_amdgcn_readlane(stack, sp);
            const Box b = nodes[node];
            const uint64_t mask = __ballot(box_d2(b, qx, qy, qz) <= tau);
            if (mask == 0) continue;
            if (node >= nleaf_pad) {
                ++n_leaf;
                const int start = b.lo_w;
                const int count = b.hi_w;
                const float4 c = (j < count) ? spos[start + j] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
                const float tc = (j < count) ? tau_sorted[start + j] : 0.0f;
                collect_leaf<M, P, kAll, KERN>(L, start, count, mask, c, tc, lane, qx, qy, qz, tau, hinv, dens64,
                                               n_pairs);
            } else {
                stack = (lane == sp) ? 2 * node + 1 : stack;
                ++sp;
                stack = (lane == sp) ? 2 * node : stack;
                ++sp;
            }
        }
    }
    wave_lds_fence();
    const int cnt = static_cast<int>(L.lcnt[lane]);
    const int c_eq = static_cast<int>(L.ecnt[lane]);
    if (dens64 && valid) dens_pull[q] = L.pacc[lane];

    if (stats && lane == 0) {
        KnnStats* const st = stats + (wave & (kStatSlots - 1));
        atomicAdd(&st->leaf_scans, n_leaf);
        atomicAdd(&st->dist_evals, n_pairs);
    }
    const int n_list = kAll ? cnt : c_eq;
    const bool over = wave_any(n_list > M);
    if (lane == 0) wave_flags[wave] = over ? kRedoTies : 0u;
    if (over || !valid) return;

    rk_sorted[q] = h;
    // Entries below tau are slots 0 .. c_lt-1 (c_lt <= 32); the c_eq >= 1
    // entries at tau follow in index order.
    const int c_lt = cnt - c_eq;
    uint64_t kk = make_key(tau, L.tidx[lane]);
    if (c_eq > 1) {
        // The first 32 - c_lt ties are scattered, the one at slot 32 is the
        // 33rd key.
        for (int i = 0; i < n_list; ++i) {
            const uint32_t si = L.list[lane][i];
            const float4 c = spos[si];
            const float d2c = norm2(qx, qy, qz, c.x, c.y, c.z);
            if (d2c < tau) continue;
            const uint32_t idx = __float_as_uint(c.w);
            int r = c_lt;
            for (int k = 0; k < n_list; ++k) {
                const float4 o = spos[L.list[lane][k]];
                if (norm2(qx, qy, qz, o.x, o.y, o.z) == tau && __float_as_uint(o.w) < idx) ++r;
            }
            if (r < K - 1 && dens64)
                unsafeAtomicAdd(&dens64[si],
                                static_cast<double>(dens_weight(KERN, qx, qy, qz, c.x, c.y, c.z, hinv, d2c)));
            if (r == K - 1) kk = make_key(tau, idx);
        }
    }
    kkey_sorted[q] = kk;
    if (kAll && nbr_out) {
        const int64_t oq = static_cast<int64_t>(__float_as_uint(me.w));
        for (int i = 0; i < cnt; ++i) {
            const float4 c = spos[L.list[lane][i]];
            const uint64_t ki = make_key(norm2(qx, qy, qz, c.x, c.y, c.z), __float_as_uint(c.w));
            int r = 0;
            for (int k = 0; k < cnt; ++k) {
                const float4 o = spos[L.list[lane][k]];
                if (make_key(norm2(qx, qy, qz, o.x, o.y, o.z), __float_as_uint(o.w)) < ki) ++r;
            }
            if (r < K) nbr_out[oq * K + r] = static_cast<int32_t>(__float_as_uint(c.w));
        }
    }
}

// Density from the selection walk's lists (wavefronts whose flag is clear;
// the collect walk takes the others).  Each particle q has a row of 8 lanes
// (eight particles per wavefront), lane l taking entries l, l + 8, ... of
// q's list: every particle c with d2 <= tau_q.  (Rows of 8 lanes leave fewer
// lanes idle in a list's last round than rows of 16: 3.2 against 3.9 ms for
// the 256^3 workload; rows of 4 take 3.7 ms.  A lane issues the loads of
// all its entries before using any: 2.5 against 3.2 ms.)  The terms are settled as in
// collect_leaf, with q as the owner: c's term on q is pulled when d2 < tau_c
// (summed over the row in FP64); q's term on c is scattered with an FP64
// atomic when d2 < tau_q and d2 > tau_c (c does not list q), and at
// d2 = tau_q when c's key is below q's 33rd key (a tie in q's first 32).
constexpr int kListThreads = 256;
constexpr int kListRow = 8;

template <int KERN>
__global__ __launch_bounds__(kListThreads) void knn_list_dens_kernel(
    const float4* __restrict__ spos, int64_t n, const float* __restrict__ tau_sorted,
    const uint64_t* __restrict__ kkey_sorted, const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_n,
    const uint32_t* __restrict__ wave_flags, double* __restrict__ dens64, double* __restrict__ dens_pull)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kListThreads + threadIdx.x;
    const int64_t q = t / kListRow;
    const int lr = static_cast<int>(t % kListRow);
    // A row shares q: rows past the end and rows of flagged
    // wavefronts leave together, so the row sum below runs on whole rows.
    if (q >= n || wave_flags[q / kWave] != 0u) return;
    const float4 me = spos[q];
    const float tau = tau_sorted[q];
    const float hinv = 1.0f / sqrtf(tau);
    const uint64_t kk = kkey_sorted[q];
    const int cnt = static_cast<int>(list_n[q]);
    // The lane's entries, then their neighbours' positions and tau, each
    // set of loads issued together (entries past the list read q itself).
    constexpr int kPer = kListCap / kListRow;
    static_assert(kListCap % kListRow == 0, "whole rounds of a row");
    uint32_t js[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int e = k * kListRow + lr;
        js[k] = e < cnt ? list[list_slot(q, e)] : static_cast<uint32_t>(q);
    }
    float4 cs[kPer];
    float tcs[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        cs[k] = spos[js[k]];
        tcs[k] = tau_sorted[js[k]];
    }
    double pacc = 0.0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (k * kListRow + lr >= cnt) break;
        const uint32_t j = js[k];
        const float4 c = cs[k];
        const float tc = tcs[k];
        const float d2 = norm2(me.x, me.y, me.z, c.x, c.y, c.z);
        if (d2 < tc) {
            const float hc = 1.0f / sqrtf(tc);   // c's 1/h, as c's own lane computes it
            pacc += static_cast<double>(dens_weight(KERN, c.x, c.y, c.z, me.x, me.y, me.z, hc, d2));
        }
        const bool give = d2 < tau ? d2 > tc : make_key(d2, __float_as_uint(c.w)) < kk;
        if (give)
            unsafeAtomicAdd(&dens64[j], static_cast<double>(dens_weight(KERN, me.x, me.y, me.z, c.x, c.y, c.z, hinv, d2)));
    }
#pragma unroll
    for (int o = kListRow / 2; o > 0; o >>= 1) pacc += __shfl_xor(pacc, o, kListRow);
    if (lr == 0) dens_pull[q] = pacc;
}

// Walk 1 with a learned bound, as a range walk.  Each lane's learned bound
// tau_init is an upper estimate of its 33rd squared distance; the walk
// evaluates only the (lane, particle) pairs of leaves whose box distance is
// within the lane's current bound, packed onto the wavefront as in the
// collect walk (lane l tests particle l % P of the leaf for the
// participating lanes of rank l / P, l / P + G, ...).  Accepted squared
// distances go to the owner's list in LDS; a drain inserts every lane's list
// into its 33 registers (med3_insert, as walk 1) and lowers the lane's bound
// to t[32].  A drain runs before any leaf that could overflow a list and at
// the end, so a list never holds more than kSeedSlots entries.  Every
// particle with d2 below the final t[32] lies in a leaf within every bound the
// lane had, so it was inserted: when t[32] < tau_init the 33 smallest are
// exact; otherwise the wavefront is flagged in seed_redo for walk 1 without a
// bound, as with the bounded walk 1.  Leaves outside the seed range of the
// collect walk are recorded for it, as walk 1 records them.
template <int P>
struct SeedLds {
    static constexpr int kSlots = P + 16;
    float list[kWave][kSlots + 1];   // row stride kSlots + 1 keeps the rows in distinct banks
    uint32_t cnt[kWave];
    float4 qtab[kWave];              // participating lanes of a leaf: (x, y, z, bound) ...
    uint32_t otab[kWave];            // ... and their lane ids
};

template <int K, int P>
__global__ __launch_bounds__(kCollectThreads) void knn_seed_walk_kernel(
    const float4* __restrict__ spos, int64_t n, const Box* __restrict__ nodes, int64_t nleaf_pad, int64_t nleaf,
    int leaf_size, int init_leaves, const float* __restrict__ tau_init, float* __restrict__ tau_out,
    uint32_t* __restrict__ wleaf, uint32_t* __restrict__ wleaf_n, uint32_t* __restrict__ seed_redo,
    unsigned long long* __restrict__ seed_fallbacks, KnnStats* __restrict__ stats)
{
    constexpr int G = kWave / P;
    constexpr int C = SeedLds<P>::kSlots;
    __shared__ SeedLds<P> lds[kCollectWaves];
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    SeedLds<P>& L = lds[wid];
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kCollectWaves + wid) * kWave;
    if (wave0 >= n) return;
    const int64_t wave = wave0 / kWave;
    const int64_t q = wave0 + lane;
    const bool valid = q < n;
    const float4 me = spos[valid ? q : n - 1];
    const float qx = me.x, qy = me.y, qz = me.z;
    // A past-the-end lane has bound -1: box distances are >= 0, so it takes
    // part in nothing.
    const float t0 = valid ? tau_init[q] : -1.0f;
    float t[K];
#pragma unroll
    for (int i = 0; i < K; ++i) t[i] = t0;
    float bound = t0;
    L.cnt[lane] = 0u;
    wave_lds_fence();

    const int j = lane % P;
    const int grp = lane / P;
    const int64_t last = (n < wave0 + kWave ? n : wave0 + kWave) - 1;
    int64_t l0 = wave0 / leaf_size - init_leaves;
    int64_t l1 = last / leaf_size + init_leaves;
    if (l0 < 0) l0 = 0;
    if (l1 > nleaf - 1) l1 = nleaf - 1;
    uint32_t rec = 0;
    int n_rec = 0;
    uint32_t* const my_leaves = wleaf + wave * kLeafCap;
    unsigned long long n_leaf = 0, n_pairs = 0, n_node = 0;

    // Insert every lane's list into its registers and lower its bound.
    const auto drain = [&]() {
        const int mine = static_cast<int>(L.cnt[lane]);
        int mx = mine;
        for (int o = kWave / 2; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
        for (int i = 0; i < mx; ++i) med3_insert<K>(t, i < mine ? L.list[lane][i] : INFINITY);
        bound = t[K - 1];
        wave_lds_fence();
        L.cnt[lane] = 0u;
        wave_lds_fence();
    };

    int stack = (lane == 0) ? 1 : 0;
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int64_t node = __builtin_amdgcn_readlane(stack, sp);
        ++n_node;
        if (node < nleaf_pad) {
            // Near child (by lane vote) on top, visited first; a child is
            // pushed only when some lane's bound reaches its box.
            const Box b0 = nodes[2 * node];
            const Box b1 = nodes[2 * node + 1];
            const float d0 = box_d2(b0, qx, qy, qz);
            const float d1 = box_d2(b1, qx, qy, qz);
            const bool need0 = wave_any(d0 <= bound);
            const bool need1 = wave_any(d1 <= bound);
            const bool near0 = 2 * __popcll(wave_ballot(d0 <= d1)) >= kWave;
            const int nearc = static_cast<int>(near0 ? 2 * node : 2 * node + 1);
            const bool need_near = near0 ? need0 : need1;
            const bool need_far = near0 ? need1 : need0;
            stack = (lane == sp) ? (need_far ? (nearc ^ 1) : nearc) : stack;
            stack = (lane == sp + 1) ? nearc : stack;
            sp += static_cast<int>(need_far) + static_cast<int>(need_near);
            continue;
        }
        const Box b = nodes[node];
        const int start = b.lo_w;
        const int count = b.hi_w;
        if (!wave_any(box_d2(b, qx, qy, qz) <= bound)) continue;
        // Room for one more entry per particle of the leaf in every list.
        if (wave_any(static_cast<int>(L.cnt[lane]) > C - count)) drain();
        const uint64_t mask = wave_ballot(box_d2(b, qx, qy, qz) <= bound);
        if (mask == 0) continue;
        const int64_t leaf = node - nleaf_pad;
        ++n_leaf;
        if (leaf < l0 || leaf > l1) {
            if (n_rec < kLeafCap) {
                rec = (lane == (n_rec & (kWave - 1))) ? static_cast<uint32_t>(leaf) : rec;
                if ((n_rec & (kWave - 1)) == kWave - 1) my_leaves[(n_rec & ~(kWave - 1)) + lane] = rec;
            }
            ++n_rec;
        }
        const float4 c = (j < count) ? spos[start + j] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
        const int a = __popcll(mask);
        n_pairs += static_cast<unsigned long long>(a) * static_cast<unsigned long long>(count);
        if ((mask >> lane) & 1u) {
            const int k = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
            L.qtab[k] = make_float4(qx, qy, qz, bound);
            L.otab[k] = static_cast<uint32_t>(lane);
        }
        wave_lds_fence();
        for (int k = grp; k < a; k += G) {
            const float4 o = L.qtab[k];
            const float d2 = norm2(o.x, o.y, o.z, c.x, c.y, c.z);
            if (d2 < o.w) {
                const uint32_t own = L.otab[k];
                const uint32_t slot = atomicAdd(&L.cnt[own], 1u);
                L.list[own][slot] = d2;
            }
        }
        wave_lds_fence();
    }
    drain();

    if (n_rec <= kLeafCap && (n_rec & (kWave - 1)) != 0 && lane < (n_rec & (kWave - 1)))
        my_leaves[(n_rec & ~(kWave - 1)) + lane] = rec;
    if (lane == 0) wleaf_n[wave] = static_cast<uint32_t>(n_rec);
    if (valid) tau_out[q] = t[K - 1];
    const uint64_t fb = wave_ballot(valid && !(t[K - 1] < t0));
    if (lane == 0) {
        seed_redo[wave] = fb ? 1u : 0u;
        if (fb) atomicAdd(&seed_fallbacks[wave & (kStatSlots - 1)], static_cast<unsigned long long>(__popcll(fb)));
    }
    if (stats && lane == 0) {
        KnnStats* const st = stats + (wave & (kStatSlots - 1));
        atomicAdd(&st->node_tests, n_node);
        atomicAdd(&st->leaf_scans, n_leaf);
        atomicAdd(&st->dist_evals, n_pairs);
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
    const Box* __restrict__ nodes, const float* __restrict__ kmax, int64_t nleaf_pad, float* __restrict__ dens, int kern)
{
    __shared__ float4 stage_p[kWavesPerBlock][kWave];
    __shared__ float4 stage_k[kWavesPerBlock][kWave];   // (kkey lo, kkey hi, hinv, -)
    const int lane = threadIdx.x % kWave;
    const int wid = wave_in_block();
    const int64_t wave0 = (static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock + wid) * kWave;
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
                    acc += static_cast<double>(dens_weight(kern, p.x, p.y, p.z, qx, qy, qz, pk.z, d2));
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

__global__ void dens_finalize_kernel(const double* __restrict__ d64, const double* __restrict__ pull, int64_t n,
                                     float* __restrict__ d)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = static_cast<float>(pull ? d64[i] + pull[i] : d64[i]);
}

}  // namespace

hipError_t launch_dens_finalize(const double* dens64, const double* dens_pull, int64_t n, float* dens, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dens_finalize_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(dens64, dens_pull, n, dens);
    return hipGetLastError();
}

hipError_t launch_knn(const float4* spos, int64_t n, const Box* nodes, int64_t nleaf_pad, int64_t nleaf,
                      int leaf_size, float* rk_sorted, uint64_t* kkey_sorted, const int32_t* inv_perm, double* dens64,
                      int32_t* nbr_out_orig, KnnStats* stats, const KnnScratch& scratch, int method, hipStream_t s,
                      const KnnMarks* marks)
{
    if (n <= 0) return hipSuccess;
    auto mark_begin = [&](int st) {
        if (marks && marks->begin) marks->begin(marks->user, st);
    };
    auto mark_end = [&](int st) {
        if (marks && marks->end) marks->end(marks->user, st);
    };
    const int64_t waves = (n + kWave - 1) / kWave;
    const int init_leaves = leaf_size >= kWave ? 1 : kWave / leaf_size;
    const unsigned ins_blocks = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (method == kKnnInsertion) {
        knn_dens_kernel<kSlots><<<ins_blocks, kKnnThreads, 0, s>>>(spos, n, nodes, nleaf_pad, nleaf, leaf_size,
                                                                  init_leaves, rk_sorted, kkey_sorted, inv_perm,
                                                                  dens64, nbr_out_orig, stats, nullptr,
                                                                  scratch.kernel);
        return hipGetLastError();
    }
    if (leaf_size < 1 || leaf_size > kWave) return hipErrorInvalidValue;
    if (n > 0x7fffffffLL - 2 * kWave) return hipErrorInvalidValue;   // 32-bit indices in the walks
    // Group boxes line up with leaves (and with the 64-particle staging
    // batches) only when the leaf size is a multiple of kGroup.
    const Box* const groups = (leaf_size % kGroup == 0) ? scratch.groups : nullptr;
    const unsigned bound_blocks = static_cast<unsigned>((waves + kBoundWaves - 1) / kBoundWaves);
    const unsigned col_blocks = static_cast<unsigned>((waves + kCollectWaves - 1) / kCollectWaves);
    // The selection walk lists each particle's neighbours when no ordered
    // lists are kept; the collect walk then runs only where it could not.
    // The selection walk runs without a seed model, or with one when the
    // learned bounds are to bound it (seed_walk 2).
    const bool select = scratch.bound == kBoundSelect && groups && (!scratch.tau_init || scratch.seed_walk == 2);
    const bool lists = select && scratch.list && scratch.list_n && !nbr_out_orig;
    mark_begin(0);
    if (scratch.tau_init && scratch.seed_walk == 1) {
#define ASM_SEED_WALK(PAD)                                                                                        \
    knn_seed_walk_kernel<kSlots, PAD><<<col_blocks, kCollectThreads, 0, s>>>(                                     \
        spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau_init, scratch.tau, scratch.wleaf,   \
        scratch.wleaf_n, scratch.seed_redo, scratch.seed_fallbacks, stats)
        if (leaf_size <= 8) ASM_SEED_WALK(8);
        else if (leaf_size <= 16) ASM_SEED_WALK(16);
        else if (leaf_size <= 32) ASM_SEED_WALK(32);
        else ASM_SEED_WALK(64);
#undef ASM_SEED_WALK
    } else if (select) {
        const unsigned sel_blocks = static_cast<unsigned>((waves + kSelWaves - 1) / kSelWaves);
        knn_select_kernel<kSlots><<<sel_blocks, kSelThreads, 0, s>>>(
            spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau, scratch.wleaf, scratch.wleaf_n,
            scratch.wave_flags, stats, groups, lists ? scratch.list : nullptr, scratch.list_n, rk_sorted, kkey_sorted,
            scratch.tau_init, nullptr, scratch.seed_redo, scratch.seed_fallbacks);
        ASM_HIP_CHECK(hipGetLastError());
        if (scratch.tau_init) {
            // Wavefronts with a query whose learned bound held fewer than 33:
            // those queries again, with the seed-range bound.
            knn_select_kernel<kSlots><<<sel_blocks, kSelThreads, 0, s>>>(
                spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau, scratch.wleaf,
                scratch.wleaf_n, scratch.wave_flags, stats, groups, lists ? scratch.list : nullptr, scratch.list_n,
                rk_sorted, kkey_sorted, scratch.tau_init, scratch.seed_redo, nullptr, nullptr);
            ASM_HIP_CHECK(hipGetLastError());
        }
        // Wavefronts the selection flagged (union past kLeafCap leaves, or a
        // seed range under 33 particles): the packet walk; every other
        // wavefront returns at once.  The collect walk rewrites the flags.
        knn_bound_kernel<kSlots><<<bound_blocks, kBoundThreads, 0, s>>>(
            spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau, scratch.wleaf, scratch.wleaf_n,
            stats, nullptr, scratch.wave_flags, nullptr, nullptr, groups);
    } else {
        knn_bound_kernel<kSlots><<<bound_blocks, kBoundThreads, 0, s>>>(
            spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau, scratch.wleaf, scratch.wleaf_n,
            stats, scratch.tau_init, nullptr, scratch.seed_redo, scratch.seed_fallbacks, groups);
    }
    ASM_HIP_CHECK(hipGetLastError());
    if (scratch.tau_init && !select) {
        // Wavefronts with a lane whose learned bound was too small: again, unbounded.
        knn_bound_kernel<kSlots><<<bound_blocks, kBoundThreads, 0, s>>>(
            spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.tau, scratch.wleaf, scratch.wleaf_n,
            stats, nullptr, scratch.seed_redo, nullptr, nullptr, groups);
        ASM_HIP_CHECK(hipGetLastError());
    }
    mark_end(0);
    mark_begin(1);
    if (lists && dens64) {
        const unsigned lb = static_cast<unsigned>((n * kListRow + kListThreads - 1) / kListThreads);
        if (scratch.kernel == kKernSpline)
            knn_list_dens_kernel<kKernSpline><<<lb, kListThreads, 0, s>>>(spos, n, scratch.tau, kkey_sorted, scratch.list,
                                                                         scratch.list_n, scratch.wave_flags, dens64,
                                                                         scratch.dens_pull);
        else
            knn_list_dens_kernel<kKernTent><<<lb, kListThreads, 0, s>>>(spos, n, scratch.tau, kkey_sorted, scratch.list,
                                                                       scratch.list_n, scratch.wave_flags, dens64,
                                                                       scratch.dens_pull);
        ASM_HIP_CHECK(hipGetLastError());
    }
    // The density kernel is a template argument of the collect walk: the
    // default tent path carries no test of the kernel in its inner loop.
#define ASM_COLLECT_K(M, PAD, ALL, KERN)                                                                        \
    knn_collect_kernel<kSlots, M, PAD, ALL, KERN><<<col_blocks, kCollectThreads, 0, s>>>(                       \
        spos, n, nodes, nleaf_pad, nleaf, leaf_size, init_leaves, scratch.wleaf, scratch.wleaf_n, scratch.tau,  \
        rk_sorted, kkey_sorted, dens64, scratch.dens_pull, nbr_out_orig, scratch.wave_flags, stats, lists)
#define ASM_COLLECT(M, PAD, ALL)                                                                                \
    do {                                                                                                        \
        if (scratch.kernel == kKernSpline) ASM_COLLECT_K(M, PAD, ALL, kKernSpline);                             \
        else ASM_COLLECT_K(M, PAD, ALL, kKernTent);                                                             \
    } while (0)
#define ASM_COLLECT_PAD(M, ALL)                  \
    if (leaf_size <= 8) ASM_COLLECT(M, 8, ALL);  \
    else if (leaf_size <= 16) ASM_COLLECT(M, 16, ALL); \
    else if (leaf_size <= 32) ASM_COLLECT(M, 32, ALL); \
    else ASM_COLLECT(M, 64, ALL);
    if (nbr_out_orig) {
        ASM_COLLECT_PAD(kCollectSlots, true)
    } else {
        ASM_COLLECT_PAD(kTieSlots, false)
    }
#undef ASM_COLLECT_PAD
#undef ASM_COLLECT
#undef ASM_COLLECT_K
    ASM_HIP_CHECK(hipGetLastError());
    mark_end(1);
    // Wavefronts flagged by the collect walk are redone by the insertion
    // kernel; every other wavefront returns at once.
    mark_begin(2);
    knn_dens_kernel<kSlots><<<ins_blocks, kKnnThreads, 0, s>>>(spos, n, nodes, nleaf_pad, nleaf, leaf_size,
                                                              init_leaves, rk_sorted, kkey_sorted, inv_perm, dens64,
                                                              nbr_out_orig, stats, scratch.wave_flags, scratch.kernel);
    ASM_HIP_CHECK(hipGetLastError());
    mark_end(2);
    return hipSuccess;
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
                              const float* kmax, int64_t nleaf_pad, float* dens, int kern, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t waves = (n + kWave - 1) / kWave;
    const unsigned blocks = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    dens_gather_kernel<<<blocks, kKnnThreads, 0, s>>>(spos, kkey, rk, n, nodes, kmax, nleaf_pad, dens, kern);
    return hipGetLastError();
}

}  // namespace asm_hip
