// Mesh deposition of the adaptive tent kernel, mesh reductions, normalisation.
//
// Replaces clear_mesh / assign_on_mesh / mesh_total (coarse_mesh.cpp:24-127),
// the periodic variant assign_on_global_mesh (global_mesh.cpp:25-123) and the
// normalisation in write_mesh_separate (coarse_mesh.cpp:150-171).
//
// The mesh is cut into cubic tiles of `tile` cells.  Every particle is listed
// in each tile its cell range overlaps (count, scan, fill).  One workgroup then
// owns one tile: it accumulates every (particle, cell) contribution into FP64
// accumulators in LDS with native ds_add_f64, and flushes the tile once with
// plain stores - no global atomics, every cell written exactly once.  The
// cell range and each cell's weight ker3 = (wx * wy) * wz are the reference's
// FP32 arithmetic, bit for bit (coarse_mesh.cpp:72-75, weights :80-93).  The
// weight's products with the fields (:95-122) are the reference's FP32
// products in the periodic variant and the FP32 study modes; the node-mesh
// tile kernel forms them exactly, as FP64 FMAs of the FP32 weight and field,
// instead of rounding each to FP32 (a sum moves by at most 2^-24 of the sum
// of its terms' magnitudes, well under the reference's own FP32
// accumulation).  The sums are carried in FP64, so each stored cell is the
// FP32 rounding of an FP64 sum.
#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace asm_hip {
namespace {

constexpr int kTileMax = 8;                 // cells per tile side (LDS sized for 8^3)
constexpr int kTileCells = kTileMax * kTileMax * kTileMax;
// Padded accumulator layout of the packed kernel: cell (x, y, z) of a tile
// at x + kPadY * y + kPadZ * z.  With rows of 8 doubles the five or so rows
// of one particle's box start on the same two of the 32 banks, so the lanes
// of a sweep pile onto a few banks; an odd row pitch spreads them.
constexpr int kPadY = kTileMax + 1;
constexpr int kPadZ = kTileMax * kPadY;
constexpr int kPadCells = kTileMax * kPadZ;
constexpr int kDepThreads = 256;

__device__ __forceinline__ void wave_lds_fence_dep()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// coarse_mesh.cpp:10-15.
__device__ __forceinline__ float tent_f(float x)
{
    return x >= 1.0f ? 0.5f : x - 0.5f * x * x;
}

// coarse_mesh.cpp:17-22: integral of the unit tent over [x1, x2].
__device__ __forceinline__ float ker_integ(float x1, float x2)
{
    const float fac1 = (x1 > 0.0f) ? 1.0f : -1.0f;
    const float fac2 = (x2 > 0.0f) ? 1.0f : -1.0f;
    return fac2 * tent_f(fac2 * x2) - fac1 * tent_f(fac1 * x1);
}

// The cubic spline option on the mesh: its separable form, the product over
// the three axes of the 1-D cubic B-spline of half-width h normalised to one,
//   w(t) = 4/3 (1 - 6 t^2 + 6 |t|^3) for |t| < 1/2, 8/3 (1 - |t|)^3 for
//   1/2 <= |t| <= 1,
// integrated exactly over each cell as the tent is (coarse_mesh.cpp:10-22):
// G(u) = int_0^u w, G(u) = 4/3 (u - 2 u^3 + 3/2 u^4) below 1/2,
// 1/2 - 2/3 (1 - u)^4 up to 1, then 1/2; the cell weight is
// F(x2) - F(x1) with F(t) = sign(t) G(|t|).  The support and so the cell
// ranges are the tent's.
__device__ __forceinline__ float spline_g(float u)
{
    if (u >= 1.0f) return 0.5f;
    if (u < 0.5f) {
        const float u2 = u * u;
        const float a = 2.0f * (u2 * u);
        const float b = 1.5f * (u2 * u2);
        return (4.0f / 3.0f) * ((u - a) + b);
    }
    const float r1 = 1.0f - u;
    const float r2 = r1 * r1;
    return 0.5f - (2.0f / 3.0f) * (r2 * r2);
}

__device__ __forceinline__ float spline_integ(float x1, float x2)
{
    const float fac1 = (x1 > 0.0f) ? 1.0f : -1.0f;
    const float fac2 = (x2 > 0.0f) ? 1.0f : -1.0f;
    return fac2 * spline_g(fac2 * x2) - fac1 * spline_g(fac1 * x1);
}

// The cell integral of the context's kernel (uniform across a launch).
__device__ __forceinline__ float cell_integ(int kern, float x1, float x2)
{
    return kern == kKernSpline ? spline_integ(x1, x2) : ker_integ(x1, x2);
}

// Cell range [lo, hi) of one axis.  Node mesh: coarse_mesh.cpp:72-75
// (truncation toward zero, clipped to [0, nc)).  Periodic variant:
// global_mesh.cpp:66-69 (floor / ceil, unclipped, wrapped at use).
__device__ __forceinline__ void axis_range(float x, float h, float boxsize, float ncf, int nc,
                                           int periodic, int& lo, int& hi)
{
    if (!periodic) {
        const int l = static_cast<int>((x - h) / boxsize * ncf);
        const int r = 1 + static_cast<int>((x + h) / boxsize * ncf);
        lo = l > 0 ? l : 0;
        hi = r < nc ? r : nc;
    } else {
        lo = static_cast<int>(floorf((x - h) / boxsize * ncf));
        hi = 1 + static_cast<int>(ceilf((x + h) / boxsize * ncf));
    }
}

struct PartRange {
    float x[3];
    float h;
    int lo[3];
    int hi[3];
    bool active;
};

// Position as deposited and cell ranges for particle i.
__device__ __forceinline__ PartRange part_range(const DepositArgs& a, int64_t i)
{
    PartRange r;
    const float4 p = a.spos[i];
    r.h = a.rk[i];
    r.x[0] = p.x;
    r.x[1] = p.y;
    r.x[2] = p.z;
    r.active = true;
    if (a.periodic) {
        // Local particles only, shifted by the node offset (global_mesh.cpp:43-56).
        if (p.x < 0.0f || p.x >= a.boxsize || p.y < 0.0f || p.y >= a.boxsize || p.z < 0.0f ||
            p.z >= a.boxsize)
            r.active = false;
        r.x[0] = p.x + a.boxsize * static_cast<float>(a.coord[0]);
        r.x[1] = p.y + a.boxsize * static_cast<float>(a.coord[1]);
        r.x[2] = p.z + a.boxsize * static_cast<float>(a.coord[2]);
    }
    const float ncf = static_cast<float>(a.nc);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        axis_range(r.x[d], r.h, a.boxsize, ncf, a.nc, a.periodic, r.lo[d], r.hi[d]);
        if (r.lo[d] >= r.hi[d]) r.active = false;
    }
    return r;
}

// Tiles whose cells a particle's range [lo, hi) touches, per axis, as one or
// two disjoint tile intervals.  Node mesh: one interval.  Periodic variant:
// the unwrapped range (around coord * nc, global_mesh.cpp:43-44, 66-69) is
// taken modulo nc (:85-87): the whole axis when it spans nc cells or more,
// else an arc that may wrap past nc - 1 into two pieces, merged into one
// interval when they meet, so that no tile is listed twice (nc need not be a
// multiple of the tile side).
struct AxisTiles {
    int a[2], b[2];  // inclusive tile intervals
    int n;
};

__device__ __forceinline__ int floor_div(int x, int m)
{
    const int q = x / m;
    return (x % m != 0 && x < 0) ? q - 1 : q;
}

__device__ __forceinline__ AxisTiles axis_tiles(const DepositArgs& a, int lo, int hi)
{
    AxisTiles t;
    t.n = 1;
    if (!a.periodic) {
        t.a[0] = lo / a.tile;
        t.b[0] = (hi - 1) / a.tile;
        return t;
    }
    const int last = a.ntx - 1;
    if (hi - lo >= a.nc) {
        t.a[0] = 0;
        t.b[0] = last;
        return t;
    }
    const int l = lo - floor_div(lo, a.nc) * a.nc;   // lo modulo nc, in [0, nc)
    const int r = l + (hi - lo) - 1;                   // last cell, unwrapped
    if (r < a.nc) {
        t.a[0] = l / a.tile;
        t.b[0] = r / a.tile;
        return t;
    }
    const int b0 = (r - a.nc) / a.tile;   // wrapped piece [0, b0]
    const int a1 = l / a.tile;            // upper piece [a1, last]
    if (b0 + 1 >= a1) {
        t.a[0] = 0;
        t.b[0] = last;
        return t;
    }
    t.a[0] = 0;
    t.b[0] = b0;
    t.a[1] = a1;
    t.b[1] = last;
    t.n = 2;
    return t;
}

__device__ __forceinline__ int axis_count(const AxisTiles& t)
{
    int c = 0;
    for (int m = 0; m < t.n; ++m) c += t.b[m] - t.a[m] + 1;
    return c;
}

__device__ __forceinline__ int axis_at(const AxisTiles& t, int k)
{
    for (int m = 0; m < t.n; ++m) {
        const int len = t.b[m] - t.a[m] + 1;
        if (k < len) return t.a[m] + k;
        k -= len;
    }
    return 0;
}

// Add 1 to counters[t] for every active lane, one atomic per distinct t in
// the wavefront (Morton-sorted particles share tiles, so per-lane atomics on
// the same counter would serialise).  Each active lane gets its own slot:
// the counter's old value plus its rank among the lanes sharing t.
__device__ __forceinline__ uint32_t wave_aggregated_inc(uint32_t* counters, int t, bool active)
{
    const int lane = threadIdx.x % kWave;
    uint64_t pending = __ballot(active);
    uint32_t mine = 0;
    while (pending) {
        const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
        const int tl = __shfl(t, leader);
        const uint64_t grp = pending & __ballot(t == tl);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[tl], static_cast<uint32_t>(__popcll(grp)));
        base = __shfl(base, leader);
        if ((grp >> lane) & 1ull) mine = base + static_cast<uint32_t>(__popcll(grp & ((1ull << lane) - 1ull)));
        pending &= ~grp;
    }
    return mine;
}

// Count (list == nullptr) or fill the per-tile particle lists.  Every lane of
// a wavefront walks the same number of (particle, tile) steps so the
// aggregated atomics see the whole wavefront.
__device__ __forceinline__ void tile_pass(const DepositArgs& a, uint32_t* counters, const uint32_t* tile_start,
                                          uint32_t* list)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool active = i < a.n;
    PartRange r;
    AxisTiles tx, ty, tz;
    int cx = 0, cy = 0, total = 0;
    if (active) {
        r = part_range(a, i);
        active = r.active;
    }
    if (active) {
        tx = axis_tiles(a, r.lo[0], r.hi[0]);
        ty = axis_tiles(a, r.lo[1], r.hi[1]);
        tz = axis_tiles(a, r.lo[2], r.hi[2]);
        cx = axis_count(tx);
        cy = axis_count(ty);
        total = cx * cy * axis_count(tz);
    }
    int maxk = total;
    for (int o = kWave / 2; o > 0; o >>= 1) maxk = max(maxk, __shfl_xor(maxk, o));
    for (int k = 0; k < maxk; ++k) {
        const bool act = k < total;
        int t = 0;
        if (act) {
            const int kx = k % cx, ky = (k / cx) % cy, kz = k / (cx * cy);
            t = (axis_at(tz, kz) * a.ntx + axis_at(ty, ky)) * a.ntx + axis_at(tx, kx);
        }
        const uint32_t slot = wave_aggregated_inc(counters, t, act);
        if (list && act && tile_start[t] + slot < a.list_cap) list[tile_start[t] + slot] = static_cast<uint32_t>(i);
    }
}

__global__ void deposit_count_kernel(DepositArgs a, uint32_t* __restrict__ tile_cnt)
{
    tile_pass(a, tile_cnt, nullptr, nullptr);
}

// Node-mesh lists, aggregated per workgroup.  The 256 particles of a
// workgroup are consecutive in Morton order, so the tiles they touch lie in a
// small box of tiles: each (particle, tile) pair is counted in LDS, and one
// global atomic per touched tile adds the workgroup's count (count pass) or
// reserves the workgroup's slots of that tile's list (fill pass, then each
// pair takes its slot within them).  A workgroup whose tile box
// exceeds kListTable falls back to one global atomic per pair.  Both passes
// see the same pairs, so the fill uses exactly the counted slots.
constexpr int kListThreads = 256;
constexpr int kListTable = 1024;
// A wavefront whose particles touch at most this many tiles counts and fills
// them by ballots (one LDS atomic per tile), instead of one LDS atomic per
// (particle, tile) pair: Morton-consecutive particles share tiles, and
// same-word LDS atomics serialise.
constexpr int kListWaveTiles = 16;

// A node-mesh particle's (x -+ h) / boxsize per axis: the part of its cell
// range (axis_range, coarse_mesh.cpp:72-75) that does not depend on nc, so
// computed once for every mesh of an nc list.
struct ListPart {
    float qlo[3], qhi[3];
    bool valid;
};

__device__ __forceinline__ ListPart list_part(const DepositArgs& a, int64_t i)
{
    ListPart q;
    q.valid = i < a.n;
    const float4 p = a.spos[q.valid ? i : 0];
    const float h = a.rk[q.valid ? i : 0];
    const float x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q.qlo[d] = (x[d] - h) / a.boxsize;
        q.qhi[d] = (x[d] + h) / a.boxsize;
    }
    return q;
}

// n / d for 0 <= n < 2^22 and d >= 1, from the reciprocal rounded to float
// and one correction step (the product is within one of the quotient).
__device__ __forceinline__ int small_div(int n, int d, float inv_d)
{
    int q = static_cast<int>(static_cast<float>(n) * inv_d);
    if (q * d > n) --q;
    else if ((q + 1) * d <= n) ++q;
    return q;
}

// One mesh's pass of a workgroup; cnt, base and bb are the workgroup's LDS.
// Every barrier inside is reached by all threads or by none (the early
// returns are uniform, or taken where no barrier follows).
template <bool kFill>
__device__ __forceinline__ void node_list_body(const DepositArgs& a, const ListPart& P,
                                               uint32_t* __restrict__ counters,
                                               const uint32_t* __restrict__ tile_start, uint32_t* __restrict__ list,
                                               uint32_t* cnt, uint32_t* base, int* bb)
{
    const int tid = threadIdx.x;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kListThreads + tid;
    int t0[3] = {0, 0, 0}, t1[3] = {-1, -1, -1};
    if (P.valid) {
        // The cell range of axis_range's node-mesh branch, from the shared
        // quotients: the same operations in the same order.
        const float ncf = static_cast<float>(a.nc);
        const float inv_tile = 1.0f / static_cast<float>(a.tile);
        bool act = true;
        int lo[3], hi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int l = static_cast<int>(P.qlo[d] * ncf);
            const int r = 1 + static_cast<int>(P.qhi[d] * ncf);
            lo[d] = l > 0 ? l : 0;
            hi[d] = r < a.nc ? r : a.nc;
            if (lo[d] >= hi[d]) act = false;
        }
        if (act) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                t0[d] = small_div(lo[d], a.tile, inv_tile);
                t1[d] = small_div(hi[d] - 1, a.tile, inv_tile);
            }
        }
    }
    const bool active = t1[0] >= 0;
    if (tid < 6) bb[tid] = (tid < 3) ? 0x7fffffff : -1;
    __syncthreads();
    int v[6];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v[d] = active ? t0[d] : 0x7fffffff;
        v[3 + d] = active ? t1[d] : -1;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            v[d] = min(v[d], __shfl_xor(v[d], o));
            v[3 + d] = max(v[3 + d], __shfl_xor(v[3 + d], o));
        }
    }
    if (tid % kWave == 0 && v[3] >= 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            atomicMin(&bb[d], v[d]);
            atomicMax(&bb[3 + d], v[3 + d]);
        }
    }
    // The wavefront's own box of tiles (v is reduced over the wavefront, so
    // uniform: held in scalar registers).
    int wb[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) wb[d] = __builtin_amdgcn_readfirstlane(v[d]);
    const int wvol = wb[3] >= 0 ? (wb[3] - wb[0] + 1) * (wb[4] - wb[1] + 1) * (wb[5] - wb[2] + 1) : 0;
    const bool per_wave = wvol <= kListWaveTiles;
    const int lane = tid % kWave;
    __syncthreads();
    if (bb[3] < 0) return;   // no active particle in the workgroup (uniform)
    const int bx = bb[3] - bb[0] + 1, by = bb[4] - bb[1] + 1, bz = bb[5] - bb[2] + 1;
    const int64_t vol = static_cast<int64_t>(bx) * by * bz;
    const int ntx = a.ntx;

    if (vol > kListTable) {
        // Spread-out workgroup: one global atomic per pair.
        if (!active) return;
        for (int tz = t0[2]; tz <= t1[2]; ++tz)
            for (int ty = t0[1]; ty <= t1[1]; ++ty)
                for (int tx = t0[0]; tx <= t1[0]; ++tx) {
                    const int64_t t = (static_cast<int64_t>(tz) * ntx + ty) * ntx + tx;
                    if (kFill) {
                        const uint32_t pos = tile_start[t] + atomicAdd(&counters[t], 1u);
                        if (pos < a.list_cap) list[pos] = static_cast<uint32_t>(i);
                    } else {
                        atomicAdd(&counters[t], 1u);
                    }
                }
        return;
    }
    for (int k = tid; k < vol; k += kListThreads) cnt[k] = 0u;
    __syncthreads();
    if (per_wave) {
        // The lanes covering each tile of the wavefront's box, counted by
        // one ballot and added by one lane.
        for (int tz = wb[2]; tz <= wb[5]; ++tz)
            for (int ty = wb[1]; ty <= wb[4]; ++ty)
                for (int tx = wb[0]; tx <= wb[3]; ++tx) {
                    const bool cov = active && tx >= t0[0] && tx <= t1[0] && ty >= t0[1] && ty <= t1[1] &&
                                     tz >= t0[2] && tz <= t1[2];
                    const uint64_t m = __ballot(cov);
                    if (m != 0ull && lane == 0)
                        atomicAdd(&cnt[((tz - bb[2]) * by + (ty - bb[1])) * bx + (tx - bb[0])],
                                  static_cast<uint32_t>(__popcll(m)));
                }
    } else if (active) {
        for (int tz = t0[2]; tz <= t1[2]; ++tz)
            for (int ty = t0[1]; ty <= t1[1]; ++ty)
                for (int tx = t0[0]; tx <= t1[0]; ++tx)
                    atomicAdd(&cnt[((tz - bb[2]) * by + (ty - bb[1])) * bx + (tx - bb[0])], 1u);
    }
    __syncthreads();
    for (int k = tid; k < vol; k += kListThreads) {
        const uint32_t c = cnt[k];
        if (c == 0u) continue;
        const int tx = bb[0] + k % bx, ty = bb[1] + (k / bx) % by, tz = bb[2] + k / (bx * by);
        const int64_t t = (static_cast<int64_t>(tz) * ntx + ty) * ntx + tx;
        if (kFill) {
            base[k] = tile_start[t] + atomicAdd(&counters[t], c);
            cnt[k] = 0u;
        } else {
            atomicAdd(&counters[t], c);
        }
    }
    if (!kFill) return;
    __syncthreads();
    if (per_wave) {
        // One LDS atomic per tile reserves the wavefront's run of slots; the
        // covering lanes take them in lane (Morton) order.
        for (int tz = wb[2]; tz <= wb[5]; ++tz)
            for (int ty = wb[1]; ty <= wb[4]; ++ty)
                for (int tx = wb[0]; tx <= wb[3]; ++tx) {
                    const bool cov = active && tx >= t0[0] && tx <= t1[0] && ty >= t0[1] && ty <= t1[1] &&
                                     tz >= t0[2] && tz <= t1[2];
                    const uint64_t m = __ballot(cov);
                    if (m == 0ull) continue;
                    const int k = ((tz - bb[2]) * by + (ty - bb[1])) * bx + (tx - bb[0]);
                    uint32_t run = 0u;
                    if (lane == 0) run = atomicAdd(&cnt[k], static_cast<uint32_t>(__popcll(m)));
                    run = __shfl(run, 0);
                    if (cov) {
                        const uint32_t pos =
                            base[k] + run + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
                        if (pos < a.list_cap) list[pos] = static_cast<uint32_t>(i);
                    }
                }
    } else if (active) {
        for (int tz = t0[2]; tz <= t1[2]; ++tz)
            for (int ty = t0[1]; ty <= t1[1]; ++ty)
                for (int tx = t0[0]; tx <= t1[0]; ++tx) {
                    const int k = ((tz - bb[2]) * by + (ty - bb[1])) * bx + (tx - bb[0]);
                    const uint32_t pos = base[k] + atomicAdd(&cnt[k], 1u);
                    if (pos < a.list_cap) list[pos] = static_cast<uint32_t>(i);
                }
    }
}

template <bool kFill>
__global__ __launch_bounds__(kListThreads) void node_list_kernel(DepositArgs a, uint32_t* __restrict__ counters,
                                                                 const uint32_t* __restrict__ tile_start,
                                                                 uint32_t* __restrict__ list)
{
    __shared__ uint32_t cnt[kListTable];
    __shared__ uint32_t base[kFill ? kListTable : 1];
    __shared__ int bb[6];
    const ListPart P = list_part(a, static_cast<int64_t>(blockIdx.x) * kListThreads + threadIdx.x);
    node_list_body<kFill>(a, P, counters, tile_start, list, cnt, base, bb);
}

// Every mesh of an nc list in one pass over the particles: the workgroup's
// particles (already in registers and cache after the first mesh) are listed
// in each mesh's tiles in turn.  Tile t of mesh j is t + J.tile_base[j] in
// the concatenated counters, starts and list.
template <bool kFill>
__global__ __launch_bounds__(kListThreads) void multi_list_kernel(MultiDeposit J, uint32_t* __restrict__ counters,
                                                                  const uint32_t* __restrict__ tile_start,
                                                                  uint32_t* __restrict__ list)
{
    __shared__ uint32_t cnt[kListTable];
    __shared__ uint32_t base[kFill ? kListTable : 1];
    __shared__ int bb[6];
    const ListPart P = list_part(J.a[0], static_cast<int64_t>(blockIdx.x) * kListThreads + threadIdx.x);
    for (int j = 0; j < J.n; ++j) {
        __syncthreads();   // the previous mesh's readers of bb, cnt and base are done
        const uint32_t off = J.tile_base[j];
        node_list_body<kFill>(J.a[j], P, counters + off, kFill ? tile_start + off : nullptr, list, cnt, base, bb);
    }
}

__global__ void deposit_fill_kernel(DepositArgs a, const uint32_t* __restrict__ tile_start,
                                    uint32_t* __restrict__ tile_cur, uint32_t* __restrict__ list)
{
    tile_pass(a, tile_cur, tile_start, list);
}

// Write one tile's accumulators to the mesh (FP32), every cell once.
template <typename Acc, int kCells = kTileCells>
__device__ __forceinline__ void flush_tile(const DepositArgs& a, const Acc (*acc)[kCells], int cx0, int cy0,
                                           int cz0, float* __restrict__ mesh, int accumulate)
{
    static_assert(kCells == kTileCells || kCells == kPadCells, "tile layout");
    const int tid = threadIdx.x;
    const int T = a.tile;
    const int ncell = T * T * T;
    const int64_t nc = a.nc;
    const int64_t nmesh = nc * nc * nc;
    for (int c = tid; c < ncell; c += static_cast<int>(blockDim.x)) {
        const int lx = c % T, ly = (c / T) % T, lz = c / (T * T);
        const int gx = cx0 + lx, gy = cy0 + ly, gz = cz0 + lz;
        if (gx >= a.nc || gy >= a.nc || gz >= a.nc) continue;
        const int64_t idx = (static_cast<int64_t>(gz) * nc + gy) * nc + gx;
        const int ac = kCells == kPadCells ? lz * kPadZ + ly * kPadY + lx : c;
        float vals[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) vals[f] = static_cast<float>(acc[f][ac]);
        if (accumulate) {
            mesh[idx] += vals[0];
            mesh[nmesh + idx] += vals[1];
            mesh[2 * nmesh + idx] += vals[2];
            mesh[3 * nmesh + 3 * idx + 0] += vals[3];
            mesh[3 * nmesh + 3 * idx + 1] += vals[4];
            mesh[3 * nmesh + 3 * idx + 2] += vals[5];
        } else {
            mesh[idx] = vals[0];
            mesh[nmesh + idx] = vals[1];
            mesh[2 * nmesh + idx] = vals[2];
            mesh[3 * nmesh + 3 * idx + 0] = vals[3];
            mesh[3 * nmesh + 3 * idx + 1] = vals[4];
            mesh[3 * nmesh + 3 * idx + 2] = vals[5];
        }
    }
}

// One workgroup per tile: FP64 LDS accumulation, then one flush of the tile.
__global__ __launch_bounds__(kDepThreads) void deposit_tile_kernel(
    DepositArgs a, const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_cnt,
    const uint32_t* __restrict__ list, float* __restrict__ mesh, int accumulate)
{
    __shared__ double acc[6][kTileCells];
    __shared__ float wxs[kDepThreads][kTileMax + 1];
    const int tid = threadIdx.x;
    const int T = a.tile;
    const int64_t tile = blockIdx.x;
    const int tx = static_cast<int>(tile % a.ntx);
    const int ty = static_cast<int>((tile / a.ntx) % a.ntx);
    const int tz = static_cast<int>(tile / (static_cast<int64_t>(a.ntx) * a.ntx));
    const int cx0 = tx * T, cy0 = ty * T, cz0 = tz * T;
    for (int c = tid; c < 6 * kTileCells; c += kDepThreads) (&acc[0][0])[c] = 0.0;
    __syncthreads();

    const float dxc = a.boxsize / static_cast<float>(a.nc);
    const uint32_t begin = tile_start[tile];
    const uint32_t count = min(tile_cnt[tile], begin < a.list_cap ? a.list_cap - begin : 0u);
    // Lanes of a wavefront take list entries `stride` apart: the list is in
    // Morton order, so neighbouring entries share cells, and LDS atomics of
    // one instruction to the same address would serialise.
    const uint32_t stride = (count + kWave - 1) / kWave;
    const uint32_t span = stride * kWave;
    for (uint32_t li = tid; li < span; li += kDepThreads) {
        const uint32_t e = (li % kWave) * stride + li / kWave;
        if (e >= count) continue;
        const int64_t i = list[begin + e];
        const PartRange r = part_range(a, i);
        const float hinv = 1.0f / r.h;
        const float ds = dxc * hinv;
        const float4 at = a.sattr[i];
        const float dens = a.dens[i];
        const int igrp = __float_as_int(at.w);
        const bool igm = (a.sign * igrp <= 0);
        // Periodic variant: the unwrapped cell range lies around coord * nc and
        // may run past either end of the node's cells; cell i lands on tile
        // cell i - k*nc - c0 for every shift k that puts it inside the tile
        // (global_mesh.cpp:85-87 takes i modulo nc).  Per axis those are
        // k in [kmin, kmax], from [c0 + k nc, c0 + k nc + T) meeting [lo, hi).
        int kmin[3] = {0, 0, 0}, kmax[3] = {0, 0, 0};
        if (a.periodic) {
            const int c0[3] = {cx0, cy0, cz0};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                kmin[d] = floor_div(r.lo[d] - c0[d] - T, a.nc) + 1;
                kmax[d] = floor_div(r.hi[d] - 1 - c0[d], a.nc);
            }
        }
        for (int kz = kmin[2]; kz <= kmax[2]; ++kz) {
            const int sz = kz * a.nc;
            const int z0 = max(r.lo[2], cz0 + sz), z1 = min(r.hi[2], cz0 + sz + T);
            if (z0 >= z1) continue;
            for (int ky = kmin[1]; ky <= kmax[1]; ++ky) {
                const int sy = ky * a.nc;
                const int y0 = max(r.lo[1], cy0 + sy), y1 = min(r.hi[1], cy0 + sy + T);
                if (y0 >= y1) continue;
                for (int kx = kmin[0]; kx <= kmax[0]; ++kx) {
                    const int sx = kx * a.nc;
                    const int x0 = max(r.lo[0], cx0 + sx), x1 = min(r.hi[0], cx0 + sx + T);
                    if (x0 >= x1) continue;
                    // x weights of this row segment, reused for every (z, y).
                    for (int ix = x0; ix < x1; ++ix) {
                        const float s1 = (static_cast<float>(ix) * dxc - r.x[0]) * hinv;
                        wxs[tid][ix - x0] = cell_integ(a.kern, s1, s1 + ds);
                    }
                    for (int iz = z0; iz < z1; ++iz) {
                        const float s1z = (static_cast<float>(iz) * dxc - r.x[2]) * hinv;
                        const float kzw = cell_integ(a.kern, s1z, s1z + ds);
                        for (int iy = y0; iy < y1; ++iy) {
                            const float s1y = (static_cast<float>(iy) * dxc - r.x[1]) * hinv;
                            const float kyw = cell_integ(a.kern, s1y, s1y + ds);
                            const int rowc = ((iz - sz - cz0) * T + (iy - sy - cy0)) * T - cx0 - sx;
                            for (int ix = x0; ix < x1; ++ix) {
                                const float kxw = wxs[tid][ix - x0];
                                float ker3 = kxw * kyw;
                                ker3 = ker3 * kzw;
                                const int c = rowc + ix;
                                if (igm) {
                                    atomicAdd(&acc[0][c], static_cast<double>(ker3));
                                    atomicAdd(&acc[2][c], static_cast<double>(ker3 * dens));
                                    atomicAdd(&acc[3][c], static_cast<double>(ker3 * at.x));
                                    atomicAdd(&acc[4][c], static_cast<double>(ker3 * at.y));
                                    atomicAdd(&acc[5][c], static_cast<double>(ker3 * at.z));
                                } else {
                                    atomicAdd(&acc[1][c], static_cast<double>(ker3));
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    flush_tile(a, acc, cx0, cy0, cz0, mesh, accumulate);
}

// Node mesh, one workgroup per tile: every (particle, cell) term of the
// tile's list goes into FP64 accumulators in LDS with ds_add_f64, and the
// tile is flushed once with plain stores, every cell exactly once.  The time
// goes to the LDS atomics (five per cell of a mesh particle) and to the
// vector work that forms the terms, so list-adjacent entries are taken in
// groups of G whose terms for one cell are summed in registers first.  The
// tile's list holds the particles of each workgroup of the list pass in
// Morton order, so G adjacent entries are nearly always neighbouring
// particles whose boxes of cells almost coincide.
//
// A wavefront takes 64 / G groups at a time (groups far apart in the list,
// so that their cells rarely coincide; entry g of group q on lane G * q + g).
// Each lane computes its entry's box and writes the entry's per-axis weights
// in the frame of the group's union box, zero outside its own box.  An
// exclusive scan lays the union boxes' (x, y) columns end to end, and the
// wavefront sweeps them 64 at a time: the groups whose columns start inside
// the sweep mark their start in a per-wavefront LDS map (tagged with the
// sweep, so it never needs clearing), and a ballot of the marks gives every
// lane its group.  The lane forms wx * wy of every member for its column once
// and walks the column's cells: each member's weight is the reference's FP32
// product (wx * wy) * wz (coarse_mesh.cpp:93), its field terms are FP64 FMAs
// of that weight and the fields (exact products, :95-122), the members' terms
// are added in FP64 in member order, and each field takes one ds_add_f64 per
// cell that some member reaches.  Lanes of one group cover consecutive x, so one instruction's
// atomics hit neighbouring words.
//
// Eight wavefronts share a tile's accumulators: the tile's padded FP64 sums
// (27 KB) plus 6.2 KB of weights and marks per wavefront let two workgroups,
// 16 wavefronts, share a CU's 160 KB of LDS.
constexpr int kTileWaves = 8;
constexpr int kTileGroup = 4;                // list-adjacent entries summed per atomic
constexpr int kColWStride = 3 * kTileMax;    // wx[8] wy[8] wz[8] per entry: 96 B, 16-B aligned

template <typename Acc, int kWaves, int G>
__device__ __forceinline__ void tile_group_body(const DepositArgs& a, int64_t tile,
                                                const uint32_t* __restrict__ tile_start,
                                                const uint32_t* __restrict__ tile_cnt,
                                                const uint32_t* __restrict__ list, float* __restrict__ mesh,
                                                int accumulate)
{
    static_assert(G == 1 || G == 2 || G == 4, "group size");
    constexpr bool kFmaFields = std::is_same<Acc, double>::value;
    constexpr int kThreads = kWaves * kWave;
    constexpr int kGroups = kWave / G;   // groups per batch
    __shared__ Acc acc[6][kPadCells];
    __shared__ __align__(16) float wts[kWaves][kWave * kColWStride];
    __shared__ int smap[kWaves][kWave];
    __shared__ uint32_t next_batch;   // batches are handed out in turn to the wavefronts that ask
    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int wid = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int g = lane % G;
    const int qi = lane / G;
    float* const W = wts[wid];
    int* const map = smap[wid];
    const int T = a.tile;
    const int tx = static_cast<int>(tile % a.ntx);
    const int ty = static_cast<int>((tile / a.ntx) % a.ntx);
    const int tz = static_cast<int>(tile / (static_cast<int64_t>(a.ntx) * a.ntx));
    const int cx0 = tx * T, cy0 = ty * T, cz0 = tz * T;
    for (int c = tid; c < 6 * kPadCells; c += kThreads) (&acc[0][0])[c] = Acc(0);
    map[lane] = -1;
    if (tid == 0) next_batch = kWaves;
    __syncthreads();

    const float dxc = a.boxsize / static_cast<float>(a.nc);
    const uint32_t begin = tile_start[tile];
    const uint32_t count = min(tile_cnt[tile], begin < a.list_cap ? a.list_cap - begin : 0u);
    const uint32_t ngrp = (count + G - 1) / G;
    const uint32_t nb = (ngrp + kGroups - 1) / kGroups;   // batches; batch b holds groups q * nb + b
    int tag = 0;                                          // sweep counter, tags the map marks
    // Wavefront w starts with batch w; each later batch goes to whichever
    // wavefront asks first (batches differ in work), one batch ahead, so that
    // the list entry of the wavefront's next batch is read while it works on
    // this one and only the particle loads wait at the top of each batch.
    const auto entry_of = [&](uint32_t bb) { return (static_cast<uint32_t>(qi) * nb + bb) * G + g; };
    const auto claim = [&]() {
        uint32_t v = 0;
        if (lane == 0) v = atomicAdd(&next_batch, 1u);
        return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
    };
    uint32_t next_idx = 0;
    if (wid < nb && entry_of(wid) < count) next_idx = list[begin + entry_of(wid)];
    uint32_t b_next = static_cast<uint32_t>(wid) < nb ? claim() : nb;
    for (uint32_t b = wid; b < nb; b = b_next, b_next = b_next < nb ? claim() : nb) {
        const uint32_t e = entry_of(b);
        const bool valid = e < count;
        const uint32_t cur_idx = next_idx;
        if (b_next < nb && entry_of(b_next) < count) next_idx = list[begin + entry_of(b_next)];
        // The entry's box [lo, hi) of tile cells per axis (empty: lo 8, hi 0).
        int lo[3] = {kTileMax, kTileMax, kTileMax}, hi[3] = {0, 0, 0};
        float x[3] = {0.0f, 0.0f, 0.0f};
        float hinv = 0.0f, ds = 0.0f;
        float4 fe = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // dens, vx, vy, vz
        int igm = 0;
        const int c0[3] = {cx0, cy0, cz0};
        if (valid) {
            const int64_t i = cur_idx;
            const PartRange r = part_range(a, i);
            hinv = 1.0f / r.h;
            ds = dxc * hinv;
            const float4 at = a.sattr[i];
            fe = make_float4(a.dens[i], at.x, at.y, at.z);
            igm = (a.sign * __float_as_int(at.w) <= 0) ? 1 : 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                x[d] = r.x[d];
                const int l = max(r.lo[d], c0[d]) - c0[d];
                const int h = min(r.hi[d], c0[d] + T) - c0[d];
                if (h > l) {
                    lo[d] = l;
                    hi[d] = h;
                }
            }
        }
        // The group's union box and its members' mesh/halo bits.
        int ulo[3] = {lo[0], lo[1], lo[2]}, uhi[3] = {hi[0], hi[1], hi[2]};
        int gm = igm << g;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                ulo[d] = min(ulo[d], __shfl_xor(ulo[d], o));
                uhi[d] = max(uhi[d], __shfl_xor(uhi[d], o));
            }
            gm |= __shfl_xor(gm, o);
        }
        int un[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) un[d] = uhi[d] > ulo[d] ? uhi[d] - ulo[d] : 0;
        // Weights in the union frame, zero outside the entry's own box.  A
        // weight or field that is not finite (h = 0 gives infinite 1/h)
        // would turn the zeros into NaN; such batches take the walk that
        // tests every member's own box instead.
        // The row is cleared whole (six 16-byte stores), then the entry's own
        // cells are written over it: the loop runs over the own box only.
        bool bad = valid && !(isfinite(fe.x) && isfinite(fe.y) && isfinite(fe.z) && isfinite(fe.w));
        {
            float4* const row = reinterpret_cast<float4*>(W + lane * kColWStride);
#pragma unroll
            for (int k = 0; k < kColWStride / 4; ++k) row[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
            for (int c = lo[d]; c < hi[d]; ++c) {
                const float s1 = (static_cast<float>(c0[d] + c) * dxc - x[d]) * hinv;
                const float w = cell_integ(a.kern, s1, s1 + ds);
                bad = bad || !isfinite(w);
                W[lane * kColWStride + d * kTileMax + (c - ulo[d])] = w;
            }
        const bool halo_batch = __any(valid && !igm);
        const bool clean_batch = !__any(bad);
        const int obox = lo[0] | (hi[0] << 4) | (lo[1] << 8) | (hi[1] << 12) | (lo[2] << 16) | (hi[2] << 20);
        const int cols = (g == 0 && un[2] > 0) ? un[0] * un[1] : 0;
        const uint32_t gpk = static_cast<uint32_t>(ulo[0] | (ulo[1] << 4) | (ulo[2] << 8) | (un[0] << 12) |
                                                   (un[1] << 16) | (un[2] << 20) | (gm << 24));
        // Exclusive scan of the column counts (only group leaders have any).
        int incl = cols;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const int pre = incl - cols;
        const int total = __shfl(incl, kWave - 1);
        int first = -1;  // group holding the column just before the sweep
        for (int base = 0; base < total; base += kWave, ++tag) {
            if (cols > 0 && pre >= base && pre < base + kWave) map[pre - base] = tag;
            wave_lds_fence_dep();
            const uint64_t sb = __ballot(map[lane] == tag);   // columns where a group starts
            const int pos = base + lane;
            const int p = first + __popcll(sb & ((2ull << lane) - 1ull));
            first = __shfl(p, kWave - 1);
            const int lead = G * p;
            const uint32_t pk = static_cast<uint32_t>(__shfl(static_cast<int>(gpk), lead));
            const int ppre = __shfl(pre, lead);
            float4 f[G];
#pragma unroll
            for (int m = 0; m < G; ++m)
                f[m] = make_float4(__shfl(fe.x, lead + m), __shfl(fe.y, lead + m), __shfl(fe.z, lead + m),
                                   __shfl(fe.w, lead + m));
            // The fields in FP64 once per sweep (the FMA walk's operands).
            double fd[kFmaFields ? G : 1][4];
            if constexpr (kFmaFields) {
#pragma unroll
                for (int m = 0; m < G; ++m) {
                    fd[m][0] = f[m].x;
                    fd[m][1] = f[m].y;
                    fd[m][2] = f[m].z;
                    fd[m][3] = f[m].w;
                }
            }
            int nz = 0, cb = 0, ux = 0, uy = 0, uz = 0;
            float wxy[G];
            float wz[G][kTileMax];
#pragma unroll
            for (int m = 0; m < G; ++m) {
                wxy[m] = 0.0f;
#pragma unroll
                for (int k = 0; k < kTileMax; ++k) wz[m][k] = 0.0f;
            }
            const int gbits = (pk >> 24) & 15;
            if (pos < total) {
                const int bx = pk & 15, by = (pk >> 4) & 15, bz = (pk >> 8) & 15;
                const int nx = (pk >> 12) & 15;
                nz = (pk >> 20) & 15;
                const int off = pos - ppre;
                const int iy = static_cast<int>((static_cast<float>(off) + 0.5f) *
                                                __builtin_amdgcn_rcpf(static_cast<float>(nx)));
                const int ix = off - iy * nx;
#pragma unroll
                for (int m = 0; m < G; ++m) {
                    const float* const wp = W + (lead + m) * kColWStride;
                    wxy[m] = wp[ix] * wp[kTileMax + iy];
                    const float4 z0 = *reinterpret_cast<const float4*>(wp + 2 * kTileMax);
                    const float4 z1 = *reinterpret_cast<const float4*>(wp + 2 * kTileMax + 4);
                    wz[m][0] = z0.x; wz[m][1] = z0.y; wz[m][2] = z0.z; wz[m][3] = z0.w;
                    wz[m][4] = z1.x; wz[m][5] = z1.y; wz[m][6] = z1.z; wz[m][7] = z1.w;
                }
                ux = bx + ix;
                uy = by + iy;
                uz = bz;
                cb = bz * kPadZ + uy * kPadY + ux;
            }
            // The column's cells.  The members' terms are summed in FP64 in
            // member order, and each field takes one atomic per cell that
            // some member reaches.
            const auto walk = [&](auto halo_tag, auto clean_tag) {
                constexpr bool kHalo = decltype(halo_tag)::value;
                constexpr bool kClean = decltype(clean_tag)::value;
                int mbox[G];
#pragma unroll
                for (int m = 0; m < G; ++m) mbox[m] = kClean ? 0 : __shfl(obox, lead + m);
                bool inxy[G];
#pragma unroll
                for (int m = 0; m < G; ++m)
                    inxy[m] = kClean || (ux >= (mbox[m] & 15) && ux < ((mbox[m] >> 4) & 15) &&
                                         uy >= ((mbox[m] >> 8) & 15) && uy < ((mbox[m] >> 12) & 15));
#pragma unroll
                for (int iz = 0; iz < kTileMax; ++iz) {
                    const bool on = iz < nz;
                    if (!__any(on)) break;
                    if (!on) continue;
                    Acc sn = Acc(0), sh = Acc(0), sc = Acc(0), sx = Acc(0), sy = Acc(0), sz = Acc(0);
                    bool any_n = false, any_h = false;
#pragma unroll
                    for (int m = 0; m < G; ++m) {
                        bool in = true;
                        if (!kClean)
                            in = inxy[m] && uz + iz >= ((mbox[m] >> 16) & 15) && uz + iz < ((mbox[m] >> 20) & 15);
                        // coarse_mesh.cpp:93: ker3 = (wx * wy) * wz
                        const float ker3 = wxy[m] * wz[m][iz];
                        const bool mi = !kHalo || ((gbits >> m) & 1);
                        const bool use_n = in && mi;
                        any_n = any_n || (use_n && ker3 != 0.0f);
                        if constexpr (kFmaFields) {
                            // FP64 sums: the weight as the reference forms
                            // it, its products with the fields exact (FMA)
                            // rather than rounded to FP32.  A clean batch has
                            // finite fields, so a zero weight adds zero; the
                            // other walk selects the products instead.
                            const double w = use_n ? static_cast<double>(ker3) : 0.0;
                            const double fx = fd[m][0], fy = fd[m][1], fz = fd[m][2], fw = fd[m][3];
                            if (m == 0) {
                                sn = w;
                                sc = kClean ? w * fx : (use_n ? w * fx : 0.0);
                                sx = kClean ? w * fy : (use_n ? w * fy : 0.0);
                                sy = kClean ? w * fz : (use_n ? w * fz : 0.0);
                                sz = kClean ? w * fw : (use_n ? w * fw : 0.0);
                            } else if (kClean) {
                                sn += w;
                                sc = fma(w, fx, sc);
                                sx = fma(w, fy, sx);
                                sy = fma(w, fz, sy);
                                sz = fma(w, fw, sz);
                            } else {
                                sn += w;
                                sc += use_n ? w * fx : 0.0;
                                sx += use_n ? w * fy : 0.0;
                                sy += use_n ? w * fz : 0.0;
                                sz += use_n ? w * fw : 0.0;
                            }
                        } else {
                            const Acc tn = use_n ? static_cast<Acc>(ker3) : Acc(0);
                            const Acc tc = use_n ? static_cast<Acc>(ker3 * f[m].x) : Acc(0);
                            const Acc tx = use_n ? static_cast<Acc>(ker3 * f[m].y) : Acc(0);
                            const Acc ty = use_n ? static_cast<Acc>(ker3 * f[m].z) : Acc(0);
                            const Acc tz = use_n ? static_cast<Acc>(ker3 * f[m].w) : Acc(0);
                            if (m == 0) {
                                sn = tn; sc = tc; sx = tx; sy = ty; sz = tz;
                            } else {
                                sn += tn; sc += tc; sx += tx; sy += ty; sz += tz;
                            }
                        }
                        if (kHalo) {
                            const bool use_h = in && !mi;
                            sh += use_h ? static_cast<Acc>(ker3) : Acc(0);
                            any_h = any_h || (use_h && ker3 != 0.0f);
                        }
                    }
                    const int c = cb + iz * kPadZ;
                    if (any_n) {
                        atomicAdd(&acc[0][c], sn);
                        atomicAdd(&acc[2][c], sc);
                        atomicAdd(&acc[3][c], sx);
                        atomicAdd(&acc[4][c], sy);
                        atomicAdd(&acc[5][c], sz);
                    }
                    if (kHalo && any_h) atomicAdd(&acc[1][c], sh);
                }
            };
            if (!clean_batch)
                walk(std::true_type{}, std::false_type{});
            else if (halo_batch)
                walk(std::true_type{}, std::true_type{});
            else
                walk(std::false_type{}, std::true_type{});
        }
        wave_lds_fence_dep();
    }
    __syncthreads();
    flush_tile(a, acc, cx0, cy0, cz0, mesh, accumulate);
}

template <typename Acc, int kWaves, int G, bool kAccumulate>
__global__ __launch_bounds__(kWaves * kWave, 4) void deposit_tile_group_kernel(
    DepositArgs a, const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_cnt,
    const uint32_t* __restrict__ list, float* __restrict__ mesh)
{
    tile_group_body<Acc, kWaves, G>(a, blockIdx.x, tile_start, tile_cnt, list, mesh, kAccumulate);
}

// The node-mesh tile kernel with the flush mode as a template argument.
template <typename Acc, int G>
void launch_group_tiles(const DepositArgs& a, unsigned ntiles, const uint32_t* tile_start, const uint32_t* tile_cnt,
                        const uint32_t* list, float* mesh, int accumulate, hipStream_t s)
{
    if (accumulate)
        deposit_tile_group_kernel<Acc, kTileWaves, G, true><<<ntiles, kTileWaves * kWave, 0, s>>>(
            a, tile_start, tile_cnt, list, mesh);
    else
        deposit_tile_group_kernel<Acc, kTileWaves, G, false><<<ntiles, kTileWaves * kWave, 0, s>>>(
            a, tile_start, tile_cnt, list, mesh);
}

// The tiles of every mesh of an nc list in one launch (FP64 sums): workgroup
// w runs tile w - J.tile_base[j] of the mesh j whose range holds w.  The
// meshes come in the caller's order; with coarse meshes first, their few
// long tiles start before the many short ones of the fine meshes.
template <int kWaves, int G>
__global__ __launch_bounds__(kWaves * kWave, 4) void multi_group_kernel(MultiDeposit J,
                                                                     const uint32_t* __restrict__ tile_start,
                                                                     const uint32_t* __restrict__ tile_cnt,
                                                                     const uint32_t* __restrict__ list)
{
    const uint32_t w = blockIdx.x;
    int j = 0;
    while (j + 1 < J.n && w >= J.tile_base[j + 1]) ++j;
    const uint32_t off = J.tile_base[j];
    tile_group_body<double, kWaves, G>(J.a[j], w - off, tile_start + off, tile_cnt + off, list, J.mesh[j], 0);
}

// Precision study: the tile's sums in FP32 with Kahan compensation.  One
// wavefront owns a tile and takes its list entries one after another, each
// entry's box of cells spread over the lanes, so that no two updates of one
// instruction touch the same cell and the compensated update (sum, carry)
// needs no atomics.  The stored cell is sum - carry.
constexpr int kKahanThreads = kWave;

__global__ __launch_bounds__(kKahanThreads) void deposit_tile_kahan_kernel(
    DepositArgs a, const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_cnt,
    const uint32_t* __restrict__ list, float* __restrict__ mesh, int accumulate)
{
    __shared__ float sum[6][kTileCells];
    __shared__ float carry[6][kTileCells];
    __shared__ float wts[3 * kTileMax];
    const int lane = threadIdx.x;
    const int T = a.tile;
    const int64_t tile = blockIdx.x;
    const int tx = static_cast<int>(tile % a.ntx);
    const int ty = static_cast<int>((tile / a.ntx) % a.ntx);
    const int tz = static_cast<int>(tile / (static_cast<int64_t>(a.ntx) * a.ntx));
    const int cx0 = tx * T, cy0 = ty * T, cz0 = tz * T;
    for (int c = lane; c < 6 * kTileCells; c += kKahanThreads) {
        (&sum[0][0])[c] = 0.0f;
        (&carry[0][0])[c] = 0.0f;
    }
    const float dxc = a.boxsize / static_cast<float>(a.nc);
    const uint32_t begin = tile_start[tile];
    const uint32_t count = min(tile_cnt[tile], begin < a.list_cap ? a.list_cap - begin : 0u);
    for (uint32_t e = 0; e < count; ++e) {
        const int64_t i = list[begin + e];
        const PartRange r = part_range(a, i);
        const float hinv = 1.0f / r.h;
        const float ds = dxc * hinv;
        const float4 at = a.sattr[i];
        const float dens = a.dens[i];
        const bool igm = (a.sign * __float_as_int(at.w) <= 0);
        const int c0[3] = {cx0, cy0, cz0};
        int b0[3], n3[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int lo = max(r.lo[d], c0[d]);
            const int hi = min(r.hi[d], c0[d] + T);
            b0[d] = lo - c0[d];
            n3[d] = hi > lo ? hi - lo : 0;
        }
        __syncthreads();   // previous entry's weights are no longer read
        if (lane < 3 * kTileMax) {
            const int d = lane / kTileMax, t = lane % kTileMax;
            if (t < n3[d]) {
                const float s1 = (static_cast<float>(c0[d] + b0[d] + t) * dxc - r.x[d]) * hinv;
                wts[lane] = cell_integ(a.kern, s1, s1 + ds);
            }
        }
        __syncthreads();
        const int nxy = n3[0] * n3[1];
        const int size = nxy * n3[2];
        for (int off = lane; off < size; off += kKahanThreads) {
            const int iz = off / nxy, rem = off % nxy, iy = rem / n3[0], ix = rem % n3[0];
            float ker3 = wts[ix] * wts[kTileMax + iy];   // coarse_mesh.cpp:93
            ker3 = ker3 * wts[2 * kTileMax + iz];
            const int c = ((b0[2] + iz) * T + (b0[1] + iy)) * T + b0[0] + ix;
            float terms[6];
            int fields[6];
            int nf = 0;
            if (igm) {
                terms[0] = ker3; fields[0] = 0;
                terms[1] = ker3 * dens; fields[1] = 2;
                terms[2] = ker3 * at.x; fields[2] = 3;
                terms[3] = ker3 * at.y; fields[3] = 4;
                terms[4] = ker3 * at.z; fields[4] = 5;
                nf = 5;
            } else {
                terms[0] = ker3; fields[0] = 1;
                nf = 1;
            }
            for (int f = 0; f < nf; ++f) {
                const int fi = fields[f];
                const float y = terms[f] - carry[fi][c];
                const float t = sum[fi][c] + y;
                carry[fi][c] = (t - sum[fi][c]) - y;
                sum[fi][c] = t;
            }
        }
    }
    __syncthreads();
    for (int c = lane; c < 6 * kTileCells; c += kKahanThreads) (&sum[0][0])[c] -= (&carry[0][0])[c];
    __syncthreads();
    flush_tile(a, sum, cx0, cy0, cz0, mesh, accumulate);
}

// Sets *flag when the list total counted on the device exceeds the list's
// capacity (a recorded graph replayed on denser inputs than it was sized for).
__global__ void flag_overflow_kernel(const uint64_t* __restrict__ total, uint64_t cap, uint32_t* __restrict__ flag)
{
    if (threadIdx.x == 0 && *total > cap) *flag = 1u;
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

// Stage 1: per-workgroup FP64 partial sums of n, halo, n^2, c over a
// grid-stride range; stage 2 adds the partials in a fixed order.
__global__ __launch_bounds__(kRedThreads) void mesh_reduce_stage1(const float* __restrict__ mesh,
                                                                  int64_t nmesh,
                                                                  double* __restrict__ partial)
{
    __shared__ double sh[4][kRedThreads];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * kRedThreads + threadIdx.x; j < nmesh;
         j += static_cast<int64_t>(gridDim.x) * kRedThreads) {
        const double nv = mesh[j];
        s0 += nv;
        s1 += static_cast<double>(mesh[nmesh + j]);
        s2 += nv * nv;
        s3 += static_cast<double>(mesh[2 * nmesh + j]);
    }
    sh[0][threadIdx.x] = s0;
    sh[1][threadIdx.x] = s1;
    sh[2][threadIdx.x] = s2;
    sh[3][threadIdx.x] = s3;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int f = 0; f < 4; ++f) sh[f][threadIdx.x] += sh[f][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[threadIdx.x * kRedBlocks + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kRedThreads) void mesh_reduce_stage2(const double* __restrict__ partial,
                                                                  double* __restrict__ out)
{
    __shared__ double sh[kRedThreads];
    for (int f = 0; f < 4; ++f) {
        double s = 0;
        for (int b = threadIdx.x; b < kRedBlocks; b += kRedThreads) s += partial[f * kRedBlocks + b];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = kRedThreads / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[f] = sh[0];
        __syncthreads();
    }
}

// coarse_mesh.cpp:158-171 in FP32, operation for operation.
__global__ void normalize_kernel(float* __restrict__ mesh, int64_t nmesh, float vol)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nmesh) return;
    float* n = mesh;
    float* halo = mesh + nmesh;
    float* c = halo + nmesh;
    float* v = c + nmesh;
    const float nj = n[j];
    if (nj > 0.0f) {
        c[j] = c[j] * (vol / (nj * nj));
        v[3 * j] = v[3 * j] / nj;
        v[3 * j + 1] = v[3 * j + 1] / nj;
        v[3 * j + 2] = v[3 * j + 2] / nj;
    } else {
        c[j] = 0.0f;
        v[3 * j] = 0.0f;
        v[3 * j + 1] = 0.0f;
        v[3 * j + 2] = 0.0f;
    }
    n[j] = nj * 8.0f;
    halo[j] = halo[j] * 8.0f;
}

// One node's normalised mesh placed into a global array of side G = nc * nn
// at cell offset (ox, oy, oz), x fastest, as the stitching tools lay it out
// (tools/reformat.cpp:81-107).  Fields: 0 n, 1 halo, 2 c, 3 vel (3
// components interleaved, as in the node mesh), 4 ntot = h + n and 5 the
// momentum vel * (n / 8) (tools/reformat2.cpp:118-126).  A row of nc cells
// is contiguous in both arrays, so a workgroup copies whole rows: the row's
// offsets are computed once and consecutive threads move consecutive words.
template <int kField>
__global__ __launch_bounds__(256) void stitch_kernel(const float* __restrict__ mesh, int nc, int64_t G, int ox,
                                                     int oy, int oz, float* __restrict__ dst)
{
    constexpr int kComps = (kField == 3 || kField == 5) ? 3 : 1;
    const int64_t nmesh = static_cast<int64_t>(nc) * nc * nc;
    const int64_t rows = static_cast<int64_t>(nc) * nc;
    const int rowlen = nc * kComps;
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int64_t ly = r % nc;
        const int64_t lz = r / nc;
        const int64_t c0 = r * nc;   // node cell (0, ly, lz)
        float* const out = dst + (((oz + lz) * G + (oy + ly)) * G + ox) * kComps;
        for (int e = threadIdx.x; e < rowlen; e += blockDim.x) {
            float v;
            if (kField == 0)
                v = mesh[c0 + e];
            else if (kField == 1)
                v = mesh[nmesh + c0 + e];
            else if (kField == 2)
                v = mesh[2 * nmesh + c0 + e];
            else if (kField == 3)
                v = mesh[3 * nmesh + 3 * c0 + e];
            else if (kField == 4)
                v = mesh[nmesh + c0 + e] + mesh[c0 + e];
            else
                v = mesh[3 * nmesh + 3 * c0 + e] * (mesh[c0 + e / 3] / 8.0f);
            out[e] = v;
        }
    }
}

}  // namespace

hipError_t launch_deposit_count(const DepositArgs& a, uint32_t* tile_cnt, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (!a.periodic) {
        node_list_kernel<false><<<static_cast<unsigned>((a.n + kListThreads - 1) / kListThreads), kListThreads, 0,
                                  s>>>(a, tile_cnt, nullptr, nullptr);
        return hipGetLastError();
    }
    const int th = 256;
    deposit_count_kernel<<<static_cast<unsigned>((a.n + th - 1) / th), th, 0, s>>>(a, tile_cnt);
    return hipGetLastError();
}

hipError_t launch_deposit_fill(const DepositArgs& a, const uint32_t* tile_start, uint32_t* tile_cur,
                               uint32_t* list, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (!a.periodic) {
        node_list_kernel<true><<<static_cast<unsigned>((a.n + kListThreads - 1) / kListThreads), kListThreads, 0,
                                 s>>>(a, tile_cur, tile_start, list);
        return hipGetLastError();
    }
    const int th = 256;
    deposit_fill_kernel<<<static_cast<unsigned>((a.n + th - 1) / th), th, 0, s>>>(a, tile_start, tile_cur, list);
    return hipGetLastError();
}

hipError_t launch_multi_count(const MultiDeposit& J, uint32_t* tile_cnt, hipStream_t s)
{
    if (J.n < 1 || J.n > kMaxMeshes || J.a[0].n <= 0) return J.n < 1 || J.n > kMaxMeshes ? hipErrorInvalidValue
                                                                                          : hipSuccess;
    multi_list_kernel<false><<<static_cast<unsigned>((J.a[0].n + kListThreads - 1) / kListThreads), kListThreads, 0,
                               s>>>(J, tile_cnt, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_multi_fill(const MultiDeposit& J, const uint32_t* tile_start, uint32_t* tile_cur, uint32_t* list,
                             hipStream_t s)
{
    if (J.n < 1 || J.n > kMaxMeshes || J.a[0].n <= 0) return J.n < 1 || J.n > kMaxMeshes ? hipErrorInvalidValue
                                                                                          : hipSuccess;
    multi_list_kernel<true><<<static_cast<unsigned>((J.a[0].n + kListThreads - 1) / kListThreads), kListThreads, 0,
                              s>>>(J, tile_cur, tile_start, list);
    return hipGetLastError();
}

hipError_t launch_multi_tiles(const MultiDeposit& J, const uint32_t* tile_start, const uint32_t* tile_cnt,
                              const uint32_t* list, hipStream_t s)
{
    if (J.n < 1 || J.n > kMaxMeshes) return hipErrorInvalidValue;
    for (int j = 0; j < J.n; ++j)
        if (J.a[j].tile < 1 || J.a[j].tile > kTileMax || J.a[j].periodic) return hipErrorInvalidValue;
    const uint32_t ntiles = J.tile_base[J.n];
    if (ntiles == 0) return hipSuccess;
    multi_group_kernel<kTileWaves, kTileGroup><<<ntiles, kTileWaves * kWave, 0, s>>>(J, tile_start, tile_cnt, list);
    return hipGetLastError();
}

hipError_t launch_deposit_tiles(const DepositArgs& a, const uint32_t* tile_start, const uint32_t* tile_cnt,
                                const uint32_t* list, float* mesh, int accumulate, hipStream_t s)
{
    if (a.tile < 1 || a.tile > kTileMax) return hipErrorInvalidValue;
    const int64_t ntiles = static_cast<int64_t>(a.ntx) * a.ntx * a.ntx;
    if (a.periodic)
        deposit_tile_kernel<<<static_cast<unsigned>(ntiles), kDepThreads, 0, s>>>(a, tile_start, tile_cnt, list,
                                                                                  mesh, accumulate);
    else if (a.accum == kAccumKahan32)
        deposit_tile_kahan_kernel<<<static_cast<unsigned>(ntiles), kKahanThreads, 0, s>>>(a, tile_start, tile_cnt,
                                                                                          list, mesh, accumulate);
    else if (a.accum == kAccumAtomic32)   // one atomic per term, as the reference's FP32 sums
        launch_group_tiles<float, 1>(a, static_cast<unsigned>(ntiles), tile_start, tile_cnt, list, mesh, accumulate, s);
    else if (accumulate)
        launch_group_tiles<double, kTileGroup>(a, static_cast<unsigned>(ntiles), tile_start, tile_cnt, list, mesh,
                                               accumulate, s);
    else {
        // The one-mesh case of the nc-list launch (its register allocation
        // spills nothing).
        MultiDeposit J{};
        J.a[0] = a;
        J.mesh[0] = mesh;
        J.tile_base[0] = 0;
        J.tile_base[1] = static_cast<uint32_t>(ntiles);
        J.n = 1;
        multi_group_kernel<kTileWaves, kTileGroup><<<static_cast<unsigned>(ntiles), kTileWaves * kWave, 0, s>>>(
            J, tile_start, tile_cnt, list);
    }
    return hipGetLastError();
}

hipError_t launch_flag_overflow(const uint64_t* total, uint64_t cap, uint32_t* flag, hipStream_t s)
{
    flag_overflow_kernel<<<1, 64, 0, s>>>(total, cap, flag);
    return hipGetLastError();
}

hipError_t launch_fill_zero_u32(uint32_t* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    fill_zero_kernel<uint32_t><<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_fill_zero_f32(float* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    fill_zero_kernel<float><<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_fill_zero_f64(double* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    fill_zero_kernel<double><<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_mesh_reduce(const float* mesh, int nc, double* partial, double* out, hipStream_t s)
{
    const int64_t nmesh = static_cast<int64_t>(nc) * nc * nc;
    mesh_reduce_stage1<<<kRedBlocks, kRedThreads, 0, s>>>(mesh, nmesh, partial);
    ASM_HIP_CHECK(hipGetLastError());
    mesh_reduce_stage2<<<1, kRedThreads, 0, s>>>(partial, out);
    return hipGetLastError();
}

hipError_t launch_normalize(float* mesh, int nc, float vol, hipStream_t s)
{
    const int64_t nmesh = static_cast<int64_t>(nc) * nc * nc;
    normalize_kernel<<<static_cast<unsigned>((nmesh + 255) / 256), 256, 0, s>>>(mesh, nmesh, vol);
    return hipGetLastError();
}

hipError_t launch_stitch(const float* mesh, int nc, int nn, const int coord[3], int field, float* dst, hipStream_t s)
{
    const int64_t rows = static_cast<int64_t>(nc) * nc;
    const unsigned blocks = static_cast<unsigned>(rows < (1 << 20) ? rows : (1 << 20));
    const int64_t G = static_cast<int64_t>(nc) * nn;
    const int ox = coord[0] * nc, oy = coord[1] * nc, oz = coord[2] * nc;
    switch (field) {
    case 0: stitch_kernel<0><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    case 1: stitch_kernel<1><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    case 2: stitch_kernel<2><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    case 3: stitch_kernel<3><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    case 4: stitch_kernel<4><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    default: stitch_kernel<5><<<blocks, 256, 0, s>>>(mesh, nc, G, ox, oy, oz, dst); break;
    }
    return hipGetLastError();
}

}  // namespace asm_hip
