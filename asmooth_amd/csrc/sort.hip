// LSD radix sort of 32-bit keys with 32-bit payloads, and an exclusive scan
// (one workgroup for short arrays, many for long ones).  Used to put
// particles in Morton order (the tree build) and to turn per-tile counts into
// list offsets (mesh deposition).
//
// Each pass handles kRadixBits (4) key bits.  A workgroup owns a tile of 4096
// keys; thread t owns 16 consecutive keys, so the in-tile order is (thread, item) and the
// per-(digit, thread) counts scanned in digit-major order give every key a
// stable destination.  Per-tile digit totals are scanned across tiles in
// digit-major order, which makes each pass stable across the whole array.
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kRadixBits = 4;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kDigitThreads = kSortThreads / kRadix;   // threads per digit row in the count scan
constexpr int kCntStride = kSortThreads + 1;          // padded digit rows

__global__ __launch_bounds__(kSortThreads) void rs_hist(const uint32_t* __restrict__ keys,
                                                        int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int nb)
{
    __shared__ uint32_t h[kSortThreads / kWave][kRadix];
    const int t = threadIdx.x;
    const int w = t / kWave;
    if (t < (kSortThreads / kWave) * kRadix) (&h[0][0])[t] = 0u;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t idx = base + static_cast<int64_t>(i) * kSortThreads + t;
        if (idx < n) atomicAdd(&h[w][(keys[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    if (t < kRadix) {
        uint32_t s = 0;
        for (int ww = 0; ww < kSortThreads / kWave; ++ww) s += h[ww][t];
        hist[static_cast<int64_t>(t) * nb + blockIdx.x] = s;
    }
}

// Exclusive scan of one value per thread across a 256-thread workgroup.
__device__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* wsum /* [4] */, uint32_t* total)
{
    const int lane = threadIdx.x % kWave;
    const int w = threadIdx.x / kWave;
    uint32_t inc = v;
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int ww = 0; ww < kSortThreads / kWave; ++ww) {
        if (ww < w) pre += wsum[ww];
        tot += wsum[ww];
    }
    if (total) *total = tot;
    return pre + inc - v;
}

__global__ __launch_bounds__(kSortThreads) void rs_scatter(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout,
                                                           uint32_t* __restrict__ vout,
                                                           int64_t n, int shift,
                                                           const uint32_t* __restrict__ hist, int nb)
{
    __shared__ uint32_t cnt[kRadix * kCntStride];
    __shared__ uint32_t wsum[kSortThreads / kWave];
    __shared__ uint32_t gbase[kRadix];
    __shared__ int64_t gofs[kRadix];
    __shared__ uint32_t sk[kSortTile];   // the tile in digit order, staged for coalesced stores
    __shared__ uint32_t sv[kSortTile];
    const int t = threadIdx.x;
#pragma unroll
    for (int d = 0; d < kRadix; ++d) cnt[d * kCntStride + t] = 0u;
    if (t < kRadix) gbase[t] = hist[static_cast<int64_t>(t) * nb + blockIdx.x];

    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile + static_cast<int64_t>(t) * kSortItems;
    uint32_t k[kSortItems], v[kSortItems];
    int nvalid = 0;
    if (base + kSortItems <= n) {
        const uint4* kp = reinterpret_cast<const uint4*>(kin + base);
        const uint4* vp = reinterpret_cast<const uint4*>(vin + base);
#pragma unroll
        for (int q = 0; q < kSortItems / 4; ++q) {
            const uint4 a = kp[q];
            const uint4 b = vp[q];
            k[4 * q + 0] = a.x; k[4 * q + 1] = a.y; k[4 * q + 2] = a.z; k[4 * q + 3] = a.w;
            v[4 * q + 0] = b.x; v[4 * q + 1] = b.y; v[4 * q + 2] = b.z; v[4 * q + 3] = b.w;
        }
        nvalid = kSortItems;
    } else {
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const int64_t idx = base + j;
            const bool ok = idx < n;
            k[j] = ok ? kin[idx] : 0u;
            v[j] = ok ? vin[idx] : 0u;
        }
        const int64_t rem = n - base;
        nvalid = rem <= 0 ? 0 : static_cast<int>(rem < kSortItems ? rem : kSortItems);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j)
        if (j < nvalid) cnt[((k[j] >> shift) & (kRadix - 1)) * kCntStride + t] += 1u;
    __syncthreads();

    // Digit-major exclusive scan of the kRadix x 256 counts: thread t scans
    // the kRadix consecutive entries e = kRadix t .. kRadix t + kRadix - 1
    // (digit t / kDigitThreads, threads kRadix (t % kDigitThreads) + i).
    const int d0 = t / kDigitThreads;
    const int t0 = (t % kDigitThreads) * kRadix;
    uint32_t s[kRadix];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kRadix; ++i) {
        s[i] = cnt[d0 * kCntStride + t0 + i];
        sum += s[i];
    }
    uint32_t run = block_exclusive_scan_256(sum, wsum, nullptr);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRadix; ++i) {
        cnt[d0 * kCntStride + t0 + i] = run;
        run += s[i];
    }
    __syncthreads();
    // cnt now holds the in-tile start of each (digit, thread) run.  A key of
    // digit d goes to gbase[d] + (its in-tile rank among digit-d keys), i.e.
    // gofs[d] + cnt[d][t] with gofs[d] = gbase[d] - (in-tile start of digit d),
    // taken before any thread advances its own counter.
    if (t < kRadix) gofs[t] = static_cast<int64_t>(gbase[t]) - static_cast<int64_t>(cnt[t * kCntStride]);
    __syncthreads();
    // Place every key at its in-tile rank in LDS, then store the tile in
    // rank order: consecutive ranks of one digit go to consecutive
    // addresses, so the stores coalesce.
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        if (j >= nvalid) break;
        const uint32_t d = (k[j] >> shift) & (kRadix - 1);
        const uint32_t r = cnt[d * kCntStride + t];
        cnt[d * kCntStride + t] = r + 1u;
        sk[r] = k[j];
        sv[r] = v[j];
    }
    __syncthreads();
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
    const int tile_n = static_cast<int>(n - tile0 < kSortTile ? n - tile0 : kSortTile);
    for (int r = t; r < tile_n; r += kSortThreads) {
        const uint32_t key = sk[r];
        const int64_t pos = gofs[(key >> shift) & (kRadix - 1)] + r;
        kout[pos] = key;
        vout[pos] = sv[r];
    }
}

constexpr int kScanThreads = 1024;

// In-place exclusive scan of m uint32 values by one workgroup; writes the
// grand total to *total (if non-null).
__global__ __launch_bounds__(kScanThreads) void scan_exclusive_1wg(uint32_t* __restrict__ a, int64_t m,
                                                                   uint64_t* __restrict__ total)
{
    __shared__ uint32_t wsum[kScanThreads / kWave];
    __shared__ uint64_t wsum64[kScanThreads / kWave];
    __shared__ uint32_t carry_s;
    __shared__ uint64_t carry64_s;   // the same running total without wrap-around
    const int t = threadIdx.x;
    const int lane = t % kWave;
    const int w = t / kWave;
    if (t == 0) {
        carry_s = 0;
        carry64_s = 0;
    }
    __syncthreads();
    constexpr int kPer = 4;
    for (int64_t chunk = 0; chunk < m; chunk += static_cast<int64_t>(kScanThreads) * kPer) {
        const int64_t b = chunk + static_cast<int64_t>(t) * kPer;
        uint32_t x[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            x[i] = (b + i < m) ? a[b + i] : 0u;
            sum += x[i];
        }
        uint64_t sum64 = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) sum64 += x[i];
        for (int o = kWave / 2; o > 0; o >>= 1) sum64 += __shfl_xor(sum64, o);
        uint32_t inc = sum;
        for (int o = 1; o < kWave; o <<= 1) {
            uint32_t y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == kWave - 1) {
            wsum[w] = inc;
            wsum64[w] = sum64;
        }
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int ww = 0; ww < kScanThreads / kWave; ++ww) {
            if (ww < w) pre += wsum[ww];
            tot += wsum[ww];
        }
        const uint32_t carry = carry_s;
        uint32_t run = carry + pre + inc - sum;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            if (b + i < m) a[b + i] = run;
            run += x[i];
        }
        __syncthreads();
        if (t == 0) {
            carry_s = carry + tot;
            uint64_t tot64 = 0;
            for (int ww = 0; ww < kScanThreads / kWave; ++ww) tot64 += wsum64[ww];
            carry64_s += tot64;
        }
        __syncthreads();
    }
    if (t == 0 && total) *total = carry64_s;
}

// Multi-workgroup form of the same scan for long arrays: chunks of 4096
// values get their 64-bit sums (one workgroup each), one workgroup scans the
// sums, and every chunk is then scanned locally from its offset.  The 32-bit
// results equal the single-workgroup scan's (sums modulo 2^32).
constexpr int kScanPer = kSortItems;
constexpr int kScanChunk = kSortThreads * kScanPer;   // 4096

__global__ __launch_bounds__(kSortThreads) void scan_chunk_sums(const uint32_t* __restrict__ a, int64_t m,
                                                                uint64_t* __restrict__ sums)
{
    __shared__ uint64_t ws[kSortThreads / kWave];
    const int t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk;
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        const int64_t idx = base + static_cast<int64_t>(i) * kSortThreads + t;
        if (idx < m) v += a[idx];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (t % kWave == 0) ws[t / kWave] = v;
    __syncthreads();
    if (t == 0) {
        uint64_t sum = 0;
        for (int w = 0; w < kSortThreads / kWave; ++w) sum += ws[w];
        sums[blockIdx.x] = sum;
    }
}

// Exclusive scan of the chunk sums in place (one workgroup); grand total.
__global__ __launch_bounds__(kScanThreads) void scan_chunk_offsets(uint64_t* __restrict__ sums, int64_t nchunks,
                                                                   uint64_t* __restrict__ total)
{
    __shared__ uint64_t wsum[kScanThreads / kWave];
    __shared__ uint64_t carry_s;
    const int t = threadIdx.x;
    const int lane = t % kWave;
    const int w = t / kWave;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < nchunks; c0 += kScanThreads) {
        const int64_t i = c0 + t;
        const uint64_t x = i < nchunks ? sums[i] : 0ull;
        uint64_t inc = x;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint64_t y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == kWave - 1) wsum[w] = inc;
        __syncthreads();
        uint64_t pre = 0, tot = 0;
        for (int ww = 0; ww < kScanThreads / kWave; ++ww) {
            if (ww < w) pre += wsum[ww];
            tot += wsum[ww];
        }
        const uint64_t carry = carry_s;
        if (i < nchunks) sums[i] = carry + pre + inc - x;
        __syncthreads();
        if (t == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (t == 0 && total) *total = carry_s;
}

// Each chunk scanned from its offset: thread t holds values 16t .. 16t+15.
__global__ __launch_bounds__(kSortThreads) void scan_chunk_apply(uint32_t* __restrict__ a, int64_t m,
                                                                 const uint64_t* __restrict__ offs)
{
    __shared__ uint32_t wsum[kSortThreads / kWave];
    const int t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk + static_cast<int64_t>(t) * kScanPer;
    uint32_t x[kScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        x[i] = (base + i < m) ? a[base + i] : 0u;
        sum += x[i];
    }
    uint32_t run = static_cast<uint32_t>(offs[blockIdx.x]) + block_exclusive_scan_256(sum, wsum, nullptr);
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        if (base + i < m) a[base + i] = run;
        run += x[i];
    }
}

}  // namespace

int64_t scan_scratch_entries(int64_t m)
{
    return (m + kScanChunk - 1) / kScanChunk;
}

hipError_t launch_scan_exclusive(uint32_t* a, int64_t m, uint64_t* total, uint64_t* scratch, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    if (!scratch || m <= 2 * kScanChunk) {
        scan_exclusive_1wg<<<1, kScanThreads, 0, s>>>(a, m, total);
        return hipGetLastError();
    }
    const int64_t nchunks = scan_scratch_entries(m);
    if (nchunks > 0x7fffffffLL) return hipErrorInvalidValue;
    scan_chunk_sums<<<static_cast<unsigned>(nchunks), kSortThreads, 0, s>>>(a, m, scratch);
    ASM_HIP_CHECK(hipGetLastError());
    scan_chunk_offsets<<<1, kScanThreads, 0, s>>>(scratch, nchunks, total);
    ASM_HIP_CHECK(hipGetLastError());
    scan_chunk_apply<<<static_cast<unsigned>(nchunks), kSortThreads, 0, s>>>(a, m, scratch);
    return hipGetLastError();
}

int64_t radix_hist_entries(int64_t n)
{
    const int64_t nb = (n + kSortTile - 1) / kSortTile;
    return nb * kRadix;
}

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                             int64_t n, int bits, uint32_t* hist, uint64_t* scan_scratch, hipStream_t s,
                             bool* result_in_tmp)
{
    *result_in_tmp = false;
    if (n <= 1) return hipSuccess;
    const int nb = static_cast<int>((n + kSortTile - 1) / kSortTile);
    uint32_t* kin = keys;
    uint32_t* vin = vals;
    uint32_t* kout = keys_tmp;
    uint32_t* vout = vals_tmp;
    for (int shift = 0; shift < bits; shift += kRadixBits) {
        rs_hist<<<nb, kSortThreads, 0, s>>>(kin, n, shift, hist, nb);
        ASM_HIP_CHECK(hipGetLastError());
        ASM_HIP_CHECK(launch_scan_exclusive(hist, static_cast<int64_t>(nb) * kRadix, nullptr, scan_scratch, s));
        rs_scatter<<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, hist, nb);
        ASM_HIP_CHECK(hipGetLastError());
        uint32_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    *result_in_tmp = (kin == keys_tmp);
    return hipSuccess;
}

}  // namespace asm_hip
