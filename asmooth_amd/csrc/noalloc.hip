// The smoothing path for callers that cannot allocate device memory.
//
// asm_smooth_noalloc does the work of asm_set_particles_device +
// asm_tree_construct + asm_neighbor_search (+ asm_assign_on_mesh for one nc)
// with the same kernels, but every work buffer is carved from `scratch`, one
// file-scope device array of kScratchBytes that each device holds from the
// library's load until the process ends.  The caller passes device inputs and
// outputs; nothing here calls hipMalloc or hipFree.
//
// Every call on a device shares that device's array: two calls that overlap
// on one device (two host threads, or streams) race on it.  The array keeps
// the last call's contents; buffers that must start at zero (the FP64 density
// sums, the tile counters) are cleared by fill kernels in each call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/asmooth.h"
#include "common.hpp"
#include "kernels.hpp"

using namespace asm_hip;

namespace {

constexpr int64_t kMaxN = ASM_NOALLOC_MAX_PARTICLES;
constexpr int kMaxNc = ASM_NOALLOC_MAX_NC;
constexpr int64_t kMaxList = ASM_NOALLOC_MAX_LIST;
constexpr int kLeaf = 32;   // the context's default leaf size
constexpr int kTile = 8;    // the context's default tile side
constexpr size_t kScratchBytes = static_cast<size_t>(96) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current (each device holds its own copy of the array).
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

int set_err(char* err, int err_len, int code, const char* fmt, ...)
{
    if (err && err_len > 0) {
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(err, static_cast<size_t>(err_len), fmt, ap);
        va_end(ap);
    }
    return code;
}

// Consecutive 256-byte aligned pieces of the scratch array.
struct Carver {
    unsigned char* base;
    size_t used = 0;
    template <class T>
    T* take(int64_t count)
    {
        T* p = reinterpret_cast<T*>(base + used);
        used += (static_cast<size_t>(count > 0 ? count : 1) * sizeof(T) + 255) & ~static_cast<size_t>(255);
        return p;
    }
};

// Sorted-order h and density back to the caller's order.
__global__ void unsort_rk_dens_kernel(const float4* __restrict__ spos, const float* __restrict__ rk,
                                      const float* __restrict__ dens, int64_t n, float* __restrict__ rk_o,
                                      float* __restrict__ dens_o)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = __float_as_uint(spos[i].w);
    rk_o[o] = rk[i];
    dens_o[o] = dens[i];
}

int64_t next_pow2(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

#define NAC(expr)                                                                                              \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess)                                                                                  \
            return set_err(err, err_len, ASM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));         \
    } while (0)

extern "C" int asm_smooth_noalloc(int device, const asm_particle* d_particles, int64_t n, int64_t np_local,
                                  float boxsize, float left, float right, int nc, int sign, float* d_rk,
                                  float* d_dens, float* d_mesh, char* err, int err_len)
{
    if (err && err_len > 0) err[0] = '\0';
    if (device < 0 || device >= kMaxDevices) return set_err(err, err_len, ASM_ERR_ARG, "device %d out of range", device);
    if (!d_particles || !d_rk || !d_dens || (nc > 0 && !d_mesh))
        return set_err(err, err_len, ASM_ERR_ARG, "null argument");
    if (n < kSlots) return set_err(err, err_len, ASM_ERR_ARG, "need at least %d particles, got %lld", kSlots,
                                   static_cast<long long>(n));
    if (n > kMaxN)
        return set_err(err, err_len, ASM_ERR_CAPACITY, "%lld particles exceed the no-allocation maximum of %lld",
                       static_cast<long long>(n), static_cast<long long>(kMaxN));
    if (np_local < 0 || np_local > n) return set_err(err, err_len, ASM_ERR_ARG, "np_local out of range");
    if (!(boxsize > 0.0f) || !(right > left)) return set_err(err, err_len, ASM_ERR_ARG, "empty box or tree cube");
    if (nc < 0 || nc > kMaxNc)
        return set_err(err, err_len, ASM_ERR_CAPACITY, "nc %d outside 0..%d (the no-allocation maximum)", nc, kMaxNc);

    NAC(hipSetDevice(device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[device]) NAC(hipGetSymbolAddress(&g_ptr[device], HIP_SYMBOL(scratch)));
        base = g_ptr[device];
    }

    // Work buffers, as the context sizes them (api.hip) for n particles,
    // leaf size kLeaf and one mesh of side nc at tile kTile.
    const int64_t nleaf = (n + kLeaf - 1) / kLeaf;
    const int64_t nleaf_pad = next_pow2(nleaf);
    const int64_t nwaves = (n + kWave - 1) / kWave;
    const int ntx = nc > 0 ? (nc + kTile - 1) / kTile : 0;
    const int64_t ntiles = static_cast<int64_t>(ntx) * ntx * ntx;
    const int64_t hist_n = radix_hist_entries(n);
    const int64_t scan_n = std::max(scan_scratch_entries(hist_n), scan_scratch_entries(ntiles));
    Carver cv{static_cast<unsigned char*>(base)};
    uint32_t* keys = cv.take<uint32_t>(n);
    uint32_t* vals = cv.take<uint32_t>(n);
    uint32_t* keys_tmp = cv.take<uint32_t>(n);
    uint32_t* vals_tmp = cv.take<uint32_t>(n);
    uint32_t* hist = cv.take<uint32_t>(hist_n);
    uint64_t* scan_tmp = cv.take<uint64_t>(scan_n);
    float4* spos = cv.take<float4>(n);
    float4* sattr = cv.take<float4>(n);
    int32_t* inv_perm = cv.take<int32_t>(n);
    Box* nodes = cv.take<Box>(2 * nleaf_pad);
    Box* groups = cv.take<Box>((n + kGroup - 1) / kGroup);
    float* rk = cv.take<float>(n);
    float* dens = cv.take<float>(n);
    uint64_t* kkey = cv.take<uint64_t>(n);
    double* dens64 = cv.take<double>(n);
    double* dens_pull = cv.take<double>(n);
    float* tau = cv.take<float>(n);
    uint32_t* wave_flags = cv.take<uint32_t>(nwaves);
    uint32_t* wleaf_n = cv.take<uint32_t>(nwaves);
    uint32_t* wleaf = cv.take<uint32_t>(nwaves * kLeafCap);
    uint32_t* tile_cnt = cv.take<uint32_t>(ntiles);
    uint32_t* tile_start = cv.take<uint32_t>(ntiles);
    uint32_t* tile_cur = cv.take<uint32_t>(ntiles);
    uint64_t* scan_total = cv.take<uint64_t>(1);
    uint32_t* list = cv.take<uint32_t>(nc > 0 ? kMaxList : 1);
    if (cv.used > kScratchBytes)
        return set_err(err, err_len, ASM_ERR_CAPACITY, "work buffers of %zu bytes exceed the %zu-byte scratch array",
                       cv.used, kScratchBytes);

    hipStream_t s = nullptr;   // the device's null stream; the call synchronizes it before returning
    // Tree (asm_tree_construct).
    NAC(launch_morton_keys(reinterpret_cast<const ParticleRec*>(d_particles), n, left, right, keys, vals, s));
    bool in_tmp = false;
    NAC(launch_radix_sort(keys, vals, keys_tmp, vals_tmp, n, kMortonBits, hist, scan_tmp, s, &in_tmp));
    const uint32_t* perm = in_tmp ? vals_tmp : vals;
    NAC(launch_gather_sorted(reinterpret_cast<const ParticleRec*>(d_particles), perm, n, spos, sattr, inv_perm, s));
    NAC(launch_build_tree(spos, n, kLeaf, nleaf, nleaf_pad, nodes, groups, s));
    // Search, h and density (asm_neighbor_search, bound + collect walks).
    KnnScratch kscr{wleaf, wleaf_n, wave_flags, tau, dens_pull, nullptr, nullptr, nullptr, groups, 0, kKernTent, kBoundSelect, nullptr, nullptr};
    NAC(launch_fill_zero_f64(dens64, n, s));
    NAC(launch_knn(spos, n, nodes, nleaf_pad, nleaf, kLeaf, rk, kkey, inv_perm, dens64, nullptr, nullptr, kscr,
                   kKnnBoundCollect, s));
    NAC(launch_dens_finalize(dens64, dens_pull, n, dens, s));
    unsort_rk_dens_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(spos, rk, dens, n, d_rk, d_dens);
    NAC(hipGetLastError());
    // One mesh (asm_assign_on_mesh): tile lists, then one workgroup per tile.
    if (nc > 0) {
        DepositArgs a{};
        a.spos = spos;
        a.sattr = sattr;
        a.rk = rk;
        a.dens = dens;
        a.n = n;
        a.boxsize = boxsize;
        a.nc = nc;
        a.sign = sign;
        a.tile = kTile;
        a.ntx = ntx;
        a.periodic = 0;
        a.accum = kAccumFp64;
        a.kern = kKernTent;
        a.list_cap = static_cast<uint32_t>(kMaxList);
        NAC(launch_fill_zero_u32(tile_cnt, ntiles, s));
        NAC(launch_deposit_count(a, tile_cnt, s));
        NAC(hipMemcpyAsync(tile_start, tile_cnt, sizeof(uint32_t) * ntiles, hipMemcpyDeviceToDevice, s));
        NAC(launch_scan_exclusive(tile_start, ntiles, scan_total, scan_tmp, s));
        NAC(launch_fill_zero_u32(tile_cur, ntiles, s));
        NAC(launch_deposit_fill(a, tile_start, tile_cur, list, s));
        NAC(launch_deposit_tiles(a, tile_start, tile_cnt, list, d_mesh, 0, s));
    }
    NAC(hipStreamSynchronize(s));
    if (nc > 0) {
        uint64_t total = 0;
        NAC(hipMemcpy(&total, scan_total, sizeof(total), hipMemcpyDeviceToHost));
        if (total > static_cast<uint64_t>(kMaxList))
            return set_err(err, err_len, ASM_ERR_CAPACITY,
                           "nc %d needs %llu tile-list entries, more than the no-allocation maximum of %lld; the "
                           "mesh is incomplete (h and density are complete)",
                           nc, static_cast<unsigned long long>(total), static_cast<long long>(kMaxList));
    }
    return ASM_OK;
}
