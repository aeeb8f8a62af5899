// C ABI of the device path: context, device buffers, call sequencing, timers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/asmooth.h"
#include "common.hpp"
#include "kernels.hpp"

using namespace asm_hip;

namespace {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t bytes = (n > 0 ? n : 1) * sizeof(T);
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DevBuf() { release(); }
};

enum TimerId {
    T_SORT = 0,
    T_TREE,
    T_KNN,
    T_DENS,
    T_DEPOSIT_LIST,
    T_DEPOSIT_TILES,
    T_REDUCE,
    T_NORMALIZE,
    T_HALO,
    T_SEED,
    T_KNN_BOUND,
    T_KNN_COLLECT,
    T_KNN_FIXUP,
    T_STITCH,
    T_COUNT
};
const char* kTimerNames[T_COUNT] = {"sort",      "tree",      "knn",       "dens",        "deposit_lists",
                                    "deposit_tiles", "mesh_reduce", "normalize", "mark_halo", "seed",
                                    "knn_bound", "knn_collect", "knn_fixup", "stitch"};

struct MeshDev {
    DevBuf<float> mesh;
};

}  // namespace

struct asm_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int keep_neighbors = 0;
    int profile = 0;
    int leaf_size = 32;
    int tile = 8;
    std::map<int, int> tile_for_nc;   // set by asm_autotune
    float tree_left = 0.0f, tree_right = 0.0f;

    // particle state
    int64_t n = 0;
    int64_t np_local = 0;
    float boxsize = 0.0f;
    bool have_particles = false;
    bool have_tree = false;
    bool have_search = false;
    // Set while a batch records its work into a HIP graph: no host reads,
    // buffers already sized by an eager run of the same work.
    bool capturing = false;
    // In a batch: a device word set when a deposition's tile list needs more
    // entries than its buffer holds (a replay on denser inputs than the
    // eager run that sized it); null outside batches.
    uint32_t* overflow_flag = nullptr;
    // Outside batches, a deposition whose tile lists fit the list buffer of
    // an earlier call reads nothing back: it compares its entry total with
    // the buffer's capacity on the device (one flag word per call) and is
    // recorded here.  The next call that reads a mesh or changes what a
    // deposition reads (settle_depositions) reads the flags and redoes, with
    // the list sized from its total, every deposition that did not fit.
    struct PendingDeposit {
        std::vector<int> ncs;   // the meshes of the call
        int sign = 0;
        int periodic = 0;
        int coord[3] = {0, 0, 0};
        bool multi = false;     // asm_assign_on_meshes
    };
    std::vector<PendingDeposit> dep_pending;
    DevBuf<uint32_t> dep_flags;
    DevBuf<ParticleRec> input;   // caller order
    // sort / tree
    DevBuf<uint32_t> keys, vals, keys_tmp, vals_tmp, hist;
    DevBuf<float4> spos, sattr;
    DevBuf<int32_t> inv_perm;
    DevBuf<uint32_t> perm_sorted;  // sorted position -> original index (points into vals or vals_tmp)
    uint32_t* perm = nullptr;
    DevBuf<Box> nodes;
    DevBuf<Box> groups;   // boxes of kGroup consecutive sorted particles
    int64_t nleaf = 0, nleaf_pad = 0;
    // search results (sorted order)
    DevBuf<float> rk, dens, kmax;
    DevBuf<uint64_t> kkey;
    DevBuf<double> dens64;
    int deterministic = 0;
    int knn_method = kKnnBoundCollect;
    int knn_bound = kBoundSelect;   // ASM_OPT_KNN_BOUND
    int deposit_accum = kAccumFp64;
    int kernel = kKernTent;   // ASM_OPT_KERNEL
    DevBuf<uint32_t> knn_flags, knn_leaf_n, knn_leaves;   // per wavefront of 64 sorted particles
    DevBuf<uint32_t> knn_list, knn_list_n;               // selection walk's neighbour lists
    DevBuf<float> knn_tau;                                // per sorted particle
    DevBuf<double> dens_pull;
    // learned seed radius
    bool seed_on = false;
    float seed_margin = 0.0f;
    int seed_walk = 2;   // ASM_OPT_SEED_WALK
    DevBuf<float> seed_params, seed_grid, seed_rcell, tau_init;
    DevBuf<uint32_t> seed_redo;
    DevBuf<unsigned long long> seed_fb;
    bool seed_used = false;   // last search used the model
    DevBuf<int32_t> nbr;
    DevBuf<KnnStats> kstats;   // kStatSlots copies
    // deposition
    DevBuf<uint32_t> tile_cnt, tile_start, tile_cur, list;
    DevBuf<uint64_t> scan_total, scan_tmp;
    DevBuf<double> red_partial, red_out;
    std::map<int, MeshDev> meshes;
    // halos
    DevBuf<HaloRec> halos;
    DevBuf<unsigned long long> halo_count;
    // timers
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    double timer_ms[T_COUNT] = {0};
    int64_t timer_n[T_COUNT] = {0};
};

namespace {

int fail(asm_context* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPC(expr)                                                                                  \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            return fail(ctx, ASM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                        __FILE__, __LINE__);                                                        \
    } while (0)

hipEvent_t get_event(asm_context* c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Scoped timer: records an event pair around a stage when profiling is on.
struct StageTimer {
    asm_context* c;
    int id;
    hipEvent_t a = nullptr;
    StageTimer(asm_context* c_, int id_) : c(c_), id(id_)
    {
        if (c->profile) {
            a = get_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~StageTimer()
    {
        if (c->profile && a) {
            hipEvent_t b = get_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->pending.push_back({id, a, b});
            }
        }
    }
};

// KnnMarks callbacks: one StageTimer-like event pair per kNN kernel.
struct KnnMarkState {
    asm_context* c;
    hipEvent_t a[3] = {nullptr, nullptr, nullptr};
};
void knn_mark_begin(void* user, int st);
void knn_mark_end(void* user, int st);

hipError_t resolve_timers(asm_context* c)
{
    if (c->pending.empty()) return hipSuccess;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    for (auto& p : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->timer_ms[p.id] += ms;
            c->timer_n[p.id] += 1;
        }
        c->free_events.push_back(p.a);
        c->free_events.push_back(p.b);
    }
    c->pending.clear();
    return hipSuccess;
}

void knn_mark_begin(void* user, int st)
{
    auto* m = static_cast<KnnMarkState*>(user);
    if (!m->c->profile) return;
    m->a[st] = get_event(m->c);
    if (m->a[st]) (void)hipEventRecord(m->a[st], m->c->stream);
}

void knn_mark_end(void* user, int st)
{
    auto* m = static_cast<KnnMarkState*>(user);
    if (!m->c->profile || !m->a[st]) return;
    hipEvent_t b = get_event(m->c);
    if (!b) return;
    (void)hipEventRecord(b, m->c->stream);
    m->c->pending.push_back({T_KNN_BOUND + st, m->a[st], b});
}

int64_t next_pow2(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

int load_common(asm_context* ctx, int64_t n, int64_t np_local, float boxsize)
{
    if (n < kSlots) return fail(ctx, ASM_ERR_ARG, "need at least %d particles, got %lld", kSlots, (long long)n);
    if (n > 0x7fffffffLL) return fail(ctx, ASM_ERR_CAPACITY, "np_with_buffers %lld exceeds 2^31-1", (long long)n);
    if (np_local < 0 || np_local > n) return fail(ctx, ASM_ERR_ARG, "np_local %lld out of range", (long long)np_local);
    if (!(boxsize > 0.0f)) return fail(ctx, ASM_ERR_ARG, "boxsize must be positive");
    ctx->n = n;
    ctx->np_local = np_local;
    ctx->boxsize = boxsize;
    ctx->have_particles = true;
    ctx->have_tree = false;
    ctx->have_search = false;
    return ASM_OK;
}

int settle_depositions(asm_context* ctx);

// Redo the pending depositions that did not fit before the context is read
// or changed (a no-op when none is pending).
#define SETTLE(ctx)                                  \
    do {                                             \
        const int settle_rc_ = settle_depositions(ctx); \
        if (settle_rc_) return settle_rc_;           \
    } while (0)

}  // namespace

extern "C" {

const char* asm_version(void)
{
    return "asmooth_amd 0.1 (gfx950)";
}

int asm_create(asm_context** out, int device)
{
    if (!out) return ASM_ERR_ARG;
    *out = nullptr;
    asm_context* ctx = new asm_context();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return ASM_ERR_HIP;
    }
    *out = ctx;
    return ASM_OK;
}

void asm_destroy(asm_context* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* asm_last_error(const asm_context* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int asm_set_option(asm_context* ctx, int option, int value)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    switch (option) {
    case ASM_OPT_KEEP_NEIGHBORS: ctx->keep_neighbors = value ? 1 : 0; return ASM_OK;
    case ASM_OPT_PROFILE: ctx->profile = value ? 1 : 0; return ASM_OK;
    case ASM_OPT_LEAF_SIZE:
        if (value < 8 || value > kWave) return fail(ctx, ASM_ERR_ARG, "leaf size %d outside 8..64", value);
        ctx->leaf_size = value;
        ctx->have_tree = false;
        ctx->have_search = false;
        return ASM_OK;
    case ASM_OPT_DETERMINISTIC: ctx->deterministic = value ? 1 : 0; return ASM_OK;
    case ASM_OPT_SEED_WALK:
        if (value < 0 || value > 2) return fail(ctx, ASM_ERR_ARG, "ASM_OPT_SEED_WALK %d outside 0..2", value);
        ctx->seed_walk = value;
        return ASM_OK;
    case ASM_OPT_KNN_METHOD:
        if (value != kKnnBoundCollect && value != kKnnInsertion)
            return fail(ctx, ASM_ERR_ARG, "kNN method %d is not 0 or 1", value);
        ctx->knn_method = value;
        return ASM_OK;
    case ASM_OPT_KNN_BOUND:
        if (value != kBoundSelect && value != kBoundPacket)
            return fail(ctx, ASM_ERR_ARG, "kNN bound walk %d is not 0 or 1", value);
        ctx->knn_bound = value;
        return ASM_OK;
    case ASM_OPT_DEPOSIT_ACCUM:
        if (value != kAccumFp64 && value != kAccumKahan32 && value != kAccumAtomic32)
            return fail(ctx, ASM_ERR_ARG, "deposition accumulation %d is not 0, 1 or 2", value);
        ctx->deposit_accum = value;
        return ASM_OK;
    case ASM_OPT_KERNEL:
        if (value != kKernTent && value != kKernSpline)
            return fail(ctx, ASM_ERR_ARG, "kernel %d is not 0 (tent) or 1 (cubic spline)", value);
        ctx->kernel = value;
        ctx->have_search = false;   // densities depend on the kernel: search again
        return ASM_OK;
    case ASM_OPT_TILE:
        if (value < 1 || value > 8) return fail(ctx, ASM_ERR_ARG, "tile %d outside 1..8", value);
        ctx->tile = value;
        ctx->tile_for_nc.clear();
        return ASM_OK;
    default: return fail(ctx, ASM_ERR_ARG, "unknown option %d", option);
    }
}

int asm_synchronize(asm_context* ctx)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    HIPC(hipSetDevice(ctx->device));
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(resolve_timers(ctx));
    return ASM_OK;
}

int asm_set_particles(asm_context* ctx, const asm_particle* particles, int64_t n, int64_t np_local,
                      float boxsize)
{
    if (!ctx || !particles) return fail(ctx, ASM_ERR_ARG, "null argument");
    SETTLE(ctx);
    int rc = load_common(ctx, n, np_local, boxsize);
    if (rc) return rc;
    HIPC(hipSetDevice(ctx->device));
    HIPC(ctx->input.ensure(static_cast<size_t>(n)));
    HIPC(hipMemcpyAsync(ctx->input.p, particles, sizeof(ParticleRec) * static_cast<size_t>(n),
                        hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ASM_OK;
}

int asm_set_particles_device(asm_context* ctx, const asm_particle* d_particles, int64_t n, int64_t np_local,
                             float boxsize)
{
    if (!ctx || !d_particles) return fail(ctx, ASM_ERR_ARG, "null argument");
    SETTLE(ctx);
    int rc = load_common(ctx, n, np_local, boxsize);
    if (rc) return rc;
    HIPC(hipSetDevice(ctx->device));
    HIPC(ctx->input.ensure(static_cast<size_t>(n)));
    HIPC(hipMemcpyAsync(ctx->input.p, d_particles, sizeof(ParticleRec) * static_cast<size_t>(n),
                        hipMemcpyDeviceToDevice, ctx->stream));
    return ASM_OK;
}

int asm_tree_construct(asm_context* ctx, float left, float right, int64_t* nnode)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!ctx->have_particles) return fail(ctx, ASM_ERR_STATE, "asm_tree_construct before asm_set_particles");
    if (!(right > left)) return fail(ctx, ASM_ERR_ARG, "tree cube [%g, %g] is empty", left, right);
    HIPC(hipSetDevice(ctx->device));
    ctx->tree_left = left;
    ctx->tree_right = right;
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    HIPC(ctx->keys.ensure(n));
    HIPC(ctx->vals.ensure(n));
    HIPC(ctx->keys_tmp.ensure(n));
    HIPC(ctx->vals_tmp.ensure(n));
    HIPC(ctx->hist.ensure(static_cast<size_t>(radix_hist_entries(n))));
    HIPC(ctx->scan_tmp.ensure(static_cast<size_t>(scan_scratch_entries(radix_hist_entries(n)))));
    HIPC(ctx->spos.ensure(n));
    HIPC(ctx->sattr.ensure(n));
    HIPC(ctx->inv_perm.ensure(n));
    {
        StageTimer t(ctx, T_SORT);
        HIPC(launch_morton_keys(ctx->input.p, n, left, right, ctx->keys.p, ctx->vals.p, s));
        bool in_tmp = false;
        HIPC(launch_radix_sort(ctx->keys.p, ctx->vals.p, ctx->keys_tmp.p, ctx->vals_tmp.p, n, kMortonBits, ctx->hist.p,
                               ctx->scan_tmp.p, s, &in_tmp));
        ctx->perm = in_tmp ? ctx->vals_tmp.p : ctx->vals.p;
    }
    ctx->nleaf = (n + ctx->leaf_size - 1) / ctx->leaf_size;
    ctx->nleaf_pad = next_pow2(ctx->nleaf);
    HIPC(ctx->nodes.ensure(static_cast<size_t>(2 * ctx->nleaf_pad)));
    const bool use_groups = ctx->leaf_size % kGroup == 0;
    if (use_groups) HIPC(ctx->groups.ensure(static_cast<size_t>((n + kGroup - 1) / kGroup)));
    {
        StageTimer t(ctx, T_TREE);
        HIPC(launch_gather_sorted(ctx->input.p, ctx->perm, n, ctx->spos.p, ctx->sattr.p, ctx->inv_perm.p, s));
        HIPC(launch_build_tree(ctx->spos.p, n, ctx->leaf_size, ctx->nleaf, ctx->nleaf_pad, ctx->nodes.p,
                               use_groups ? ctx->groups.p : nullptr, s));
    }
    ctx->have_tree = true;
    ctx->have_search = false;
    if (nnode) *nnode = 2 * ctx->nleaf_pad - 1;
    return ASM_OK;
}

int asm_neighbor_search(asm_context* ctx, int k)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (k != kSlots - 1) return fail(ctx, ASM_ERR_ARG, "k must be %d (33 slots incl. self), got %d", kSlots - 1, k);
    if (!ctx->have_tree) return fail(ctx, ASM_ERR_STATE, "asm_neighbor_search before asm_tree_construct");
    HIPC(hipSetDevice(ctx->device));
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    HIPC(ctx->rk.ensure(n));
    HIPC(ctx->dens.ensure(n));
    HIPC(ctx->kkey.ensure(n));
    HIPC(ctx->kmax.ensure(static_cast<size_t>(2 * ctx->nleaf_pad)));
    HIPC(ctx->kstats.ensure(kStatSlots));
    const size_t nwaves = static_cast<size_t>((n + kWave - 1) / kWave);
    HIPC(ctx->knn_flags.ensure(nwaves));
    HIPC(ctx->knn_leaf_n.ensure(nwaves));
    const bool bound_collect = ctx->knn_method == kKnnBoundCollect;
    if (bound_collect) {
        HIPC(ctx->knn_leaves.ensure(nwaves * kLeafCap));
        HIPC(ctx->knn_tau.ensure(n));
        if (!ctx->deterministic) HIPC(ctx->dens_pull.ensure(n));
    }
    // Lists of the selection walk (not with kept ordered lists, which the
    // collect walk writes)
    const bool sel_lists = bound_collect && ctx->knn_bound == kBoundSelect && !ctx->keep_neighbors;
    if (sel_lists) {
        HIPC(ctx->knn_list.ensure(nwaves * kWave * kListCap));
        HIPC(ctx->knn_list_n.ensure(n));
    }
    KnnScratch kscr{ctx->knn_leaves.p, ctx->knn_leaf_n.p, ctx->knn_flags.p, ctx->knn_tau.p, ctx->dens_pull.p,
                    nullptr, nullptr, nullptr, ctx->leaf_size % kGroup == 0 ? ctx->groups.p : nullptr, 0,
                    ctx->kernel, ctx->knn_bound, sel_lists ? ctx->knn_list.p : nullptr,
                    sel_lists ? ctx->knn_list_n.p : nullptr};
    ctx->seed_used = false;
    if (bound_collect && ctx->seed_on) {
        HIPC(ctx->seed_grid.ensure(static_cast<size_t>(kMlpGrid) * kMlpGrid * kMlpGrid));
        HIPC(ctx->seed_rcell.ensure(static_cast<size_t>(kMlpGrid) * kMlpGrid * kMlpGrid));
        HIPC(ctx->tau_init.ensure(n));
        HIPC(ctx->seed_redo.ensure(nwaves));
        HIPC(ctx->seed_fb.ensure(kStatSlots));
        const float side = ctx->tree_right - ctx->tree_left;
        const float spacing = std::cbrt(side * side * side / static_cast<float>(n));
        StageTimer t(ctx, T_SEED);
        HIPC(launch_fill_zero_u32(reinterpret_cast<uint32_t*>(ctx->seed_fb.p), 2 * kStatSlots, s));
        HIPC(launch_seed_grid(ctx->spos.p, n, ctx->tree_left, ctx->tree_right, ctx->seed_grid.p, s));
        HIPC(launch_seed_bound(ctx->spos.p, n, ctx->seed_grid.p, ctx->tree_left, ctx->tree_right, ctx->seed_params.p,
                               spacing * (1.0f + ctx->seed_margin), ctx->seed_rcell.p, ctx->tau_init.p, s));
        kscr.tau_init = ctx->tau_init.p;
        kscr.seed_redo = ctx->seed_redo.p;
        kscr.seed_fallbacks = ctx->seed_fb.p;
        kscr.seed_walk = ctx->seed_walk;
        ctx->seed_used = true;
    }
    int32_t* nbr = nullptr;
    if (ctx->keep_neighbors) {
        HIPC(ctx->nbr.ensure(static_cast<size_t>(n) * kSlots));
        nbr = ctx->nbr.p;
    }
    if (!ctx->deterministic) HIPC(ctx->dens64.ensure(n));
    {
        StageTimer t(ctx, T_KNN);
        HIPC(launch_fill_zero_u32(reinterpret_cast<uint32_t*>(ctx->kstats.p), kStatSlots * sizeof(KnnStats) / 4, s));
        if (!ctx->deterministic) HIPC(launch_fill_zero_f64(ctx->dens64.p, n, s));
        KnnMarkState ms{ctx};
        const KnnMarks marks{&ms, knn_mark_begin, knn_mark_end};
        HIPC(launch_knn(ctx->spos.p, n, ctx->nodes.p, ctx->nleaf_pad, ctx->nleaf, ctx->leaf_size, ctx->rk.p,
                        ctx->kkey.p, ctx->inv_perm.p, ctx->deterministic ? nullptr : ctx->dens64.p, nbr, ctx->kstats.p,
                        kscr, ctx->knn_method, s, &marks));
    }
    {
        StageTimer t(ctx, T_DENS);
        if (ctx->deterministic) {
            HIPC(launch_kmax(ctx->kkey.p, n, ctx->leaf_size, ctx->nleaf, ctx->nleaf_pad, ctx->kmax.p, s));
            HIPC(launch_dens_gather(ctx->spos.p, ctx->kkey.p, ctx->rk.p, n, ctx->nodes.p, ctx->kmax.p,
                                    ctx->nleaf_pad, ctx->dens.p, ctx->kernel, s));
        } else {
            HIPC(launch_dens_finalize(ctx->dens64.p, bound_collect ? ctx->dens_pull.p : nullptr, n, ctx->dens.p, s));
        }
    }
    ctx->have_search = true;
    return ASM_OK;
}

int asm_mark_halo_particles(asm_context* ctx, const asm_halo* halos, int64_t n_halo, float left, float right,
                            int64_t* np_halo, double* m_halo)
{
    if (!ctx || (n_halo > 0 && !halos)) return fail(ctx, ASM_ERR_ARG, "null argument");
    SETTLE(ctx);
    if (!ctx->have_tree) return fail(ctx, ASM_ERR_STATE, "asm_mark_halo_particles before asm_tree_construct");
    HIPC(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    HIPC(ctx->halo_count.ensure(1));
    HIPC(launch_fill_zero_u32(reinterpret_cast<uint32_t*>(ctx->halo_count.p), 2, s));
    double mass = 0.0;
    for (int64_t i = 0; i < n_halo; ++i) {
        const asm_halo& h = halos[i];
        if (left < h.x[0] && h.x[0] < right && left < h.x[1] && h.x[1] < right && left < h.x[2] && h.x[2] < right)
            mass += static_cast<double>(h.mass / 8.0f);
    }
    if (n_halo > 0) {
        HIPC(ctx->halos.ensure(static_cast<size_t>(n_halo)));
        HIPC(hipMemcpyAsync(ctx->halos.p, halos, sizeof(HaloRec) * static_cast<size_t>(n_halo),
                            hipMemcpyHostToDevice, s));
        StageTimer t(ctx, T_HALO);
        HIPC(launch_mark_halo(ctx->spos.p, ctx->sattr.p, ctx->n, ctx->nodes.p, ctx->nleaf_pad, ctx->halos.p, n_halo,
                              left, right, ctx->halo_count.p, s));
    }
    unsigned long long cnt = 0;
    HIPC(hipMemcpyAsync(&cnt, ctx->halo_count.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (np_halo) *np_halo = static_cast<int64_t>(cnt);
    if (m_halo) *m_halo = mass;
    return ASM_OK;
}

}  // extern "C"

namespace {

constexpr size_t kMaxPendingDeposits = 16;

// Whether a deposition of meshes `ncs` may skip the read-back of its list
// size: a caller's own call (defer_ok: not the autotuner's, a batch's or a
// redo), outside batches and capture, with a list buffer from an earlier
// call.  A pending deposition of one of the same meshes is settled first,
// so that a later redo never overwrites a newer mesh.
int may_defer(asm_context* ctx, const std::vector<int>& ncs, bool defer_ok, bool* defer)
{
    *defer = false;
    if (!defer_ok || ctx->capturing || ctx->overflow_flag || ctx->list.cap == 0) return ASM_OK;
    bool overlap = false;
    for (const auto& p : ctx->dep_pending)
        for (int nc : p.ncs)
            if (std::find(ncs.begin(), ncs.end(), nc) != ncs.end()) overlap = true;
    if (overlap || ctx->dep_pending.size() >= kMaxPendingDeposits) {
        const int rc = settle_depositions(ctx);
        if (rc) return rc;
    }
    HIPC(ctx->dep_flags.ensure(kMaxPendingDeposits));
    *defer = true;
    return ASM_OK;
}

// Size the list buffer from the counted total (read back), with an eighth
// more for the next call on slightly denser inputs.
int size_list(asm_context* ctx, const char* what, int nc, int tile)
{
    uint64_t total = 0;
    HIPC(hipMemcpyAsync(&total, ctx->scan_total.p, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    // List offsets are 32-bit: more (particle, tile) entries than that
    // would wrap them.
    if (total > 0xffffffffull)
        return fail(ctx, ASM_ERR_CAPACITY, "%s (nc %d, tile %d) needs %llu tile-list entries, more than 2^32 - 1; use larger tiles",
                    what, nc, tile, static_cast<unsigned long long>(total));
    const uint64_t want = total + total / 8 + 1;
    HIPC(ctx->list.ensure(want < 0xffffffffull ? want : 0xffffffffull));
    return ASM_OK;
}

int deposit_one(asm_context* ctx, int nc, int sign, int periodic, const int* coord, bool defer_ok = false)
{
    if (nc <= 0 || nc > 2048) return fail(ctx, ASM_ERR_ARG, "nc %d outside 1..2048", nc);
    if (!ctx->have_search) return fail(ctx, ASM_ERR_STATE, "mesh assignment before asm_neighbor_search");
    HIPC(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t nmesh = static_cast<int64_t>(nc) * nc * nc;
    MeshDev& m = ctx->meshes[nc];
    HIPC(m.mesh.ensure(static_cast<size_t>(6 * nmesh)));

    DepositArgs a{};
    a.spos = ctx->spos.p;
    a.sattr = ctx->sattr.p;
    a.rk = ctx->rk.p;
    a.dens = ctx->dens.p;
    a.n = ctx->n;
    a.boxsize = ctx->boxsize;
    a.nc = nc;
    a.sign = sign;
    {
        const auto it = ctx->tile_for_nc.find(nc);
        a.tile = (it != ctx->tile_for_nc.end() && !periodic) ? it->second : ctx->tile;
    }
    a.ntx = (nc + a.tile - 1) / a.tile;
    a.periodic = periodic;
    a.accum = ctx->deposit_accum;
    a.kern = ctx->kernel;
    a.coord[0] = coord ? coord[0] : 0;
    a.coord[1] = coord ? coord[1] : 0;
    a.coord[2] = coord ? coord[2] : 0;
    const int64_t ntiles = static_cast<int64_t>(a.ntx) * a.ntx * a.ntx;
    if (ntiles > 0x7fffffffLL) return fail(ctx, ASM_ERR_CAPACITY, "nc %d with tile %d gives too many tiles", nc, a.tile);
    HIPC(ctx->tile_cnt.ensure(ntiles));
    HIPC(ctx->tile_start.ensure(ntiles));
    HIPC(ctx->tile_cur.ensure(ntiles));
    HIPC(ctx->scan_total.ensure(1));
    HIPC(ctx->scan_tmp.ensure(static_cast<size_t>(scan_scratch_entries(ntiles))));
    bool defer = false;
    {
        const int rc = may_defer(ctx, std::vector<int>{nc}, defer_ok, &defer);
        if (rc) return rc;
    }
    {
        StageTimer t(ctx, T_DEPOSIT_LIST);
        HIPC(launch_fill_zero_u32(ctx->tile_cnt.p, ntiles, s));
        HIPC(launch_deposit_count(a, ctx->tile_cnt.p, s));
        HIPC(hipMemcpyAsync(ctx->tile_start.p, ctx->tile_cnt.p, sizeof(uint32_t) * ntiles, hipMemcpyDeviceToDevice, s));
        HIPC(launch_scan_exclusive(ctx->tile_start.p, ntiles, ctx->scan_total.p, ctx->scan_tmp.p, s));
    }
    if (!ctx->capturing && !defer) {
        const int rc = size_list(ctx, "the mesh", nc, a.tile);
        if (rc) return rc;
    }
    a.list_cap = static_cast<uint32_t>(ctx->list.cap < 0xffffffffull ? ctx->list.cap : 0xffffffffull);
    if (ctx->overflow_flag) HIPC(launch_flag_overflow(ctx->scan_total.p, a.list_cap, ctx->overflow_flag, s));
    if (defer) {
        uint32_t* const flag = ctx->dep_flags.p + ctx->dep_pending.size();
        HIPC(launch_fill_zero_u32(flag, 1, s));
        HIPC(launch_flag_overflow(ctx->scan_total.p, a.list_cap, flag, s));
        asm_context::PendingDeposit pd;
        pd.ncs = {nc};
        pd.sign = sign;
        pd.periodic = periodic;
        for (int d = 0; d < 3; ++d) pd.coord[d] = a.coord[d];
        ctx->dep_pending.push_back(pd);
    }
    {
        StageTimer t(ctx, T_DEPOSIT_LIST);
        HIPC(launch_fill_zero_u32(ctx->tile_cur.p, ntiles, s));
        HIPC(launch_deposit_fill(a, ctx->tile_start.p, ctx->tile_cur.p, ctx->list.p, s));
    }
    {
        StageTimer t(ctx, T_DEPOSIT_TILES);
        HIPC(launch_deposit_tiles(a, ctx->tile_start.p, ctx->tile_cnt.p, ctx->list.p, m.mesh.p, 0, s));
    }
    return ASM_OK;
}

// Every node mesh of an nc list together (FP64 sums, the default): one count
// and one fill pass over the particles for all meshes, one scan of the
// concatenated tile counters, one read-back of the list total (none while a
// batch records) and one launch over all meshes' tiles, coarse meshes first.
// Other accumulation modes take the per-mesh path.
int deposit_meshes(asm_context* ctx, const int* ncs, int n_nc, int sign, bool defer_ok = false)
{
    std::vector<int> list_nc;
    for (int i = 0; i < n_nc; ++i) {
        if (ncs[i] <= 0 || ncs[i] > 2048) return fail(ctx, ASM_ERR_ARG, "nc %d outside 1..2048", ncs[i]);
        if (std::find(list_nc.begin(), list_nc.end(), ncs[i]) == list_nc.end()) list_nc.push_back(ncs[i]);
    }
    if (ctx->deposit_accum != kAccumFp64 || static_cast<int>(list_nc.size()) > kMaxMeshes) {
        for (int nc : list_nc) {
            const int rc = deposit_one(ctx, nc, sign, 0, nullptr, defer_ok);
            if (rc) return rc;
        }
        return ASM_OK;
    }
    if (!ctx->have_search) return fail(ctx, ASM_ERR_STATE, "mesh assignment before asm_neighbor_search");
    HIPC(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::sort(list_nc.begin(), list_nc.end());   // coarse meshes (long tiles) first
    MultiDeposit J{};
    J.n = static_cast<int>(list_nc.size());
    uint64_t tiles = 0;
    for (int j = 0; j < J.n; ++j) {
        const int nc = list_nc[j];
        const int64_t nmesh = static_cast<int64_t>(nc) * nc * nc;
        MeshDev& m = ctx->meshes[nc];
        HIPC(m.mesh.ensure(static_cast<size_t>(6 * nmesh)));
        DepositArgs& a = J.a[j];
        a.spos = ctx->spos.p;
        a.sattr = ctx->sattr.p;
        a.rk = ctx->rk.p;
        a.dens = ctx->dens.p;
        a.n = ctx->n;
        a.boxsize = ctx->boxsize;
        a.nc = nc;
        a.sign = sign;
        const auto it = ctx->tile_for_nc.find(nc);
        a.tile = it != ctx->tile_for_nc.end() ? it->second : ctx->tile;
        a.ntx = (nc + a.tile - 1) / a.tile;
        a.periodic = 0;
        a.accum = kAccumFp64;
        a.kern = ctx->kernel;
        J.mesh[j] = m.mesh.p;
        J.tile_base[j] = static_cast<uint32_t>(tiles);
        tiles += static_cast<uint64_t>(a.ntx) * a.ntx * a.ntx;
        if (tiles > 0x7fffffffull) return fail(ctx, ASM_ERR_CAPACITY, "the nc list gives too many tiles");
    }
    J.tile_base[J.n] = static_cast<uint32_t>(tiles);
    const int64_t ntiles = static_cast<int64_t>(tiles);
    HIPC(ctx->tile_cnt.ensure(ntiles));
    HIPC(ctx->tile_start.ensure(ntiles));
    HIPC(ctx->tile_cur.ensure(ntiles));
    HIPC(ctx->scan_total.ensure(1));
    HIPC(ctx->scan_tmp.ensure(static_cast<size_t>(scan_scratch_entries(ntiles))));
    bool defer = false;
    {
        const int rc = may_defer(ctx, list_nc, defer_ok, &defer);
        if (rc) return rc;
    }
    {
        StageTimer t(ctx, T_DEPOSIT_LIST);
        HIPC(launch_fill_zero_u32(ctx->tile_cnt.p, ntiles, s));
        HIPC(launch_multi_count(J, ctx->tile_cnt.p, s));
        HIPC(hipMemcpyAsync(ctx->tile_start.p, ctx->tile_cnt.p, sizeof(uint32_t) * ntiles, hipMemcpyDeviceToDevice, s));
        HIPC(launch_scan_exclusive(ctx->tile_start.p, ntiles, ctx->scan_total.p, ctx->scan_tmp.p, s));
    }
    if (!ctx->capturing && !defer) {
        const int rc = size_list(ctx, "the nc list", list_nc.back(), J.a[J.n - 1].tile);
        if (rc) return rc;
    }
    const uint32_t cap = static_cast<uint32_t>(ctx->list.cap < 0xffffffffull ? ctx->list.cap : 0xffffffffull);
    for (int j = 0; j < J.n; ++j) J.a[j].list_cap = cap;
    if (ctx->overflow_flag) HIPC(launch_flag_overflow(ctx->scan_total.p, cap, ctx->overflow_flag, s));
    if (defer) {
        uint32_t* const flag = ctx->dep_flags.p + ctx->dep_pending.size();
        HIPC(launch_fill_zero_u32(flag, 1, s));
        HIPC(launch_flag_overflow(ctx->scan_total.p, cap, flag, s));
        asm_context::PendingDeposit pd;
        pd.ncs = list_nc;
        pd.sign = sign;
        pd.multi = true;
        ctx->dep_pending.push_back(pd);
    }
    {
        StageTimer t(ctx, T_DEPOSIT_LIST);
        HIPC(launch_fill_zero_u32(ctx->tile_cur.p, ntiles, s));
        HIPC(launch_multi_fill(J, ctx->tile_start.p, ctx->tile_cur.p, ctx->list.p, s));
    }
    {
        StageTimer t(ctx, T_DEPOSIT_TILES);
        HIPC(launch_multi_tiles(J, ctx->tile_start.p, ctx->tile_cnt.p, ctx->list.p, s));
    }
    return ASM_OK;
}

int settle_depositions(asm_context* ctx)
{
    if (ctx->dep_pending.empty()) return ASM_OK;
    std::vector<asm_context::PendingDeposit> pend;
    pend.swap(ctx->dep_pending);
    std::vector<uint32_t> flags(pend.size(), 0u);
    HIPC(hipSetDevice(ctx->device));
    HIPC(hipMemcpyAsync(flags.data(), ctx->dep_flags.p, sizeof(uint32_t) * pend.size(), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    // The pending calls cover disjoint meshes (may_defer), so their order
    // does not matter.
    for (size_t i = 0; i < pend.size(); ++i) {
        if (!flags[i]) continue;
        const auto& p = pend[i];
        const int rc = p.multi ? deposit_meshes(ctx, p.ncs.data(), static_cast<int>(p.ncs.size()), p.sign)
                               : deposit_one(ctx, p.ncs[0], p.sign, p.periodic, p.periodic ? p.coord : nullptr);
        if (rc) return rc;
    }
    return ASM_OK;
}

}  // namespace

extern "C" {

int asm_assign_on_mesh(asm_context* ctx, int nc, int sign)
{
    if (!ctx) return ASM_ERR_ARG;
    return deposit_one(ctx, nc, sign, 0, nullptr, true);
}

int asm_assign_on_meshes(asm_context* ctx, const int* ncs, int n_nc, int sign)
{
    if (!ctx || !ncs || n_nc <= 0) return fail(ctx, ASM_ERR_ARG, "empty nc list");
    return deposit_meshes(ctx, ncs, n_nc, sign, true);
}

int asm_assign_on_global_mesh(asm_context* ctx, int nc, int sign, const int coord[3])
{
    if (!ctx || !coord) return fail(ctx, ASM_ERR_ARG, "null argument");
    return deposit_one(ctx, nc, sign, 1, coord, true);
}

}  // extern "C"

namespace {

// Milliseconds of fn() on the context stream (one warm-up call, then one timed).
template <typename F>
int time_on_stream(asm_context* ctx, F fn, float* ms)
{
    int rc = fn();
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, ctx->stream));
    rc = fn();
    HIPC(hipEventRecord(e1, ctx->stream));
    HIPC(hipEventSynchronize(e1));
    HIPC(hipEventElapsedTime(ms, e0, e1));
    HIPC(hipEventDestroy(e0));
    HIPC(hipEventDestroy(e1));
    return rc;
}

}  // namespace

extern "C" {

int asm_autotune(asm_context* ctx, const int* leaf_sizes, int n_leaf, const int* ncs, int n_nc, int sign,
                 int* leaf_out, int* tiles_out)
{
    if (!ctx || (n_leaf > 0 && !leaf_sizes) || (n_nc > 0 && !ncs)) return fail(ctx, ASM_ERR_ARG, "null argument");
    SETTLE(ctx);
    if (!ctx->have_search) return fail(ctx, ASM_ERR_STATE, "asm_autotune before asm_neighbor_search");
    HIPC(hipSetDevice(ctx->device));
    const int profile = ctx->profile;
    ctx->profile = 0;   // keep the stage timers for the caller's own runs
    int rc = ASM_OK;
    if (n_leaf > 0) {
        int best = ctx->leaf_size;
        float best_ms = INFINITY;
        for (int i = 0; i < n_leaf && rc == ASM_OK; ++i) {
            const int ls = leaf_sizes[i];
            if (ls < 8 || ls > kWave) {
                rc = fail(ctx, ASM_ERR_ARG, "leaf size %d outside 8..64", ls);
                break;
            }
            ctx->leaf_size = ls;
            float ms = 0.0f;
            rc = time_on_stream(ctx, [&]() {
                int r = asm_tree_construct(ctx, ctx->tree_left, ctx->tree_right, nullptr);
                return r ? r : asm_neighbor_search(ctx, kSlots - 1);
            }, &ms);
            if (rc == ASM_OK && ms < best_ms) {
                best_ms = ms;
                best = ls;
            }
        }
        ctx->leaf_size = best;
        if (rc == ASM_OK) rc = asm_tree_construct(ctx, ctx->tree_left, ctx->tree_right, nullptr);
        if (rc == ASM_OK) rc = asm_neighbor_search(ctx, kSlots - 1);
    }
    if (leaf_out) *leaf_out = ctx->leaf_size;
    for (int i = 0; i < n_nc && rc == ASM_OK; ++i) {
        const int nc = ncs[i];
        int best = ctx->tile;
        float best_ms = INFINITY;
        for (int tile = 4; tile <= 8 && rc == ASM_OK; ++tile) {
            ctx->tile_for_nc[nc] = tile;
            float ms = 0.0f;
            rc = time_on_stream(ctx, [&]() { return deposit_one(ctx, nc, sign, 0, nullptr); }, &ms);
            if (rc == ASM_OK && ms < best_ms) {
                best_ms = ms;
                best = tile;
            }
        }
        ctx->tile_for_nc[nc] = best;
        if (tiles_out) tiles_out[i] = best;
    }
    ctx->profile = profile;
    return rc;
}

int asm_set_seed_model(asm_context* ctx, const asm_mlp* m, float margin)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!m) {
        ctx->seed_on = false;
        return ASM_OK;
    }
    if (!(margin >= 0.0f)) return fail(ctx, ASM_ERR_ARG, "seed margin %g is negative", margin);
    HIPC(hipSetDevice(ctx->device));
    std::vector<float> p(kMlpParams);
    const int rc = asm_mlp_get_params(const_cast<asm_mlp*>(m), p.data());
    if (rc) return fail(ctx, rc, "cannot read the seed model's parameters");
    HIPC(ctx->seed_params.ensure(kMlpParams));
    HIPC(hipMemcpyAsync(ctx->seed_params.p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->seed_on = true;
    ctx->seed_margin = margin;
    return ASM_OK;
}

int asm_get_seed_stats(asm_context* ctx, int64_t out[3])
{
    if (!ctx || !out) return ASM_ERR_ARG;
    SETTLE(ctx);
    out[0] = out[1] = 0;
    out[2] = ctx->n;
    if (!ctx->seed_used) return ASM_OK;
    HIPC(hipSetDevice(ctx->device));
    const size_t nwaves = static_cast<size_t>((ctx->n + kWave - 1) / kWave);
    std::vector<unsigned long long> fb(kStatSlots);
    std::vector<uint32_t> redo(nwaves);
    HIPC(hipMemcpyAsync(fb.data(), ctx->seed_fb.p, sizeof(unsigned long long) * kStatSlots, hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipMemcpyAsync(redo.data(), ctx->seed_redo.p, sizeof(uint32_t) * nwaves, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    for (unsigned long long v : fb) out[0] += static_cast<int64_t>(v);
    for (uint32_t v : redo) out[1] += v ? 1 : 0;
    return ASM_OK;
}

int asm_seed_samples(asm_context* ctx, const int64_t* idx, int64_t m, float* x, float* target)
{
    if (!ctx || m < 0 || (m > 0 && (!idx || !x || !target))) return fail(ctx, ASM_ERR_ARG, "null argument");
    SETTLE(ctx);
    if (!ctx->have_search) return fail(ctx, ASM_ERR_STATE, "asm_seed_samples before asm_neighbor_search");
    if (m == 0) return ASM_OK;
    for (int64_t b = 0; b < m; ++b)
        if (idx[b] < 0 || idx[b] >= ctx->n) return fail(ctx, ASM_ERR_ARG, "index %lld outside the particles",
                                                       static_cast<long long>(idx[b]));
    HIPC(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    HIPC(ctx->seed_grid.ensure(static_cast<size_t>(kMlpGrid) * kMlpGrid * kMlpGrid));
    DevBuf<int64_t> didx;
    DevBuf<int32_t> sidx;
    DevBuf<float> dx, dt;
    HIPC(didx.ensure(m));
    HIPC(sidx.ensure(m));
    HIPC(dx.ensure(static_cast<size_t>(m) * kMlpIn));
    HIPC(dt.ensure(m));
    const float side = ctx->tree_right - ctx->tree_left;
    const float spacing = std::cbrt(side * side * side / static_cast<float>(ctx->n));
    HIPC(hipMemcpyAsync(didx.p, idx, sizeof(int64_t) * m, hipMemcpyHostToDevice, s));
    HIPC(launch_seed_grid(ctx->spos.p, ctx->n, ctx->tree_left, ctx->tree_right, ctx->seed_grid.p, s));
    HIPC(launch_seed_samples(ctx->spos.p, ctx->rk.p, ctx->inv_perm.p, didx.p, m, ctx->seed_grid.p, ctx->tree_left,
                             ctx->tree_right, 1.0f / spacing, sidx.p, dx.p, dt.p, s));
    HIPC(hipMemcpyAsync(x, dx.p, sizeof(float) * m * kMlpIn, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(target, dt.p, sizeof(float) * m, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return ASM_OK;
}

int asm_mesh_stats(asm_context* ctx, int nc, double out[4])
{
    if (!ctx || !out) return ASM_ERR_ARG;
    SETTLE(ctx);
    auto it = ctx->meshes.find(nc);
    if (it == ctx->meshes.end()) return fail(ctx, ASM_ERR_STATE, "no mesh for nc %d", nc);
    HIPC(hipSetDevice(ctx->device));
    HIPC(ctx->red_partial.ensure(4 * 1024));
    HIPC(ctx->red_out.ensure(4));
    {
        StageTimer t(ctx, T_REDUCE);
        HIPC(launch_mesh_reduce(it->second.mesh.p, nc, ctx->red_partial.p, ctx->red_out.p, ctx->stream));
    }
    HIPC(hipMemcpyAsync(out, ctx->red_out.p, sizeof(double) * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ASM_OK;
}

int asm_mesh_total(asm_context* ctx, int nc, double totals[2])
{
    double out[4];
    int rc = asm_mesh_stats(ctx, nc, out);
    if (rc) return rc;
    totals[0] = out[0];
    totals[1] = out[1];
    return ASM_OK;
}

int asm_normalize_mesh(asm_context* ctx, int nc)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    auto it = ctx->meshes.find(nc);
    if (it == ctx->meshes.end()) return fail(ctx, ASM_ERR_STATE, "no mesh for nc %d", nc);
    HIPC(hipSetDevice(ctx->device));
    // coarse_mesh.cpp:151: vol = pow(boxsize/nc, 3) in FP32.
    const float vol = std::pow(ctx->boxsize / static_cast<float>(nc), 3.0f);
    StageTimer t(ctx, T_NORMALIZE);
    HIPC(launch_normalize(it->second.mesh.p, nc, vol, ctx->stream));
    return ASM_OK;
}

int asm_stitch_mesh(asm_context* ctx, int nc, int nn, const int coord[3], int field, float* dst)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!coord || !dst) return fail(ctx, ASM_ERR_ARG, "null argument");
    if (nn <= 0 || field < 0 || field > 5) return fail(ctx, ASM_ERR_ARG, "nn %d or field %d out of range", nn, field);
    for (int a = 0; a < 3; ++a)
        if (coord[a] < 0 || coord[a] >= nn) return fail(ctx, ASM_ERR_ARG, "node coordinate outside 0..%d", nn - 1);
    if (static_cast<int64_t>(nc) * nn > 0x7fffffffLL)
        return fail(ctx, ASM_ERR_CAPACITY, "global side %lld beyond int32", static_cast<long long>(nc) * nn);
    auto it = ctx->meshes.find(nc);
    if (it == ctx->meshes.end()) return fail(ctx, ASM_ERR_STATE, "no mesh for nc %d", nc);
    HIPC(hipSetDevice(ctx->device));
    StageTimer t(ctx, T_STITCH);
    HIPC(launch_stitch(it->second.mesh.p, nc, nn, coord, field, dst, ctx->stream));
    return ASM_OK;
}

int asm_get_mesh(asm_context* ctx, int nc, float* mesh)
{
    if (!ctx || !mesh) return ASM_ERR_ARG;
    SETTLE(ctx);
    auto it = ctx->meshes.find(nc);
    if (it == ctx->meshes.end()) return fail(ctx, ASM_ERR_STATE, "no mesh for nc %d", nc);
    HIPC(hipSetDevice(ctx->device));
    const size_t nmesh = static_cast<size_t>(nc) * nc * nc;
    HIPC(hipMemcpyAsync(mesh, it->second.mesh.p, sizeof(float) * 6 * nmesh, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ASM_OK;
}

const float* asm_mesh_device_ptr(asm_context* ctx, int nc)
{
    if (!ctx || settle_depositions(ctx)) return nullptr;
    auto it = ctx->meshes.find(nc);
    return it == ctx->meshes.end() ? nullptr : it->second.mesh.p;
}

}  // extern "C"

namespace {

// Sorted-order per-particle values back to the caller's order.
__global__ void unsort_kernel(const float4* __restrict__ spos, const float4* __restrict__ sattr,
                              const float* __restrict__ rk, const float* __restrict__ dens, int64_t n,
                              float* __restrict__ rk_o, float* __restrict__ dens_o, int32_t* __restrict__ igrp_o)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = __float_as_uint(spos[i].w);
    rk_o[o] = rk ? rk[i] : 0.0f;
    dens_o[o] = dens ? dens[i] : 0.0f;
    igrp_o[o] = __float_as_int(sattr[i].w);
}

}  // namespace

extern "C" {

int asm_get_rk_dens(asm_context* ctx, float* rk, float* dens, int32_t* igrp)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!ctx->have_tree) return fail(ctx, ASM_ERR_STATE, "no results before asm_tree_construct");
    HIPC(hipSetDevice(ctx->device));
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    DevBuf<float> r, d;
    DevBuf<int32_t> g;
    HIPC(r.ensure(n));
    HIPC(d.ensure(n));
    HIPC(g.ensure(n));
    unsort_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
        ctx->spos.p, ctx->sattr.p, ctx->have_search ? ctx->rk.p : nullptr,
        ctx->have_search ? ctx->dens.p : nullptr, n, r.p, d.p, g.p);
    HIPC(hipGetLastError());
    if (rk) HIPC(hipMemcpyAsync(rk, r.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (dens) HIPC(hipMemcpyAsync(dens, d.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (igrp) HIPC(hipMemcpyAsync(igrp, g.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return ASM_OK;
}

int asm_get_particles(asm_context* ctx, asm_particle* particles)
{
    if (!ctx || !particles) return ASM_ERR_ARG;
    SETTLE(ctx);
    const int64_t n = ctx->n;
    std::vector<float> rk(n), dens(n);
    std::vector<int32_t> igrp(n);
    int rc = asm_get_rk_dens(ctx, rk.data(), dens.data(), igrp.data());
    if (rc) return rc;
    HIPC(hipMemcpy(particles, ctx->input.p, sizeof(ParticleRec) * static_cast<size_t>(n), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) {
        particles[i].rk = rk[i];
        particles[i].dens = dens[i];
        particles[i].igrp = igrp[i];
    }
    return ASM_OK;
}

int asm_get_neighbors(asm_context* ctx, int32_t* nbr)
{
    if (!ctx || !nbr) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!ctx->have_search || !ctx->keep_neighbors || !ctx->nbr.p)
        return fail(ctx, ASM_ERR_STATE, "neighbour lists not kept (set ASM_OPT_KEEP_NEIGHBORS before the search)");
    HIPC(hipSetDevice(ctx->device));
    HIPC(hipMemcpyAsync(nbr, ctx->nbr.p, sizeof(int32_t) * kSlots * static_cast<size_t>(ctx->n),
                        hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ASM_OK;
}

int asm_get_tree(asm_context* ctx, float* spos, int32_t* nodes, int64_t* nleaf, int64_t* nleaf_pad,
                 int32_t* leaf_size)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!ctx->have_tree) return fail(ctx, ASM_ERR_STATE, "no tree before asm_tree_construct");
    HIPC(hipSetDevice(ctx->device));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (spos) HIPC(hipMemcpy(spos, ctx->spos.p, sizeof(float4) * static_cast<size_t>(ctx->n), hipMemcpyDeviceToHost));
    if (nodes)
        HIPC(hipMemcpy(nodes, ctx->nodes.p, sizeof(Box) * static_cast<size_t>(2 * ctx->nleaf_pad),
                       hipMemcpyDeviceToHost));
    if (nleaf) *nleaf = ctx->nleaf;
    if (nleaf_pad) *nleaf_pad = ctx->nleaf_pad;
    if (leaf_size) *leaf_size = ctx->leaf_size;
    return ASM_OK;
}

int asm_get_sorted_results(asm_context* ctx, float* rk, float* dens)
{
    if (!ctx) return ASM_ERR_ARG;
    SETTLE(ctx);
    if (!ctx->have_search) return fail(ctx, ASM_ERR_STATE, "no search results");
    HIPC(hipSetDevice(ctx->device));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (rk) HIPC(hipMemcpy(rk, ctx->rk.p, sizeof(float) * static_cast<size_t>(ctx->n), hipMemcpyDeviceToHost));
    if (dens) HIPC(hipMemcpy(dens, ctx->dens.p, sizeof(float) * static_cast<size_t>(ctx->n), hipMemcpyDeviceToHost));
    return ASM_OK;
}

int asm_timer_count(void)
{
    return T_COUNT;
}

const char* asm_timer_name(int i)
{
    return (i >= 0 && i < T_COUNT) ? kTimerNames[i] : "";
}

int asm_get_timer(asm_context* ctx, int i, double* total_ms, int64_t* launches)
{
    if (!ctx || i < 0 || i >= T_COUNT) return ASM_ERR_ARG;
    HIPC(hipSetDevice(ctx->device));
    HIPC(resolve_timers(ctx));
    if (total_ms) *total_ms = ctx->timer_ms[i];
    if (launches) *launches = ctx->timer_n[i];
    return ASM_OK;
}

int asm_reset_timers(asm_context* ctx)
{
    if (!ctx) return ASM_ERR_ARG;
    HIPC(hipSetDevice(ctx->device));
    HIPC(resolve_timers(ctx));
    for (int i = 0; i < T_COUNT; ++i) {
        ctx->timer_ms[i] = 0.0;
        ctx->timer_n[i] = 0;
    }
    return ASM_OK;
}

int asm_get_search_stats(asm_context* ctx, uint64_t out[4])
{
    if (!ctx || !out) return ASM_ERR_ARG;
    SETTLE(ctx);
    HIPC(hipSetDevice(ctx->device));
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!ctx->have_search || !ctx->kstats.p) return ASM_OK;
    std::vector<KnnStats> st(kStatSlots);
    HIPC(hipMemcpyAsync(st.data(), ctx->kstats.p, kStatSlots * sizeof(KnnStats), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    for (const KnnStats& x : st) {
        out[0] += x.node_tests;
        out[1] += x.leaf_scans;
        out[2] += x.dist_evals;
        out[3] += x.inserts;
    }
    return ASM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of independent node shares (one context per share), recorded once
// into a HIP graph: every share's tree build, search, depositions, FP64 mesh
// totals and normalisation on its own stream, the streams forked from and
// joined to one launch stream, so that small shares run side by side.

#define HIPC2(b, expr)                                                                        \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return batch_fail((b), ASM_ERR_HIP, "%s", hipGetErrorString(_e)); \
    } while (0)

int batch_fail(struct asm_batch* b, int code, const char* fmt, ...);

struct asm_batch {
    int device = 0;
    hipStream_t main = nullptr;
    std::string err;
    struct Cube {
        asm_context* ctx = nullptr;
        float left = 0.0f, right = 0.0f;
        int sign = -1;
        std::vector<int> ncs;
        int64_t n = 0, np_local = 0;
    };
    std::vector<Cube> cubes;
    DevBuf<double> totals;          // per (cube, nc): sum n, sum halo, sum n^2, sum c (raw)
    DevBuf<uint32_t> overflow;      // per cube: tile lists overflowed in the last run
    std::vector<size_t> total_off;  // first totals slot of each cube
    int order = 0;                  // ASM_OPT_BATCH_ORDER
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int batch_fail(asm_batch* b, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

namespace {

// One share's work on its context stream; totals at `tot` (4 per nc).
int run_cube(asm_batch::Cube& c, double* tot)
{
    asm_context* ctx = c.ctx;
    if (ctx->overflow_flag) HIPC(launch_fill_zero_u32(ctx->overflow_flag, 1, ctx->stream));
    int rc = asm_tree_construct(ctx, c.left, c.right, nullptr);
    if (!rc) rc = asm_neighbor_search(ctx, kSlots - 1);
    if (!rc) rc = deposit_meshes(ctx, c.ncs.data(), static_cast<int>(c.ncs.size()), c.sign);
    for (size_t i = 0; i < c.ncs.size() && !rc; ++i) {
        const int nc = c.ncs[i];
        MeshDev& m = ctx->meshes[nc];
        HIPC(ctx->red_partial.ensure(4 * 1024));
        HIPC(launch_mesh_reduce(m.mesh.p, nc, ctx->red_partial.p, tot + 4 * i, ctx->stream));
        // coarse_mesh.cpp:151: vol = pow(boxsize/nc, 3) in FP32 (as asm_normalize_mesh)
        const float vol = std::pow(ctx->boxsize / static_cast<float>(nc), 3.0f);
        HIPC(launch_normalize(m.mesh.p, nc, vol, ctx->stream));
    }
    return rc;
}

// After the batch's work has finished: ASM_ERR_CAPACITY when a share's
// tile lists overflowed in the last launch (its meshes are then incomplete).
int batch_check_overflow(asm_batch* b)
{
    if (!b->overflow.p || b->cubes.empty()) return ASM_OK;
    std::vector<uint32_t> f(b->cubes.size());
    HIPC2(b, hipMemcpy(f.data(), b->overflow.p, sizeof(uint32_t) * f.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < f.size(); ++i)
        if (f[i])
            return batch_fail(b, ASM_ERR_CAPACITY,
                              "share %zu: the inputs need more tile-list entries than the graph was recorded with; "
                              "its meshes are incomplete (capture again after changing the inputs this way)",
                              i);
    return ASM_OK;
}

int batch_release_graph(asm_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return ASM_OK;
}

}  // namespace

extern "C" {

int asm_batch_create(asm_batch** out, int device)
{
    if (!out) return ASM_ERR_ARG;
    *out = nullptr;
    asm_batch* b = new asm_batch();
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return ASM_ERR_HIP;
    }
    *out = b;
    return ASM_OK;
}

void asm_batch_destroy(asm_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    batch_release_graph(b);
    for (auto& c : b->cubes) asm_destroy(c.ctx);
    b->totals.release();
    if (b->main) (void)hipStreamDestroy(b->main);
    delete b;
}

const char* asm_batch_last_error(const asm_batch* b)
{
    return b ? b->err.c_str() : "null batch";
}

int asm_batch_add(asm_batch* b, const asm_particle* particles, int64_t n, int64_t np_local, float boxsize,
                  float left, float right, const int* ncs, int n_nc, int sign)
{
    if (!b || !particles || !ncs || n_nc <= 0) return batch_fail(b, ASM_ERR_ARG, "null argument");
    batch_release_graph(b);
    asm_batch::Cube c;
    int rc = asm_create(&c.ctx, b->device);
    if (rc) return batch_fail(b, rc, "cannot create a context for share %zu", b->cubes.size());
    rc = asm_set_particles(c.ctx, particles, n, np_local, boxsize);
    if (rc) {
        b->err = asm_last_error(c.ctx);
        asm_destroy(c.ctx);
        return rc;
    }
    c.left = left;
    c.right = right;
    c.sign = sign;
    c.ncs.assign(ncs, ncs + n_nc);
    c.n = n;
    c.np_local = np_local;
    b->cubes.push_back(c);
    return static_cast<int>(b->cubes.size() - 1);
}

int asm_batch_set_option(asm_batch* b, int option, int value)
{
    if (!b) return ASM_ERR_ARG;
    batch_release_graph(b);
    if (option == ASM_OPT_BATCH_ORDER) {
        if (value != 0 && value != 1) return batch_fail(b, ASM_ERR_ARG, "batch order %d is not 0 or 1", value);
        b->order = value;
        return ASM_OK;
    }
    for (auto& c : b->cubes) {
        const int rc = asm_set_option(c.ctx, option, value);
        if (rc) {
            b->err = asm_last_error(c.ctx);
            return rc;
        }
    }
    return ASM_OK;
}

// Autotune on the share with the most particles, applied to every share.
int asm_batch_autotune(asm_batch* b, const int* leaf_sizes, int n_leaf, int* leaf_out, int* tiles_out)
{
    if (!b || b->cubes.empty()) return batch_fail(b, ASM_ERR_STATE, "empty batch");
    batch_release_graph(b);
    size_t big = 0;
    for (size_t i = 1; i < b->cubes.size(); ++i)
        if (b->cubes[i].n > b->cubes[big].n) big = i;
    asm_batch::Cube& c = b->cubes[big];
    int rc = asm_tree_construct(c.ctx, c.left, c.right, nullptr);
    if (!rc) rc = asm_neighbor_search(c.ctx, kSlots - 1);
    std::vector<int> tiles(c.ncs.size());
    int leaf = 0;
    if (!rc) rc = asm_autotune(c.ctx, leaf_sizes, n_leaf, c.ncs.data(), static_cast<int>(c.ncs.size()), c.sign,
                               &leaf, tiles.data());
    if (rc) {
        b->err = asm_last_error(c.ctx);
        return rc;
    }
    for (auto& o : b->cubes) {
        if (o.ctx->leaf_size != leaf) {
            o.ctx->leaf_size = leaf;
            o.ctx->have_tree = o.ctx->have_search = false;
        }
        for (size_t i = 0; i < c.ncs.size(); ++i) o.ctx->tile_for_nc[c.ncs[i]] = tiles[i];
    }
    if (leaf_out) *leaf_out = leaf;
    if (tiles_out)
        for (size_t i = 0; i < tiles.size(); ++i) tiles_out[i] = tiles[i];
    return ASM_OK;
}

// Runs every share once eagerly (sizing all buffers), then records the same
// work into a graph.  n_streams: how many streams the shares are spread over.
int asm_batch_capture(asm_batch* b, int n_streams)
{
    if (!b || b->cubes.empty()) return batch_fail(b, ASM_ERR_STATE, "empty batch");
    if (n_streams < 1) return batch_fail(b, ASM_ERR_ARG, "n_streams %d < 1", n_streams);
    HIPC2(b, hipSetDevice(b->device));
    batch_release_graph(b);
    b->total_off.clear();
    size_t slots = 0;
    for (auto& c : b->cubes) {
        b->total_off.push_back(slots);
        slots += 4 * c.ncs.size();
    }
    HIPC2(b, b->totals.ensure(slots));
    HIPC2(b, b->overflow.ensure(b->cubes.size()));
    for (size_t i = 0; i < b->cubes.size(); ++i) b->cubes[i].ctx->overflow_flag = b->overflow.p + i;
    // Eager run: every buffer reaches its size.
    for (size_t i = 0; i < b->cubes.size(); ++i) {
        asm_batch::Cube& c = b->cubes[i];
        c.ctx->profile = 0;
        const int rc = run_cube(c, b->totals.p + b->total_off[i]);
        if (rc) {
            b->err = asm_last_error(c.ctx);
            return rc;
        }
    }
    for (auto& c : b->cubes) HIPC2(b, hipStreamSynchronize(c.ctx->stream));
    // Capture: fork from main to the share streams, join back.
    const int ns = std::min<int>(n_streams, static_cast<int>(b->cubes.size()));
    std::vector<hipEvent_t> ev(ns + 1, nullptr);
    const auto destroy_events = [&ev]() {
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
    };
    for (auto& x : ev) {
        const hipError_t ce = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (ce != hipSuccess) {
            destroy_events();
            return batch_fail(b, ASM_ERR_HIP, "%s", hipGetErrorString(ce));
        }
    }
    int rc = ASM_OK;
    const hipError_t cb = hipStreamBeginCapture(b->main, hipStreamCaptureModeRelaxed);
    if (cb != hipSuccess) {
        destroy_events();
        return batch_fail(b, ASM_ERR_HIP, "%s", hipGetErrorString(cb));
    }
    hipError_t e = hipEventRecord(ev[ns], b->main);
    // Which stream runs which share, and in what order: order 0 puts share i
    // on stream i % ns in the order added; order 1 takes the shares largest
    // first, each onto the stream with the fewest particles so far.
    const size_t nc = b->cubes.size();
    std::vector<size_t> order(nc);
    std::vector<int> stream_of(nc);
    for (size_t i = 0; i < nc; ++i) order[i] = i;
    if (b->order == 1) {
        std::stable_sort(order.begin(), order.end(),
                         [b](size_t x, size_t y) { return b->cubes[x].n > b->cubes[y].n; });
        std::vector<int64_t> load(ns, 0);
        for (size_t i : order) {
            int s = 0;
            for (int t = 1; t < ns; ++t)
                if (load[t] < load[s]) s = t;
            stream_of[i] = s;
            load[s] += b->cubes[i].n;
        }
    } else {
        for (size_t i = 0; i < nc; ++i) stream_of[i] = static_cast<int>(i % ns);
    }
    std::vector<hipStream_t> lane_stream(ns);
    for (int s = 0; s < ns && e == hipSuccess; ++s) {
        lane_stream[s] = b->cubes[s].ctx->stream;
        e = hipStreamWaitEvent(lane_stream[s], ev[ns], 0);
    }
    for (size_t k = 0; k < nc && e == hipSuccess && rc == ASM_OK; ++k) {
        const size_t i = order[k];
        asm_batch::Cube& c = b->cubes[i];
        hipStream_t own = c.ctx->stream;
        c.ctx->stream = lane_stream[stream_of[i]];
        c.ctx->capturing = true;
        rc = run_cube(c, b->totals.p + b->total_off[i]);
        c.ctx->capturing = false;
        c.ctx->stream = own;
        if (rc) b->err = asm_last_error(c.ctx);
    }
    for (int s = 0; s < ns && e == hipSuccess; ++s) {
        e = hipEventRecord(ev[s], lane_stream[s]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->main, ev[s], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->main, &g);
    destroy_events();
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return batch_fail(b, ASM_ERR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    HIPC2(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return ASM_OK;
}

// The same work without the graph: every share in order, each on its own
// context stream, waiting for each before the next (the one-share-at-a-time
// baseline of the batched configuration).
int asm_batch_run_sequential(asm_batch* b)
{
    if (!b || b->cubes.empty() || b->total_off.size() != b->cubes.size())
        return batch_fail(b, ASM_ERR_STATE, "asm_batch_run_sequential before asm_batch_capture");
    HIPC2(b, hipSetDevice(b->device));
    for (size_t i = 0; i < b->cubes.size(); ++i) {
        asm_batch::Cube& c = b->cubes[i];
        const int rc = run_cube(c, b->totals.p + b->total_off[i]);
        if (rc) {
            b->err = asm_last_error(c.ctx);
            return rc;
        }
        HIPC2(b, hipStreamSynchronize(c.ctx->stream));
    }
    return ASM_OK;
}

int asm_batch_launch(asm_batch* b)
{
    if (!b || !b->exec) return batch_fail(b, ASM_ERR_STATE, "asm_batch_launch before asm_batch_capture");
    HIPC2(b, hipSetDevice(b->device));
    HIPC2(b, hipGraphLaunch(b->exec, b->main));
    return ASM_OK;
}

int asm_batch_synchronize(asm_batch* b)
{
    if (!b) return ASM_ERR_ARG;
    HIPC2(b, hipSetDevice(b->device));
    HIPC2(b, hipStreamSynchronize(b->main));
    return batch_check_overflow(b);
}

// Share i's particles replaced in place: same count, so its buffers and the
// recorded graph stay valid and the next launch smooths the new records.
int asm_batch_set_particles(asm_batch* b, int i, const asm_particle* particles, int64_t n, int64_t np_local)
{
    if (!b || !particles || i < 0 || i >= static_cast<int>(b->cubes.size()))
        return batch_fail(b, ASM_ERR_ARG, "no share %d", i);
    asm_batch::Cube& c = b->cubes[i];
    if (n != c.n) return batch_fail(b, ASM_ERR_ARG, "share %d holds %lld particles, not %lld", i,
                                    static_cast<long long>(c.n), static_cast<long long>(n));
    HIPC2(b, hipSetDevice(b->device));
    HIPC2(b, hipStreamSynchronize(b->main));   // no replay may be reading the old records
    const int rc = asm_set_particles(c.ctx, particles, n, np_local, c.ctx->boxsize);
    if (rc) {
        b->err = asm_last_error(c.ctx);
        return rc;
    }
    c.np_local = np_local;
    return ASM_OK;
}

int asm_batch_size(const asm_batch* b)
{
    return b ? static_cast<int>(b->cubes.size()) : 0;
}

// Normalised mesh (6 nc^3 floats) and raw FP64 totals {sum n, sum halo,
// sum n^2, sum c} of share i after the last launch.
int asm_batch_get_mesh(asm_batch* b, int i, int nc, float* mesh, double totals[4])
{
    if (!b || i < 0 || i >= static_cast<int>(b->cubes.size())) return batch_fail(b, ASM_ERR_ARG, "no share %d", i);
    asm_batch::Cube& c = b->cubes[i];
    size_t k = 0;
    while (k < c.ncs.size() && c.ncs[k] != nc) ++k;
    if (k == c.ncs.size()) return batch_fail(b, ASM_ERR_ARG, "share %d has no mesh nc %d", i, nc);
    HIPC2(b, hipSetDevice(b->device));
    HIPC2(b, hipStreamSynchronize(b->main));
    const int orc = batch_check_overflow(b);
    if (orc) return orc;
    if (mesh) {
        const int rc = asm_get_mesh(c.ctx, nc, mesh);
        if (rc) {
            b->err = asm_last_error(c.ctx);
            return rc;
        }
    }
    if (totals && !b->total_off.empty())
        HIPC2(b, hipMemcpy(totals, b->totals.p + b->total_off[i] + 4 * k, sizeof(double) * 4, hipMemcpyDeviceToHost));
    return ASM_OK;
}

// h and density of share i (caller order) after the last launch.
int asm_batch_get_rk_dens(asm_batch* b, int i, float* rk, float* dens)
{
    if (!b || i < 0 || i >= static_cast<int>(b->cubes.size())) return batch_fail(b, ASM_ERR_ARG, "no share %d", i);
    HIPC2(b, hipSetDevice(b->device));
    HIPC2(b, hipStreamSynchronize(b->main));
    const int rc = asm_get_rk_dens(b->cubes[i].ctx, rk, dens, nullptr);
    if (rc) b->err = asm_last_error(b->cubes[i].ctx);
    return rc;
}

}  // extern "C"
