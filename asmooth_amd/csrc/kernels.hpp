// Host-side launchers of the gfx950 kernels (one translation unit per stage).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace asm_hip {

// ---- sort.hip --------------------------------------------------------------
// In-place exclusive scan of m 32-bit counts; *total (if given) receives
// their sum in 64 bits, so a caller can tell when the 32-bit offsets wrapped.
// With scratch (scan_scratch_entries(m) 64-bit words) long arrays are
// scanned by many workgroups; without it, by one.
int64_t scan_scratch_entries(int64_t m);
hipError_t launch_scan_exclusive(uint32_t* a, int64_t m, uint64_t* total, uint64_t* scratch, hipStream_t s);
int64_t radix_hist_entries(int64_t n);
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                             int64_t n, int bits, uint32_t* hist, uint64_t* scan_scratch, hipStream_t s,
                             bool* result_in_tmp);

// ---- tree.hip --------------------------------------------------------------
// The caller's particle record, byte-compatible with the reference's
// `Particle` (basic_types.h:10-15): x[3], v[3], dens, rk, igrp.
struct ParticleRec {
    float x[3];
    float v[3];
    float dens;
    float rk;
    int igrp;
};
static_assert(sizeof(ParticleRec) == 36, "Particle record must be 36 bytes");

hipError_t launch_morton_keys(const ParticleRec* p, int64_t n, float left, float right,
                              uint32_t* keys, uint32_t* vals, hipStream_t s);
// spos = (x, y, z, original index bits); sattr = (vx, vy, vz, igrp bits);
// inv_perm[orig] = sorted position.
hipError_t launch_gather_sorted(const ParticleRec* p, const uint32_t* perm, int64_t n,
                                float4* spos, float4* sattr, int32_t* inv_perm, hipStream_t s);
// nodes: 2*nleaf_pad boxes (implicit binary tree, leaves at nleaf_pad..);
// groups: ceil(n / kGroup) boxes of kGroup consecutive sorted particles
// (lo_w = first particle, hi_w = count), or null.
hipError_t launch_build_tree(const float4* spos, int64_t n, int leaf_size, int64_t nleaf,
                             int64_t nleaf_pad, Box* nodes, Box* groups, hipStream_t s);

// ---- knn.hip ---------------------------------------------------------------
// Smoothing kernels (ASM_OPT_KERNEL): the reference's separable tent
// (nbr_search.cpp:27-37, coarse_mesh.cpp:10-22) or the cubic spline
// (nbr_search.cpp:39-53 for the density; its separable form on the mesh).
constexpr int kKernTent = 0;
constexpr int kKernSpline = 1;
struct KnnStats {
    unsigned long long node_tests;
    unsigned long long leaf_scans;
    unsigned long long dist_evals;
    unsigned long long inserts;
};
// Morton keys of the tree build: kMortonAxisBits bits per axis of the tree
// cube, interleaved into kMortonBits.  Nine bits (cells of 1/512 of the cube)
// order the particles as well for the walks as ten and save a radix pass.
constexpr int kMortonAxisBits = 9;
constexpr int kMortonBits = 3 * kMortonAxisBits;

// Search counters are spread over kStatSlots copies (wavefront w adds to
// copy w % kStatSlots), so the per-wavefront atomics do not serialize on one
// address; the host sums the copies.
constexpr int kStatSlots = 1024;
// kNN: rk = sqrtf(d2 of the 33rd key) and that key, per sorted particle.
// With dens64 non-null the search also scatters the density terms (FP64
// atomics).  method kKnnBoundCollect: bound walk + collect walk, with the
// insertion kernel redoing flagged wavefronts (wave_flags: one uint32 per 64
// particles); kKnnInsertion: the insertion kernel for every wavefront.
constexpr int kKnnBoundCollect = 0;
constexpr int kKnnInsertion = 1;
constexpr int kBoundSelect = 0;
constexpr int kBoundPacket = 1;
constexpr int kLeafCap = 256;
// Per-particle neighbour lists of the selection walk (every particle within
// tau: 33 plus ties), kListCap sorted positions per particle; a particle with
// more ties sends its wavefront to the collect walk.
constexpr int kListCap = 40;   // a multiple of 4 (rows read 16 bytes at a time)   // leaf ids recorded per wavefront by the bound walk
// Device work buffers of the bound/collect method: per wavefront of 64
// sorted particles, wleaf kLeafCap ids, wleaf_n and wave_flags one each;
// per sorted particle, tau (33rd squared distance) and dens_pull (density
// terms gathered by the particle itself; the rest arrive in dens64).
struct KnnScratch {
    uint32_t* wleaf;
    uint32_t* wleaf_n;
    uint32_t* wave_flags;
    float* tau;
    double* dens_pull;
    // Learned seed (all null without it): per sorted particle an initial
    // bound on the 33rd squared distance; per wavefront a flag set when a
    // lane's bound held fewer than 33 particles; fallback lanes counted in
    // seed_fallbacks[wave % kStatSlots].
    const float* tau_init;
    uint32_t* seed_redo;
    unsigned long long* seed_fallbacks;
    // Boxes of kGroup consecutive sorted particles (launch_build_tree), used
    // by the bound walk to skip groups of an opened leaf; null when the leaf
    // size is not a multiple of kGroup.
    const Box* groups;
    // With tau_init: 1 runs walk 1 as a range walk within the learned bounds
    // (knn_seed_walk_kernel), 0 as the bounded walk 1.
    int seed_walk;
    // Density kernel of the scatter: kKernTent or kKernSpline.
    int kernel;
    // Walk 1 without a learned bound: kBoundSelect (per-query selection,
    // needs group boxes) or kBoundPacket (the packet walk).
    int bound;
    // With the selection walk and no kept lists: per particle a row of
    // kListCap sorted positions and the list length.  The density is
    // then settled from the lists, and the collect walk runs only for the
    // wavefronts the selection could not list.  Null: the collect walk for
    // every wavefront.
    uint32_t* list;
    uint32_t* list_n;
};
// Optional marks around the kernels of the bound/collect method (stage 0:
// bound walk, 1: collect walk, 2: insertion kernel for flagged wavefronts).
struct KnnMarks {
    void* user;
    void (*begin)(void* user, int stage);
    void (*end)(void* user, int stage);
};
hipError_t launch_knn(const float4* spos, int64_t n, const Box* nodes, int64_t nleaf_pad, int64_t nleaf,
                      int leaf_size, float* rk_sorted, uint64_t* kkey_sorted, const int32_t* inv_perm, double* dens64,
                      int32_t* nbr_out_orig, KnnStats* stats, const KnnScratch& scratch, int method, hipStream_t s,
                      const KnnMarks* marks = nullptr);
// dens = dens64 (+ dens_pull when non-null), rounded to FP32.
hipError_t launch_dens_finalize(const double* dens64, const double* dens_pull, int64_t n, float* dens, hipStream_t s);
// Per-node maximum of the 33rd squared distance (2*nleaf_pad floats).
hipError_t launch_kmax(const uint64_t* kkey, int64_t n, int leaf_size, int64_t nleaf, int64_t nleaf_pad, float* kmax,
                       hipStream_t s);
// Density by gather (sorted order, float).
hipError_t launch_dens_gather(const float4* spos, const uint64_t* kkey, const float* rk, int64_t n, const Box* nodes,
                              const float* kmax, int64_t nleaf_pad, float* dens, int kern, hipStream_t s);

// ---- mlp.hip -----------------------------------------------------------------
// Learned seed radius: 27 inputs (ln CIC density over 3x3x3 cells of a
// kMlpGrid^3 mesh of the tree cube), hidden layers 64-64 (GELU), output
// ln(h / spacing).
constexpr int kMlpIn = 27;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;
constexpr int kMlpGrid = 64;
constexpr float kMlpLnFloor = 1e-3f;   // ln(density / mean + floor)
// grid (kMlpGrid^3 floats) <- ln(CIC density / mean + floor) over [left, right)^3.
hipError_t launch_seed_grid(const float4* spos, int64_t n, float left, float right, float* grid, hipStream_t s);
// x[b] (kMlpIn floats) = features of the particle at sorted position sidx[b].
hipError_t launch_seed_features(const float4* spos, const int32_t* sidx, int64_t m, const float* lngrid, float left,
                                float right, float* x, hipStream_t s);
// Training samples of caller indices idx[0..m): features and ln(h / spacing).
hipError_t launch_seed_samples(const float4* spos, const float* rk_sorted, const int32_t* inv_perm,
                               const int64_t* idx, int64_t m, const float* lngrid, float left, float right,
                               float inv_spacing, int32_t* sidx, float* x, float* target, hipStream_t s);
// rcell[c] = exp(f(x_c)) * scale per mesh cell (kMlpGrid^3 floats), scale =
// spacing * (1 + margin); tau_init[q] = rcell[cell of q]^2.
hipError_t launch_seed_bound(const float4* spos, int64_t n, const float* lngrid, float left, float right,
                             const float* params, float scale, float* rcell, float* tau_init, hipStream_t s);

// ---- deposit.hip -------------------------------------------------------------
struct DepositArgs {
    const float4* spos;
    const float4* sattr;
    const float* rk;
    const float* dens;
    int64_t n;
    float boxsize;
    int nc;
    int sign;
    int tile;        // tile side in cells
    int ntx;         // tiles per axis
    int periodic;    // 0: node mesh (coarse_mesh.cpp), 1: periodic variant (global_mesh.cpp)
    int coord[3];    // node coordinate for the periodic variant
    uint32_t list_cap;   // entries the tile list holds (writes and reads past it are dropped)
    int accum;           // kAccumFp64 (default), kAccumKahan32, kAccumAtomic32 (node mesh only)
    int kern;            // kKernTent (default) or kKernSpline: the cell integral of the kernel
};
constexpr int kAccumFp64 = 0;      // FP64 LDS atomics
constexpr int kAccumKahan32 = 1;   // FP32 with Kahan compensation, one wavefront per tile
constexpr int kAccumAtomic32 = 2;  // plain FP32 LDS atomics
// Every node mesh of an nc list at once (FP64 sums): mesh j's tiles are
// tile_base[j] .. tile_base[j+1]-1 of one concatenated tile space, whose
// counters, starts and list are shared; the particle fields of every a[j]
// are the same.
constexpr int kMaxMeshes = 8;
struct MultiDeposit {
    DepositArgs a[kMaxMeshes];
    float* mesh[kMaxMeshes];
    uint32_t tile_base[kMaxMeshes + 1];
    int n;
};
hipError_t launch_multi_count(const MultiDeposit& J, uint32_t* tile_cnt, hipStream_t s);
hipError_t launch_multi_fill(const MultiDeposit& J, const uint32_t* tile_start, uint32_t* tile_cur, uint32_t* list,
                             hipStream_t s);
hipError_t launch_multi_tiles(const MultiDeposit& J, const uint32_t* tile_start, const uint32_t* tile_cnt,
                              const uint32_t* list, hipStream_t s);
hipError_t launch_deposit_count(const DepositArgs& a, uint32_t* tile_cnt, hipStream_t s);
hipError_t launch_deposit_fill(const DepositArgs& a, const uint32_t* tile_start, uint32_t* tile_cur,
                               uint32_t* list, hipStream_t s);
hipError_t launch_deposit_tiles(const DepositArgs& a, const uint32_t* tile_start, const uint32_t* tile_cnt,
                                const uint32_t* list, float* mesh, int accumulate, hipStream_t s);
// *flag = 1 when *total > cap (both on the device).
hipError_t launch_flag_overflow(const uint64_t* total, uint64_t cap, uint32_t* flag, hipStream_t s);
hipError_t launch_fill_zero_u32(uint32_t* p, int64_t n, hipStream_t s);
hipError_t launch_fill_zero_f32(float* p, int64_t n, hipStream_t s);
hipError_t launch_fill_zero_f64(double* p, int64_t n, hipStream_t s);
// FP64 reductions of a mesh: out[0] = sum n, out[1] = sum halo, out[2] = sum n^2,
// out[3] = sum c (raw).  partial must hold 4 * 1024 doubles.
hipError_t launch_mesh_reduce(const float* mesh, int nc, double* partial, double* out, hipStream_t s);
hipError_t launch_normalize(float* mesh, int nc, float vol, hipStream_t s);
// A node's normalised mesh (node coord of nn^3) into a global array of side
// nc * nn: field 0 n, 1 halo, 2 c, 3 vel, 4 h + n, 5 vel * (n / 8).
hipError_t launch_stitch(const float* mesh, int nc, int nn, const int coord[3], int field, float* dst, hipStream_t s);

// ---- halo.hip ----------------------------------------------------------------
struct HaloRec {
    float x[3];
    float r;
    float mass;
};
hipError_t launch_mark_halo(const float4* spos, float4* sattr, int64_t n, const Box* nodes,
                            int64_t nleaf_pad, const HaloRec* halos, int64_t nh, float left,
                            float right, unsigned long long* count, hipStream_t s);

}  // namespace asm_hip
