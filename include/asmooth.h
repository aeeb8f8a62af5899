/*
 * asmooth_amd - C ABI of the MI355X (gfx950) adaptive-smoothing path.
 *
 * The reference (junkoda/asmooth) has no plugin or FFI layer: its hot path is a
 * fixed sequence of C++ calls made by for_each_redshift
 * (clumping_tree_main.cpp:390-533).  Each entry point below replaces one of
 * those calls, cited next to it; the caller (a C++ driver, or Python through
 * ctypes) makes them in the same order:
 *
 *   asm_set_particles         Particles after exchange_buffer + remove_extra
 *   asm_tree_construct        KDTree::construct
 *   asm_neighbor_search       the neighbor_search loop
 *   asm_mark_halo_particles   the mark_halo_particles loop (HALO_EXCISE)
 *   asm_assign_on_mesh        clear_mesh + assign_on_mesh, per nc
 *   asm_mesh_total            mesh_total (n and halo)
 *   asm_normalize_mesh        the normalisation inside write_mesh_separate
 *   asm_get_mesh / asm_write_mesh_separate   the mesh files
 *
 * Conventions: every function returns ASM_OK (0) or a negative status and
 * never aborts; asm_last_error() describes the last failure on a context.
 * The caller owns host buffers; the context owns device buffers (no globals),
 * reuses them across calls, and runs everything on one HIP stream of its own.
 * One context per host thread.  Arrays of particles are in the caller's order
 * (the device copy is Morton-sorted internally and never exposed).
 */
#ifndef ASMOOTH_AMD_H
#define ASMOOTH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Particle record, byte-identical to the reference's Particle
 * (basic_types.h:10-15): position, velocity, density, smoothing length, group. */
typedef struct asm_particle {
    float x[3];
    float v[3];
    float dens;
    float rk;
    int32_t igrp;
} asm_particle;

/* Halo record, byte-identical to the reference's Halo (basic_types.h:17-20). */
typedef struct asm_halo {
    float x[3];
    float r;
    float mass;
} asm_halo;

typedef struct asm_context asm_context;

enum {
    ASM_OK = 0,
    ASM_ERR_ARG = -1,      /* invalid argument */
    ASM_ERR_HIP = -2,      /* HIP runtime failure (message holds the HIP error) */
    ASM_ERR_STATE = -3,    /* call out of order (e.g. search before construct) */
    ASM_ERR_CAPACITY = -4, /* size beyond what the library supports */
    ASM_ERR_IO = -5        /* file could not be read or written */
};

/* Options for asm_set_option. */
enum {
    ASM_OPT_KEEP_NEIGHBORS = 1, /* 1: asm_neighbor_search keeps the 33-slot lists */
    ASM_OPT_PROFILE = 2,        /* 1: time every kernel with HIP events */
    ASM_OPT_LEAF_SIZE = 3,      /* particles per tree leaf, 8..64 (default 32) */
    ASM_OPT_TILE = 4,           /* deposition tile side in cells, 1..8 (default 8) */
    ASM_OPT_DETERMINISTIC = 5,  /* 1: density by an ordered gather, bitwise reproducible
                                      run to run (default 0: FP64 atomic scatter) */
    ASM_OPT_KNN_METHOD = 6,     /* 0: bound walk + collect walk (default); 1: sorted-key
                                      insertion lists for every wavefront.  Same results. */
    ASM_OPT_DEPOSIT_ACCUM = 7,  /* node-mesh cell sums: 0 FP64 (default); 1 FP32 with Kahan
                                      compensation; 2 plain FP32 atomics (precision study) */
    ASM_OPT_SEED_WALK = 8,      /* with a seed model: 2 (default) bounds the selection walk by
                                      the learned bounds instead of its seed-range histogram,
                                      settling the queries whose bound holds fewer than 33
                                      particles in a second pass with the histogram bound; 0
                                      starts the packet walk from the learned bounds; 1 searches
                                      within them by a range walk (studies: slower on the 256^3
                                      workload).  Results are the same. */
    ASM_OPT_KERNEL = 9,         /* smoothing kernel: 0 (default) the reference's separable tent
                                      (ker, nbr_search.cpp:27-37; coarse_mesh.cpp:10-22); 1 the
                                      cubic spline: spline_ker (nbr_search.cpp:39-53, which the
                                      reference defines but does not call) for the density, and
                                      the separable cubic B-spline of half-width h, integrated
                                      exactly over each cell, on the mesh.  Setting it requires
                                      a new asm_neighbor_search. */
    ASM_OPT_KNN_BOUND = 10,     /* first walk of ASM_OPT_KNN_METHOD 0 (the 33rd distance of
                                      every particle): 0 (default) selection, one query at a
                                      time over its wavefront with a bitonic sort of the
                                      candidates within a seed-range bound; 1 the packet walk
                                      (64 queries share every candidate, med3 insertion).
                                      Same results.  With a seed model the packet walk runs. */
    ASM_OPT_BATCH_ORDER = 11    /* asm_batch_set_option only: 0 (default) share i on stream
                                      i % n_streams in the order added; 1 largest share first,
                                      each onto the stream with the fewest particles so far */
};

const char* asm_version(void);

/* Context on HIP device `device` (ordinal). */
int asm_create(asm_context** ctx, int device);
void asm_destroy(asm_context* ctx);
const char* asm_last_error(const asm_context* ctx);
int asm_set_option(asm_context* ctx, int option, int value);
int asm_synchronize(asm_context* ctx);

/* The buffered particle set of one node (Particles, basic_types.h:26-37, as
 * it stands after exchange_buffer and remove_extra_particles,
 * clumping_tree_main.cpp:376-385): np_with_buffers records whose first
 * np_local are the node's own.  Copied to the device; dens is reset to 0 on
 * the device (the reference relies on the reader, file_reader2.cpp:80). */
int asm_set_particles(asm_context* ctx, const asm_particle* particles, int64_t np_with_buffers,
                      int64_t np_local, float boxsize);
/* Same, from records already in device memory on the context's device. */
int asm_set_particles_device(asm_context* ctx, const asm_particle* d_particles,
                             int64_t np_with_buffers, int64_t np_local, float boxsize);

/* KDTree::construct (kdtree_balanced.h:18; called at clumping_tree_main.cpp:392-394
 * with left = -bf*L, right = (1+bf)*L).  Builds the device tree; *nnode
 * receives the number of tree nodes. */
int asm_tree_construct(asm_context* ctx, float left, float right, int64_t* nnode);

/* The neighbour loop (clumping_tree_main.cpp:418-430 calling neighbor_search,
 * nbr_search.h:9): rk = distance to the k-th nearest other particle and the
 * tent-kernel density scatter to the k nearest (self included).  k must be 32
 * (KNeighbors::k = 33 slots, k_neighbors.h:47).  Ties are broken by index. */
int asm_neighbor_search(asm_context* ctx, int k);

/* mark_halo_particles over a halo catalogue (nbr_search.h:11; loop
 * clumping_tree_main.cpp:465-485): igrp = -1 for particles with d2 < r*r of a
 * halo whose centre lies strictly inside (left, right)^3.  *np_halo = number
 * of (halo, particle) marks, *m_halo = sum of mass/8 over those halos. */
int asm_mark_halo_particles(asm_context* ctx, const asm_halo* halos, int64_t n_halo, float left,
                            float right, int64_t* np_halo, double* m_halo);

/* clear_mesh + assign_on_mesh (coarse_mesh.h:4,6; clumping_tree_main.cpp:525-526)
 * into the context's device mesh for `nc` (6*nc^3 floats: n | halo | c | v
 * with v xyz-interleaved).  sign as in the reference (-1 for the SO pass).
 * The first deposition of a context reads its tile-list size back to size
 * the list buffer (with an eighth to spare); later ones (this call, the
 * list and the periodic variant) return without waiting on the device: the
 * next call that reads a mesh or changes the context's inputs or options
 * redoes, before anything else, any deposition whose list outgrew the
 * buffer, so results never depend on it. */
int asm_assign_on_mesh(asm_context* ctx, int nc, int sign);
/* Every nc of a -nc list in one call (clumping_tree_main.cpp:521-533). */
int asm_assign_on_meshes(asm_context* ctx, const int* ncs, int n_nc, int sign);
/* Periodic node-mesh variant (assign_on_global_mesh, global_mesh.cpp:25-123):
 * local particles only, node offset coord*L, cell indices wrapped. */
int asm_assign_on_global_mesh(asm_context* ctx, int nc, int sign, const int coord[3]);

/* mesh_total (coarse_mesh.h:5) of n and of halo, in FP64 (check_total_mass,
 * clumping_tree_main.cpp:578-600). */
int asm_mesh_total(asm_context* ctx, int nc, double totals[2]);
/* FP64 reductions of the raw mesh: {sum n, sum halo, sum n^2, sum c}. */
int asm_mesh_stats(asm_context* ctx, int nc, double out[4]);
/* The normalisation of write_mesh_separate (coarse_mesh.cpp:150-171), in place. */
int asm_normalize_mesh(asm_context* ctx, int nc);
/* Copy the mesh for nc (6*nc^3 floats) to host memory. */
int asm_get_mesh(asm_context* ctx, int nc, float* mesh);
/* Device address of the mesh for nc (valid until the context is destroyed). */
const float* asm_mesh_device_ptr(asm_context* ctx, int nc);
/* The stitching of node meshes into one global mesh (tools/reformat.cpp:
 * 55-106, tools/reformat2.cpp:82-142, which do it offline from the node files)
 * done on the device: this context's mesh for nc, already normalised
 * (asm_normalize_mesh), as node `coord` of an nn^3 decomposition, is written
 * into dst, a device array of (nc*nn)^3 cells with x fastest, of one field:
 *   ASM_FIELD_N, _HALO, _C (1 float per cell), _VEL (3), and the two fields
 *   reformat2 derives: _NTOT = h + n (1) and _MOMENTUM = vel * (n / 8) (3).
 * Runs on the context's stream (asm_synchronize before reading dst). */
enum {
    ASM_FIELD_N = 0,
    ASM_FIELD_HALO = 1,
    ASM_FIELD_C = 2,
    ASM_FIELD_VEL = 3,
    ASM_FIELD_NTOT = 4,
    ASM_FIELD_MOMENTUM = 5
};
int asm_stitch_mesh(asm_context* ctx, int nc, int nn, const int coord[3], int field, float* dst);

/* Results per particle in the caller's order: rk, dens and igrp are written,
 * x and v are copied back unchanged. */
int asm_get_particles(asm_context* ctx, asm_particle* particles);
int asm_get_rk_dens(asm_context* ctx, float* rk, float* dens, int32_t* igrp);
/* Neighbour lists (np_with_buffers x 33 original indices, ascending (d2, index))
 * kept by the last search when ASM_OPT_KEEP_NEIGHBORS is set. */
int asm_get_neighbors(asm_context* ctx, int32_t* nbr);

/* The device tree as built, for inspection: sorted positions (n x 4 floats:
 * x, y, z, original index as int32 bits) and node records ((2*nleaf_pad) x 8
 * words: lo xyz, first particle, hi xyz, particle count; node 1 is the root,
 * node i has children 2i and 2i+1, leaves are nleaf_pad .. 2*nleaf_pad-1). */
int asm_get_tree(asm_context* ctx, float* spos, int32_t* nodes, int64_t* nleaf, int64_t* nleaf_pad,
                 int32_t* leaf_size);
/* Search results in sorted (tree) order: rk and dens. */
int asm_get_sorted_results(asm_context* ctx, float* rk, float* dens);

/* Host steps either side of the path. */
/* Buffered set of a single periodic rank: the 26 periodic images of
 * exchange_buffer (exchange_buffer.cpp:13-167, one rank) followed by
 * remove_extra_particles (clumping_tree_main.cpp:550-576).  xv: np x 6 floats.
 * Returns the buffered count, or a negative status. */
int64_t asm_build_buffered_set(const float* xv, int64_t np, float boxsize, float bf,
                               asm_particle* out, int64_t cap, int64_t* np_local);
/* Node share of a periodic box cut into nn^3 node cubes: node `coord` gets its
 * own particles and every periodic image within bf*Lnode of its cube, in node
 * coordinates (what the reference holds after the exchange on that rank). */
int64_t asm_build_node_share(const float* xv, int64_t np, float box_total, int nn,
                             const int coord[3], float bf, asm_particle* out, int64_t cap,
                             int64_t* np_local);
/* Periodic images of a single rank's halo catalogue (exchange_buffer_template.h:
 * 18-122 with one rank; halos are not trimmed).  Returns the count. */
int64_t asm_build_buffered_halos(const asm_halo* halos, int64_t n, float boxsize, float bf,
                                 asm_halo* out, int64_t cap);
/* Halo catalogue reader (halo_file2.cpp:88-138): 3 x int32 header, then
 * records of 22 float32 (position at 0..2, mass/radius at 3/5, or 4/6 with
 * m200); positions are taken modulo boxsize minus shift; records whose
 * overdensity differs from Delta (Bryan-Norman at z, omega_m, or 200) by 10
 * or more are an error; pid_flag skips the 50 x 32-byte particle block after
 * each record.  *halos is malloc'ed; free with asm_free. */
int asm_read_halo_file(const char* path, float boxsize, const float shift[3], float z, float omega_m,
                       int m200, int pid_flag, asm_halo** halos, int64_t* n_halo);
/* Bryan & Norman (1998) virial overdensity (halo_file2.cpp:47-61). */
double asm_halo_delta_vir(double z, double omega_m);
/* cubep3m xv reader (file_reader2.cpp:24-86): header = int32 np + 11 float32,
 * then np x {x,y,z,vx,vy,vz}.  *xv is malloc'ed; free with asm_free. */
int asm_read_xv(const char* path, float** xv, int64_t* np, float header[11]);
/* cubep3m zip reader (file_reader2.cpp:90-272): the four files
 * <filebase><inode>/<redshift>zip{0,1,2,3}_<inode>.dat of one node
 * (per-particle cell offsets in 1/256ths, int16 velocities, per-cell counts
 * with 255 escaping to int32) decoded into np x {x,y,z,vx,vy,vz}:
 * x = mesh_scale ((b + 0.5) / 256 + i - 1) - shift, v = s / v_r2i, in FP32.
 * The node cube is mesh_scale * nc_node_dim.  header: the 16 words of zip0's
 * header (np_local, a, t, tau, ..., v_r2i at word 12).  *xv is malloc'ed. */
int asm_read_zip(const char* filebase, const char* redshift, int inode, int nc_node_dim, int mesh_scale,
                 const float shift[3], float** xv, int64_t* np, float header[16]);
int asm_write_xv(const char* path, const float* xv, int64_t np, const float header[11]);
void asm_free(void* p);
/* The four files of write_mesh_separate (coarse_mesh.cpp:173-201) from a mesh
 * that is already normalised: <base>n<inode>.dat, h, c (int32 nc + nc^3
 * floats) and vel (int32 nc + 3 nc^3 floats). */
int asm_write_mesh_separate(const char* base, int inode, const float* mesh, int nc);
/* The offline stitching tools, with status codes instead of aborts.
 * asm_reformat (tools/reformat.cpp:11-123): node files <file_base><inode>.dat
 * (inode = ix + nn iy + nn^2 iz; int32 nc, then nc^3 cells of n_components
 * floats) into <out_dir>/<file_base>_all.dat (3 x int32 nc*nn, then the
 * global cells, x fastest); *total = sum of the first nc floats of each row,
 * as the tool prints it.  asm_reformat2 (tools/reformat2.cpp:10-171): the
 * n, h and vel files <in_dir>/<redshift>{n,h,vel}<inode>.dat into
 * <out_dir>/<redshift>{n,ntot,v}_all.dat holding n, h + n and vel * (n / 8);
 * totals = {sum n / 8, sum (h + n) / 8}. */
int asm_reformat(int nn, const char* out_dir, const char* file_base, int n_components, double* total);
int asm_reformat2(int nn, const char* out_dir, const char* redshift, const char* in_dir, double totals[2]);

/* Autotuner (no counterpart in the reference; the reference has fixed
 * choices: leaf quota 16, kdtree_balanced.cpp:245, and no tiling).  After a
 * search, times the tree build + search for each leaf size in leaf_sizes
 * (n_leaf may be 0 to keep the current one) and the deposition of each nc in
 * ncs for every tile side 4..8, on the context's own particles, and keeps the
 * fastest: later trees use *leaf_out and later depositions of nc use
 * tiles_out[i] (until ASM_OPT_TILE or ASM_OPT_LEAF_SIZE is set).  The tree
 * and search are redone with the chosen leaf size, so the context is ready
 * for asm_assign_on_mesh.  Results do not depend on either choice. */
int asm_autotune(asm_context* ctx, const int* leaf_sizes, int n_leaf, const int* ncs, int n_nc, int sign,
                 int* leaf_out, int* tiles_out);

/* Learned seed radius (no counterpart in the reference).  A 27-64-64-1
 * multilayer perceptron with GELU activations maps ln(CIC density / mean +
 * 1e-3) over the 3x3x3 cells around a particle, on a 64^3 mesh of the tree
 * cube, to ln(h / s), s the mean particle spacing (tree cube volume over
 * particle count, cube root).  Parameters, asm_mlp_param_count() floats:
 * W1[27][64], b1[64], W2[64][64], b2[64], w3[64], b3.  The model is trained
 * on the device (mean squared error, Adam). */
typedef struct asm_mlp asm_mlp;
int asm_mlp_param_count(void);
int asm_mlp_create(asm_mlp** m, int device);
void asm_mlp_destroy(asm_mlp* m);
const char* asm_mlp_last_error(const asm_mlp* m);
int asm_mlp_set_params(asm_mlp* m, const float* params);   /* also resets the optimiser state */
int asm_mlp_get_params(asm_mlp* m, float* params);
/* y[i] = f(x[i]), x: n x 27 floats (host). */
int asm_mlp_forward(asm_mlp* m, const float* x, int64_t n, float* y);
/* Mean squared error of f(x) against t and its gradient (host arrays). */
int asm_mlp_loss_grad(asm_mlp* m, const float* x, const float* t, int64_t n, double* loss, float* grad);
/* Adam (beta 0.9, 0.999) over shuffled batches; epoch_loss[e] = mean loss
 * of epoch e before each step. */
int asm_mlp_train(asm_mlp* m, const float* x, const float* t, int64_t n, int epochs, int batch, float lr,
                  uint64_t seed, float* epoch_loss);
/* Features (m x 27) and targets ln(h / s) of particles idx[0..m) (caller
 * order), after a search: training samples. */
int asm_seed_samples(asm_context* ctx, const int64_t* idx, int64_t m, float* x, float* target);
/* Use the model in later searches: a particle's search starts bounded by the
 * radius (1 + margin) * s * exp(f(x)), and a particle whose radius holds
 * fewer than 33 particles is searched again without it.  Results are those
 * of the plain search, bit for bit.  m = NULL switches it off. */
int asm_set_seed_model(asm_context* ctx, const asm_mlp* m, float margin);
/* Of the last search with a seed model: out[0] particles searched again,
 * out[1] wavefronts (64 sorted particles) searched again, out[2] particles. */
int asm_get_seed_stats(asm_context* ctx, int64_t out[3]);

/* Batches of independent node shares (no counterpart in the reference,
 * whose ranks each run one share).  Each share gets its own context; its tree
 * build, search, the depositions of its nc list, the FP64 mesh totals and
 * the normalisation are recorded once into one HIP graph (shares spread over
 * n_streams streams forked from one launch stream), which is then launched
 * as a whole.  asm_batch_add returns the share's index. */
typedef struct asm_batch asm_batch;
int asm_batch_create(asm_batch** b, int device);
void asm_batch_destroy(asm_batch* b);
const char* asm_batch_last_error(const asm_batch* b);
int asm_batch_add(asm_batch* b, const asm_particle* particles, int64_t n, int64_t np_local, float boxsize,
                  float left, float right, const int* ncs, int n_nc, int sign);
int asm_batch_size(const asm_batch* b);
/* asm_set_option on every share. */
int asm_batch_set_option(asm_batch* b, int option, int value);
/* asm_autotune on the share with the most particles; its choice is used by
 * every share.  tiles_out: one per nc of that share. */
int asm_batch_autotune(asm_batch* b, const int* leaf_sizes, int n_leaf, int* leaf_out, int* tiles_out);
/* Runs every share once (sizing its buffers) and records the graph. */
int asm_batch_capture(asm_batch* b, int n_streams);
int asm_batch_launch(asm_batch* b);
/* Waits for the last launch.  Returns ASM_ERR_CAPACITY when a share's inputs
 * (changed by asm_batch_set_particles after the capture) needed more tile-list
 * entries than the recorded graph holds: that share's meshes are incomplete,
 * and asm_batch_get_mesh fails the same way until a launch fits (capture
 * again to size the buffers for the new inputs). */
int asm_batch_synchronize(asm_batch* b);
/* Share i's particles replaced in place (n must equal the share's count): the
 * graph stays recorded and the next launch smooths the new records. */
int asm_batch_set_particles(asm_batch* b, int i, const asm_particle* particles, int64_t n, int64_t np_local);
/* The same work without the graph, one share after another (a baseline). */
int asm_batch_run_sequential(asm_batch* b);
/* Normalised mesh (6 nc^3 floats) and raw FP64 totals {sum n, sum halo,
 * sum n^2, sum c} of share i after the last launch (either may be NULL). */
int asm_batch_get_mesh(asm_batch* b, int i, int nc, float* mesh, double totals[4]);
/* h and density of share i, in its particles' order. */
int asm_batch_get_rk_dens(asm_batch* b, int i, float* rk, float* dens);

/* Timers (HIP events on the context stream) when ASM_OPT_PROFILE is set. */
int asm_timer_count(void);
const char* asm_timer_name(int i);
/* Total milliseconds and number of launches of timer i since the last reset. */
int asm_get_timer(asm_context* ctx, int i, double* total_ms, int64_t* launches);
int asm_reset_timers(asm_context* ctx);
/* Work counters of the last neighbour search: node tests, leaf scans, distance
 * evaluations, list insertions. */
int asm_get_search_stats(asm_context* ctx, uint64_t out[4]);

/* No counterpart in the reference: an entry point that allocates nothing, for
 * callers that cannot call hipMalloc (tight serving loops, harnesses that
 * forbid device allocation).  It does the work of asm_set_particles_device +
 * asm_tree_construct(left, right) + asm_neighbor_search(32), and with nc > 0
 * asm_assign_on_mesh(nc, sign), on device `device`, with the same kernels,
 * leaf size 32 and tile 8, but every work buffer lies in one fixed device
 * array of the library (96 MiB per device, held from the library's load until
 * the process ends).  d_particles: n device records (caller order); outputs,
 * in device memory: d_rk and d_dens (n floats, caller order) and, with
 * nc > 0, d_mesh (6 nc^3 floats, the raw sums of asm_assign_on_mesh).
 * Limits: n <= ASM_NOALLOC_MAX_PARTICLES, nc <= ASM_NOALLOC_MAX_NC, and at most
 * ASM_NOALLOC_MAX_LIST (particle, tile) pairs for the mesh; beyond them it
 * returns ASM_ERR_CAPACITY (and, for the pairs, after h and density are
 * written).  err (may be NULL) receives a message of at most err_len bytes.
 * The call runs on the device's null stream and returns when the work is
 * done.  Every call on one device shares that device's array: calls that
 * overlap on one device race. */
#define ASM_NOALLOC_MAX_PARTICLES (1 << 19)
#define ASM_NOALLOC_MAX_NC 128
#define ASM_NOALLOC_MAX_LIST (1 << 23)
int asm_smooth_noalloc(int device, const asm_particle* d_particles, int64_t n, int64_t np_local, float boxsize,
                       float left, float right, int nc, int sign, float* d_rk, float* d_dens, float* d_mesh,
                       char* err, int err_len);

/* Measurement helpers (no counterpart in the reference): a plain device copy
 * of `bytes` (multiple of 16) whose hardware byte counters calibrate those of
 * the smoothing kernels, and raw device allocation for it. */
int asm_calibration_copy(void* dst, const void* src, int64_t bytes);
int asm_device_alloc(void** p, int64_t bytes);
int asm_device_free(void* p);
/* Synchronous copy of `bytes` from device memory to host memory, and back. */
int asm_device_to_host(void* host, const void* dev, int64_t bytes);
int asm_host_to_device(void* dev, const void* host, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
