/*
 * CPU oracle for the asmooth adaptive-smoothing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library never links or calls it.
 *
 * Every function restates the reference algorithm with exact FP32 arithmetic
 * (no contraction, correctly rounded div/sqrt, the reference's operation order)
 * and FP64 reductions where the reference's own summation order is undefined.
 * Reference citations are relative to the asmooth source tree.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_buffered_set    - the compiled reference's exchange_buffer on one rank
 *                         (tests/golden/reference_mpi.npz), order included;
 *                         closed form (1+2bf)^3 ratio.
 *   orc_knn             - pinned by orc_knn_brute (O(N^2)), by scipy cKDTree
 *                         (library kNN, FP64 metric) and by the compiled
 *                         reference (oracle/_ref) neighbour lists and h.
 *   orc_density         - closed forms (W(0)=hinv^3, self-term), reference dens
 *                         from the compiled reference run single-threaded.
 *   orc_assign_on_mesh  - partition of unity, single-particle product of 1-D
 *                         integrals, numerical quadrature of the tent, mass
 *                         conservation, the compiled reference mesh.
 *   orc_assign_on_global_mesh - the compiled reference's assign_on_global_mesh
 *                         at every node coordinate of a 2^3 decomposition
 *                         (tests/golden/reference_mpi.npz); a wrapped
 *                         single-particle outer product; equality with
 *                         orc_assign_on_mesh for particles whose support lies
 *                         inside the cube.
 *   orc_mark_halo       - compiled reference mark_halo_particles.
 *   orc_spline          - the compiled reference's spline_ker (nbr_search.cpp
 *                         compiled in place), bit for bit; W(0) = 8/(pi h^3),
 *                         continuity at r = 1/2, unit integral by quadrature.
 *   orc_spline_integ    - the integral of its 1-D B-spline: scipy quadrature,
 *                         partition of unity, symmetry.
 *   orc_normalize_*     - algebraic identities; compiled reference writer.
 */
#ifndef ASMOOTH_ORACLE_H
#define ASMOOTH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Single periodic rank buffered set (exchange_buffer.cpp:13-167 with one rank,
 * then remove_extra_particles, clumping_tree_main.cpp:550-576).
 * xv: np x {x,y,z,vx,vy,vz}.  out: cap x 6.  Returns the buffered count or -1
 * if cap is too small.  *np_local_out receives the local count after removal. */
int64_t orc_buffered_set(const float* xv, int64_t np, float boxsize, float bf,
                         float* out, int64_t cap, int64_t* np_local_out);

/* Exact k nearest (canonical order: ascending (d2, index)), k_slots entries per
 * query including the query itself.  pos: n x 3.  queries: nq indices or NULL
 * for all.  nbr/d2: nq x k_slots; rk: nq (sqrtf of the last slot's d2). */
int orc_knn(const float* pos, int64_t n, const int64_t* queries, int64_t nq,
            int k_slots, int32_t* nbr, float* d2, float* rk, int nthreads);
int orc_knn_brute(const float* pos, int64_t n, const int64_t* queries, int64_t nq,
                  int k_slots, int32_t* nbr, float* d2, float* rk, int nthreads);

/* Squared distance (nbr_search.cpp:19-25) and tent kernel (nbr_search.cpp:27-37). */
float orc_norm2(const float* x, const float* y);
float orc_tent(const float* x, const float* y, float hinv);

/* Density scatter (nbr_search.cpp:73-84): for every particle p, slots
 * 0..k_slots-2 of its list get the FP32 tent term, accumulated in FP64. */
void orc_density(const float* pos, int64_t n, const int32_t* nbr, int k_slots,
                 const float* rk, double* dens64);

/* Cell integral of the unit tent (coarse_mesh.cpp:10-22). */
float orc_ker_integ(float x1, float x2);
/* The cubic spline option: the density kernel of nbr_search.cpp:39-53, the
 * cell integral of the separable cubic B-spline, and the selection (0 tent,
 * 1 spline) used by orc_density and the two depositions. */
float orc_spline(const float* x, const float* y, float hinv);
float orc_spline_integ(float x1, float x2);
void orc_set_kernel(int kernel);

/* Mesh deposition (coarse_mesh.cpp:45-127), FP32 terms, FP64 sums.
 * mesh64: 6*nc^3 = n | halo | c | v(xyz interleaved). Accumulates (caller zeroes). */
int orc_assign_on_mesh(const float* pos, const float* vel, const float* dens,
                       const float* rk, const int32_t* igrp, int64_t n,
                       float boxsize, int nc, int sign, double* mesh64);

/* Periodic node-mesh variant (global_mesh.cpp:25-123). */
int orc_assign_on_global_mesh(const float* pos, const float* vel, const float* dens,
                              const float* rk, const int32_t* igrp, int64_t n,
                              float boxsize, int nc, int sign, const int* coord,
                              double* mesh64);

/* FP64 totals of n and halo (coarse_mesh.cpp:35-43). */
void orc_mesh_total_f32(const float* mesh, int nc, double* out2);
void orc_mesh_total_f64(const double* mesh, int nc, double* out2);

/* Normalisation (coarse_mesh.cpp:150-171): FP32 as the reference, and FP64. */
void orc_normalize_f32(float* mesh, int nc, float boxsize);
void orc_normalize_f64(double* mesh, int nc, double boxsize);

/* Halo excision (nbr_search.cpp:140-165; loop clumping_tree_main.cpp:465-485).
 * halos: nh x {x,y,z,r,mass}. igrp is updated in place. */
void orc_mark_halo(const float* pos, int64_t n, const float* halos, int64_t nh,
                   float left, float right, int32_t* igrp, int64_t* np_halo,
                   double* m_halo);

#ifdef __cplusplus
}
#endif
#endif
