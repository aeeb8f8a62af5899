// Harness around the reference's own hot-path functions.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (kdtree_balanced.cpp, nbr_search.cpp,
// coarse_mesh.cpp, basic_types.cpp, open.cpp) into oracle/_ref/libasmref.so.
// Tests use it to pin the oracle; bench.py's cpu_baseline leg times it.
// It calls, in the reference's own order, the functions the reference driver
// calls (clumping_tree_main.cpp:390-430, 521-533) and exposes plain C entry
// points so that Python can drive it through ctypes.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "basic_types.h"
#include "kdtree_balanced.h"
#include "k_neighbors.h"
#include "nbr_search.h"
#include "coarse_mesh.h"

// Defined with external linkage in nbr_search.cpp:91-127 (declared there only).
void find_nbr_recursive(Node const* const, const float x[], KNeighbors* const knbrs);

namespace {

struct RefState {
    Particles* particles = nullptr;
    KDTreeSimple* tree = nullptr;
    int n = 0;
    int nnode = 0;
    std::vector<int>* perm = nullptr;  // perm[i] = original index of tree slot i
};

double now_s()
{
    struct timeval tp;
    gettimeofday(&tp, nullptr);
    return (double)tp.tv_sec + (double)tp.tv_usec * 1e-6;
}

void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

}  // namespace

extern "C" {

// Fill a Particles set (basic_types.h:10-37).  The original index is stashed in
// the bits of rk, which the reference does not read before neighbor_search
// writes it (nbr_search.cpp:73), so the tree permutation can be recovered.
void* ref_create(const float* pos, const float* vel, const int32_t* igrp, int n,
                 int np_local, float boxsize)
{
    RefState* s = new RefState();
    s->particles = new Particles();
    s->particles->allocate(n + 1);
    s->particles->np_local = np_local;
    s->particles->np_with_buffers = n;
    s->particles->np_total = np_local;
    s->particles->boxsize = boxsize;
    Particle* p = s->particles->particle;
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            p[i].x[a] = pos[3 * i + a];
            p[i].v[a] = vel ? vel[3 * i + a] : 0.0f;
        }
        p[i].dens = 0.0f;
        int32_t bits = i;
        std::memcpy(&p[i].rk, &bits, sizeof(float));
        p[i].igrp = igrp ? igrp[i] : 0;
    }
    s->n = n;
    s->perm = new std::vector<int>(n);
    return s;
}

void ref_destroy(void* h)
{
    RefState* s = (RefState*)h;
    if (!s) return;
    delete s->tree;
    // Particles::~Particles uses delete on malloc'ed memory; free it here.
    std::free(s->particles->particle);
    s->particles->particle = nullptr;
    delete s->particles;
    delete s->perm;
    delete s;
}

// KDTreeSimple::allocate + construct (clumping_tree_main.cpp:191-199, 392-394).
int ref_construct(void* h, float left, float right, double* seconds)
{
    RefState* s = (RefState*)h;
    s->tree = new KDTreeSimple();
    const int np_max = s->n * 2 + 1024;
    s->tree->allocate(np_max, 16);
    double t0 = now_s();
    s->nnode = s->tree->construct(s->particles, left, right);
    if (seconds) *seconds = now_s() - t0;
    set_nbr_search(s->tree->root(), s->particles->particle);
    Particle* p = s->particles->particle;
    for (int i = 0; i < s->n; ++i) {
        int32_t bits;
        std::memcpy(&bits, &p[i].rk, sizeof(float));
        (*s->perm)[i] = bits;
        p[i].rk = 0.0f;
    }
    return s->nnode;
}

// The neighbour loop of clumping_tree_main.cpp:418-430.
int ref_neighbor_search_all(void* h, int nthreads, double* seconds)
{
    RefState* s = (RefState*)h;
    set_threads(nthreads);
    set_nbr_search(s->tree->root(), s->particles->particle);
    KNeighbors knbrs;
    Particle* const p = s->particles->particle;
    const int np = s->particles->np_with_buffers;
    double t0 = now_s();
#pragma omp parallel for private(knbrs)
    for (int i = 0; i < np; ++i) neighbor_search(p + i, &knbrs);
    if (seconds) *seconds = now_s() - t0;
    return 0;
}

// The reference's own neighbour list for every particle (find_nbr_recursive),
// returned ascending (slot 0 = nearest) in original indices.  rk is sqrtf of
// the farthest slot as in nbr_search.cpp:73.  Each particle's search is
// independent (the loop of clumping_tree_main.cpp:425-430 without its density
// scatter), so nthreads > 1 gives the same lists.
int ref_knn_lists(void* h, int32_t* nbr, float* d2, float* rk, int nthreads)
{
    RefState* s = (RefState*)h;
    set_threads(nthreads);
    set_nbr_search(s->tree->root(), s->particles->particle);
    Particle* const p = s->particles->particle;
    const int k = KNeighbors::k;
    KNeighbors knbrs;
#pragma omp parallel for private(knbrs) schedule(dynamic, 4096)
    for (int i = 0; i < s->n; ++i) {
        knbrs.clear();
        find_nbr_recursive(s->tree->root(), p[i].x, &knbrs);
        const int oi = (*s->perm)[i];
        for (int slot = 0; slot < k; ++slot) {
            // descending array: index k-1 is the nearest
            Particle* q = knbrs.nbr(k - 1 - slot);
            nbr[(int64_t)oi * k + slot] = q ? (*s->perm)[q - p] : -1;
            d2[(int64_t)oi * k + slot] = knbrs.r2(k - 1 - slot);
        }
        rk[oi] = std::sqrt(knbrs.get_kth_value());
    }
    return k;
}

// rk, dens, igrp in original order.
int ref_get_particles(void* h, float* rk, float* dens, int32_t* igrp)
{
    RefState* s = (RefState*)h;
    Particle* const p = s->particles->particle;
    for (int i = 0; i < s->n; ++i) {
        const int oi = (*s->perm)[i];
        if (rk) rk[oi] = p[i].rk;
        if (dens) dens[oi] = p[i].dens;
        if (igrp) igrp[oi] = p[i].igrp;
    }
    return 0;
}

// Set dens (original order) before deposition, e.g. to deposit the oracle's
// dens through the reference's assign_on_mesh.
int ref_set_dens(void* h, const float* dens, const float* rk)
{
    RefState* s = (RefState*)h;
    Particle* const p = s->particles->particle;
    for (int i = 0; i < s->n; ++i) {
        const int oi = (*s->perm)[i];
        if (dens) p[i].dens = dens[oi];
        if (rk) p[i].rk = rk[oi];
    }
    return 0;
}

// clear_mesh + assign_on_mesh + mesh_total (clumping_tree_main.cpp:521-528).
int ref_assign_on_mesh(void* h, float* mesh, int nc, int sign, int nthreads,
                       double* totals2, double* seconds)
{
    RefState* s = (RefState*)h;
    set_threads(nthreads);
    double t0 = now_s();
    clear_mesh(mesh, nc);
    assign_on_mesh(s->particles, mesh, nc, sign);
    if (seconds) *seconds = now_s() - t0;
    if (totals2) {
        totals2[0] = mesh_total(mesh, nc);
        totals2[1] = mesh_total(mesh + (int64_t)nc * nc * nc, nc);
    }
    return 0;
}

// mark_halo_particles loop (clumping_tree_main.cpp:465-485).
int ref_mark_halo(void* h, const float* halos, int nh, float left, float right,
                  int nthreads, int64_t* np_halo_out, double* m_halo_out)
{
    RefState* s = (RefState*)h;
    set_threads(nthreads);
    set_nbr_search(s->tree->root(), s->particles->particle);
    long np_halo = 0;
    double m_halo = 0.0;
#pragma omp parallel for reduction(+ : np_halo, m_halo)
    for (int i = 0; i < nh; ++i) {
        const float* hh = halos + 5 * i;
        if (left < hh[0] && hh[0] < right && left < hh[1] && hh[1] < right &&
            left < hh[2] && hh[2] < right) {
            int nph = mark_halo_particles(s->tree->root(), hh, hh[3]);
            np_halo += nph;
            m_halo += hh[4] / 8.0f;
        }
    }
    *np_halo_out = np_halo;
    *m_halo_out = m_halo;
    return 0;
}

// write_mesh_separate (coarse_mesh.cpp:147-208): normalises in place, writes 4 files.
int ref_write_mesh_separate(const char* base, int inode, float* mesh, int nc, float boxsize)
{
    write_mesh_separate(base, inode, mesh, nc, boxsize);
    return 0;
}

// The timed hot path of one snapshot as the reference driver runs it:
// construct (clumping_tree_main.cpp:392-394), the neighbour loop (:425-430) and
// per nc clear + assign (:521-526).  seconds3 = {construction, neighbors,
// mesh_assign} as the reference Logger would charge them.
int ref_time_path(const float* pos, const float* vel, int n, int np_local,
                  float boxsize, float bf, const int* ncs, int n_nc, int nthreads,
                  double* seconds3, double* totals /* 2*n_nc */)
{
    void* h = ref_create(pos, vel, nullptr, n, np_local, boxsize);
    double tc = 0, tn = 0, tm = 0;
    ref_construct(h, -bf * boxsize, (1.0f + bf) * boxsize, &tc);
    ref_neighbor_search_all(h, nthreads, &tn);
    int ncmax = 0;
    for (int i = 0; i < n_nc; ++i) ncmax = ncs[i] > ncmax ? ncs[i] : ncmax;
    float* mesh = (float*)std::malloc(sizeof(float) * 6 * (size_t)ncmax * ncmax * ncmax);
    for (int i = 0; i < n_nc; ++i) {
        double t = 0;
        ref_assign_on_mesh(h, mesh, ncs[i], -1, nthreads, totals ? totals + 2 * i : nullptr, &t);
        tm += t;
    }
    std::free(mesh);
    ref_destroy(h);
    seconds3[0] = tc;
    seconds3[1] = tn;
    seconds3[2] = tm;
    return 0;
}

}  // extern "C"
