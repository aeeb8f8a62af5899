/*
 * CPU oracle for the asmooth adaptive-smoothing path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Built with -O2 -ffp-contract=off (no FMA contraction); on x86-64 float
 * arithmetic is IEEE single precision and sqrtf/division are correctly rounded.
 * See oracle.h for what pins each function.
 */
#include "oracle.h"

#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846   /* the double <math.h> gives under POSIX */
#endif
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Buffered set for one periodic rank.                                       */
/* ------------------------------------------------------------------------ */

/* Region index 0..26 of a position (exchange_buffer.cpp:6-11). */
static int region_of(const float* x, float a, float b)
{
    return ((x[0] >= a) + (x[0] > b))
         + 3 * ((x[1] >= a) + (x[1] > b))
         + 9 * ((x[2] >= a) + (x[2] > b));
}

int64_t orc_buffered_set(const float* xv, int64_t np, float boxsize, float bf,
                         float* out, int64_t cap, int64_t* np_local_out)
{
    /* Thresholds (exchange_buffer.cpp:23-24). */
    const float a = bf * boxsize;
    const float b = (1.0f - bf) * boxsize;

    /* Copy phase (exchange_buffer.cpp:36-47): buffer candidates are stored
     * from the end of the allocation downwards, i.e. q[] ends up in reverse
     * file order. */
    int64_t nbuf_total = 0;
    for (int64_t i = 0; i < np; ++i)
        if (region_of(xv + 6 * i, a, b) != 13) nbuf_total++;
    float* q = (float*)malloc(sizeof(float) * 6 * (size_t)(nbuf_total > 0 ? nbuf_total : 1));
    int64_t nbuf[27];
    memset(nbuf, 0, sizeof(nbuf));
    {
        int64_t w = nbuf_total - 1;
        for (int64_t i = 0; i < np; ++i) {
            int r = region_of(xv + 6 * i, a, b);
            if (r != 13) {
                memcpy(q + 6 * w, xv + 6 * i, sizeof(float) * 6);
                w--;
                nbuf[r]++;
            }
        }
    }
    /* In-place cycle sort by region (exchange_buffer.cpp:60-93). */
    int64_t ibuf[27], ibuf_begin[27], ibuf_end[27];
    ibuf[0] = ibuf_begin[0] = 0;
    ibuf_end[0] = nbuf[0];
    for (int i = 1; i < 27; ++i) {
        ibuf[i] = ibuf[i - 1] + nbuf[i - 1];
        ibuf_begin[i] = ibuf[i];
        ibuf_end[i] = ibuf[i] + nbuf[i];
    }
    for (int i = 0; i < 27; ++i) {
        for (int64_t j = ibuf[i]; j < ibuf_end[i]; ++j) {
            int index = region_of(q + 6 * j, a, b);
            while (index != i) {
                int64_t k = ibuf[index]++;
                float tmp[6];
                memcpy(tmp, q + 6 * k, sizeof(tmp));
                memcpy(q + 6 * k, q + 6 * j, sizeof(tmp));
                memcpy(q + 6 * j, tmp, sizeof(tmp));
                index = region_of(tmp, a, b);
            }
        }
    }

    /* Assemble: local particles in file order, then images per (source region
     * i, direction j) (exchange_buffer.cpp:110-151), shifted by -(dn-1)*L. */
    int64_t cnt = 0;
    float* tmp_all;
    int64_t n_images = 0;
    for (int i = 0; i < 27; ++i) {
        if (i == 13) continue;
        int bn0 = i % 3, bn1 = (i / 3) % 3, bn2 = i / 9;
        for (int j = 0; j < 27; ++j) {
            if (j == 13) continue;
            int dn0 = j % 3, dn1 = (j / 3) % 3, dn2 = j / 9;
            if ((dn0 == 1 || dn0 == bn0) && (dn1 == 1 || dn1 == bn1) && (dn2 == 1 || dn2 == bn2))
                n_images += nbuf[i];
        }
    }
    tmp_all = (float*)malloc(sizeof(float) * 6 * (size_t)(np + n_images + 1));
    memcpy(tmp_all, xv, sizeof(float) * 6 * (size_t)np);
    cnt = np;
    for (int i = 0; i < 27; ++i) {
        if (i == 13) continue;
        int bn0 = i % 3, bn1 = (i / 3) % 3, bn2 = i / 9;
        for (int j = 0; j < 27; ++j) {
            if (j == 13) continue;
            int dn0 = j % 3, dn1 = (j / 3) % 3, dn2 = j / 9;
            if (!((dn0 == 1 || dn0 == bn0) && (dn1 == 1 || dn1 == bn1) && (dn2 == 1 || dn2 == bn2)))
                continue;
            const float s0 = (float)(dn0 - 1) * boxsize;
            const float s1 = (float)(dn1 - 1) * boxsize;
            const float s2 = (float)(dn2 - 1) * boxsize;
            for (int64_t k = ibuf_begin[i]; k < ibuf_begin[i] + nbuf[i]; ++k) {
                float* o = tmp_all + 6 * cnt;
                memcpy(o, q + 6 * k, sizeof(float) * 6);
                o[0] -= s0;
                o[1] -= s1;
                o[2] -= s2;
                cnt++;
            }
        }
    }
    free(q);

    /* remove_extra_particles (clumping_tree_main.cpp:550-576). */
    const float dx = bf * boxsize;
    const float left = -dx;
    const float right = boxsize + dx;
    int64_t nkeep = 0, nlocal = np;
    for (int64_t i = 0; i < cnt; ++i) {
        const float* p = tmp_all + 6 * i;
        if (left < p[0] && p[0] < right && left < p[1] && p[1] < right &&
            left < p[2] && p[2] < right) {
            if (nkeep >= cap) {
                free(tmp_all);
                return -1;
            }
            memcpy(out + 6 * nkeep, p, sizeof(float) * 6);
            nkeep++;
        }
    }
    nlocal -= (cnt - nkeep);
    if (np_local_out) *np_local_out = nlocal;
    free(tmp_all);
    return nkeep;
}

/* ------------------------------------------------------------------------ */
/* Distances and kernels.                                                    */
/* ------------------------------------------------------------------------ */

/* nbr_search.cpp:19-25: dx*dx + dy*dy + dz*dz, left to right, no FMA. */
float orc_norm2(const float* x, const float* y)
{
    const float dx = x[0] - y[0];
    const float dy = x[1] - y[1];
    const float dz = x[2] - y[2];
    float s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

/* nbr_search.cpp:27-37: hinv^3 (1-|dx| hinv)(1-|dy| hinv)(1-|dz| hinv). */
float orc_tent(const float* x, const float* y, float hinv)
{
    float w = hinv * hinv;
    w = w * hinv;
    w = w * (1.0f - fabsf(x[0] - y[0]) * hinv);
    w = w * (1.0f - fabsf(x[1] - y[1]) * hinv);
    w = w * (1.0f - fabsf(x[2] - y[2]) * hinv);
    return w;
}

/* nbr_search.cpp:39-53 (spline_ker, defined there but not called): the cubic
 * spline 8/(pi h^3) (1 - 6 r^2 + 6 r^3) for r < 1/2, 8/(pi h^3) 2 (1 - r)^3
 * beyond, r = |x - y| / h.  The factor is evaluated in FP64 (M_PI is a double)
 * and rounded once; r = sqrtf(d2) * hinv; the pieces in FP32, left to right. */
float orc_spline(const float* x, const float* y, float hinv)
{
    const double hd = (double)hinv;
    const float fac = (float)(((8.0 / M_PI) * hd) * hd * hd);
    const float r = sqrtf(orc_norm2(x, y)) * hinv;
    if (r < 0.5f) {
        const float r6 = 6.0f * r;
        return fac * ((1.0f - r6 * r) + (r6 * r) * r);
    }
    const float r1 = 1.0f - r;
    return fac * (((2.0f * r1) * r1) * r1);
}

/* Density and mesh kernel of the oracle: 0 the tent (the reference's), 1 the
 * cubic spline (orc_spline for the density; the separable cubic B-spline,
 * orc_spline_integ, on the mesh). */
static int g_kernel = 0;

void orc_set_kernel(int kernel)
{
    g_kernel = kernel;
}

/* The separable cubic B-spline of half-width 1 normalised to one,
 * w(t) = 4/3 (1 - 6 t^2 + 6 |t|^3) for |t| < 1/2 and 8/3 (1 - |t|)^3 up to 1:
 * its antiderivative on [0, inf), G(u) = 4/3 (u - 2 u^3 + 3/2 u^4) below 1/2,
 * 1/2 - 2/3 (1 - u)^4 up to 1, then 1/2. */
static float spline_g(float u)
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

/* Its integral over [x1, x2], as orc_ker_integ is the tent's. */
float orc_spline_integ(float x1, float x2)
{
    const float fac1 = (float)(-1 + 2 * (x1 > 0.0f));
    const float fac2 = (float)(-1 + 2 * (x2 > 0.0f));
    return fac2 * spline_g(fac2 * x2) - fac1 * spline_g(fac1 * x1);
}

/* coarse_mesh.cpp:10-15: antiderivative of the unit tent on [0, inf). */
static float tent_f(float x)
{
    return x >= 1.0f ? 0.5f : x - 0.5f * x * x;
}

/* coarse_mesh.cpp:17-22. */
float orc_ker_integ(float x1, float x2)
{
    const float fac1 = (float)(-1 + 2 * (x1 > 0.0f));
    const float fac2 = (float)(-1 + 2 * (x2 > 0.0f));
    return fac2 * tent_f(fac2 * x2) - fac1 * tent_f(fac1 * x1);
}

/* The cell integral of the selected kernel. */
static float cell_integ(float x1, float x2)
{
    return g_kernel == 1 ? orc_spline_integ(x1, x2) : orc_ker_integ(x1, x2);
}

/* ------------------------------------------------------------------------ */
/* Exact kNN in canonical order.                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    int k;
    float* d2;
    int32_t* id;
} klist;

static void klist_init(klist* L)
{
    for (int i = 0; i < L->k; ++i) {
        L->d2[i] = FLT_MAX;
        L->id[i] = INT32_MAX;
    }
}

/* Insert (d, i) keeping ascending (d2, id); the list keeps its k smallest. */
static void klist_push(klist* L, float d, int32_t i)
{
    const int k = L->k;
    if (!(d < L->d2[k - 1] || (d == L->d2[k - 1] && i < L->id[k - 1]))) return;
    int j = k - 1;
    while (j > 0 && (d < L->d2[j - 1] || (d == L->d2[j - 1] && i < L->id[j - 1]))) {
        L->d2[j] = L->d2[j - 1];
        L->id[j] = L->id[j - 1];
        --j;
    }
    L->d2[j] = d;
    L->id[j] = i;
}

int orc_knn_brute(const float* pos, int64_t n, const int64_t* queries, int64_t nq,
                  int k_slots, int32_t* nbr, float* d2, float* rk, int nthreads)
{
    if (n < k_slots || k_slots <= 0) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        float* dbuf = (float*)malloc(sizeof(float) * k_slots);
        int32_t* ibuf = (int32_t*)malloc(sizeof(int32_t) * k_slots);
        klist L = {k_slots, dbuf, ibuf};
#pragma omp for schedule(dynamic, 64)
        for (int64_t qi = 0; qi < nq; ++qi) {
            const int64_t q = queries ? queries[qi] : qi;
            klist_init(&L);
            for (int64_t j = 0; j < n; ++j)
                klist_push(&L, orc_norm2(pos + 3 * q, pos + 3 * j), (int32_t)j);
            for (int s = 0; s < k_slots; ++s) {
                if (nbr) nbr[qi * k_slots + s] = L.id[s];
                if (d2) d2[qi * k_slots + s] = L.d2[s];
            }
            if (rk) rk[qi] = sqrtf(L.d2[k_slots - 1]);
        }
        free(dbuf);
        free(ibuf);
    }
    return 0;
}

/* Grid-accelerated exact search.  Shells of cells at growing Chebyshev radius
 * are scanned until every unscanned point is provably farther (FP64 gap with
 * a 1e-5 relative margin, far above the FP32 rounding of d2) than the current
 * last slot, so ties at the last slot are still resolved by index. */
int orc_knn(const float* pos, int64_t n, const int64_t* queries, int64_t nq,
            int k_slots, int32_t* nbr, float* d2, float* rk, int nthreads)
{
    if (n < k_slots || k_slots <= 0) return -1;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            if (pos[3 * i + a] < lo[a]) lo[a] = pos[3 * i + a];
            if (pos[3 * i + a] > hi[a]) hi[a] = pos[3 * i + a];
        }
    double ext[3], vol = 1.0;
    for (int a = 0; a < 3; ++a) {
        ext[a] = (double)hi[a] - (double)lo[a];
        if (ext[a] <= 0.0) ext[a] = 1e-3;
        vol *= ext[a];
    }
    double cs = cbrt(vol * 4.0 / (double)n);
    int g[3];
    for (;;) {
        int64_t tot = 1;
        for (int a = 0; a < 3; ++a) {
            g[a] = (int)(ext[a] / cs) + 1;
            if (g[a] < 1) g[a] = 1;
            tot *= g[a];
        }
        if (tot <= 4 * n + 64) break;
        cs *= 1.25;
    }
    const int64_t ncell = (int64_t)g[0] * g[1] * g[2];
    int64_t* start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cell_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int32_t* items = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (int)(((double)pos[3 * i + a] - lo[a]) / cs);
            if (c[a] >= g[a]) c[a] = g[a] - 1;
            if (c[a] < 0) c[a] = 0;
        }
        cell_of[i] = ((int64_t)c[2] * g[1] + c[1]) * g[0] + c[0];
        start[cell_of[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
        memcpy(fill, start, sizeof(int64_t) * (size_t)ncell);
        for (int64_t i = 0; i < n; ++i) items[fill[cell_of[i]]++] = (int32_t)i;
        free(fill);
    }
    free(cell_of);

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        float* dbuf = (float*)malloc(sizeof(float) * k_slots);
        int32_t* ibuf = (int32_t*)malloc(sizeof(int32_t) * k_slots);
        klist L = {k_slots, dbuf, ibuf};
#pragma omp for schedule(dynamic, 256)
        for (int64_t qi = 0; qi < nq; ++qi) {
            const int64_t q = queries ? queries[qi] : qi;
            const float* x = pos + 3 * q;
            int c[3];
            for (int a = 0; a < 3; ++a) {
                c[a] = (int)(((double)x[a] - lo[a]) / cs);
                if (c[a] >= g[a]) c[a] = g[a] - 1;
                if (c[a] < 0) c[a] = 0;
            }
            klist_init(&L);
            for (int r = 0;; ++r) {
                const int z0 = c[2] - r, z1 = c[2] + r;
                const int y0 = c[1] - r, y1 = c[1] + r;
                const int x0 = c[0] - r, x1 = c[0] + r;
                for (int cz = z0; cz <= z1; ++cz) {
                    if (cz < 0 || cz >= g[2]) continue;
                    for (int cy = y0; cy <= y1; ++cy) {
                        if (cy < 0 || cy >= g[1]) continue;
                        const int onface = (cz == z0 || cz == z1 || cy == y0 || cy == y1);
                        for (int cx = x0; cx <= x1; cx += (onface ? 1 : (x1 - x0 > 0 ? x1 - x0 : 1))) {
                            if (cx < 0 || cx >= g[0]) continue;
                            const int64_t cell = ((int64_t)cz * g[1] + cy) * g[0] + cx;
                            for (int64_t t = start[cell]; t < start[cell + 1]; ++t) {
                                const int32_t j = items[t];
                                klist_push(&L, orc_norm2(x, pos + 3 * (int64_t)j), j);
                            }
                        }
                    }
                }
                /* Distance from x to the outside of the scanned cell cube. */
                double gap = INFINITY;
                int all_boundary = 1;
                for (int a = 0; a < 3; ++a) {
                    if (c[a] - r > 0) {
                        double face = lo[a] + (double)(c[a] - r) * cs;
                        double d = (double)x[a] - face;
                        if (d < gap) gap = d;
                        all_boundary = 0;
                    }
                    if (c[a] + r < g[a] - 1) {
                        double face = lo[a] + (double)(c[a] + r + 1) * cs;
                        double d = face - (double)x[a];
                        if (d < gap) gap = d;
                        all_boundary = 0;
                    }
                }
                if (all_boundary) break;
                if (L.d2[k_slots - 1] < FLT_MAX) {
                    if (gap < 0) gap = 0;
                    double kth = (double)L.d2[k_slots - 1];
                    if (gap * gap > kth * (1.0 + 1e-5) + 1e-30) break;
                }
            }
            for (int s = 0; s < k_slots; ++s) {
                if (nbr) nbr[qi * k_slots + s] = L.id[s];
                if (d2) d2[qi * k_slots + s] = L.d2[s];
            }
            if (rk) rk[qi] = sqrtf(L.d2[k_slots - 1]);
        }
        free(dbuf);
        free(ibuf);
    }
    free(start);
    free(items);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Density scatter.                                                          */
/* ------------------------------------------------------------------------ */

void orc_density(const float* pos, int64_t n, const int32_t* nbr, int k_slots,
                 const float* rk, double* dens64)
{
    for (int64_t p = 0; p < n; ++p) {
        const float hinv = 1.0f / rk[p];
        /* The last slot (the k-th other particle, which sets h) is excluded
         * (nbr_search.cpp:78: slots 1..k-1 of the descending list). */
        for (int s = 0; s < k_slots - 1; ++s) {
            const int32_t q = nbr[p * k_slots + s];
            const float* x = pos + 3 * p;
            const float* y = pos + 3 * (int64_t)q;
            dens64[q] += (double)(g_kernel == 1 ? orc_spline(x, y, hinv) : orc_tent(x, y, hinv));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Mesh deposition.                                                          */
/* ------------------------------------------------------------------------ */

static inline void deposit_cell(double* mesh64, int64_t nmesh, int64_t index, float ker3,
                                int igm, float dens, const float* v)
{
    double* nn = mesh64;
    double* halo = nn + nmesh;
    double* c = halo + nmesh;
    double* vv = c + nmesh;
    if (igm) {
        nn[index] += (double)ker3;
        c[index] += (double)(ker3 * dens);
        vv[3 * index + 0] += (double)(ker3 * v[0]);
        vv[3 * index + 1] += (double)(ker3 * v[1]);
        vv[3 * index + 2] += (double)(ker3 * v[2]);
    } else {
        halo[index] += (double)ker3;
    }
}

int orc_assign_on_mesh(const float* pos, const float* vel, const float* dens,
                       const float* rk, const int32_t* igrp, int64_t n,
                       float boxsize, int nc, int sign, double* mesh64)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    const float dx = boxsize / (float)nc;
    const int64_t nc2 = (int64_t)nc * nc;
    const float ncf = (float)nc;
    for (int64_t j = 0; j < n; ++j) {
        const float* x = pos + 3 * j;
        const float h = rk[j];
        if (!(h > 0.0f)) return -1;
        const float hinv = 1.0f / h;
        int left[3], right[3];
        for (int i = 0; i < 3; ++i) {
            int l = (int)((x[i] - h) / boxsize * ncf);
            int r = 1 + (int)((x[i] + h) / boxsize * ncf);
            left[i] = l > 0 ? l : 0;
            right[i] = r < nc ? r : nc;
        }
        const float ds = dx * hinv;
        const int igm = (sign * igrp[j] <= 0);
        for (int iz = left[2]; iz < right[2]; ++iz) {
            float s1 = ((float)iz * dx - x[2]) * hinv;
            const float kz = cell_integ(s1, s1 + ds);
            for (int iy = left[1]; iy < right[1]; ++iy) {
                s1 = ((float)iy * dx - x[1]) * hinv;
                const float ky = cell_integ(s1, s1 + ds);
                for (int ix = left[0]; ix < right[0]; ++ix) {
                    s1 = ((float)ix * dx - x[0]) * hinv;
                    const float kx = cell_integ(s1, s1 + ds);
                    const int64_t index = iz * nc2 + (int64_t)iy * nc + ix;
                    float ker3 = kx * ky;
                    ker3 = ker3 * kz;
                    deposit_cell(mesh64, nmesh, index, ker3, igm, dens[j], vel + 3 * j);
                }
            }
        }
    }
    return 0;
}

int orc_assign_on_global_mesh(const float* pos, const float* vel, const float* dens,
                              const float* rk, const int32_t* igrp, int64_t n,
                              float boxsize, int nc, int sign, const int* coord,
                              double* mesh64)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    const float dx = boxsize / (float)nc;
    const int64_t nc2 = (int64_t)nc * nc;
    const float ncf = (float)nc;
    const float offset[3] = {boxsize * (float)coord[0], boxsize * (float)coord[1],
                             boxsize * (float)coord[2]};
    for (int64_t j = 0; j < n; ++j) {
        const float* y = pos + 3 * j;
        /* Local particles only (global_mesh.cpp:52-56). */
        if (y[0] < 0.0f || y[0] >= boxsize || y[1] < 0.0f || y[1] >= boxsize ||
            y[2] < 0.0f || y[2] >= boxsize)
            continue;
        float x[3] = {y[0] + offset[0], y[1] + offset[1], y[2] + offset[2]};
        const float h = rk[j];
        if (!(h > 0.0f)) return -1;
        const float hinv = 1.0f / h;
        int left[3], right[3];
        for (int i = 0; i < 3; ++i) {
            left[i] = (int)floorf((x[i] - h) / boxsize * ncf);
            right[i] = 1 + (int)ceilf((x[i] + h) / boxsize * ncf);
            if (left[i] < -nc) return -2;
        }
        const float ds = dx * hinv;
        const int igm = (sign * igrp[j] <= 0);
        for (int iz = left[2]; iz < right[2]; ++iz) {
            float s1 = ((float)iz * dx - x[2]) * hinv;
            const float kz = cell_integ(s1, s1 + ds);
            for (int iy = left[1]; iy < right[1]; ++iy) {
                s1 = ((float)iy * dx - x[1]) * hinv;
                const float ky = cell_integ(s1, s1 + ds);
                for (int ix = left[0]; ix < right[0]; ++ix) {
                    s1 = ((float)ix * dx - x[0]) * hinv;
                    const float kx = cell_integ(s1, s1 + ds);
                    const int64_t index = (int64_t)((iz + nc) % nc) * nc2 +
                                          (int64_t)((iy + nc) % nc) * nc + ((ix + nc) % nc);
                    float ker3 = kx * ky;
                    ker3 = ker3 * kz;
                    deposit_cell(mesh64, nmesh, index, ker3, igm, dens[j], vel + 3 * j);
                }
            }
        }
    }
    return 0;
}

void orc_mesh_total_f32(const float* mesh, int nc, double* out2)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    double a = 0.0, b = 0.0;
    for (int64_t j = 0; j < nmesh; ++j) a += mesh[j];
    for (int64_t j = 0; j < nmesh; ++j) b += mesh[nmesh + j];
    out2[0] = a;
    out2[1] = b;
}

void orc_mesh_total_f64(const double* mesh, int nc, double* out2)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    double a = 0.0, b = 0.0;
    for (int64_t j = 0; j < nmesh; ++j) a += mesh[j];
    for (int64_t j = 0; j < nmesh; ++j) b += mesh[nmesh + j];
    out2[0] = a;
    out2[1] = b;
}

void orc_normalize_f32(float* mesh, int nc, float boxsize)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    const float vol = powf(boxsize / (float)nc, 3.0f);
    float* n = mesh;
    float* halo = mesh + nmesh;
    float* c = halo + nmesh;
    float* v = c + nmesh;
    for (int64_t j = 0; j < nmesh; ++j) {
        if (n[j] > 0.0f) {
            c[j] *= vol / (n[j] * n[j]);
            v[3 * j] /= n[j];
            v[3 * j + 1] /= n[j];
            v[3 * j + 2] /= n[j];
        } else {
            c[j] = v[3 * j] = v[3 * j + 1] = v[3 * j + 2] = 0.0f;
        }
        n[j] *= 8.0f;
        halo[j] *= 8.0f;
    }
}

void orc_normalize_f64(double* mesh, int nc, double boxsize)
{
    const int64_t nmesh = (int64_t)nc * nc * nc;
    const double d = boxsize / nc;
    const double vol = d * d * d;
    double* n = mesh;
    double* halo = mesh + nmesh;
    double* c = halo + nmesh;
    double* v = c + nmesh;
    for (int64_t j = 0; j < nmesh; ++j) {
        if (n[j] > 0.0) {
            c[j] *= vol / (n[j] * n[j]);
            v[3 * j] /= n[j];
            v[3 * j + 1] /= n[j];
            v[3 * j + 2] /= n[j];
        } else {
            c[j] = v[3 * j] = v[3 * j + 1] = v[3 * j + 2] = 0.0;
        }
        n[j] *= 8.0;
        halo[j] *= 8.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Halo excision.                                                            */
/* ------------------------------------------------------------------------ */

void orc_mark_halo(const float* pos, int64_t n, const float* halos, int64_t nh,
                   float left, float right, int32_t* igrp, int64_t* np_halo,
                   double* m_halo)
{
    int64_t count = 0;
    double mass = 0.0;
    for (int64_t i = 0; i < nh; ++i) {
        const float* h = halos + 5 * i;
        if (!(left < h[0] && h[0] < right && left < h[1] && h[1] < right &&
              left < h[2] && h[2] < right))
            continue;
        const float r = h[3];
        const float r2 = r * r;
        for (int64_t j = 0; j < n; ++j) {
            if (orc_norm2(h, pos + 3 * j) < r2) {
                igrp[j] = -1;
                count++;
            }
        }
        mass += (double)(h[4] / 8.0f);
    }
    *np_halo = count;
    *m_halo = mass;
}
