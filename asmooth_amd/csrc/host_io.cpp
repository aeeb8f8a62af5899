// Host-side steps either side of the device path: the buffered particle set
// of a node, the cubep3m xv reader/writer, the mesh file writer and the
// stitching of node mesh files into global ones.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/asmooth.h"

namespace {

// Region of one axis: 0 below a, 2 above b, 1 between (exchange_buffer.cpp:6-11).
inline int axis_region(float u, float a, float b)
{
    return (u >= a) + (u > b);
}

inline int region_index(const float* u, float a, float b)
{
    return axis_region(u[0], a, b) + 3 * axis_region(u[1], a, b) + 9 * axis_region(u[2], a, b);
}

// Direction j (0..26, 13 = none) may carry a particle of region i when, on
// every axis, it either does not move or moves toward the face the particle
// is near (exchange_buffer.cpp:123-126).
inline bool direction_carries(int i, int j)
{
    if (i == 13 || j == 13) return false;
    const int bn[3] = {i % 3, (i / 3) % 3, i / 9};
    const int dn[3] = {j % 3, (j / 3) % 3, j / 9};
    for (int a = 0; a < 3; ++a)
        if (!(dn[a] == 1 || dn[a] == bn[a])) return false;
    return true;
}

inline asm_particle make_rec(const float* u, const float* v)
{
    asm_particle p;
    for (int a = 0; a < 3; ++a) {
        p.x[a] = u[a];
        p.v[a] = v[a];
    }
    p.dens = 0.0f;
    p.rk = 0.0f;
    p.igrp = 0;
    return p;
}

// The buffer records of one node in the order its exchange sends them
// (exchange_buffer.cpp:36-93, exchange_buffer_template.h:38-79): the records
// of `idx` (the node's own, in file order) outside the central region are
// taken in reverse order (the copy loop writes them from the end of the
// allocation down, :36-47), then cycle-sorted in place by region (:60-93).
// q receives record indices; region r occupies q[begin[r], begin[r] + nbuf[r]).
template <typename Region>
void send_order(const std::vector<int64_t>& idx, Region&& region, std::vector<int64_t>& q, int64_t begin[27],
                int64_t nbuf[27])
{
    q.clear();
    for (int r = 0; r < 27; ++r) nbuf[r] = 0;
    for (size_t t = idx.size(); t-- > 0;) {
        const int r = region(idx[t]);
        if (r != 13) {
            q.push_back(idx[t]);
            nbuf[r]++;
        }
    }
    int64_t next[27];
    next[0] = begin[0] = 0;
    for (int r = 1; r < 27; ++r) next[r] = begin[r] = begin[r - 1] + nbuf[r - 1];
    for (int r = 0; r < 27; ++r) {
        const int64_t end = begin[r] + nbuf[r];
        for (int64_t j = next[r]; j < end; ++j) {
            int g = region(q[j]);
            while (g != r) {
                const int64_t k = next[g]++;
                std::swap(q[k], q[j]);
                g = region(q[j]);
            }
        }
    }
}

// Local records followed by their periodic images, in the order the
// reference's single-rank exchange produces them (exchange_buffer.cpp:13-151,
// exchange_buffer_template.h:18-122): the records in file order, then, per
// (source region i, direction j) pair that carries region i (:110-151), the
// region-i records in send order (send_order), shifted by -(dn-1) L.
// emit(rec, shift) receives each record with the shift to subtract from its
// position (zero for the local copies).
template <typename Emit>
void periodic_images(const float* recs, int width, int64_t n, float boxsize, float bf, Emit&& emit)
{
    const float a = bf * boxsize;            // exchange_buffer.cpp:23
    const float b = (1.0f - bf) * boxsize;   // exchange_buffer.cpp:24
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    std::vector<int64_t> idx(static_cast<size_t>(n)), q;
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    int64_t begin[27], nbuf[27];
    send_order(idx, [&](int64_t i) { return region_index(recs + width * i, a, b); }, q, begin, nbuf);
    for (int64_t i = 0; i < n; ++i) emit(recs + width * i, zero);
    for (int i = 0; i < 27; ++i) {
        for (int j = 0; j < 27; ++j) {
            if (!direction_carries(i, j)) continue;
            const float s[3] = {static_cast<float>(j % 3 - 1) * boxsize,
                                static_cast<float>((j / 3) % 3 - 1) * boxsize,
                                static_cast<float>(j / 9 - 1) * boxsize};
            for (int64_t k = begin[i]; k < begin[i] + nbuf[i]; ++k) emit(recs + width * q[k], s);
        }
    }
}

// remove_extra_particles (clumping_tree_main.cpp:550-576): keep records
// strictly inside (-bf L, L + bf L)^3, order preserved.
int64_t keep_inside(std::vector<asm_particle>& all, int64_t nlocal_in, float boxsize, float bf,
                    asm_particle* out, int64_t cap, int64_t* np_local)
{
    const float dx = bf * boxsize;
    const float left = -dx;
    const float right = boxsize + dx;
    int64_t nkeep = 0;
    for (const asm_particle& p : all) {
        if (left < p.x[0] && p.x[0] < right && left < p.x[1] && p.x[1] < right && left < p.x[2] &&
            p.x[2] < right) {
            if (nkeep >= cap) return ASM_ERR_CAPACITY;
            out[nkeep++] = p;
        }
    }
    if (np_local) *np_local = nlocal_in - (static_cast<int64_t>(all.size()) - nkeep);
    return nkeep;
}

}  // namespace

extern "C" {

int64_t asm_build_buffered_set(const float* xv, int64_t np, float boxsize, float bf,
                               asm_particle* out, int64_t cap, int64_t* np_local)
{
    if (!xv || !out || np < 0 || !(boxsize > 0.0f) || !(bf > 0.0f)) return ASM_ERR_ARG;
    std::vector<asm_particle> all;
    periodic_images(xv, 6, np, boxsize, bf, [&](const float* rec, const float* shift) {
        const float u[3] = {rec[0] - shift[0], rec[1] - shift[1], rec[2] - shift[2]};
        all.push_back(make_rec(u, rec + 3));
    });
    return keep_inside(all, np, boxsize, bf, out, cap, np_local);
}

int64_t asm_build_buffered_halos(const asm_halo* halos, int64_t n, float boxsize, float bf, asm_halo* out,
                                 int64_t cap)
{
    if ((!halos && n > 0) || !out || n < 0 || !(boxsize > 0.0f) || !(bf > 0.0f)) return ASM_ERR_ARG;
    static_assert(sizeof(asm_halo) == 5 * sizeof(float), "halo record is 5 floats");
    int64_t cnt = 0;
    bool overflow = false;
    periodic_images(reinterpret_cast<const float*>(halos), 5, n, boxsize, bf,
                    [&](const float* rec, const float* shift) {
                        if (cnt >= cap) {
                            overflow = true;
                            return;
                        }
                        asm_halo h;
                        for (int a = 0; a < 3; ++a) h.x[a] = rec[a] - shift[a];
                        h.r = rec[3];
                        h.mass = rec[4];
                        out[cnt++] = h;
                    });
    return overflow ? ASM_ERR_CAPACITY : cnt;
}

int64_t asm_build_node_share(const float* xv, int64_t np, float box_total, int nn, const int coord[3],
                             float bf, asm_particle* out, int64_t cap, int64_t* np_local)
{
    if (!xv || !out || np < 0 || nn < 1 || !(box_total > 0.0f) || !(bf > 0.0f)) return ASM_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (coord[a] < 0 || coord[a] >= nn) return ASM_ERR_ARG;
    const float L = box_total / static_cast<float>(nn);
    const float a = bf * L;
    const float b = (1.0f - bf) * L;
    // Every particle's node and node-local position (the node files of a
    // cubep3m run hold node-local positions).
    std::vector<float> u(3 * static_cast<size_t>(np));
    std::vector<int32_t> node(static_cast<size_t>(np));
    for (int64_t i = 0; i < np; ++i) {
        const float* p = xv + 6 * i;
        int home[3];
        for (int d = 0; d < 3; ++d) {
            int h = static_cast<int>(p[d] / L);
            if (h < 0) h = 0;
            if (h >= nn) h = nn - 1;
            home[d] = h;
            u[3 * i + d] = p[d] - static_cast<float>(h) * L;
        }
        node[i] = home[0] + nn * (home[1] + nn * home[2]);
    }
    const auto node_at = [nn](int x, int y, int z) {
        return ((x % nn + nn) % nn) + nn * (((y % nn + nn) % nn) + nn * ((z % nn + nn) % nn));
    };
    const int me = node_at(coord[0], coord[1], coord[2]);
    // The nodes that send to this one: the neighbour in direction -dn for
    // every direction dn (mpi_interface.cpp:31-32, 54-68), each with its own
    // records in file order.
    std::vector<int> sender(27, -1);
    std::vector<int> slot(static_cast<size_t>(nn) * nn * nn, -1);
    std::vector<std::vector<int64_t>> members;
    for (int j = 0; j < 27; ++j) {
        if (j == 13) continue;
        const int s = node_at(coord[0] - (j % 3 - 1), coord[1] - ((j / 3) % 3 - 1), coord[2] - (j / 9 - 1));
        if (slot[s] < 0) {
            slot[s] = static_cast<int>(members.size());
            members.emplace_back();
        }
        sender[j] = slot[s];
    }
    std::vector<asm_particle> all;
    for (int64_t i = 0; i < np; ++i) {
        if (node[i] == me) all.push_back(make_rec(&u[3 * i], xv + 6 * i + 3));
        if (slot[node[i]] >= 0) members[slot[node[i]]].push_back(i);
    }
    const int64_t nlocal = static_cast<int64_t>(all.size());
    std::vector<std::vector<int64_t>> q(members.size());
    std::vector<std::array<int64_t, 27>> begin(members.size()), nbuf(members.size());
    const auto region = [&](int64_t i) { return region_index(&u[3 * i], a, b); };
    for (size_t m = 0; m < members.size(); ++m) send_order(members[m], region, q[m], begin[m].data(), nbuf[m].data());
    // Receives in the exchange's order (exchange_buffer.cpp:112-151): per
    // source region i and direction j, the sender's region-i records, shifted
    // by -(dn-1) L.
    for (int i = 0; i < 27; ++i) {
        for (int j = 0; j < 27; ++j) {
            if (!direction_carries(i, j)) continue;
            const int m = sender[j];
            const float s[3] = {static_cast<float>(j % 3 - 1) * L, static_cast<float>((j / 3) % 3 - 1) * L,
                                static_cast<float>(j / 9 - 1) * L};
            for (int64_t k = begin[m][i]; k < begin[m][i] + nbuf[m][i]; ++k) {
                const int64_t r = q[m][k];
                const float w[3] = {u[3 * r] - s[0], u[3 * r + 1] - s[1], u[3 * r + 2] - s[2]};
                all.push_back(make_rec(w, xv + 6 * r + 3));
            }
        }
    }
    return keep_inside(all, nlocal, L, bf, out, cap, np_local);
}

int asm_read_xv(const char* path, float** xv, int64_t* np, float header[11])
{
    if (!path || !xv || !np) return ASM_ERR_ARG;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return ASM_ERR_IO;
    int32_t n = 0;
    float hdr[11];
    if (std::fread(&n, sizeof(int32_t), 1, fp) != 1 || std::fread(hdr, sizeof(float), 11, fp) != 11 || n < 0) {
        std::fclose(fp);
        return ASM_ERR_IO;
    }
    float* buf = static_cast<float*>(std::malloc(sizeof(float) * 6 * static_cast<size_t>(n > 0 ? n : 1)));
    if (!buf) {
        std::fclose(fp);
        return ASM_ERR_CAPACITY;
    }
    if (std::fread(buf, sizeof(float), 6 * static_cast<size_t>(n), fp) != 6 * static_cast<size_t>(n)) {
        std::free(buf);
        std::fclose(fp);
        return ASM_ERR_IO;
    }
    std::fclose(fp);
    *xv = buf;
    *np = n;
    if (header) std::memcpy(header, hdr, sizeof(hdr));
    return ASM_OK;
}

// cubep3m zip reader (file_reader2.cpp:90-272).  Four files per node,
// <filebase><inode>/<redshift>zip{0,1,2,3}_<inode>.dat:
//   zip0: 16-word header (int32 np_local, ..., float v_r2i at word 12), then
//         3 bytes per particle, its offset in its coarse cell in 1/256ths;
//   zip1: 16-word header, then 3 int16 velocities per particle;
//   zip2: one byte per coarse cell (i fastest over nc_node_dim^3 cells), the
//         particle count, 255 meaning "read it from zip3";
//   zip3: int32 counts of the cells marked 255, in cell order.
// Position x = mesh_scale * ((b + 0.5) / 256 + i - 1) - shift and velocity
// v = s / v_r2i, in FP32 and in that order (:226-235).  zip0 and zip1 must end
// exactly after the last particle and the counts must add up to np_local
// (:240-256).  No byte swapping: little-endian files, as the default build.
int asm_read_zip(const char* filebase, const char* redshift, int inode, int nc_node_dim, int mesh_scale,
                 const float shift[3], float** xv, int64_t* np, float header[16])
{
    if (!filebase || !redshift || !xv || !np || nc_node_dim <= 0 || mesh_scale <= 0) return ASM_ERR_ARG;
    // The four files are read whole, then decoded from memory.
    std::vector<unsigned char> zf[4];
    for (int z = 0; z < 4; ++z) {
        std::vector<char> name(std::strlen(filebase) + std::strlen(redshift) + 64);
        std::snprintf(name.data(), name.size(), "%s%d/%szip%d_%d.dat", filebase, inode, redshift, z, inode);
        FILE* fp = std::fopen(name.data(), "rb");
        if (!fp) return ASM_ERR_IO;
        bool ok = std::fseek(fp, 0, SEEK_END) == 0;
        const long size = ok ? std::ftell(fp) : -1;
        ok = ok && size >= 0 && std::fseek(fp, 0, SEEK_SET) == 0;
        if (ok) {
            zf[z].resize(static_cast<size_t>(size));
            ok = std::fread(zf[z].data(), 1, zf[z].size(), fp) == zf[z].size();
        }
        std::fclose(fp);
        if (!ok) return ASM_ERR_IO;
    }
    if (zf[0].size() < 64 || zf[1].size() < 64) return ASM_ERR_IO;
    int32_t h0[16];
    std::memcpy(h0, zf[0].data(), sizeof(h0));
    const int64_t n = h0[0];
    // zip0 and zip1 end exactly after the last particle (:240-248)
    if (n < 0 || zf[0].size() != 64 + 3 * static_cast<size_t>(n) || zf[1].size() != 64 + 6 * static_cast<size_t>(n))
        return ASM_ERR_IO;
    float v_r2i;
    std::memcpy(&v_r2i, &h0[12], sizeof(float));
    const float sh[3] = {shift ? shift[0] : 0.0f, shift ? shift[1] : 0.0f, shift ? shift[2] : 0.0f};
    const int64_t nd = nc_node_dim;
    if (static_cast<int64_t>(zf[2].size()) < nd * nd * nd) return ASM_ERR_IO;
    float* out = static_cast<float*>(std::malloc(sizeof(float) * 6 * static_cast<size_t>(n > 0 ? n : 1)));
    if (!out) return ASM_ERR_CAPACITY;
    const unsigned char* b = zf[0].data() + 64;
    const unsigned char* vs = zf[1].data() + 64;
    const unsigned char* c8 = zf[2].data();
    size_t esc = 0;   // next int32 of zip3
    const float ms = static_cast<float>(mesh_scale);
    int64_t k_out = 0;
    int rc = ASM_OK;
    for (int64_t k = 1; k <= nd && rc == ASM_OK; ++k)
        for (int64_t j = 1; j <= nd && rc == ASM_OK; ++j)
            for (int64_t i = 1; i <= nd; ++i) {
                int32_t cnt = *c8++;
                if (cnt == 255) {
                    if (zf[3].size() < 4 * (esc + 1)) {
                        rc = ASM_ERR_IO;
                        break;
                    }
                    std::memcpy(&cnt, zf[3].data() + 4 * esc, sizeof(int32_t));
                    ++esc;
                }
                if (cnt < 0 || k_out + cnt > n) {
                    rc = ASM_ERR_IO;
                    break;
                }
                const float cell[3] = {static_cast<float>(i), static_cast<float>(j), static_cast<float>(k)};
                for (int32_t l = 0; l < cnt; ++l, ++k_out) {
                    int16_t v[3];
                    std::memcpy(v, vs + 6 * k_out, sizeof(v));
                    float* o = out + 6 * k_out;
                    for (int a = 0; a < 3; ++a) {
                        const float frac = (static_cast<float>(b[3 * k_out + a]) + 0.5f) / 256.0f;
                        o[a] = ms * ((frac + cell[a]) - 1.0f) - sh[a];
                        o[3 + a] = static_cast<float>(v[a]) / v_r2i;
                    }
                }
            }
    if (rc == ASM_OK && k_out != n) rc = ASM_ERR_IO;
    if (rc != ASM_OK) {
        std::free(out);
        return rc;
    }
    *xv = out;
    *np = n;
    if (header) std::memcpy(header, h0, sizeof(h0));
    return ASM_OK;
}

int asm_write_xv(const char* path, const float* xv, int64_t np, const float header[11])
{
    if (!path || !xv || np < 0 || np > 0x7fffffff) return ASM_ERR_ARG;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return ASM_ERR_IO;
    const int32_t n = static_cast<int32_t>(np);
    float hdr[11] = {0};
    if (header) std::memcpy(hdr, header, sizeof(hdr));
    bool ok = std::fwrite(&n, sizeof(int32_t), 1, fp) == 1 && std::fwrite(hdr, sizeof(float), 11, fp) == 11 &&
              std::fwrite(xv, sizeof(float), 6 * static_cast<size_t>(np), fp) == 6 * static_cast<size_t>(np);
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? ASM_OK : ASM_ERR_IO;
}


// Bryan & Norman virial overdensity relative to the mean (halo_file2.cpp:47-61).
double asm_halo_delta_vir(double z, double omega_m)
{
    const double a = 1.0 / (1.0 + z);
    const double omega_l = 1.0 - omega_m;
    const double omega_z = omega_m / (omega_m + omega_l * a * a * a);
    const double x = omega_z - 1.0;
    return (18.0 * M_PI * M_PI + 82.0 * x - 39.0 * x * x) * ((omega_m + omega_l * a * a * a) / omega_m);
}

int asm_read_halo_file(const char* path, float boxsize, const float shift[3], float z, float omega_m, int m200,
                       int pid_flag, asm_halo** halos, int64_t* n_halo)
{
    if (!path || !halos || !n_halo || !(boxsize > 0.0f)) return ASM_ERR_ARG;
    // Record layout (halo_file2.cpp:17-30): 22 floats; position at 0..2; mass
    // at 3 and radius at 5 (Bryan-Norman), or 4 and 6 (M200).
    const int ndat = 22, off_m = m200 ? 4 : 3, off_r = m200 ? 6 : 5;
    if (!m200 && !(omega_m > 0.0f)) return ASM_ERR_ARG;
    const float delta = m200 ? 200.0f : static_cast<float>(asm_halo_delta_vir(z, omega_m));
    const float torr = 10.0f;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return ASM_ERR_IO;
    int32_t header[3];
    if (std::fread(header, sizeof(int32_t), 3, fp) != 3) {
        std::fclose(fp);
        return ASM_ERR_IO;
    }
    std::vector<asm_halo> out;
    float buf[22];
    const float sh[3] = {shift ? shift[0] : 0.0f, shift ? shift[1] : 0.0f, shift ? shift[2] : 0.0f};
    while (std::fread(buf, sizeof(float), ndat, fp) == static_cast<size_t>(ndat)) {
        asm_halo h;
        for (int a = 0; a < 3; ++a) h.x[a] = std::fmod(buf[a], boxsize) - sh[a];
        const float r = buf[off_r];
        const float m = buf[off_m];
        h.r = r;
        h.mass = m;
        // overdensity sanity check (halo_file2.cpp:114-121), an error here
        if (std::fabs(m / (4.0f / 3.0f * M_PI * r * r * r) - delta) >= torr) {
            std::fclose(fp);
            return ASM_ERR_IO;
        }
        // 50 particle IDs (8 bytes) and their x,v (6 floats) follow each record
        // when the catalogue carries them (halo_file2.cpp:123-127).
        if (pid_flag && std::fseek(fp, 50 * (8 + 4 * 6), SEEK_CUR) != 0) {
            std::fclose(fp);
            return ASM_ERR_IO;
        }
        out.push_back(h);
    }
    std::fclose(fp);
    asm_halo* mem = static_cast<asm_halo*>(std::malloc(sizeof(asm_halo) * (out.empty() ? 1 : out.size())));
    if (!mem) return ASM_ERR_CAPACITY;
    if (!out.empty()) std::memcpy(mem, out.data(), sizeof(asm_halo) * out.size());
    *halos = mem;
    *n_halo = static_cast<int64_t>(out.size());
    return ASM_OK;
}

void asm_free(void* p)
{
    std::free(p);
}

int asm_write_mesh_separate(const char* base, int inode, const float* mesh, int nc)
{
    if (!base || !mesh || nc <= 0) return ASM_ERR_ARG;
    const size_t nmesh = static_cast<size_t>(nc) * nc * nc;
    const char* tags[4] = {"n", "h", "c", "vel"};
    const float* parts[4] = {mesh, mesh + nmesh, mesh + 2 * nmesh, mesh + 3 * nmesh};
    const size_t counts[4] = {nmesh, nmesh, nmesh, 3 * nmesh};
    for (int f = 0; f < 4; ++f) {
        std::vector<char> name(std::strlen(base) + 32);
        std::snprintf(name.data(), name.size(), "%s%s%d.dat", base, tags[f], inode);
        FILE* fp = std::fopen(name.data(), "wb");
        if (!fp) return ASM_ERR_IO;
        const int32_t n = nc;
        bool ok = std::fwrite(&n, sizeof(int32_t), 1, fp) == 1 &&
                  std::fwrite(parts[f], sizeof(float), counts[f], fp) == counts[f];
        ok = (std::fclose(fp) == 0) && ok;
        if (!ok) return ASM_ERR_IO;
    }
    return ASM_OK;
}

}  // extern "C"

namespace {

// Files of one z-slab of nodes, closed on scope exit.
struct NodeFiles {
    std::vector<FILE*> fp;
    explicit NodeFiles(size_t n) : fp(n, nullptr) {}
    ~NodeFiles()
    {
        for (FILE* f : fp)
            if (f) std::fclose(f);
    }
};

std::string node_name(const char* fmt, const char* a, const char* b, int inode)
{
    std::vector<char> buf(std::strlen(a) + (b ? std::strlen(b) : 0) + std::strlen(fmt) + 16);
    if (b)
        std::snprintf(buf.data(), buf.size(), fmt, a, b, inode);
    else
        std::snprintf(buf.data(), buf.size(), fmt, a, inode);
    return std::string(buf.data());
}

// Opens one node file and reads its int32 cell count; every file of a
// stitch must hold the same count.
int open_node(const std::string& name, FILE** fp, int* nc)
{
    *fp = std::fopen(name.c_str(), "rb");
    if (!*fp) return ASM_ERR_IO;
    int32_t n = 0;
    if (std::fread(&n, sizeof(int32_t), 1, *fp) != 1 || n <= 0) return ASM_ERR_IO;
    if (*nc == 0) *nc = n;
    return n == *nc ? ASM_OK : ASM_ERR_IO;
}

int write_header(FILE* fp, int64_t global)
{
    if (global > 0x7fffffff) return ASM_ERR_CAPACITY;
    const int32_t h[3] = {static_cast<int32_t>(global), static_cast<int32_t>(global), static_cast<int32_t>(global)};
    return std::fwrite(h, sizeof(int32_t), 3, fp) == 3 ? ASM_OK : ASM_ERR_IO;
}

}  // namespace

extern "C" {

// tools/reformat.cpp:11-123.  Node files <file_base><inode>.dat (inode =
// ix + nn iy + nn^2 iz), each int32 nc then nc^3 cells of n_components
// floats with x fastest, become <out_dir>/<file_base>_all.dat: three int32
// nc*nn, then the global cells with x fastest.  One z-slab of nodes is open
// at a time and rows of nc cells are copied in global order (:55-106).
// *total is the sum of the first nc floats of every row copied, accumulated
// in long double in that order (:104-105; for one component, the sum of the
// field).  Node files longer than nc^3 cells are read only that far, as
// there.
int asm_reformat(int nn, const char* out_dir, const char* file_base, int n_components, double* total)
{
    if (nn <= 0 || !out_dir || !file_base || n_components <= 0) return ASM_ERR_ARG;
    FILE* out = std::fopen(node_name("%s/%s_all.dat", out_dir, file_base, 0).c_str(), "wb");
    if (!out) return ASM_ERR_IO;
    long double sum = 0.0L;
    int nc = 0;
    int rc = ASM_OK;
    std::vector<float> row;
    for (int iz_node = 0; iz_node < nn && rc == ASM_OK; ++iz_node) {
        NodeFiles f(static_cast<size_t>(nn) * nn);
        for (int k = 0; k < nn * nn && rc == ASM_OK; ++k)
            rc = open_node(node_name("%s%d.dat", file_base, nullptr, nn * nn * iz_node + k), &f.fp[k], &nc);
        if (rc != ASM_OK) break;
        const size_t ndata = static_cast<size_t>(nc) * n_components;
        row.resize(ndata);
        if (iz_node == 0) rc = write_header(out, static_cast<int64_t>(nc) * nn);
        for (int iz = 0; iz < nc && rc == ASM_OK; ++iz)
            for (int iy_node = 0; iy_node < nn && rc == ASM_OK; ++iy_node)
                for (int iy = 0; iy < nc && rc == ASM_OK; ++iy)
                    for (int ix_node = 0; ix_node < nn; ++ix_node) {
                        FILE* in = f.fp[iy_node * nn + ix_node];
                        if (std::fread(row.data(), sizeof(float), ndata, in) != ndata ||
                            std::fwrite(row.data(), sizeof(float), ndata, out) != ndata) {
                            rc = ASM_ERR_IO;
                            break;
                        }
                        for (int i = 0; i < nc; ++i) sum += row[i];
                    }
    }
    if (std::fclose(out) != 0 && rc == ASM_OK) rc = ASM_ERR_IO;
    if (total) *total = static_cast<double>(sum);
    return rc;
}

// tools/reformat2.cpp:10-171.  From the n, h and vel node files of
// write_mesh_separate, <in_dir>/<redshift>{n,h,vel}<inode>.dat, writes
// <out_dir>/<redshift>n_all.dat (n as stored), ntot_all.dat (h + n) and
// v_all.dat (vel * (n / 8), the momentum), stitched as asm_reformat does.
// totals[0] = sum of n / 8 and totals[1] = sum of (h + n) / 8, in long
// double in file order (:118-126).
int asm_reformat2(int nn, const char* out_dir, const char* redshift, const char* in_dir, double totals[2])
{
    if (nn <= 0 || !out_dir || !redshift || !in_dir) return ASM_ERR_ARG;
    const char* tags_out[3] = {"%s/%sn_all.dat", "%s/%sntot_all.dat", "%s/%sv_all.dat"};
    const char* tags_in[3] = {"/%sn%d.dat", "/%sh%d.dat", "/%svel%d.dat"};
    FILE* out[3] = {nullptr, nullptr, nullptr};
    int rc = ASM_OK;
    for (int f = 0; f < 3 && rc == ASM_OK; ++f) {
        out[f] = std::fopen(node_name(tags_out[f], out_dir, redshift, 0).c_str(), "wb");
        if (!out[f]) rc = ASM_ERR_IO;
    }
    long double t1 = 0.0L, t2 = 0.0L;
    int nc = 0;
    std::vector<float> bn, bh, bv;
    for (int iz_node = 0; iz_node < nn && rc == ASM_OK; ++iz_node) {
        NodeFiles f(3 * static_cast<size_t>(nn) * nn);
        for (int k = 0; k < nn * nn && rc == ASM_OK; ++k)
            for (int g = 0; g < 3 && rc == ASM_OK; ++g) {
                const std::string fmt = std::string("%s") + tags_in[g];
                rc = open_node(node_name(fmt.c_str(), in_dir, redshift, nn * nn * iz_node + k), &f.fp[3 * k + g],
                               &nc);
            }
        if (rc != ASM_OK) break;
        const size_t m = static_cast<size_t>(nc);
        bn.resize(m);
        bh.resize(m);
        bv.resize(3 * m);
        if (iz_node == 0)
            for (int g = 0; g < 3 && rc == ASM_OK; ++g) rc = write_header(out[g], static_cast<int64_t>(nc) * nn);
        for (int iz = 0; iz < nc && rc == ASM_OK; ++iz)
            for (int iy_node = 0; iy_node < nn && rc == ASM_OK; ++iy_node)
                for (int iy = 0; iy < nc && rc == ASM_OK; ++iy)
                    for (int ix_node = 0; ix_node < nn; ++ix_node) {
                        FILE** in = &f.fp[3 * (iy_node * nn + ix_node)];
                        if (std::fread(bn.data(), sizeof(float), m, in[0]) != m ||
                            std::fread(bh.data(), sizeof(float), m, in[1]) != m ||
                            std::fread(bv.data(), sizeof(float), 3 * m, in[2]) != 3 * m) {
                            rc = ASM_ERR_IO;
                            break;
                        }
                        for (size_t j = 0; j < m; ++j) {
                            const float n = bn[j] / 8.0f;
                            t1 += n;
                            bh[j] += bn[j];
                            t2 += bh[j] / 8.0f;
                            bv[3 * j] *= n;
                            bv[3 * j + 1] *= n;
                            bv[3 * j + 2] *= n;
                        }
                        if (std::fwrite(bn.data(), sizeof(float), m, out[0]) != m ||
                            std::fwrite(bh.data(), sizeof(float), m, out[1]) != m ||
                            std::fwrite(bv.data(), sizeof(float), 3 * m, out[2]) != 3 * m) {
                            rc = ASM_ERR_IO;
                            break;
                        }
                    }
    }
    for (int f = 0; f < 3; ++f)
        if (out[f] && std::fclose(out[f]) != 0 && rc == ASM_OK) rc = ASM_ERR_IO;
    if (totals) {
        totals[0] = static_cast<double>(t1);
        totals[1] = static_cast<double>(t2);
    }
    return rc;
}

}  // extern "C"
