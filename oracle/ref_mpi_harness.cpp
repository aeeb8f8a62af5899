// Harness around the reference's MPI-dependent functions.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `mpi`) compiles this file
// with MPICH's mpicxx together with the reference sources where they lie
// (exchange_buffer.cpp, mpi_interface.cpp, global_mesh.cpp, halo_file2.cpp,
// file_reader2.cpp, basic_types.cpp) into oracle/_ref/ref_mpi.  It is run
// only in the container that holds the reference, under mpiexec, by
// tests/golden/make_golden_mpi.py, which turns its outputs into the fixture
// tests/golden/reference_mpi.npz.  Nothing else runs it.
//
// Every rank writes one little binary file <out>/<mode><inode>.bin:
//   int32 magic 0x61736d31, int32 inode, int32 coords[3], int32 count, int32 width,
//   then count * width float32 words (records as the reference holds them).
//
// Modes (argv[1]):
//   exchange <node_dir> <redshift> <buffer_factor> <alloc> <boxsize> <out>
//       read_pm_file_xv (file_reader2.cpp:24-86) of this rank's node, then
//       exchange_buffer (exchange_buffer.cpp:13-167): the rank's particles
//       after the exchange, np_with_buffers records of 9 words (Particle).
//       The node cube side is set as the driver sets it (boxsize = -xv L,
//       clumping_tree_main.cpp:350-352) from argv[6].
//   zip <node_dir> <redshift> <nc_node_dim> <mesh_scale> <alloc> <out>
//       read_pm_file_zip (file_reader2.cpp:90-272): np_local records of 6
//       words (x, v).
//   global_mesh <particle_file> <nc> <sign> <out>
//       assign_on_global_mesh (global_mesh.cpp:25-123) with this rank's
//       coordinate, into a zeroed mesh: 6 nc^3 words.  The particle file is
//       float32 boxsize, int32 n, then n Particle records (9 words).
//   halo <halo_file> <boxsize> <z> <omega_m> <buffer_factor> <alloc> <out>
//       (halos->boxsize as clumping_tree_main.cpp:362 sets it)
//       read_and_exchage_halo (halo_file2.cpp:63-86, reader :88-138, periodic
//       exchange exchange_buffer_template.h:18-122), no shift: np_with_buffers
//       Halo records of 5 words.
#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "basic_types.h"
#include "mpi_interface.h"
#include "exchange_buffer.h"
#include "file_reader2.h"
#include "halo_file.h"

// Defined with external linkage in global_mesh.cpp:25 (global_mesh.h declares nothing).
void assign_on_global_mesh(Particles const* const particles, MpiInterface const* const mpi, float mesh[],
                           const int nc, const int sign);

namespace {

// Particles::allocate takes its records from malloc (basic_types.cpp:24-31);
// they are freed here so that the destructor's delete never sees them.
void release(Particles& particles)
{
    std::free(particles.particle);
    particles.particle = nullptr;
}

int write_out(const std::string& dir, const char* mode, const MpiInterface& mpi, const float* data,
              int64_t count, int width)
{
    char path[1024];
    std::snprintf(path, sizeof(path), "%s/%s%d.bin", dir.c_str(), mode, mpi.index());
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
        std::fprintf(stderr, "cannot write %s\n", path);
        return 1;
    }
    const int32_t head[7] = {0x61736d31, mpi.index(), mpi.coordinate()[0], mpi.coordinate()[1],
                             mpi.coordinate()[2], static_cast<int32_t>(count), width};
    std::fwrite(head, sizeof(int32_t), 7, fp);
    std::fwrite(data, sizeof(float), static_cast<size_t>(count) * width, fp);
    std::fclose(fp);
    return 0;
}

int run(int argc, char** argv, const MpiInterface& mpi)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    const float zero_shift[3] = {0.0f, 0.0f, 0.0f};
    if (mode == "exchange" && argc == 8) {
        Particles particles;
        particles.allocate(std::atoi(argv[5]));
        particles.boxsize = static_cast<float>(std::atof(argv[6]));
        const float bf = static_cast<float>(std::atof(argv[4]));
        read_pm_file_xv(argv[2], argv[3], mpi.index(), bf, zero_shift, &particles);
        exchange_buffer(&mpi, &particles, bf);
        static_assert(sizeof(Particle) == 9 * sizeof(float), "Particle is 9 words");
        const int rc = write_out(argv[7], "exchange", mpi, reinterpret_cast<const float*>(particles.particle),
                                 particles.np_with_buffers, 9);
        release(particles);
        return rc;
    }
    if (mode == "zip" && argc == 8) {
        Particles particles;
        particles.allocate(std::atoi(argv[6]));
        read_pm_file_zip(argv[2], argv[3], mpi.index(), 0.0f, std::atoi(argv[4]), std::atoi(argv[5]), zero_shift,
                         &particles);
        std::vector<float> xv(6 * static_cast<size_t>(particles.np_local));
        for (int i = 0; i < particles.np_local; ++i)
            for (int a = 0; a < 3; ++a) {
                xv[6 * i + a] = particles.particle[i].x[a];
                xv[6 * i + 3 + a] = particles.particle[i].v[a];
            }
        release(particles);
        return write_out(argv[7], "zip", mpi, xv.data(), static_cast<int64_t>(xv.size() / 6), 6);
    }
    if (mode == "global_mesh" && argc == 6) {
        FILE* fp = std::fopen(argv[2], "rb");
        if (!fp) return 3;
        float boxsize = 0.0f;
        int32_t n = 0;
        if (std::fread(&boxsize, sizeof(float), 1, fp) != 1 || std::fread(&n, sizeof(int32_t), 1, fp) != 1) return 3;
        Particles particles;
        particles.allocate(n + 1);
        if (std::fread(particles.particle, sizeof(Particle), n, fp) != static_cast<size_t>(n)) return 3;
        std::fclose(fp);
        particles.boxsize = boxsize;
        particles.np_local = n;
        particles.np_with_buffers = n;
        const int nc = std::atoi(argv[3]);
        const int sign = std::atoi(argv[4]);
        std::vector<float> mesh(6 * static_cast<size_t>(nc) * nc * nc, 0.0f);
        assign_on_global_mesh(&particles, &mpi, mesh.data(), nc, sign);
        release(particles);
        return write_out(argv[5], "global_mesh", mpi, mesh.data(), static_cast<int64_t>(mesh.size()), 1);
    }
    if (mode == "halo" && argc == 9) {
        ParticleSet<Halo> halos;
        halos.allocate(std::atoi(argv[7]));
        halos.boxsize = static_cast<float>(std::atof(argv[3]));
        read_and_exchage_halo(argv[2], static_cast<float>(std::atof(argv[4])),
                              static_cast<float>(std::atof(argv[5])), &mpi, &halos,
                              static_cast<float>(std::atof(argv[6])), zero_shift);
        static_assert(sizeof(Halo) == 5 * sizeof(float), "Halo is 5 words");
        return write_out(argv[8], "halo", mpi, reinterpret_cast<const float*>(halos.particle),
                         halos.np_with_buffers, 5);
    }
    std::fprintf(stderr, "usage: ref_mpi exchange|zip|global_mesh|halo ... (see the source header)\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rc = 0;
    {
        MpiInterface mpi;
        rc = run(argc, argv, mpi);
    }
    MPI_Finalize();
    return rc;
}
