// clumping_tree_gpu: host driver of the device smoothing path with the
// reference driver's command line and output files (clumping_tree_main.cpp).
//
//   clumping_tree_gpu (-xv <L> | -nc_node_dim <n> [-mesh_scale 4]) -nc 32,64
//                     [-buffer_factor 0.05] [-node_dir results/node]
//                     [-pm_redshift redshifts.txt] [-allocate MB] [-device 0]
//                     [-halo_excise 0|1]
//
// One process handles one node share on one GPU (node index 0, single
// periodic rank: the 26 neighbours are the node itself, as with `mpiexec -n 1`
// of the reference).  Per redshift it reads shift/<z>shift.txt if present
// (clumping_tree_main.cpp:323-347), then <node_dir>0/<z>xv0.dat
// (file_reader2.cpp:24-86) less the shift, or with -nc_node_dim the zip
// files <node_dir>0/<z>zip{0,1,2,3}_0.dat (file_reader2.cpp:90-272), builds
// the buffered set, runs the device path and writes
// so/nc<nc>/<z>{n,h,c,vel}0.dat (coarse_mesh.cpp:173-201).
#include <sys/stat.h>
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "asmooth.h"

namespace {

double now_s()
{
    struct timeval tp;
    gettimeofday(&tp, nullptr);
    return static_cast<double>(tp.tv_sec) + static_cast<double>(tp.tv_usec) * 1e-6;
}

// "-key value" pairs (option.cpp:11-19) with the reference's defaults
// (clumping_tree_main.cpp:89-99).
std::map<std::string, std::string> parse_options(int argc, char** argv)
{
    std::map<std::string, std::string> op = {
        {"-pm_redshift", "redshifts.txt"}, {"-node_dir", "results/node"}, {"-allocate", "512"},
        {"-xv", "-1"}, {"-buffer_factor", "0.05"}, {"-nc_node_dim", "0"}, {"-l", "0.2"},
        {"-nc", "256"}, {"-mesh_scale", "4"}, {"-omegam", "0"}, {"-device", "0"}, {"-halo_excise", "0"},
        {"-m200", "0"}, {"-pid_flag", "1"}};
    for (int j = 1; j + 1 < argc; j += 2) {
        if (argv[j][0] != '-')
            std::fprintf(stderr, "Error: %s must be an option\n", argv[j]);
        else
            op[argv[j]] = argv[j + 1];
    }
    return op;
}

// Comma list of nc (clumping_tree_main.cpp:43-52,130-147).
std::vector<int> parse_nc_list(const std::string& s)
{
    std::vector<int> out;
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t q = s.find(',', pos);
        std::string tok = s.substr(pos, q == std::string::npos ? std::string::npos : q - pos);
        out.push_back(std::atoi(tok.c_str()));
        if (q == std::string::npos) break;
        pos = q + 1;
    }
    return out;
}

bool make_dirs(const std::string& path)
{
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur.push_back(path[i]);
        if (path[i] == '/' || i + 1 == path.size()) {
            struct stat st;
            if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0) return false;
        }
    }
    return true;
}

int die(asm_context* ctx, const char* what, int rc)
{
    std::fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, ctx ? asm_last_error(ctx) : "");
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc == 1) {
        std::printf("clumping_tree_gpu [options]\n"
                    "\t-pm_redshift <redshifts.txt>; file with list of redshifts\n"
                    "\t-node_dir <node directory>; -node_dir results/node\n"
                    "\t-allocate <MB for particles> (accepted; device buffers size themselves)\n"
                    "\t-buffer_factor 0.05\n"
                    "\t-xv <local boxsize> (xv particle format)\n"
                    "\t-nc_node_dim 32 (for zip particle format)\n"
                    "\t-mesh_scale 4   (for zip particle format)\n"
                    "\t-nc <nc1,nc2,...> list of number of mesh\n"
                    "\t-device <hip device ordinal>\n"
                    "\t-halo_excise 1 (excise SO halos of <node_dir>0/<z>halo0.dat)\n"
                    "\t-omegam 0.308; (for Bryan-Norman over density)\n"
                    "\t-m200 0|1 (M200/r200 columns instead of Bryan-Norman)\n"
                    "\t-pid_flag 0|1 (halo records followed by particle-ID blocks; default 1)\n");
        return 0;
    }
    auto op = parse_options(argc, argv);
    // -xv <L>: xv files of cube side L; otherwise the zip format, whose cube
    // is mesh_scale * nc_node_dim (clumping_tree_main.cpp:103-127, file_reader2.cpp:268).
    const int nc_node_dim = std::atoi(op["-nc_node_dim"].c_str());
    const int mesh_scale = std::atoi(op["-mesh_scale"].c_str());
    const bool read_xv_file = std::atof(op["-xv"].c_str()) > 0.0;
    const float boxsize = read_xv_file ? static_cast<float>(std::atof(op["-xv"].c_str()))
                                       : static_cast<float>(mesh_scale) * static_cast<float>(nc_node_dim);
    const float bf = static_cast<float>(std::atof(op["-buffer_factor"].c_str()));
    if (!(boxsize > 0.0f)) {
        std::fprintf(stderr, "Error: option -nc_node_dim not given (or give -xv <boxsize>)\n");
        return 1;
    }
    if (!(bf > 0.0f)) {
        std::fprintf(stderr, "Error: -buffer_factor must be positive\n");
        return 1;
    }
    std::vector<int> ncs = parse_nc_list(op["-nc"]);
    for (int nc : ncs)
        if (nc <= 0) {
            std::fprintf(stderr, "Error: bad -nc list %s\n", op["-nc"].c_str());
            return 1;
        }
    std::vector<std::string> redshifts;
    {
        FILE* fp = std::fopen(op["-pm_redshift"].c_str(), "r");
        if (!fp) {
            std::fprintf(stderr, "Error: Redshift file %s does not exist\n", op["-pm_redshift"].c_str());
            return 1;
        }
        char buf[128], z[64];
        while (std::fgets(buf, sizeof(buf), fp))
            if (std::sscanf(buf, "%63s", z) == 1) redshifts.push_back(z);
        std::fclose(fp);
    }
    for (int nc : ncs) {
        char d[64];
        std::snprintf(d, sizeof(d), "so/nc%d/", nc);
        if (!make_dirs(d)) {
            std::fprintf(stderr, "Error: Failed to make new directory: %s\n", d);
            return 1;
        }
    }
    asm_context* ctx = nullptr;
    int rc = asm_create(&ctx, std::atoi(op["-device"].c_str()));
    if (rc) return die(nullptr, "asm_create", rc);
    asm_set_option(ctx, ASM_OPT_PROFILE, 1);
    const int inode = 0;
    const std::string node_dir = op["-node_dir"];

    for (const std::string& z : redshifts) {
        std::printf("z %s\n", z.c_str());
        double t_io0 = now_s();
        float shift[3] = {0, 0, 0};
        {
            char sf[256];
            std::snprintf(sf, sizeof(sf), "shift/%sshift.txt", z.c_str());
            FILE* fp = std::fopen(sf, "r");
            if (fp) {
                float zz;
                if (std::fscanf(fp, "%e %e %e %e", &zz, shift, shift + 1, shift + 2) != 4) {
                    std::fprintf(stderr, "Error: bad shift file %s\n", sf);
                    std::fclose(fp);
                    return 1;
                }
                std::fclose(fp);
                std::printf("shift %g %g %g\n", shift[0], shift[1], shift[2]);
            } else {
                std::printf("No shift file\n");
            }
        }
        float* xv = nullptr;
        int64_t np = 0;
        if (read_xv_file) {
            char fname[512];
            std::snprintf(fname, sizeof(fname), "%s%d/%sxv%d.dat", node_dir.c_str(), inode, z.c_str(), inode);
            float hdr[11];
            rc = asm_read_xv(fname, &xv, &np, hdr);
            if (rc) {
                std::fprintf(stderr, "Unable to read xv file: %s\n", fname);
                asm_destroy(ctx);
                return 1;
            }
            for (int64_t i = 0; i < np; ++i)
                for (int a = 0; a < 3; ++a) xv[6 * i + a] -= shift[a];
        } else {
            float hdr[16];
            rc = asm_read_zip(node_dir.c_str(), z.c_str(), inode, nc_node_dim, mesh_scale, shift, &xv, &np, hdr);
            if (rc) {
                std::fprintf(stderr, "Unable to read zip files %s%d/%szip*_%d.dat (%d)\n", node_dir.c_str(), inode,
                             z.c_str(), inode, rc);
                asm_destroy(ctx);
                return 1;
            }
        }
        std::printf("np_local %lld\n", static_cast<long long>(np));
        const int64_t cap = static_cast<int64_t>(static_cast<double>(np) * (1.0 + 2.0 * bf) * (1.0 + 2.0 * bf) *
                                                 (1.0 + 2.0 * bf) * 1.3) + 4096;
        std::vector<asm_particle> parts(static_cast<size_t>(cap));
        int64_t np_local = 0;
        double t_comm0 = now_s();
        const int64_t nbuf = asm_build_buffered_set(xv, np, boxsize, bf, parts.data(), cap, &np_local);
        asm_free(xv);
        if (nbuf < 0) return die(ctx, "asm_build_buffered_set", static_cast<int>(nbuf));
        std::printf("np_with_buffers %lld\n", static_cast<long long>(nbuf));
        double t_path0 = now_s();
        if ((rc = asm_set_particles(ctx, parts.data(), nbuf, np_local, boxsize))) return die(ctx, "asm_set_particles", rc);
        int64_t nnode = 0;
        asm_synchronize(ctx);
        double t0 = now_s();
        if ((rc = asm_tree_construct(ctx, -bf * boxsize, (1.0f + bf) * boxsize, &nnode)))
            return die(ctx, "asm_tree_construct", rc);
        asm_synchronize(ctx);
        double t1 = now_s();
        std::printf("nnode %lld\n", static_cast<long long>(nnode));
        if ((rc = asm_neighbor_search(ctx, 32))) return die(ctx, "asm_neighbor_search", rc);
        asm_synchronize(ctx);
        double t2 = now_s();

        if (std::atoi(op["-halo_excise"].c_str())) {
            // read_and_exchage_halo (halo_file2.cpp:63-86) + the marking loop
            // (clumping_tree_main.cpp:452-511)
            char hf[512];
            std::snprintf(hf, sizeof(hf), "%s%d/%shalo%d.dat", node_dir.c_str(), inode, z.c_str(), inode);
            asm_halo* halos = nullptr;
            int64_t nh = 0;
            const float omegam = static_cast<float>(std::atof(op["-omegam"].c_str()));
            if ((rc = asm_read_halo_file(hf, boxsize, shift, static_cast<float>(std::atof(z.c_str())), omegam,
                                         std::atoi(op["-m200"].c_str()), std::atoi(op["-pid_flag"].c_str()),
                                         &halos, &nh))) {
                std::fprintf(stderr, "Unable to read halo data: %s (%d)\n", hf, rc);
                return 1;
            }
            std::vector<asm_halo> hbuf(static_cast<size_t>(8 * nh + 16));
            const int64_t nhb = asm_build_buffered_halos(halos, nh, boxsize, bf, hbuf.data(),
                                                         static_cast<int64_t>(hbuf.size()));
            asm_free(halos);
            if (nhb < 0) return die(ctx, "asm_build_buffered_halos", static_cast<int>(nhb));
            std::printf("nhalo_local= %lld\n", static_cast<long long>(nh));
            int64_t np_halo = 0;
            double m_halo = 0.0;
            if ((rc = asm_mark_halo_particles(ctx, hbuf.data(), nhb, -boxsize * bf, boxsize + boxsize * bf,
                                              &np_halo, &m_halo)))
                return die(ctx, "asm_mark_halo_particles", rc);
            std::printf("n_so_halo %lld %lld\n", static_cast<long long>(nh), static_cast<long long>(nhb));
            std::printf("so_halo_local %lld %g\n", static_cast<long long>(np_halo), m_halo);
            std::printf("check halo_mass %s %lld %g\n", z.c_str(), static_cast<long long>(np_halo), m_halo);
            // clumping_tree_main.cpp:507-509 asserts this (the run stops)
            if (m_halo > 100.0 && !(0.80 < np_halo / m_halo && np_halo / m_halo < 1.20)) {
                std::fprintf(stderr, "Error: halo particle count / mass ratio %g outside (0.8, 1.2)\n",
                             np_halo / m_halo);
                asm_destroy(ctx);
                return 1;
            }
        } else {
            std::printf("haloes not excised\n");
        }

        double t_mesh = 0.0, t_out = 0.0;
        for (int nc : ncs) {
            double a = now_s();
            if ((rc = asm_assign_on_mesh(ctx, nc, -1))) return die(ctx, "asm_assign_on_mesh", rc);
            double tot[2];
            if ((rc = asm_mesh_total(ctx, nc, tot))) return die(ctx, "asm_mesh_total", rc);
            // check_total_mass (clumping_tree_main.cpp:578-600): against
            // np_total, the node's particle count before the buffer was
            // trimmed (exchange_buffer.cpp:156-163); the reference asserts
            // (:597-598), so the run stops before this mesh is written.
            const double m_total = tot[0] + tot[1];
            std::printf("check total_mass %4d %15.1f %15.1f %lld\n", nc, tot[0], m_total, static_cast<long long>(np));
            const double ratio = m_total / static_cast<double>(np);
            if (!(0.999 < ratio && ratio < 1.001)) {
                std::fprintf(stderr, "Error: total mass %.1f deviates from np_total %lld by more than 1e-3 (nc %d)\n",
                             m_total, static_cast<long long>(np), nc);
                asm_destroy(ctx);
                return 1;
            }
            if ((rc = asm_normalize_mesh(ctx, nc))) return die(ctx, "asm_normalize_mesh", rc);
            asm_synchronize(ctx);
            double b = now_s();
            t_mesh += b - a;
            std::vector<float> mesh(static_cast<size_t>(6) * nc * nc * nc);
            if ((rc = asm_get_mesh(ctx, nc, mesh.data()))) return die(ctx, "asm_get_mesh", rc);
            char base[256];
            std::snprintf(base, sizeof(base), "so/nc%d/%s", nc, z.c_str());
            if ((rc = asm_write_mesh_separate(base, inode, mesh.data(), nc))) return die(ctx, "asm_write_mesh_separate", rc);
            t_out += now_s() - b;
        }
        const double total = now_s() - t_io0;
        std::printf("time %20s %8.2e\n", "io", t_comm0 - t_io0);
        std::printf("time %20s %8.2e\n", "communication", t_path0 - t_comm0);
        std::printf("time %20s %8.2e\n", "construction", t1 - t0);
        std::printf("time %20s %8.2e\n", "neighbors", t2 - t1);
        std::printf("time %20s %8.2e\n", "mesh_assign", t_mesh);
        std::printf("time %20s %8.2e\n", "output", t_out);
        std::printf("time %20s %10.4e\n", "total", total);
    }
    std::printf("boxsize %g\ndone all\n", boxsize);
    asm_destroy(ctx);
    return 0;
}
