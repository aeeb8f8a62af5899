"""Write tests/golden/reference_mpi.npz from the reference's MPI-dependent
functions, compiled from the reference sources (oracle/Makefile, target
_ref/ref_mpi and _ref/ref_mpi_m200, built with MPICH's mpicxx) and run under
mpiexec in this container:

  exchange  8 ranks (a 2^3 node decomposition, nc_node_dim 2): each rank reads
            its node's xv file (read_pm_file_xv, file_reader2.cpp:24-86) and
            runs exchange_buffer (exchange_buffer.cpp:13-167); stored: every
            rank's records after the exchange, in order (x, v).
            Input: `ex_xv`, a global set in a periodic box of side 2 * EX_L,
            cut into node files in global order, node-local positions.
            The same on 1 rank (`ex1_*`, a single periodic rank, whose 26
            neighbours are itself).
  global    8 ranks: assign_on_global_mesh (global_mesh.cpp:25-123) of the
            same particle set with each rank's node coordinate; stored: every
            rank's raw mesh (6 nc^3 floats) and its coordinate.  `gm1_*`: the
            same for one particle whose kernel wraps on every axis.
  halo      1 rank, both halo layouts (default -DPID_FLAG and -DM200):
            read_and_exchage_halo (halo_file2.cpp:63-138 and
            exchange_buffer_template.h:18-122); stored: the halos after the
            exchange, in order.  The catalogue bytes are stored too.
  zip       1 rank: read_pm_file_zip (file_reader2.cpp:90-272) of a node
            with counts above 254 (escaped through zip3); stored: x, v.

Each input is seeded, small and stored with its outputs, so the tests read
only the fixture.

    python tests/golden/make_golden_mpi.py
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_MPI = os.path.join(ROOT, "oracle", "_ref", "ref_mpi")
REF_MPI_M200 = os.path.join(ROOT, "oracle", "_ref", "ref_mpi_m200")
MPIEXEC = "/opt/conda/bin/mpiexec"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_mpi.npz")

EX_L, EX_NN, EX_BF, EX_N = 16.0, 2, 0.125, 1500       # exchange: node side, nodes per axis
GM_L, GM_NC, GM_N = 16.0, 8, 400                      # global mesh
HALO_L, HALO_BF, HALO_Z, HALO_OM = 32.0, 0.125, 0.5, 0.308
ZIP_ND, ZIP_MS = 4, 4


def _mpirun(n, args):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    cmd = [MPIEXEC, "-n", str(n), *args]
    res = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    if res.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{res.stdout}")


def _read_out(path):
    raw = np.fromfile(path, np.int32)
    assert raw[0] == 0x61736D31, path
    inode, cx, cy, cz, count, width = (int(v) for v in raw[1:7])
    data = raw[7:].view(np.float32).reshape(count, width) if width > 1 else raw[7:].view(np.float32)
    return inode, (cx, cy, cz), data


def _write_xv(path, xv):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        np.array([len(xv)], np.int32).tofile(f)
        np.zeros(11, np.float32).tofile(f)
        np.asarray(xv, np.float32).tofile(f)


def exchange_case(d, nn=EX_NN):
    """A periodic box of side 2 EX_L: uniform particles, plus particles exactly
    on the region thresholds bf L and (1 - bf) L and on node faces, so the
    >= / > comparisons of exchange_buffer.cpp:6-11 matter."""
    rng = np.random.default_rng(11)
    box = nn * EX_L
    pos = rng.uniform(0.0, box, (EX_N // (1 if nn > 1 else 3), 3)).astype(np.float32)
    a = np.float32(EX_BF) * np.float32(EX_L)
    b = (np.float32(1.0) - np.float32(EX_BF)) * np.float32(EX_L)
    special = np.float32([0.0, a, b, EX_L - 1e-3, EX_L, EX_L + a, EX_L + b])[: 7 if nn > 1 else 4]
    extra = rng.choice(special, (200, 3)).astype(np.float32)
    mix = rng.uniform(0.0, box, (200, 3)).astype(np.float32)
    keep = rng.integers(0, 2, (200, 3)).astype(bool)
    extra = np.where(keep, extra, mix)
    pos = np.concatenate([pos, extra]).astype(np.float32)
    vel = rng.normal(0.0, 1.0, pos.shape).astype(np.float32)
    xv = np.concatenate([pos, vel], 1).astype(np.float32)
    # node files in global order, node-local positions (the same split as
    # asm_build_node_share: home = (int)(x / L) clipped, u = x - home * L)
    L = np.float32(EX_L)
    home = np.clip((pos / L).astype(np.int32), 0, nn - 1)
    u = (pos - home.astype(np.float32) * L).astype(np.float32)
    inode = home[:, 0] + nn * home[:, 1] + nn * nn * home[:, 2]
    d = os.path.join(d, f"ex{nn}")
    for k in range(nn ** 3):
        sel = inode == k
        _write_xv(os.path.join(d, f"node{k}", f"0.000xv{k}.dat"), np.concatenate([u[sel], vel[sel]], 1))
    out = os.path.join(d, "ex_out")
    os.makedirs(out)
    _mpirun(nn ** 3, [REF_MPI, "exchange", os.path.join(d, "node"), "0.000", str(EX_BF), str(4 * EX_N),
                      str(EX_L), out])
    shares = {}
    for k in range(nn ** 3):
        i, c, data = _read_out(os.path.join(out, f"exchange{k}.bin"))
        shares[i] = (c, np.ascontiguousarray(data[:, :6]), int((inode == i).sum()))
    return xv, shares


def global_case(d, single=False):
    """Particles of one node (side GM_L) with buffer particles outside [0, L)
    that the function must skip, halo and IGM particles, and corner and face
    particles whose kernels wrap, deposited at every node coordinate.  With
    single, one IGM particle near a corner whose kernel wraps on every axis."""
    rng = np.random.default_rng(12)
    L = GM_L
    n = 1 if single else GM_N
    pos = rng.uniform(-0.15 * L, 1.15 * L, (GM_N, 3)).astype(np.float32)
    corners = np.float32([[0.3, 0.2, 0.1], [15.8, 15.9, 15.7], [0.0, 8.0, 15.99], [15.99, 0.0, 4.0],
                          [8.0, 8.0, 8.0], [1e-3, 15.5, 0.5]])
    pos[: len(corners)] = corners
    rk = rng.uniform(0.4, 6.0, GM_N).astype(np.float32)
    rk[: len(corners)] = np.float32([3.0, 2.5, 5.5, 4.0, 7.5, 1.3])
    rec = np.zeros((GM_N, 9), np.float32)
    rec[:, :3] = pos
    rec[:, 3:6] = rng.normal(0.0, 100.0, (GM_N, 3)).astype(np.float32)
    rec[:, 6] = rng.lognormal(0.0, 1.0, GM_N).astype(np.float32)   # dens
    rec[:, 7] = rk
    igrp = rng.choice(np.int32([0, 0, -1, 1, 0]), GM_N).astype(np.int32)
    rec[:, 8] = igrp.view(np.float32)
    if single:
        rec = rec[:1].copy()
        rec[0, :3] = np.float32([15.7, 0.2, 15.9])
        rec[0, 7] = np.float32(3.3)
        rec[0, 8] = np.int32(0).view(np.float32)
        igrp = igrp[:1] * 0
    path = os.path.join(d, f"gm_particles{n}.bin")
    with open(path, "wb") as f:
        np.array([L], np.float32).tofile(f)
        np.array([n], np.int32).tofile(f)
        rec.tofile(f)
    out = os.path.join(d, f"gm_out{n}")
    os.makedirs(out)
    _mpirun(8, [REF_MPI, "global_mesh", path, str(GM_NC), "-1", out])
    meshes, coords = [], []
    for k in range(8):
        i, c, data = _read_out(os.path.join(out, f"global_mesh{k}.bin"))
        meshes.append(np.asarray(data, np.float32))
        coords.append(c)
    return rec, igrp, np.stack(meshes), np.int32(coords)


def _halo_delta_vir(z, om):
    a = 1.0 / (1.0 + z)
    ol = 1.0 - om
    oz = om / (om + ol * a ** 3)
    x = oz - 1.0
    return (18.0 * math.pi ** 2 + 82.0 * x - 39.0 * x * x) * ((om + ol * a ** 3) / om)


def halo_case(d, m200):
    """A catalogue of 22-float records with the 50-particle ID block after each
    (-DPID_FLAG), positions partly beyond the box (taken modulo) and near
    faces, at the overdensity the reader checks."""
    rng = np.random.default_rng(13 + int(m200))
    n = 40
    delta = 200.0 if m200 else _halo_delta_vir(HALO_Z, HALO_OM)
    off_m, off_r = (4, 6) if m200 else (3, 5)
    x = rng.uniform(0.0, HALO_L, (n, 3)).astype(np.float32)
    x[:10] = rng.uniform(HALO_L, 1.5 * HALO_L, (10, 3)).astype(np.float32)
    x[10:20, 0] = rng.choice(np.float32([0.5, 1.0, 4.0, 28.0, 31.0, 31.9]), 10)
    r = rng.uniform(0.3, 2.0, n).astype(np.float32)
    path = os.path.join(d, f"halo{int(m200)}.dat")
    with open(path, "wb") as f:
        np.array([n, 7, 9], np.int32).tofile(f)
        for i in range(n):
            rec = rng.normal(0.0, 1.0, 22).astype(np.float32)
            rec[:3] = x[i]
            rec[off_r] = r[i]
            rec[off_m] = np.float32(delta * 4.0 / 3.0 * math.pi * float(r[i]) ** 3)
            rec.tofile(f)
            np.full(50 * 32, 0xAB, np.uint8).tofile(f)   # particle IDs and xv, skipped
    out = os.path.join(d, f"halo_out{int(m200)}")
    os.makedirs(out)
    _mpirun(1, [REF_MPI_M200 if m200 else REF_MPI, "halo", path, str(HALO_L), str(HALO_Z), str(HALO_OM),
                str(HALO_BF), str(8 * n), out])
    _, _, data = _read_out(os.path.join(out, "halo0.bin"))
    return np.fromfile(path, np.uint8), np.ascontiguousarray(data)


def zip_case(d):
    """One node's zip files: Poisson counts per coarse cell plus cells of 255
    and more particles (escaped through zip3)."""
    rng = np.random.default_rng(14)
    nd = ZIP_ND
    counts = rng.poisson(3.0, nd ** 3).astype(np.int64)
    counts[[3, 17, 40]] = [255, 300, 254]
    n = int(counts.sum())
    b = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    s = rng.integers(-32768, 32768, (n, 3)).astype(np.int16)
    v_r2i = np.float32(37.5)
    h0 = np.zeros(16, np.float32)
    h0[12] = v_r2i
    h0 = h0.view(np.int32)
    h0[0] = n
    os.makedirs(os.path.join(d, "znode0"))
    base = os.path.join(d, "znode0", "7.221zip")
    files = {}
    files[0] = np.concatenate([h0.view(np.uint8), b.ravel()])
    files[1] = np.concatenate([np.arange(16, dtype=np.float32).view(np.uint8), s.ravel().view(np.uint8)])
    files[2] = np.minimum(counts, 255).astype(np.uint8)
    files[3] = counts[counts >= 255].astype(np.int32).view(np.uint8)
    for k, v in files.items():
        v.tofile(f"{base}{k}_0.dat")
    out = os.path.join(d, "zip_out")
    os.makedirs(out)
    _mpirun(1, [REF_MPI, "zip", os.path.join(d, "znode"), "7.221", str(nd), str(ZIP_MS), str(2 * n), out])
    _, _, data = _read_out(os.path.join(out, "zip0.bin"))
    return files, np.ascontiguousarray(data)


def main() -> int:
    for exe in (REF_MPI, REF_MPI_M200, MPIEXEC):
        if not os.path.exists(exe):
            print(f"{exe} missing: build it with `make -C oracle ref` where the reference tree is")
            return 1
    arrays = {}
    with tempfile.TemporaryDirectory() as d:
        xv1, shares1 = exchange_case(d, 1)
        arrays.update(ex1_xv=xv1, ex1_share0=shares1[0][1])
        xv, shares = exchange_case(d)
        arrays.update(ex_xv=xv, ex_L=np.float32(EX_L), ex_nn=np.int32(EX_NN), ex_bf=np.float32(EX_BF))
        for k, (c, data, nlocal) in shares.items():
            arrays[f"ex_share{k}"] = data
            arrays[f"ex_coord{k}"] = np.int32(c)
            arrays[f"ex_nlocal{k}"] = np.int64(nlocal)
        rec, igrp, meshes, coords = global_case(d)
        arrays.update(gm_rec=rec, gm_igrp=igrp, gm_mesh=meshes, gm_coord=coords, gm_L=np.float32(GM_L),
                      gm_nc=np.int32(GM_NC), gm_sign=np.int32(-1))
        rec1, _, meshes1, coords1 = global_case(d, single=True)
        arrays.update(gm1_rec=rec1, gm1_mesh=meshes1, gm1_coord=coords1)
        for m200 in (0, 1):
            raw, halos = halo_case(d, bool(m200))
            arrays[f"halo_file{m200}"] = raw
            arrays[f"halo_out{m200}"] = halos
        arrays.update(halo_L=np.float32(HALO_L), halo_bf=np.float32(HALO_BF), halo_z=np.float32(HALO_Z),
                      halo_om=np.float32(HALO_OM))
        files, zxv = zip_case(d)
        for k, v in files.items():
            arrays[f"zip_file{k}"] = v
        arrays.update(zip_xv=zxv, zip_nd=np.int32(ZIP_ND), zip_ms=np.int32(ZIP_MS))
    np.savez_compressed(OUT, **arrays)
    sizes = {k: len(v[1]) for k, v in shares.items()}
    print(f"wrote {OUT}: exchange shares {sizes}, global meshes {meshes.shape}, zip {len(zxv)} particles")
    return 0


if __name__ == "__main__":
    sys.exit(main())
