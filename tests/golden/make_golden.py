"""Write tests/golden/reference_blobs10.npz from the reference's own functions.

Runs kdtree_balanced.cpp / nbr_search.cpp / coarse_mesh.cpp, compiled from
the reference sources into oracle/_ref/libasmref.so (oracle/Makefile), on a
small seeded set, single-threaded, and stores their outputs as fixtures:

    input:  buffered set (x, v) of blobs_uniform(n_side=10, L=20, seed=5), bf=0.25
    rk      reference smoothing length of every particle (neighbor_search)
    dens    reference density, one thread (FP32 sums in tree order)
    nbr     reference neighbour lists (33 slots, ascending) of 200 particles
    mesh    raw reference mesh for nc=10 (assign_on_mesh, sign -1) from the
            reference's own rk and dens
    norm    the same mesh after write_mesh_separate's normalisation

    python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from asmooth_amd import synth  # noqa: E402

N_SIDE, L, BF, SEED, NC = 10, 20.0, 0.25, 5, 10
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_blobs10.npz")


def main() -> int:
    xv = synth.blobs_uniform(n_side=N_SIDE, L=L, seed=SEED)
    buf, nl = O.buffered_set(xv, L, BF)
    pos, vel = buf[:, :3], buf[:, 3:]
    R = O.Ref(pos, vel, boxsize=L, np_local=nl)
    R.construct(-BF * L, (1 + BF) * L)
    nbr, _, _ = R.knn_lists()
    R.neighbor_search_all(1)
    rk, dens, _ = R.particles()
    mesh, tot = R.assign_on_mesh(NC, -1, nthreads=1)
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "g")
        O.ref_write_mesh_separate(base, 0, mesh, NC, L)
        m = NC ** 3
        parts = []
        for tag in ("n", "h", "c", "vel"):
            raw = open(f"{base}{tag}0.dat", "rb").read()
            parts.append(np.frombuffer(raw[4:], np.float32))
        norm = np.concatenate(parts)
        assert len(norm) == 6 * m
    sel = np.arange(0, len(pos), max(1, len(pos) // 200))[:200]
    np.savez_compressed(OUT, pos=pos, vel=vel, np_local=nl, boxsize=L, bf=BF, nc=NC, rk=rk, dens=dens,
                        nbr_index=sel, nbr=nbr[sel], mesh=mesh, mesh_total=tot, norm=norm)
    print(f"wrote {OUT}: {len(pos)} particles, {nl} local")
    return 0


if __name__ == "__main__":
    sys.exit(main())
