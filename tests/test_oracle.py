"""Pins of the CPU oracle (no GPU): against the compiled reference functions
(oracle/_ref, built from the reference sources), against library routines
(scipy cKDTree, numerical quadrature), brute force, and closed forms.  A
plausible slip in the oracle (a dropped term, a wrong sign or index, a
transposed operand) fails one of these."""
import numpy as np
import pytest

import asmooth_amd as A
from asmooth_amd import synth
import oracle as O
from tests import helpers
from tests.helpers import mesh_fields, small_set

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference tree absent: oracle/_ref not built")


def _buf(kind="blobs", n_side=12, seed=0, bf=0.05):
    xv, L, bf = small_set(kind, n_side, seed, bf)
    b, nl = O.buffered_set(xv, L, bf)
    return b, nl, L, bf


# ------------------------------------------------------------------ kNN ---

@pytest.mark.parametrize("kind,seed", [("blobs", 0), ("lattice", 1), ("uniform", 2)])
def test_grid_knn_equals_brute_force(kind, seed):
    b, _, _, _ = _buf(kind, 10, seed)
    g = O.knn(b[:, :3])
    f = O.knn(b[:, :3], brute=True)
    for x, y in zip(g, f):
        assert np.array_equal(x, y)


def test_knn_handles_duplicates_by_index():
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 4, (60, 3)).astype(np.float32)
    pos[10:50] = pos[3]          # 41 coincident particles: the 33 nearest are all at d2 = 0
    nbr, d2, rk = O.knn(pos, brute=True)
    assert rk[3] == 0.0
    dup = np.r_[3, np.arange(10, 50)]
    assert list(nbr[3]) == sorted(dup)[:33]


def test_knn_vs_scipy_ckdtree():
    from scipy.spatial import cKDTree
    b, _, _, _ = _buf("blobs", 14, 5)
    pos = b[:, :3]
    nbr, d2, rk = O.knn(pos)
    dd, ii = cKDTree(pos.astype(np.float64)).query(pos.astype(np.float64), k=33)
    # h agrees to FP32 rounding of the distance
    assert np.max(np.abs(rk - dd[:, -1]) / dd[:, -1]) < 1e-6
    # sets agree except where the FP64 and FP32 metrics order near-ties differently
    same = np.array([set(a) == set(c) for a, c in zip(nbr, ii)])
    assert same.mean() > 0.999
    for i in np.where(~same)[0]:
        gap = abs(dd[i, -1] - np.sqrt(np.float64(d2[i, -1])))
        assert gap < 1e-5 * dd[i, -1]


@needs_ref
@pytest.mark.parametrize("kind,seed", [("blobs", 0), ("zeldovich", 3), ("uniform", 4)])
def test_knn_vs_reference(kind, seed):
    b, nl, L, bf = _buf(kind, 14, seed)
    pos = b[:, :3]
    nbr, d2, rk = O.knn(pos)
    R = O.Ref(pos, b[:, 3:], boxsize=L, np_local=nl)
    R.construct(-bf * L, (1 + bf) * L)
    rn, rd2, rrk = R.knn_lists()
    # The reference keeps a candidate only if d2 < fl(sqrt(K))^2 (nbr_search.cpp:94,103);
    # that can differ from the exact rule on near-ties.  Count them; they are rare.
    diff_h = np.sum(rrk != rk)
    assert diff_h <= max(2, len(rk) // 20000)
    same = np.array([set(a) == set(c) for a, c in zip(rn, nbr)])
    assert same.mean() >= 0.9999
    # where sets agree, the sorted distances agree exactly
    assert np.array_equal(np.sort(rd2[same], axis=1), np.sort(d2[same], axis=1))


# -------------------------------------------------------------- density ---

def test_tent_closed_forms():
    x = np.array([1.0, 2.0, 3.0], np.float32)
    for hinv in (0.25, 0.3, 1.7):
        # W(0) = hinv^3
        h = np.float32(hinv)
        assert O.tent(x, x, hinv) == (h * h) * h
    # separable: W(dx,0,0) = hinv^3 (1 - |dx| hinv)
    y = x + np.array([0.5, 0, 0], np.float32)
    assert O.tent(x, y, 0.5) == pytest.approx(0.125 * 0.75, rel=1e-7)
    # unit normalisation of the tent: hinv^3 * integral over [-h,h]^3 = 1
    h = 2.0
    g = (np.arange(400) + 0.5) / 400 * 2 * h - h
    w1 = np.sum(1.0 - np.abs(g) / h) * (2 * h / 400)
    assert (1 / h) ** 3 * w1 ** 3 == pytest.approx(1.0, rel=1e-4)


@needs_ref
def test_density_vs_reference_single_thread():
    b, nl, L, bf = _buf("blobs", 12, 6)
    pos = b[:, :3]
    nbr, _, rk = O.knn(pos)
    d64 = O.density(pos, nbr, rk)
    R = O.Ref(pos, b[:, 3:], boxsize=L, np_local=nl)
    R.construct(-bf * L, (1 + bf) * L)
    R.neighbor_search_all(1)
    rrk, rdens, _ = R.particles()
    ok = rrk == rk
    assert ok.mean() > 0.999
    # FP32 sequential sums (reference) vs FP64 sums of the same FP32 terms
    rel = np.abs(rdens - d64) / d64
    assert np.median(rel) < 2e-7 and np.max(rel[ok]) < 1e-5
    # every particle gets at least its own term hinv^3
    assert np.all(d64 >= (1.0 / rk.astype(np.float64)) ** 3 * (1 - 1e-6))


# ------------------------------------------------------------ deposition ---

def _F(t):
    t = np.asarray(t, np.float64)
    a = np.minimum(np.abs(t), 1.0)
    return np.sign(t) * (a - 0.5 * a * a)


def test_ker_integ_closed_form_and_quadrature():
    from scipy.integrate import quad
    rng = np.random.default_rng(0)
    for _ in range(300):
        x1 = np.float32(rng.uniform(-2, 2))
        x2 = np.float32(x1 + rng.uniform(0, 2))
        got = O.ker_integ(x1, x2)
        q, _ = quad(lambda t: max(0.0, 1.0 - abs(t)), float(x1), float(x2), points=[-1, 0, 1])
        assert abs(got - q) < 2e-7 + 1e-6 * abs(q)
        assert abs(got - (_F(x2) - _F(x1))) < 2e-7
    # sgn(0) = -1 branch (coarse_mesh.cpp:19-20) still integrates correctly
    assert O.ker_integ(0.0, 0.5) == pytest.approx(0.375, abs=1e-7)
    assert O.ker_integ(-0.5, 0.0) == pytest.approx(0.375, abs=1e-7)


def test_single_particle_is_outer_product_and_partition_of_unity():
    L, nc = 16.0, 16
    pos = np.array([[7.3, 8.1, 6.6]], np.float32)
    vel = np.array([[1.0, -2.0, 0.5]], np.float32)
    h = np.array([2.7], np.float32)
    m = O.assign_on_mesh(pos, vel, np.array([3.0], np.float32), h, np.zeros(1, np.int32), L, nc, -1)
    f = mesh_fields(m, nc)
    n = f["n"].reshape(nc, nc, nc)       # (z, y, x)
    dx = L / nc
    edges = np.arange(nc + 1) * dx
    w = [(_F((edges[1:] - pos[0, a]) / h[0]) - _F((edges[:-1] - pos[0, a]) / h[0])) for a in range(3)]
    want = w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    assert np.max(np.abs(n - want)) < 1e-6
    assert n.sum() == pytest.approx(1.0, abs=1e-6)
    assert f["c"].sum() == pytest.approx(3.0, abs=1e-5)
    v = f["v"].reshape(-1, 3).sum(0)
    assert np.allclose(v, vel[0], atol=1e-5)
    assert f["halo"].sum() == 0.0


def test_halo_particles_go_to_the_halo_field():
    L, nc = 16.0, 8
    pos = np.array([[7.3, 8.1, 6.6]], np.float32)
    m = O.assign_on_mesh(pos, np.ones((1, 3), np.float32), np.ones(1, np.float32), np.array([2.0], np.float32),
                         np.array([-1], np.int32), L, nc, -1)
    f = mesh_fields(m, nc)
    assert f["n"].sum() == 0 and f["halo"].sum() == pytest.approx(1.0, abs=1e-6)
    # sign = +1 flips the selection (coarse_mesh.cpp:95,117)
    m = O.assign_on_mesh(pos, np.ones((1, 3), np.float32), np.ones(1, np.float32), np.array([2.0], np.float32),
                         np.array([-1], np.int32), L, nc, 1)
    assert mesh_fields(m, nc)["n"].sum() == pytest.approx(1.0, abs=1e-6)


def test_exact_lattice_gives_h_2a_and_unit_density():
    # spacing a = 2: shells 1+6+12+8+6 = 33 points within 2a, so h = 2a = 4
    # exactly, and the tent of half-width 2a sums to a constant over a lattice:
    # every cell of side a holds exactly one particle's worth of weight.
    xv, L, bf = small_set("lattice", 10, 0, bf=0.5)
    b, nl = O.buffered_set(xv, L, bf)
    nbr, d2, rk = O.knn(b[:, :3])
    assert np.all(rk[:nl] == 4.0)
    dens = O.density(b[:, :3], nbr, rk)
    m = O.assign_on_mesh(b[:, :3], b[:, 3:], dens.astype(np.float32), rk, np.zeros(len(b), np.int32), L, 10, -1)
    n = mesh_fields(m, 10)["n"]
    assert np.max(np.abs(n - 1.0)) < 1e-6
    # interior density is the same for every local particle (a lattice is homogeneous)
    assert np.ptp(dens[:nl]) < 1e-9 * dens[:nl].mean()


@needs_ref
def test_mesh_vs_reference_single_thread():
    # buffer (11.2) wider than every h plus the reach of the truncated
    # neighbourhoods at its edge: periodic images then conserve mass exactly
    b, nl, L, bf = _buf("zeldovich", 16, 2, bf=0.35)
    pos, vel = b[:, :3], b[:, 3:]
    nbr, _, rk = O.knn(pos)
    dens = O.density(pos, nbr, rk).astype(np.float32)
    R = O.Ref(pos, vel, boxsize=L, np_local=nl)
    R.construct(-bf * L, (1 + bf) * L)
    R.set_dens(dens, rk)
    for nc in (8, 16, 32):
        rm, rtot = R.assign_on_mesh(nc, -1, nthreads=1)
        om = O.assign_on_mesh(pos, vel, dens, rk, np.zeros(len(pos), np.int32), L, nc, -1)
        fr, fo = mesh_fields(rm, nc), mesh_fields(om, nc)
        assert np.max(np.abs(fr["n"] - fo["n"]) / np.maximum(fo["n"], 1e-6)) < 1e-5
        assert np.max(np.abs(fr["c"] - fo["c"]) / np.maximum(fo["c"], 1e-6)) < 1e-5
        assert np.max(np.abs(fr["v"] - fo["v"])) < 1e-4 * np.max(np.abs(fo["v"]))
        # mesh_total and the reference's mass check (clumping_tree_main.cpp:597)
        on, oh = O.mesh_total(om, nc)
        assert rtot[0] == pytest.approx(on, rel=1e-6)
        assert 0.999 < (on + oh) / nl < 1.001
        assert abs((on + oh) / nl - 1.0) < 1e-6


@needs_ref
def test_normalisation_vs_reference_writer(tmp_path):
    b, nl, L, bf = _buf("blobs", 8, 3)
    nbr, _, rk = O.knn(b[:, :3])
    dens = O.density(b[:, :3], nbr, rk).astype(np.float32)
    nc = 8
    raw = O.assign_on_mesh(b[:, :3], b[:, 3:], dens, rk, np.zeros(len(b), np.int32), L, nc, -1).astype(np.float32)
    base = str(tmp_path / "0.000")
    O.ref_write_mesh_separate(base, 0, raw, nc, L)
    ours = O.normalize_f32(raw, nc, L)
    m = nc ** 3
    for tag, lo, cnt in (("n", 0, m), ("h", m, m), ("c", 2 * m, m), ("vel", 3 * m, 3 * m)):
        data = open(f"{base}{tag}0.dat", "rb").read()
        assert np.frombuffer(data[:4], np.int32)[0] == nc
        assert np.array_equal(np.frombuffer(data[4:], np.float32), ours[lo:lo + cnt])
    # and the product writer emits byte-identical files
    base2 = str(tmp_path / "p0.000")
    A.write_mesh_separate(base2, 0, ours, nc)
    for tag in ("n", "h", "c", "vel"):
        assert open(f"{base}{tag}0.dat", "rb").read() == open(f"{base2}{tag}0.dat", "rb").read()
    # FP64 normalisation agrees with FP32 to rounding
    o64 = O.normalize_f64(raw.astype(np.float64), nc, L)
    assert np.allclose(ours, o64, rtol=1e-6, atol=1e-6)


@needs_ref
def test_mark_halo_vs_reference():
    b, nl, L, bf = _buf("blobs", 12, 8)
    pos = b[:, :3]
    rng = np.random.default_rng(2)
    halos = np.zeros((30, 5), np.float32)
    halos[:, :3] = rng.uniform(-2, L + 2, (30, 3))
    halos[:, 3] = rng.uniform(0.5, 4.0, 30)
    halos[:, 4] = rng.uniform(10, 500, 30)
    g, nph, mh = O.mark_halo(pos, halos, -bf * L, L + bf * L)
    R = O.Ref(pos, b[:, 3:], boxsize=L, np_local=nl)
    R.construct(-bf * L, (1 + bf) * L)
    rnph, rmh = R.mark_halo(halos, -bf * L, L + bf * L)
    _, _, rg = R.particles()
    assert np.array_equal(rg, g)
    assert rnph == nph and rmh == pytest.approx(mh, rel=1e-12)
    assert nph > 0


def test_global_mesh_equals_node_mesh_for_interior_particles():
    L, nc = 16.0, 8
    rng = np.random.default_rng(4)
    pos = rng.uniform(5, 11, (50, 3)).astype(np.float32)     # supports stay inside [0, L)
    vel = rng.standard_normal((50, 3)).astype(np.float32)
    h = rng.uniform(0.5, 3.0, 50).astype(np.float32)
    d = rng.uniform(0.1, 1, 50).astype(np.float32)
    g = np.zeros(50, np.int32)
    a = O.assign_on_mesh(pos, vel, d, h, g, L, nc, -1)
    b = O.assign_on_mesh(pos, vel, d, h, g, L, nc, -1, periodic_coord=(0, 0, 0))
    assert np.allclose(a, b, rtol=0, atol=1e-12)
    # near the edge the periodic variant wraps: full weight stays on the mesh
    pos2 = np.array([[0.3, 15.8, 8.0]], np.float32)
    w = O.assign_on_mesh(pos2, vel[:1], d[:1], np.array([2.5], np.float32), g[:1], L, nc, -1, periodic_coord=(0, 0, 0))
    assert mesh_fields(w, nc)["n"].sum() == pytest.approx(1.0, abs=1e-6)


# ---------------------------------------------------------- buffered set ---

@pytest.mark.parametrize("kind,bf", [("blobs", 0.05), ("uniform", 0.1), ("lattice", 0.2)])
def test_product_buffered_set_equals_oracle(kind, bf):
    xv, L, _ = small_set(kind, 12, 3, bf)
    ob, onl = O.buffered_set(xv, L, bf)
    pb, pnl = A.build_buffered_set(xv, L, bf)
    assert onl == pnl == len(xv)
    assert np.array_equal(np.concatenate([pb["x"], pb["v"]], 1), ob)
    # every particle in the buffer layer is a periodic image of a local one
    lo, hi = -bf * L, L + bf * L
    assert np.all((ob[:, :3] > lo) & (ob[:, :3] < hi))
    assert np.array_equal(ob[:len(xv)], xv)


def test_buffered_count_matches_closed_form():
    xv, L, bf = small_set("uniform", 32, 0, 0.05)
    b, nl = O.buffered_set(xv, L, bf)
    # uniform: N_buf / N = (1 + 2 bf)^3 = 1.331 (exchange_buffer.cpp:166-167)
    assert len(b) / nl == pytest.approx(1.331, rel=5e-3)


def test_node_share_single_node_is_the_buffered_set():
    xv, L, bf = small_set("blobs", 10, 1)
    a, na = A.build_node_share(xv, L, 1, (0, 0, 0), bf)
    b, nb = A.build_buffered_set(xv, L, bf)
    assert na == nb and len(a) == len(b)
    key = lambda p: np.lexsort((p["x"][:, 2], p["x"][:, 1], p["x"][:, 0]))
    assert np.array_equal(a["x"][key(a)], b["x"][key(b)])


def test_node_shares_partition_the_local_particles():
    xv, L, bf = small_set("uniform", 16, 2)
    tot = 0
    for c in [(i, j, k) for i in range(2) for j in range(2) for k in range(2)]:
        s, nl = A.build_node_share(xv, L, 2, c, bf)
        assert np.all((s["x"][:nl] >= 0) & (s["x"][:nl] < L / 2))
        tot += nl
    assert tot == len(xv)


@pytest.mark.parametrize("centre", [(24.0, 40.0, 60.0), (1.0, 2.0, 94.0)])
def test_region_oracle_equals_whole_set_oracle(centre):
    # The sampled-region check used at full size gives, for its particles and
    # cells, exactly what the oracle gives on the whole set.
    b, nl, L, bf = _buf("zeldovich", 48, 5)
    ncs = [16, 48]
    pos, vel = b[:, :3], b[:, 3:6]
    nbr, _, rk = O.knn(pos)
    dens64 = O.density(pos, nbr, rk)
    hmax = float(rk.max())
    r = helpers.region_oracle(b, L, np.array(centre), half=8.0, hmax=hmax, ncs=ncs)
    i = r["index"]
    assert 0 < len(i) < len(b)
    assert np.array_equal(r["rk"].view(np.int32), rk[i].view(np.int32))
    assert np.array_equal(r["nbr"], nbr[i])
    assert np.array_equal(r["dens64"], dens64[i])
    dens32 = dens64.astype(np.float32)
    for nc in ncs:
        cells, o6 = r["cells"][nc]
        assert len(cells) > 0
        whole = O.assign_on_mesh(pos, vel, dens32, rk, np.zeros(len(b), np.int32), L, nc, -1)
        assert np.array_equal(o6, helpers.mesh_at(whole, nc, cells))


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (reference tree absent)")
def test_reference_disagreements_on_the_64cubed_blob_set():
    """The 64^3 blob set whole (349 k buffered particles): the reference's own
    search (kdtree_balanced.cpp + nbr_search.cpp, compiled here) against the
    oracle's canonical kNN.  Its d2 < fl(sqrt(K))^2 filter
    (nbr_search.cpp:94,103) changes no h and no neighbour set here; only the
    order of equal distances differs (discovery order, k_neighbors.h:26-28),
    and every such row holds the same distances slot by slot."""
    L, bf = 128.0, 0.05
    xv = synth.blobs_uniform(n_side=64, L=L, seed=0)
    buf, nl = O.buffered_set(xv, L, bf)
    pos, vel = buf[:, :3], buf[:, 3:]
    onbr, od2, ork = O.knn(pos)
    R = O.Ref(pos, vel, boxsize=L, np_local=nl)
    try:
        R.construct(-bf * L, (1 + bf) * L)
        rnbr, rd2, rrk = R.knn_lists(nthreads=0)
    finally:
        R.close()
    d = O.compare_with_reference(onbr, ork, rnbr, rrk)
    assert d["particles"] == len(pos)
    assert d["h"] == 0 and d["set33"] == 0 and d["set32"] == 0, d
    rows = np.nonzero(np.any(onbr != rnbr, axis=1))[0]
    assert len(rows) == d["order"] and d["order"] < 100
    assert np.array_equal(od2[rows].view(np.int32), rd2[rows].view(np.int32))
