"""ctypes bindings of the CPU oracle (liboracle.so) and of the compiled
reference harness (_ref/libasmref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libasmref.so")
K_SLOTS = 33

_o = None
_r = None


def _p(a):
    return None if a is None else a.ctypes.data


def olib() -> C.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_LIB)
        vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
        P = C.POINTER
        L.orc_buffered_set.restype = i64
        L.orc_buffered_set.argtypes = [vp, i64, f32, f32, vp, i64, P(i64)]
        for f in (L.orc_knn, L.orc_knn_brute):
            f.restype = i32
            f.argtypes = [vp, i64, vp, i64, i32, vp, vp, vp, i32]
        L.orc_norm2.restype = f32
        L.orc_norm2.argtypes = [vp, vp]
        L.orc_tent.restype = f32
        L.orc_tent.argtypes = [vp, vp, f32]
        L.orc_density.restype = None
        L.orc_density.argtypes = [vp, i64, vp, i32, vp, vp]
        L.orc_ker_integ.restype = f32
        L.orc_ker_integ.argtypes = [f32, f32]
        L.orc_spline.restype = f32
        L.orc_spline.argtypes = [vp, vp, f32]
        L.orc_spline_integ.restype = f32
        L.orc_spline_integ.argtypes = [f32, f32]
        L.orc_set_kernel.restype = None
        L.orc_set_kernel.argtypes = [i32]
        L.orc_assign_on_mesh.restype = i32
        L.orc_assign_on_mesh.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, vp]
        L.orc_assign_on_global_mesh.restype = i32
        L.orc_assign_on_global_mesh.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, vp, vp]
        L.orc_mesh_total_f32.restype = None
        L.orc_mesh_total_f32.argtypes = [vp, i32, vp]
        L.orc_mesh_total_f64.restype = None
        L.orc_mesh_total_f64.argtypes = [vp, i32, vp]
        L.orc_normalize_f32.restype = None
        L.orc_normalize_f32.argtypes = [vp, i32, f32]
        L.orc_normalize_f64.restype = None
        L.orc_normalize_f64.argtypes = [vp, i32, C.c_double]
        L.orc_mark_halo.restype = None
        L.orc_mark_halo.argtypes = [vp, i64, vp, i64, f32, f32, vp, P(i64), P(C.c_double)]
        _o = L
    return _o


def have_ref() -> bool:
    return os.path.exists(REF_LIB)


def rlib() -> C.CDLL:
    global _r
    if _r is None:
        if not have_ref():
            raise RuntimeError(f"{REF_LIB} missing (built only where the reference tree exists)")
        L = C.CDLL(REF_LIB)
        vp, i32, i64, f32, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
        P = C.POINTER
        L.ref_create.restype = vp
        L.ref_create.argtypes = [vp, vp, vp, i32, i32, f32]
        L.ref_destroy.restype = None
        L.ref_destroy.argtypes = [vp]
        L.ref_construct.restype = i32
        L.ref_construct.argtypes = [vp, f32, f32, P(f64)]
        L.ref_neighbor_search_all.restype = i32
        L.ref_neighbor_search_all.argtypes = [vp, i32, P(f64)]
        L.ref_knn_lists.restype = i32
        L.ref_knn_lists.argtypes = [vp, vp, vp, vp, i32]
        L.ref_get_particles.restype = i32
        L.ref_get_particles.argtypes = [vp, vp, vp, vp]
        L.ref_set_dens.restype = i32
        L.ref_set_dens.argtypes = [vp, vp, vp]
        L.ref_assign_on_mesh.restype = i32
        L.ref_assign_on_mesh.argtypes = [vp, vp, i32, i32, i32, vp, P(f64)]
        L.ref_mark_halo.restype = i32
        L.ref_mark_halo.argtypes = [vp, vp, i32, f32, f32, i32, P(i64), P(f64)]
        L.ref_write_mesh_separate.restype = i32
        L.ref_write_mesh_separate.argtypes = [C.c_char_p, i32, vp, i32, f32]
        L.ref_time_path.restype = i32
        L.ref_time_path.argtypes = [vp, vp, i32, i32, f32, f32, vp, i32, i32, vp, vp]
        _r = L
    return _r


# ---------------------------------------------------------------- oracle ---

def buffered_set(xv: np.ndarray, L: float, bf: float):
    xv = np.ascontiguousarray(xv, np.float32).reshape(-1, 6)
    cap = int(len(xv) * (1 + 2 * bf) ** 3 * 1.3) + 4096
    out = np.empty((cap, 6), np.float32)
    nl = C.c_int64()
    n = olib().orc_buffered_set(_p(xv), len(xv), float(L), float(bf), _p(out), cap, C.byref(nl))
    if n < 0:
        raise RuntimeError("orc_buffered_set: capacity")
    return out[:n].copy(), nl.value


def knn(pos: np.ndarray, queries=None, k_slots: int = K_SLOTS, brute: bool = False, nthreads: int = 0):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    q = None if queries is None else np.ascontiguousarray(queries, np.int64)
    nq = len(pos) if q is None else len(q)
    nbr = np.empty((nq, k_slots), np.int32)
    d2 = np.empty((nq, k_slots), np.float32)
    rk = np.empty(nq, np.float32)
    f = olib().orc_knn_brute if brute else olib().orc_knn
    rc = f(_p(pos), len(pos), _p(q), nq, k_slots, _p(nbr), _p(d2), _p(rk), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"orc_knn failed ({rc})")
    return nbr, d2, rk


class kernel:
    """Context manager selecting the oracle's smoothing kernel for density()
    and assign_on_mesh(): 0 the tent (default), 1 the cubic spline."""

    def __init__(self, k: int):
        self.k = int(k)

    def __enter__(self):
        olib().orc_set_kernel(self.k)
        return self

    def __exit__(self, *exc):
        olib().orc_set_kernel(0)


def density(pos: np.ndarray, nbr: np.ndarray, rk: np.ndarray, kern: int = 0) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    nbr = np.ascontiguousarray(nbr, np.int32)
    rk = np.ascontiguousarray(rk, np.float32)
    d = np.zeros(len(pos), np.float64)
    with kernel(kern):
        olib().orc_density(_p(pos), len(pos), _p(nbr), nbr.shape[1], _p(rk), _p(d))
    return d


def spline(x, y, hinv) -> float:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return olib().orc_spline(_p(x), _p(y), float(hinv))


def spline_integ(x1: float, x2: float) -> float:
    return olib().orc_spline_integ(float(x1), float(x2))


_rs = None


def ref_spline_ker(x, y, hinv) -> float:
    """The reference's own spline_ker (nbr_search.cpp:39-53), compiled from its
    source into _ref/libasmref_spline.so (where the reference tree exists)."""
    global _rs
    if _rs is None:
        path = os.path.join(HERE, "_ref", "libasmref_spline.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing (built only where the reference tree exists)")
        _rs = C.CDLL(path)
        _rs.ref_spline_ker.restype = C.c_float
        _rs.ref_spline_ker.argtypes = [C.c_void_p, C.c_void_p, C.c_float]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return _rs.ref_spline_ker(_p(x), _p(y), float(hinv))


def have_ref_spline() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libasmref_spline.so"))


def ker_integ(x1: float, x2: float) -> float:
    return olib().orc_ker_integ(float(x1), float(x2))


def tent(x, y, hinv) -> float:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return olib().orc_tent(_p(x), _p(y), float(hinv))


def norm2(x, y) -> float:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return olib().orc_norm2(_p(x), _p(y))


def assign_on_mesh(pos, vel, dens, rk, igrp, L: float, nc: int, sign: int = -1, periodic_coord=None, kern: int = 0):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    dens = np.ascontiguousarray(dens, np.float32)
    rk = np.ascontiguousarray(rk, np.float32)
    igrp = np.ascontiguousarray(igrp, np.int32)
    mesh = np.zeros(6 * nc ** 3, np.float64)
    with kernel(kern):
        if periodic_coord is None:
            rc = olib().orc_assign_on_mesh(_p(pos), _p(vel), _p(dens), _p(rk), _p(igrp), len(pos), float(L),
                                           int(nc), int(sign), _p(mesh))
        else:
            c3 = np.ascontiguousarray(periodic_coord, np.int32)
            rc = olib().orc_assign_on_global_mesh(_p(pos), _p(vel), _p(dens), _p(rk), _p(igrp), len(pos),
                                                  float(L), int(nc), int(sign), _p(c3), _p(mesh))
    if rc != 0:
        raise RuntimeError(f"orc_assign_on_mesh failed ({rc})")
    return mesh


def normalize_f32(mesh: np.ndarray, nc: int, L: float) -> np.ndarray:
    m = np.ascontiguousarray(mesh, np.float32).copy()
    olib().orc_normalize_f32(_p(m), int(nc), float(L))
    return m


def normalize_f64(mesh: np.ndarray, nc: int, L: float) -> np.ndarray:
    m = np.ascontiguousarray(mesh, np.float64).copy()
    olib().orc_normalize_f64(_p(m), int(nc), float(L))
    return m


def mesh_total(mesh: np.ndarray, nc: int):
    out = np.zeros(2, np.float64)
    if mesh.dtype == np.float32:
        olib().orc_mesh_total_f32(_p(np.ascontiguousarray(mesh)), int(nc), _p(out))
    else:
        olib().orc_mesh_total_f64(_p(np.ascontiguousarray(mesh, np.float64)), int(nc), _p(out))
    return float(out[0]), float(out[1])


def mark_halo(pos, halos, left: float, right: float, igrp=None):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    h = np.ascontiguousarray(halos, np.float32).reshape(-1, 5)
    g = np.zeros(len(pos), np.int32) if igrp is None else np.ascontiguousarray(igrp, np.int32).copy()
    nph, mh = C.c_int64(), C.c_double()
    olib().orc_mark_halo(_p(pos), len(pos), _p(h), len(h), float(left), float(right), _p(g), C.byref(nph),
                         C.byref(mh))
    return g, nph.value, mh.value


# ------------------------------------------------------------- reference ---

class Ref:
    """The reference's own functions on one particle set (single rank)."""

    def __init__(self, pos, vel=None, igrp=None, np_local=None, boxsize: float = 1.0):
        self.L = rlib()
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self.igrp = None if igrp is None else np.ascontiguousarray(igrp, np.int32)
        n = len(self.pos)
        self.h = self.L.ref_create(_p(self.pos), _p(self.vel), _p(self.igrp), n,
                                   n if np_local is None else int(np_local), float(boxsize))
        self.n = n

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def construct(self, left: float, right: float) -> int:
        t = C.c_double()
        return self.L.ref_construct(self.h, float(left), float(right), C.byref(t))

    def neighbor_search_all(self, nthreads: int = 1) -> float:
        t = C.c_double()
        self.L.ref_neighbor_search_all(self.h, int(nthreads), C.byref(t))
        return t.value

    def knn_lists(self, nthreads: int = 1):
        nbr = np.empty((self.n, K_SLOTS), np.int32)
        d2 = np.empty((self.n, K_SLOTS), np.float32)
        rk = np.empty(self.n, np.float32)
        self.L.ref_knn_lists(self.h, _p(nbr), _p(d2), _p(rk), int(nthreads))
        return nbr, d2, rk

    def particles(self):
        rk = np.empty(self.n, np.float32)
        dens = np.empty(self.n, np.float32)
        igrp = np.empty(self.n, np.int32)
        self.L.ref_get_particles(self.h, _p(rk), _p(dens), _p(igrp))
        return rk, dens, igrp

    def set_dens(self, dens=None, rk=None):
        d = None if dens is None else np.ascontiguousarray(dens, np.float32)
        r = None if rk is None else np.ascontiguousarray(rk, np.float32)
        self.L.ref_set_dens(self.h, _p(d), _p(r))

    def assign_on_mesh(self, nc: int, sign: int = -1, nthreads: int = 1):
        mesh = np.empty(6 * nc ** 3, np.float32)
        tot = np.zeros(2, np.float64)
        t = C.c_double()
        self.L.ref_assign_on_mesh(self.h, _p(mesh), int(nc), int(sign), int(nthreads), _p(tot), C.byref(t))
        return mesh, tot

    def mark_halo(self, halos, left: float, right: float, nthreads: int = 1):
        h = np.ascontiguousarray(halos, np.float32).reshape(-1, 5)
        nph, mh = C.c_int64(), C.c_double()
        self.L.ref_mark_halo(self.h, _p(h), len(h), float(left), float(right), int(nthreads), C.byref(nph),
                             C.byref(mh))
        return nph.value, mh.value


def ref_write_mesh_separate(base: str, inode: int, mesh: np.ndarray, nc: int, L: float) -> None:
    m = np.ascontiguousarray(mesh, np.float32).copy()
    rlib().ref_write_mesh_separate(base.encode(), int(inode), _p(m), int(nc), float(L))


def compare_with_reference(nbr, rk, ref_nbr, ref_rk):
    """Particles whose result differs from the reference's own search: h
    (bit pattern), the 33-slot neighbour set, the set of the 32 nearest
    (the particle itself and the 31 its density term goes to,
    nbr_search.cpp:78-84) and the order of the 33 slots (the reference ranks
    equal distances by discovery, k_neighbors.h:26-28; the canonical order
    ranks them by index)."""
    nbr = np.asarray(nbr)
    ref_nbr = np.asarray(ref_nbr)
    return {
        "h": int(np.count_nonzero(np.asarray(rk).view(np.int32) != np.asarray(ref_rk).view(np.int32))),
        "set33": int(np.count_nonzero(np.any(np.sort(nbr, 1) != np.sort(ref_nbr, 1), axis=1))),
        "set32": int(np.count_nonzero(np.any(np.sort(nbr[:, :32], 1) != np.sort(ref_nbr[:, :32], 1), axis=1))),
        "order": int(np.count_nonzero(np.any(nbr != ref_nbr, axis=1))),
        "particles": int(len(nbr)),
    }


def ref_time_path(pos, vel, np_local: int, L: float, bf: float, ncs, nthreads: int):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    nc_arr = np.ascontiguousarray(ncs, np.int32)
    sec = np.zeros(3, np.float64)
    tot = np.zeros(2 * len(nc_arr), np.float64)
    rlib().ref_time_path(_p(pos), _p(vel), len(pos), int(np_local), float(L), float(bf), _p(nc_arr), len(nc_arr),
                         int(nthreads), _p(sec), _p(tot))
    return sec, tot


# ------------------------------------------------------------- stitching ---

FIELD_COMPONENTS = (1, 1, 1, 3, 1, 3)   # n, halo, c, vel, h + n, vel * (n / 8)


def stitch_field(meshes: dict, nc: int, nn: int, field: int) -> np.ndarray:
    """One field of the global mesh stitched from nn^3 normalised node meshes
    (dict inode -> 6 nc^3 floats, inode = ix + nn iy + nn^2 iz), x fastest
    over nc*nn cells per axis (tools/reformat.cpp:81-107).  Fields 4 and 5
    are tools/reformat2.cpp:118-126 in float32: h + n and vel * (n / 8)."""
    comps = FIELD_COMPONENTS[field]
    G = nc * nn
    m = nc ** 3
    out = np.zeros((G, G, G, comps), np.float32)
    for inode, mesh in meshes.items():
        mesh = np.asarray(mesh, np.float32)
        ix, iy, iz = inode % nn, (inode // nn) % nn, inode // (nn * nn)
        n, h, c = mesh[:m], mesh[m:2 * m], mesh[2 * m:3 * m]
        v = mesh[3 * m:6 * m].reshape(m, 3)
        f = [n[:, None], h[:, None], c[:, None], v, (h + n)[:, None], v * (n / np.float32(8.0))[:, None]][field]
        out[iz * nc:(iz + 1) * nc, iy * nc:(iy + 1) * nc, ix * nc:(ix + 1) * nc] = f.reshape(nc, nc, nc, comps)
    return out.ravel()


def ref_tool(name: str) -> str:
    """Path of a stitching tool compiled from the reference (_ref/reformat,
    _ref/reformat2), or '' when the reference tree was absent at build time."""
    p = os.path.join(HERE, "_ref", name)
    return p if os.path.exists(p) else ""


# ---------------------------------------------------------- zip particles ---

def read_zip(filebase: str, redshift: str, inode: int, nc_node_dim: int, mesh_scale: int, shift):
    """cubep3m zip reader (file_reader2.cpp:90-272) restated with numpy:
    (xv (np, 6) float32, np_local).  Counts per coarse cell from zip2 (i
    fastest), 255 escaping to the next int32 of zip3; the particles of a cell
    are consecutive in zip0 (3 offset bytes) and zip1 (3 int16); x = ms *
    ((b + 0.5) / 256 + i - 1) - shift, v = s / v_r2i, each step in float32."""
    path = lambda z: f"{filebase}{inode}/{redshift}zip{z}_{inode}.dat"
    z0 = np.fromfile(path(0), np.uint8)
    z1 = np.fromfile(path(1), np.uint8)
    z2 = np.fromfile(path(2), np.uint8)
    z3 = np.fromfile(path(3), np.uint8)
    n = int(z0[:4].view(np.int32)[0])
    v_r2i = z0[:64].view(np.float32)[12]
    nd = nc_node_dim
    cnt = z2[:nd ** 3].astype(np.int64)
    esc = np.nonzero(cnt == 255)[0]
    cnt[esc] = z3[:4 * len(esc)].view(np.int32)
    assert cnt.sum() == n and len(z0) == 64 + 3 * n and len(z1) == 64 + 6 * n
    cell = np.repeat(np.arange(nd ** 3), cnt)
    ijk = np.stack([cell % nd + 1, (cell // nd) % nd + 1, cell // (nd * nd) + 1], 1).astype(np.float32)
    b = z0[64:].reshape(n, 3).astype(np.float32)
    s = z1[64:].view(np.int16).reshape(n, 3).astype(np.float32)
    f32 = np.float32
    frac = (b + f32(0.5)) / f32(256.0)
    x = f32(mesh_scale) * ((frac + ijk) - f32(1.0)) - np.asarray(shift, np.float32)
    v = s / v_r2i
    return np.concatenate([x, v], 1).astype(np.float32), n
