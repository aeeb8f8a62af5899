"""asmooth_amd: MI355X (gfx950) adaptive-smoothing path of asmooth.

Python mirror of the reference's hot-path interface over the C ABI of
``include/asmooth.h`` (``libasmooth.so``, loaded with ctypes).  The method
names follow the reference functions they replace:

    ctx = Context()                           # one HIP stream, device buffers
    ctx.set_particles(p, np_local, boxsize)   # Particles after the exchange
    ctx.construct(-bf*L, (1+bf)*L)            # KDTree::construct
    ctx.neighbor_search(32)                   # neighbor_search loop
    ctx.mark_halo_particles(halos, lo, hi)    # mark_halo_particles loop
    ctx.assign_on_mesh(nc, sign=-1)           # clear_mesh + assign_on_mesh
    ctx.mesh_total(nc)                        # mesh_total for n and halo
    ctx.normalize_mesh(nc); ctx.get_mesh(nc)  # write_mesh_separate's normalisation

Errors raise AsmoothError with the library's message; nothing falls back to
a CPU path.  A missing library raises at import of this package's Context.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

__all__ = [
    "PARTICLE_DTYPE", "HALO_DTYPE", "K_SLOTS", "AsmoothError", "Context", "lib",
    "build_buffered_set", "build_node_share", "build_buffered_halos", "read_halo_file", "halo_delta_vir", "read_xv", "write_xv", "write_mesh_separate",
    "particles_from_xv", "timer_names", "Mlp", "mlp_init_params", "Batch",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libasmooth.so")

# basic_types.h:10-15 (36 bytes) and :17-20 (20 bytes).
PARTICLE_DTYPE = np.dtype([("x", "<f4", (3,)), ("v", "<f4", (3,)), ("dens", "<f4"),
                           ("rk", "<f4"), ("igrp", "<i4")])
HALO_DTYPE = np.dtype([("x", "<f4", (3,)), ("r", "<f4"), ("mass", "<f4")])
assert PARTICLE_DTYPE.itemsize == 36 and HALO_DTYPE.itemsize == 20

K_SLOTS = 33  # neighbour slots including the particle itself (k_neighbors.h:47)

OPT_KEEP_NEIGHBORS = 1
OPT_PROFILE = 2
OPT_LEAF_SIZE = 3
OPT_TILE = 4
OPT_DETERMINISTIC = 5
OPT_KNN_METHOD = 6
OPT_DEPOSIT_ACCUM = 7
OPT_SEED_WALK = 8
OPT_KERNEL = 9
OPT_KNN_BOUND = 10
OPT_BATCH_ORDER = 11
KERNEL_TENT = 0
KERNEL_SPLINE = 1
# Limits of asm_smooth_noalloc (include/asmooth.h).
NOALLOC_MAX_PARTICLES = 1 << 19
NOALLOC_MAX_NC = 128
NOALLOC_MAX_LIST = 1 << 23
# fields of asm_stitch_mesh (FIELD_NTOT and FIELD_MOMENTUM as tools/reformat2.cpp derives them)
FIELD_N, FIELD_HALO, FIELD_C, FIELD_VEL, FIELD_NTOT, FIELD_MOMENTUM = range(6)
FIELD_COMPONENTS = (1, 1, 1, 3, 1, 3)


class AsmoothError(RuntimeError):
    """A C-ABI call returned a negative status."""


_lib = None


def lib() -> C.CDLL:
    """Load libasmooth.so (built by asmooth_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AsmoothError(f"{LIB_PATH} is missing: run `python -m asmooth_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
    P = C.POINTER
    sigs = {
        "asm_version": (C.c_char_p, []),
        "asm_create": (i32, [P(vp), i32]),
        "asm_destroy": (None, [vp]),
        "asm_last_error": (C.c_char_p, [vp]),
        "asm_set_option": (i32, [vp, i32, i32]),
        "asm_synchronize": (i32, [vp]),
        "asm_set_particles": (i32, [vp, vp, i64, i64, f32]),
        "asm_set_particles_device": (i32, [vp, vp, i64, i64, f32]),
        "asm_tree_construct": (i32, [vp, f32, f32, P(i64)]),
        "asm_neighbor_search": (i32, [vp, i32]),
        "asm_mark_halo_particles": (i32, [vp, vp, i64, f32, f32, P(i64), P(f64)]),
        "asm_assign_on_mesh": (i32, [vp, i32, i32]),
        "asm_assign_on_meshes": (i32, [vp, P(i32), i32, i32]),
        "asm_assign_on_global_mesh": (i32, [vp, i32, i32, P(i32)]),
        "asm_mesh_total": (i32, [vp, i32, P(f64)]),
        "asm_mesh_stats": (i32, [vp, i32, P(f64)]),
        "asm_normalize_mesh": (i32, [vp, i32]),
        "asm_get_mesh": (i32, [vp, i32, vp]),
        "asm_mesh_device_ptr": (vp, [vp, i32]),
        "asm_stitch_mesh": (i32, [vp, i32, i32, P(i32), i32, vp]),
        "asm_reformat": (i32, [i32, C.c_char_p, C.c_char_p, i32, P(f64)]),
        "asm_reformat2": (i32, [i32, C.c_char_p, C.c_char_p, C.c_char_p, P(f64)]),
        "asm_device_to_host": (i32, [vp, vp, i64]),
        "asm_host_to_device": (i32, [vp, vp, i64]),
        "asm_smooth_noalloc": (i32, [i32, vp, i64, i64, f32, f32, f32, i32, i32, vp, vp, vp, C.c_char_p, i32]),
        "asm_get_particles": (i32, [vp, vp]),
        "asm_get_rk_dens": (i32, [vp, vp, vp, vp]),
        "asm_get_neighbors": (i32, [vp, vp]),
        "asm_build_buffered_set": (i64, [vp, i64, f32, f32, vp, i64, P(i64)]),
        "asm_build_node_share": (i64, [vp, i64, f32, i32, P(i32), f32, vp, i64, P(i64)]),
        "asm_build_buffered_halos": (i64, [vp, i64, f32, f32, vp, i64]),
        "asm_read_halo_file": (i32, [C.c_char_p, f32, vp, f32, f32, i32, i32, P(vp), P(i64)]),
        "asm_halo_delta_vir": (f64, [f64, f64]),
        "asm_read_xv": (i32, [C.c_char_p, P(vp), P(i64), vp]),
        "asm_read_zip": (i32, [C.c_char_p, C.c_char_p, i32, i32, i32, vp, P(vp), P(i64), vp]),
        "asm_write_xv": (i32, [C.c_char_p, vp, i64, vp]),
        "asm_free": (None, [vp]),
        "asm_write_mesh_separate": (i32, [C.c_char_p, i32, vp, i32]),
        "asm_get_tree": (i32, [vp, vp, vp, P(i64), P(i64), P(i32)]),
        "asm_get_sorted_results": (i32, [vp, vp, vp]),
        "asm_calibration_copy": (i32, [vp, vp, i64]),
        "asm_device_alloc": (i32, [P(vp), i64]),
        "asm_device_free": (i32, [vp]),
        "asm_timer_count": (i32, []),
        "asm_timer_name": (C.c_char_p, [i32]),
        "asm_get_timer": (i32, [vp, i32, P(f64), P(i64)]),
        "asm_reset_timers": (i32, [vp]),
        "asm_get_search_stats": (i32, [vp, P(C.c_uint64)]),
        "asm_autotune": (i32, [vp, P(i32), i32, P(i32), i32, i32, P(i32), P(i32)]),
        "asm_mlp_param_count": (i32, []),
        "asm_mlp_create": (i32, [P(vp), i32]),
        "asm_mlp_destroy": (None, [vp]),
        "asm_mlp_last_error": (C.c_char_p, [vp]),
        "asm_mlp_set_params": (i32, [vp, vp]),
        "asm_mlp_get_params": (i32, [vp, vp]),
        "asm_mlp_forward": (i32, [vp, vp, i64, vp]),
        "asm_mlp_loss_grad": (i32, [vp, vp, vp, i64, P(f64), vp]),
        "asm_mlp_train": (i32, [vp, vp, vp, i64, i32, i32, C.c_float, C.c_uint64, vp]),
        "asm_seed_samples": (i32, [vp, vp, i64, vp, vp]),
        "asm_set_seed_model": (i32, [vp, vp, C.c_float]),
        "asm_get_seed_stats": (i32, [vp, P(i64)]),
        "asm_batch_create": (i32, [P(vp), i32]),
        "asm_batch_destroy": (None, [vp]),
        "asm_batch_last_error": (C.c_char_p, [vp]),
        "asm_batch_add": (i32, [vp, vp, i64, i64, C.c_float, C.c_float, C.c_float, P(i32), i32, i32]),
        "asm_batch_size": (i32, [vp]),
        "asm_batch_set_option": (i32, [vp, i32, i32]),
        "asm_batch_autotune": (i32, [vp, P(i32), i32, P(i32), P(i32)]),
        "asm_batch_capture": (i32, [vp, i32]),
        "asm_batch_launch": (i32, [vp]),
        "asm_batch_run_sequential": (i32, [vp]),
        "asm_batch_synchronize": (i32, [vp]),
        "asm_batch_set_particles": (i32, [vp, i32, vp, i64, i64]),
        "asm_batch_get_mesh": (i32, [vp, i32, i32, vp, P(f64)]),
        "asm_batch_get_rk_dens": (i32, [vp, i32, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def timer_names() -> list[str]:
    L = lib()
    return [L.asm_timer_name(i).decode() for i in range(L.asm_timer_count())]


class Context:
    """One device context: the replacement of the reference's per-process
    Particles / KDTree / mesh buffers (clumping_tree_main.cpp:178-216)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.asm_create(C.byref(h), int(device))
        if rc != 0:
            raise AsmoothError(f"asm_create(device={device}) failed with status {rc}")
        self._h = h
        self.n = 0
        self.np_local = 0
        self.boxsize = 0.0

    # -- plumbing -------------------------------------------------------
    def _check(self, rc: int, what: str) -> None:
        if rc < 0:
            msg = self._L.asm_last_error(self._h).decode(errors="replace")
            raise AsmoothError(f"{what}: status {rc}: {msg}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.asm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option: int, value: int) -> None:
        self._check(self._L.asm_set_option(self._h, option, int(value)), "asm_set_option")

    def synchronize(self) -> None:
        self._check(self._L.asm_synchronize(self._h), "asm_synchronize")

    # -- the path ---------------------------------------------------------
    def set_particles(self, particles: np.ndarray, np_local: int, boxsize: float) -> None:
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._check(self._L.asm_set_particles(self._h, _ptr(p), len(p), int(np_local), float(boxsize)),
                    "asm_set_particles")
        self.n, self.np_local, self.boxsize = len(p), int(np_local), float(boxsize)

    def set_particles_device(self, dev_ptr: int, n: int, np_local: int, boxsize: float) -> None:
        self._check(self._L.asm_set_particles_device(self._h, C.c_void_p(dev_ptr), int(n), int(np_local),
                                                     float(boxsize)), "asm_set_particles_device")
        self.n, self.np_local, self.boxsize = int(n), int(np_local), float(boxsize)

    def construct(self, left: float, right: float) -> int:
        nn = C.c_int64()
        self._check(self._L.asm_tree_construct(self._h, float(left), float(right), C.byref(nn)),
                    "asm_tree_construct")
        return nn.value

    def neighbor_search(self, k: int = 32) -> None:
        self._check(self._L.asm_neighbor_search(self._h, int(k)), "asm_neighbor_search")

    def mark_halo_particles(self, halos: np.ndarray, left: float, right: float) -> tuple[int, float]:
        h = np.ascontiguousarray(halos, dtype=HALO_DTYPE)
        nph, mh = C.c_int64(), C.c_double()
        self._check(self._L.asm_mark_halo_particles(self._h, _ptr(h) if len(h) else None, len(h),
                                                    float(left), float(right), C.byref(nph), C.byref(mh)),
                    "asm_mark_halo_particles")
        return nph.value, mh.value

    def assign_on_mesh(self, nc: int, sign: int = -1) -> None:
        self._check(self._L.asm_assign_on_mesh(self._h, int(nc), int(sign)), "asm_assign_on_mesh")

    def assign_on_meshes(self, ncs: Sequence[int], sign: int = -1) -> None:
        arr = (C.c_int * len(ncs))(*[int(v) for v in ncs])
        self._check(self._L.asm_assign_on_meshes(self._h, arr, len(ncs), int(sign)), "asm_assign_on_meshes")

    def assign_on_global_mesh(self, nc: int, sign: int = -1, coord=(0, 0, 0)) -> None:
        arr = (C.c_int * 3)(*[int(v) for v in coord])
        self._check(self._L.asm_assign_on_global_mesh(self._h, int(nc), int(sign), arr),
                    "asm_assign_on_global_mesh")

    def mesh_total(self, nc: int) -> tuple[float, float]:
        out = (C.c_double * 2)()
        self._check(self._L.asm_mesh_total(self._h, int(nc), out), "asm_mesh_total")
        return out[0], out[1]

    def mesh_stats(self, nc: int) -> tuple[float, float, float, float]:
        out = (C.c_double * 4)()
        self._check(self._L.asm_mesh_stats(self._h, int(nc), out), "asm_mesh_stats")
        return tuple(out)  # type: ignore[return-value]

    def normalize_mesh(self, nc: int) -> None:
        self._check(self._L.asm_normalize_mesh(self._h, int(nc)), "asm_normalize_mesh")

    def get_mesh(self, nc: int) -> np.ndarray:
        m = np.empty(6 * nc ** 3, dtype=np.float32)
        self._check(self._L.asm_get_mesh(self._h, int(nc), _ptr(m)), "asm_get_mesh")
        return m

    def mesh_device_ptr(self, nc: int) -> int:
        return int(self._L.asm_mesh_device_ptr(self._h, int(nc)) or 0)

    def stitch_mesh(self, nc: int, nn: int, coord, field: int, dst_ptr: int) -> None:
        """Write this context's normalised mesh for nc, as node `coord` of an
        nn^3 decomposition, into the global device array at dst_ptr (one of
        the FIELD_* fields; runs on the context's stream)."""
        c3 = (C.c_int * 3)(*[int(v) for v in coord])
        self._check(self._L.asm_stitch_mesh(self._h, int(nc), int(nn), c3, int(field), C.c_void_p(dst_ptr)),
                    "asm_stitch_mesh")

    def get_particles(self) -> np.ndarray:
        out = np.empty(self.n, dtype=PARTICLE_DTYPE)
        self._check(self._L.asm_get_particles(self._h, _ptr(out)), "asm_get_particles")
        return out

    def get_rk_dens(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        rk = np.empty(self.n, np.float32)
        dens = np.empty(self.n, np.float32)
        igrp = np.empty(self.n, np.int32)
        self._check(self._L.asm_get_rk_dens(self._h, _ptr(rk), _ptr(dens), _ptr(igrp)), "asm_get_rk_dens")
        return rk, dens, igrp

    def get_neighbors(self) -> np.ndarray:
        nbr = np.empty((self.n, K_SLOTS), np.int32)
        self._check(self._L.asm_get_neighbors(self._h, _ptr(nbr)), "asm_get_neighbors")
        return nbr

    def get_tree(self):
        """(spos (n, 4) float32 with the original index in column 3 as int bits,
        nodes (2*nleaf_pad, 8) int32 words, nleaf, nleaf_pad, leaf_size)."""
        nl, npad, ls = C.c_int64(), C.c_int64(), C.c_int32()
        self._check(self._L.asm_get_tree(self._h, None, None, C.byref(nl), C.byref(npad), C.byref(ls)), "asm_get_tree")
        spos = np.empty((self.n, 4), np.float32)
        nodes = np.empty((2 * npad.value, 8), np.int32)
        self._check(self._L.asm_get_tree(self._h, _ptr(spos), _ptr(nodes), None, None, None), "asm_get_tree")
        return spos, nodes, nl.value, npad.value, ls.value

    def get_sorted_results(self):
        rk = np.empty(self.n, np.float32)
        d64 = np.empty(self.n, np.float32)
        self._check(self._L.asm_get_sorted_results(self._h, _ptr(rk), _ptr(d64)), "asm_get_sorted_results")
        return rk, d64

    # -- instrumentation -------------------------------------------------
    def timers(self) -> dict[str, tuple[float, int]]:
        out = {}
        for i, name in enumerate(timer_names()):
            ms, n = C.c_double(), C.c_int64()
            self._check(self._L.asm_get_timer(self._h, i, C.byref(ms), C.byref(n)), "asm_get_timer")
            out[name] = (ms.value, n.value)
        return out

    def reset_timers(self) -> None:
        self._check(self._L.asm_reset_timers(self._h), "asm_reset_timers")

    def search_stats(self) -> dict[str, int]:
        out = (C.c_uint64 * 4)()
        self._check(self._L.asm_get_search_stats(self._h, out), "asm_get_search_stats")
        return dict(zip(["node_tests", "leaf_scans", "dist_evals", "inserts"], list(out)))

    # -- learned seed radius -----------------------------------------------
    def seed_samples(self, idx: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Features (m, 27) and targets ln(h / spacing) of particles idx (caller order)."""
        idx = np.ascontiguousarray(idx, np.int64)
        x = np.empty((len(idx), MLP_INPUTS), np.float32)
        t = np.empty(len(idx), np.float32)
        self._check(self._L.asm_seed_samples(self._h, _ptr(idx), len(idx), _ptr(x), _ptr(t)), "asm_seed_samples")
        return x, t

    def set_seed_model(self, model: "Mlp | None", margin: float = 0.15) -> None:
        self._check(self._L.asm_set_seed_model(self._h, model._h if model is not None else None, float(margin)),
                    "asm_set_seed_model")

    def seed_stats(self) -> dict[str, int]:
        out = (C.c_int64 * 3)()
        self._check(self._L.asm_get_seed_stats(self._h, out), "asm_get_seed_stats")
        return {"fallback_particles": out[0], "fallback_wavefronts": out[1], "particles": out[2]}

    def autotune(self, leaf_sizes=(16, 24, 32), ncs=(), sign: int = -1) -> tuple[int, dict[int, int]]:
        """Time leaf sizes and per-nc deposition tiles on this context's
        particles and keep the fastest (asm_autotune).  Returns (leaf size,
        {nc: tile})."""
        ls = (C.c_int32 * max(len(leaf_sizes), 1))(*leaf_sizes)
        nn = (C.c_int32 * max(len(ncs), 1))(*ncs)
        tiles = (C.c_int32 * max(len(ncs), 1))()
        leaf = C.c_int32()
        self._check(self._L.asm_autotune(self._h, ls, len(leaf_sizes), nn, len(ncs), int(sign), C.byref(leaf), tiles),
                    "asm_autotune")
        return leaf.value, {int(nc): int(tiles[i]) for i, nc in enumerate(ncs)}


MLP_INPUTS = 27
MLP_HIDDEN = 64


class Mlp:
    """The learned seed radius model (asm_mlp): a 27-64-64-1 GELU network,
    trained and evaluated on the device."""

    def __init__(self, device: int = 0, params: np.ndarray | None = None):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.asm_mlp_create(C.byref(h), int(device))
        if rc != 0:
            raise AsmoothError(f"asm_mlp_create(device={device}) failed with status {rc}")
        self._h = h
        self.n_params = int(self._L.asm_mlp_param_count())
        if params is not None:
            self.set_params(params)

    def _check(self, rc: int, what: str) -> None:
        if rc < 0:
            msg = self._L.asm_mlp_last_error(self._h).decode(errors="replace")
            raise AsmoothError(f"{what}: status {rc}: {msg}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.asm_mlp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise AsmoothError(f"expected {self.n_params} parameters, got {p.size}")
        self._check(self._L.asm_mlp_set_params(self._h, _ptr(p)), "asm_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self._check(self._L.asm_mlp_get_params(self._h, _ptr(p)), "asm_mlp_get_params")
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        y = np.empty(len(x), np.float32)
        self._check(self._L.asm_mlp_forward(self._h, _ptr(x), len(x), _ptr(y)), "asm_mlp_forward")
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        self._check(self._L.asm_mlp_loss_grad(self._h, _ptr(x), _ptr(t), len(x), C.byref(loss), _ptr(g)),
                    "asm_mlp_loss_grad")
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 5, batch: int = 4096, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.empty(epochs, np.float32)
        self._check(self._L.asm_mlp_train(self._h, _ptr(x), _ptr(t), len(x), int(epochs), int(batch), float(lr),
                                          int(seed), _ptr(losses)), "asm_mlp_train")
        return losses


class Batch:
    """Independent node shares run as one recorded HIP graph (asm_batch)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.asm_batch_create(C.byref(h), int(device))
        if rc != 0:
            raise AsmoothError(f"asm_batch_create(device={device}) failed with status {rc}")
        self._h = h
        self.shares = []

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = self._L.asm_batch_last_error(self._h).decode(errors="replace")
            raise AsmoothError(f"{what}: status {rc}: {msg}")
        return rc

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.asm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, particles: np.ndarray, np_local: int, boxsize: float, left: float, right: float,
            ncs: Sequence[int], sign: int = -1) -> int:
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        nn = (C.c_int32 * len(ncs))(*ncs)
        i = self._check(self._L.asm_batch_add(self._h, _ptr(p), len(p), int(np_local), float(boxsize), float(left),
                                              float(right), nn, len(ncs), int(sign)), "asm_batch_add")
        self.shares.append({"n": len(p), "np_local": int(np_local), "ncs": list(ncs)})
        return i

    def set_option(self, option: int, value: int) -> None:
        self._check(self._L.asm_batch_set_option(self._h, option, int(value)), "asm_batch_set_option")

    def autotune(self, leaf_sizes=(16, 24, 32)) -> tuple[int, list[int]]:
        ls = (C.c_int32 * len(leaf_sizes))(*leaf_sizes)
        big = max(range(len(self.shares)), key=lambda i: self.shares[i]["n"])
        tiles = (C.c_int32 * len(self.shares[big]["ncs"]))()
        leaf = C.c_int32()
        self._check(self._L.asm_batch_autotune(self._h, ls, len(leaf_sizes), C.byref(leaf), tiles),
                    "asm_batch_autotune")
        return leaf.value, list(tiles)

    def capture(self, n_streams: int = 8) -> None:
        self._check(self._L.asm_batch_capture(self._h, int(n_streams)), "asm_batch_capture")

    def launch(self) -> None:
        self._check(self._L.asm_batch_launch(self._h), "asm_batch_launch")

    def run_sequential(self) -> None:
        self._check(self._L.asm_batch_run_sequential(self._h), "asm_batch_run_sequential")

    def synchronize(self) -> None:
        self._check(self._L.asm_batch_synchronize(self._h), "asm_batch_synchronize")

    def set_particles(self, i: int, particles: np.ndarray, np_local: int) -> None:
        """Replace share i's records in place (same count; the graph stays)."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._check(self._L.asm_batch_set_particles(self._h, int(i), _ptr(p), len(p), int(np_local)),
                    "asm_batch_set_particles")
        self.shares[i]["np_local"] = int(np_local)

    def get_mesh(self, i: int, nc: int) -> tuple[np.ndarray, np.ndarray]:
        mesh = np.empty(6 * nc ** 3, np.float32)
        tot = (C.c_double * 4)()
        self._check(self._L.asm_batch_get_mesh(self._h, int(i), int(nc), _ptr(mesh), tot), "asm_batch_get_mesh")
        return mesh, np.array(list(tot))

    def get_rk_dens(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.shares[i]["n"]
        rk = np.empty(n, np.float32)
        dens = np.empty(n, np.float32)
        self._check(self._L.asm_batch_get_rk_dens(self._h, int(i), _ptr(rk), _ptr(dens)), "asm_batch_get_rk_dens")
        return rk, dens


def mlp_init_params(seed: int = 0) -> np.ndarray:
    """He-style random initialisation in the asm_mlp parameter layout
    (W1[27][64], b1, W2[64][64], b2, w3[64], b3)."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((MLP_INPUTS, MLP_HIDDEN)) * np.sqrt(2.0 / MLP_INPUTS)
    w2 = rng.standard_normal((MLP_HIDDEN, MLP_HIDDEN)) * np.sqrt(2.0 / MLP_HIDDEN)
    w3 = rng.standard_normal(MLP_HIDDEN) * np.sqrt(1.0 / MLP_HIDDEN)
    return np.concatenate([w1.ravel(), np.zeros(MLP_HIDDEN), w2.ravel(), np.zeros(MLP_HIDDEN), w3,
                           np.zeros(1)]).astype(np.float32)


# -- host steps either side of the path ------------------------------------

def particles_from_xv(xv: np.ndarray) -> np.ndarray:
    """Particle records from an (n, 6) x/v array (dens = rk = igrp = 0)."""
    xv = np.asarray(xv, np.float32).reshape(-1, 6)
    p = np.zeros(len(xv), PARTICLE_DTYPE)
    p["x"] = xv[:, :3]
    p["v"] = xv[:, 3:]
    return p


def build_buffered_set(xv: np.ndarray, boxsize: float, bf: float) -> tuple[np.ndarray, int]:
    """Single periodic rank: exchange_buffer images + remove_extra_particles."""
    xv = np.ascontiguousarray(xv, np.float32).reshape(-1, 6)
    cap = int(len(xv) * (1 + 2 * bf) ** 3 * 1.2) + 1024
    out = np.empty(cap, PARTICLE_DTYPE)
    nl = C.c_int64()
    n = lib().asm_build_buffered_set(_ptr(xv), len(xv), float(boxsize), float(bf), _ptr(out), cap, C.byref(nl))
    if n < 0:
        raise AsmoothError(f"asm_build_buffered_set failed with status {n}")
    return out[:n].copy(), nl.value


def build_node_share(xv: np.ndarray, box_total: float, nn: int, coord, bf: float) -> tuple[np.ndarray, int]:
    """Node `coord` of an nn^3 decomposition of a periodic box (node coordinates)."""
    xv = np.ascontiguousarray(xv, np.float32).reshape(-1, 6)
    cap = int(len(xv) / nn ** 3 * (1 + 2 * bf) ** 3 * 1.5) + 4096
    out = np.empty(cap, PARTICLE_DTYPE)
    nl = C.c_int64()
    c3 = (C.c_int * 3)(*[int(v) for v in coord])
    n = lib().asm_build_node_share(_ptr(xv), len(xv), float(box_total), int(nn), c3, float(bf), _ptr(out), cap,
                                   C.byref(nl))
    if n < 0:
        raise AsmoothError(f"asm_build_node_share failed with status {n}")
    return out[:n].copy(), nl.value


def build_buffered_halos(halos: np.ndarray, boxsize: float, bf: float) -> np.ndarray:
    """Halo catalogue with the periodic images of a single rank (not trimmed)."""
    h = np.ascontiguousarray(halos, HALO_DTYPE)
    cap = 8 * len(h) + 16
    out = np.empty(cap, HALO_DTYPE)
    n = lib().asm_build_buffered_halos(_ptr(h) if len(h) else None, len(h), float(boxsize), float(bf), _ptr(out), cap)
    if n < 0:
        raise AsmoothError(f"asm_build_buffered_halos failed with status {n}")
    return out[:n].copy()


def read_halo_file(path: str, boxsize: float, shift=(0.0, 0.0, 0.0), z: float = 0.0, omega_m: float = 0.3,
                   m200: bool = False, pid_flag: bool = True) -> np.ndarray:
    """SO halo catalogue (22-float records, optional particle-ID blocks)."""
    L = lib()
    sh = np.ascontiguousarray(shift, np.float32)
    buf, n = C.c_void_p(), C.c_int64()
    rc = L.asm_read_halo_file(path.encode(), float(boxsize), _ptr(sh), float(z), float(omega_m), int(m200),
                              int(pid_flag), C.byref(buf), C.byref(n))
    if rc < 0:
        raise AsmoothError(f"asm_read_halo_file({path}) failed with status {rc}")
    try:
        raw = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_float)), shape=(max(n.value, 1) * 5,)).copy()
    finally:
        L.asm_free(buf)
    return raw[: n.value * 5].view(HALO_DTYPE).copy()


def halo_delta_vir(z: float, omega_m: float) -> float:
    return float(lib().asm_halo_delta_vir(float(z), float(omega_m)))


def read_xv(path: str) -> tuple[np.ndarray, np.ndarray]:
    """cubep3m xv file -> ((n, 6) float32, header[11])."""
    L = lib()
    buf, n = C.c_void_p(), C.c_int64()
    hdr = np.zeros(11, np.float32)
    rc = L.asm_read_xv(path.encode(), C.byref(buf), C.byref(n), _ptr(hdr))
    if rc < 0:
        raise AsmoothError(f"asm_read_xv({path}) failed with status {rc}")
    try:
        arr = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_float)), shape=(n.value * 6,)).copy()
    finally:
        L.asm_free(buf)
    return arr.reshape(-1, 6), hdr


def read_zip(filebase: str, redshift: str, inode: int, nc_node_dim: int, mesh_scale: int = 4,
             shift=(0.0, 0.0, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """cubep3m zip reader (file_reader2.cpp:90-272): (xv (np, 6) float32,
    zip0 header as 16 float32 words)."""
    p, n = C.c_void_p(), C.c_int64()
    hdr = np.zeros(16, np.float32)
    sh = np.ascontiguousarray(shift, np.float32)
    rc = lib().asm_read_zip(filebase.encode(), redshift.encode(), int(inode), int(nc_node_dim), int(mesh_scale),
                            _ptr(sh), C.byref(p), C.byref(n), _ptr(hdr))
    if rc < 0:
        raise AsmoothError(f"asm_read_zip({filebase}{inode}/{redshift}) failed with status {rc}")
    try:
        xv = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n.value * 6,)).copy() \
            if n.value else np.zeros(0, np.float32)
    finally:
        lib().asm_free(p)
    return xv.reshape(-1, 6), hdr


def write_xv(path: str, xv: np.ndarray, header=None) -> None:
    xv = np.ascontiguousarray(xv, np.float32).reshape(-1, 6)
    hdr = np.zeros(11, np.float32) if header is None else np.ascontiguousarray(header, np.float32)
    rc = lib().asm_write_xv(path.encode(), _ptr(xv), len(xv), _ptr(hdr))
    if rc < 0:
        raise AsmoothError(f"asm_write_xv({path}) failed with status {rc}")


def write_mesh_separate(base: str, inode: int, mesh: np.ndarray, nc: int) -> None:
    m = np.ascontiguousarray(mesh, np.float32)
    rc = lib().asm_write_mesh_separate(base.encode(), int(inode), _ptr(m), int(nc))
    if rc < 0:
        raise AsmoothError(f"asm_write_mesh_separate({base}) failed with status {rc}")


def reformat(nn: int, out_dir: str, file_base: str, n_components: int) -> float:
    """tools/reformat.cpp: stitch nn^3 node files <file_base><inode>.dat into
    <out_dir>/<file_base>_all.dat; returns the total the tool prints."""
    tot = C.c_double()
    rc = lib().asm_reformat(int(nn), out_dir.encode(), file_base.encode(), int(n_components), C.byref(tot))
    if rc < 0:
        raise AsmoothError(f"asm_reformat({file_base}) failed with status {rc}")
    return tot.value


def reformat2(nn: int, out_dir: str, redshift: str, in_dir: str) -> tuple[float, float]:
    """tools/reformat2.cpp: n, ntot and momentum global files from the node
    files of write_mesh_separate; returns the two totals the tool prints."""
    tot = (C.c_double * 2)()
    rc = lib().asm_reformat2(int(nn), out_dir.encode(), redshift.encode(), in_dir.encode(), tot)
    if rc < 0:
        raise AsmoothError(f"asm_reformat2({redshift}) failed with status {rc}")
    return tot[0], tot[1]


def smooth_noalloc(device: int, d_particles: int, n: int, np_local: int, boxsize: float, left: float,
                   right: float, nc: int, sign: int, d_rk: int, d_dens: int, d_mesh: int | None) -> None:
    """asm_smooth_noalloc: tree, search, h and density (and one raw mesh when
    nc > 0) of n device particle records, into device outputs, with no device
    allocation (the library's fixed work array)."""
    err = C.create_string_buffer(512)
    rc = lib().asm_smooth_noalloc(int(device), C.c_void_p(d_particles), int(n), int(np_local), float(boxsize),
                                  float(left), float(right), int(nc), int(sign), C.c_void_p(d_rk),
                                  C.c_void_p(d_dens), C.c_void_p(d_mesh) if d_mesh else None, err, len(err))
    if rc != 0:
        raise AsmoothError(f"asm_smooth_noalloc failed ({rc}): {err.value.decode(errors='replace')}")


class DeviceArray:
    """Raw device allocation of n float32 (for asm_stitch_mesh destinations)."""

    def __init__(self, n: int):
        self.n = int(n)
        p = C.c_void_p()
        if lib().asm_device_alloc(C.byref(p), 4 * self.n) != 0:
            raise AsmoothError("device allocation failed")
        self.ptr = p.value

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, np.float32)
        if lib().asm_device_to_host(_ptr(out), C.c_void_p(self.ptr), 4 * self.n) != 0:
            raise AsmoothError("device to host copy failed")
        return out

    def from_host(self, a: np.ndarray) -> None:
        """Copy the bytes of a (at most 4 n) to the start of the array."""
        a = np.ascontiguousarray(a)
        if a.nbytes > 4 * self.n:
            raise AsmoothError("host array larger than the device array")
        if lib().asm_host_to_device(C.c_void_p(self.ptr), _ptr(a), a.nbytes) != 0:
            raise AsmoothError("host to device copy failed")

    def close(self) -> None:
        if self.ptr:
            lib().asm_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

