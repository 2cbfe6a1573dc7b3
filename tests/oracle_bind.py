"""ctypes bindings of the oracle (oracle/liboracle.so, the CPU restatement) and
of the reference harness (oracle/_ref/libptref.so, the reference's own code).
Test infrastructure: used only as the checker."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libptref.so")
THREADS = max(1, min(16, os.cpu_count() or 1))

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32p = ctypes.POINTER(ctypes.c_uint32)

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_PATH)
        lib.or_trace.argtypes = [_vp, _vp, _sz, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        lib.or_trace.restype = None
        lib.or_trace_f64.argtypes = [_vp, _vp, _sz, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int]
        lib.or_trace_f64.restype = None
        lib.or_sample_2d.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        lib.or_sample_2d.restype = None
        lib.or_path_key.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
        lib.or_path_key.restype = ctypes.c_uint64
        lib.or_render.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int]
        lib.or_render.restype = ctypes.c_uint64
        lib.or_texture_sample.argtypes = [_vp, ctypes.c_uint32, _vp, _sz, _vp]
        lib.or_texture_sample.restype = None
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_last_error.restype = ctypes.c_char_p
        lib.ref_free.argtypes = [_vp]
        lib.ref_world_create.argtypes = [_vp, _sz]
        lib.ref_world_create.restype = _vp
        lib.ref_world_destroy.argtypes = [_vp]
        lib.ref_world_pack.argtypes = [_vp, ctypes.POINTER(_u32p), ctypes.POINTER(_sz)]
        lib.ref_world_tree.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_u32p), ctypes.POINTER(_sz),
                                       ctypes.POINTER(_u32p), ctypes.POINTER(_sz)]
        lib.ref_world_camera.argtypes = [_vp, _vp]
        lib.ref_world_camera.restype = None
        lib.ref_scene.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_double,
                                  ctypes.POINTER(_u32p), ctypes.POINTER(_sz), ctypes.POINTER(_u32p), ctypes.POINTER(_sz)]
        lib.ref_world_trace.argtypes = [_vp, _vp, _sz, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        lib.ref_world_trace.restype = ctypes.c_long
        lib.ref_world_render.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
        lib.ref_sample_2d.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        lib.ref_hardware_threads.restype = ctypes.c_int
        lib.ref_texture_sample.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _sz, _vp]
        lib.ref_texture_sample.restype = None
        lib.ref_decode_image.argtypes = [ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]
        _ref = lib
    return _ref


def _take(lib, p, n) -> np.ndarray:
    try:
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)
    finally:
        lib.ref_free(p)


# ---- oracle --------------------------------------------------------------------

def or_trace(ptgp: np.ndarray, rays: np.ndarray, any_hit: bool = False, threads: int = THREADS):
    """Returns (t, u, v, kind, prim, inst, counts[6])."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(rays)
    t, u, v = (np.zeros(n, np.float32) for _ in range(3))
    kind, prim, inst = (np.zeros(n, np.uint32) for _ in range(3))
    counts = np.zeros(6, np.uint64)
    w = np.ascontiguousarray(ptgp, np.uint32)
    oracle().or_trace(w.ctypes.data, rays.ctypes.data, n, 1 if any_hit else 0, t.ctypes.data, u.ctypes.data,
                      v.ctypes.data, kind.ctypes.data, prim.ctypes.data, inst.ctypes.data, counts.ctypes.data, threads)
    return t, u, v, kind, prim, inst, counts


def or_trace_f64(ptgp: np.ndarray, rays: np.ndarray, threads: int = THREADS, float_ray: bool = False):
    """Closest hits with every candidate primitive tested in double. float_ray: the ray is moved into an
    instance's object space in float, as the reference moves it (geometry.hpp:213-217), before the double
    tests; otherwise that transform is in double too."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(rays)
    t = np.zeros(n, np.float64)
    kind, prim, inst = (np.zeros(n, np.uint32) for _ in range(3))
    w = np.ascontiguousarray(ptgp, np.uint32)
    oracle().or_trace_f64(w.ctypes.data, rays.ctypes.data, n, t.ctypes.data, kind.ctypes.data, prim.ctypes.data,
                          inst.ctypes.data, threads, 1 if float_ray else 0)
    return t, kind, prim, inst


def or_sample_2d(seed, i, j, index, pair):
    a, b = ctypes.c_float(), ctypes.c_float()
    oracle().or_sample_2d(seed, i, j, index, pair, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def or_render(ptgp, width, height, spp, depth=10, seed=0, first_sample=0, nee=True, rows=None, aux=False,
              threads: int = THREADS):
    j0, j1 = rows if rows else (0, height)
    rgb = np.zeros((j1 - j0, width, 3), np.float32)
    ax = np.zeros((j1 - j0, width, 7), np.float32) if aux else None
    w = np.ascontiguousarray(ptgp, np.uint32)
    rays = oracle().or_render(w.ctypes.data, width, height, spp, depth, seed, first_sample, 1 if nee else 0, j0, j1,
                              rgb.ctypes.data, ax.ctypes.data if aux else None, threads)
    return rgb, ax, int(rays)


# ---- the reference harness ------------------------------------------------------

REF_SCENES = {"materials": 0, "field": 1, "mesh": 2, "room": 3, "crowd": 4}


def ref_scene(name: str, crowd_side: int = 8, flatten: bool = False, obj_path: str = "", turn_deg: float = 0.0):
    """(PTSD, PTGP) of one of the reference's own scenes, built and packed by the reference."""
    lib = ref()
    a, na, b, nb = _u32p(), _sz(0), _u32p(), _sz(0)
    if lib.ref_scene(REF_SCENES[name], crowd_side, 1 if flatten else 0, obj_path.encode(), float(turn_deg), ctypes.byref(a),
                     ctypes.byref(na), ctypes.byref(b), ctypes.byref(nb)):
        raise RuntimeError(lib.ref_last_error().decode())
    return _take(lib, a, na), _take(lib, b, nb)


class RefWorld:
    """The reference's geometry built (geometry::build(true)) from a PTSD scene."""

    def __init__(self, ptsd):
        self.lib = ref()
        self._w = np.ascontiguousarray(ptsd, np.uint32)
        self.h = self.lib.ref_world_create(self._w.ctypes.data, self._w.size)
        if not self.h:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def close(self):
        if self.h:
            self.lib.ref_world_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *e):
        self.close()

    def pack(self):
        """serialize(pack_for_gpu(...)), or None where the reference refuses (too deep)."""
        p, n = _u32p(), _sz(0)
        if self.lib.ref_world_pack(self.h, ctypes.byref(p), ctypes.byref(n)):
            return None
        return _take(self.lib, p, n)

    def tree(self, which: int):
        a, na, b, nb = _u32p(), _sz(0), _u32p(), _sz(0)
        if self.lib.ref_world_tree(self.h, which, ctypes.byref(a), ctypes.byref(na), ctypes.byref(b), ctypes.byref(nb)):
            raise RuntimeError(self.lib.ref_last_error().decode())
        return _take(self.lib, a, na), _take(self.lib, b, nb)

    def camera(self):
        out = np.zeros(12, np.float32)
        self.lib.ref_world_camera(self.h, out.ctypes.data)
        return out.reshape(4, 3)

    def trace(self, rays, any_hit=False, threads: int = THREADS):
        """(kind, index, sub, t, u, v, disagreements with geometry::hit/occluded)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        kind, index, sub = (np.zeros(n, np.uint32) for _ in range(3))
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        bad = self.lib.ref_world_trace(self.h, rays.ctypes.data, n, 1 if any_hit else 0, kind.ctypes.data,
                                       index.ctypes.data, sub.ctypes.data, t.ctypes.data, u.ctypes.data,
                                       v.ctypes.data, threads)
        return kind, index, sub, t, u, v, int(bad)

    def render(self, width, height, spp, depth=10, seed=0, first_sample=0, nee=True, threads: int = THREADS,
               rows=None, aux=False):
        j0, j1 = rows if rows else (0, height)
        rgb = np.zeros((j1 - j0, width, 3), np.float32)
        ax = np.zeros((j1 - j0, width, 7), np.float32) if aux else None
        rays, secs = ctypes.c_uint64(0), ctypes.c_double(0)
        self.lib.ref_world_render(self.h, width, height, spp, depth, seed, first_sample, 1 if nee else 0, threads,
                                  j0, j1, rgb.ctypes.data, ax.ctypes.data if aux else None, ctypes.byref(rays),
                                  ctypes.byref(secs))
        return rgb, ax, int(rays.value), float(secs.value)


def ref_denoise(rgb, aux, spp, threads: int = THREADS):
    """The reference's --denoise step (main.cpp:172-186, denoise.hpp) on (h, w, 3) rgb and (h, w, 7) aux.
    Returns (image, seconds)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    aux = np.ascontiguousarray(aux, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.zeros_like(rgb)
    secs = ctypes.c_double(0)
    lib = ref()
    lib.ref_denoise.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.ref_denoise(w, h, int(spp), rgb.ctypes.data, aux.ctypes.data, out.ctypes.data, threads, ctypes.byref(secs))
    return out, secs.value


def or_texture_sample(ptgp, t: int, uv) -> np.ndarray:
    """The oracle's image_texture::sample of texture t of a PTGP buffer at (n, 2) uv: (n, 3) floats."""
    p = np.ascontiguousarray(ptgp, np.uint32)
    q = np.ascontiguousarray(uv, np.float32)
    out = np.zeros((len(q), 3), np.float32)
    oracle().or_texture_sample(p.ctypes.data, t, q.ctypes.data, len(q), out.ctypes.data)
    return out


def ref_texture_sample(rgb, uv) -> np.ndarray:
    """The reference's image_texture::sample over (h, w, 3) sRGB bytes at (n, 2) uv: (n, 3) floats."""
    img = np.ascontiguousarray(rgb, np.uint8)
    q = np.ascontiguousarray(uv, np.float32)
    out = np.zeros((len(q), 3), np.float32)
    ref().ref_texture_sample(img.ctypes.data, img.shape[1], img.shape[0], q.ctypes.data, len(q), out.ctypes.data)
    return out


def ref_decode_image(data: bytes) -> np.ndarray:
    """The reference's decode_image of a file's bytes: (h, w, 3) uint8; RuntimeError with its message."""
    lib = ref()
    w, h, p = ctypes.c_int(), ctypes.c_int(), ctypes.POINTER(ctypes.c_uint8)()
    if lib.ref_decode_image(data, len(data), ctypes.byref(w), ctypes.byref(h), ctypes.byref(p)) != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    try:
        n = 3 * w.value * h.value
        return np.ctypeslib.as_array(p, (max(n, 1),))[:n].copy().reshape(h.value, w.value, 3)
    finally:
        lib.ref_free(p)


def ref_sample_2d(seed, i, j, index, pair):
    a, b = ctypes.c_float(), ctypes.c_float()
    ref().ref_sample_2d(seed, i, j, index, pair, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def ptgp_to_ref_ids(ptgp: np.ndarray, ptsd: np.ndarray, kind, prim, inst):
    """Maps PTGP-numbered hits (kind 1-3, prim, inst) to the reference's
    (kind 1 large / 2 sphere / 3 flat triangle / 4 instance, index, sub)."""
    n_small = int(ptgp[20])
    n_flat_tris = int(ptsd[13])
    n_mesh = int(ptsd[14])
    hdr = 32 + 12 * int(ptsd[10]) + 5 * (int(ptsd[11]) + int(ptsd[12])) + 28 * n_flat_tris
    sizes = ptsd[hdr:hdr + n_mesh].astype(np.int64)
    base = n_flat_tris + np.concatenate([[0], np.cumsum(sizes)])[:-1] if n_mesh else np.zeros(0, np.int64)
    inst_mesh = np.zeros(0, np.int64)
    if int(ptsd[15]):
        start = hdr + n_mesh + 28 * int(ptsd[16])
        inst_mesh = ptsd[start::16][: int(ptsd[15])].astype(np.int64)
    kind = np.asarray(kind)
    prim = np.asarray(prim).astype(np.int64)
    inst = np.asarray(inst).astype(np.int64)
    rk = np.zeros_like(kind)
    ri = np.zeros(len(kind), np.int64)
    rs = np.zeros(len(kind), np.int64)
    m = kind == 1
    rk[m], ri[m] = 1, prim[m] - n_small
    m = kind == 2
    rk[m], ri[m] = 2, prim[m]
    m = (kind == 3) & (inst == 0xFFFFFFFF)
    rk[m], ri[m] = 3, prim[m]
    m = (kind == 3) & (inst != 0xFFFFFFFF)
    if m.any():
        rk[m], ri[m] = 4, inst[m]
        rs[m] = prim[m] - base[inst_mesh[inst[m]]]
    return rk, ri, rs
