"""photon_tracer_amd: an MI355X path-tracing core behind a C ABI.

The native library (``libphoton_tracer_amd.so``, built in this directory from
``csrc/``) does all the work: scene generation and packing on the host, BVH
traversal and the wavefront path loop as gfx950 HIP kernels. This module only
binds its C ABI (``include/photon_tracer.h``) with ctypes and moves numpy
arrays across it. There is no fallback: if the library is missing, importing
the bindings raises.

Reference correspondence (apollo-2006/photon_tracer):
  generate / pack   scene builders + geometry::build + pack_for_gpu
                    (renderer.hpp:158-343, geometry.hpp:74-104, gpu_pack.hpp:84-203)
  Scene.trace       geometry::hit / geometry::occluded (geometry.hpp:111-190)
  Scene.render      render_row_linear over a frame or band of rows
                    (renderer.hpp:366-416)
  Batch.render      many render_row_linear calls (jobs of several scenes,
                    views, sizes and spp) as one launch sequence
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphoton_tracer_amd.so")

HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("kind", "<u4"), ("prim", "<u4"), ("inst", "<u4")])
MISS, LARGE_SPHERE, SPHERE, TRIANGLE = 0, 1, 2, 3
NONE = 0xFFFFFFFF

RENDER_NO_NEE = 1
RENDER_NO_GRAPH = 2
RENDER_TIMING = 4
RENDER_COUNT = 8
RENDER_DEVICE_OUTPUT = 16
RENDER_NO_TAIL = 32
RENDER_GUIDE = 64
RENDER_TRAIN = 128
RENDER_PER_RAY = 256
GUIDE_PARAMS = 18720
GUIDE_OUTPUTS = 32
JOB_CAMERA = 0x100
BATCH_STREAMS = 0x200

# Algorithmic bytes of one traversal query (the ray in, the hit out, and every
# record the reference's walk must read): 32 + 16 per query, 128 per node slab
# test, 4 per leaf reference, 48 per triangle test, 16 per sphere test, 52 per
# instance entry (to_object rows + root), 16 per large-sphere test.
VISIT_BYTES = (128, 4, 48, 16, 52, 16)
QUERY_BYTES = 32 + 16


class PTError(RuntimeError):
    """A failure reported by the library (status code and pt_last_error())."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class _RenderParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_int32),
        ("max_bounces", ctypes.c_int32), ("seed", ctypes.c_uint64), ("first_sample", ctypes.c_int32),
        ("row_begin", ctypes.c_int32), ("row_end", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("samples_per_batch", ctypes.c_int32), ("stack_entries", ctypes.c_int32),
    ]


class _RenderStats(ctypes.Structure):
    _fields_ = [
        ("rays", ctypes.c_uint64), ("primary_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("render_ms", ctypes.c_double), ("extend_ms", ctypes.c_double), ("shadow_ms", ctypes.c_double),
        ("visits", ctypes.c_uint64 * 7), ("batches", ctypes.c_int32), ("samples_per_batch", ctypes.c_int32),
        ("guide_ms", ctypes.c_double), ("guide_queries", ctypes.c_uint64),
    ]


class _GuideParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("learning_rate", ctypes.c_float), ("box_lo", ctypes.c_float * 3),
                ("box_hi", ctypes.c_float * 3), ("steps_per_batch", ctypes.c_int32), ("minibatch", ctypes.c_int32),
                ("weight_clamp", ctypes.c_float), ("bounces", ctypes.c_int32)]


class _BatchPart(ctypes.Structure):
    _fields_ = [("root", ctypes.c_uint32), ("n_large", ctypes.c_uint32), ("large_off", ctypes.c_uint32),
                ("n_lights", ctypes.c_uint32), ("lights_off", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("camera", ctypes.c_float * 16)]


class _BatchJob(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("spp", ctypes.c_int32), ("seed", ctypes.c_uint64), ("first_sample", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("camera", ctypes.c_float * 16)]


class _BatchParams(ctypes.Structure):
    _fields_ = [("max_bounces", ctypes.c_int32), ("flags", ctypes.c_int32), ("stack_entries", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("streams", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("max_paths", ctypes.c_uint64)]


class _Tried(ctypes.Structure):
    _fields_ = [("stack_entries", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("max_paths", ctypes.c_uint64), ("mrays_per_s", ctypes.c_double)]


class _TuneResult(ctypes.Structure):
    _fields_ = [("stack_entries", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("n_tried", ctypes.c_int32), ("max_paths", ctypes.c_uint64), ("tried", _Tried * 32)]


class _Image(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("rgb", ctypes.c_void_p)]


class _JobDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_int32),
        ("primitives", ctypes.c_uint32), ("seed", ctypes.c_uint64),
    ]


_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def library() -> ctypes.CDLL:
    """The loaded native library (loaded on first use)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C photon_tracer_amd")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    lib.pt_last_error.restype = ctypes.c_char_p
    lib.pt_version.restype = ctypes.c_char_p
    lib.pt_free.argtypes = [vp]
    lib.pt_device_info.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_int)]
    lib.pt_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(_u32p),
                                ctypes.POINTER(sz)]
    lib.pt_generate_job.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(_JobDesc),
                                    ctypes.POINTER(_u32p), ctypes.POINTER(sz)]
    lib.pt_pack.argtypes = [vp, sz, ctypes.POINTER(_u32p), ctypes.POINTER(sz)]
    lib.pt_ptsd_add_obj.argtypes = [vp, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_double, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(_u32p),
                                    ctypes.POINTER(sz)]
    lib.pt_ptsd_add_obj_textured.argtypes = [vp, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz, vp, sz,
                                             ctypes.POINTER(ctypes.c_float), ctypes.c_double, ctypes.c_uint32,
                                             ctypes.c_double, ctypes.POINTER(_u32p), ctypes.POINTER(sz)]
    lib.pt_decode_image.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]
    lib.pt_pack_ex.argtypes = [vp, sz, ctypes.c_int, ctypes.POINTER(_u32p), ctypes.POINTER(sz)]
    lib.pt_bvh_build.argtypes = [vp, sz, ctypes.c_int, ctypes.POINTER(_u32p), ctypes.POINTER(sz), ctypes.POINTER(_u32p),
                                 ctypes.POINTER(ctypes.c_double)]
    lib.pt_ptgp_depth.argtypes = [vp, sz, ctypes.POINTER(ctypes.c_int)]
    lib.pt_scene_create.argtypes = [vp, sz, ctypes.POINTER(vp)]
    lib.pt_scene_destroy.argtypes = [vp]
    lib.pt_trace.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    lib.pt_trace_device.argtypes = [vp, vp, sz, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_uint64)]
    lib.pt_trace_device_noalloc.argtypes = [vp, vp, sz, ctypes.c_int, vp, vp]
    lib.pt_render.argtypes = [vp, ctypes.POINTER(_RenderParams), vp, vp, ctypes.POINTER(_RenderStats)]
    lib.pt_denoise.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double)]
    lib.pt_progressive_create.argtypes = [vp, ctypes.POINTER(_RenderParams), ctypes.POINTER(vp)]
    lib.pt_progressive_step.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(_RenderStats)]
    lib.pt_progressive_destroy.argtypes = [vp]
    lib.pt_device_alloc.argtypes = [sz, ctypes.POINTER(vp)]
    lib.pt_device_free.argtypes = [vp]
    lib.pt_device_copy.argtypes = [vp, vp, sz, ctypes.c_int]
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pt_guide_create.argtypes = [vp, ctypes.POINTER(_GuideParams), ctypes.POINTER(vp)]
    lib.pt_guide_destroy.argtypes = [vp]
    lib.pt_scene_set_guide.argtypes = [vp, vp]
    lib.pt_guide_get_weights.argtypes = [vp, vp, sz]
    lib.pt_guide_set_weights.argtypes = [vp, vp, sz]
    lib.pt_guide_eval.argtypes = [vp, vp, vp, sz, vp]
    lib.pt_guide_pdf.argtypes = [vp, vp, vp, vp, sz, vp]
    lib.pt_guide_sample.argtypes = [vp, vp, vp, vp, sz, vp]
    lib.pt_guide_gradient.argtypes = [vp, vp, sz, vp, fp]
    lib.pt_guide_step.argtypes = [vp, vp, sz, fp]
    lib.pt_guide_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), fp]
    lib.pt_batch_merge.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(_u32p), ctypes.POINTER(sz), vp]
    lib.pt_batch_create.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.pt_batch_destroy.argtypes = [vp]
    lib.pt_render_batch.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(_BatchParams), vp,
                                    ctypes.POINTER(_RenderStats)]
    lib.pt_batch_autotune.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_TuneResult)]
    lib.pt_scene_autotune.argtypes = [vp, vp, ctypes.POINTER(_TuneResult)]
    lib.pt_scene_autotune.restype = ctypes.c_int
    for name in ("pt_batch_merge", "pt_batch_create", "pt_batch_destroy", "pt_render_batch", "pt_batch_autotune"):
        getattr(lib, name).restype = ctypes.c_int
    for name in ("pt_device_alloc", "pt_device_free", "pt_device_copy", "pt_device_info", "pt_generate",
                 "pt_generate_job", "pt_pack", "pt_pack_ex", "pt_bvh_build", "pt_ptsd_add_obj", "pt_ptsd_add_obj_textured", "pt_decode_image", "pt_ptgp_depth", "pt_scene_create",
                 "pt_scene_destroy", "pt_trace",
                 "pt_trace_device", "pt_trace_device_noalloc", "pt_render", "pt_guide_create", "pt_guide_destroy", "pt_scene_set_guide",
                 "pt_guide_get_weights", "pt_guide_set_weights", "pt_guide_eval", "pt_guide_pdf", "pt_guide_sample",
                 "pt_guide_gradient", "pt_guide_step", "pt_guide_stats"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        raise PTError(status, library().pt_last_error().decode(errors="replace"))


def _take_words(ptr, n) -> np.ndarray:
    try:
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()
    finally:
        library().pt_free(ptr)


def version() -> str:
    return library().pt_version().decode()


def device_info() -> tuple[str, int]:
    """(gfx arch name, compute units) of the current HIP device."""
    name = ctypes.create_string_buffer(128)
    cus = ctypes.c_int(0)
    _check(library().pt_device_info(name, 128, ctypes.byref(cus)))
    return name.value.decode(), cus.value


def generate(config: int, seed: int = 0, scale: float = 1.0) -> np.ndarray:
    """A benchmark scene (config 0, 1, 2 or 4) as a PTSD word array."""
    p, n = _u32p(), ctypes.c_size_t(0)
    _check(library().pt_generate(int(config), int(seed), float(scale), ctypes.byref(p), ctypes.byref(n)))
    return _take_words(p, n)


@dataclass
class Job:
    kind: int
    width: int
    height: int
    spp: int
    primitives: int
    seed: int
    ptsd: np.ndarray


def generate_job(seed: int, scale: float = 1.0) -> Job:
    """One job of the batched workload (config 3)."""
    d = _JobDesc()
    p, n = _u32p(), ctypes.c_size_t(0)
    _check(library().pt_generate_job(int(seed), float(scale), ctypes.byref(d), ctypes.byref(p), ctypes.byref(n)))
    return Job(d.kind, d.width, d.height, d.spp, d.primitives, d.seed, _take_words(p, n))


def _words(a) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype != np.uint32:
        a = np.frombuffer(a.tobytes(), dtype=np.uint32)
    return a


def pack(ptsd, device_bvh: bool = False) -> np.ndarray:
    """Builds the BVHs of a PTSD scene and packs it as PTGP v1 words; with
    device_bvh, the trees are built on the GPU (same words)."""
    w = _words(ptsd)
    p, n = _u32p(), ctypes.c_size_t(0)
    if device_bvh:
        _check(library().pt_pack_ex(w.ctypes.data, w.size, 1, ctypes.byref(p), ctypes.byref(n)))
    else:
        _check(library().pt_pack(w.ctypes.data, w.size, ctypes.byref(p), ctypes.byref(n)))
    return _take_words(p, n)


def decode_image(data: bytes) -> np.ndarray:
    """decode_image: a PNG or PPM file's bytes as an (h, w, 3) uint8 sRGB array, rows top first."""
    data = bytes(data)
    w, h, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.POINTER(ctypes.c_uint8)()
    _check(library().pt_decode_image(data, len(data), ctypes.byref(w), ctypes.byref(h), ctypes.byref(p)))
    try:
        n = 3 * w.value * h.value
        return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(h.value, w.value, 3)
    finally:
        library().pt_free(p)


def add_obj(ptsd, obj_text, base, height: float, material: int, turn_deg: float = 0.0, mtl_text=None,
            images=None) -> np.ndarray:
    """load_obj: the OBJ text's triangles appended to a PTSD scene, with
    materials from mtl_text (the library every mtllib names) or the given
    material index; images: {name: image} for the MTL's map_Kd names
    (textures of matte materials), each image an (h, w, 3) uint8 sRGB array
    or a PNG/PPM file's bytes (decode_image). Returns the new PTSD words."""
    w = _words(ptsd)
    data = obj_text.encode() if isinstance(obj_text, str) else bytes(obj_text)
    mtl = None if mtl_text is None else (mtl_text.encode() if isinstance(mtl_text, str) else bytes(mtl_text))
    b = (ctypes.c_float * 3)(*[float(x) for x in base])
    p, n = _u32p(), ctypes.c_size_t(0)
    if images is None:
        _check(library().pt_ptsd_add_obj(w.ctypes.data, w.size, data, len(data), mtl, 0 if mtl is None else len(mtl),
                                         b, float(height), int(material), float(turn_deg), ctypes.byref(p),
                                         ctypes.byref(n)))
    else:
        keep = [(k.encode(), decode_image(v) if isinstance(v, (bytes, bytearray, memoryview))
                 else np.ascontiguousarray(v, dtype=np.uint8)) for k, v in images.items()]
        arr = (_Image * max(1, len(keep)))()
        for i, (name, img) in enumerate(keep):
            if img.ndim != 3 or img.shape[2] != 3:
                raise ValueError(f"image {name.decode()!r}: expected a (height, width, 3) array of sRGB bytes")
            arr[i] = _Image(name, img.shape[1], img.shape[0], img.ctypes.data)
        _check(library().pt_ptsd_add_obj_textured(w.ctypes.data, w.size, data, len(data), mtl,
                                                  0 if mtl is None else len(mtl), arr, len(keep), b, float(height),
                                                  int(material), float(turn_deg), ctypes.byref(p), ctypes.byref(n)))
    return _take_words(p, n)


def bvh_build(boxes, device: bool = False):
    """bvh::build over boxes of shape (n, 6) float32 (min xyz, max xyz).
    Returns (nodes as an (k, 32) uint32 array of 4-wide nodes, order, ms)."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    nodes, order = _u32p(), _u32p()
    k, ms = ctypes.c_size_t(0), ctypes.c_double(0)
    _check(library().pt_bvh_build(b.ctypes.data, b.shape[0], 1 if device else 0, ctypes.byref(nodes), ctypes.byref(k),
                                  ctypes.byref(order), ctypes.byref(ms)))
    words = _take_words(nodes, ctypes.c_size_t(32 * k.value)).reshape(-1, 32)
    return words, _take_words(order, ctypes.c_size_t(b.shape[0])), ms.value


def ptgp_depth(ptgp) -> int:
    w = _words(ptgp)
    d = ctypes.c_int(0)
    _check(library().pt_ptgp_depth(w.ctypes.data, w.size, ctypes.byref(d)))
    return d.value


@dataclass
class RenderStats:
    rays: int
    primary_rays: int
    shadow_rays: int
    render_ms: float
    extend_ms: float
    shadow_ms: float
    visits: tuple = (0, 0, 0, 0, 0, 0, 0)
    batches: int = 0
    samples_per_batch: int = 0
    guide_ms: float = 0.0
    guide_queries: int = 0

    @property
    def mrays_per_s(self) -> float:
        return self.rays / (self.render_ms * 1e-3) / 1e6 if self.render_ms > 0 else 0.0


def rays_array(o, d, tmin=0.001, tmax=1000.0) -> np.ndarray:
    """n x 8 float32 rays (o, tmin, d, tmax) from origin and direction arrays."""
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = max(len(o), len(d))
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


class Scene:
    """A packed scene resident on the GPU (pt_scene)."""

    def __init__(self, ptgp):
        self._words = _words(ptgp)
        self._guides = []
        h = ctypes.c_void_p()
        _check(library().pt_scene_create(self._words.ctypes.data, self._words.size, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            for g in self._guides:   # a guide lives on its scene's stream: it goes first
                g.close()
            _check(library().pt_scene_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def words(self) -> np.ndarray:
        return self._words

    def trace(self, rays, any_hit: bool = False) -> np.ndarray:
        """Closest (or any) hit of n rays given as n x 8 float32 (o, tmin, d, tmax)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros(len(r), dtype=HIT_DTYPE)
        _check(library().pt_trace(self._h, r.ctypes.data, len(r), 1 if any_hit else 0, hits.ctypes.data))
        return hits

    def trace_device(self, d_rays: int, n: int, d_hits: int, any_hit: bool = False, count: bool = False):
        """Traversal over device buffers (addresses). Returns (kernel ms, visit totals or None)."""
        ms = ctypes.c_float(0)
        vis = (ctypes.c_uint64 * 6)()
        _check(library().pt_trace_device(self._h, ctypes.c_void_p(d_rays), n, 1 if any_hit else 0,
                                         ctypes.c_void_p(d_hits), ctypes.byref(ms), vis if count else None))
        return ms.value, (tuple(vis) if count else None)

    def set_guide(self, guide: "Guide | None") -> None:
        """Attaches a learned guide for guided / training renders (None detaches)."""
        _check(library().pt_scene_set_guide(self._h, guide._h if guide is not None else None))

    def render(self, width: int, height: int, spp: int, max_bounces: int = 10, seed: int = 0,
               first_sample: int = 0, rows: tuple[int, int] | None = None, nee: bool = True, aux: bool = False,
               graph: bool = True, timing: bool = False, count: bool = False, samples_per_batch: int = 0,
               stack_entries: int = 0, device_out: "DeviceBuffer | None" = None, tail: bool = True,
               guide: bool = False, train: bool = False, per_ray: bool = False):
        """Renders rows [j0, j1) (from the bottom; default all). Returns (rgb, aux, stats) with rgb of shape
        (rows, width, 3), the band's top row first, as render_row_linear writes rows. guide / train: sample
        (and train) the attached guide at diffuse bounces. per_ray: camera rays walk one per lane in their own
        order, not as 64-ray packets (which pick another of two tied hits now and then)."""
        j0, j1 = rows if rows else (0, height)
        p = _RenderParams(width, height, spp, max_bounces, seed, first_sample, j0, j1,
                          (0 if nee else RENDER_NO_NEE) | (0 if graph else RENDER_NO_GRAPH) |
                          (RENDER_TIMING if timing else 0) | (RENDER_COUNT if count else 0) |
                          (RENDER_DEVICE_OUTPUT if device_out is not None else 0) |
                          (0 if tail else RENDER_NO_TAIL) | (RENDER_GUIDE if guide else 0) |
                          (RENDER_TRAIN if train else 0) | (RENDER_PER_RAY if per_ray else 0), samples_per_batch,
                          stack_entries)
        nrows = j1 - j0
        st = _RenderStats()
        if device_out is not None:
            # The image stays in HBM, in device_out; no aux in this mode.
            if device_out.nbytes < 12 * max(nrows, 0) * width or aux:
                raise ValueError("device_out must hold 12 bytes a pixel, and takes no aux")
            _check(library().pt_render(self._h, ctypes.byref(p), ctypes.c_void_p(device_out.ptr), None,
                                       ctypes.byref(st)))
            rgb, ax = None, None
        else:
            rgb = np.zeros((max(nrows, 0), width, 3), np.float32)
            ax = np.zeros((max(nrows, 0), width, 7), np.float32) if aux else None
            _check(library().pt_render(self._h, ctypes.byref(p), rgb.ctypes.data, ax.ctypes.data if aux else None,
                                       ctypes.byref(st)))
        stats = RenderStats(st.rays, st.primary_rays, st.shadow_rays, st.render_ms, st.extend_ms, st.shadow_ms,
                            tuple(st.visits), st.batches, st.samples_per_batch, st.guide_ms, st.guide_queries)
        return rgb, ax, stats


    def autotune(self, width: int, height: int, spp: int, max_bounces: int = 10, seed: int = 0,
                 rows: tuple[int, int] | None = None, nee: bool = True, tail: bool = True,
                 guide: bool = False) -> "TuneResult":
        """pt_scene_autotune: times this render under candidate launch shapes (stack entries x traversal blocks
        per CU, then smaller batches) and keeps the fastest for the scene's later renders."""
        j0, j1 = rows if rows else (0, height)
        p = _RenderParams(width, height, spp, max_bounces, seed, 0, j0, j1,
                          (0 if nee else RENDER_NO_NEE) | (0 if tail else RENDER_NO_TAIL) |
                          (RENDER_GUIDE if guide else 0), 0, 0)
        r = _TuneResult()
        _check(library().pt_scene_autotune(self._h, ctypes.byref(p), ctypes.byref(r)))
        tried = [(t.stack_entries, t.blocks_per_cu, t.streams, t.max_paths, t.mrays_per_s) for t in r.tried[:r.n_tried]]
        return TuneResult(r.stack_entries, r.blocks_per_cu, r.streams, r.max_paths, tried)


class Progressive:
    """A progressive render (pt_progressive_*): `spp` samples in adaptive steps (1, 2, 4 ... up to 64 while a
    step stays under 20 ms), as the reference's browser demo renders. step() returns (done, step_spp, image of
    the running average)."""

    def __init__(self, scene: "Scene", width: int, height: int, spp: int, max_bounces: int = 10, seed: int = 0,
                 first_sample: int = 0, nee: bool = True, per_ray: bool = False):
        self.width, self.height, self.spp = width, height, spp
        p = _RenderParams(width, height, spp, max_bounces, seed, first_sample, 0, 0,
                          (0 if nee else RENDER_NO_NEE) | (RENDER_PER_RAY if per_ray else 0), 0, 0)
        h = ctypes.c_void_p()
        _check(library().pt_progressive_create(scene._h, ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.done = 0

    def step(self):
        img = np.zeros((self.height, self.width, 3), np.float32)
        done, n = ctypes.c_int(0), ctypes.c_int(0)
        _check(library().pt_progressive_step(self._h, img.ctypes.data, ctypes.byref(done), ctypes.byref(n), None))
        self.done = done.value
        return done.value, n.value, img

    def close(self):
        if self._h:
            library().pt_progressive_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def denoise(rgb: np.ndarray, aux: np.ndarray, spp: int) -> tuple[np.ndarray, float]:
    """The reference's denoiser (denoise.hpp, as main.cpp's --denoise step feeds it) on the device, over a
    render's rgb (h, w, 3) and aux (h, w, 7) outputs from `spp` samples. Returns (image, device_ms)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    aux = np.ascontiguousarray(aux, dtype=np.float32)
    h, w = rgb.shape[:2]
    assert rgb.shape == (h, w, 3) and aux.shape == (h, w, 7)
    out = np.zeros_like(rgb)
    ms = ctypes.c_double(0)
    _check(library().pt_denoise(w, h, int(spp), rgb.ctypes.data, aux.ctypes.data, out.ctypes.data, 0,
                                ctypes.byref(ms)))
    return out, ms.value


def _f32(a, cols: int) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, cols)


class Guide:
    """A learned path guide on the GPU (pt_guide): an MLP giving an 8-lobe vMF mixture per shading point.
    Training samples are rows of 12 floats: position, weight, normal, 0, direction, 0."""

    def __init__(self, scene: Scene, seed: int = 0, learning_rate: float = 0.0, box=None,
                 steps_per_batch: int = 0, minibatch: int = 0, weight_clamp: float = 0.0, bounces: int = 0):
        p = _GuideParams()
        p.seed = int(seed)
        p.learning_rate = float(learning_rate)
        if box is not None:
            lo, hi = box
            p.box_lo[:] = [float(v) for v in lo]
            p.box_hi[:] = [float(v) for v in hi]
        p.steps_per_batch = int(steps_per_batch)
        p.minibatch = int(minibatch)
        p.weight_clamp = float(weight_clamp)
        p.bounces = int(bounces)
        h = ctypes.c_void_p()
        _check(library().pt_guide_create(scene._h, ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.scene = scene
        scene._guides.append(self)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(library().pt_guide_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def weights(self) -> np.ndarray:
        w = np.zeros(GUIDE_PARAMS, np.float32)
        _check(library().pt_guide_get_weights(self._h, w.ctypes.data, w.size))
        return w

    def set_weights(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32).ravel()
        _check(library().pt_guide_set_weights(self._h, w.ctypes.data, w.size))

    def eval(self, pos, nrm) -> np.ndarray:
        x, n = _f32(pos, 3), _f32(nrm, 3)
        out = np.zeros((len(x), GUIDE_OUTPUTS), np.float32)
        _check(library().pt_guide_eval(self._h, x.ctypes.data, n.ctypes.data, len(x), out.ctypes.data))
        return out

    def pdf(self, pos, nrm, dirs) -> np.ndarray:
        x, n, d = _f32(pos, 3), _f32(nrm, 3), _f32(dirs, 3)
        out = np.zeros(len(x), np.float32)
        _check(library().pt_guide_pdf(self._h, x.ctypes.data, n.ctypes.data, d.ctypes.data, len(x), out.ctypes.data))
        return out

    def sample(self, pos, nrm, u) -> np.ndarray:
        x, n, uu = _f32(pos, 3), _f32(nrm, 3), _f32(u, 3)
        out = np.zeros((len(x), 4), np.float32)
        _check(library().pt_guide_sample(self._h, x.ctypes.data, n.ctypes.data, uu.ctypes.data, len(x),
                                         out.ctypes.data))
        return out

    def gradient(self, samples) -> tuple[np.ndarray, float]:
        s = _f32(samples, 12)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = ctypes.c_float(0)
        _check(library().pt_guide_gradient(self._h, s.ctypes.data, len(s), g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def step(self, samples) -> float:
        s = _f32(samples, 12)
        loss = ctypes.c_float(0)
        _check(library().pt_guide_step(self._h, s.ctypes.data, len(s), ctypes.byref(loss)))
        return loss.value

    def stats(self) -> tuple[int, int, float]:
        st, sa, lo = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_float(0)
        _check(library().pt_guide_stats(self._h, ctypes.byref(st), ctypes.byref(sa), ctypes.byref(lo)))
        return st.value, sa.value, lo.value


class DeviceBuffer:
    """A block of HBM owned by the library's allocator (pt_device_alloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(library().pt_device_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(library().pt_device_copy(ctypes.c_void_p(self.ptr), a.ctypes.data, a.nbytes, 1))

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        _check(library().pt_device_copy(a.ctypes.data, ctypes.c_void_p(self.ptr), a.nbytes, 0))
        return a

    def free(self) -> None:
        if self.ptr:
            _check(library().pt_device_free(ctypes.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def write_pfm(path: str, rgb: np.ndarray) -> None:
    """Little-endian PFM, bottom row first, as image_write.hpp:15-28 writes it."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(rgb[::-1], dtype="<f4").tobytes())


def read_pfm(path: str) -> np.ndarray:
    """(h, w, 3) float32, top row first."""
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF", "not a color PFM"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, 3)[::-1].astype(np.float32)


# ---- batched renders ------------------------------------------------------------

@dataclass
class BatchJob:
    """One job of a batched render: scene number in the batch, image size, spp, seed, and optionally its
    own camera rows (4 x 4 floats: origin, corner, across, up) instead of its scene's."""
    scene: int
    width: int
    height: int
    spp: int
    seed: int = 0
    first_sample: int = 0
    nee: bool = True
    camera: "np.ndarray | None" = None


@dataclass
class BatchPart:
    root: int
    n_large: int
    large_off: int
    n_lights: int
    lights_off: int
    flags: int
    camera: np.ndarray


def _scene_arrays(ptgps):
    ws = [_words(p) for p in ptgps]
    ptrs = (ctypes.c_void_p * len(ws))(*[w.ctypes.data for w in ws])
    sizes = (ctypes.c_size_t * len(ws))(*[w.size for w in ws])
    return ws, ptrs, sizes


def batch_merge(ptgps) -> tuple[np.ndarray, list[BatchPart]]:
    """Host only: several flat PTGP scenes merged into one PTGP buffer, and where each scene went."""
    ws, ptrs, sizes = _scene_arrays(ptgps)
    parts = (_BatchPart * len(ws))()
    p, n = _u32p(), ctypes.c_size_t(0)
    _check(library().pt_batch_merge(ptrs, sizes, len(ws), ctypes.byref(p), ctypes.byref(n), parts))
    out = [BatchPart(q.root, q.n_large, q.large_off, q.n_lights, q.lights_off, q.flags,
                     np.array(q.camera[:], np.float32).reshape(4, 4)) for q in parts]
    return _take_words(p, n), out


def _job_array(jobs):
    arr = (_BatchJob * len(jobs))()
    for k, j in enumerate(jobs):
        a = arr[k]
        a.scene, a.width, a.height, a.spp = int(j.scene), int(j.width), int(j.height), int(j.spp)
        a.seed, a.first_sample = int(j.seed), int(j.first_sample)
        a.flags = (0 if j.nee else RENDER_NO_NEE) | (JOB_CAMERA if j.camera is not None else 0)
        if j.camera is not None:
            a.camera[:] = [float(v) for v in np.asarray(j.camera, np.float32).ravel()[:16]]
    return arr


@dataclass
class TuneResult:
    stack_entries: int
    blocks_per_cu: int
    streams: int
    max_paths: int
    tried: list


class Batch:
    """Several flat scenes merged on the GPU (pt_batch), rendered many jobs at a time."""

    def __init__(self, ptgps):
        self._ws, ptrs, sizes = _scene_arrays(ptgps)
        h = ctypes.c_void_p()
        _check(library().pt_batch_create(ptrs, sizes, len(self._ws), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(library().pt_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, jobs, max_bounces: int = 10, graph: bool = True, tail: bool = True, stack_entries: int = 0,
               blocks_per_cu: int = 0, streams: int = 0, max_paths: int = 0, device_out: "DeviceBuffer | None" = None,
               shared: bool = True, per_ray: bool = False):
        """Renders every job. Returns (images, stats): one (height, width, 3) array a job, top row first
        (None with device_out: the images stay in that device buffer, one after another). shared: the jobs'
        paths in shared-queue passes; False: each job's own bounce loops over `streams` HIP streams. per_ray:
        camera rays walked one per lane (each job's image then equals pt_render's with per_ray bit for bit)."""
        arr = _job_array(jobs)
        flags = (0 if graph else RENDER_NO_GRAPH) | (0 if tail else RENDER_NO_TAIL) | \
            (RENDER_DEVICE_OUTPUT if device_out is not None else 0) | (0 if shared else BATCH_STREAMS) | \
            (RENDER_PER_RAY if per_ray else 0)
        p = _BatchParams(int(max_bounces), flags, int(stack_entries), int(blocks_per_cu), int(streams), 0,
                         int(max_paths))
        sizes = [int(j.width) * int(j.height) for j in jobs]
        st = _RenderStats()
        if device_out is not None:
            if device_out.nbytes < 12 * sum(sizes):
                raise ValueError("device_out holds fewer than 12 bytes per pixel of the jobs")
            _check(library().pt_render_batch(self._h, arr, len(jobs), ctypes.byref(p), ctypes.c_void_p(device_out.ptr),
                                             ctypes.byref(st)))
            stats = RenderStats(st.rays, st.primary_rays, st.shadow_rays, st.render_ms, st.extend_ms, st.shadow_ms,
                                tuple(st.visits), st.batches, st.samples_per_batch)
            return None, stats
        rgb = np.zeros(3 * sum(sizes), np.float32)
        _check(library().pt_render_batch(self._h, arr, len(jobs), ctypes.byref(p), rgb.ctypes.data, ctypes.byref(st)))
        out, at = [], 0
        for j, n in zip(jobs, sizes):
            out.append(rgb[3 * at:3 * (at + n)].reshape(int(j.height), int(j.width), 3))
            at += n
        stats = RenderStats(st.rays, st.primary_rays, st.shadow_rays, st.render_ms, st.extend_ms, st.shadow_ms,
                            tuple(st.visits), st.batches, st.samples_per_batch)
        return out, stats

    def autotune(self, jobs, max_bounces: int = 10) -> TuneResult:
        """Times short renders of the jobs under candidate launch shapes and keeps the fastest for this batch."""
        arr = _job_array(jobs)
        r = _TuneResult()
        _check(library().pt_batch_autotune(self._h, arr, len(jobs), int(max_bounces), ctypes.byref(r)))
        tried = [(t.stack_entries, t.blocks_per_cu, t.streams, t.max_paths, t.mrays_per_s) for t in r.tried[:r.n_tried]]
        return TuneResult(r.stack_entries, r.blocks_per_cu, r.streams, r.max_paths, tried)
