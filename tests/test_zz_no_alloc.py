"""pt_trace_device_noalloc (traversal with its work counter and stack spill
area in the library's per-device scratch array) against pt_trace_device on
the same device rays: identical hits, bit for bit, on two calls in a row,
the second with fewer rays, for flat trees, deep trees (a reduced grid) and
instanced scenes; closest and any hit; on the scene's stream and on a
caller's stream."""
import ctypes

import numpy as np
import pytest

import photon_tracer_amd as pt
import rays as R

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime the library itself links (already loaded with it), for a stream of the caller's own."""
    pt.library()
    hip = ctypes.CDLL("libamdhip64.so.7")
    for name in ("hipStreamCreate", "hipStreamSynchronize", "hipStreamDestroy"):
        getattr(hip, name).restype = ctypes.c_int
    return hip


def run(lib, scene, rays, any_hit, noalloc, stream=None):
    n = len(rays)
    dr = pt.DeviceBuffer(max(rays.nbytes, 16))
    dh = pt.DeviceBuffer(max(n * pt.HIT_DTYPE.itemsize, 16))
    try:
        dr.upload(np.ascontiguousarray(rays, np.float32))
        # Stale hits from an earlier call must not show through.
        dh.upload(np.full(n * 6, 0x7F, np.uint32))
        if noalloc:
            pt._check(lib.pt_trace_device_noalloc(scene._h, ctypes.c_void_p(dr.ptr), n, int(any_hit),
                                                  ctypes.c_void_p(dh.ptr), ctypes.c_void_p(stream)))
            if stream is not None:
                assert hip_runtime().hipStreamSynchronize(ctypes.c_void_p(stream)) == 0
        else:
            pt._check(lib.pt_trace_device(scene._h, ctypes.c_void_p(dr.ptr), n, int(any_hit), ctypes.c_void_p(dh.ptr),
                                          None, None))
        return dh.download(np.zeros(n, pt.HIT_DTYPE))
    finally:
        dr.free()
        dh.free()


def scenes():
    return [("config0", pt.pack(pt.generate(0, 2, 1.0))),
            ("config1 1M (reduced grid)", pt.pack(pt.generate(1, 1, 1.0))),
            ("config2 2%", pt.pack(pt.generate(2, 3, 0.02)))]


@pytest.mark.parametrize("case", scenes(), ids=lambda c: c[0])
def test_noalloc_trace_equals_trace_device(case):
    name, ptgp = case
    lib = pt.library()
    big = R.mixed_batch(ptgp, 40000, 3)
    small = R.mixed_batch(ptgp, 3000, 4)
    with pt.Scene(ptgp) as s:
        for any_hit in (False, True):
            for label, rays in (("first", big), ("second, smaller", small)):
                want = run(lib, s, rays, any_hit, noalloc=False)
                got = run(lib, s, rays, any_hit, noalloc=True)
                same = np.array_equal(want.view(np.uint32), got.view(np.uint32))
                print(f"{name}: {label} call, {len(rays)} rays, any_hit={any_hit}: identical={same}", flush=True)
                assert same


def test_noalloc_on_a_callers_stream():
    ptgp = pt.pack(pt.generate(1, 2, 0.05))
    lib = pt.library()
    hip = hip_runtime()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        rays = R.mixed_batch(ptgp, 20000, 9)
        with pt.Scene(ptgp) as s:
            want = run(lib, s, rays, False, noalloc=False)
            got = run(lib, s, rays, False, noalloc=True, stream=stream.value)
            got2 = run(lib, s, rays[:777], False, noalloc=True, stream=stream.value)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32))
        assert np.array_equal(want[:777].view(np.uint32), got2.view(np.uint32))
    finally:
        assert hip.hipStreamDestroy(stream) == 0


def test_noalloc_rejects_bad_arguments():
    ptgp = pt.pack(pt.generate(0, 2, 0.25))
    lib = pt.library()
    with pt.Scene(ptgp) as s:
        assert lib.pt_trace_device_noalloc(s._h, None, 10, 0, None, None) != 0
        assert lib.pt_trace_device_noalloc(None, None, 10, 0, None, None) != 0
        assert lib.pt_trace_device_noalloc(s._h, None, 0, 0, None, None) == 0
