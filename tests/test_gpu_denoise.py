"""The denoiser (pt_denoise) against the reference's own (denoise.hpp, fed as
main.cpp's --denoise step feeds it, compiled from the reference's sources in
oracle/_ref): same inputs, same outputs up to the rounding of the device's
exponential, on rendered images (render color and aux from the GPU) and on
random ones with sky pixels (zero normals) and black albedo."""
import ctypes

import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc

pytestmark = pytest.mark.gpu


def close(a, b):
    """Element-wise |a - b| within 1e-4 of max(|b|, 1e-3): the filter's arithmetic in the reference's order,
    only the exponential's last bit differs, carried through five passes (measured: at most 2e-5)."""
    return np.abs(a - b) <= 1e-4 * np.maximum(np.abs(b), 1e-3)


@pytest.mark.parametrize("name,spp", [("room", 16), ("field", 4), ("materials", 64)])
def test_denoise_matches_reference_on_renders(name, spp, ref_lib):
    ptgp = rc.load_scene(name)[1]
    W, H = 200, 120
    with pt.Scene(ptgp) as s:
        rgb, aux, _ = s.render(W, H, spp, seed=5, aux=True)
    ours, ms = pt.denoise(rgb, aux, spp)
    ref, secs = ob.ref_denoise(rgb, aux, spp)
    ok = close(ours, ref)
    rel = np.abs(ours - ref) / np.maximum(np.abs(ref), 1e-3)
    print(f"{name}: {ok.mean():.6f} within tolerance, max rel {rel.max():.2e}; device {ms:.3f} ms, "
          f"reference {secs * 1e3:.1f} ms; mean squared change {np.mean((ours - rgb) ** 2):.2e}", flush=True)
    assert ok.all()
    assert np.var(ours) < np.var(rgb)   # it did smooth


def test_denoise_matches_reference_on_random_inputs(ref_lib):
    rng = np.random.default_rng(3)
    H, W = 70, 90
    rgb = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    albedo = rng.uniform(0, 1, (H, W, 3))
    albedo[10:20, 10:30] = 0                                   # black surfaces: the 0.01 floor
    normal = rng.normal(size=(H, W, 3))
    normal[40:55, 50:80] = 0                                   # sky
    normal[5:8] *= 1e-7                                        # below the 1e-6 length: treated as sky
    lum2 = rng.uniform(0, 4, (H, W, 1))
    aux = np.concatenate([albedo, normal, lum2], 2).astype(np.float32)
    ours, _ = pt.denoise(rgb, aux, 7)
    ref, _ = ob.ref_denoise(rgb, aux, 7)
    assert close(ours, ref).all()


def test_device_pointers_and_bad_arguments():
    rng = np.random.default_rng(4)
    H, W = 33, 47
    rgb = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    aux = np.concatenate([rng.uniform(0, 1, (H, W, 3)), rng.normal(size=(H, W, 3)), rng.uniform(0, 1, (H, W, 1))],
                         2).astype(np.float32)
    host, _ = pt.denoise(rgb, aux, 9)
    bufs = [pt.DeviceBuffer(a.nbytes) for a in (rgb, aux, rgb)]
    bufs[0].upload(rgb)
    bufs[1].upload(aux)
    lib = pt.library()
    ms = ctypes.c_double(0)
    assert lib.pt_denoise(W, H, 9, ctypes.c_void_p(bufs[0].ptr), ctypes.c_void_p(bufs[1].ptr),
                          ctypes.c_void_p(bufs[2].ptr), 1, ctypes.byref(ms)) == 0
    dev = bufs[2].download(np.zeros_like(rgb))
    assert np.array_equal(dev, host)
    for b in bufs:
        b.free()
    p = rgb.ctypes.data
    assert lib.pt_denoise(0, H, 9, p, aux.ctypes.data, p, 0, None) != 0          # size
    assert lib.pt_denoise(W, H, 0, p, aux.ctypes.data, host.ctypes.data, 0, None) != 0   # spp
    assert lib.pt_denoise(W, H, 9, p, aux.ctypes.data, p, 0, None) != 0          # out aliases rgb
    assert lib.pt_denoise(W, H, 9, p, aux.ctypes.data, host.ctypes.data, 6, None) != 0   # unknown flag
