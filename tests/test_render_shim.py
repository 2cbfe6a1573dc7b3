"""include/photon_tracer_render.hpp, the reference-side binding of
render_row_linear (renderer.hpp:411-416) over pt_render, compiled against the
reference's own headers into oracle/_ref/render_shim_check (oracle/Makefile)
where the reference is present. On the GPU: every row rendered by its own
render_row_linear call equals the whole frame from render_frame and the
library's own render of the reference's packed scene, bit for bit, with the
same ray count; without a GPU the binding reports the library's error."""
import os
import subprocess

import numpy as np
import pytest

import photon_tracer_amd as pt
import render_check as rc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "oracle", "_ref", "render_shim_check")
SCENES = {"materials": 0, "field": 1, "room": 3}   # build_world's scene ids (renderer.hpp:151)


def _need_shim():
    if not os.path.exists(SHIM):
        pytest.skip("oracle/_ref/render_shim_check is built only where the reference's headers are")


def test_binding_compiles_links_and_reports_errors():
    _need_shim()
    run = subprocess.run([SHIM], capture_output=True, text=True, timeout=60)
    assert run.returncode == 1 and "usage" in run.stderr
    try:
        pt.device_info()
        return   # a GPU is here: the renders are the GPU test's
    except pt.PTError:
        pass
    run = subprocess.run([SHIM, "0", "16", "9", "1", "1", os.devnull], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and run.stderr.strip()   # the library's message, thrown as runtime_error


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCENES))
def test_row_calls_equal_the_frame_and_the_library_render(name, tmp_path):
    _need_shim()
    W, H, spp, seed = 64, 36, 8, 5
    out = tmp_path / "img.bin"
    run = subprocess.run([SHIM, str(SCENES[name]), str(W), str(H), str(spp), str(seed), str(out)],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    by_rows, by_frame = (int(x) for x in run.stdout.split()[1:3])
    img = np.fromfile(out, np.float32).reshape(2, H, W, 3)
    with pt.Scene(rc.load_scene(name)[1]) as s:
        want, _, st = s.render(W, H, spp, seed=seed)
    print(f"{name}: rays by rows {by_rows}, by frame {by_frame}, library {st.rays}", flush=True)
    assert np.array_equal(img[0], img[1]) and np.array_equal(img[1], want)
    assert by_rows == by_frame == st.rays
