"""Textured mattes on the GPU (image_texture, texture.hpp:13-33; the albedo
times the texture at the hit's texture coordinates, material.hpp:72): the
path loop against the oracle integrator pixel by pixel, and the reference's
own statistical render fixture for its textured "spot crowd" scene
(tests/render_test.py:55, tests/reference.json)."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc
from scenes import textured_quads

pytestmark = pytest.mark.gpu


def _scene(name):
    if name == "spot crowd":
        return pt.pack(rc.load_spot())
    rng = np.random.default_rng(4)
    return pt.pack(textured_quads(rng.integers(0, 256, (37, 23, 3), dtype=np.uint8),
                                  rng.integers(0, 256, (8, 64, 3), dtype=np.uint8)))


@pytest.mark.parametrize("name", ["spot crowd", "quads"])
def test_textured_render_matches_oracle_per_pixel(name):
    """Same per-path random streams and arithmetic as the oracle, texture
    lookups included: 1e-4 relative on 99.9% of values (the others are paths
    that part where the device's sinf/cosf round a direction differently)."""
    ptgp = _scene(name)
    W, H, spp = 96, 54, 8
    with pt.Scene(ptgp) as s:
        rgb, _, st = s.render(W, H, spp, seed=11)
    orgb, _, orays = ob.or_render(ptgp, W, H, spp, seed=11)
    frac, note = rc.per_value_agreement(rgb, orgb)
    print(f"{name}: {frac:.5f} of values within 1e-4 ({note}); rays {st.rays} vs oracle {orays}", flush=True)
    assert frac >= 0.999
    assert abs(st.rays - orays) <= 0.002 * orays + 10
    # the texture is read: the same scene with its textures switched off renders differently
    plain = ptgp.copy()
    m0, ml = int(plain[12]) // 4, int(plain[13]) // 4
    plain[m0 + 9: m0 + ml: 12] = 0
    with pt.Scene(plain) as s:
        untextured, _, _ = s.render(W, H, spp, seed=11)
    assert rc.block_mre(untextured, rgb, block=6) > 0.02


def test_spot_crowd_reference_render_fixture():
    """The reference's statistical check on the GPU image of its textured
    scene: 960x540, 200 spp, seed 1, against the fixture's per-block mean and
    sigma over 16 seeds of the reference's CPU renderer."""
    ref = rc.load_reference()
    with pt.Scene(_scene("spot crowd")) as s:
        rgb, _, st = s.render(rc.WIDTH, rc.HEIGHT, rc.SPP, seed=1)
    ok, msg = rc.passes(rgb, ref["scenes"]["spot crowd"])
    print(f"spot crowd: {msg}; {st.mrays_per_s:.0f} Mrays/s", flush=True)
    assert ok, msg
