"""Pins the oracle (oracle/liboracle.so) before anything is checked against it:
its traversal against the reference's own traversal code (oracle/_ref) bit
for bit, against FP64 intersection, its sampler against the reference's, and
its integrator against the reference renderer and against closed forms."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import rays as R
import render_check as rc

SCENES = ["field", "room", "teapot", "crowd"]


def generated(config, scale, seed=3):
    ptsd = pt.generate(config, seed, scale)
    return ptsd, pt.pack(ptsd)


def all_cases():
    out = [(n, *rc.load_scene(n)) for n in SCENES]
    for cfg, sc in [(0, 1.0), (1, 0.05), (2, 0.02), (4, 0.02)]:
        out.append((f"config{cfg}", *generated(cfg, sc)))
    return out


@pytest.mark.parametrize("case", all_cases(), ids=lambda c: c[0])
def test_oracle_trace_is_the_reference_traversal(case, ref_lib):
    name, ptsd, ptgp = case
    batch = R.mixed_batch(ptgp, 8000, 7)
    t, u, v, kind, prim, inst, _ = ob.or_trace(ptgp, batch)
    with ob.RefWorld(ptsd) as w:
        rk, ri, rs, rt, ru, rv, bad = w.trace(batch)
        ref_any = w.trace(batch, any_hit=True)[0]
    assert bad == 0  # the harness's restated hit agrees with geometry::hit
    ok, oi, os_ = ob.ptgp_to_ref_ids(ptgp, ptsd, kind, prim, inst)
    assert np.array_equal(ok, rk)
    assert np.array_equal(oi, ri)
    assert np.array_equal(os_, rs)
    hit = kind != 0
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(u[hit].view(np.uint32), ru[hit].view(np.uint32))
    assert np.array_equal(v[hit].view(np.uint32), rv[hit].view(np.uint32))
    any_kind = ob.or_trace(ptgp, batch, any_hit=True)[3]
    assert np.array_equal(any_kind, ref_any)


def test_oracle_shadow_rays_match_reference(ref_lib):
    ptsd, ptgp = generated(1, 0.05)
    prim = R.camera_rays(ptgp, 6000, 1)
    t = ob.or_trace(ptgp, prim)[0]
    origins = prim[:, 0:3] + prim[:, 4:7] * t[:, None]
    sh = R.shadow_rays(ptgp, origins.astype(np.float32), 2)
    with ob.RefWorld(ptsd) as w:
        ref_any = w.trace(sh, any_hit=True)[0]
    ours = ob.or_trace(ptgp, sh, any_hit=True)[3]
    assert np.array_equal(ours, ref_any)
    assert 0.05 < ours.mean() < 0.95


@pytest.mark.parametrize("case", all_cases(), ids=lambda c: c[0])
def test_oracle_against_fp64(case):
    """The FP32 path against FP64 intersection of the same candidate
    primitives: the hit primitive agrees on >= 99.99% of rays away from
    self-intersection; where both hit the same primitive with t well above
    t_min, t agrees to 1e-5 relative except for rays grazing a sphere."""
    name, ptsd, ptgp = case
    batch = R.camera_rays(ptgp, 20000, 17)
    t, _, _, kind, prim, inst, _ = ob.or_trace(ptgp, batch)
    t64, k64, p64, i64 = ob.or_trace_f64(ptgp, batch)
    same = (k64 == kind) & (p64 == prim) & (i64 == inst)
    assert same.mean() >= 0.9999, (name, same.mean())
    both = same & (kind == 3)
    if both.any():
        rel = np.abs(t[both] - t64[both]) / t64[both]
        assert np.quantile(rel, 0.9999) <= 1e-5, (name, rel.max())


def test_sampler_is_the_reference_sampler(ref_lib):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        seed = int(rng.integers(0, 2**63))
        i, j = int(rng.integers(0, 4000)), int(rng.integers(0, 4000))
        index = int(rng.integers(0, 2**32))
        pair = int(rng.integers(0, 5))
        assert ob.or_sample_2d(seed, i, j, index, pair) == ob.ref_sample_2d(seed, i, j, index, pair)


def test_sampler_is_stratified():
    # The first 2^k samples of pixel (i, j), pair 0, fall one per cell of a
    # 2^(k/2) x 2^(k/2) grid: Sobol (0,2)-sequence under Owen scrambling.
    pts = np.array([ob.or_sample_2d(1, 3, 4, s, 0) for s in range(256)])
    cells = (pts * 16).astype(int)
    assert len({(a, b) for a, b in cells}) == 256


@pytest.mark.parametrize("name", ["materials", "room", "field"])
def test_oracle_render_matches_reference_renderer(name, ref_lib):
    ptsd, ptgp = rc.load_scene(name)
    W, H, spp = 160, 90, 128
    o, _, orays = ob.or_render(ptgp, W, H, spp, seed=5)
    with ob.RefWorld(ptsd) as w:
        r, _, rrays, _ = w.render(W, H, spp, seed=9)
        r2, _, _, _ = w.render(W, H, spp, seed=10)
    # Two reference seeds give the noise floor of the block statistic.
    noise = rc.block_mre(r2, r, block=10)
    assert rc.block_mre(o, r, block=10) < 1.6 * noise + 1e-3, (rc.block_mre(o, r, block=10), noise)
    assert abs(orays / rrays - 1) < 0.01
    assert np.allclose(o.mean(axis=(0, 1)), r.mean(axis=(0, 1)), rtol=0.01)


def _sphere_light_scene(nee: bool):
    """A matte ground (large sphere, albedo 0.5) under one small sphere light,
    no sky; camera straight down onto the point below the light's side."""
    w = np.zeros(32, np.uint32)
    f = lambda x: np.float32(x).view(np.uint32)
    w[0], w[1], w[2] = 0x44535450, 1, 2 if nee else 0
    w[3:6] = [f(0), f(3), f(0.001)]   # lookfrom
    w[6:9] = [f(0), f(0), f(0)]       # lookat
    w[9] = f(10)
    w[10], w[11], w[12] = 2, 1, 1
    mats = np.zeros(24, np.uint32)
    mats[0:3] = [f(0.5)] * 3
    mats[3] = 0
    mats[8] = f(1.5)
    mats[12 + 3] = 3
    mats[16:19] = [f(4.0)] * 3
    mats[20] = f(1.5)
    sph = np.array([f(1.0), f(1.0), f(0.0), f(0.2), 1], np.uint32)      # light: r 0.2 at (1, 1, 0)
    ground = np.array([f(0.0), f(-1000.0), f(0.0), f(1000.0), 0], np.uint32)
    return np.concatenate([w, mats, sph, ground])


@pytest.mark.parametrize("nee", [True, False])
def test_oracle_direct_light_closed_form(nee):
    """Radiance of the ground at the camera's centre pixel: albedo * E *
    sin^2(theta_max) * cos(theta) for a sphere light fully above the horizon
    (one bounce after the camera ray: depth 2)."""
    ptsd = _sphere_light_scene(nee)
    ptgp = pt.pack(ptsd)
    spp = 4096 if nee else 65536
    rgb, _, _ = ob.or_render(ptgp, 16, 9, spp, depth=2, seed=1, nee=nee)
    centre = rgb[3:6, 6:10].mean(axis=(0, 1))
    # The closed form averaged over those pixels' footprints on the ground
    # (pixel i covers u in [i, i+1] / (W - 1), camera.hpp:45-47).
    cam = R.camera(ptgp)
    u = (6 + np.arange(4 * 64 + 0.5) / 64)[: 4 * 64] / 15
    row_j = 8 - np.arange(3, 6)   # rows 3..5 counted from the bottom
    v = (row_j.min() + np.arange(3 * 64) / 64) / 8
    U, Vv = np.meshgrid(u, v)
    d = cam[1] + cam[2] * U[..., None] + cam[3] * Vv[..., None] - cam[0]
    t = -cam[0][1] / d[..., 1]
    p = cam[0] + d * t[..., None]            # on the plane y = 0 (the r = 1000 ground, to 1e-5)
    c = np.array([1.0, 1.0, 0.0])
    dist2 = ((c - p) ** 2).sum(-1)
    expect = float(np.mean(0.5 * 4.0 * (0.2 ** 2 / dist2) * (1.0 / np.sqrt(dist2))))
    tol = 0.01 if nee else 0.05
    assert np.allclose(centre, expect, rtol=tol), (centre, expect)
