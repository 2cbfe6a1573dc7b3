"""The GPU path loop (pt_render) against the oracle integrator, the
reference's own render fixtures, and its determinism guarantees."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc

pytestmark = pytest.mark.gpu


def scene_of(name):
    if name.startswith("config"):
        cfg, sc = {"config0": (0, 0.25), "config1": (1, 0.05), "config2": (2, 0.02), "config4": (4, 0.02)}[name]
        return pt.pack(pt.generate(cfg, 3, sc))
    return rc.load_scene(name)[1]


@pytest.mark.parametrize("name", ["materials", "field", "room", "teapot", "crowd", "config0", "config1", "config2",
                                  "config4"])
def test_render_matches_oracle_per_pixel(name):
    """Same per-path random streams and the same arithmetic: the pixels equal
    the oracle's to float rounding, except where a path parts because the
    device's sinf/cosf round a sample direction differently from the host's
    libm (tolerance: 1e-4 relative on 99.9% of values; the others printed)."""
    ptgp = scene_of(name)
    W, H, spp = 96, 54, 8
    with pt.Scene(ptgp) as s:
        rgb, _, st = s.render(W, H, spp, seed=11)
    orgb, _, orays = ob.or_render(ptgp, W, H, spp, seed=11)
    frac, note = rc.per_value_agreement(rgb, orgb)
    print(f"{name}: {frac:.5f} of values within 1e-4 ({note}); rays {st.rays} vs oracle {orays}", flush=True)
    assert frac >= 0.999
    assert abs(st.rays - orays) <= 0.002 * orays + 10
    assert abs(rgb.mean() / max(orgb.mean(), 1e-9) - 1) < 0.02


def test_determinism_and_batching():
    """Camera packets are 64 consecutive paths of a batch, so a render repeats bit for bit; walked per ray,
    paths are keyed by (pixel, sample) alone and the batch size changes no bit either."""
    ptgp = scene_of("field")
    with pt.Scene(ptgp) as s:
        a, _, sa = s.render(128, 72, 16, seed=5)
        b, _, sb = s.render(128, 72, 16, seed=5)
        d, _, _ = s.render(128, 72, 16, seed=5, graph=False)
        e, _, _ = s.render(128, 72, 16, seed=6)
        r, _, sr = s.render(128, 72, 16, seed=5, per_ray=True)
        c, _, sc = s.render(128, 72, 16, seed=5, samples_per_batch=1, per_ray=True)
    assert np.array_equal(a, b) and sa.rays == sb.rays
    assert np.array_equal(a, d)       # graph replay == eager launches
    assert not np.array_equal(a, e)
    assert np.array_equal(r, c) and sr.rays == sc.rays   # batch size does not matter


@pytest.mark.parametrize("name", ["field", "room", "crowd", "config1"])
def test_tail_launch_equals_stage_by_stage(name):
    """Finishing the last, small queues path by path in one launch gives the
    same bytes and ray counts as running every bounce stage by stage."""
    ptgp = scene_of(name)
    with pt.Scene(ptgp) as s:
        a, _, sa = s.render(160, 90, 16, seed=21)
        b, _, sb = s.render(160, 90, 16, seed=21, tail=False)
    assert np.array_equal(a, b)
    assert (sa.rays, sa.shadow_rays, sa.primary_rays) == (sb.rays, sb.shadow_rays, sb.primary_rays)


@pytest.mark.parametrize("name,scale,size", [("field", None, (192, 108, 16)), ("room", None, (192, 108, 16)),
                                             ("materials", None, (192, 108, 16)), ("config0", None, (192, 108, 16)),
                                             ("config1", 1.0, (480, 270, 16)), ("teapot", None, (192, 108, 16))])
def test_camera_packets_agree_with_per_ray_walk(name, scale, size):
    """Camera rays walked 64 at a time as packets (the default for scenes without instances) find each ray's
    own nearest hit; only a tie, two candidates at one t, can go the other way. The images then agree bit for
    bit on at least 99.99% of pixels (the tie bound for primitive ids) and the ray counts to 1e-4. config1 at
    scale 1 is the 1,003,520-triangle mesh, teapot the reference's teapot mesh."""
    ptgp = pt.pack(pt.generate(1, 1, scale)) if scale else scene_of(name)
    W, H, spp = size
    with pt.Scene(ptgp) as s:
        pk, _, sp = s.render(W, H, spp, seed=13)
        pr, _, sr = s.render(W, H, spp, seed=13, per_ray=True)
        pk2, _, _ = s.render(W, H, spp, seed=13, stack_entries=16)
    same = np.all(pk.view(np.uint32) == pr.view(np.uint32), axis=2)
    print(f"{name}: {int((~same).sum())} of {same.size} pixels differ; rays {sp.rays} vs {sr.rays}", flush=True)
    assert same.mean() >= 0.9999
    assert abs(sp.rays - sr.rays) <= 1e-4 * sr.rays
    assert sp.primary_rays == sr.primary_rays
    assert np.array_equal(pk, pk2)    # the launch shape does not regroup packets


def test_device_output_equals_host_output():
    ptgp = scene_of("field")
    with pt.Scene(ptgp) as s:
        host, _, st = s.render(96, 54, 8, seed=9)
        buf = pt.DeviceBuffer(96 * 54 * 12)
        _, _, sd = s.render(96, 54, 8, seed=9, device_out=buf)
        dev = buf.download(np.zeros((54, 96, 3), np.float32))
        buf.free()
    assert np.array_equal(host, dev) and st.rays == sd.rays


def test_graph_reuse_across_renders():
    """The cached graph reads seed and first sample from HBM: a second seed
    through the cached graph equals a fresh eager render of that seed."""
    ptgp = scene_of("room")
    with pt.Scene(ptgp) as s:
        s.render(80, 45, 8, seed=1)
        a, _, _ = s.render(80, 45, 8, seed=2, first_sample=8)
        b, _, _ = s.render(80, 45, 8, seed=2, first_sample=8, graph=False)
    assert np.array_equal(a, b)


def test_row_band_equals_full_frame():
    ptgp = scene_of("room")
    with pt.Scene(ptgp) as s:
        full, _, _ = s.render(80, 45, 8, seed=2, per_ray=True)
        band, _, _ = s.render(80, 45, 8, seed=2, rows=(10, 20), per_ray=True)
    # full is top row first: j = 19 is row 45 - 1 - 19 = 25
    assert np.array_equal(band, full[45 - 20:45 - 10])


def test_progressive_passes_continue_the_sequence():
    """Two passes of 8 spp starting at samples 0 and 8 average to the 16 spp image."""
    ptgp = scene_of("materials")
    with pt.Scene(ptgp) as s:
        full, _, _ = s.render(64, 36, 16, seed=4, per_ray=True)
        p0, _, _ = s.render(64, 36, 8, seed=4, first_sample=0, per_ray=True)
        p1, _, _ = s.render(64, 36, 8, seed=4, first_sample=8, per_ray=True)
    assert np.allclose((p0 + p1) / 2, full, rtol=1e-5, atol=1e-6)


def test_aux_matches_oracle():
    ptgp = scene_of("room")
    with pt.Scene(ptgp) as s:
        rgb, aux, _ = s.render(64, 36, 4, seed=8, aux=True)
        rgb2, _, _ = s.render(64, 36, 4, seed=8)
    _, oaux, _ = ob.or_render(ptgp, 64, 36, 4, seed=8, aux=True)
    assert np.array_equal(rgb, rgb2)
    frac, note = rc.per_value_agreement(aux, oaux)
    print(f"aux: {frac:.5f} of values within 1e-4 ({note})", flush=True)
    assert frac >= 0.999


def test_no_nee_matches_oracle():
    ptgp = scene_of("room")
    with pt.Scene(ptgp) as s:
        rgb, _, st = s.render(64, 36, 8, seed=3, nee=False)
    orgb, _, orays = ob.or_render(ptgp, 64, 36, 8, seed=3, nee=False)
    frac, note = rc.per_value_agreement(rgb, orgb)
    print(f"no NEE: {frac:.5f} of values within 1e-4 ({note})", flush=True)
    assert frac >= 0.999
    assert st.shadow_rays == 0


@pytest.mark.parametrize("name", ["materials", "field", "teapot", "room", "crowd"])
def test_reference_render_fixture(name):
    """The reference's own statistical check (tests/render_test.py) on the
    GPU image: 960x540, 200 spp, seed 1, against tests/reference.json."""
    ref = rc.load_reference()
    ptgp = scene_of(name)
    with pt.Scene(ptgp) as s:
        rgb, _, st = s.render(rc.WIDTH, rc.HEIGHT, rc.SPP, seed=1)
    ok, msg = rc.passes(rgb, ref["scenes"][name])
    print(f"{name}: {msg}; {st.mrays_per_s:.0f} Mrays/s", flush=True)
    assert ok, msg


def test_bad_render_parameters_are_refused():
    ptgp = scene_of("materials")
    with pt.Scene(ptgp) as s:
        for kw in ({"width": 1, "height": 10, "spp": 1}, {"width": 10, "height": 10, "spp": 0},
                   {"width": 10, "height": 10, "spp": 4, "max_bounces": 0},
                   {"width": 10, "height": 10, "spp": 4, "rows": (5, 20)},
                   {"width": 10, "height": 10, "spp": 4, "samples_per_batch": 3},
                   {"width": 10, "height": 10, "spp": 4, "stack_entries": 12}):
            with pytest.raises(pt.PTError):
                s.render(**kw)


def test_progressive_driver_equals_one_render():
    """pt_progressive_*: the samples of a progressive render are those of one render of the same total (same
    seed, pixel and sample numbers), so the running average after the last step equals that render up to the
    rounding of summing per-step averages; steps start at one sample, are powers of two up to 64, and add up
    to the total; a step after the last is refused."""
    ptgp = rc.load_scene("field")[1]
    W, H, spp = 96, 64, 100
    with pt.Scene(ptgp) as s:
        want, _, _ = s.render(W, H, spp, seed=7, per_ray=True)
        steps, img = [], None
        with pt.Progressive(s, W, H, spp, seed=7, per_ray=True) as prog:
            while prog.done < spp:
                done, n, img = prog.step()
                steps.append(n)
            with pytest.raises(pt.PTError):
                prog.step()
    print(f"steps {steps}", flush=True)
    assert steps[0] == 1 and sum(steps) == spp
    assert all(n & (n - 1) == 0 and n <= 64 for n in steps[:-1])
    assert np.allclose(img, want, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("name", ["crowd", "config1"])
def test_scene_autotune_keeps_the_image(name):
    """pt_scene_autotune times the frame under candidate launch shapes and keeps the fastest; the shape
    changes no bit of the image, before or after tuning (camera rays walked per ray: the tuned batch size
    regroups camera packets), and tuning a training render is refused."""
    ptgp = scene_of(name)
    with pt.Scene(ptgp) as s:
        before, _, sb = s.render(160, 90, 16, seed=3, per_ray=True)
        r = s.autotune(160, 90, 16, seed=3)
        after, _, sa = s.render(160, 90, 16, seed=3, per_ray=True)
        print(f"{name}: tuned stack {r.stack_entries}, blocks/CU {r.blocks_per_cu}, paths {r.max_paths}; "
              f"tried {[(t[0], t[1], t[3], round(t[4])) for t in r.tried]}", flush=True)
    assert r.stack_entries in (8, 16, 32) and r.blocks_per_cu >= 1 and r.streams == 1
    assert len(r.tried) >= 6 and (r.stack_entries, r.blocks_per_cu, r.max_paths) in [(t[0], t[1], t[3]) for t in r.tried]
    assert np.array_equal(before, after) and sb.rays == sa.rays
