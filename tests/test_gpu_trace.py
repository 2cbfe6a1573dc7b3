"""GPU traversal (pt_trace) against the oracle on identical ray batches:
bit-identical t, u, v and the same primitive as the reference's traversal."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import rays as R
import render_check as rc

pytestmark = pytest.mark.gpu


def cases():
    out = [(n, *rc.load_scene(n)) for n in ["materials", "field", "room", "teapot", "crowd"]]
    for cfg, sc in [(0, 1.0), (1, 0.05), (2, 0.02), (4, 0.02)]:
        ptsd = pt.generate(cfg, 3, sc)
        out.append((f"config{cfg}", ptsd, pt.pack(ptsd)))
    return out


def compare(ptgp, batch, label):
    with pt.Scene(ptgp) as s:
        h = s.trace(batch)
        a = s.trace(batch, any_hit=True)
    t, u, v, kind, prim, inst, _ = ob.or_trace(ptgp, batch)
    hit = kind != 0
    print(f"{label}: {len(batch)} rays, {hit.mean():.3f} hit", flush=True)
    assert np.array_equal(h["kind"], kind)
    assert np.array_equal(h["prim"][hit], prim[hit])
    assert np.array_equal(h["inst"][hit], inst[hit])
    assert np.array_equal(h["t"][hit].view(np.uint32), t[hit].view(np.uint32))
    tri = kind == 3
    assert np.array_equal(h["u"][tri].view(np.uint32), u[tri].view(np.uint32))
    assert np.array_equal(h["v"][tri].view(np.uint32), v[tri].view(np.uint32))
    any_kind = ob.or_trace(ptgp, batch, any_hit=True)[3]
    assert np.array_equal(a["kind"], any_kind)
    return h


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_trace_matches_oracle(case):
    name, ptsd, ptgp = case
    compare(ptgp, R.mixed_batch(ptgp, 20000, 7), name)


def test_trace_full_size_mesh_against_oracle_and_fp64():
    """The 1,003,520-triangle mesh (11 levels, beyond the reference's GPU
    stack): identical to the oracle; >= 99.99% the FP64 primitive."""
    ptsd = pt.generate(1, 1, 1.0)
    ptgp = pt.pack(ptsd)
    batch = R.mixed_batch(ptgp, 100000, 21)
    h = compare(ptgp, batch, "config1 full")
    t64, k64, p64, i64 = ob.or_trace_f64(ptgp, batch)
    same = (h["kind"] == k64) & (h["prim"] == p64)
    assert same.mean() >= 0.9999, same.mean()


def test_shadow_rays_against_oracle():
    ptsd = pt.generate(1, 2, 0.1)
    ptgp = pt.pack(ptsd)
    prim = R.camera_rays(ptgp, 30000, 3)
    t = ob.or_trace(ptgp, prim)[0]
    sh = R.shadow_rays(ptgp, (prim[:, 0:3] + prim[:, 4:7] * t[:, None]).astype(np.float32), 4)
    with pt.Scene(ptgp) as s:
        a = s.trace(sh, any_hit=True)
    assert np.array_equal(a["kind"], ob.or_trace(ptgp, sh, any_hit=True)[3])


def test_small_stack_spills_to_hbm():
    """With an 8-entry LDS ring, deep trees spill; results must not change."""
    ptsd = pt.generate(1, 1, 0.3)
    ptgp = pt.pack(ptsd)
    batch = R.mixed_batch(ptgp, 20000, 5)
    with pt.Scene(ptgp) as s:
        rgb16, _, st16 = s.render(96, 54, 4, seed=1, stack_entries=16)
        rgb8, _, st8 = s.render(96, 54, 4, seed=1, stack_entries=8)
        rgb32, _, _ = s.render(96, 54, 4, seed=1, stack_entries=32)
    assert np.array_equal(rgb8, rgb16) and np.array_equal(rgb32, rgb16)
    assert st8.rays == st16.rays
    compare(ptgp, batch, "config1 0.3")


def test_empty_and_odd_batches():
    ptgp = pt.pack(pt.generate(0, 0, 0.25))
    with pt.Scene(ptgp) as s:
        assert len(s.trace(np.zeros((0, 8), np.float32))) == 0
        for n in (1, 63, 65, 1000):
            compare(ptgp, R.camera_rays(ptgp, n, n), f"n={n}")
