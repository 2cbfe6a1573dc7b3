"""Guided rendering: one-sample MIS of the cosine lobe and the learned guide
at diffuse bounces. The estimator must stay unbiased for any guide (the
guided image converges to the plain one), be deterministic for fixed
weights, train online from the path loop's own samples, and, trained, beat
the plain tracer on the hard-lighting room (config 4) at equal samples."""
import numpy as np
import pytest

import photon_tracer_amd as pt
import render_check as rc

pytestmark = pytest.mark.gpu


def rel_mse(img, ref):
    """mean over pixels and channels of (I - R)^2 / (R^2 + 0.01)."""
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


def passes(s, W, H, spp, n, seed0, **kw):
    """n independent renders (seeds seed0 ..): their mean and the standard error of the mean."""
    imgs = np.stack([s.render(W, H, spp, seed=seed0 + k, **kw)[0] for k in range(n)])
    return imgs.mean(0), imgs.std(0, ddof=1) / np.sqrt(n)


def block(a, k):
    h, w = a.shape[0] // k * k, a.shape[1] // k * k
    return a[:h, :w].reshape(h // k, k, w // k, k, -1).mean((1, 3))


@pytest.mark.parametrize("scene,trained", [("room", False), ("room", True), ("hard_room", True)])
def test_guided_render_is_unbiased(scene, trained):
    """Guided and plain images agree block by block within their noise: 8x8-pixel blocks, z-scores from
    the spread of 12 independent renders each; at most 2% of block channels beyond |z| = 4, and the
    image means within 4 standard errors. Scenes: the reference's room, and the config-4 room (lamps behind
    a partition; its mesh at 2% of its triangles)."""
    ptgp = rc.load_scene("room")[1] if scene == "room" else pt.pack(pt.generate(4, 1, 0.02))
    W, H = 96, 72
    with pt.Scene(ptgp) as s:
        g = pt.Guide(s, seed=21, learning_rate=3e-3)
        s.set_guide(g)
        if trained:
            for k in range(8):
                s.render(W, H, 16, seed=500 + k, train=True)
            steps, used, loss = g.stats()
            assert steps > 0 and used > 0 and np.isfinite(loss)
        mp, ep = passes(s, W, H, 32, 12, 100)
        mg, eg = passes(s, W, H, 32, 12, 200, guide=True)
    bp, bg = block(mp, 8), block(mg, 8)
    # standard error of a block mean from the per-pixel standard errors (pixels' errors are independent)
    sp = np.sqrt(block(ep ** 2, 8) / 64)
    sg = np.sqrt(block(eg ** 2, 8) / 64)
    z = (bg - bp) / np.sqrt(sp ** 2 + sg ** 2 + 1e-12)
    frac = float(np.mean(np.abs(z) > 4))
    zi = (mg.mean() - mp.mean()) / np.sqrt((ep ** 2).sum() / ep.size ** 2 + (eg ** 2).sum() / eg.size ** 2)
    print(f"{scene}, trained={trained}: {frac:.4f} of block channels beyond |z|=4; image mean z = {zi:.2f}; "
          f"means {mp.mean():.5f} plain, {mg.mean():.5f} guided", flush=True)
    assert frac <= 0.02
    assert abs(zi) < 4


def test_guided_render_is_deterministic():
    ptgp = rc.load_scene("room")[1]
    with pt.Scene(ptgp) as s:
        g = pt.Guide(s, seed=5)
        s.set_guide(g)
        a, _, sa = s.render(80, 60, 8, seed=3, guide=True)
        b, _, sb = s.render(80, 60, 8, seed=3, guide=True)
        c, _, _ = s.render(80, 60, 8, seed=3, guide=True, graph=False)
        e, _, _ = s.render(80, 60, 8, seed=3)
        r, _, _ = s.render(80, 60, 8, seed=3, guide=True, per_ray=True)
        d, _, _ = s.render(80, 60, 8, seed=3, guide=True, samples_per_batch=2, per_ray=True)
    assert np.array_equal(a, b) and sa.rays == sb.rays
    assert np.array_equal(a, c)     # graph replay == eager
    assert not np.array_equal(a, e)
    assert np.array_equal(r, d)     # camera rays walked per ray: batch size does not matter


def test_guide_stats():
    """guide_queries (PT_RENDER_COUNT) counts the diffuse hits the network was evaluated for: none in a plain
    render; at one bounce in the closed config-4 room, nearly every camera ray (the room's surfaces are matte,
    the lamps tiny); more with more bounces; the same on a second render. guide_ms (PT_RENDER_TIMING) is the
    inference launches' share of the render."""
    ptgp = pt.pack(pt.generate(4, 1, 0.02))
    W, H, spp = 64, 48, 4
    with pt.Scene(ptgp) as s:
        g = pt.Guide(s, seed=3)
        s.set_guide(g)
        _, _, plain = s.render(W, H, spp, seed=1, count=True)
        _, _, one = s.render(W, H, spp, seed=1, guide=True, count=True, max_bounces=1)
        _, _, full = s.render(W, H, spp, seed=1, guide=True, count=True)
        _, _, again = s.render(W, H, spp, seed=1, guide=True, count=True)
        _, _, timed = s.render(W, H, spp, seed=1, guide=True, timing=True)
        g.close()
    print(f"guide queries: plain {plain.guide_queries}, one bounce {one.guide_queries} of {W * H * spp} camera rays, "
          f"ten bounces {full.guide_queries}; inference {timed.guide_ms:.3f} of {timed.render_ms:.3f} ms", flush=True)
    assert plain.guide_queries == 0 and plain.guide_ms == 0
    assert 0.9 * W * H * spp <= one.guide_queries <= W * H * spp
    assert full.guide_queries > one.guide_queries and full.guide_queries == again.guide_queries
    assert 0 < timed.guide_ms < timed.render_ms


def test_guided_render_needs_a_guide():
    with pt.Scene(rc.load_scene("room")[1]) as s:
        with pytest.raises(pt.PTError):
            s.render(32, 32, 4, guide=True)


def test_trained_guide_beats_plain_on_the_hard_room():
    """Config 4 (the lamps hide behind a partition; light reaches the room through a 0.5 x 0.5 opening):
    after online training (one call: 1-spp batches, 16 Adam steps after each), guided renders have
    clearly lower relMSE than plain ones at equal spp, against a 4096-spp plain reference (measured
    about 0.2x; asserted below 0.5x over four seeds)."""
    ptgp = pt.pack(pt.generate(4, 1, 0.02))
    W, H, spp = 192, 144, 64
    with pt.Scene(ptgp) as s:
        ref, _, _ = s.render(W, H, 4096, seed=9999)
        g = pt.Guide(s, seed=2, learning_rate=3e-3, steps_per_batch=16)
        s.set_guide(g)
        s.render(W, H, 64, seed=1000, train=True, samples_per_batch=1)
        steps, used, loss = g.stats()
        plain = [rel_mse(s.render(W, H, spp, seed=10 + k)[0], ref) for k in range(4)]
        guided = [rel_mse(s.render(W, H, spp, seed=20 + k, guide=True)[0], ref) for k in range(4)]
    print(f"trained {steps} steps on {used} samples (loss {loss:.3f}); relMSE at {spp} spp: "
          f"plain {np.mean(plain):.4f}, guided {np.mean(guided):.4f} ({np.mean(guided) / np.mean(plain):.3f}x)", flush=True)
    assert steps == 64 * 16
    assert np.mean(guided) < 0.5 * np.mean(plain)
