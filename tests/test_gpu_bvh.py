"""The device BVH builder (pt_bvh_build with PT_BVH_DEVICE, pt_pack_ex with
PT_PACK_DEVICE_BVH) against the host builder and the reference's packed
scenes. The host builder reproduces the reference's bvh::build node for node
(the reference's own serialize(pack_for_gpu(...)) fixtures pack byte for byte,
test_host.py), so the bar here is bit-exact: the same 4-wide nodes, the same
primitive order, the same PTGP words.

The box sets aim at the builder's branches: one wave per small segment (n up
to 1024) and the level passes above it; every centroid in one place (no axis
to bin, halving); a geometric run of centroids (a degenerate tree that stops
at the reference's depth limit of 64, with leaves over 8 primitives);
centroids in a plane (an axis of zero extent); a coarse grid (many equal
bin costs, so ties decide); boxes of very different sizes."""
import numpy as np
import pytest

import photon_tracer_amd as pt
import render_check as rc

pytestmark = pytest.mark.gpu


def make_boxes(kind, n, seed=1):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        c, h = rng.uniform(0, 1, (n, 3)), rng.uniform(0, 0.01, (n, 3))
    elif kind == "clusters":
        k = np.arange(n) % 7
        c = np.stack([k * 3 + 0.1 * rng.normal(size=n), 0.1 * rng.normal(size=n) - k, 0.05 * rng.normal(size=n)], 1)
        h = np.broadcast_to([0.001, 0.002, 0.001], (n, 3))
    elif kind == "same":
        c, h = np.broadcast_to([1.0, 2.0, 3.0], (n, 3)), np.full((n, 3), 0.5)
    elif kind == "geometric":
        x = 1.3 ** (np.arange(n) % 300)
        c = np.stack([x, np.zeros(n), np.zeros(n)], 1)
        h = np.stack([1e-3 * x, np.ones(n), np.ones(n)], 1)
    elif kind == "planar":
        c = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.zeros(n)], 1)
        h = np.full((n, 3), 0.01)
    elif kind == "grid":
        c = np.stack([np.floor(rng.uniform(0, 4, n)), np.floor(rng.uniform(0, 3, n)), np.full(n, 0.5)], 1)
        h = np.full((n, 3), 0.1)
    elif kind == "signed zeros":
        # flat boxes on the plane z = 0, their z bounds -0.0 or +0.0 by turns (OBJ files write "-0.000000"):
        # the reference's std::min/std::max keep whichever comes first
        c = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.zeros(n)], 1)
        h = np.stack([np.full(n, 0.01), np.full(n, 0.01), np.zeros(n)], 1)
        b = np.concatenate([c - h, c + h], 1).astype(np.float32)
        b[::2, 2], b[1::2, 2], b[::3, 5] = -0.0, 0.0, -0.0
        return b
    elif kind == "mixed":
        c = 100 * rng.normal(size=(n, 3))
        r = np.where(rng.uniform(size=n) < 0.01, 50 * rng.uniform(size=n), 0.2 * rng.uniform(size=n))
        h = np.repeat(r[:, None], 3, 1)
    else:
        raise ValueError(kind)
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def leaf_primitives(nodes, order):
    """Every primitive index the leaves reference, in leaf order."""
    counts = nodes[:, 28:32]
    firsts = nodes[:, 24:28]
    out = [order[f:f + k] for f, k in zip(firsts[counts > 0], counts[counts > 0])]
    return np.concatenate(out) if out else np.zeros(0, np.uint32)


CASES = [("uniform", n) for n in (1, 2, 3, 8, 9, 64, 500, 1024, 1025, 5000, 100_000)] + [
    ("clusters", 12_000), ("same", 8), ("same", 9), ("same", 3000), ("geometric", 300), ("geometric", 3000),
    ("planar", 9000), ("grid", 6000), ("mixed", 20_000), ("signed zeros", 5000), ("uniform", 1_000_000)]


@pytest.mark.parametrize("kind,n", CASES)
def test_device_build_equals_host_build(kind, n):
    boxes = make_boxes(kind, n)
    hn, ho, hms = pt.bvh_build(boxes, device=False)
    dn, do, dms = pt.bvh_build(boxes, device=True)
    print(f"{kind} n={n}: {len(hn)} nodes; host {hms:.1f} ms, device {dms:.1f} ms", flush=True)
    assert dn.shape == hn.shape
    assert np.array_equal(do, ho)
    assert np.array_equal(dn, hn)
    if kind == "signed zeros":   # both zeros reach the node words
        assert {0x80000000, 0} <= set(dn[:, 8:12].reshape(-1).tolist()) | set(dn[:, 20:24].reshape(-1).tolist())
    # and a sanity check that does not lean on the host builder: each primitive in exactly one leaf
    assert np.array_equal(np.sort(leaf_primitives(dn, do)), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("name", ["materials", "field", "teapot", "room", "crowd"])
def test_device_pack_matches_reference_pack_for_gpu(name):
    ptsd, ptgp = rc.load_scene(name)
    ours = pt.pack(ptsd, device_bvh=True)
    assert np.array_equal(ours, ptgp)


@pytest.mark.parametrize("config,scale", [(0, 1.0), (1, 1.0), (2, 0.1), (4, 0.1)])
def test_device_pack_equals_host_pack_on_generated_scenes(config, scale):
    ptsd = pt.generate(config, 3, scale)
    assert np.array_equal(pt.pack(ptsd, device_bvh=True), pt.pack(ptsd))


def test_device_build_repeats_and_edge_arguments():
    boxes = make_boxes("uniform", 4000, seed=9)
    a = pt.bvh_build(boxes, device=True)
    b = pt.bvh_build(boxes, device=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    nodes, order, _ = pt.bvh_build(np.zeros((0, 6), np.float32), device=True)
    assert nodes.shape == (0, 32) and order.size == 0
