"""The merged scene of a batch, on the CPU: each scene, seen through the
merged buffer from its own root (with its large spheres, lights, flags and
camera), must trace and render exactly as the scene alone, in the oracle.
That pins the index rebasing (nodes, leaf ranges, spheres, triangles,
materials, lights, large spheres) without a GPU: a wrong offset changes a hit
or a pixel. The reference has no batching; the single-scene oracle is pinned
against the reference's own code and fixtures elsewhere."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc


def scenes():
    """Mixed inputs: the reference's sphere field, a displaced-mesh job, a sphere job, and a lit room."""
    out = [rc.load_scene("field")[1]]
    for seed in (3, 11):
        out.append(pt.pack(pt.generate_job(seed, 0.01).ptsd))
    out.append(pt.pack(pt.generate(4, 2, 0.004)))
    return out


def job_view(merged: np.ndarray, part: pt.BatchPart) -> np.ndarray:
    """The merged buffer seen as one scene: root, flags and camera of `part`, and its runs of the merged
    large-sphere and light lists appended as the view's own sections."""
    w = merged.copy()
    def sec(k):
        return int(w[2 + 2 * k]) // 4, int(w[3 + 2 * k]) // 4
    l_at, _ = sec(6)
    g_at, _ = sec(7)
    lights = merged[l_at + part.lights_off:l_at + part.lights_off + part.n_lights]
    large = merged[g_at + part.large_off:g_at + part.large_off + part.n_large]
    tail = []
    for k, arr in ((6, lights), (7, large)):
        at = w.size + sum(len(t) for t in tail)
        w[2 + 2 * k] = 4 * at
        w[3 + 2 * k] = 4 * len(arr)
        tail.append(np.concatenate([arr, np.zeros((-len(arr)) % 4, np.uint32)]).astype(np.uint32))
    w = np.concatenate([w] + tail).astype(np.uint32)
    w[18] = part.root
    w[21] = part.flags
    w[22:38] = np.frombuffer(part.camera.astype(np.float32).tobytes(), np.uint32)
    return w


def counts(w):
    return int(w[20]), int(w[7]) // 32, int(w[9]) // 112   # small spheres, spheres, triangles


@pytest.fixture(scope="module")
def merged():
    sc = scenes()
    m, parts = pt.batch_merge(sc)
    return sc, m, parts


def test_merged_buffer_is_a_valid_scene(merged):
    sc, m, parts = merged
    assert pt.ptgp_depth(m) <= max(pt.ptgp_depth(s) for s in sc) + 1   # four scenes: one level above them
    for s, p in zip(sc, parts):
        assert p.flags == s[21]
        assert np.array_equal(p.camera.ravel(), np.frombuffer(s[22:38].tobytes(), np.float32))
        assert pt.ptgp_depth(job_view(m, p)) == pt.ptgp_depth(s)


def test_each_scene_traces_as_alone(merged):
    sc, m, parts = merged
    ns_all = sum(counts(s)[0] for s in sc)
    small_off = big_off = tri_off = 0
    rng = np.random.default_rng(7)
    for s, p in zip(sc, parts):
        ns, nsph, ntri = counts(s)
        cam = np.frombuffer(s[22:38].tobytes(), np.float32).reshape(4, 4)
        n = 4000
        o = np.repeat(cam[0:1, :3], n, 0)
        uv = rng.uniform(0, 1, (n, 2)).astype(np.float32)
        d = cam[1, :3] + uv[:, :1] * cam[2, :3] + uv[:, 1:] * cam[3, :3] - cam[0, :3]
        rays = pt.rays_array(o, d)
        view = job_view(m, p)
        for any_hit in (False, True):
            a = ob.or_trace(s, rays, any_hit)
            b = ob.or_trace(view, rays, any_hit)
            assert np.array_equal(a[3], b[3])                        # kind
            if not any_hit:
                assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))   # t, bit for bit
                assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
                kind, prim = a[3], a[4]
                want = prim.astype(np.int64).copy()
                sph = (kind == 1) | (kind == 2)
                want[sph] = np.where(prim[sph] < ns, small_off + prim[sph], ns_all + big_off + prim[sph] - ns)
                tri = kind == 3
                want[tri] = tri_off + prim[tri]
                assert np.array_equal(want[kind != 0], b[4][kind != 0].astype(np.int64))
        assert (a[3] != 0).any()
        small_off += ns
        big_off += nsph - ns
        tri_off += ntri


def test_each_scene_renders_as_alone(merged):
    sc, m, parts = merged
    for k, (s, p) in enumerate(zip(sc, parts)):
        a, _, ra = ob.or_render(s, 24, 16, 2, seed=k)
        b, _, rb = ob.or_render(job_view(m, p), 24, 16, 2, seed=k)
        assert ra == rb
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_instanced_scenes_are_refused():
    inst = pt.pack(pt.generate(2, 0, 0.002))
    flat = pt.pack(pt.generate(0, 0, 0.1))
    with pytest.raises(pt.PTError):
        pt.batch_merge([flat, inst])


def test_single_scene_merge_keeps_its_root():
    s = pt.pack(pt.generate(0, 4, 0.1))
    m, parts = pt.batch_merge([s])
    assert parts[0].root == s[18] and m[18] == s[18]
    assert pt.ptgp_depth(m) == pt.ptgp_depth(s)
