"""Host side of the library, no GPU: the C ABI loads and exports what the
header declares; scenes pack byte for byte as the reference packs them; bad
buffers are refused with an error, not a crash."""
import ctypes
import os
import re

import numpy as np
import pytest

import photon_tracer_amd as pt
import render_check as rc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "photon_tracer.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pt.library()
    names = header_symbols()
    assert "pt_render" in names and "pt_trace" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version():
    assert "gfx950" in pt.version()


@pytest.mark.parametrize("name", ["materials", "field", "teapot", "room", "crowd"])
def test_pack_matches_reference_pack_for_gpu(name):
    # The fixture holds the reference's own serialize(pack_for_gpu(...)) of
    # its scene; our builder + packer must reproduce it byte for byte.
    ptsd, ptgp = rc.load_scene(name)
    ours = pt.pack(ptsd)
    assert ours.size == ptgp.size
    assert np.array_equal(ours, ptgp)


@pytest.mark.parametrize("config,scale", [(0, 1.0), (1, 0.01), (2, 0.01), (4, 0.005)])
def test_generated_scenes_are_deterministic(config, scale):
    a = pt.generate(config, 5, scale)
    b = pt.generate(config, 5, scale)
    c = pt.generate(config, 6, scale)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert np.array_equal(pt.pack(a), pt.pack(b))


def test_config_sizes():
    w = pt.generate(0, 0, 1.0)
    assert w[11] == 22 * 22 + 4 and w[12] == 1          # 484 small + 3 big + 1 light; the ground
    w = pt.generate(1, 0, 1.0)
    assert w[13] == 20 * 224 * 224                      # 1,003,520 triangles
    w = pt.generate(2, 0, 1.0)
    assert w[14] == 3 and w[15] == 4096 and w[16] == 50_000 + 200_000 + 500_000


def test_deep_tree_is_packed():
    # 1M triangles: 11 levels, deeper than the reference's WebGPU packer allows.
    ptgp = pt.pack(pt.generate(1, 1, 1.0))
    assert pt.ptgp_depth(ptgp) > 10


def test_job_generator():
    jobs = [pt.generate_job(s, 0.01) for s in range(8)]
    assert {j.kind for j in jobs} == {0, 1}
    for j in jobs:
        assert (j.width, j.height) in {(256, 256), (512, 512), (1024, 576), (1024, 1024)}
        assert j.spp in (4, 16, 64)
        assert pt.pack(j.ptsd).size > 48


def test_bad_buffers_are_refused():
    ptgp = pt.pack(pt.generate(0, 0, 0.25))
    assert pt.ptgp_depth(ptgp) >= 1
    bad = ptgp.copy()
    bad[0] = 0
    with pytest.raises(pt.PTError, match="magic"):
        pt.ptgp_depth(bad)
    bad = ptgp.copy()
    bad[3] += 64   # nodes section runs past the end
    with pytest.raises(pt.PTError):
        pt.ptgp_depth(bad)
    bad = ptgp.copy()
    node0 = int(bad[2]) // 4 + int(bad[18]) * 32
    bad[node0 + 24: node0 + 28] = 0xFFFFFF   # child links out of range
    bad[node0 + 28: node0 + 32] = 0
    with pytest.raises(pt.PTError):
        pt.ptgp_depth(bad)
    with pytest.raises(pt.PTError):
        pt.ptgp_depth(ptgp[:20])
    with pytest.raises(pt.PTError, match="PTSD"):
        pt.pack(np.zeros(40, np.uint32))
    with pytest.raises(pt.PTError):
        pt.generate(9, 0, 1.0)


def test_bvh_build_host_path():
    # pt_bvh_build without PT_BVH_DEVICE runs the host builder that the pack
    # test above pins to the reference; here the C ABI around it: every
    # primitive in exactly one leaf, children inside their parents' slots,
    # empty slots +inf, and argument errors.
    rng = np.random.default_rng(2)
    c = rng.uniform(-5, 5, (3000, 3))
    boxes = np.concatenate([c - 0.05, c + 0.05], 1).astype(np.float32)
    nodes, order, _ = pt.bvh_build(boxes)
    lo = nodes[:, 0:12].view(np.float32).reshape(-1, 3, 4)
    hi = nodes[:, 12:24].view(np.float32).reshape(-1, 3, 4)
    first, count = nodes[:, 24:28], nodes[:, 28:32]
    seen = []
    for k in range(len(nodes)):
        for s in range(4):
            if lo[k, 0, s] == np.inf:
                assert hi[k, 0, s] == np.inf and first[k, s] == 0 and count[k, s] == 0
                continue
            if count[k, s]:
                prims = order[first[k, s]:first[k, s] + count[k, s]]
                seen.append(prims)
                assert (boxes[prims, :3] >= lo[k, :, s]).all() and (boxes[prims, 3:] <= hi[k, :, s]).all()
            else:
                assert first[k, s] > k   # preorder: children after their parent
    assert np.array_equal(np.sort(np.concatenate(seen)), np.arange(3000))
    again = pt.bvh_build(boxes)
    assert np.array_equal(again[0], nodes) and np.array_equal(again[1], order)
    empty = pt.bvh_build(np.zeros((0, 6), np.float32))
    assert empty[0].shape == (0, 32) and empty[1].size == 0
    lib = pt.library()
    n_nodes = ctypes.c_size_t(0)
    p1, p2 = ctypes.POINTER(ctypes.c_uint32)(), ctypes.POINTER(ctypes.c_uint32)()
    assert lib.pt_bvh_build(boxes.ctypes.data, 3000, 4, ctypes.byref(p1), ctypes.byref(n_nodes), ctypes.byref(p2),
                            None) != 0
    assert lib.pt_bvh_build(None, 3, 0, ctypes.byref(p1), ctypes.byref(n_nodes), ctypes.byref(p2), None) != 0
    w = pt.generate(0, 0, 0.25)
    assert lib.pt_pack_ex(w.ctypes.data, w.size, 2, ctypes.byref(p1), ctypes.byref(n_nodes)) != 0
    bad = boxes.copy()
    bad[7, 4] = np.nan   # refused before any device call
    with pytest.raises(pt.PTError, match="NaN"):
        pt.bvh_build(bad, device=True)


def test_cycle_is_refused():
    ptgp = pt.pack(pt.generate(0, 0, 0.25))
    bad = ptgp.copy()
    root = int(bad[18])
    node = int(bad[2]) // 4 + root * 32
    # Point the root's first interior child back at the root.
    for c in range(4):
        if bad[node + 28 + c] == 0 and np.frombuffer(bad[node + c:node + c + 1].tobytes(), np.float32)[0] != np.inf:
            bad[node + 24 + c] = root
            break
    with pytest.raises(pt.PTError, match="cycle"):
        pt.ptgp_depth(bad)


def test_pfm_round_trip(tmp_path):
    img = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    p = str(tmp_path / "x.pfm")
    pt.write_pfm(p, img)
    assert np.array_equal(pt.read_pfm(p), img)
    raw = open(p, "rb").read()
    assert raw.startswith(b"PF\n7 5\n-1.0\n")
    # bottom row first (image_write.hpp:24-28)
    first = np.frombuffer(raw[len(b"PF\n7 5\n-1.0\n"):][: 7 * 12], "<f4").reshape(7, 3)
    assert np.array_equal(first, img[-1])


GOLDEN = os.path.join(ROOT, "tests", "golden")


def ptsd_triangles(w):
    """The flat triangle records of a PTSD buffer: 28 words each, after the
    32-word header, 12-word materials and 5-word spheres and large spheres."""
    start = 32 + 12 * int(w[10]) + 5 * (int(w[11]) + int(w[12]))
    return w[start:start + 28 * int(w[13])].reshape(-1, 28)


def test_obj_loader_matches_reference_load_obj():
    # The teapot scene's fixture holds the reference's own load_obj output
    # (renderer.hpp:230: the Utah teapot, base (0, -0.5, -1.2), height 0.75,
    # copper, no file reader, no turn). Our loader's triangles of the same
    # OBJ text, appended to that scene, must repeat them word for word.
    import gzip
    ptsd, _ = rc.load_scene("teapot")
    text = gzip.open(os.path.join(GOLDEN, "teapot.obj.gz")).read()
    ref = ptsd_triangles(ptsd)
    copper = int(ref[0, 24])
    assert (ref[:, 24] == copper).all() and len(ref) > 1000
    out = pt.add_obj(ptsd, text, (0.0, -0.5, -1.2), 0.75, copper)
    tris = ptsd_triangles(out)
    assert len(tris) == 2 * len(ref)
    assert np.array_equal(tris[:len(ref)], ref)
    assert np.array_equal(tris[len(ref):], ref)
    assert pt.pack(out).size > pt.pack(ptsd).size


def test_obj_loader_faces_normals_and_errors():
    # A quad (a fan of two triangles) with file normals on one face only, a
    # negative index, texture coordinates, and a degenerate face that is
    # dropped; turned 90 degrees and scaled to height 2.
    ptsd = pt.generate(0, 0, 0.05)
    n0 = int(ptsd[13])
    text = b"""v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 2 2 2
vt 0.25 0.5
vn 0 0 1
f 1/1/1 2/1/1 3/1/1 4/1/1
f -1 -1 -1
f 1 2 5
"""
    out = pt.add_obj(ptsd, text, (0.0, 0.0, 0.0), 2.0, 0, turn_deg=90)
    tris = ptsd_triangles(out)[n0:]
    assert len(tris) == 3                     # the degenerate face is dropped
    f = tris.view(np.float32)
    # turned 90 degrees about y ((x, z) -> (z, -x)), scaled to height 2 and set down on the origin: y spans [0, 2]
    assert np.isclose(f[:, [1, 4, 7]].min(), 0) and np.isclose(f[:, [1, 4, 7]].max(), 2)
    # the quad's corners carry the file normal (0, 0, 1) turned to (1, 0, 0) and the texture coordinate
    assert np.allclose(f[0, 9:18].reshape(3, 3), [[1, 0, 0]] * 3, atol=1e-6)
    assert np.allclose(f[0, 18:24], [0.25, 0.5] * 3)
    # the third face has no file normals: unit area-weighted vertex normals
    assert np.allclose(np.linalg.norm(f[2, 9:18].reshape(3, 3), axis=1), 1, atol=1e-6)
    with pytest.raises(pt.PTError, match="out of range"):
        pt.add_obj(ptsd, b"v 0 0 0\nf 1 2 3\n", (0, 0, 0), 1.0, 0)
    with pytest.raises(pt.PTError):
        pt.add_obj(ptsd, text, (0, 0, 0), 1.0, 10_000)        # no such material
    assert np.array_equal(pt.add_obj(ptsd, b"# nothing\n", (0, 0, 0), 1.0, 0), ptsd)


def test_obj_loader_with_mtl_matches_reference_load_obj():
    # The reference's mesh scene loading Spot through its file reader: the OBJ
    # names spot.mtl, a matte textured by spot_texture.png. Our loader, given
    # the same MTL text and the image the reference decoded, must add the same
    # triangles, a material with the same words and the same texture.
    import gzip
    ref_ptsd = rc.load_spot("spot")
    text = gzip.open(os.path.join(GOLDEN, "spot.obj.gz")).read()
    mtl = open(os.path.join(GOLDEN, "spot.mtl"), "rb").read()
    ref = ptsd_triangles(ref_ptsd)
    n_mat = int(ref_ptsd[10])
    (image,) = rc.ptsd_textures(ref_ptsd)
    out = pt.add_obj(ref_ptsd, text, (0.0, -0.5, -1.2), 0.75, 0, mtl_text=mtl, images={"spot_texture.png": image})
    assert int(out[10]) == n_mat + 1                       # one material made, appended
    tris = ptsd_triangles(out)[len(ref):]
    assert len(tris) == len(ref) and (tris[:, 24] == n_mat).all()
    assert np.array_equal(np.delete(tris, 24, axis=1), np.delete(ref, 24, axis=1))
    mats = out[32:32 + 12 * (n_mat + 1)].reshape(-1, 12)
    theirs = mats[int(ref[0, 24])]
    assert mats[n_mat, 3] == 0 and np.array_equal(mats[n_mat, :3].view(np.float32), [1, 1, 1])   # matte, Kd 1 1 1
    assert np.array_equal(np.delete(mats[n_mat], 9), np.delete(theirs, 9))
    # word 9: texture number + 1; ours is the scene's second texture, holding the same image
    assert theirs[9] == 1 and mats[n_mat, 9] == 2 and int(out[17]) == 2
    textures = rc.ptsd_textures(out)
    assert np.array_equal(textures[1], image) and np.array_equal(textures[0], image)
    # Without the image the material is an untextured matte, as the reference makes it when its
    # reader finds no file (obj.hpp:100-104).
    plain_tex = pt.add_obj(ref_ptsd, text, (0.0, -0.5, -1.2), 0.75, 0, mtl_text=mtl)
    assert plain_tex[32 + 12 * n_mat + 9] == 0 and int(plain_tex[17]) == 1
    # Without the MTL text every face takes the given material, as the browser demo loads.
    plain = ptsd_triangles(pt.add_obj(ref_ptsd, text, (0.0, -0.5, -1.2), 0.75, 2))[len(ref):]
    assert (plain[:, 24] == 2).all()


def test_obj_loader_mtl_kinds_match_reference_make_material():
    # Every branch of make_material (obj.hpp:89-112), from an OBJ and MTL of
    # our own that the reference's mesh scene loaded through its file reader
    # (tests/golden/make_golden.py): the same triangles and, material by
    # material, the same words.
    z = np.load(os.path.join(GOLDEN, "mtl_kinds.npz"))
    ref_ptsd = z["ptsd"]
    ref = ptsd_triangles(ref_ptsd)
    n_mat = int(ref_ptsd[10])
    fallback = int(ref[-1, 24])   # the unknown name's faces: the scene's copper, the reference's fallback
    out = pt.add_obj(ref_ptsd, z["obj"].tobytes(), (0.0, -0.5, -1.2), 0.75, fallback, mtl_text=z["mtl"].tobytes())
    tris = ptsd_triangles(out)[len(ref):]
    assert len(tris) == len(ref) == 18
    assert np.array_equal(np.delete(tris, 24, axis=1), np.delete(ref, 24, axis=1))
    mats = out[32:32 + 12 * int(out[10])].reshape(-1, 12)
    ref_mats = ref_ptsd[32:32 + 12 * n_mat].reshape(-1, 12)
    for ours, theirs in zip(tris[:, 24], ref[:, 24]):
        assert np.array_equal(mats[ours], ref_mats[theirs])
    assert len(set(tris[:16, 24])) == 8 and set(mats[tris[:16:2, 24], 3]) == {0, 1, 2, 3}
