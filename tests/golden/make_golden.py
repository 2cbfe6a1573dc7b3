"""Writes the golden fixtures from the reference, in this container (where
/root/reference exists), through the reference harness oracle/_ref:

  ref_scenes.npz      for each of the reference's own render-test scenes
                      (materials, field, teapot, room, crowd 8; the scenes of
                      tests/render_test.py:43-56), the PTSD scene description
                      exported from its build_world() and the PTGP buffer
                      serialize(pack_for_gpu(...)) made by the reference itself.
  reference.json.gz   the reference's statistical render fixture
                      (tests/reference.json: per 20x20 block mean and sigma at
                      960x540, 200 spp, 16 seeds), copied unchanged.
  teapot.obj.gz       the Utah teapot the reference's mesh scene loads
                      (models/teapot.obj, renderer.hpp:221-231), copied
                      unchanged: the OBJ loader's input.
  spot_scene.npz,     the reference's mesh scene loading Keenan Crane's Spot
  spot.obj.gz,        through its file reader (models/spot: the OBJ names
  spot.mtl,           spot.mtl, a matte textured by spot_texture.png),
  spot_texture.png    exported as PTSD with the texture as the reference's
                      decoder read it (spot_ptsd), the reference's "spot
                      crowd" render-test scene (tests/render_test.py:55:
                      crowd of 6 x 6, turned 150 degrees) exported the same
                      way (spot_crowd_ptsd), and the three input files copied
                      unchanged: the MTL, texture and image decoder fixture.
  mtl_kinds.npz       a small OBJ and MTL of our own (one quad per material:
                      a light, glass by dissolve, by Tr and by illum 7, metals
                      from illum 3 and 5 with and without Ks, a gray Kd from
                      one value, an unknown name) loaded by the reference's
                      mesh scene through its file reader, as PTSD, with the
                      two texts: every branch of make_material.

Run: python tests/golden/make_golden.py
"""
import gzip
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

REF = "/root/reference"
SCENES = {
    "materials": ("materials", {}),
    "field": ("field", {}),
    "teapot": ("mesh", {"obj_path": f"{REF}/models/teapot.obj"}),
    "room": ("room", {}),
    "crowd": ("crowd", {"crowd_side": 8, "obj_path": f"{REF}/models/teapot.obj"}),
}


def main():
    arrays = {}
    for name, (ref_name, kw) in SCENES.items():
        ptsd, ptgp = ob.ref_scene(ref_name, **kw)
        arrays[f"{name}_ptsd"] = ptsd
        arrays[f"{name}_ptgp"] = ptgp
        print(name, ptsd.size, ptgp.size)
    np.savez_compressed(os.path.join(HERE, "ref_scenes.npz"), **arrays)
    with open(f"{REF}/tests/reference.json", "rb") as src, gzip.open(os.path.join(HERE, "reference.json.gz"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    with open(f"{REF}/models/teapot.obj", "rb") as src, gzip.open(os.path.join(HERE, "teapot.obj.gz"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    write_spot()
    write_mtl_kinds()


def write_spot():
    obj = f"{REF}/models/spot/spot.obj"
    ptsd, _ = ob.ref_scene("mesh", obj_path=obj)
    crowd, _ = ob.ref_scene("crowd", crowd_side=6, obj_path=obj, turn_deg=150)
    np.savez_compressed(os.path.join(HERE, "spot_scene.npz"), spot_ptsd=ptsd, spot_crowd_ptsd=crowd)
    with open(obj, "rb") as src, gzip.open(os.path.join(HERE, "spot.obj.gz"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    shutil.copyfile(f"{REF}/models/spot/spot.mtl", os.path.join(HERE, "spot.mtl"))
    shutil.copyfile(f"{REF}/models/spot/spot_texture.png", os.path.join(HERE, "spot_texture.png"))
    print("spot", ptsd.size, "spot crowd", crowd.size)


MTL_KINDS = """newmtl lamp
Ke 4 3 2
newmtl clear
d 0.5
Ni 1.33
newmtl frosted
Tr 0.25
newmtl crystal
illum 7
Ni 0.9
newmtl copper
illum 3
Ks 0.9 0.6 0.5
Ns 40
newmtl steel
illum 5
Kd 0.6 0.6 0.7
newmtl chalk
Kd 0.7
newmtl rough
illum 3
Ns 0
"""


def write_mtl_kinds():
    import tempfile
    names = ["lamp", "clear", "frosted", "crystal", "copper", "steel", "chalk", "rough", "nosuch"]
    lines = ["mtllib kinds.mtl"]
    for k, name in enumerate(names):
        x = 2.0 * k
        lines += [f"v {x} 0 0", f"v {x + 1} 0 0", f"v {x + 1} 1 0.5", f"v {x} 1 0.25", f"usemtl {name}",
                  f"f {4 * k + 1} {4 * k + 2} {4 * k + 3} {4 * k + 4}"]
    obj = "\n".join(lines) + "\n"
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "kinds.obj"), "w").write(obj)
        open(os.path.join(d, "kinds.mtl"), "w").write(MTL_KINDS)
        ptsd, _ = ob.ref_scene("mesh", obj_path=os.path.join(d, "kinds.obj"))
    np.savez_compressed(os.path.join(HERE, "mtl_kinds.npz"), ptsd=ptsd,
                        obj=np.frombuffer(obj.encode(), np.uint8), mtl=np.frombuffer(MTL_KINDS.encode(), np.uint8))
    print("mtl kinds", ptsd.size)


if __name__ == "__main__":
    main()
