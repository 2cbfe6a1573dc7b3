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


if __name__ == "__main__":
    main()
