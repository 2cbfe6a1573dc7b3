"""Small PTSD scenes written out word by word for the tests (the layout is
the one photon_tracer_amd/csrc/host.hpp documents)."""
import numpy as np


def textured_quads(*images):
    """A PTSD scene: one matte quad per image (material 0 textured by the
    first, material 2 by the second), a light (material 1), camera at +z."""
    f = lambda x: int(np.float32(x).view(np.uint32))
    head = np.zeros(32, np.uint32)
    head[0], head[1], head[2] = 0x44535450, 1, 1
    head[3:10] = [f(0), f(0), f(3), f(0), f(0), f(0), f(40)]
    mats = np.zeros((3, 12), np.uint32)
    mats[0, :3] = [f(0.9)] * 3
    mats[0, 9] = 1
    mats[1, 3], mats[1, 4:7] = 3, [f(2.0)] * 3
    mats[2, :3] = [f(0.5), f(0.7), f(0.9)]
    mats[2, 9] = 2 if len(images) > 1 else 0
    mats[:, 8] = f(1.5)
    tris = []
    for k, (y0, m) in enumerate(((-1.0, 0), (0.0, 2))):
        quad = [(-1, y0, 0), (1, y0, 0), (1, y0 + 1, 0), (-1, y0 + 1, 0)]
        uvq = [(0, 0), (1, 0), (1, 1), (0, 1)]
        for a, b, c in ((0, 1, 2), (0, 2, 3)):
            t = np.zeros(28, np.uint32)
            t[0:9] = [f(x) for p in (quad[a], quad[b], quad[c]) for x in p]
            t[9:18] = [f(x) for _ in range(3) for x in (0, 0, 1)]
            t[18:24] = [f(x) for q in (uvq[a], uvq[b], uvq[c]) for x in (2 * q[0] - 0.5, q[1])]
            t[24] = m
            tris.append(t)
    light = np.array([f(0), f(3), f(2), f(0.5), 1], np.uint32)
    head[10], head[11], head[13], head[17] = 3, 1, len(tris), len(images)
    tex = []
    for im in images:
        h, w, _ = im.shape
        raw = np.zeros(((3 * w * h + 3) // 4) * 4, np.uint8)
        raw[:3 * w * h] = im.reshape(-1)
        tex += [np.array([w, h], np.uint32), raw.view(np.uint32)]
    return np.concatenate([head, mats.reshape(-1), light, np.concatenate(tris)] + tex)
