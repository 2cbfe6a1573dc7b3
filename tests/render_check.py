"""The reference's statistical render check (tests/render_test.py:36-40,74,79,
104-182), restated: 20x20 block means of a linear image against the fixture's
per-block mean and sigma (16 seeds), with |z| limits per block, per block row
and for the whole image."""
from __future__ import annotations

import gzip
import json
import os

import numpy as np

WIDTH, HEIGHT, SPP, BLOCK, SEEDS = 960, 540, 200, 20, 16   # render_test.py:36-40
Z_BLOCK, Z_ROW, Z_IMAGE = 10.0, 6.0, 5.0                   # render_test.py:74
SIGMA_FLOOR = 2e-4                                         # render_test.py:79

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_reference() -> dict:
    with gzip.open(os.path.join(GOLDEN, "reference.json.gz"), "rt") as f:
        return json.load(f)


def load_scene(name: str):
    """(PTSD, PTGP) of a reference scene, as the reference built and packed it."""
    z = np.load(os.path.join(GOLDEN, "ref_scenes.npz"))
    return z[f"{name}_ptsd"], z[f"{name}_ptgp"]


def load_spot(name: str = "spot_crowd"):
    """A PTSD scene of the reference's textured Spot, as the reference built it
    and exported it with its decoded texture: "spot" (its mesh scene loading
    models/spot/spot.obj) or "spot_crowd" (tests/render_test.py:55: a crowd
    of 6 x 6 turned 150 degrees)."""
    return np.load(os.path.join(GOLDEN, "spot_scene.npz"))[f"{name}_ptsd"]


def ptsd_textures(w: np.ndarray) -> list:
    """The textures of a PTSD buffer as (h, w, 3) uint8 sRGB arrays: after the
    header, materials, spheres, triangles, mesh sizes and triangles and the
    instances, each texture's width and height then its bytes, word padded."""
    pos = 32 + 12 * int(w[10]) + 5 * (int(w[11]) + int(w[12])) + 28 * int(w[13])
    pos += int(w[14]) + 28 * int(w[16]) + 16 * int(w[15])
    out = []
    for _ in range(int(w[17])):
        tw, th = int(w[pos]), int(w[pos + 1])
        nb = 3 * tw * th
        out.append(w[pos + 2:pos + 2 + (nb + 3) // 4].view(np.uint8)[:nb].reshape(th, tw, 3))
        pos += 2 + (nb + 3) // 4
    assert pos == len(w)
    return out


def block_means(img: np.ndarray, block: int = BLOCK) -> np.ndarray:
    """(bh, bw, 3) block means of an (h, w, 3) image, top row first."""
    h, w, _ = img.shape
    bh, bw = h // block, w // block
    return img[: bh * block, : bw * block].astype(np.float64).reshape(bh, block, bw, block, 3).mean(axis=(1, 3))


def z_scores(img: np.ndarray, ref: dict):
    """(worst block |z|, worst row |z|, image z per channel) against one scene's fixture."""
    bw, bh = ref["blocks"]
    means = block_means(img)
    assert means.shape[:2] == (bh, bw), "block grid differs from the fixture"
    m = means.reshape(-1)
    r = np.asarray(ref["mean"], np.float64)
    s = np.asarray(ref["sigma"], np.float64)
    scale = np.sqrt(1 + 1 / SEEDS)
    diff = m - r
    var = (np.maximum(s, SIGMA_FLOOR) * scale) ** 2
    worst_block = float(np.max(np.abs(diff) / np.sqrt(var)))
    d3 = diff.reshape(bh, bw, 3)
    v3 = var.reshape(bh, bw, 3)
    worst_row = float(np.max(np.abs(d3.sum(axis=1) / np.sqrt(v3.sum(axis=1)))))
    image = d3.sum(axis=(0, 1)) / np.sqrt(v3.sum(axis=(0, 1)))
    return worst_block, worst_row, image


def passes(img: np.ndarray, ref: dict) -> tuple[bool, str]:
    b, r, im = z_scores(img, ref)
    ok = b <= Z_BLOCK and r <= Z_ROW and float(np.max(np.abs(im))) <= Z_IMAGE
    return ok, f"block z {b:.2f}, row z {r:.2f}, image z {', '.join(f'{z:+.2f}' for z in im)}"


def block_mre(a: np.ndarray, b: np.ndarray, block: int = BLOCK, floor: float = 1e-3) -> float:
    """Mean over blocks and channels of |A - B| / max(B, floor)."""
    ma, mb = block_means(a, block), block_means(b, block)
    return float(np.mean(np.abs(ma - mb) / np.maximum(mb, floor)))


def per_value_agreement(a: np.ndarray, b: np.ndarray, rel: float = 1e-4) -> tuple[float, str]:
    """Fraction of values of a within rel * max(1, |b|) of b, and a note on
    the ones that are not (how many, the largest difference and where)."""
    diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
    tol = rel * np.maximum(1.0, np.abs(b.astype(np.float64)))
    close = diff <= tol
    bad = np.argwhere(~close)
    if len(bad) == 0:
        return 1.0, "all values agree"
    worst = tuple(int(x) for x in np.unravel_index(np.argmax(diff - tol), diff.shape))
    return float(close.mean()), (f"{len(bad)} of {diff.size} values differ, the largest by {diff[worst]:.4g} at "
                                 f"{worst} ({float(a[worst]):.6g} against {float(b[worst]):.6g})")
