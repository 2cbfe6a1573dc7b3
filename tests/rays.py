"""Seeded ray batches for the parity tests: camera rays through the packed
camera (camera.hpp:45-47), incoherent secondary rays from their hit points,
and shadow rays towards the light spheres."""
from __future__ import annotations

import numpy as np

import oracle_bind as ob


def camera(ptgp: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(ptgp[22:38]).view(np.float32).reshape(4, 4)[:, :3].astype(np.float32)


def camera_rays(ptgp: np.ndarray, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    cam = camera(ptgp)
    u = rng.random(n, dtype=np.float32)
    v = rng.random(n, dtype=np.float32)
    d = (cam[1][None, :] + cam[2][None, :] * u[:, None] + cam[3][None, :] * v[:, None] - cam[0][None, :]).astype(np.float32)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = cam[0]
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = 1000.0
    return r


def secondary_rays(ptgp: np.ndarray, primary: np.ndarray, seed: int) -> np.ndarray:
    """From each primary hit point, a uniformly random direction."""
    t, _, _, kind, _, _, _ = ob.or_trace(ptgp, primary)
    hit = kind != 0
    o = (primary[hit, 0:3] + primary[hit, 4:7] * t[hit, None]).astype(np.float32)
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(len(o), 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = 1000.0
    return r


def shadow_rays(ptgp: np.ndarray, origins: np.ndarray, seed: int) -> np.ndarray:
    """From each origin toward a random point on a random light sphere, with
    t_max just short of it (renderer.hpp:69-73)."""
    sec_off, sec_len = int(ptgp[2 + 2 * 6]) // 4, int(ptgp[3 + 2 * 6]) // 4
    lights = ptgp[sec_off:sec_off + sec_len]
    if len(lights) == 0:
        raise ValueError("scene has no lights")
    sph_off = int(ptgp[2 + 2 * 2]) // 4
    spheres = np.ascontiguousarray(ptgp[sph_off:sph_off + 8 * (int(ptgp[3 + 2 * 2]) // 32)]).view(np.float32).reshape(-1, 8)
    rng = np.random.default_rng(seed)
    pick = lights[rng.integers(0, len(lights), len(origins))]
    c = spheres[pick, 0:3]
    rad = spheres[pick, 3:4]
    jitter = rng.normal(size=(len(origins), 3)).astype(np.float32)
    jitter /= np.linalg.norm(jitter, axis=1, keepdims=True)
    target = c + jitter * rad * 0.5
    d = (target - origins).astype(np.float32)
    r = np.empty((len(origins), 8), np.float32)
    r[:, 0:3] = origins
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = 0.9999
    return r


def mixed_batch(ptgp: np.ndarray, n_primary: int, seed: int) -> np.ndarray:
    p = camera_rays(ptgp, n_primary, seed)
    s = secondary_rays(ptgp, p, seed + 1)
    return np.concatenate([p, s])
