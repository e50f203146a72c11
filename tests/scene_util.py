"""Helpers shared by the tests: scenes, cameras, oracle <-> device setup."""
from __future__ import annotations

import functools

import numpy as np

from lisa_amd import scenes


@functools.lru_cache(maxsize=None)
def cornell(seg, rings, width, height, spp, bounces):
    return scenes.cornell(seg, rings, width, height, spp, bounces)


def camera(sc, capi_or_oracle):
    W, H = sc["width"], sc["height"]
    cam = sc["camera"]
    return capi_or_oracle.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)


def random_rays(n, seed=0, inside=True):
    """Seeded rays from points inside the Cornell box in uniform directions."""
    rs = np.random.default_rng(seed)
    O = rs.uniform(-0.49, 0.49, size=(n, 3)).astype(np.float32)
    if not inside:
        O[:, 2] = 1.5
    D = rs.normal(size=(n, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    return O, D.astype(np.float32)
