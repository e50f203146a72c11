"""Helpers shared by the tests: scenes, cameras, oracle <-> device setup."""
from __future__ import annotations

import ctypes as C
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


class DeviceArray:
    """A float32 array copied into device memory (hipMalloc / hipMemcpy)."""

    def __init__(self, a):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        a = np.ascontiguousarray(a, dtype=np.float32)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), max(a.nbytes, 4)) == 0
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
