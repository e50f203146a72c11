"""ctypes view of the CPU oracle (oracle/lisa_oracle.cpp).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker.  The oracle is
a scalar C++ restatement of src/LiSA/src/shader.cu, maths.cu,
bsdfs/lambertian.cu and src/cuda/random.h (see lisa_oracle.h for the list).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LISA_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of `make sanitize`)
LIB_PATH = os.environ.get("LISA_ORACLE_LIB") or os.path.join(HERE, "_build", "liblisa_oracle.so")
NCOUNTERS = 8
COUNTER_NAMES = ("rays_radiance", "rays_occlusion", "node_visits", "tri_tests",
                 "radiance_hits", "samples", "invalid_dir_rays", "pixels")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Material(C.Structure):
    _fields_ = [("roughness", C.c_float), ("alpha", C.c_float), ("n", C.c_float),
                ("diffuse", C.c_float * 3), ("emit", C.c_int32), ("emission", C.c_float * 3)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("bounces", C.c_uint32), ("subframe", C.c_uint32),
                ("eye", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3),
                ("W", C.c_float * 3), ("record_bounces", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32p, u32p, i32p, u64p, u8p = (C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint8))
        L.orc_tea16.restype = C.c_uint32
        L.orc_tea16.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_lcg.restype = C.c_uint32
        L.orc_lcg.argtypes = [u32p]
        L.orc_rnd.restype = C.c_float
        L.orc_rnd.argtypes = [u32p]
        L.orc_rng.restype = C.c_float
        L.orc_rng.argtypes = [u32p]
        L.orc_hemisphere.argtypes = [f32p, u32p, f32p]
        L.orc_fresnel.restype = C.c_float
        L.orc_fresnel.argtypes = [C.c_float, C.c_float]
        L.orc_refract.argtypes = [C.c_float, f32p, f32p, C.c_float, f32p]
        L.orc_reflect.argtypes = [f32p, f32p, f32p]
        L.orc_brdf.restype = C.c_float
        L.orc_brdf.argtypes = [f32p, f32p]
        L.orc_bounce.argtypes = [f32p, f32p, u32p, C.c_float, f32p]
        L.orc_bounce_many.argtypes = [f32p, f32p, u32p, C.c_float, C.c_int32, f32p]
        L.orc_bsdf_pdf.argtypes = [f32p, f32p, f32p, C.c_int32, C.c_float, f32p]
        L.orc_bary_normal.argtypes = [f32p] * 8
        L.orc_camera_uvw.argtypes = [f32p, f32p, C.c_float, C.c_float, f32p, f32p, f32p]
        L.orc_srgb_quantize.argtypes = [f32p, C.c_int, u8p]
        L.orc_save_ppm.restype = C.c_int
        L.orc_save_ppm.argtypes = [C.c_char_p, u8p, C.c_int, C.c_int]
        L.orc_morton3.restype = C.c_uint32
        L.orc_morton3.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_nee_scripted.restype = C.c_int32
        L.orc_nee_scripted.argtypes = [i32p, C.c_int32, C.c_int32, C.c_int32, i32p, i32p]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [f32p, f32p, i32p, C.c_int32, C.POINTER(Material), C.c_int32]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_update_vertices.argtypes = [C.c_void_p, f32p, f32p]
        L.orc_num_tris.restype = C.c_int32
        L.orc_num_tris.argtypes = [C.c_void_p]
        L.orc_get_bvh.restype = C.c_int
        L.orc_get_bvh.argtypes = [C.c_void_p, C.c_int32, u32p, u32p, u32p, f32p]
        L.orc_num_emitters.restype = C.c_int32
        L.orc_num_emitters.argtypes = [C.c_void_p]
        L.orc_scene_set_emitter_cull.argtypes = [C.c_void_p, C.c_int32]
        L.orc_raw_cull.argtypes = [C.c_void_p, f32p, f32p, C.c_int32, i32p]
        L.orc_occlusion.argtypes = [C.c_void_p, f32p, f32p, C.c_int32, i32p, i32p, i32p, u64p]
        L.orc_closest_hit.argtypes = [C.c_void_p, f32p, f32p, C.c_int32, C.c_float, C.c_float,
                                      i32p, f32p, u64p, C.c_int]
        L.orc_closest_hit_brute.argtypes = [C.c_void_p, f32p, f32p, C.c_int32, C.c_float,
                                            C.c_float, i32p, f32p]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), i32p, C.c_int32, f32p, f32p,
                                 f32p, i32p, i32p, u32p, i32p, u64p, C.c_int]
        L.orc_cache_bounds.argtypes = [C.c_void_p, f32p, f32p]
        L.orc_cache_queries.restype = C.c_int
        L.orc_cache_queries.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.c_uint32, i32p,
                                        C.c_int32, f32p, f32p, f32p, i32p, C.c_int32, i32p, C.c_int,
                                        C.c_int, C.c_int, C.c_int]
        L.orc_render_full_paths.restype = C.c_int
        L.orc_render_full_paths.argtypes = [C.c_void_p, C.POINTER(RenderParams), i32p, C.c_int32,
                                            f32p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_cache_train_rows.restype = C.c_int
        L.orc_cache_train_rows.argtypes = [C.c_void_p, C.POINTER(RenderParams), i32p, C.c_int32,
                                           C.c_uint32, f32p, f32p, i32p, C.c_int32, i32p, C.c_int]
        _lib = L
    return _lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _vec(v):
    return f32(v).ctypes.data_as(C.POINTER(C.c_float))


# ----- scalar restatements (for known-answer tests) -----
def tea16(a, b):
    return lib().orc_tea16(a & 0xFFFFFFFF, b & 0xFFFFFFFF)


def lcg_seq(seed, n):
    s = C.c_uint32(seed)
    return [lib().orc_lcg(C.byref(s)) for _ in range(n)], s.value


def rnd_seq(seed, n):
    s = C.c_uint32(seed)
    return [lib().orc_rnd(C.byref(s)) for _ in range(n)], s.value


def rng_seq(seed, n):
    s = C.c_uint32(seed)
    return [lib().orc_rng(C.byref(s)) for _ in range(n)], s.value


def hemisphere(N, seed):
    s = C.c_uint32(seed)
    out = np.zeros(3, np.float32)
    n = f32(N)
    lib().orc_hemisphere(_p(n, C.c_float), C.byref(s), _p(out, C.c_float))
    return out, s.value


def fresnel(cosT, eta):
    return lib().orc_fresnel(cosT, eta)


def refract(cosI, d, N, eta):
    out = np.zeros(3, np.float32)
    d, N = f32(d), f32(N)
    lib().orc_refract(cosI, _p(d, C.c_float), _p(N, C.c_float), eta, _p(out, C.c_float))
    return out


def reflect(i, n):
    out = np.zeros(3, np.float32)
    i, n = f32(i), f32(n)
    lib().orc_reflect(_p(i, C.c_float), _p(n, C.c_float), _p(out, C.c_float))
    return out


def brdf(N, L):
    N, L = f32(N), f32(L)
    return lib().orc_brdf(_p(N, C.c_float), _p(L, C.c_float))


def bounce(d, N, seed, roughness):
    s = C.c_uint32(seed)
    out = np.zeros(3, np.float32)
    d, N = f32(d), f32(N)
    lib().orc_bounce(_p(d, C.c_float), _p(N, C.c_float), C.byref(s), roughness, _p(out, C.c_float))
    return out, s.value


def bounce_many(d, N, seed, roughness, n):
    """n bounce directions (not normalised) from one RNG stream."""
    s = C.c_uint32(seed)
    out = np.zeros((n, 3), np.float32)
    d, N = f32(d), f32(N)
    lib().orc_bounce_many(_p(d, C.c_float), _p(N, C.c_float), C.byref(s), roughness, n,
                          _p(out, C.c_float))
    return out


def bsdf_pdf(d, N, L, roughness):
    """Density per steradian of bounce(d, N, ., roughness) at the unit directions L (n x 3)."""
    L = f32(L).reshape(-1, 3)
    out = np.zeros(L.shape[0], np.float32)
    d, N = f32(d), f32(N)
    lib().orc_bsdf_pdf(_p(d, C.c_float), _p(N, C.c_float), _p(L, C.c_float), L.shape[0], roughness,
                       _p(out, C.c_float))
    return out


def bary_normal(P, n1, n2, n3, v1, v2, v3):
    args = [f32(a) for a in (P, n1, n2, n3, v1, v2, v3)]
    out = np.zeros(3, np.float32)
    lib().orc_bary_normal(*[_p(a, C.c_float) for a in args], _p(out, C.c_float))
    return out


def camera_uvw(eye, lookat, fov, aspect):
    U, V, W = (np.zeros(3, np.float32) for _ in range(3))
    e, l = f32(eye), f32(lookat)
    lib().orc_camera_uvw(_p(e, C.c_float), _p(l, C.c_float), fov, aspect,
                         _p(U, C.c_float), _p(V, C.c_float), _p(W, C.c_float))
    return U, V, W


def srgb_quantize(rgba):
    rgba = f32(rgba).reshape(-1, 4)
    out = np.zeros((rgba.shape[0], 4), np.uint8)
    lib().orc_srgb_quantize(_p(rgba, C.c_float), rgba.shape[0], _p(out, C.c_uint8))
    return out


def save_ppm(path, rgba8, width, height):
    a = np.ascontiguousarray(rgba8, dtype=np.uint8)
    return lib().orc_save_ppm(path.encode(), _p(a, C.c_uint8), width, height)


def morton3(x, y, z):
    return lib().orc_morton3(x, y, z)


def nee_scripted(outcomes, hit_in=False, light_in=-1):
    o = np.ascontiguousarray(outcomes, dtype=np.int32)
    h, l = C.c_int32(), C.c_int32()
    r = lib().orc_nee_scripted(_p(o, C.c_int32) if o.size else None, o.size, int(hit_in),
                               light_in, C.byref(h), C.byref(l))
    return r, bool(h.value), l.value


def materials_array(mats):
    arr = (Material * max(len(mats), 1))()
    for i, m in enumerate(mats):
        arr[i].roughness = m.get("roughness", 0.0)
        arr[i].alpha = m.get("alpha", 1.0)
        arr[i].n = m.get("n", 1.0)
        arr[i].diffuse[:] = list(m.get("diffuse", (0.0, 0.0, 0.0)))
        arr[i].emit = 1 if m.get("emit", False) else 0
        arr[i].emission[:] = list(m.get("emission", (0.0, 0.0, 0.0)))
    return arr


class Scene:
    """Triangle soup + LBVH held by the oracle."""

    def __init__(self, verts, normals, mat_idx, materials):
        L = lib()
        self.verts = f32(verts).reshape(-1, 9)
        self.normals = f32(normals).reshape(-1, 9)
        self.mat_idx = np.ascontiguousarray(mat_idx, dtype=np.int32)
        self.T = int(self.verts.shape[0])
        self._mats = materials_array(materials)
        self.h = L.orc_scene_create(_p(self.verts, C.c_float), _p(self.normals, C.c_float),
                                    _p(self.mat_idx, C.c_int32), self.T, self._mats,
                                    len(materials))

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_vertices(self, verts, normals=None):
        self.verts = f32(verts).reshape(-1, 9)
        assert self.verts.shape[0] == self.T
        n = None if normals is None else f32(normals).reshape(-1, 9)
        if n is not None:
            self.normals = n
        lib().orc_scene_update_vertices(self.h, _p(self.verts, C.c_float), _p(n, C.c_float))

    def raw_cull(self, O, d):
        """The NEE loop's cull of raw hemisphere draws: 1 = the line certainly
        misses every emitter."""
        O, d = f32(O).reshape(-1, 3), f32(d).reshape(-1, 3)
        out = np.zeros(O.shape[0], np.int32)
        lib().orc_raw_cull(self.h, _p(O, C.c_float), _p(d, C.c_float), O.shape[0], _p(out, C.c_int32))
        return out

    def set_emitter_cull(self, enabled):
        """Switch the emitter-query culls off or on (results unchanged)."""
        lib().orc_scene_set_emitter_cull(self.h, 1 if enabled else 0)

    def bvh(self, which=0):
        T = self.T if which == 0 else lib().orc_num_emitters(self.h)
        nI = max(T - 1, 1)
        m = np.zeros(max(T, 1), np.uint32)
        p = np.zeros(max(T, 1), np.uint32)
        r = np.zeros(2 * nI, np.uint32)
        b = np.zeros(12 * nI, np.float32)
        lib().orc_get_bvh(self.h, which, _p(m, C.c_uint32), _p(p, C.c_uint32),
                          _p(r, C.c_uint32), _p(b, C.c_float))
        return {"morton": m[:T], "prim": p[:T], "refs": r.reshape(nI, 2),
                "boxes": b.reshape(nI, 12)}

    def occlusion(self, O, D, hit_before):
        O, D = f32(O).reshape(-1, 3), f32(D).reshape(-1, 3)
        hb = np.ascontiguousarray(hit_before, dtype=np.int32)
        n = O.shape[0]
        fast = np.zeros(n, np.int32)
        ref = np.zeros(n, np.int32)
        cnt = np.zeros(NCOUNTERS, np.uint64)
        lib().orc_occlusion(self.h, _p(O, C.c_float), _p(D, C.c_float), n, _p(hb, C.c_int32),
                            _p(fast, C.c_int32), _p(ref, C.c_int32), _p(cnt, C.c_uint64))
        return fast, ref, cnt

    def closest_hit(self, O, D, tmin=1e-4, tmax=1e16, threads=1, brute=False):
        O, D = f32(O).reshape(-1, 3), f32(D).reshape(-1, 3)
        n = O.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        cnt = np.zeros(NCOUNTERS, np.uint64)
        if brute:
            lib().orc_closest_hit_brute(self.h, _p(O, C.c_float), _p(D, C.c_float), n, tmin, tmax,
                                        _p(prim, C.c_int32), _p(t, C.c_float))
        else:
            lib().orc_closest_hit(self.h, _p(O, C.c_float), _p(D, C.c_float), n, tmin, tmax,
                                  _p(prim, C.c_int32), _p(t, C.c_float), _p(cnt, C.c_uint64),
                                  threads)
        return prim, t, cnt

    def cache_bounds(self):
        """(bmin, 1 / extent) of the cache's position features."""
        bmin = np.zeros(3, np.float32)
        binv = np.zeros(3, np.float32)
        lib().orc_cache_bounds(self.h, _p(bmin, C.c_float), _p(binv, C.c_float))
        return bmin, binv

    @staticmethod
    def _params(width, height, spp, bounces, eye, U, V, W, subframe):
        p = RenderParams()
        p.width, p.height, p.spp, p.bounces, p.subframe = width, height, spp, bounces, subframe
        p.eye[:], p.U[:], p.V[:], p.W[:] = list(eye), list(U), list(V), list(W)
        p.record_bounces = 0
        return p

    def cache_queries(self, width, height, spp, prefix_bounces, eye, U, V, W, pixels, subframe=0,
                      threads=1, emit_expect=False, nee_loops=1, nee_loops_deep=None):
        """Prefix paths ending in cache queries (orc_cache_queries); emit_expect
        takes the prefix bounces' emission in expectation, nee_loops averages
        that many NEE runs at the camera ray's hit and nee_loops_deep (default:
        nee_loops) at later vertices."""
        if nee_loops_deep is None:
            nee_loops_deep = nee_loops
        p = self._params(width, height, spp, prefix_bounces + 1, eye, U, V, W, subframe)
        px = np.ascontiguousarray(pixels, dtype=np.int32)
        n = px.shape[0]
        cap = n * spp
        acc = np.zeros((n, 4), np.float32)
        f = np.zeros((cap, 12), np.float32)
        at = np.zeros((cap, 3), np.float32)
        it = np.zeros(cap, np.int32)
        nq = C.c_int32()
        rc = lib().orc_cache_queries(self.h, C.byref(p), prefix_bounces, _p(px, C.c_int32), n,
                                     _p(acc, C.c_float), _p(f, C.c_float), _p(at, C.c_float),
                                     _p(it, C.c_int32), cap, C.byref(nq), threads,
                                     1 if emit_expect else 0, nee_loops, nee_loops_deep)
        if rc != 0:
            raise RuntimeError(f"orc_cache_queries failed: {rc}")
        k = nq.value
        return {"accum": acc, "feat": f[:k], "att": at[:k], "item": it[:k], "n": k}

    def render_full_paths(self, width, height, spp, bounces, eye, U, V, W, pixels, subframe=0,
                          threads=1, emit_expect=True, nee_loops=1, nee_loops_deep=None):
        """The variance-reduced plain tracer (orc_render_full_paths): whole paths with
        the cache prefix's estimators and no query; per-pixel means (n, 4)."""
        if nee_loops_deep is None:
            nee_loops_deep = nee_loops
        p = self._params(width, height, spp, bounces, eye, U, V, W, subframe)
        px = np.ascontiguousarray(pixels, dtype=np.int32)
        acc = np.zeros((px.shape[0], 4), np.float32)
        rc = lib().orc_render_full_paths(self.h, C.byref(p), _p(px, C.c_int32), px.shape[0],
                                         _p(acc, C.c_float), threads, 1 if emit_expect else 0,
                                         nee_loops, nee_loops_deep)
        if rc != 0:
            raise RuntimeError(f"orc_render_full_paths failed: {rc}")
        return acc

    def cache_train_rows(self, width, height, spp, bounces, eye, U, V, W, pixels, subframe=0,
                         depth=0, threads=1):
        """Training rows of plain paths (orc_cache_train_rows): every surface hit, or the hits
        of radiance ray number `depth` only."""
        p = self._params(width, height, spp, bounces, eye, U, V, W, subframe)
        px = np.ascontiguousarray(pixels, dtype=np.int32)
        n = px.shape[0]
        cap = n * spp * bounces
        f = np.zeros((cap, 12), np.float32)
        t = np.zeros((cap, 3), np.float32)
        path = np.zeros(cap, np.int32)
        nr = C.c_int32()
        rc = lib().orc_cache_train_rows(self.h, C.byref(p), _p(px, C.c_int32), n, depth,
                                        _p(f, C.c_float),
                                        _p(t, C.c_float), _p(path, C.c_int32), cap, C.byref(nr),
                                        threads)
        if rc != 0:
            raise RuntimeError(f"orc_cache_train_rows failed: {rc}")
        k = nr.value
        return {"feat": f[:k], "tgt": t[:k], "path": path[:k], "n": k}

    def render(self, width, height, spp, bounces, eye, U, V, W, subframe=0, pixels=None,
               accum_prev=None, record=False, record_bounces=None, threads=1):
        p = RenderParams()
        p.width, p.height, p.spp, p.bounces, p.subframe = width, height, spp, bounces, subframe
        p.eye[:], p.U[:], p.V[:], p.W[:] = list(eye), list(U), list(V), list(W)
        rb = bounces if record_bounces is None else record_bounces
        p.record_bounces = rb if record else 0
        px = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.int32)
        n = width * height if px is None else px.shape[0]
        accum = np.zeros((n, 4), np.float32)
        prev = None if accum_prev is None else f32(accum_prev).reshape(n, 4)
        cnt = np.zeros(NCOUNTERS, np.uint64)
        out = {"accum": accum, "counters": cnt}
        srgb = sb = so = sh = rh = None
        if record:
            srgb = np.zeros((n, spp, 3), np.float32)
            sb = np.zeros((n, spp), np.int32)
            so = np.zeros((n, spp), np.int32)
            sh = np.zeros((n, spp), np.uint32)
            rh = np.zeros((n, spp, max(rb, 1)), np.int32) if rb > 0 else None
            out.update(sample_rgb=srgb, sample_bounces=sb, sample_occl=so, sample_hash=sh,
                       rad_hits=rh)
        rc = lib().orc_render(self.h, C.byref(p), _p(px, C.c_int32), n, _p(accum, C.c_float),
                              _p(prev, C.c_float), _p(srgb, C.c_float), _p(sb, C.c_int32),
                              _p(so, C.c_int32), _p(sh, C.c_uint32), _p(rh, C.c_int32),
                              _p(cnt, C.c_uint64), threads)
        if rc != 0:
            raise RuntimeError(f"orc_render failed: {rc}")
        return out
