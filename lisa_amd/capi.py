"""ctypes binding of liblisa.so (include/lisa.h).

The library is the product: this module only marshals arrays.  It never
falls back to anything else -- if liblisa.so is missing or a device call
fails, it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LISA_LIBRARY selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("LISA_LIBRARY") or os.path.join(HERE, "_build", "liblisa.so")
CLI_PATH = os.path.join(HERE, "_build", "lisa")

OK = 0
ERR_INVALID_ARG = -1
ERR_HIP = -2
ERR_ALLOC = -3
ERR_NO_SCENE = -4
ERR_PARSE = -5
ERR_IO = -6
ERR_UNSUPPORTED = -7

MODE_PLAIN = 0
FLAG_FP64_ACCUM = 1
FLAG_COUNT = 2
FLAG_PERSISTENT = 4
FLAG_WAVEFRONT = 8
NCOUNTERS = 8
COUNTER_NAMES = ("rays_radiance", "rays_occlusion", "node_visits", "tri_tests",
                 "radiance_hits", "samples", "invalid_dir_rays", "pixels")

# Every function declared in include/lisa.h.
EXPORTS = (
    "lisa_create", "lisa_destroy", "lisa_last_error", "lisa_set_scene", "lisa_update_vertices",
    "lisa_update_vertices_noalloc",
    "lisa_render", "lisa_render_progressive",
    "lisa_copy_accum", "lisa_accum64_save", "lisa_accum64_load", "lisa_synchronize", "lisa_get_stats",
    "lisa_set_wide_threshold", "lisa_get_counters",
    "lisa_get_stream", "lisa_get_bvh", "lisa_trace_rays", "lisa_trace_subset",
    "lisa_batch_create", "lisa_batch_destroy", "lisa_batch_set_config", "lisa_batch_get_config",
    "lisa_batch_render", "lisa_batch_autotune",
    "lisa_cache_create", "lisa_cache_destroy", "lisa_cache_last_error", "lisa_cache_set_weights",
    "lisa_cache_get_weights", "lisa_cache_infer", "lisa_cache_gradients", "lisa_cache_train",
    "lisa_cache_last_ms", "lisa_render_cached", "lisa_cache_trace_queries", "lisa_cache_trace_rows",
    "lisa_camera_uvw", "lisa_srgb_quantize", "lisa_save_ppm", "lisa_scene_file_load",
    "lisa_scene_file_free", "lisa_scene_file_info", "lisa_scene_file_vertices",
    "lisa_scene_file_normals", "lisa_scene_file_mat_indices", "lisa_scene_file_materials",
    "lisa_scene_file_output_image",
)


class LisaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"lisa status {status}: {msg}")
        self.status = status


class Material(C.Structure):
    _fields_ = [("roughness", C.c_float), ("alpha", C.c_float), ("n", C.c_float),
                ("diffuse", C.c_float * 3), ("emit", C.c_int32), ("emission", C.c_float * 3)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("bounces", C.c_uint32), ("subframe", C.c_uint32),
                ("eye", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3),
                ("W", C.c_float * 3), ("mode", C.c_int32), ("flags", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("render_ms", C.c_float), ("bvh_build_ms", C.c_float), ("num_tris", C.c_int32),
                ("num_nodes", C.c_int32), ("num_emitters", C.c_int32),
                ("device_bytes", C.c_uint64), ("wide_tree", C.c_int32), ("reserved", C.c_int32)]


class BatchConfig(C.Structure):
    _fields_ = [("grid_blocks", C.c_int32), ("chunk", C.c_int32), ("form", C.c_int32),
                ("graph", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [("num_tris", C.c_int32), ("num_materials", C.c_int32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("num_samples", C.c_uint32), ("num_bounces", C.c_uint32),
                ("eye", C.c_float * 3), ("look_at", C.c_float * 3), ("fov", C.c_float)]


_lib = None


def lib():
    """Load liblisa.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m lisa_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, f32p, u32p, i32p, u64p, u8p = (C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint8))
    st = C.c_int32
    sig = {
        "lisa_create": (st, [C.POINTER(C.c_void_p), C.c_int32]),
        "lisa_destroy": (None, [vp]),
        "lisa_last_error": (C.c_char_p, [vp]),
        "lisa_set_scene": (st, [vp, f32p, f32p, i32p, C.c_int32, C.POINTER(Material), C.c_int32]),
        "lisa_update_vertices": (st, [vp, f32p, f32p]),
        "lisa_update_vertices_noalloc": (st, [vp, vp, vp]),
        "lisa_render": (st, [vp, C.POINTER(RenderParams), f32p, u8p]),
        "lisa_render_progressive": (st, [vp, C.POINTER(RenderParams), C.c_int32]),
        "lisa_copy_accum": (st, [vp, f32p, C.c_int64]),
        "lisa_accum64_save": (st, [vp, C.POINTER(C.c_double), C.c_int64]),
        "lisa_accum64_load": (st, [vp, C.POINTER(C.c_double), C.c_int64]),
        "lisa_synchronize": (st, [vp]),
        "lisa_get_stats": (st, [vp, C.POINTER(Stats)]),
        "lisa_set_wide_threshold": (st, [vp, C.c_int64]),
        "lisa_get_counters": (st, [vp, u64p]),
        "lisa_get_stream": (vp, [vp]),
        "lisa_get_bvh": (st, [vp, C.c_int32, u32p, u32p, u32p, f32p]),
        "lisa_trace_rays": (st, [vp, f32p, f32p, C.c_int32, C.c_float, C.c_float, i32p, f32p]),
        "lisa_trace_subset": (st, [vp, C.POINTER(RenderParams), i32p, C.c_int32, f32p, f32p,
                                   i32p, i32p, u32p, i32p, C.c_int32]),
        "lisa_batch_create": (st, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(RenderParams), C.c_int32]),
        "lisa_batch_destroy": (None, [vp]),
        "lisa_batch_set_config": (st, [vp, C.POINTER(BatchConfig)]),
        "lisa_batch_get_config": (st, [vp, C.POINTER(BatchConfig)]),
        "lisa_batch_render": (st, [vp, f32p]),
        "lisa_batch_autotune": (st, [vp, C.c_int32, C.POINTER(BatchConfig)]),
        "lisa_cache_create": (st, [C.POINTER(C.c_void_p), C.c_int32, C.c_uint64, C.c_float]),
        "lisa_cache_destroy": (None, [vp]),
        "lisa_cache_last_error": (C.c_char_p, [vp]),
        "lisa_cache_set_weights": (st, [vp, f32p]),
        "lisa_cache_get_weights": (st, [vp, f32p]),
        "lisa_cache_infer": (st, [vp, f32p, C.c_int64, f32p]),
        "lisa_cache_gradients": (st, [vp, f32p, f32p, C.c_int64, f32p, f32p]),
        "lisa_cache_train": (st, [vp, f32p, f32p, C.c_int64, f32p]),
        "lisa_cache_last_ms": (C.c_float, [vp]),
        "lisa_render_cached": (st, [vp, vp, C.POINTER(RenderParams), C.POINTER(CacheFrameParams),
                                    C.POINTER(CacheFrameStats)]),
        "lisa_cache_trace_queries": (st, [vp, C.POINTER(RenderParams), C.POINTER(CacheFrameParams), i32p,
                                          C.c_int32, f32p, f32p, f32p, i32p, C.c_int32, i32p]),
        "lisa_cache_trace_rows": (st, [vp, C.POINTER(RenderParams), i32p, C.c_int32, C.c_uint32,
                                       f32p, f32p, C.c_int32, i32p]),
        "lisa_camera_uvw": (st, [f32p, f32p, C.c_float, C.c_float, f32p, f32p, f32p]),
        "lisa_srgb_quantize": (st, [f32p, C.c_int64, u8p]),
        "lisa_save_ppm": (st, [C.c_char_p, u8p, C.c_int32, C.c_int32]),
        "lisa_scene_file_load": (st, [C.c_char_p, C.POINTER(C.c_void_p)]),
        "lisa_scene_file_free": (None, [vp]),
        "lisa_scene_file_info": (st, [vp, C.POINTER(SceneInfo)]),
        "lisa_scene_file_vertices": (f32p, [vp]),
        "lisa_scene_file_normals": (f32p, [vp]),
        "lisa_scene_file_mat_indices": (i32p, [vp]),
        "lisa_scene_file_materials": (C.POINTER(Material), [vp]),
        "lisa_scene_file_output_image": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _check(status, ctx=None):
    if status != OK:
        msg = lib().lisa_last_error(ctx)
        raise LisaError(status, msg.decode() if msg else "")


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


def camera_uvw(eye, look_at, fov, aspect):
    U, V, W = (np.zeros(3, np.float32) for _ in range(3))
    e, l = f32(eye), f32(look_at)
    _check(lib().lisa_camera_uvw(_p(e, C.c_float), _p(l, C.c_float), fov, aspect,
                                 _p(U, C.c_float), _p(V, C.c_float), _p(W, C.c_float)))
    return U, V, W


def srgb_quantize(rgba):
    rgba = f32(rgba).reshape(-1, 4)
    out = np.zeros((rgba.shape[0], 4), np.uint8)
    _check(lib().lisa_srgb_quantize(_p(rgba, C.c_float), rgba.shape[0], _p(out, C.c_uint8)))
    return out


def save_ppm(path, rgba8, width, height):
    a = np.ascontiguousarray(rgba8, dtype=np.uint8)
    _check(lib().lisa_save_ppm(path.encode(), _p(a, C.c_uint8), width, height))


def render_params(width, height, spp, bounces, eye, U, V, W, subframe=0, flags=0,
                  mode=MODE_PLAIN):
    p = RenderParams()
    p.width, p.height, p.spp, p.bounces, p.subframe = width, height, spp, bounces, subframe
    p.eye[:], p.U[:], p.V[:], p.W[:] = [list(map(float, v)) for v in (eye, U, V, W)]
    p.mode, p.flags = mode, flags
    return p


class SceneFile:
    """A parsed `.rto` scene (SceneParser, src/LiSA/src/scene_parser.cc)."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().lisa_scene_file_load(str(path).encode(), C.byref(h)))
        self.h = h
        info = SceneInfo()
        _check(lib().lisa_scene_file_info(self.h, C.byref(info)))
        T, M = info.num_tris, info.num_materials
        L = lib()
        self.num_tris, self.num_materials = T, M
        self.width, self.height = info.width, info.height
        self.num_samples, self.num_bounces = info.num_samples, info.num_bounces
        self.eye = np.array(info.eye[:], np.float32)
        self.look_at = np.array(info.look_at[:], np.float32)
        self.fov = float(info.fov)
        self.verts = np.ctypeslib.as_array(L.lisa_scene_file_vertices(self.h), (max(T, 1) * 9,))[
            : 9 * T].reshape(T, 9).copy() if T else np.zeros((0, 9), np.float32)
        self.normals = np.ctypeslib.as_array(L.lisa_scene_file_normals(self.h), (max(T, 1) * 9,))[
            : 9 * T].reshape(T, 9).copy() if T else np.zeros((0, 9), np.float32)
        self.mat_idx = np.ctypeslib.as_array(L.lisa_scene_file_mat_indices(self.h), (max(T, 1),))[
            :T].copy() if T else np.zeros(0, np.int32)
        mp = L.lisa_scene_file_materials(self.h)
        self.materials = []
        for i in range(M):
            m = mp[i]
            self.materials.append({"roughness": m.roughness, "alpha": m.alpha, "n": m.n,
                                   "diffuse": tuple(m.diffuse), "emit": bool(m.emit),
                                   "emission": tuple(m.emission)})
        self.output_image = L.lisa_scene_file_output_image(self.h).decode()
        L.lisa_scene_file_free(self.h)
        self.h = None


class Context:
    """One device context (lisa_ctx): scene, LBVH and accum buffer in HBM."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().lisa_create(C.byref(h), device))
        self.h = h
        self.T = 0

    def close(self):
        if getattr(self, "h", None):
            lib().lisa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, status):
        _check(status, self.h)

    def set_scene(self, verts, normals, mat_idx, materials):
        v = f32(verts).reshape(-1, 9)
        n = f32(normals).reshape(-1, 9)
        m = np.ascontiguousarray(mat_idx, dtype=np.int32)
        mats = materials_array(materials)
        self.T = int(v.shape[0])
        self._ck(lib().lisa_set_scene(self.h, _p(v, C.c_float), _p(n, C.c_float),
                                      _p(m, C.c_int32), self.T, mats, len(materials)))

    def update_vertices(self, verts, normals=None):
        v = f32(verts).reshape(-1, 9)
        if v.shape[0] != self.T:
            raise ValueError("update_vertices: triangle count differs from the scene")
        n = None if normals is None else f32(normals).reshape(-1, 9)
        self._ck(lib().lisa_update_vertices(self.h, _p(v, C.c_float), _p(n, C.c_float)))

    def update_vertices_noalloc(self, d_verts, d_normals=None):
        """lisa_update_vertices_noalloc: new positions (and normals) already in device
        memory, given as integer device addresses; no allocation by the library."""
        self._ck(lib().lisa_update_vertices_noalloc(self.h, d_verts, d_normals))

    def render(self, params, want_accum=True, want_rgba8=False):
        npx = params.width * params.height
        acc = np.zeros((npx, 4), np.float32) if want_accum else None
        rgb = np.zeros((npx, 4), np.uint8) if want_rgba8 else None
        self._ck(lib().lisa_render(self.h, C.byref(params), _p(acc, C.c_float),
                                   _p(rgb, C.c_uint8)))
        return acc, rgb

    def render_progressive(self, params, n_subframes):
        """Subframes params.subframe .. + n_subframes - 1 in one graph launch
        (lisa_render_progressive); returns the accum buffer."""
        self._ck(lib().lisa_render_progressive(self.h, C.byref(params), n_subframes))
        return self.copy_accum(params.width * params.height)

    def copy_accum(self, n_pixels):
        acc = np.zeros((n_pixels, 4), np.float32)
        self._ck(lib().lisa_copy_accum(self.h, _p(acc, C.c_float), n_pixels))
        return acc

    def accum64_save(self, n_pixels):
        """The FP64 progressive state (n_pixels x 4: r, g, b sums, samples)."""
        out = np.zeros((n_pixels, 4), np.float64)
        self._ck(lib().lisa_accum64_save(self.h, _p(out, C.c_double), n_pixels))
        return out

    def accum64_load(self, sums):
        a = np.ascontiguousarray(sums, dtype=np.float64).reshape(-1, 4)
        self._ck(lib().lisa_accum64_load(self.h, _p(a, C.c_double), a.shape[0]))

    def synchronize(self):
        self._ck(lib().lisa_synchronize(self.h))

    def set_wide_threshold(self, min_triangles):
        """Scenes set from now on with at least min_triangles triangles get the 4-wide
        tree that the uncounted path kernels walk (lisa_set_wide_threshold)."""
        self._ck(lib().lisa_set_wide_threshold(self.h, int(min_triangles)))

    def stats(self):
        s = Stats()
        self._ck(lib().lisa_get_stats(self.h, C.byref(s)))
        return {"render_ms": s.render_ms, "bvh_build_ms": s.bvh_build_ms,
                "num_tris": s.num_tris, "num_nodes": s.num_nodes,
                "num_emitters": s.num_emitters, "device_bytes": s.device_bytes,
                "wide_tree": bool(s.wide_tree)}

    def counters(self):
        c = np.zeros(NCOUNTERS, np.uint64)
        self._ck(lib().lisa_get_counters(self.h, _p(c, C.c_uint64)))
        return c

    def stream(self):
        return lib().lisa_get_stream(self.h)

    def bvh(self, which=0):
        T = self.T if which == 0 else self.stats()["num_emitters"]
        nI = max(T - 1, 1)
        m = np.zeros(max(T, 1), np.uint32)
        p = np.zeros(max(T, 1), np.uint32)
        r = np.zeros(2 * nI, np.uint32)
        b = np.zeros(12 * nI, np.float32)
        self._ck(lib().lisa_get_bvh(self.h, which, _p(m, C.c_uint32), _p(p, C.c_uint32),
                                    _p(r, C.c_uint32), _p(b, C.c_float)))
        return {"morton": m[:T], "prim": p[:T], "refs": r.reshape(nI, 2),
                "boxes": b.reshape(nI, 12)}

    def trace_rays(self, O, D, tmin=1e-4, tmax=1e16):
        O, D = f32(O).reshape(-1, 3), f32(D).reshape(-1, 3)
        n = O.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        self._ck(lib().lisa_trace_rays(self.h, _p(O, C.c_float), _p(D, C.c_float), n, tmin, tmax,
                                       _p(prim, C.c_int32), _p(t, C.c_float)))
        return prim, t

    def trace_subset(self, params, pixels, record_bounces=None, records=True):
        """The plain path loop on a pixel list; `records=False` returns the means only."""
        px = np.ascontiguousarray(pixels, dtype=np.int32)
        n, spp = px.shape[0], params.spp
        rb = params.bounces if record_bounces is None else record_bounces
        if not records:
            rb = 0
        acc = np.zeros((n, 4), np.float32)
        srgb = np.zeros((n, spp, 3), np.float32) if records else None
        sb = np.zeros((n, spp), np.int32) if records else None
        so = np.zeros((n, spp), np.int32) if records else None
        sh = np.zeros((n, spp), np.uint32) if records else None
        rh = np.zeros((n, spp, max(rb, 1)), np.int32) if rb > 0 else None
        self._ck(lib().lisa_trace_subset(self.h, C.byref(params), _p(px, C.c_int32), n,
                                         _p(acc, C.c_float), _p(srgb, C.c_float),
                                         _p(sb, C.c_int32), _p(so, C.c_int32),
                                         _p(sh, C.c_uint32), _p(rh, C.c_int32), rb))
        return {"accum": acc, "sample_rgb": srgb, "sample_bounces": sb, "sample_occl": so,
                "sample_hash": sh, "rad_hits": rh, "counters": self.counters()}


class Batch:
    """The batched wavefront form (lisa_batch_*): several contexts' scenes in
    one persistent, graph-captured launch."""

    def __init__(self, ctxs, params):
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(ctxs))(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h
                                          for c in ctxs])
        parr = (RenderParams * len(params))(*params)
        self.params = list(params)
        h = C.c_void_p()
        _check(lib().lisa_batch_create(C.byref(h), arr, parr, len(ctxs)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lisa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        """(grid_blocks, chunk, form, graph)."""
        c = BatchConfig()
        _check(lib().lisa_batch_get_config(self.h, C.byref(c)))
        return c.grid_blocks, c.chunk, c.form, c.graph

    def set_config(self, grid_blocks, chunk, form=0, graph=1):
        c = BatchConfig(grid_blocks, chunk, form, graph)
        _check(lib().lisa_batch_set_config(self.h, C.byref(c)))

    def render(self):
        ms = C.c_float()
        _check(lib().lisa_batch_render(self.h, C.byref(ms)))
        return ms.value

    def autotune(self, tune_spp=2):
        """Measured pick of (grid_blocks, chunk, form, graph)."""
        c = BatchConfig()
        _check(lib().lisa_batch_autotune(self.h, tune_spp, C.byref(c)))
        return c.grid_blocks, c.chunk, c.form, c.graph

    def accum(self, i):
        p = self.params[i]
        return self.ctxs[i].copy_accum(p.width * p.height)


CACHE_PARAMS = 23552
CACHE_INPUTS = 12


CACHE_PERSISTENT = 1
CACHE_EMISSION = 2  # prefix bounces' emission in expectation (same mean, less variance)
CACHE_SERIAL = 8  # train before the frame's paths (training kernels timed alone)
CACHE_FULL_PATHS = 16  # no network: whole paths with the prefix estimators (variance-reduced plain)


class CacheFrameParams(C.Structure):
    _fields_ = [("prefix_bounces", C.c_uint32), ("train_paths", C.c_uint32),
                ("train_bounces", C.c_uint32), ("train_max", C.c_uint32),
                ("train_steps", C.c_uint32), ("fallback_loss", C.c_float),
                ("flags", C.c_uint32), ("nee_loops", C.c_uint32), ("nee_loops_deep", C.c_uint32),
                ("judge_pixels", C.c_uint32), ("judge_rmse", C.c_float), ("seed", C.c_uint64)]


class CacheFrameStats(C.Structure):
    _fields_ = [("trace_train_ms", C.c_float), ("train_ms", C.c_float), ("render_ms", C.c_float),
                ("infer_ms", C.c_float), ("total_ms", C.c_float), ("loss", C.c_float),
                ("queries", C.c_int32), ("train_samples", C.c_int32), ("fell_back", C.c_int32),
                ("train_kernel_ms", C.c_float), ("infer_kernel_ms", C.c_float),
                ("judge_rmse_est", C.c_float), ("judge_plain", C.c_int32)]


def cache_frame_params(prefix_bounces=2, train_paths=8192, train_bounces=12, train_max=1 << 16,
                       train_steps=1, fallback_loss=0.0, seed=1, flags=0, nee_loops=1,
                       nee_loops_deep=None, judge_pixels=0, judge_rmse=0.0):
    """nee_loops_deep defaults to nee_loops (the same NEE runs at every prefix vertex)."""
    return CacheFrameParams(prefix_bounces, train_paths, train_bounces, train_max, train_steps,
                            fallback_loss, flags, nee_loops,
                            nee_loops if nee_loops_deep is None else nee_loops_deep, judge_pixels,
                            judge_rmse, seed)


def render_cached(ctx, cache, params, cparams):
    """One frame of the cached renderer (lisa_render_cached); returns its stats as a dict."""
    s = CacheFrameStats()
    st = lib().lisa_render_cached(ctx.h, cache.h, C.byref(params), C.byref(cparams), C.byref(s))
    if st != OK:
        msg = lib().lisa_last_error(ctx.h)
        raise LisaError(st, msg.decode() if msg else "")
    return {k: getattr(s, k) for k, _ in CacheFrameStats._fields_}


def cache_trace_queries(ctx, params, cparams, pixels, cap=None):
    """Prefix paths on a pixel list (lisa_cache_trace_queries) with the prefix settings of
    cparams (prefix_bounces, CACHE_EMISSION, nee_loops, nee_loops_deep): accum, features,
    throughput, item
    (at most `cap` queries copied back, default all; 0: the per-pixel accum only)."""
    px = np.ascontiguousarray(pixels, dtype=np.int32)
    n = px.shape[0]
    cap = n * params.spp if cap is None else cap
    acc = np.zeros((n, 4), np.float32)
    f = np.zeros((cap, CACHE_INPUTS), np.float32) if cap > 0 else None
    at = np.zeros((cap, 3), np.float32) if cap > 0 else None
    it = np.zeros(cap, np.int32) if cap > 0 else None
    nq = C.c_int32()
    ctx._ck(lib().lisa_cache_trace_queries(ctx.h, C.byref(params), C.byref(cparams), _p(px, C.c_int32), n,
                                           _p(acc, C.c_float), _p(f, C.c_float), _p(at, C.c_float),
                                           _p(it, C.c_int32), cap, C.byref(nq)))
    k = min(nq.value, cap)
    if cap == 0:
        return {"accum": acc, "feat": None, "att": None, "item": None, "n": nq.value}
    return {"accum": acc, "feat": f[:k], "att": at[:k], "item": it[:k], "n": nq.value}


def cache_trace_rows(ctx, params, pixels, cap=None, depth=0):
    """Training rows of plain paths on a pixel list (lisa_cache_trace_rows): features, targets
    of every surface hit (depth 0) or of radiance ray number `depth`'s hits."""
    px = np.ascontiguousarray(pixels, dtype=np.int32)
    n = px.shape[0]
    cap = n * params.spp * params.bounces if cap is None else cap
    f = np.zeros((cap, CACHE_INPUTS), np.float32)
    t = np.zeros((cap, 3), np.float32)
    nr = C.c_int32()
    ctx._ck(lib().lisa_cache_trace_rows(ctx.h, C.byref(params), _p(px, C.c_int32), n, depth,
                                        _p(f, C.c_float),
                                        _p(t, C.c_float), cap, C.byref(nr)))
    k = min(nr.value, cap)
    return {"feat": f[:k], "tgt": t[:k], "n": nr.value}


class Cache:
    """The learned radiance cache (lisa_cache_*): MLP weights + Adam state on a device."""

    def __init__(self, device=0, seed=0, lr=1e-3):
        h = C.c_void_p()
        _check(lib().lisa_cache_create(C.byref(h), device, seed, lr))
        self.h = h

    def _ck(self, status):
        if status != OK:
            msg = lib().lisa_cache_last_error(self.h)
            raise LisaError(status, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            lib().lisa_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self):
        w = np.zeros(CACHE_PARAMS, np.float32)
        self._ck(lib().lisa_cache_get_weights(self.h, _p(w, C.c_float)))
        return w

    def set_weights(self, w):
        w = f32(w)
        assert w.size == CACHE_PARAMS
        self._ck(lib().lisa_cache_set_weights(self.h, _p(w, C.c_float)))

    def infer(self, feat):
        x = f32(feat).reshape(-1, CACHE_INPUTS)
        out = np.zeros((x.shape[0], 3), np.float32)
        self._ck(lib().lisa_cache_infer(self.h, _p(x, C.c_float), x.shape[0], _p(out, C.c_float)))
        return out

    def gradients(self, feat, target):
        x = f32(feat).reshape(-1, CACHE_INPUTS)
        t = f32(target).reshape(-1, 3)
        g = np.zeros(CACHE_PARAMS, np.float32)
        loss = C.c_float()
        self._ck(lib().lisa_cache_gradients(self.h, _p(x, C.c_float), _p(t, C.c_float),
                                            x.shape[0], _p(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def train(self, feat, target):
        x = f32(feat).reshape(-1, CACHE_INPUTS)
        t = f32(target).reshape(-1, 3)
        loss = C.c_float()
        self._ck(lib().lisa_cache_train(self.h, _p(x, C.c_float), _p(t, C.c_float), x.shape[0],
                                        C.byref(loss)))
        return loss.value

    def last_ms(self):
        return lib().lisa_cache_last_ms(self.h)
