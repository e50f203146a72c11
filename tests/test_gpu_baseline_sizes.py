"""GPU parity at the full sizes of the benchmark workloads.

The other GPU tests use small scenes; these repeat the bit-exact checks on
the scenes the benchmarks time, at their full triangle counts:

* the displaced dielectric icosphere, subdivision 9 (5,242,880 triangles,
  1920x1080, 12 bounces): LBVH (Morton codes, order, topology, boxes) of
  both trees, closest hits of 65,536 seeded rays, and a 256-pixel x 4-spp
  path subset (per-sample records, counters);
* the 2048^2 fBm island (8,380,418 triangles): the same, plus the cached
  renderer's prefix paths with the benchmark's settings (prefix 2, the
  variance-reduced prefix with 8 NEE runs at the camera ray's hit, 1 later);
* for both (scenes of more than 2^20 triangles, whose renders walk the 4-wide
  form of the tree): a full-image render bit-identical to a context that
  walks the binary tree;
* four scenes of the 64-scene batch (heightfields, spheres and boxes, 15k to
  1.1M triangles, random Lambertian / dielectric n in [1.3, 1.7] / emissive
  materials, 256^2..768^2, 32 spp, 8 bounces) rendered together in one
  batched launch, each image checked against the oracle on seeded pixels;
* the 65k-triangle Cornell box at its full 256 spp on 128 pixels (the
  per-pixel RNG stream runs through all 256 samples).

Bars as in test_gpu_parity: bit-exact integers and hit distances, radiance
within 1e-5 relative.
"""
import numpy as np
import pytest

from lisa_amd import scenes
from tests.scene_util import camera

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _close(a, b, rtol=RTOL):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-30) + 1e-30)


def _subset_compare(g, r):
    assert np.array_equal(g["sample_bounces"], r["sample_bounces"])
    assert np.array_equal(g["sample_occl"], r["sample_occl"])
    assert np.array_equal(g["sample_hash"], r["sample_hash"])
    assert np.array_equal(g["rad_hits"], r["rad_hits"])
    assert _close(g["sample_rgb"], r["sample_rgb"])
    assert _close(g["accum"], r["accum"])
    for k in range(8):
        assert int(g["counters"][k]) == int(r["counters"][k]), k


def _tree_depth(refs):
    """Depth of the binary tree given by child refs (bit 31 = leaf)."""
    refs = refs.reshape(-1, 2)
    depth, frontier = 0, np.zeros(1, np.int64)
    while frontier.size:
        depth += 1
        ch = refs[frontier].reshape(-1)
        frontier = (ch[(ch & 0x80000000) == 0]).astype(np.int64)
    return depth


def _bvh_and_rays(ctx, o, sc, seed):
    for which in (0, 1):
        g, r = ctx.bvh(which), o.bvh(which)
        for key in ("morton", "prim", "refs"):
            assert np.array_equal(g[key], r[key]), (which, key)
        assert np.array_equal(g["boxes"].view(np.uint32), r["boxes"].view(np.uint32)), which
        if which == 0:
            depth = _tree_depth(g["refs"])
            print(f"tree depth {depth}", flush=True)
            assert depth > 24
    v = np.asarray(sc["verts"], np.float32).reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    rs = np.random.default_rng(seed)
    n = 1 << 16
    O = (lo + (hi - lo) * rs.uniform(0.05, 0.95, (n, 3))).astype(np.float32)
    D = rs.normal(size=(n, 3))
    D = (D / np.linalg.norm(D, axis=1, keepdims=True)).astype(np.float32)
    gp, gt = ctx.trace_rays(O, D)
    rp, rt, _ = o.closest_hit(O, D, threads=16)
    print(f"hits {(gp >= 0).mean():.3f}", flush=True)
    assert (rp >= 0).mean() > 0.1  # open scenes: many rays leave the box
    assert np.array_equal(gp, rp)
    assert np.array_equal(gt.view(np.uint32), rt.view(np.uint32))


def _path_subset(capi, ctx, o, sc, npx, spp, bounces, seed):
    W, H = sc["width"], sc["height"]
    cam = sc["camera"]
    U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
    px = scenes.seeded_pixels(W, H, npx, seed=seed)
    p = capi.render_params(W, H, spp, bounces, cam["eye"], U, V, Wv)
    g = ctx.trace_subset(p, px)
    r = o.render(W, H, spp, bounces, cam["eye"], U, V, Wv, pixels=px, record=True, threads=16)
    print(f"rays {int(r['counters'][0] + r['counters'][1])}", flush=True)
    _subset_compare(g, r)
    return U, V, Wv


def _wide_equals_binary(capi, ctx, sc, U, V, Wv):
    """A full-image render through the 4-wide walk equals the binary walk's."""
    assert ctx.stats()["wide_tree"]
    W, H = sc["width"], sc["height"]
    p = capi.render_params(W, H, 2, sc["bounces"], sc["camera"]["eye"], U, V, Wv)
    with capi.Context(0) as plain:
        plain.set_wide_threshold(1 << 62)
        plain.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        assert not plain.stats()["wide_tree"]
        a, _ = ctx.render(p)
        b, _ = plain.render(p)
    print("full image: wide walk == binary walk", flush=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_icosphere_level9_full_size(capi, oracle_mod):
    sc = scenes.displaced_icosphere(level=9)
    assert sc["verts"].shape[0] > 5_242_880
    print("scene built", flush=True)
    with capi.Context(0) as ctx:
        ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        _bvh_and_rays(ctx, o, sc, seed=9)
        U, V, Wv = _path_subset(capi, ctx, o, sc, 256, 4, 12, seed=9)
        _wide_equals_binary(capi, ctx, sc, U, V, Wv)


def test_island_2048_full_size(capi, oracle_mod):
    sc = scenes.island()
    assert sc["verts"].shape[0] > 8_380_000
    print("scene built", flush=True)
    with capi.Context(0) as ctx:
        ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        _bvh_and_rays(ctx, o, sc, seed=5)
        U, V, Wv = _path_subset(capi, ctx, o, sc, 256, 4, 12, seed=5)
        _wide_equals_binary(capi, ctx, sc, U, V, Wv)
        # the cached renderer's prefix paths with the benchmark's settings
        W, H = sc["width"], sc["height"]
        eye = sc["camera"]["eye"]
        px = scenes.seeded_pixels(W, H, 2048, seed=6)
        p = capi.render_params(W, H, 4, 3, eye, U, V, Wv, subframe=3)
        cp = capi.cache_frame_params(prefix_bounces=2, flags=capi.CACHE_EMISSION, nee_loops=8,
                                     nee_loops_deep=1)
        g = capi.cache_trace_queries(ctx, p, cp, px)
        r = o.cache_queries(W, H, 4, 2, eye, U, V, Wv, px, subframe=3, threads=16,
                            emit_expect=True, nee_loops=8, nee_loops_deep=1)
        print(f"queries {g['n']}", flush=True)
        assert g["n"] == r["n"] > 1000
        assert np.array_equal(g["accum"].view(np.uint32), r["accum"].view(np.uint32))
        # in the oracle's order (listed pixel, then sample)
        assert np.array_equal(g["item"], r["item"])
        assert np.array_equal(g["att"], r["att"])
        exact = [0, 1, 2, 7, 8, 9, 10, 11]
        assert np.array_equal(g["feat"][:, exact], r["feat"][:, exact])
        assert np.max(np.abs(g["feat"][:, 3:7] - r["feat"][:, 3:7])) <= 2e-6


BATCH = (2, 3, 8, 9)


def test_batch_scenes_against_oracle(capi, oracle_mod):
    """Four of the 64 batch scenes in one batched launch (the lane and wave
    forms, the ray-level wavefront form, and the autotuner's pick): every
    image equals the oracle's on 512 seeded pixels."""
    ctxs, params, scs = [], [], []
    try:
        for i in BATCH:
            sc = scenes.batch_scene(i)
            c = capi.Context(0)
            c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
            U, V, Wv = camera(sc, capi)
            params.append(capi.render_params(sc["width"], sc["height"], sc["spp"], sc["bounces"],
                                             sc["camera"]["eye"], U, V, Wv))
            ctxs.append(c)
            scs.append(sc)
        refs = []
        for sc, p in zip(scs, params):
            o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
            px = scenes.seeded_pixels(p.width, p.height, 512, seed=11)
            r = o.render(p.width, p.height, p.spp, p.bounces, p.eye, p.U, p.V, p.W, pixels=px,
                         threads=16)
            refs.append((px, r["accum"]))
            print(f"{sc['name']}: {sc['verts'].shape[0]} triangles, {p.width}^2", flush=True)
        batch = capi.Batch(ctxs, params)
        try:
            for cfg in ((4096, 1, 0), (4096, 1, 1), (8192, 2, 2), None):
                if cfg is None:
                    batch.autotune(tune_spp=2)
                else:
                    batch.set_config(*cfg)
                batch.render()
                for k, (px, ref) in enumerate(refs):
                    got = batch.accum(k)[px]
                    assert _close(got, ref), (cfg, k)
        finally:
            batch.close()
    finally:
        for c in ctxs:
            c.close()


def test_cornell65k_full_spp(capi, oracle_mod):
    """The benchmark's own configuration, 256 spp x 7 bounces, on 128
    seeded pixels: the per-pixel LCG stream runs through all 256 samples."""
    sc = scenes.cornell_65k()
    with capi.Context(0) as ctx:
        ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        _path_subset(capi, ctx, o, sc, 128, 256, 7, seed=13)
