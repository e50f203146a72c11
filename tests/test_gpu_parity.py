"""GPU parity: liblisa.so (HIP, gfx950) against the CPU oracle.

Bar (bit-exact unless stated): Morton order, BVH topology and boxes;
closest-hit triangle ids and t; per-sample bounce counts, occlusion-ray
counts, per-sample hash of every ray's hit id, radiance hit ids; per-sample
radiance within 1e-5 relative (FP32); accum buffers within 1e-5 relative.
Everything goes through the C ABI (lisa_amd.capi).
"""
import numpy as np
import pytest

from lisa_amd import scenes
from tests.scene_util import camera, cornell, random_rays

pytestmark = pytest.mark.gpu

RTOL = 1e-5


@pytest.fixture(scope="module")
def ctx(capi):
    c = capi.Context(0)
    yield c
    c.close()


def _pair(capi, oracle_mod, ctx, sc):
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    return o


def _close(a, b, rtol=RTOL):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-30) + 1e-30)


@pytest.mark.parametrize("seg,rings", [(32, 16), (256, 128)])
def test_bvh_bit_exact(capi, oracle_mod, ctx, seg, rings):
    sc = cornell(seg, rings, 64, 64, 1, 1)
    o = _pair(capi, oracle_mod, ctx, sc)
    g = ctx.bvh()
    r = o.bvh()
    print("T", sc["verts"].shape[0], flush=True)
    assert np.array_equal(g["morton"], r["morton"])
    assert np.array_equal(g["prim"], r["prim"])
    assert np.array_equal(g["refs"], r["refs"])
    assert np.array_equal(g["boxes"].view(np.uint32), r["boxes"].view(np.uint32))


@pytest.mark.parametrize("seg,rings", [(32, 16), (256, 128)])
def test_hit_ids_bit_exact(capi, oracle_mod, ctx, seg, rings):
    sc = cornell(seg, rings, 64, 64, 1, 1)
    o = _pair(capi, oracle_mod, ctx, sc)
    O, D = random_rays(1 << 16, seed=seg)
    gp, gt = ctx.trace_rays(O, D)
    rp, rt, _ = o.closest_hit(O, D, threads=8)
    assert (rp >= 0).mean() > 0.8
    assert np.array_equal(gp, rp)
    assert np.array_equal(gt.view(np.uint32), rt.view(np.uint32))


def _subset_compare(g, r, rad_exact=True):
    assert np.array_equal(g["sample_bounces"], r["sample_bounces"])
    assert np.array_equal(g["sample_occl"], r["sample_occl"])
    assert np.array_equal(g["sample_hash"], r["sample_hash"])
    assert np.array_equal(g["rad_hits"], r["rad_hits"])
    assert _close(g["sample_rgb"], r["sample_rgb"])
    assert _close(g["accum"], r["accum"])
    for k in range(8):  # ray / node / triangle counters
        assert int(g["counters"][k]) == int(r["counters"][k]), k


def test_path_subset_small_cornell(capi, oracle_mod, ctx):
    sc = cornell(32, 16, 256, 256, 16, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    px = scenes.seeded_pixels(256, 256, 1024, seed=1)
    p = capi.render_params(256, 256, 16, 7, sc["camera"]["eye"], U, V, W)
    g = ctx.trace_subset(p, px)
    r = o.render(256, 256, 16, 7, sc["camera"]["eye"], U, V, W, pixels=px, record=True, threads=8)
    _subset_compare(g, r)


def test_path_subset_65k_cornell(capi, oracle_mod, ctx):
    sc = cornell(256, 128, 1024, 1024, 256, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    px = scenes.seeded_pixels(1024, 1024, 4096, seed=2)
    p = capi.render_params(1024, 1024, 4, 7, sc["camera"]["eye"], U, V, W)
    g = ctx.trace_subset(p, px)
    r = o.render(1024, 1024, 4, 7, sc["camera"]["eye"], U, V, W, pixels=px, record=True, threads=8)
    _subset_compare(g, r)


def test_full_render_small_cornell(capi, oracle_mod, ctx):
    sc = cornell(32, 16, 128, 128, 8, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    p = capi.render_params(128, 128, 8, 7, sc["camera"]["eye"], U, V, W, flags=capi.FLAG_COUNT)
    acc, rgba8 = ctx.render(p, want_accum=True, want_rgba8=True)
    cnt = ctx.counters()
    r = o.render(128, 128, 8, 7, sc["camera"]["eye"], U, V, W, threads=8)
    assert _close(acc, r["accum"])
    assert np.array_equal(rgba8, oracle_mod.srgb_quantize(r["accum"]))
    assert np.array_equal(cnt, r["counters"])
    # the uncounted product kernel gives the same image
    p.flags = 0
    acc2, _ = ctx.render(p)
    assert np.array_equal(acc2.view(np.uint32), acc.view(np.uint32))
