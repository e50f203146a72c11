"""GPU: the 4-wide form of the tree (lisa_set_wide_threshold).

A context whose threshold is 0 gives every scene the wide tree, so the small
scenes of the oracle tests run through the wide walk (closest_wide) that
large scenes take by default.  Bar: hit ids and t bit-exact against the
oracle; every image, FP64 sum, cache query, training row and trained cached
frame bit-identical to a context that walks the binary tree; the node-visit
counters stay the binary walk's.
"""
import numpy as np
import pytest

from lisa_amd import scenes
from tests.scene_util import DeviceArray, camera, cornell, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(capi):
    wide, plain = capi.Context(0), capi.Context(0)
    wide.set_wide_threshold(0)
    plain.set_wide_threshold(1 << 62)
    yield wide, plain
    wide.close()
    plain.close()


def _set(pair, sc):
    for c in pair:
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    assert pair[0].stats()["wide_tree"] and not pair[1].stats()["wide_tree"]


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_threshold_picks_the_tree(capi):
    sc = cornell(256, 128, 16, 16, 1, 1)  # 65,036 triangles
    with capi.Context(0) as c:
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        assert not c.stats()["wide_tree"]  # default: 2^20 triangles
        b = c.stats()["device_bytes"]
        c.set_wide_threshold(sc["verts"].shape[0])
        assert not c.stats()["wide_tree"]  # from the next scene on
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        assert c.stats()["wide_tree"]
        assert c.stats()["device_bytes"] == b + 128 * c.stats()["num_nodes"]
        c.set_wide_threshold(sc["verts"].shape[0] + 1)
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        assert not c.stats()["wide_tree"]
        with pytest.raises(capi.LisaError, match="min_triangles"):
            c.set_wide_threshold(-1)


@pytest.mark.parametrize("seg,rings", [(32, 16), (256, 128)])
def test_wide_hit_ids_match_oracle(capi, oracle_mod, pair, seg, rings):
    wide, plain = pair
    sc = cornell(seg, rings, 64, 64, 1, 1)
    _set(pair, sc)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    for inside in (True, False):
        O, D = random_rays(1 << 16, seed=seg + 7, inside=inside)
        gp, gt = wide.trace_rays(O, D)
        rp, rt, _ = o.closest_hit(O, D, threads=8)
        assert (rp >= 0).mean() > (0.8 if inside else 0.02)
        assert np.array_equal(gp, rp)
        assert _same(gt, rt)
        # the counters of lisa_trace_rays stay the binary walk's
        cw = wide.counters()
        plain.trace_rays(O, D)
        assert np.array_equal(cw, plain.counters())


def test_wide_edges_and_axis_rays(capi, oracle_mod, pair):
    """Rays through the shared edges and the centre vertex of a triangle fan
    and axis-parallel rays (equal-t hits on several leaves of one wide node)."""
    wide, _ = pair
    k = 8
    ang = np.linspace(0, 2 * np.pi, k + 1)
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1).astype(np.float32)
    v = np.concatenate([np.zeros((k, 3), np.float32), rim[:-1], rim[1:]], axis=1)
    n = np.tile(np.float32([0, 0, 1]), (k, 3))
    mats = [{"diffuse": (1, 1, 1), "alpha": 1.0, "roughness": 1.0}]
    wide.set_scene(v, n, np.zeros(k, np.int32), mats)
    assert wide.stats()["wide_tree"]
    o = oracle_mod.Scene(v, n, np.zeros(k, np.int32), mats)
    pts = [np.float32([0, 0, 0])] + [(rim[i] * np.float32(s)).astype(np.float32)
                                     for i in range(k) for s in np.linspace(0.05, 0.95, 19)]
    P = np.array(pts, np.float32)
    for d in (np.float32([0, 0, -1]), np.float32([0.3, -0.2, -1]), np.float32([1e-3, 0, -1])):
        D = np.tile(d, (len(P), 1))
        gp, gt = wide.trace_rays(P - d, D)
        rp, rt, _ = o.closest_hit(P - d, D, brute=True)
        assert np.all(gp >= 0)
        assert np.array_equal(gp, rp) and _same(gt, rt)


def _scenes():
    yield "cornell", cornell(32, 16, 72, 40, 4, 7)
    yield "cornell65k", cornell(256, 128, 48, 32, 2, 7)
    yield "icosphere", scenes.displaced_icosphere(level=4, width=48, height=32, spp=2, bounces=12)
    yield "island", scenes.island(n=64, width=48, height=32, spp=2, bounces=12)


@pytest.mark.parametrize("name", ["cornell", "cornell65k", "icosphere", "island"])
def test_wide_renders_same_bits(capi, pair, name):
    """lisa_render through the wide walk (tiled, persistent, FP64 sums,
    progressive subframes and the graph-replayed run) against the binary
    walk's images."""
    sc = dict(_scenes())[name]
    wide, plain = pair
    _set(pair, sc)
    W, H = sc["width"], sc["height"]
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    for flags in (0, capi.FLAG_PERSISTENT, capi.FLAG_FP64_ACCUM):
        for f in range(2):
            p = capi.render_params(W, H, sc["spp"], sc["bounces"], eye, U, V, Wv, subframe=f, flags=flags)
            a, _ = wide.render(p)
            b, _ = plain.render(p)
            assert _same(a, b), (name, flags, f)
    for f in range(3):
        last, _ = plain.render(capi.render_params(W, H, sc["spp"], sc["bounces"], eye, U, V, Wv, subframe=f))
    got = wide.render_progressive(capi.render_params(W, H, sc["spp"], sc["bounces"], eye, U, V, Wv), 3)
    assert _same(got, last)


@pytest.mark.parametrize("emit", [False, True])
def test_wide_cache_queries_and_rows_match_oracle(capi, oracle_mod, pair, emit):
    """The cache's path modes through the wide walk: queries in the oracle's
    order and bit-exact, training rows bit-identical to the binary walk's."""
    wide, plain = pair
    W = H = 64
    sc = cornell(32, 16, W, H, 4, 7)
    _set(pair, sc)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    px = np.sort(np.random.default_rng(3).choice(W * H, 1200, replace=False)).astype(np.int32)
    p = capi.render_params(W, H, 2, 3, eye, U, V, Wv, subframe=4)
    cp = capi.cache_frame_params(prefix_bounces=2, flags=capi.CACHE_EMISSION if emit else 0,
                                 nee_loops=3 if emit else 1, nee_loops_deep=2 if emit else 1)
    g = capi.cache_trace_queries(wide, p, cp, px)
    r = o.cache_queries(W, H, 2, 2, eye, U, V, Wv, px, subframe=4, threads=8, emit_expect=emit,
                        nee_loops=3 if emit else 1, nee_loops_deep=2 if emit else 1)
    assert g["n"] == r["n"] > 0
    assert np.array_equal(g["item"], r["item"])
    assert np.array_equal(g["att"], r["att"])
    assert np.array_equal(g["feat"][:, [0, 1, 2, 7, 8, 9, 10, 11]], r["feat"][:, [0, 1, 2, 7, 8, 9, 10, 11]])
    h = capi.cache_trace_queries(plain, p, cp, px)
    assert _same(g["feat"], h["feat"]) and _same(g["accum"], h["accum"])
    pr = capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=5)
    a, b = capi.cache_trace_rows(wide, pr, px), capi.cache_trace_rows(plain, pr, px)
    assert a["n"] == b["n"] > 0
    assert np.array_equal(a["feat"], b["feat"]) and np.array_equal(a["tgt"], b["tgt"], equal_nan=True)


def test_wide_trained_cached_frames_same_bits(capi, pair):
    """Trained cached frames (training paths, prefix paths with the light
    samples, fallback frames) through the wide walk: the binary walk's
    images and weights."""
    wide, plain = pair
    W, H = 64, 48
    sc = cornell(32, 16, W, H, 2, 7)
    _set(pair, sc)
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    a, b = capi.Cache(0, seed=1), capi.Cache(0, seed=1)
    try:
        kw = dict(prefix_bounces=2, train_paths=4000, train_bounces=7, train_steps=2, seed=5,
                  flags=capi.CACHE_EMISSION, nee_loops=4)
        for f in range(3):
            p = capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=f)
            cp = capi.cache_frame_params(fallback_loss=1e-9 if f == 2 else 0.0, **kw)
            s1, s2 = capi.render_cached(wide, a, p, cp), capi.render_cached(plain, b, p, cp)
            assert s1["queries"] == s2["queries"] and s1["train_samples"] == s2["train_samples"] > 0
            assert s1["fell_back"] == s2["fell_back"] == (1 if f == 2 else 0)
            assert _same(wide.copy_accum(W * H), plain.copy_accum(W * H)), f
        assert _same(a.weights(), b.weights())
    finally:
        a.close()
        b.close()


def test_wide_tree_follows_refits(capi, pair):
    """lisa_update_vertices and lisa_update_vertices_noalloc rebuild the wide
    form from the refitted tree: renders after either equal the binary
    walk's after the same refit."""
    wide, plain = pair
    sc = cornell(64, 32, 48, 40, 2, 7)
    _set(pair, sc)
    U, V, Wv = camera(sc, capi)
    p = capi.render_params(48, 40, 2, 7, sc["camera"]["eye"], U, V, Wv)
    rs = np.random.default_rng(4)
    for noalloc in (False, True):
        v2 = (sc["verts"] + rs.normal(scale=0.02, size=sc["verts"].shape)).astype(np.float32)
        plain.update_vertices(v2)
        if noalloc:
            dv = DeviceArray(v2)
            try:
                st = capi.lib().lisa_update_vertices_noalloc(wide.h, dv.ptr, None)
                assert st == capi.OK, capi.lib().lisa_last_error(wide.h)
            finally:
                dv.free()
        else:
            wide.update_vertices(v2)
        a, _ = wide.render(p)
        b, _ = plain.render(p)
        assert _same(a, b), noalloc


def test_wide_edge_case_scenes(capi, pair):
    """One and two triangles and a scene with NaN vertices through the wide
    walk: the binary walk's images."""
    wide, plain = pair
    sc = cornell(32, 16, 32, 32, 4, 5)
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    mats = sc["materials"]
    cases = [sc["verts"][:1], sc["verts"][:2]]
    nan = np.array(sc["verts"], np.float32).reshape(-1, 3, 3).copy()
    nan[::7, 0] = np.nan
    cases.append(nan.reshape(-1, 9))
    for v in cases:
        T = v.shape[0]
        for c in pair:
            c.set_scene(v, sc["normals"][:T], sc["mat_idx"][:T], mats)
        assert wide.stats()["wide_tree"]
        p = capi.render_params(32, 32, 4, 5, eye, U, V, Wv)
        a, _ = wide.render(p)
        b, _ = plain.render(p)
        assert _same(a, b), T
