"""Pins of the oracle's radiance-cache path forms (orc_cache_queries,
orc_cache_train_rows) through the plain path loop's own per-sample records,
so that a wrong query vertex, a wrong target or a wrong position feature
fails without retyping the feature formula:

* the query of a sample is at the surface hit of radiance ray P+1: a sample
  queries iff the plain loop with P+1 bounces records a non-emitting
  triangle at ray P+1 (one sample per pixel, where the random streams
  agree), and the queried position lies on that triangle;
* the radiance gathered before the query equals the plain radiance of the
  samples that did not query;
* a training path's first row has throughput 1 and nothing gathered before
  it, so its target is the plain sample radiance, bit for bit; the number of
  rows of a path is its number of non-emitting surface hits.
"""
from __future__ import annotations

import numpy as np

from tests.scene_util import camera, cornell

W = H = 48


def _setup(oracle_mod):
    sc = cornell(32, 16, W, H, 4, 7)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    U, V, Wv = camera(sc, oracle_mod)
    emissive = np.array([bool(m.get("emit", False)) for m in sc["materials"]])
    tri_emit = emissive[np.asarray(sc["mat_idx"])]
    return sc, o, (U, V, Wv), tri_emit


def _pixels(n, seed=3):
    return np.sort(np.random.default_rng(seed).choice(W * H, n, replace=False)).astype(np.int32)


def test_bounds_are_the_root_boxes(oracle_mod):
    sc, o, _, _ = _setup(oracle_mod)
    bmin, binv = o.cache_bounds()
    v = np.asarray(sc["verts"], np.float32).reshape(-1, 3)
    # every vertex inside, and the box tight to one ulp-widened leaf box
    assert np.all(v >= bmin)
    bmax = bmin + 1.0 / binv
    assert np.allclose(bmin, v.min(0), atol=1e-6)
    assert np.allclose(bmax, v.max(0), atol=1e-5)


def test_queries_follow_plain_records(oracle_mod):
    sc, o, (U, V, Wv), tri_emit = _setup(oracle_mod)
    eye = sc["camera"]["eye"]
    # one sample per pixel: the plain loop goes on drawing random numbers at
    # the query vertex (NEE, bounce), so only the first sample of a pixel
    # corresponds sample for sample
    px = _pixels(900)
    spp, P = 1, 2
    q = o.cache_queries(W, H, spp, P, eye, U, V, Wv, px, threads=4)
    r = o.render(W, H, spp, P + 1, eye, U, V, Wv, pixels=px, record=True, threads=4)
    hits = r["rad_hits"][:, :, P]  # hit id of radiance ray P+1 (-3: not traced)
    want = (hits >= 0) & ~tri_emit[np.maximum(hits, 0)]
    assert q["n"] == int(want.sum()) and q["n"] > 300
    # query items: the pixel list index of each querying sample, (pixel, sample) order
    assert np.array_equal(q["item"], np.nonzero(want)[0].astype(np.int32))
    # the queried position lies on the triangle that ray P+1 hit
    bmin, binv = o.cache_bounds()
    pos = bmin + q["feat"][:, :3].astype(np.float64) / binv
    tri = np.asarray(sc["verts"], np.float64).reshape(-1, 3, 3)[hits[want]]
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    dist = np.abs(np.einsum("ij,ij->i", pos - tri[:, 0], nrm))
    assert dist.max() < 1e-5
    # albedo / roughness / dielectric features are the hit triangle's material
    mats = sc["materials"]
    m_of = np.asarray(sc["mat_idx"])[hits[want]]
    alb = np.array([mats[k]["diffuse"] for k in m_of], np.float32)
    assert np.array_equal(q["feat"][:, 7:10], np.clip(alb, 0, 1))
    diel = np.array([mats[k].get("alpha", 1.0) < 1.0 for k in m_of])
    assert np.array_equal(q["feat"][:, 11], diel.astype(np.float32))
    assert np.all((q["feat"] >= 0) & (q["feat"] <= 1))
    # pixels whose samples never query have the plain mean of P+1 bounces
    none = ~want.any(axis=1)
    assert none.sum() > 0
    assert np.array_equal(q["accum"][none], r["accum"][none])
    # first-vertex throughput is 1; all throughputs in [0, 1]
    assert np.all((q["att"] >= 0) & (q["att"] <= 1))


def _unit(th01, ph01):
    th = th01.astype(np.float64) * np.pi
    ph = ph01.astype(np.float64) * 2 * np.pi - np.pi
    return np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], 1)


def test_direction_and_normal_features(oracle_mod):
    """Along a training path, the direction feature of vertex k+1 points from
    vertex k to it (vertex 0: from the eye); the normal feature is the hit
    triangle's normal up to the smooth-shading difference."""
    sc, o, (U, V, Wv), _ = _setup(oracle_mod)
    eye = np.asarray(sc["camera"]["eye"], np.float64)
    px = _pixels(200, seed=5)
    t = o.cache_train_rows(W, H, 1, 7, sc["camera"]["eye"], U, V, Wv, px, threads=4)
    bmin, binv = o.cache_bounds()
    pos = bmin + t["feat"][:, :3].astype(np.float64) / binv
    d = _unit(t["feat"][:, 3], t["feat"][:, 4])
    prev = np.empty_like(pos)
    first = np.r_[True, t["path"][1:] != t["path"][:-1]]
    prev[first] = eye
    prev[~first] = pos[np.nonzero(~first)[0] - 1]
    seg = pos - prev
    ln = np.linalg.norm(seg, axis=1)
    ok = ln > 0.05
    assert ok.sum() > 200
    cosang = np.einsum("ij,ij->i", d[ok], seg[ok] / ln[ok, None])
    assert cosang.min() > 1 - 1e-4
    n = _unit(t["feat"][:, 5], t["feat"][:, 6])
    v = np.asarray(sc["verts"], np.float64).reshape(-1, 3, 3)
    vn = np.asarray(sc["normals"], np.float64).reshape(-1, 3, 3)
    # the nearest triangle plane through the position: pick by smallest plane distance
    # among triangles whose vertex normals all agree with n to 0.3 rad
    cos_best = []
    for i in np.nonzero(ok)[0][:150]:
        g = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        g /= np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30)
        dist = np.abs(np.einsum("ij,ij->i", pos[i] - v[:, 0], g))
        k = np.argmin(dist + 10 * (np.linalg.norm(v.mean(1) - pos[i], axis=1) > 0.2))
        cos_best.append(abs(np.dot(vn[k].mean(0) / np.linalg.norm(vn[k].mean(0)), n[i])))
    assert np.median(cos_best) > 0.99


def test_training_rows_follow_plain_samples(oracle_mod):
    sc, o, (U, V, Wv), tri_emit = _setup(oracle_mod)
    eye = sc["camera"]["eye"]
    px = _pixels(256, seed=11)
    B = 7
    t = o.cache_train_rows(W, H, 1, B, eye, U, V, Wv, px, subframe=9, threads=4)
    r = o.render(W, H, 1, B, eye, U, V, Wv, pixels=px, subframe=9, record=True, threads=4)
    hits = r["rad_hits"][:, 0, :]
    surf = (hits >= 0) & ~tri_emit[np.maximum(hits, 0)]
    per_path = np.bincount(t["path"], minlength=len(px))
    assert np.array_equal(per_path, surf.sum(1))
    first = np.r_[0, np.cumsum(per_path)[:-1]][per_path > 0]
    # first row: target = sample radiance (throughput 1, nothing gathered before)
    assert np.array_equal(t["tgt"][first], r["sample_rgb"][per_path > 0, 0, :])
    # every finite target is non-negative
    fin = np.isfinite(t["tgt"])
    assert np.all(t["tgt"][fin] >= 0)
    # the position feature of row k lies on the k-th surface hit of its path
    bmin, binv = o.cache_bounds()
    pos = bmin + t["feat"][:, :3].astype(np.float64) / binv
    ids = hits[surf]  # row order: path-major, vertex order
    tri = np.asarray(sc["verts"], np.float64).reshape(-1, 3, 3)[ids]
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    assert np.abs(np.einsum("ij,ij->i", pos - tri[:, 0], nrm)).max() < 1e-5


def test_depth_rows_are_the_kth_vertex(oracle_mod):
    """Rows at depth d are, path by path, the d-th row of the every-vertex
    rows (each radiance ray before the end of a path ends at a recorded
    surface hit)."""
    sc, o, (U, V, Wv), _ = _setup(oracle_mod)
    eye = sc["camera"]["eye"]
    px = _pixels(256, seed=13)
    allr = o.cache_train_rows(W, H, 1, 7, eye, U, V, Wv, px, subframe=2, threads=4)
    for d in (1, 3):
        t = o.cache_train_rows(W, H, 1, 7, eye, U, V, Wv, px, subframe=2, depth=d, threads=4)
        starts = np.r_[0, np.nonzero(allr["path"][1:] != allr["path"][:-1])[0] + 1]
        counts = np.diff(np.r_[starts, allr["n"]])
        pick = starts[counts >= d] + d - 1
        assert t["n"] == pick.size > 0
        assert np.array_equal(t["path"], allr["path"][pick])
        assert np.array_equal(t["feat"], allr["feat"][pick])
        assert np.array_equal(t["tgt"], allr["tgt"][pick], equal_nan=True)
