"""Pins for the oracle's LBVH and traversal (no GPU).

* Morton codes: an independent bit interleave of the quantised coordinates.
* LBVH (Karras, HPG 2012): sorted keys non-decreasing, stable for equal keys,
  every internal node spans a contiguous range of sorted leaves whose split
  sits at the highest differing bit of the index-augmented keys, every leaf
  reached exactly once, depth < 64, refit boxes = brute-force unions.
* Traversal: the BVH closest hit equals a brute-force loop over every
  triangle (same watertight test, same tie rule), bit for bit.
* Watertightness (Woop, Benthin & Wald, JCGT 2013): rays through shared
  edges and vertices of a triangle fan never fall through.
"""
import numpy as np
import pytest

from tests.scene_util import cornell, random_rays

LEAF = 0x80000000
EMPTY = 0xFFFFFFFF


def interleave(qx, qy, qz):
    code = 0
    for b in range(10):
        code |= ((qx >> b) & 1) << (3 * b + 2)
        code |= ((qy >> b) & 1) << (3 * b + 1)
        code |= ((qz >> b) & 1) << (3 * b)
    return code


def test_morton_interleave(oracle_mod):
    rs = np.random.default_rng(0)
    for _ in range(500):
        u = rs.uniform(-0.1, 1.1, size=3).astype(np.float32)
        q = [min(max(int(np.float32(x) * np.float32(1024.0)), 0), 1023) if x >= 0 else 0 for x in u]
        assert oracle_mod.morton3(*map(float, u)) == interleave(*q)
    assert oracle_mod.morton3(1.0, 1.0, 1.0) == (1 << 30) - 1
    assert oracle_mod.morton3(0.0, 0.0, 0.0) == 0


def _tri_boxes(sc, prim):
    # triangle bounds widened outward by one ulp
    v = sc["verts"].reshape(-1, 3, 3)[prim]
    return (np.nextafter(v.min(axis=1), np.float32(-np.inf)),
            np.nextafter(v.max(axis=1), np.float32(np.inf)))


def _check_tree(sc, bvh):
    T = sc["verts"].shape[0]
    morton, prim, refs, boxes = bvh["morton"], bvh["prim"], bvh["refs"], bvh["boxes"]
    assert np.all(np.diff(morton.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(prim), np.arange(T))
    same = np.diff(morton.astype(np.int64)) == 0
    assert np.all(np.diff(prim.astype(np.int64))[same] > 0)  # stable
    if T == 1:
        assert refs[0, 0] == LEAF and refs[0, 1] == EMPTY
        return
    lo, hi = _tri_boxes(sc, prim)  # per sorted leaf

    def aug(i):
        return (int(morton[i]) << 32) | i

    seen = np.zeros(T, np.int32)
    stack = [(0, 0)]
    ranges = {}
    maxdepth = 0

    def span(ref):
        if ref & LEAF:
            k = ref & ~LEAF
            return k, k
        return ranges[ref]

    # post-order: compute ranges and boxes bottom up
    order = []
    while stack:
        node, depth = stack.pop()
        maxdepth = max(maxdepth, depth)
        order.append(node)
        for r in refs[node]:
            r = int(r)
            if r & LEAF:
                seen[r & ~LEAF] += 1
            else:
                stack.append((r, depth + 1))
    assert np.all(seen == 1)
    assert maxdepth < 64
    for node in reversed(order):
        r0, r1 = int(refs[node, 0]), int(refs[node, 1])
        a0, b0 = span(r0)
        a1, b1 = span(r1)
        assert b0 + 1 == a1  # contiguous, left range then right range
        ranges[node] = (a0, b1)
        # split at the highest differing bit of the augmented keys of the range
        first, last = aug(a0), aug(b1)
        if first != last:
            bit = (first ^ last).bit_length() - 1
            assert (aug(b0) >> bit) & 1 == 0 and (aug(a1) >> bit) & 1 == 1
        for ch, (a, b) in enumerate(((a0, b0), (a1, b1))):
            want_lo = lo[a:b + 1].min(axis=0)
            want_hi = hi[a:b + 1].max(axis=0)
            got = boxes[node, 6 * ch:6 * ch + 6]
            assert np.array_equal(got[:3], want_lo) and np.array_equal(got[3:], want_hi)
    assert ranges[0] == (0, T - 1)


@pytest.mark.parametrize("seg,rings", [(32, 16), (64, 32)])
def test_lbvh_invariants(oracle_mod, seg, rings):
    sc = cornell(seg, rings, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    _check_tree(sc, o.bvh())


def test_lbvh_duplicate_keys_and_tiny(oracle_mod):
    # many triangles with the same centroid: equal Morton keys, index-augmented split
    rs = np.random.default_rng(3)
    base = rs.uniform(-1, 1, size=(1, 9)).astype(np.float32)
    v = np.repeat(base, 37, axis=0)
    n = np.tile(np.float32([0, 1, 0]), (37, 3))
    mats = [{"diffuse": (1, 1, 1), "alpha": 1.0, "roughness": 1.0}]
    for T in (1, 2, 3, 37):
        sc = {"verts": v[:T], "normals": n[:T], "mat_idx": np.zeros(T, np.int32),
              "materials": mats}
        o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], mats)
        _check_tree(sc, o.bvh())


@pytest.mark.parametrize("seg,rings,inside", [(32, 16, True), (32, 16, False), (64, 32, True)])
def test_traversal_equals_brute_force(oracle_mod, seg, rings, inside):
    sc = cornell(seg, rings, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    O, D = random_rays(4000, seed=seg + inside, inside=inside)
    # plus unnormalised and axis-parallel directions (dielectric / lerp paths)
    D[:500] *= 3.7
    D[500:600] = np.float32([0, -1, 0])
    D[600:700] = np.float32([1, 0, 0])
    p1, t1, cnt = o.closest_hit(O, D, threads=4)
    p2, t2, _ = o.closest_hit(O, D, brute=True)
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert cnt[2] > 0 and cnt[3] > 0


def test_nan_vertices_are_never_hit(oracle_mod):
    """Triangles with one, two or three NaN vertices in the Cornell box: the
    LBVH over them still answers every ray as the brute-force loop does (NaN
    boxes drop out of the unions, a NaN edge function or t never accepts),
    no NaN triangle is ever the closest hit, and the other triangles' hits
    are those of the scene without them."""
    sc = cornell(32, 16, 8, 8, 1, 1)
    v = np.array(sc["verts"][:30], np.float32).reshape(-1, 3, 3)
    for i in range(v.shape[0]):
        v[i, : 1 + i % 3] = np.nan
    T0 = sc["verts"].shape[0]
    verts = np.concatenate([sc["verts"], v.reshape(-1, 9)])
    o = oracle_mod.Scene(verts, np.concatenate([sc["normals"], sc["normals"][:30]]),
                         np.concatenate([sc["mat_idx"], sc["mat_idx"][:30]]), sc["materials"])
    ref = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    O, D = random_rays(4000, seed=9)
    p1, t1, _ = o.closest_hit(O, D, threads=4)
    p2, t2, _ = o.closest_hit(O, D, brute=True)
    p3, t3, _ = ref.closest_hit(O, D, brute=True)
    assert np.array_equal(p1, p2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert np.all(p1 < T0) and (p1 >= 0).mean() > 0.8
    assert np.array_equal(p1, p3) and np.array_equal(t1.view(np.uint32), t3.view(np.uint32))


def test_zero_and_nan_directions_miss(oracle_mod):
    sc = cornell(32, 16, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    O = np.zeros((3, 3), np.float32)
    D = np.float32([[0, 0, 0], [np.nan, 0, 1], [np.inf, 0, 0]])
    p, _, cnt = o.closest_hit(O, D)
    assert np.all(p == -1) and cnt[2] == 0


def test_watertight_fan(oracle_mod):
    # a fan of 8 triangles around the origin in the z = 0 plane; rays straight
    # down through every shared edge and the shared centre vertex must hit
    k = 8
    ang = np.linspace(0, 2 * np.pi, k + 1)
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1).astype(np.float32)
    v = np.concatenate([np.zeros((k, 3), np.float32), rim[:-1], rim[1:]], axis=1)
    n = np.tile(np.float32([0, 0, 1]), (k, 3))
    mats = [{"diffuse": (1, 1, 1), "alpha": 1.0, "roughness": 1.0}]
    o = oracle_mod.Scene(v, n, np.zeros(k, np.int32), mats)
    pts = [np.float32([0, 0, 0])]
    for i in range(k):
        for s in np.linspace(0.05, 0.95, 19):
            pts.append((rim[i] * np.float32(s)).astype(np.float32))  # on shared spokes
    P = np.array(pts, np.float32)
    O = P + np.float32([0, 0, 1])
    for d in (np.float32([0, 0, -1]), np.float32([0.3, -0.2, -1]), np.float32([1e-3, 0, -1])):
        D = np.tile(d, (len(P), 1))
        Oo = P - d  # ray passes exactly through P at t = 1
        prim, t, _ = o.closest_hit(Oo, D)
        assert np.all(prim >= 0), np.nonzero(prim < 0)


def test_emitter_bvh_invariants(oracle_mod):
    sc = cornell(32, 16, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    emits = np.array([m.get("emit", False) for m in sc["materials"]])
    gids = np.nonzero(emits[sc["mat_idx"]])[0]
    b = o.bvh(which=1)
    assert np.array_equal(np.sort(b["prim"]), gids)
    sub = {"verts": sc["verts"][b["prim"]]}
    # the subset tree's leaves refer to sorted positions; check it with the subset soup
    local = dict(b, prim=np.arange(len(gids), dtype=np.uint32))
    _check_tree(sub, local)


@pytest.mark.parametrize("seed", [0, 1])
def test_occlusion_query_equals_closest_hit(oracle_mod, seed):
    """NEE's emitter-BVH + any-hit query gives the same outcome as one full
    closest-hit query, for both values of the (stale) hit flag."""
    sc = cornell(64, 32, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    rs = np.random.default_rng(seed)
    n = 20000
    O = rs.uniform(-0.49, 0.49, size=(n, 3)).astype(np.float32)
    # half the rays aimed at points of the light quad, half in random directions
    L = np.stack([rs.uniform(-0.125, 0.125, n), np.full(n, 0.499), rs.uniform(-0.125, 0.125, n)], 1)
    D = (L - O).astype(np.float32)
    D[n // 2:] = rs.normal(size=(n - n // 2, 3)).astype(np.float32)
    D[:100] = 0.0  # zero directions
    for hb in (0, 1):
        fast, ref, cnt = o.occlusion(O, D, np.full(n, hb))
        assert np.array_equal(fast, ref)
        assert (fast >= 0).sum() > 1000  # lights are found
        if hb:
            assert (fast == -4).sum() > 1000 and (fast == -1).sum() > 100


@pytest.mark.parametrize("scale,shift", [(1.0, (0.0, 0.0, 0.0)), (40.0, (3000.0, -700.0, 250.0))])
def test_emitter_cull_is_conservative(oracle_mod, scale, shift):
    """The division-free cull in front of NEE's emitter query never drops a
    ray that the full query would answer with an emitter: rays aimed at the
    light quad's corners and at points within 1e-7 .. 1e-2 (relative) of its
    edges, inside and outside, from points all over the box, in the unit
    scene and in a scaled, translated copy (the cull's margin scales with
    the largest coordinate).  Rays in random directions are culled: they
    cost no node visit unless they pass near a light."""
    sc = cornell(32, 16, 8, 8, 1, 1)
    shift = np.asarray(shift, np.float64)
    V = (np.asarray(sc["verts"], np.float64).reshape(-1, 3) * scale + shift)
    verts = V.astype(np.float32).reshape(np.asarray(sc["verts"]).shape)
    o = oracle_mod.Scene(verts, sc["normals"], sc["mat_idx"], sc["materials"])
    emits = np.array([bool(m.get("emit", False)) for m in sc["materials"]])
    light = verts.reshape(-1, 3, 3)[emits[np.asarray(sc["mat_idx"])]].reshape(-1, 3).astype(np.float64)
    lo, hi = light.min(0), light.max(0)
    rs = np.random.default_rng(7)
    n = 40000
    O = (rs.uniform(-0.49, 0.49, size=(n, 3)) * scale + shift)
    # targets on the light's edges and corners, pushed in or out along the
    # edge normal in the light's plane by a relative offset
    flat = int(np.argmin(hi - lo))
    ax = [a for a in range(3) if a != flat]
    T = np.empty((n, 3))
    T[:, flat] = lo[flat]
    u = rs.uniform(0, 1, n)
    side = rs.integers(0, 4, n)
    off = 10.0 ** rs.uniform(-7, -2, n) * scale * rs.choice([-1.0, 1.0], n)
    for k in range(4):
        m = side == k
        a_run, a_fix = (ax[0], ax[1]) if k < 2 else (ax[1], ax[0])
        T[m, a_run] = lo[a_run] + u[m] * (hi[a_run] - lo[a_run])
        T[m, a_fix] = (lo[a_fix] if k % 2 == 0 else hi[a_fix]) + off[m] * (1 if k % 2 else -1)
    corners = rs.random(n) < 0.1
    T[corners, ax[0]] = np.where(rs.random(corners.sum()) < 0.5, lo[ax[0]], hi[ax[0]])
    T[corners, ax[1]] = np.where(rs.random(corners.sum()) < 0.5, lo[ax[1]], hi[ax[1]])
    O32 = O.astype(np.float32)
    D = (T - O32).astype(np.float32)
    for hb in (0, 1):
        fast, ref, _ = o.occlusion(O32, D, np.full(n, hb))
        assert np.array_equal(fast, ref)
        assert (fast >= 0).sum() > 5000
    # the raw-draw culls (any length, either sign) never drop a line that the
    # full query answers with an emitter in one of its two directions
    scale_d = (10.0 ** rs.uniform(-0.3, 0.3, n))[:, None] * rs.choice([-1.0, 1.0], n)[:, None]
    raw = (D * scale_d).astype(np.float32)
    cull = o.raw_cull(O32, raw)
    _, fwd, _ = o.occlusion(O32, D, np.zeros(n, np.int32))
    _, bwd, _ = o.occlusion(O32, -D, np.zeros(n, np.int32))
    assert not np.any(cull.astype(bool) & ((fwd >= 0) | (bwd >= 0)))
    # random directions: culled unless they pass near the light
    R = rs.normal(size=(n, 3)).astype(np.float32)
    fast, ref, cnt = o.occlusion(O32, R, np.zeros(n, np.int32))
    assert np.array_equal(fast, ref)
    assert int(cnt[2]) < 0.25 * n
    _, bwd, _ = o.occlusion(O32, -R, np.zeros(n, np.int32))
    cull = o.raw_cull(O32, R).astype(bool)
    assert cull.sum() > 0.5 * n and not np.any(cull & ((fast >= 0) | (bwd >= 0)))


def test_refit_update_keeps_topology(oracle_mod):
    """Refit-only update: new positions, same Morton order and hierarchy;
    boxes are the unions of the new triangle boxes and traversal still
    equals the brute-force loop."""
    sc = cornell(32, 16, 8, 8, 1, 1)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    before = o.bvh()
    rs = np.random.default_rng(5)
    v2 = (sc["verts"] + rs.normal(scale=0.02, size=sc["verts"].shape)).astype(np.float32)
    o.update_vertices(v2)
    after = o.bvh()
    assert np.array_equal(before["morton"], after["morton"])
    assert np.array_equal(before["prim"], after["prim"])
    assert np.array_equal(before["refs"], after["refs"])
    assert not np.array_equal(before["boxes"], after["boxes"])
    _check_tree({"verts": v2}, after)
    O, D = random_rays(3000, seed=9)
    p1, t1, _ = o.closest_hit(O, D)
    p2, t2, _ = o.closest_hit(O, D, brute=True)
    assert np.array_equal(p1, p2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
