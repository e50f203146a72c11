"""CPU: the prefix-emission estimator of the cached renderer's prefix paths
(orc_cache_queries with emit_expect, the oracle of LISA_CACHE_EMISSION).

At a roughness-1 Lambertian vertex with a bounce left, the plain tracer's
bounce (shader.cu:251-252 -> closest hit 211-223) adds Le * att when it hits
an emitter and ends the path.  The estimator takes that term in expectation
from one light sample and draws the bounce on the condition that it hits no
emitter, with the throughput scaled by (1 - the sample's hit estimate).  Its
expectation is the plain one, so both quantities that the cache's image is
built from keep their means: the radiance gathered before the queries
(accum) and the throughput that reaches the queries (what a network that
answers 1 everywhere would add).  The check is a paired z-test over pixels
(same seeds, independent pixels) on two scenes: the island's sun seen from
the ground and the Cornell box's ceiling light.  The bound 4 leaves room for
chance; leaving out the emission term gives z = -10 and -19, leaving out the
(1 - p) throughput factor z = 7.4 on the Cornell box, so both mistakes fail.
"""
import numpy as np
import pytest

from lisa_amd import scenes
from tests.scene_util import cornell

Z_BOUND = 4.0


def _pixel_sums(o, om, sc, W, H, spp, prefix, subframe, emit, loops=1, deep=None):
    cam = sc["camera"]
    U, V, Wv = om.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
    px = np.arange(W * H, dtype=np.int32)
    r = o.cache_queries(W, H, spp, prefix, cam["eye"], U, V, Wv, px, subframe=subframe, threads=8,
                        emit_expect=emit, nee_loops=loops, nee_loops_deep=deep)
    acc = r["accum"][:, :3].astype(np.float64).sum(1)
    q = np.zeros(W * H)
    np.add.at(q, r["item"], r["att"].astype(np.float64).sum(1))
    return acc, q / spp, r


def _scene(om, sc):
    return om.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])


CASES = {
    "island": (lambda: scenes.island(n=128, width=64, height=36), 64, 36, 128, 2),
    "cornell": (lambda: cornell(32, 16, 48, 48, 4, 7), 48, 48, 256, 1),
}


@pytest.mark.parametrize("loops,deep", [(1, 1), (3, 3), (8, 1)])
@pytest.mark.parametrize("case", sorted(CASES))
def test_prefix_emission_keeps_the_expectation(oracle_mod, case, loops, deep):
    """loops / deep: NEE runs averaged at the camera ray's hit / at later
    prefix vertices (each from the vertex's entry hit flag and light; the
    last run's flag goes on), which keeps the expectation of the reference's
    NEE at that vertex and of what follows.  Starting the later runs from
    the flag the first run left instead gives z = +25 on the Cornell box."""
    make, W, H, spp, prefix = CASES[case]
    sc = make()
    o = _scene(oracle_mod, sc)
    a0, q0, r0 = _pixel_sums(o, oracle_mod, sc, W, H, spp, prefix, 3, False)
    a1, q1, r1 = _pixel_sums(o, oracle_mod, sc, W, H, spp, prefix, 3, True, loops, deep)
    assert r0["n"] > 0 and r1["n"] > 0
    for name, x, y in (("accum", a0, a1), ("query throughput", q0, q1)):
        d = y - x
        z = d.mean() / (d.std() / np.sqrt(d.size))
        print(f"{case} loops {loops}/{deep} {name}: plain {x.mean():.5f} estimator {y.mean():.5f} "
              f"z {z:+.2f}", flush=True)
        assert abs(z) < Z_BOUND, (case, name, z)


def test_prefix_emission_lowers_the_island_variance(oracle_mod):
    """On the island most of the prefix's noise is the sun hit by a diffuse
    bounce; taken in expectation it halves the per-pixel Monte Carlo
    variance of the prefix radiance (measured: 0.050 -> 0.024 at 16 spp)."""
    W, H, spp = 96, 54, 16
    sc = scenes.island(n=256, width=W, height=H)
    o = _scene(oracle_mod, sc)
    var = {}
    for emit in (False, True):
        x = _pixel_sums(o, oracle_mod, sc, W, H, spp, 2, 5, emit)[0]
        y = _pixel_sums(o, oracle_mod, sc, W, H, spp, 2, 6, emit)[0]
        var[emit] = float(np.var(x - y) / 2)
    print("Monte Carlo variance of the prefix radiance:", var, flush=True)
    assert var[True] < 0.7 * var[False]


def test_nee_loops_lower_the_island_variance(oracle_mod):
    """With the sun's hits taken in expectation, the reference's NEE (one
    light or none after up to 30 hemisphere tries) is most of what is left
    of the prefix's noise on the island: the mean of 4 NEE runs per vertex
    cuts the per-pixel variance to about a third (measured here: 0.024 ->
    0.0074 at 16 spp)."""
    W, H, spp = 96, 54, 16
    sc = scenes.island(n=256, width=W, height=H)
    o = _scene(oracle_mod, sc)
    var = {}
    for loops in (1, 4):
        x = _pixel_sums(o, oracle_mod, sc, W, H, spp, 2, 5, True, loops)[0]
        y = _pixel_sums(o, oracle_mod, sc, W, H, spp, 2, 6, True, loops)[0]
        var[loops] = float(np.var(x - y) / 2)
    print("Monte Carlo variance of the prefix radiance by NEE runs:", var, flush=True)
    assert var[4] < 0.5 * var[1]


def test_nee_loops_need_the_estimator(oracle_mod):
    sc = cornell(16, 8, 16, 16, 1, 7)
    o = _scene(oracle_mod, sc)
    with pytest.raises(RuntimeError):
        _pixel_sums(o, oracle_mod, sc, 16, 16, 1, 2, 1, False, 2)


def test_no_emitter_is_the_plain_prefix(oracle_mod):
    """Without emitters there is nothing to sample: the estimator's paths
    are the plain prefix paths, bit for bit."""
    sc = cornell(16, 8, 32, 32, 2, 7)
    mats = [dict(m) for m in sc["materials"]]
    for m in mats:
        m["emit"] = False
        m["emission"] = (0.0, 0.0, 0.0)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], mats)
    sc2 = dict(sc)
    a = _pixel_sums(o, oracle_mod, sc2, 32, 32, 2, 2, 1, False)[2]
    b = _pixel_sums(o, oracle_mod, sc2, 32, 32, 2, 2, 1, True)[2]
    assert a["n"] == b["n"] > 0
    for k in ("accum", "feat", "att", "item"):
        assert np.array_equal(a[k], b[k]), k


def test_prefix_light_sample_is_bounded_in_the_cornell_box(oracle_mod):
    """In the closed Cornell box many vertices lie close to the ceiling light,
    where the light sample's hit estimate p^ = q |cos| A / r^2 could exceed 1
    (and the throughput factor 1 - p^ turn negative).  The estimator is used
    only where a bound on p^ from the vertex's distance to the emitters' box
    keeps it below 1/2 (elsewhere the reference's bounce), so every query
    throughput and pixel mean stays non-negative, and the estimator's
    per-pixel variance is no larger than the plain prefix's.  Without the
    bound this scene gives negative throughputs."""
    W = H = 48
    sc = cornell(32, 16, W, H, 4, 7)
    o = _scene(oracle_mod, sc)
    var = {}
    for emit in (False, True):
        x, _, r = _pixel_sums(o, oracle_mod, sc, W, H, 64, 2, 5, emit)
        y = _pixel_sums(o, oracle_mod, sc, W, H, 64, 2, 6, emit)[0]
        assert r["n"] > 0
        assert np.all(r["att"] >= 0) and np.all(r["accum"] >= 0), emit
        var[emit] = float(np.var(x - y) / 2)
    print("Cornell prefix variance (plain, estimator):", var, flush=True)
    assert var[True] <= 1.05 * var[False]


@pytest.mark.parametrize("case", sorted(CASES))
def test_full_paths_variance_reduced_plain(oracle_mod, case):
    """The variance-reduced plain tracer (orc_render_full_paths: the prefix
    estimators on whole 7-bounce paths, no network): without the estimators
    it is the plain tracer bit for bit; with them (light samples, 8 NEE runs
    at the camera ray's hit, 1 later) its pixel means keep the plain ones'
    expectation (paired z-test) at less variance on the island."""
    make, W, H, spp, _ = CASES[case]
    sc = make()
    o = _scene(oracle_mod, sc)
    cam = sc["camera"]
    U, V, Wv = oracle_mod.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
    px = np.arange(W * H, dtype=np.int32)
    spp = 64
    plain = o.render(W, H, spp, 7, cam["eye"], U, V, Wv, subframe=3, pixels=px, threads=8)["accum"]
    same = o.render_full_paths(W, H, spp, 7, cam["eye"], U, V, Wv, px, subframe=3, threads=8,
                               emit_expect=False)
    assert np.array_equal(same.view(np.uint32), plain.view(np.uint32))
    vr = o.render_full_paths(W, H, spp, 7, cam["eye"], U, V, Wv, px, subframe=3, threads=8,
                             emit_expect=True, nee_loops=8, nee_loops_deep=1)
    vr2 = o.render_full_paths(W, H, spp, 7, cam["eye"], U, V, Wv, px, subframe=4, threads=8,
                              emit_expect=True, nee_loops=8, nee_loops_deep=1)
    plain2 = o.render(W, H, spp, 7, cam["eye"], U, V, Wv, subframe=4, pixels=px, threads=8)["accum"]
    d = vr[:, :3].astype(np.float64).sum(1) - plain[:, :3].astype(np.float64).sum(1)
    z = d.mean() / (d.std() / np.sqrt(d.size))
    v_vr = float(np.var(vr[:, :3].astype(np.float64) - vr2[:, :3]) / 2)
    v_pl = float(np.var(plain[:, :3].astype(np.float64) - plain2[:, :3]) / 2)
    print(f"{case}: z {z:+.2f}, per-pixel variance plain {v_pl:.3g} vr {v_vr:.3g}", flush=True)
    assert abs(z) < Z_BOUND
    assert np.all(vr >= 0)
    if case == "island":
        assert v_vr < 0.7 * v_pl
