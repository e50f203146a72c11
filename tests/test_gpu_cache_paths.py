"""GPU: the radiance-cache path forms of lisa_render_cached against the CPU
oracle (orc_cache_queries, orc_cache_train_rows) and the plain tracer.

Bars: query items, throughputs, position / albedo / roughness / dielectric
features and training targets are bit-exact (the path loop is, and the
features are FP32 arithmetic without transcendental functions); the
direction and normal features go through acosf / atan2f, whose GPU and glibc
results may differ by a few ulps: |diff| <= 2e-6 (the features lie in
[0, 1]).  Per-pixel means: 1e-5 relative, as in the plain path parity.
Queries and rows come out of the GPU in atomic order, so both sides are
sorted by (item, exact features) before the comparison.
"""
import numpy as np
import pytest

from tests.scene_util import camera, cornell

pytestmark = pytest.mark.gpu

EXACT = [0, 1, 2, 7, 8, 9, 10, 11]
TRIG = [3, 4, 5, 6]


@pytest.fixture(scope="module")
def ctx(capi):
    c = capi.Context(0)
    yield c
    c.close()


def _scene(capi, oracle_mod, ctx, W=64, H=64):
    sc = cornell(32, 16, W, H, 4, 7)
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    return sc, o, camera(sc, capi)


def _close(a, b, rtol=1e-5):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-30)


@pytest.mark.parametrize("spp,prefix,emit,loops,deep", [
    (1, 2, False, 1, 1), (4, 2, False, 1, 1), (2, 1, False, 1, 1), (2, 0, False, 1, 1),
    (4, 2, True, 1, 1), (2, 1, True, 1, 1), (8, 3, True, 1, 1), (4, 2, True, 3, 3),
    (2, 3, True, 16, 16), (4, 2, True, 8, 1), (2, 3, True, 1, 5)])
def test_queries_match_oracle(capi, oracle_mod, ctx, spp, prefix, emit, loops, deep):
    """emit: the prefix-emission estimator (LISA_CACHE_EMISSION), whose light
    samples, conditioned bounces and throughputs are bit-exact too; loops /
    deep: NEE runs averaged at the camera ray's hit / at later vertices."""
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    px = np.sort(np.random.default_rng(spp).choice(W * H, 1500, replace=False)).astype(np.int32)
    p = capi.render_params(W, H, spp, prefix + 1, eye, U, V, Wv, subframe=3)
    cp = capi.cache_frame_params(prefix_bounces=prefix, flags=capi.CACHE_EMISSION if emit else 0,
                                 nee_loops=loops, nee_loops_deep=deep)
    g = capi.cache_trace_queries(ctx, p, cp, px)
    r = o.cache_queries(W, H, spp, prefix, eye, U, V, Wv, px, subframe=3, threads=8, emit_expect=emit,
                        nee_loops=loops, nee_loops_deep=deep)
    print("queries", g["n"], r["n"], flush=True)
    assert g["n"] == r["n"] > 0
    # the oracle's order: listed pixel, then sample
    assert np.array_equal(g["item"], r["item"])
    assert np.array_equal(g["feat"][:, EXACT], r["feat"][:, EXACT])
    assert np.abs(g["feat"][:, TRIG] - r["feat"][:, TRIG]).max() <= 2e-6
    assert np.array_equal(g["att"], r["att"])
    assert _close(g["accum"], r["accum"])


def test_training_rows_match_oracle(capi, oracle_mod, ctx):
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    px = np.random.default_rng(9).integers(0, W * H, 1024).astype(np.int32)  # repeats allowed
    p = capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=5)
    g = capi.cache_trace_rows(ctx, p, px)
    r = o.cache_train_rows(W, H, 2, 7, eye, U, V, Wv, px, subframe=5, threads=8)
    print("rows", g["n"], r["n"], flush=True)
    assert g["n"] == r["n"] > 1024
    # the oracle's order: listed pixel, then sample, then vertex
    assert np.array_equal(g["feat"][:, EXACT], r["feat"][:, EXACT])
    assert np.abs(g["feat"][:, TRIG] - r["feat"][:, TRIG]).max() <= 2e-6
    # targets bit-exact, including the non-finite ones of zero-throughput channels
    assert np.array_equal(g["tgt"], r["tgt"], equal_nan=True)
    # a cap keeps the first rows only and reports the full count
    g2 = capi.cache_trace_rows(ctx, p, px, cap=100)
    assert g2["n"] == g["n"] and g2["feat"].shape[0] == 100
    assert np.array_equal(g2["feat"], g["feat"][:100])
    assert np.array_equal(g2["tgt"], g["tgt"][:100], equal_nan=True)
    # rows at one depth (the query depth of lisa_render_cached)
    g3 = capi.cache_trace_rows(ctx, p, px, depth=3)
    r3 = o.cache_train_rows(W, H, 2, 7, eye, U, V, Wv, px, subframe=5, depth=3, threads=8)
    assert g3["n"] == r3["n"] > 0
    assert np.array_equal(g3["feat"][:, EXACT], r3["feat"][:, EXACT])
    assert np.array_equal(g3["tgt"], r3["tgt"], equal_nan=True)


def test_forced_fallback_is_the_plain_progressive_render(capi, oracle_mod, ctx):
    """With a fallback threshold below any training loss every frame is the
    plain tracer, and the running mean over frames is lisa_render's
    progressive subframe mean, bit for bit."""
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    cache = capi.Cache(0, seed=1, lr=1e-3)
    plain = capi.Context(0)
    try:
        plain.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        cp = capi.cache_frame_params(prefix_bounces=2, train_paths=512, train_bounces=7,
                                     train_steps=1, fallback_loss=1e-30, seed=2)
        for f in range(3):
            p = capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=f)
            st = capi.render_cached(ctx, cache, p, cp)
            assert st["fell_back"] == 1 and st["train_samples"] > 0
            want, _ = plain.render(p)
            got = ctx.copy_accum(W * H)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f
    finally:
        plain.close()
        cache.close()


def _cached_rmse(capi, ctx, cache, cp, params_of, frames, ref):
    for f in range(frames):
        st = capi.render_cached(ctx, cache, params_of(f), cp)
        assert st["fell_back"] == 0
    img = ctx.copy_accum(ref.shape[0])
    assert np.all(np.isfinite(img)) and np.all(img[:, :3] >= 0)
    return float(np.sqrt(np.mean((img[:, :3].astype(np.float64) - ref) ** 2))), st


def test_trained_cache_gains_over_truncated_paths(capi, oracle_mod, ctx):
    """What the trained network gains: the same prefix paths (2 bounces, then
    the cache) with the network's answer replaced by 0 are plain paths cut
    after 3 radiance rays, and the Cornell box loses its indirect light.
    Trained online for 96 frames, the network recovers most of it: the RMSE
    against an 8192-spp plain render drops well below the cut paths' and the
    untrained network's."""
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    plain = capi.Context(0)
    try:
        plain.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        ref = plain.trace_subset(capi.render_params(W, H, 8192, 7, eye, U, V, Wv, subframe=77),
                                 np.arange(W * H, dtype=np.int32), records=False)["accum"]
        ref = ref[:, :3].astype(np.float64)
    finally:
        plain.close()

    def params_of(f):
        return capi.render_params(W, H, 4, 7, eye, U, V, Wv, subframe=f)

    frames = 96
    trained = capi.Cache(0, seed=1, lr=1e-3)
    untrained = capi.Cache(0, seed=1, lr=1e-3)
    zero = capi.Cache(0, seed=1, lr=1e-3)
    try:
        zero.set_weights(np.zeros(capi.CACHE_PARAMS, np.float32))
        cp = capi.cache_frame_params(prefix_bounces=2, train_paths=32768, train_bounces=7,
                                     train_steps=4, seed=2)
        off = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_bounces=7,
                                      train_steps=0, seed=2)
        e_tr, st = _cached_rmse(capi, ctx, trained, cp, params_of, frames, ref)
        assert st["queries"] > 0 and st["train_samples"] > 0
        e_un, _ = _cached_rmse(capi, ctx, untrained, off, params_of, frames, ref)
        e_0, _ = _cached_rmse(capi, ctx, zero, off, params_of, frames, ref)
        print(f"RMSE vs 8192-spp plain: trained {e_tr:.5f}, untrained {e_un:.5f}, "
              f"cut paths (zero cache) {e_0:.5f}", flush=True)
        assert e_tr < 0.6 * e_0
        assert e_tr < 0.25 * e_un
    finally:
        for c in (trained, untrained, zero):
            c.close()


@pytest.mark.parametrize("emission", [False, True])
def test_persistent_cached_frames_same_image(capi, oracle_mod, ctx, emission):
    """LISA_CACHE_PERSISTENT runs a frame's paths in one persistent launch.
    With a network that answers 0 (and no training) the frame is the prefix
    radiance alone, which both launch forms compute identically."""
    W, H = 100, 60
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    other = capi.Context(0)
    a, b = capi.Cache(0, seed=1), capi.Cache(0, seed=1)
    try:
        other.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        for c in (a, b):
            c.set_weights(np.zeros(capi.CACHE_PARAMS, np.float32))
        em = capi.CACHE_EMISSION if emission else 0
        tiled = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0, flags=em)
        pers = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0,
                                       flags=capi.CACHE_PERSISTENT | em)
        for f in range(3):
            p = capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=f)
            s1 = capi.render_cached(ctx, a, p, tiled)
            s2 = capi.render_cached(other, b, p, pers)
            assert s1["queries"] == s2["queries"] > 0
            g1, g2 = ctx.copy_accum(W * H), other.copy_accum(W * H)
            assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)), f
    finally:
        other.close()
        a.close()
        b.close()


def test_trained_cached_frames_are_reproducible(capi, oracle_mod, ctx):
    """Queries and training rows take fixed slots (pixel, sample) and each
    pixel's queries are summed in sample order, so a trained cached sequence
    is the same bits whatever the launch form and stream order: tiled paths
    with training beside the frame on a second stream, and a persistent
    launch with training first, give the same images and weights."""
    W, H = 96, 80
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    other = capi.Context(0)
    a, b = capi.Cache(0, seed=1), capi.Cache(0, seed=1)
    try:
        other.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        kw = dict(prefix_bounces=2, train_paths=6000, train_bounces=7, train_steps=3, seed=4)
        overlap = capi.cache_frame_params(**kw)
        serial = capi.cache_frame_params(flags=capi.CACHE_PERSISTENT | capi.CACHE_SERIAL, **kw)
        for f in range(4):
            p = capi.render_params(W, H, 3, 7, eye, U, V, Wv, subframe=f)
            s1 = capi.render_cached(ctx, a, p, overlap)
            s2 = capi.render_cached(other, b, p, serial)
            assert s1["queries"] == s2["queries"] > W * H
            assert s1["train_samples"] == s2["train_samples"] > 0
            assert s1["loss"] == s2["loss"]
            g1, g2 = ctx.copy_accum(W * H), other.copy_accum(W * H)
            assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)), f
        assert np.array_equal(a.weights().view(np.uint32), b.weights().view(np.uint32))
    finally:
        other.close()
        a.close()
        b.close()


def test_variance_reduced_prefix_on_the_gpu(capi, oracle_mod, ctx):
    """The prefix radiance of a cached frame (the image with a cache that
    answers 0) from the GPU's two prefix forms on a small island: the
    variance-reduced prefix (LISA_CACHE_EMISSION, 3 NEE runs) converges to
    the plain prefix's mean (paired over pixels at 4096 spp, |z| < 4) and at
    64 spp lands much closer to it (RMSE below 0.75x the plain prefix's)."""
    from lisa_amd import scenes
    W, H = 192, 108
    sc = scenes.island(n=256, width=W, height=H)
    other = capi.Context(0)
    try:
        other.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        cam = sc["camera"]
        U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
        px = scenes.seeded_pixels(W, H, 2048, seed=3)
        plain = capi.cache_frame_params(prefix_bounces=2)
        vr = capi.cache_frame_params(prefix_bounces=2, flags=capi.CACHE_EMISSION, nee_loops=3)

        def prefix(cp, spp, sub, calls=1):
            acc = np.zeros((px.size, 3))
            for c in range(calls):
                p = capi.render_params(W, H, spp, 3, cam["eye"], U, V, Wv, subframe=sub + c)
                acc += capi.cache_trace_queries(other, p, cp, px, cap=0)["accum"][:, :3] / calls
            return acc.sum(1)

        ref = prefix(plain, 1024, 1000, 4)
        hi = prefix(vr, 1024, 2000, 4)
        d = hi - ref
        z = d.mean() / (d.std() / np.sqrt(d.size))
        e_plain = float(np.sqrt(np.mean((prefix(plain, 64, 7) - ref) ** 2)))
        e_vr = float(np.sqrt(np.mean((prefix(vr, 64, 7) - ref) ** 2)))
        print(f"prefix mean plain {ref.mean():.5f} reduced {hi.mean():.5f} z {z:+.2f}; "
              f"RMSE at 64 spp: plain {e_plain:.5f} reduced {e_vr:.5f}", flush=True)
        assert abs(z) < 4.0
        assert e_vr < 0.75 * e_plain
    finally:
        other.close()



def test_emission_prefix_after_the_emitters_are_gone(capi, oracle_mod, ctx):
    """A scene without emitters set on a context that had some: the
    estimator has nothing to sample, and its prefix paths (one NEE run)
    are the plain prefix's bit for bit."""
    W = H = 32
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)  # with its ceiling light
    mats = [dict(m, emit=False, emission=(0.0, 0.0, 0.0)) for m in sc["materials"]]
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], mats)
    px = np.arange(W * H, dtype=np.int32)
    p = capi.render_params(W, H, 2, 3, sc["camera"]["eye"], U, V, Wv, subframe=1)
    a = capi.cache_trace_queries(ctx, p, capi.cache_frame_params(prefix_bounces=2), px)
    b = capi.cache_trace_queries(ctx, p, capi.cache_frame_params(prefix_bounces=2,
                                                                 flags=capi.CACHE_EMISSION), px)
    assert a["n"] == b["n"] > 0
    assert np.array_equal(a["accum"].view(np.uint32), b["accum"].view(np.uint32))
    assert np.array_equal(a["item"], b["item"])
    assert np.array_equal(a["feat"], b["feat"])
    assert np.array_equal(a["att"], b["att"])


def test_plain_tracer_judges_the_cache(capi, oracle_mod, ctx):
    """judge_pixels: the plain tracer's own samples on seeded pixels judge
    the cached running mean.  A network that answers 0 (the frame is the
    prefix radiance alone, biased dark) is found out at the first frame it
    may judge (frame 19: from the third block of 8 frames, with 4 frames of
    the block in), and the sequence goes on with plain frames whose mean
    restarts, so the image ends near the plain reference.  With a bound no
    error reaches, the judge never takes over."""
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    plain = capi.Context(0)
    try:
        plain.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        ref = plain.trace_subset(capi.render_params(W, H, 4096, 7, eye, U, V, Wv, subframe=77),
                                 np.arange(W * H, dtype=np.int32), records=False)["accum"][:, :3]
    finally:
        plain.close()
    for bound, expect in ((0.01, True), (1e9, False)):
        cache = capi.Cache(0, seed=1)
        try:
            cache.set_weights(np.zeros(capi.CACHE_PARAMS, np.float32))
            cp = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0,
                                         judge_pixels=1024, judge_rmse=bound)
            took = None
            for f in range(40):
                st = capi.render_cached(ctx, cache, capi.render_params(W, H, 4, 7, eye, U, V, Wv, subframe=f),
                                        cp)
                if st["judge_plain"] and took is None:
                    took = f
                if not st["judge_plain"]:
                    assert st["judge_rmse_est"] >= 0
            img = ctx.copy_accum(W * H)[:, :3]
            e = float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))
            print(f"bound {bound}: plain frames from {took}, RMSE {e:.4f}", flush=True)
            if expect:
                assert took == 20
                assert e < 0.05
            else:
                assert took is None and e > 0.05
        finally:
            cache.close()


def test_cached_sequence_must_continue_frame0(capi, oracle_mod, ctx):
    """Frames after the first continue the sequence a cached frame 0 started:
    the judge's buffers are sized at frame 0, so a frame > 0 with another
    judge_pixels, after a plain render wrote the accum buffer, or with no
    cached frame 0 at all is refused before anything runs on the device."""
    W = H = 32
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    cache = capi.Cache(0, seed=1)
    try:
        cp = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0,
                                     judge_pixels=64, judge_rmse=1e9)
        p = lambda f: capi.render_params(W, H, 2, 7, eye, U, V, Wv, subframe=f)  # noqa: E731
        fresh = capi.Context(0)
        try:
            fresh.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
            with pytest.raises(capi.LisaError, match="frame > 0"):
                capi.render_cached(fresh, cache, p(1), cp)
        finally:
            fresh.close()
        capi.render_cached(ctx, cache, p(0), cp)
        capi.render_cached(ctx, cache, p(1), cp)
        more = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0,
                                       judge_pixels=128, judge_rmse=1e9)
        with pytest.raises(capi.LisaError, match="judge_pixels changed"):
            capi.render_cached(ctx, cache, p(2), more)
        none = capi.cache_frame_params(prefix_bounces=2, train_paths=0, train_steps=0)
        with pytest.raises(capi.LisaError, match="judge_pixels changed"):
            capi.render_cached(ctx, cache, p(2), none)
        capi.render_cached(ctx, cache, p(2), cp)  # the sequence itself goes on
        ctx.render(p(0))  # a plain render takes the accum buffer
        with pytest.raises(capi.LisaError, match="no cached frame 0"):
            capi.render_cached(ctx, cache, p(1), cp)
        capi.render_cached(ctx, cache, p(0), none)  # a new sequence without the judge
        capi.render_cached(ctx, cache, p(1), none)
        huge = capi.cache_frame_params(prefix_bounces=2, train_paths=(1 << 24) + 1, train_steps=1)
        with pytest.raises(capi.LisaError, match="train_paths"):
            capi.render_cached(ctx, cache, p(2), huge)
    finally:
        cache.close()


def test_full_paths_match_oracle(capi, oracle_mod, ctx):
    """LISA_CACHE_FULL_PATHS (the variance-reduced plain tracer, no network):
    without the estimators the per-pixel means are the plain tracer's bit for
    bit; with them (light samples, 8 / 1 NEE runs) they equal the oracle's
    orc_render_full_paths bit for bit, on the Cornell box and a small island."""
    from lisa_amd import scenes
    W = H = 64
    sc, o, (U, V, Wv) = _scene(capi, oracle_mod, ctx, W, H)
    eye = sc["camera"]["eye"]
    px = np.sort(np.random.default_rng(5).choice(W * H, 1024, replace=False)).astype(np.int32)
    p = capi.render_params(W, H, 8, 7, eye, U, V, Wv, subframe=2)
    cp = capi.cache_frame_params(prefix_bounces=2, flags=capi.CACHE_FULL_PATHS)
    g = capi.cache_trace_queries(ctx, p, cp, px, cap=0)
    plain = ctx.trace_subset(p, px, records=False)["accum"]
    assert g["n"] == 0
    assert np.array_equal(g["accum"].view(np.uint32), plain.view(np.uint32))
    isl = scenes.island(n=128, width=96, height=64)
    for sc_, W_, H_ in ((sc, W, H), (isl, 96, 64)):
        ctx.set_scene(sc_["verts"], sc_["normals"], sc_["mat_idx"], sc_["materials"])
        o_ = oracle_mod.Scene(sc_["verts"], sc_["normals"], sc_["mat_idx"], sc_["materials"])
        cam = sc_["camera"]
        U_, V_, W2 = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W_ / H_)
        px_ = np.sort(np.random.default_rng(6).choice(W_ * H_, 1024, replace=False)).astype(np.int32)
        p_ = capi.render_params(W_, H_, 8, 12, cam["eye"], U_, V_, W2, subframe=5)
        cp_ = capi.cache_frame_params(prefix_bounces=2, flags=capi.CACHE_FULL_PATHS | capi.CACHE_EMISSION,
                                      nee_loops=8, nee_loops_deep=1)
        g = capi.cache_trace_queries(ctx, p_, cp_, px_, cap=0)
        r = o_.render_full_paths(W_, H_, 8, 12, cam["eye"], U_, V_, W2, px_, subframe=5, threads=8,
                                 emit_expect=True, nee_loops=8, nee_loops_deep=1)
        assert np.array_equal(g["accum"].view(np.uint32), r.view(np.uint32))
