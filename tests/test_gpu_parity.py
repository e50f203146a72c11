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
def test_emitter_bvh_bit_exact(capi, oracle_mod, ctx, seg, rings):
    sc = cornell(seg, rings, 64, 64, 1, 1)
    o = _pair(capi, oracle_mod, ctx, sc)
    g = ctx.bvh(which=1)
    r = o.bvh(which=1)
    assert len(r["prim"]) == 2
    for key in ("morton", "prim", "refs"):
        assert np.array_equal(g[key], r[key]), key
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


def test_path_subset_scaled_translated_and_island(capi, oracle_mod, ctx):
    """Per-sample parity away from the unit scene: the Cornell box scaled by 40
    and moved to (3000, -700, 250) (the emitter-query culls' margin scales
    with the largest coordinate, computed on the host from the BVH roots),
    and a small island (open scene, sun quad, dielectric water)."""
    base = cornell(32, 16, 128, 128, 8, 7)
    shift = np.array([3000.0, -700.0, 250.0])
    v = (np.asarray(base["verts"], np.float64).reshape(-1, 3) * 40.0 + shift).astype(np.float32)
    sc = dict(base, verts=v.reshape(np.asarray(base["verts"]).shape),
              camera=dict(base["camera"],
                          eye=tuple(np.asarray(base["camera"]["eye"]) * 40.0 + shift),
                          look_at=tuple(np.asarray(base["camera"]["look_at"]) * 40.0 + shift)))
    isl = scenes.island(n=128, width=96, height=64)
    for sc_, W, H in ((sc, 128, 128), (isl, 96, 64)):
        o = _pair(capi, oracle_mod, ctx, sc_)
        cam = sc_["camera"]
        U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
        px = scenes.seeded_pixels(W, H, 768, seed=4)
        p = capi.render_params(W, H, 8, 7, cam["eye"], U, V, Wv)
        g = ctx.trace_subset(p, px)
        r = o.render(W, H, 8, 7, cam["eye"], U, V, Wv, pixels=px, record=True, threads=8)
        assert int(r["counters"][4]) > 1000  # surface hits, so NEE runs
        _subset_compare(g, r)


def test_path_subset_displaced_icosphere(capi, oracle_mod, ctx):
    """The traversal-bound scene's kind at a smaller level: a displaced
    dielectric icosphere (level 6, 81,920 triangles) in the Cornell walls,
    12 bounces, so most paths refract through the glass many times."""
    sc = scenes.displaced_icosphere(level=6, width=96, height=64, spp=4, bounces=12)
    o = _pair(capi, oracle_mod, ctx, sc)
    cam = sc["camera"]
    U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], 96 / 64)
    px = scenes.seeded_pixels(96, 64, 1024, seed=6)
    p = capi.render_params(96, 64, 4, 12, cam["eye"], U, V, Wv)
    g = ctx.trace_subset(p, px)
    r = o.render(96, 64, 4, 12, cam["eye"], U, V, Wv, pixels=px, record=True, threads=8)
    assert int(r["counters"][0]) > 4 * 1024 * 3  # several radiance rays per sample
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


def test_hit_ids_edges_and_axis_rays(capi, oracle_mod, ctx):
    """Rays through shared edges / vertices of a triangle fan and axis-parallel
    rays: the device agrees with the oracle's brute-force loop bit for bit."""
    k = 8
    ang = np.linspace(0, 2 * np.pi, k + 1)
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1).astype(np.float32)
    v = np.concatenate([np.zeros((k, 3), np.float32), rim[:-1], rim[1:]], axis=1)
    n = np.tile(np.float32([0, 0, 1]), (k, 3))
    mats = [{"diffuse": (1, 1, 1), "alpha": 1.0, "roughness": 1.0}]
    ctx.set_scene(v, n, np.zeros(k, np.int32), mats)
    o = oracle_mod.Scene(v, n, np.zeros(k, np.int32), mats)
    pts = [np.float32([0, 0, 0])] + [(rim[i] * np.float32(s)).astype(np.float32)
                                     for i in range(k) for s in np.linspace(0.05, 0.95, 19)]
    P = np.array(pts, np.float32)
    for d in (np.float32([0, 0, -1]), np.float32([0.3, -0.2, -1]), np.float32([1e-3, 0, -1])):
        D = np.tile(d, (len(P), 1))
        gp, gt = ctx.trace_rays(P - d, D)
        rp, rt, _ = o.closest_hit(P - d, D, brute=True)
        assert np.all(gp >= 0)
        assert np.array_equal(gp, rp) and np.array_equal(gt.view(np.uint32), rt.view(np.uint32))


def _nan_tris(v):
    v = np.array(v, np.float32).reshape(-1, 3, 3)
    for i in range(v.shape[0]):
        v[i, : 1 + i % 3] = np.nan
    return v.reshape(-1, 9)


def test_edge_case_scenes(capi, oracle_mod, ctx):
    sc = cornell(32, 16, 32, 32, 4, 5)
    U, V, W = camera(sc, capi)
    eye = sc["camera"]["eye"]
    mats = sc["materials"]
    px = np.arange(32 * 32, dtype=np.int32)
    cases = {
        "empty": (np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros(0, np.int32)),
        "one": (sc["verts"][:1], sc["normals"][:1], sc["mat_idx"][:1]),
        "two": (sc["verts"][:2], sc["normals"][:2], sc["mat_idx"][:2]),
        # degenerate (zero-area) triangles mixed into the box
        "degenerate": (np.concatenate([sc["verts"], np.repeat(sc["verts"][:5, :3], 3, 1)]),
                       np.concatenate([sc["normals"], sc["normals"][:5]]),
                       np.concatenate([sc["mat_idx"], sc["mat_idx"][:5]])),
        # triangles with NaN vertices (one, two or all three) mixed into the box
        "nan": (np.concatenate([sc["verts"], _nan_tris(sc["verts"][:6])]),
                np.concatenate([sc["normals"], sc["normals"][:6]]),
                np.concatenate([sc["mat_idx"], sc["mat_idx"][:6]])),
    }
    for name, (v, n, m) in cases.items():
        ctx.set_scene(v, n, m, mats)
        o = oracle_mod.Scene(v, n, m, mats)
        p = capi.render_params(32, 32, 4, 5, eye, U, V, W)
        g = ctx.trace_subset(p, px)
        r = o.render(32, 32, 4, 5, eye, U, V, W, pixels=px, record=True, threads=4)
        print(name, flush=True)
        _subset_compare(g, r)
        acc, _ = ctx.render(p)
        assert _close(acc, r["accum"]), name


def test_progressive_subframes_and_determinism(capi, oracle_mod, ctx):
    sc = cornell(32, 16, 64, 64, 4, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    eye = sc["camera"]["eye"]
    prev = None
    for sub in range(3):  # shader.cu:160-164 running mean over subframes
        p = capi.render_params(64, 64, 4, 7, eye, U, V, W, subframe=sub)
        acc, _ = ctx.render(p)
        r = o.render(64, 64, 4, 7, eye, U, V, W, subframe=sub, accum_prev=prev, threads=8)
        assert _close(acc, r["accum"]), sub
        prev = r["accum"]
    p = capi.render_params(64, 64, 4, 7, eye, U, V, W)
    a1, _ = ctx.render(p)
    a2, _ = ctx.render(p)
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))  # run-to-run identical


def test_fp64_accumulation(capi, oracle_mod, ctx):
    sc = cornell(32, 16, 64, 64, 16, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    eye = sc["camera"]["eye"]
    px = np.arange(64 * 64, dtype=np.int32)
    p = capi.render_params(64, 64, 16, 7, eye, U, V, W, flags=capi.FLAG_FP64_ACCUM)
    acc, _ = ctx.render(p)
    g = ctx.trace_subset(capi.render_params(64, 64, 16, 7, eye, U, V, W), px)
    want = g["sample_rgb"].astype(np.float64).mean(axis=1)  # exact mean of the samples
    assert np.allclose(acc[:, :3], want, rtol=1e-6, atol=1e-7)


def test_cli_renders_rto(capi, oracle_mod, tmp_path):
    import subprocess
    from lisa_amd import scenes
    sc = scenes.cornell(32, 16, 48, 32, 4, 5)
    out = tmp_path / "out.ppm"
    rto = scenes.write_rto(sc, str(tmp_path / "scene"), output_image=str(out))
    res = subprocess.run([capi.CLI_PATH, "-s", rto], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Starting rendering..." in res.stdout and "Rendering finished in" in res.stdout
    f = capi.SceneFile(rto)
    assert np.array_equal(f.verts, sc["verts"]) and np.array_equal(f.mat_idx, sc["mat_idx"])
    cam = sc["camera"]
    U, V, W = oracle_mod.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], 48 / 32)
    o = oracle_mod.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    r = o.render(48, 32, 4, 5, cam["eye"], U, V, W, threads=4)
    ref = tmp_path / "ref.ppm"
    oracle_mod.save_ppm(str(ref), oracle_mod.srgb_quantize(r["accum"]), 48, 32)
    assert out.read_bytes() == ref.read_bytes()


def test_cli_cache_mode(capi, tmp_path):
    """`lisa -s scene.rto --cache`: progressive frames of the learned cache
    (4 samples each up to the scene's count) through the same CLI; the PPM
    and the float dump hold a finite, non-negative image of the right size
    whose mean is within 30 % of the plain render's."""
    import subprocess
    from lisa_amd import scenes
    sc = scenes.cornell(32, 16, 48, 32, 64, 5)
    out = tmp_path / "out.ppm"
    pfm = tmp_path / "out.pfm"
    rto = scenes.write_rto(sc, str(tmp_path / "scene"), output_image=str(out))
    res = subprocess.run([capi.CLI_PATH, "-s", rto, "--cache", "--dump-float", str(pfm)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert res.stdout.count("nb sample") == 16
    data = out.read_bytes()
    assert data.startswith(b"P6\n48 32\n255\n") and len(data) == len(b"P6\n48 32\n255\n") + 48 * 32 * 3
    raw = pfm.read_bytes()
    img = np.frombuffer(raw[raw.index(b"-1.0\n") + 5:], dtype="<f4").reshape(32, 48, 3)
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
    plain = tmp_path / "plain.pfm"
    res = subprocess.run([capi.CLI_PATH, "-s", rto, "--dump-float", str(plain)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = plain.read_bytes()
    ref = np.frombuffer(raw[raw.index(b"-1.0\n") + 5:], dtype="<f4")
    print("cache mean", img.mean(), "plain mean", ref.mean(), flush=True)
    assert abs(img.mean() - ref.mean()) <= 0.3 * ref.mean()


def test_refit_update_bit_exact(capi, oracle_mod, ctx):
    """lisa_update_vertices (refit with the build's topology) matches the
    oracle's refit bit for bit, and the path loop stays in parity after it."""
    sc = cornell(64, 32, 64, 64, 4, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    rs = np.random.default_rng(11)
    v2 = (sc["verts"] + rs.normal(scale=0.01, size=sc["verts"].shape)).astype(np.float32)
    n2 = sc["normals"][::-1].copy()
    ctx.update_vertices(v2, n2)
    o.update_vertices(v2, n2)
    for which in (0, 1):
        g, r = ctx.bvh(which), o.bvh(which)
        for key in ("morton", "prim", "refs"):
            assert np.array_equal(g[key], r[key]), (which, key)
        assert np.array_equal(g["boxes"].view(np.uint32), r["boxes"].view(np.uint32)), which
    U, V, W = camera(sc, capi)
    px = scenes.seeded_pixels(64, 64, 1024, seed=3)
    p = capi.render_params(64, 64, 4, 7, sc["camera"]["eye"], U, V, W)
    g = ctx.trace_subset(p, px)
    r = o.render(64, 64, 4, 7, sc["camera"]["eye"], U, V, W, pixels=px, record=True, threads=8)
    _subset_compare(g, r)


def test_batched_wavefront_bit_exact(capi, oracle_mod):
    """Three scenes of different sizes, spp and bounces in one persistent,
    graph-captured launch: every image equals lisa_render's bit for bit,
    for both forms (lanes pulling pixels, waves pulling tiles), several
    queue configurations and after autotuning."""
    specs = [(32, 16, 64, 48, 4, 7), (64, 32, 40, 72, 3, 5), (16, 8, 96, 40, 2, 9)]
    ctxs, params, refs = [], [], []
    for seg, rings, W, H, spp, B in specs:
        sc = cornell(seg, rings, W, H, spp, B)
        c = capi.Context(0)
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        U, V, Wv = camera(sc, capi)
        p = capi.render_params(W, H, spp, B, sc["camera"]["eye"], U, V, Wv)
        refs.append(c.render(p)[0])
        ctxs.append(c)
        params.append(p)
    batch = capi.Batch(ctxs, params)
    for cfg in [None, (64, 1), (1024, 4), (300, 16), (64, 1, 1), (1024, 1, 1), (7, 1, 1),
                (256, 4, 0, 0), (256, 1, 1, 0),
                # the ray-level wavefront form: slots for every pixel, few slots
                # (pixels refill freed slots), plain launches
                (4096, 2, 2), (16, 1, 2), (300, 3, 2, 0)]:
        if cfg:
            batch.set_config(*cfg)
        for rep in range(2):  # graph replays
            batch.render()
            for i in range(3):
                got = batch.accum(i)
                assert np.array_equal(got.view(np.uint32), refs[i].view(np.uint32)), (cfg, rep, i)
    best = batch.autotune(tune_spp=1)
    assert best[0] > 0 and best[1] > 0 and best[2] in (0, 1, 2)
    batch.render()
    for i in range(3):
        assert np.array_equal(batch.accum(i).view(np.uint32), refs[i].view(np.uint32))
    # a context whose scene, vertices or accum buffer changed makes the
    # captured launch stale: the batch refuses to run until re-created
    sc = cornell(*specs[1][:2], *specs[1][2:])
    for change in ("vertices", "scene", "accum"):
        b2 = capi.Batch(ctxs, params)
        b2.render()
        if change == "vertices":
            ctxs[1].update_vertices(sc["verts"], sc["normals"])
        elif change == "scene":
            ctxs[1].set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        else:
            p = params[1]
            ctxs[1].render(capi.render_params(8, 8, 1, 1, sc["camera"]["eye"], p.U, p.V, p.W))
            ctxs[1].render(p)
        with pytest.raises(capi.LisaError, match="re-create the batch"):
            b2.render()
        b2.close()
    batch.close()
    for c in ctxs:
        c.close()


def test_progressive_graph_same_images(capi, ctx):
    """lisa_render_progressive: runs of subframes as one graph launch leave
    the accum buffer bit-identical to the same subframes rendered one
    lisa_render call at a time -- the FP32 running mean and the FP64 sums,
    across runs of different lengths (the graph rebuilt), a run at another
    image size and a repeated run (the graph replayed with new subframes)."""
    W, H, spp, B = 40, 24, 3, 5
    sc = cornell(16, 8, W, H, spp, B)
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    other = capi.Context(0)
    try:
        other.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        for flags in (0, capi.FLAG_FP64_ACCUM):
            par = lambda sub, w=W, h=H: capi.render_params(w, h, spp, B, eye, U, V, Wv,  # noqa: E731
                                                           subframe=sub, flags=flags)
            seq = [ctx.render(par(k))[0] for k in range(9)]
            runs = [(0, 3), (3, 3), (6, 2), (8, 1)]
            for first, n in runs:
                got = other.render_progressive(par(first), n)
                want = seq[first + n - 1]
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (flags, first, n)
            # another size (a new graph), then the first size again from subframe 0
            small = other.render_progressive(par(0, 16, 8), 2)
            ctx.render(par(0, 16, 8))
            assert np.array_equal(small.view(np.uint32), ctx.render(par(1, 16, 8))[0].view(np.uint32))
            got = other.render_progressive(par(0), 3)
            assert np.array_equal(got.view(np.uint32), seq[2].view(np.uint32))
        with pytest.raises(capi.LisaError, match="n_subframes"):
            other.render_progressive(capi.render_params(W, H, spp, B, eye, U, V, Wv), 0)
        with pytest.raises(capi.LisaError, match="no flag but"):
            other.render_progressive(capi.render_params(W, H, spp, B, eye, U, V, Wv,
                                                        flags=capi.FLAG_COUNT), 2)
        with pytest.raises(capi.LisaError, match="another size"):
            other.render_progressive(capi.render_params(8, 8, spp, B, eye, U, V, Wv, subframe=1), 2)
    finally:
        other.close()


def test_batch_replay_ends_fp64_and_cached_sequences(capi):
    """A batch replay overwrites each context's accum buffer: the FP64 sums
    of an earlier LISA_FLAG_FP64_ACCUM render no longer describe it (saving
    them and an FP64 subframe 1 are refused), and a cached sequence started
    on the context cannot go on (its frame 1 is refused)."""
    W, H, spp, B = 32, 24, 2, 3
    sc = cornell(16, 8, W, H, spp, B)
    c = capi.Context(0)
    cache = capi.Cache(0, seed=1, lr=1e-3)
    try:
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        U, V, Wv = camera(sc, capi)
        eye = sc["camera"]["eye"]
        params = lambda sub=0, flags=0: capi.render_params(W, H, spp, B, eye, U, V, Wv,  # noqa: E731
                                                           subframe=sub, flags=flags)
        c.render(params(flags=capi.FLAG_FP64_ACCUM))
        assert c.accum64_save(W * H).shape == (W * H, 4)
        b = capi.Batch([c], [params()])
        try:
            b.render()
            with pytest.raises(capi.LisaError, match="no FP64 sums"):
                c.accum64_save(W * H)
            with pytest.raises(capi.LisaError, match="FP64 sums"):
                c.render(params(sub=1, flags=capi.FLAG_FP64_ACCUM))
            cp = capi.cache_frame_params(prefix_bounces=1, train_paths=256, train_bounces=B,
                                         train_max=256)
            capi.render_cached(c, cache, params(sub=0), cp)
            b.render()
            with pytest.raises(capi.LisaError, match="no cached frame 0"):
                capi.render_cached(c, cache, params(sub=1), cp)
        finally:
            b.close()
    finally:
        cache.close()
        c.close()


def test_converged_within_3_sigma(capi, oracle_mod, ctx):
    """Independent seeds: the GPU's 256-spp pixel means and the oracle's
    (another subframe, so another RNG stream) agree within the per-pixel
    3-sigma Monte-Carlo bound, sigma from each side's own per-sample spread.
    Pixels with zero spread on both sides must agree exactly."""
    W = H = 64
    spp = 256
    sc = cornell(32, 16, W, H, spp, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    px = np.arange(W * H, dtype=np.int32)
    g = ctx.trace_subset(capi.render_params(W, H, spp, 7, eye, U, V, Wv, subframe=0), px,
                         record_bounces=0)
    r = o.render(W, H, spp, 7, eye, U, V, Wv, subframe=1, pixels=px, record=True,
                 record_bounces=0, threads=8)
    m1 = g["sample_rgb"].astype(np.float64).mean(1)
    m2 = r["sample_rgb"].astype(np.float64).mean(1)
    s1 = g["sample_rgb"].astype(np.float64).std(1, ddof=1) / np.sqrt(spp)
    s2 = r["sample_rgb"].astype(np.float64).std(1, ddof=1) / np.sqrt(spp)
    sig = np.sqrt(s1 ** 2 + s2 ** 2)
    flat = sig == 0
    assert np.array_equal(m1[flat], m2[flat])
    z = np.abs(m1 - m2)[~flat] / sig[~flat]
    frac = float(np.mean(z > 3))
    print("3-sigma outliers", frac, "of", z.size, flush=True)
    assert frac < 0.01


def test_rmse_against_16k_spp_oracle(capi, oracle_mod, ctx):
    """RMSE of GPU renders (64 .. 1024 spp) against a 16384-spp oracle render
    of 128 seeded pixels with an independent seed: it falls as 1/sqrt(spp)
    within the bound set by the measured per-sample variance."""
    W = H = 64
    sc = cornell(32, 16, W, H, 16, 7)
    o = _pair(capi, oracle_mod, ctx, sc)
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    px = np.sort(np.random.default_rng(17).choice(W * H, 128, replace=False)).astype(np.int32)
    ref = o.render(W, H, 16384, 7, eye, U, V, Wv, subframe=99, pixels=px, threads=8)["accum"]
    ref = ref[:, :3].astype(np.float64)
    var = None
    errs = {}
    for spp in (64, 256, 1024):
        g = ctx.trace_subset(capi.render_params(W, H, spp, 7, eye, U, V, Wv, subframe=0), px,
                             record_bounces=0)
        if var is None:
            var = float(np.mean(g["sample_rgb"].astype(np.float64).var(1, ddof=1)))
        errs[spp] = float(np.sqrt(np.mean((g["accum"][:, :3] - ref) ** 2)))
    print("RMSE vs 16k-spp oracle:", errs, "per-sample variance", var, flush=True)
    for spp, e in errs.items():
        expect = np.sqrt(var / spp + var / 16384)
        assert 0.5 * expect < e < 1.6 * expect, (spp, e, expect)
    assert errs[1024] < errs[64] / 2.5


def test_persistent_launch_same_image(capi, oracle_mod, ctx):
    """LISA_FLAG_PERSISTENT (lanes pull pixels from a queue) renders the
    same image as the tiled launch, bit for bit, on both progressive
    subframes and a partial edge tile (100 x 60)."""
    sc = cornell(32, 16, 100, 60, 4, 7)
    _pair(capi, oracle_mod, ctx, sc)
    U, V, W = camera(sc, capi)
    eye = sc["camera"]["eye"]
    plain = capi.Context(0)
    try:
        plain.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        for f in range(2):
            p = capi.render_params(100, 60, 4, 7, eye, U, V, W, subframe=f)
            want, _ = plain.render(p)
            p.flags = capi.FLAG_PERSISTENT
            got, _ = ctx.render(p)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f
        p = capi.render_params(100, 60, 4, 7, eye, U, V, W, flags=capi.FLAG_PERSISTENT | capi.FLAG_COUNT)
        with pytest.raises(capi.LisaError):
            ctx.render(p)
    finally:
        plain.close()


def test_wavefront_form_same_image(capi, oracle_mod, ctx):
    """LISA_FLAG_WAVEFRONT (path states in HBM, rounds of the traversal
    kernel with per-wave request queues and the shading kernel with ballot /
    prefix-sum compaction) renders every image bit for bit as the
    megakernel: partial edge tiles, progressive subframes, an open scene with
    dielectric water, a displaced glass icosphere at 12 bounces, and the
    edge-case scenes (empty, one triangle, NaN vertices)."""
    isl = scenes.island(n=128, width=96, height=64)
    ico = scenes.displaced_icosphere(level=5, width=72, height=40, spp=3, bounces=12)
    base = cornell(32, 16, 100, 60, 4, 7)
    cases = [("cornell", base, 100, 60, 4, 7), ("island", isl, 96, 64, 4, 12),
             ("icosphere", ico, 72, 40, 3, 12)]
    mats = base["materials"]
    cases.append(("empty", dict(base, verts=np.zeros((0, 9), np.float32),
                                normals=np.zeros((0, 9), np.float32),
                                mat_idx=np.zeros(0, np.int32)), 40, 24, 2, 5))
    cases.append(("nan", dict(base, verts=np.concatenate([base["verts"], _nan_tris(base["verts"][:6])]),
                              normals=np.concatenate([base["normals"], base["normals"][:6]]),
                              mat_idx=np.concatenate([base["mat_idx"], base["mat_idx"][:6]])),
                  40, 24, 2, 5))
    plain = capi.Context(0)
    try:
        for name, sc, W, H, spp, B in cases:
            for c in (ctx, plain):
                c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc.get("materials", mats))
            cam = sc["camera"]
            U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
            for sub in range(2):
                p = capi.render_params(W, H, spp, B, cam["eye"], U, V, Wv, subframe=sub)
                want, _ = plain.render(p)
                p.flags = capi.FLAG_WAVEFRONT
                got, _ = ctx.render(p)
                print(name, sub, flush=True)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, sub)
        p = capi.render_params(8, 8, 1, 1, (0, 0, 1), U, V, Wv, flags=capi.FLAG_WAVEFRONT | capi.FLAG_COUNT)
        with pytest.raises(capi.LisaError):
            ctx.render(p)
    finally:
        plain.close()


def test_fp64_progressive_and_resume(capi, oracle_mod, ctx):
    """LISA_FLAG_FP64_ACCUM across subframes: subframe k's accum is the exact
    FP64 mean of every sample of subframes 0..k (the kernel's per-pixel FP64
    sums in sample order, checked bit for bit against the same sums of the
    per-sample records); lisa_accum64_save after subframe 1 and
    lisa_accum64_load into another context make its subframe 2 bit-identical
    to the uninterrupted sequence (a resumable reference render)."""
    W, H, spp = 48, 40, 8
    sc = cornell(32, 16, W, H, spp, 7)
    _pair(capi, oracle_mod, ctx, sc)
    U, V, Wv = camera(sc, capi)
    eye = sc["camera"]["eye"]
    px = np.arange(W * H, dtype=np.int32)
    sums = np.zeros((W * H, 3))
    n = 0
    saved = None
    accs = []
    for sub in range(3):
        rec = ctx.trace_subset(capi.render_params(W, H, spp, 7, eye, U, V, Wv, subframe=sub), px,
                               record_bounces=0)["sample_rgb"].astype(np.float64)
        part = np.zeros((W * H, 3))
        for k in range(spp):  # the kernel's order: samples in turn, then onto the sums so far
            part = part + rec[:, k]
        sums = part if sub == 0 else sums + part
        n += spp
        p = capi.render_params(W, H, spp, 7, eye, U, V, Wv, subframe=sub, flags=capi.FLAG_FP64_ACCUM)
        acc, _ = ctx.render(p)
        want = (sums * (1.0 / n)).astype(np.float32)
        assert np.array_equal(acc[:, :3].view(np.uint32), want.view(np.uint32)), sub
        accs.append(acc)
        if sub == 1:
            saved = ctx.accum64_save(W * H)
            assert np.all(saved[:, 3] == 2 * spp)
    other = capi.Context(0)
    try:
        other.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        p = capi.render_params(W, H, spp, 7, eye, U, V, Wv, subframe=2, flags=capi.FLAG_FP64_ACCUM)
        other.render(capi.render_params(W, H, spp, 7, eye, U, V, Wv))  # a plain render: no FP64 sums
        with pytest.raises(capi.LisaError, match="FP64 sums"):
            other.render(p)
        other.accum64_load(saved)
        assert np.array_equal(other.copy_accum(W * H).view(np.uint32), accs[1].view(np.uint32))
        acc, _ = other.render(p)
        assert np.array_equal(acc.view(np.uint32), accs[2].view(np.uint32))
        assert np.array_equal(other.accum64_save(W * H), ctx.accum64_save(W * H))
    finally:
        other.close()
