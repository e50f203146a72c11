"""Throughput of the plain path-tracing megakernel on one MI355X.

A step is one full render (subframe 0, every sample in one launch, as
render.cc:133-148) of the 65k-triangle Cornell box at 1024x1024, 256 spp,
7 bounces, with the scene and LBVH already resident in HBM.  The
traversal-bound 5.2M-triangle displaced icosphere (1920x1080, 64 spp,
12 bounces) is `--workload icosphere`.

    python bench.py [--steps K] [--warmup W] [--spp S] [--no-cpu] [--workload W]

Prints one JSON line.  `value` is rays traced (radiance + occlusion, exact
counts from the counting kernel, which is bit-identical in work to the
timed one) per second of wall time over the K timed steps.  `roofline` uses
the average kernel duration from HIP events recorded on the context's own
stream; `cpu_baseline` times the CPU oracle on a seeded pixel subset of the
same workload on this host's cores.  The default run also times the other
half of the metric, the learned cache against the plain tracer to RMSE <=
1e-2 on the island (`bench_rmse`), and reports it under `cache_vs_plain`
(`--no-cache-leg` skips it; `--workload island` prints it on its own).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s; time to RMSE≤1e-2 vs 16k-spp ref, cache vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes: one 64-byte node line per BVH node visit, one 48-byte
# triangle record per triangle test, 96 bytes (vertices + normals) per
# shaded radiance hit, one float4 accum store per pixel.
S_NODE, S_TRI, S_SHADE, S_PIXEL = 64, 48, 96, 16


# name -> (scene factory, default spp, default CPU-baseline pixel count)
WORKLOADS = {
    "cornell": ("cornell_small", 16, 65536),
    "cornell65k": ("cornell_65k", 256, 131072),
    "icosphere": ("displaced_icosphere", 64, 2048),
    "batch64": ("batch_scene", 32, 0),
    "island": ("island", 0, 0),
    "cornell_cache": ("cornell_65k", 0, 0),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-pixels", type=int, default=None)
    ap.add_argument("--workload", default="cornell65k", choices=sorted(WORKLOADS),
                    help="cornell: the 256^2 case (CPU oracle on the whole image); "
                         "cornell65k: the headline configuration; icosphere: the "
                         "traversal-bound 5.2M-triangle scene; batch64: the batched "
                         "wavefront form; island / cornell_cache: time to RMSE <= 1e-2, "
                         "learned cache vs plain")
    ap.add_argument("--cache-spp", type=int, default=4, help="spp per cached frame")
    ap.add_argument("--prefix", type=int, default=2, help="radiance rays before the cache")
    ap.add_argument("--train-paths", type=int, default=262144,
                    help="training paths per frame (their rows at the query depth: ~2^16 on "
                         "the island)")
    ap.add_argument("--train-steps", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3, help="the cache's Adam learning rate")
    ap.add_argument("--rmse-limit-s", type=float, default=30.0,
                    help="give up a method after this much render time")
    ap.add_argument("--persistent", action="store_true",
                    help="time the persistent launch (LISA_FLAG_PERSISTENT) of the plain path")
    ap.add_argument("--wavefront", action="store_true",
                    help="time the ray-level wavefront form (LISA_FLAG_WAVEFRONT) of the plain path")
    ap.add_argument("--cache-persistent", action="store_true",
                    help="cached frames' paths in one persistent launch (LISA_CACHE_PERSISTENT)")
    ap.add_argument("--cache-emission", action=argparse.BooleanOptionalAction, default=True,
                    help="prefix bounces' emission in expectation (LISA_CACHE_EMISSION)")
    ap.add_argument("--nee-loops", type=int, default=8,
                    help="with --cache-emission: NEE runs averaged at the camera ray's hit")
    ap.add_argument("--nee-loops-deep", type=int, default=1,
                    help="with --cache-emission: NEE runs averaged at later prefix vertices")
    ap.add_argument("--judge-pixels", type=int, default=0,
                    help="cached frames judged by the plain tracer on this many pixels "
                         "(plain frames once the bias estimate passes --judge-rmse)")
    ap.add_argument("--judge-rmse", type=float, default=0.02)
    ap.add_argument("--compare-emission", action="store_true",
                    help="island / cornell_cache: also time the cache with --cache-emission flipped")
    ap.add_argument("--cache-leg", action=argparse.BooleanOptionalAction, default=True,
                    help="default workload: also time the island's cache against plain "
                         "(time to RMSE <= 1e-2) and report it under cache_vs_plain")
    ap.add_argument("--hbm-leg", action=argparse.BooleanOptionalAction, default=True,
                    help="default workload: also time the 5.2M-triangle icosphere (two steps) "
                         "and report its roofline under roofline_hbm_case")
    ap.add_argument("--vr-spp", type=int, default=16,
                    help="spp per frame of the variance-reduced plain tracer in the time-to-RMSE "
                         "comparison (0: skip it)")
    ap.add_argument("--cache-overlap", action="store_true",
                    help="--cache-frames: train beside the frame's paths as the time-to-RMSE "
                         "runs do (default there: train first, so the training kernels are "
                         "timed alone)")
    ap.add_argument("--cache-frames", type=int, default=0,
                    help="island / cornell_cache: only run this many cached frames (no "
                         "reference, no plain sweep) and report their timings, for profiling")
    return ap.parse_args()


def profile_record(workload, spp, bvh):
    """The plain megakernel's rocprofv3 figures recorded in
    profiles/k_path_<workload>_pmc.json for the same spp and tree walk
    ("binary" or "4-wide"; a record without the key is of the binary walk), or {}:
    hbm_bytes_per_launch (two --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled
    by the copy-kernel calibration) and valu_lane_utilization
    (SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU))."""
    path = os.path.join(ROOT, "profiles", f"k_path_{workload}_pmc.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
        return rec if int(rec["spp"]) == spp and rec.get("bvh", "binary") == bvh else {}
    except (OSError, KeyError, ValueError):
        return {}


def bench_plain(a, workload, steps, warmup, cpu=True):
    """One plain-megakernel workload: exact ray counts from the counting
    instantiation, `steps` timed renders after `warmup`, the kernel's
    roofline (algorithmic bytes over the HIP-event kernel time; DRAM traffic
    from the committed rocprof counters) and, with `cpu`, the CPU oracle on a
    seeded pixel subset.  Returns (result dict, ctx triangle count)."""
    from lisa_amd import capi, scenes

    factory, def_spp, def_cpu_px = WORKLOADS[workload]
    sc = getattr(scenes, factory)()
    W, H, B = sc["width"], sc["height"], sc["bounces"]
    spp = (a.spp if workload == a.workload else None) or def_spp
    cpu_pixels = a.cpu_pixels or def_cpu_px
    cam = sc["camera"]
    U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
    ctx = capi.Context(0)
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    bvh_ms = ctx.stats()["bvh_build_ms"]
    wide = ctx.stats()["wide_tree"]  # scenes of >= 2^20 triangles: the 4-wide walk
    print(f"[bench] scene {sc['name']}: {ctx.T} triangles, LBVH build {bvh_ms:.2f} ms"
          + (" (with its 4-wide form)" if wide else ""), file=sys.stderr, flush=True)

    # exact work per step (counting instantiation of the same kernel)
    p = capi.render_params(W, H, spp, B, cam["eye"], U, V, Wv, flags=capi.FLAG_COUNT)
    ctx.render(p, want_accum=False)
    cnt = [int(x) for x in ctx.counters()]
    counts = dict(zip(capi.COUNTER_NAMES, cnt))
    rays = counts["rays_radiance"] + counts["rays_occlusion"]
    print(f"[bench] counted step: {counts}", file=sys.stderr, flush=True)

    p.flags = (capi.FLAG_PERSISTENT if a.persistent else 0) | (capi.FLAG_WAVEFRONT if a.wavefront else 0)
    for _ in range(warmup):
        ctx.render(p, want_accum=False)
    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.render(p, want_accum=False)
        kernel_ms.append(ctx.stats()["render_ms"])
        print(f"[bench] step {i}: kernel {kernel_ms[-1]:.1f} ms", file=sys.stderr, flush=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    avg_kernel_ms = sum(kernel_ms) / len(kernel_ms)

    alg_bytes = (S_NODE * counts["node_visits"] + S_TRI * counts["tri_tests"]
                 + S_SHADE * counts["radiance_hits"] + S_PIXEL * W * H)
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    rec = {} if (a.persistent or a.wavefront) else profile_record(workload, spp,
                                                                  "4-wide" if wide else "binary")
    traffic = rec.get("hbm_bytes_per_launch")
    resident = workload in ("cornell", "cornell65k")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "dram_frac": (None if traffic is None else
                              round(traffic / (avg_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
                # fraction of a wave's 64 lanes active per VALU instruction
                "lane_utilization": rec.get("valu_lane_utilization"),
                "limiter": ("instruction issue under divergence; the scene is L2 / Infinity "
                            "Cache resident, so `frac` counts bytes served mostly from cache"
                            if resident else
                            "instruction issue under divergence, with the scene's nodes and "
                            "triangles streamed from HBM"),
                "kernel": ("k_path_pq<0>" if a.persistent else
                           "k_wf_trace + k_wf_shade rounds" if a.wavefront else
                           "k_path<false,false,false,true> (4-wide walk)" if wide else
                           "k_path<false,false,false>"),
                "alg_bytes_per_launch": alg_bytes,
                "avg_kernel_ms": round(avg_kernel_ms, 3)}

    cpu_res = None
    if cpu:
        from oracle import oracle
        threads = len(os.sched_getaffinity(0))
        o = oracle.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        px = scenes.seeded_pixels(W, H, cpu_pixels, seed=7)
        t = time.perf_counter()
        r = o.render(W, H, spp, B, cam["eye"], U, V, Wv, pixels=px, threads=threads)
        dt = time.perf_counter() - t
        crays = int(r["counters"][0] + r["counters"][1])
        # one thread on a sixteenth of the sample (about the same wall time)
        px1 = px[:: max(threads, 1)]
        t = time.perf_counter()
        r1 = o.render(W, H, spp, B, cam["eye"], U, V, Wv, pixels=px1, threads=1)
        dt1 = time.perf_counter() - t
        crays1 = int(r1["counters"][0] + r1["counters"][1])
        whole = " (the whole image)" if len(px) == W * H else ""
        cpu_res = {"value": round(crays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                   "kind": "port",
                   "sample": f"{len(px)} seeded pixels of the same {W}x{H} render{whole} at "
                             f"{spp} spp x {B} bounces ({crays} rays, {dt:.1f} s, OpenMP over "
                             f"pixels)",
                   "samples_per_s": round(len(px) * spp / dt, 1),
                   "single_thread": {"value": round(crays1 / dt1 / 1e6, 3), "unit": "Mrays/s",
                                     "cores": 1,
                                     "sample": f"every {max(threads, 1)}th pixel of that sample "
                                               f"({len(px1)} pixels, {crays1} rays, {dt1:.1f} s)"}}
        print(f"[bench] cpu oracle: {cpu_res}", file=sys.stderr, flush=True)

    out = {"metric": METRIC, "value": round(rays * steps / wall / 1e6, 2), "unit": "Mrays/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"{sc['name']} plain "
                                  + ("wavefront form" if a.wavefront else "megakernel")
                                  + (" (persistent launch)" if a.persistent else ""),
                      "width": W, "height": H,
                      "spp": spp, "bounces": B, "triangles": ctx.T,
                      "bvh": "binary + 4-wide form" if wide else "binary",
                      "parallelism": "single GPU"},
           "samples_per_s": round(W * H * spp * steps / wall, 1),
           "rays_per_step": rays, "bvh_build_ms": round(bvh_ms, 3),
           "roofline": roofline, "cpu_baseline": cpu_res}
    ctx.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "more GPUs would shard the image by pixel tiles (no exchange during a render)",
              file=sys.stderr)
        sys.exit(2)

    if a.workload == "batch64":
        return bench_batch(a)
    if a.workload in ("island", "cornell_cache"):
        return bench_rmse(a)
    out = bench_plain(a, a.workload, a.steps, a.warmup, cpu=not a.no_cpu)
    if a.workload == "cornell65k" and a.hbm_leg:
        # the HBM-facing case: the 5.2M-triangle icosphere, whose nodes and
        # triangles stream from HBM (the headline scene is cache-resident)
        try:
            r = bench_plain(a, "icosphere", 2, 1, cpu=False)
            out["roofline_hbm_case"] = dict(r["roofline"], workload=r["config"]["workload"],
                                            mrays_per_s=r["value"], ms_per_step=r["ms_per_step"],
                                            triangles=r["config"]["triangles"])
        except Exception as e:  # the headline line is still reported
            out["roofline_hbm_case"] = {"error": f"{type(e).__name__}: {e}"}
            print(f"[bench] icosphere leg failed: {e!r}", file=sys.stderr, flush=True)
    if a.workload == "cornell65k" and a.cache_leg:
        # the metric's second half: time to RMSE <= 1e-2 on the island, cache vs plain
        try:
            r = bench_rmse(a, workload="island", emit=False)
            out["cache_vs_plain"] = {
                k: r[k] for k in ("metric", "plain_time_fit_s", "plain_progressive_time_s",
                                  "plain_time_s", "speedup_vs_plain_fit",
                                  "speedup_vs_plain_progressive", "speedup_vs_plain",
                                  "vr_plain_time_s", "vr_plain_frames", "speedup_vs_vr_plain",
                                  "cache_frames", "cache_rmse", "rmse_sample", "config",
                                  "mlp_roofline")}
            out["cache_vs_plain"]["cache_time_s"] = r["value"]
        except Exception as e:  # the plain line is still reported
            out["cache_vs_plain"] = {"error": f"{type(e).__name__}: {e}"}
            print(f"[bench] island leg failed: {e!r}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


RMSE_TARGET = 1e-2
REF_PIXELS = 65536
MLP_WEIGHTS = 21696  # 144*64 + 3*64*64 + 64*3 (no biases)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_rmse(a, workload=None, emit=True):
    """Time to RMSE <= 1e-2 against a 16k-spp plain render: learned cache vs plain.
    Returns the result line (printed unless emit is False); `workload`
    overrides `--workload`.

    Reference: 16 x 1024 spp with FP64 accumulation on 65536 seeded pixels
    (the plain loop on a pixel list; subframes 1000..1015 seed it apart from
    every measured render); the RMSE of an image is taken over those pixels
    (an unbiased estimate of the full-image mean squared error).  Plain: the
    spp sweep over powers of two, each a single full-frame render; its time
    is the first render that reaches the target.  Plain progressive: 16-spp
    subframes, time at the first subframe that reaches it.  Cache: frames of
    lisa_render_cached (online training + `--cache-spp` samples of paths cut
    after `--prefix` radiance rays, whose diffuse bounces' emission is taken
    in expectation unless `--no-cache-emission`; `--compare-emission` also
    times the other form), time summed over frames up to the first
    frame that reaches it (checked after every frame, like the progressive
    plain render).  `plain_time_fit_s` is the sweep's time scaled to the spp
    that reaches the target exactly, taking RMSE proportional to
    1/sqrt(spp) from the last sweep point above the target.  Times are wall
    time around the synchronous render calls; reading the image back for the
    RMSE is not timed."""
    import numpy as np
    from lisa_amd import capi, scenes

    workload = workload or a.workload
    factory = WORKLOADS[workload][0]
    sc = getattr(scenes, factory)()
    W, H = sc["width"], sc["height"]
    B = 12 if workload == "island" else sc["bounces"]
    cam = sc["camera"]
    U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], W / H)
    ctx = capi.Context(0)
    ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    if a.cache_frames > 0:
        return bench_cache_frames(a, capi, ctx, sc, W, H, B, U, V, Wv)
    pix = scenes.seeded_pixels(W, H, REF_PIXELS, seed=7)
    t = time.perf_counter()
    ref = np.zeros((REF_PIXELS, 3))
    for k in range(16):
        p = capi.render_params(W, H, 1024, B, cam["eye"], U, V, Wv, subframe=1000 + k,
                               flags=capi.FLAG_FP64_ACCUM)
        ref += ctx.trace_subset(p, pix, records=False)["accum"][:, :3].astype(np.float64) / 16
        print(f"[bench] reference chunk {k + 1}/16", file=sys.stderr, flush=True)
    ref_s = time.perf_counter() - t

    def rmse(img):
        return float(np.sqrt(np.mean((img[pix, :3].astype(np.float64) - ref) ** 2)))

    # plain: power-of-two spp sweep, one render each
    sweep = []
    plain_t = None
    spp = 1
    while spp <= 65536:
        p = capi.render_params(W, H, spp, B, cam["eye"], U, V, Wv)
        ctx.synchronize()
        t = time.perf_counter()
        acc, _ = ctx.render(p)
        dt = time.perf_counter() - t
        e = rmse(acc)
        sweep.append({"spp": spp, "s": round(dt, 4), "rmse": round(e, 6)})
        print(f"[bench] plain {spp} spp: {dt:.3f} s, rmse {e:.5f}", file=sys.stderr, flush=True)
        if e <= RMSE_TARGET:
            plain_t = dt
            break
        if dt > a.rmse_limit_s:
            break
        spp *= 2
    # plain progressive, 16 spp per subframe
    prog_t, tot = None, 0.0
    for f in range(100000):
        p = capi.render_params(W, H, 16, B, cam["eye"], U, V, Wv, subframe=f)
        t = time.perf_counter()
        ctx.render(p, want_accum=False)
        ctx.synchronize()
        tot += time.perf_counter() - t
        if rmse(ctx.copy_accum(W * H)) <= RMSE_TARGET:
            prog_t = tot
            break
        if tot > a.rmse_limit_s:
            break
    print(f"[bench] plain progressive: {prog_t}", file=sys.stderr, flush=True)
    # cache
    def run_cache(flags, frame_spp=None):
        frame_spp = frame_spp or a.cache_spp
        cache = capi.Cache(0, seed=1, lr=a.lr)
        cp = capi.cache_frame_params(prefix_bounces=a.prefix, train_paths=a.train_paths,
                                     train_bounces=B, train_max=1 << 16, train_steps=a.train_steps,
                                     seed=3, flags=flags,
                                     nee_loops=a.nee_loops if flags & capi.CACHE_EMISSION else 1,
                                     nee_loops_deep=a.nee_loops_deep if flags & capi.CACHE_EMISSION else 1,
                                     judge_pixels=a.judge_pixels, judge_rmse=a.judge_rmse)
        cache_t, tot, frames, last = None, 0.0, 0, None
        for f in range(1 << 20):
            p = capi.render_params(W, H, frame_spp, B, cam["eye"], U, V, Wv, subframe=f)
            t = time.perf_counter()
            last = capi.render_cached(ctx, cache, p, cp)
            tot += time.perf_counter() - t
            frames = f + 1
            if rmse(ctx.copy_accum(W * H)) <= RMSE_TARGET:
                cache_t = tot
                break
            if tot > a.rmse_limit_s:
                break
        e_cache = rmse(ctx.copy_accum(W * H))
        cache.close()
        print(f"[bench] cache (flags {flags}): {cache_t} s after {frames} frames, "
              f"rmse {e_cache:.5f}, last {last}", file=sys.stderr, flush=True)
        return cache_t, tot, frames, last, e_cache

    other = None
    if a.compare_emission:
        o_t, _, o_frames, _, o_e = run_cache(cache_flags(a, capi) ^ capi.CACHE_EMISSION)
        other = {"cache_emission": not a.cache_emission,
                 "time_s": None if o_t is None else round(o_t, 4), "frames": o_frames,
                 "rmse": round(o_e, 6)}
    cache_t, tot, frames, last, e_cache = run_cache(cache_flags(a, capi))
    # the same estimators on whole paths without the network (the
    # variance-reduced plain tracer): what the network itself gains
    vr_t, vr_frames = None, 0
    if a.vr_spp > 0 and a.cache_emission:
        vr_t, _, vr_frames, _, _ = run_cache(capi.CACHE_FULL_PATHS | capi.CACHE_EMISSION,
                                             frame_spp=a.vr_spp)
    mlp = mlp_roofline(last, a.train_steps, overlapped=True)
    fit_t = None
    above = [r for r in sweep if r["rmse"] > RMSE_TARGET]
    if plain_t is not None and above:
        r = above[-1]
        fit_t = r["s"] * (r["rmse"] / RMSE_TARGET) ** 2
    out = {"metric": "time to RMSE<=1e-2 vs 16k-spp plain reference (s)",
           "value": None if cache_t is None else round(cache_t, 4), "unit": "s", "n_gpus": 1,
           "steps": frames, "warmup": 0,
           "ms_per_step": round(tot * 1e3 / max(frames, 1), 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"{sc['name']} learned cache vs plain", "width": W, "height": H,
                      "bounces": B, "triangles": ctx.T, "cache_prefix_bounces": a.prefix,
                      "cache_spp_per_frame": a.cache_spp, "train_paths": a.train_paths,
                      "train_rows_max": 1 << 16, "train_steps": a.train_steps, "lr": a.lr,
                      "persistent_frames": a.cache_persistent,
                      "cache_emission": a.cache_emission,
                      "nee_loops": a.nee_loops if a.cache_emission else 1,
                      "nee_loops_deep": a.nee_loops_deep if a.cache_emission else 1,
                      "judge_pixels": a.judge_pixels, "judge_rmse": a.judge_rmse,
                      "parallelism": "single GPU"},
           "plain_time_s": None if plain_t is None else round(plain_t, 4),
           "plain_progressive_time_s": None if prog_t is None else round(prog_t, 4),
           "plain_time_fit_s": None if fit_t is None else round(fit_t, 4),
           "speedup_vs_plain_fit": (None if (fit_t is None or cache_t is None)
                                    else round(fit_t / cache_t, 3)),
           "speedup_vs_plain": (None if (plain_t is None or cache_t is None)
                                else round(plain_t / cache_t, 3)),
           "speedup_vs_plain_progressive": (None if (prog_t is None or cache_t is None)
                                            else round(prog_t / cache_t, 3)),
           "vr_plain_time_s": None if vr_t is None else round(vr_t, 4),
           "vr_plain_frames": vr_frames,
           "vr_plain_spp_per_frame": a.vr_spp,
           "speedup_vs_vr_plain": (None if (vr_t is None or cache_t is None)
                                   else round(vr_t / cache_t, 3)),
           "rmse_sample": (f"{REF_PIXELS} seeded pixels of {W * H}: the mean squared error over "
                           "them estimates the full image's without bias"),
           "cache_frames": frames, "cache_rmse": round(e_cache, 6), "cache_other_form": other,
           "plain_sweep": sweep,
           "reference": {"spp": 16384, "pixels": REF_PIXELS, "fp64": True,
                         "seconds": round(ref_s, 2)},
           "last_frame": last, "mlp_roofline": mlp, "roofline": None, "cpu_baseline": None}
    if emit:
        print(json.dumps(out), flush=True)
    ctx.close()
    return out


def mlp_roofline(st, train_steps, overlapped=False):
    """MFMA roofline of the cache MLP from one frame's HIP-event kernel times.
    Algorithmic FLOPs: 2 per real weight per query forward; 3x that per
    training row (forward, backward data pass, weight gradients).  Training
    that ran beside the frame's paths has no kernel time of its own (its
    events include the wait for the paths), so it is left out then."""
    if st is None or st["queries"] <= 0 or st["infer_kernel_ms"] <= 0:
        return None
    f_q = 2 * MLP_WEIGHTS
    inf_tf = st["queries"] * f_q / (st["infer_kernel_ms"] * 1e-3) / 1e12
    out = {"bound": "mfma", "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
           "infer_queries": st["queries"], "infer_kernel_ms": round(st["infer_kernel_ms"], 4),
           "infer_achieved": round(inf_tf, 2),
           "infer_frac": round(inf_tf / MFMA_BF16_PEAK_TFLOPS, 5)}
    if st["train_kernel_ms"] > 0 and not overlapped:
        tr_tf = (st["train_samples"] * 3 * f_q * train_steps / (st["train_kernel_ms"] * 1e-3)
                 / 1e12)
        out.update({"train_rows": st["train_samples"], "train_steps": train_steps,
                    "train_kernel_ms": round(st["train_kernel_ms"], 4),
                    "train_achieved": round(tr_tf, 2),
                    "train_frac": round(tr_tf / MFMA_BF16_PEAK_TFLOPS, 5)})
    return out


def cache_flags(a, capi, serial=False):
    return ((capi.CACHE_PERSISTENT if a.cache_persistent else 0)
            | (capi.CACHE_EMISSION if a.cache_emission else 0)
            | (capi.CACHE_SERIAL if serial else 0))


def bench_cache_frames(a, capi, ctx, sc, W, H, B, U, V, Wv):
    """`--cache-frames N`: N cached frames after `--warmup` untimed ones; per-frame
    wall time and the MLP's MFMA roofline (kernel times from HIP events)."""
    cam = sc["camera"]
    cache = capi.Cache(0, seed=1, lr=a.lr)
    serial = not a.cache_overlap  # by default the training kernels are timed alone here
    cp = capi.cache_frame_params(prefix_bounces=a.prefix, train_paths=a.train_paths,
                                 train_bounces=B, train_max=1 << 16, train_steps=a.train_steps,
                                 seed=3, flags=cache_flags(a, capi, serial),
                                 nee_loops=a.nee_loops if a.cache_emission else 1,
                                 nee_loops_deep=a.nee_loops_deep if a.cache_emission else 1)
    stats = []
    for f in range(a.warmup + a.cache_frames):
        p = capi.render_params(W, H, a.cache_spp, B, cam["eye"], U, V, Wv, subframe=f)
        t = time.perf_counter()
        st = capi.render_cached(ctx, cache, p, cp)
        st["wall_ms"] = (time.perf_counter() - t) * 1e3
        if f >= a.warmup:
            stats.append(st)
        print(f"[bench] frame {f}: {st['wall_ms']:.2f} ms, {st['queries']} queries, "
              f"{st['train_samples']} rows", file=sys.stderr, flush=True)
    med = sorted(stats, key=lambda x: x["infer_kernel_ms"])[len(stats) // 2]
    ms = sum(x["wall_ms"] for x in stats) / len(stats)
    out = {"metric": "cached frame time (ms)", "value": round(ms, 3), "unit": "ms",
           "n_gpus": 1, "steps": a.cache_frames, "warmup": a.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"{sc['name']} learned cache frames", "width": W, "height": H,
                      "bounces": B, "cache_prefix_bounces": a.prefix,
                      "cache_spp_per_frame": a.cache_spp, "train_paths": a.train_paths,
                      "train_steps": a.train_steps, "cache_emission": a.cache_emission,
                      "training_beside_paths": not serial,
                      "nee_loops": a.nee_loops if a.cache_emission else 1,
                      "nee_loops_deep": a.nee_loops_deep if a.cache_emission else 1,
                      "judge_pixels": a.judge_pixels, "judge_rmse": a.judge_rmse,
                      "parallelism": "single GPU"},
           "frame": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in med.items()},
           "mlp_roofline": mlp_roofline(med, a.train_steps, overlapped=not serial), "roofline": None,
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)
    cache.close()
    ctx.close()


def bench_batch(a):
    """64 seeded scenes (10k-2M triangles each, 256^2-1024^2, 32 spp, 8
    bounces) in one persistent graph-captured launch after autotuning,
    beside the same scenes rendered one lisa_render launch after another."""
    from lisa_amd import capi, scenes

    n = 64
    ctxs, params, rays, tris = [], [], 0, 0
    for i in range(n):
        sc = scenes.batch_scene(i)
        c = capi.Context(0)
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        cam = sc["camera"]
        U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"],
                                   sc["width"] / sc["height"])
        p = capi.render_params(sc["width"], sc["height"], a.spp or sc["spp"], sc["bounces"],
                               cam["eye"], U, V, Wv, flags=capi.FLAG_COUNT)
        c.render(p, want_accum=False)
        cnt = c.counters()
        rays += int(cnt[0] + cnt[1])
        tris += c.T
        p.flags = 0
        ctxs.append(c)
        params.append(p)
    print(f"[bench] batch: {n} scenes, {tris} triangles, {rays} rays per step",
          file=sys.stderr, flush=True)
    # one launch per scene, back to back
    for c, p in zip(ctxs, params):
        c.render(p, want_accum=False)
    t = time.perf_counter()
    for c, p in zip(ctxs, params):
        c.render(p, want_accum=False)
    seq_s = time.perf_counter() - t
    batch = capi.Batch(ctxs, params)
    # each form once at a fixed configuration, then the autotuner's pick
    forms = {}
    cus = 256
    for name, cfg in (("lane queue", (cus * 16, 1, 0)), ("wave tiles", (cus * 16, 1, 1)),
                      ("ray wavefront", (cus * 32, 2, 2))):
        batch.set_config(*cfg)
        batch.render()
        forms[name] = round(batch.render(), 2)
        print(f"[bench] form {name} {cfg}: {forms[name]} ms", file=sys.stderr, flush=True)
    cfg = batch.autotune(tune_spp=2)
    print(f"[bench] autotuned grid/chunk/form/graph: {cfg}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        batch.render()
    t0 = time.perf_counter()
    ms = [batch.render() for _ in range(a.steps)]
    wall = time.perf_counter() - t0
    out = {"metric": METRIC, "value": round(rays * a.steps / wall / 1e6, 2), "unit": "Mrays/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(wall * 1e3 / a.steps, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batch64: 64 scenes in one batched launch (hipGraph)", "scenes": n,
                      "triangles": tris, "spp": params[0].spp, "bounces": params[0].bounces,
                      "grid_blocks": cfg[0], "chunk": cfg[1],
                      "form": ("lane queue", "wave tiles", "ray wavefront")[cfg[2]],
                      "parallelism": "single GPU"},
           "rays_per_step": rays, "graph_replay_ms": [round(x, 2) for x in ms],
           "form_ms_at_fixed_config": forms,
           "sequential_plain_ms": round(seq_s * 1e3, 2),
           "speedup_vs_sequential_plain": round(seq_s / (wall / a.steps), 3)}
    print(json.dumps(out), flush=True)
    batch.close()
    for c in ctxs:
        c.close()


if __name__ == "__main__":
    main()
