"""Parity of the traversal's private overflow stack (run as a script by
tests/test_gpu_stack_overflow.py with LISA_LIBRARY pointing at the build
whose LDS stack holds only 2 entries, so nearly every push of a traversal on
the 65k-triangle Cornell box goes to the overflow array): closest hits of
65,536 seeded rays and a 512-pixel x 4-spp path subset (per-sample records
and counters) equal the oracle's; through the 4-wide walk (forced on the
same scene), whose stack takes up to three entries per visit, the hits equal
the oracle's and a render equals the binary walk's bit for bit.  Prints OK on
success."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from lisa_amd import capi, scenes  # noqa: E402
from oracle import oracle  # noqa: E402
from tests.scene_util import random_rays  # noqa: E402


def main():
    assert os.path.samefile(capi.LIB_PATH, sys.argv[1]), capi.LIB_PATH
    sc = scenes.cornell_65k()
    o = oracle.Scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
    with capi.Context(0) as ctx:
        ctx.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        O, D = random_rays(1 << 16, seed=21)
        gp, gt = ctx.trace_rays(O, D)
        rp, rt, _ = o.closest_hit(O, D, threads=16)
        assert np.array_equal(gp, rp) and np.array_equal(gt.view(np.uint32), rt.view(np.uint32))
        print("hits ok", flush=True)
        W, H = 256, 256
        cam = sc["camera"]
        U, V, Wv = capi.camera_uvw(cam["eye"], cam["look_at"], cam["fov"], 1.0)
        px = scenes.seeded_pixels(W, H, 512, seed=22)
        p = capi.render_params(W, H, 4, 7, cam["eye"], U, V, Wv)
        g = ctx.trace_subset(p, px)
        r = o.render(W, H, 4, 7, cam["eye"], U, V, Wv, pixels=px, record=True, threads=16)
        for k in ("sample_bounces", "sample_occl", "sample_hash", "rad_hits"):
            assert np.array_equal(g[k], r[k]), k
        for k in ("sample_rgb", "accum"):
            a, b = g[k].astype(np.float64), r[k].astype(np.float64)
            assert np.all(np.abs(a - b) <= 1e-5 * np.abs(b) + 1e-30), k
        assert np.array_equal(g["counters"], r["counters"])
        print("binary walk ok", flush=True)
        p = capi.render_params(W, H, 4, 7, cam["eye"], U, V, Wv)
        want, _ = ctx.render(p)
    with capi.Context(0) as wide:
        wide.set_wide_threshold(0)
        wide.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        assert wide.stats()["wide_tree"]
        gp, gt = wide.trace_rays(O, D)
        assert np.array_equal(gp, rp) and np.array_equal(gt.view(np.uint32), rt.view(np.uint32))
        got, _ = wide.render(p)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        print("4-wide walk ok", flush=True)
    print("OK", flush=True)


if __name__ == "__main__":
    main()
