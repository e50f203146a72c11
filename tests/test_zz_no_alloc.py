"""GPU: lisa_update_vertices_noalloc, the refit for callers that cannot
allocate (its union tables live in a device array of the library), against
lisa_update_vertices on the same scene and vertices.

Two calls in a row, the second on a smaller scene: both LBVHs (Morton codes,
order, topology, every box) and a render after the update are bit-identical
to the allocating entry point's.  A scene above the array's 1,048,576
triangles is refused with a status before anything runs.  The vertices are
handed over in device memory the test allocates itself through the HIP
runtime the library uses.
"""
import numpy as np
import pytest

from tests.scene_util import DeviceArray, camera, cornell

pytestmark = pytest.mark.gpu


def _noalloc(capi):
    return capi.lib().lisa_update_vertices_noalloc


def _same_trees(a, b):
    for which in (0, 1):
        g, r = a.bvh(which), b.bvh(which)
        for key in ("morton", "prim", "refs"):
            assert np.array_equal(g[key], r[key]), (which, key)
        assert np.array_equal(g["boxes"].view(np.uint32), r["boxes"].view(np.uint32)), which


def test_noalloc_refit_equals_allocating_refit(capi):
    fn = _noalloc(capi)
    ref, dut = capi.Context(0), capi.Context(0)
    try:
        for seg, rings, seed in ((256, 128, 1), (32, 16, 2)):  # the second call: a smaller scene
            sc = cornell(seg, rings, 48, 40, 2, 7)
            for c in (ref, dut):
                c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
            rs = np.random.default_rng(seed)
            v2 = (sc["verts"] + rs.normal(scale=0.01, size=sc["verts"].shape)).astype(np.float32)
            n2 = sc["normals"][::-1].copy()
            ref.update_vertices(v2, n2)
            dv, dn = DeviceArray(v2), DeviceArray(n2)
            try:
                st = fn(dut.h, dv.ptr, dn.ptr)
                assert st == capi.OK, capi.lib().lisa_last_error(dut.h)
            finally:
                dv.free()
                dn.free()
            print(f"T {sc['verts'].shape[0]}: trees", flush=True)
            _same_trees(ref, dut)
            U, V, W = camera(sc, capi)
            p = capi.render_params(48, 40, 2, 7, sc["camera"]["eye"], U, V, W)
            a, _ = ref.render(p)
            b, _ = dut.render(p)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ref.close()
        dut.close()


def test_noalloc_refit_refuses_large_scenes(capi):
    fn = _noalloc(capi)
    sc = cornell(1024, 520, 16, 16, 1, 1)
    assert sc["verts"].shape[0] > (1 << 20)
    with capi.Context(0) as c:
        c.set_scene(sc["verts"], sc["normals"], sc["mat_idx"], sc["materials"])
        before = c.bvh(0)["boxes"].copy()
        dv = DeviceArray(sc["verts"])
        try:
            assert fn(c.h, dv.ptr, None) == capi.ERR_INVALID_ARG
        finally:
            dv.free()
        assert b"at most 1048576 triangles" in capi.lib().lisa_last_error(c.h)
        assert np.array_equal(c.bvh(0)["boxes"].view(np.uint32), before.view(np.uint32))
