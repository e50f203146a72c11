"""Pins for the CPU oracle's scalar functions (no GPU).

Each check is chosen so a plausible transcription mistake (swapped
operands, a dropped term, a wrong sign or index) fails it:
* RNG (random.h:31-67, maths.cu:6-8): golden vectors from the published TEA
  and Numerical Recipes LCG (tests/golden/make_rng_kat.py);
* vector algebra / BSDF (vec_math.h, maths.cu, lambertian.cu): closed forms
  and invariants -- Snell's law, total internal reflection, Schlick at
  normal and grazing incidence, reflect as an involution, 1/pi for BRDF(N,N);
* barycentric normal (maths.cu:33-56): vertex normals at the vertices,
  linear interpolation inside;
* camera (Camera.cpp:34-45): the closed-form frame;
* sRGB + PPM (helpers.h:35-66, sutil.cpp:97-117,523-600): sRGB curve values
  and the byte layout of a 2x2 image.
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rng_kat.json")


@pytest.fixture(scope="module")
def kat():
    with open(GOLD) as fh:
        return json.load(fh)


def test_tea16_golden(oracle_mod, kat):
    for a, b, want in kat["tea16"]:
        assert oracle_mod.tea16(a, b) == want


def test_lcg_golden(oracle_mod, kat):
    for seed, states in kat["lcg_states"]:
        vals, final = oracle_mod.lcg_seq(seed, len(states))
        assert vals == [s & 0xFFFFFF for s in states]
        assert final == states[-1]


def test_rnd_and_rng_are_exact_dyadics(oracle_mod, kat):
    for seed, want in kat["rnd"]:
        got, _ = oracle_mod.rnd_seq(seed, len(want))
        assert got == want  # exact: k / 2^24 is representable in FP32
        r, _ = oracle_mod.rng_seq(seed, len(want))
        assert r == [w * 2.0 - 1.0 for w in want]
        assert all(-1.0 <= x < 1.0 for x in r)


def test_reflect_involution_and_mirror(oracle_mod):
    rs = np.random.default_rng(0)
    for _ in range(100):
        n = rs.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        i = rs.normal(size=3).astype(np.float32)
        r = oracle_mod.reflect(i, n)
        # tangential part kept, normal part negated
        assert np.allclose(np.dot(r, n), -np.dot(i, n), atol=1e-5)
        assert np.allclose(r - np.dot(r, n) * n, i - np.dot(i, n) * n, atol=1e-5)
        assert np.allclose(oracle_mod.reflect(r, n), i, atol=1e-5)
    # axis-aligned: exact
    assert np.array_equal(oracle_mod.reflect([1, -1, 0], [0, 1, 0]), np.float32([1, 1, 0]))


def test_refract_snell_and_tir(oracle_mod):
    N = np.float32([0, 1, 0])
    for eta in (1 / 1.5, 1 / 1.33, 1.5):
        for th in np.linspace(0.05, 1.5, 12):
            d = np.float32([math.sin(th), -math.cos(th), 0])
            cosI = float(-np.dot(d, N))
            t = oracle_mod.refract(cosI, d, N, eta)
            sin_t2 = eta * eta * (1 - cosI * cosI)
            if sin_t2 >= 1.0:  # total internal reflection -> zero vector (maths.cu:29)
                assert np.all(t == 0.0)
            else:
                tn = t / np.linalg.norm(t)
                assert abs(np.linalg.norm(t) - 1.0) < 1e-5  # unit in, unit out
                assert abs(abs(tn[0]) - eta * math.sin(th)) < 1e-5  # eta sin(i) = sin(t)
                assert tn[1] < 0  # continues through the surface


def test_fresnel_schlick(oracle_mod):
    # fresnel(cos, eta): R0 at normal incidence, 1 at grazing, R0 = ((1-eta)/(1+eta))^2
    for eta in (1.5, 1 / 1.5, 1.33):
        r0 = ((1 - eta) / (1 + eta)) ** 2
        assert oracle_mod.fresnel(1.0, eta) == pytest.approx(r0, rel=1e-6)
        assert oracle_mod.fresnel(0.0, eta) == pytest.approx(1.0, rel=1e-6)
        c = 0.3
        assert oracle_mod.fresnel(c, eta) == pytest.approx(r0 + (1 - r0) * (1 - c) ** 5, rel=1e-5)
    assert oracle_mod.fresnel(1.0, 1.5) == pytest.approx(0.04, rel=1e-6)


def test_brdf_closed_forms(oracle_mod):
    N = np.float32([0, 0, 1])
    assert oracle_mod.brdf(N, N) == np.float32(np.float64(1.0) / math.pi)
    assert oracle_mod.brdf(N, -N) == 0.0  # clamped
    L = np.float32([0, math.sqrt(0.5), math.sqrt(0.5)])
    c = float(np.float32(np.dot(N, L)))
    assert oracle_mod.brdf(N, L) == pytest.approx(c * c / math.pi, rel=1e-6)


def test_hemisphere_side_unit_and_three_draws(oracle_mod):
    rs = np.random.default_rng(1)
    for k in range(200):
        n = rs.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        seed = int(rs.integers(0, 2**32))
        v, state = oracle_mod.hemisphere(n, seed)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-6
        assert np.dot(v, n) >= 0.0
        _, s3 = oracle_mod.lcg_seq(seed, 3)
        assert state == s3
        # direction is +-normalize(rng_x, rng_y, rng_z) of the next three draws
        r, _ = oracle_mod.rng_seq(seed, 3)
        u = np.array(r, np.float64)
        u /= np.linalg.norm(u)
        assert np.allclose(np.abs(v), np.abs(u), atol=1e-6)


def test_bounce_lerp_endpoints(oracle_mod):
    d = np.float32([0.6, -0.8, 0.0])
    N = np.float32([0, 1, 0])
    mirror, s0 = oracle_mod.bounce(d, N, 5, 0.0)
    assert np.allclose(mirror, [0.6, 0.8, 0.0], atol=1e-7)
    diffuse, s1 = oracle_mod.bounce(d, N, 5, 1.0)
    h, sh = oracle_mod.hemisphere(N, 5)
    # lerp(a, b, 1) = a + (b - a) is b up to rounding (vec_math.h:494-497)
    assert np.allclose(diffuse, h, atol=1e-7) and s1 == sh == s0


def test_bary_normal(oracle_mod):
    v1, v2, v3 = np.float32([0, 0, 0]), np.float32([1, 0, 0]), np.float32([0, 1, 0])
    n1, n2, n3 = np.float32([0, 0, 1]), np.float32([1, 0, 0]), np.float32([0, 1, 0])
    for v, n in ((v1, n1), (v2, n2), (v3, n3)):
        assert np.allclose(oracle_mod.bary_normal(v, n1, n2, n3, v1, v2, v3), n, atol=1e-6)
    c = (v1 + v2 + v3) / 3
    want = (n1 + n2 + n3) / np.linalg.norm(n1 + n2 + n3)
    assert np.allclose(oracle_mod.bary_normal(c, n1, n2, n3, v1, v2, v3), want, atol=1e-6)
    P = np.float32([0.25, 0.5, 0])  # u = 0.25, v = 0.25, w = 0.5
    w = 0.25 * n1 + 0.25 * n2 + 0.5 * n3
    assert np.allclose(oracle_mod.bary_normal(P, n1, n2, n3, v1, v2, v3), w / np.linalg.norm(w),
                       atol=1e-6)


def test_camera_frame_closed_form(oracle_mod, capi):
    eye, look, fov = (0.0, 0.0, 1.85), (0.0, 0.0, 0.0), 40.0
    for aspect in (1.0, 1920 / 1080):
        U, V, W = oracle_mod.camera_uvw(eye, look, fov, aspect)
        vlen = 1.85 * math.tan(math.radians(20))
        assert np.allclose(W, [0, 0, -1.85])
        assert np.allclose(V, [0, vlen, 0], rtol=1e-6)
        assert np.allclose(U, [vlen * aspect, 0, 0], rtol=1e-6)
        # the product's host code gives the same bits
        U2, V2, W2 = capi.camera_uvw(eye, look, fov, aspect)
        for a, b in ((U, U2), (V, V2), (W, W2)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _srgb(c):
    return 12.92 * c if c < 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055


def test_srgb_quantize(oracle_mod, capi):
    vals = np.float32([0.0, 1.0, 2.0, -1.0, 0.5, 0.002, 0.1, 0.25, 0.75, 0.9])
    rgba = np.zeros((len(vals), 4), np.float32)
    rgba[:, 0] = vals
    q = oracle_mod.srgb_quantize(rgba)
    for v, got in zip(vals, q[:, 0]):
        want = min(int(256 * _srgb(min(max(float(v), 0.0), 1.0))), 255)
        assert got == want
    assert np.all(q[:, 3] == 255)
    assert np.array_equal(capi.srgb_quantize(rgba), q)


def test_ppm_bytes(oracle_mod, capi, tmp_path):
    rgba8 = np.array([[1, 2, 3, 255], [4, 5, 6, 255],      # launch row y=0 (bottom)
                      [7, 8, 9, 255], [10, 11, 12, 255]], np.uint8)  # y=1 (top)
    want = b"P6\n2 2\n255\n" + bytes([7, 8, 9, 10, 11, 12, 1, 2, 3, 4, 5, 6])
    p1, p2 = str(tmp_path / "a.ppm"), str(tmp_path / "b.ppm")
    oracle_mod.save_ppm(p1, rgba8, 2, 2)
    capi.save_ppm(p2, rgba8, 2, 2)
    assert open(p1, "rb").read() == want
    assert open(p2, "rb").read() == want
    with pytest.raises(capi.LisaError):
        capi.save_ppm(str(tmp_path / "x.png"), rgba8, 2, 2)
