"""CPU: the density of the Lambertian sampler (`bsdf_pdf`, the interface's
pdf; lambertian.cu has none) against histograms of the sampler itself
(`bsdf_bounce`, lambertian.cu:7-13).

2M bounce directions per roughness are binned; the expected count of a bin
is the density integrated over it (24 x 24 midpoint sub-grid).  Bins are
latitude-longitude at roughness 1 and, below it, polar about the mirror
direction r (the bounce lies in the cone of half-angle asin(t / (1 - t))
around r when t < 1/2, where each direction has two preimages on the
hemisphere and the density has an integrable fold at the cone's edge; the
two outer rings are left out there).  A right density gives a chi-square
per bin near 1 (0.87-0.95 here); a density missing the (s / t)^2 distance
factor fails at every roughness below 1 (chi-square per bin 10^3-10^5), one
missing the second preimage fails at roughness 0.2 (197).
"""
import numpy as np
import pytest

N = np.array([0.3, 0.9, 0.2]) / np.linalg.norm([0.3, 0.9, 0.2])
D = np.array([0.5, -0.8, 0.1]) / np.linalg.norm([0.5, -0.8, 0.1])
SAMPLES = 2_000_000


def _counts(om, t, frame):
    v = om.bounce_many(D, N, 999, t, SAMPLES).astype(np.float64)
    w = v / np.linalg.norm(v, axis=1, keepdims=True)
    r = D - 2 * np.dot(N, D) * N
    if frame == "latlong":
        ax, ay, az, amax = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]), np.pi
        nb_a, nb_b, skip = 24, 48, 0
    else:
        az = r
        ax = np.cross(r, [1.0, 0, 0])
        ax /= np.linalg.norm(ax)
        ay = np.cross(az, ax)
        amax = np.arcsin(t / (1 - t)) if t < 0.5 else np.pi
        nb_a, nb_b, skip = 20, 24, (2 if t < 0.5 else 0)
    al = np.arccos(np.clip(w @ az, -1, 1))
    be = np.arctan2(w @ ay, w @ ax)
    H, _, _ = np.histogram2d(al, be, bins=[nb_a, nb_b], range=[[0, amax], [-np.pi, np.pi]])
    k = 24
    aa = (np.arange(nb_a * k) + 0.5) * amax / (nb_a * k)
    bb = -np.pi + (np.arange(nb_b * k) + 0.5) * 2 * np.pi / (nb_b * k)
    A, B = np.meshgrid(aa, bb, indexing="ij")
    L = (np.cos(A)[..., None] * az + np.sin(A)[..., None]
         * (np.cos(B)[..., None] * ax + np.sin(B)[..., None] * ay)).reshape(-1, 3)
    pdf = om.bsdf_pdf(D, N, L, t).astype(np.float64).reshape(A.shape)
    dA = np.sin(A) * (amax / (nb_a * k)) * (2 * np.pi / (nb_b * k))
    E = (pdf * dA).reshape(nb_a, k, nb_b, k).sum((1, 3)) * SAMPLES
    keep = E > 50
    if skip:
        keep[nb_a - skip:] = False
    return H[keep], E[keep], float((pdf * dA).sum())


@pytest.mark.parametrize("t,frame", [(1.0, "latlong"), (0.7, "latlong"), (0.6, "cone"), (0.4, "cone"),
                                     (0.2, "cone")])
def test_sampler_density_matches_its_histogram(oracle_mod, t, frame):
    h, e, total = _counts(oracle_mod, t, frame)
    chi = float(((h - e) ** 2 / e).sum() / h.size)
    print(f"roughness {t}: {h.size} bins, chi2 per bin {chi:.3f}, density integral {total:.4f}",
          flush=True)
    assert h.size > 80
    assert chi < 1.5
    assert abs(total - 1.0) < (0.005 if t >= 0.5 else 0.03)


def test_density_edge_cases(oracle_mod):
    L = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]], np.float32)
    up = np.array([0.0, 1.0, 0.0])
    d = np.array([0.0, -1.0, 0.0])
    # roughness 1: q = 1 / (12 max|w|^3) on N's side, 0 below it
    assert np.array_equal(oracle_mod.bsdf_pdf(d, up, L, 1.0), np.array([1 / 12, 0.0], np.float32))
    # a mirror has no density; roughness outside (0, 1] neither
    assert np.all(oracle_mod.bsdf_pdf(d, up, L, 0.0) == 0)
    assert np.all(oracle_mod.bsdf_pdf(d, up, L, 1.5) == 0)
