"""Synthetic scenes for LiSA's `.rto` + `.obj` input (seeded, deterministic).

The reference ships no scene assets (its README example, README.md:45-131,
names .obj files that are absent), so every workload here is generated:

* ``cornell(sphere_seg, sphere_rings)`` -- unit Cornell box [-0.5, 0.5]^3
  with five Lambertian walls (white floor/ceiling/back, red left, green right,
  open at +z), a tall white Lambertian box, a small dielectric box (n = 1.5),
  a white UV sphere and a 2-triangle ceiling quad light emitting (1, 1, 1);
  camera and materials as in README.md:45-131.  32x16 segments gives ~1k
  triangles, 256x128 gives 65k.
* ``displaced_icosphere(level)`` -- an icosphere (20 * 4^level triangles,
  level 9 = 5,242,880) displaced by 3-octave value noise (amplitude 0.1),
  dielectric n = 1.5, inside the Cornell walls under the quad light.

A scene is a dict holding the de-indexed triangle soup exactly as
``parse_obj`` builds it (src/LiSA/src/parse_obj.cc:24-69): ``verts`` and
``normals`` float32 [T, 9], ``mat_idx`` int32 [T], plus ``materials`` (dicts
with the fields of ``Material``, src/LiSA/include/structs.hh:16-23), camera
and render settings.  ``write_rto`` writes the same scene as `.rto` + `.obj`
files whose parse gives the identical soup (floats printed with 9
significant digits round-trip float32 exactly).
"""
from __future__ import annotations

import os

import numpy as np

WHITE = {"name": "white", "diffuse": (1.0, 1.0, 1.0), "alpha": 1.0, "roughness": 1.0, "n": 1.0}
GLASS = {"name": "transparent", "diffuse": (1.0, 1.0, 1.0), "alpha": 0.0, "roughness": 0.0,
         "n": 1.5}
RED = {"name": "red", "diffuse": (1.0, 0.0, 0.0), "alpha": 1.0, "roughness": 1.0, "n": 1.0}
GREEN = {"name": "green", "diffuse": (0.0, 1.0, 0.0), "alpha": 1.0, "roughness": 1.0, "n": 1.0}
LIGHT = {"name": "light", "diffuse": (0.0, 0.0, 0.0), "alpha": 1.0, "emit": True,
         "emission": (1.0, 1.0, 1.0)}

CAMERA = {"eye": (0.0, 0.0, 1.85), "look_at": (0.0, 0.0, 0.0), "fov": 40.0}


def _quad(a, b, c, d, n):
    """Two triangles (a, b, c), (a, c, d) with constant normal n."""
    v = np.array([[a, b, c], [a, c, d]], np.float32).reshape(2, 9)
    nn = np.tile(np.asarray(n, np.float32), 6).reshape(2, 9)
    return v, nn


def _box(lo, hi):
    """Axis-aligned box, 12 triangles, outward face normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-1, 0, 0)),
        ((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0)),
        ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
        ((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (0, 1, 0)),
        ((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -1)),
        ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1)),
    ]
    vs, ns = zip(*[_quad(*f) for f in faces])
    return np.concatenate(vs), np.concatenate(ns)


def _uv_sphere(center, radius, seg, rings):
    """UV sphere: `seg` segments around, `rings` rings pole to pole, smooth normals.

    Per ring i and segment j, triangles (a, c, b) unless i is the top ring and
    (a, d, c) unless i is the bottom ring, in that order."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, seg + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    pos = np.asarray(center, np.float64) + radius * nrm
    i, j = np.meshgrid(np.arange(rings), np.arange(seg), indexing="ij")
    corner = {"a": (i, j), "b": (i, j + 1), "c": (i + 1, j + 1), "d": (i + 1, j)}
    def tri(arr, k1, k2, k3):
        return np.stack([arr[corner[k][0], corner[k][1]] for k in (k1, k2, k3)], 2)
    tv = np.stack([tri(pos, "a", "c", "b"), tri(pos, "a", "d", "c")], 2)  # [rings, seg, 2, 3, 3]
    tn = np.stack([tri(nrm, "a", "c", "b"), tri(nrm, "a", "d", "c")], 2)
    keep = np.ones((rings, seg, 2), bool)
    keep[0, :, 0] = False
    keep[rings - 1, :, 1] = False
    v = tv[keep].astype(np.float32).reshape(-1, 9)
    n = tn[keep].astype(np.float32).reshape(-1, 9)
    return v, n


def _walls_and_light():
    h = 0.5
    parts = [
        # floor, ceiling, back: white; left red; right green (inward normals)
        (_quad((-h, -h, -h), (h, -h, -h), (h, -h, h), (-h, -h, h), (0, 1, 0)), "white"),
        (_quad((-h, h, -h), (-h, h, h), (h, h, h), (h, h, -h), (0, -1, 0)), "white"),
        (_quad((-h, -h, -h), (-h, h, -h), (h, h, -h), (h, -h, -h), (0, 0, 1)), "white"),
        (_quad((-h, -h, -h), (-h, -h, h), (-h, h, h), (-h, h, -h), (1, 0, 0)), "red"),
        (_quad((h, -h, -h), (h, h, -h), (h, h, h), (h, -h, h), (-1, 0, 0)), "green"),
    ]
    ly, s = 0.499, 0.125
    light = (_quad((-s, ly, -s), (-s, ly, s), (s, ly, s), (s, ly, -s), (0, -1, 0)), "light")
    return parts, light


def _assemble(groups, materials):
    names = [m["name"] for m in materials]
    vs, ns, ms, meshes = [], [], [], []
    for (v, n), mname in groups:
        vs.append(v)
        ns.append(n)
        ms.append(np.full(v.shape[0], names.index(mname), np.int32))
        meshes.append((mname, v, n))
    return (np.ascontiguousarray(np.concatenate(vs)), np.ascontiguousarray(np.concatenate(ns)),
            np.ascontiguousarray(np.concatenate(ms)), meshes)


def _scene(groups, materials, width, height, spp, bounces, name):
    v, n, m, meshes = _assemble(groups, materials)
    return {"name": name, "verts": v, "normals": n, "mat_idx": m,
            "materials": [dict(x) for x in materials], "meshes": meshes,
            "camera": dict(CAMERA), "width": width, "height": height,
            "spp": spp, "bounces": bounces}


def cornell(sphere_seg=32, sphere_rings=16, width=256, height=256, spp=16, bounces=7):
    """Cornell box scene; defaults are the 256^2 x 16 spp x 7 bounce small case."""
    materials = [WHITE, GLASS, RED, GREEN, LIGHT]
    walls, light = _walls_and_light()
    groups = list(walls)
    groups.append((_box((-0.35, -0.5, -0.35), (-0.05, 0.1, -0.05)), "white"))
    groups.append((_box((0.05, -0.5, 0.0), (0.35, -0.2, 0.3)), "transparent"))
    groups.append((_uv_sphere((0.22, -0.35, -0.22), 0.15, sphere_seg, sphere_rings), "white"))
    groups.append(light)
    return _scene(groups, materials, width, height, spp, bounces,
                  f"cornell_{sphere_seg}x{sphere_rings}")


def cornell_small():
    """Cornell box, 32x16 sphere (~1k triangles), 256x256, 16 spp, 7 bounces."""
    return cornell(32, 16, 256, 256, 16, 7)


def cornell_65k():
    """Cornell box, 256x128 sphere (~65k triangles), 1024x1024, 256 spp, 7 bounces."""
    return cornell(256, 128, 1024, 1024, 256, 7)


# ---------------------------------------------------------------------------
# Seeded value noise (splitmix64 lattice hash, smoothstep trilinear blend).
# ---------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


def _lattice(ix, iy, iz, seed):
    with np.errstate(over="ignore"):
        k = (ix.astype(np.int64).astype(np.uint64) * np.uint64(73856093)
             ^ iy.astype(np.int64).astype(np.uint64) * np.uint64(19349663)
             ^ iz.astype(np.int64).astype(np.uint64) * np.uint64(83492791)
             ^ np.uint64(seed) * np.uint64(0x632BE59BD9B4E019))
        h = _splitmix64(k)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def value_noise(p, freq, seed):
    q = p * freq
    i0 = np.floor(q)
    f = q - i0
    w = f * f * (3.0 - 2.0 * f)
    i0 = i0.astype(np.int64)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = _lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += val * wx * wy * wz
    return out


def fbm(p, octaves=3, base_freq=4.0, seed=0):
    total, amp, norm = np.zeros(p.shape[0]), 1.0, 0.0
    for o in range(octaves):
        total += amp * value_noise(p, base_freq * (2 ** o), seed + o)
        norm += amp
        amp *= 0.5
    return total / norm


def icosphere(level):
    """Unit icosphere as (vertices float64 [V,3], faces int64 [F,3]); F = 20 * 4^level."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (1 << 32) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk >> 32, uk & 0xFFFFFFFF
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1) + base
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _flat_soup(pos, faces, center):
    tri = pos[faces]  # [F,3,3]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    flip = np.einsum("ij,ij->i", n, tri.mean(1) - center) < 0
    n[flip] *= -1.0
    v = tri.astype(np.float32).reshape(-1, 9)
    nn = np.repeat(n[:, None, :], 3, axis=1).astype(np.float32).reshape(-1, 9)
    return v, nn


def displaced_icosphere(level=9, width=1920, height=1080, spp=64, bounces=12, seed=0,
                        amplitude=0.1, radius=0.3, center=(0.0, -0.15, -0.05)):
    """Displaced icosphere (dielectric n=1.5) inside the Cornell walls under the light."""
    materials = [WHITE, GLASS, RED, GREEN, LIGHT]
    u, f = icosphere(level)
    r = radius * (1.0 + amplitude * fbm(u, octaves=3, base_freq=4.0, seed=seed))
    c = np.asarray(center, np.float64)
    pos = c + u * r[:, None]
    walls, light = _walls_and_light()
    groups = list(walls)
    groups.append((_flat_soup(pos, f, c), "transparent"))
    groups.append(light)
    return _scene(groups, materials, width, height, spp, bounces, f"icosphere_l{level}")


# ---------------------------------------------------------------------------
# .rto / .obj writer (README.md:45-131 syntax; parse_obj.cc face format).
# ---------------------------------------------------------------------------
def _fmt(x):
    return "%.9g" % float(x)


def write_obj(path, v, n):
    """De-indexed OBJ: 3 `v` + 3 `vn` per triangle, faces `f a//a b//b c//c`."""
    T = v.shape[0]
    vv = v.reshape(-1, 3)
    nn = n.reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write("".join(f"v {_fmt(a)} {_fmt(b)} {_fmt(c)}\n" for a, b, c in vv))
        fh.write("".join(f"vn {_fmt(a)} {_fmt(b)} {_fmt(c)}\n" for a, b, c in nn))
        fh.write("".join(f"f {3*k+1}//{3*k+1} {3*k+2}//{3*k+2} {3*k+3}//{3*k+3}\n"
                         for k in range(T)))


def write_rto(scene, directory, output_image="out.ppm"):
    """Write `scene` as directory/scene.rto + one .obj per mesh; returns the .rto path.

    Mesh paths are written relative to the current directory when the renderer
    runs (as parse_obj opens them), so they are emitted as absolute paths.
    """
    os.makedirs(directory, exist_ok=True)
    directory = os.path.abspath(directory)
    lines = []
    for m in scene["materials"]:
        lines.append(f"material {m['name']} {{")
        if m.get("emit"):
            e = m["emission"]
            lines.append("  emit = true")
            lines.append(f"  color = ({_fmt(e[0])}, {_fmt(e[1])}, {_fmt(e[2])}, 1)")
        else:
            d = m["diffuse"]
            lines.append(f"  color = ({_fmt(d[0])}, {_fmt(d[1])}, {_fmt(d[2])}, "
                         f"{_fmt(m['alpha'])})")
            if m["alpha"] < 1.0:
                lines.append(f"  n = {_fmt(m['n'])}")
            else:
                lines.append(f"  roughness = {_fmt(m['roughness'])}")
        lines.append("}")
        lines.append("")
    for i, (mname, v, n) in enumerate(scene["meshes"]):
        obj = os.path.join(directory, f"mesh{i:03d}.obj")
        write_obj(obj, v, n)
        lines += ["mesh {", f"  obj_file = {obj}", f"  material = {mname}", "}", ""]
    cam = scene["camera"]
    e, l = cam["eye"], cam["look_at"]
    lines += ["camera {",
              f"  position = ({_fmt(e[0])}, {_fmt(e[1])}, {_fmt(e[2])})",
              f"  look_at = ({_fmt(l[0])}, {_fmt(l[1])}, {_fmt(l[2])})",
              f"  fov = {_fmt(cam['fov'])}",
              "}", "",
              f"num_samples = {scene['spp']}",
              f"num_bounces = {scene['bounces']}",
              "",
              f"width = {scene['width']}",
              f"height = {scene['height']}",
              "",
              f"output_image = {output_image}", ""]
    path = os.path.join(directory, "scene.rto")
    with open(path, "w") as fh:
        fh.write("\n".join(lines))
    return path


def seeded_pixels(width, height, n, seed=0):
    """`n` distinct pixel ids drawn with a seeded generator (sorted)."""
    rs = np.random.default_rng(seed)
    n = min(n, width * height)
    return np.sort(rs.choice(width * height, size=n, replace=False)).astype(np.int32)


# ---------------------------------------------------------------------------
# Batched multi-scene workload: 64 seeded scenes with log-uniform triangle
# counts in [10k, 2M] built from a sine-noise heightfield, boxes and UV
# spheres; materials 70% Lambertian / 20% dielectric (n in [1.3, 1.7]) /
# 10% emissive, plus an emissive quad; square images of 256..1024 pixels.
# ---------------------------------------------------------------------------
def _heightfield(n, size, amp, rs):
    """n x n vertex grid over [-size, size]^2 with a sine-noise height field."""
    xs = np.linspace(-size, size, n)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = np.zeros_like(X)
    for _ in range(4):
        fx, fz = rs.uniform(0.5, 6.0, 2) * np.pi / size
        ph = rs.uniform(0, 2 * np.pi, 2)
        Y += rs.uniform(0.2, 1.0) * np.sin(fx * X + ph[0]) * np.sin(fz * Z + ph[1])
    Y *= amp / 4.0
    P = np.stack([X, Y, Z], -1)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a, b, c, d = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
    tri = np.concatenate([np.stack([a, c, b], 2), np.stack([a, d, c], 2)]).reshape(-1, 3, 3)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm[nrm[:, 1] < 0] *= -1.0
    return (tri.astype(np.float32).reshape(-1, 9),
            np.repeat(nrm[:, None], 3, 1).astype(np.float32).reshape(-1, 9))


def _random_material(rs, name):
    u = rs.uniform()
    if u < 0.7:
        return {"name": name, "diffuse": tuple(rs.uniform(0.2, 0.9, 3)), "alpha": 1.0,
                "roughness": float(rs.uniform(0.0, 1.0)), "n": 1.0}
    if u < 0.9:
        return {"name": name, "diffuse": (1.0, 1.0, 1.0), "alpha": 0.0, "roughness": 0.0,
                "n": float(rs.uniform(1.3, 1.7))}
    return {"name": name, "diffuse": (0.0, 0.0, 0.0), "alpha": 1.0, "emit": True,
            "emission": tuple(rs.uniform(1.0, 5.0, 3))}


def batch_scene(index, seed=0, spp=32, bounces=8):
    """Scene `index` (0..63) of the seeded multi-scene batch."""
    rs = np.random.default_rng([seed, index])
    target = int(np.exp(rs.uniform(np.log(10_000), np.log(2_000_000))))
    mats = [_random_material(rs, f"m{k}") for k in range(8)]
    mats.append({"name": "sun", "diffuse": (0, 0, 0), "alpha": 1.0, "emit": True,
                 "emission": (4.0, 4.0, 4.0)})
    groups = []
    size = 1.0
    n = int(np.clip(np.sqrt(0.6 * target / 2.0) + 1, 8, 2048))
    groups.append((_heightfield(n, size, 0.15, rs), f"m{rs.integers(0, 8)}"))
    budget = target - 2 * (n - 1) ** 2
    while budget > 12:
        if rs.uniform() < 0.5 or budget < 200:
            c = rs.uniform(-0.8, 0.8, 3) * np.array([1, 0, 1]) + np.array([0, 0.2, 0])
            h = rs.uniform(0.03, 0.2, 3)
            groups.append((_box(tuple(c - h), tuple(c + h)), f"m{rs.integers(0, 8)}"))
            budget -= 12
        else:
            seg = int(np.clip(np.sqrt(budget / 4.0), 8, 512))
            rings = max(seg // 2, 3)
            c = (rs.uniform(-0.7, 0.7), rs.uniform(0.1, 0.5), rs.uniform(-0.7, 0.7))
            groups.append((_uv_sphere(c, rs.uniform(0.05, 0.3), seg, rings), f"m{rs.integers(0, 8)}"))
            budget -= 2 * seg * (rings - 1)
    s = 0.4
    groups.append((_quad((-s, 1.5, -s), (-s, 1.5, s), (s, 1.5, s), (s, 1.5, -s), (0, -1, 0)), "sun"))
    res = int(rs.choice([256, 384, 512, 768, 1024]))
    sc = _scene(groups, mats, res, res, spp, bounces, f"batch{index:02d}")
    sc["camera"] = {"eye": (0.0, 1.2, 2.2), "look_at": (0.0, 0.0, 0.0), "fov": 45.0}
    return sc


# ---------------------------------------------------------------------------
# Learned-radiance-cache workload: a seeded fBm heightfield island on a
# 2048^2 vertex grid (2 * 2047^2 = 8,380,418 triangles, smooth normals), its
# patches Lambertian with albedo U[0.2, 0.8], a dielectric water plane at
# n = 1.33 and a quad sun light; 1920x1080.
# ---------------------------------------------------------------------------
def _fbm_grid(n, octaves, base, seed):
    """fBm of smoothstep-interpolated value noise on an n x n grid over [0,1]^2."""
    rs = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n)
    out = np.zeros((n, n))
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        f = base * (2 ** o)
        lat = rs.uniform(-1.0, 1.0, (f + 2, f + 2))
        q = u * f
        i0 = np.minimum(np.floor(q).astype(np.int64), f)
        t = q - i0
        w = t * t * (3.0 - 2.0 * t)
        a = lat[i0[:, None], i0[None, :]]
        b = lat[i0[:, None] + 1, i0[None, :]]
        c = lat[i0[:, None], i0[None, :] + 1]
        d = lat[i0[:, None] + 1, i0[None, :] + 1]
        wx, wz = w[:, None], w[None, :]
        out += amp * ((a * (1 - wx) + b * wx) * (1 - wz) + (c * (1 - wx) + d * wx) * wz)
        norm += amp
        amp *= 0.5
    return out / norm


def island(n=2048, width=1920, height=1080, spp=64, bounces=12, seed=0, patches=16):
    rs = np.random.default_rng(seed)
    size = 1.0
    h = _fbm_grid(n, 7, 4, seed)
    xs = np.linspace(-size, size, n)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    r2 = (X * X + Z * Z) / (size * size)
    Y = 0.35 * (h + 0.55) * np.clip(1.0 - r2, 0.0, 1.0) - 0.08
    # smooth vertex normals from central differences
    dx = 2 * size / (n - 1)
    gx = np.gradient(Y, dx, axis=0)
    gz = np.gradient(Y, dx, axis=1)
    N = np.stack([-gx, np.ones_like(Y), -gz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    P = np.stack([X, Y, Z], -1)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    corners = ((i, j), (i + 1, j + 1), (i + 1, j)), ((i, j), (i, j + 1), (i + 1, j + 1))
    tv = np.concatenate([np.stack([P[a] for a in c], 2) for c in corners]).reshape(-1, 9)
    tn = np.concatenate([np.stack([N[a] for a in c], 2) for c in corners]).reshape(-1, 9)
    # albedo patches: 32x32-quad blocks pick one of `patches` materials
    blk = (i // 32) * ((n - 2) // 32 + 1) + (j // 32)
    pick = rs.integers(0, patches, blk.max() + 1)[blk]
    mat = np.concatenate([pick, pick]).reshape(-1).astype(np.int32)
    materials = [{"name": f"ground{k}", "diffuse": tuple(rs.uniform(0.2, 0.8, 3)), "alpha": 1.0,
                  "roughness": 1.0, "n": 1.0} for k in range(patches)]
    materials.append({"name": "water", "diffuse": (1.0, 1.0, 1.0), "alpha": 0.0,
                      "roughness": 0.0, "n": 1.33})
    materials.append({"name": "sun", "diffuse": (0.0, 0.0, 0.0), "alpha": 1.0, "emit": True,
                      "emission": (12.0, 11.0, 9.0)})
    w = 1.6
    wv, wn = _quad((-w, 0.0, -w), (-w, 0.0, w), (w, 0.0, w), (w, 0.0, -w), (0, 1, 0))
    s, sy = 0.5, 2.5
    lv, ln = _quad((-s + 0.6, sy, -s), (-s + 0.6, sy, s), (s + 0.6, sy, s), (s + 0.6, sy, -s),
                   (0, -1, 0))
    verts = np.ascontiguousarray(np.concatenate([tv.astype(np.float32), wv, lv]))
    norms = np.ascontiguousarray(np.concatenate([tn.astype(np.float32), wn, ln]))
    mat_idx = np.concatenate([mat, np.full(2, patches, np.int32), np.full(2, patches + 1, np.int32)])
    return {"name": f"island_{n}", "verts": verts, "normals": norms,
            "mat_idx": np.ascontiguousarray(mat_idx, dtype=np.int32),
            "materials": materials, "meshes": [],
            "camera": {"eye": (0.0, 1.1, 2.3), "look_at": (0.0, 0.0, 0.0), "fov": 40.0},
            "width": width, "height": height, "spp": spp, "bounces": bounces}
