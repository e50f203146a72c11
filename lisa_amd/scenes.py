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
    """UV sphere: `seg` segments around, `rings` rings pole to pole, smooth normals."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, seg + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    pos = np.asarray(center, np.float64) + radius * nrm
    tris_v, tris_n = [], []
    for i in range(rings):
        for j in range(seg):
            a, b = (i, j), (i, j + 1)
            c, d = (i + 1, j + 1), (i + 1, j)
            if i != 0:
                tris_v.append([pos[a], pos[c], pos[b]])
                tris_n.append([nrm[a], nrm[c], nrm[b]])
            if i != rings - 1:
                tris_v.append([pos[a], pos[d], pos[c]])
                tris_n.append([nrm[a], nrm[d], nrm[c]])
    v = np.asarray(tris_v, np.float32).reshape(-1, 9)
    n = np.asarray(tris_n, np.float32).reshape(-1, 9)
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
