"""Seeded procedural scenes for the per-ray path (synthetic inputs).

The reference renders a hand-modelled indoor OBJ (``src/main.rs:70-88``); the
workloads here are procedural so that every size is reproducible:

* ``box_with_light``: the closed box [-1,1]^3 of axis-aligned quads.  Five
  faces are whole quads; the z = -1 face is four quads around a central
  0.5 x 0.5 hole that holds the light, itself 2 x 2 sub-quads (8 triangles
  with ``SMaterial::sky()``, since direct sampling draws one of exactly 8
  light triangles, ``src/scene.rs:143-161``).  Walls face inward, wound
  counter-clockwise seen from inside (the geometric normal of
  ``src/triangle.rs:69-80`` is not flipped toward the ray).
* an icosphere of radius 0.4 at the origin, displaced radially, with the
  glossy ``SMaterial::white().with_glossiness(4)`` (``src/main.rs:73``);
* optionally free equilateral triangles scattered in the box.

All randomness comes from splitmix64 counter streams and float64 arithmetic
using only +, -, *, /, sqrt (correctly rounded), with exp/log only for the
log-uniform edge lengths; coordinates are rounded to float32 at the end.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)

# SMaterial encodings (src/material.rs:10-30, 84-136).
MAT_SKY = 0xC0000000  # emissive + direct-sample bits (material.rs:85-90)


def mat_diffuse(r: float, g: float, b: float) -> int:
    """SMaterial::diffuse (material.rs:98-101): channels truncated from f32*255."""
    f = np.float32
    ri = int(np.float32(f(r) * f(255.0)))
    gi = int(np.float32(f(g) * f(255.0)))
    bi = int(np.float32(f(b) * f(255.0)))
    return (bi << 16) | (gi << 8) | ri


def mat_diffuse_array(rgb: np.ndarray) -> np.ndarray:
    """mat_diffuse of every row of an (n, 3) array."""
    q = (rgb.astype(np.float32) * np.float32(255.0)).astype(np.float32).astype(np.uint32)
    return ((q[:, 2] << 16) | (q[:, 1] << 8) | q[:, 0]).astype(np.uint32)


def mat_with_glossiness(mat: int, gloss: int) -> int:
    """SMaterial::with_glossiness (material.rs:111-118)."""
    assert 0 <= gloss <= 6
    return (mat & 0b11100011_11111111_11111111_11111111) | (gloss << 26)


MAT_WHITE = mat_diffuse(1.0, 1.0, 1.0)
MAT_GLOSSY_WHITE = mat_with_glossiness(MAT_WHITE, 4)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class Stream:
    """splitmix64 counter stream: draw k of seed s is mix64(s + (k+1)*golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            z = self.seed + k * GOLDEN
            out = _mix64(z)
        self.count += n
        return out

    def uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        """n float64 values in [lo, hi) from the top 53 bits."""
        x = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * x


def _lattice_value(seed: int, ix: np.ndarray, iy: np.ndarray, iz: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        h = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)
        h = _mix64(h + ix.astype(np.int64).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        h = _mix64(h + iy.astype(np.int64).astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
        h = _mix64(h + iz.astype(np.int64).astype(np.uint64) * np.uint64(0x165667B19E3779F9))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, f0: float = 4.0) -> np.ndarray:
    """Value noise in [-1, 1]: trilinear over a seeded lattice with smoothstep
    weights, octave k at frequency f0*2^k with gain 2^-k, normalised by the
    gain sum."""
    total = np.zeros(p.shape[0])
    gain_sum = 0.0
    for k in range(octaves):
        f = f0 * (2.0 ** k)
        g = 2.0 ** (-k)
        q = p * f
        i0 = np.floor(q)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        i0 = i0.astype(np.int64)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = w[:, 0] if dx else 1.0 - w[:, 0]
            for dy in (0, 1):
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                for dz in (0, 1):
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    val = _lattice_value(seed + k, i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
                    acc += (wx * wy) * wz * val
        total += g * acc
        gain_sum += g
    return total / gain_sum


_ICO_FACES = np.array(
    [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
     (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
     (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)


def icosphere(subdivisions: int):
    """Unit icosphere (float64 vertices, int64 faces wound outward).

    Each level splits every face (a,b,c) into (a,ab,ca), (b,bc,ab), (c,ca,bc),
    (ab,bc,ca); new vertices are the normalised edge midpoints, numbered in
    ascending order of their (min, max) vertex pair."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 dtype=np.float64)
    v /= np.sqrt((v * v).sum(axis=1))[:, None]
    f = _ICO_FACES.copy()
    for _ in range(subdivisions):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1)
        e = e.reshape(-1, 2)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * nv + hi
        uniq, inv = np.unique(key, return_inverse=True)
        ulo = uniq // nv
        uhi = uniq % nv
        mid = (v[ulo] + v[uhi]) * 0.5
        mid /= np.sqrt((mid * mid).sum(axis=1))[:, None]
        v = np.concatenate([v, mid], 0)
        m = (inv + nv).reshape(-1, 3)
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                      np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3)
    return v, f


def _quad(center, u, v):
    """Corners of the quad center +- u +- v, counter-clockwise seen from the
    side u x v points to; returned as two fan triangles (wavefront.rs:125-140)."""
    c = np.asarray(center, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    p = [c - u - v, c + u - v, c + u + v, c - u + v]
    return [(p[0], p[1], p[2]), (p[0], p[2], p[3])]


@dataclasses.dataclass
class SceneData:
    """A triangle soup with per-triangle materials and a camera."""

    tris: np.ndarray  # (n, 9) float32: v0, v1, v2
    materials: np.ndarray  # (n,) uint32
    camera_pos: tuple = (0.0, 0.9, 0.0)
    camera_dpos: tuple = (0.0, 0.0, 0.0)
    camera_q: tuple = (0.70710677, -0.70710677, 0.0, 0.0)
    camera_dq: tuple = (0.0, 0.0, 0.0, 0.0)
    screen_distance: float = float(np.float32(1.0 / math.sin(math.pi / 5.0)))
    emission_override: Optional[tuple] = None
    uv: Optional[np.ndarray] = None  # (n, 6) float32 or None
    name: str = ""

    @property
    def ntris(self) -> int:
        return int(self.tris.shape[0])

    def camera_args(self):
        return (self.camera_pos, self.camera_dpos, self.camera_q, self.camera_dq,
                self.screen_distance)


def box_with_light(stream: Stream):
    """The 9 wall quads (18 triangles) and the 2x2 light (8 triangles)."""
    tris, mats = [], []
    X, Y, Z = np.eye(3)
    walls = [
        ((-1, 0, 0), Y, Z),    # x = -1, normal +x
        ((1, 0, 0), Z, Y),     # x = +1, normal -x
        ((0, -1, 0), Z, X),    # y = -1, normal +y
        ((0, 1, 0), X, Z),     # y = +1, normal -y
        ((0, 0, 1), Y, X),     # z = +1, normal -z
    ]
    quads = [_quad(c, u, v) for c, u, v in walls]
    # z = -1 face (normal +z) around the hole [-0.25, 0.25]^2
    quads.append(_quad((0, -0.625, -1), X, 0.375 * Y))
    quads.append(_quad((0, 0.625, -1), X, 0.375 * Y))
    quads.append(_quad((-0.625, 0, -1), 0.375 * X, 0.25 * Y))
    quads.append(_quad((0.625, 0, -1), 0.375 * X, 0.25 * Y))
    for q in quads:
        rgb = stream.uniform(3, 0.2, 0.8)
        m = mat_diffuse(*rgb)
        for t in q:
            tris.append(np.concatenate(t))
            mats.append(m)
    for cx in (-0.125, 0.125):
        for cy in (-0.125, 0.125):
            for t in _quad((cx, cy, -1), 0.125 * X, 0.125 * Y):
                tris.append(np.concatenate(t))
                mats.append(MAT_SKY)
    return np.array(tris, dtype=np.float64), np.array(mats, dtype=np.uint32)


def _sphere_tris(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    t = np.concatenate([v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]], 1)
    return t


def random_rotations(stream: Stream, n: int) -> np.ndarray:
    """Uniform unit quaternions by rejection from the 4-ball (float64)."""
    out = np.zeros((0, 4))
    while out.shape[0] < n:
        q = stream.uniform(4 * 2 * n, -1.0, 1.0).reshape(-1, 4)
        r2 = (q * q).sum(1)
        q = q[(r2 <= 1.0) & (r2 > 1e-6)]
        q /= np.sqrt((q * q).sum(1))[:, None]
        out = np.concatenate([out, q], 0)
    return out[:n]


def _rotate(q: np.ndarray, p: np.ndarray) -> np.ndarray:
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)
    return np.einsum("nij,nj->ni", r, p)


def free_triangles(stream: Stream, n: int):
    """n equilateral triangles: centroids uniform in [-0.95,0.95]^3, edge
    log-uniform in [0.005, 0.1], uniformly random orientation, diffuse albedo
    U(0.2, 0.8) per channel."""
    c = stream.uniform(3 * n, -0.95, 0.95).reshape(n, 3)
    edge = np.exp(stream.uniform(n, math.log(0.005), math.log(0.1)))
    q = random_rotations(stream, n)
    s3 = math.sqrt(3.0)
    local = np.array([[1.0, 0.0, 0.0], [-0.5, s3 / 2, 0.0], [-0.5, -s3 / 2, 0.0]])
    tris = np.zeros((n, 9))
    for k in range(3):
        p = np.repeat(local[k][None, :], n, 0) * (edge / s3)[:, None]
        tris[:, 3 * k:3 * k + 3] = c + _rotate(q, p)
    rgb = stream.uniform(3 * n, 0.2, 0.8).reshape(n, 3)
    return tris, mat_diffuse_array(rgb)


def _finish(parts, **kw) -> SceneData:
    tris = np.concatenate([p[0] for p in parts], 0).astype(np.float32)
    tris[tris == 0.0] = 0.0  # canonical +0.0
    mats = np.concatenate([p[1] for p in parts], 0).astype(np.uint32)
    return SceneData(tris=np.ascontiguousarray(tris), materials=mats, **kw)


def config0(subdivisions: int = 5, seed: int = 1, emission=(10.0, 10.0, 10.0)) -> SceneData:
    """Closed box + light + icosphere (default subdivision 5 = 20,480 faces)
    with vertices displaced radially by 0.05*u, u ~ U(-1, 1), one draw per
    vertex in index order."""
    st = Stream(seed)
    box = box_with_light(Stream(seed ^ 0x5EED0B0C))
    v, f = icosphere(subdivisions)
    u = st.uniform(v.shape[0], -1.0, 1.0)
    v = v * (0.4 + 0.05 * u)[:, None]
    sph = (_sphere_tris(v, f), np.full(f.shape[0], MAT_GLOSSY_WHITE, dtype=np.uint32))
    return _finish([box, sph], emission_override=emission, name=f"config0-sub{subdivisions}")


def config1(subdivisions: int = 8, n_free: int = 20000, seed: int = 2,
            emission=(10.0, 10.0, 10.0)) -> SceneData:
    """Box + light + icosphere (default subdivision 8 = 1,310,720 faces)
    displaced by 4-octave value noise of amplitude 0.08, plus n_free free
    triangles."""
    box = box_with_light(Stream(seed ^ 0x5EED0B0C))
    v, f = icosphere(subdivisions)
    noise = value_noise(v, seed)
    v = v * (0.4 + 0.08 * noise)[:, None]
    sph = (_sphere_tris(v, f), np.full(f.shape[0], MAT_GLOSSY_WHITE, dtype=np.uint32))
    free = free_triangles(Stream(seed * 7919 + 17), n_free)
    return _finish([box, sph, free], emission_override=emission,
                   name=f"config1-sub{subdivisions}-free{n_free}")


def procedural_scene(n_tris: int, seed: int, emission=(10.0, 10.0, 10.0)) -> SceneData:
    """A scene of the small configuration's kind with about n_tris
    triangles: the box with its light (26 triangles), the largest icosphere
    (20 * 4^k faces, radius 0.4, radially displaced by 0.05*u) that fits,
    and free triangles for the rest."""
    budget = max(int(n_tris) - 26, 20)
    k = 0
    while 20 * 4 ** (k + 1) <= budget and k < 9:
        k += 1
    st = Stream(seed)
    box = box_with_light(Stream(seed ^ 0x5EED0B0C))
    v, f = icosphere(k)
    u = st.uniform(v.shape[0], -1.0, 1.0)
    v = v * (0.4 + 0.05 * u)[:, None]
    sph = (_sphere_tris(v, f), np.full(f.shape[0], MAT_GLOSSY_WHITE, dtype=np.uint32))
    parts = [box, sph]
    n_free = budget - f.shape[0]
    if n_free > 0:
        parts.append(free_triangles(Stream(seed * 7919 + 17), n_free))
    return _finish(parts, emission_override=emission, name=f"procedural-{n_tris}-s{seed}")


def batch_jobs(n_jobs: int = 64, seed0: int = 1000, tri_range=(1e3, 2e6),
               width_range=(64, 1920), height_range=(64, 1080), max_spp: int = 64):
    """The many-jobs configuration: job j uses seed seed0 + j to draw its
    triangle count (log-uniform in tri_range), its width and height
    (log-uniform, rounded down to multiples of 16) and its samples per pixel
    (uniform in {1, 2, 4, ..., max_spp}).  Returns (n_tris, seed, width,
    height, spp) tuples; procedural_scene(n_tris, seed) makes the scene."""
    out = []
    n_spp = int(math.log2(max_spp)) + 1
    for j in range(n_jobs):
        st = Stream(seed0 + j)
        a, b, c, d = st.uniform(4, 0.0, 1.0)
        n = int(math.exp(math.log(tri_range[0]) + a * (math.log(tri_range[1]) - math.log(tri_range[0]))))
        w = int(math.exp(math.log(width_range[0]) + b * (math.log(width_range[1]) - math.log(width_range[0]))))
        h = int(math.exp(math.log(height_range[0]) + c * (math.log(height_range[1]) - math.log(height_range[0]))))
        spp = 1 << min(int(d * n_spp), n_spp - 1)
        out.append((n, seed0 + j, max(16, w // 16 * 16), max(16, h // 16 * 16), spp))
    return out


def mesh_from_obj_arrays(vertices: np.ndarray, faces: np.ndarray, materials: np.ndarray,
                         **kw) -> SceneData:
    tris = np.concatenate([vertices[faces[:, 0]], vertices[faces[:, 1]],
                           vertices[faces[:, 2]]], 1).astype(np.float32)
    return SceneData(tris=np.ascontiguousarray(tris), materials=materials.astype(np.uint32),
                     **kw)


def write_obj(path: str, scene: SceneData) -> dict:
    """Writes the soup as OBJ (3 vertices per face, %.9g so every float32
    round-trips) with one `usemtl` per material run; returns the
    name -> material table that Mesh::load_with_materials needs."""
    names = {}
    lines = []
    t = scene.tris
    for i in range(t.shape[0]):
        for k in range(3):
            lines.append("v %.9g %.9g %.9g" % (t[i, 3 * k], t[i, 3 * k + 1], t[i, 3 * k + 2]))
    cur = None
    for i in range(t.shape[0]):
        m = int(scene.materials[i])
        if m != cur:
            nm = "m%08x" % m
            names[nm] = m
            lines.append("usemtl " + nm)
            cur = m
        lines.append("f %d %d %d" % (3 * i + 1, 3 * i + 2, 3 * i + 3))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return names


def rays_inward(n: int, seed: int, radius: float = 10.0, coherent: bool = False,
                target_scale: float = 1.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Rays from a sphere of the given radius toward random points on the unit
    sphere (bench.rs:195-223 shapes, seeded).  Returns (n, 6) float32."""
    st = Stream(seed)

    def sphere_pts(m):
        z = st.uniform(m, -1.0, 1.0)
        phi = st.uniform(m, 0.0, 2.0 * math.pi)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)

    if coherent:
        o = np.repeat(sphere_pts((n + 7) // 8), 8, 0)[:n] * radius
        dst = sphere_pts(n) * 0.5 * target_scale
    else:
        o = sphere_pts(n) * radius
        dst = sphere_pts(n) * target_scale
    o = o + np.asarray(center)
    dst = dst + np.asarray(center)
    d = dst - o
    d /= np.sqrt((d * d).sum(1))[:, None]
    return np.ascontiguousarray(np.concatenate([o, d], 1).astype(np.float32))
