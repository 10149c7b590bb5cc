"""Test-side helpers: an independent restatement of the reference OBJ reader
(wavefront.rs:30-153) for checking the product's loader, plus scene
builders shared by the GPU tests."""
from __future__ import annotations

import ctypes
import math

import numpy as np

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]

# the material table of the reference's own scene (main.rs:70-79), encoded as
# SMaterial does (material.rs:84-136)


def _diffuse(r, g, b):
    f = np.float32
    return (int(f(b) * f(255)) << 16) | (int(f(g) * f(255)) << 8) | int(f(r) * f(255))


def _gloss(m, g):
    return (m & 0b11100011_11111111_11111111_11111111) | (g << 26)


def _tex(m, t):
    return (m & 0b11111100_11111111_11111111_11111111) | (t << 24)


INDOOR_MATERIALS = {
    "baseboard": _gloss(_diffuse(1, 1, 1), 4),
    "ceiling": _gloss(_diffuse(1, 1, 1), 1),
    "fauteuil": _diffuse(1.0, 0.1, 0.4),
    "floor": _tex(_gloss(_diffuse(0.569, 0.494, 0.345), 4), 1),
    "glass": 0xC0000000,
    "wall": _gloss(_diffuse(0.65, 0.7, 0.9), 1),
    "wood_light": _tex(_gloss(_diffuse(0.6, 0.533, 0.455), 3), 2),
}


def strtof(s: str) -> np.float32:
    return np.float32(_libc.strtof(s.encode(), None))


def read_obj(path, materials=None):
    """Mesh::load_with_materials restated: returns (v9, uv6, mat) float32/uint32."""
    materials = materials or {}
    verts, tex, tris, uvs, mats = [], [], [], [], []
    mat = _diffuse(1.0, 1.0, 1.0)

    def idx(s):
        parts = s.split("/")
        v = int(parts[0]) - 1
        t = int(parts[1]) - 1 if len(parts) > 1 else None
        return v, t

    def push(i0, i1, i2):
        v0, v1, v2 = verts[i0[0]], verts[i1[0]], verts[i2[0]]
        tris.append(np.concatenate([v0, v1, v2]))
        if i0[1] is not None and i1[1] is not None and i2[1] is not None:
            uvs.append(np.concatenate([tex[i0[1]], tex[i1[1]], tex[i2[1]]]))
        else:
            uvs.append(np.zeros(6, np.float32))
        mats.append(mat)

    with open(path) as f:
        for line in f.read().split("\n"):
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                verts.append(np.array([strtof(x) for x in p[1:4]], np.float32))
            elif p[0] == "vt":
                tex.append(np.array([strtof(x) for x in p[1:3]], np.float32))
            elif p[0] == "usemtl":
                mat = materials[p[1]]
            elif p[0] == "f":
                ids = [idx(x) for x in p[1:]]
                i0, i1, i2 = ids[0], ids[1], ids[2]
                push(i0, i1, i2)
                for i3 in ids[3:]:
                    push(i0, i2, i3)
                    i2 = i3
    return (np.array(tris, np.float32).reshape(-1, 9), np.array(uvs, np.float32).reshape(-1, 6),
            np.array(mats, np.uint32))


def indoor_camera(time_s: float = 0.0):
    """Renderer::update_scene + Camera::set_rotation (renderer.rs:141-148,
    scene.rs:66-76) at a fixed time with zero deltas, as float32."""
    f = np.float32
    alpha = f(f(time_s) * f(-0.02) + f(0.1))
    pos = (f(f(-3.8) * f(math.sin(alpha))), f(1.6), f(f(3.0) * f(math.cos(alpha))))
    q = (f(math.cos(f(alpha * f(0.5)))), f(0.0), f(-f(math.sin(f(alpha * f(0.5))))), f(0.0))
    return pos, (0.0, 0.0, 0.0), q, (0.0, 0.0, 0.0, 0.0)


def indoor_scene(models_dir):
    from convector_amd.scenes import SceneData

    v, uv, m = read_obj(str(models_dir / "indoor.obj"), INDOOR_MATERIALS)
    pos, dpos, q, dq = indoor_camera()
    return SceneData(tris=v, materials=m, uv=uv, camera_pos=pos, camera_dpos=dpos,
                     camera_q=q, camera_dq=dq, emission_override=None, name="indoor")


MODEL_MATERIALS = {"indoor": INDOOR_MATERIALS,
                   "box_walls": {"wall": _diffuse(1, 1, 1), "glass": 0xC0000000},
                   "box_windows": {"wall": _diffuse(1, 1, 1), "glass": 0xC0000000}}


def model_scene(models_dir, name):
    from convector_amd.scenes import SceneData

    v, uv, m = read_obj(str(models_dir / (name + ".obj")), MODEL_MATERIALS.get(name))
    return SceneData(tris=v, materials=m, uv=uv, name=name)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# --- exact float32 arithmetic on rationals (independent of any FPU) ---------
from fractions import Fraction  # noqa: E402


def f32r(q) -> float:
    """Nearest float32 (ties to even) of an exact rational."""
    q = Fraction(q)
    if q == 0:
        return 0.0
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    while Fraction(2) ** e > q:
        e -= 1
    while Fraction(2) ** (e + 1) <= q:
        e += 1
    e = max(e, -126)
    scale = Fraction(2) ** (e - 23)
    m = q / scale
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * float(fl * scale)


def xmul(a, b):
    return f32r(Fraction(a) * Fraction(b))


def xadd(a, b):
    return f32r(Fraction(a) + Fraction(b))


def xfma(a, b, c):
    return f32r(Fraction(a) * Fraction(b) + Fraction(c))


def f32c(x):
    """A Python float constant rounded to float32 (as a Rust f32 literal)."""
    return float(np.float32(x))


def edge_distance64(tris, tri_id, rays):
    v = tris[tri_id].astype(np.float64).reshape(-1, 3, 3)
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    t = ((v[:, 0] - o) * n).sum(1) / (d * n).sum(1)
    p = o + t[:, None] * d
    dist = np.full(len(p), np.inf)
    for i, j in ((0, 1), (1, 2), (2, 0)):
        e = v[:, j] - v[:, i]
        dist = np.minimum(dist, np.linalg.norm(np.cross(e, p - v[:, i]), axis=1)
                          / np.linalg.norm(e, axis=1))
    return dist


def check_tolerance(sd, hits, ref, rays):
    """Hits of a traversal that may differ from the reference's only at exact
    ties in t or within 1e-6 of a triangle edge (FP64 distance): at most one
    mismatch in 10,000 rays, and where the ids agree t, u, v are the same
    bits."""
    gid, oid = hits["tri_id"], ref["tri"]
    mism = np.nonzero(gid != oid)[0]
    assert len(mism) <= max(1, len(rays) // 10000), len(mism)
    for k in mism:
        tg, to = hits["t"][k], ref["tuv"][k, 0]
        if tg == to:
            continue  # an exact tie: two triangles at the same distance
        near = [float(edge_distance64(sd.tris, np.array([t]), rays[k:k + 1])[0])
                for t in (gid[k], oid[k]) if t != 0xFFFFFFFF]
        assert near and min(near) <= 1e-6, (k, near)
    same = gid == oid
    assert (bits(hits["t"][same]) == bits(ref["tuv"][same, 0])).all()
    assert (bits(hits["u"][same]) == bits(ref["tuv"][same, 1])).all()
    assert (bits(hits["v"][same]) == bits(ref["tuv"][same, 2])).all()
