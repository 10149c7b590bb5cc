"""Oracle geometry against the reference's known answers (vector3.rs:446-540,
quaternion.rs:167-234, aabb.rs:186-254, triangle.rs:188-259) and FP64
re-intersection."""
import math

import numpy as np
import pytest


def _unit_vectors(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1)[:, None]).astype(np.float32)


def rot_hemi(oracle, v, n, approx=None):
    approx = oracle.APPROX_PORTABLE if approx is None else approx
    v = np.ascontiguousarray(np.broadcast_to(np.float32(v), np.shape(n) if np.ndim(n) > 1 else (1, 3)),
                             np.float32)
    n = np.ascontiguousarray(np.broadcast_to(np.float32(n), v.shape), np.float32)
    out = np.zeros_like(v)
    oracle.lib().or_rotate_hemisphere(oracle.ptr(v), oracle.ptr(n), len(v), approx, oracle.ptr(out))
    return out


X, Y, Z = np.eye(3, dtype=np.float32)


@pytest.mark.parametrize("approx", [0, 1])
def test_rotate_hemisphere_known(oracle, approx):
    # vector3.rs:446-469
    def close(a, b):
        return np.sum((np.float32(a) - np.float32(b)) ** 2) <= 1e-6
    assert close(rot_hemi(oracle, X, Y[None], approx)[0], X)
    assert close(rot_hemi(oracle, Y, X[None], approx)[0], Y)
    assert close(rot_hemi(oracle, Y, Y[None], approx)[0], -Z)
    assert close(rot_hemi(oracle, X, X[None], approx)[0], -Z)
    assert close(rot_hemi(oracle, X, Z[None], approx)[0], X)
    assert close(rot_hemi(oracle, Y, Z[None], approx)[0], Y)


def test_rotate_hemisphere_extrema(oracle):
    # vector3.rs:495-512: exact equality at normal -z (the flip path)
    for v in (X, Y, Z):
        out = rot_hemi(oracle, v, (-Z)[None])[0]
        assert (out == -v).all()


@pytest.mark.parametrize("approx", [0, 1])
def test_rotate_hemisphere_orthogonal_linear(oracle, approx):
    # vector3.rs:471-493
    ns, xs, ys = _unit_vectors(4096, 1), _unit_vectors(4096, 2), _unit_vectors(4096, 3)
    s = (xs + ys).astype(np.float32)
    sm = rot_hemi(oracle, s, ns, approx)
    n2b = (s.astype(np.float64) ** 2).sum(1)
    n2a = (sm.astype(np.float64) ** 2).sum(1)
    assert ((n2b - n2a) ** 2).max() <= 1e-5
    sa = rot_hemi(oracle, xs, ns, approx) + rot_hemi(oracle, ys, ns, approx)
    assert ((sm - sa) ** 2).sum(1).max() <= 1e-6


def test_rotate_hemisphere_random_coverage(oracle):
    # vector3.rs:514-540: hemisphere samples rotated to +x keep x >= 0 and
    # cover every sign of y and z
    h = np.zeros(4096 * 8 * 3, np.float32)
    oracle.lib().or_rng_ref_hemisphere(1, 2, 3, 4096, oracle.ptr(h))
    v = h.reshape(-1, 3)
    w = rot_hemi(oracle, v, np.repeat(X[None], len(v), 0))
    eps = 1e-4
    assert (w[:, 0] + eps >= 0).all()
    assert (w[:, 1] < -eps).any() and (w[:, 1] > eps).any()
    assert (w[:, 2] < -eps).any() and (w[:, 2] > eps).any()


def qrot(oracle, v, q):
    v = np.ascontiguousarray(v, np.float32)
    q = np.ascontiguousarray(np.broadcast_to(np.float32(q), (len(v), 4)), np.float32)
    out = np.zeros_like(v)
    oracle.lib().or_quat_rotate(oracle.ptr(v), oracle.ptr(q), len(v), oracle.ptr(out))
    return out


def test_quaternion_rotations(oracle):
    # quaternion.rs:167-219
    v = _unit_vectors(256, 7)
    h = np.float32(0.5) * np.float32(math.sqrt(2.0))
    assert (((qrot(oracle, v, (1, 0, 0, 0)) - v) ** 2).sum(1) <= 1e-14).all()
    ex = np.stack([v[:, 0], -v[:, 2], v[:, 1]], 1)
    assert (((qrot(oracle, v, (h, h, 0, 0)) - ex) ** 2).sum(1) <= 1e-12).all()
    ey = np.stack([v[:, 2], v[:, 1], -v[:, 0]], 1)
    assert (((qrot(oracle, v, (h, 0, h, 0)) - ey) ** 2).sum(1) <= 1e-12).all()
    ez = np.stack([-v[:, 1], v[:, 0], v[:, 2]], 1)
    assert (((qrot(oracle, v, (h, 0, 0, h)) - ez) ** 2).sum(1) <= 1e-12).all()


def test_quaternion_interpolate(oracle):
    # quaternion.rs:221-234
    h = np.float32(0.5) * np.float32(math.sqrt(2.0))
    q = np.array([[1, 0, 0, 0]], np.float32)
    dq = np.array([[h - np.float32(1), 0, 0, h]], np.float32)
    t = np.array([0.5], np.float32)
    out = np.zeros((1, 4), np.float32)
    oracle.lib().or_quat_interpolate(oracle.ptr(q), oracle.ptr(dq), oracle.ptr(t), 1,
                                     oracle.ptr(out))
    r = qrot(oracle, np.array([[1, 0, 0]], np.float32), out[0])
    assert ((r[0] - np.array([h, h, 0])) ** 2).sum() <= 1e-12


def test_aabb_area(oracle):
    # aabb.rs:195-200
    b = np.array([1, 2, 3, 5, 7, 9], np.float32)
    assert oracle.lib().or_aabb_area(oracle.ptr(b)) == 148.0


def _aabb_any(oracle, box, o, d, approx=0):
    rays = np.array([list(o) + list(d)], np.float32)
    tmin, tmax = np.zeros(1, np.float32), np.zeros(1, np.float32)
    mask = np.zeros(1, np.uint32)
    oracle.lib().or_aabb_intersect(oracle.ptr(box), oracle.ptr(rays), 1, approx,
                                   oracle.ptr(tmin), oracle.ptr(tmax), oracle.ptr(mask))
    # MAabbIntersection::any: mask and tmax's sign bit 0
    return bool(mask[0] >> 31) and not (tmax.view(np.uint32)[0] >> 31)


@pytest.mark.parametrize("approx", [0, 1])
def test_aabb_intersect_known(oracle, approx):
    # aabb.rs:202-254
    box = np.array([0, 1, 2, 1, 3, 5], np.float32)

    def nrm(v):
        v = np.float32(v)
        return v / np.float32(math.sqrt(float((v.astype(np.float64) ** 2).sum())))

    cases = [((0, 0, 0), nrm((2, 3, 5)), True, False),
             ((0, 0, 0), nrm((1, 4, 5)), True, False),
             ((0, 0, 0), nrm((2, 3, 0)), False, False),
             ((0.2, 1.2, 2.2), nrm((1, 1, 0)), True, True),
             ((0.01, 2.0, 3.5), nrm((0, 0, 1)), True, True)]
    for o, d, fwd, bwd in cases:
        assert _aabb_any(oracle, box, o, d, approx) == fwd
        assert _aabb_any(oracle, box, o, [np.float32(0) - x for x in d], approx) == bwd


def test_aabb_against_f64_slab(oracle):
    rng = np.random.default_rng(9)
    box = np.array([-1, -2, -0.5, 1.5, 0.5, 2.0], np.float32)
    o = rng.uniform(-4, 4, (2000, 3)).astype(np.float32)
    tgt = rng.uniform(-2.5, 2.5, (2000, 3))
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    hits = []
    for i in range(2000):
        od, dd = o[i].astype(np.float64), d[i].astype(np.float64)
        t1 = (box[:3] - od) / dd
        t2 = (box[3:] - od) / dd
        tmin = np.minimum(t1, t2).max()
        tmax = np.maximum(t1, t2).min()
        exp = tmax >= tmin and tmax >= 0
        margin = min(abs(tmax - tmin), abs(tmax)) > 1e-4
        if margin:
            assert _aabb_any(oracle, box, o[i], d[i]) == exp
        hits.append(exp)
    assert 0.1 < np.mean(hits) < 0.9


def test_triangle_known_answer(oracle):
    # triangle.rs:188-224
    tri = np.array([0, 1, 1, -1, -1, 1, 1, -1, 1], np.float32)
    out = np.zeros(4, np.float32)
    r1 = np.array([0, 0, 0, 0, 0, 1], np.float32)
    r2 = np.array([-1, 0, 0, 0, 0, 1], np.float32)
    assert oracle.lib().or_triangle_intersect(oracle.ptr(tri), oracle.ptr(r1), 1e5, oracle.ptr(out))
    assert 0.99 < out[0] < 1.01
    assert not oracle.lib().or_triangle_intersect(oracle.ptr(tri), oracle.ptr(r2), 1e5, None)


@pytest.mark.parametrize("approx", [0, 1])
def test_triangle_direct_known_answer(oracle, approx):
    # triangle.rs:226-259
    tri = np.array([0, 1, 1, -1, -1, 1, 1, -1, 1], np.float32)
    for r in ([0, 0, 0, 0, 0, 1], [-1, 0, 0, 0, 0, 1]):
        out = np.zeros(6, np.float32)
        oracle.lib().or_triangle_intersect_direct(oracle.ptr(tri), oracle.ptr(np.float32(r)),
                                                  approx, oracle.ptr(out))
        assert 0.99 < out[4] < 1.01
        assert 0.99 < float((out[:3].astype(np.float64) ** 2).sum()) < 1.01
        assert abs(out[3] - 2.0) < 2e-3  # area of the triangle


def test_triangle_fp64_reintersection(oracle):
    """t, u, v within 1e-5 relative of the same formula in FP64, and the hit
    decision equals the textbook Moller-Trumbore test away from edges."""
    rng = np.random.default_rng(11)
    n = 3000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    tgt = (tris[:, 0:3] + tris[:, 3:6] + tris[:, 6:9]) / 3 + rng.normal(0, 0.3, (n, 3))
    d = (tgt - o)
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    checked = 0
    for i in range(n):
        out = np.zeros(4, np.float32)
        ray = np.concatenate([o[i], d[i]]).astype(np.float32)
        hit = oracle.lib().or_triangle_intersect(oracle.ptr(tris[i]), oracle.ptr(ray), 1e5,
                                                 oracle.ptr(out))
        v0, v1, v2 = (tris[i, 3 * k:3 * k + 3].astype(np.float64) for k in range(3))
        od, dd = o[i].astype(np.float64), d[i].astype(np.float64)
        e1, e2 = v0 - v2, v1 - v0
        nn = np.cross(e1, e2)
        den = dd @ nn
        t = ((v0 - od) @ nn) / den
        c = np.cross(dd, v0 - od)
        u, v = (c @ e2) / den, (c @ e1) / den
        w = 1 - u - v
        cosang = abs(den) / (np.linalg.norm(dd) * np.linalg.norm(nn))
        if cosang > 0.1 and abs(t) < 100:
            # well-conditioned rays: within 1e-5 relative (absolute near 0)
            for a, b in ((out[0], t), (out[1], u), (out[2], v)):
                assert abs(a - b) <= 1e-5 * max(1.0, abs(b))
        # Moller-Trumbore on (v0, v1, v2)
        edge1, edge2 = v1 - v0, v2 - v0
        p = np.cross(dd, edge2)
        det = edge1 @ p
        s = od - v0
        bu = (s @ p) / det
        q = np.cross(s, edge1)
        bv = (dd @ q) / det
        tt = (edge2 @ q) / det
        mt = bu >= 0 and bv >= 0 and bu + bv <= 1 and tt >= 0
        if min(abs(u), abs(v), abs(w), abs(t)) > 1e-5:
            assert bool(hit) == mt
            checked += 1
    assert checked > 2500
