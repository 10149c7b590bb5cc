"""The oracle's build with 16 and 32 centroid bins (bvh.rs:275 with
bins.len() changed).  No reference output exists for these bin counts, so
the trees are pinned by what the build must satisfy, checked here in FP64
from the triangles alone: every inner node's two children are one of the
candidate splits of its binning (centroid bins over the barycenter bounds,
or stably sorted chunks when the node is small or a bin holds more than
n/8 of n > bins), that candidate is the cheapest by the TreeSAH cost (to
FP32 rounding), a node splits exactly when the cheapest candidate is not
dearer than n * 120, and the leaves cover every triangle once.  At 64 bins
the same checks hold for the reference's own trees."""
import math

import numpy as np
import pytest

from convector_amd import scenes

P = 0.8


def _tree_sah64(parent_area, box, n):
    """TreeSurfaceAreaHeuristic::aabb_cost (bvh.rs:452-469) in FP64."""
    if n == 0:
        return 0.0
    ac_ap = _area(box) / parent_area
    m = math.log2(n)
    aabb_term = 1.0 + ac_ap * (2.0 * P - n * P ** m) / (P - 2.0 * P * P)
    tri_term = n * P ** (m - 1.0) * ac_ap
    return 40.0 * aabb_term + 120.0 * tri_term


def _area(box):
    d = box[3:] - box[:3]
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _enclose(lo, hi, ids):
    return np.concatenate([lo[ids].min(0), hi[ids].max(0)]).astype(np.float64)


def _candidates(ids, cen, nbins):
    """Every candidate (axis, left id set) of InterimNode::split's binning
    of these triangles (bvh.rs:142-261)."""
    n = len(ids)
    out = []
    for axis in range(3):
        c = cen[ids, axis]
        uniform = n <= nbins // 2
        if not uniform:
            mn = np.float32(c.min())
            size = np.float32(c.max()) - mn
            with np.errstate(divide="ignore", invalid="ignore"):
                f = np.floor((np.float32(nbins) * (c - mn)) / size)
            b = np.where((f >= 0) & (f < nbins), f, nbins - 1).astype(np.int64)
            if np.bincount(b, minlength=nbins).max() > n // 8 and n > nbins:
                uniform = True
        if uniform:
            srt = np.argsort(c, kind="stable")
            tpb = max(1, -(-n // nbins))
            b = np.empty(n, np.int64)
            b[srt] = np.arange(n) // tpb
        used = np.unique(b)
        if len(used) < 2:
            continue
        for cut in used[1:]:
            out.append((axis, frozenset(ids[b < cut].tolist())))
    return out


def _walk(nodes, order):
    """(node triangle ids, left child ids, right child ids or None) per node."""
    res = []

    def ids_of(i):
        nd = nodes[i]
        if nd["len"]:
            return order[nd["index"]:nd["index"] + nd["len"]]
        return np.concatenate([ids_of(nd["index"]), ids_of(nd["index"] + 1)])

    def visit(i):
        nd = nodes[i]
        if nd["len"]:
            res.append((ids_of(i), None, None))
            return
        res.append((ids_of(i), ids_of(nd["index"]), ids_of(nd["index"] + 1)))
        visit(nd["index"])
        visit(nd["index"] + 1)

    root = np.concatenate([ids_of(0), ids_of(1)])
    res.append((root, ids_of(0), ids_of(1)))
    visit(0)
    visit(1)
    return res


def _check_tree(oracle, tris, nbins):
    st, nodes, order, _ = oracle.bvh_build(tris, nbins)
    assert st == 0
    t = tris.reshape(-1, 3, 3).astype(np.float32)
    lo, hi = t.min(1), t.max(1)
    third = np.float32(1.0) / np.float32(3.0)
    cen = ((t[:, 0] + t[:, 1]) + t[:, 2]) * third
    assert sorted(order.tolist()) == list(range(len(tris)))
    stats = {"inner": 0, "leaf": 0}
    for ids, left, right in _walk(nodes, order):
        n = len(ids)
        pbox = _enclose(lo, hi, ids)
        pa = _area(pbox)
        cands = _candidates(ids, cen, nbins)
        costs = {}
        for axis, lset in cands:
            L = np.array(sorted(lset))
            R = np.array(sorted(set(ids.tolist()) - lset))
            costs[(axis, lset)] = (_tree_sah64(pa, _enclose(lo, hi, L), len(L))
                                   + _tree_sah64(pa, _enclose(lo, hi, R), len(R)))
        best = min(costs.values()) if costs else math.inf
        if left is None:
            stats["leaf"] += 1
            if n > 1:
                # no split: n * 120 was cheaper than every candidate
                assert best >= n * 120.0 * (1.0 - 1e-5), (n, best)
            continue
        stats["inner"] += 1
        assert len(left) + len(right) == n
        ls, rs = frozenset(left.tolist()), frozenset(right.tolist())
        chosen = [c for (a, s), c in costs.items() if s in (ls, rs)]
        assert chosen, "children are not a candidate split of the node's bins"
        assert min(chosen) <= best * (1.0 + 1e-5) + 1e-9, (n, min(chosen), best)
        assert min(chosen) <= n * 120.0 * (1.0 + 1e-5)
        # the child of the smaller surface area comes first (bvh.rs:337-345)
        assert _area(_enclose(lo, hi, left)) <= _area(_enclose(lo, hi, right)) * (1 + 1e-6)
    return stats


def _clustered(n, seed):
    """Triangles packed in a few tight clusters plus scattered ones, so that
    one bin takes more than n/8 and the sorted fallback runs."""
    rng = np.random.default_rng(seed)
    centres = np.concatenate([rng.normal(0, 0.01, (n * 3 // 4, 3)),
                              rng.uniform(-1, 1, (n - n * 3 // 4, 3))])
    return (centres[:, None, :] + rng.normal(0, 0.02, (n, 3, 3))).reshape(n, 9).astype(np.float32)


@pytest.mark.parametrize("nbins", [16, 32, 64])
@pytest.mark.parametrize("case", ["clustered", "procedural"])
def test_binned_build_is_the_cheapest_candidate_tree(oracle, nbins, case):
    if case == "clustered":
        tris = _clustered(600, 5)
    else:
        tris = scenes.procedural_scene(700, 42).tris
    stats = _check_tree(oracle, tris, nbins)
    assert stats["inner"] > 20


def test_bin_counts_change_the_tree(oracle):
    tris = scenes.procedural_scene(3000, 8).tris
    trees = {b: oracle.bvh_build(tris, b) for b in (16, 32, 64)}
    assert all(t[0] == 0 for t in trees.values())
    assert not np.array_equal(trees[16][1], trees[64][1])
    assert not np.array_equal(trees[32][1], trees[64][1])
    # 64 bins is the reference build itself
    st, nodes, order, _ = oracle.bvh_build(tris)
    assert np.array_equal(nodes, trees[64][1]) and np.array_equal(order, trees[64][2])


def test_bad_bin_count(oracle):
    st, _, _, _ = oracle.bvh_build(scenes.procedural_scene(100, 1).tris, 48)
    assert st == -4
