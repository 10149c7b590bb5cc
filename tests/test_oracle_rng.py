"""The oracle's random streams against golden vectors written by an
independent big-integer restatement of random.rs (tests/golden/
make_rng_vectors.py), and the reference's own range tests
(random.rs:198-258)."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rng_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_ref_u32_stream(oracle, golden):
    for case in golden["ref"]:
        x, y, i = case["seed"]
        out = np.zeros(64 * 8, np.uint32)
        oracle.lib().or_rng_ref_draws(x, y, i, 64, oracle.ptr(out))
        assert out.tolist() == [u for row in case["u32"] for u in row]


@pytest.mark.parametrize("kind,key", [(0, "unit"), (1, "biunit"), (2, "angle")])
def test_ref_float_conversions(oracle, golden, kind, key):
    for case in golden["ref"]:
        x, y, i = case["seed"]
        out = np.zeros(16 * 8, np.float32)
        oracle.lib().or_rng_ref_floats(x, y, i, 16, kind, oracle.ptr(out))
        exp = np.array(case[key], np.float32).reshape(-1)
        assert out.view(np.uint32).tolist() == exp.view(np.uint32).tolist()


def test_pixel_stream(oracle, golden):
    for case in golden["pixel"]:
        out = np.zeros(40, np.uint32)
        oracle.lib().or_rng_pixel_draws(case["pixel"], case["frame"], 40, oracle.ptr(out))
        assert out.tolist() == case["u32"]


def test_sample_ranges(oracle):
    # random.rs:198-231: unit in [0,1], biunit in [-1,1], angle in [-pi,pi]
    for kind, lo, hi in [(0, 0.0, 1.0), (1, -1.0, 1.0), (2, -math.pi, math.pi)]:
        out = np.zeros(4096 * 8, np.float32)
        oracle.lib().or_rng_ref_floats(2, 5, 7, 4096, kind, oracle.ptr(out))
        assert out.min() >= np.float32(lo) and out.max() <= np.float32(hi)


def test_unit_closed_interval_endpoints():
    # fma(i32 -> f32, -2^-32, 0.5): i32::MIN gives exactly 1, i32::MAX gives 0
    from helpers import xfma
    assert xfma(float(np.float32(-2147483648)), -2.0 ** -32, 0.5) == 1.0
    assert xfma(float(np.float32(2147483647)), -2.0 ** -32, 0.5) == 0.0


def test_hemisphere_unit_norm(oracle):
    # random.rs:233-240: |v| within [0.991, 1.009]; and z >= 0
    out = np.zeros(4096 * 8 * 3, np.float32)
    oracle.lib().or_rng_ref_hemisphere(2, 5, 7, 4096, oracle.ptr(out))
    v = out.reshape(-1, 3).astype(np.float64)
    r = np.sqrt((v * v).sum(1))
    assert r.min() >= 0.991 and r.max() <= 1.009
    assert v[:, 2].min() >= 0.0


def test_hemisphere_cosine_weighted(oracle):
    # cosine-weighted: E[z] = 2/3, E[z^2] = 1/2 for the exact distribution
    out = np.zeros(4096 * 8 * 3, np.float32)
    oracle.lib().or_rng_ref_hemisphere(1, 2, 3, 4096, oracle.ptr(out))
    z = out.reshape(-1, 3)[:, 2].astype(np.float64)
    assert abs(z.mean() - 2 / 3) < 0.01
    assert abs((z * z).mean() - 0.5) < 0.01


def test_u32_nonzero_xor(oracle):
    # random.rs:243-258
    out = np.zeros(4096 * 8, np.uint32)
    oracle.lib().or_rng_ref_draws(2, 5, 7, 4096, oracle.ptr(out))
    x = np.bitwise_xor.reduce(out.reshape(-1, 8), axis=0)
    assert (x != 0).all()
