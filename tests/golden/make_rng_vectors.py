#!/usr/bin/env python3
"""Writes tests/golden/rng_vectors.json: golden draws of the reference Rng.

Independent restatement of ruuda/convector src/random.rs with Python
integers (exact wrapping u64 arithmetic) and exact rational arithmetic for
the float conversions, so it shares no code with the C oracle:

  Rng::with_seed       random.rs:43-70
  Rng::next            random.rs:73-90  (returns the old state)
  sample_u32           random.rs:97-101 (Mu64 -> 8 x u32, little endian)
  sample_unit/biunit/angle  random.rs:104-136

Also the PIXEL stream (splitmix64 premix of pixel ^ frame << 32, then
xorshift64 13/7/17, high 32 bits) that this project defines.
"""
from __future__ import annotations

import json
import os
import struct
from fractions import Fraction

M64 = (1 << 64) - 1


def with_seed(x, y, i):
    a = (x * 12276630456901467871) & M64
    b = (y * 7661526868048087387) & M64
    c = (i * 2268244495640532043) & M64
    seed = (a + b + c) & M64
    seed = (seed + seed % 9358246936573323101) & M64
    primes = [14491630826648200009, 13149596372461506851, 6119410235796056053,
              14990141545859273719]
    return [(seed * p) & M64 for p in primes]


STEP = [3 * 1073243692214514217, 5 * 3335100457702756523, 7 * 8789056573444181,
        11 * 781436371140792079]


def draws(x, y, i, n):
    st = with_seed(x, y, i)
    out = []
    for _ in range(n):
        lanes = []
        for k in range(4):
            lanes += [st[k] & 0xFFFFFFFF, st[k] >> 32]
        out.append(lanes)
        st = [(st[k] * STEP[k]) & M64 for k in range(4)]
    return out


def f32_round(q: Fraction) -> float:
    """Rounds an exact rational to the nearest float32 (ties to even)."""
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


def i32(u):
    return u - (1 << 32) if u >= (1 << 31) else u


def to_f32(v: int) -> Fraction:
    """cvtdq2ps: the i32 rounded to float32."""
    return Fraction(f32_round(Fraction(v)))


PI_F32 = Fraction(struct.unpack("<f", struct.pack("<f", 3.14159265358979))[0])


def unit(u):
    # fma(f32(i), 0.5 / i32::MIN, 0.5), one rounding
    return f32_round(to_f32(i32(u)) * Fraction(1, -(1 << 32)) + Fraction(1, 2))


def biunit(u):
    if i32(u) == 0:
        return -0.0  # +0 times a negative constant
    return f32_round(to_f32(i32(u)) * Fraction(1, -(1 << 31)))


def angle(u):
    if i32(u) == 0:
        return -0.0
    k = Fraction(f32_round(PI_F32 / -(1 << 31)))
    return f32_round(to_f32(i32(u)) * k)


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def pixel_draws(pixel, frame, n):
    s = splitmix64(pixel ^ (frame << 32)) or 0x9E3779B97F4A7C15
    out = []
    for _ in range(n):
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        out.append(s >> 32)
    return out


def main():
    seeds = [(2, 5, 7), (1, 2, 3), (0, 0, 0), (32, 64, 1), (1248, 704, 16), (4294967295, 7, 99)]
    data = {"ref": [], "pixel": []}
    for x, y, i in seeds:
        d = draws(x, y, i, 64)
        data["ref"].append({
            "seed": [x, y, i], "u32": d,
            "unit": [[unit(u) for u in row] for row in d[:16]],
            "biunit": [[biunit(u) for u in row] for row in d[:16]],
            "angle": [[angle(u) for u in row] for row in d[:16]],
        })
    for pixel, frame in [(0, 1), (1, 1), (1279, 3), (942079, 16), (2073599, 1)]:
        data["pixel"].append({"pixel": pixel, "frame": frame, "u32": pixel_draws(pixel, frame, 40)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(data, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
