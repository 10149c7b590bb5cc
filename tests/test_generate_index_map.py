"""The camera-ray generation walks a job's rectangle in bands of 8 rows,
column by column, so that a wave's 64 paths form an 8 x 8 tile; the last
band holds the rows left over (rect_h % 8).  Restated from k_generate, the
map from path index to pixel must reach every pixel exactly once for any
rectangle."""
import numpy as np
import pytest


def generate_map(w, h):
    """(tx, ty) of path k = 0 .. w*h-1 as k_generate computes them."""
    k = np.arange(w * h, dtype=np.int64)
    full = h & ~7
    tx = np.empty_like(k)
    ty = np.empty_like(k)
    band = k < full * w
    kb = k[band]
    b = kb // (8 * w)
    r = kb - b * (8 * w)
    tx[band] = r >> 3
    ty[band] = b * 8 + (r & 7)
    if (~band).any():
        hb = h - full
        r = k[~band] - full * w
        tx[~band] = r // hb
        ty[~band] = full + r % hb
    return tx, ty


@pytest.mark.parametrize("w", list(range(1, 65)) + [1280, 1920])
def test_every_pixel_exactly_once(w):
    for h in list(range(1, 24)) + [736, 1080]:
        tx, ty = generate_map(w, h)
        assert tx.min() >= 0 and tx.max() < w and ty.min() >= 0 and ty.max() < h
        seen = np.zeros((h, w), np.int64)
        np.add.at(seen, (ty, tx), 1)
        assert (seen == 1).all(), (w, h)


def test_waves_are_8x8_tiles():
    tx, ty = generate_map(64, 16)
    for wave in range(len(tx) // 64):
        x, y = tx[64 * wave:64 * wave + 64], ty[64 * wave:64 * wave + 64]
        assert x.max() - x.min() == 7 and y.max() - y.min() == 7
