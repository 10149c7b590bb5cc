"""The queue-partition map of the lane-refill traversal
(convector_amd/csrc/kernels.hip, k_extend_refill, cv_tuning.queue_parts),
enumerated on the CPU: with chunks of 2^glog rays (glog = floor(log2(n /
parts)) clamped to 12..18) dealt to the partitions in turn, partition x
holds part_size(x) rays, and its local positions l < part_size(x) map to
queue positions ((l >> glog) * parts + x) << glog | (l & (2^glog - 1)).
Every queue position below n must be reached exactly once, and none at or
above n, or a ray would be traversed twice, skipped, or read out of
bounds."""
import numpy as np
import pytest


def _glog(n, parts):
    g = max(n // parts, 1).bit_length() - 1
    return min(max(g, 12), 18)


def _part_size(n, parts, glog, x):
    full = n >> glog
    c = (full - x + parts - 1) // parts if full > x else 0
    return (c << glog) + ((n & ((1 << glog) - 1)) if (full & (parts - 1)) == x else 0)


@pytest.mark.parametrize("parts", [1, 8])
@pytest.mark.parametrize("n", [0, 1, 100, 4095, 4096, 4097, 8 * 4096 - 1, 8 * 4096, 8 * 4096 + 1,
                               9 * 4096 + 5, 65536 * 8 + 4095, 122880, 942080, 2073600,
                               (1 << 21) * 8 + 3, 33177600])
def test_partition_map_is_a_bijection(n, parts):
    glog = _glog(n, parts)
    mask = (1 << glog) - 1
    seen = np.zeros(n + 1, np.int64)
    total = 0
    for x in range(parts):
        size = _part_size(n, parts, glog, x)
        total += size
        l = np.arange(size, dtype=np.int64)
        q = (((l >> glog) * parts + x) << glog) | (l & mask)
        assert (q < n).all()
        np.add.at(seen, q, 1)
    assert total == n
    assert (seen[:n] == 1).all()


def test_chunk_sizes():
    """Chunks of 2^12 rays for small queues, 2^18 for large ones, and about
    n / parts between."""
    assert _glog(100, 8) == 12
    assert _glog(942080, 8) == 16
    assert _glog(33177600, 8) == 18
