"""The all-cores CPU baseline computes the frame it claims to time: the
oracle split over pixel bands (fragments bucketed by band in draw order,
targets renumbered per band) equals the full-frame oracle run bit for bit,
for band counts that do and do not divide the height, with empty bands and
with the tail engaged."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("bands", [1, 7, 16, 64])
def test_banded_frame_equals_full_frame(oracle, bands):
    sys.path.insert(0, ROOT)
    import bench

    rng = np.random.default_rng(bands)
    W, H, K = 48, 40, 4
    n = 6000
    tg = rng.integers(0, W * (H // 2), n).astype(np.uint32)  # the lower half of the frame stays empty
    d = rng.uniform(-0.1, 1.1, n).astype(np.float32)
    a = rng.uniform(0.1, 0.7, n).astype(np.float32)
    c = np.concatenate([rng.uniform(0, 1, (n, 3)) * a[:, None], a[:, None]], 1).astype(np.float32)
    full = oracle.run(oracle.make_config(oracle.LINKEDLIST, K, 1, W, H, ll_capacity=n + 1), tg, d, c)["out"]
    with ThreadPoolExecutor(4) as ex:
        got = bench.oracle_frame_banded(oracle, tg, d, c, K, W, H, bands, ex)
    assert np.array_equal(got.view(np.uint32), full.view(np.uint32))
