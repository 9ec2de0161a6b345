"""The 256x256 workload's committed fixture (tests/golden/small_256_ll.json,
written by tests/golden/make_golden.py): the scene generator and the oracle
reproduce it on the CPU, and the HIP Linked List reproduces it bit for bit
on the GPU, at K = 8 and K = 4."""
import hashlib
import json
import os

import numpy as np
import pytest

from vk_order_independent_transparency_amd import scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_256_ll.json")


def _fixture():
    with open(GOLDEN) as f:
        return json.load(f)


def _check(name, out, want):
    out = np.ascontiguousarray(out, dtype=np.float32)
    for i, px in want["pixels"].items():
        assert out[int(i)].tolist() == px, (name, i)
    assert hashlib.sha256(out.tobytes()).hexdigest() == want["sha256"], name


def test_oracle_reproduces_the_fixture(oracle):
    g = _fixture()
    p = scene.SMALL_256
    tg, d, c = scene.generate(p)
    assert len(tg) == g["fragments"]
    assert hashlib.sha256(tg.tobytes() + d.tobytes() + c.tobytes()).hexdigest() == g["fragments_sha256"]
    for K in (8, 4):
        cfg = oracle.make_config(oracle.LINKEDLIST, K, 1, p.width, p.height, ll_capacity=len(tg) + 1)
        _check(f"oracle K={K}", oracle.run(cfg, tg, d, c)["out"], g["images"][f"K{K}"])


@pytest.mark.gpu
@pytest.mark.parametrize("K", [8, 4])
def test_hip_linked_list_reproduces_the_fixture(K):
    from gpu_run import run_gpu

    g = _fixture()
    p = scene.SMALL_256
    tg, d, c = scene.generate(p)
    print("golden 256x256 K", K, len(tg), flush=True)
    r = run_gpu(1, K, 1, p.width, p.height, tg, d, c, ll_pool_nodes=len(tg) + 1)
    _check(f"HIP K={K}", r["out"], g["images"][f"K{K}"])
