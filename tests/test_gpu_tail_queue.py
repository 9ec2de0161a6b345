"""Colour-pass tail queue: a tail fragment's record goes to the region of
1024 fragment ids that holds its id, and the link pass pushes the records
onto unordered per-target lists (Loop64's classification pushes its own);
the composites fold each list in id order. Frames whose tails fill many
regions (K = 1: nearly every fragment),
a partial last region, appends whose first ids fall inside a region, and a
Linked List pool that overflows inside a region: bit-exact with the oracle,
with the same tail count as one append of the whole stream."""
import numpy as np
import pytest

from conftest import deep_scene
from gpu_run import run_gpu

pytestmark = pytest.mark.gpu

LL, LOOP32, LOOP64, SPIN = 1, 2, 3, 4


def _same(name, got, want):
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        raise AssertionError(f"{name}: {len(bad)} mismatches, first at {bad[:3].tolist()}")


def _bounds(F):
    # first ids 1, 1023, 1025, 3000, 4097, ...: none of them on a region boundary
    cut = [0, 1, 1023, 1025, 3000, 4097, 20011, F - 5, F]
    return sorted(set(min(max(x, 0), F) for x in cut))


@pytest.mark.parametrize("alg", [LOOP64, SPIN, LOOP32])
@pytest.mark.parametrize("K", [1, 2, 8])
def test_tail_queue_uneven_appends(oracle, alg, K):
    p, tg, d, c = deep_scene(seed=42, width=64, height=48, spheres=200)
    F = len(tg)
    assert F % 1024 != 0 and F > 20011
    print("alg", alg, "K", K, "F", F, flush=True)
    want = oracle.run(oracle.make_config(alg, K, 1, p.width, p.height), tg, d, c)
    one = run_gpu(alg, K, 1, p.width, p.height, tg, d, c)
    many = run_gpu(alg, K, 1, p.width, p.height, tg, d, c, bounds=_bounds(F))
    _same("out, one append", one["out"], want["out"])
    _same("out, eight appends", many["out"], want["out"])
    assert one["tail_fragments"] == many["tail_fragments"] > 0


@pytest.mark.parametrize("extra", [0, 777])
def test_linked_list_overflow_inside_a_region(oracle, extra):
    p, tg, d, c = deep_scene(seed=43, width=64, height=48, spheres=200)
    F = len(tg)
    cap = (F // 2 // 1024) * 1024 + extra  # overflow from node cap - 1 on
    print("F", F, "cap", cap, flush=True)
    want = oracle.run(oracle.make_config(LL, 4, 1, p.width, p.height, ll_capacity=cap), tg, d, c)
    one = run_gpu(LL, 4, 1, p.width, p.height, tg, d, c, ll_pool_nodes=cap, max_fragments=F)
    many = run_gpu(LL, 4, 1, p.width, p.height, tg, d, c, ll_pool_nodes=cap, max_fragments=F, bounds=_bounds(F))
    _same("out, one append", one["out"], want["out"])
    _same("out, eight appends", many["out"], want["out"])
    assert one["ll_overflow"] == many["ll_overflow"] > 0
    assert one["tail_fragments"] == many["tail_fragments"] > 0
