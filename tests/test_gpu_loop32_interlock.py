"""Loop32 (two geometry passes, oitLoop.frag.glsl) and the ordered interlock
(oitInterlock.frag.glsl with pixel_interlock_ordered) on the GPU against the
oracle, bit for bit: every K, tail blending on and off, shared depths,
negative and signed-zero depths, several appends per frame (Loop32 replays
its colour pass over them; the interlock's pool grows mid-frame), coverage
shading for the interlock, and frames that reuse the context."""
import numpy as np
import pytest

from conftest import deep_scene
from gpu_run import run_gpu
from vk_order_independent_transparency_amd import Renderer

pytestmark = pytest.mark.gpu

LOOP, INTERLOCK = 2, 5


def _say(*a):
    print(*a, flush=True)


def _same(name, got, want):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    if not np.array_equal(g, w):
        bad = np.argwhere(g != w)
        raise AssertionError(f"{name}: {len(bad)} mismatches, first {bad[:3].tolist()}")


def _cfg(oracle, alg, K, tail, p, **kw):
    return oracle.make_config(alg, K, tail, p.width, p.height, samples=p.samples, **kw)


@pytest.mark.parametrize("K", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("tail", [0, 1])
def test_loop32_bitexact(oracle, K, tail):
    p, tg, d, c = deep_scene(seed=81, spheres=160 if K <= 16 else 400)
    _say("Loop32", K, tail, len(tg))
    g = run_gpu(LOOP, K, tail, p.width, p.height, tg, d, c)
    o = oracle.run(_cfg(oracle, oracle.LOOP, K, tail, p), tg, d, c)
    _same("count", g["count"], o["count"])
    _same("frontk_key", g["frontk_key"], o["frontk_key"])
    _same("frontk_val(colour)", g["frontk_val"], o["frontk_val"])
    _same("out", g["out"], o["out"])


@pytest.mark.parametrize("K", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("tail", [0, 1])
def test_interlock_bitexact(oracle, K, tail):
    p, tg, d, c = deep_scene(seed=83, spheres=160 if K <= 16 else 400)
    _say("Interlock", K, tail, len(tg))
    g = run_gpu(INTERLOCK, K, tail, p.width, p.height, tg, d, c)
    o = oracle.run(_cfg(oracle, oracle.INTERLOCK, K, tail, p), tg, d, c)
    _same("count", g["count"], o["count"])  # deterministic: the critical sections run in primitive order
    _same("frontk_key", g["frontk_key"], o["frontk_key"])
    _same("frontk_val(id)", g["frontk_val"], o["frontk_val"])
    _same("out", g["out"], o["out"])


@pytest.mark.parametrize("alg", [LOOP, INTERLOCK])
def test_shared_and_signed_depths(oracle, alg):
    rng = np.random.default_rng(87)
    base_d = np.array([-0.5, -0.0, 0.0, 0.25, 0.25, 0.5, 0.5, 0.5, 0.75, 1.5], np.float32)
    tg = np.repeat(np.arange(6, dtype=np.uint32), len(base_d))
    d = np.tile(base_d, 6)
    c = rng.uniform(0.05, 0.4, (len(tg), 4)).astype(np.float32)  # distinct colours: the winner of a shared depth shows
    order = rng.permutation(len(tg))
    tg, d, c = tg[order], d[order], c[order]
    for K in (2, 4, 8):
        g = run_gpu(alg, K, 1, 3, 2, tg, d, c)
        o = oracle.run(oracle.make_config(alg, K, 1, 3, 2), tg, d, c)
        _same(f"out K={K}", g["out"], o["out"])
        _same(f"frontk_val K={K}", g["frontk_val"], o["frontk_val"])


@pytest.mark.parametrize("alg", [LOOP, INTERLOCK])
def test_several_appends_per_frame(oracle, alg):
    p, tg, d, c = deep_scene(seed=89)
    o = oracle.run(_cfg(oracle, alg, 8, 1, p), tg, d, c)
    for splits in (1, 3, 7):
        g = run_gpu(alg, 8, 1, p.width, p.height, tg, d, c, splits=splits)
        _same(f"out splits={splits}", g["out"], o["out"])


@pytest.mark.parametrize("S", [2, 4, 8])
def test_interlock_coverage(oracle, S):
    p, tg, d, c = deep_scene(seed=91)
    m = np.random.default_rng(91 + S).integers(1, 1 << S, len(tg)).astype(np.uint32)
    for K in (1, 4, 16, 32):
        g = run_gpu(INTERLOCK, K, 1, p.width, p.height, tg, d, c, coverage_samples=S, mask=m)
        o = oracle.run(oracle.make_config(oracle.INTERLOCK, K, 1, p.width, p.height, coverage_samples=S), tg, d, c,
                       mask=m)
        _same(f"out K={K}", g["out"], o["out"])


def test_frames_reuse_the_context(oracle):
    """Several frames on one context in a row, switching algorithms and sizes: the interlock's
    pool is left clean by its resolve, Loop32's planes are cleared per frame."""
    with Renderer(0) as r:
        for seed, alg, K, w, h, sp in ((93, INTERLOCK, 4, 48, 40, 160), (94, INTERLOCK, 8, 64, 48, 300),
                                       (95, LOOP, 4, 48, 40, 160), (96, INTERLOCK, 16, 40, 30, 120),
                                       (97, LOOP, 16, 64, 48, 300), (98, LOOP, 8, 32, 24, 90)):
            p, tg, d, c = deep_scene(seed=seed, width=w, height=h, spheres=sp)
            g = run_gpu(alg, K, 1, w, h, tg, d, c, renderer=r)
            o = oracle.run(oracle.make_config(alg, K, 1, w, h), tg, d, c)
            _same(f"out seed={seed}", g["out"], o["out"])
