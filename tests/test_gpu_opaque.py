"""Opaque pass and depth-test culling (oit_draw_opaque) on the GPU against the
oracle: the z-buffer alone (ties, the clear depth), every algorithm with an
opaque draw in front of part of the scene (bit-exact; Simple where no pixel
holds more than K surviving fragments; WBOIT within 1e-5 relative), coverage
shading with per-sample planes, several opaque draws and appends per frame,
and the frame rules (opaque before transparent, a frame without an opaque
draw is not culled)."""
import numpy as np
import pytest

from conftest import deep_scene
from gpu_run import run_gpu
from vk_order_independent_transparency_amd import Renderer, _abi, scene

pytestmark = pytest.mark.gpu

SIMPLE, LL, LOOP, LOOP64, SPIN, INTERLOCK, WB = 0, 1, 2, 3, 4, 5, 6


def _say(*a):
    print(*a, flush=True)


def _same(name, got, want):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    if not np.array_equal(g, w):
        bad = np.argwhere(g != w)
        raise AssertionError(f"{name}: {len(bad)} mismatches, first {bad[:3].tolist()}")


def _opaque_scene(p, seed, spheres=12, samples=1):
    """Opaque spheres (alpha 1) over the same frame, some repeated so that equal depths tie."""
    q = scene.SceneParams(seed=seed, width=p.width, height=p.height, spheres=spheres, radius=(0.08, 0.25),
                          alpha=(1.0, 1.0), samples=samples)
    tg, d, c = scene.generate(q)
    n = len(tg) // 10
    tg = np.concatenate([tg, tg[:n]])
    d = np.concatenate([d, d[:n]])
    c = np.concatenate([c, c[::-1][:n]])  # the repeats carry other colours: the first must stay
    return tg, d, c


def _live(oracle, opq, tg, d):
    k = np.array([oracle.depth_key(float(x)) for x in d], np.uint32)
    return k < opq[0][tg]


def test_opaque_planes_alone(oracle):
    p, _, _, _ = deep_scene(seed=101)
    otg, od, oc = _opaque_scene(p, 5)
    _say("opaque only", len(otg))
    for splits in (1, 4):
        # one transparent fragment behind the clear depth: culled everywhere
        g = run_gpu(LL, 4, 1, p.width, p.height, np.uint32([0]), np.float32([2.0]),
                    np.float32([[0.1, 0.1, 0.1, 0.5]]), opaque=(otg, od, oc, None), opaque_splits=splits)
        cfg = oracle.make_config(LL, 4, 1, p.width, p.height, ll_capacity=4)
        keys, cols = oracle.opaque_pass(cfg, otg, od, oc)
        _same(f"out = opaque colour, splits={splits}", g["out"], cols)
        assert (keys != oracle.depth_key(1.0)).sum() > 0


@pytest.mark.parametrize("alg", [LL, LOOP, LOOP64, SPIN, INTERLOCK])
@pytest.mark.parametrize("K", [4, 16])
@pytest.mark.parametrize("tail", [0, 1])
def test_culled_frame_bitexact(oracle, alg, K, tail):
    p, tg, d, c = deep_scene(seed=103 + K)
    otg, od, oc = _opaque_scene(p, 7 + K)
    cap = len(tg) * 2 if alg in (LL, INTERLOCK) else 0
    _say("opaque", alg, K, tail, len(tg), len(otg))
    g = run_gpu(alg, K, tail, p.width, p.height, tg, d, c, ll_pool_nodes=cap, opaque=(otg, od, oc, None))
    base = oracle.make_config(alg, K, tail, p.width, p.height, ll_capacity=cap or len(tg) + 1)
    opq = oracle.opaque_pass(base, otg, od, oc)
    live = _live(oracle, opq, tg, d)
    assert 0 < live.sum() < len(tg)
    o = oracle.run(oracle.make_config(alg, K, tail, p.width, p.height, ll_capacity=cap or len(tg) + 1, opaque=opq),
                   tg, d, c)
    if alg == SPIN:  # its counter counts early-depth passes, arrival dependent beyond K
        _same("min(count,K)", np.minimum(g["count"], K), np.minimum(o["count"], K))
    else:
        _same("count", g["count"], o["count"])
    _same("frontk_key", g["frontk_key"], o["frontk_key"])
    _same("out", g["out"], o["out"])


def test_simple_culled(oracle):
    p, tg, d, c = deep_scene(seed=107)
    otg, od, oc = _opaque_scene(p, 11)
    K = 8
    g = run_gpu(SIMPLE, K, 1, p.width, p.height, tg, d, c, opaque=(otg, od, oc, None))
    opq = oracle.opaque_pass(oracle.make_config(SIMPLE, K, 1, p.width, p.height), otg, od, oc)
    o = oracle.run(oracle.make_config(SIMPLE, K, 1, p.width, p.height, opaque=opq), tg, d, c)
    live = _live(oracle, opq, tg, d)
    n = np.bincount(tg[live], minlength=p.targets)
    _same("count", g["count"], n.astype(np.uint32))
    small = n <= K
    assert small.sum() > 0 and (~small).sum() > 0
    _same("out (n<=K)", g["out"][small], o["out"][small])


def test_weighted_culled(oracle):
    p, tg, d, c = deep_scene(seed=109)
    otg, od, oc = _opaque_scene(p, 13)
    g = run_gpu(WB, 1, 1, p.width, p.height, tg, d, c, opaque=(otg, od, oc, None))
    opq = oracle.opaque_pass(oracle.make_config(WB, 1, 1, p.width, p.height), otg, od, oc)
    cfg = oracle.make_config(WB, 1, 1, p.width, p.height, opaque=opq)
    o64 = oracle.run_fp64(cfg, tg, d, c)["out"]
    err = np.abs(g["out"].astype(np.float64) - o64) / np.maximum(np.abs(o64), 1e-3)
    _say("WBOIT+opaque max rel err vs fp64", err.max())
    assert err.max() <= 1e-5


@pytest.mark.parametrize("alg", [LL, SIMPLE, SPIN, INTERLOCK])
@pytest.mark.parametrize("S", [2, 4])
def test_coverage_culled(oracle, alg, S):
    p, tg, d, c = deep_scene(seed=111 + S)
    otg, od, oc = _opaque_scene(p, 17 + S)
    rng = np.random.default_rng(113 + S)
    m = rng.integers(1, 1 << S, len(tg)).astype(np.uint32)
    om = rng.integers(1, 1 << S, len(otg)).astype(np.uint32)
    K = 32 if alg == SIMPLE else 8
    cap = len(tg) * 2 if alg in (LL, INTERLOCK) else 0
    g = run_gpu(alg, K, 1, p.width, p.height, tg, d, c, coverage_samples=S, mask=m, ll_pool_nodes=cap,
                opaque=(otg, od, oc, om))
    base = oracle.make_config(alg, K, 1, p.width, p.height, coverage_samples=S, ll_capacity=cap or len(tg) + 1)
    opq = oracle.opaque_pass(base, otg, od, oc, mask=om)
    o = oracle.run(oracle.make_config(alg, K, 1, p.width, p.height, coverage_samples=S,
                                      ll_capacity=cap or len(tg) + 1, opaque=opq), tg, d, c, mask=m)
    if alg == SIMPLE:
        small = o["count"] <= K
        _same("count", g["count"], o["count"])
        _same("out (n<=K)", g["out"][small], o["out"][small])
    else:
        _same("out", g["out"], o["out"])


@pytest.mark.parametrize("alg", [LL, LOOP, LOOP64])
def test_several_draws_per_frame(oracle, alg):
    p, tg, d, c = deep_scene(seed=117)
    otg, od, oc = _opaque_scene(p, 19)
    cap = len(tg) * 2 if alg == LL else 0
    opq = oracle.opaque_pass(oracle.make_config(alg, 8, 1, p.width, p.height), otg, od, oc)
    o = oracle.run(oracle.make_config(alg, 8, 1, p.width, p.height, ll_capacity=cap or len(tg) + 1, opaque=opq),
                   tg, d, c)
    for osp, sp in ((3, 1), (1, 3), (5, 7)):
        g = run_gpu(alg, 8, 1, p.width, p.height, tg, d, c, ll_pool_nodes=cap, opaque=(otg, od, oc, None),
                    opaque_splits=osp, splits=sp)
        _same(f"out opaque splits={osp} splits={sp}", g["out"], o["out"])


def test_frame_rules(oracle):
    p, tg, d, c = deep_scene(seed=119, width=32, height=24, spheres=60)
    from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=LL, layers=8, tail_blend=True, width=p.width, height=p.height,
                                ll_pool_nodes=2 * len(tg)))
        t_, d_, c_ = (DeviceBuffer.from_numpy(a) for a in (tg, d, c))
        out = DeviceBuffer((p.targets, 4), np.float32)
        r.clear()
        r.append(t_, d_, c_)
        with pytest.raises(_abi.OitError) as e:  # opaque after transparent in the same frame
            r.draw_opaque(t_, d_, c_)
        assert e.value.status == _abi.OIT_ERR_STATE
        r.resolve(out)
        # a zero-fragment opaque draw still gives the frame the depth test against 1.0
        d2 = d.copy()
        d2[::3] = 1.0 + d2[::3]
        d2_ = DeviceBuffer.from_numpy(d2)
        r.clear()
        r.draw_opaque(t_, d2_, c_, count=0)
        r.append(t_, d2_, c_)
        r.resolve(out)
        opq = (np.full(p.targets, oracle.depth_key(1.0), np.uint32),
               np.tile(np.float32([0.5, 0.5, 0.5, 1.0]), (p.targets, 1)))
        want = oracle.run(oracle.make_config(LL, 8, 1, p.width, p.height, ll_capacity=2 * len(tg), opaque=opq),
                          tg, d2, c)["out"]
        _same("depth 1.0 test", out.numpy(), want)
        # the next frame has no opaque draw: nothing is culled
        r.clear()
        r.append(t_, d2_, c_)
        r.resolve(out)
        want = oracle.run(oracle.make_config(LL, 8, 1, p.width, p.height, ll_capacity=2 * len(tg)), tg, d2, c)["out"]
        _same("no opaque draw", out.numpy(), want)
