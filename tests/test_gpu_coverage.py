"""Coverage shading (OIT_COVERAGE_SHADING: per-pixel fragments with sample
masks) and the display finalize on the GPU against the oracle, bit for bit:
Linked List, Spinlock and Simple (where the arrival-ordered set does not
matter), colour-pass tails per sample, dropped zero-mask fragments, and
the MSAA resolve / supersample downsample / sRGB8 encode."""
import numpy as np
import pytest

from conftest import deep_scene, per_target_counts
from gpu_run import run_gpu
from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig, Renderer, _abi

pytestmark = pytest.mark.gpu

LL, SIMPLE, SPIN = 1, 0, 4


def _say(*a):
    print(*a, flush=True)


def _masks(n, S, seed, zero_fraction=0.0):
    rng = np.random.default_rng(seed)
    m = rng.integers(1, 1 << S, n).astype(np.uint32)
    if zero_fraction:
        m[rng.uniform(size=n) < zero_fraction] = 0
    return m


def _same(name, got, want):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    if not np.array_equal(g, w):
        bad = np.argwhere(g != w)
        raise AssertionError(f"{name}: {len(bad)} mismatches, first {bad[:3].tolist()}")


@pytest.mark.parametrize("S", [2, 4, 8])
@pytest.mark.parametrize("K,tail,pool", [(4, 1, 1.5), (16, 1, 1.5), (8, 0, 1.5), (4, 1, 0.5)])
def test_linkedlist_coverage(oracle, S, K, tail, pool):
    p, tg, d, c = deep_scene(seed=41 + S)
    m = _masks(len(tg), S, S * 7 + K)
    cap = int(len(tg) * pool) + 1
    _say("LL coverage", S, K, tail, pool, len(tg))
    g = run_gpu(LL, K, tail, p.width, p.height, tg, d, c, ll_pool_nodes=cap, coverage_samples=S, mask=m)
    o = oracle.run(oracle.make_config(oracle.LINKEDLIST, K, tail, p.width, p.height, ll_capacity=cap,
                                      coverage_samples=S), tg, d, c, mask=m)
    _same("out", g["out"], o["out"])
    _same("frontk_key", g["frontk_key"], o["frontk_key"])
    _same("frontk_val", g["frontk_val"], o["frontk_val"])


@pytest.mark.parametrize("K", [2, 8])
def test_spinlock_coverage(oracle, K):
    S = 4
    p, tg, d, c = deep_scene(seed=53)
    m = _masks(len(tg), S, 53 + K)
    _say("Spinlock coverage", K, len(tg))
    g = run_gpu(SPIN, K, 1, p.width, p.height, tg, d, c, coverage_samples=S, mask=m)
    o = oracle.run(oracle.make_config(oracle.SPINLOCK, K, 1, p.width, p.height, coverage_samples=S), tg, d, c, mask=m)
    _same("out", g["out"], o["out"])
    _same("frontk_key", g["frontk_key"], o["frontk_key"])


def test_simple_coverage_with_replay(oracle):
    S, K = 4, 4
    p, tg, d, c = deep_scene(seed=59)
    m = _masks(len(tg), S, 59)
    _say("Simple coverage", len(tg))
    g = run_gpu(SIMPLE, K, 1, p.width, p.height, tg, d, c, coverage_samples=S, mask=m)
    cfg = oracle.make_config(oracle.SIMPLE, K, 1, p.width, p.height, coverage_samples=S)
    o = oracle.run(cfg, tg, d, c, mask=m)
    small = per_target_counts(tg, p.targets) <= K
    _same("out (n<=K)", g["out"][small], o["out"][small])
    # the stored set depends on arrival order: replay the GPU's set through the oracle's composite
    T = p.targets
    ids = g["frontk_val"]
    n = np.minimum(g["count"], K)
    valid = np.arange(K)[None, :] < n[:, None]
    safe = np.where(valid, ids, 0)
    from vk_order_independent_transparency_amd import scene  # noqa: F401
    col = np.array([oracle.encode(c[i]) for i in range(len(tg))], np.uint32)
    keys = np.array([oracle.depth_key(x) for x in d], np.uint32)
    ab_col, ab_key, ab_id, ab_mask = col[safe], keys[safe], safe.astype(np.uint32), m[safe]
    stored = np.zeros(len(tg), bool)
    stored[safe[valid]] = True
    tail_flag = (~stored).astype(np.uint8)
    rep = oracle.replay_kbuffer(cfg, tg, c, g["count"], ab_col, ab_key, ab_id, tail_flag, False, mask=m,
                                ab_mask=ab_mask)
    _same("out (replay)", g["out"], rep)


def test_zero_mask_fragments_are_dropped(oracle):
    S, K = 4, 4
    p, tg, d, c = deep_scene(seed=61)
    m = _masks(len(tg), S, 61, zero_fraction=0.3)
    cap = len(tg) + 1
    g = run_gpu(LL, K, 1, p.width, p.height, tg, d, c, ll_pool_nodes=cap, coverage_samples=S, mask=m)
    o = oracle.run(oracle.make_config(oracle.LINKEDLIST, K, 1, p.width, p.height, ll_capacity=cap,
                                      coverage_samples=S), tg, d, c, mask=m)
    _same("out", g["out"], o["out"])


@pytest.mark.parametrize("W,H,S,ss", [(37, 23, 1, 1), (37, 23, 4, 1), (20, 13, 1, 2), (16, 9, 8, 2)])
def test_finalize_srgb8(oracle, W, H, S, ss):
    rng = np.random.default_rng(W * H + S + ss)
    img = rng.uniform(-0.05, 1.05, (S * H * ss * W * ss, 4)).astype(np.float32)
    want = oracle.finalize_srgb8(img, W, H, S, ss)
    with Renderer(0) as r:
        src = DeviceBuffer.from_numpy(img)
        out = DeviceBuffer((H, W, 4), np.uint8)
        r.finalize_srgb8(src, out, W, H, S, ss)
        got = out.numpy()
    assert np.array_equal(got, want)


def test_coverage_rejections():
    with Renderer(0) as r:
        for alg in (3, 6):
            with pytest.raises(_abi.OitError) as e:
                r.configure(FrameConfig(algorithm=alg, layers=4, tail_blend=True, width=8, height=8,
                                        coverage_samples=4))
            assert e.value.status == _abi.OIT_ERR_UNSUPPORTED
        with pytest.raises(_abi.OitError):
            r.configure(FrameConfig(algorithm=LL, layers=4, tail_blend=True, width=8, height=8, samples=4,
                                    coverage_samples=4))
        r.configure(FrameConfig(algorithm=LL, layers=4, tail_blend=True, width=8, height=8, coverage_samples=4,
                                ll_pool_nodes=64))
        t = DeviceBuffer.from_numpy(np.zeros(4, np.uint32))
        dd = DeviceBuffer.from_numpy(np.zeros(4, np.float32))
        cc = DeviceBuffer.from_numpy(np.zeros((4, 4), np.float32))
        r.clear()
        with pytest.raises(_abi.OitError) as e:
            r.append(t, dd, cc)  # no masks
        assert e.value.status == _abi.OIT_ERR_INVALID_ARG
