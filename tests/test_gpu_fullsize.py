"""The two full-size frames, checked by properties that do not need the
oracle to run the whole frame.

1920x1080, 1000 spheres: one device-generated stream feeds the Linked List,
Loop64 and Spinlock at K = 8 and 16; wherever a pixel holds at most K
fragments the three front-K depth-key lists agree bit for bit (and the
Linked List's and Spinlock's fragment ids too), and on sampled pixels the
Linked List, Loop64 and Spinlock images (colour-pass tail included, where
a pixel holds more than K) equal the oracle run on those pixels' own
fragments.

3840x2160 with 4 samples per pixel (sample shading), 4000 spheres: WBOIT and
the Linked List at K = 32 over the same stream; on 4096 sampled sample
targets WBOIT is within 1e-5 relative of the FP64 oracle and the Linked List
image equals the oracle bit for bit (a target's result depends only on its
own fragments in draw order, so the oracle runs on the sampled targets'
fragments alone)."""
import numpy as np
import pytest

from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig, Renderer, Stream, scene

pytestmark = pytest.mark.gpu

SIMPLE, LL, LOOP, LOOP64, SPIN, INTERLOCK, WB = 0, 1, 2, 3, 4, 5, 6


def _say(*a):
    print(*a, flush=True)


def _device_stream(p):
    st = Stream()
    g = scene.DeviceScene(p)
    t, d, c = g.buffers()
    g.emit(t, d, c, st)
    st.synchronize()
    F = g.fragments
    g.close()
    return st, F, t, d, c


def _subset(tg, d, c, targets):
    """Fragments of the given targets, in draw order, with targets renumbered 0..m-1."""
    targets = np.unique(targets)
    sel = np.nonzero(np.isin(tg, targets))[0]
    return targets, np.searchsorted(targets, tg[sel]).astype(np.uint32), d[sel], c[sel]


def test_1080p_front_k_agree_across_variants(oracle):
    p = scene.SceneParams(seed=1, width=1920, height=1080, spheres=1000, radius=(0.01, 0.05), alpha=(0.1, 0.5))
    st, F, t, d, c = _device_stream(p)
    T = p.targets
    _say("1080p", F)
    for K in (8, 16):
        res = {}
        with Renderer(0) as r:
            for alg in (LL, LOOP64, SPIN):
                r.configure(FrameConfig(algorithm=alg, layers=K, tail_blend=True, width=p.width, height=p.height,
                                        ll_pool_nodes=F + F // 2 if alg == LL else 0))
                out = DeviceBuffer((T, 4), np.float32)
                fk = DeviceBuffer((T, K), np.uint32)
                fv = DeviceBuffer((T, K), np.uint32)
                cnt = DeviceBuffer((T,), np.uint32)
                r.clear(st)
                r.append(t, d, c, st, count=F)
                r.resolve_debug(out, fk, fv, cnt, st)
                st.synchronize()
                res[alg] = (out.numpy(), fk.numpy(), fv.numpy(), cnt.numpy())
        n = res[LL][3]  # list length = fragments of the pixel
        small = n <= K
        assert small.sum() > 0.5 * T
        for alg in (LOOP64, SPIN):
            assert np.array_equal(res[alg][1][small], res[LL][1][small]), (K, alg)
        assert np.array_equal(res[SPIN][2][small], res[LL][2][small]), K  # the same fragment ids
        # the Linked List image on sampled pixels against the oracle on their fragments alone
        rng = np.random.default_rng(K)
        pick = rng.choice(T, 4096, replace=False)
        tg_h = t.numpy()[:F]
        tgts, sub_t, sub_d, sub_c = _subset(tg_h, d.numpy()[:F], c.numpy()[:F], pick)
        cfg = oracle.make_config(LL, K, 1, len(tgts), 1, ll_capacity=len(sub_t) + 1)
        want = oracle.run(cfg, sub_t, sub_d, sub_c)["out"]
        assert np.array_equal(res[LL][0][tgts].view(np.uint32), want.view(np.uint32)), K
        # Loop64 and Spinlock on the same sampled pixels, tail included (canonical order), with the
        # pixels over K counted so that the tail is known to be engaged
        deep = int((n[tgts] > K).sum())
        _say("K", K, "sampled pixels holding more than K fragments:", deep)
        assert deep > 0
        for alg in (LOOP64, SPIN):
            want = oracle.run(oracle.make_config(alg, K, 1, len(tgts), 1), sub_t, sub_d, sub_c)["out"]
            assert np.array_equal(res[alg][0][tgts].view(np.uint32), want.view(np.uint32)), (K, alg)
        # The whole frame: the canonical tail order bit for bit, and the error of the canonical
        # image against the sequential reading of the shaders (each evicted colour blended at the
        # evicting fragment's position, oitLoop64.frag.glsl:124-140, oitSpinlock.frag.glsl:109-128)
        tg_all, d_all, c_all = tg_h, d.numpy()[:F], c.numpy()[:F]
        for alg in (LOOP64, SPIN):
            want = oracle.run(oracle.make_config(alg, K, 1, p.width, p.height), tg_all, d_all, c_all)["out"]
            assert np.array_equal(res[alg][0].view(np.uint32), want.view(np.uint32)), (K, alg)
            seq = oracle.run(oracle.make_config(alg, K, 1, p.width, p.height, tail_order=oracle.TAIL_SEQUENTIAL),
                             tg_all, d_all, c_all)["out"]
            err = np.abs(res[alg][0].astype(np.float64) - seq.astype(np.float64))
            diff = (err > 0).any(axis=1)
            _say(f"{'loop64' if alg == LOOP64 else 'spinlock'} K={K}: canonical vs sequential tail order: "
                 f"max abs error per channel {err.max(axis=0).tolist()}, {int(diff.sum())} of {T} pixels differ, "
                 f"{int((err > 2.0 ** -12).any(axis=1).sum())} by more than 2^-12")
            assert not diff[small].any()  # no tail where a pixel holds at most K fragments


def test_4k_4x_wboit_and_linked_list_k32(oracle):
    p = scene.SceneParams(seed=3, width=3840, height=2160, samples=4, spheres=4000, radius=(0.005, 0.03),
                          alpha=(0.05, 0.4))
    st, F, t, d, c = _device_stream(p)
    T = p.targets
    _say("4K 4x", F, "fragment-samples over", T, "sample targets")
    outs = {}
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=WB, layers=1, tail_blend=False, width=p.width, height=p.height, samples=4))
        o = DeviceBuffer((T, 4), np.float32)
        r.render(t, d, c, o, st, count=F)
        st.synchronize()
        outs[WB] = o.numpy()
        del o
        r.configure(FrameConfig(algorithm=LL, layers=32, tail_blend=True, width=p.width, height=p.height, samples=4,
                                ll_pool_nodes=F + F // 2))
        o = DeviceBuffer((T, 4), np.float32)
        r.render(t, d, c, o, st, count=F)
        st.synchronize()
        outs[LL] = o.numpy()
        st_ll = r.stats()
        del o
    assert st_ll.ll_overflow == 0
    rng = np.random.default_rng(44)
    pick = rng.choice(T, 4096, replace=False)
    tg_h = t.numpy()[:F]
    tgts, sub_t, sub_d, sub_c = _subset(tg_h, d.numpy()[:F], c.numpy()[:F], pick)
    _say("sampled", len(tgts), "targets,", len(sub_t), "fragments")
    cfg = oracle.make_config(WB, 1, 1, len(tgts), 1)
    o64 = oracle.run_fp64(cfg, sub_t, sub_d, sub_c)["out"]
    got = outs[WB][tgts].astype(np.float64)
    err = np.abs(got - o64) / np.maximum(np.abs(o64), 1e-3)
    _say("WBOIT max rel err vs FP64 on the sample", err.max())
    assert err.max() <= 1e-5
    cfg = oracle.make_config(LL, 32, 1, len(tgts), 1, ll_capacity=len(sub_t) + 1)
    want = oracle.run(cfg, sub_t, sub_d, sub_c)["out"]
    assert np.array_equal(outs[LL][tgts].view(np.uint32), want.view(np.uint32))
    # WBOIT against the exact image (reported, not a bound: WBOIT is an approximation)
    _say("WBOIT vs Linked List K=32, mean abs difference", float(np.abs(outs[WB] - outs[LL]).mean()))
