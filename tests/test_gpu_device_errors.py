"""Bounded device waits: a corrupted structure planted from the host (a head
that points at a node whose link is never written, a Spinlock lock that is
never released, a list that loops back on itself) ends the frame's kernels
with OIT_ERR_DEVICE from oit_get_stats, with the bit of the wait that gave
up, instead of hanging the queue; the next frame on the same context is
bit-exact against the oracle again. Every kernel here exits normally: each
wait gives up after its trip limit, each walk at its first non-descending
link."""
import ctypes

import numpy as np
import pytest

from vk_order_independent_transparency_amd import (OIT_LINKEDLIST, OIT_SPINLOCK, DeviceBuffer, FrameConfig, Renderer,
                                                   _abi, scene)

pytestmark = pytest.mark.gpu

ERR_LIST_WAIT, ERR_LOCK_WAIT, ERR_WALK = 1, 2, 4


def _say(*a):
    print(*a, flush=True)


def _scene():
    p = scene.SceneParams(seed=2, width=64, height=48, spheres=96, radius=(0.04, 0.16), alpha=(0.1, 0.6))
    tg, d, c = scene.generate(p)
    return p, tg, d, c


def _stats(r):
    out = _abi.OitStats()
    rc = _abi.lib().oit_get_stats(r._ctx, ctypes.byref(out), None)
    return rc, out


def _poke(r, what, index, value):
    _abi.check("oit_debug_set_word", _abi.lib().oit_debug_set_word(r._ctx, what, index, value, None))


def _clean_frame_matches_oracle(r, oracle, alg, K, p, bufs, cap):
    tg_d, d_d, c_d, tg, d, c = bufs
    out = DeviceBuffer((p.targets, 4), np.float32)
    r.render(tg_d, d_d, c_d, out)
    rc, st = _stats(r)
    assert rc == 0 and st.device_errors == 0
    cfg = oracle.make_config(alg, K, 1, p.width, p.height, ll_capacity=cap) if alg == OIT_LINKEDLIST else \
        oracle.make_config(alg, K, 1, p.width, p.height)
    want = oracle.run(cfg, tg, d, c)["out"]
    assert np.array_equal(out.numpy().view(np.uint32), want.view(np.uint32))


def _bufs(tg, d, c):
    return (DeviceBuffer.from_numpy(tg), DeviceBuffer.from_numpy(d), DeviceBuffer.from_numpy(c), tg, d, c)


def test_list_insert_and_walk_give_up_on_an_unwritten_link(oracle):
    p, tg, d, c = _scene()
    F = len(tg)
    cap = F + 1 + 64  # nodes F+1 .. cap-1 are never appended: their links stay kUnset
    K = 8
    bufs = _bufs(tg, d, c)
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=K, tail_blend=True, width=p.width, height=p.height,
                                ll_pool_nodes=cap))
        _clean_frame_matches_oracle(r, oracle, OIT_LINKEDLIST, K, p, bufs, cap)
        t0 = int(tg[0])
        _say("planting head", t0, "->", cap - 1)
        out = DeviceBuffer((p.targets, 4), np.float32)
        r.clear()
        _poke(r, _abi.DEBUG_LL_HEAD, t0, cap - 1)  # inserts at t0 wait behind node cap-1's link forever
        r.append(bufs[0], bufs[1], bufs[2])
        r.resolve(out)
        rc, st = _stats(r)
        _say("status", rc, "device_errors", st.device_errors)
        assert rc == _abi.OIT_ERR_DEVICE
        assert st.device_errors & ERR_LIST_WAIT and st.device_errors & ERR_WALK
        rc, st = _stats(r)  # reported once
        assert rc == 0 and st.device_errors == 0
        _clean_frame_matches_oracle(r, oracle, OIT_LINKEDLIST, K, p, bufs, cap)


def test_list_walk_gives_up_on_a_cycle(oracle):
    p, tg, d, c = _scene()
    F = len(tg)
    cap = F + 1
    K = 4
    bufs = _bufs(tg, d, c)
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=K, tail_blend=True, width=p.width, height=p.height,
                                ll_pool_nodes=cap))
        out = DeviceBuffer((p.targets, 4), np.float32)
        r.clear()
        r.append(bufs[0], bufs[1], bufs[2])
        node = 7  # fragment 6's node: its link now points at itself
        _poke(r, _abi.DEBUG_LL_NEXT, node, node)
        r.resolve(out)
        rc, st = _stats(r)
        _say("status", rc, "device_errors", st.device_errors)
        assert rc == _abi.OIT_ERR_DEVICE and st.device_errors == ERR_WALK
        _clean_frame_matches_oracle(r, oracle, OIT_LINKEDLIST, K, p, bufs, cap)


def test_spinlock_gives_up_on_a_lock_never_released(oracle):
    p, tg, d, c = _scene()
    K = 4
    bufs = _bufs(tg, d, c)
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=OIT_SPINLOCK, layers=K, tail_blend=True, width=p.width, height=p.height))
        # fragments below the K-th of a pixel are stored without the lock: the
        # planted lock is on the deepest pixel, whose later fragments wait for it
        per = np.bincount(tg.astype(np.int64), minlength=p.targets)
        t0 = int(per.argmax())
        assert per[t0] > K
        out = DeviceBuffer((p.targets, 4), np.float32)
        r.clear()
        _poke(r, _abi.DEBUG_SPIN_LOCK, t0, 1)  # held, count 0
        r.append(bufs[0], bufs[1], bufs[2])
        r.resolve(out)
        rc, st = _stats(r)
        _say("status", rc, "device_errors", st.device_errors)
        assert rc == _abi.OIT_ERR_DEVICE and st.device_errors == ERR_LOCK_WAIT
        _clean_frame_matches_oracle(r, oracle, OIT_SPINLOCK, K, p, bufs, 0)


def test_debug_hook_checks_its_index():
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=4, tail_blend=True, width=8, height=8,
                                ll_pool_nodes=100))
        assert _abi.lib().oit_debug_set_word(r._ctx, _abi.DEBUG_LL_HEAD, 1 << 40, 0, None) == _abi.OIT_ERR_INVALID_ARG
        assert _abi.lib().oit_debug_set_word(r._ctx, _abi.DEBUG_SPIN_LOCK, 0, 0, None) == _abi.OIT_ERR_INVALID_ARG
        assert _abi.lib().oit_debug_set_word(r._ctx, 9, 0, 0, None) == _abi.OIT_ERR_INVALID_ARG
