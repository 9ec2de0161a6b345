"""The request loop (oit_server_*): every request's image is the same when
it runs alone and when it runs in a mix of concurrent requests (bit-exact
for the Linked List, Loop64 and the ordered interlock, within 1e-5
relative for WBOIT, whose
float atomics accumulate in arrival order), the Linked List result matches
the oracle, the HBM budget is respected (idle contexts trimmed, oversized
requests refused) and bad requests are rejected at submission."""
import numpy as np
import pytest

from vk_order_independent_transparency_amd import DeviceBuffer, _abi, scene, serve

pytestmark = pytest.mark.gpu

SMALL_RES = ((128, 96), (200, 150), (256, 256), (97, 61))


def _say(*a):
    print(*a, flush=True)


def _requests(n=14, seed=3):
    reqs, _ = serve.served_workload(seed, n, 1000.0, resolutions=SMALL_RES, spheres=(40, 400),
                                    algorithms=(_abi.OIT_LINKEDLIST, _abi.OIT_LOOP64, _abi.OIT_WEIGHTED,
                                                _abi.OIT_SIMPLE, _abi.OIT_LOOP, _abi.OIT_INTERLOCK),
                                    layers=(4, 8, 16, 32))
    return reqs


def _outs(reqs):
    return [DeviceBuffer((r.scene.targets, 4), "float32") for r in reqs]


def _alone(reqs):
    outs = _outs(reqs)
    with serve.Server(0, slots=1) as s:
        for i, r in enumerate(reqs):
            s.submit(r, outs[i], tag=i)
            s.drain()
        done = s.poll(1024)
    assert len(done) == len(reqs) and all(c.status == 0 for c in done)
    return [o.numpy() for o in outs]


def test_mix_equals_alone():
    reqs = _requests()
    _say("requests", [(r.scene.width, r.scene.spheres, r.algorithm, r.layers) for r in reqs])
    alone = _alone(reqs)
    outs = _outs(reqs)
    with serve.Server(0, slots=4) as s:
        done = s.replay(reqs, np.zeros(len(reqs)), outs)  # one burst: all slots busy at once
        st = s.stats()
    assert all(c.status == 0 for c in done), [c.status for c in done]
    assert st.peak_in_flight == 4
    for i, r in enumerate(reqs):
        got = outs[i].numpy()
        if r.algorithm == _abi.OIT_WEIGHTED:
            assert np.allclose(got, alone[i], rtol=1e-5, atol=1e-7), i
        elif r.algorithm in (_abi.OIT_SIMPLE, _abi.OIT_LOOP):
            # Simple keeps the first K arrivals (README.md:23): pixels with at most K fragments do
            # not depend on arrival order. Loop32's depth slots hold the K nearest distinct depths:
            # pixels whose fragments all have distinct depths do not depend on it either.
            tg, d, _ = scene.generate(r.scene)
            T = r.scene.targets
            n = np.bincount(tg, minlength=T)
            if r.algorithm == _abi.OIT_SIMPLE:
                keep = n <= r.layers
            else:
                pairs = np.unique(np.stack([tg.astype(np.uint64), d.view(np.uint32).astype(np.uint64)]), axis=1)
                keep = np.bincount(pairs[0].astype(np.int64), minlength=T) == n
            assert keep.sum() > T // 2, i
            assert np.array_equal(got[keep].view(np.uint32), alone[i][keep].view(np.uint32)), i
        else:
            assert np.array_equal(got.view(np.uint32), alone[i].view(np.uint32)), i
    for c, r in zip(done, reqs):
        assert c.fragments == scene.count(r.scene)


def test_linked_list_request_matches_oracle(oracle):
    r = serve.Request(scene.SceneParams(seed=9, width=120, height=90, spheres=200, radius=(0.02, 0.1),
                                        alpha=(0.1, 0.6)), _abi.OIT_LINKEDLIST, 8)
    out = DeviceBuffer((r.scene.targets, 4), "float32")
    with serve.Server(0, slots=2) as s:
        s.submit(r, out)
        s.drain()
        (c,) = s.poll()
    assert c.status == 0
    tg, d, col = scene.generate(r.scene)
    cfg = oracle.make_config(oracle.LINKEDLIST, 8, 1, r.scene.width, r.scene.height, ll_capacity=len(tg) + 1)
    want = oracle.run(cfg, tg, d, col)["out"]
    assert np.array_equal(out.numpy().view(np.uint32), want.view(np.uint32))


def test_budget_is_respected():
    reqs = _requests(n=10, seed=5)
    need = []
    for r in reqs:
        F = scene.count(r.scene)
        T = r.scene.targets
        if r.algorithm == _abi.OIT_LINKEDLIST:
            need.append(4 * T + 12 * (F + 1) + 4 * T)
        elif r.algorithm == _abi.OIT_LOOP64:
            need.append(8 * r.layers * T + 4 * T + 8 * (-(-F // 64)) + 4 * T + 28 * F + 4 * (-(-F // 1024)))
        elif r.algorithm == _abi.OIT_SIMPLE:
            need.append(4 * T + 12 * r.layers * T + 4 * T + 28 * F + 4 * (-(-F // 1024)))
        elif r.algorithm == _abi.OIT_LOOP:
            need.append(12 * r.layers * T + 4 * T + 28 * F + 4 * (-(-F // 1024)))
        elif r.algorithm == _abi.OIT_INTERLOCK:
            need.append(4 * T + 28 * (F + 1))
        else:
            need.append(20 * T)
        need[-1] += (24 * F + 16 * T) * 5 // 4  # the slot's stream and output buffers (grown by a quarter)
    budget = int(sorted(need)[-1] * 1.2)  # room for the largest request, not for all at once
    big = serve.Request(scene.SceneParams(seed=1, width=2048, height=2048, spheres=300, radius=(0.01, 0.05),
                                          alpha=(0.1, 0.5)), _abi.OIT_LOOP64, 32)
    with serve.Server(0, slots=4, budget_bytes=budget) as s:
        done = s.replay(reqs + [big], np.zeros(len(reqs) + 1))
        st = s.stats()
    assert all(c.status == 0 for c in done[:-1]), [c.status for c in done]
    assert done[-1].status == _abi.OIT_ERR_OVER_BUDGET
    assert st.peak_held_bytes <= budget


def test_bad_requests_are_rejected():
    r = serve.Request(scene.SceneParams(seed=1, width=64, height=64, spheres=10, radius=(0.1, 0.2),
                                        alpha=(0.1, 0.2)), _abi.OIT_LINKEDLIST, 3)
    with serve.Server(0, slots=1) as s:
        with pytest.raises(_abi.OitError) as e:
            s.submit(r)
        assert e.value.status == _abi.OIT_ERR_INVALID_ARG
        with pytest.raises(_abi.OitError) as e:
            s.submit(serve.Request(r.scene, 9, 8))  # no such algorithm
        assert e.value.status == _abi.OIT_ERR_INVALID_ARG


def test_replay_keeps_other_callers_completions_and_reports_device_time():
    """A completion of a request submitted with oit_server_submit (tag 0, not yet polled) is not
    taken by a replay call, whose own requests carry tags 0.. as well; every completion carries
    its device time, which lies within its arrival-to-done interval."""
    reqs = _requests(n=3, seed=11)
    with serve.Server(0, slots=2) as s:
        s.submit(reqs[0], tag=0)
        s.drain()
        done = s.replay(reqs[1:], np.zeros(2))
        assert [c.tag for c in done] == [0, 1]
        assert [c.fragments for c in done] == [scene.count(r.scene) for r in reqs[1:]]
        (mine,) = s.poll()
        assert mine.tag == 0 and mine.fragments == scene.count(reqs[0].scene)
        for c in done + [mine]:
            assert c.status == 0
            assert 0 < c.device_us <= c.t_done_us - c.t_submit_us
