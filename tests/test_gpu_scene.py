"""The device scene generator (oit_scene_gen_*) against the host generator:
the same fragment stream, bit for bit and in the same order, for every
sample count, clustered scenes, edge-clipped spheres and the 1920x1080
1000-sphere workload; plus its error returns."""
import numpy as np
import pytest

from vk_order_independent_transparency_amd import _abi, scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _device_stream(p, stream=None):
    g = scene.DeviceScene(p)
    t, d, c = g.buffers()
    g.emit(t, d, c, stream)
    n = g.fragments
    out = t.numpy()[:n], d.numpy()[:n], c.numpy()[:n]
    g.close()
    return out


def _assert_identical(p):
    ht, hd, hc = scene.generate(p)
    _say("scene", p, len(ht))
    dt, dd, dc = _device_stream(p)
    assert len(dt) == len(ht)
    assert np.array_equal(dt, ht), f"target differs at {np.argmax(dt != ht)}"
    assert np.array_equal(dd.view(np.uint32), hd.view(np.uint32)), f"depth differs at {np.argmax(dd != hd)}"
    assert np.array_equal(dc.view(np.uint32), hc.view(np.uint32))


@pytest.mark.parametrize("samples", [1, 2, 4, 8])
def test_device_stream_equals_host_stream(samples):
    _assert_identical(scene.SceneParams(seed=11, width=97, height=61, spheres=120, radius=(0.02, 0.3),
                                        alpha=(0.1, 0.9), samples=samples))


def test_device_stream_clustered_and_wide_rows():
    # radius 0.6 spheres span rows wider than a wave (more than 64 columns)
    _assert_identical(scene.SceneParams(seed=5, width=300, height=40, spheres=30, radius=(0.1, 0.6),
                                        alpha=(0.2, 0.5), cluster_fraction=0.5, cluster_extent=0.1))


def test_device_stream_small_256():
    _assert_identical(scene.SMALL_256)


def test_device_stream_hd_1080p():
    _assert_identical(scene.HD_1080P)


def test_device_emit_repeats_and_checks_capacity():
    p = scene.SceneParams(seed=2, width=64, height=64, spheres=40, radius=(0.02, 0.1), alpha=(0.2, 0.6))
    g = scene.DeviceScene(p)
    assert g.fragments == scene.count(p)
    t, d, c = g.buffers()
    g.emit(t, d, c)
    first = t.numpy().copy(), d.numpy().copy()
    g.emit(t, d, c)
    assert np.array_equal(first[0], t.numpy()) and np.array_equal(first[1].view(np.uint32),
                                                                 d.numpy().view(np.uint32))
    with pytest.raises(_abi.OitError) as e:
        g.emit(t, d, c, capacity=g.fragments - 1)
    assert e.value.status == _abi.OIT_ERR_CAPACITY
    with pytest.raises(_abi.OitError) as e:
        g.emit(t.ptr + 4, d, c)
    assert e.value.status == _abi.OIT_ERR_INVALID_ARG
    g.close()


def test_device_scene_without_spheres_is_empty():
    g = scene.DeviceScene(scene.SceneParams(seed=1, width=16, height=16, spheres=0, radius=(0.1, 0.2),
                                            alpha=(0.1, 0.2)))
    assert g.fragments == 0
    g.emit(0, 0, 0)
    g.close()
