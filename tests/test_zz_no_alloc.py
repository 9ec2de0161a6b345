"""oit_ll_render_noalloc (one Linked List frame in the library's own device
array, no allocation) against the context path it varies: oit_clear +
oit_append + oit_resolve of an OIT_LINKEDLIST context with a pool of
count + 1 nodes. Bit for bit, on two calls in a row, the second with a
smaller frame (the array keeps the first call's contents), and once more
with the first frame."""
import ctypes

import numpy as np
import pytest

from vk_order_independent_transparency_amd import (OIT_LINKEDLIST, BatchFrame, DeviceBuffer, FrameConfig, OitError,
                                                   Renderer, Stream, _abi, render_noalloc, scene)


def _say(*a):
    print(*a, flush=True)


def _frame(p, layers, tail_blend, stream):
    g = scene.DeviceScene(p)
    t, d, c = g.buffers()
    g.emit(t, d, c, stream)
    stream.synchronize()
    n = g.fragments
    g.close()
    return BatchFrame(width=p.width, height=p.height, layers=layers, target=t, depth=d, rgba=c, count=n,
                      out=DeviceBuffer((p.targets, 4), "float32"), samples=p.samples, tail_blend=tail_blend)


def _context_image(f, stream):
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=f.layers, tail_blend=f.tail_blend, width=f.width,
                                height=f.height, samples=f.samples, ll_pool_nodes=f.count + 1,
                                background=f.background))
        o = DeviceBuffer((f.width * f.height * f.samples, 4), "float32")
        r.render(f.target, f.depth, f.rgba, o, stream, count=f.count)
        stream.synchronize()
        return o.numpy()


def _noalloc_image(f, stream):
    f.out.upload(np.full((f.width * f.height * f.samples, 4), np.nan, np.float32), stream)  # ordered before the frame
    render_noalloc(f, stream)
    stream.synchronize()
    return f.out.numpy()


@pytest.mark.gpu
def test_noalloc_equals_context_path_twice_then_smaller():
    st = Stream()
    big = _frame(scene.SceneParams(seed=1, width=1920, height=1080, spheres=1000, radius=(0.01, 0.05),
                                   alpha=(0.1, 0.5)), 16, True, st)
    small = _frame(scene.SceneParams(seed=1, width=256, height=256, spheres=64, radius=(0.02, 0.08),
                                     alpha=(0.2, 0.6)), 8, True, st)
    deep = _frame(scene.SceneParams(seed=3, width=96, height=80, spheres=400, radius=(0.05, 0.2), alpha=(0.1, 0.6),
                                    samples=4), 4, False, st)
    for name, f in (("1920x1080 K=16", big), ("256x256 K=8", small), ("96x80x4 K=4 no tail", deep),
                    ("1920x1080 K=16 again", big)):
        _say("noalloc frame", name, f.count, "fragments")
        want = _context_image(f, st)
        got = _noalloc_image(f, st)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name


@pytest.mark.gpu
def test_noalloc_rejects_frames_beyond_the_array():
    st = Stream()
    # 23 M fragments need 12 B each of the 256 MiB array: rejected on the host before any launch;
    # the buffers are real so that nothing could be read out of bounds
    F = 23_000_000
    t = DeviceBuffer((F,), "uint32")
    d = DeviceBuffer((F,), "float32")
    c = DeviceBuffer((F, 4), "float32")
    f = BatchFrame(width=64, height=64, layers=8, target=t, depth=d, rgba=c, count=F,
                   out=DeviceBuffer((64 * 64, 4), "float32"))
    with pytest.raises(OitError) as e:
        render_noalloc(f, st)
    assert e.value.status == _abi.OIT_ERR_CAPACITY


def test_noalloc_host_checks_without_gpu():
    """Sizes and arguments are checked on the host before any HIP call, so these
    run without a GPU: nothing here reaches the device (the pointers are never used)."""
    lib = _abi.lib()

    def call(**kw):
        f = _abi.OitBatchFrame()
        f.width, f.height, f.samples, f.layers, f.tail_blend = 64, 64, 1, 8, 1
        f.frags.target, f.frags.depth, f.frags.rgba, f.frags.count = 0x1000, 0x2000, 0x3000, 1000
        f.out = 0x4000
        for k, v in kw.items():
            if k in ("count", "coverage"):
                setattr(f.frags, k, v)
            else:
                setattr(f, k, v)
        return lib.oit_ll_render_noalloc(0, ctypes.byref(f), None)

    # 4 KiB + 4 * 4096 heads + 12 * (F + 1) nodes against 256 MiB
    limit = (_abi.NOALLOC_SCRATCH_BYTES - 4096 - 256 - 4 * 4096) // 12 - 256
    assert call(count=23_000_000) == _abi.OIT_ERR_CAPACITY
    assert call(count=limit + 1024) == _abi.OIT_ERR_CAPACITY
    assert call(width=8192, height=8192, count=0) == _abi.OIT_ERR_CAPACITY  # 256 MiB of heads alone
    assert call(layers=3) == _abi.OIT_ERR_INVALID_ARG
    assert call(width=0) == _abi.OIT_ERR_INVALID_ARG
    assert call(out=0x4004) == _abi.OIT_ERR_INVALID_ARG  # float4 stores need 16-byte alignment
    assert call(coverage=0x5000) == _abi.OIT_ERR_UNSUPPORTED
    assert lib.oit_ll_render_noalloc(0, None, None) == _abi.OIT_ERR_INVALID_ARG
    assert call(count=100, **{"reserved": 1}) == _abi.OIT_ERR_INVALID_ARG
