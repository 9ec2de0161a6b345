"""Reconfiguration reuses device buffers: a context that serves frames of
changing size, K and algorithm keeps its allocations while they fit, stays
bit-exact with the oracle on every frame (the Linked List pool, the Loop64
and Loop32 slots, the WBOIT accumulators and the colour-pass tail queue are kept clean across sizes,
algorithms and frames cleared before their resolve), and oit_trim releases
everything."""
import numpy as np
import pytest

from conftest import deep_scene
from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig, Renderer

pytestmark = pytest.mark.gpu

LL, LOOP32, LOOP64, SPIN, WB = 1, 2, 3, 4, 6


def _frame(r, oracle, alg, K, tail, seed, w, h, spheres, pool_factor=1.5, unresolved_first=False):
    p, tg, d, c = deep_scene(seed=seed, width=w, height=h, spheres=spheres)
    F = len(tg)
    cap = int(F * pool_factor) + 1 if alg == LL else 0
    print("frame", alg, K, tail, w, h, F, flush=True)
    r.configure(FrameConfig(algorithm=alg, layers=K, tail_blend=bool(tail), width=w, height=h, ll_pool_nodes=cap))
    bt, bd, bc = DeviceBuffer.from_numpy(tg), DeviceBuffer.from_numpy(d), DeviceBuffer.from_numpy(c)
    out = DeviceBuffer((w * h, 4), "float32")
    if unresolved_first:  # leaves the pool dirty: the next clear must restore it
        r.clear()
        r.append(bt, bd, bc)
    r.render(bt, bd, bc, out)
    got = out.numpy()
    want = oracle.run(oracle.make_config(alg, K, tail, w, h, ll_capacity=cap), tg, d, c)["out"]
    return got, want


def _exact(got, want):
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_reconfigure_sequence_stays_exact(oracle):
    with Renderer(0) as r:
        seq = [
            (LL, 8, 1, 11, 64, 48, 160, 1.5, False),
            (LL, 16, 1, 12, 96, 80, 300, 1.5, False),   # grows
            (LL, 4, 1, 13, 40, 30, 100, 1.5, False),    # fits: reused, heads beyond T stay 0
            (LL, 8, 1, 14, 96, 80, 300, 0.5, True),     # overflowing pool, dirty frame first
            (LL, 8, 1, 15, 48, 40, 160, 1.5, False),    # after a dirty frame
            (LOOP64, 8, 1, 16, 64, 48, 160, 0, False),  # other algorithm: LL buffers released
            (LL, 16, 1, 17, 64, 48, 160, 1.5, False),   # LL again: fresh pool filled clean
            (SPIN, 4, 1, 18, 48, 40, 120, 0, False),
            (WB, 1, 0, 19, 48, 40, 120, 0, False),
            # tail queue: every clear empties it with the tail lists; an
            # unresolved frame leaves records the next frame must not link
            (SPIN, 4, 1, 20, 64, 48, 200, 0, True),
            (SPIN, 8, 1, 21, 48, 40, 160, 0, False),
            (LL, 8, 1, 22, 48, 40, 160, 0.5, True),
            (SPIN, 4, 1, 23, 64, 48, 200, 0, False),
            (LOOP64, 4, 1, 24, 64, 48, 200, 0, True),
            (SPIN, 2, 1, 25, 64, 48, 200, 0, False),
            # Loop64 slots: the composite empties them for the next frame of
            # the same shape; a reconfiguration or an unresolved frame refills
            (LOOP64, 8, 1, 26, 64, 48, 200, 0, False),
            (LOOP64, 8, 1, 27, 64, 48, 160, 0, False),
            (LOOP64, 8, 0, 28, 64, 48, 240, 0, False),
            (LOOP64, 16, 1, 29, 48, 40, 160, 0, False),
            (LOOP64, 16, 1, 30, 48, 40, 200, 0, True),
            (LOOP64, 4, 1, 31, 63, 37, 150, 0, False),
            # targets not a multiple of four: the clear's fills end in single words
            (SPIN, 4, 1, 32, 63, 37, 150, 0, False),
            (SPIN, 4, 1, 33, 63, 37, 150, 0, True),
            # Loop32 depth and colour slots, as the Loop64 slots
            (LOOP32, 8, 1, 34, 64, 48, 200, 0, False),
            (LOOP32, 8, 1, 35, 64, 48, 160, 0, False),
            (LOOP32, 16, 1, 36, 63, 37, 150, 0, True),
            (LOOP32, 4, 0, 37, 48, 40, 120, 0, False),
            (LOOP32, 4, 1, 38, 48, 40, 200, 0, False),
            # WBOIT accumulators: emptied by the composite, refilled after a
            # reconfiguration or an unresolved frame
            (WB, 1, 0, 39, 64, 48, 200, 0, False),
            (WB, 1, 0, 40, 64, 48, 160, 0, False),
            (WB, 1, 0, 41, 64, 48, 200, 0, True),
            (WB, 1, 0, 42, 63, 37, 150, 0, False),
            (WB, 1, 0, 43, 63, 37, 150, 0, False),
        ]
        resident = []
        for alg, K, tail, seed, w, h, sp, pf, dirty in seq:
            got, want = _frame(r, oracle, alg, K, tail, seed, w, h, sp, pf, dirty)
            if alg == WB:
                assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
            else:
                _exact(got, want)
            resident.append(r.sizes().resident_bytes)
        assert resident[2] == resident[1]  # the smaller Linked List frame reused the larger buffers
        r.trim()
        with pytest.raises(RuntimeError):
            r.clear()
