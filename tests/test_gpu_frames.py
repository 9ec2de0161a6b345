"""Frames back to back on one configured context, one of them abandoned
(cleared and appended, never resolved): every composite leaves the state it
read at its clear value (Linked List heads and links, Loop64 and Loop32
slots and counts, the Simple count, the Spinlock lock words and bounds, the
WBOIT accumulators, the colour-pass tail lists), so a frame's clear fills
only what an abandoned frame or a reconfiguration left behind. Every
resolved frame must equal the oracle: bit for bit (Simple where its stored
set is not arrival dependent, and by replay elsewhere), WBOIT within 1e-5
relative of the FP64 oracle."""
import numpy as np
import pytest

from conftest import deep_scene
from gpu_run import simple_replay
from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig, Renderer

pytestmark = pytest.mark.gpu

SIMPLE, LL, LOOP, LOOP64, SPIN, INTERLOCK, WB = 0, 1, 2, 3, 4, 5, 6

# (seed, spheres, resolved): different depth complexity frame to frame
FRAMES = [(51, 200, True), (52, 90, True), (53, 260, False), (54, 140, True), (55, 60, True), (56, 220, True)]


@pytest.mark.parametrize("alg", [SIMPLE, LL, LOOP, LOOP64, SPIN, INTERLOCK, WB])
@pytest.mark.parametrize("K", [4, 16])
def test_frames_on_one_context(oracle, alg, K):
    scenes = [deep_scene(seed=s, spheres=n) for s, n, _ in FRAMES]
    W, H = scenes[0][0].width, scenes[0][0].height
    T = W * H
    cap = max(len(x[1]) for x in scenes) + 1 if alg == LL else 0
    tail = alg != WB
    layers = 1 if alg == WB else K
    with Renderer(0) as r:
        r.configure(FrameConfig(algorithm=alg, layers=layers, tail_blend=tail, width=W, height=H, ll_pool_nodes=cap,
                                max_fragments=max(len(x[1]) for x in scenes)))
        out = DeviceBuffer((T, 4), np.float32)
        fk_val = DeviceBuffer((T, layers), np.uint32)
        for (p, tg, d, c), (seed, _, resolved) in zip(scenes, FRAMES):
            print("frame", alg, K, seed, len(tg), "resolved" if resolved else "abandoned", flush=True)
            bufs = [DeviceBuffer.from_numpy(np.ascontiguousarray(a)) for a in (tg, d, c)]
            r.clear()
            r.append(*bufs)
            if not resolved:
                continue
            if alg == SIMPLE:
                r.resolve_debug(out, frontk_val=fk_val)  # the stored set, for the replay
            else:
                r.resolve(out)
            got = out.numpy()
            cfg = oracle.make_config(alg, layers, int(tail), W, H, ll_capacity=cap)
            if alg == WB:
                want = oracle.run_fp64(cfg, tg, d, c)["out"]
                err = np.abs(got.astype(np.float64) - want) / np.maximum(np.abs(want), 1e-3)
                assert err.max() <= 1e-5, (seed, err.max())
                continue
            want = oracle.run(cfg, tg, d, c)["out"]
            if alg == SIMPLE:  # the first K to arrive are stored: replayed through the oracle's composite
                want = simple_replay(oracle, cfg, fk_val.numpy(), tg, d, c, K, T)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), seed
