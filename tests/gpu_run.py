"""Runs one frame through the HIP path (C ABI via the package) and returns
numpy arrays shaped like the oracle's outputs. Used by the GPU tests only.
Device memory comes from the library's own C ABI (no other HIP runtime is
loaded into the test process)."""
from __future__ import annotations

import numpy as np


def run_gpu(algorithm, layers, tail_blend, width, height, target, depth, rgba, samples=1, ll_pool_nodes=0,
            splits=1, background=(0.5, 0.5, 0.5, 1.0), max_fragments=0, renderer=None, variant=0,
            coverage_samples=0, mask=None, opaque=None, opaque_splits=1, bounds=None):
    from vk_order_independent_transparency_amd import DeviceBuffer, FrameConfig, Renderer

    T = width * height * samples
    K = layers
    own = renderer is None
    r = renderer or Renderer(0)
    try:
        cfg = FrameConfig(algorithm=algorithm, layers=layers, tail_blend=bool(tail_blend), width=width,
                          height=height, samples=samples, ll_pool_nodes=ll_pool_nodes,
                          max_fragments=max_fragments, background=background, variant=variant,
                          coverage_samples=coverage_samples)
        r.configure(cfg)
        F = int(len(target))
        tg = DeviceBuffer.from_numpy(np.ascontiguousarray(target, dtype=np.uint32).reshape(F))
        dp = DeviceBuffer.from_numpy(np.ascontiguousarray(depth, dtype=np.float32).reshape(F))
        cl = DeviceBuffer.from_numpy(np.ascontiguousarray(rgba, dtype=np.float32).reshape(F, 4))
        mk = None if mask is None else DeviceBuffer.from_numpy(np.ascontiguousarray(mask, dtype=np.uint32).reshape(F))
        out = DeviceBuffer((T, 4), np.float32)
        fk_key = DeviceBuffer((T, K), np.uint32)
        fk_val = DeviceBuffer((T, K), np.uint32)
        counts = DeviceBuffer((T,), np.uint32)
        r.clear()
        keep = []
        if opaque is not None:  # (target, depth, rgba, mask or None): the frame's opaque draws
            otg, od, oc, om = opaque
            G = int(len(otg))
            ob = [DeviceBuffer.from_numpy(np.ascontiguousarray(otg, dtype=np.uint32).reshape(G)),
                  DeviceBuffer.from_numpy(np.ascontiguousarray(od, dtype=np.float32).reshape(G)),
                  DeviceBuffer.from_numpy(np.ascontiguousarray(oc, dtype=np.float32).reshape(G, 4))]
            ob.append(None if om is None else
                      DeviceBuffer.from_numpy(np.ascontiguousarray(om, dtype=np.uint32).reshape(G)))
            keep.append(ob)
            obounds = np.linspace(0, G, opaque_splits + 1).astype(np.int64)
            for a, b in zip(obounds[:-1], obounds[1:]):
                n = int(b - a)
                r.draw_opaque(ob[0].view(a, n), ob[1].view(a, n), ob[2].view(a, n), count=n,
                              coverage=None if ob[3] is None else ob[3].view(a, n))
        # appends: `splits` equal parts, or the given first ids (`bounds`, from 0 to F)
        bounds = np.linspace(0, F, splits + 1).astype(np.int64) if bounds is None else np.asarray(bounds, np.int64)
        for a, b in zip(bounds[:-1], bounds[1:]):
            n = int(b - a)
            r.append(tg.view(a, n), dp.view(a, n), cl.view(a, n), count=n,
                     coverage=None if mk is None else mk.view(a, n))
        r.resolve_debug(out, fk_key, fk_val, counts)
        st = r.stats()
        return {
            "out": out.numpy(),
            "frontk_key": fk_key.numpy(),
            "frontk_val": fk_val.numpy(),
            "count": counts.numpy(),
            "ll_overflow": int(st.ll_overflow),
            "tail_fragments": int(st.tail_fragments),
            "fragments": int(st.fragments),
        }
    finally:
        if own:
            r.close()


def simple_replay(oracle, cfg, stored, tg, d, c, K, T):
    """Simple's stored set is the first K fragments to arrive: the GPU's stored
    set goes through the oracle's composite, the other fragments as its tail."""
    n = np.bincount(np.asarray(tg, np.int64), minlength=T)
    tail_flag = np.ones(len(tg), np.uint8)
    ids = stored[stored != 0xFFFFFFFF].astype(np.int64)
    assert len(np.unique(ids)) == len(ids)
    assert np.array_equal(np.sort(tg[ids]), np.sort(np.repeat(np.arange(T), np.minimum(n, K))))
    tail_flag[ids] = 0
    ab_col = np.zeros((T, K), np.uint32)
    ab_key = np.zeros((T, K), np.uint32)
    ab_id = np.zeros((T, K), np.uint32)
    for t in np.nonzero(n > 0)[0]:
        for k in range(min(n[t], K)):
            i = int(stored[t, k])
            ab_col[t, k] = oracle.encode(c[i])
            ab_key[t, k] = oracle.depth_key(float(d[i]))
            ab_id[t, k] = i
    return oracle.replay_kbuffer(cfg, tg, c, n.astype(np.uint32), ab_col, ab_key, ab_id, tail_flag,
                                 tail_quantized=False)
