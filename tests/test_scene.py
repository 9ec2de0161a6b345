"""The synthetic sphere stream: an independent numpy restatement of the
generator's definition (splitmix64, binary32 arithmetic) must reproduce the
native generator bit for bit on small scenes; plus invariants of the stream."""
import numpy as np
import pytest

from vk_order_independent_transparency_amd import scene

M64 = (1 << 64) - 1
POS = {
    1: [(0.5, 0.5)],
    4: [(0.375, 0.125), (0.875, 0.375), (0.125, 0.625), (0.625, 0.875)],
}


class SplitMix:
    def __init__(self, seed):
        self.s = seed & M64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def uniform(self):
        return np.float32(self.next() >> 40) * np.float32(2.0 ** -24)

    def range(self, lo, hi):
        lo, hi = np.float32(lo), np.float32(hi)
        return np.float32(lo + self.uniform() * np.float32(hi - lo))


def reference_stream(p: scene.SceneParams):
    f32 = np.float32
    rng = SplitMix(p.seed)
    ext = f32(p.cluster_extent)
    ox = rng.range(0.0, f32(1.0) - ext)
    oy = rng.range(0.0, f32(1.0) - ext)
    ncl = int(np.floor(p.spheres * float(np.float32(p.cluster_fraction))))
    W, H, S = p.width, p.height, p.samples
    xs = np.arange(W, dtype=np.float32)
    ys = np.arange(H, dtype=np.float32)
    out_t, out_d, out_c = [], [], []
    for i in range(p.spheres):
        cx, cy, cz = rng.uniform(), rng.uniform(), rng.uniform()
        r = rng.range(p.radius[0], p.radius[1])
        R, G, B = rng.uniform(), rng.uniform(), rng.uniform()
        a = rng.range(p.alpha[0], p.alpha[1])
        if i < ncl:
            cx = f32(ox + cx * ext)
            cy = f32(oy + cy * ext)
        r2 = f32(r * r)
        col = np.float32([R * a, G * a, B * a, a])
        # all pixels and samples, in (y, x, s) order
        px = np.stack([(xs + f32(sx)) / f32(W) for sx, _ in POS[S]], axis=-1)  # [W, S]
        py = np.stack([(ys + f32(sy)) / f32(H) for _, sy in POS[S]], axis=-1)  # [H, S]
        dx = (px - cx).astype(np.float32)
        dy = (py - cy).astype(np.float32)
        d2 = (dx[None, :, :] * dx[None, :, :] + dy[:, None, :] * dy[:, None, :]).astype(np.float32)  # [H, W, S]
        hit = d2 < r2
        yy, xx, ss = np.nonzero(hit)
        h = np.sqrt((r2 - d2[hit]).astype(np.float32)).astype(np.float32)
        t = (ss * W * H + yy * W + xx).astype(np.uint32)
        front = (cz - h).astype(np.float32)
        back = (cz + h).astype(np.float32)
        out_t.append(np.repeat(t, 2))
        out_d.append(np.stack([front, back], axis=1).reshape(-1))
        out_c.append(np.tile(col, (2 * len(t), 1)))
    return (np.concatenate(out_t), np.concatenate(out_d).astype(np.float32),
            np.concatenate(out_c).astype(np.float32))


@pytest.mark.parametrize("p", [
    scene.SceneParams(seed=1, width=64, height=64, spheres=16, radius=(0.02, 0.08), alpha=(0.2, 0.6)),
    scene.SceneParams(seed=9, width=40, height=24, spheres=30, radius=(0.01, 0.2), alpha=(0.1, 0.5)),
    scene.SceneParams(seed=5, width=32, height=32, spheres=20, radius=(0.03, 0.1), alpha=(0.05, 0.4), samples=4),
    scene.SceneParams(seed=7, width=48, height=48, spheres=50, radius=(0.01, 0.05), alpha=(0.1, 0.5),
                      cluster_fraction=0.2, cluster_extent=0.1),
])
def test_native_generator_matches_restatement(p):
    tg, d, c = scene.generate(p)
    rt, rd, rc = reference_stream(p)
    assert len(tg) == len(rt) == scene.count(p)
    assert np.array_equal(tg, rt)
    assert np.array_equal(d.view(np.uint32), rd.view(np.uint32))
    assert np.array_equal(c.view(np.uint32), rc.view(np.uint32))


def test_stream_invariants():
    p = scene.SMALL_256
    tg, d, c = scene.generate(p)
    assert len(tg) % 2 == 0
    front, back = d[0::2], d[1::2]
    assert (tg[0::2] == tg[1::2]).all()
    assert (front < back).all()
    assert (tg < p.targets).all()
    assert (c[:, 3] >= 0.2).all() and (c[:, 3] < 0.6).all()
    assert (c[:, :3] <= c[:, 3:4]).all()
    # some front depths fall below 0 (z - h with z near 0): the depth key must order them
    assert (front < 0).any()


def test_seeds_differ_and_repeat():
    a = scene.generate(scene.SMALL_256)[1]
    b = scene.generate(scene.SMALL_256)[1]
    other = scene.generate(scene.SceneParams(**{**scene.SMALL_256.__dict__, "seed": 2}))[1]
    assert np.array_equal(a, b)
    assert len(other) != len(a) or not np.array_equal(a, other)


def test_fragment_count_near_expectation():
    """E[F] = 2 * P * sum of disc areas (edge clipping aside) for the 1080p workload."""
    p = scene.HD_1080P
    F = scene.count(p)
    r0, r1 = p.radius
    mean_r2 = (r1 ** 3 - r0 ** 3) / (3 * (r1 - r0))
    expect = 2 * p.width * p.height * p.spheres * np.pi * mean_r2
    assert 0.85 * expect < F < 1.05 * expect


def test_invalid_params_rejected():
    from vk_order_independent_transparency_amd import OitError

    with pytest.raises(OitError):
        scene.count(scene.SceneParams(seed=1, width=0, height=4, spheres=1, radius=(0.1, 0.2), alpha=(0.1, 0.2)))
    with pytest.raises(OitError):
        scene.count(scene.SceneParams(seed=1, width=4, height=4, spheres=1, radius=(0.1, 0.2), alpha=(0.1, 0.2),
                                      samples=3))
