"""The opaque pass and the depth-test culling of transparent fragments in the
oracle (oitRender.cpp:68-122, oitColorDepthDefines.glsl:24-37), pinned by
brute force and by what the mathematics fixes:
- the z-buffer equals, per sample target, the (depth key, id)-smallest
  fragment nearer than the clear depth 1.0 (a lexicographic sort, not a
  sequential loop): LESS keeps the earliest of equal depths;
- a culled fragment contributes nothing: the composite of a frame with an
  opaque pass equals the composite of the same frame without one over the
  stream with the culled fragments removed;
- without tail blending the colour buffer is the opaque colour;
- coverage shading: post_depth_coverage culls per sample, so the front
  composite and the colour buffer equal the averages over samples of
  per-sample frames whose sample s is tested against opaque plane s, and the
  image is that averaged composite blended over the averaged colour buffer:
  the composite pass runs once per pixel and its output is blended onto
  every sample before the MSAA resolve (oitLinkedList.frag.glsl:147-172,
  main.cpp:700-714), so it is not the average of per-sample images once the
  opaque colour differs between samples."""
import numpy as np
import pytest

BG = np.float32([0.5, 0.5, 0.5, 1.0])


def _stream(rng, T, n, dmax=1.0):
    tg = rng.integers(0, T, n).astype(np.uint32)
    d = rng.uniform(0, dmax, n).astype(np.float32)
    a = rng.uniform(0.1, 0.7, n).astype(np.float32)
    rgb = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    c = np.concatenate([rgb * a[:, None], a[:, None]], 1).astype(np.float32)
    return tg, d, c


def _opaque_stream(rng, T, n):
    tg = rng.integers(0, T, n).astype(np.uint32)
    d = rng.uniform(0, 1.2, n).astype(np.float32)  # some at or behind the clear depth
    d[: n // 8] = d[n // 8: 2 * (n // 8)]  # repeated depths: ties
    tg[: n // 8] = tg[n // 8: 2 * (n // 8)]
    c = np.concatenate([rng.uniform(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    return tg, d, c


def test_opaque_pass_is_the_nearest_earliest_fragment(oracle):
    rng = np.random.default_rng(3)
    W, H = 9, 7
    T = W * H
    tg, d, c = _opaque_stream(rng, T, 400)
    cfg = oracle.make_config(oracle.LINKEDLIST, 4, 1, W, H)
    keys, cols = oracle.opaque_pass(cfg, tg, d, c)
    k = np.array([oracle.depth_key(float(x)) for x in d], np.uint32)
    clear = oracle.depth_key(1.0)
    for t in range(T):
        idx = np.nonzero((tg == t) & (k < clear))[0]
        if len(idx) == 0:
            assert keys[t] == clear and np.array_equal(cols[t], BG)
            continue
        order = np.lexsort((idx, k[idx]))  # by key, then id
        w = idx[order[0]]
        assert keys[t] == k[w] and np.array_equal(cols[t], c[w]), t
    # a tie: the first of two equal depths stays
    cfg1 = oracle.make_config(oracle.LINKEDLIST, 4, 1, 1, 1)
    two = np.float32([[0.1, 0.2, 0.3, 1.0], [0.9, 0.8, 0.7, 1.0]])
    _, cc = oracle.opaque_pass(cfg1, np.uint32([0, 0]), np.float32([0.25, 0.25]), two)
    assert np.array_equal(cc[0], two[0])
    # at the clear depth nothing is written
    kk, cc = oracle.opaque_pass(cfg1, np.uint32([0]), np.float32([1.0]), two[:1])
    assert kk[0] == oracle.depth_key(1.0) and np.array_equal(cc[0], BG)


@pytest.mark.parametrize("alg", [0, 1, 2, 3, 4, 5, 6])
def test_culled_fragments_contribute_nothing(oracle, alg):
    rng = np.random.default_rng(11 + alg)
    W, H, K = 6, 5, 8
    T = W * H
    otg, od, oc = _opaque_stream(rng, T, 60)
    tg, d, c = _stream(rng, T, 500, dmax=1.1)
    base = oracle.make_config(alg, K, 0, W, H, ll_capacity=10000)
    opq = oracle.opaque_pass(base, otg, od, oc)
    k = np.array([oracle.depth_key(float(x)) for x in d], np.uint32)
    live = k < opq[0][tg]
    assert 0 < live.sum() < len(tg)
    with_opq = oracle.run(oracle.make_config(alg, K, 0, W, H, ll_capacity=10000, opaque=opq), tg, d, c)
    filtered = oracle.run(base, tg[live], d[live], c[live])
    if alg == oracle.WEIGHTED:
        assert np.array_equal(with_opq["wb_accum"], filtered["wb_accum"])
        assert np.array_equal(with_opq["wb_reveal"], filtered["wb_reveal"])
        r = with_opq["wb_reveal"][:, None]
        comp = with_opq["composite"]
        want = (comp * (np.float32(1) - r) + opq[1] * r).astype(np.float32)
        assert np.array_equal(with_opq["out"][:, :3], want[:, :3])
    else:
        assert np.array_equal(with_opq["composite"], filtered["composite"])
        assert np.array_equal(with_opq["colorbuf"], opq[1])  # no tail blending: the opaque colour
        assert np.array_equal(with_opq["count"], filtered["count"])


@pytest.mark.parametrize("alg", [0, 1, 3, 4])
def test_tail_fragments_blend_over_the_opaque_colour(oracle, alg):
    """K=1 and two fragments in front of the opaque surface: the second goes to
    a colour-pass tail (or is evicted) and blends over the opaque colour; one
    behind it does not appear anywhere."""
    opq_c = np.float32([0.2, 0.3, 0.1, 1.0])
    opq = (np.uint32([oracle.depth_key(0.6)]), opq_c[None, :])
    c = np.float32([[0.2, 0.1, 0.05, 0.4], [0.05, 0.3, 0.1, 0.5], [0.3, 0.3, 0.3, 0.6]])
    d = np.float32([0.5, 0.2, 0.7])  # the third is behind the opaque depth
    cfg = oracle.make_config(alg, 1, 1, 1, 1, ll_capacity=2, opaque=opq)  # Linked List: node 2 overflows
    r = oracle.run(cfg, np.uint32([0, 0, 0]), d, c)
    r2 = oracle.run(cfg, np.uint32([0, 0]), d[:2], c[:2])
    assert np.array_equal(r["out"], r2["out"]) and r["tail_fragments"] == r2["tail_fragments"]
    assert not np.array_equal(r["colorbuf"][0], opq_c)  # something went to the tail over the opaque colour


@pytest.mark.parametrize("alg,S", [(1, 2), (1, 4), (0, 4), (4, 4), (5, 4), (1, 8)])
def test_coverage_culls_per_sample(oracle, alg, S):
    rng = np.random.default_rng(41 + S + alg)
    W, H, K = 5, 4, 32
    P = W * H
    # opaque fragments with masks, and their per-sample expansion: same planes
    otg, od, oc = _opaque_stream(rng, P, 40)
    om = rng.integers(1, 1 << S, len(otg)).astype(np.uint32)
    cfg_cov = oracle.make_config(alg, K, 1, W, H, coverage_samples=S, ll_capacity=100000)
    opq = oracle.opaque_pass(cfg_cov, otg, od, oc, mask=om)
    ids, tgs = [], []
    for i in range(len(otg)):
        for j in range(S):
            if (om[i] >> j) & 1:
                ids.append(i)
                tgs.append(j * P + otg[i])
    cfg_s = oracle.make_config(alg, K, 1, W, H, samples=S, ll_capacity=100000)
    opq_s = oracle.opaque_pass(cfg_s, np.array(tgs, np.uint32), od[ids], oc[ids])
    assert np.array_equal(opq[0], opq_s[0]) and np.array_equal(opq[1], opq_s[1])
    # transparent frame: coverage with the opaque planes vs the average of per-sample frames
    tg, d, c = _stream(rng, P, 120, dmax=1.1)
    m = rng.integers(1, 1 << S, len(tg)).astype(np.uint32)
    r = oracle.run(oracle.make_config(alg, K, 1, W, H, coverage_samples=S, ll_capacity=100000, opaque=opq),
                   tg, d, c, mask=m)
    ids, tgs = [], []
    for i in range(len(tg)):
        for j in range(S):
            if (m[i] >> j) & 1:
                ids.append(i)
                tgs.append(j * P + tg[i])
    ids = np.array(ids)
    rs = oracle.run(oracle.make_config(alg, K, 1, W, H, samples=S, ll_capacity=100000, opaque=opq_s),
                    np.array(tgs, np.uint32), d[ids], c[ids])
    comp = rs["composite"].reshape(S, P, 4).sum(axis=0, dtype=np.float64) / S
    cbuf = rs["colorbuf"].reshape(S, P, 4).sum(axis=0, dtype=np.float64) / S
    assert np.allclose(r["composite"], comp, rtol=0, atol=2e-6)
    assert np.allclose(r["colorbuf"], cbuf, rtol=0, atol=2e-6)
    assert np.allclose(r["out"], comp + (1.0 - comp[:, 3:]) * cbuf, rtol=0, atol=3e-6)
    per_sample_images = rs["out"].reshape(S, P, 4).sum(axis=0, dtype=np.float64) / S
    assert not np.allclose(r["out"], per_sample_images, rtol=0, atol=1e-4)
    # a fragment behind every covered sample's opaque depth does not exist
    k = np.array([oracle.depth_key(float(x)) for x in d], np.uint32)
    vis = np.zeros(len(tg), bool)
    for j in range(S):
        vis |= ((m >> j) & 1).astype(bool) & (k < opq[0][j * P + tg])
    assert (~vis).sum() > 0
    r2 = oracle.run(oracle.make_config(alg, K, 1, W, H, coverage_samples=S, ll_capacity=100000, opaque=opq),
                    tg[vis], d[vis], c[vis], mask=m[vis])
    assert np.array_equal(r["composite"], r2["composite"])


def test_fp64_mirror_with_opaque(oracle):
    rng = np.random.default_rng(71)
    W, H = 6, 5
    T = W * H
    otg, od, oc = _opaque_stream(rng, T, 50)
    tg, d, c = _stream(rng, T, 300, dmax=1.1)
    opq = oracle.opaque_pass(oracle.make_config(1, 32, 1, W, H), otg, od, oc)
    cfg = oracle.make_config(1, 32, 1, W, H, ll_capacity=10000, opaque=opq)
    f32 = oracle.run(cfg, tg, d, c)["out"]
    f64 = oracle.run_fp64(cfg, tg, d, c)["out"]
    assert np.allclose(f32, f64, rtol=0, atol=1e-5)  # K = 32 holds every pixel here
    assert np.bincount(tg, minlength=T).max() <= 32
