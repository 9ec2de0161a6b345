"""Coverage shading (OIT_COVERAGE_SHADING) and the display finalize in the
oracle, pinned by closed forms and by the per-sample reading: where no
pixel holds more than K fragments, the coverage composite equals the
average over samples of sample-shaded frames whose sample s holds exactly
the fragments covering s (up to the rounding of averaging before vs after
the final blend)."""
import numpy as np
import pytest

BG = np.float32([0.5, 0.5, 0.5, 1.0])


def over(top, base):
    t = np.float32(1.0) - top[3]
    return (top + t * base).astype(np.float32)


def q(oracle, c):
    return oracle.decode_premult(oracle.encode(c))


def _stream(rng, W, H, n, S):
    tg = rng.integers(0, W * H, n).astype(np.uint32)
    d = rng.uniform(0, 1, n).astype(np.float32)
    a = rng.uniform(0.1, 0.7, n).astype(np.float32)
    rgb = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    c = np.concatenate([rgb * a[:, None], a[:, None]], 1).astype(np.float32)
    m = rng.integers(1, 1 << S, n).astype(np.uint32)
    return tg, d, c, m


@pytest.mark.parametrize("alg", [0, 1, 4])
def test_single_fragment_covers_part_of_the_pixel(oracle, alg):
    S = 4
    c = np.float32([[0.2, 0.1, 0.05, 0.4]])
    for m in (0b0001, 0b0101, 0b1110, 0b1111):
        cfg = oracle.make_config(alg, 4, 1, 1, 1, coverage_samples=S, ll_capacity=8)
        r = oracle.run(cfg, np.uint32([0]), np.float32([0.5]), c, mask=np.uint32([m]))
        qc = q(oracle, c[0])
        s = np.zeros(4, np.float32)
        for j in range(S):
            s = (s + (qc if (m >> j) & 1 else np.zeros(4, np.float32))).astype(np.float32)
        comp = (s / np.float32(S)).astype(np.float32)
        assert np.array_equal(r["composite"][0], comp), m
        assert np.array_equal(r["out"][0], over(comp, BG))


def test_zero_mask_fragments_do_not_exist(oracle):
    cfg = oracle.make_config(1, 4, 1, 1, 1, coverage_samples=4, ll_capacity=8)
    c = np.float32([[0.2, 0.1, 0.05, 0.4], [0.3, 0.3, 0.3, 0.5]])
    r = oracle.run(cfg, np.uint32([0, 0]), np.float32([0.5, 0.2]), c, mask=np.uint32([0b0011, 0]))
    r1 = oracle.run(cfg, np.uint32([0]), np.float32([0.5]), c[:1], mask=np.uint32([0b0011]))
    assert np.array_equal(r["out"], r1["out"]) and r["count"][0] == 1


def test_colour_pass_tail_is_per_sample(oracle):
    """Simple K=1: the second fragment goes to the ROP tail and blends only into its samples."""
    S = 4
    c = np.float32([[0.2, 0.1, 0.05, 0.4], [0.05, 0.3, 0.1, 0.5]])
    cfg = oracle.make_config(0, 1, 1, 1, 1, coverage_samples=S)
    r = oracle.run(cfg, np.uint32([0, 0]), np.float32([0.5, 0.2]), c, mask=np.uint32([0b1111, 0b0110]))
    s = np.zeros(4, np.float32)
    for j in range(S):
        s = (s + (over(c[1], BG) if (0b0110 >> j) & 1 else BG)).astype(np.float32)
    cb = (s / np.float32(S)).astype(np.float32)
    assert np.array_equal(r["colorbuf"][0], cb)


@pytest.mark.parametrize("alg,S", [(1, 2), (1, 4), (1, 8), (0, 4), (4, 4)])
def test_coverage_equals_average_of_per_sample_frames_when_nothing_is_evicted(oracle, alg, S):
    rng = np.random.default_rng(31 + S + alg)
    W, H, K = 6, 5, 32
    tg, d, c, m = _stream(rng, W, H, 200, S)
    cfg = oracle.make_config(alg, K, 1, W, H, coverage_samples=S, ll_capacity=1000)
    r = oracle.run(cfg, tg, d, c, mask=m)
    # per-sample frames: sample s of pixel p is target s*W*H + p, one fragment copy per covered sample
    ids, tgs = [], []
    for i in range(len(tg)):
        for j in range(S):
            if (m[i] >> j) & 1:
                ids.append(i)
                tgs.append(j * W * H + tg[i])
    ids = np.array(ids)
    cfg_s = oracle.make_config(alg, K, 1, W, H, samples=S, ll_capacity=100000)
    rs = oracle.run(cfg_s, np.array(tgs, np.uint32), d[ids], c[ids])
    n_per_pixel = np.bincount(tg, minlength=W * H)
    ok = n_per_pixel <= K
    planes = rs["composite"].reshape(S, W * H, 4)
    avg = planes.sum(axis=0, dtype=np.float64) / S
    assert np.allclose(r["composite"][ok], avg[ok], rtol=0, atol=2e-6)
    assert ok.sum() > 0


def test_finalize_srgb8(oracle):
    rng = np.random.default_rng(5)
    W, H, S, ss = 7, 5, 4, 2
    img = rng.uniform(0, 1, (S * H * ss * W * ss, 4)).astype(np.float32)
    out = oracle.finalize_srgb8(img, W, H, S, ss)
    planes = img.reshape(S, H * ss, W * ss, 4)
    thr, _, _ = oracle.tables()
    for y in range(H):
        for x in range(W):
            acc = np.zeros(4, np.float32)
            for s in range(S):
                for dy in range(ss):
                    for dx in range(ss):
                        acc = (acc + planes[s, y * ss + dy, x * ss + dx]).astype(np.float32)
            v = (acc / np.float32(S * ss * ss)).astype(np.float32)
            for k in range(3):
                code = int(np.searchsorted(thr[1:256], v[k], side="right"))  # thresholds <= v
                assert out[y, x, k] == code
            assert out[y, x, 3] == int(np.floor(np.float32(min(max(v[3], 0), 1)) * np.float32(255) + np.float32(0.5)))
    # a uniform image keeps its value: 0.5 linear -> sRGB 188, alpha 1 -> 255
    flat = np.tile(np.float32([0.5, 0.5, 0.5, 1.0]), (W * H, 1))
    assert (oracle.finalize_srgb8(flat, W, H) == np.uint8([188, 188, 188, 255])).all()


def test_unsupported_coverage_configs(oracle):
    c = np.float32([[0.2, 0.1, 0.05, 0.4]])
    for alg in (3, 6):  # Loop64, WBOIT
        with pytest.raises(RuntimeError):
            oracle.run(oracle.make_config(alg, 4, 1, 1, 1, coverage_samples=4), np.uint32([0]), np.float32([0.5]), c,
                       mask=np.uint32([1]))
    with pytest.raises(RuntimeError):  # coverage needs masks
        oracle.run(oracle.make_config(1, 4, 1, 1, 1, coverage_samples=4, ll_capacity=4), np.uint32([0]),
                   np.float32([0.5]), c)
