"""The WBOIT weight function and resolve, pinned in 50-digit decimal arithmetic.

Expected values are computed here from the text of the reference shader, not
from the oracle's code:
  colour pass, oitWeighted.frag.glsl:53-79
    color.rgb *= color.a                                       (:55-56)
    depthZ      = -IN.depth * 10                               (:65; IN.depth = view-space z,
                                                                 object.vert.glsl:35)
    distWeight  = clamp(0.03 / (1e-5 + pow(depthZ / 200, 4)), 1e-2, 3e3)        (:67)
    alphaWeight = min(1, max(max(r, g), max(b, a)) * 40 + 0.01); alphaWeight^2   (:69-70)
    outColor    = color * alphaWeight * distWeight; outReveal = a                (:72-78)
  blends, main.cpp:559-588
    accum  += outColor                      (ONE, ONE)
    reveal *= 1 - a                         (ZERO, ONE_MINUS_SRC_COLOR)
    dst = src * (1 - src.a) + dst * src.a   (ONE_MINUS_SRC_ALPHA, SRC_ALPHA) with
    src = (accum.rgb / max(accum.a, 1e-5), reveal)                                (:98-109)
The view distance -IN.depth stands in as near + depth * span (the synthetic
scene's mapping; defaults 2 and 8).

The fragments below put distWeight on its lower clamp, at its upper end
(depthZ = 0 gives exactly 0.03 / 1e-5 = 3e3) and in between, and alphaWeight
saturated and not, so that changing the exponent 4, the /200, the *10, the
clamp bounds or the 40 * max + 0.01 term moves the oracle's output far beyond
the tolerance; test_each_mutation_is_caught shows that for each such change.
"""
from decimal import Decimal as D, getcontext

import numpy as np
import pytest

getcontext().prec = 50

NEAR, SPAN = 2.0, 8.0
BG = (0.5, 0.5, 0.5, 1.0)

# (depth, unpremultiplied rgb, alpha): distWeight regime / alphaWeight regime
FRAGS = [
    (0.30, (0.9, 0.2, 0.1), 0.50),      # view 4.4: unclamped; saturated alpha weight
    (0.05, (0.010, 0.004, 0.002), 0.012),  # view 2.4: unclamped; max * 40 + 0.01 < 1
    (5.00, (0.1, 0.8, 0.3), 0.40),      # view 42: 0.03 / (1e-5 + 2.1^4) < 1e-2, lower clamp
    (-0.25, (0.2, 0.3, 0.9), 0.30),     # view 0: distWeight = 3e3 exactly, the upper clamp
    (-0.1875, (0.6, 0.6, 0.2), 0.20),   # view 0.5: just under the upper clamp
    (0.90, (0.015, 0.01, 0.005), 0.02),  # view 9.2: unclamped; unsaturated
]


def dec(x):
    return D(repr(float(x)))


def f32(x):
    return float(np.float32(x))


def premult(rgb, a):
    a32 = np.float32(a)
    return np.float32([np.float32(rgb[0]) * a32, np.float32(rgb[1]) * a32, np.float32(rgb[2]) * a32, a32])


def weight(depth, c, *, exponent=4, div=200, scale=10, lo=D("1e-2"), hi=D("3e3"), slope=40, bias=D("0.01")):
    """distWeight * alphaWeight of oitWeighted.frag.glsl:65-72 in decimal, for
    the premultiplied colour c (float32 values taken exactly)."""
    view = dec(np.float32(NEAR)) + dec(np.float32(depth)) * dec(np.float32(SPAN))
    depth_z = view * scale
    x = depth_z / div
    dist = D("0.03") / (D("1e-5") + x ** exponent)
    dist = min(max(dist, lo), hi)
    m = max(dec(v) for v in c)
    aw = min(D(1), m * slope + bias)
    return aw * aw * dist


def resolve(frags, bg=BG, **mut):
    """accum, reveal and the composited pixel of oitWeighted.frag.glsl:98-109
    with the WEIGHTED_COMPOSITE blend (main.cpp:578-588), in decimal."""
    acc = [D(0)] * 4
    rev = D(1)
    for depth, c in frags:
        w = weight(depth, c, **mut)
        acc = [acc[k] + dec(c[k]) * w for k in range(4)]
        rev = rev * (1 - dec(c[3]))
    den = max(acc[3], D("1e-5"))
    src = [acc[0] / den, acc[1] / den, acc[2] / den, rev]
    out = [src[k] * (1 - rev) + dec(bg[k]) * rev for k in range(4)]
    return acc, rev, out


def _stream(frags):
    n = len(frags)
    tg = np.zeros(n, np.uint32)
    d = np.float32([f[0] for f in frags])
    c = np.stack([f[1] for f in frags]).astype(np.float32)
    return tg, d, c


def _frags(idx):
    return [(FRAGS[i][0], premult(FRAGS[i][1], FRAGS[i][2])) for i in idx]


def _rel(got, want):
    want = float(want)
    return abs(float(got) - want) / max(abs(want), 1e-30)


def test_regimes_are_the_ones_named():
    """The fixture really reaches each regime of the weight function."""
    def parts(i):
        depth, rgb, a = FRAGS[i]
        c = premult(rgb, a)
        view = dec(np.float32(NEAR)) + dec(np.float32(depth)) * dec(np.float32(SPAN))
        x = view * 10 / 200
        raw = D("0.03") / (D("1e-5") + x ** 4)
        return raw, max(dec(v) for v in c) * 40 + D("0.01")
    raw, aw = parts(0)
    assert D("1e-2") < raw < D("3e3") and aw > 1
    raw, aw = parts(1)
    assert D("1e-2") < raw < D("3e3") and aw < 1
    raw, _ = parts(2)
    assert raw < D("1e-2")
    raw, _ = parts(3)
    assert raw == D("3e3")
    raw, _ = parts(4)
    assert D("2.8e3") < raw < D("3e3")
    raw, aw = parts(5)
    assert D("1e-2") < raw < D("3e3") and aw < 1


@pytest.mark.parametrize("i", range(len(FRAGS)))
def test_single_fragment_accum_is_color_times_weight(oracle, i):
    """One fragment: accum = premultiplied colour * weight, reveal = 1 - a."""
    frags = _frags([i])
    tg, d, c = _stream(frags)
    cfg = oracle.make_config(oracle.WEIGHTED, 1, 1, 1, 1, wboit_view_near=NEAR, wboit_view_span=SPAN)
    r = oracle.run(cfg, tg, d, c)
    acc, rev, out = resolve(frags)
    for k in range(4):
        assert _rel(r["wb_accum"][0, k], acc[k]) < 2e-6, (k, r["wb_accum"][0], acc)
    assert r["wb_reveal"][0] == np.float32(1) - c[0, 3]
    r64 = oracle.run_fp64(cfg, tg, d, c)
    for k in range(4):
        assert _rel(r64["wb_accum"][0, k], acc[k]) < 1e-12


@pytest.mark.parametrize("idx", [(0, 1), (1, 2), (2, 3), (3, 4), (0, 1, 2), (1, 3, 5), (2, 4, 5), (0, 2, 3, 4, 5),
                                 (0, 1, 2, 3, 4, 5)])
def test_multi_fragment_resolve(oracle, idx):
    """Two to six fragments on one pixel: the accumulation, revealage and the
    composited pixel against the decimal evaluation (float32 oracle within
    2e-6 relative / 1e-6 absolute, FP64 mirror within 1e-12)."""
    frags = _frags(idx)
    tg, d, c = _stream(frags)
    cfg = oracle.make_config(oracle.WEIGHTED, 1, 1, 1, 1, wboit_view_near=NEAR, wboit_view_span=SPAN)
    r = oracle.run(cfg, tg, d, c)
    acc, rev, out = resolve(frags)
    for k in range(4):
        assert _rel(r["wb_accum"][0, k], acc[k]) < 2e-6
        assert abs(float(r["out"][0, k]) - float(out[k])) < 1e-6, (k, r["out"][0], out)
    assert _rel(r["wb_reveal"][0], rev) < 1e-6
    r64 = oracle.run_fp64(cfg, tg, d, c)
    for k in range(4):
        assert abs(float(r64["out"][0, k]) - float(out[k])) < 1e-12
        assert _rel(r64["wb_accum"][0, k], acc[k]) < 1e-12


MUTATIONS = {
    "exponent 3": dict(exponent=3),
    "exponent 2": dict(exponent=2),
    "/100": dict(div=100),
    "/400": dict(div=400),
    "no *10": dict(scale=1),
    "*100": dict(scale=100),
    "lower clamp 1e-3": dict(lo=D("1e-3")),
    "upper clamp 1e3": dict(hi=D("1e3")),
    "slope 4": dict(slope=4),
    "slope 400": dict(slope=400),
    "bias 0": dict(bias=D(0)),
    "bias 0.1": dict(bias=D("0.1")),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_each_mutation_is_caught(name):
    """Each plausible slip in the weight function moves the expected values
    of the fixtures above by far more than the tolerances the oracle is held
    to, so an oracle (and a GPU kernel checked against it) carrying that slip
    fails test_single_fragment_accum_is_color_times_weight or
    test_multi_fragment_resolve."""
    mut = MUTATIONS[name]
    worst_acc = 0.0
    worst_out = 0.0
    for i in range(len(FRAGS)):
        a0, _, _ = resolve(_frags([i]))
        a1, _, _ = resolve(_frags([i]), **mut)
        worst_acc = max(worst_acc, max(_rel(a1[k], a0[k]) for k in range(4)))
    for idx in [(0, 1), (1, 2), (2, 3), (3, 4), (0, 1, 2), (1, 3, 5), (2, 4, 5), (0, 2, 3, 4, 5)]:
        _, _, o0 = resolve(_frags(idx))
        _, _, o1 = resolve(_frags(idx), **mut)
        worst_out = max(worst_out, max(abs(float(o1[k] - o0[k])) for k in range(4)))
    assert worst_acc > 1e-3, (name, worst_acc)
    assert worst_out > 1e-4, (name, worst_out)
