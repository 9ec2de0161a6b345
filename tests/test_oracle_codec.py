"""Pins the oracle's colour codec and depth key against what the mathematics
fixes, independently of the oracle's own code:
 - the encode thresholds against the sRGB curve of shaderCommon.glsl:60-69
   evaluated with 60-digit decimal arithmetic (correct rounding at every
   one of the 255 code boundaries);
 - the decode table against shaderCommon.glsl:84-94 in 60-digit decimals
   (each entry is the binary32 nearest to the exact value);
 - alpha against the binary32 quotient u / 255;
 - the depth key's ordering against numpy's float ordering.
"""
from decimal import Decimal, getcontext

import numpy as np

getcontext().prec = 60
D = Decimal
INV24 = D(1) / D("2.4")
BRANCH_ENC = D(float(np.float32(0.0031308)))
BRANCH_DEC = D(float(np.float32(0.04045)))


def s_exact(c: float) -> D:
    x = D(float(c))
    if x < BRANCH_ENC:
        return x * D("12.92")
    return (x ** INV24) * D("1.055") - D("0.055")


def q_exact(c: float) -> int:
    s = s_exact(c)
    s = max(D(0), min(D(1), s))
    return int((s * 255 + D("0.5")).to_integral_value(rounding="ROUND_FLOOR"))


def lin_exact(u: int) -> D:
    f = D(u) / D(255)
    if f < BRANCH_DEC:
        return f / D("12.92")
    return ((f + D("0.055")) / D("1.055")) ** D("2.4")


def test_encode_thresholds_are_exact_code_boundaries(oracle):
    thr, _, _ = oracle.tables()
    assert thr[0] == -np.inf
    assert np.all(np.diff(thr[1:].astype(np.float64)) > 0)
    for q in range(1, 256):
        c = np.float32(thr[q])
        prev = np.nextafter(c, np.float32(-np.inf), dtype=np.float32)
        assert q_exact(float(c)) >= q, q
        assert q_exact(float(prev)) < q, q


def test_decode_table_is_correctly_rounded(oracle):
    _, dec, alpha = oracle.tables()
    for u in range(256):
        exact = lin_exact(u)
        f = np.float32(dec[u])
        lo = np.nextafter(f, np.float32(-np.inf), dtype=np.float32)
        hi = np.nextafter(f, np.float32(np.inf), dtype=np.float32)
        err = abs(D(float(f)) - exact)
        assert err <= abs(D(float(lo)) - exact), u
        assert err <= abs(D(float(hi)) - exact), u
        assert alpha[u] == np.float32(u) / np.float32(255)


def test_encode_matches_curve_on_random_values(oracle):
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.uniform(0, 1, 3000), rng.uniform(0, 0.01, 500), [0.0, 1.0, 0.5, 1.5, -0.25]])
    for v in vals.astype(np.float32):
        code = oracle.encode(np.float32([v, v, v, 1.0]))
        want = q_exact(float(v))
        assert code & 255 == want
        assert (code >> 8) & 255 == want
        assert (code >> 16) & 255 == want
        assert code >> 24 == 255


def test_encode_unpremultiplies_and_quantizes_alpha(oracle):
    for a in (0.2, 0.37, 0.5, 0.6):
        a32 = np.float32(a)
        rgb = np.float32([0.25, 0.5, 0.75])
        code = oracle.encode(np.concatenate([rgb * a32, [a32]]).astype(np.float32))
        unpremult = (rgb * a32) / a32
        assert code & 255 == q_exact(float(unpremult[0]))
        assert (code >> 16) & 255 == q_exact(float(unpremult[2]))
        assert code >> 24 == int(np.floor(np.float32(a32 * np.float32(255)) + np.float32(0.5)))
    # alpha 0 and NaN give code 0 channels
    assert oracle.encode(np.float32([0.3, 0.3, 0.3, 0.0])) == 0
    assert oracle.encode(np.float32([np.nan, 0.0, 0.0, 1.0])) & 255 == 0


def test_decode_premultiplies(oracle):
    _, dec, alpha = oracle.tables()
    for packed in (0xFF000000, 0x80FF7F01, 0x33102030, 0xFFFFFFFF):
        o = oracle.decode_premult(packed)
        a = alpha[packed >> 24]
        assert o[3] == a
        assert o[0] == np.float32(dec[packed & 255] * a)
        assert o[2] == np.float32(dec[(packed >> 16) & 255] * a)


def test_round_trip_error_bounded(oracle):
    """decode(encode(x)) is within half an 8-bit step of x in sRGB space."""
    rng = np.random.default_rng(1)
    for v in rng.uniform(0, 1, 400).astype(np.float32):
        o = oracle.decode_premult(oracle.encode(np.float32([v, v, v, 1.0])))
        s_in = float(s_exact(float(v)))
        s_out = float(s_exact(float(o[0])))
        assert abs(s_in - s_out) <= 0.5 / 255 + 1e-6


def test_depth_key_orders_every_float(oracle):
    rng = np.random.default_rng(2)
    vals = np.concatenate([
        rng.uniform(-2, 2, 500), rng.normal(0, 1e-30, 50), [0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45, 3.4e38, -3.4e38],
    ]).astype(np.float32)
    keys = np.array([oracle.depth_key(float(v)) for v in vals], dtype=np.uint64)
    assert (keys != 0xFFFFFFFF).all()
    order = np.argsort(vals, kind="stable")
    sv, sk = vals[order], keys[order]
    for i in range(len(sv) - 1):
        if sv[i] < sv[i + 1]:
            assert sk[i] < sk[i + 1]
        elif sv[i] == sv[i + 1] and not (sv[i] == 0):
            assert sk[i] == sk[i + 1]
    # non-negative floats keep the reference's uint ordering (floatBitsToUint)
    pos = np.abs(vals[np.isfinite(vals)])
    pk = np.array([oracle.depth_key(float(v)) for v in pos], dtype=np.uint64)
    bits = pos.view(np.uint32).astype(np.uint64)
    assert np.array_equal(np.argsort(pk, kind="stable"), np.argsort(bits, kind="stable"))
    # NaN sorts after +inf and never takes the empty sentinel, not even the
    # all-ones-mantissa NaN whose raw key would be 0xFFFFFFFF
    assert oracle.depth_key(float("nan")) > oracle.depth_key(float("inf"))
    all_ones_nan = float(np.uint32(0x7FFFFFFF).view(np.float32))
    assert oracle.depth_key(all_ones_nan) == 0xFFFFFFFE
