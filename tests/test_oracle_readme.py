"""The reference README's worked examples (README.md:45, 53-66, 72, 80, 86-96)
and closed forms of the blend, run through the oracle.

Colours c1..c4 are distinct; "(cX, 0.X)" is (colour, depth). Fragment ids are
arrival order. Expected colours are built from the oracle's codec (pinned in
test_oracle_codec.py) and the blend equation written out here from
main.cpp:548-558: dst = src + (1 - src.a) * dst.
"""
import numpy as np
import pytest

BG = np.float32([0.5, 0.5, 0.5, 1.0])
C = {
    1: np.float32([0.9, 0.1, 0.1, 0.5]),
    2: np.float32([0.1, 0.9, 0.1, 0.4]),
    3: np.float32([0.1, 0.1, 0.9, 0.3]),
    4: np.float32([0.7, 0.7, 0.1, 0.6]),
}


def premult(c):
    return np.float32([c[0] * c[3], c[1] * c[3], c[2] * c[3], c[3]])


def over(top, base):
    t = np.float32(1) - top[3]
    return (top + t * base).astype(np.float32)


def q(oracle, c):
    """Colour after the sRGB8 round trip of the A-buffer."""
    return oracle.decode_premult(oracle.encode(premult(c)))


def stream(pairs, target=None):
    n = len(pairs)
    tg = np.zeros(n, np.uint32) if target is None else np.asarray(target, np.uint32)
    d = np.float32([p[1] for p in pairs])
    c = np.stack([premult(C[p[0]]) for p in pairs]).astype(np.float32)
    return tg, d, c


def test_simple_readme_example(oracle):
    """README.md:45: K=2, arrivals (c4,.4),(c2,.2),(c1,.1),(c3,.3): the A-buffer keeps the first
    two, (c4,.4),(c2,.2), sorts them to (c2,.2),(c4,.4); (c1,.1) then (c3,.3) are tail-blended."""
    tg, d, c = stream([(4, .4), (2, .2), (1, .1), (3, .3)])
    r = oracle.run(oracle.make_config(oracle.SIMPLE, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_val"][0]) == [1, 0]
    assert list(r["tail_flag"]) == [0, 0, 1, 1]
    assert r["count"][0] == 4
    # c1 first, c3 on top (primitive order); the canonical fold associates newest first
    colorbuf = over(over(premult(C[3]), premult(C[1])), BG)
    comp = over(q(oracle, C[2]), q(oracle, C[4]))
    assert np.array_equal(r["colorbuf"][0], colorbuf)
    assert np.array_equal(r["composite"][0], comp)
    assert np.array_equal(r["out"][0], over(comp, colorbuf))


def test_linkedlist_readme_pool_table(oracle):
    """README.md:53-66: K=2, a 4-element A-buffer including the terminator; pixel 1 gets
    (c4,.4), (c2,.2), (c3,.3) and pixel 2 (c1,.1); nodes 1..3 hold c4, c1, c2 and (c3,.3)
    is tail-blended. Pixel 1 composites (c2,.2),(c4,.4); pixel 2 (c1,.1)."""
    tg, d, c = stream([(4, .4), (1, .1), (2, .2), (3, .3)], target=[0, 1, 0, 0])
    r = oracle.run(oracle.make_config(oracle.LINKEDLIST, 2, 1, 2, 1, ll_capacity=4), tg, d, c)
    assert r["ll_overflow"] == 1
    assert list(r["tail_flag"]) == [0, 0, 0, 1]
    assert list(r["count"]) == [2, 1]
    assert list(r["frontk_val"][0]) == [2, 0]
    assert list(r["frontk_val"][1]) == [1, 0xFFFFFFFF]
    assert np.array_equal(r["colorbuf"][0], over(premult(C[3]), BG))
    assert np.array_equal(r["out"][1], over(q(oracle, C[1]), BG))
    assert np.array_equal(r["out"][0], over(over(q(oracle, C[2]), q(oracle, C[4])), over(premult(C[3]), BG)))


def test_linkedlist_composite_tail(oracle):
    """K=2 with a large pool: the list is walked newest first; the two nearest are kept and
    the rest blended as the composite tail, first pushed out on top (oitLinkedList.frag.glsl:131-143)."""
    tg, d, c = stream([(4, .4), (2, .2), (1, .1), (3, .3)])
    r = oracle.run(oracle.make_config(oracle.LINKEDLIST, 2, 1, 1, 1, ll_capacity=100), tg, d, c)
    assert list(r["frontk_val"][0]) == [2, 1]  # (c1,.1) id 2, (c2,.2) id 1
    # walk order ids 3,2,1,0: gather (c3,.3),(c1,.1); then (c2,.2) pushes out c3; then (c4,.4) goes itself
    tail = over(q(oracle, C[3]), q(oracle, C[4]))
    comp = over(over(q(oracle, C[1]), q(oracle, C[2])), tail)
    assert np.array_equal(r["composite"][0], comp)
    assert np.array_equal(r["out"][0], over(comp, BG))
    # without tail blending the pushed-out fragments vanish (insertionSort, :140)
    r0 = oracle.run(oracle.make_config(oracle.LINKEDLIST, 2, 0, 1, 1, ll_capacity=100), tg, d, c)
    assert np.array_equal(r0["out"][0], over(over(q(oracle, C[1]), q(oracle, C[2])), BG))


def test_loop64_readme_example(oracle):
    """README.md:80: K=2, the four fragments of the Simple example: the A-buffer ends as
    ((c1,.1),(c2,.2)); (c4,.4) and (c3,.3) are tail-blended."""
    tg, d, c = stream([(4, .4), (2, .2), (1, .1), (3, .3)])
    r = oracle.run(oracle.make_config(oracle.LOOP64, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_key"][0]) == [oracle.depth_key(.1), oracle.depth_key(.2)]
    assert list(r["frontk_val"][0]) == [oracle.encode(premult(C[1])), oracle.encode(premult(C[2]))]
    assert list(r["tail_flag"]) == [1, 0, 0, 1]
    # canonical tail: evicted fragments at their own draw position, packed colour
    assert np.array_equal(r["colorbuf"][0], over(over(q(oracle, C[3]), q(oracle, C[4])), BG))
    # sequential reading of the shader: c4 is evicted by c1 (position 2), c3 rejects itself
    # (position 3): the same order, one ROP blend per event
    rs = oracle.run(oracle.make_config(oracle.LOOP64, 2, 1, 1, 1, tail_order=oracle.TAIL_SEQUENTIAL), tg, d, c)
    assert np.array_equal(rs["colorbuf"][0], over(q(oracle, C[3]), over(q(oracle, C[4]), BG)))
    assert np.allclose(rs["colorbuf"][0], r["colorbuf"][0], rtol=0, atol=2 ** -22)


def test_spinlock_readme_trace(oracle):
    """README.md:86-96: K=2; the critical section is entered by (c3,.3), (c2,.2), (c4,.4) and
    (c1,.1) in that order; (c4,.4) is behind both and tail-blended, (c1,.1) evicts (c3,.3).
    The A-buffer ends as (c1,.1),(c2,.2)."""
    tg, d, c = stream([(3, .3), (2, .2), (4, .4), (1, .1)])
    r = oracle.run(oracle.make_config(oracle.SPINLOCK, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_val"][0]) == [3, 1]
    assert list(r["tail_flag"]) == [1, 0, 1, 0]
    assert r["count"][0] == 4
    # sequential order: c4 tail-blends itself at position 2 (full-precision colour),
    # c3's packed colour is blended at c1's position 3
    rs = oracle.run(oracle.make_config(oracle.SPINLOCK, 2, 1, 1, 1, tail_order=oracle.TAIL_SEQUENTIAL), tg, d, c)
    assert np.array_equal(rs["colorbuf"][0], over(q(oracle, C[3]), over(premult(C[4]), BG)))
    # canonical: both at their own ids (c3 id 0, c4 id 2), packed colours (whether c4 is rejected
    # or stored and evicted later depends on the arrival order)
    assert np.array_equal(r["colorbuf"][0], over(over(q(oracle, C[4]), q(oracle, C[3])), BG))


@pytest.mark.parametrize("alg", [0, 1, 3, 4])
def test_single_fragment_closed_form(oracle, alg):
    tg, d, c = stream([(2, .5)])
    r = oracle.run(oracle.make_config(alg, 4, 1, 1, 1, ll_capacity=8), tg, d, c)
    assert np.array_equal(r["out"][0], over(q(oracle, C[2]), BG))


def test_weighted_single_fragment_is_over(oracle):
    """oitWeighted.frag.glsl with one fragment: accum = w*c, reveal = 1-a; the resolve gives
    c.rgb + (1-a)*bg, i.e. exactly premultiplied over, up to rounding."""
    tg, d, c = stream([(1, .3)])
    r = oracle.run(oracle.make_config(oracle.WEIGHTED, 1, 1, 1, 1), tg, d, c)
    want = over(premult(C[1]), BG)
    assert np.allclose(r["out"][0, :3], want[:3], rtol=0, atol=2e-7)
    assert r["wb_reveal"][0] == np.float32(1) - C[1][3]


@pytest.mark.parametrize("alg", [0, 1, 3, 4])
@pytest.mark.parametrize("n", [1, 3, 8])
def test_identical_layers_closed_form(oracle, alg, n):
    """n layers of one colour, K >= n: composite alpha = 1 - (1 - a)^n (a = packed alpha)."""
    pairs = [(2, 0.1 + 0.05 * i) for i in range(n)]
    tg, d, c = stream(pairs)
    r = oracle.run(oracle.make_config(alg, 8, 1, 1, 1, ll_capacity=32), tg, d, c)
    a = float(q(oracle, C[2])[3])
    assert abs(float(r["composite"][0, 3]) - (1 - (1 - a) ** n)) < 1e-6
    rgb = float(q(oracle, C[2])[1]) / a
    assert abs(float(r["composite"][0, 1]) - rgb * (1 - (1 - a) ** n)) < 1e-6


@pytest.mark.parametrize("alg", [0, 1, 3, 4, 6])
def test_empty_frame_is_background(oracle, alg):
    tg = np.zeros(0, np.uint32)
    r = oracle.run(oracle.make_config(alg, 4, 1, 3, 2, ll_capacity=4), tg, np.zeros(0, np.float32),
                   np.zeros((0, 4), np.float32))
    assert np.array_equal(r["out"], np.tile(BG, (6, 1)))


def test_equal_depth_ties_by_fragment_id(oracle):
    """Equal depths keep draw order (smaller id nearer) in every algorithm that stores ids."""
    tg, d, c = stream([(1, .5), (2, .5), (3, .5)])
    for alg in (oracle.LINKEDLIST, oracle.SPINLOCK, oracle.SIMPLE):
        r = oracle.run(oracle.make_config(alg, 4, 1, 1, 1, ll_capacity=8), tg, d, c)
        assert list(r["frontk_val"][0][:3]) == [0, 1, 2]
        comp = over(over(q(oracle, C[1]), q(oracle, C[2])), q(oracle, C[3]))
        assert np.array_equal(r["composite"][0], comp)


def test_loop32_readme_example(oracle):
    """README.md:72: K=2, the Simple example's four fragments: the depth pass keeps (0.1, 0.2),
    the colour pass matches c1 and c2 to them, and (c4,.4), (c3,.3) tail-blend in primitive
    order with their shaded colours."""
    tg, d, c = stream([(4, .4), (2, .2), (1, .1), (3, .3)])
    r = oracle.run(oracle.make_config(oracle.LOOP, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_key"][0]) == [oracle.depth_key(.1), oracle.depth_key(.2)]
    assert list(r["frontk_val"][0]) == [oracle.encode(premult(C[1])), oracle.encode(premult(C[2]))]
    assert list(r["tail_flag"]) == [1, 0, 0, 1]
    assert np.array_equal(r["colorbuf"][0], over(over(premult(C[3]), premult(C[4])), BG))
    assert np.array_equal(r["composite"][0], over(q(oracle, C[1]), q(oracle, C[2])))


def test_loop32_shared_depth_keeps_the_last_colour(oracle):
    """Two fragments at one depth share one depth slot; the colour store of the later one
    wins (primitive order) and the earlier colour is neither stored nor tail-blended."""
    tg, d, c = stream([(1, .3), (2, .3), (3, .6)])
    r = oracle.run(oracle.make_config(oracle.LOOP, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_key"][0]) == [oracle.depth_key(.3), oracle.depth_key(.6)]
    assert list(r["frontk_val"][0]) == [oracle.encode(premult(C[2])), oracle.encode(premult(C[3]))]
    assert list(r["tail_flag"]) == [0, 0, 0]


def test_interlock_is_the_spinlock_in_primitive_order(oracle):
    """README.md:86-96 trace as an ordered interlock: the critical sections run in primitive
    order, and each invocation's tail output is blended at its own position."""
    tg, d, c = stream([(3, .3), (2, .2), (4, .4), (1, .1)])
    r = oracle.run(oracle.make_config(oracle.INTERLOCK, 2, 1, 1, 1), tg, d, c)
    assert list(r["frontk_val"][0]) == [3, 1]
    assert list(r["tail_flag"]) == [1, 0, 1, 0]
    assert np.array_equal(r["colorbuf"][0], over(q(oracle, C[3]), over(premult(C[4]), BG)))
    rs = oracle.run(oracle.make_config(oracle.SPINLOCK, 2, 1, 1, 1, tail_order=oracle.TAIL_SEQUENTIAL), tg, d, c)
    assert np.array_equal(rs["out"], r["out"])
