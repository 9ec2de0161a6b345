"""Algorithmic (compulsory) HBM bytes of each kernel and of a whole frame.

One place for the formulas that bench.py's `roofline` block uses.
F = fragments of the frame, P = targets (pixels x samples), K = layers,
N = fragments that end in the colour-pass tail.

Linked List (this build's layout: 8-byte (depth key, colour) node word plus
a 4-byte link word per node, 4-byte head per target):
  clear    : 4P                      (head words)
  store    : 20F in (depth 4 + premultiplied RGBA 16) + 12F out (node + kUnset link)
  link     : 4F in (target) + 8F (head read-modify-write, one per fragment
             without run merging) + 4F out (link word)
  resolve  : 4P (heads) + 12F (nodes) + 16P (RGBA out)
Frame total 60F + 24P. The reference's layout (16-byte uvec4 nodes,
oitLinkedList.frag.glsl:77-82) needs 64F + 24P for the same work; frame
fractions are reported against that figure so the two layouts are compared
on the same work.
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def ll_stage_bytes(F: int, P: int, fused: bool = True) -> dict:
    """Per-launch algorithmic bytes. Fused append (default): stream in 24F,
    node word 8F, link word 4F, head read-modify-write 8F. The resetting
    resolve: heads 4P read + 4P reset, nodes 12F read + 4F link reset, 16P out.
    Split append: store 32F (stream 20F in, node 8F + kUnset link 4F out) and
    link 16F (target 4F, head read-modify-write 8F, link 4F)."""
    if fused:
        return {"clear": 0, "store": 44 * F, "resolve": 8 * P + 16 * F + 16 * P}
    return {"clear": 0, "store": 32 * F, "link": 16 * F, "resolve": 8 * P + 16 * F + 16 * P}


def ll_frame_bytes_reference(F: int, P: int) -> int:
    """64F + 24P: the reference node layout's compulsory traffic for one frame."""
    return 64 * F + 24 * P


def loop64_stage_bytes(F: int, P: int, K: int, occupied: int | None = None, tail: bool = True) -> dict:
    """Loop64: no clear (the composite empties the slots it reads, and only
    a reconfigured or unresolved A-buffer is filled); store: stream in 20F +
    at least one 8-byte slot read-modify-write 16F; with tail blending the
    classify pass reads (target, depth) 8F and one K-th slot per target 8P;
    composite: the occupied slots read and set back to empty 16 x occupied,
    the first empty slot 8P, RGBA out 16P. `occupied` = sum over targets of
    min(fragments, K); min(F, KP) bounds it from above."""
    occ = min(F, K * P) if occupied is None else occupied
    b = {"clear": 0, "store": 20 * F + 16 * F}
    if tail:
        b["classify"] = 8 * F + 8 * P
    b["resolve"] = 16 * occ + 8 * P + 16 * P
    return b


def kbuffer_stage_bytes(F: int, P: int, K: int, spin: bool) -> dict:
    aux = 12 if spin else 4
    return {
        "clear": aux * P,
        "store": 24 * F + 12 * min(F, K * P) + 8 * F,
        "resolve": 4 * P + 12 * min(F, K * P) + 16 * P,
    }


def wboit_stage_bytes(F: int, P: int) -> dict:
    """WBOIT: no clear (the composite sets the accumulators it read back to
    the clear values where they are not already; counted here for every
    target, an upper bound); store: stream in 24F + accumulation
    read-modify-write 40F; composite: accumulators read 20P and reset 20P,
    RGBA out 16P."""
    return {
        "clear": 0,
        "store": 24 * F + 40 * F,  # stream in + accumulation read-modify-write
        "resolve": 20 * P + 20 * P + 16 * P,
    }


def loop32_stage_bytes(F: int, P: int, K: int, occupied: int | None = None) -> dict:
    """Loop32 (two geometry passes): no clear (the composite empties the
    depth and colour slots it reads); depth pass 8F in (target, depth) + at
    least one 4-byte slot read-modify-write 8F; colour pass 24F in + the K-th
    depth 4F + one 8-byte slot read-modify-write 16F; composite: occupied
    depth and colour slots read and set back to empty 24 x occupied, the
    first empty depth 4P, RGBA out 16P. `occupied` as for Loop64 (distinct
    depths here; fragments bound them from above)."""
    occ = min(F, K * P) if occupied is None else occupied
    return {
        "clear": 0,
        "store": 8 * F + 8 * F,
        "classify": 24 * F + 4 * F + 16 * F,
        "resolve": 24 * occ + 4 * P + 16 * P,
    }


def interlock_stage_bytes(F: int, P: int) -> dict:
    """Ordered interlock: the Linked List append plus the shaded colour per
    node (16F); the resolve walks the nodes like the Linked List composite."""
    return {"clear": 0, "store": 44 * F + 16 * F, "resolve": 8 * P + 16 * F + 16 * P}


def frac(bytes_: float, seconds: float) -> tuple[float, float]:
    """Returns (GB/s, fraction of the HBM peak)."""
    gbs = bytes_ / seconds / 1e9
    return gbs, gbs / HBM_PEAK_GBS
