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


def loop64_stage_bytes(F: int, P: int, K: int) -> dict:
    return {
        "clear": 8 * K * P,
        "store": 20 * F + 16 * F,  # stream in + at least one slot read-modify-write
        "resolve": 8 * K * P + 16 * P,
    }


def kbuffer_stage_bytes(F: int, P: int, K: int, spin: bool) -> dict:
    aux = 12 if spin else 4
    return {
        "clear": aux * P,
        "store": 24 * F + 12 * min(F, K * P) + 8 * F,
        "resolve": 4 * P + 12 * min(F, K * P) + 16 * P,
    }


def wboit_stage_bytes(F: int, P: int) -> dict:
    return {
        "clear": 20 * P,
        "store": 24 * F + 40 * F,  # stream in + accumulation read-modify-write
        "resolve": 20 * P + 16 * P,
    }


def loop32_stage_bytes(F: int, P: int, K: int) -> dict:
    """Loop32 (two geometry passes): depth planes 4KP + (id, colour) planes 8KP
    cleared; depth pass 8F in (target, depth) + at least one 4-byte slot
    read-modify-write 8F; colour pass 24F in + the K-th depth 4F + one 8-byte
    slot read-modify-write 16F; composite 12KP + 16P."""
    return {
        "clear": 12 * K * P,
        "store": 8 * F + 8 * F,
        "classify": 24 * F + 4 * F + 16 * F,
        "resolve": 12 * K * P + 16 * P,
    }


def interlock_stage_bytes(F: int, P: int) -> dict:
    """Ordered interlock: the Linked List append plus the shaded colour per
    node (16F); the resolve walks the nodes like the Linked List composite."""
    return {"clear": 0, "store": 44 * F + 16 * F, "resolve": 8 * P + 16 * F + 16 * P}


def frac(bytes_: float, seconds: float) -> tuple[float, float]:
    """Returns (GB/s, fraction of the HBM peak)."""
    gbs = bytes_ / seconds / 1e9
    return gbs, gbs / HBM_PEAK_GBS
