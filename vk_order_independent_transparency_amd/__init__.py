"""MI355X (gfx950) order-independent-transparency fragment-store and
composite path: Simple, Linked List, Loop64, Spinlock-style k-buffer and
weighted blended OIT as hand-written HIP kernels behind the C ABI of
include/oit.h (liboit_hip.so, built in-tree)."""
from ._abi import (  # noqa: F401
    AA_MSAA_4X,
    AA_MSAA_8X,
    AA_NONE,
    AA_SSAA_4X,
    AA_SSAA_8X,
    AA_SUPER_4X,
    OIT_INTERLOCK,
    OIT_LINKEDLIST,
    OIT_LOOP,
    OIT_LOOP64,
    OIT_SIMPLE,
    OIT_SPINLOCK,
    OIT_WEIGHTED,
    OitError,
    lib,
)
from .renderer import FrameConfig, Renderer, State  # noqa: F401
from . import scene  # noqa: F401
from .device import DeviceBuffer, Event, Stream, device_count  # noqa: F401
from .batch import Batch, BatchFrame, render_noalloc  # noqa: F401
from .serve import Request, Server  # noqa: F401
