"""ctypes view of include/oit.h and loader of the in-tree liboit_hip.so.

The library is the product: there is no fallback. If liboit_hip.so is not
built, importing the package's renderer raises ImportError naming the build
step (``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes
import os

LIB_NAME = "liboit_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# enum oit_algorithm: the reference's shaders/common.h:44-51
OIT_SIMPLE = 0
OIT_LINKEDLIST = 1
OIT_LOOP = 2
OIT_LOOP64 = 3
OIT_SPINLOCK = 4
OIT_INTERLOCK = 5
OIT_WEIGHTED = 6
NUM_ALGORITHMS = 7

# AA types: the reference's shaders/common.h:58-64
AA_NONE = 0
AA_MSAA_4X = 1
AA_SSAA_4X = 2
AA_SUPER_4X = 3
AA_MSAA_8X = 4
AA_SSAA_8X = 5

OIT_OK = 0
OIT_ERR_INVALID_ARG = -1
OIT_ERR_OOM = -2
OIT_ERR_OVER_BUDGET = -3
OIT_ERR_DEVICE = -4
OIT_ERR_UNSUPPORTED = -5
OIT_ERR_CAPACITY = -6
OIT_ERR_STATE = -7
OIT_ERR_NOT_READY = -8

VARIANT_LL_SPLIT_APPEND = 0x1
VARIANT_LL_SINGLE_WALK = 0x2
VARIANT_LL_APPEND_8 = 0x4
VARIANT_LL_KEYS_ONLY = 0x8
VARIANT_STORE_CHUNKS_1 = 0x10
VARIANT_STORE_CHUNKS_4 = 0x20

BATCH_CONCURRENT = 0x1
BATCH_GRAPH = 0x2
BATCH_FRAME_ORDER = 0x4

OIT_PROFILE_STAGES = 6
DEBUG_LL_HEAD, DEBUG_LL_NEXT, DEBUG_SPIN_LOCK = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 256 << 20
PROFILE_STAGE_NAMES = ("clear", "store", "link", "classify", "tail_link", "resolve")


class OitConfig(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_uint32),
        ("layers", ctypes.c_uint32),
        ("tail_blend", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("ll_nodes_per_pixel", ctypes.c_uint32),
        ("variant", ctypes.c_uint32),
        ("ll_pool_nodes", ctypes.c_uint64),
        ("max_fragments", ctypes.c_uint64),
        ("background", ctypes.c_float * 4),
        ("wboit_view_near", ctypes.c_float),
        ("wboit_view_span", ctypes.c_float),
        ("coverage_samples", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class OitFragments(ctypes.Structure):
    _fields_ = [
        ("target", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("rgba", ctypes.c_void_p),
        ("count", ctypes.c_uint64),
        ("coverage", ctypes.c_void_p),
    ]


class OitSizes(ctypes.Structure):
    _fields_ = [
        ("abuffer_bytes", ctypes.c_uint64),
        ("aux_bytes", ctypes.c_uint64),
        ("tail_bytes", ctypes.c_uint64),
        ("total_bytes", ctypes.c_uint64),
        ("ll_capacity", ctypes.c_uint64),
        ("bytes_per_element", ctypes.c_uint32),
        ("targets", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("resident_bytes", ctypes.c_uint64),
    ]


class OitStats(ctypes.Structure):
    _fields_ = [
        ("fragments", ctypes.c_uint64),
        ("ll_overflow", ctypes.c_uint64),
        ("tail_fragments", ctypes.c_uint64),
        ("appends", ctypes.c_uint64),
        ("device_errors", ctypes.c_uint64),
    ]


class OitProfile(ctypes.Structure):
    _fields_ = [
        ("ms", ctypes.c_double * OIT_PROFILE_STAGES),
        ("launches", ctypes.c_uint64 * OIT_PROFILE_STAGES),
    ]


class OitBatchFrame(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("layers", ctypes.c_uint32),
        ("tail_blend", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("background", ctypes.c_float * 4),
        ("frags", OitFragments),
        ("out", ctypes.c_void_p),
    ]


class OitBatchTuning(ctypes.Structure):
    _fields_ = [
        ("append_chunks", ctypes.c_uint32),
        ("resolve_variant", (ctypes.c_uint32 * 2) * 6),
        ("resolve_blocks", (ctypes.c_uint32 * 2) * 6),
        ("bin_density", ctypes.c_float),
    ]


class OitBatchSizes(ctypes.Structure):
    _fields_ = [
        ("pool_nodes", ctypes.c_uint64),
        ("head_entries", ctypes.c_uint64),
        ("bytes", ctypes.c_uint64),
        ("groups", ctypes.c_uint32),
        ("frames", ctypes.c_uint32),
    ]


class OitSceneParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("spheres", ctypes.c_uint32),
        ("radius_min", ctypes.c_float),
        ("radius_max", ctypes.c_float),
        ("alpha_min", ctypes.c_float),
        ("alpha_max", ctypes.c_float),
        ("cluster_fraction", ctypes.c_float),
        ("cluster_extent", ctypes.c_float),
    ]


class OitServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("budget_bytes", ctypes.c_uint64)]


class OitRequest(ctypes.Structure):
    _fields_ = [
        ("scene", OitSceneParams),
        ("algorithm", ctypes.c_uint32),
        ("layers", ctypes.c_uint32),
        ("tail_blend", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("out", ctypes.c_void_p),
        ("tag", ctypes.c_uint64),
    ]


class OitCompletion(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("slot", ctypes.c_uint32),
        ("fragments", ctypes.c_uint64),
        ("abuffer_bytes", ctypes.c_uint64),
        ("out", ctypes.c_void_p),
        ("t_submit_us", ctypes.c_double),
        ("t_start_us", ctypes.c_double),
        ("t_done_us", ctypes.c_double),
        ("device_us", ctypes.c_double),
    ]


class OitServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("held_bytes", ctypes.c_uint64),
        ("peak_held_bytes", ctypes.c_uint64),
        ("peak_in_flight", ctypes.c_uint32),
        ("slots", ctypes.c_uint32),
    ]


# name -> (restype, argtypes); every symbol include/oit.h declares
_P = ctypes.c_void_p
SIGNATURES = {
    "oit_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "oit_destroy": (ctypes.c_int, [_P]),
    "oit_configure": (ctypes.c_int, [_P, ctypes.POINTER(OitConfig)]),
    "oit_query_sizes": (ctypes.c_int, [_P, ctypes.POINTER(OitSizes)]),
    "oit_trim": (ctypes.c_int, [_P]),
    "oit_clear": (ctypes.c_int, [_P, _P]),
    "oit_append": (ctypes.c_int, [_P, ctypes.POINTER(OitFragments), _P]),
    "oit_draw_opaque": (ctypes.c_int, [_P, ctypes.POINTER(OitFragments), _P]),
    "oit_resolve": (ctypes.c_int, [_P, _P, _P]),
    "oit_resolve_debug": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "oit_debug_set_word": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _P]),
    "oit_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(OitStats), _P]),
    "oit_finalize_srgb8": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          _P, _P]),
    "oit_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "oit_profile_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "oit_profile_read": (ctypes.c_int, [_P, ctypes.POINTER(OitProfile)]),
    "oit_profile_reset": (ctypes.c_int, [_P]),
    "oit_codec_tables": (ctypes.c_int, [_P, _P, _P]),
    "oit_selftest_codec": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint32)]),
    "oit_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "oit_set_device": (ctypes.c_int, [ctypes.c_int]),
    "oit_device_alloc": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_uint64]),
    "oit_device_free": (ctypes.c_int, [_P]),
    "oit_memcpy_h2d": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P]),
    "oit_memcpy_d2h": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P]),
    "oit_memcpy_d2d": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P]),
    "oit_stream_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "oit_stream_destroy": (ctypes.c_int, [_P]),
    "oit_stream_synchronize": (ctypes.c_int, [_P]),
    "oit_event_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "oit_event_destroy": (ctypes.c_int, [_P]),
    "oit_event_record": (ctypes.c_int, [_P, _P]),
    "oit_event_synchronize": (ctypes.c_int, [_P]),
    "oit_event_elapsed_ms": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), _P, _P]),
    "oit_scene_count": (ctypes.c_int, [ctypes.POINTER(OitSceneParams), ctypes.POINTER(ctypes.c_uint64)]),
    "oit_scene_generate_host": (
        ctypes.c_int,
        [ctypes.POINTER(OitSceneParams), _P, _P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)],
    ),
    "oit_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_uint32]),
    "oit_batch_destroy": (ctypes.c_int, [_P]),
    "oit_batch_query": (ctypes.c_int, [_P, ctypes.POINTER(OitBatchFrame), ctypes.c_uint32,
                                       ctypes.POINTER(OitBatchSizes)]),
    "oit_batch_reserve": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64]),
    "oit_batch_run": (ctypes.c_int, [_P, ctypes.POINTER(OitBatchFrame), ctypes.c_uint32, _P]),
    "oit_batch_tune": (ctypes.c_int, [_P, ctypes.POINTER(OitBatchFrame), ctypes.c_uint32, _P]),
    "oit_batch_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(OitBatchTuning)]),
    "oit_batch_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(OitBatchTuning)]),
    "oit_ll_render_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(OitBatchFrame), _P]),
    "oit_server_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(OitServerConfig)]),
    "oit_server_destroy": (ctypes.c_int, [_P]),
    "oit_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(OitRequest)]),
    "oit_server_poll": (ctypes.c_int, [_P, ctypes.POINTER(OitCompletion), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]),
    "oit_server_drain": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "oit_server_replay": (ctypes.c_int, [_P, ctypes.POINTER(OitRequest), ctypes.POINTER(ctypes.c_double),
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(OitCompletion)]),
    "oit_server_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(OitServerStats)]),
    "oit_scene_gen_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(OitSceneParams), ctypes.c_int]),
    "oit_scene_gen_destroy": (ctypes.c_int, [_P]),
    "oit_scene_gen_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "oit_scene_gen_emit": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint64, _P]),
}

_lib = None


class OitError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, fn: str, status: int):
        self.status = status
        msg = lib().oit_status_string(status)
        super().__init__(f"{fn} failed: {msg.decode() if msg else status} ({status})")


def lib() -> ctypes.CDLL:
    """Loads liboit_hip.so once; raises ImportError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -c \"import __graft_entry__ as g; g.build()\"`"
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(fn: str, status: int) -> None:
    if status != OIT_OK:
        raise OitError(fn, status)
