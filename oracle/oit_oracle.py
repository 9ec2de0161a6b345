"""ctypes binding of the CPU oracle (oracle/liboit_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use. The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# OIT_ORACLE_LIB selects another build of the same source (the sanitizer
# build of `make -C oracle sanitize`)
LIB_PATH = os.environ.get("OIT_ORACLE_LIB") or os.path.join(HERE, "liboit_oracle.so")

SIMPLE, LINKEDLIST, LOOP, LOOP64, SPINLOCK, INTERLOCK, WEIGHTED = 0, 1, 2, 3, 4, 5, 6
TAIL_CANONICAL, TAIL_SEQUENTIAL = 0, 1


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_uint32),
        ("layers", ctypes.c_uint32),
        ("tail_blend", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("tail_order", ctypes.c_uint32),
        ("ll_capacity", ctypes.c_uint64),
        ("background", ctypes.c_float * 4),
        ("wboit_view_near", ctypes.c_float),
        ("wboit_view_span", ctypes.c_float),
        ("coverage_samples", ctypes.c_uint32),
        ("opaque_key", ctypes.c_void_p),
        ("opaque_rgba", ctypes.c_void_p),
    ]


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("out", ctypes.c_void_p),
        ("composite", ctypes.c_void_p),
        ("colorbuf", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("frontk_key", ctypes.c_void_p),
        ("frontk_val", ctypes.c_void_p),
        ("tail_flag", ctypes.c_void_p),
        ("wb_accum", ctypes.c_void_p),
        ("wb_reveal", ctypes.c_void_p),
        ("ll_overflow", ctypes.c_uint64),
        ("tail_fragments", ctypes.c_uint64),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_tables.argtypes = [P, P, P]
        L.orc_tables.restype = None
        L.orc_depth_key.argtypes = [ctypes.c_float]
        L.orc_depth_key.restype = ctypes.c_uint32
        L.orc_encode.argtypes = [P]
        L.orc_encode.restype = ctypes.c_uint32
        L.orc_decode_premult.argtypes = [ctypes.c_uint32, P]
        L.orc_decode_premult.restype = None
        L.orc_srgb_encode_exact.argtypes = [ctypes.c_double]
        L.orc_srgb_encode_exact.restype = ctypes.c_double
        L.orc_srgb_decode_exact.argtypes = [ctypes.c_double]
        L.orc_srgb_decode_exact.restype = ctypes.c_double
        L.orc_run.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, ctypes.POINTER(OrcResult)]
        L.orc_run.restype = ctypes.c_int
        L.orc_run_fp64.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, ctypes.c_int, P, P, P]
        L.orc_run_fp64.restype = ctypes.c_int
        L.orc_replay_kbuffer.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, P, P, P, P,
                                         ctypes.c_int, P]
        L.orc_replay_kbuffer.restype = ctypes.c_int
        L.orc_run_cov.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, P, ctypes.POINTER(OrcResult)]
        L.orc_run_cov.restype = ctypes.c_int
        L.orc_replay_kbuffer_cov.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, P, P, P, P, P, P,
                                             ctypes.c_int, P]
        L.orc_replay_kbuffer_cov.restype = ctypes.c_int
        L.orc_finalize_srgb8.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, P]
        L.orc_finalize_srgb8.restype = ctypes.c_int
        L.orc_opaque_pass.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_uint64, P, P, P, P, P, P]
        L.orc_opaque_pass.restype = ctypes.c_int
        _lib = L
    return _lib


def tables() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    thr = np.zeros(256, np.float32)
    dec = np.zeros(256, np.float32)
    alpha = np.zeros(256, np.float32)
    lib().orc_tables(thr.ctypes.data, dec.ctypes.data, alpha.ctypes.data)
    return thr, dec, alpha


def depth_key(d: float) -> int:
    return int(lib().orc_depth_key(ctypes.c_float(d)))


def encode(premult) -> int:
    a = np.ascontiguousarray(premult, dtype=np.float32)
    return int(lib().orc_encode(a.ctypes.data))


def decode_premult(packed: int) -> np.ndarray:
    o = np.zeros(4, np.float32)
    lib().orc_decode_premult(ctypes.c_uint32(packed), o.ctypes.data)
    return o


def make_config(algorithm, layers, tail_blend, width, height, samples=1, tail_order=TAIL_CANONICAL,
                ll_capacity=0, background=(0.5, 0.5, 0.5, 1.0), wboit_view_near=2.0, wboit_view_span=8.0,
                coverage_samples=0, opaque=None):
    """`opaque`: the frame's opaque pass, (depth keys, colours) per sample
    target as returned by opaque_pass()."""
    c = OrcConfig()
    c.coverage_samples = coverage_samples
    c.algorithm = algorithm
    c.layers = layers
    c.tail_blend = 1 if tail_blend else 0
    c.width, c.height, c.samples = width, height, samples
    c.tail_order = tail_order
    c.ll_capacity = ll_capacity
    for i in range(4):
        c.background[i] = background[i]
    c.wboit_view_near = wboit_view_near
    c.wboit_view_span = wboit_view_span
    if opaque is not None:
        keys = np.ascontiguousarray(opaque[0], dtype=np.uint32)
        cols = np.ascontiguousarray(opaque[1], dtype=np.float32).reshape(-1, 4)
        assert keys.shape[0] == cols.shape[0]
        c._opaque = (keys, cols)  # kept alive with the config
        c.opaque_key = keys.ctypes.data
        c.opaque_rgba = cols.ctypes.data
    return c


def _arrays(target, depth, rgba):
    target = np.ascontiguousarray(target, dtype=np.uint32)
    depth = np.ascontiguousarray(depth, dtype=np.float32)
    rgba = np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1, 4)
    assert target.shape[0] == depth.shape[0] == rgba.shape[0]
    return target, depth, rgba


def opaque_pass(cfg: OrcConfig, target, depth, rgba, mask=None) -> tuple[np.ndarray, np.ndarray]:
    """The frame's opaque pass (depth test LESS with depth write over the
    opaque stream in draw order): depth keys [n] and colours [n, 4] per sample
    target, n = W*H*samples, or coverage_samples planes of W*H."""
    target, depth, rgba = _arrays(target, depth, rgba)
    planes = cfg.coverage_samples if cfg.coverage_samples > 1 else 1
    n = cfg.width * cfg.height * cfg.samples * planes
    keys = np.zeros(n, np.uint32)
    cols = np.zeros((n, 4), np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint32)
    rc = lib().orc_opaque_pass(ctypes.byref(cfg), target.shape[0], target.ctypes.data, depth.ctypes.data,
                               rgba.ctypes.data, None if m is None else m.ctypes.data, keys.ctypes.data,
                               cols.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_opaque_pass failed: {rc}")
    return keys, cols


def run(cfg: OrcConfig, target, depth, rgba, mask=None) -> dict:
    """Sequential run of one frame; returns numpy arrays of every output.
    `mask`: per-fragment sample masks (coverage shading)."""
    target, depth, rgba = _arrays(target, depth, rgba)
    mask_arr = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint32)
    F = target.shape[0]
    T = cfg.width * cfg.height * cfg.samples
    K = max(cfg.layers, 1)
    out = {
        "out": np.zeros((T, 4), np.float32),
        "composite": np.zeros((T, 4), np.float32),
        "colorbuf": np.zeros((T, 4), np.float32),
        "count": np.zeros(T, np.uint32),
        "frontk_key": np.zeros((T, K), np.uint32),
        "frontk_val": np.zeros((T, K), np.uint32),
        "tail_flag": np.zeros(max(F, 1), np.uint8),
        "wb_accum": np.zeros((T, 4), np.float32),
        "wb_reveal": np.zeros(T, np.float32),
    }
    r = OrcResult()
    for name in ("out", "composite", "colorbuf", "count", "frontk_key", "frontk_val", "tail_flag", "wb_accum",
                 "wb_reveal"):
        setattr(r, name, out[name].ctypes.data)
    rc = lib().orc_run_cov(ctypes.byref(cfg), F, target.ctypes.data, depth.ctypes.data, rgba.ctypes.data,
                           None if mask_arr is None else mask_arr.ctypes.data, ctypes.byref(r))
    if rc != 0:
        raise RuntimeError(f"orc_run failed: {rc}")
    out["tail_flag"] = out["tail_flag"][:F]
    out["ll_overflow"] = int(r.ll_overflow)
    out["tail_fragments"] = int(r.tail_fragments)
    return out


def run_fp64(cfg: OrcConfig, target, depth, rgba, quantize=True) -> dict:
    target, depth, rgba = _arrays(target, depth, rgba)
    F = target.shape[0]
    T = cfg.width * cfg.height * cfg.samples
    out = np.zeros((T, 4), np.float64)
    acc = np.zeros((T, 4), np.float64)
    rev = np.zeros(T, np.float64)
    rc = lib().orc_run_fp64(ctypes.byref(cfg), F, target.ctypes.data, depth.ctypes.data, rgba.ctypes.data,
                            1 if quantize else 0, out.ctypes.data, acc.ctypes.data, rev.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_run_fp64 failed: {rc}")
    return {"out": out, "wb_accum": acc, "wb_reveal": rev}


def replay_kbuffer(cfg: OrcConfig, target, rgba, count, ab_color, ab_key, ab_id, tail_flag, tail_quantized,
                   mask=None, ab_mask=None):
    """Composite of given k-buffer contents ([T, K] arrays) and a given tail set
    (with coverage shading: fragment masks and the A-buffer's mask plane)."""
    target = np.ascontiguousarray(target, dtype=np.uint32)
    rgba = np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1, 4)
    T = cfg.width * cfg.height * cfg.samples
    out = np.zeros((T, 4), np.float32)
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (count, ab_color, ab_key, ab_id)]
    tf = np.ascontiguousarray(tail_flag, dtype=np.uint8)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint32)
    am = None if ab_mask is None else np.ascontiguousarray(ab_mask, dtype=np.uint32)
    rc = lib().orc_replay_kbuffer_cov(ctypes.byref(cfg), target.shape[0], target.ctypes.data, rgba.ctypes.data,
                                      None if m is None else m.ctypes.data, arrs[0].ctypes.data,
                                      arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data,
                                      None if am is None else am.ctypes.data, tf.ctypes.data,
                                      1 if tail_quantized else 0, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_replay_kbuffer failed: {rc}")
    return out


def finalize_srgb8(image, width, height, samples=1, supersample=1) -> np.ndarray:
    """Display image (MSAA resolve and supersample box filter, then sRGB8):
    uint8 [height, width, 4] RGBA."""
    img = np.ascontiguousarray(image, dtype=np.float32).reshape(-1, 4)
    assert img.shape[0] == width * height * samples * supersample * supersample
    out = np.zeros((height, width, 4), np.uint8)
    rc = lib().orc_finalize_srgb8(img.ctypes.data, width, height, samples, supersample, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_finalize_srgb8 failed: {rc}")
    return out
