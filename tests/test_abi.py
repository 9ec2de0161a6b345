"""The C ABI boundary without a GPU: the library loads, exports every function
include/oit.h declares, and the ctypes structures have the C layout (checked
against offsetof/sizeof from a C program compiled from the header)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from vk_order_independent_transparency_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oit.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oit_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported_and_bound():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_abi.SIGNATURES) == set(names)


STRUCTS = {
    "oit_config": _abi.OitConfig,
    "oit_fragments": _abi.OitFragments,
    "oit_sizes": _abi.OitSizes,
    "oit_stats": _abi.OitStats,
    "oit_profile": _abi.OitProfile,
    "oit_scene_params": _abi.OitSceneParams,
    "oit_batch_frame": _abi.OitBatchFrame,
    "oit_batch_tuning": _abi.OitBatchTuning,
    "oit_batch_sizes": _abi.OitBatchSizes,
    "oit_server_config": _abi.OitServerConfig,
    "oit_request": _abi.OitRequest,
    "oit_completion": _abi.OitCompletion,
    "oit_server_stats": _abi.OitServerStats,
}


def test_struct_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oit.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)]).decode().splitlines():
        k, v = line.rsplit(" ", 1)
        got[k] = int(v)
    for cname, cls in STRUCTS.items():
        assert got[f"{cname} size"] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[f"{cname}.{fname}"] == getattr(cls, fname).offset, f"{cname}.{fname}"


def test_enums_match_reference_values():
    # shaders/common.h:44-51 and :58-64 of the reference
    assert (_abi.OIT_SIMPLE, _abi.OIT_LINKEDLIST, _abi.OIT_LOOP, _abi.OIT_LOOP64, _abi.OIT_SPINLOCK,
            _abi.OIT_INTERLOCK, _abi.OIT_WEIGHTED) == tuple(range(7))
    assert (_abi.AA_NONE, _abi.AA_MSAA_4X, _abi.AA_SSAA_4X, _abi.AA_SUPER_4X, _abi.AA_MSAA_8X,
            _abi.AA_SSAA_8X) == tuple(range(6))
    text = open(HEADER).read()
    for name, value in (("OIT_SIMPLE", 0), ("OIT_LINKEDLIST", 1), ("OIT_LOOP64", 3), ("OIT_SPINLOCK", 4),
                        ("OIT_WEIGHTED", 6)):
        assert re.search(rf"{name}\s*=\s*{value}\b", text)


def test_status_strings():
    lib = _abi.lib()
    assert lib.oit_status_string(0) == b"ok"
    assert lib.oit_status_string(-6) == b"capacity exceeded"
    assert lib.oit_status_string(12345) == b"unknown status"


def test_null_arguments_are_errors_not_crashes():
    lib = _abi.lib()
    assert lib.oit_create(None, 0) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_destroy(None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_configure(None, None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_clear(None, None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_append(None, None, None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_resolve(None, None, None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_device_alloc(None, 16) == _abi.OIT_ERR_INVALID_ARG


def test_create_without_a_gpu_reports_device_error():
    from vk_order_independent_transparency_amd import device_count

    if device_count() > 0:
        pytest.skip("a GPU is present")
    ctx = ctypes.c_void_p()
    assert _abi.lib().oit_create(ctypes.byref(ctx), 0) == _abi.OIT_ERR_DEVICE
    assert not ctx.value


def test_state_mirrors_reference_antialiasing_rules():
    from vk_order_independent_transparency_amd import FrameConfig, State

    s = State()
    assert (s.algorithm, s.oitLayers, s.linkedListAllocatedPerElement, s.tailBlend) == (_abi.OIT_SPINLOCK, 8, 10,
                                                                                       True)
    for aa, (msaa, ss, sup) in {0: (1, False, 1), 1: (4, False, 1), 2: (4, True, 1), 3: (1, False, 2),
                                4: (8, False, 1), 5: (8, True, 1)}.items():
        s.aaType = aa
        s.recomputeAntialiasingSettings()
        assert (s.msaa, s.sampleShading, s.supersample) == (msaa, ss, sup)
    s.aaType = _abi.AA_SSAA_4X
    cfg = FrameConfig.from_state(s, 64, 32)
    assert (cfg.width, cfg.height, cfg.samples) == (64, 32, 4)
    s.aaType = _abi.AA_SUPER_4X
    cfg = FrameConfig.from_state(s, 64, 32)
    assert (cfg.width, cfg.height, cfg.samples) == (128, 64, 1)
    s.aaType = _abi.AA_MSAA_4X  # coverage shading: per-pixel fragments with 4-sample masks
    cfg = FrameConfig.from_state(s, 64, 32)
    assert (cfg.width, cfg.height, cfg.samples, cfg.coverage_samples) == (64, 32, 1, 4)


def test_library_codec_tables_equal_the_oracle(oracle):
    """The product computes its sRGB8 tables with its own host code; they must equal the
    oracle's (which test_oracle_codec.py pins to the exact curve)."""
    import numpy as np

    thr = np.zeros(257, np.float32)
    dec = np.zeros(256, np.float32)
    alpha = np.zeros(256, np.float32)
    assert _abi.lib().oit_codec_tables(thr.ctypes.data, dec.ctypes.data, alpha.ctypes.data) == 0
    othr, odec, oalpha = oracle.tables()
    assert thr[0] == -np.inf and thr[256] == np.inf
    assert np.array_equal(thr[1:256].view(np.uint32), othr[1:256].view(np.uint32))
    assert np.array_equal(dec.view(np.uint32), odec.view(np.uint32))
    assert np.array_equal(alpha.view(np.uint32), oalpha.view(np.uint32))


def test_device_scene_generator_argument_checks():
    lib = _abi.lib()
    from vk_order_independent_transparency_amd import device_count, scene

    g = ctypes.c_void_p()
    assert lib.oit_scene_gen_create(None, None, 0) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_scene_gen_create(ctypes.byref(g), None, 0) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_scene_gen_destroy(None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_scene_gen_emit(None, None, None, None, 0, None) == _abi.OIT_ERR_INVALID_ARG
    bad = scene.SceneParams(seed=1, width=8, height=8, spheres=1, radius=(0.1, 0.2), alpha=(0.1, 0.2),
                            samples=3).to_c()
    assert lib.oit_scene_gen_create(ctypes.byref(g), ctypes.byref(bad), 0) == _abi.OIT_ERR_INVALID_ARG
    if device_count() == 0:
        ok = scene.SMALL_256.to_c()
        assert lib.oit_scene_gen_create(ctypes.byref(g), ctypes.byref(ok), 0) == _abi.OIT_ERR_DEVICE
        assert not g.value


def test_server_argument_checks():
    lib = _abi.lib()
    from vk_order_independent_transparency_amd import device_count

    s = ctypes.c_void_p()
    cfg = _abi.OitServerConfig(slots=0, reserved=0, budget_bytes=1 << 30)
    assert lib.oit_server_create(ctypes.byref(s), 0, ctypes.byref(cfg)) == _abi.OIT_ERR_INVALID_ARG
    cfg.slots = 17
    assert lib.oit_server_create(ctypes.byref(s), 0, ctypes.byref(cfg)) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_server_submit(None, None) == _abi.OIT_ERR_INVALID_ARG
    assert lib.oit_server_destroy(None) == _abi.OIT_ERR_INVALID_ARG
    if device_count() == 0:
        cfg.slots = 2
        assert lib.oit_server_create(ctypes.byref(s), 0, ctypes.byref(cfg)) == _abi.OIT_ERR_DEVICE


def test_served_workload_is_deterministic():
    from vk_order_independent_transparency_amd import serve

    a, ta = serve.served_workload(7, 50, 200.0)
    b, tb = serve.served_workload(7, 50, 200.0)
    assert a == b and (ta == tb).all()
    assert (np.diff(ta) > 0).all()
    gaps = np.diff(np.concatenate([[0.0], ta]))
    assert 2000 < gaps.mean() < 8000  # mean gap 1/rate = 5000 us
    assert {r.algorithm for r in a} <= {1, 3, 6} and {r.layers for r in a} <= {8, 16, 32}
    assert all((r.scene.width, r.scene.height) in serve.RESOLUTIONS for r in a)
    assert all(200 <= r.scene.spheres <= 8000 for r in a)
