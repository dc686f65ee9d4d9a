"""The C-ABI library loads and exports every entry point include/pepc_gpu.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pepc_gpu.h")).read()
    return sorted(set(re.findall(r"\b(pepc_gpu_\w+)\s*\(", src)))


def test_library_exports_header_symbols():
    from pepc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libpepc_gpu.so not built: run python -m pepc_amd.build")
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 17
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def test_config_default_without_device():
    from pepc_amd import _lib
    cfg = _lib.default_config()
    assert cfg.theta2 == 0.36 and cfg.eps2 == 0.0 and cfg.group_size == 16
    assert cfg.precision == _lib.FP32_KAHAN and cfg.node_factor == 2.5


def test_create_rejects_bad_config_before_touching_device():
    from pepc_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    for kw in ({"group_size": 0}, {"group_size": 65}, {"sign": 2.0},
               {"theta2": -1.0}, {"precision": 7}, {"list_tile": 32}, {"list_tile": 100},
               {"precision": 1, "list_tile": 128}, {"group_pack": 65}):
        cfg = _lib.default_config(**kw)
        assert L.pepc_gpu_create(ctypes.byref(h), ctypes.byref(cfg)) == _lib.EINVAL

