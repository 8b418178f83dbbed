"""The C-ABI library loads, exports every entry point include/pk.h declares,
and fails loudly (no CPU fallback) where a device is needed but absent."""
import ctypes
import os
import re

import numpy as np
import pytest

import pike_amd
from pike_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pk_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_ffi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    assert sorted(_ffi.SIGNATURES) == declared_functions()


def test_version_and_error_string():
    lib = pike_amd.lib()
    assert b"gfx950" in lib.pk_version()
    rc = lib.pk_bvh_build(None, 0, None, 0, 1, ctypes.byref(ctypes.c_void_p()))
    assert rc == _ffi.PK_E_INVALID
    assert b"pk_bvh_build" in lib.pk_last_error()


def _gpu_visible():
    return os.path.exists("/dev/kfd") and any(e.startswith("renderD") for e in os.listdir("/dev/dri")) \
        if os.path.isdir("/dev/dri") else False


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is visible here")
def test_context_without_device_fails_loudly():
    with pytest.raises(pike_amd.PikeError) as e:
        pike_amd.Context(0)
    assert e.value.code == _ffi.PK_E_NO_DEVICE


def test_null_handles_are_rejected():
    lib = pike_amd.lib()
    assert lib.pk_sync(None) == _ffi.PK_E_INVALID
    assert lib.pk_frame_begin(None, 0) == _ffi.PK_E_INVALID
    assert lib.pk_trace_closest(None, None, 0) == _ffi.PK_E_INVALID
    assert lib.pk_render(None, None, None) == _ffi.PK_E_INVALID
    assert lib.pk_scene_upload(None, None, None, ctypes.byref(ctypes.c_void_p())) == _ffi.PK_E_INVALID
    assert lib.pk_read_image(None, None, 0) == _ffi.PK_E_INVALID
    info = np.zeros(6, np.int64)
    assert lib.pk_bvh_info(None, info.ctypes.data) == _ffi.PK_E_INVALID
    assert lib.pk_env_tables(None, info.ctypes.data, None, None, None) == _ffi.PK_E_INVALID
