"""The C ABI library without a GPU: it loads, exports exactly what
include/rtm_hip.h declares, and fails cleanly before initialisation."""
import ctypes
import os
import re

import pytest

from rtm3d_demo_amd import hip

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(REPO, "include", "rtm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rtmhip_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = hip.load()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(hip.EXPORTED) == names


def test_calls_before_init_fail_cleanly():
    lib = hip.load()
    p = ctypes.c_void_p()
    rc = lib.rtmhip_alloc(ctypes.byref(p), 16)
    if rc == hip.OK:  # a GPU is present and someone initialised already
        lib.rtmhip_free(p)
        pytest.skip("library already initialised in this process")
    assert rc == hip.ERR_NOT_INIT
    assert b"rtmhip_init" in lib.rtmhip_last_error()
    assert lib.rtmhip_step(None, None, None, None, 8, 8, 8, 8) == hip.ERR_NOT_INIT


def test_version_and_device_count():
    assert hip.load().rtmhip_version() >= 10000
    assert hip.device_count() >= 0


def test_host_library_exports():
    from rtm3d_demo_amd import host
    L = host.load()
    for n in ("rtm3d_stencil_coefs", "rtm3d_ricker", "rtm3d_taper_table",
              "rtm3d_extend_velocity", "rtm3d_v2dt2", "rtm3d_run_json"):
        assert hasattr(L, n)
