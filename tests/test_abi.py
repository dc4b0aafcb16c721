"""The C ABI: librdp_mi.so loads without a GPU and exports every function
include/rdp_mi.h declares; the Python binding covers the same set."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rdp_mi.h")
LIB = os.path.join(ROOT, "parallel-rdp_amd", "librdp_mi.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(rdp_mi_[a-z0-9_]+)\s*\(", text))
    types = set(re.findall(r"\(\s*\*\s*(rdp_mi_[a-z0-9_]+)\s*\)", text))
    return sorted(names - types)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import importlib.util
        spec = importlib.util.spec_from_file_location("rdp_build", os.path.join(ROOT, "parallel-rdp_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    return ctypes.CDLL(LIB)


def test_header_declares_the_command_processor_surface():
    names = declared_functions()
    for required in ("rdp_mi_create", "rdp_mi_destroy", "rdp_mi_enqueue_command", "rdp_mi_flush",
                     "rdp_mi_signal_timeline", "rdp_mi_wait_for_timeline", "rdp_mi_idle",
                     "rdp_mi_begin_read_rdram", "rdp_mi_set_vi_register", "rdp_mi_scanout_sync"):
        assert required in names


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_wraps_declared_functions():
    src = open(os.path.join(ROOT, "parallel-rdp_amd", "rdp_mi.py")).read()
    unwrapped = [n for n in declared_functions() if n not in src]
    assert not unwrapped, f"binding does not wrap: {unwrapped}"


def test_create_rejects_bad_arguments_without_touching_the_device(lib):
    lib.rdp_mi_create.restype = ctypes.c_int
    lib.rdp_mi_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                  ctypes.c_size_t, ctypes.c_uint32]
    lib.rdp_mi_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    assert lib.rdp_mi_create(ctypes.byref(h), None, 0, 3 << 20, 3 << 19, 0) != 0
    assert b"power of two" in lib.rdp_mi_last_error()
    assert lib.rdp_mi_create(ctypes.byref(h), None, 0, 8 << 20, 1 << 20, 0) != 0
    assert not h.value


def test_server_rejects_bad_arguments_without_touching_the_device(lib):
    vp = ctypes.c_void_p
    lib.rdp_mi_server_create.restype = ctypes.c_int
    lib.rdp_mi_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_uint, ctypes.c_size_t, ctypes.c_uint,
                                         ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(vp)]
    lib.rdp_mi_server_start.restype = ctypes.c_int
    lib.rdp_mi_server_start.argtypes = [vp]
    lib.rdp_mi_server_reset.restype = ctypes.c_int
    lib.rdp_mi_server_reset.argtypes = [vp, ctypes.c_double]
    lib.rdp_mi_last_error.restype = ctypes.c_char_p
    out = vp()
    assert lib.rdp_mi_server_create(None, 0, 0, 0, 0, 0, ctypes.byref(out)) != 0
    assert not out.value
    procs = (vp * 2)(None, None)
    assert lib.rdp_mi_server_create(procs, 2, 0, 0, 0, 0, ctypes.byref(out)) != 0
    assert b"null processor" in lib.rdp_mi_last_error()
    assert lib.rdp_mi_server_start(None) != 0
    assert lib.rdp_mi_server_reset(None, 0.0) != 0


def test_launch_config_accepts_any_shape_without_a_device(lib):
    lib.rdp_mi_set_launch_config.restype = ctypes.c_int
    lib.rdp_mi_set_launch_config.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    for chunk, cls, tile, pers in ((-1, 0, 8, 0), (0, 1, 16, 0), (128, 2, 8, 1), (64, 9, 16, 0)):
        assert lib.rdp_mi_set_launch_config(chunk, cls, tile, pers) == 0
    assert lib.rdp_mi_set_launch_config(64, 0, 12, 0) != 0
    assert lib.rdp_mi_set_launch_config(64, 0, 16, 1) != 0
    assert lib.rdp_mi_set_launch_config(64, 0, 8, 0) == 0
