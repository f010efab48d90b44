"""The C-ABI library loads without a GPU and exports every entry point that
include/gs_hip.h declares; the Python binding names the same set."""
import ctypes
import os
import re

from helpers import ROOT, gs


def declared():
    text = open(os.path.join(ROOT, "include", "gs_hip.h")).read()
    return sorted(set(re.findall(r"^\s*int\s+(gs_hip_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gs.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert sorted(gs.EXPORTED) == declared()


def test_frame_struct_layout():
    # the C struct: camera (9+3+3+4 floats + 2 u32 = 84 B), 6 u32, 3 floats bg, 2+1+6+5+6 floats, 3 exposure
    assert ctypes.sizeof(gs.Camera) == 84
    assert ctypes.sizeof(gs.Frame) == 84 + 6 * 4 + 3 * 4 + 3 * 4 + 6 * 4 + 5 * 4 + 6 * 4 + 3 * 4
    assert gs.Frame.exposure.offset == ctypes.sizeof(gs.Frame) - 12


def test_errors_surface_as_exceptions_without_a_gpu():
    lib = gs.library()
    buf = ctypes.create_string_buffer(256)
    rc = lib.gs_hip_workspace_create(None, 16, 0, 64, 48)
    assert rc == 2  # GS_HIP_E_ARGUMENT, before any device call
    assert lib.gs_hip_last_error(buf, 256) == 2 and b"null" in buf.value
