"""The C-ABI library loads without a GPU and exports every entry point that
include/gs_hip.h declares; the Python binding names the same set."""
import ctypes
import os
import re
import subprocess

import pytest

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


def test_binding_structs_match_the_header(tmp_path):
    """Every field of the binding's structs sits at the offset the C compiler
    gives it (gcc on include/gs_hip.h)."""
    structs = [gs.TrainerOptions, gs.TrainerInfo, gs.TrainerStats, gs.PassTimes, gs.TrainerEval, gs.Tuning,
               gs.Frame, gs.Camera]
    c_names = {gs.TrainerOptions: "gs_hip_trainer_options", gs.TrainerInfo: "gs_hip_trainer_info",
               gs.TrainerStats: "gs_hip_trainer_stats", gs.PassTimes: "gs_hip_pass_times",
               gs.TrainerEval: "gs_hip_trainer_eval", gs.Tuning: "gs_hip_tuning", gs.Frame: "gs_hip_frame",
               gs.Camera: "gs_hip_camera"}
    lines = []
    for s in structs:
        c = c_names[s]
        lines.append(f'printf("{c} %zu\\n", sizeof({c}));')
        for name, _ in s._fields_:
            lines.append(f'printf("{c}.{name} %zu\\n", offsetof({c}, {name}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"gs_hip.h\"\nint main(void) {\n" +
                   "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for s in structs:
        c = c_names[s]
        assert int(got[c]) == ctypes.sizeof(s), c
        for name, _ in s._fields_:
            assert int(got[f"{c}.{name}"]) == getattr(s, name).offset, (c, name)


def test_trainer_create_reports_a_missing_folder_without_a_gpu(tmp_path):
    """The folder is read before the device is touched: a missing one is an
    I/O error with the path in the message, and no handle comes back."""
    lib = gs.library()
    o = gs.TrainerOptions(0, 64, 100, 1000, 1, 1, 0, 0, 1, 1)
    h = ctypes.c_void_p(1)
    missing = str(tmp_path / "no_such_scene")
    rc = lib.gs_hip_trainer_create(missing.encode(), ctypes.byref(o), ctypes.byref(h), None)
    buf = ctypes.create_string_buffer(512)
    assert rc == 4 and h.value is None  # GS_HIP_E_IO
    assert lib.gs_hip_last_error(buf, 512) == 4 and b"no_such_scene" in buf.value
    o.sh_degree = 4
    assert lib.gs_hip_trainer_create(missing.encode(), ctypes.byref(o), ctypes.byref(h), None) == 2
    with pytest.raises(gs.GsHipError, match="I/O|no_such_scene|transforms"):
        gs.NativeTrainer(missing, width=64, iterations=100, max_gaussians=1000, sh_degree=1)
    assert lib.gs_hip_trainer_destroy(None) == 0


def test_every_entry_point_has_argument_types():
    """Without argtypes ctypes passes a Python int as a C int: a size_t or a
    pointer above 2^31 would arrive truncated."""
    lib = gs.library()
    for name in gs.EXPORTED:
        assert getattr(lib, name).argtypes is not None, name
