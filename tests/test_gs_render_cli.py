"""gs_render end to end (the reference's command surface,
engine/cli/gs_render.cpp:1-254): a 3DGS PLY written here, rendered through
--views cameras.txt (two output sizes) and as one orbit view.  Every --views
image equals, byte for byte, the same camera through gs_hip_view_render with
the PLY's values and SH rest rounded to fp16 (the viewer's storage)."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from helpers import ROOT, gs, make_camera, make_scene

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "on-divice-gsplate_amd", "gs_render")
CLEAR = (0.08, 0.09, 0.11, 1.0)


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    yield name
    gs.synchronize()


def write_ply(path, params, sh):
    n = params.shape[0]
    coeffs = sh.shape[1]
    names = ["x", "y", "z"] + [f"f_dc_{c}" for c in range(3)] + [f"f_rest_{i}" for i in range(3 * coeffs)] + \
        ["opacity"] + [f"scale_{k}" for k in range(3)] + [f"rot_{k}" for k in range(4)]
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    head += "".join(f"property float {p}\n" for p in names) + "end_header\n"
    cols = [params[:, 0:3], params[:, 12:15], sh.transpose(0, 2, 1).reshape(n, -1), params[:, 3:4], params[:, 4:7],
            params[:, 8:12]]
    body = np.concatenate(cols, axis=1).astype("<f4")
    with open(path, "wb") as f:
        f.write(head.encode() + body.tobytes())


def read_ppm(path):
    data = open(path, "rb").read()
    m = re.match(rb"P6\s+(\d+)\s+(\d+)\s+255\s", data)
    w, h = int(m.group(1)), int(m.group(2))
    return np.frombuffer(data[m.end():], np.uint8).reshape(h, w, 3)


def test_gs_render_views_and_orbit(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    print("[case] gs_render --views (2 sizes) and an orbit view", flush=True)
    n, deg = 800, 2
    params, sh = make_scene(n, deg, 77)
    ply = tmp_path / "scene.ply"
    write_ply(ply, params, sh)
    cams = [make_camera(128, 96, 0.3), make_camera(96, 64, 1.7), make_camera(128, 96, 2.9)]
    with open(tmp_path / "cameras.txt", "w") as f:
        for c in cams:
            R = list(c.rotation)
            t = list(c.translation)
            rows = [R[0], R[1], R[2], t[0], R[3], R[4], R[5], t[1], R[6], R[7], R[8], t[2]]
            f.write(" ".join(repr(float(v)) for v in [c.width, c.height, c.fx, c.fy, c.cx, c.cy] + rows) + "\n\n")
    out = tmp_path / "views"
    out.mkdir()
    r = subprocess.run([EXE, str(ply), "--views", str(tmp_path / "cameras.txt"), str(out)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    sh16 = sh.astype(np.float16).astype(np.float32)
    for i, c in enumerate(cams):
        got = read_ppm(out / f"view_{i:05d}.ppm")
        with gs.Workspace(n, deg, c.width, c.height) as ws:
            ws.load(params, sh16)
            want, vis = ws.view_render(gs.make_frame(c, n, deg), CLEAR)
        assert got.shape == (c.height, c.width, 3)
        np.testing.assert_array_equal(got, want[..., :3])
        assert vis > 0
    r = subprocess.run([EXE, str(ply), str(tmp_path / "orbit.ppm"), "160", "120", "30", "10"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0 and re.search(r"rendered 160x120: \d+ visible", r.stdout)
    img = read_ppm(tmp_path / "orbit.ppm")
    assert img.shape == (120, 160, 3)
    assert (np.abs(img.astype(int) - np.array([20, 23, 28])).sum(-1) > 30).mean() > 0.02  # something drawn
