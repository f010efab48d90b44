"""gs_train end to end on a transforms.json folder (the reference's command
surface, engine/cli/gs_train.cpp): PNG frames and subject masks (coverage
1 - T) rendered by the oracle's fp64 renderer from a seeded scene, loaded,
initialised at random around the cameras, trained with masked targets over
random backgrounds, 4 views per optimizer step and an autotuned launch
shape, evaluated on the held-out frames, written as a 3DGS PLY."""
import json
import math
import os
import re
import struct
import subprocess
import zlib

import numpy as np
import pytest

from helpers import ROOT, gs, make_scene, orc

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "on-divice-gsplate_amd", "gs_train")


def png(img):
    h, w = img.shape[:2]
    colour = 2 if img.ndim == 3 else 0
    rows = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, colour, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(rows)) + chunk(b"IEND", b""))


def write_dataset(root, n, deg, w, h, views, seed, dims=None, opacity=0.0):
    """dims: {frame: brightness factor} applied to those photos (a camera's
    exposure); opacity: added to every opacity logit."""
    params, sh = make_scene(n, deg, seed, spread=0.6)
    params[:, 3] += opacity
    frames = []
    os.makedirs(os.path.join(root, "images"), exist_ok=True)
    os.makedirs(os.path.join(root, "masks"), exist_ok=True)
    for k in range(views):
        a = 2 * math.pi * k / views
        eye = (3.0 * math.sin(a), 0.6 + 0.3 * math.sin(3 * a), 3.0 * math.cos(a))
        cam = gs.look_at_camera(eye, (0.0, 0.0, 0.0), w, h)
        rgb, T = orc.render64(params, sh, cam, deg)
        cover = np.clip(1.0 - T.reshape(h, w), 1e-6, 1.0)
        # the photo: colour un-premultiplied where covered; the mask: coverage
        dim = (dims or {}).get(k, 1.0)
        img = np.clip(np.rint(rgb.reshape(h, w, 3) / cover[..., None] * 255.0 * dim), 0, 255).astype(np.uint8)
        mask = np.clip(np.rint((1.0 - T.reshape(h, w)) * 255.0), 0, 255).astype(np.uint8)
        name, mname = f"images/frame_{k:03d}.png", f"masks/frame_{k:03d}.png"
        with open(os.path.join(root, name), "wb") as f:
            f.write(png(img))
        with open(os.path.join(root, mname), "wb") as f:
            f.write(png(mask))
        R = np.array(cam.rotation, np.float64).reshape(3, 3)
        c2w = np.eye(4)
        c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = R[0], -R[1], -R[2], eye
        frames.append({"file_path": name, "mask_path": mname, "transform_matrix": c2w.tolist()})
    with open(os.path.join(root, "transforms.json"), "w") as f:
        json.dump({"w": w, "h": h, "fl_x": cam.fx, "fl_y": cam.fy, "cx": cam.cx, "cy": cam.cy, "frames": frames}, f)


def test_gs_train_on_a_transforms_json_folder(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    print("[case] gs_train dataset folder with masks, 4 views per step, autotune", flush=True)
    data = tmp_path / "scene"
    write_dataset(str(data), 300, 1, 96, 72, 16, 61)
    out = tmp_path / "out.ply"
    cmd = [EXE, str(data), str(out), "--width", "72", "--iters", "1600", "--sh", "1", "--init-gaussians", "1000",
           "--max-gaussians", "6000", "--views-per-step", "4", "--autotune", "--eval-every", "400"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "dataset scene: 16 frames (2 held out, 0 skipped), 96x72, masks 1" in r.stdout
    assert "init from random" in r.stdout
    evals = [float(x) for x in re.findall(r"eval: \d+ views psnr ([0-9.]+)", r.stdout)]
    assert len(evals) >= 4 and evals[-1] > evals[0] + 1.0 and evals[-1] > 20.0, evals
    assert re.search(r"trained 1600 views in 400 optimizer steps", r.stdout)
    head = out.read_bytes()[:2048].decode("latin-1")
    count = int(re.search(r"element vertex (\d+)", head).group(1))
    assert count > 100


def test_gs_train_refines_poses_and_exposure(tmp_path):
    """--refine-poses / --refine-exposure / --pose-out (cli/gs_train.cpp:107-124,
    284-302): frames 3 and 11 darkened as a camera's exposure would; the run
    reports pose and gain updates, aligns the held-out views before
    evaluating, and writes every frame's refined camera-to-world."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    print("[case] gs_train --refine-poses --refine-exposure --pose-out", flush=True)
    data = tmp_path / "scene"
    # opaque enough that each view has > 500 pixels for the gain fit (trainer.cpp:1702)
    write_dataset(str(data), 600, 1, 96, 72, 16, 62, dims={3: 0.8, 11: 0.85}, opacity=1.5)
    out, poses = tmp_path / "out.ply", tmp_path / "poses.json"
    cmd = [EXE, str(data), str(out), "--width", "96", "--iters", "800", "--sh", "1", "--init-gaussians", "2000",
           "--max-gaussians", "6000", "--views-per-step", "4", "--eval-every", "800", "--refine-poses",
           "--pose-from", "200", "--pose-lr", "5e-4,5e-4", "--refine-exposure", "--pose-out", str(poses)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    m = re.search(r"poses: (\d+) updates", r.stdout)
    assert m and int(m.group(1)) == (800 - 200) // 4 * 4  # every trained view from step 200
    m = re.search(r"exposure: (\d+) updates, brightness gains ([0-9.]+)\.\.([0-9.]+)", r.stdout)
    assert m and int(m.group(1)) > 0 and float(m.group(2)) < 0.97 < 1.0 <= float(m.group(3))
    assert re.search(r"eval: 2 views psnr [0-9.]+ ssim [0-9.]+ \(poses aligned; [0-9.]+ dB at the given poses\)",
                     r.stdout)
    doc = json.loads(poses.read_text())
    assert len(doc["frames"]) == 16
    assert sum(f["holdout"] for f in doc["frames"]) == 2
    given = json.loads((data / "transforms.json").read_text())["frames"]
    for f, g in zip(doc["frames"], given):
        M, G = np.array(f["transform_matrix"]), np.array(g["transform_matrix"])
        assert np.abs(M[:3, :3] @ M[:3, :3].T - np.eye(3)).max() < 1e-5
        assert np.abs(M[:3, 3] - G[:3, 3]).max() < 0.1  # corrections stay small on exact poses


def test_gs_train_with_the_stored_state_backward(tmp_path):
    """The same training with the stored-state backward (--stored-state 1,
    16 px bins, two views per step): the held-out PSNR rises as with the
    re-walk."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    print("[case] gs_train --stored-state 1 --bin-shift 0, 2 views per step", flush=True)
    data = tmp_path / "scene"
    write_dataset(str(data), 300, 1, 96, 72, 16, 63)
    out = tmp_path / "out.ply"
    cmd = [EXE, str(data), str(out), "--width", "72", "--iters", "1200", "--sh", "1", "--init-gaussians", "1000",
           "--max-gaussians", "6000", "--views-per-step", "2", "--stored-state", "1", "--bin-shift", "0",
           "--eval-every", "400"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    evals = [float(x) for x in re.findall(r"eval: \d+ views psnr ([0-9.]+)", r.stdout)]
    assert len(evals) >= 3 and evals[-1] > evals[0] + 1.0 and evals[-1] > 20.0, evals
    assert re.search(r"trained 1200 views in 600 optimizer steps", r.stdout)
