"""The trainer handle through ctypes (gs_hip_trainer_*, the reference's
NativeTrainer: prepare / step / preview / evaluate / exportPly / release,
app/src/main/cpp/jni_bridge.cpp:815-1108) on a transforms.json folder of
PNG frames and subject masks rendered by the oracle's fp64 renderer."""
import re
import struct
import zlib

import numpy as np
import pytest

from helpers import gs
from test_gs_train_cli import write_dataset

pytestmark = pytest.mark.gpu


def test_trainer_handle_trains_previews_evaluates_and_exports(tmp_path):
    print("[case] trainer handle: 16 views 96x72, 1600 views at 4 per step", flush=True)
    data = tmp_path / "scene"
    write_dataset(str(data), 300, 1, 96, 72, 16, 63)
    t = gs.NativeTrainer(str(data), width=72, iterations=1600, max_gaussians=6000, sh_degree=1, masked=True,
                         views_per_step=4, seed=3)
    info = t.info
    assert (info.frames, info.holdout, info.skipped, info.width, info.height) == (16, 2, 0, 96, 72)
    assert info.masks == 1 and info.depth == 0 and info.init_source == b"random"
    assert 0 < info.init_gaussians <= 3000  # half the slots at most
    before = t.evaluate()
    assert before.views == 2 and not before.aligned

    done, st, times = t.step(8, profile=True)
    assert done == 8 and st.step == 8 * 4 and st.optimizer_steps == 8 and not st.finished
    assert times is not None and times.rasterize > 0 and times.backward > 0
    total = done
    while not t.stats().finished:
        n, st, _ = t.step(100)
        assert n > 0
        total += n
    assert total == 400 and st.step == 1600 and st.iterations == 1600
    assert st.gaussians > 100 and st.pairs > 0 and np.isfinite(st.loss)
    after = t.evaluate()
    print(f"held-out psnr {before.psnr:.2f} -> {after.psnr:.2f}", flush=True)
    assert after.psnr > before.psnr + 1.0 and after.psnr > 20.0

    # preview: the render, then the photo over the same background through its mask
    render, photo = t.preview(1)  # the second held-out view: frame 8
    assert render.shape == photo.shape == (72, 96, 4)
    assert (photo[..., 3] == 255).all()
    img = read_png(data / "images" / "frame_008.png")
    mask = read_png(data / "masks" / "frame_008.png").astype(np.float32) / 255.0
    bg = np.array([0.08, 0.09, 0.11], np.float32)
    want = img.astype(np.float32) * mask[..., None] + bg * 255.0 * (1.0 - mask[..., None]) + 0.5
    assert np.abs(photo[..., :3].astype(np.int32) - np.floor(want).astype(np.int32)).max() <= 1
    # the render covers the subject where the mask does, and is the background elsewhere
    covered = mask > 0.9
    empty = mask < 1e-3
    assert covered.sum() > 200 and empty.sum() > 200
    assert np.abs(render[covered][:, :3].astype(np.int32) - photo[covered][:, :3].astype(np.int32)).mean() < 20
    assert np.abs(render[empty][:, :3].astype(np.int32) - np.rint(bg * 255).astype(np.int32)).mean() < 8

    # the PLY: the header's count is the alive gaussians, the file size their rows
    path = tmp_path / "out.ply"
    count, nbytes = t.export_ply(str(path))
    blob = path.read_bytes()
    assert len(blob) == nbytes and count == t.stats().gaussians
    head = blob[:blob.index(b"end_header\n") + len(b"end_header\n")].decode()
    assert int(re.search(r"element vertex (\d+)", head).group(1)) == count
    assert len(blob) - len(head) == count * (17 + 3 * 3) * 4  # SH degree 1: 3 rest coefficients per channel
    t.release()
    t.release()  # a second release is a no-op


def read_png(path):
    """The 8-bit grey or RGB PNGs write_dataset writes (one IDAT, filter 0 on every row)."""
    blob = open(path, "rb").read()
    w, h = struct.unpack(">II", blob[16:24])
    colour = blob[25]
    at, idat = 8, b""
    while at < len(blob):
        n = struct.unpack(">I", blob[at:at + 4])[0]
        if blob[at + 4:at + 8] == b"IDAT":
            idat += blob[at + 8:at + 8 + n]
        at += 12 + n
    ch = 3 if colour == 2 else 1
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * ch)
    assert (rows[:, 0] == 0).all()
    img = rows[:, 1:].reshape(h, w, ch)
    return img if ch == 3 else img[..., 0]
