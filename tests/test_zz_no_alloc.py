"""gs_hip_loss_noalloc (the loss stage on caller-owned buffers, its scratch
in the library's file-scope device array) against gs_hip_loss on the same
image and target: dL/dC bit for bit and the loss sums equal, on two calls in
a row, the second smaller than the first (the array keeps the first call's
contents), with a masked target and a background; and a size beyond the
array's limit is refused with GS_HIP_E_CAPACITY before anything runs."""
import numpy as np
import pytest

from helpers import gs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    print(f"\n[gpu] {gs.init(0)}", flush=True)
    yield
    gs.synchronize()


def case_inputs(w, h, seed):
    rng = np.random.default_rng(seed)
    image = np.zeros((h * w, 4), np.float32)
    image[:, :3] = rng.uniform(0, 1, (h * w, 3))
    image[:, 3] = rng.uniform(0, 1, h * w)
    target = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    return image, target


@pytest.mark.parametrize("masked", [False, True])
def test_loss_noalloc_equals_the_workspace_loss(masked):
    flags = gs.FLAG_MASKED_TARGET if masked else 0
    for (w, h, seed) in ((333, 250, 1), (97, 61, 2)):  # the second call smaller than the first
        print(f"[case] loss_noalloc {w}x{h} masked={masked}", flush=True)
        image, target = case_inputs(w, h, seed + 10 * masked)
        cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
        frame = gs.make_frame(cam, 1, 0, background=(0.2, 0.1, 0.3), flags=flags,
                              tp=gs.TrainParams(lambda_dssim=0.3))
        with gs.Workspace(1, 0, w, h) as ws:
            ws.upload(gs.BUF_IMAGE, image)
            ws.set_target(target)
            want = ws.loss(frame)
            want_dldc = ws.download(gs.BUF_DLDC, np.float32).reshape(-1, 4)
            img_ptr, _ = ws.buffer(gs.BUF_IMAGE)
            tgt_ptr, _ = ws.buffer(gs.BUF_TARGET)
            out = gs.DeviceBuffer(np.full((h * w, 4), np.nan, np.float32))
            try:
                got = gs.loss_noalloc(frame, img_ptr, tgt_ptr, out.ptr)
                got_dldc = np.empty((h * w, 4), np.float32)
                gs.library().gs_hip_download(got_dldc.ctypes.data, out.ptr, got_dldc.nbytes)
            finally:
                out.free()
        np.testing.assert_array_equal(got_dldc[:, :3], want_dldc[:, :3])
        assert got == want, (got, want)
        print(f"[ok] {w}x{h}: losses {got}", flush=True)


def test_loss_noalloc_refuses_sizes_beyond_its_scratch():
    w, h = 2048, 1100  # 2.25M pixels > 2^21
    cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
    frame = gs.make_frame(cam, 1, 0)
    bufs = [gs.DeviceBuffer(np.zeros(4, np.float32)) for _ in range(3)]  # refused before any launch
    try:
        with pytest.raises(gs.GsHipError) as e:
            gs.loss_noalloc(frame, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr)
    finally:
        for b in bufs:
            b.free()
    assert e.value.code == 3 and "scratch limit" in str(e.value)
