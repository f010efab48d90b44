"""The viewer renderer (gs_hip_view_render) against the oracle's per-pixel
restatement, through the C ABI: the same projection arithmetic with
contraction off, the same exp, fp16 colour and opacity, the same drawing
order (16-bit key, ties by index) and 8-bit rounding at every blend, so the
RGBA8 images and the visible counts are bit-exact.  SH rest goes in rounded
to fp16 on both sides, as the viewer stores it."""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    print(f"\n[gpu] {name}", flush=True)
    yield name
    gs.synchronize()


def fp16(a):
    return a.astype(np.float16).astype(np.float32)


def view_pair(params, sh, deg, cam, clear, antialiased, reserve=None):
    n = params.shape[0]
    sh16 = fp16(sh) if sh is not None and sh.size else None
    with gs.Workspace(n, deg, cam.width, cam.height) as ws:
        if reserve is not None:
            ws.reserve_pairs(reserve)
        ws.load(params, sh16)
        img, vis = ws.view_render(gs.make_frame(cam, n, deg), clear, antialiased)
        again, vis2 = ws.view_render(gs.make_frame(cam, n, deg), clear, antialiased)
        overflows = ws.overflows()
    np.testing.assert_array_equal(img, again)  # a second render of the same view: same image
    assert vis == vis2
    want, want_vis = orc.view_render32(params, sh16, cam, deg, antialiased, clear)
    return img, vis, want, want_vis, overflows


VIEW_CASES = [
    # name, n, degree, w, h, antialiased, clear, seed, log-scale offset
    ("sh0-300-64x48", 300, 0, 64, 48, False, (0.0, 0.0, 0.0, 1.0), 91, 0.0),
    ("sh3-2000-160x120-aa", 2000, 3, 160, 120, True, (0.08, 0.09, 0.11, 1.0), 92, 0.0),
    ("sh1-1500-100x75-ragged", 1500, 1, 100, 75, False, (0.2, 0.5, 0.9, 0.0), 93, -0.5),
    ("sh2-800-200x150-large", 800, 2, 200, 150, True, (0.0, 0.0, 0.0, 1.0), 94, 1.2),
]


@pytest.mark.parametrize("name,n,deg,w,h,aa,clear,seed,grow", VIEW_CASES, ids=[c[0] for c in VIEW_CASES])
def test_view_render_matches_oracle_bit_exact(name, n, deg, w, h, aa, clear, seed, grow):
    print(f"[case] viewer {name}", flush=True)
    params, sh = make_scene(n, deg, seed)
    params[:, 4:7] += grow
    cam = make_camera(w, h, 0.5 + seed * 0.1)
    img, vis, want, want_vis, _ = view_pair(params, sh, deg, cam, clear, aa)
    lit = int((img[..., :3].astype(int).sum(-1) > 30).sum())
    diff = np.abs(img.astype(int) - want.astype(int))
    print(f"[info] {name}: visible {vis} / {want_vis}, lit {lit}, max diff {diff.max()}, "
          f"differing values {(diff > 0).sum()}", flush=True)
    assert vis == want_vis and vis > 0
    assert lit > w * h // 10
    np.testing.assert_array_equal(img, want)


def test_view_render_pair_overflow_retry():
    """Large quads over a small initial pair buffer: the render grows the
    buffers, runs again and gives the image a render with room gives."""
    print("[case] viewer pair overflow", flush=True)
    n, deg, w, h = 1200, 0, 256, 192
    params, _ = make_scene(n, deg, 95)
    params[:, 4:7] += 1.8
    cam = make_camera(w, h, 0.8)
    img, vis, want, want_vis, overflows = view_pair(params, None, deg, cam, (0.0, 0.0, 0.0, 1.0), False)
    print(f"[info] visible {vis}, overflow retries {overflows}", flush=True)
    assert overflows >= 1
    assert vis == want_vis
    np.testing.assert_array_equal(img, want)


def test_view_render_nothing_visible_gives_the_clear_colour():
    print("[case] viewer empty view", flush=True)
    n, deg, w, h = 64, 0, 48, 40
    params, _ = make_scene(n, deg, 96)
    params[:, 7] = 0.0  # every slot dead
    cam = make_camera(w, h, 0.0)
    img, vis, want, want_vis, _ = view_pair(params, None, deg, cam, (0.08, 0.09, 0.11, 1.0), False)
    assert vis == 0 and want_vis == 0
    np.testing.assert_array_equal(img.reshape(-1, 4), np.tile(np.uint8([20, 23, 28, 255]), (w * h, 1)))
