"""Coarse pair bins (gs_hip_tuning.bin_shift 1, 2: pairs binned by 32x32 or
64x64 pixel squares, each 16x16 raster tile walking its bin's list) through
the C ABI.

* Binning is bit-exact with the oracle's at that bin size: every slot's bin
  count, the sorted pairs and the bins' ranges (orc_bin32 with `side`).
* A bin lists, in depth order, every splat of each of its tiles' own lists
  (tests/test_oracle.py pins that on the CPU), and a splat the tile does not
  reach draws nothing at its pixels, so a tile composites the same splats in
  the same order as with bin_shift 0: the image, the loss sums and -- with
  GS_HIP_FLAG_DETERMINISTIC -- every 2D gradient are bitwise those of the
  reference's 16x16 binning, and so are whole training steps."""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, orc, smooth_target

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    print(f"\n[gpu] {gs.init(0)}", flush=True)
    yield
    gs.synchronize()


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("shift", [1, 2])
@pytest.mark.parametrize("depth16", [False, True], ids=["float-key", "depth16"])
def test_bins_are_bit_exact_with_the_oracle(shift, depth16):
    n, deg, w, h = 20000, 2, 640, 360
    case(f"bins of {16 << shift} px: {n} gaussians SH{deg} {w}x{h}, depth16 {depth16}")
    params, sh = make_scene(n, deg, 81)
    params[::5, 4:7] += 1.0  # some large splats: pairs of many bins
    cam = make_camera(w, h, 0.9)
    side = 16 << shift
    rec, foot, tiles, keys = orc.project32(params, sh, cam, deg, depth16=depth16)
    counts = orc.bin_counts32(rec, foot, tiles, w, h, side)
    o_pt, o_pg, o_ranges = orc.bin32(rec, foot, counts, keys, w, h, side)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.set_tuning(bin_shift=shift)
        ws.load(params, sh)
        frame = gs.make_frame(cam, n, deg, flags=gs.FLAG_DEPTH16 if depth16 else 0)
        ws.project(frame)
        pairs = ws.bin(frame)
        g_counts = ws.download(gs.BUF_TILES, np.uint32, n)
        pt = ws.download(gs.BUF_PAIR_TILES, np.uint32, pairs)
        pg = ws.download(gs.BUF_PAIR_GAUSSIANS, np.uint32, pairs)
        B = len(o_ranges)
        ranges = ws.download(gs.BUF_RANGES, np.uint32, 2 * B).reshape(B, 2)
    print(f"[info] {pairs} pairs in {B - 1} bins (oracle 16 px tiles: {int(tiles.sum())} pairs)", flush=True)
    np.testing.assert_array_equal(g_counts, counts)
    assert pairs == len(o_pt) and pairs < int(tiles.sum())
    np.testing.assert_array_equal(pt, o_pt)
    np.testing.assert_array_equal(pg, o_pg)
    np.testing.assert_array_equal(ranges, o_ranges)


def step_state(params, sh, cam, target, n, deg, w, h, tuning, flags):
    with gs.Workspace(n, deg, w, h) as ws:
        ws.set_tuning(**tuning)
        ws.load(params, sh)
        ws.set_target(target)
        _, losses = ws.train_step(gs.make_frame(cam, n, deg, flags=gs.FLAG_NO_UPDATE | flags))
        return (ws.download(gs.BUF_IMAGE, np.float32), ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9),
                np.array(losses))


@pytest.mark.parametrize("shift", [1, 2])
def test_bins_draw_the_same_image_and_gradients(shift):
    n, deg, w, h = 200_000, 3, 1280, 720
    case(f"bins of {16 << shift} px against 16 px tiles: {n} gaussians SH{deg} {w}x{h}")
    params, sh = gs.synthetic_scene(n, deg, 83)
    cam = gs.orbit_cameras(1, w, h)[0]
    target = smooth_target(w, h, 84)
    base = dict(raster_groups_per_cu=0, graph=0, bin_shift=0)
    ref = step_state(params, sh, cam, target, n, deg, w, h, base, gs.FLAG_DETERMINISTIC)
    for tuning in (dict(base, bin_shift=shift), dict(raster_groups_per_cu=4, graph=1, bin_shift=shift)):
        got = step_state(params, sh, cam, target, n, deg, w, h, tuning, gs.FLAG_DETERMINISTIC)
        np.testing.assert_array_equal(got[0].view(np.uint32), ref[0].view(np.uint32))
        np.testing.assert_array_equal(got[2], ref[2])
        np.testing.assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
    atomic = step_state(params, sh, cam, target, n, deg, w, h, dict(base, bin_shift=shift), 0)
    np.testing.assert_array_equal(atomic[0].view(np.uint32), ref[0].view(np.uint32))
    err = (np.abs(atomic[1] - ref[1]) / (np.abs(ref[1]).max(axis=0) + 1e-30)).max()
    print(f"[info] float-atomic sums with bins against the deterministic ones: {err:.2e} of the plane maxima",
          flush=True)
    assert err <= 1e-5


def test_batched_training_with_bins_equals_16px_binning():
    n, deg, w, h, views = 20000, 2, 320, 240, 4
    case(f"4 deterministic optimizer steps of {views} views: 32 px bins, persistent, graph vs 16 px tiles")
    params, sh = make_scene(n, deg, 85)
    cams = [make_camera(w, h, 0.2 + 1.1 * v) for v in range(views)]
    bufs = [gs.DeviceBuffer(smooth_target(w, h, 86 + v)) for v in range(views)]
    out = []
    try:
        for tuning in (dict(raster_groups_per_cu=0, graph=0, bin_shift=0),
                       dict(raster_groups_per_cu=4, graph=1, bin_shift=1)):
            with gs.Workspace(n, deg, w, h) as ws:
                ws.set_views(views)
                ws.set_tuning(**tuning)
                ws.load(params, sh)
                losses = []
                for s in range(4):
                    frames = [gs.make_frame(c, n, deg, step=s, flags=gs.FLAG_DETERMINISTIC) for c in cams]
                    _, ls = ws.train_views(frames, [b.ptr for b in bufs])
                    losses.append(ls)
                    ws.download(gs.BUF_TOTALS, np.uint32, 16)  # eager work between graph replays
                out.append((ws.store(n), np.array(losses)))
    finally:
        for b in bufs:
            b.free()
    (p0, s0), l0 = out[0]
    (p1, s1), l1 = out[1]
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(p0.view(np.uint32), p1.view(np.uint32))
    np.testing.assert_array_equal(s0.view(np.uint32), s1.view(np.uint32))
