"""GS_HIP_FLAG_DETERMINISTIC through the C ABI: the backward's 2D gradient
sums do not depend on the order in which tiles and waves run, so a view's
GRAD2D is bitwise the same whatever the launch shape, on every run, and as
one view of a batched step as on its own — the reference's requirement for
its batched steps against sequential ones (max difference 0.0,
engine/tests/test_trainer.cpp:784-812), whose pair-sum pass has no float
atomics (engine/shaders/train_pair_sum.comp:17-32).  The deterministic sums
stay within 1e-5 of each plane's largest value of the default (float
atomic) backward, which the parity tests hold to the oracle."""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, smooth_target

pytestmark = pytest.mark.gpu

N, DEG, W, H = 20000, 2, 320, 240
SHAPES = [dict(raster_groups_per_cu=0, graph=0), dict(raster_groups_per_cu=4, graph=0),
          dict(raster_groups_per_cu=2, graph=1)]


@pytest.fixture(scope="module", autouse=True)
def device():
    print(f"\n[gpu] {gs.init(0)}", flush=True)
    yield
    gs.synchronize()


def case(msg):
    print(f"[case] {msg}", flush=True)


def single_view_grad2d(params, sh, cam, target, flags, tuning):
    with gs.Workspace(N, DEG, W, H) as ws:
        ws.set_tuning(**tuning)
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, N, DEG, flags=gs.FLAG_NO_UPDATE | flags)
        ws.train_step(frame)
        ws.train_step(frame)  # the second step starts from non-zero sums of the first in no buffer
        return ws.download(gs.BUF_GRAD2D, np.float32, 9 * N).reshape(N, 9)


def test_single_view_gradients_are_bitwise_repeatable_across_launch_shapes():
    case(f"deterministic backward: {N} gaussians SH{DEG} {W}x{H}, 3 launch shapes x 2 runs")
    params, sh = make_scene(N, DEG, 91)
    cam = make_camera(W, H, 0.7)
    target = smooth_target(W, H, 92)
    runs = [single_view_grad2d(params, sh, cam, target, gs.FLAG_DETERMINISTIC, t) for t in SHAPES for _ in range(2)]
    for r in runs[1:]:
        np.testing.assert_array_equal(runs[0].view(np.uint32), r.view(np.uint32))
    atomic = single_view_grad2d(params, sh, cam, target, 0, SHAPES[0])
    scale = np.abs(atomic).max(axis=0) + 1e-30
    err = (np.abs(runs[0] - atomic) / scale).max()
    print(f"deterministic vs float-atomic sums: {err:.2e} of the plane maxima; "
          f"{int((np.abs(runs[0]).sum(axis=1) > 0).sum())} gaussians with gradients", flush=True)
    assert err <= 1e-5
    assert (np.abs(runs[0]).sum(axis=1) > 0).sum() > N // 4


def test_batched_view_gradients_equal_single_view_gradients_bitwise():
    views = 8
    case(f"deterministic backward: {views}-view batched step vs each view alone, {N} gaussians {W}x{H}")
    params, sh = make_scene(N, DEG, 93)
    cams = [make_camera(W, H, 0.3 + 0.8 * v) for v in range(views)]
    targets = [smooth_target(W, H, 94 + v) for v in range(views)]
    bufs = [gs.DeviceBuffer(t) for t in targets]
    try:
        with gs.Workspace(N, DEG, W, H) as ws:
            ws.set_views(views)
            ws.set_tuning(raster_groups_per_cu=4)
            ws.load(params, sh)
            frames = [gs.make_frame(c, N, DEG, flags=gs.FLAG_NO_UPDATE | gs.FLAG_DETERMINISTIC) for c in cams]
            ws.train_views(frames, [b.ptr for b in bufs])
            batched = [ws.download(gs.BUF_GRAD2D, np.float32, 9 * N, view=v).reshape(N, 9) for v in range(views)]
    finally:
        for b in bufs:
            b.free()
    for v in range(views):
        alone = single_view_grad2d(params, sh, cams[v], targets[v], gs.FLAG_DETERMINISTIC, SHAPES[0])
        np.testing.assert_array_equal(batched[v].view(np.uint32), alone.view(np.uint32), err_msg=f"view {v}")


@pytest.mark.parametrize("views", [1, 4])
def test_training_steps_are_bitwise_repeatable(views):
    case(f"deterministic training: 4 optimizer steps of {views} view(s), two launch shapes")
    params, sh = make_scene(N, DEG, 95)
    cams = [make_camera(W, H, 0.5 + 1.3 * v) for v in range(views)]
    bufs = [gs.DeviceBuffer(smooth_target(W, H, 96 + v)) for v in range(views)]
    out = []
    try:
        for tuning in SHAPES[:2]:
            with gs.Workspace(N, DEG, W, H) as ws:
                ws.set_views(views)
                ws.set_tuning(**tuning)
                ws.load(params, sh)
                for s in range(4):
                    frames = [gs.make_frame(c, N, DEG, step=s, flags=gs.FLAG_DETERMINISTIC) for c in cams]
                    ws.train_views(frames, [b.ptr for b in bufs])
                out.append(ws.store(N))
    finally:
        for b in bufs:
            b.free()
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
    assert np.abs(out[0][0] - params).max() > 0.0  # the steps moved the model
