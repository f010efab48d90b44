"""The stored-state (bucket) backward through the C ABI: the forward stores
each pixel's (C, T) at every 64th position of its tile's walk and one wave
per 64-splat bucket sweeps the tile's pixels front to back, one splat per
lane (gs_hip_tuning.stored_state = 1).

* against the oracle at the single-view kernel-check size (200k gaussians,
  SH3, 1280x720): binning bit-exact, forward, loss and backward within the
  bars of parity_check.py, the adjoint against the fp64 adjoint;
* against the recompute backward on the same steps: the forward's image and
  LAST bit-equal (storing the state changes nothing it computes), the 2D
  gradients within 2e-5 of each value's largest (float-atomic order and the
  two walks' rounding), three training steps' loss sums within 1e-5, for 16,
  32 and 64 px pair bins, one and four views per step, graph replay and the
  persistent forward.  (Parameters after the steps are not compared value by
  value: Adam's first steps move a parameter by about its rate times the
  sign of its gradient, so a gradient that is rounding noise in both walks
  moves it by +-rate either way.)
* GS_HIP_FLAG_DETERMINISTIC steps take the re-walk: bitwise equal to the
  recompute backward's deterministic step.
The benchmark-size cases (1M gaussians in 8 views, the 16-scene batch) are
in test_gpu_headline.py.
"""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, orc, smooth_target
from parity_check import check_adjoint, check_view, first_step_gradients, frozen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    print(f"\n[gpu] {name}", flush=True)
    yield name
    gs.synchronize()


def case(msg):
    print(f"[case] {msg}", flush=True)


def test_stored_state_against_the_oracle_at_200k():
    n, deg, w, h = 200_000, 3, 1280, 720
    case(f"stored-state backward, {n} gaussians SH{deg}, {w}x{h}, one view")
    cam = gs.orbit_cameras(64, w, h)[3]
    gt_p, gt_sh = gs.synthetic_scene(n, deg, 1001)
    with gs.Workspace(n, deg, w, h) as gt:
        gt.load(gt_p, gt_sh)
        target = gt.render(gs.make_frame(cam, n, deg))
        target[..., 3] = 255
    params, sh = gs.synthetic_scene(n, deg, 2002)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.set_tuning(stored_state=1)
        ws.load(params, sh)
        ws.set_target(target)
        pairs, losses = ws.train_step(gs.make_frame(cam, n, deg, tp=frozen()))
        g2 = check_view(ws, 0, pairs, params, sh, cam, target, losses, False, "200k stored-state", deg=deg)
        gp, gsh = first_step_gradients(ws, n)
    check_adjoint(gp, gsh, params, sh, [cam], deg, [g2], "200k stored-state")


def run(n, deg, w, h, views, steps, seed, tuning, flags=0):
    """`steps` training steps from one seeded start; returns the first step's
    images, LAST and 2D gradients per view, the loss sums of every step and
    the parameters after the last."""
    params, sh = make_scene(n, deg, seed)
    cams = [make_camera(w, h, 0.3 + 0.9 * v) for v in range(views)]
    targets = [gs.DeviceBuffer(smooth_target(w, h, seed + v)) for v in range(views)]
    try:
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_views(views)
            ws.set_tuning(**tuning)
            ws.load(params, sh)
            first, losses = None, []
            for s in range(steps):
                frames = [gs.make_frame(c, n, deg, tp=gs.TrainParams(lambda_dssim=0.2), step=s + 1, flags=flags)
                          for c in cams]
                _, ls = ws.train_views(frames, [t.ptr for t in targets])
                losses.append(ls)
                if first is None:
                    first = [(ws.download(gs.BUF_IMAGE, np.float32, view=v),
                              ws.download(gs.BUF_LAST, np.uint32, view=v),
                              ws.download(gs.BUF_GRAD2D, np.float32, 9 * n, view=v).reshape(n, 9))
                             for v in range(views)]
                ws.download(gs.BUF_TOTALS, np.uint32, 16)  # eager work between (graph) replays
            after = ws.store(n)[0]
    finally:
        for t in targets:
            t.free()
    return first, np.asarray(losses), after


BASE = dict(raster_groups_per_cu=0, depth_digit_bits=8, tile_digit_bits=8, graph=0, bin_shift=0, stored_state=0)
CASES = [
    ("16px-1view", 1, dict(BASE)),
    ("16px-4views", 4, dict(BASE)),
    ("32px-bins-4views", 4, dict(BASE, bin_shift=1)),
    ("64px-bins-1view", 1, dict(BASE, bin_shift=2)),
    ("graph-4views", 4, dict(BASE, graph=1)),
    ("persistent-forward-4views", 4, dict(BASE, raster_groups_per_cu=2)),
]


@pytest.mark.parametrize("name,views,tuning", CASES, ids=[c[0] for c in CASES])
def test_stored_state_equals_recompute(name, views, tuning):
    n, deg, w, h = 3000, 3, 200, 152
    case(f"stored-state vs recompute backward: {name}, {n} gaussians, {w}x{h}, 3 steps")
    ref_first, ref_losses, _ = run(n, deg, w, h, views, 3, 91, dict(tuning, stored_state=0))
    first, losses, _ = run(n, deg, w, h, views, 3, 91, dict(tuning, stored_state=1))
    for v, ((ri, rl, rg), (i, l, g)) in enumerate(zip(ref_first, first)):
        np.testing.assert_array_equal(i.view(np.uint32), ri.view(np.uint32))
        np.testing.assert_array_equal(l, rl)
        scale = np.abs(rg).max(axis=0) + 1e-30
        err = (np.abs(g - rg) / scale).max()
        print(f"[info] {name} view {v}: 2D gradients within {err:.2e} of each value's largest", flush=True)
        assert err <= 2e-5, f"view {v}: {err:.3g}"
    np.testing.assert_array_equal(losses[0], ref_losses[0])  # the same forward
    np.testing.assert_allclose(losses, ref_losses, rtol=1e-5)


def test_deterministic_steps_take_the_rewalk():
    n, deg, w, h = 2000, 2, 160, 120
    case("GS_HIP_FLAG_DETERMINISTIC with stored_state = 1: the recompute walk, bit for bit")
    a_first, _, a_after = run(n, deg, w, h, 2, 2, 17, dict(BASE, stored_state=0), flags=gs.FLAG_DETERMINISTIC)
    b_first, _, b_after = run(n, deg, w, h, 2, 2, 17, dict(BASE, stored_state=1), flags=gs.FLAG_DETERMINISTIC)
    for (ai, al, ag), (bi, bl, bg) in zip(a_first, b_first):
        np.testing.assert_array_equal(ag.view(np.uint32), bg.view(np.uint32))
    np.testing.assert_array_equal(a_after.view(np.uint32), b_after.view(np.uint32))


def test_stored_state_overflow_retry_matches_a_step_with_room():
    """A step whose pairs overflow the buffers runs again with larger pair
    buffers and with bucket states grown to match (the forward of the first
    attempt wrote states for a smaller reserve): the result is that of a step
    that had room from the start."""
    case("stored-state backward, pair overflow: the first step outgrows its buffers")
    n, deg, w, h = 6000, 0, 512, 384
    params, _ = make_scene(n, deg, 73)
    params = params.copy()
    params[:, 4:7] += 1.5  # large splats: many tiles each
    cam = make_camera(w, h, 1.2)
    target = gs.DeviceBuffer(smooth_target(w, h, 73))
    results = []
    try:
        for reserve in (None, 4_000_000):
            with gs.Workspace(n, deg, w, h) as ws:
                ws.set_tuning(stored_state=1)
                if reserve:
                    ws.reserve_pairs(reserve)
                ws.load(params, None)
                frame = gs.make_frame(cam, n, deg, tp=gs.TrainParams(lambda_dssim=0.2), step=1)
                pairs, losses = ws.train_views([frame], [target.ptr])
                results.append((pairs[0], losses[0], ws.overflows(),
                                ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9)))
    finally:
        target.free()
    (p0, l0, o0, g0), (p1, l1, o1, g1) = results
    print(f"[info] pairs {p0}, overflow steps {o0} / {o1}", flush=True)
    assert p0 == p1 and o0 >= 1 and o1 == 0
    np.testing.assert_allclose(l0, l1, rtol=1e-6)
    scale = np.abs(g1).max(axis=0) + 1e-30
    assert (np.abs(g0 - g1) / scale).max() <= 2e-5
