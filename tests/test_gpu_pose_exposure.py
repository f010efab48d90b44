"""Camera pose gradient (GS_HIP_FLAG_POSE_GRAD) and exposure (GS_HIP_FLAG_EXPOSURE)
through the C ABI against the oracle.

Pose: the per-256-slot partials summed in double equal the fp64 pose adjoint
of the GPU's own GRAD2D within 1e-4 of the summed magnitudes, and satisfy the
reference's identity with the per-gaussian gradients (test_trainer.cpp:
404-464) within 1e-3 of the largest component; a pose-only step
(NO_UPDATE | POSE_GRAD) gives the same gradient and leaves the model alone;
in a batched step each flagged view gets its own view's gradient.
Exposure: the loss and dL/dC with the photo scaled by exposure[] within 1e-4
of the fp64 loss of the GPU's own render; the gain fit sums within 1e-5 of
the oracle's sums of the same render.
"""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, orc, smooth_target

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    print(f"\n[gpu] {name}", flush=True)
    yield name
    gs.synchronize()


def frozen(lam=0.2):
    return gs.TrainParams(lambda_dssim=lam, lr_position=0.0, lr_scale=0.0, lr_rotation=0.0, lr_opacity=0.0,
                          lr_sh_dc=0.0, lr_sh_rest=0.0)


def case(msg):
    print(f"[case] {msg}", flush=True)


def masked_target(w, h, seed):
    """A smooth photo whose alpha covers opaque (255), edge (200, 243) and empty (0) pixels."""
    t = smooth_target(w, h, seed)
    rng = np.random.default_rng(seed)
    t[..., 3] = rng.choice(np.array([255, 255, 255, 243, 200, 0], np.uint8), size=(h, w))
    return t


@pytest.mark.parametrize("deg,lam", [(0, 0.2), (3, 0.0), (1, 1.0)], ids=["sh0", "sh3-l1", "sh1-dssim"])
def test_pose_gradient_matches_oracle(deg, lam):
    case(f"pose gradient, SH {deg}, lambda {lam}")
    w, h, n = 96, 72, 500
    params, sh = make_scene(n, deg, 61 + deg)
    cam = make_camera(w, h, 1.3)
    target = smooth_target(w, h, 61)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, n, deg, background=(0.2, 0.3, 0.4), flags=gs.FLAG_POSE_GRAD, tp=frozen(lam))
        ws.train_step(frame)
        got = ws.pose_gradient()
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9).astype(np.float64)
        m = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16) / (1 - np.float32(0.9))
        # a pose-only step: the same gradient, the model and moments left alone
        p_before = ws.store(n)[0]
        m_before = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16)
        frame.flags |= gs.FLAG_NO_UPDATE
        ws.train_step(frame)
        again = ws.pose_gradient()
        np.testing.assert_array_equal(ws.store(n)[0], p_before)
        np.testing.assert_array_equal(ws.download(gs.BUF_M_PARAMS, np.float32, n * 16), m_before)
    per = orc.pose_adjoint64(params.astype(np.float64), sh.astype(np.float64), cam, deg, g2)
    want = per.sum(0)
    scale = np.abs(per).sum(0)
    print(f"[info] pose gradient {np.array2string(got, precision=5)} oracle {np.array2string(want, precision=5)}",
          flush=True)
    assert np.all(np.abs(got - want) <= 1e-4 * scale + 1e-12), (got, want, scale)
    assert np.abs(want).max() > 1e-4
    assert np.all(np.abs(again - got) <= 1e-5 * scale + 1e-12)  # GRAD2D's float-atomic order may differ
    if deg == 0:
        # the reference's identity: W sum dL/dp and W sum((p - c) x dL/dp + tau)
        R = np.array(cam.rotation, np.float64).reshape(3, 3)
        c = np.array(cam.position, np.float64)
        p = params.astype(np.float64)
        q = p[:, 8:12]
        ln = np.linalg.norm(q, axis=1, keepdims=True)
        qw, u = q[:, :1] / ln, q[:, 1:] / ln
        gw, gu = m[:, 8:9] * ln, m[:, 9:12] * ln
        tau = 0.5 * (-gw * u + qw * gu + np.cross(u, gu))
        ident = np.concatenate([R @ m[:, 0:3].sum(0), R @ (np.cross(p[:, 0:3] - c, m[:, 0:3]) + tau).sum(0)])
        assert np.abs(got - ident).max() <= 1e-3 * np.abs(ident).max(), (got, ident)


def test_pose_gradient_per_view_in_a_batched_step():
    case("pose gradient per view in a 3-view step (views 0 and 2 flagged)")
    w, h, n, deg = 80, 64, 600, 1
    params, sh = make_scene(n, deg, 71)
    cams = [make_camera(w, h, a) for a in (0.4, 1.6, 2.8)]
    targets = [smooth_target(w, h, 71 + v) for v in range(3)]
    bufs = [gs.DeviceBuffer(t) for t in targets]
    try:
        single = []
        for v in (0, 2):
            with gs.Workspace(n, deg, w, h) as ws:
                ws.load(params, sh)
                ws.set_target(targets[v])
                ws.train_step(gs.make_frame(cams[v], n, deg, flags=gs.FLAG_POSE_GRAD, tp=frozen()))
                single.append(ws.pose_gradient())
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_views(3)
            ws.load(params, sh)
            flags = [gs.FLAG_POSE_GRAD, 0, gs.FLAG_POSE_GRAD]
            frames = [gs.make_frame(c, n, deg, flags=f, tp=frozen()) for c, f in zip(cams, flags)]
            ws.train_views(frames, [b.ptr for b in bufs])
            batched = [ws.pose_gradient(0), ws.pose_gradient(2)]
    finally:
        for b in bufs:
            b.free()
    for s, b in zip(single, batched):
        print(f"[info] single {np.array2string(s, precision=5)} batched {np.array2string(b, precision=5)}", flush=True)
        assert np.abs(s - b).max() <= 1e-4 * np.abs(s).max()


@pytest.mark.parametrize("masked", [False, True], ids=["plain", "masked"])
def test_exposure_loss_and_fit_sums(masked):
    case(f"exposure: loss, dL/dC and gain fit sums (masked {masked})")
    w, h, n, deg = 96, 72, 1500, 1
    params, sh = make_scene(n, deg, 81)
    params[:, 3] += 2.0  # mostly opaque, so the fit has pixels
    cam = make_camera(w, h, 0.9)
    target = masked_target(w, h, 81)
    exposure = (0.8, 1.1, 1.25)
    bg = (0.25, 0.5, 0.125)
    lam = 0.2
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        flags = gs.FLAG_EXPOSURE | (gs.FLAG_MASKED_TARGET if masked else 0)
        frame = gs.make_frame(cam, n, deg, background=bg, flags=flags, tp=frozen(lam))
        frame.exposure[:] = list(exposure)
        _, losses = ws.train_step(frame)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        dldc = ws.download(gs.BUF_DLDC, np.float32).reshape(-1, 4)
        sums = ws.exposure_sums()
        tiles = ws.tiles_x * ws.tiles_y
        fit = ws.download(gs.BUF_TILE_FIT, np.float32, tiles * 8).reshape(tiles, 8)
    x = img[:, :3].astype(np.float64) + img[:, 3:4].astype(np.float64) * np.asarray(bg)
    y = orc.target_colour64(target, bg, masked, exposure)
    L_ref, d_ref = orc.loss_image64(x, y, w, h, lam, grad=True)
    e = x - y
    assert losses[0] == pytest.approx(np.abs(e).sum(), rel=1e-4)
    assert losses[1] == pytest.approx((e * e).sum(), rel=1e-4)
    sign_ok = np.abs(e) > 1e-6
    assert np.abs(dldc[:, :3] - d_ref)[sign_ok].max() <= 1e-4 * np.abs(d_ref).max()
    want = orc.exposure_sums64(img, target)
    print(f"[info] fit sums {np.array2string(sums, precision=4)} oracle {np.array2string(want, precision=4)}",
          flush=True)
    assert want[3] > 500 and want[7] > want[3]
    np.testing.assert_array_equal(sums[[3, 7]], want[[3, 7]])
    np.testing.assert_allclose(sums, want, rtol=1e-5)
    np.testing.assert_allclose(fit.astype(np.float64).sum(0), sums, rtol=1e-6)


def test_pose_gradient_in_a_batch_of_scenes():
    """Pose gradients of views in a batch of scenes (gs_hip_batch_train) equal
    each scene's own step's, read from the scene's workspace."""
    case("pose gradient in a batch of 2 scenes (SH 2 120x90 x2, SH 0 64x48 x1)")
    specs = [(900, 2, 120, 90, 2, 91), (400, 0, 64, 48, 1, 92)]
    inputs = []
    for n, deg, w, h, views, seed in specs:
        params, sh = make_scene(n, deg, seed)
        cams = [make_camera(w, h, 0.5 + 1.1 * v + 0.1 * seed) for v in range(views)]
        targets = [smooth_target(w, h, seed + v) for v in range(views)]
        inputs.append((params, sh, cams, targets))
    results = []
    for batched in (False, True):
        spaces, bufs = [], []
        try:
            for (n, deg, w, h, views, _), (params, sh, cams, targets) in zip(specs, inputs):
                ws = gs.Workspace(n, deg, w, h)
                ws.set_views(views)
                ws.load(params, sh)
                spaces.append(ws)
                bufs.append([gs.DeviceBuffer(t) for t in targets])
            fps = [[gs.make_frame(c, s[0], s[1], flags=gs.FLAG_POSE_GRAD, tp=frozen()) for c in inp[2]]
                   for s, inp in zip(specs, inputs)]
            if batched:
                with gs.Batch(spaces) as batch:
                    batch.train(fps, [[b.ptr for b in bs] for bs in bufs])
            else:
                for ws, fs, bs in zip(spaces, fps, bufs):
                    ws.train_views(fs, [b.ptr for b in bs])
            results.append([[ws.pose_gradient(v) for v in range(s[4])] for ws, s in zip(spaces, specs)])
        finally:
            for ws in spaces:
                ws.close()
            for bs in bufs:
                for b in bs:
                    b.free()
    for own, bat in zip(*results):
        for a, b in zip(own, bat):
            print(f"[info] own {np.array2string(a, precision=5)} batch {np.array2string(b, precision=5)}", flush=True)
            assert np.abs(a).max() > 0 and np.abs(a - b).max() <= 1e-4 * np.abs(a).max()
