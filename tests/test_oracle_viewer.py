"""Pins the oracle's viewer restatement (orc_half, orc_view_render32)
against what fixes it independently: numpy's IEEE fp16 conversion, the
closed forms of the reference's viewer tests (engine/tests/test_splats.cpp:
188-278: single-splat falloff, depth order, Y-flip equivalence with SH
parity signs), and a float64 evaluation of the quad contract from the
oracle's fp64 projection (conic of the dilated covariance, eigen-axes from
numpy).  No GPU."""
import math

import numpy as np

from helpers import gs, make_scene, orc

SH_C0 = 0.28209479177387814


def test_half_matches_numpy_fp16_exhaustively():
    bits = np.arange(0, 0x10000, dtype=np.uint32).astype(np.uint16)
    vals = bits.view(np.float16)
    finite = np.isfinite(vals)
    v32 = vals[finite].astype(np.float32)
    np.testing.assert_array_equal(orc.half(v32).view(np.uint32), v32.view(np.uint32))
    # midpoints between neighbours (exact in fp32) and one fp32 ulp either side: ties to even
    pos = np.sort(v32[v32 > 0])
    mid = ((pos[:-1].astype(np.float64) + pos[1:]) / 2).astype(np.float32)
    probes = np.concatenate([mid, np.nextafter(mid, np.float32(0)), np.nextafter(mid, np.float32(np.inf)),
                             np.float32([65519.0, 65520.0, 65536.0, 1e-9, 2.9802322e-08, 2.9802326e-08])])
    probes = np.concatenate([probes, -probes])
    with np.errstate(over="ignore"):
        want = probes.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(orc.half(probes).view(np.uint32), want.view(np.uint32))


def splats(spec):
    """Scene rows from (pos, scale, rgb, opacity) as the reference's viewer
    tests build them: identity rotation, DC = (rgb - 0.5) / C0."""
    p = np.zeros((len(spec), 16), np.float32)
    for i, (pos, scale, rgb, op) in enumerate(spec):
        p[i, 0:3] = pos
        p[i, 3] = math.log(op / (1.0 - op))
        p[i, 4:7] = math.log(scale)
        p[i, 7] = 1.0
        p[i, 8] = 1.0
        p[i, 12:15] = [(c - 0.5) / SH_C0 for c in rgb]
    return p


def orbit_camera(w, h, yaw=0.0, pitch=0.0, distance=3.0, flip=False):
    """The viewer's orbit camera (Y-up display, 50 degrees on the short side;
    engine/src/camera.cpp:83-107); flip = the scene turned half about X."""
    ce, se = math.cos(pitch), math.sin(pitch)
    eye = (distance * ce * math.sin(yaw), distance * se, distance * ce * math.cos(yaw))
    short = min(w, h)
    fov_y = 2.0 * math.degrees(math.atan(math.tan(math.radians(25.0)) * h / short))
    cam = gs.look_at_camera(eye, (0.0, 0.0, 0.0), w, h, fov_y_deg=fov_y)
    if flip:
        for r in range(3):
            cam.rotation[r * 3 + 1] = -cam.rotation[r * 3 + 1]
            cam.rotation[r * 3 + 2] = -cam.rotation[r * 3 + 2]
        cam.position[1] = -cam.position[1]
        cam.position[2] = -cam.position[2]
    return cam


BLACK = (0.0, 0.0, 0.0, 1.0)


def test_single_splat_falloff_closed_form():
    w = h = 64
    opacity = 0.95
    img, visible = orc.view_render32(splats([((0, 0, 0), 0.1, (1, 0, 0), opacity)]), None, orbit_camera(w, h), 0,
                                     clear=BLACK)
    assert visible == 1
    focal = 0.5 * 64.0 / math.tan(math.radians(25.0))
    var = (focal * 0.1 / 3.0) ** 2 + 0.3
    for dx in (0, 2, 4):
        px, py = dx + 0.5, 0.5
        alpha = min(opacity * math.exp(-0.5 * (px * px + py * py) / var), 0.99)
        p = img[32, 32 + dx]
        assert abs(int(p[0]) - 255.0 * alpha) <= 1.0, (dx, p, 255.0 * alpha)  # the reference's own bar is 6
        assert p[1] == 0 and p[2] == 0 and p[3] == 255
    assert tuple(img[2, 2, :3]) == (0, 0, 0)


def test_depth_order_closed_form():
    for red_in_front in (True, False):
        z = 0.5 if red_in_front else -0.5  # the camera sits at +z: larger z is nearer
        scene = splats([((0, 0, z), 0.1, (1, 0, 0), 0.9), ((0, 0, -z), 0.1, (0, 0, 1), 0.9)])
        img, visible = orc.view_render32(scene, None, orbit_camera(64, 64), 0, clear=BLACK)
        p = img[32, 32]
        assert visible == 2
        if red_in_front:
            assert p[0] > 180 and p[2] < 60, p
        else:
            assert p[2] > 180 and p[0] < 60, p


def flip_scenes():
    rng = np.random.default_rng(3)
    u = lambda *s: rng.uniform(-1.0, 1.0, size=s)
    n = 400
    up = np.zeros((n, 16), np.float32)
    up[:, 0:3] = 0.8 * u(n, 3)
    up[:, 4:7] = np.log(0.03 + 0.05 * np.abs(u(n, 3)))
    up[:, 12:15] = 1.5 * u(n, 3)
    up[:, 8:12] = u(n, 4)
    up[:, 3] = 2.0 * u(n)
    up[:, 7] = 1.0
    sh_up = (0.3 * u(n, 15, 3)).astype(np.float32)
    sign = np.float32([-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1])
    down = up.copy()
    down[:, 1:3] *= -1.0
    down[:, 10:12] *= -1.0
    sh_down = sh_up * sign[None, :, None]
    return up, sh_up, down, sh_down


def test_flip_equivalence_with_sh_parity():
    up, sh_up, down, sh_down = flip_scenes()
    img_up, _ = orc.view_render32(up, orc.half(sh_up), orbit_camera(96, 96, 0.7, 0.4), 3, clear=BLACK)
    img_down, _ = orc.view_render32(down, orc.half(sh_down), orbit_camera(96, 96, 0.7, 0.4, flip=True), 3,
                                    clear=BLACK)
    lit = int(((img_up[..., :3].astype(int).sum(-1)) > 30).sum())
    worst = int(np.abs(img_up.astype(int) - img_down.astype(int)).max())
    assert lit > 500
    assert worst <= 3  # the reference's bar


def test_quad_contract_against_float64_gaussian():
    """A rotated anisotropic splat alone over black: inside the quad
    (|d . e_i| <= 2 sqrt(2 lambda_i)) red = round(255 min(o exp(-d^T S^-1 d / 2),
    0.99)) with S the dilated covariance, 0 outside; lambda_i, e_i from numpy."""
    w, h = 96, 80
    p = splats([((0.05, -0.03, 0.1), 0.1, (1, 0, 0), 0.8)])
    p[0, 4:7] = np.log([0.25, 0.06, 0.1])
    q = np.array([0.9, 0.2, -0.3, 0.25])
    p[0, 8:12] = q / np.linalg.norm(q)
    cam = orbit_camera(w, h, 0.3, 0.2)
    img, visible = orc.view_render32(p, None, cam, 0, clear=BLACK)
    assert visible == 1
    pr = orc.project64(p, None, cam, 0)[0]
    mx, my, op = pr[1], pr[2], pr[4]
    cov = np.array([[pr[11], pr[12]], [pr[12], pr[13]]])
    conic = np.linalg.inv(cov)
    np.testing.assert_allclose(conic, [[pr[5], pr[6]], [pr[6], pr[7]]], rtol=1e-9)
    lam, vec = np.linalg.eigh(cov)
    op16 = float(orc.half(np.float32(1.0 / (1.0 + math.exp(-float(p[0, 3]))))))
    ys, xs = np.mgrid[0:h, 0:w]
    d = np.stack([xs + 0.5 - mx, ys + 0.5 - my], -1)
    proj = d @ vec / np.sqrt(2.0 * lam)  # local quad coordinates
    alpha = np.minimum(op16 * np.exp(-0.5 * np.einsum("...i,ij,...j->...", d, conic, d)), 0.99)
    inside = (np.abs(proj) <= 2.0).all(-1) & (alpha >= 1.0 / 255.0)
    want = np.where(inside, np.rint(255.0 * alpha), 0.0)
    # pixels at the quad edge or the alpha cut are decided by the last bits
    edge = (np.abs(np.abs(proj) - 2.0) < 1e-4).any(-1) | (np.abs(alpha - 1.0 / 255.0) < 1e-6)
    assert inside.sum() > 200 and (~inside & (alpha >= 1.0 / 255.0)).sum() > 0  # the quad cuts the tails
    diff = np.abs(img[..., 0].astype(float) - want)[~edge]
    assert diff.max() <= 1.0
    assert (img[..., 1] == 0).all() and (img[..., 2] == 0).all()


def test_clear_colour_alpha_and_empty_view():
    p = splats([((0, 0, 5.0), 0.1, (1, 1, 1), 0.9)])  # behind the camera
    clear = (0.08, 0.09, 0.11, 1.0)
    img, visible = orc.view_render32(p, None, orbit_camera(40, 30), 0, clear=clear)
    assert visible == 0
    np.testing.assert_array_equal(img.reshape(-1, 4), np.tile(np.uint8([20, 23, 28, 255]), (40 * 30, 1)))


def test_premultiplied_over_into_transparent_target():
    """Clear alpha 0: one splat leaves alpha = round(255 a) and colour
    round(255 c a) (premultiplied), the UNORM target's values."""
    img, _ = orc.view_render32(splats([((0, 0, 0), 0.1, (0.5, 0.25, 1.0), 0.6)]), None, orbit_camera(64, 64), 0,
                               clear=(0, 0, 0, 0))
    a = img[32, 32, 3] / 255.0
    assert 0.5 < a < 0.61
    for c, col in enumerate((0.5, 0.25, 1.0)):
        assert abs(int(img[32, 32, c]) - 255.0 * col * a) <= 1.5


def test_antialiased_scales_opacity_by_the_determinant_ratio():
    p = splats([((0, 0, 0), 0.02, (1, 0, 0), 0.9)])  # small: the dilation matters
    cam = orbit_camera(64, 64)
    plain, _ = orc.view_render32(p, None, cam, 0, clear=BLACK)
    aa, _ = orc.view_render32(p, None, cam, 0, antialiased=True, clear=BLACK)
    pr = orc.project64(p, None, cam, 0)[0]
    det = pr[11] * pr[13] - pr[12] ** 2
    det_raw = (pr[11] - 0.3) * (pr[13] - 0.3) - pr[12] ** 2
    ratio = math.sqrt(det_raw / det)
    assert ratio < 0.8
    assert abs(int(aa[32, 32, 0]) - plain[32, 32, 0] * ratio) <= 2.0
