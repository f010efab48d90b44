"""Stage-by-stage comparison of one view of a HIP step with the oracle, on
the state the GPU left in a workspace (shared by the parity tests of every
size).  Bars, each stated where it is asserted:

* bit-exact: raster records, footprints, tile counts, depth keys and order,
  sorted pairs, tile ranges;
* forward: the oracle's fp32 tile compositor on the GPU's records and pairs;
  the last contributor equal on >= 99.9% of the pixels that meet no splat at
  a skip / clamp / stop threshold (the GPU's hardware exp and evaluation
  order may decide those the other way), and their image within 2e-5 plus
  twice the pixel's evaluation spread (the oracle's estimate of how far
  another fp32 evaluation moves it: tiny except where needle splats far
  along their axis cancel large terms); the ambiguous pixels within one
  splat's share;
* loss: the sums and dL/dC against the fp64 loss of the GPU's own image
  (1e-4 relative);
* backward: the oracle's tile backward on the GPU's forward state; every
  gaussian's nine sums within 1e-4 of that value's largest magnitude plus the
  gaussian's share (the magnitudes of its terms at ambiguous pixels, where
  another evaluation's decisions can move them, and every term's magnitude
  times its pixel's conditioning -- the walk divides by 1 - alpha, so 8 ulp of
  alpha move the recovered T by 8 ulp x alpha / (1 - alpha) -- and 32 ulp of
  fp32 accumulation, whose rounding follows the sum of the terms' magnitudes
  rather than the sum); at most 2% of the visible gaussians may have an error
  above the strict 1e-4 bound (within their share).
"""
import numpy as np

from helpers import gs, orc


def view_state(ws, n, pairs, view=0):
    d = lambda which, dt, cnt: ws.download(which, dt, cnt, view=view)
    rec = d(gs.BUF_RECORDS, np.float32, n * 12).reshape(n, 12)
    foot8 = d(gs.BUF_FOOTPRINT, np.float32, n * 8).reshape(n, 8)
    tiles = d(gs.BUF_TILES, np.uint32, n)
    keys = d(gs.BUF_DEPTH_KEYS, np.uint32, n)
    order = d(gs.BUF_DEPTH_ORDER, np.uint32, n)
    pt = d(gs.BUF_PAIR_TILES, np.uint32, pairs)
    pg = d(gs.BUF_PAIR_GAUSSIANS, np.uint32, pairs)
    T = ws.tiles_x * ws.tiles_y
    ranges = d(gs.BUF_RANGES, np.uint32, (T + 1) * 2).reshape(T + 1, 2)
    np.testing.assert_array_equal(foot8[:, 4:6].view(np.uint32), rec[:, 0:2].view(np.uint32))
    np.testing.assert_array_equal(foot8[:, 6].view(np.uint32), tiles)
    return rec, foot8[:, :4].copy(), tiles, keys, order, pt, pg, ranges


def check_project_bin(state, params, sh, cam, deg, depth16, label):
    """Bit-exact projection, depth order and binning; returns the pair count."""
    rec, foot, tiles, keys, order, pt, pg, ranges = state
    w, h = cam.width, cam.height
    r_rec, r_foot, r_tiles, r_keys = orc.project32(params, sh, cam, deg, depth16=depth16)
    np.testing.assert_array_equal(tiles, r_tiles)
    if depth16:  # 16-bit keys carry the slot's tile count in their upper half
        np.testing.assert_array_equal(keys & 0xFFFF, r_keys[order] & 0xFFFF)
        np.testing.assert_array_equal(keys >> 16, tiles[order])
    else:
        np.testing.assert_array_equal(keys, r_keys[order])
    np.testing.assert_array_equal(order, np.argsort(r_keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(rec.view(np.uint32), r_rec.view(np.uint32))
    np.testing.assert_array_equal(foot.view(np.uint32), r_foot.view(np.uint32))
    o_pt, o_pg, o_ranges = orc.bin32(r_rec, r_foot, r_tiles, r_keys, w, h)
    assert len(pt) == len(o_pt) and len(pt) > 0
    np.testing.assert_array_equal(pt, o_pt)
    np.testing.assert_array_equal(pg, o_pg)
    np.testing.assert_array_equal(ranges, o_ranges)
    print(f"[ok] {label}: bit-exact projection and binning, {int((tiles > 0).sum())} visible, {len(pt)} pairs",
          flush=True)
    return len(pt)


def check_forward(rec, pg, ranges, img, last, w, h, label, max_ambiguous=0.01):
    """max_ambiguous: the largest fraction of pixels that may meet a splat at a
    decision threshold (1% for ordinary scenes; needle splats far along their
    axis make fp32 evaluation ill-conditioned over much of the image)."""
    o_img, o_last, amb, spread = orc.forward32(rec, pg, ranges, w, h, flags=True, spread=True)
    same = last == o_last
    assert same[~amb].mean() > 0.999, f"{label}: last contributor differs on {(~same & ~amb).sum()} clear pixels"
    assert amb.mean() < max_ambiguous, f"{label}: {amb.sum()} threshold-ambiguous pixels"
    perr = np.abs(img - o_img).max(axis=1)
    clear = same & ~amb
    over = perr - 2.0 * spread  # each pixel may differ by twice its evaluation spread
    print(f"[info] {label}: forward: {amb.sum()} ambiguous pixels, {(~same).sum()} with another last contributor "
          f"({(~same & ~amb).sum()} of them clear); max error clear {perr[clear].max():.3g} (beyond twice the "
          f"evaluation spread {over[clear].max():.3g}; spread above 1e-5 at {(spread > 1e-5).sum()} pixels), "
          f"ambiguous {perr[same & amb].max(initial=0):.3g}", flush=True)
    if over[clear].max() >= 2e-5:
        for p in np.argsort(-np.where(clear, over, -1))[:3]:
            describe_pixel(rec, pg, ranges, img, o_img, last, o_last, int(p), w)
    assert over[clear].max() < 2e-5
    assert perr[same & amb].max(initial=0) <= 2.0 / 255.0 * max(1.0, float(np.abs(rec[:, 6:9]).max())) or \
        max_ambiguous > 0.01


def describe_pixel(rec, pg, ranges, img, o_img, last, o_last, p, w):
    """Prints one pixel's splats in its tile list: the oracle's power, alpha
    and whether the splat's bbox (the rasterizers' first test) holds it."""
    py, px = divmod(p, w)
    t = (py // 16) * ((w + 15) // 16) + px // 16
    print(f"[diag] pixel {px},{py}: gpu {img[p]} last {last[p]}, oracle {o_img[p]} last {o_last[p]}", flush=True)
    T = 1.0
    for k in range(ranges[t, 0], ranges[t, 1]):
        r = rec[pg[k]]
        dx, dy = np.float32(r[0] - (px + 0.5)), np.float32(r[1] - (py + 0.5))
        power = np.float32(-0.5) * (r[2] * dx * dx + r[4] * dy * dy) - r[3] * dx * dy
        raw = r[5] * np.exp(np.float64(power))
        lo, hi = r[9:11].view(np.uint32)
        inside = (lo & 0xffff) <= px <= (hi & 0xffff) and (lo >> 16) <= py <= (hi >> 16)
        mag = 0.5 * (abs(r[2]) * dx * dx + abs(r[4]) * dy * dy) + abs(r[3] * dx * dy)
        if raw >= 1 / 255 or not inside and raw > 1e-3:
            print(f"[diag]   #{k - ranges[t, 0]} g{pg[k]}: power {power:.6g} alpha {raw:.6g} mag {mag:.3g} "
                  f"bbox-holds {inside} T {T:.4g}", flush=True)
        if raw >= 1 / 255 and power <= 0:
            T *= 1 - min(raw, 0.99)
            if T < 1e-4:
                break


def check_loss(img, dldc, losses, target, bg, lam, w, h, label):
    x = img[:, :3].astype(np.float64) + img[:, 3:4].astype(np.float64) * np.asarray(bg)
    y = target.reshape(-1, 4)[:, :3] / 255.0
    L_ref, d_ref = orc.loss_image64(x, y, w, h, lam, grad=True)
    e = x - y
    assert abs(losses[0] - np.abs(e).sum()) <= 1e-4 * np.abs(e).sum()
    assert abs(losses[1] - (e * e).sum()) <= 1e-4 * (e * e).sum()
    ssim_ref = orc.ssim64(x, y, w, h)
    assert abs(losses[2] - ssim_ref) <= 1e-4 * abs(ssim_ref)
    L_gpu = (1 - lam) * losses[0] / (3 * w * h) + lam * (1 - losses[2] / (3 * w * h))
    assert abs(L_gpu - L_ref) < 1e-4 * max(1.0, abs(L_ref))
    sign_ok = np.abs(e) > 1e-6
    dd = np.abs(dldc[:, :3] - d_ref)
    assert dd[sign_ok].max() <= 1e-4 * np.abs(d_ref).max()
    print(f"[ok] {label}: loss {L_gpu:.6f} (fp64 {L_ref:.6f})", flush=True)
    return L_gpu


def describe(rec, pg, ranges, img, last, g2, o_g2, share, i, w, h):
    """Prints what a failing gaussian looks like: its record, both sums, its
    share, and the pixels of its tiles whose GPU last contributor differs
    from the oracle's forward."""
    np.set_printoptions(precision=6, linewidth=160)
    r = rec[i]
    lo, hi = r[9:11].view(np.uint32)
    print(f"[diag] gaussian {i}: mean {r[0]:.3f} {r[1]:.3f} conic {r[2]:.6g} {r[3]:.6g} {r[4]:.6g} opacity {r[5]:.6g} "
          f"rgb {r[6:9]} bbox x {lo & 0xffff}..{hi & 0xffff} y {lo >> 16}..{hi >> 16}", flush=True)
    print(f"[diag]   gpu    {g2[i]}\n[diag]   oracle {o_g2[i]}\n[diag]   share  {share[i]}", flush=True)
    o_img, o_last = orc.forward32(rec, pg, ranges, w, h)
    tiles_x = (w + 15) // 16
    for t in np.nonzero(np.isin(np.arange(len(ranges) - 1), [k for k in range(len(ranges) - 1)
                                                            if i in pg[ranges[k, 0]:ranges[k, 1]]]))[0]:
        ty, tx = divmod(int(t), tiles_x)
        for py in range(ty * 16, min(h, ty * 16 + 16)):
            for px in range(tx * 16, min(w, tx * 16 + 16)):
                p = py * w + px
                if last[p] != o_last[p]:
                    print(f"[diag]   pixel {px},{py} (tile {t}): last gpu {last[p]} oracle {o_last[p]}, image "
                          f"gpu {img[p]} oracle {o_img[p]}", flush=True)


def check_backward(rec, pg, ranges, img, last, dldc, bg, w, h, g2, tiles, label, cap=0.02):
    assert np.all(np.isfinite(g2)), f"{label}: {int((~np.isfinite(g2)).sum())} non-finite 2D gradient values"
    o_g2, tainted, share = orc.backward32(rec, pg, ranges, img, last, dldc, bg, w, h, flags=True, share=True)
    visible = tiles > 0
    loosened = np.zeros(len(tiles), bool)
    needed = np.zeros(len(tiles), bool)
    worst = 0.0
    for c in range(9):
        scale = np.abs(o_g2[:, c]).max()
        err = np.abs(g2[:, c] - o_g2[:, c])
        allowed = 1e-4 * scale + share[:, c]
        bad = err > allowed
        if bad.any():
            describe(rec, pg, ranges, img, last, g2, o_g2, share, int(np.nonzero(bad)[0][np.argmax(err[bad])]), w, h)
        assert not bad.any(), (f"{label}: grad2d[{c}]: {bad.sum()} gaussians off, worst err "
                               f"{err[bad].max():.3g} allowed {allowed[bad][np.argmax(err[bad])]:.3g}, max {scale:.3g}")
        loosened |= share[:, c] > 1e-4 * scale
        needed |= err > 1e-4 * scale
        worst = max(worst, float((err / (scale + 1e-300)).max()))
    nv = max(1, visible.sum())
    print(f"[info] {label}: backward: {visible.sum()} visible gaussians; {tainted.sum()} ({100 * tainted.sum() / nv:.2f}%) "
          f"with a term at an ambiguous pixel, {loosened.sum()} ({100 * loosened.sum() / nv:.3f}%) whose share loosens "
          f"the 1e-4 bound, {needed.sum()} ({100 * needed.sum() / nv:.3f}%) whose error needs it; worst error "
          f"{worst:.3g} of a value's largest", flush=True)
    assert needed.sum() / nv <= cap, f"{label}: {100 * needed.sum() / nv:.2f}% of the gaussians rely on their share"
    return o_g2


def frozen(lam=0.2):
    """All learning rates zero: Adam's first moment then holds (1 - beta1) g."""
    return gs.TrainParams(lambda_dssim=lam, lr_position=0.0, lr_scale=0.0, lr_rotation=0.0, lr_opacity=0.0,
                          lr_sh_dc=0.0, lr_sh_rest=0.0)


def first_step_gradients(ws, n):
    m = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16) / (1 - np.float32(0.9))
    coeffs = ws.sh_coeffs
    msh = ws.download(gs.BUF_M_SH, np.float32, coeffs * ws.capacity * 3).reshape(coeffs, ws.capacity, 3)
    msh = np.transpose(msh[:, :n, :], (1, 0, 2)) / (1 - np.float32(0.9))
    return m, msh


GROUPS = {"position": [0, 1, 2], "opacity": [3], "scale": [4, 5, 6], "rotation": [8, 9, 10, 11],
          "dc": [12, 13, 14]}


def check_adjoint(gp, gsh, params, sh, cams, deg, g2_views, label, tol=1e-4):
    """The optimizer's gradient (Adam's first moment) against the fp64
    projection adjoint of the GPU's 2D gradients, averaged over the views:
    within tol of the largest value of the group in any one view (the mean of
    V views, each within tol of its own largest, is within tol of the
    largest of them)."""
    r_gp, r_gsh = None, None
    view_max = {name: 0.0 for name in GROUPS}
    view_max_sh = 0.0
    for cam_v, g2 in zip(cams, g2_views):
        a, b = orc.adjoint64(params, sh, cam_v, deg, g2.astype(np.float64))
        for name, cols in GROUPS.items():
            view_max[name] = max(view_max[name], float(np.abs(a[:, cols]).max()))
        view_max_sh = max(view_max_sh, float(np.abs(b).max()) if b.size else 0.0)
        r_gp = a if r_gp is None else r_gp + a
        r_gsh = b if r_gsh is None else r_gsh + b
    r_gp /= len(g2_views)
    r_gsh /= len(g2_views)
    for gname, cols in GROUPS.items():
        scale = view_max[gname]
        err = np.abs(gp[:, cols] - r_gp[:, cols]).max()
        assert scale > 0 and err <= tol * scale, f"{label} {gname}: err {err:.3g} vs {scale:.3g}"
    if r_gsh.size:
        assert np.abs(gsh - r_gsh).max() <= tol * view_max_sh
    print(f"[ok] {label}: adjoint within {tol:g} of each group's largest", flush=True)


def check_view(ws, view, pairs, params, sh, cam, target, losses, depth16, label, bg=(0.0, 0.0, 0.0), lam=0.2,
               deg=3):
    n = params.shape[0]
    w, h = cam.width, cam.height
    state = view_state(ws, n, pairs, view)
    check_project_bin(state, params, sh, cam, deg, depth16, label)
    rec, foot, tiles, keys, order, pt, pg, ranges = state
    img = ws.download(gs.BUF_IMAGE, np.float32, view=view).reshape(-1, 4)
    last = ws.download(gs.BUF_LAST, np.uint32, w * h, view=view)
    dldc = ws.download(gs.BUF_DLDC, np.float32, view=view).reshape(-1, 4)
    g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n, view=view).reshape(n, 9)
    check_forward(rec, pg, ranges, img, last, w, h, label)
    check_loss(img, dldc, losses, target, bg, lam, w, h, label)
    check_backward(rec, pg, ranges, img, last, dldc, bg, w, h, g2, tiles, label)
    return g2
