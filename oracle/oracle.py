"""ctypes binding of liboracle.so — TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference's training-step algorithm (see
oracle.c's header for what each function follows and what pins it).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as
the checker; nothing in the product imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcCamera(ctypes.Structure):
    _fields_ = [("rotation", ctypes.c_float * 9), ("translation", ctypes.c_float * 3),
                ("position", ctypes.c_float * 3), ("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        f32, f64, u32, u64, i32 = ctypes.c_float, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        P = ctypes.POINTER
        cam_p = P(OrcCamera)
        sig = {
            "orc_expf": ([f32], f32), "orc_logf": ([f32], f32),
            "orc_sh_basis64": ([P(f64), i32, P(f64)], None),
            "orc_footprint_tiles32": ([f32, f32, f32, f32, f32, f32, u32, u32, P(u32), u32], u32),
            "orc_project32": ([P(f32), P(f32), u32, u32, cam_p, u32, i32, P(f32), P(f32), P(u32), P(u32)], None),
            "orc_bin32": ([P(f32), P(f32), P(u32), P(u32), u32, u32, u32, P(u32), P(u32), u64, P(u32)], u64),
            "orc_bin32_bins": ([P(f32), P(f32), P(u32), P(u32), u32, u32, u32, u32, P(u32), P(u32), u64, P(u32)], u64),
            "orc_footprint_bins32": ([f32, f32, f32, f32, f32, f32, u32, u32, u32, P(u32), u32], u32),
            "orc_bin_counts32": ([P(f32), P(f32), P(u32), u32, u32, u32, u32, P(u32)], None),
            "orc_forward32": ([P(f32), P(u32), P(u32), u32, u32, P(f32), P(u32), P(ctypes.c_uint8), P(f32)], None),
            "orc_project64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64)], None),
            "orc_render64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(f64), P(f64)], None),
            "orc_ssim64": ([P(f64), P(f64), u32, u32, f64, f64, P(f64)], f64),
            "orc_loss_image64": ([P(f64), P(f64), u32, u32, f64, P(f64)], f64),
            "orc_loss64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(ctypes.c_uint8), i32, f64], f64),
            "orc_backward64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(ctypes.c_uint8), i32, f64, P(f64),
                                P(f64), P(f64)], f64),
            "orc_backward32": ([P(f32), P(u32), P(u32), P(f32), P(u32), P(f32), P(f32), u32, u32, u32, P(f64),
                                P(ctypes.c_uint8), P(f64)], None),
            "orc_render64_brute": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(f64)], None),
            "orc_fd_frozen64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(ctypes.c_uint8), f64, u32, P(u32),
                                 P(u32), f64, P(f64)], f64),
            "orc_adjoint64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(f64), P(f64)], None),
            "orc_pose_adjoint64": ([P(f64), P(f64), u32, u32, cam_p, i32, P(f64), P(f64)], None),
            "orc_adam": ([P(f32), P(f32), P(f32), P(f32), P(f32), u64, f32, f32, f32, f32, f32], None),
            "orc_densify_mark": ([P(f32), P(f32), u32, f32, f32, f32, f32, f32, P(u32), P(u32), P(u32)], None),
            "orc_half": ([f32], f32),
            "orc_view_render32": ([P(f32), P(f32), u32, u32, cam_p, u32, i32, P(f32), P(ctypes.c_uint8), P(u32)],
                                  None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def camera(cam) -> OrcCamera:
    """An OrcCamera from any object with gs::Camera's fields."""
    c = OrcCamera()
    c.rotation[:] = list(cam.rotation)
    c.translation[:] = list(cam.translation)
    c.position[:] = list(cam.position)
    c.fx, c.fy, c.cx, c.cy = cam.fx, cam.fy, cam.cx, cam.cy
    c.width, c.height = cam.width, cam.height
    return c


def expf(x: float) -> float:
    return library().orc_expf(x)


def logf(x: float) -> float:
    return library().orc_logf(x)


def sh_basis64(d, degree: int) -> np.ndarray:
    d = np.ascontiguousarray(d, np.float64)
    out = np.zeros(16, np.float64)
    library().orc_sh_basis64(_p(d, ctypes.c_double), degree, _p(out, ctypes.c_double))
    return out


def footprint_tiles32(mx, my, a, b, c, t, width, height, side: int = 16) -> np.ndarray:
    """The bins of `side` pixels (16: the reference's tiles) a footprint reaches."""
    lib = library()
    n = lib.orc_footprint_bins32(mx, my, a, b, c, t, width, height, side, None, 0)
    out = np.zeros(max(n, 1), np.uint32)
    lib.orc_footprint_bins32(mx, my, a, b, c, t, width, height, side, _p(out, ctypes.c_uint32), n)
    return out[:n]


def bin_counts32(records, foot, tiles, width: int, height: int, side: int) -> np.ndarray:
    """Bins of `side` pixels each slot's footprint reaches (0 where tiles is 0)."""
    n = records.shape[0]
    out = np.zeros(n, np.uint32)
    library().orc_bin_counts32(_p(np.ascontiguousarray(records, np.float32), ctypes.c_float),
                               _p(np.ascontiguousarray(foot, np.float32), ctypes.c_float),
                               _p(np.ascontiguousarray(tiles, np.uint32), ctypes.c_uint32), n, width, height, side,
                               _p(out, ctypes.c_uint32))
    return out


def project32(params, sh, cam, sh_degree: int, depth16: bool = False):
    params = np.ascontiguousarray(params, np.float32).reshape(-1, 16)
    n = params.shape[0]
    sh = np.ascontiguousarray(sh if sh is not None else np.zeros((n, 0, 3)), np.float32).reshape(n, -1, 3)
    coeffs = sh.shape[1]
    records = np.zeros((n, 12), np.float32)
    foot = np.zeros((n, 4), np.float32)
    tiles = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint32)
    shp = sh if sh.size else np.zeros(1, np.float32)
    library().orc_project32(_p(params, ctypes.c_float), _p(shp, ctypes.c_float), n, coeffs,
                            ctypes.byref(camera(cam)), sh_degree, int(depth16), _p(records, ctypes.c_float),
                            _p(foot, ctypes.c_float), _p(tiles, ctypes.c_uint32), _p(keys, ctypes.c_uint32))
    return records, foot, tiles, keys


def bin32(records, foot, tiles, keys, width: int, height: int, side: int = 16):
    """Pairs sorted by (bin, key, index) and the bins' ranges; `tiles` holds
    each slot's count of bins of `side` pixels (16: the reference's tiles)."""
    n = records.shape[0]
    total = int(np.sum(tiles, dtype=np.uint64))
    pt = np.zeros(max(total, 1), np.uint32)
    pg = np.zeros(max(total, 1), np.uint32)
    T = ((width + side - 1) // side) * ((height + side - 1) // side)
    ranges = np.zeros((T + 1, 2), np.uint32)
    got = library().orc_bin32_bins(_p(np.ascontiguousarray(records, np.float32), ctypes.c_float),
                                   _p(np.ascontiguousarray(foot, np.float32), ctypes.c_float),
                                   _p(np.ascontiguousarray(tiles, np.uint32), ctypes.c_uint32),
                                   _p(np.ascontiguousarray(keys, np.uint32), ctypes.c_uint32), n, width, height, side,
                                   _p(pt, ctypes.c_uint32), _p(pg, ctypes.c_uint32), total,
                                   _p(ranges, ctypes.c_uint32))
    assert got == total
    return pt[:total], pg[:total], ranges


def forward32(records, pair_gaussians, ranges, width: int, height: int, flags: bool = False, spread: bool = False):
    """(image [HW,4], last [HW]) of the fp32 tile compositor; with flags=True also
    ambiguous [HW] (bool): pixels that met a splat at a skip/clamp/stop threshold;
    with spread=True also spread [HW]: how far another fp32 evaluation of the same
    splats can move each pixel's values (oracle.c orc_forward32)."""
    image = np.zeros((height * width, 4), np.float32)
    last = np.zeros(height * width, np.uint32)
    amb = np.zeros(height * width, np.uint8)
    spr = np.zeros(height * width, np.float32)
    pg = np.ascontiguousarray(pair_gaussians, np.uint32)
    if pg.size == 0:
        pg = np.zeros(1, np.uint32)
    library().orc_forward32(_p(np.ascontiguousarray(records, np.float32), ctypes.c_float), _p(pg, ctypes.c_uint32),
                            _p(np.ascontiguousarray(ranges, np.uint32), ctypes.c_uint32), width, height,
                            _p(image, ctypes.c_float), _p(last, ctypes.c_uint32), _p(amb, ctypes.c_uint8),
                            _p(spr, ctypes.c_float))
    out = (image, last)
    if flags:
        out += (amb.astype(bool),)
    if spread:
        out += (spr,)
    return out


def _scene64(params, sh):
    params = np.ascontiguousarray(params, np.float64).reshape(-1, 16)
    n = params.shape[0]
    sh = np.ascontiguousarray(sh if sh is not None else np.zeros((n, 0, 3)), np.float64).reshape(n, -1, 3)
    shp = sh if sh.size else np.zeros(1, np.float64)
    return params, sh, shp, n, sh.shape[1]


def project64(params, sh, cam, degree: int) -> np.ndarray:
    params, sh, shp, n, coeffs = _scene64(params, sh)
    out = np.zeros((n, 16), np.float64)
    library().orc_project64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                            ctypes.byref(camera(cam)), degree, _p(out, ctypes.c_double))
    return out


def render64(params, sh, cam, degree: int, bg=(0.0, 0.0, 0.0)):
    params, sh, shp, n, coeffs = _scene64(params, sh)
    rgb = np.zeros((cam.height * cam.width, 3), np.float64)
    T = np.zeros(cam.height * cam.width, np.float64)
    bgv = np.asarray(bg, np.float64)
    library().orc_render64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                           ctypes.byref(camera(cam)), degree, _p(bgv, ctypes.c_double), _p(rgb, ctypes.c_double),
                           _p(T, ctypes.c_double))
    return rgb, T


def render64_brute(params, sh, cam, degree: int, bg=(0.0, 0.0, 0.0)) -> np.ndarray:
    """rgb [HW, 3] of the loop over every visible gaussian at every pixel (the
    check of render64's tile lists)."""
    params, sh, shp, n, coeffs = _scene64(params, sh)
    rgb = np.zeros((cam.height * cam.width, 3), np.float64)
    bgv = np.asarray(bg, np.float64)
    library().orc_render64_brute(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                                 ctypes.byref(camera(cam)), degree, _p(bgv, ctypes.c_double), _p(rgb, ctypes.c_double))
    return rgb


def fd_frozen64(params, sh, cam, degree: int, target, lam: float, gaussians, columns, h: float = 1e-4,
                bg=(0.0, 0.0, 0.0)):
    """(loss, fd [k]): central differences of the fp64 loss (unmasked target)
    by parameter columns[k] of gaussian gaussians[k] (0..15 the slot, 16 + 3 j + c
    SH rest coefficient j channel c), with the compositor's branches frozen as
    in the unperturbed render (oracle.c orc_fd_frozen64)."""
    params, sh, shp, n, coeffs = _scene64(params, sh)
    t = np.ascontiguousarray(target, np.uint8).reshape(-1)
    gi = np.ascontiguousarray(gaussians, np.uint32)
    ci = np.ascontiguousarray(columns, np.uint32)
    fd = np.zeros(gi.size, np.float64)
    bgv = np.asarray(bg, np.float64)
    L = library().orc_fd_frozen64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                                  ctypes.byref(camera(cam)), degree, _p(bgv, ctypes.c_double), _p(t, ctypes.c_uint8),
                                  lam, gi.size, _p(gi, ctypes.c_uint32), _p(ci, ctypes.c_uint32), h,
                                  _p(fd, ctypes.c_double))
    return L, fd


def ssim64(x, y, width: int, height: int, grad: bool = False):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 3)
    g = np.zeros_like(x) if grad else None
    s = library().orc_ssim64(_p(x, ctypes.c_double), _p(y, ctypes.c_double), width, height, 1e-4, 9e-4,
                             _p(g, ctypes.c_double) if grad else None)
    return (s, g) if grad else s


def loss_image64(x, y, width: int, height: int, lam: float, grad: bool = False):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 3)
    g = np.zeros_like(x) if grad else None
    L = library().orc_loss_image64(_p(x, ctypes.c_double), _p(y, ctypes.c_double), width, height, lam,
                                   _p(g, ctypes.c_double) if grad else None)
    return (L, g) if grad else L


def loss64(params, sh, cam, degree: int, target, lam: float, bg=(0.0, 0.0, 0.0), masked: bool = False) -> float:
    params, sh, shp, n, coeffs = _scene64(params, sh)
    t = np.ascontiguousarray(target, np.uint8).reshape(-1)
    bgv = np.asarray(bg, np.float64)
    return library().orc_loss64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                                ctypes.byref(camera(cam)), degree, _p(bgv, ctypes.c_double),
                                _p(t, ctypes.c_uint8), int(masked), lam)


def backward64(params, sh, cam, degree: int, target, lam: float, bg=(0.0, 0.0, 0.0), masked: bool = False):
    """(loss, grad2d [n,9], gparams [n,16], gsh [n,coeffs,3])."""
    params, sh, shp, n, coeffs = _scene64(params, sh)
    t = np.ascontiguousarray(target, np.uint8).reshape(-1)
    bgv = np.asarray(bg, np.float64)
    g2 = np.zeros((n, 9), np.float64)
    gp = np.zeros((n, 16), np.float64)
    gs = np.zeros((n, max(coeffs, 0), 3), np.float64)
    gsp = gs if gs.size else np.zeros(1, np.float64)
    L = library().orc_backward64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                                 ctypes.byref(camera(cam)), degree, _p(bgv, ctypes.c_double), _p(t, ctypes.c_uint8),
                                 int(masked), lam, _p(g2, ctypes.c_double), _p(gp, ctypes.c_double),
                                 _p(gsp, ctypes.c_double))
    return L, g2, gp, gs


def adam(p, g, m, v, lr, beta1=0.9, beta2=0.999, eps=1e-15, bias1=10.0, bias2=1000.0):
    p, g, m, v, lr = (np.ascontiguousarray(a, np.float32).copy() for a in (p, g, m, v, lr))
    library().orc_adam(_p(p, ctypes.c_float), _p(g, ctypes.c_float), _p(m, ctypes.c_float), _p(v, ctypes.c_float),
                       _p(lr, ctypes.c_float), p.size, beta1, beta2, eps, bias1, bias2)
    return p, m, v


def densify_mark(params, stats, grad_threshold, clone_max_scale, prune_opacity, prune_max_scale=0.0,
                 prune_max_radius=0.0):
    params = np.ascontiguousarray(params, np.float32).reshape(-1, 16)
    stats = np.ascontiguousarray(stats, np.float32).reshape(-1, 4)
    n = params.shape[0]
    action = np.zeros(n, np.uint32)
    release = np.zeros(n, np.uint32)
    prune = np.zeros(n, np.uint32)
    library().orc_densify_mark(_p(params, ctypes.c_float), _p(stats, ctypes.c_float), n, grad_threshold,
                               clone_max_scale, prune_opacity, prune_max_scale, prune_max_radius,
                               _p(action, ctypes.c_uint32), _p(release, ctypes.c_uint32), _p(prune, ctypes.c_uint32))
    return action, release, prune


def backward32(records, pair_gaussians, ranges, image, last, dldc, bg, width: int, height: int,
               flags: bool = False, share: bool = False):
    """grad2d [n, 9] (double sums) of the tile backward on the given forward state;
    with flags=True also tainted [n] (bool): gaussians that have a term at a pixel
    whose walk meets a splat at a skip/clamp threshold; with share=True also
    share [n, 9]: how far another fp32 evaluation can move each sum (those
    terms' magnitudes, plus every term's magnitude times its pixel's
    conditioning; oracle.c orc_backward32)."""
    records = np.ascontiguousarray(records, np.float32).reshape(-1, 12)
    n = records.shape[0]
    pg = np.ascontiguousarray(pair_gaussians, np.uint32)
    if pg.size == 0:
        pg = np.zeros(1, np.uint32)
    out = np.zeros((n, 9), np.float64)
    tainted = np.zeros(max(n, 1), np.uint8)
    sh9 = np.zeros((max(n, 1), 9), np.float64)
    bgv = np.asarray(bg, np.float32)
    library().orc_backward32(_p(records, ctypes.c_float), _p(pg, ctypes.c_uint32),
                             _p(np.ascontiguousarray(ranges, np.uint32), ctypes.c_uint32),
                             _p(np.ascontiguousarray(image, np.float32), ctypes.c_float),
                             _p(np.ascontiguousarray(last, np.uint32), ctypes.c_uint32),
                             _p(np.ascontiguousarray(dldc, np.float32), ctypes.c_float), _p(bgv, ctypes.c_float),
                             n, width, height, _p(out, ctypes.c_double),
                             _p(tainted, ctypes.c_uint8) if (flags or share) else None,
                             _p(sh9, ctypes.c_double) if share else None)
    res = (out,)
    if flags:
        res += (tainted[:n].astype(bool),)
    if share:
        res += (sh9[:n],)
    return res if len(res) > 1 else out


def adjoint64(params, sh, cam, degree: int, grad2d):
    """(gparams [n,16], gsh [n,coeffs,3]) of the projection adjoint of grad2d [n,9]."""
    params, sh, shp, n, coeffs = _scene64(params, sh)
    g2 = np.ascontiguousarray(grad2d, np.float64).reshape(n, 9)
    gp = np.zeros((n, 16), np.float64)
    gs = np.zeros((n, max(coeffs, 0), 3), np.float64)
    gsp = gs if gs.size else np.zeros(1, np.float64)
    library().orc_adjoint64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                            ctypes.byref(camera(cam)), degree, _p(g2, ctypes.c_double), _p(gp, ctypes.c_double),
                            _p(gsp, ctypes.c_double))
    return gp, gs


def pose_adjoint64(params, sh, cam, degree: int, grad2d):
    """[n, 6]: each gaussian's share of the camera pose gradient (dL/dd xyz,
    dL/dw xyz in camera axes) from grad2d [n, 9] (train_step.glsl:196-201);
    the view's gradient is the column sum (trainer.cpp:1319-1330)."""
    params, sh, shp, n, coeffs = _scene64(params, sh)
    g2 = np.ascontiguousarray(grad2d, np.float64).reshape(n, 9)
    out = np.zeros((n, 6), np.float64)
    library().orc_pose_adjoint64(_p(params, ctypes.c_double), _p(shp, ctypes.c_double), n, coeffs,
                                 ctypes.byref(camera(cam)), degree, _p(g2, ctypes.c_double), _p(out, ctypes.c_double))
    return out


def target_colour64(target, bg=(0.0, 0.0, 0.0), masked: bool = False, exposure=None) -> np.ndarray:
    """[pixels, 3]: the colour the render is compared with (train_forward.comp:
    122-127): the photo times the view's exposure (1 / its gain), over the
    background by the photo's alpha when masked."""
    t = np.asarray(target, np.uint8).reshape(-1, 4).astype(np.float64) / 255.0
    photo = t[:, :3] * (np.asarray(exposure, np.float64) if exposure is not None else 1.0)
    if masked:
        return np.asarray(bg, np.float64) * (1.0 - t[:, 3:4]) + photo * t[:, 3:4]
    return photo


def exposure_sums64(image, target) -> np.ndarray:
    """The exposure gain's least-squares sums of one view (train_forward.comp:
    142-151, summed as trainer.cpp:786-792): over pixels whose photo alpha is
    above 0.95 and whose final transmittance is below 0.05, sum C^2 per
    channel, the pixel count, sum C x photo per channel; then the pixels with
    any alpha.  image [pixels, 4] = (C rgb before the background, T)."""
    im = np.asarray(image, np.float64).reshape(-1, 4)
    t = np.asarray(target, np.uint8).reshape(-1, 4).astype(np.float64) / 255.0
    fit = (t[:, 3] > 0.95) & (im[:, 3] < 0.05)
    C, P = im[fit, :3], t[fit, :3]
    return np.concatenate([(C * C).sum(0), [fit.sum()], (C * P).sum(0), [(t[:, 3] > 0).sum()]])


def half(x) -> np.ndarray:
    """Round to the nearest fp16 value (ties to even), as float32."""
    f = library().orc_half
    a = np.asarray(x, np.float32)
    return np.array([f(float(v)) for v in a.reshape(-1)], np.float32).reshape(a.shape)


def view_render32(params, sh, cam, degree: int, antialiased: bool = False, clear=(0.0, 0.0, 0.0, 0.0)):
    """(rgba [H, W, 4] uint8, visible) of the viewer renderer; `sh` is used as
    given (round it to fp16 first for the viewer's storage precision)."""
    params = np.ascontiguousarray(params, np.float32).reshape(-1, 16)
    n = params.shape[0]
    sh = np.ascontiguousarray(sh if sh is not None else np.zeros((n, 0, 3)), np.float32).reshape(n, -1, 3)
    coeffs = sh.shape[1]
    shp = sh if sh.size else np.zeros(1, np.float32)
    out = np.zeros((cam.height, cam.width, 4), np.uint8)
    cl = np.ascontiguousarray(clear, np.float32)
    vis = ctypes.c_uint32(0)
    library().orc_view_render32(_p(params, ctypes.c_float), _p(shp, ctypes.c_float), n, coeffs,
                                ctypes.byref(camera(cam)), degree, int(antialiased), _p(cl, ctypes.c_float),
                                _p(out, ctypes.c_uint8), ctypes.byref(vis))
    return out, int(vis.value)


def adam_rates(n: int, lr_position: float, lr_opacity: float, lr_scale: float, lr_rotation: float,
               lr_sh_dc: float) -> np.ndarray:
    """[n, 16] per-parameter learning rates in the slot layout
    (train_step.glsl:232-240): position, opacity | scale, - | rotation | DC, -."""
    lr = np.zeros(16, np.float32)
    lr[0:3] = lr_position
    lr[3] = lr_opacity
    lr[4:7] = lr_scale
    lr[8:12] = lr_rotation
    lr[12:15] = lr_sh_dc
    return np.tile(lr, (n, 1))


def train(params, sh, deg: int, cams, targets, steps: int, lr, lr_sh_rest: float, lam: float = 0.2, order=None,
          bg=(0.0, 0.0, 0.0), step0: int = 0, state=None):
    """`steps` training steps of the CPU restatement, one view each (view
    order[k], or k mod views): fp32 projection, binning, compositing and tile
    backward; fp64 loss, SSIM and projection adjoint; fp32 Adam with
    beta 0.9 / 0.999, eps 1e-15 and the bias corrections of step step0 + k + 1.
    lr: [n, 16] rates (adam_rates).  state: the Adam moments (m, v, m_sh,
    v_sh) to continue from, kept in this dict across calls when given (a new
    dict starts at zero).  Returns (params, sh) after the steps."""
    p, s = np.array(params, np.float32, copy=True), np.array(sh, np.float32, copy=True)
    if state is None:
        state = {}
    if not state:
        state.update(m=np.zeros_like(p), v=np.zeros_like(p), msh=np.zeros_like(s), vsh=np.zeros_like(s))
    m, v, msh, vsh = state["m"], state["v"], state["msh"], state["vsh"]
    lr_rest = np.full(s.shape, lr_sh_rest, np.float32)
    for k in range(steps):
        view = int(order[k]) if order is not None else k % len(cams)
        cam, target = cams[view], targets[view]
        w, h = cam.width, cam.height
        rec, foot, tiles, keys = project32(p, s, cam, deg)
        pt, pg, ranges = bin32(rec, foot, tiles, keys, w, h)
        img, last = forward32(rec, pg, ranges, w, h)
        x = img[:, :3].astype(np.float64) + img[:, 3:4].astype(np.float64) * np.asarray(bg, np.float64)
        y = np.asarray(target).reshape(-1, 4)[:, :3] / 255.0
        _, dldc = loss_image64(x, y, w, h, lam, grad=True)
        d4 = np.zeros((w * h, 4), np.float32)
        d4[:, :3] = dldc
        g2 = backward32(rec, pg, ranges, img, last, d4, bg, w, h)
        gp, gsh = adjoint64(p, s, cam, deg, g2)
        b1, b2 = 1.0 / (1.0 - 0.9 ** (step0 + k + 1)), 1.0 / (1.0 - 0.999 ** (step0 + k + 1))
        _adam_in_place(p, gp.astype(np.float32), m, v, lr, b1, b2)
        if s.size:
            _adam_in_place(s, gsh.astype(np.float32), msh, vsh, lr_rest, b1, b2)
    state.update(m=m, v=v, msh=msh, vsh=vsh)
    return p, s


def _adam_in_place(p, g, m, v, lr, bias1, bias2, beta1=0.9, beta2=0.999, eps=1e-15):
    """orc_adam on the arrays themselves (contiguous float32)."""
    library().orc_adam(_p(p, ctypes.c_float), _p(g, ctypes.c_float), _p(m, ctypes.c_float), _p(v, ctypes.c_float),
                       _p(lr, ctypes.c_float), p.size, beta1, beta2, eps, bias1, bias2)


def mean_psnr(params, sh, deg: int, cams, targets, bg=(0.0, 0.0, 0.0)) -> float:
    """Mean over the views of the PSNR of the fp64 render against the RGBA8 targets' colour."""
    out = []
    for cam, t in zip(cams, targets):
        rgb, _ = render64(params, sh, cam, deg, bg)
        y = np.asarray(t).reshape(-1, 4)[:, :3] / 255.0
        out.append(-10.0 * np.log10(max(float(np.mean((rgb - y) ** 2)), 1e-12)))
    return float(np.mean(out))


def render_targets64(params, sh, deg: int, cams):
    """RGBA8 targets (alpha 255) from the fp64 render, rounded to nearest."""
    out = []
    for cam in cams:
        rgb, _ = render64(params, sh, cam, deg)
        t = np.zeros((cam.height, cam.width, 4), np.uint8)
        t[..., :3] = np.clip(np.rint(rgb.reshape(cam.height, cam.width, 3) * 255.0), 0, 255).astype(np.uint8)
        t[..., 3] = 255
        out.append(t)
    return out
