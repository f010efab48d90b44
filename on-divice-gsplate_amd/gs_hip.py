"""ctypes binding of libgs_hip.so (the C ABI in include/gs_hip.h).

The Python-side mirror of the trainer's stage interface, used by the tests
and the benchmark.  It loads the in-tree library built by the Makefile next
to this file and fails loudly when it is missing: there is no CPU fallback.
Device memory is owned by the library's workspace; Python moves numpy arrays
in and out of it.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgs_hip.so")

TILE = 16
FLAG_MASKED_TARGET = 1
FLAG_SPARSE_ADAM = 2
FLAG_DEPTH16 = 4
FLAG_NO_UPDATE = 8
FLAG_POSE_GRAD = 16
FLAG_EXPOSURE = 32
FLAG_DETERMINISTIC = 64
VIEW_ANTIALIASED = 1

(BUF_PARAMS, BUF_SH, BUF_M_PARAMS, BUF_V_PARAMS, BUF_M_SH, BUF_V_SH, BUF_STATS, BUF_RECORDS, BUF_FOOTPRINT,
 BUF_TILES, BUF_DEPTH_KEYS, BUF_DEPTH_ORDER, BUF_OFFSETS, BUF_GRAD2D, BUF_PAIR_TILES, BUF_PAIR_GAUSSIANS,
 BUF_RANGES, BUF_IMAGE, BUF_LAST, BUF_TARGET, BUF_DLDC, BUF_TILE_LOSS, BUF_TOTALS, BUF_TILE_ORDER, BUF_TILE_FIT,
 BUF_POSE_GRAD) = range(26)
MAX_VIEWS = 8

EXPORTED = [
    "gs_hip_init", "gs_hip_device_name", "gs_hip_malloc", "gs_hip_free", "gs_hip_upload", "gs_hip_download",
    "gs_hip_memset", "gs_hip_copy", "gs_hip_synchronize", "gs_hip_last_error", "gs_hip_workspace_create",
    "gs_hip_workspace_destroy", "gs_hip_workspace_buffer", "gs_hip_workspace_views",
    "gs_hip_workspace_view_buffer", "gs_hip_workspace_reserve_pairs",
    "gs_hip_workspace_load", "gs_hip_workspace_store", "gs_hip_project", "gs_hip_bin", "gs_hip_forward",
    "gs_hip_set_target", "gs_hip_set_target_device", "gs_hip_loss", "gs_hip_backward", "gs_hip_adam", "gs_hip_densify", "gs_hip_reset_opacity", "gs_hip_train_step", "gs_hip_train_views",
    "gs_hip_render", "gs_hip_workspace_get_tuning", "gs_hip_workspace_set_tuning", "gs_hip_autotune",
    "gs_hip_workspace_overflows", "gs_hip_workspace_timings", "gs_hip_view_render",
    "gs_hip_batch_create", "gs_hip_batch_destroy", "gs_hip_batch_train", "gs_hip_batch_get_tuning",
    "gs_hip_batch_set_tuning", "gs_hip_batch_timings", "gs_hip_pose_gradient", "gs_hip_exposure_sums",
    "gs_hip_spz_info", "gs_hip_workspace_load_spz", "gs_hip_export_ply", "gs_hip_loss_noalloc",
    "gs_hip_trainer_create", "gs_hip_trainer_destroy", "gs_hip_trainer_steps", "gs_hip_trainer_get_stats",
    "gs_hip_trainer_evaluate", "gs_hip_trainer_preview", "gs_hip_trainer_export_ply",
]


class Camera(ctypes.Structure):
    _fields_ = [("rotation", ctypes.c_float * 9), ("translation", ctypes.c_float * 3),
                ("position", ctypes.c_float * 3), ("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("raster_groups_per_cu", ctypes.c_uint32), ("depth_digit_bits", ctypes.c_uint32),
                ("tile_digit_bits", ctypes.c_uint32), ("graph", ctypes.c_uint32), ("bin_shift", ctypes.c_uint32),
                ("stored_state", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class Frame(ctypes.Structure):
    _fields_ = [("cam", Camera), ("tiles_x", ctypes.c_uint32), ("tiles_y", ctypes.c_uint32),
                ("slots", ctypes.c_uint32), ("sh_degree", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("background", ctypes.c_float * 3), ("lambda_dssim", ctypes.c_float),
                ("ssim_c1", ctypes.c_float), ("ssim_c2", ctypes.c_float), ("lr_position", ctypes.c_float),
                ("lr_scale", ctypes.c_float), ("lr_rotation", ctypes.c_float), ("lr_opacity", ctypes.c_float),
                ("lr_sh_dc", ctypes.c_float), ("lr_sh_rest", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("epsilon", ctypes.c_float), ("bias1", ctypes.c_float),
                ("bias2", ctypes.c_float), ("grad_threshold", ctypes.c_float), ("clone_max_scale", ctypes.c_float),
                ("prune_opacity", ctypes.c_float), ("prune_max_scale", ctypes.c_float),
                ("prune_max_radius", ctypes.c_float), ("split_divisor", ctypes.c_float),
                ("exposure", ctypes.c_float * 3)]


_lib = None


def library() -> ctypes.CDLL:
    """The loaded libgs_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               "g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
        fp, dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        u32p, u64p = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)
        frame_p = ctypes.POINTER(Frame)
        sig = {
            "gs_hip_init": [i32], "gs_hip_device_name": [ctypes.c_char_p, sz],
            "gs_hip_malloc": [ctypes.POINTER(vp), sz], "gs_hip_free": [vp], "gs_hip_upload": [vp, vp, sz],
            "gs_hip_download": [vp, vp, sz], "gs_hip_memset": [vp, i32, sz], "gs_hip_copy": [vp, vp, sz],
            "gs_hip_synchronize": [],
            "gs_hip_last_error": [ctypes.c_char_p, sz],
            "gs_hip_workspace_create": [ctypes.POINTER(vp), u32, i32, u32, u32],
            "gs_hip_workspace_destroy": [vp],
            "gs_hip_workspace_buffer": [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)],
            "gs_hip_workspace_reserve_pairs": [vp, u64], "gs_hip_workspace_views": [vp, u32],
            "gs_hip_workspace_view_buffer": [vp, u32, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)],
            "gs_hip_train_views": [vp, frame_p, u32, ctypes.POINTER(vp), u64p, dp],
            "gs_hip_workspace_load": [vp, fp, fp, u32, u32], "gs_hip_workspace_store": [vp, fp, fp, u32],
            "gs_hip_project": [vp, frame_p], "gs_hip_bin": [vp, frame_p, u64p], "gs_hip_forward": [vp, frame_p],
            "gs_hip_set_target": [vp, ctypes.POINTER(ctypes.c_uint8)], "gs_hip_set_target_device": [vp, vp],
            "gs_hip_loss": [vp, frame_p, dp],
            "gs_hip_loss_noalloc": [frame_p, vp, vp, vp, vp, dp],
            "gs_hip_backward": [vp, frame_p], "gs_hip_adam": [vp, frame_p],
            "gs_hip_densify": [vp, frame_p, u32p, u32p],
            "gs_hip_reset_opacity": [vp, frame_p, ctypes.c_float], "gs_hip_train_step": [vp, frame_p, u64p, dp],
            "gs_hip_render": [vp, frame_p, ctypes.POINTER(ctypes.c_uint8)],
            "gs_hip_view_render": [vp, frame_p, fp, u32, ctypes.POINTER(ctypes.c_uint8), u32p],
            "gs_hip_workspace_get_tuning": [vp, ctypes.POINTER(Tuning)],
            "gs_hip_workspace_set_tuning": [vp, ctypes.POINTER(Tuning)],
            "gs_hip_autotune": [vp, frame_p, u32, ctypes.POINTER(vp), u32, ctypes.POINTER(Tuning), dp],
            "gs_hip_workspace_overflows": [vp, u64p],
            "gs_hip_workspace_timings": [vp, i32, ctypes.POINTER(ctypes.c_char_p), dp, i32,
                                         ctypes.POINTER(ctypes.c_int)],
            "gs_hip_batch_create": [ctypes.POINTER(vp), ctypes.POINTER(vp), u32],
            "gs_hip_batch_destroy": [vp],
            "gs_hip_batch_train": [vp, frame_p, u32p, ctypes.POINTER(vp), u64p, dp],
            "gs_hip_batch_get_tuning": [vp, ctypes.POINTER(Tuning)],
            "gs_hip_batch_set_tuning": [vp, ctypes.POINTER(Tuning)],
            "gs_hip_batch_timings": [vp, i32, ctypes.POINTER(ctypes.c_char_p), dp, i32, ctypes.POINTER(ctypes.c_int)],
            "gs_hip_pose_gradient": [vp, u32, dp], "gs_hip_exposure_sums": [vp, u32, dp],
            "gs_hip_spz_info": [ctypes.POINTER(ctypes.c_uint8), sz, u32p, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int)],
            "gs_hip_workspace_load_spz": [vp, ctypes.POINTER(ctypes.c_uint8), sz, u32p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)],
            "gs_hip_export_ply": [vp, u32, i32, ctypes.POINTER(ctypes.c_uint8), sz, u32p],
            "gs_hip_trainer_create": [ctypes.c_char_p, ctypes.POINTER(TrainerOptions), ctypes.POINTER(vp),
                                      ctypes.POINTER(TrainerInfo)],
            "gs_hip_trainer_destroy": [vp],
            "gs_hip_trainer_steps": [vp, u32, i32, u32p, ctypes.POINTER(TrainerStats), ctypes.POINTER(PassTimes)],
            "gs_hip_trainer_get_stats": [vp, ctypes.POINTER(TrainerStats)],
            "gs_hip_trainer_evaluate": [vp, ctypes.POINTER(TrainerEval)],
            "gs_hip_trainer_preview": [vp, u32, ctypes.POINTER(ctypes.c_uint8), sz, u32p, u32p],
            "gs_hip_trainer_export_ply": [vp, ctypes.c_char_p, u32p, u64p],
        }
        missing = sorted(set(EXPORTED) - set(sig))
        if missing:  # an unsigned call would pass Python ints as C int, truncating sizes and pointers
            raise RuntimeError(f"gs_hip.py declares no argument types for {missing}")
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


class GsHipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"gs_hip error {code}: {message}")
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        library().gs_hip_last_error(buf, 1024)
        raise GsHipError(rc, buf.value.decode(errors="replace"))


def init(device: int = 0) -> str:
    lib = library()
    check(lib.gs_hip_init(device))
    buf = ctypes.create_string_buffer(256)
    check(lib.gs_hip_device_name(buf, 256))
    return buf.value.decode()


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ---- cameras -----------------------------------------------------------------
def look_at_camera(eye, target, width: int, height: int, fov_y_deg: float = 60.0, up=(0.0, 1.0, 0.0)) -> Camera:
    """Pinhole camera at `eye` looking at `target`, scene -> camera in OpenCV
    axes (x right, y down, z forward), as dataset_camera builds from an OpenGL
    camera-to-world matrix (engine/src/dataset.cpp:226-250)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f]).astype(np.float32)
    t = (-(R.astype(np.float64) @ eye)).astype(np.float32)
    fy = (height / 2.0) / math.tan(math.radians(fov_y_deg) / 2.0)
    cam = Camera()
    cam.rotation[:] = [float(v) for v in R.reshape(-1)]
    cam.translation[:] = [float(v) for v in t]
    cam.position[:] = [float(v) for v in eye.astype(np.float32)]
    cam.fx = cam.fy = fy
    cam.cx = width / 2.0
    cam.cy = height / 2.0
    cam.width, cam.height = width, height
    return cam


@dataclass
class TrainParams:
    """Per-step knobs (the reference's TrainOptions values for one step)."""
    lambda_dssim: float = 0.2
    lr_position: float = 1.6e-4
    lr_scale: float = 5e-3
    lr_rotation: float = 5e-3
    lr_opacity: float = 5e-3
    lr_sh_dc: float = 5e-3
    lr_sh_rest: float = 5e-3
    grad_threshold: float = 2e-4
    clone_max_scale: float = 0.01
    prune_opacity: float = 0.005
    seed: int = 1


def make_frame(cam: Camera, slots: int, sh_degree: int, step: int = 0, background=(0.0, 0.0, 0.0), flags: int = 0,
               tp: TrainParams | None = None, batch_scaling: int = 1) -> Frame:
    """One view's frame at optimizer step `step` (0-based).  batch_scaling = B
    > 1: Adam's hyperparameters for a step on the mean gradient of B views,
    scaled to stand in for B single-view steps (rates x sqrt(B), beta1^B and
    beta2^B, bias correction with the scaled betas; the rule of Zhao et al.,
    "On Scaling Up 3D Gaussian Splatting Training", 2024)."""
    tp = tp or TrainParams()
    k = max(1, int(batch_scaling))
    rate = math.sqrt(k)
    f = Frame()
    f.cam = cam
    f.tiles_x = (cam.width + TILE - 1) // TILE
    f.tiles_y = (cam.height + TILE - 1) // TILE
    f.slots = slots
    f.sh_degree = sh_degree
    f.flags = flags
    f.seed = (tp.seed * 9781 + step * 6271) & 0xffffffff
    f.background[:] = [float(b) for b in background]
    f.lambda_dssim = tp.lambda_dssim
    f.ssim_c1, f.ssim_c2 = 0.01 ** 2, 0.03 ** 2
    f.lr_position, f.lr_scale, f.lr_rotation = tp.lr_position * rate, tp.lr_scale * rate, tp.lr_rotation * rate
    f.lr_opacity, f.lr_sh_dc, f.lr_sh_rest = tp.lr_opacity * rate, tp.lr_sh_dc * rate, tp.lr_sh_rest * rate
    beta1, beta2 = 0.9 ** k, 0.999 ** k
    f.beta1, f.beta2, f.epsilon = beta1, beta2, 1e-15
    n = step + 1.0
    f.bias1 = 1.0 / (1.0 - beta1 ** n)
    f.bias2 = 1.0 / (1.0 - beta2 ** n)
    f.grad_threshold, f.clone_max_scale, f.prune_opacity = tp.grad_threshold, tp.clone_max_scale, tp.prune_opacity
    f.prune_max_scale = 0.0
    f.prune_max_radius = 0.0
    f.split_divisor = 1.6
    f.exposure[:] = [1.0, 1.0, 1.0]
    return f


# ---- workspace ---------------------------------------------------------------
class Workspace:
    """Every device buffer of one trainer (gs_hip_workspace)."""

    def __init__(self, capacity: int, sh_degree: int, width: int, height: int):
        self.lib = library()
        self.handle = ctypes.c_void_p()
        check(self.lib.gs_hip_workspace_create(ctypes.byref(self.handle), capacity, sh_degree, width, height))
        self.capacity, self.sh_degree, self.width, self.height = capacity, sh_degree, width, height
        self.sh_coeffs = 0 if sh_degree == 0 else (sh_degree + 1) ** 2 - 1
        self.tiles_x = (width + TILE - 1) // TILE
        self.tiles_y = (height + TILE - 1) // TILE

    def close(self) -> None:
        if self.handle:
            check(self.lib.gs_hip_workspace_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buffer(self, which: int, view: int = 0):
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t()
        check(self.lib.gs_hip_workspace_view_buffer(self.handle, view, which, ctypes.byref(ptr),
                                                    ctypes.byref(size)))
        return ptr, size.value

    def reserve_pairs(self, pairs: int) -> None:
        check(self.lib.gs_hip_workspace_reserve_pairs(self.handle, pairs))

    def set_views(self, count: int) -> None:
        check(self.lib.gs_hip_workspace_views(self.handle, count))

    def train_views(self, frames, targets=None, want_losses: bool = True):
        """One batched optimizer step over len(frames) views; targets: device
        pointers (or None).  Returns (pairs per view, [3 losses per view])."""
        n = len(frames)
        arr = (Frame * n)(*frames)
        tarr = None
        if targets is not None:
            tarr = (ctypes.c_void_p * n)(*[t if isinstance(t, int) else (t.value if t else None) for t in targets])
        pairs = (ctypes.c_uint64 * n)()
        losses = (ctypes.c_double * (3 * n))()
        check(self.lib.gs_hip_train_views(self.handle, arr, n, tarr, pairs, losses if want_losses else None))
        ls = [tuple(losses[3 * v:3 * v + 3]) for v in range(n)] if want_losses else None
        return list(pairs), ls

    def download(self, which: int, dtype, count: int | None = None, view: int = 0) -> np.ndarray:
        ptr, size = self.buffer(which, view)
        item = np.dtype(dtype).itemsize
        n = size // item if count is None else count
        out = np.empty(n, dtype)
        if n:
            check(self.lib.gs_hip_download(out.ctypes.data_as(ctypes.c_void_p), ptr, n * item))
        return out

    def upload(self, which: int, data: np.ndarray) -> None:
        ptr, size = self.buffer(which)
        data = np.ascontiguousarray(data)
        if data.nbytes > size:
            raise ValueError("upload larger than the buffer")
        check(self.lib.gs_hip_upload(ptr, data.ctypes.data_as(ctypes.c_void_p), data.nbytes))

    def load(self, params: np.ndarray, sh_rest: np.ndarray | None) -> None:
        """params [n,16] float32; sh_rest [n, coeffs, 3] float32 (or None)."""
        params = np.ascontiguousarray(params, np.float32).reshape(-1, 16)
        n = params.shape[0]
        if sh_rest is None or sh_rest.size == 0:
            check(self.lib.gs_hip_workspace_load(self.handle, _fptr(params), None, n, 0))
        else:
            sh_rest = np.ascontiguousarray(sh_rest, np.float32).reshape(n, -1, 3)
            check(self.lib.gs_hip_workspace_load(self.handle, _fptr(params), _fptr(sh_rest), n, sh_rest.shape[1]))

    def store(self, count: int):
        params = np.empty((count, 16), np.float32)
        sh = np.empty((count, self.sh_coeffs, 3), np.float32)
        check(self.lib.gs_hip_workspace_store(self.handle, _fptr(params), _fptr(sh) if self.sh_coeffs else None,
                                              count))
        return params, sh

    def project(self, frame: Frame) -> None:
        check(self.lib.gs_hip_project(self.handle, ctypes.byref(frame)))

    def bin(self, frame: Frame) -> int:
        pairs = ctypes.c_uint64()
        check(self.lib.gs_hip_bin(self.handle, ctypes.byref(frame), ctypes.byref(pairs)))
        return pairs.value

    def forward(self, frame: Frame) -> None:
        check(self.lib.gs_hip_forward(self.handle, ctypes.byref(frame)))

    def set_target(self, rgba8: np.ndarray) -> None:
        rgba8 = np.ascontiguousarray(rgba8, np.uint8)
        assert rgba8.size == self.width * self.height * 4
        check(self.lib.gs_hip_set_target(self.handle, rgba8.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def set_target_device(self, device_ptr) -> None:
        check(self.lib.gs_hip_set_target_device(self.handle, device_ptr))

    def loss(self, frame: Frame, want_losses: bool = True):
        out = (ctypes.c_double * 3)()
        check(self.lib.gs_hip_loss(self.handle, ctypes.byref(frame), out if want_losses else None))
        return tuple(out) if want_losses else None

    def backward(self, frame: Frame) -> None:
        check(self.lib.gs_hip_backward(self.handle, ctypes.byref(frame)))

    def adam(self, frame: Frame) -> None:
        check(self.lib.gs_hip_adam(self.handle, ctypes.byref(frame)))

    def densify(self, frame: Frame):
        slots = ctypes.c_uint32(frame.slots)
        counts = (ctypes.c_uint32 * 3)()
        check(self.lib.gs_hip_densify(self.handle, ctypes.byref(frame), ctypes.byref(slots), counts))
        return slots.value, tuple(counts)

    def train_step(self, frame: Frame, want_losses: bool = True):
        pairs = ctypes.c_uint64()
        out = (ctypes.c_double * 3)()
        check(self.lib.gs_hip_train_step(self.handle, ctypes.byref(frame), ctypes.byref(pairs),
                                         out if want_losses else None))
        return pairs.value, (tuple(out) if want_losses else None)

    def pose_gradient(self, view: int = 0) -> np.ndarray:
        """dL/dd xyz, dL/dw xyz of the view's latest step with FLAG_POSE_GRAD."""
        out = (ctypes.c_double * 6)()
        check(self.lib.gs_hip_pose_gradient(self.handle, view, out))
        return np.array(out[:], np.float64)

    def exposure_sums(self, view: int = 0) -> np.ndarray:
        """The gain fit sums of the view's latest step with FLAG_EXPOSURE:
        sum C^2 rgb, fitted pixels, sum C photo rgb, subject pixels."""
        out = (ctypes.c_double * 8)()
        check(self.lib.gs_hip_exposure_sums(self.handle, view, out))
        return np.array(out[:], np.float64)

    def load_spz(self, data: bytes):
        """Decodes an SPZ file into slots [0, n) on the device; returns (n, SH degree, antialiased)."""
        n, deg, aa = ctypes.c_uint32(), ctypes.c_int(), ctypes.c_int()
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        check(self.lib.gs_hip_workspace_load_spz(self.handle, buf, len(data), ctypes.byref(n), ctypes.byref(deg),
                                                 ctypes.byref(aa)))
        return n.value, deg.value, bool(aa.value)

    def export_ply_rows(self, slots: int, rdf: bool = True) -> np.ndarray:
        """The alive slots' 3DGS PLY vertex rows, float32 [alive, 17 + 3 K]."""
        count = ctypes.c_uint32()
        check(self.lib.gs_hip_export_ply(self.handle, slots, int(rdf), None, 0, ctypes.byref(count)))
        per = 17 + 3 * self.sh_coeffs
        out = np.empty((count.value, per), np.float32)
        check(self.lib.gs_hip_export_ply(self.handle, slots, int(rdf), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                         out.nbytes, ctypes.byref(count)))
        return out

    def render(self, frame: Frame) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.uint8)
        check(self.lib.gs_hip_render(self.handle, ctypes.byref(frame),
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def view_render(self, frame: Frame, clear=(0.0, 0.0, 0.0, 0.0), antialiased: bool = False):
        """(RGBA8 [H, W, 4], splats drawn): the viewer renderer's image of the
        loaded model (back-to-front quads into an 8-bit target)."""
        out = np.empty((self.height, self.width, 4), np.uint8)
        cl = (ctypes.c_float * 4)(*[float(c) for c in clear])
        vis = ctypes.c_uint32()
        check(self.lib.gs_hip_view_render(self.handle, ctypes.byref(frame), cl, VIEW_ANTIALIASED if antialiased else 0,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(vis)))
        return out, vis.value

    def tuning(self) -> Tuning:
        t = Tuning()
        check(self.lib.gs_hip_workspace_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def set_tuning(self, raster_groups_per_cu=None, depth_digit_bits=None, tile_digit_bits=None, graph=None,
                   bin_shift=None, stored_state=None) -> Tuning:
        t = self.tuning()
        for name, value in (("raster_groups_per_cu", raster_groups_per_cu), ("depth_digit_bits", depth_digit_bits),
                            ("tile_digit_bits", tile_digit_bits), ("graph", graph), ("bin_shift", bin_shift),
                            ("stored_state", stored_state)):
            if value is not None:
                setattr(t, name, int(value))
        check(self.lib.gs_hip_workspace_set_tuning(self.handle, ctypes.byref(t)))
        return t

    def autotune(self, frames, targets=None, repeats: int = 3):
        """Times the candidate launch shapes on these views (model untouched);
        keeps and returns (the fastest Tuning, its ms per step)."""
        count = len(frames)
        arr = (Frame * count)(*frames)
        tp = None
        if targets is not None:
            tp = (ctypes.c_void_p * count)(*[t if isinstance(t, int) else (t.value if t else None)
                                             for t in targets])
        best = Tuning()
        ms = ctypes.c_double()
        check(self.lib.gs_hip_autotune(self.handle, arr, count, tp, repeats, ctypes.byref(best), ctypes.byref(ms)))
        return best, ms.value

    def overflows(self) -> int:
        n = ctypes.c_uint64()
        check(self.lib.gs_hip_workspace_overflows(self.handle, ctypes.byref(n)))
        return n.value

    def timings(self, enable: bool = True):
        cap = 1 << 16
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        count = ctypes.c_int()
        check(self.lib.gs_hip_workspace_timings(self.handle, int(enable), names, ms, cap, ctypes.byref(count)))
        return [(names[i].decode(), ms[i]) for i in range(count.value)]


class Batch:
    """Several workspaces trained by one launch per stage (gs_hip_batch); the
    workspaces must outlive the batch."""

    def __init__(self, workspaces):
        self.lib = library()
        self.scenes = list(workspaces)
        self.handle = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(self.scenes))(*[w.handle.value for w in self.scenes])
        check(self.lib.gs_hip_batch_create(ctypes.byref(self.handle), arr, len(self.scenes)))

    def close(self) -> None:
        if self.handle:
            check(self.lib.gs_hip_batch_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, frames_per_scene, targets_per_scene=None, want_losses: bool = True):
        """One optimizer step of every scene: frames_per_scene[s] is the list of
        scene s's frames, targets_per_scene[s] their device RGBA8 pointers (or
        None).  Returns (pairs per view, [3 losses per view]), views in order."""
        frames = [f for fs in frames_per_scene for f in fs]
        n = len(frames)
        counts = (ctypes.c_uint32 * len(frames_per_scene))(*[len(fs) for fs in frames_per_scene])
        arr = (Frame * n)(*frames)
        tarr = None
        if targets_per_scene is not None:
            flat = []
            for fs, ts in zip(frames_per_scene, targets_per_scene):
                flat += list(ts) if ts is not None else [None] * len(fs)
            tarr = (ctypes.c_void_p * n)(*[t if isinstance(t, int) else (t.value if t else None) for t in flat])
        pairs = (ctypes.c_uint64 * n)()
        losses = (ctypes.c_double * (3 * n))()
        check(self.lib.gs_hip_batch_train(self.handle, arr, counts, tarr, pairs, losses if want_losses else None))
        ls = [tuple(losses[3 * v:3 * v + 3]) for v in range(n)] if want_losses else None
        return list(pairs), ls

    def tuning(self) -> Tuning:
        t = Tuning()
        check(self.lib.gs_hip_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def set_tuning(self, **kw) -> Tuning:
        t = self.tuning()
        for name, value in kw.items():
            if value is not None:
                setattr(t, name, int(value))
        check(self.lib.gs_hip_batch_set_tuning(self.handle, ctypes.byref(t)))
        return t

    def timings(self, enable: bool = True):
        cap = 1 << 16
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        count = ctypes.c_int()
        check(self.lib.gs_hip_batch_timings(self.handle, int(enable), names, ms, cap, ctypes.byref(count)))
        return [(names[i].decode(), ms[i]) for i in range(count.value)]


def spz_info(data: bytes):
    """(count, SH degree, antialiased) from an SPZ file's header."""
    n, deg, aa = ctypes.c_uint32(), ctypes.c_int(), ctypes.c_int()
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    check(library().gs_hip_spz_info(buf, len(data), ctypes.byref(n), ctypes.byref(deg), ctypes.byref(aa)))
    return n.value, deg.value, bool(aa.value)


def synchronize() -> None:
    check(library().gs_hip_synchronize())


def loss_noalloc(frame: Frame, image_ptr, target_ptr, dldc_ptr, want_losses: bool = True, stream=None):
    """gs_hip_loss_noalloc: the loss stage on caller-owned device buffers
    (float4 image, RGBA8 target, float4 dL/dC out), allocating nothing.
    Returns the three loss sums (or None)."""
    losses = (ctypes.c_double * 3)()
    check(library().gs_hip_loss_noalloc(ctypes.byref(frame), _vptr(image_ptr), _vptr(target_ptr), _vptr(dldc_ptr),
                                        _vptr(stream), losses if want_losses else None))
    return tuple(losses) if want_losses else None


# ---- the trainer handle (NativeTrainer.kt:8-47) ---------------------------------
class TrainerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("width", ctypes.c_uint32), ("iterations", ctypes.c_uint32),
                ("max_gaussians", ctypes.c_uint32), ("sh_degree", ctypes.c_int), ("masked", ctypes.c_int),
                ("refine_poses", ctypes.c_int), ("refine_exposure", ctypes.c_int),
                ("views_per_step", ctypes.c_uint32), ("seed", ctypes.c_uint32)]


class TrainerInfo(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_uint32), ("holdout", ctypes.c_uint32), ("skipped", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("masks", ctypes.c_int),
                ("depth", ctypes.c_int), ("points", ctypes.c_uint32), ("init_points", ctypes.c_uint64),
                ("init_gaussians", ctypes.c_uint32), ("init_source", ctypes.c_char * 16)]


class TrainerStats(ctypes.Structure):
    _fields_ = [("step", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("gaussians", ctypes.c_uint32),
                ("optimizer_steps", ctypes.c_uint32), ("sh_degree", ctypes.c_int),
                ("densify_events", ctypes.c_uint32), ("overflow_steps", ctypes.c_uint32), ("pairs", ctypes.c_uint64),
                ("l1", ctypes.c_double), ("psnr", ctypes.c_double), ("ssim", ctypes.c_double),
                ("loss", ctypes.c_double), ("step_ms", ctypes.c_double), ("finished", ctypes.c_int)]


class PassTimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("project", "sort", "rasterize", "ssim", "backward", "optimize")] + \
        [("valid", ctypes.c_int)]


class TrainerEval(ctypes.Structure):
    _fields_ = [("psnr", ctypes.c_double), ("ssim", ctypes.c_double), ("l1", ctypes.c_double),
                ("psnr_unaligned", ctypes.c_double), ("views", ctypes.c_uint32), ("aligned", ctypes.c_int)]


class NativeTrainer:
    """A training session on a nerfstudio folder: the reference's
    NativeTrainer (prepare / step / preview / evaluate / exportPly / release)
    over the gs_hip_trainer_* calls."""

    def __init__(self, dataset_dir: str, width: int = 512, iterations: int = 3000, max_gaussians: int = 300000,
                 sh_degree: int = 3, masked: bool = True, refine_poses: bool = False, refine_exposure: bool = False,
                 views_per_step: int = 1, seed: int = 1, device: int = 0):
        self.lib = library()
        o = TrainerOptions(device, width, iterations, max_gaussians, sh_degree, int(masked), int(refine_poses),
                           int(refine_exposure), views_per_step, seed)
        self.info = TrainerInfo()
        h = ctypes.c_void_p()
        check(self.lib.gs_hip_trainer_create(os.fsencode(dataset_dir), ctypes.byref(o), ctypes.byref(h),
                                             ctypes.byref(self.info)))
        self.handle = h

    def step(self, count: int, profile: bool = False):
        """(steps run, TrainerStats, PassTimes or None)."""
        done, st, times = ctypes.c_uint32(), TrainerStats(), PassTimes()
        check(self.lib.gs_hip_trainer_steps(self.handle, count, int(profile), ctypes.byref(done), ctypes.byref(st),
                                            ctypes.byref(times)))
        return done.value, st, (times if times.valid else None)

    def stats(self) -> TrainerStats:
        st = TrainerStats()
        check(self.lib.gs_hip_trainer_get_stats(self.handle, ctypes.byref(st)))
        return st

    def evaluate(self) -> TrainerEval:
        e = TrainerEval()
        check(self.lib.gs_hip_trainer_evaluate(self.handle, ctypes.byref(e)))
        return e

    def preview(self, index: int) -> tuple[np.ndarray, np.ndarray]:
        """(render, photo), each uint8[H][W][4], over the preview background."""
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.gs_hip_trainer_preview(self.handle, index, None, 0, ctypes.byref(w), ctypes.byref(h)))
        out = np.zeros((2, h.value, w.value, 4), np.uint8)
        check(self.lib.gs_hip_trainer_preview(self.handle, index, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                              out.nbytes, ctypes.byref(w), ctypes.byref(h)))
        return out[0], out[1]

    def export_ply(self, path: str) -> tuple[int, int]:
        """(gaussians, file bytes)."""
        n, b = ctypes.c_uint32(), ctypes.c_uint64()
        check(self.lib.gs_hip_trainer_export_ply(self.handle, os.fsencode(path), ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def release(self) -> None:
        if getattr(self, "handle", None):
            check(self.lib.gs_hip_trainer_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _vptr(p):
    if p is None or isinstance(p, int):
        return p
    return p.value if isinstance(p, ctypes.c_void_p) else int(p)


def device_alloc(nbytes: int) -> int:
    ptr = ctypes.c_void_p()
    check(library().gs_hip_malloc(ctypes.byref(ptr), nbytes))
    return ptr.value


def device_free(ptr: int) -> None:
    check(library().gs_hip_free(ptr))


def device_copy(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy by the library's streaming kernel (k_copy16)."""
    check(library().gs_hip_copy(dst, src, nbytes))


# ---- seeded synthetic scenes (the benchmark configurations) -----------------
def synthetic_scene(n: int, sh_degree: int, seed: int):
    """Gaussians of the benchmark configurations: means uniform in the unit
    ball, log-scales N(-4, 0.5), quaternions normalise(N(0,1)^4) (w, x, y, z),
    opacity logits N(0,1), DC uniform[0,1], SH rest N(0, 0.2).  Returns
    (params [n,16] float32, sh_rest [n, coeffs, 3] float32)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(size=(n, 1)) ** (1.0 / 3.0)
    pos = d * r
    log_scale = rng.normal(-4.0, 0.5, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    logit = rng.normal(size=(n, 1))
    dc = rng.uniform(size=(n, 3))
    coeffs = 0 if sh_degree == 0 else (sh_degree + 1) ** 2 - 1
    sh = rng.normal(0.0, 0.2, size=(n, coeffs, 3)).astype(np.float32)
    params = np.zeros((n, 16), np.float32)
    params[:, 0:3] = pos
    params[:, 3:4] = logit
    params[:, 4:7] = log_scale
    params[:, 7] = 1.0
    params[:, 8:12] = q
    params[:, 12:15] = dc
    return params, sh


def orbit_cameras(count: int, width: int, height: int, radius: float = 3.0, height_y: float = 0.5):
    """`count` cameras evenly spaced on a circle around the origin."""
    cams = []
    for k in range(count):
        th = 2.0 * math.pi * k / count
        cams.append(look_at_camera((radius * math.sin(th), height_y, radius * math.cos(th)), (0, 0, 0), width,
                                   height))
    return cams


def jitter_cameras(count: int, width: int, height: int, seed: int):
    """Cameras at seeded azimuth U[0, 2pi), radius U[2.5, 3.5], height U[-0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(count):
        th = rng.uniform(0.0, 2.0 * math.pi)
        r = rng.uniform(2.5, 3.5)
        h = rng.uniform(-0.5, 1.5)
        cams.append(look_at_camera((r * math.sin(th), h, r * math.cos(th)), (0, 0, 0), width, height))
    return cams


class DeviceBuffer:
    """A device allocation (gs_hip_malloc) holding a copy of a numpy array."""

    def __init__(self, data: np.ndarray):
        data = np.ascontiguousarray(data)
        self.nbytes = data.nbytes
        self.ptr = ctypes.c_void_p()
        check(library().gs_hip_malloc(ctypes.byref(self.ptr), self.nbytes))
        check(library().gs_hip_upload(self.ptr, data.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def free(self) -> None:
        if self.ptr:
            check(library().gs_hip_free(self.ptr))
            self.ptr = ctypes.c_void_p()
