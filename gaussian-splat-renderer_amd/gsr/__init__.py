"""Python binding of libgsr.so, the MI355X (gfx950) splat-tiles forward renderer.

The C ABI (include/gsr.h) is the product; this module is a thin ctypes layer over it that
mirrors the reference's SplatTileRenderer host interface (render/splat_tile_renderer.h:143-196):

    ctx = gsr.Context(0)
    scene = gsr.Scene(ctx, pos_alpha, scales, rotations, sh, sh_degree=3)        # GpuScene::upload
    r = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8")                  # init
    img = r.render_to_host(cam)                                                   # render + readBack
    keys, depth, values, ranges, control = r.read_back()                          # readBack
    splats, counts = r.read_back_splats()                                         # readBackSplats

Device memory is owned by the library (no torch on the path). The library must be built
(``__graft_entry__.build()``); importing this module without it raises, there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from typing import Optional

import numpy as np

from . import camera, scenes  # noqa: F401  (re-exported helpers)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsr.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libgsr.so is not built ({LIB_PATH}); run __graft_entry__.build()")
_lib = ctypes.CDLL(LIB_PATH)

# --- enums (gsr.h) ----------------------------------------------------------------------------
SCENE_F32_SOA, SCENE_F16 = 0, 1
BOUNDS = {"rect": 0, "tight": 1, "cutoff-only": 2, "cutoff": 2}
TARGETS = {"rgba8": 0, "rgba32f": 1}
SORT_ALGORITHMS = {"reduce-then-scan": 0, "onesweep": 1}
RENDER_LOOPS = {"culled": 0, "divergent": 1, "predicated": 2}
SORT_MODES = {"binning": 0, "global": 1}
RENDER_SPLATS = {"full": 0, "p32": 1, "p24": 2}
TILE_STATS = ("list_length", "fetched", "evaluated", "blended", "terminated", "lane_slots", "max_depth", "pixels")
STAGES = ("preprocess", "scan", "depth_sort", "duplicate", "sort", "ranges", "render", "total")
CULL_VISIBLE, CULL_NEAR, CULL_SINGULAR, CULL_OFFSCREEN, CULL_LOW_OPACITY = 0, 1, 2, 3, 4


class GsrError(RuntimeError):
    pass


class Camera(ctypes.Structure):
    """gsr_camera = SplatCamera (splat_math.h:43-50)."""

    _fields_ = [
        ("view_inria", ctypes.c_float * 16),
        ("position", ctypes.c_float * 3),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("tan_half_fov_x", ctypes.c_float),
        ("tan_half_fov_y", ctypes.c_float),
        ("focal_x", ctypes.c_float),
        ("focal_y", ctypes.c_float),
        ("tiles_x", ctypes.c_uint32),
        ("tiles_y", ctypes.c_uint32),
    ]

    @property
    def view(self) -> np.ndarray:
        return np.array(self.view_inria[:], dtype=np.float32)

    @property
    def tiles(self) -> int:
        return int(self.tiles_x) * int(self.tiles_y)


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("sh_degree", ctypes.c_int32),
        ("format", ctypes.c_int32),
        ("device_pointers", ctypes.c_int32),
        ("pos_alpha", ctypes.c_void_p),
        ("scales", ctypes.c_void_p),
        ("rotations", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("geometry", ctypes.c_void_p),
        ("color", ctypes.c_void_p),
    ]


class Settings(ctypes.Structure):
    _fields_ = [
        ("bounds", ctypes.c_int32),
        ("budget_mb", ctypes.c_uint32),
        ("max_sh_degree", ctypes.c_int32),
        ("format", ctypes.c_int32),
        ("profile_stages", ctypes.c_int32),
        ("render_pixels_per_thread", ctypes.c_int32),
        ("render_exact_alpha", ctypes.c_int32),
        ("sort_algorithm", ctypes.c_int32),
        ("render_loop", ctypes.c_int32),
        ("sort_mode", ctypes.c_int32),
        ("full_key_sort", ctypes.c_int32),
        ("render_splat", ctypes.c_int32),
        ("sort_window", ctypes.c_uint32),
        ("adaptive_list", ctypes.c_int32),
        ("adaptive_initial_entries", ctypes.c_uint32),
        ("frame_stats", ctypes.c_int32),
        ("differentiable", ctypes.c_int32),
        ("tile_stats", ctypes.c_int32),
        ("tile_size", ctypes.c_int32),
        ("render_batch", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 1),
    ]


class Control(ctypes.Structure):
    """gsr_control = TileControl (splat_tile_renderer.h:116-128), 80 bytes."""

    _fields_ = [
        ("gaussian_count", ctypes.c_uint32),
        ("visible", ctypes.c_uint32),
        ("duplicates", ctypes.c_uint32),
        ("entries", ctypes.c_uint32),
        ("cull_counts", ctypes.c_uint32 * 8),
        ("overflow", ctypes.c_uint32),
        ("ranges_groups", ctypes.c_uint32 * 3),
        ("coverage_mismatch", ctypes.c_uint32),
        ("sort_timeouts", ctypes.c_uint32),
        ("pad", ctypes.c_uint32 * 2),
    ]

    def as_dict(self) -> dict:
        return {
            "gaussian_count": self.gaussian_count,
            "visible": self.visible,
            "duplicates": self.duplicates,
            "entries": self.entries,
            "cull_counts": list(self.cull_counts),
            "overflow": self.overflow,
            "coverage_mismatch": self.coverage_mismatch,
            "sort_timeouts": self.sort_timeouts,
        }


assert ctypes.sizeof(Control) == 80
assert ctypes.sizeof(Camera) == 108

# 48-byte ProjectedSplat as a numpy record.
SPLAT_DTYPE = np.dtype(
    [
        ("center", np.float32, 2),
        ("depth", np.float32),
        ("cull_reason", np.uint32),
        ("conic_opacity", np.float32, 4),
        ("color", np.float32, 3),
        ("radius", np.float32),
    ]
)
assert SPLAT_DTYPE.itemsize == 48

_vp = ctypes.c_void_p
_u32, _i32, _u64, _f32, _sz = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_float, ctypes.c_size_t


def _sig(name, restype, argtypes):
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


_lib_last_error = _sig("gsr_last_error", ctypes.c_char_p, [])
gsr_version = _sig("gsr_version", ctypes.c_char_p, [])
_context_create = _sig("gsr_context_create", _i32, [_i32, ctypes.POINTER(_vp)])
_context_sync = _sig("gsr_context_synchronize", _i32, [_vp])
_context_destroy = _sig("gsr_context_destroy", None, [_vp])
_buffer_create = _sig("gsr_buffer_create", _i32, [_vp, _sz, ctypes.POINTER(_vp)])
_buffer_ptr = _sig("gsr_buffer_device_ptr", _vp, [_vp])
_buffer_read = _sig("gsr_buffer_read", _i32, [_vp, _vp, _sz])
_buffer_write = _sig("gsr_buffer_write", _i32, [_vp, _vp, _sz])
_buffer_destroy = _sig("gsr_buffer_destroy", None, [_vp])
_scene_create = _sig("gsr_scene_create", _i32, [_vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)])
_scene_destroy = _sig("gsr_scene_destroy", None, [_vp])
_encode_f16 = _sig("gsr_encode_f16", _i32, [_u32, _i32, _vp, _vp, _vp, _vp, _vp, _vp])
_camera_make = _sig("gsr_camera_make", _i32, [_vp, _u32, _u32, _f32, _f32, ctypes.POINTER(Camera)])
_camera_from_viewer = _sig("gsr_camera_from_viewer", _i32, [_vp, _vp, _vp, _u32, _u32, ctypes.POINTER(Camera)])
_renderer_create = _sig("gsr_renderer_create", _i32, [_vp, _vp, ctypes.POINTER(Settings), ctypes.POINTER(_vp)])
_renderer_destroy = _sig("gsr_renderer_destroy", None, [_vp])
_renderer_capacity = _sig("gsr_renderer_capacity", _u32, [_vp])
_renderer_total_bytes = _sig("gsr_renderer_total_bytes", _u64, [_vp])
_render = _sig("gsr_render", _i32, [_vp, ctypes.POINTER(Camera), _vp, _vp])
_read_control = _sig("gsr_read_control", _i32, [_vp, ctypes.POINTER(Control)])
_read_lists = _sig("gsr_read_lists", _i32, [_vp, _vp, _vp, _vp, _vp, _u32])
_read_splats = _sig("gsr_read_splats", _i32, [_vp, _vp, _vp])
_read_stage_times = _sig("gsr_read_stage_times", _i32, [_vp, _vp, _u32])
_read_tile_stats = _sig("gsr_read_tile_stats", _i32, [_vp, _vp, _u32])
_read_kernel_times = _sig("gsr_read_kernel_times", _i32, [_vp, _vp, _vp, _u32, _vp])
_read_packed = _sig("gsr_read_packed_splats", _i32, [_vp, _vp])
_renderer_max_capacity = _sig("gsr_renderer_max_capacity", _u32, [_vp])
_renderer_list_resizes = _sig("gsr_renderer_list_resizes", _u32, [_vp])
_renderer_last_stats = _sig("gsr_renderer_last_stats", _i32, [_vp, _vp, _vp])
_sort_pairs = _sig("gsr_sort_pairs", _i32, [_vp, _vp, _vp, _u32, _vp, _u32, _i32, _vp])
_sort_pairs_noalloc = _sig("gsr_sort_pairs_noalloc", _i32, [_vp, _vp, _vp, _u32, _vp, _u32, _vp])


class PlySceneStruct(ctypes.Structure):
    """gsr_ply_scene (include/gsr.h)."""

    _fields_ = [
        ("count", ctypes.c_uint32),
        ("sh_degree", ctypes.c_int32),
        ("pos_alpha", ctypes.POINTER(ctypes.c_float)),
        ("scales", ctypes.POINTER(ctypes.c_float)),
        ("rotations", ctypes.POINTER(ctypes.c_float)),
        ("sh", ctypes.POINTER(ctypes.c_float)),
        ("raw_opacity", ctypes.POINTER(ctypes.c_float)),
        ("raw_log_scales", ctypes.POINTER(ctypes.c_float)),
        ("raw_rotations", ctypes.POINTER(ctypes.c_float)),
        ("non_finite_values", ctypes.c_uint64),
        ("zero_quaternions", ctypes.c_uint32),
        ("warning_count", ctypes.c_uint32),
        ("warnings", ctypes.POINTER(ctypes.c_char_p)),
        ("comment_count", ctypes.c_uint32),
        ("comments", ctypes.POINTER(ctypes.c_char_p)),
        ("internal", ctypes.c_void_p),
    ]


_ply_load = _sig("gsr_ply_load", _i32, [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(PlySceneStruct))])
_ply_load_memory = _sig("gsr_ply_load_memory", _i32,
                        [_vp, ctypes.c_uint64, ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(PlySceneStruct))])
_ply_free = _sig("gsr_ply_free", None, [ctypes.POINTER(PlySceneStruct)])
_exclusive_scan = _sig("gsr_exclusive_scan", _i32, [_vp, _vp, _vp, _u32, _vp, _vp])
_batch_create = _sig("gsr_batch_create", _i32, [_vp, _u32, ctypes.POINTER(_vp)])
_batch_destroy = _sig("gsr_batch_destroy", None, [_vp])
_batch_branches = _sig("gsr_batch_branches", _u32, [_vp])
_batch_render = _sig("gsr_batch_render", _i32, [_vp, _vp, _u32, _vp, _vp, _vp])
_batch_capture = _sig("gsr_batch_capture", _i32, [_vp, _vp, _u32, _vp, _vp])
_batch_replay = _sig("gsr_batch_replay", _i32, [_vp, _vp])
_batch_set_persistent = _sig("gsr_batch_set_persistent", _i32, [_vp, _i32])
_cm_param_count = _sig("gsr_colour_model_param_count", _u32, [])
_cm_create = _sig("gsr_colour_model_create", _i32, [_vp, _vp, _u32, ctypes.POINTER(_vp)])
_cm_destroy = _sig("gsr_colour_model_destroy", None, [_vp])
_cm_train = _sig("gsr_colour_model_train", _i32, [_vp, _u32, _u32, _f32, _f32, _vp])
_cm_loss = _sig("gsr_colour_model_loss", _i32, [_vp, _u32, _u32, _vp])
_cm_gradients = _sig("gsr_colour_model_gradients", _i32, [_vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp])
_cm_get = _sig("gsr_colour_model_get", _i32, [_vp, _vp, _vp])
_cm_set = _sig("gsr_colour_model_set", _i32, [_vp, _vp, _vp])
_cm_evaluate = _sig("gsr_colour_model_evaluate", _i32, [_vp, _vp, _vp, _u32, _vp])
_cm_fallback = _sig("gsr_colour_model_fallback", _i32, [_vp, _f32, _vp, _vp])
_renderer_set_colour_model = _sig("gsr_renderer_set_colour_model", _i32, [_vp, _vp])


class Gradients(ctypes.Structure):
    """gsr_gradients: device pointers (float32), each may be null."""

    _fields_ = [("means", _vp), ("log_scales", _vp), ("quats", _vp), ("opacity_logits", _vp), ("sh", _vp)]


_backward = _sig("gsr_backward", _i32, [_vp, _vp, ctypes.POINTER(Gradients), _vp])

# Every symbol include/gsr.h declares (checked by the CPU tests).
EXPORTED = (
    "gsr_last_error gsr_version gsr_context_create gsr_context_synchronize gsr_context_destroy "
    "gsr_buffer_create gsr_buffer_device_ptr gsr_buffer_read gsr_buffer_write gsr_buffer_destroy "
    "gsr_scene_create gsr_scene_destroy gsr_encode_f16 gsr_camera_make gsr_camera_from_viewer "
    "gsr_renderer_create gsr_renderer_destroy gsr_renderer_capacity gsr_renderer_total_bytes gsr_render "
    "gsr_renderer_max_capacity gsr_renderer_list_resizes gsr_renderer_last_stats "
    "gsr_read_control gsr_read_lists gsr_read_splats gsr_read_packed_splats gsr_read_stage_times gsr_read_tile_stats gsr_read_kernel_times gsr_sort_pairs "
    "gsr_sort_pairs_noalloc gsr_exclusive_scan "
    "gsr_batch_create gsr_batch_destroy gsr_batch_branches gsr_batch_render gsr_batch_capture gsr_batch_replay "
    "gsr_batch_set_persistent "
    "gsr_colour_model_param_count gsr_colour_model_create gsr_colour_model_destroy gsr_colour_model_train "
    "gsr_colour_model_loss gsr_colour_model_gradients gsr_colour_model_get gsr_colour_model_set "
    "gsr_colour_model_evaluate gsr_colour_model_fallback gsr_renderer_set_colour_model gsr_backward "
    "gsr_ply_load gsr_ply_load_memory gsr_ply_free"
).split()


def _check(status: int, what: str) -> None:
    if status != 0:
        raise GsrError(f"{what} failed (status {status}): {_lib_last_error().decode()}")


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def _f32c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# --- camera (host) ------------------------------------------------------------------------------


def camera_make(view_inria, width: int, height: int, tan_half_fov_x: float, tan_half_fov_y: float) -> Camera:
    """makeSplatCamera (splat_math.cpp:9-27)."""
    v = _f32c(view_inria).reshape(16)
    cam = Camera()
    _check(_camera_make(_ptr(v), width, height, float(tan_half_fov_x), float(tan_half_fov_y), ctypes.byref(cam)),
           "gsr_camera_make")
    return cam


def camera_from_viewer(view, proj, model, width: int, height: int) -> Camera:
    """splatCameraFromViewer (splat_math.cpp:29-42); matrices column-major, flattened m[col*4+row]."""
    v, p, m = (_f32c(x).reshape(16) for x in (view, proj, model))
    cam = Camera()
    _check(_camera_from_viewer(_ptr(v), _ptr(p), _ptr(m), width, height, ctypes.byref(cam)), "gsr_camera_from_viewer")
    return cam


def encode_f16(pos_alpha, scales, rotations, sh, sh_degree: int):
    """The fp16 scene records (scene_compression.cpp:34-49): geometry (N, 8) u32, colour (N, H) u16."""
    pa, sc, ro, s = (_f32c(x) for x in (pos_alpha, scales, rotations, sh))
    n = pa.shape[0]
    K = (sh_degree + 1) ** 2
    halves = (K * 3 + 7) // 8 * 8
    geom = np.zeros((n, 8), np.uint32)
    color = np.zeros((n, halves), np.uint16)
    _check(_encode_f16(n, sh_degree, _ptr(pa), _ptr(sc), _ptr(ro), _ptr(s), _ptr(geom), _ptr(color)), "gsr_encode_f16")
    return geom, color


# --- PLY scenes (host) --------------------------------------------------------------------------

GSR_ERROR_FORMAT, GSR_ERROR_IO = 5, 6


class PlyError(GsrError):
    """A bad .ply file (the reference's PlyError): the message names the problem."""


class PlyScene:
    """loadGaussianPly (scene/ply_loader.h:86-95): the activated f32 SoA arrays (the inputs of
    Scene(..., fmt="f32")), the raw file values, and the loader's counters, warnings, comments."""

    def __init__(self, p: PlySceneStruct):
        n, K = int(p.count), (int(p.sh_degree) + 1) ** 2

        def arr(ptr, cols):
            return np.ctypeslib.as_array(ptr, shape=(n * cols,)).reshape(n, cols).copy() if n else \
                np.zeros((0, cols), np.float32)

        self.count, self.sh_degree = n, int(p.sh_degree)
        self.pos_alpha, self.scales, self.rotations = arr(p.pos_alpha, 4), arr(p.scales, 4), arr(p.rotations, 4)
        self.sh = arr(p.sh, K * 3)
        self.raw_opacity = arr(p.raw_opacity, 1).reshape(n)
        self.raw_log_scales, self.raw_rotations = arr(p.raw_log_scales, 3), arr(p.raw_rotations, 4)
        self.non_finite_values, self.zero_quaternions = int(p.non_finite_values), int(p.zero_quaternions)
        self.warnings = [p.warnings[i].decode() for i in range(p.warning_count)]
        self.comments = [p.comments[i].decode() for i in range(p.comment_count)]

    @property
    def K(self) -> int:
        return (self.sh_degree + 1) ** 2


def _ply_result(status: int, out) -> PlyScene:
    if status != 0:
        msg = _lib_last_error().decode()
        raise (PlyError if status == GSR_ERROR_FORMAT else GsrError)(msg)
    try:
        return PlyScene(out.contents)
    finally:
        _ply_free(out)


def load_ply(path: str) -> PlyScene:
    """Loads and activates a binary little-endian INRIA 3DGS .ply file (host)."""
    out = ctypes.POINTER(PlySceneStruct)()
    return _ply_result(_ply_load(os.fsencode(path), ctypes.byref(out)), out)


def load_ply_bytes(data: bytes, name: str = "<memory>") -> PlyScene:
    """The same from bytes in memory; `name` prefixes the error messages."""
    buf = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = ctypes.POINTER(PlySceneStruct)()
    return _ply_result(_ply_load_memory(buf.ctypes.data, len(data), name.encode(), ctypes.byref(out)), out)


# --- objects ------------------------------------------------------------------------------------


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        _check(_context_create(device, ctypes.byref(h)), "gsr_context_create")
        self.handle = h
        self.device = device

    def synchronize(self) -> None:
        _check(_context_sync(self.handle), "gsr_context_synchronize")

    def close(self) -> None:
        if self.handle:
            _context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation owned by the library (the caller-owned colour target)."""

    def __init__(self, ctx: Context, nbytes: int):
        h = _vp()
        _check(_buffer_create(ctx.handle, nbytes, ctypes.byref(h)), "gsr_buffer_create")
        self.handle, self.nbytes, self.ctx = h, nbytes, ctx

    @property
    def ptr(self) -> int:
        return _buffer_ptr(self.handle)

    def read(self, out: np.ndarray) -> np.ndarray:
        _check(_buffer_read(self.handle, _ptr(out), out.nbytes), "gsr_buffer_read")
        return out

    def write(self, src: np.ndarray) -> None:
        src = np.ascontiguousarray(src)
        _check(_buffer_write(self.handle, _ptr(src), src.nbytes), "gsr_buffer_write")

    def close(self) -> None:
        if self.handle:
            _buffer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """GpuScene: the scene buffers on the device, f32 SoA or fp16 records."""

    def __init__(self, ctx: Context, pos_alpha=None, scales=None, rotations=None, sh=None, sh_degree: int = 3,
                 fmt: str = "f32", geometry=None, color=None, device_pointers: bool = False):
        """device_pointers: the arrays are first copied into device buffers (kept by this
        object) and the scene is created from those device pointers."""
        desc = SceneDesc()
        desc.sh_degree = sh_degree
        desc.device_pointers = 0
        self._keep = []
        self._device_arrays = []
        if fmt == "f32":
            arrays = [_f32c(x) for x in (pos_alpha, scales, rotations, sh)]
            n = arrays[0].shape[0]
            K = (sh_degree + 1) ** 2
            if arrays[0].shape != (n, 4) or arrays[1].shape != (n, 4) or arrays[2].shape != (n, 4):
                raise ValueError("pos_alpha, scales, rotations must be (N, 4) float32")
            if arrays[3].size != n * K * 3:
                raise ValueError(f"sh must hold N*K*3 = {n * K * 3} floats")
            desc.format = SCENE_F32_SOA
            desc.count = n
            desc.pos_alpha, desc.scales, desc.rotations, desc.sh = (_ptr(a) for a in arrays)
            self._keep = arrays
        elif fmt == "f16":
            if geometry is None:
                geometry, color = encode_f16(pos_alpha, scales, rotations, sh, sh_degree)
            g = np.ascontiguousarray(geometry, np.uint32)
            c = np.ascontiguousarray(color, np.uint16)
            desc.format = SCENE_F16
            desc.count = g.shape[0]
            desc.geometry, desc.color = _ptr(g), _ptr(c)
            self._keep = [g, c]
        else:
            raise ValueError(f"unknown scene format {fmt!r}")
        if device_pointers:
            names = ("pos_alpha", "scales", "rotations", "sh") if fmt == "f32" else ("geometry", "color")
            for name, arr in zip(names, self._keep):
                buf = DeviceBuffer(ctx, max(arr.nbytes, 1))
                if arr.nbytes:
                    buf.write(arr)
                setattr(desc, name, buf.ptr)
                self._device_arrays.append(buf)
            desc.device_pointers = 1
        h = _vp()
        _check(_scene_create(ctx.handle, ctypes.byref(desc), ctypes.byref(h)), "gsr_scene_create")
        self._keep = []  # copied to the device
        self.handle, self.count, self.sh_degree, self.format, self.ctx = h, int(desc.count), sh_degree, fmt, ctx
        self._renderers = weakref.WeakSet()  # open renderers bound to this scene

    @classmethod
    def from_device(cls, ctx: Context, count: int, pos_alpha_ptr: int, scales_ptr: int, rotations_ptr: int,
                    sh_ptr: int, sh_degree: int) -> "Scene":
        """An f32 scene bound to caller-owned device arrays (pos_alpha, scales, rotations: N x 4
        floats, bound in place and read by every frame; sh: N x K x 3, copied at creation). The
        caller keeps the arrays alive and unchanged while the scene is in use."""
        desc = SceneDesc()
        desc.sh_degree = sh_degree
        desc.format = SCENE_F32_SOA
        desc.count = int(count)
        desc.device_pointers = 1
        desc.pos_alpha, desc.scales, desc.rotations, desc.sh = pos_alpha_ptr, scales_ptr, rotations_ptr, sh_ptr
        h = _vp()
        _check(_scene_create(ctx.handle, ctypes.byref(desc), ctypes.byref(h)), "gsr_scene_create")
        self = cls.__new__(cls)
        self._keep, self._device_arrays = [], []
        self.handle, self.count, self.sh_degree, self.format, self.ctx = h, int(count), sh_degree, "f32", ctx
        self._renderers = weakref.WeakSet()
        return self

    def close(self) -> None:
        if self.handle:
            if any(r.handle for r in list(self._renderers)):
                raise GsrError("Scene.close: a renderer bound to this scene is still open (the scene must outlive "
                               "its renderers; close them first)")
            _scene_destroy(self.handle)
            self.handle = None
        for b in self._device_arrays:
            b.close()
        self._device_arrays = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """SplatTileRenderer for one scene (init + render + the test readbacks)."""

    def __init__(self, ctx: Context, scene: Scene, bounds: str = "tight", target: str = "rgba8", budget_mb: int = 256,
                 max_sh_degree: int = 3, profile_stages: int = 0, pixels_per_thread: int = 0,
                 exact_alpha: bool = False, sort_algorithm: str = "reduce-then-scan", render_loop: str = "culled",
                 sort_mode: str = "binning", full_key_sort: bool = False, render_splat: str = "full",
                 sort_window: int = 0, adaptive_list: bool = False, adaptive_initial_entries: int = 0,
                 frame_stats: bool = False, differentiable: bool = False, tile_stats: bool = False,
                 tile_size: int = 16, render_batch: int = 0):
        s = Settings()
        s.bounds = BOUNDS[bounds]
        s.budget_mb = budget_mb
        s.max_sh_degree = max_sh_degree
        s.format = TARGETS[target]
        s.profile_stages = int(profile_stages)  # False/True = 0/1; 2 = also per-kernel events
        s.render_pixels_per_thread = pixels_per_thread
        s.render_exact_alpha = 1 if exact_alpha else 0
        s.sort_algorithm = SORT_ALGORITHMS[sort_algorithm]
        s.render_loop = RENDER_LOOPS[render_loop]
        s.sort_mode = SORT_MODES[sort_mode]
        s.full_key_sort = 1 if full_key_sort else 0
        s.render_splat = RENDER_SPLATS[render_splat]
        s.sort_window = sort_window
        s.adaptive_list = 1 if adaptive_list else 0
        s.adaptive_initial_entries = adaptive_initial_entries
        s.frame_stats = 1 if frame_stats else 0
        s.differentiable = 1 if differentiable else 0
        s.tile_stats = 1 if tile_stats else 0
        s.tile_size = int(tile_size)
        s.render_batch = int(render_batch)
        h = _vp()
        _check(_renderer_create(ctx.handle, scene.handle, ctypes.byref(s), ctypes.byref(h)), "gsr_renderer_create")
        self.handle, self.ctx, self.scene, self.target, self.bounds = h, ctx, scene, target, bounds
        scene._renderers.add(self)
        self.sort_mode = sort_mode
        self.render_splat = render_splat
        self.tile_size = int(tile_size) or 16
        self._colour: Optional["ColourModel"] = None
        self._image: Optional[DeviceBuffer] = None
        self._cam: Optional[Camera] = None

    @property
    def capacity(self) -> int:
        return int(_renderer_capacity(self.handle))

    @property
    def total_bytes(self) -> int:
        return int(_renderer_total_bytes(self.handle))

    @property
    def max_capacity(self) -> int:
        return int(_renderer_max_capacity(self.handle))

    @property
    def list_resizes(self) -> int:
        return int(_renderer_list_resizes(self.handle))

    def last_stats(self):
        """lastStats: (control, frame index) of the newest completed frame, or None."""
        c = Control()
        idx = ctypes.c_uint64()
        if _renderer_last_stats(self.handle, ctypes.byref(c), ctypes.byref(idx)) != 0:
            return None
        return c, int(idx.value)

    def image_bytes(self, cam: Camera) -> int:
        return int(cam.width) * int(cam.height) * (4 if self.target == "rgba8" else 16)

    def backward(self, grad_image_ptr: int, means: Optional[int] = None, log_scales: Optional[int] = None,
                 quats: Optional[int] = None, opacity_logits: Optional[int] = None, sh: Optional[int] = None,
                 stream: Optional[int] = None) -> None:
        """gsr_backward on device pointers: dL/d(image) of the last frame (RGBA32F) -> the gradients
        of the scene parameters (asynchronous)."""
        g = Gradients(means, log_scales, quats, opacity_logits, sh)
        _check(_backward(self.handle, grad_image_ptr, ctypes.byref(g), stream), "gsr_backward")

    def backward_to_host(self, grad_image: np.ndarray):
        """gsr_backward with host arrays: grad_image (H, W, 3 or 4) float32 -> dict of numpy
        gradients (means, log_scales, quats, opacity_logits, sh)."""
        cam = self._cam
        if cam is None:
            raise GsrError("Renderer.backward_to_host: no frame rendered yet")
        H, W = int(cam.height), int(cam.width)
        gi = np.zeros((H, W, 4), np.float32)
        gi[..., :3] = np.asarray(grad_image, np.float32).reshape(H, W, -1)[..., :3]
        n = self.scene.count
        K = (self.scene.sh_degree + 1) ** 2
        shapes = {"means": (n, 3), "log_scales": (n, 3), "quats": (n, 4), "opacity_logits": (n,), "sh": (n, K * 3)}
        bufs = {k: DeviceBuffer(self.ctx, max(4 * int(np.prod(v)), 4)) for k, v in shapes.items()}
        gbuf = DeviceBuffer(self.ctx, gi.nbytes)
        try:
            gbuf.write(gi)
            self.backward(gbuf.ptr, **{k: b.ptr for k, b in bufs.items()})
            self.ctx.synchronize()
            return {k: bufs[k].read(np.empty(shapes[k], np.float32)) for k in shapes}
        finally:
            gbuf.close()
            for b in bufs.values():
                b.close()

    def set_colour_model(self, model: Optional["ColourModel"]) -> None:
        """Learned colour for the gaussians the model's last fallback pass marked (None: SH only)."""
        _check(_renderer_set_colour_model(self.handle, None if model is None else model.handle),
               "gsr_renderer_set_colour_model")
        self._colour = model
        if model is not None:
            model._bound.add(self)

    def render(self, cam: Camera, out_ptr: Optional[int] = None, stream: Optional[int] = None) -> int:
        """Record one frame (asynchronous). Returns the device pointer written."""
        if out_ptr is None:
            need = self.image_bytes(cam)
            if self._image is None or self._image.nbytes < need:
                self._image = DeviceBuffer(self.ctx, need)
            out_ptr = self._image.ptr
        _check(_render(self.handle, ctypes.byref(cam), out_ptr, stream), "gsr_render")
        self._cam = cam
        return out_ptr

    def read_image(self, cam: Optional[Camera] = None) -> np.ndarray:
        cam = cam or self._cam
        h, w = int(cam.height), int(cam.width)
        if self.target == "rgba8":
            out = np.empty((h, w, 4), np.uint8)
        else:
            out = np.empty((h, w, 4), np.float32)
        return self._image.read(out)

    def render_to_host(self, cam: Camera) -> np.ndarray:
        self.render(cam)
        return self.read_image(cam)

    def read_control(self) -> Control:
        c = Control()
        _check(_read_control(self.handle, ctypes.byref(c)), "gsr_read_control")
        return c

    def bins(self, cam: Optional[Camera] = None):
        """The lists' grid for a camera: (bins_x, bins_y) of tile_size px."""
        cam = cam or self._cam
        S = self.tile_size
        return (int(cam.width) + S - 1) // S, (int(cam.height) + S - 1) // S

    def read_back(self):
        """readBack: sorted tile keys, depth keys, values (entries each), ranges (bins, 2), control."""
        c = self.read_control()
        n = int(c.entries)
        bx, by = self.bins()
        tiles = bx * by
        keys = np.empty(n, np.uint32)
        depth = np.empty(n, np.uint32)
        vals = np.empty(n, np.uint32)
        ranges = np.empty((tiles, 2), np.uint32)
        _check(_read_lists(self.handle, _ptr(keys), _ptr(depth), _ptr(vals), _ptr(ranges), n), "gsr_read_lists")
        return keys, depth, vals, ranges, c

    def read_back_splats(self):
        """readBackSplats: (N,) SPLAT_DTYPE records and tile counts, index order."""
        n = self.scene.count
        splats = np.empty(n, SPLAT_DTYPE)
        counts = np.empty(n, np.uint32)
        _check(_read_splats(self.handle, _ptr(splats), _ptr(counts)), "gsr_read_splats")
        return splats, counts

    def read_back_packed(self) -> np.ndarray:
        """readBackPackedSplats: (N, 8) (p32) or (N, 6) (p24) render-splat words."""
        words = np.empty((self.scene.count, 8 if self.render_splat == "p32" else 6), np.uint32)
        _check(_read_packed(self.handle, _ptr(words)), "gsr_read_packed_splats")
        return words

    def stage_times(self) -> dict:
        ms = (ctypes.c_float * len(STAGES))()
        _check(_read_stage_times(self.handle, ctypes.cast(ms, _vp), len(STAGES)), "gsr_read_stage_times")
        return {k: float(v) for k, v in zip(STAGES, ms)}

    def kernel_times(self):
        """profile_stages=2: [(kernel name, ms)] of the last frame, in launch order."""
        n = 96
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_float * n)()
        cnt = ctypes.c_uint32()
        _check(_read_kernel_times(self.handle, ctypes.cast(names, _vp), ctypes.cast(ms, _vp), n, ctypes.byref(cnt)),
               "gsr_read_kernel_times")
        return [(names[i].decode(), float(ms[i])) for i in range(min(cnt.value, n))]

    def tile_stats(self) -> np.ndarray:
        """tile_stats=True: (tiles, 8) uint32 per-tile counters of the last frame (TILE_STATS order)."""
        bx, by = self.bins()
        tiles = bx * by
        out = np.zeros((tiles, len(TILE_STATS)), np.uint32)
        _check(_read_tile_stats(self.handle, _ptr(out), tiles), "gsr_read_tile_stats")
        return out

    def close(self) -> None:
        if self.handle:
            _renderer_destroy(self.handle)
            self.handle = None
        if self._image is not None:
            self._image.close()
            self._image = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_HIP_RUNTIME = None


def hip_runtime():
    """The HIP runtime libgsr.so runs on (its SONAME, already loaded by the library), for callers
    that need their own streams or events next to the library's."""
    global _HIP_RUNTIME
    if _HIP_RUNTIME is None:
        try:
            _HIP_RUNTIME = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            _HIP_RUNTIME = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    return _HIP_RUNTIME


class Stream:
    """A non-blocking HIP stream of default priority on the context's device (frames on several
    streams, frames in flight). `handle` is what Renderer.render and Batch take as `stream`."""

    def __init__(self, ctx: Context):
        hip = hip_runtime()
        if hip.hipSetDevice(ctx.device) != 0:
            raise GsrError("hipSetDevice failed")
        h = ctypes.c_void_p()
        if hip.hipStreamCreateWithFlags(ctypes.byref(h), 1) != 0:
            raise GsrError("hipStreamCreateWithFlags failed")
        self.handle = h.value

    def synchronize(self) -> None:
        if hip_runtime().hipStreamSynchronize(ctypes.c_void_p(self.handle)) != 0:
            raise GsrError("hipStreamSynchronize failed")

    def close(self) -> None:
        if self.handle:
            hip = hip_runtime()
            hip.hipStreamSynchronize(ctypes.c_void_p(self.handle))
            hip.hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many views of one scene on `branches` streams at once (gsr_batch_*): view v is rendered by
    renderers[v % branches] on branch stream v % branches, forked from and joined back to the
    caller's stream; capture() records the batch into a HIP graph, replay() launches it.
    persistent=True: one renderer per view, the views' frames run up to the tile sort on the
    branches and one render launch draws every view's tiles from a shared work counter."""

    def __init__(self, ctx: Context, branches: int, persistent: bool = False):
        h = _vp()
        _check(_batch_create(ctx.handle, int(branches), ctypes.byref(h)), "gsr_batch_create")
        self.handle, self.ctx, self.branches = h, ctx, int(branches)
        self._captured = None  # keeps the captured arrays and objects alive
        self.persistent = False
        if persistent:
            self.set_persistent(True)

    def set_persistent(self, on: bool) -> None:
        _check(_batch_set_persistent(self.handle, int(bool(on))), "gsr_batch_set_persistent")
        if bool(on) != self.persistent:
            self._captured = None
        self.persistent = bool(on)

    @staticmethod
    def _arrays(renderers, cams, outs):
        rs = (_vp * len(renderers))(*[r.handle for r in renderers])
        cs = (Camera * len(cams))(*cams)
        os_ = (_vp * len(outs))(*[int(o) for o in outs])
        return rs, cs, os_

    def render(self, renderers, cams, outs, stream: Optional[int] = None) -> None:
        """cams: Camera per view; outs: device pointers (one image per view, the renderer's format)."""
        if len(cams) != len(outs):
            raise GsrError("Batch.render: one output per camera")
        rs, cs, os_ = self._arrays(renderers, cams, outs)
        _check(_batch_render(self.handle, rs, len(cams), cs, os_, stream), "gsr_batch_render")

    def capture(self, renderers, cams, outs) -> None:
        if len(cams) != len(outs):
            raise GsrError("Batch.capture: one output per camera")
        rs, cs, os_ = self._arrays(renderers, cams, outs)
        _check(_batch_capture(self.handle, rs, len(cams), cs, os_), "gsr_batch_capture")
        self._captured = (list(renderers), rs, cs, os_)

    def replay(self, stream: Optional[int] = None) -> None:
        _check(_batch_replay(self.handle, stream), "gsr_batch_replay")

    def close(self) -> None:
        if self.handle:
            _batch_destroy(self.handle)
            self.handle = None
        self._captured = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ColourModel:
    """Learned colour (gsr_colour_model_*): a 16-value latent per gaussian and one shared MLP
    ([latent, direction, 1] -> 64 -> ReLU -> 64 -> ReLU -> 3, bf16 MFMA), trained towards the
    scene's SH colours; fallback() marks the gaussians it reproduces within a threshold, and
    renderers bound with Renderer.set_colour_model colour those from the MLP."""

    LATENT = 16

    def __init__(self, ctx: Context, scene: Scene, seed: int = 0):
        h = _vp()
        _check(_cm_create(ctx.handle, scene.handle, int(seed) & 0xFFFFFFFF, ctypes.byref(h)), "gsr_colour_model_create")
        self.handle, self.ctx, self.scene, self.count = h, ctx, scene, scene.count
        scene._renderers.add(self)  # the scene must outlive the model
        self._bound = weakref.WeakSet()

    @staticmethod
    def param_count() -> int:
        return int(_cm_param_count())

    def train(self, steps: int, samples_per_step: int = 1 << 20, lr_weights: float = 1e-3,
              lr_latents: float = 1e-2) -> float:
        """Adam steps; returns the last step's mean squared error (per channel)."""
        loss = ctypes.c_float()
        _check(_cm_train(self.handle, int(steps), int(samples_per_step), float(lr_weights), float(lr_latents),
                         ctypes.byref(loss)), "gsr_colour_model_train")
        return float(loss.value)

    def loss(self, seed: int, samples: int) -> float:
        out = ctypes.c_float()
        _check(_cm_loss(self.handle, int(seed) & 0xFFFFFFFF, int(samples), ctypes.byref(out)), "gsr_colour_model_loss")
        return float(out.value)

    def gradients(self, seed: int, samples: int):
        """(grad_params, grad_latents (N, 16), loss, gaussians, dirs (samples, 3)) of one batch."""
        gp = np.zeros(self.param_count(), np.float32)
        gl = np.zeros((self.count, self.LATENT), np.float32)
        g = np.zeros(samples, np.uint32)
        d = np.zeros((samples, 3), np.float32)
        loss = ctypes.c_float()
        _check(_cm_gradients(self.handle, int(seed) & 0xFFFFFFFF, int(samples), _ptr(gp), _ptr(gl),
                             ctypes.byref(loss), _ptr(g), _ptr(d)), "gsr_colour_model_gradients")
        return gp, gl, float(loss.value), g, d

    def get(self):
        p = np.zeros(self.param_count(), np.float32)
        lat = np.zeros((self.count, self.LATENT), np.float32)
        _check(_cm_get(self.handle, _ptr(p), _ptr(lat)), "gsr_colour_model_get")
        return p, lat

    def set(self, params=None, latents=None) -> None:
        p = None if params is None else _f32c(params).reshape(-1)
        lat = None if latents is None else _f32c(latents).reshape(-1)
        if p is not None and p.size != self.param_count():
            raise ValueError(f"params must hold {self.param_count()} floats")
        if lat is not None and lat.size != self.count * self.LATENT:
            raise ValueError(f"latents must hold {self.count * self.LATENT} floats")
        _check(_cm_set(self.handle, _ptr(p), _ptr(lat)), "gsr_colour_model_set")

    def evaluate(self, gaussians, dirs) -> np.ndarray:
        """y + 0.5 (unclamped) at (gaussian, direction) pairs: (n, 3)."""
        g = np.ascontiguousarray(gaussians, np.uint32).reshape(-1)
        d = _f32c(dirs).reshape(-1, 3)
        if d.shape[0] != g.size:
            raise ValueError("one direction per gaussian index")
        out = np.zeros((g.size, 3), np.float32)
        _check(_cm_evaluate(self.handle, _ptr(g), _ptr(d), g.size, _ptr(out)), "gsr_colour_model_evaluate")
        return out

    def fallback(self, threshold: float):
        """Marks the gaussians within `threshold` of their SH colour; returns (count, errors)."""
        n = ctypes.c_uint32()
        err = np.zeros(self.count, np.float32)
        _check(_cm_fallback(self.handle, float(threshold), ctypes.byref(n), _ptr(err)), "gsr_colour_model_fallback")
        return int(n.value), err

    def close(self) -> None:
        if self.handle:
            if any(r.handle for r in list(self._bound) if getattr(r, "_colour", None) is self):
                raise GsrError("ColourModel.close: a renderer bound to this model is still open (unbind or close it)")
            _cm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_batch(ctx: Context, make_renderer, cams, outs, candidates=(1, 2, 4, 8), replays: int = 5, timer=None,
               tile_sizes=(16,), batches=(0,), persistent=(False,)):
    """The batch's autotuner: for every (tile size, render batch, branch count, persistent mode) in
    tile_sizes x batches x candidates x persistent, renders the views once (the renderers' buffers
    for every size), captures the batch into a HIP graph and times `replays` replays after one
    warm-up; returns the fastest configuration as {"tile_size", "render_batch", "branches",
    "persistent"} and {(tile size, batch, branches, persistent): ms per replay}.
    make_renderer(tile_size=, render_batch=) makes one renderer of the scene (batch 0 = the default
    of the tile size); batches is one tuple for every tile size or a dict {tile size: tuple}. The
    persistent mode takes one renderer per view and the default culled loop, so it is tried for
    16-px tiles and the default batch only. timer(fn) times fn on the context's stream, in ms."""
    results = {}
    for tile in tile_sizes:
        for rb in (batches[tile] if isinstance(batches, dict) else batches):
            for pers in persistent:
                if pers and (tile != 16 or rb not in (0, 256)):
                    continue
                pool = []
                try:
                    for b in candidates:
                        while len(pool) < (len(cams) if pers else b):
                            pool.append(make_renderer(tile_size=tile, render_batch=rb))
                        use = pool if pers else pool[:b]
                        batch = Batch(ctx, b, persistent=pers)
                        try:
                            batch.render(use, cams, outs)
                            ctx.synchronize()
                            batch.capture(use, cams, outs)
                            batch.replay()
                            ctx.synchronize()

                            def run():
                                for _ in range(replays):
                                    batch.replay()
                            results[(tile, rb, b, bool(pers))] = timer(run) / replays
                        finally:
                            batch.close()
                finally:
                    for r in pool:
                        r.close()
    best = min(results, key=results.get)
    return {"tile_size": best[0], "render_batch": best[1], "branches": best[2], "persistent": best[3]}, results


NOALLOC_SORT_MAX_KEYS = 1 << 20  # the fixed scratch of gsr_sort_pairs_noalloc


def sort_pairs_device(ctx: Context, keys_ptr: int, values_ptr: int, max_count: int, key_bits: int = 64,
                      count_ptr: Optional[int] = None, algorithm: str = "reduce-then-scan", noalloc: bool = False,
                      stream: Optional[int] = None) -> None:
    """GpuRadixSort in place on device pointers (asynchronous). noalloc: gsr_sort_pairs_noalloc (no
    allocation; at most NOALLOC_SORT_MAX_KEYS keys, reduce-then-scan only)."""
    if noalloc:
        if algorithm != "reduce-then-scan":
            raise GsrError("sort_pairs_device: the allocation-free sort is reduce-then-scan")
        _check(_sort_pairs_noalloc(ctx.handle, keys_ptr, values_ptr, int(max_count), count_ptr, int(key_bits), stream),
               "gsr_sort_pairs_noalloc")
    else:
        _check(_sort_pairs(ctx.handle, keys_ptr, values_ptr, int(max_count), count_ptr, int(key_bits),
                           SORT_ALGORITHMS[algorithm], stream), "gsr_sort_pairs")


def sort_pairs(ctx: Context, keys: np.ndarray, values: np.ndarray, key_bits: int = 64, count: Optional[int] = None,
               algorithm: str = "reduce-then-scan"):
    """GpuRadixSort on host arrays: returns (sorted keys, values) by the low key_bits bits, stable."""
    keys = np.ascontiguousarray(keys, np.uint64)
    values = np.ascontiguousarray(values, np.uint32)
    n = keys.shape[0]
    kb = DeviceBuffer(ctx, max(8 * n, 8))
    vb = DeviceBuffer(ctx, max(4 * n, 4))
    cb = DeviceBuffer(ctx, 4)
    if n:
        kb.write(keys)
        vb.write(values)
    cnt_ptr = None
    if count is not None:
        cb.write(np.array([count], np.uint32))
        cnt_ptr = cb.ptr
    _check(_sort_pairs(ctx.handle, kb.ptr, vb.ptr, n, cnt_ptr, key_bits, SORT_ALGORITHMS[algorithm], None),
           "gsr_sort_pairs")
    ok, ov = np.empty_like(keys), np.empty_like(values)
    if n:
        kb.read(ok)
        vb.read(ov)
    return ok, ov


def exclusive_scan(ctx: Context, values: np.ndarray):
    """GpuScan: exclusive prefix sum of u32 values and the total."""
    v = np.ascontiguousarray(values, np.uint32)
    n = v.shape[0]
    ib, ob, tb = DeviceBuffer(ctx, max(4 * n, 4)), DeviceBuffer(ctx, max(4 * n, 4)), DeviceBuffer(ctx, 4)
    if n:
        ib.write(v)
    _check(_exclusive_scan(ctx.handle, ib.ptr, ob.ptr, n, tb.ptr, None), "gsr_exclusive_scan")
    out = np.empty_like(v)
    if n:
        ob.read(out)
    total = tb.read(np.empty(1, np.uint32))[0]
    return out, int(total)
