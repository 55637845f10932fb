"""ctypes binding of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker or as the
timed CPU baseline; never by the product.

The oracle is a C++ restatement of the reference's CPU forward pass (src/splat/*); see
gsr_oracle.hpp for the file-by-file mapping.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


if not os.path.exists(LIB_PATH):
    build()
_lib = ctypes.CDLL(LIB_PATH)

_vp, _u32, _i32, _f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_float

SPLAT_DTYPE = np.dtype([("center", np.float32, 2), ("depth", np.float32), ("cull_reason", np.uint32),
                        ("conic_opacity", np.float32, 4), ("color", np.float32, 3), ("radius", np.float32)])
RECT, TIGHT, CUTOFF = 0, 1, 2
RENDER_SPLAT = {"full": 0, "p32": 1, "p24": 2}
EXACT_PIXEL_SORT = 0xFFFFFFFF


class RenderOptions(ctypes.Structure):
    _fields_ = [("maxShDegree", _i32), ("earlyTermination", _i32), ("bound", _u32), ("bounds", _u32),
                ("threads", _i32), ("background", _f32 * 3), ("renderSplat", _u32), ("sortWindow", _u32)]


class Stats(ctypes.Structure):
    _fields_ = [("gaussians", _u32), ("visible", _u32), ("culledNear", _u32), ("culledSingular", _u32),
                ("culledOffScreen", _u32), ("threads", _u32), ("tileInstances", ctypes.c_uint64),
                ("evaluatedFragments", ctypes.c_uint64), ("blendedFragments", ctypes.c_uint64),
                ("terminatedPixels", ctypes.c_uint64), ("preprocessMs", ctypes.c_double),
                ("sortMs", ctypes.c_double), ("binMs", ctypes.c_double), ("rasterMs", ctypes.c_double),
                ("totalMs", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, restype, argtypes):
    f = getattr(_lib, name)
    f.restype, f.argtypes = restype, argtypes
    return f


_camera_from_viewer = _sig("oracle_camera_from_viewer", None, [_vp] * 3 + [_u32, _u32, _vp, _vp, _vp])
_make_camera = _sig("oracle_make_camera", None, [_vp, _u32, _u32, _f32, _f32, _vp, _vp, _vp])
det_logf = _sig("oracle_det_logf", _f32, [_f32])
float_to_half = _sig("oracle_float_to_half", ctypes.c_uint16, [_f32])
half_to_float = _sig("oracle_half_to_float", _f32, [ctypes.c_uint16])
_splat_math = _sig("oracle_splat_math", None, [_vp, _vp, _vp, _vp, _u32, _u32, _f32, _f32, _vp, _vp, _vp, _vp, _vp,
                                                _vp])
_splat_math_f64 = _sig("oracle_splat_math_f64", None, [_vp, _vp, _vp, _vp, _u32, _u32, _f32, _f32, _vp, _vp, _vp,
                                                        _vp, _vp, _vp])
_tile_rect = _sig("oracle_tile_rect", None, [_f32, _f32, _f32, _u32, _u32, _vp])
_eval_sh = _sig("oracle_eval_sh", None, [_i32, _vp, _vp, _vp])
_alpha_at = _sig("oracle_splat_alpha_at", _f32, [_vp, _f32, _f32])
_preprocess = _sig("oracle_preprocess", None, [_u32, _vp, _vp, _vp, _vp, _u32, _i32, _vp, _u32, _u32, _f32, _f32, _i32,
                                               _u32, _u32, _vp, _vp, _i32])
_encode_render_splats = _sig("oracle_encode_render_splats", None, [_u32, _vp, _u32, _vp])
_ray_depth = _sig("oracle_ray_depth", None, [_vp, _vp, _vp, _vp, _vp])
_preprocess_f64 = _sig("oracle_preprocess_f64", None, [_u32, _vp, _vp, _vp, _vp, _u32, _i32, _vp, _u32, _u32, _f32,
                                                       _f32, _i32, _vp, _i32])
_coverage_tiles = _sig("oracle_coverage_tiles", _u32, [_vp, _u32, _i32, _u32, _u32, _i32, _i32, _vp, _u32, _vp])
_preprocess_bins = _sig("oracle_preprocess_bins", None, [_u32, _vp, _vp, _vp, _vp, _u32, _i32, _vp, _u32, _u32, _f32,
                                                         _f32, _i32, _u32, _u32, _i32, _vp, _vp, _i32])
_tile_lists_bins = _sig("oracle_tile_lists_bins", _u32, [_vp, _vp, _u32, _u32, _u32, _i32, _u32, _u32, _u32, _u32, _u32,
                                                         _vp, _vp, _vp, _vp, _vp, _vp])
_tile_lists = _sig("oracle_tile_lists", _u32, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp,
                                               _vp])
_radix_sort_pairs = _sig("oracle_radix_sort_pairs", _u32, [_vp, _vp, _u32])
_encode_f16 = _sig("oracle_encode_f16", None, [_u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp])
_decode_f16 = _sig("oracle_decode_f16", None, [_u32, _vp, _vp, _u32, _vp, _vp, _vp, _vp])
_render = _sig("oracle_render", None, [_u32, _vp, _vp, _vp, _vp, _u32, _i32, _vp, _u32, _u32, _f32, _f32, _vp, _vp,
                                       _vp, _vp, _vp, ctypes.c_uint64])


def default_threads() -> int:
    """Host threads for the oracle: OMP_NUM_THREADS when set (the process's share of the host's
    cores), else every core up to 16."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env if env > 0 else min(os.cpu_count() or 1, 16)


def _p(a):
    return None if a is None else a.ctypes.data


def _f32c(a):
    return np.ascontiguousarray(a, np.float32)


class OracleCamera:
    """The oracle's SplatCamera: viewInria (column-major) + size + tan(fov/2)."""

    def __init__(self, view_inria, width, height, tan_x, tan_y):
        self.view = _f32c(view_inria).reshape(16).copy()
        self.width, self.height = int(width), int(height)
        self.tan_x, self.tan_y = float(np.float32(tan_x)), float(np.float32(tan_y))
        pos = np.zeros(3, np.float32)
        focal = np.zeros(2, np.float32)
        tiles = np.zeros(2, np.uint32)
        _make_camera(_p(self.view), self.width, self.height, self.tan_x, self.tan_y, _p(pos), _p(focal), _p(tiles))
        self.position, self.focal, self.tiles_x, self.tiles_y = pos, focal, int(tiles[0]), int(tiles[1])

    @classmethod
    def from_viewer(cls, view, proj, model, width, height):
        vi = np.zeros(16, np.float32)
        tx, ty = ctypes.c_float(), ctypes.c_float()
        v, p, m = (_f32c(x).reshape(16) for x in (view, proj, model))
        _camera_from_viewer(_p(v), _p(p), _p(m), width, height, _p(vi), ctypes.byref(tx), ctypes.byref(ty))
        return cls(vi, width, height, tx.value, ty.value)

    @classmethod
    def from_gsr(cls, cam):
        return cls(np.array(cam.view_inria[:], np.float32), cam.width, cam.height, cam.tan_half_fov_x,
                   cam.tan_half_fov_y)

    def args(self):
        return (_p(self.view), self.width, self.height, self.tan_x, self.tan_y)


def axis_camera(w, h, f):
    """A camera at the origin looking down +z (INRIA convention), focal f (test_splat_math.cpp:52-55)."""
    return OracleCamera(np.eye(4, dtype=np.float32), w, h, np.float32(w) / (np.float32(2.0) * np.float32(f)),
                        np.float32(h) / (np.float32(2.0) * np.float32(f)))


def splat_math(t, sigma, q, cam: OracleCamera, f64: bool = False):
    """cov3D (3x3 [col][row]), cov2D (a, b, c), conic, radius, invertible, centre."""
    dt = np.float64 if f64 else np.float32
    t, sigma, q = (np.ascontiguousarray(x, dt) for x in (t, sigma, q))
    cov3 = np.zeros(9, dt)
    cov2 = np.zeros(3, dt)
    conic = np.zeros(3, dt)
    radius = (ctypes.c_double if f64 else ctypes.c_float)()
    inv = ctypes.c_int32()
    center = np.zeros(2, dt)
    fn = _splat_math_f64 if f64 else _splat_math
    fn(_p(t), _p(sigma), _p(q), *cam.args(), _p(cov3), _p(cov2), _p(conic), ctypes.byref(radius), ctypes.byref(inv),
       _p(center))
    return dict(cov3d=cov3.reshape(3, 3), cov2d=cov2, conic=conic, radius=radius.value, invertible=bool(inv.value),
                center=center)


def tile_rect(cx, cy, radius, tiles_x, tiles_y):
    out = np.zeros(4, np.int32)
    _tile_rect(cx, cy, radius, tiles_x, tiles_y, _p(out))
    return tuple(int(v) for v in out)  # minX, minY, maxX, maxY


def eval_sh(degree, sh, direction):
    sh = _f32c(sh).reshape(-1)
    if sh.size < 48:
        sh = np.concatenate([sh, np.zeros(48 - sh.size, np.float32)])
    d = _f32c(direction)
    out = np.zeros(3, np.float32)
    _eval_sh(degree, _p(sh), _p(d), _p(out))
    return out


def splat_alpha_at(splat, px, py):
    s = np.ascontiguousarray(np.asarray(splat, SPLAT_DTYPE).reshape(1))
    return float(_alpha_at(_p(s), px, py))


def preprocess(scene, cam: OracleCamera, degree=None, gpu_cull=True, bounds=TIGHT, threads=None, render_splat="full",
               bin_size=16):
    """preprocessGaussian over the scene: (N,) splats and the per-gaussian tile counts (bins of
    bin_size = 16 or 8 px). render_splat "p24": visible splats hold the conic and opacity as the p24
    record gives them back (the tiles are counted with those)."""
    threads = default_threads() if threads is None else threads
    n = scene.count
    out = np.zeros(n, SPLAT_DTYPE)
    counts = np.zeros(n, np.uint32)
    deg = scene.sh_degree if degree is None else min(degree, scene.sh_degree)
    _preprocess_bins(n, _p(scene.pos_alpha), _p(scene.scales), _p(scene.rotations), _p(scene.sh), scene.K, deg,
                     *cam.args(), 1 if gpu_cull else 0, bounds, RENDER_SPLAT[render_splat], int(bin_size), _p(out),
                     _p(counts), threads)
    return out, counts


def ray_depth(sigma, q, t_cam, view_inria):
    """computeRayDepthParams: (p00, p01, p02, p11, p12, p22, b0, b1, b2) as float32."""
    a = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (sigma, q, t_cam, view_inria)]
    out = np.zeros(9, np.float32)
    _ray_depth(*(_p(x) for x in a), _p(out))
    return out


def encode_render_splats(splats, mode: str):
    """encodeRenderSplat (render_splat.h:29-48): (N, 8) words for "p32", (N, 6) for "p24"."""
    splats = np.ascontiguousarray(splats, SPLAT_DTYPE)
    words = np.zeros((splats.shape[0], 8 if mode == "p32" else 6), np.uint32)
    _encode_render_splats(splats.shape[0], _p(splats), RENDER_SPLAT[mode], _p(words))
    return words


def preprocess_f64(scene, cam: OracleCamera, degree=None, gpu_cull=True, threads=None):
    """FP64 twin: (N, 12) doubles: cx, cy, depth, cull, conic a b c, opacity, r g b, radius."""
    threads = default_threads() if threads is None else threads
    n = scene.count
    out = np.zeros((n, 12), np.float64)
    deg = scene.sh_degree if degree is None else min(degree, scene.sh_degree)
    _preprocess_f64(n, _p(scene.pos_alpha), _p(scene.scales), _p(scene.rotations), _p(scene.sh), scene.K, deg,
                    *cam.args(), 1 if gpu_cull else 0, _p(out), threads)
    return out


def coverage_tiles(splat, bounds, tiles_x, tiles_y, S=16, grid=None):
    """The bin tiles a splat is binned into, its count and whether it is borderline."""
    s = np.ascontiguousarray(np.asarray(splat, SPLAT_DTYPE).reshape(1))
    gx, gy = grid if grid is not None else (tiles_x, tiles_y)
    cap = 1 << 16
    buf = np.zeros(cap, np.uint32)
    border = ctypes.c_int32()
    cnt = _coverage_tiles(_p(s), bounds, S, tiles_x, tiles_y, gx, gy, _p(buf), cap, ctypes.byref(border))
    return buf[: min(cnt, cap)].copy(), int(cnt), bool(border.value)


def tile_lists(splats, counts, tiles_x, tiles_y, bounds=TIGHT, capacity=None, mode=0, bin_size=16, grid=None):
    """The GPU path's sorted list: tile keys, depth keys, values, ranges (tiles, 2), duplicates, mismatches.
    tiles_x/y: the 16-px tiles; bin_size 8: the list is of 8-px bins on grid = (bins_x, bins_y)."""
    splats = np.ascontiguousarray(splats, SPLAT_DTYPE)
    counts = np.ascontiguousarray(counts, np.uint32)
    n = splats.shape[0]
    dup = ctypes.c_uint64()
    mism = ctypes.c_uint32()
    total = int(counts.astype(np.uint64).sum())
    cap = total if capacity is None else min(int(capacity), total)
    hi = np.zeros(max(cap, 1), np.uint32)
    lo = np.zeros(max(cap, 1), np.uint32)
    vals = np.zeros(max(cap, 1), np.uint32)
    gx, gy = grid if grid is not None else (tiles_x, tiles_y)
    ranges = np.zeros((gx * gy, 2), np.uint32)
    cap_arg = cap if capacity is not None else 0xFFFFFFFF
    e = _tile_lists_bins(_p(splats), _p(counts), n, tiles_x, tiles_y, int(bin_size), gx, gy, bounds, cap_arg, mode,
                         _p(hi), _p(lo), _p(vals), _p(ranges), ctypes.byref(dup), ctypes.byref(mism))
    return hi[:e], lo[:e], vals[:e], ranges, int(dup.value), int(mism.value)


def radix_sort_pairs(keys, values):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(values, np.uint32).copy()
    passes = _radix_sort_pairs(_p(k), _p(v), k.shape[0])
    return k, v, int(passes)


def encode_f16(scene):
    n, K = scene.count, scene.K
    halves = (K * 3 + 7) // 8 * 8
    geom = np.zeros((n, 8), np.uint32)
    color = np.zeros((n, halves), np.uint16)
    _encode_f16(n, _p(scene.pos_alpha), _p(scene.scales), _p(scene.rotations), _p(scene.sh), K, _p(geom), _p(color))
    return geom, color


def decode_f16(geom, color, sh_degree):
    """decodeGaussian, f16 records: pos_alpha, scales, rotations, sh as float32 arrays."""
    geom = np.ascontiguousarray(geom, np.uint32)
    color = np.ascontiguousarray(color, np.uint16)
    n = geom.shape[0]
    K = (sh_degree + 1) ** 2
    pa, sc, ro = (np.zeros((n, 4), np.float32) for _ in range(3))
    sh = np.zeros((n, K * 3), np.float32)
    _decode_f16(n, _p(geom), _p(color), K, _p(pa), _p(sc), _p(ro), _p(sh))
    return pa, sc, ro, sh


def render(scene, cam: OracleCamera, max_sh_degree=3, early_termination=True, bound=0, bounds=RECT, threads=None,
           background=(0.0, 0.0, 0.0), want_lists=False, render_splat="full", sort_window=0):
    """renderReference: (H, W, 3) float32 image (unclamped) and ReferenceStats (+ per-tile lists).
    render_splat "p32"/"p24": every splat as the packed record gives it back (fp16 colour; p24
    also fp16 conic and opacity). sort_window K > 0: StopThePop's per-pixel window of K splats
    ordered by ray depth; EXACT_PIXEL_SORT: every pixel sorts all its splats."""
    threads = default_threads() if threads is None else threads
    opt = RenderOptions(max_sh_degree, 1 if early_termination else 0, bound, bounds, threads,
                        (ctypes.c_float * 3)(*background), RENDER_SPLAT[render_splat], sort_window)
    rgb = np.zeros((cam.height, cam.width, 3), np.float32)
    st = Stats()
    tiles = cam.tiles_x * cam.tiles_y
    start = np.zeros(tiles + 1, np.uint32) if want_lists else None
    lists = None
    if want_lists:
        # tile instances are not known in advance: one pass to get them.
        _render(scene.count, _p(scene.pos_alpha), _p(scene.scales), _p(scene.rotations), _p(scene.sh), scene.K,
                scene.sh_degree, *cam.args(), ctypes.byref(opt), _p(rgb), ctypes.byref(st), _p(start), None, 0)
        lists = np.zeros(max(int(st.tileInstances), 1), np.uint32)
    _render(scene.count, _p(scene.pos_alpha), _p(scene.scales), _p(scene.rotations), _p(scene.sh), scene.K,
            scene.sh_degree, *cam.args(), ctypes.byref(opt), _p(rgb), ctypes.byref(st), _p(start), _p(lists),
            0 if lists is None else lists.size)
    if want_lists:
        return rgb, st.as_dict(), start, lists[: int(st.tileInstances)]
    return rgb, st.as_dict()


_diff_render = _sig("oracle_diff_render_f64", ctypes.c_double,
                    [ctypes.c_uint32] + [ctypes.c_void_p] * 5 + [ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                                                 ctypes.c_uint32, ctypes.c_uint32, _f32, _f32,
                                                                 ctypes.c_void_p, ctypes.c_void_p])
_diff_fd = _sig("oracle_diff_fd_f64", None,
                [ctypes.c_uint32] + [ctypes.c_void_p] * 5 + [ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                                             ctypes.c_uint32, ctypes.c_uint32, _f32, _f32,
                                                             ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                                             ctypes.c_double, ctypes.c_void_p, ctypes.c_int32])

_diff_grad = _sig("oracle_diff_grad_f64", None,
                  [ctypes.c_uint32] + [ctypes.c_void_p] * 5 + [ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                                               ctypes.c_uint32, ctypes.c_uint32, _f32, _f32,
                                                               ctypes.c_void_p, ctypes.c_void_p])


class RawParams:
    """Raw parameters of the differentiable render (float64): means (N, 3), log_scales (N, 3),
    quats (N, 4) w-first, used unnormalised, logits (N,), sh (N, K * 3) coefficient-major RGB."""

    def __init__(self, means, log_scales, quats, logits, sh, sh_degree):
        f = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
        self.means, self.log_scales, self.quats, self.logits, self.sh = (f(a) for a in
                                                                          (means, log_scales, quats, logits, sh))
        self.sh_degree = int(sh_degree)
        self.count = self.means.shape[0]
        self.K = (self.sh_degree + 1) ** 2

    def _args(self):
        return [self.count, _p(self.means), _p(self.log_scales), _p(self.quats), _p(self.logits), _p(self.sh),
                self.K, self.sh_degree]


def diff_render_f64(params: RawParams, cam: OracleCamera, target=None):
    """FP64 render of the raw parameters (rect bounds, black background): (loss, (H, W, 3) image);
    loss = mean |image - target| (0 without a target)."""
    img = np.zeros((cam.height, cam.width, 3), np.float64)
    t = None if target is None else np.ascontiguousarray(np.asarray(target, np.float32)[..., :3])
    loss = _diff_render(*params._args(), *cam.args(), None if t is None else _p(t), _p(img))
    return float(loss), img


def diff_fd_f64(params: RawParams, cam: OracleCamera, target, which, h=1e-4, threads=None):
    """Central differences of the L1 loss for the gaussians `which`: (len(which), 11 + 3K) in the
    order mean (3), log-scale (3), quaternion (4), logit (1), SH (K x 3)."""
    threads = default_threads() if threads is None else threads
    w = np.ascontiguousarray(which, np.uint32)
    t = np.ascontiguousarray(np.asarray(target, np.float32)[..., :3])
    out = np.zeros((w.size, 11 + 3 * params.K), np.float64)
    _diff_fd(*params._args(), *cam.args(), _p(t), w.size, _p(w), float(h), _p(out), int(threads))
    return out


def diff_grad_f64(params: RawParams, cam: OracleCamera, grad_image):
    """The exact gradient of a loss L whose derivative in the image is grad_image (H, W, 3):
    (N, 11 + 3K) in the order of diff_fd_f64; zero rows for culled gaussians."""
    g = np.ascontiguousarray(np.asarray(grad_image, np.float32)[..., :3])
    assert g.shape == (cam.height, cam.width, 3)
    out = np.zeros((params.count, 11 + 3 * params.K), np.float64)
    _diff_grad(*params._args(), *cam.args(), _p(g), _p(out))
    return out


_diff_local = _sig("oracle_diff_local_f64", None,
                   [ctypes.c_uint32] + [ctypes.c_void_p] * 5 + [ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                                                ctypes.c_uint32, ctypes.c_uint32, _f32, _f32,
                                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                                ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                                ctypes.c_int32])


def diff_local_f64(params: RawParams, cam: OracleCamera, target=None, grad_image=None, which=(), h=1e-4,
                   want_image=False, want_fd=True, threads=None):
    """The tiled FP64 references for large scenes (each pixel walks only its tile's splats): dict of
    loss (L1 against target) and image (if want_image), and for the gaussians `which` the exact
    gradient rows for dL/dimage = grad_image ("grad") and the central differences with step h
    ("fd", from re-renders of the tiles the moved gaussian reaches), both (len(which), 11 + 3K)
    in the order of diff_fd_f64."""
    threads = default_threads() if threads is None else threads
    w = np.ascontiguousarray(which, np.uint32)
    P = 11 + 3 * params.K
    t = None if target is None else np.ascontiguousarray(np.asarray(target, np.float32)[..., :3])
    g = None if grad_image is None else np.ascontiguousarray(np.asarray(grad_image, np.float32)[..., :3])
    img = np.zeros((cam.height, cam.width, 3), np.float64) if want_image else None
    grad = np.zeros((w.size, P), np.float64) if g is not None and w.size else None
    fd = np.zeros((w.size, P), np.float64) if want_fd and t is not None and w.size else None
    loss = ctypes.c_double(0.0)
    _diff_local(*params._args(), *cam.args(), None if t is None else _p(t), None if g is None else _p(g), w.size,
                _p(w) if w.size else None, float(h), None if img is None else _p(img), ctypes.byref(loss),
                None if grad is None else _p(grad), None if fd is None else _p(fd), int(threads))
    return {"loss": float(loss.value), "image": img, "grad": grad, "fd": fd}


def quantize8(rgb):
    """writePngRgb (image_io.cpp:51-57): clamp to [0, 1], NaN -> 0, lround(v * 255)."""
    v = np.nan_to_num(np.clip(np.asarray(rgb, np.float32), 0.0, 1.0), nan=0.0).astype(np.float32)
    x = (v * np.float32(255.0)).astype(np.float32)  # the float product, then round half away from zero
    return np.floor(x.astype(np.float64) + 0.5).astype(np.uint8)


def psnr(a, b):
    """imgdiff (tools/imgdiff.cpp:108-127): clamp both to [0, 1], PSNR with peak 1."""
    a = np.clip(np.asarray(a, np.float64), 0.0, 1.0)
    b = np.clip(np.asarray(b, np.float64), 0.0, 1.0)
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)
