"""The CPU oracle against the reference's own known answers (tests/test_splat_math.cpp) and
against closed forms that a plausible slip (dropped term, wrong sign, transposed product)
would break. CPU only."""
import math

import numpy as np
import pytest


def quat_axis_angle(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.array([math.cos(angle / 2), *(a * math.sin(angle / 2))], np.float32)


def rodrigues(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def near(a, b, rel, abs_=0.0):
    return abs(a - b) <= abs_ + rel * max(abs(a), abs(b))


def test_cov3d_known_answers(oracle):
    cam = oracle.axis_camera(64, 64, 64.0)
    # test_splat_math.cpp:58-84. cov3d is [col][row].
    a = oracle.splat_math([0, 0, 5], [2.0, 1.0, 0.5], [1, 0, 0, 0], cam)["cov3d"]
    assert near(a[0, 0], 4.0, 1e-6) and near(a[1, 1], 1.0, 1e-6) and near(a[2, 2], 0.25, 1e-6)
    assert abs(a[0, 1]) < 1e-6 and abs(a[0, 2]) < 1e-6 and abs(a[1, 2]) < 1e-6
    b = oracle.splat_math([0, 0, 5], [2.0, 1.0, 0.5], quat_axis_angle([0, 0, 1], math.pi / 2), cam)["cov3d"]
    assert near(b[0, 0], 1.0, 1e-5) and near(b[1, 1], 4.0, 1e-5) and near(b[2, 2], 0.25, 1e-5)
    # Arbitrary rotation vs an independent axis-angle formula (Rodrigues), R S^2 R^T.
    axis, angle, sigma = [0.3, -0.8, 0.5], 1.1, np.array([0.7, 0.2, 0.05])
    R = rodrigues(axis, angle)
    expected = R @ np.diag(sigma ** 2) @ R.T
    got = oracle.splat_math([0, 0, 5], sigma, quat_axis_angle(axis, angle), cam)["cov3d"]
    assert np.max(np.abs(got.T - expected)) < 1e-6
    assert got[0, 1] == got[1, 0] and got[0, 2] == got[2, 0] and got[1, 2] == got[2, 1]


def test_cov2d_on_axis_radius_31_centre(oracle):
    # test_splat_math.cpp:87-105: sigma 0.1 at depth 5, focal 500 -> cov2D = 100.3, radius 31.
    cam = oracle.axis_camera(1000, 800, 500.0)
    m = oracle.splat_math([0, 0, 5], [0.1, 0.1, 0.1], [1, 0, 0, 0], cam)
    cov = m["cov2d"]
    assert near(cov[0], 100.3, 1e-5) and near(cov[2], 100.3, 1e-5) and abs(cov[1]) < 1e-6
    assert m["invertible"]
    assert near(m["conic"][0], 1 / 100.3, 1e-5) and near(m["conic"][2], 1 / 100.3, 1e-5) and m["conic"][1] == 0
    assert m["radius"] == 31.0
    assert near(m["center"][0], 499.5, 1e-6) and near(m["center"][1], 399.5, 1e-6)


def _project_d(p, cam):
    return np.array([cam.focal[0] * p[0] / p[2] + (cam.width - 1) * 0.5,
                     cam.focal[1] * p[1] / p[2] + (cam.height - 1) * 0.5], np.float64)


def test_cov2d_numerical_jacobian(oracle):
    # test_splat_math.cpp:112-155: a rotated, translated camera; cov2D = Jn W Sigma W^T Jn^T + 0.3 I.
    R = rodrigues([0.2, 1.0, 0.1], 0.4)
    view = np.eye(4)
    view[:3, :3] = R
    view = view @ np.array([[1, 0, 0, 0.3], [0, 1, 0, -0.2], [0, 0, 1, 0.5], [0, 0, 0, 1]])
    view_cm = view.T.astype(np.float32).reshape(16)  # column-major
    cam = oracle.OracleCamera(view_cm, 640, 480, 0.6, 0.45)
    t = np.array([0.9, -0.5, 3.0], np.float32)
    sigma, q = [0.3, 0.08, 0.02], quat_axis_angle([1, 2, 3], 0.7)
    m = oracle.splat_math(t, sigma, q, cam)
    S = m["cov3d"].T.astype(np.float64)  # [row][col]
    h = 1e-5
    Jn = np.zeros((2, 3))
    for k in range(3):
        dp = np.zeros(3)
        dp[k] = h
        Jn[:, k] = (_project_d(t + dp, cam) - _project_d(t - dp, cam)) / (2 * h)
    W = np.asarray(view_cm, np.float64).reshape(4, 4).T[:3, :3]
    exp = (Jn @ W) @ S @ (Jn @ W).T
    got = m["cov2d"]
    assert near(got[0], exp[0, 0] + 0.3, 1e-4)
    assert near(got[1], exp[0, 1], 1e-4, 1e-4)
    assert near(got[2], exp[1, 1] + 0.3, 1e-4)
    c = m["center"]
    ce = _project_d(t, cam)
    assert near(c[0], ce[0], 1e-5, 1e-3) and near(c[1], ce[1], 1e-5, 1e-3)
    cov = np.array([[got[0], got[1]], [got[1], got[2]]], np.float64)
    con = np.array([[m["conic"][0], m["conic"][1]], [m["conic"][1], m["conic"][2]]], np.float64)
    assert np.allclose(cov @ con, np.eye(2), atol=1e-4)
    mid = 0.5 * (float(got[0]) + float(got[2]))
    det = float(got[0]) * float(got[2]) - float(got[1]) ** 2
    lmax = mid + math.sqrt(max(0.1, mid * mid - det))
    assert m["radius"] == float(math.ceil(3 * math.sqrt(lmax)))


def test_cov2d_frustum_clamp(oracle):
    # test_splat_math.cpp:157-172: J evaluated at x/z = -1.3 tan, centre not clamped.
    cam = oracle.axis_camera(400, 300, 200.0)
    m = oracle.splat_math([-10.0, 0.0, 2.0], [0.2, 0.1, 0.3], [1, 0, 0, 0], cam)
    tx, tz, f = -1.3 * 1.0 * 2.0, 2.0, 200.0
    a = (f / tz) ** 2 * 0.04 + (f * tx / tz ** 2) ** 2 * 0.09 + 0.3
    assert near(m["cov2d"][0], a, 1e-5)
    assert near(m["cov2d"][2], (f / tz) ** 2 * 0.01 + 0.3, 1e-5)
    assert near(m["center"][0], 200.0 * -5.0 + 199.5, 1e-5)


def test_tile_rect(oracle):
    # test_splat_math.cpp:175-187.
    assert oracle.tile_rect(100.5, 40.0, 10.0, 10, 10) == (5, 1, 7, 4)
    r = oracle.tile_rect(-50.0, 10.0, 10.0, 10, 10)
    assert r[2] <= r[0] or r[3] <= r[1]
    r = oracle.tile_rect(1000.0, 10.0, 30.0, 10, 10)
    assert r[2] <= r[0] or r[3] <= r[1]
    assert oracle.tile_rect(5.0, 5.0, 100.0, 10, 10) == (0, 0, 7, 7)


def _scene(means, colors, opacities, sigma, degree=0):
    from gsr.scenes import SyntheticScene

    n = len(means)
    pa = np.array([[*m, o] for m, o in zip(means, opacities)], np.float32)
    sc = np.array([[sigma, sigma, sigma, 0.0]] * n, np.float32)
    ro = np.array([[1, 0, 0, 0]] * n, np.float32)
    dc = (np.asarray(colors, np.float32) - np.float32(0.5)) / np.float32(0.28209479177387814)
    return SyntheticScene(pa, sc, ro, dc.astype(np.float32).reshape(n, 3), degree, 0)


def test_culls(oracle):
    # test_splat_math.cpp:188-200 (reference culls, gpu_cull = False).
    cam = oracle.axis_camera(64, 64, 64.0)

    def pre(p):
        s, _ = oracle.preprocess(_scene([p], [(0.5, 0.5, 0.5)], [0.5], 0.1), cam, gpu_cull=False, bounds=0)
        return s[0]

    assert pre((0, 0, 0.2))["cull_reason"] == 1
    assert pre((0, 0, -1.0))["cull_reason"] == 1
    assert pre((0, 0, 0.25))["cull_reason"] == 0
    assert pre((50, 0, 1.0))["cull_reason"] == 3
    s = pre((0, 0, 2.0))
    assert s["radius"] > 0 and s["depth"] == 2.0 and s["conic_opacity"][3] == np.float32(0.5)
    assert near(s["color"][0], 0.5, 1e-6)
    # The GPU twin's extra cull: opacity < 1/255 (splat_preprocess.comp:306-309).
    lo, _ = oracle.preprocess(_scene([(0, 0, 2.0)], [(0.5, 0.5, 0.5)], [0.003], 0.1), cam, gpu_cull=True)
    assert lo[0]["cull_reason"] == 4


def test_spherical_harmonics(oracle):
    # test_splat_math.cpp:203-245.
    C0, C1 = 0.28209479177387814, 0.4886025119029199
    sh = np.zeros((16, 3), np.float32)
    sh[0] = (1.0, -1.0, 0.0)
    d0 = oracle.eval_sh(3, sh, [0, 0, 1])
    assert near(d0[0], 0.5 + C0, 1e-6) and near(d0[1], 0.5 - C0, 1e-6) and near(d0[2], 0.5, 1e-6)
    sh[0] = -5.0
    assert np.all(oracle.eval_sh(0, sh, [0, 0, 1]) == 0.0)
    d = np.array([0.3, -0.5, 0.8])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    exp1 = [-C1 * d[1], C1 * d[2], -C1 * d[0]]
    for k in (1, 2, 3):
        c = np.zeros((16, 3), np.float32)
        c[k] = (1, 0, 0)
        assert near(oracle.eval_sh(1, c, d)[0], 0.5 + exp1[k - 1], 1e-6)
        assert near(oracle.eval_sh(0, c, d)[0], 0.5, 1e-6)
    c6 = np.zeros((16, 3), np.float32)
    c6[6] = 1.0
    y6 = 0.31539156525252005 * (2 * d[2] ** 2 - d[0] ** 2 - d[1] ** 2)
    assert near(oracle.eval_sh(2, c6, d)[1], 0.5 + y6, 1e-6)
    assert near(oracle.eval_sh(1, c6, d)[1], 0.5, 1e-6)
    c15 = np.zeros((16, 3), np.float32)
    c15[15] = 1.0
    y15 = -0.5900435899266435 * d[0] * (d[0] ** 2 - 3 * d[1] ** 2)
    assert near(oracle.eval_sh(3, c15, d)[2], 0.5 + y15, 1e-6)


def test_sh_all_basis_functions_orthonormal(oracle):
    # All 16 real SH basis functions of evalSH are orthonormal over the sphere (4 pi x the
    # mean over a Fibonacci lattice): a wrong constant, sign-flipped or swapped term breaks
    # the Gram matrix. Y_k = evalSH(coefficient k = 1, DC offset 10) - evalSH(DC offset 10).
    C0 = 0.28209479177387814
    n = 4000
    i = np.arange(n) + 0.5
    zc = 1 - 2 * i / n
    phi = i * math.pi * (3 - math.sqrt(5))
    rr = np.sqrt(1 - zc * zc)
    dirs = np.stack([rr * np.cos(phi), rr * np.sin(phi), zc], 1).astype(np.float32)
    Y = np.zeros((n, 16))
    groups = [list(range(g, min(g + 3, 16))) for g in range(0, 16, 3)]
    for di, d in enumerate(dirs):
        base_c = np.zeros((16, 3), np.float32)
        base_c[0, :] = np.float32(10.0 / C0)
        base = oracle.eval_sh(3, base_c, d).astype(np.float64)
        for grp in groups:
            c = base_c.copy()
            for ch, k in enumerate(grp):
                c[k, ch] += np.float32(1.0)
            out = oracle.eval_sh(3, c, d).astype(np.float64)
            for ch, k in enumerate(grp):
                Y[di, k] = out[ch] - base[ch]
    G = 4 * math.pi * (Y.T @ Y) / n
    assert np.max(np.abs(G - np.eye(16))) < 1e-2, np.round(G, 3)


def test_depth_sort_matches_stable_sort(oracle):
    # test_splat_math.cpp:248-268: radix order == stable_sort order, with exact ties.
    rng = np.random.default_rng(42)
    z = rng.uniform(-1.0, 30.0, 20000).astype(np.float32)
    z[::3] = np.floor(z[::3])
    keep = np.nonzero(z > np.float32(0.2))[0].astype(np.uint32)
    keys = z[keep].view(np.uint32)
    k, v, passes = oracle.radix_sort_pairs(keys, keep)
    expected = keep[np.argsort(z[keep], kind="stable")]
    assert np.array_equal(v, expected)
    assert passes >= 1
    k0, v0, p0 = oracle.radix_sort_pairs(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert k0.size == 0 and p0 == 0


def test_reference_single_gaussian(oracle):
    # test_splat_math.cpp:286-322.
    cam = oracle.axis_camera(64, 64, 64.0)
    sc = _scene([(0, 0, 4)], [(1.0, 0.5, 0.25)], [0.8], 0.25)
    img, st = oracle.render(sc, cam, threads=2)
    assert st["visible"] == 1
    for x, y in ((31, 31), (35, 30), (20, 40)):
        dx, dy = 31.5 - x, 31.5 - y
        a = min(0.99, 0.8 * math.exp(-0.5 * (dx * dx + dy * dy) / 16.3))
        e = a if a >= 1 / 255 else 0.0
        assert near(img[y, x, 0], e, 1e-4, 1e-6)
        assert near(img[y, x, 1], 0.5 * e, 1e-4, 1e-6)
    edge = _scene([(0.01875, 0, 4)], [(1.0, 1.0, 1.0)], [0.99], 0.25)
    tile_img, _ = oracle.render(edge, cam, bound=0)
    rad_img, _ = oracle.render(edge, cam, bound=1)
    e = 0.99 * math.exp(-0.5 * (13.2 * 13.2 + 0.5 * 0.5) / 16.3)
    assert near(tile_img[31, 45, 0], e, 1e-3)
    assert rad_img[31, 45, 0] == 0.0
    assert tile_img[31, 44, 0] == rad_img[31, 44, 0] and tile_img[31, 44, 0] > 0


def test_reference_compositing_and_early_termination(oracle):
    # test_splat_math.cpp:324-375.
    cam = oracle.axis_camera(32, 32, 32.0)
    means = [(0, 0, 2), (0, 0, 3), (0, 0, 4), (0, 0, 5)]
    colors = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
    sc = _scene(means, colors, [0.5, 0.6, 0.7, 0.8], 1.0)
    img, _ = oracle.render(sc, cam, threads=1, early_termination=False)
    splats, _ = oracle.preprocess(sc, cam, degree=0, gpu_cull=False, bounds=0)
    front = np.zeros(3)
    T = 1.0
    for s in splats:
        a = oracle.splat_alpha_at(s, 16.0, 16.0)
        front += np.asarray(s["color"], np.float64) * a * T
        T *= 1 - a
    back = np.zeros(3)
    for s in splats[::-1]:
        a = oracle.splat_alpha_at(s, 16.0, 16.0)
        back = np.asarray(s["color"], np.float64) * a + (1 - a) * back
    assert np.linalg.norm(front - back) < 1e-12
    assert np.allclose(img[16, 16], front, rtol=1e-5)
    opaque = _scene(means, colors, [0.99, 0.95, 0.95, 0.95], 4.0)
    e, st = oracle.render(opaque, cam, early_termination=True)
    assert e[16, 16, 2] == 0.0
    assert st["terminatedPixels"] > 0
    ne, _ = oracle.render(opaque, cam, early_termination=False)
    assert 0.0 < ne[16, 16, 2] < 2e-3


def test_camera_from_viewer_matches_closed_form(oracle):
    # splatCameraFromViewer: viewInria = diag(1,-1,-1) view; position = eye; focal = size / (2 tan).
    from gsr import camera

    eye = np.array([1.0, 2.0, 7.0])
    view = camera.look_at(eye, [0.2, -0.1, 0.0])
    proj = camera.perspective(60.0, 1280 / 720)
    c = oracle.OracleCamera.from_viewer(view, proj, camera.IDENTITY, 1280, 720)
    V = view.reshape(4, 4).T.astype(np.float64)
    assert np.allclose(c.view.reshape(4, 4).T, np.diag([1, -1, -1, 1]) @ V, atol=1e-6)
    assert np.allclose(c.position, eye, atol=1e-5)
    assert near(c.tan_y, math.tan(math.radians(30.0)), 1e-6)
    assert near(c.tan_x, math.tan(math.radians(30.0)) * 1280 / 720, 1e-6)
    assert near(c.focal[1], 720 / (2 * math.tan(math.radians(30.0))), 1e-5)
    assert (c.tiles_x, c.tiles_y) == (80, 45)


def test_fp32_preprocess_error_vs_fp64_small_workload(oracle):
    """The FP32 preprocess error against the FP64 twin on the small workload (100k gaussians, seed 0,
    means ~ N(0, 2^2 I), 1280x720 camera at radius 8 looking at the origin), stated per visible
    splat as max / median: centre (px, worst axis), conic (relative to the largest entry),
    colour (absolute, worst channel), depth (relative), radius (px). The GPU preprocess gives the
    FP32 oracle's bits (the parity tests), so these are the kernels' errors too. The bounds below
    are the stated figures rounded up: 1.6e-4 px, 5.1e-5, 2.0e-7, radius 0."""
    from gsr.scenes import make_scene
    from gsr import camera

    sc = make_scene(100_000, seed=0, mean_sigma=2.0)
    cam = oracle.OracleCamera.from_viewer(camera.orbit_view(8.0, 0.0, 0.0), camera.perspective(60.0, 16 / 9),
                                          camera.IDENTITY, 1280, 720)
    s32, _ = oracle.preprocess(sc, cam)
    s64 = oracle.preprocess_f64(sc, cam)
    assert np.array_equal(s32["cull_reason"], s64[:, 3].astype(np.uint32))
    vis = s32["cull_reason"] == 0
    centre = np.abs(s32["center"][vis] - s64[vis, 0:2]).max(axis=1)
    conic = (np.abs(s32["conic_opacity"][vis, :3] - s64[vis, 4:7]) /
             np.abs(s64[vis, 4:7]).max(axis=1, keepdims=True)).max(axis=1)
    colour = np.abs(s32["color"][vis] - s64[vis, 8:11]).max(axis=1)
    depth = np.abs(s32["depth"][vis] - s64[vis, 2]) / s64[vis, 2]
    radius = np.abs(s32["radius"][vis] - s64[vis, 11])
    for name, a in (("centre px", centre), ("conic rel", conic), ("colour abs", colour), ("depth rel", depth),
                    ("radius px", radius)):
        print(f"{name}: max {a.max():.3e}, median {np.median(a):.3e} over {vis.sum()} visible splats")
    assert centre.max() < 2e-4 and conic.max() < 6e-5 and colour.max() < 3e-7 and radius.max() == 0


def test_fp64_twin_agrees_with_fp32(oracle):
    from gsr.scenes import make_scene
    from gsr import camera

    sc = make_scene(2000, seed=3)
    view = camera.orbit_view(8.0, 0.0, 0.0)
    cam = oracle.OracleCamera.from_viewer(view, camera.perspective(60.0, 16 / 9), camera.IDENTITY, 1280, 720)
    s32, _ = oracle.preprocess(sc, cam)
    s64 = oracle.preprocess_f64(sc, cam)
    vis = (s32["cull_reason"] == 0) & (s64[:, 3] == 0)
    assert vis.sum() > 1000
    assert np.max(np.abs(s32["center"][vis] - s64[vis, 0:2])) < 1e-2
    rel = np.abs(s32["conic_opacity"][vis, :3] - s64[vis, 4:7]) / (np.abs(s64[vis, 4:7]) + 1e-3)
    assert np.max(rel) < 1e-2
    assert np.max(np.abs(s32["color"][vis] - s64[vis, 8:11])) < 1e-5
