"""Oracle pins for StopThePop (reference_renderer.cpp:61-70, 143-196; splat_math.cpp:195-205).

- computeRayDepthParams: P and b against numpy's float64 inverse of the camera-space
  covariance W R S^2 R^T W^T (an independent formula), and t* = (v.b) / (v^T P v) against a
  brute-force maximum of the gaussian along the ray;
- the window: with K at least the number of splats any pixel sees it never fills, so the image
  equals the exact per-pixel sort; K = 0 is INRIA's order; the exact per-pixel sort equals a
  numpy brute force (ray depths, stable sort, INRIA's blend) on a small scene.
"""
import numpy as np


def _rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_ray_depth_params_and_peak(oracle):
    from gsr import camera

    rng = np.random.default_rng(3)
    view = camera.look_at([1.0, 2.0, 7.0], [0, 0, 0])
    cam = oracle.OracleCamera.from_viewer(view, camera.perspective(60.0, 16 / 9), camera.IDENTITY, 640, 360)
    V = cam.view.reshape(4, 4).T.astype(np.float64)  # row-major viewInria
    W = V[:3, :3]
    for _ in range(200):
        sigma = np.exp(rng.uniform(-3, 0, 3)).astype(np.float32)
        q = rng.normal(size=4).astype(np.float32)
        q /= np.linalg.norm(q)
        mean = rng.normal(0, 1.5, 3)
        t = (V @ np.append(mean, 1.0))[:3].astype(np.float32)
        got = oracle.ray_depth(sigma, q, t, cam.view)
        R = _rot(q.astype(np.float64))
        cov = W @ R @ np.diag(sigma.astype(np.float64) ** 2) @ R.T @ W.T
        P = np.linalg.inv(cov)
        b = P @ t.astype(np.float64)
        want = np.array([P[0, 0], P[0, 1], P[0, 2], P[1, 1], P[1, 2], P[2, 2], *b])
        assert np.allclose(got, want, rtol=2e-3, atol=1e-3 * np.abs(want).max())
        # t* maximises the gaussian along the ray t (vx, vy, 1)
        v = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.3, 0.3), 1.0])
        p00, p01, p02, p11, p12, p22, b0, b1, b2 = got.astype(np.float64)
        tstar = (b0 * v[0] + b1 * v[1] + b2) / (p00 * v[0] ** 2 + 2 * p01 * v[0] * v[1] + 2 * p02 * v[0] +
                                               p11 * v[1] ** 2 + 2 * p12 * v[1] + p22)
        ts = np.linspace(tstar - 1.0, tstar + 1.0, 20001)
        d = ts[:, None] * v[None, :] - t.astype(np.float64)[None, :]
        maha = np.einsum("ni,ij,nj->n", d, P, d)
        assert abs(ts[np.argmin(maha)] - tstar) < 2e-4


def _small(oracle, n=300, W=96, H=64):
    from gsr import camera, scenes

    sc = scenes.make_scene(n, seed=5, mean_sigma=0.6, log_scale_mean=-2.5)
    cam = oracle.OracleCamera.from_viewer(camera.look_at([0.5, 0.8, 4.0], [0, 0, 0]), camera.perspective(60.0, W / H),
                                          camera.IDENTITY, W, H)
    return sc, cam


def test_window_limits(oracle):
    sc, cam = _small(oracle)
    inria, _ = oracle.render(sc, cam, bounds=oracle.TIGHT)
    k0, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, sort_window=0)
    assert np.array_equal(inria, k0)
    exact, st = oracle.render(sc, cam, bounds=oracle.TIGHT, sort_window=oracle.EXACT_PIXEL_SORT)
    big, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, sort_window=4096)
    assert np.array_equal(exact, big), "a window that never fills is the exact sort"
    k4, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, sort_window=4)
    assert not np.array_equal(exact, inria)
    assert oracle.psnr(k4, exact) > oracle.psnr(inria, exact)


def test_exact_sort_matches_brute_force(oracle):
    sc, cam = _small(oracle, n=120, W=48, H=32)
    exact, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, sort_window=oracle.EXACT_PIXEL_SORT)
    sp, counts = oracle.preprocess(sc, cam, bounds=oracle.TIGHT, gpu_cull=False)
    V = cam.view.reshape(4, 4).T.astype(np.float64)
    vis = np.nonzero(sp["cull_reason"] == 0)[0]
    params = {}
    for i in vis:
        t = (V @ np.append(sc.pos_alpha[i, :3].astype(np.float64), 1.0))[:3]
        params[i] = oracle.ray_depth(sc.scales[i, :3], sc.rotations[i], t.astype(np.float32), cam.view)
    f = np.array(cam.focal, np.float64)
    img = np.zeros((cam.height, cam.width, 3))
    for y in range(cam.height):
        for x in range(cam.width):
            vx = (x - 0.5 * (cam.width - 1)) / f[0]
            vy = (y - 0.5 * (cam.height - 1)) / f[1]
            frags = []
            for i in vis:
                s = sp[i]
                dx, dy = s["center"][0] - x, s["center"][1] - y
                a, b, c, o = s["conic_opacity"].astype(np.float64)
                power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
                if power > 0:
                    continue
                alpha = min(0.99, o * np.exp(power))
                if alpha < 1 / 255:
                    continue
                p00, p01, p02, p11, p12, p22, b0, b1, b2 = params[i].astype(np.float64)
                depth = (b0 * vx + b1 * vy + b2) / (p00 * vx * vx + 2 * p01 * vx * vy + 2 * p02 * vx + p11 * vy * vy +
                                                   2 * p12 * vy + p22)
                frags.append((depth, alpha, s["color"].astype(np.float64)))
            frags.sort(key=lambda e: e[0])
            T, C = 1.0, np.zeros(3)
            for _, alpha, col in frags:
                if T * (1 - alpha) < 1e-4:
                    break
                C += col * alpha * T
                T *= 1 - alpha
            img[y, x] = C
    assert oracle.psnr(exact, img) > 70.0
