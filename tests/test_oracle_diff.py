"""The FP64 differentiable-render reference (oracle/gsr_grad_oracle.cpp), the checker of the GPU
backward. Pinned here, on the CPU:
  - its image against the FP32 oracle render (renderReference restated, rect bounds) of the
    same scene: the FP64 forward is the same function (differences at FP32 rounding, except at
    pixels where the two precisions take a threshold differently);
  - its central differences against two calls of the FP64 render with one parameter moved by
    +-h (the parameter order and the difference formula);
  - a culled gaussian has zero differences; a quadratic-free parameter (an SH coefficient enters
    the loss piecewise linearly) gives the same difference for h and h / 10;
  - its exact gradient (dual-number Jacobians of the per-gaussian math, the composite's
    derivative per pixel) against the central differences wherever those are derivatives (the
    differences at h and h / 5 agree), is linear in the image gradient, and is zero for a culled
    gaussian."""
import numpy as np


def _camera(oracle, W=96, H=64):
    from gsr import camera

    import gsr

    cam = gsr.camera_from_viewer(camera.orbit_view(4.0, 10.0, 20.0), camera.perspective(60.0, W / H),
                                 camera.IDENTITY, W, H)
    return oracle.OracleCamera.from_gsr(cam)


def raw_scene(n=120, seed=5, degree=3):
    rng = np.random.default_rng(seed)
    K = (degree + 1) ** 2
    means = rng.normal(0, 0.6, (n, 3))
    log_scales = rng.normal(-1.9, 0.3, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    logits = rng.normal(0, 1.0, n)
    sh = np.concatenate([rng.normal(0, 0.5, (n, 3)), rng.normal(0, 0.1, (n, (K - 1) * 3))], axis=1)
    return means, log_scales, q, logits, sh


def f32_scene(raw, degree=3):
    from gsr.scenes import SyntheticScene

    means, log_scales, q, logits, sh = raw
    n = means.shape[0]
    alpha = (1.0 / (1.0 + np.exp(-logits))).astype(np.float32)
    pa = np.concatenate([means.astype(np.float32), alpha[:, None]], 1)
    sc = np.concatenate([np.exp(log_scales).astype(np.float32), np.zeros((n, 1), np.float32)], 1)
    return SyntheticScene(np.ascontiguousarray(pa), np.ascontiguousarray(sc), np.ascontiguousarray(q, np.float32),
                          np.ascontiguousarray(sh, np.float32), degree, 0)


def params_of(oracle, raw, degree=3):
    """The raw parameters as the FP32 scene holds them (so both renders see the same values)."""
    means, log_scales, q, logits, sh = raw
    sc = f32_scene(raw, degree)
    # log / logit of the float32 values: exp / sigmoid of them give the float32 values back to ~1 ulp
    return oracle.RawParams(sc.pos_alpha[:, :3].astype(np.float64), np.log(sc.scales[:, :3].astype(np.float64)),
                            sc.rotations.astype(np.float64),
                            np.log(sc.pos_alpha[:, 3].astype(np.float64) / (1 - sc.pos_alpha[:, 3].astype(np.float64))),
                            sc.sh.astype(np.float64), degree)


def test_fp64_forward_matches_fp32_oracle(oracle):
    ocam = _camera(oracle)
    raw = raw_scene()
    p = params_of(oracle, raw)
    loss, img64 = oracle.diff_render_f64(p, ocam)
    assert loss == 0.0
    img32, st = oracle.render(f32_scene(raw), ocam, bounds=oracle.RECT)
    d = np.abs(img64 - img32)
    print(f"FP64 vs FP32 oracle: median {np.median(d):.2e}, p99.9 {np.quantile(d, 0.999):.2e}, max {d.max():.2e}")
    assert img64.max() > 0.2, "the scene covers the image"
    assert np.median(d) < 1e-6 and np.quantile(d, 0.999) < 1e-4 and d.max() < 0.02


def test_fd_matches_manual_differences(oracle):
    ocam = _camera(oracle)
    raw = raw_scene()
    p = params_of(oracle, raw)
    _, target = oracle.diff_render_f64(params_of(oracle, raw_scene(seed=6)), ocam)
    target = target.astype(np.float32)
    which = [3, 17]
    h = 1e-4
    fd = oracle.diff_fd_f64(p, ocam, target, which, h=h, threads=4)
    assert fd.shape == (2, 11 + 48)
    # parameter k of gaussian g: mean (0-2), log-scale (3-5), quaternion (6-9), logit (10), SH (11+)
    fields = [("means", 3), ("log_scales", 3), ("quats", 4), ("logits", 1), ("sh", 48)]
    for row, g in enumerate(which):
        for k in (0, 4, 7, 10, 11, 30):
            off = 0
            for name, width in fields:
                if k < off + width:
                    break
                off += width
            arr = getattr(p, name)
            flat = arr.reshape(arr.shape[0], -1)
            x0 = flat[g, k - off]
            flat[g, k - off] = x0 + h
            lp, _ = oracle.diff_render_f64(p, ocam, target)
            flat[g, k - off] = x0 - h
            lm, _ = oracle.diff_render_f64(p, ocam, target)
            flat[g, k - off] = x0
            assert fd[row, k] == (lp - lm) / (2 * h), (g, k)


def test_fd_culled_zero_and_linear_sh(oracle):
    ocam = _camera(oracle)
    raw = raw_scene()
    means = raw[0].copy()
    means[7] = [0.0, 0.0, 40.0]  # behind the camera at (orbit radius 4): near-plane culled
    raw = (means,) + raw[1:]
    p = params_of(oracle, raw)
    _, target = oracle.diff_render_f64(params_of(oracle, raw_scene(seed=6)), ocam)
    target = target.astype(np.float32)
    fd = oracle.diff_fd_f64(p, ocam, target, [7, 2], h=1e-4, threads=4)
    assert np.all(fd[0] == 0.0)
    fd2 = oracle.diff_fd_f64(p, ocam, target, [2], h=1e-5, threads=4)
    # the DC coefficients enter linearly between kinks of |.|: the same difference for both h
    assert np.allclose(fd[1, 11:14], fd2[0, 11:14], rtol=1e-4, atol=1e-9)


def test_exact_gradient_matches_fd(oracle):
    ocam = _camera(oracle)
    raw = raw_scene()
    means = raw[0].copy()
    means[7] = [0.0, 0.0, 40.0]  # culled
    raw = (means,) + raw[1:]
    p = params_of(oracle, raw)
    _, target = oracle.diff_render_f64(params_of(oracle, raw_scene(seed=6)), ocam)
    target = target.astype(np.float32)
    _, img = oracle.diff_render_f64(p, ocam, target)
    G = (np.sign(img - target) / (3 * ocam.width * ocam.height)).astype(np.float32)
    exact = oracle.diff_grad_f64(p, ocam, G)
    assert exact.shape == (120, 59)
    assert np.all(exact[7] == 0)
    which = [g for g in range(120) if g != 7][:24]
    fd = oracle.diff_fd_f64(p, ocam, target, which, h=1e-4, threads=4)
    fd5 = oracle.diff_fd_f64(p, ocam, target, which, h=2e-5, threads=4)
    a = exact[which]
    scale = np.abs(fd).max()
    for name, lo, hi in [("means", 0, 3), ("log_scales", 3, 6), ("quats", 6, 10), ("logits", 10, 11),
                         ("sh", 11, 59)]:
        A, B, B5 = a[:, lo:hi].ravel(), fd[:, lo:hi].ravel(), fd5[:, lo:hi].ravel()
        keep = np.abs(B) > 1e-6 * scale
        stable = keep & (np.abs(B - B5) <= 0.01 * np.abs(B))
        rel = np.abs(A[stable] - B[stable]) / np.abs(B[stable])
        print(f"{name:10s}: {stable.sum()} of {keep.sum()} stable, median rel {np.median(rel):.1e}, "
              f"max {rel.max():.1e}")
        assert stable.sum() >= 0.9 * keep.sum(), name
        assert np.median(rel) < 1e-6 and np.quantile(rel, 0.95) < 1e-4 and rel.max() < 1e-2, name


def test_exact_gradient_is_linear_in_image_gradient(oracle):
    ocam = _camera(oracle)
    p = params_of(oracle, raw_scene())
    rng = np.random.default_rng(2)
    shape = (ocam.height, ocam.width, 3)
    # small integers: G1 + 2 G2 is exact in float32, so only the FP64 sums' rounding remains
    G1 = rng.integers(-8, 9, size=shape).astype(np.float32)
    G2 = rng.integers(-8, 9, size=shape).astype(np.float32)
    g1, g2 = oracle.diff_grad_f64(p, ocam, G1), oracle.diff_grad_f64(p, ocam, G2)
    g12 = oracle.diff_grad_f64(p, ocam, G1 + 2 * G2)
    assert np.abs(g12 - (g1 + 2 * g2)).max() <= 1e-12 * np.abs(g12).max()
    assert np.all(oracle.diff_grad_f64(p, ocam, np.zeros(shape, np.float32)) == 0)


def test_tiled_fp64_references_equal_the_full_ones(oracle):
    """The tiled FP64 references (per-tile lists; central differences from re-renders of the tiles the
    moved gaussian reaches) give the full references' image, loss, gradient rows and central
    differences: the same arithmetic per pixel, only the splats that cannot reach the pixel's tile
    (and the pixels a move cannot change) are skipped."""
    raw = list(raw_scene(n=300, seed=21))
    cam = _camera(oracle, 128, 80)
    raw[0] = raw[0].copy()
    raw[0][:2] = 3.0 * np.asarray(cam.position, np.float64)  # behind the camera: culled at the near plane
    p = params_of(oracle, raw)
    rng = np.random.default_rng(4)
    target = rng.uniform(0, 0.6, (cam.height, cam.width, 3)).astype(np.float32)
    G = (rng.normal(size=(cam.height, cam.width, 3)) / (3 * cam.width * cam.height)).astype(np.float32)
    loss, img = oracle.diff_render_f64(p, cam, target)
    exact = oracle.diff_grad_f64(p, cam, G)
    which = np.nonzero(np.abs(exact).sum(1) > 0)[0][:24]
    s32, _ = oracle.preprocess(f32_scene(raw), cam)
    assert np.all(s32["cull_reason"][:2] == 1)
    which = np.concatenate([which[which > 1], [0, 1]])  # the two culled ones too
    fd = oracle.diff_fd_f64(p, cam, target, which, h=1e-4)
    loc = oracle.diff_local_f64(p, cam, target, G, which, h=1e-4, want_image=True)
    assert np.array_equal(loc["image"], img)
    assert abs(loc["loss"] - loss) <= 1e-12 * loss
    assert np.allclose(loc["grad"], exact[which], rtol=1e-10, atol=1e-14 * np.abs(exact).max())
    # the full differences subtract two sums over every pixel (rounding ~1e-9 of the largest)
    assert np.allclose(loc["fd"], fd, rtol=1e-7, atol=1e-8 * np.abs(fd).max())
    assert np.all(loc["grad"][-2:] == 0) and np.all(loc["fd"][-2:] == 0)
