"""Differentiable rendering (gsr_backward; no counterpart in the reference). The GPU gradients of
an L1 loss against a target image, with respect to means, log-scales, quaternions, opacity
logits and SH, against FP64 central finite differences (h = 1e-4) of the FP64 reference forward
(oracle/gsr_grad_oracle.cpp, pinned on the CPU by tests/test_oracle_diff.py) on 64 seeded
visible gaussians, and for every gaussian against that reference's exact gradient. The target is the render of the scene with means moved by N(0, 0.01^2) and
opacity logits by N(0, 0.1^2). The forward is discontinuous where a pixel's alpha crosses 1/255
or a splat's tile rect changes, and a central difference across such a point is not a
derivative (it then changes with the step): the comparison takes the parameters whose
differences at h = 1e-4 and 2e-5 agree within 1 % (at least 75 % of them), and on those asks
for 98 % within 2 % relative and a cosine above 0.999 per parameter group."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, __import__("os").path.dirname(__file__))
from test_oracle_diff import f32_scene, params_of, raw_scene  # noqa: E402


def log(msg):
    print(msg, flush=True)


def _camera(gsr, W=96, H=64):
    from gsr import camera

    return gsr.camera_from_viewer(camera.orbit_view(4.0, 10.0, 20.0), camera.perspective(60.0, W / H),
                                  camera.IDENTITY, W, H)


def _target(oracle, raw, ocam, seed=9):
    rng = np.random.default_rng(seed)
    means, log_scales, q, logits, sh = raw
    moved = (means + rng.normal(0, 0.01, means.shape), log_scales, q, logits + rng.normal(0, 0.1, logits.shape), sh)
    _, img = oracle.diff_render_f64(params_of(oracle, moved), ocam)
    return img.astype(np.float32)


GROUPS = [("means", 0, 3), ("log_scales", 3, 6), ("quats", 6, 10), ("opacity_logits", 10, 11), ("sh", 11, 59)]


@pytest.mark.parametrize("sort_mode,bounds,W,H", [("global", "rect", 96, 64), ("binning", "rect", 96, 64),
                                                  ("global", "tight", 96, 64), ("global", "tight", 90, 50)])
def test_backward_matches_fp64_finite_differences(gsr_ctx, oracle, sort_mode, bounds, W, H):
    """The FP64 reference reaches a splat's rect tiles; tight bounds leave out only tiles where its
    alpha stays below 1/255, so the same reference holds for them. 90x50: partial tiles and
    quadrants past the image edge."""
    import gsr

    log(f"case backward 120 gaussians {W}x{H} sort_mode={sort_mode} bounds={bounds}")
    raw = raw_scene()
    sc = f32_scene(raw)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, 3)
    cam = _camera(gsr, W, H)
    ocam = oracle.OracleCamera.from_gsr(cam)
    target = _target(oracle, raw, ocam)
    r = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target="rgba32f", sort_mode=sort_mode, differentiable=True)
    img = r.render_to_host(cam)[..., :3]
    p = params_of(oracle, raw)
    loss64, img64 = oracle.diff_render_f64(p, ocam, target)
    d = np.abs(img - img64)
    log(f"  forward vs FP64: max {d.max():.2e}, median {np.median(d):.2e}")
    assert d.max() < 1e-3
    # The exact-alpha render of a plain renderer is the same image up to the exponential's
    # rounding (the training forward takes exp as v_exp_f32 of power log2(e)).
    plain = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target="rgba32f", sort_mode=sort_mode, exact_alpha=True)
    dp = np.abs(plain.render_to_host(cam)[..., :3] - img)
    log(f"  training forward vs exact-alpha render: max {dp.max():.2e}")
    assert dp.max() <= 1e-4
    plain.close()
    r.render(cam)
    H, W = img.shape[:2]
    grad_img = (np.sign(img - target) / np.float32(3 * W * H)).astype(np.float32)
    g = r.backward_to_host(grad_img)
    splats, counts = r.read_back_splats()
    visible = np.nonzero((splats["cull_reason"] == 0) & (counts > 0))[0]
    culled = np.nonzero(splats["cull_reason"] != 0)[0]
    assert len(visible) >= 64
    for name, _, _ in GROUPS:
        assert np.all(g[name][culled] == 0), f"culled gaussians get no {name} gradient"
    which = np.random.default_rng(3).choice(visible, 64, replace=False)
    fd = oracle.diff_fd_f64(p, ocam, target, which, h=1e-4)
    # Where a central difference crosses a discontinuity it is not a derivative, and it then
    # changes with the step: compare where the differences at h and h / 5 agree within 1 %.
    fd5 = oracle.diff_fd_f64(p, ocam, target, which, h=2e-5)
    gpu = np.concatenate([g["means"][which], g["log_scales"][which], g["quats"][which],
                          g["opacity_logits"][which][:, None], g["sh"][which]], 1).astype(np.float64)
    scale = np.abs(fd).max()
    for name, lo, hi in GROUPS:
        a, b, b5 = gpu[:, lo:hi].ravel(), fd[:, lo:hi].ravel(), fd5[:, lo:hi].ravel()
        keep = np.abs(b) > 1e-4 * scale  # parameters with a gradient to compare
        stable = keep & (np.abs(b - b5) <= 0.01 * np.abs(b))
        rel = np.abs(a[stable] - b[stable]) / np.abs(b[stable])
        cos = float(a[stable] @ b[stable] / (np.linalg.norm(a[stable]) * np.linalg.norm(b[stable])))
        within = float((rel < 0.02).mean())
        log(f"  {name:15s}: {stable.sum()} of {keep.sum()} differences stable; on them cosine {cos:.6f}, "
            f"{within * 100:5.1f}% within 2%, median rel {np.median(rel):.2e}")
        assert stable.sum() >= 0.75 * keep.sum(), name
        assert cos > 0.999 and within >= 0.98, name
    r.close()
    scene.close()


@pytest.mark.parametrize("n,seed,sort_mode,bounds,W,H", [(120, 5, "global", "rect", 96, 64),
                                                          (120, 5, "binning", "tight", 90, 50),
                                                          (1500, 11, "global", "tight", 256, 144),
                                                          (1500, 11, "binning", "rect", 250, 130)])
def test_backward_matches_fp64_exact_gradient(gsr_ctx, oracle, n, seed, sort_mode, bounds, W, H):
    """Every gaussian's gradient against the FP64 reference's exact gradient
    (oracle_diff_grad_f64: dual-number Jacobians of the per-gaussian math and the composite's
    derivative, pinned against central differences on the CPU) for the same image gradient, a
    random one. Per parameter group: cosine above 0.99999 over all gaussians, and 99 % of the
    components within 1e-3 relative (plus 1e-5 of the group's largest). The two forwards differ
    only where a pixel's FP32 and FP64 alphas fall on either side of a cutoff (1/255, 0.99, the
    1e-4 transmittance), which moves a few components."""
    import gsr

    log(f"case exact gradient {n} gaussians seed {seed} {W}x{H} sort_mode={sort_mode} bounds={bounds}")
    raw = raw_scene(n=n, seed=seed)
    sc = f32_scene(raw)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, 3)
    cam = _camera(gsr, W, H)
    ocam = oracle.OracleCamera.from_gsr(cam)
    r = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target="rgba32f", sort_mode=sort_mode, differentiable=True)
    r.render(cam)
    G = (np.random.default_rng(seed + 1).normal(size=(H, W, 3)) / (3 * W * H)).astype(np.float32)
    g = r.backward_to_host(G)
    gpu = np.concatenate([g["means"], g["log_scales"], g["quats"], g["opacity_logits"][:, None], g["sh"]],
                         1).astype(np.float64)
    exact = oracle.diff_grad_f64(params_of(oracle, raw), ocam, G)
    splats, counts = r.read_back_splats()
    reached = np.abs(exact).sum(1) > 0
    log(f"  {int(reached.sum())} of {n} gaussians with a gradient; visible on the GPU "
        f"{int(((splats['cull_reason'] == 0) & (counts > 0)).sum())}")
    assert reached.sum() >= 0.5 * n
    for name, lo, hi in GROUPS:
        a, b = gpu[:, lo:hi].ravel(), exact[:, lo:hi].ravel()
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        err = np.abs(a - b)
        within = float((err <= 1e-3 * np.abs(b) + 1e-5 * np.abs(b).max()).mean())
        nz = np.abs(b) > 1e-3 * np.abs(b).max()
        log(f"  {name:15s}: cosine {cos:.8f}, {within * 100:6.2f}% within, median rel (>1e-3 of max) "
            f"{np.median(err[nz] / np.abs(b[nz])):.1e}, max abs err / max {err.max() / np.abs(b).max():.1e}")
        assert cos > 0.99999 and within >= 0.99, name
    r.close()
    scene.close()


def test_backward_nothing_visible(gsr_ctx):
    """A frame with every gaussian culled (the camera looks away): a black image and all-zero
    gradients."""
    import gsr
    from gsr import camera

    means, log_scales, q, logits, sh = raw_scene()
    sc = f32_scene((means + np.array([0.0, 0.0, 60.0]), log_scales, q, logits, sh))  # behind the camera
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, 3)
    cam = gsr.camera_from_viewer(camera.orbit_view(4.0, 0.0, 0.0), camera.perspective(60.0, 1.5), camera.IDENTITY,
                                 96, 64)
    r = gsr.Renderer(gsr_ctx, scene, target="rgba32f", differentiable=True)
    img = r.render_to_host(cam)
    assert np.all(img == 0)
    g = r.backward_to_host(np.ones((64, 96, 3), np.float32))
    for name, arr in g.items():
        assert np.all(arr == 0), name
    r.close()
    scene.close()


def test_backward_arguments(gsr_ctx):
    import gsr
    from gsr import scenes

    sc = scenes.make_scene(500, seed=1)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, 3)
    with pytest.raises(gsr.GsrError, match="RGBA32F"):
        gsr.Renderer(gsr_ctx, scene, target="rgba8", differentiable=True)
    plain = gsr.Renderer(gsr_ctx, scene, target="rgba32f")
    cam = _camera(gsr)
    plain.render_to_host(cam)
    with pytest.raises(gsr.GsrError, match="not created differentiable"):
        plain.backward_to_host(np.zeros((64, 96, 3), np.float32))
    r = gsr.Renderer(gsr_ctx, scene, target="rgba32f", differentiable=True)
    with pytest.raises(gsr.GsrError, match="no frame"):
        r.backward(1, None)
    model = gsr.ColourModel(gsr_ctx, scene)
    with pytest.raises(gsr.GsrError, match="differentiable"):
        r.set_colour_model(model)
    model.close()
    for x in (plain, r):
        x.close()
    scene.close()


def test_torch_autograd_bridge_and_training(gsr_ctx, oracle):
    """gsr.autograd.render as a torch.autograd.Function: the gradients of a torch L1 loss equal
    those of the C ABI called directly, and Adam steps on all five parameter tensors bring the
    render towards the target (the loss falls)."""
    import torch

    import gsr
    from gsr import autograd

    log("case torch autograd bridge 120 gaussians 96x64")
    raw = raw_scene()
    sc = f32_scene(raw)
    cam = _camera(gsr)
    ocam = oracle.OracleCamera.from_gsr(cam)
    target = _target(oracle, raw, ocam)
    dev = torch.device("cpu")  # the bridge moves parameters through host memory
    alpha = sc.pos_alpha[:, 3].astype(np.float64)
    init = [sc.pos_alpha[:, :3], np.log(sc.scales[:, :3].astype(np.float64)), sc.rotations, np.log(alpha / (1 - alpha)),
            sc.sh]
    params = [torch.tensor(np.asarray(a, np.float32), device=dev, requires_grad=True) for a in init]
    view = autograd.View(gsr_ctx, cam, 3, bounds="rect", sort_mode="global")
    tgt = torch.tensor(target, device=dev)
    img = autograd.render(*params, view)
    loss = (img - tgt).abs().mean()
    loss.backward()
    # The same frame through the C ABI directly (the scene from the same float32 values).
    lg = params[3].detach().numpy().astype(np.float64)
    pa = np.concatenate([params[0].detach().numpy(), (1.0 / (1.0 + np.exp(-lg))).astype(np.float32)[:, None]], 1)
    scl = np.concatenate([np.exp(params[1].detach().numpy()), np.zeros((120, 1), np.float32)], 1)
    scene = gsr.Scene(gsr_ctx, pa, scl, sc.rotations, sc.sh, 3)
    r = gsr.Renderer(gsr_ctx, scene, bounds="rect", target="rgba32f", sort_mode="global", differentiable=True)
    img2 = r.render_to_host(cam)[..., :3]
    assert np.array_equal(img2, img.detach().cpu().numpy())
    r.render(cam)
    direct = r.backward_to_host((np.sign(img2 - target) / np.float32(img2.size)).astype(np.float32))
    # Equal up to float rounding: the per-gaussian sums are float atomics across tiles, whose order
    # varies from run to run.
    for t, name in zip(params, ("means", "log_scales", "quats", "opacity_logits", "sh")):
        a, b = t.grad.cpu().numpy().reshape(direct[name].shape), direct[name]
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max() + 1e-12, name
    r.close()
    scene.close()
    # Training: Adam on every parameter tensor.
    opt = torch.optim.Adam(params, lr=2e-3)
    first = float(loss.detach())
    for step in range(60):
        opt.zero_grad()
        loss = (autograd.render(*params, view) - tgt).abs().mean()
        loss.backward()
        opt.step()
    last = float(loss.detach())
    log(f"  L1 to the target: {first:.5f} -> {last:.5f} after 60 Adam steps")
    assert last < 0.6 * first


def test_scene_from_device_pointers(gsr_ctx):
    """Scene.from_device binds caller-owned device arrays: the same image as the host-array scene."""
    import gsr
    from gsr import scenes

    sc = scenes.make_scene(3000, seed=4)
    bufs = [gsr.DeviceBuffer(gsr_ctx, a.nbytes) for a in (sc.pos_alpha, sc.scales, sc.rotations, sc.sh)]
    for b, a in zip(bufs, (sc.pos_alpha, sc.scales, sc.rotations, sc.sh)):
        b.write(a)
    dscene = gsr.Scene.from_device(gsr_ctx, 3000, *(b.ptr for b in bufs), 3)
    hscene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, 3)
    cam = _camera(gsr)
    images = []
    for scene in (dscene, hscene):
        r = gsr.Renderer(gsr_ctx, scene, target="rgba8")
        images.append(r.render_to_host(cam))
        r.close()
    assert np.array_equal(images[0], images[1])
    dscene.close()
    hscene.close()
    for b in bufs:
        b.close()


@pytest.mark.parametrize("sort_mode", ["global", "binning"])
def test_backward_500k_720p_workload(gsr_ctx, oracle, sort_mode):
    """The backward workload at its own size: 500k gaussians (means ~ N(0, 2^2 I), the small
    workload's distributions, seed 0), 1280x720, camera at radius 8, L1 loss against the render of
    the scene with means moved by N(0, 0.01^2) and opacity logits by N(0, 0.1^2). On 64 seeded
    gaussians that reach the image, the GPU gradients (means, log-scales, quaternions, opacity
    logits, SH) against the tiled FP64 references: the exact gradient for the same dL/d(image)
    (cosine above 0.99999 and 99 % within 1e-3 relative per group) and central differences
    (h = 1e-4) where those are derivatives (the differences at h and h / 5 agree within 1 %: at
    least half of them, fewer than on the small scenes because in this dense scene a move of 1e-4
    crosses more pixels' alpha cutoffs and terminations; on those a cosine above 0.999 and 98 %
    within 2 %)."""
    import gsr
    from gsr import camera, scenes

    log(f"case backward 500k-720p sort_mode={sort_mode}")
    n, W, H = 500_000, 1280, 720
    sc = scenes.make_scene(n, seed=0, mean_sigma=2.0)
    cam = gsr.camera_from_viewer(camera.orbit_view(8.0, 0.0, 0.0), camera.perspective(60.0, W / H),
                                 camera.IDENTITY, W, H)
    ocam = oracle.OracleCamera.from_gsr(cam)
    rng = np.random.default_rng(0)
    alpha = sc.pos_alpha[:, 3].astype(np.float64)
    moved = sc.pos_alpha.copy()
    moved[:, :3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    moved[:, 3] = (1.0 / (1.0 + np.exp(-(np.log(alpha / (1.0 - alpha)) + rng.normal(0.0, 0.1, n))))).astype(np.float32)
    tscene = gsr.Scene(gsr_ctx, moved, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    tr = gsr.Renderer(gsr_ctx, tscene, target="rgba32f", sort_mode=sort_mode, differentiable=True)
    target = tr.render_to_host(cam)[..., :3].copy()
    tr.close()
    tscene.close()
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(gsr_ctx, scene, target="rgba32f", sort_mode=sort_mode, differentiable=True)
    img = r.render_to_host(cam)[..., :3].copy()
    G = (np.sign(img - target) / np.float32(3 * W * H)).astype(np.float32)
    g = r.backward_to_host(G)
    gpu_all = np.concatenate([g["means"], g["log_scales"], g["quats"], g["opacity_logits"][:, None], g["sh"]], 1)
    splats, counts = r.read_back_splats()
    reach = np.nonzero((splats["cull_reason"] == 0) & (counts > 0) & (np.abs(g["sh"][:, :3]).sum(1) > 0))[0]
    assert reach.size > 100_000
    which = np.sort(np.random.default_rng(3).choice(reach, 64, replace=False))
    p = oracle.RawParams(sc.pos_alpha[:, :3].astype(np.float64), np.log(sc.scales[:, :3].astype(np.float64)),
                         sc.rotations.astype(np.float64), np.log(alpha / (1.0 - alpha)), sc.sh.astype(np.float64),
                         sc.sh_degree)
    ref = oracle.diff_local_f64(p, ocam, target, G, which, h=1e-4, want_image=True)
    d = np.abs(img - ref["image"])
    log(f"  forward vs FP64: max {d.max():.2e}, median {np.median(d):.2e}; L1 {ref['loss']:.6f}")
    assert np.median(d) < 1e-6 and (d < 1e-3).mean() > 0.9999
    fd5 = oracle.diff_local_f64(p, ocam, target, None, which, h=2e-5)["fd"]
    gpu = gpu_all[which].astype(np.float64)
    exact, fd = ref["grad"], ref["fd"]
    scale = np.abs(fd).max()
    for name, lo, hi in GROUPS:
        a, b = gpu[:, lo:hi].ravel(), exact[:, lo:hi].ravel()
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        within = float((np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5 * np.abs(b).max()).mean())
        f, f5 = fd[:, lo:hi].ravel(), fd5[:, lo:hi].ravel()
        keep = np.abs(f) > 1e-4 * scale
        stable = keep & (np.abs(f - f5) <= 0.01 * np.abs(f))
        rel = np.abs(a[stable] - f[stable]) / np.abs(f[stable])
        fcos = float(a[stable] @ f[stable] / (np.linalg.norm(a[stable]) * np.linalg.norm(f[stable])))
        log(f"  {name:15s}: exact cosine {cos:.8f}, {within * 100:6.2f}% within 1e-3; FD {stable.sum()} of "
            f"{keep.sum()} stable, cosine {fcos:.6f}, {(rel < 0.02).mean() * 100:5.1f}% within 2%, median rel "
            f"{np.median(rel):.1e}")
        assert cos > 0.99999 and within >= 0.99, name
        assert stable.sum() >= 0.5 * keep.sum() and stable.sum() >= min(10, keep.sum()), name
        assert fcos > 0.999 and (rel < 0.02).mean() >= 0.98, name
    r.close()
    scene.close()
