"""Benchmark of the splat-tiles forward path on one MI355X.

Workload (the headline configuration): 1.2M synthetic gaussians (means ~ N(0, 3^2 I), SH
degree 3, seed 0), 1920x1080, the 300-frame 360-degree orbit at radius 10 and elevation
15 degrees. A step is one frame: preprocess, scan, tile-list build and sort (binning mode:
compaction, 32-bit depth sort, duplicate, 32-bit tile sort; global mode: duplicate, 64-bit
(tile, depth) sort), ranges and render to RGBA8, with the scene already resident in HBM.
W warm-up frames are rendered first (the reference excludes frames/10 + 5 = 35), then K
frames are timed with HIP events on the render stream.

Prints ONE JSON line: ms/frame (value), per-stage ms, sorted keys/s, the roofline of the
dominant kernel, the fp16-scene and other-sort-mode frame times, the image quality of one
frame against the CPU oracle, and the oracle's own CPU time on the same frame (cpu_baseline);
also the batched multi-view throughput (32 cameras of mixed sizes on a 1M-gaussian scene), the
orbit's throughput with 2 and 4 frames in flight, the learned colour on a 3M-gaussian scene
(training time and error, fallback fraction at 2/255, ms/frame and PSNR against plain SH), and
the differentiable render (500k gaussians, 1280x720, L1 loss: forward and backward ms).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gaussian-splat-renderer_amd"))

METRIC = "1080p ms/frame @1.2M splats at >=58.7 dB vs CPU ref; learned vs plain SH"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
ORBIT_FRAMES = 300


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=35)
    ap.add_argument("--count", type=int, default=1_200_000)
    ap.add_argument("--bounds", default="tight")
    ap.add_argument("--sort-mode", default="global", choices=("binning", "global"))
    ap.add_argument("--render-splat", default="full", choices=("full", "p32", "p24"))
    ap.add_argument("--ply", default=None, help="render an INRIA .ply scene instead of the synthetic one")
    ap.add_argument("--radius", type=float, default=10.0, help="orbit radius")
    ap.add_argument("--elevation", type=float, default=15.0, help="orbit elevation, degrees")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-multiview", action="store_true", help="skip the batched multi-view measurement")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-colour measurement")
    ap.add_argument("--no-backward", action="store_true", help="skip the differentiable-rendering measurement")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    ap.add_argument("--profile-learned", type=int, default=0,
                    help="for rocprofv3 runs of the learned colour: render this many orbit frames of a 3M-gaussian "
                         "scene plain, train 8 steps of 32768 samples, render them with the model marking every "
                         "gaussian, and exit")
    ap.add_argument("--profile-backward", type=int, default=0,
                    help="for rocprofv3 runs of the differentiable render: this many forward + backward iterations "
                         "of the 500k-gaussian 1280x720 workload, then exit")
    ap.add_argument("--profile-frames", type=int, default=0,
                    help="for rocprofv3 counter runs: render this many orbit frames of the configured mode, then "
                         "copy 256 MiB device to device twice (the counters' scale reference) and exit")
    return ap.parse_args()


class HipTimer:
    """hipEvent pair on a given stream (libamdhip64 of the library's runtime)."""

    def __init__(self):
        try:  # the runtime libgsr.so already loaded (its SONAME)
            self.hip = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.start, self.stop = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(self.start)) == 0
        assert self.hip.hipEventCreate(ctypes.byref(self.stop)) == 0

    def record_start(self, stream=None):
        assert self.hip.hipEventRecord(self.start, stream) == 0

    def record_stop(self, stream=None):
        assert self.hip.hipEventRecord(self.stop, stream) == 0

    def elapsed_ms(self):
        assert self.hip.hipEventSynchronize(self.stop) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.start, self.stop) == 0
        return ms.value


def multiview_bench(ctx, views=32, count=1_000_000, reps=5):
    """Batched multi-view: `views` cameras of mixed sizes ({640x360, 1280x720, 1920x1080}, uniform,
    seeded) on a shell of radius 8..12 around a 1M-gaussian scene of the same distributions,
    looking at the origin. Sequential = one renderer, one gsr_render per view on one stream;
    batched = gsr_batch in the configuration the autotuner picks (tile size 16 or 8, render batch
    depth, 1, 2, 4 or 8 streams, stream or persistent mode: one renderer per view, every view's
    frame up to the tile sort on the branches, then one render launch over all views' tiles), the
    whole batch captured into one HIP graph and replayed. Both: `reps` timed batches after a
    warm-up, host clock around them with the device synchronised."""
    import numpy as np

    import gsr
    from gsr import camera, scenes

    rng = np.random.default_rng(5)
    sizes = [(640, 360), (1280, 720), (1920, 1080)]
    cams = []
    for _ in range(views):
        W, H = sizes[int(rng.integers(0, 3))]
        d = rng.normal(size=3)
        eye = d / np.linalg.norm(d) * rng.uniform(8.0, 12.0)
        cams.append(gsr.camera_from_viewer(camera.look_at(eye, [0.0, 0.0, 0.0]), camera.perspective(60.0, W / H),
                                           camera.IDENTITY, W, H))
    pixels = sum(int(c.width) * int(c.height) for c in cams)
    sc = scenes.make_scene(count, seed=0, mean_sigma=3.0)
    scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    outs = [gsr.DeviceBuffer(ctx, int(c.width) * int(c.height) * 4) for c in cams]

    def make(tile_size=16, render_batch=0):
        return gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global", tile_size=tile_size,
                            render_batch=render_batch)

    def timed(fn):
        ctx.synchronize()
        t0 = time.perf_counter()
        fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3

    seq = make()
    for c, o in zip(cams, outs):  # warm-up: every size once
        seq.render(c, o.ptr)

    def sequential():
        for _ in range(reps):
            for c, o in zip(cams, outs):
                seq.render(c, o.ptr)
    seq_ms = timed(sequential) / reps
    seq.close()
    best, by = gsr.tune_batch(ctx, make, cams, [o.ptr for o in outs], candidates=(1, 2, 4, 8), replays=reps,
                              timer=timed, tile_sizes=(16, 8), batches={16: (0, 128, 512), 8: (0, 32, 128)},
                              persistent=(False, True))
    best_ms = by[(best["tile_size"], best["render_batch"], best["branches"], best["persistent"])]
    default16 = {k[2]: v for k, v in by.items() if k[0] == 16 and k[1] == 0 and not k[3]}
    persistent = {k[2]: v for k, v in by.items() if k[3]}
    pbest = min(persistent, key=persistent.get)
    for o in outs:
        o.close()
    scene.close()
    print(f"# multiview: {views} views, {pixels / 1e6:.1f} Mpixel: sequential {seq_ms:.3f} ms, batched "
          f"{best_ms:.3f} ms with {best} ({len(by)} configurations tuned; 16-px default batch: "
          f"{', '.join(f'{k}: {v:.3f}' for k, v in default16.items())}); persistent render "
          f"{persistent[pbest]:.3f} ms with {pbest} branches "
          f"({', '.join(f'{k}: {v:.3f}' for k, v in persistent.items())})", file=sys.stderr, flush=True)
    return {"views": views, "gaussians": count, "sizes": "uniform of 640x360, 1280x720, 1920x1080",
            "megapixels": round(pixels / 1e6, 2), "sequential_ms": round(seq_ms, 3),
            "sequential_mpix_per_s": round(pixels / 1e6 / (seq_ms * 1e-3), 1), "batched_ms": round(best_ms, 3),
            "batched_mpix_per_s": round(pixels / 1e6 / (best_ms * 1e-3), 1), "autotuned": best,
            "batched_ms_by_config": {f"tile{k[0]}-batch{k[1] or 'default'}-branches{k[2]}"
                                     f"{'-persistent' if k[3] else ''}": round(v, 3) for k, v in sorted(by.items())},
            "persistent_ms": round(persistent[pbest], 3), "persistent_branches": pbest,
            "persistent_mpix_per_s": round(pixels / 1e6 / (persistent[pbest] * 1e-3), 1),
            "persistent_ms_by_branches": {str(k): round(v, 3) for k, v in persistent.items()},
            "graph": True}


PSNR_BOUND_DB = 35.0


def learned_bench(ctx, cams, frames, count=3_000_000, samples=1 << 24):
    """Learned colour on two 3M-gaussian scenes of the same distributions (seed 0, means ~ N(0, 3^2 I),
    SH degree 3): the benchmark's (independent SH coefficients per gaussian) and the low-rank one
    (each gaussian's view dependence a mix of 2 patterns shared by the scene). Per scene: a colour
    model trained on `samples` (gaussian, direction, SH RGB) samples (512 Adam steps), the 1080p orbit
    (global sort mode, full records) rendered plain and with the model at fallback thresholds swept
    over quantiles of the per-gaussian worst error (and at 2/255): ms/frame (wall clock around
    `frames` synchronised frames), PSNR of frame 50 against the plain SH render, and the colour bytes
    read per gaussian (plain: the 48 SH floats; with the model 1 + 32 f + 192 (1 - f) for a learned
    fraction f). The operating point at the PSNR bound is the form with the fewest colour bytes whose
    frame 50 stays within it (plain included)."""
    import numpy as np

    import gsr
    from gsr import scenes

    W, H = int(cams[0].width), int(cams[0].height)
    out = gsr.DeviceBuffer(ctx, W * H * 4)

    def orbit(r):
        for i in range(10):
            r.render(cams[i], out.ptr)
        ctx.synchronize()
        t0 = time.perf_counter()
        for i in range(frames):
            r.render(cams[(35 + i) % ORBIT_FRAMES], out.ptr)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / frames

    def psnr(a, b):
        mse = float(np.mean((a[..., :3].astype(np.float64) / 255.0 - b[..., :3].astype(np.float64) / 255.0) ** 2))
        return None if mse == 0 else round(10.0 * np.log10(1.0 / mse), 2)

    res = {"gaussians": count, "training_samples": samples, "psnr_bound_db": PSNR_BOUND_DB,
           "plain_colour_bytes_per_gaussian": 192, "shared_mlp_bytes": 4 * gsr.ColourModel.param_count(),
           "scenes": {}}
    for name, make in (("independent_sh", lambda: scenes.make_scene(count, seed=0, mean_sigma=3.0)),
                       ("lowrank_sh", lambda: scenes.make_lowrank_scene(count, seed=0, rank=2))):
        sc = make()
        scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
        del sc
        plain = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
        plain_ms = orbit(plain)
        ref = plain.render_to_host(cams[50])
        model = gsr.ColourModel(ctx, scene, seed=0)
        step_samples = 1 << 15
        ctx.synchronize()
        t0 = time.perf_counter()
        last = model.train(samples // step_samples, step_samples, lr_weights=4e-3, lr_latents=1e-2)
        train_ms = (time.perf_counter() - t0) * 1e3
        mse = model.loss(0x5EED, 1 << 20)
        r = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
        r.set_colour_model(model)
        _, err = model.fallback(float("inf"))
        forms = {}
        sweep = [("fallback_2_255", 2.0 / 255.0)] + [(f"quantile_{q:g}", float(np.quantile(err, q)))
                                                      for q in (0.05, 0.1, 0.25, 0.5)] + [("all_learned", float("inf"))]
        for form, threshold in sweep:
            t0 = time.perf_counter()
            n, _ = model.fallback(threshold)
            fallback_ms = (time.perf_counter() - t0) * 1e3
            ms = orbit(r)
            img = r.render_to_host(cams[50])
            f = n / count
            forms[form] = {"threshold": round(threshold, 5) if np.isfinite(threshold) else None,
                           "learned_fraction": round(f, 4), "ms_per_frame": round(ms, 4),
                           "throughput_vs_plain": round(plain_ms / ms, 3), "psnr_db_vs_plain_sh": psnr(img, ref),
                           "colour_bytes_per_gaussian": round(1.0 + 32.0 * f + 192.0 * (1.0 - f), 1),
                           "fallback_pass_ms": round(fallback_ms, 2)}
        ok = {"plain": {"learned_fraction": 0.0, "ms_per_frame": round(plain_ms, 4), "colour_bytes_per_gaussian": 192.0}}
        ok.update({k: v for k, v in forms.items() if v["psnr_db_vs_plain_sh"] is None
                   or v["psnr_db_vs_plain_sh"] >= PSNR_BOUND_DB})
        best = min(ok, key=lambda k: ok[k]["colour_bytes_per_gaussian"])
        res["scenes"][name] = {"training_ms": round(train_ms, 1), "train_last_mse": round(last, 6),
                               "eval_mse": round(mse, 6), "median_worst_error": round(float(np.median(err)), 4),
                               "plain_ms_per_frame": round(plain_ms, 4), "forms": forms,
                               "at_psnr_bound": {"form": best, **ok[best]}}
        print(f"# learned colour, {name} ({count} gaussians): trained 2^{int(np.log2(samples))} samples in "
              f"{train_ms:.0f} ms, eval MSE {mse:.5f}; plain {plain_ms:.3f} ms/frame; "
              + "; ".join(f"{k}: {v['learned_fraction'] * 100:.1f}% learned, {v['ms_per_frame']:.3f} ms/frame, "
                          f"PSNR vs plain {v['psnr_db_vs_plain_sh']}" for k, v in forms.items())
              + f"; at {PSNR_BOUND_DB} dB: {best} ({ok[best]['colour_bytes_per_gaussian']} colour bytes per gaussian)",
              file=sys.stderr, flush=True)
        for x in (r, plain, model, scene):
            x.close()
    out.close()
    return res


def backward_bench(ctx, count=500_000, W=1280, H=720, reps=20):
    """Differentiable rendering: `count` gaussians of the small workload's distributions (seed 0,
    means ~ N(0, 2^2 I)), 1280x720 camera at radius 8 looking at the origin, L1 loss against the
    render of the same scene with means moved by N(0, 0.01^2) and opacity logits by N(0, 0.1^2).
    Times (wall clock around `reps` synchronised iterations) the differentiable forward alone,
    gsr_backward alone and both, with dL/d(image) of the frame resident on the device (its
    element-wise sign is not timed)."""
    import numpy as np

    import gsr
    from gsr import camera, scenes

    sc = scenes.make_scene(count, seed=0, mean_sigma=2.0)
    cam = gsr.camera_from_viewer(camera.orbit_view(8.0, 0.0, 0.0), camera.perspective(60.0, W / H),
                                 camera.IDENTITY, W, H)
    rng = np.random.default_rng(0)
    alpha = sc.pos_alpha[:, 3].astype(np.float64)
    logit = np.log(alpha / (1.0 - alpha)) + rng.normal(0.0, 0.1, count)
    moved = sc.pos_alpha.copy()
    moved[:, :3] += rng.normal(0.0, 0.01, (count, 3)).astype(np.float32)
    moved[:, 3] = (1.0 / (1.0 + np.exp(-logit))).astype(np.float32)
    tscene = gsr.Scene(ctx, moved, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    tr = gsr.Renderer(ctx, tscene, target="rgba32f", sort_mode="global", differentiable=True)
    target = tr.render_to_host(cam)[..., :3]
    tr.close()
    tscene.close()
    scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(ctx, scene, target="rgba32f", sort_mode="global", differentiable=True)
    img = r.render_to_host(cam)[..., :3]
    loss = float(np.abs(img - target).mean())
    gimg = np.zeros((H, W, 4), np.float32)
    gimg[..., :3] = np.sign(img - target) / np.float32(img.size)
    gbuf = gsr.DeviceBuffer(ctx, gimg.nbytes)
    gbuf.write(gimg)
    out = gsr.DeviceBuffer(ctx, W * H * 16)
    K = (sc.sh_degree + 1) ** 2
    grads = {k: gsr.DeviceBuffer(ctx, count * 4 * w) for k, w in
             (("means", 3), ("log_scales", 3), ("quats", 4), ("opacity_logits", 1), ("sh", 3 * K))}
    ptrs = {k: b.ptr for k, b in grads.items()}

    def timed(fn):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    fwd = timed(lambda: r.render(cam, out.ptr))
    r.render(cam, out.ptr)
    bwd = timed(lambda: r.backward(gbuf.ptr, **ptrs))

    def both():
        r.render(cam, out.ptr)
        r.backward(gbuf.ptr, **ptrs)

    total = timed(both)
    plain = gsr.Renderer(ctx, scene, target="rgba32f", sort_mode="global")
    plain_ms = timed(lambda: plain.render(cam, out.ptr))
    g_means = grads["means"].read(np.empty((count, 3), np.float32))
    res = {"gaussians": count, "width": W, "height": H, "l1_loss": round(loss, 6),
           "forward_ms": round(fwd, 4), "backward_ms": round(bwd, 4), "forward_backward_ms": round(total, 4),
           "plain_forward_ms": round(plain_ms, 4),
           "gaussians_with_mean_gradient": int((np.abs(g_means).sum(1) > 0).sum())}
    print(f"# backward ({count} gaussians, {W}x{H}): forward {fwd:.3f} ms, backward {bwd:.3f} ms, both "
          f"{total:.3f} ms (plain forward {plain_ms:.3f} ms); L1 {loss:.5f}", file=sys.stderr, flush=True)
    for b in list(grads.values()) + [gbuf, out]:
        b.close()
    plain.close()
    r.close()
    scene.close()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this renderer runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import gsr
    from gsr import camera, scenes

    W, H = 1920, 1080
    t0 = time.time()
    if args.ply:
        from gsr.scenes import SyntheticScene

        ply = gsr.load_ply(args.ply)
        sc = SyntheticScene(ply.pos_alpha, ply.scales, ply.rotations, ply.sh, ply.sh_degree, 0)
        args.count = sc.count
        workload = f"ply:{os.path.basename(args.ply)}-1080p-orbit"
        print(f"# {args.ply}: {sc.count} gaussians, SH degree {sc.sh_degree}, loaded in {time.time() - t0:.1f} s",
              file=sys.stderr, flush=True)
    else:
        sc = scenes.make_scene(args.count, seed=0, mean_sigma=3.0)
        workload = "1.2M-1080p-orbit" if args.count == 1_200_000 else f"{args.count}-1080p-orbit"
        print(f"# scene generated in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    proj = camera.perspective(60.0, W / H)
    cams = [gsr.camera_from_viewer(v, proj, camera.IDENTITY, W, H)
            for v in camera.orbit_frames(ORBIT_FRAMES, args.radius, args.elevation)]

    ctx = gsr.Context(0)
    out = gsr.DeviceBuffer(ctx, W * H * 4)
    timer = HipTimer()
    if args.profile_frames > 0:
        scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
        r = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", sort_mode=args.sort_mode,
                         render_splat=args.render_splat)
        for i in range(args.profile_frames):
            r.render(cams[(args.warmup + 9 * i) % ORBIT_FRAMES], out.ptr)
        ctx.synchronize()
        print(f"# {args.profile_frames} frames rendered", file=sys.stderr, flush=True)
        nbytes = 256 << 20
        a, b = gsr.DeviceBuffer(ctx, nbytes), gsr.DeviceBuffer(ctx, nbytes)
        for _ in range(2):
            assert timer.hip.hipMemcpy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), ctypes.c_size_t(nbytes), 3) == 0
        ctx.synchronize()
        print("# 2 x 256 MiB copies done", file=sys.stderr, flush=True)
        return

    if args.profile_backward > 0:
        bw = backward_bench(ctx, reps=args.profile_backward)
        print(f"# {args.profile_backward} forward + backward iterations: {bw}", file=sys.stderr, flush=True)
        return
    if args.profile_learned > 0:
        big = scenes.make_scene(3_000_000, seed=0, mean_sigma=3.0)
        scene = gsr.Scene(ctx, big.pos_alpha, big.scales, big.rotations, big.sh, big.sh_degree)
        del big
        plain = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
        for i in range(args.profile_learned):
            plain.render(cams[(args.warmup + i) % ORBIT_FRAMES], out.ptr)
        ctx.synchronize()
        model = gsr.ColourModel(ctx, scene, seed=0)
        model.train(8, 1 << 15, lr_weights=4e-3, lr_latents=1e-2)  # 8 steps of the workload's training
        marked, _ = model.fallback(float("inf"))
        r = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
        r.set_colour_model(model)
        for i in range(args.profile_learned):
            r.render(cams[(args.warmup + i) % ORBIT_FRAMES], out.ptr)
        ctx.synchronize()
        print(f"# {args.profile_learned} plain and {args.profile_learned} learned frames rendered "
              f"({marked} gaussians learned)", file=sys.stderr, flush=True)
        r.close()
        plain.close()
        model.close()
        scene.close()
        return

    def timed_orbit(renderer, first, steps):
        hip_stream = ctypes.c_void_p()
        assert timer.hip.hipStreamCreateWithFlags(ctypes.byref(hip_stream), 1) == 0  # non-blocking, default priority
        for i in range(args.warmup):
            renderer.render(cams[(first + i) % ORBIT_FRAMES], out.ptr, hip_stream.value)
        assert timer.hip.hipStreamSynchronize(hip_stream) == 0
        timer.record_start(hip_stream)
        for i in range(steps):
            renderer.render(cams[(first + args.warmup + i) % ORBIT_FRAMES], out.ptr, hip_stream.value)
        timer.record_stop(hip_stream)
        ms = timer.elapsed_ms()
        assert timer.hip.hipStreamSynchronize(hip_stream) == 0
        assert timer.hip.hipStreamDestroy(hip_stream) == 0
        return ms

    n_prof = min(args.steps, 60)
    prof_cams = [cams[(args.warmup + i * (ORBIT_FRAMES // n_prof)) % ORBIT_FRAMES] for i in range(n_prof)]

    def profile_stages(scene, mode, full_key_sort=False):
        """Per-stage and per-kernel ms and list counts over n_prof frames spread over the orbit: a
        second renderer with events between the stages and after every kernel (profile_stages=2).
        Kernels launched several times per frame (the radix passes) are summed per frame, with
        their launch count."""
        rp = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", profile_stages=2, sort_mode=mode,
                          render_splat=args.render_splat, full_key_sort=full_key_sort)
        stage_sum = {k: 0.0 for k in gsr.STAGES}
        kern_sum, kern_calls = {}, {}
        entries = vis = 0
        culls = np.zeros(5, np.float64)
        for cam in prof_cams:
            rp.render(cam, out.ptr)
            st = rp.stage_times()
            ctl = rp.read_control()
            for k in stage_sum:
                stage_sum[k] += st[k]
            calls = {}
            for name, ms in rp.kernel_times():
                kern_sum[name] = kern_sum.get(name, 0.0) + ms
                calls[name] = calls.get(name, 0) + 1
            kern_calls.update(calls)
            entries += ctl.entries
            vis += ctl.visible
            culls += np.array(ctl.cull_counts[:5], np.float64)
        rp.close()
        kernels = {k: (v / n_prof, kern_calls[k]) for k, v in kern_sum.items()}
        return ({k: v / n_prof for k, v in stage_sum.items()}, entries / n_prof, vis / n_prof, culls / n_prof,
                kernels)

    def render_pairs(scene, mode):
        """The render's per-tile statistics (tile_stats build) summed per frame, mean over the same
        frames: list entries, fetched entries, evaluated and blended (pixel, splat) pairs of the
        reference's loop, terminated pixels, and the lane slots of this library's loop."""
        rs = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", sort_mode=mode,
                          render_splat=args.render_splat, tile_stats=True)
        tot = np.zeros(len(gsr.TILE_STATS), np.float64)
        for cam in prof_cams:
            rs.render(cam, out.ptr)
            tot += rs.tile_stats().astype(np.float64).sum(axis=0)
        rs.close()
        return {k: v / n_prof for k, v in zip(gsr.TILE_STATS, tot)}

    other = "global" if args.sort_mode == "binning" else "binning"
    packed = [m for m in ("p32", "p24") if m != args.render_splat]
    runs = [("f32", args.sort_mode, args.render_splat, 0), ("f16", args.sort_mode, args.render_splat, 0),
            ("f32", other, args.render_splat, 0)] + [("f32", args.sort_mode, m, 0) for m in packed] + \
        [("f32", args.sort_mode, args.render_splat, k) for k in (4, 16, 64)]
    result, modes, splat_modes, windows = {}, {}, {}, {}
    for fmt, mode, rs, window in runs:
        scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree, fmt=fmt)
        r = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", sort_mode=mode, render_splat=rs,
                         sort_window=window)
        ms = timed_orbit(r, 0, args.steps)
        result[(fmt, mode, rs, window)] = ms / args.steps
        print(f"# {fmt} {mode} {rs} window {window}: {ms / args.steps:.4f} ms/frame over {args.steps} frames",
              file=sys.stderr, flush=True)
        if window:
            windows[str(window)] = round(ms / args.steps, 4)
        elif rs != args.render_splat:
            splat_modes[rs] = round(ms / args.steps, 4)
        elif fmt == "f32":
            st, ent, vis, culls, kern = profile_stages(scene, mode)
            modes[mode] = {"ms_per_frame": round(ms / args.steps, 4), "stages_ms": {k: round(v, 4) for k, v in st.items()},
                           "mean_entries": int(ent), "mean_visible": int(vis)}
            if mode == args.sort_mode:
                stages, mean_entries, mean_visible, mean_culls, kernels = st, ent, vis, culls, kern
                pairs = render_pairs(scene, mode)
                if mode == "global":  # the full 64-bit (tile, depth) key radix sort, for its keys/s
                    fst, fent, _, _, fkern = profile_stages(scene, mode, full_key_sort=True)
                    full_sort = {"ms": fst["sort"], "entries": fent, "kernels": fkern}
                # Quality: one frame (the first timed one) vs the oracle, RGBA8.
                qcam = cams[args.warmup % ORBIT_FRAMES]
                r.render(qcam, out.ptr)
                img = out.read(np.empty((H, W, 4), np.uint8))
                qsplats, _ = r.read_back_splats()
        r.close()
        scene.close()

    # Per-kernel algorithmic bytes (one launch; the frame's mean counts). N gaussians, V visible,
    # M list entries, T tiles, G sort blocks. kPreprocess reads 16 B position + opacity per gaussian,
    # 32 B scale + rotation unless culled at the near plane or for opacity and K x 12 B of SH per
    # visible gaussian, and writes the 48-B splat + 4-B tile count per gaussian, 8 B per tile range
    # zeroed and per visible gaussian its 20-B coverage record (global mode also its 4-B depth
    # bits); kFrameSetup reads the 32-B block records; kDuplicate reads 4 B count per gaussian and
    # the coverage record per visible one and writes a 4-B tile key + 4-B value per entry; a radix
    # pass reads the 4-B keys
    # (histogram), then 8 B per entry and writes 8 B per entry (scatter); kRanges reads the keys
    # and writes 8 B per tile; the tile depth sort reads 4 B value + 4 B depth bits and writes 4 B
    # per entry; the render reads 4 B value + 48 B splat per fetched entry and 8 B per tile range
    # and writes 4 B per pixel (in global mode it also sorts its tile first: the tile depth sort's
    # 12 B per entry).
    n = args.count
    tiles = cams[0].tiles
    near, low_opacity = mean_culls[1], mean_culls[4]
    K = sc.K
    M, V = mean_entries, mean_visible
    blocks = min((int(M) + 2047) // 2048, 1024)
    key_b = 4
    pre_bytes = (n * 16 + (n - near - low_opacity) * 32 + V * K * 12 + n * 52 + tiles * 8 +
                 V * (24 if args.sort_mode == "global" else 20))
    algo = {"kPreprocess": pre_bytes,
            "kFrameSetup": (n + 255) // 256 * 32 + (n + 255) // 256 * 4 + 80,
            "kDuplicate": n * 4 + V * 20 + M * (key_b + 4),
            "kSortHistogram": M * key_b + blocks * 256 * 4,
            "kSortScan": 2 * blocks * 256 * 4,
            "kSortScatter": 2 * M * (key_b + 4) + blocks * 256 * 4,
            "kRanges": M * key_b + tiles * 8,
            "kTileDepthSort": M * 12,
            "kRenderCulled": pairs["fetched"] * 52 + tiles * 8 + W * H * 4 +
                             (M * 12 if args.sort_mode == "global" else 0)}
    traffic_by_kernel = {}
    traffic_file = os.path.join(ROOT, "profiles", "pmc_traffic_1.2M-1080p-orbit.md")
    if os.path.exists(traffic_file) and not args.ply and args.count == 1_200_000 and args.sort_mode == "global":
        for ln in open(traffic_file):
            cells = [c.strip() for c in ln.split("|")]
            if len(cells) > 5 and cells[1].startswith("k") and cells[4].replace(".", "").isdigit():
                traffic_by_kernel[cells[1].split("<")[0].split(" ")[0]] = int(float(cells[4]))
    kernel_table = {}
    for name, (ms, calls) in sorted(kernels.items(), key=lambda kv: -kv[1][0]):
        entry = {"us_per_frame": round(ms * 1e3, 2), "launches": calls}
        if name in algo:
            b = algo[name] * calls
            entry.update({"algorithmic_bytes": int(b), "gb_per_s": round(b / (ms * 1e-3) / 1e9, 1),
                          "hbm_frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)})
            if name in traffic_by_kernel:
                entry["traffic_bytes"] = traffic_by_kernel[name] * calls
        kernel_table[name] = entry

    # Roofline of the dominant kernel, the render (kRenderCulled): it is bound by VALU issue, so its
    # figure is the reference algorithm's arithmetic per launch over its launch time against the FP32
    # peak. Per evaluated (pixel, splat) pair the reference's loop (tile_render.comp:335-352) does
    # 16 operations (d = centre - pixel: 2; power: 9; power > 0: 1; exp: 1; o * exp: 1; min: 1;
    # alpha < 1/255: 1) and per blended pair 10 more (T (1 - alpha): 2; the test: 1; alpha T and
    # three colour FMAs: 7). Pair counts: the tile statistics of the same frames.
    render_ms = kernels["kRenderCulled"][0]
    ops = 16.0 * pairs["evaluated"] + 10.0 * pairs["blended"]
    achieved = ops / (render_ms * 1e-3) / 1e12
    render_bytes = algo["kRenderCulled"]
    roofline = {"bound": "valu", "kernel": "kRenderCulled<rgba8, " + args.render_splat + ">",
                "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": traffic_by_kernel.get("kRenderCulled"),
                "algorithmic_bytes": int(render_bytes),
                "hbm_gb_per_s": round(render_bytes / (render_ms * 1e-3) / 1e9, 1),
                "hbm_frac": round(render_bytes / (render_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "launch_us": round(render_ms * 1e3, 2), "ops_per_launch": int(ops),
                "pairs_per_launch": {k: int(v) for k, v in pairs.items() if k not in ("max_depth",)},
                "unpacked_fp32_peak": FP32_PEAK_TFLOPS / 2,
                "frac_of_unpacked_peak": round(achieved / (FP32_PEAK_TFLOPS / 2), 4),
                "note": "achieved = (16 ops x evaluated + 10 ops x blended pairs of the reference's tile loop, "
                        "counted by the tile-stats build on the same frames) / launch time (hipEvents on the render "
                        "stream; in global mode the launch also sorts each tile's range by depth, whose work is not "
                        "counted in ops); peak = MI355X FP32 157.3 TFLOP/s; traffic = rocprofv3 FETCH_SIZE + "
                        "WRITE_SIZE per launch scaled by a plain copy (profiles/); lane_slots = (pixel, splat) "
                        "evaluations this kernel's loop issued; the library is built without packed FP32 "
                        "(v_pk_*_f32), so one wave64 FP32 instruction takes a SIMD 4 cycles: "
                        "unpacked_fp32_peak = half the published peak is the rate its VALU code can reach"}
    if args.sort_mode == "binning":
        passes = {"depth_radix": 4, "tile_radix": (int(tiles).bit_length() + 7) // 8}
        sort_ms = stages["sort"] + stages["depth_sort"]
    else:
        passes = {"tile_radix": (int(tiles).bit_length() + 7) // 8, "per_tile_depth_sort_in_render": 1}
        sort_ms = stages["sort"]  # the tile-id radix passes and ranges; the depth order within a tile is the render's
    full_key = None
    if args.sort_mode == "global":
        fpasses = (32 + int(tiles).bit_length() + 7) // 8
        fm = full_sort["entries"]
        full_key = {"ms": round(full_sort["ms"], 4), "entries": int(fm), "passes": fpasses,
                    "sorted_keys_per_s": round(fm / (full_sort["ms"] * 1e-3), 1),
                    "key_passes_per_s": round(fm * fpasses / (full_sort["ms"] * 1e-3), 1),
                    "gb_per_s": round(fm * fpasses * (8 + 24) / (full_sort["ms"] * 1e-3) / 1e9, 1),
                    "note": "the 64-bit (tile, depth) key LSD radix sort of the frame's list, all key bits, "
                            "reduce-then-scan; bytes = passes x (8 B histogram read + 12 B read + 12 B written) "
                            "per entry"}

    multiview = None if args.no_multiview or args.ply else multiview_bench(ctx)
    learned = None if args.no_learned or args.ply else learned_bench(ctx, cams, min(args.steps, 100))
    backward = None if args.no_backward or args.ply else backward_bench(ctx)

    # Frames in flight (the reference viewer keeps two, application.h:55): the orbit rendered by
    # k renderers on k streams (gsr_batch, frame i on branch i % k, each branch its own image),
    # so one frame's latency-bound kernels overlap the next's. Throughput only: the headline
    # `value` stays the serial per-frame time.
    in_flight = {}
    scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    for k in (2, 4):
        rs = [gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", sort_mode=args.sort_mode,
                           render_splat=args.render_splat) for _ in range(k)]
        imgs = [gsr.DeviceBuffer(ctx, W * H * 4) for _ in range(k)]
        batch = gsr.Batch(ctx, k)
        orbit = [cams[(args.warmup + i) % ORBIT_FRAMES] for i in range(args.steps)]
        targets = [imgs[i % k].ptr for i in range(args.steps)]
        batch.render(rs, orbit[:k * 4], targets[:k * 4])  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        batch.render(rs, orbit, targets)
        ctx.synchronize()
        in_flight[str(k)] = round((time.perf_counter() - t0) * 1e3 / args.steps, 4)
        batch.close()
        for x in rs + imgs:
            x.close()
        print(f"# {k} frames in flight: {in_flight[str(k)]:.4f} ms/frame over {args.steps} frames", file=sys.stderr,
              flush=True)
    scene.close()

    cpu = None
    psnr = maxdiff = fp64_error = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        ocam = oracle.OracleCamera.from_gsr(qcam)
        ob = {"rect": oracle.RECT, "tight": oracle.TIGHT}[args.bounds]
        times, phase_runs = [], []
        t_start = time.time()
        ref = None
        while True:
            rgb, st = oracle.render(sc, ocam, bounds=ob, threads=threads)
            times.append(st["totalMs"])
            phase_runs.append(st)
            if ref is None:
                ref = rgb
            if time.time() - t_start > args.cpu_seconds or len(times) >= 5:
                break
        q = oracle.quantize8(ref)
        maxdiff = int(np.abs(img[..., :3].astype(np.int32) - q.astype(np.int32)).max())
        psnr = oracle.psnr(img[..., :3].astype(np.float64) / 255.0, ref)
        phases = {k: round(float(np.median([p[k] for p in phase_runs])), 2)
                  for k in ("preprocessMs", "sortMs", "binMs", "rasterMs")}
        # Single-thread DepthSorter::radixSortPairs (depth_sort.cpp:44-75) on 2^22 positive-float depth keys.
        rng = np.random.default_rng(0)
        dkeys = rng.uniform(0.2, 30.0, 1 << 22).astype(np.float32).view(np.uint32)
        t1 = time.perf_counter()
        _, _, rpasses = oracle.radix_sort_pairs(dkeys, np.arange(dkeys.size, dtype=np.uint32))
        radix_s = time.perf_counter() - t1
        cpu_model = "unknown"
        try:
            for ln in open("/proc/cpuinfo"):
                if ln.startswith("model name"):
                    cpu_model = ln.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(float(np.median(times)), 2), "unit": "ms/frame", "cores": threads, "kind": "port",
               "sample": f"{len(times)} renders of orbit frame {args.warmup % ORBIT_FRAMES} ({workload}, "
                         f"{args.bounds} bounds) by the CPU oracle (renderReference restated), median",
               "phases_ms": {"preprocess": phases["preprocessMs"], "sort": phases["sortMs"], "bin": phases["binMs"],
                             "raster": phases["rasterMs"]},
               "cpu_model": cpu_model,
               "radix_sort_pairs_1thread": {"keys": int(dkeys.size), "passes": rpasses,
                                            "keys_per_s": round(dkeys.size / radix_s, 1)}}
        # FP32 preprocess error of this GPU frame's splats against the FP64 twin of the oracle.
        s64 = oracle.preprocess_f64(sc, ocam, threads=threads)
        vis = (qsplats["cull_reason"] == 0) & (s64[:, 3] == 0)

        def stat(a):
            return {"median": float(f"{np.median(a):.3e}"), "max": float(f"{np.max(a):.3e}")}
        fp64_error = {
            "visible": int(vis.sum()),
            "cull_mismatches": int((qsplats["cull_reason"] != s64[:, 3]).sum()),
            "centre_px": stat(np.abs(qsplats["center"][vis] - s64[vis, 0:2]).max(axis=1)),
            "conic_rel": stat((np.abs(qsplats["conic_opacity"][vis, :3] - s64[vis, 4:7]) /
                               np.abs(s64[vis, 4:7]).max(axis=1, keepdims=True)).max(axis=1)),
            "colour_abs": stat(np.abs(qsplats["color"][vis] - s64[vis, 8:11]).max(axis=1)),
            "depth_rel": stat(np.abs(qsplats["depth"][vis] - s64[vis, 2]) / s64[vis, 2]),
            "radius_px": stat(np.abs(qsplats["radius"][vis] - s64[vis, 11])),
            "frame": args.warmup % ORBIT_FRAMES}

    line = {
        "metric": METRIC,
        "value": round(result[("f32", args.sort_mode, args.render_splat, 0)], 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(result[("f32", args.sort_mode, args.render_splat, 0)], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "ply file" if args.ply else "synthetic",
        "config": {"workload": workload, "gaussians": n, "width": W, "height": H,
                   "orbit": f"300 frames, radius {args.radius:g}, elevation {args.elevation:g} deg",
                   "bounds": args.bounds, "scene_format": "f32",
                   "target": "rgba8", "sort_mode": args.sort_mode, "render_splat": args.render_splat},
        "fp16_scene_ms_per_frame": round(result[("f16", args.sort_mode, args.render_splat, 0)], 4),
        "render_splat_ms_per_frame": splat_modes,
        "stopthepop_ms_per_frame": windows,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "mean_entries": int(mean_entries),
        "mean_visible": int(mean_visible),
        "sort_passes": passes,
        "sorted_keys_per_s": round(mean_entries / (sort_ms * 1e-3), 1),
        "sorted_keys_note": ("list entries / sort stage time: the two 32-bit tile-id radix passes and the ranges; "
                             "each tile's depth order is sorted inside the render (full_key_sort gives the keys/s of "
                             "the whole 64-bit sort)") if args.sort_mode == "global" else
                            "list entries / (depth sort + tile sort) stage time",
        "full_key_sort": full_key,
        "kernels": kernel_table,
        "sort_modes": modes,
        "psnr_db_vs_oracle": None if psnr is None else round(psnr, 2),
        "max_diff_levels_vs_oracle": maxdiff,
        "roofline": roofline,
        "preprocess_fp32_error_vs_fp64": fp64_error,
        "cpu_baseline": cpu,
        "multiview": multiview,
        "frames_in_flight_ms_per_frame": in_flight,
        "learned_colour": learned,
        "backward": backward,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
