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
                         "scene plain, then with a colour model marking every gaussian, and exit")
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
    batched = gsr_batch with the branch count the autotuner picks (1, 2, 4 or 8 streams), the
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

    def make():
        return gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")

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
                              timer=timed)
    for o in outs:
        o.close()
    scene.close()
    print(f"# multiview: {views} views, {pixels / 1e6:.1f} Mpixel: sequential {seq_ms:.3f} ms, batched "
          f"{by[best]:.3f} ms with {best} branches ({', '.join(f'{k}: {v:.3f}' for k, v in by.items())})",
          file=sys.stderr, flush=True)
    return {"views": views, "gaussians": count, "sizes": "uniform of 640x360, 1280x720, 1920x1080",
            "megapixels": round(pixels / 1e6, 2), "sequential_ms": round(seq_ms, 3),
            "sequential_mpix_per_s": round(pixels / 1e6 / (seq_ms * 1e-3), 1), "batched_ms": round(by[best], 3),
            "batched_mpix_per_s": round(pixels / 1e6 / (by[best] * 1e-3), 1), "branches": best,
            "batched_ms_by_branches": {str(k): round(v, 3) for k, v in by.items()}, "graph": True}


def learned_bench(ctx, cams, frames, count=3_000_000, samples=1 << 24):
    """Learned colour: a 3M-gaussian scene of the same distributions (seed 0, means ~ N(0, 3^2 I),
    SH degree 3) with a colour model trained on `samples` (gaussian, direction, SH RGB) samples
    (512 Adam steps), then the 1080p orbit (global sort mode, full records) rendered plain and with
    the model, after fallback passes at 2/255 and with every gaussian learned: ms/frame (wall
    clock around `frames` synchronised frames), PSNR of frame 50 against the plain SH render, and
    the colour bytes per gaussian."""
    import numpy as np

    import gsr
    from gsr import scenes

    sc = scenes.make_scene(count, seed=0, mean_sigma=3.0)
    scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    del sc
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

    plain = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
    plain_ms = orbit(plain)
    ref = plain.render_to_host(cams[50])
    model = gsr.ColourModel(ctx, scene, seed=0)
    step_samples = 1 << 15
    ctx.synchronize()
    t0 = time.perf_counter()
    last = model.train(samples // step_samples, step_samples, lr_weights=2e-3, lr_latents=2e-2)
    train_ms = (time.perf_counter() - t0) * 1e3
    mse = model.loss(0x5EED, 1 << 20)
    r = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
    r.set_colour_model(model)
    forms = {}
    for name, threshold in (("fallback_2_255", 2.0 / 255.0), ("all_learned", float("inf"))):
        t0 = time.perf_counter()
        n, err = model.fallback(threshold)
        fallback_ms = (time.perf_counter() - t0) * 1e3
        ms = orbit(r)
        img = r.render_to_host(cams[50])
        forms[name] = {"learned_fraction": round(n / count, 4), "ms_per_frame": round(ms, 4),
                       "throughput_vs_plain": round(plain_ms / ms, 3), "psnr_db_vs_plain_sh": psnr(img, ref),
                       "colour_bytes_per_gaussian": round(32.0 + 192.0 * (1.0 - n / count), 1),
                       "fallback_pass_ms": round(fallback_ms, 2)}
    res = {"gaussians": count, "training_samples": samples, "training_ms": round(train_ms, 1),
           "train_last_mse": round(last, 6), "eval_mse": round(mse, 6),
           "median_worst_error": round(float(np.median(err)), 4), "plain_ms_per_frame": round(plain_ms, 4),
           "plain_colour_bytes_per_gaussian": 192, "shared_mlp_bytes": 4 * gsr.ColourModel.param_count(),
           "forms": forms}
    print(f"# learned colour ({count} gaussians): trained 2^{int(np.log2(samples))} samples in {train_ms:.0f} ms, "
          f"eval MSE {mse:.5f}; plain {plain_ms:.3f} ms/frame; "
          + "; ".join(f"{k}: {v['learned_fraction'] * 100:.1f}% learned, {v['ms_per_frame']:.3f} ms/frame, "
                      f"PSNR vs plain {v['psnr_db_vs_plain_sh']}" for k, v in forms.items()),
          file=sys.stderr, flush=True)
    r.close()
    plain.close()
    model.close()
    out.close()
    scene.close()
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

    if args.profile_learned > 0:
        big = scenes.make_scene(3_000_000, seed=0, mean_sigma=3.0)
        scene = gsr.Scene(ctx, big.pos_alpha, big.scales, big.rotations, big.sh, big.sh_degree)
        del big
        plain = gsr.Renderer(ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
        for i in range(args.profile_learned):
            plain.render(cams[(args.warmup + i) % ORBIT_FRAMES], out.ptr)
        ctx.synchronize()
        model = gsr.ColourModel(ctx, scene, seed=0)
        model.train(4, 1 << 16)
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

    def profile_stages(scene, mode):
        """Per-stage ms and list counts: a second renderer with events between the stages."""
        rp = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", profile_stages=True, sort_mode=mode,
                          render_splat=args.render_splat)
        stage_sum = {k: 0.0 for k in gsr.STAGES}
        entries = vis = 0
        culls = np.zeros(5, np.float64)
        n_prof = min(args.steps, 60)
        for i in range(n_prof):
            cam = cams[(args.warmup + i * (ORBIT_FRAMES // n_prof)) % ORBIT_FRAMES]
            rp.render(cam, out.ptr)
            st = rp.stage_times()
            ctl = rp.read_control()
            for k in stage_sum:
                stage_sum[k] += st[k]
            entries += ctl.entries
            vis += ctl.visible
            culls += np.array(ctl.cull_counts[:5], np.float64)
        rp.close()
        return {k: v / n_prof for k, v in stage_sum.items()}, entries / n_prof, vis / n_prof, culls / n_prof

    other = "global" if args.sort_mode == "binning" else "binning"
    packed = [m for m in ("p32", "p24") if m != args.render_splat]
    runs = [("f32", args.sort_mode, args.render_splat, 0), ("f16", args.sort_mode, args.render_splat, 0),
            ("f32", other, args.render_splat, 0)] + [("f32", args.sort_mode, m, 0) for m in packed] + \
        [("f32", args.sort_mode, args.render_splat, k) for k in (4, 16)]
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
            st, ent, vis, culls = profile_stages(scene, mode)
            modes[mode] = {"ms_per_frame": round(ms / args.steps, 4), "stages_ms": {k: round(v, 4) for k, v in st.items()},
                           "mean_entries": int(ent), "mean_visible": int(vis)}
            if mode == args.sort_mode:
                stages, mean_entries, mean_visible, mean_culls = st, ent, vis, culls
                # Quality: one frame (the first timed one) vs the oracle, RGBA8.
                qcam = cams[args.warmup % ORBIT_FRAMES]
                r.render(qcam, out.ptr)
                img = out.read(np.empty((H, W, 4), np.uint8))
        r.close()
        scene.close()

    # Roofline of the dominant HBM-bound kernel, kPreprocess (f32 scene): algorithmic bytes per
    # launch / its launch time (hipEvents around it on the render stream, mean of the profiled
    # frames). Reads: 16 B position + opacity per gaussian, 32 B scale + rotation unless culled at
    # the near plane or for opacity, 4 x 48 B SH per gaussian that reaches the colour step
    # (visible). Writes: the 48-B splat + 4-B tile count per gaussian, 8 B per tile range zeroed.
    n = args.count
    tiles = cams[0].tiles
    near, low_opacity = mean_culls[1], mean_culls[4]
    K = sc.K
    pre_bytes = n * 16 + (n - near - low_opacity) * 32 + mean_visible * K * 12 + n * 52 + tiles * 8
    achieved = pre_bytes / (stages["preprocess"] * 1e-3) / 1e9
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "pmc_traffic_kPreprocess_1.2M-1080p-orbit.md")
    if os.path.exists(traffic_file) and not args.ply and args.count == 1_200_000:
        for line in open(traffic_file):
            if line.startswith("HBM bytes per launch:"):
                traffic = int(line.split(":")[1].split()[0])
    roofline = {"bound": "hbm", "kernel": "kPreprocess<f32>", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(pre_bytes),
                "note": "algorithmic bytes per launch / launch time (hipEvents on the render stream); traffic = "
                        "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, scaled by a plain copy (profiles/)"}
    if args.sort_mode == "binning":
        passes = {"depth_radix": 4, "tile_radix": (int(tiles).bit_length() + 7) // 8}
        sort_ms = stages["sort"] + stages["depth_sort"]
    else:
        passes = {"tile_radix": (int(tiles).bit_length() + 7) // 8, "per_tile_depth_sort": 1}
        sort_ms = stages["sort"]

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
    psnr = maxdiff = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        ocam = oracle.OracleCamera.from_gsr(qcam)
        ob = {"rect": oracle.RECT, "tight": oracle.TIGHT}[args.bounds]
        times = []
        t_start = time.time()
        ref = None
        while True:
            rgb, st = oracle.render(sc, ocam, bounds=ob, threads=threads)
            times.append(st["totalMs"])
            if ref is None:
                ref = rgb
            if time.time() - t_start > args.cpu_seconds or len(times) >= 5:
                break
        q = oracle.quantize8(ref)
        maxdiff = int(np.abs(img[..., :3].astype(np.int32) - q.astype(np.int32)).max())
        psnr = oracle.psnr(img[..., :3].astype(np.float64) / 255.0, ref)
        cpu = {"value": round(float(np.median(times)), 2), "unit": "ms/frame", "cores": threads, "kind": "port",
               "sample": f"{len(times)} renders of orbit frame {args.warmup % ORBIT_FRAMES} ({workload}, "
                         f"{args.bounds} bounds) by the CPU oracle (renderReference restated), median"}

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
        "sort_modes": modes,
        "psnr_db_vs_oracle": None if psnr is None else round(psnr, 2),
        "max_diff_levels_vs_oracle": maxdiff,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "multiview": multiview,
        "frames_in_flight_ms_per_frame": in_flight,
        "learned_colour": learned,
        "backward": backward,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
