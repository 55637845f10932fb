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
frame against the CPU oracle, and the oracle's own CPU time on the same frame (cpu_baseline).
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
    ap.add_argument("--sort-mode", default="binning", choices=("binning", "global"))
    ap.add_argument("--ply", default=None, help="render an INRIA .ply scene instead of the synthetic one")
    ap.add_argument("--radius", type=float, default=10.0, help="orbit radius")
    ap.add_argument("--elevation", type=float, default=15.0, help="orbit elevation, degrees")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
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
        rp = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", profile_stages=True, sort_mode=mode)
        stage_sum = {k: 0.0 for k in gsr.STAGES}
        entries = vis = 0
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
        rp.close()
        return {k: v / n_prof for k, v in stage_sum.items()}, entries / n_prof, vis / n_prof

    other = "global" if args.sort_mode == "binning" else "binning"
    result, modes = {}, {}
    for fmt, mode in (("f32", args.sort_mode), ("f16", args.sort_mode), ("f32", other)):
        scene = gsr.Scene(ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree, fmt=fmt)
        r = gsr.Renderer(ctx, scene, bounds=args.bounds, target="rgba8", sort_mode=mode)
        ms = timed_orbit(r, 0, args.steps)
        result[(fmt, mode)] = ms / args.steps
        print(f"# {fmt} {mode}: {ms / args.steps:.4f} ms/frame over {args.steps} frames", file=sys.stderr, flush=True)
        if fmt == "f32":
            st, ent, vis = profile_stages(scene, mode)
            modes[mode] = {"ms_per_frame": round(ms / args.steps, 4), "stages_ms": {k: round(v, 4) for k, v in st.items()},
                           "mean_entries": int(ent), "mean_visible": int(vis)}
            if mode == args.sort_mode:
                stages, mean_entries, mean_visible = st, ent, vis
                # Quality: one frame (the first timed one) vs the oracle, RGBA8.
                qcam = cams[args.warmup % ORBIT_FRAMES]
                r.render(qcam, out.ptr)
                img = out.read(np.empty((H, W, 4), np.uint8))
        r.close()
        scene.close()

    # Roofline of the dominant HBM-bound stage: algorithmic bytes / stage time.
    n = args.count
    tiles = cams[0].tiles
    tile_bits = int(tiles).bit_length()
    if args.sort_mode == "binning":
        tile_passes = (tile_bits + 7) // 8
        # per pass: key read (histogram) + key/value read + key/value write, 4 + 8 + 8 B
        sort_bytes = tile_passes * 20.0 * mean_entries
        depth_sort_bytes = 4 * 20.0 * mean_visible
        dup_bytes = mean_entries * 8.0 + mean_visible * (48 + 4 + 4)  # list written; splat, order, count read
        stage_bytes = {"depth_sort": depth_sort_bytes}
        sorted_keys = mean_entries * tile_passes + mean_visible * 4
        passes = {"depth": 4, "tile": tile_passes}
    else:
        key_passes = (32 + tile_bits + 7) // 8
        sort_bytes = key_passes * 32.0 * mean_entries  # 8 + 12 + 12 B per pass
        dup_bytes = mean_entries * 12.0 + n * 4.0
        stage_bytes = {}
        sorted_keys = mean_entries * key_passes
        passes = {"key": key_passes}
    pre_bytes = n * (16 + 32 + 192 + 48 + 4)  # upper bound: every gaussian's record read, splat + count written
    stage_bytes.update({"preprocess": pre_bytes, "sort": sort_bytes, "duplicate": dup_bytes,
                        "ranges": mean_entries * (4.0 if args.sort_mode == "binning" else 8.0) + tiles * 8.0})
    hbm_stages = {k: stages[k] for k in stage_bytes}
    dom = max(hbm_stages, key=hbm_stages.get)
    achieved = stage_bytes[dom] / (stages[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "note": "algorithmic bytes per launch / stage time (hipEvents on the render stream)"}

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
        "value": round(result[("f32", args.sort_mode)], 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(result[("f32", args.sort_mode)], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "ply file" if args.ply else "synthetic",
        "config": {"workload": workload, "gaussians": n, "width": W, "height": H,
                   "orbit": f"300 frames, radius {args.radius:g}, elevation {args.elevation:g} deg",
                   "bounds": args.bounds, "scene_format": "f32",
                   "target": "rgba8", "sort_mode": args.sort_mode},
        "fp16_scene_ms_per_frame": round(result[("f16", args.sort_mode)], 4),
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "mean_entries": int(mean_entries),
        "mean_visible": int(mean_visible),
        "sort_passes": passes,
        "sorted_keys_per_s": round(sorted_keys / ((stages["sort"] + stages["depth_sort"]) * 1e-3), 1),
        "sort_modes": modes,
        "psnr_db_vs_oracle": None if psnr is None else round(psnr, 2),
        "max_diff_levels_vs_oracle": maxdiff,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
