"""Batched multi-view (gsr_batch_*): views of mixed sizes on several streams, and the same batch
captured into one HIP graph. Every image must equal the one gsr_render gives for that view
alone (same kernels, same inputs: bit-exact). The graph is replayed several times with other
work and synchronisations between the replays, and every replay is checked."""
import numpy as np
import pytest

SIZES = [(640, 360), (320, 180), (250, 130), (1000, 563), (64, 48), (640, 360), (17, 9)]


def log(msg):
    print(msg, flush=True)


def _views(gsr, n):
    from gsr import camera

    rng = np.random.default_rng(11)
    cams = []
    for i in range(n):
        W, H = SIZES[i % len(SIZES)]
        radius = rng.uniform(5.0, 9.0)
        view = camera.orbit_view(radius, rng.uniform(-30.0, 30.0), rng.uniform(0.0, 360.0))
        cams.append(gsr.camera_from_viewer(view, camera.perspective(60.0, W / H), camera.IDENTITY, W, H))
    return cams


def _read(gsr, buf, cam):
    return buf.read(np.empty((int(cam.height), int(cam.width), 4), np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("branches,sort_mode", [(1, "global"), (3, "global"), (4, "binning")])
def test_batch_and_graph_equal_single_renders(gsr_ctx, branches, sort_mode):
    import gsr
    from gsr import scenes

    log(f"case batch branches={branches} sort_mode={sort_mode}")
    sc = scenes.make_scene(40_000, seed=2, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _views(gsr, 10)
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * 4) for c in cams]

    single = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode=sort_mode)
    want = [single.render_to_host(c) for c in cams]

    renderers = [gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode=sort_mode)
                 for _ in range(branches)]
    batch = gsr.Batch(gsr_ctx, branches)
    batch.render(renderers, cams, [o.ptr for o in outs])
    gsr_ctx.synchronize()
    for i, (o, c) in enumerate(zip(outs, cams)):
        assert np.array_equal(_read(gsr, o, c), want[i]), f"batched view {i} ({c.width}x{c.height}) differs"

    batch.capture(renderers, cams, [o.ptr for o in outs])
    other = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode=sort_mode)
    for rep in range(3):
        for o in outs:  # the replay must write every image again
            o.write(np.zeros(o.nbytes, np.uint8))
        batch.replay()
        gsr_ctx.synchronize()
        for i, (o, c) in enumerate(zip(outs, cams)):
            assert np.array_equal(_read(gsr, o, c), want[i]), f"replay {rep}: view {i} differs"
        # eager work on the same scene between the replays
        img = other.render_to_host(cams[rep])
        assert np.array_equal(img, want[rep])
    log(f"  {len(cams)} views, 3 graph replays equal the single renders")
    batch.close()
    for r in renderers + [single, other]:
        r.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
@pytest.mark.parametrize("branches,sort_mode,target,splat",
                         [(1, "global", "rgba8", "p32"), (4, "global", "rgba32f", "p24"),
                          (3, "binning", "rgba8", "full")])
def test_persistent_batch_equals_single_renders(gsr_ctx, branches, sort_mode, target, splat):
    """Persistent mode: the views' frames up to the tile sort on the branches, then one render
    over every view's tiles from a shared counter. Same tile body as the per-frame culled render:
    every image bit-equal to the view's own gsr_render, eagerly (twice in a row, which reuses the
    two alternating job tables) and in graph replays with eager work between them."""
    import gsr
    from gsr import scenes

    log(f"case persistent branches={branches} sort_mode={sort_mode} target={target} splat={splat}")
    sc = scenes.make_scene(40_000, seed=3, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _views(gsr, 9)
    px = 16 if target == "rgba32f" else 4
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * px) for c in cams]

    def read(o, c):
        if target == "rgba32f":
            return o.read(np.empty((int(c.height), int(c.width), 4), np.float32))
        return _read(gsr, o, c)

    def make():
        return gsr.Renderer(gsr_ctx, scene, bounds="tight", target=target, sort_mode=sort_mode, render_splat=splat)

    single = make()
    want = [single.render_to_host(c) for c in cams]
    renderers = [make() for _ in cams]
    batch = gsr.Batch(gsr_ctx, branches, persistent=True)
    for rep in range(2):
        for o in outs:
            o.write(np.zeros(o.nbytes, np.uint8))
        batch.render(renderers, cams, [o.ptr for o in outs])
        gsr_ctx.synchronize()
        for i, (o, c) in enumerate(zip(outs, cams)):
            assert np.array_equal(read(o, c), want[i]), f"eager {rep}: persistent view {i} ({c.width}x{c.height})"

    batch.capture(renderers, cams, [o.ptr for o in outs])
    for rep in range(3):
        for o in outs:
            o.write(np.zeros(o.nbytes, np.uint8))
        batch.replay()
        gsr_ctx.synchronize()
        for i, (o, c) in enumerate(zip(outs, cams)):
            assert np.array_equal(read(o, c), want[i]), f"replay {rep}: persistent view {i} differs"
        # eager frames of the captured renderers between the replays (on the context's stream)
        img = renderers[rep].render_to_host(cams[rep])
        assert np.array_equal(img, want[rep])
    log(f"  {len(cams)} views: 2 eager persistent batches and 3 replays equal the single renders")

    # back to stream mode on the same batch: the captured graph is dropped, images unchanged
    batch.set_persistent(False)
    batch.render(renderers[:branches], cams, [o.ptr for o in outs])
    gsr_ctx.synchronize()
    for i, (o, c) in enumerate(zip(outs, cams)):
        assert np.array_equal(read(o, c), want[i]), f"stream mode after persistent: view {i}"
    with pytest.raises(gsr.GsrError):
        batch.replay()
    batch.close()
    for r in renderers + [single]:
        r.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
def test_persistent_batch_arguments(gsr_ctx):
    import gsr
    from gsr import scenes

    log("case persistent batch arguments")
    sc = scenes.make_scene(5_000, seed=4)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _views(gsr, 3)
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * 16) for c in cams]
    ptrs = [o.ptr for o in outs]
    batch = gsr.Batch(gsr_ctx, 2, persistent=True)
    plain = [gsr.Renderer(gsr_ctx, scene, target="rgba8") for _ in cams]
    with pytest.raises(gsr.GsrError, match="same"):  # one renderer per view, not per branch
        batch.render([plain[0], plain[1], plain[0]], cams, ptrs)
    stp = gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_window=8)
    with pytest.raises(gsr.GsrError, match="persistent"):
        batch.render([plain[0], stp, plain[2]], cams, ptrs)
    f32 = gsr.Renderer(gsr_ctx, scene, target="rgba32f")
    with pytest.raises(gsr.GsrError, match="format"):
        batch.render([plain[0], f32, plain[2]], cams, ptrs)
    bins8 = gsr.Renderer(gsr_ctx, scene, target="rgba8", tile_size=8)
    with pytest.raises(gsr.GsrError, match="persistent"):
        batch.render([plain[0], plain[1], bins8], cams, ptrs)
    batch.render(plain, cams, ptrs)  # valid after the refusals
    gsr_ctx.synchronize()
    for r in plain + [stp, f32, bins8]:
        r.close()
    batch.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
def test_batch_arguments(gsr_ctx):
    import gsr
    from gsr import scenes

    sc = scenes.make_scene(2_000, seed=3)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _views(gsr, 3)
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * 4) for c in cams]
    r = gsr.Renderer(gsr_ctx, scene, target="rgba8")
    batch = gsr.Batch(gsr_ctx, 2)
    with pytest.raises(gsr.GsrError, match="same"):
        batch.render([r, r], cams, [o.ptr for o in outs])
    r2 = gsr.Renderer(gsr_ctx, scene, target="rgba8")
    with pytest.raises(gsr.GsrError, match="not rendered yet"):
        batch.capture([r, r2], cams, [o.ptr for o in outs])
    with pytest.raises(gsr.GsrError, match="nothing captured"):
        batch.replay()
    stats = gsr.Renderer(gsr_ctx, scene, target="rgba8", frame_stats=True)
    batch.render([stats, r2], cams, [o.ptr for o in outs])
    gsr_ctx.synchronize()
    with pytest.raises(gsr.GsrError, match="plain renderers"):
        batch.capture([stats, r2], cams, [o.ptr for o in outs])
    with pytest.raises(gsr.GsrError):
        gsr.Batch(gsr_ctx, 0)
    batch.close()
    for x in (r, r2, stats):
        x.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
def test_replay_refuses_changed_or_destroyed_renderers(gsr_ctx):
    """A captured graph holds raw pointers into its renderers: after a renderer reallocates (a larger
    view), changes its colour binding or is destroyed, gsr_batch_replay refuses on the host before
    launching anything; a fresh capture works again. The one-sweep sort is refused at capture."""
    import gsr
    from gsr import camera, scenes

    log("case replay refusals")
    sc = scenes.make_scene(5_000, seed=4, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = [gsr.camera_from_viewer(camera.orbit_view(6.0, 10.0, a), camera.perspective(60.0, 320 / 180),
                                   camera.IDENTITY, 320, 180) for a in (0.0, 90.0)]
    outs = [gsr.DeviceBuffer(gsr_ctx, 320 * 180 * 4) for _ in cams]
    rs = [gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_mode="global") for _ in range(2)]
    want = [rs[0].render_to_host(c).copy() for c in cams]
    batch = gsr.Batch(gsr_ctx, 2)
    batch.render(rs, cams, [o.ptr for o in outs])
    batch.capture(rs, cams, [o.ptr for o in outs])
    batch.replay()
    gsr_ctx.synchronize()
    assert np.array_equal(_read(gsr, outs[1], cams[1]), want[1])
    # a larger view reallocates renderer 1's per-tile buffers
    big = gsr.camera_from_viewer(camera.orbit_view(6.0, 10.0, 0.0), camera.perspective(60.0, 1280 / 720),
                                 camera.IDENTITY, 1280, 720)
    rs[1].render_to_host(big)
    with pytest.raises(gsr.GsrError, match="changed since the capture"):
        batch.replay()
    batch.capture(rs, cams, [o.ptr for o in outs])  # a new capture sees the new buffers
    batch.replay()
    gsr_ctx.synchronize()
    assert np.array_equal(_read(gsr, outs[1], cams[1]), want[1])
    rs[1].close()
    with pytest.raises(gsr.GsrError, match="destroyed"):
        batch.replay()
    one = gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_mode="global", sort_algorithm="onesweep")
    one.render_to_host(cams[0])
    with pytest.raises(gsr.GsrError, match="one-sweep"):
        batch.capture([rs[0], one], cams, [o.ptr for o in outs])
    batch.close()
    for x in (rs[0], one):
        x.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
def test_frames_on_alternating_streams(gsr_ctx):
    """Frames of one renderer on different streams are ordered after each other by the library:
    alternating two streams with no caller synchronisation gives every frame's own image."""
    import gsr
    from gsr import camera, scenes

    log("case alternating streams")
    sc = scenes.make_scene(60_000, seed=6, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    W, H = 640, 360
    cams = [gsr.camera_from_viewer(camera.orbit_view(7.0, 5.0, 30.0 * i), camera.perspective(60.0, W / H),
                                   camera.IDENTITY, W, H) for i in range(6)]
    r = gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_mode="global")
    want = [r.render_to_host(c).copy() for c in cams]
    outs = [gsr.DeviceBuffer(gsr_ctx, W * H * 4) for _ in cams]
    streams = [gsr.Stream(gsr_ctx) for _ in range(2)]
    for i, c in enumerate(cams):
        r.render(c, outs[i].ptr, streams[i % 2].handle)
    gsr_ctx.synchronize()
    for i, c in enumerate(cams):
        assert np.array_equal(_read(gsr, outs[i], c), want[i]), f"frame {i} differs"
    for s in streams:
        s.close()
    r.close()
    for o in outs:
        o.close()
    scene.close()


def _workload_views(gsr, views=32, seed=5):
    """The batched multi-view workload's cameras: `views` of sizes uniform over {640x360, 1280x720,
    1920x1080}, uniform directions on shells of radius 8..12, looking at the origin."""
    from gsr import camera

    rng = np.random.default_rng(seed)
    sizes = [(640, 360), (1280, 720), (1920, 1080)]
    cams = []
    for _ in range(views):
        W, H = sizes[int(rng.integers(0, 3))]
        d = rng.normal(size=3)
        eye = d / np.linalg.norm(d) * rng.uniform(8.0, 12.0)
        cams.append(gsr.camera_from_viewer(camera.look_at(eye, [0.0, 0.0, 0.0]), camera.perspective(60.0, W / H),
                                           camera.IDENTITY, W, H))
    return cams


@pytest.mark.gpu
def test_batched_multiview_workload_32_views_1m(gsr_ctx):
    """The batched multi-view workload at its own size: 32 cameras of mixed sizes around a
    1M-gaussian scene (means ~ N(0, 3^2 I), seed 0), 8 branches. The batch and three replays of its
    graph equal one gsr_render per view, bit for bit; three views (one of each size) also match the
    oracle: tile lists and ranges bit-exact, RGBA8 within one level."""
    import gsr
    import oracle
    from gsr import scenes

    log("case batch 32 views 1M gaussians")
    sc = scenes.make_scene(1_000_000, seed=0, mean_sigma=3.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _workload_views(gsr)
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * 4) for c in cams]
    single = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode="global")
    want = [single.render_to_host(c).copy() for c in cams]
    # three views against the oracle, one of each size
    checked = set()
    for i, c in enumerate(cams):
        if int(c.width) in checked:
            continue
        checked.add(int(c.width))
        img = single.render_to_host(c)
        ocam = oracle.OracleCamera.from_gsr(c)
        o_splats, o_counts = oracle.preprocess(sc, ocam, bounds=oracle.TIGHT)
        keys, depth, vals, ranges, ctl = single.read_back()
        hi, lo, ov, orng, dup, _ = oracle.tile_lists(o_splats, o_counts, c.tiles_x, c.tiles_y, bounds=oracle.TIGHT,
                                                     mode=0)
        assert ctl.duplicates == dup and np.array_equal(keys, hi) and np.array_equal(depth, lo)
        assert np.array_equal(vals, ov) and np.array_equal(ranges, orng), f"view {i}: lists differ"
        ref, _ = oracle.render(sc, ocam, bounds=oracle.TIGHT)
        diff = np.abs(img[..., :3].astype(np.int32) - oracle.quantize8(ref).astype(np.int32)).max()
        log(f"  view {i} {c.width}x{c.height}: {ctl.entries} entries bit-exact, max {diff} level")
        assert diff <= 1
    assert len(checked) == 3
    renderers = [gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode="global") for _ in range(8)]
    batch = gsr.Batch(gsr_ctx, 8)
    batch.render(renderers, cams, [o.ptr for o in outs])
    gsr_ctx.synchronize()
    for i, (o, c) in enumerate(zip(outs, cams)):
        assert np.array_equal(_read(gsr, o, c), want[i]), f"batched view {i} differs"
    batch.capture(renderers, cams, [o.ptr for o in outs])
    for rep in range(3):
        for o in outs:
            o.write(np.zeros(o.nbytes, np.uint8))
        batch.replay()
        gsr_ctx.synchronize()
        for i, (o, c) in enumerate(zip(outs, cams)):
            assert np.array_equal(_read(gsr, o, c), want[i]), f"replay {rep}: view {i} differs"
        assert np.array_equal(single.render_to_host(cams[rep]), want[rep])  # eager work between replays
    log(f"  32 views ({sum(int(c.width) * int(c.height) for c in cams) / 1e6:.1f} Mpixel): batch and 3 replays "
        "equal the single renders")
    batch.close()
    for x in renderers + [single]:
        x.close()
    for o in outs:
        o.close()
    scene.close()


@pytest.mark.gpu
def test_autotuner_tile_size_and_batch_depth(gsr_ctx):
    """tune_batch searches tile size x batch depth x branch count x persistent mode on a captured batch: every
    configuration is timed, the best is one of them, and the batch captured in that configuration
    renders every view exactly as a single renderer of that configuration."""
    import time

    import gsr
    from gsr import scenes

    log("case autotuner")
    sc = scenes.make_scene(40_000, seed=2, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cams = _views(gsr, 6)
    outs = [gsr.DeviceBuffer(gsr_ctx, int(c.width) * int(c.height) * 4) for c in cams]

    def make(tile_size=16, render_batch=0):
        return gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_mode="global", tile_size=tile_size,
                            render_batch=render_batch)

    def timer(fn):
        gsr_ctx.synchronize()
        t0 = time.perf_counter()
        fn()
        gsr_ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3

    best, by = gsr.tune_batch(gsr_ctx, make, cams, [o.ptr for o in outs], candidates=(1, 2), replays=2, timer=timer,
                              tile_sizes=(16, 8), batches={16: (0, 512), 8: (0, 32)}, persistent=(False, True))
    # 2 tile sizes x 2 batches x 2 branch counts in stream mode, + 16-px default batch persistent
    assert len(by) == 10 and all(v > 0 for v in by.values())
    assert sum(k[3] for k in by) == 2 and all(k[:2] == (16, 0) for k in by if k[3])
    assert (best["tile_size"], best["render_batch"], best["branches"], best["persistent"]) in by
    log(f"  best {best} of {len(by)}")
    single = make(best["tile_size"], best["render_batch"])
    want = [single.render_to_host(c).copy() for c in cams]
    rs = [make(best["tile_size"], best["render_batch"])
          for _ in range(len(cams) if best["persistent"] else best["branches"])]
    batch = gsr.Batch(gsr_ctx, best["branches"], persistent=best["persistent"])
    batch.render(rs, cams, [o.ptr for o in outs])
    batch.capture(rs, cams, [o.ptr for o in outs])
    batch.replay()
    gsr_ctx.synchronize()
    for i, (o, c) in enumerate(zip(outs, cams)):
        assert np.array_equal(_read(gsr, o, c), want[i])
    batch.close()
    for x in rs + [single]:
        x.close()
    for o in outs:
        o.close()
    scene.close()
