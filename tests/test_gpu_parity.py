"""GPU parity of the splat-tiles path (through the C ABI) against the CPU oracle.

Mirrors the reference's GPU suites (tests/test_tile_raster.cpp, test_tight_bounds.cpp,
test_gpu_sort.cpp):
  - per-gaussian splats and tile counts: bit-exact (FP32, same operation order);
  - the sorted (tile, depth, gaussian) list and per-tile ranges: bit-exact in both sort modes
    (binning: depth sort then stable tile sort; global: 64-bit keys), incl. overflow, where the
    two modes drop different entries (binning: the farthest, global: the highest indices);
  - control block counts (N, V, M, entries, overflow, cull counts, coverage mismatch = 0);
  - image: float PSNR > 60 dB (test_tile_raster.cpp:498) and RGBA8 within 1 level per
    channel of the quantised oracle image (>= 58.7 dB);
  - the stand-alone radix sort (stable, 64-bit keys) and scan on adversarial sizes.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _camera(gsr, view, W, H, fov=60.0):
    from gsr import camera

    return gsr.camera_from_viewer(view, camera.perspective(fov, W / H), camera.IDENTITY, W, H)


def _check_splats(gsr_splats, gsr_counts, o_splats, o_counts):
    """Bit-exact comparison of the 48-byte splats (culled ones: reason + depth only)."""
    assert np.array_equal(gsr_splats["cull_reason"], o_splats["cull_reason"]), "cull reasons differ"
    assert np.array_equal(gsr_splats["depth"].view(np.uint32), o_splats["depth"].view(np.uint32)), "depths differ"
    vis = o_splats["cull_reason"] == 0
    for f in ("center", "conic_opacity", "color", "radius"):
        a = gsr_splats[f][vis].view(np.uint32)
        b = o_splats[f][vis].view(np.uint32)
        bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
        assert bad.size == 0, f"{f}: {bad.size} visible splats differ, first {np.nonzero(vis)[0][bad[:5]]}"
    assert np.array_equal(gsr_counts, o_counts), f"tile counts differ at {np.nonzero(gsr_counts != o_counts)[0][:10]}"


def _check_lists(r, o_splats, o_counts, tiles_x, tiles_y, bounds, capacity=None):
    import oracle

    keys, depth, vals, ranges, ctl = r.read_back()
    mode = {"global": 0, "binning": 1}[r.sort_mode]
    hi, lo, ov, orng, dup, mism = oracle.tile_lists(o_splats, o_counts, tiles_x, tiles_y, bounds=bounds,
                                                    capacity=capacity, mode=mode)
    assert ctl.duplicates == dup, (ctl.duplicates, dup)
    assert ctl.entries == len(hi), (ctl.entries, len(hi))
    assert ctl.overflow == dup - len(hi)
    assert ctl.coverage_mismatch == 0 and mism == 0
    assert ctl.sort_timeouts == 0
    assert np.array_equal(keys, hi), f"tile keys differ at {np.nonzero(keys != hi)[0][:10]}"
    assert np.array_equal(depth, lo), "depth keys differ"
    assert np.array_equal(vals, ov), f"values differ at {np.nonzero(vals != ov)[0][:10]}"
    assert np.array_equal(ranges, orng), "ranges differ"
    return ctl


def _render_case(gsr_ctx, sc, view, W, H, bounds, fmt="f32", target="rgba32f", budget_mb=256, fov=60.0,
                 exact_alpha=False, ppt=0, loop="culled", sort_mode="binning", render_splat="full"):
    import gsr
    import oracle

    cam = _camera(gsr, view, W, H, fov)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree, fmt=fmt)
    r = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target=target, budget_mb=budget_mb, exact_alpha=exact_alpha,
                     pixels_per_thread=ppt, render_loop=loop, sort_mode=sort_mode, render_splat=render_splat)
    img = r.render_to_host(cam)
    if fmt == "f16":  # the oracle renders exactly what the GPU decodes (gsplat_ref.cpp:182-190)
        from gsr.scenes import SyntheticScene

        g, c = oracle.encode_f16(sc)
        sc = SyntheticScene(*oracle.decode_f16(g, c, sc.sh_degree), sc.sh_degree, sc.seed)
    ocam = oracle.OracleCamera.from_gsr(cam)
    ob = {"rect": oracle.RECT, "tight": oracle.TIGHT, "cutoff-only": oracle.CUTOFF}[bounds]
    o_splats, o_counts = oracle.preprocess(sc, ocam, bounds=ob, render_splat=render_splat)
    return r, scene, cam, img, sc, ocam, ob, o_splats, o_counts


@pytest.mark.parametrize("sort_mode", ["binning", "global"])
@pytest.mark.parametrize("bounds", ["rect", "tight"])
def test_100k_720p_splats_lists_image(gsr_ctx, bounds, sort_mode):
    """The small workload: 100k gaussians, N(0, 2^2), camera at radius 8, 1280x720."""
    import oracle
    from gsr import camera, scenes

    log(f"case 100k-720p bounds={bounds} sort_mode={sort_mode}")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    r, scene, cam, img, sc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, camera.orbit_view(8.0, 0.0, 0.0), 1280,
                                                              720, bounds, sort_mode=sort_mode)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    assert ctl.gaussian_count == 100_000
    cull = np.bincount(os_["cull_reason"], minlength=5)
    assert list(ctl.cull_counts[:5]) == list(cull[:5])
    assert ctl.visible == int((oc != 0).sum())
    ref, st = oracle.render(sc, ocam, bounds=ob)
    p = oracle.psnr(img[..., :3], ref)
    log(f"  entries {ctl.entries} (oracle tile instances {st['tileInstances']}), PSNR {p:.2f} dB")
    assert p > 60.0
    r.close()
    scene.close()


@pytest.mark.parametrize("bounds,sort_mode,W,H,radius", [
    ("cutoff-only", "global", 640, 360, 6.0),
    ("cutoff-only", "binning", 640, 360, 6.0),
    ("tight", "global", 1000, 563, 8.0),
    ("tight", "binning", 17, 9, 8.0),
    ("rect", "global", 333, 211, 3.0),
    ("tight", "global", 640, 360, 1.2),
    ("tight", "binning", 640, 360, 1.2),
])
def test_bounds_sizes_and_close_cameras(gsr_ctx, bounds, sort_mode, W, H, radius):
    """cutoff-only bounds, image sizes that are not multiples of the tile, a 17x9 image (two
    tiles, ~30k entries in one), and a camera inside a dense cloud of larger splats (thousands of
    near-plane culls, 1.5M entries, tiles over 2048 entries): splats, lists and ranges
    bit-exact, image > 60 dB."""
    import oracle
    from gsr import camera, scenes

    log(f"case 30k {W}x{H} radius={radius} bounds={bounds} sort_mode={sort_mode}")
    if radius < 2.0:  # inside the cloud
        sc = scenes.make_scene(30_000, seed=4, mean_sigma=1.0, log_scale_mean=-2.5)
    else:
        sc = scenes.make_scene(30_000, seed=4, mean_sigma=2.0)
    r, scene, cam, img, sc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, camera.orbit_view(radius, 10.0, 30.0), W, H,
                                                              bounds, sort_mode=sort_mode)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    _, _, _, ranges, _ = r.read_back()
    longest = int((ranges[:, 1] - ranges[:, 0]).max()) if ranges.size else 0
    ref, _ = oracle.render(sc, ocam, bounds=ob)
    p = oracle.psnr(img[..., :3], ref)
    log(f"  entries {ctl.entries}, visible {ctl.visible}, near-culled {ctl.cull_counts[1]}, longest tile "
        f"{longest}, PSNR {p:.2f} dB")
    assert p > 60.0
    r.close()
    scene.close()


@pytest.mark.parametrize("exact_alpha,ppt,loop", [(False, 1, "culled"), (False, 1, "predicated"), (False, 1, "divergent"),
                                                  (True, 1, "predicated"), (False, 2, "predicated"),
                                                  (True, 4, "predicated")])
def test_rgba8_within_one_level(gsr_ctx, exact_alpha, ppt, loop):
    """Every render variant: RGBA8 within one level of the quantised oracle image."""
    import oracle
    from gsr import camera, scenes

    log(f"case rgba8 100k-720p tight exact_alpha={exact_alpha} ppt={ppt}")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    r, scene, cam, img, sc, ocam, ob, *_ = _render_case(gsr_ctx, sc, camera.orbit_view(8.0, 0.0, 0.0), 1280, 720,
                                                        "tight", target="rgba8", exact_alpha=exact_alpha, ppt=ppt,
                                                        loop=loop)
    ref, _ = oracle.render(sc, ocam, bounds=ob)
    q = oracle.quantize8(ref)
    diff = np.abs(img[..., :3].astype(np.int32) - q.astype(np.int32))
    p = oracle.psnr(img[..., :3].astype(np.float64) / 255.0, ref)
    log(f"  max diff {diff.max()} levels, {int((diff > 0).sum())} channel values off by one, PSNR {p:.2f} dB")
    assert diff.max() <= 1
    assert p >= 58.7
    r.close()
    scene.close()


@pytest.mark.parametrize("case", ["100k-720p", "test-scene-640x360", "needles-giants-1280x720",
                                  "near-clamp-opacities-1280x720"])
def test_culled_loop_equals_uncompacted_loop(gsr_ctx, case):
    """The quadrant-culled loop drops only pairs below the alpha cutoff (a box, then the exact
    ellipse-vs-quadrant test with a margin sized to the FP32 rounding of q): its RGBA32F image
    equals the uncompacted predicated loop's bit for bit (same q, same blend order). The
    needles-and-giants scene stresses that margin: one long axis (up to e^0.5) and two tiny ones
    (e^-7) at every rotation, so the projected conics are nearly singular and the quadratic's terms
    cancel, 1 % large isotropic splats near the camera, opacities mostly near 1. The culled loop
    also leaves out the 0.99 alpha clamp for wave lists whose opacities are all <= 0.98, where it
    cannot apply; the uncompacted loop always clamps. The near-clamp scene puts every opacity in
    [0.975, 0.9999] (a third exactly 0.98 or 0.99) at 1-2 px splats, so both paths run side by side
    and the clamp binds at many centre pixels."""
    import gsr
    from gsr import camera, scenes

    log(f"case culled vs predicated, {case}")
    if case == "100k-720p":
        sc, view, W, H = scenes.make_scene(100_000, seed=0, mean_sigma=2.0), camera.orbit_view(8.0, 0.0, 0.0), 1280, 720
    elif case == "needles-giants-1280x720":
        sc = scenes.make_scene(20_000, seed=3, mean_sigma=1.5)
        rng = np.random.default_rng(3)
        n = sc.count
        sig = np.full((n, 3), np.exp(-7.0), np.float32)
        sig[:, 0] = np.exp(rng.uniform(-2.0, 0.5, n))
        giant = rng.random(n) < 0.01
        sig[giant] = np.exp(rng.uniform(-1.0, 0.7, (int(giant.sum()), 1))).astype(np.float32)
        sc.scales[:, :3] = sig
        sc.pos_alpha[:, 3] = (1.0 / (1.0 + np.exp(-rng.normal(3.0, 1.0, n)))).astype(np.float32)
        view, W, H = camera.orbit_view(6.0, 10.0, 30.0), 1280, 720
    elif case == "near-clamp-opacities-1280x720":
        sc = scenes.make_scene(60_000, seed=5, mean_sigma=2.0)
        rng = np.random.default_rng(5)
        n = sc.count
        o = rng.uniform(0.975, 0.9999, n)
        pick = rng.random(n)
        o[pick < 1 / 6] = 0.98
        o[(pick >= 1 / 6) & (pick < 1 / 3)] = 0.99
        sc.pos_alpha[:, 3] = o.astype(np.float32)
        sc.scales[:, :3] = np.exp(rng.normal(-5.0, 0.5, (n, 3))).astype(np.float32)
        view, W, H = camera.orbit_view(8.0, 0.0, 0.0), 1280, 720
    else:
        sc, view, W, H = scenes.make_test_scene(250_000, seed=99), camera.look_at([0.3, 0.5, 6.0], [0, 0, 0]), 640, 360
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, view, W, H)
    imgs = []
    for loop in ("culled", "predicated"):
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", render_loop=loop)
        imgs.append(r.render_to_host(cam).copy())
        r.close()
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    scene.close()


@pytest.mark.parametrize("fmt,degree", [("f32", 3), ("f32", 1), ("f32", 0), ("f16", 3), ("f16", 2)])
def test_device_pointer_scene_and_sh_degrees(gsr_ctx, fmt, degree):
    """Scenes created from device pointers render exactly what host-pointer scenes do, for the
    SH degrees whose records are not a whole number of 16-B quads too (K*3 = 3, 12, 27, 48)."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case device pointers fmt={fmt} degree={degree}")
    sc = scenes.make_scene(30_000, seed=11, mean_sigma=2.0, sh_degree=degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 10.0, 40.0), 640, 360)
    imgs = []
    for dev in (False, True):
        scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree, fmt=fmt,
                          device_pointers=dev)
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f")
        imgs.append(r.render_to_host(cam).copy())
        if not dev:
            gs, gc = r.read_back_splats()
        r.close()
        scene.close()
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    ref_scene = sc
    if fmt == "f16":
        from gsr.scenes import SyntheticScene

        g, c = oracle.encode_f16(sc)
        ref_scene = SyntheticScene(*oracle.decode_f16(g, c, sc.sh_degree), sc.sh_degree, sc.seed)
    ocam = oracle.OracleCamera.from_gsr(cam)
    os_, oc = oracle.preprocess(ref_scene, ocam, bounds=oracle.TIGHT)
    _check_splats(gs, gc, os_, oc)


def test_ply_scene_end_to_end(gsr_ctx, tmp_path):
    """A scene written as an INRIA .ply file, loaded by gsr_ply_load and rendered: splats and
    lists bit-exact against the oracle on the loaded arrays, image > 60 dB."""
    import gsr
    import oracle
    from gsr import camera, scenes
    from gsr.scenes import SyntheticScene

    log("case ply end to end, 50k 640x360")
    sc = scenes.make_scene(50_000, seed=21, mean_sigma=2.0)
    n, K = sc.count, sc.K
    a = sc.pos_alpha[:, 3].astype(np.float64).clip(1e-6, 1 - 1e-6)
    cols = [sc.pos_alpha[:, :3], np.zeros((n, 3), np.float32), sc.sh.reshape(n, K, 3)[:, 0, :],
            sc.sh.reshape(n, K, 3)[:, 1:, :].transpose(0, 2, 1).reshape(n, -1),
            np.log(a / (1 - a)).astype(np.float32)[:, None], np.log(sc.scales[:, :3]), sc.rotations]
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + \
        [f"f_rest_{i}" for i in range(3 * (K - 1))] + ["opacity", "scale_0", "scale_1", "scale_2"] + \
        [f"rot_{i}" for i in range(4)]
    header = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n + \
        "".join(f"property float {p}\n" for p in names) + "end_header\n"
    path = tmp_path / "scene.ply"
    path.write_bytes(header.encode() + np.concatenate(cols, axis=1).astype("<f4").tobytes())
    ply = gsr.load_ply(str(path))
    assert ply.count == n and ply.sh_degree == 3
    assert np.allclose(ply.pos_alpha, sc.pos_alpha, rtol=1e-5, atol=1e-6)
    loaded = SyntheticScene(ply.pos_alpha, ply.scales, ply.rotations, ply.sh, ply.sh_degree, 0)
    r, scene, cam, img, _, ocam, ob, os_, oc = _render_case(gsr_ctx, loaded, camera.orbit_view(8.0, 10.0, 40.0),
                                                             640, 360, "tight")
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    ref, _ = oracle.render(loaded, ocam, bounds=ob)
    assert oracle.psnr(img[..., :3], ref) > 60.0
    r.close()
    scene.close()


@pytest.mark.parametrize("mode", ["p32", "p24"])
def test_packed_render_splats(gsr_ctx, mode):
    """p32 / p24 render splats (test_scene_formats.cpp:248-316): the 48-byte splats (p24: decoded
    conic and opacity) and tile lists bit-exact, the packed words equal to the oracle's encoder
    for every visible splat, the image > 60 dB vs the oracle rendering the same records and
    > 75 / 70 dB vs the full-precision GPU image."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case packed render splats {mode}, 100k-720p")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    view = camera.orbit_view(8.0, 0.0, 0.0)
    r, scene, cam, img, sc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, view, 1280, 720, "tight", render_splat=mode)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    words = r.read_back_packed()
    vis = oc > 0
    assert np.array_equal(words[vis], oracle.encode_render_splats(os_[vis], mode))
    ref, _ = oracle.render(sc, ocam, bounds=ob, render_splat=mode)
    p_ref = oracle.psnr(img[..., :3], ref)
    rf = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f")
    full = rf.render_to_host(cam).copy()
    rf.close()
    p_full = oracle.psnr(img[..., :3], full[..., :3])
    log(f"  PSNR vs oracle({mode}) {p_ref:.2f} dB, vs full GPU image {p_full:.2f} dB")
    assert p_ref > 60.0
    assert p_full > (75.0 if mode == "p32" else 70.0)
    r.close()
    scene.close()


@pytest.mark.parametrize("window", [1, 4, 16, 17, 40, 64])
def test_stopthepop_window(gsr_ctx, window):
    """StopThePop window K (test_tight_bounds.cpp:472-481): the GPU image against the oracle with
    the same window > 60 dB (the lists are the INRIA-order lists, bit-exact), for windows held in
    registers (K <= 16) and in per-pixel LDS heaps (17..64, splat_tile_renderer.cpp:142);
    K > 64 is refused."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case StopThePop K={window}, test scene 250k 640x360")
    sc = scenes.make_test_scene(250_000, seed=99)
    view = camera.look_at([0.3, 0.5, 6.0], [0.0, 0.0, 0.0])
    cam = _camera(gsr, view, 640, 360)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_window=window)
    img = r.render_to_host(cam).copy()
    ocam = oracle.OracleCamera.from_gsr(cam)
    os_, oc = oracle.preprocess(sc, ocam, bounds=oracle.TIGHT)
    _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, oracle.TIGHT)
    ref, _ = oracle.render(sc, ocam, bounds=oracle.TIGHT, sort_window=window)
    inria, _ = oracle.render(sc, ocam, bounds=oracle.TIGHT)
    p, p_inria = oracle.psnr(img[..., :3], ref), oracle.psnr(img[..., :3], inria)
    log(f"  vs oracle same window {p:.2f} dB (vs INRIA order {p_inria:.2f} dB)")
    assert p > 60.0 and p > p_inria
    r.close()
    with pytest.raises(gsr.GsrError, match="sort_window"):
        gsr.Renderer(gsr_ctx, scene, sort_window=65)
    scene.close()


def test_adaptive_list_capacity(gsr_ctx):
    """adaptCapacity (splat_tile_renderer.cpp:444-486): a first frame that overflows a small guess
    grows the list to 1.5 M (then the lists are bit-exact against the oracle); 120 frames below
    0.4 capacity shrink it, never below min(256k, budget); the frame statistics arrive without a
    stall and name their frame."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log("case adaptive list, 300k 1920x1080")
    sc = scenes.make_scene(300_000, seed=2, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", adaptive_list=True, adaptive_initial_entries=4096)
    assert r.capacity == 4096 and r.max_capacity == 16267264  # 256 MB / 16.5 B, in blocks of 2048
    cam = _camera(gsr, camera.orbit_view(6.0, 10.0, 30.0), 1920, 1080)
    assert r.last_stats() is None
    r.render(cam)
    gsr_ctx.synchronize()
    ctl0 = r.read_control()
    assert ctl0.overflow > 0 and ctl0.entries == 4096
    r.render(cam)  # collects frame 0's statistics and grows before recording frame 1
    gsr_ctx.synchronize()
    st = r.last_stats()
    assert st is not None and st[1] in (0, 1) and st[0].duplicates == ctl0.duplicates
    m = ctl0.duplicates
    assert r.list_resizes == 1 and r.capacity == max(-(-int(1.5 * m) // 2048) * 2048, 256 * 1024)
    ctl1 = r.read_control()
    assert ctl1.overflow == 0 and ctl1.duplicates == m
    ocam = oracle.OracleCamera.from_gsr(cam)
    os_, oc = oracle.preprocess(sc, ocam, bounds=oracle.TIGHT)
    _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, oracle.TIGHT)
    log(f"  M {m}: capacity 4096 -> {r.capacity}")
    grown = r.capacity
    # a far view with few entries: 120 frames below 0.4 capacity shrink the list to 1.5 x their M
    far = _camera(gsr, camera.orbit_view(60.0, 10.0, 30.0), 1920, 1080)
    for _ in range(125):
        r.render(far)
        gsr_ctx.synchronize()
    m_far = r.read_control().duplicates
    want = max(-(-int(1.5 * m_far) // 2048) * 2048, 256 * 1024)
    log(f"  after 125 far frames (M {m_far}): capacity {r.capacity}, resizes {r.list_resizes}")
    assert 0 < m_far < 0.4 * grown
    assert r.list_resizes == 2 and r.capacity == want < grown
    st = r.last_stats()
    assert st is not None and st[1] >= 120
    r.render(cam)
    gsr_ctx.synchronize()
    assert r.read_control().duplicates == m  # the list still holds what fits
    r.close()
    scene.close()


def test_f16_scene_bit_exact(gsr_ctx):
    """The fp16 records: the GPU decode equals decodeGaussian, so splats and lists stay bit-exact."""
    import oracle
    from gsr import camera, scenes

    log("case f16 100k-720p tight")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    r, scene, cam, img, sc16, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, camera.orbit_view(8.0, 0.0, 30.0),
                                                                1280, 720, "tight", fmt="f16")
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    ref, _ = oracle.render(sc16, ocam, bounds=ob)
    assert oracle.psnr(img[..., :3], ref) > 60.0
    r.close()
    scene.close()


def test_reference_test_scene_ties_and_overflow(gsr_ctx):
    """test_tile_raster.cpp's renderer scene (250k, 640x360, eye (0.3, 0.5, 6)): exact depth
    ties inside tiles, near/behind/off-screen/transparent gaussians, and a 1 MB budget."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log("case test-scene 250k 640x360")
    sc = scenes.make_test_scene(250_000, seed=99)
    view = camera.look_at([0.3, 0.5, 6.0], [0.0, 0.0, 0.0])
    for bounds, mode in (("rect", "binning"), ("tight", "binning"), ("tight", "global")):
        r, scene, cam, img, sc2, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, view, 640, 360, bounds, sort_mode=mode)
        gs, gc = r.read_back_splats()
        _check_splats(gs, gc, os_, oc)
        keys, depth, vals, ranges, ctl = r.read_back()
        ties = int(((keys[1:] == keys[:-1]) & (depth[1:] == depth[:-1])).sum())
        log(f"  {bounds} {mode}: entries {ctl.entries}, adjacent equal (tile, depth): {ties}")
        assert ties > 0
        _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
        ref, st = oracle.render(sc2, ocam, bounds=ob)
        assert st["terminatedPixels"] > 0
        assert oracle.psnr(img[..., :3], ref) > 60.0
        r.close()
        scene.close()
    # Overflow: the entries past the capacity are dropped in duplication order: depth order in
    # binning mode (16.5 B per entry), gaussian-index order in global mode (24.5 B per entry).
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, view, 640, 360)
    ocam = oracle.OracleCamera.from_gsr(cam)
    os_, oc = oracle.preprocess(sc, ocam, bounds=oracle.TIGHT)
    for mode, capacity in (("binning", 63488), ("global", 40960)):
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", budget_mb=1, sort_mode=mode)
        r.render(cam)
        assert r.capacity == capacity
        ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, oracle.TIGHT, capacity=r.capacity)
        log(f"  1 MB budget, {mode}: duplicates {ctl.duplicates}, entries {ctl.entries}, overflow {ctl.overflow}")
        assert ctl.overflow > 0 and ctl.entries == r.capacity
        r.close()
    scene.close()


def test_global_tile_depth_sort_equals_full_key_sort(gsr_ctx):
    """Global mode orders the tile bits with the radix sort and each tile's range with the
    per-tile depth sort; on a dense scene (tiles of up to ~68k entries: LDS chunks plus merge
    rounds) the lists equal the full 45-bit radix sort's and the oracle's, bit for bit."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log("case global per-tile depth sort vs full-key sort, dense 200k 320x180")
    sc = scenes.make_scene(200_000, seed=7, mean_sigma=0.5)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 10.0, 20.0), 320, 180)
    out = []
    for full in (False, True):
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_mode="global", full_key_sort=full)
        img = r.render_to_host(cam).copy()
        lists = r.read_back()
        out.append((img, lists))
        if not full:
            ocam = oracle.OracleCamera.from_gsr(cam)
            os_, oc = oracle.preprocess(sc, ocam, bounds=oracle.TIGHT)
            ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, oracle.TIGHT)
            longest = int((lists[3][:, 1] - lists[3][:, 0]).max())
            log(f"  entries {ctl.entries}, longest tile range {longest}")
            assert longest > 4 * 2048
        r.close()
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    for a, b in zip(out[0][1][:4], out[1][1][:4]):
        assert np.array_equal(a, b)
    scene.close()


def test_binning_lists_equal_global(gsr_ctx):
    """Without overflow the two sort modes build the same list and the same image, bit for bit
    (splat_tile_renderer.h:8-26); binning rejects the 64-bit-only one-sweep sort."""
    import gsr
    from gsr import camera, scenes

    log("case binning vs global, 100k-720p")
    sc = scenes.make_scene(100_000, seed=3, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 20.0, 70.0), 1280, 720)
    out = []
    for mode in ("binning", "global"):
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode=mode)
        img = r.render_to_host(cam).copy()
        out.append((img, r.read_back()))
        r.close()
    assert out[0][1][4].overflow == 0 and out[0][1][4].entries > 0
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1][:4], out[1][1][:4]):
        assert np.array_equal(a, b)
    with pytest.raises(gsr.GsrError, match="one-sweep"):
        gsr.Renderer(gsr_ctx, scene, sort_mode="binning", sort_algorithm="onesweep")
    scene.close()


def test_onesweep_lists_equal_reduce_then_scan(gsr_ctx):
    """Both sort pass structures produce the same list, bit for bit, on the 100k workload."""
    import gsr
    from gsr import camera, scenes

    log("case onesweep vs reduce-then-scan, 100k-720p")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 0.0, 0.0), 1280, 720)
    out = []
    for alg in ("reduce-then-scan", "onesweep"):
        r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_algorithm=alg, sort_mode="global")
        img = r.render_to_host(cam).copy()
        out.append((img, r.read_back()))
        assert out[-1][1][4].sort_timeouts == 0
        r.close()
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1][:4], out[1][1][:4]):
        assert np.array_equal(a, b)
    scene.close()


@pytest.mark.parametrize("sort_mode", ["global", "binning"])
@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_1_2m_1080p_orbit_frame(gsr_ctx, sort_mode, fmt):
    """The headline workload at full size: 1.2M gaussians (means ~ N(0, 3^2 I), seed 0), 1920x1080,
    orbit frame 40 (48 deg, radius 10, elevation 15), tight bounds, in both sort modes and with the
    f32 and fp16 scene storage: splats, tile lists, ranges and control counts bit-exact, the float
    image > 60 dB and the RGBA8 image within one level of the quantised oracle image."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case 1.2M-1080p orbit frame 40, tight, sort_mode={sort_mode}, scene {fmt}")
    sc = scenes.make_scene(1_200_000, seed=0, mean_sigma=3.0)
    view = camera.orbit_view(10.0, 15.0, 360.0 * 40 / 300)
    r, scene, cam, img, osc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, view, 1920, 1080, "tight", fmt=fmt,
                                                              sort_mode=sort_mode)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    ref, st = oracle.render(osc, ocam, bounds=ob)
    p = oracle.psnr(img[..., :3], ref)
    r8 = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8", sort_mode=sort_mode)
    img8 = r8.render_to_host(cam)
    diff = np.abs(img8[..., :3].astype(np.int32) - oracle.quantize8(ref).astype(np.int32))
    p8 = oracle.psnr(img8[..., :3].astype(np.float64) / 255.0, ref)
    log(f"  V {ctl.visible}, M {ctl.duplicates}, float PSNR {p:.2f} dB, RGBA8 {p8:.2f} dB, max {diff.max()} level")
    assert p > 60.0 and diff.max() <= 1 and p8 >= 58.7
    assert ctl.duplicates == st["tileInstances"]
    r8.close()
    r.close()
    scene.close()


@pytest.mark.parametrize("bounds,sort_mode,render_splat", [("tight", "binning", "full"), ("tight", "global", "full"),
                                                           ("rect", "binning", "full"), ("tight", "binning", "p32"),
                                                           ("tight", "global", "p24"),
                                                           ("cutoff-only", "binning", "full")])
def test_needle_scene_lists_and_image(gsr_ctx, bounds, sort_mode, render_splat):
    """The reference's tight-bounds scene (test_tight_bounds.cpp:181-211, 150k gaussians: every 5th a
    needle with axes (3 s, 0.01 s, 0.02 s), arbitrary rotations, every 7th alpha below 1/255 x 2.55,
    near-opaque and tied-depth gaussians) seen from (0.3, 0.5, 6) at 640x360 (:218-225), checked
    against the oracle as test_tight_bounds.cpp:343-398 checks the GPU against its CPU
    construction, but with nothing left out: splats and tile counts bit-exact, lists and ranges
    bit-exact (no borderline exclusions), no overflow or coverage mismatch, and the tight and
    cutoff-only images against the oracle render with the same bounds; tight bounds drop entries
    and leave the image of rect bounds unchanged (tight_bounds.cpp:433-447)."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case needle scene {bounds} {sort_mode} {render_splat}")
    sc = scenes.make_needle_scene(150_000, seed=77)
    view = camera.look_at([0.3, 0.5, 6.0], [0.0, 0.0, 0.0])
    r, scene, cam, img, osc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, view, 640, 360, bounds, sort_mode=sort_mode,
                                                              render_splat=render_splat)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    assert ctl.overflow == 0 and ctl.coverage_mismatch == 0
    ref, st = oracle.render(osc, ocam, bounds=ob, render_splat=render_splat)
    p = oracle.psnr(img[..., :3], ref)
    rect = gsr.Renderer(gsr_ctx, scene, bounds="rect", target="rgba32f", sort_mode=sort_mode, render_splat=render_splat)
    rimg = rect.render_to_host(cam)
    rctl = rect.read_control()
    log(f"  entries {ctl.entries} (rect {rctl.entries}), needles visible "
        f"{int(((gs['cull_reason'] == 0) & (np.arange(sc.count) % 5 == 0)).sum())}, PSNR {p:.2f} dB")
    assert p > 60.0
    if bounds == "tight":
        assert ctl.entries < rctl.entries
        assert np.array_equal(img, rimg), "tight bounds changed the image"
    rect.close()
    r.close()
    scene.close()


def test_deterministic_and_empty(gsr_ctx):
    import gsr
    from gsr import camera, scenes

    log("case determinism + empty scene")
    sc = scenes.make_scene(50_000, seed=5, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba8")
    cam = _camera(gsr, camera.orbit_view(8.0, 10.0, 20.0), 640, 360)
    a = r.render_to_host(cam).copy()
    la = r.read_back()
    b = r.render_to_host(cam).copy()
    lb = r.read_back()
    assert np.array_equal(a, b)
    for x, y in zip(la[:4], lb[:4]):
        assert np.array_equal(x, y)
    r.close()
    scene.close()
    empty = gsr.Scene(gsr_ctx, np.zeros((0, 4)), np.zeros((0, 4)), np.zeros((0, 4)), np.zeros((0, 48)), 3)
    r = gsr.Renderer(gsr_ctx, empty, bounds="tight", target="rgba8")
    img = r.render_to_host(cam)
    assert np.all(img == 0)
    ctl = r.read_control()
    assert ctl.duplicates == 0 and ctl.entries == 0 and ctl.gaussian_count == 0
    r.close()
    empty.close()


SORT_SIZES = [0, 1, 31, 32, 33, 255, 256, 257, 2047, 2048, 2049, 4095, 4097, 65537, 1_000_003]


@pytest.mark.parametrize("algorithm", ["reduce-then-scan", "onesweep"])
@pytest.mark.parametrize("n", SORT_SIZES)
def test_sort_pairs_stable(gsr_ctx, n, algorithm):
    """test_gpu_sort.cpp:198-258: sizes around block boundaries, adversarial key patterns."""
    import gsr

    rng = np.random.default_rng(1234 + n)
    patterns = {
        "random44": rng.integers(0, 1 << 44, n, dtype=np.uint64),
        "all-equal": np.full(n, 0xDEADBEEF, np.uint64),
        "sorted": np.arange(n, dtype=np.uint64),
        "reversed": np.arange(n, dtype=np.uint64)[::-1].copy(),
        "4-values": rng.integers(0, 4, n, dtype=np.uint64) << np.uint64(20),
        "top-byte": rng.integers(0, 256, n, dtype=np.uint64) << np.uint64(36),
    }
    for name, keys in patterns.items():
        if n > 100_000 and name not in ("random44", "4-values"):
            continue
        vals = np.arange(n, dtype=np.uint32)
        ok, ov = gsr.sort_pairs(gsr_ctx, keys, vals, key_bits=44, algorithm=algorithm)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ok, keys[order]), f"{name}: keys"
        assert np.array_equal(ov, vals[order]), f"{name}: values (stability)"


@pytest.mark.parametrize("key_bits", [9, 13, 21, 33])
def test_sort_pairs_uneven_digits(gsr_ctx, key_bits):
    """The reduce-then-scan passes split the key bits evenly (9: 5 + 4, 13: 7 + 6, 21: 7 + 7 + 7,
    33: 7 + 7 + 7 + 6 + 6): a stable sort whatever the split, keys using every bit."""
    import gsr

    n = 100_003
    rng = np.random.default_rng(key_bits)
    keys = rng.integers(0, 1 << key_bits, n, dtype=np.uint64)
    keys[:16] = (1 << key_bits) - 1  # the top digit value of every pass
    vals = np.arange(n, dtype=np.uint32)
    ok, ov = gsr.sort_pairs(gsr_ctx, keys, vals, key_bits=key_bits, algorithm="reduce-then-scan")
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])


@pytest.mark.parametrize("algorithm", ["reduce-then-scan", "onesweep"])
def test_sort_pairs_count_below_capacity(gsr_ctx, algorithm):
    """A device-side count < the uploaded keys: only the first `count` are sorted, the tail is untouched."""
    import gsr

    n, count = 10_000, 6_001
    keys = np.random.default_rng(7).integers(0, 1 << 40, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    ok, ov = gsr.sort_pairs(gsr_ctx, keys, vals, key_bits=40, count=count, algorithm=algorithm)
    order = np.argsort(keys[:count], kind="stable")
    assert np.array_equal(ok[:count], keys[:count][order])
    assert np.array_equal(ov[:count], vals[:count][order])
    assert np.array_equal(ok[count:], keys[count:]) and np.array_equal(ov[count:], vals[count:])


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 100_000, 1_000_003])
def test_exclusive_scan(gsr_ctx, n):
    """test_tile_raster.cpp:114-170: the scan vs a CPU prefix sum."""
    import gsr

    v = np.random.default_rng(n).integers(0, 50, n, dtype=np.uint32)
    out, total = gsr.exclusive_scan(gsr_ctx, v)
    ref = np.concatenate([[0], np.cumsum(v, dtype=np.uint64)[:-1]]).astype(np.uint32) if n else v
    assert np.array_equal(out, ref)
    assert total == int(v.astype(np.uint64).sum())


@pytest.mark.parametrize("sort_mode", ["binning", "global"])
def test_tile_stats_match_oracle_counts(gsr_ctx, sort_mode):
    """The stats build of the render (tile_stats) on the small workload: per-tile list lengths and
    pixel counts exact; the evaluated / blended pair and terminated pixel totals equal those of the
    oracle's renderReference (same definitions: a pixel walks its tile's list until it terminates)
    up to the alpha cutoff's rounding; the image is the plain render's, bit for bit."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case tile stats 100k-720p sort_mode={sort_mode}")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 0.0, 0.0), 1280, 720)
    r = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_mode=sort_mode, tile_stats=True)
    plain = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_mode=sort_mode)
    img = r.render_to_host(cam).copy()
    assert np.array_equal(img, plain.render_to_host(cam)), "the stats build changed the image"
    st = r.tile_stats().astype(np.int64)
    _, _, _, ranges, _ = r.read_back()
    ranges = ranges.reshape(-1, 2).astype(np.int64)
    assert np.array_equal(st[:, 0], ranges[:, 1] - ranges[:, 0])
    tx, ty = np.meshgrid(np.arange(cam.tiles_x), np.arange(cam.tiles_y))
    w = np.minimum(16, 1280 - 16 * tx).ravel()
    h = np.minimum(16, 720 - 16 * ty).ravel()
    assert np.array_equal(st[:, 7], w * h)
    assert (st[:, 1] <= st[:, 0]).all() and (st[:, 4] <= st[:, 7]).all()
    assert (st[st[:, 4] < st[:, 7], 1] == st[st[:, 4] < st[:, 7], 0]).all(), "a tile stopped with live pixels"
    assert (st[:, 5] % 256 == 0).all() and (st[:, 3] <= st[:, 2]).all() and (st[:, 6] <= st[:, 1]).all()
    _, ost = oracle.render(sc, oracle.OracleCamera.from_gsr(cam), bounds=oracle.TIGHT)
    ev, bl, te = st[:, 2].sum(), st[:, 3].sum(), st[:, 4].sum()
    log(f"  evaluated {ev} (oracle {ost['evaluatedFragments']}), blended {bl} ({ost['blendedFragments']}), "
        f"terminated {te} ({ost['terminatedPixels']}), lane slots {st[:, 5].sum()}")
    assert abs(ev - ost["evaluatedFragments"]) <= 1e-4 * ost["evaluatedFragments"]
    assert abs(bl - ost["blendedFragments"]) <= 1e-4 * ost["blendedFragments"]
    assert abs(te - ost["terminatedPixels"]) <= 1e-3 * max(ost["terminatedPixels"], 1)
    r.close()
    plain.close()
    scene.close()


def test_kernel_times(gsr_ctx):
    """profile_stages=2: one event after every kernel of the frame; the kernels appear in launch
    order and their times add up to the frame's stage total. Global mode has no frame-setup scan
    (the duplicate pass takes its block offsets from the preprocess's superblock sums), and with
    the culled render no tile depth sort pass (the render's workgroups sort their own tiles)."""
    import gsr
    from gsr import camera, scenes

    log("case kernel times")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 0.0, 0.0), 1280, 720)
    for mode, expect in (("global", ["kPreprocess", "kDuplicate", "kSortHistogram", "kSortScan",
                                     "kSortScatter", "kRanges", "kRenderCulled"]),
                         ("binning", ["kPreprocess", "kFrameSetup", "kCompactVisible", "kGatherCounts",
                                      "kDuplicate", "kRanges", "kRenderCulled"])):
        r = gsr.Renderer(gsr_ctx, scene, target="rgba8", sort_mode=mode, profile_stages=2)
        r.render_to_host(cam)
        kt = r.kernel_times()
        names = [n for n, _ in kt]
        assert mode == "binning" or ("kFrameSetup" not in names and "kTileDepthSort" not in names), names
        pos = [names.index(k) for k in expect]
        assert pos == sorted(pos), names
        assert all(ms >= 0.0 for _, ms in kt)
        total = r.stage_times()["total"]
        assert abs(sum(ms for _, ms in kt) - total) <= 0.05 * total + 0.01, (sum(ms for _, ms in kt), total)
        r.close()
    scene.close()


def test_tile_depth_sort_long_ranges(gsr_ctx):
    """The per-tile depth sort of global mode on every size class of tile range: a 600k-gaussian
    cloud (means ~ N(0, 1.5^2 I)) at 640x360 from radius 7 has ranges up to 2048 entries (one wave,
    registers), 2049..4096 and 4097..8192 (2 or 4 waves, registers with cross-wave stages through
    LDS) and over 8192 (LDS chunks and merge rounds): lists and ranges bit-exact against the oracle
    and equal to the full 64-bit key sort, both for the sort inside the culled render (the default)
    and for the tile depth sort pass (the other render loops)."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log("case tile depth sort: long ranges")
    sc = scenes.make_scene(600_000, seed=4, mean_sigma=1.5)
    view = camera.orbit_view(7.0, 10.0, 30.0)
    r, scene, cam, img, osc, ocam, ob, os_, oc = _render_case(gsr_ctx, sc, view, 640, 360, "tight", sort_mode="global")
    ctl = _check_lists(r, os_, oc, cam.tiles_x, cam.tiles_y, ob)
    _, _, _, ranges, _ = r.read_back()
    L = ranges.reshape(-1, 2)[:, 1].astype(np.int64) - ranges.reshape(-1, 2)[:, 0]
    bands = [int(((L > a) & (L <= b)).sum()) for a, b in ((0, 2048), (2048, 4096), (4096, 8192), (8192, 1 << 30))]
    log(f"  {ctl.entries} entries; tiles per range band <=2048 / <=4096 / <=8192 / more: {bands}")
    assert all(b > 0 for b in bands)
    full = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_mode="global", full_key_sort=True)
    fimg = full.render_to_host(cam)
    assert np.array_equal(fimg, img)
    for a, b in zip(full.read_back()[:4], r.read_back()[:4]):
        assert np.array_equal(a, b)
    full.close()
    sep = gsr.Renderer(gsr_ctx, scene, bounds="tight", target="rgba32f", sort_mode="global", render_loop="predicated")
    sep.render_to_host(cam)
    for a, b in zip(sep.read_back()[:4], r.read_back()[:4]):
        assert np.array_equal(a, b)
    sep.close()
    r.close()
    scene.close()


@pytest.mark.parametrize("bounds,sort_mode,render_splat,case", [
    ("tight", "binning", "full", "100k-720p"), ("tight", "global", "full", "100k-720p"),
    ("rect", "binning", "full", "100k-720p"), ("tight", "binning", "full", "needles"),
    ("tight", "global", "p32", "needles"), ("rect", "binning", "full", "needles"),
    ("cutoff-only", "binning", "full", "needles")])
def test_tile_size_8(gsr_ctx, bounds, sort_mode, render_splat, case):
    """tileSize 8 (splat_tile_renderer.cpp:139, 537-541; the cases of test_tight_bounds.cpp:413-423):
    getRect stays on 16-px tiles and the lists are built on 8-px bins. Tile counts, lists and ranges
    over the bin grid bit-exact against the oracle with 8-px bins; the image > 60 dB against the
    oracle render, and with tight bounds bit-identical to the rect image of the same renderer form
    (fewer entries, nothing visible dropped)."""
    import gsr
    import oracle
    from gsr import camera, scenes

    log(f"case tile size 8, {case}, {bounds} {sort_mode} {render_splat}")
    if case == "needles":
        sc, view, W, H = scenes.make_needle_scene(150_000, seed=77), camera.look_at([0.3, 0.5, 6.0], [0, 0, 0]), 640, 360
    else:
        sc, view, W, H = scenes.make_scene(100_000, seed=0, mean_sigma=2.0), camera.orbit_view(8.0, 0.0, 0.0), 1280, 720
    cam = _camera(gsr, view, W, H)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    r = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target="rgba32f", sort_mode=sort_mode, render_splat=render_splat,
                     tile_size=8)
    img = r.render_to_host(cam).copy()
    ob = {"rect": oracle.RECT, "tight": oracle.TIGHT, "cutoff-only": oracle.CUTOFF}[bounds]
    ocam = oracle.OracleCamera.from_gsr(cam)
    os_, oc = oracle.preprocess(sc, ocam, bounds=ob, render_splat=render_splat, bin_size=8)
    gs, gc = r.read_back_splats()
    _check_splats(gs, gc, os_, oc)
    bx, by = r.bins()
    assert (bx, by) == ((W + 7) // 8, (H + 7) // 8)
    keys, depth, vals, ranges, ctl = r.read_back()
    hi, lo, ov, orng, dup, mism = oracle.tile_lists(os_, oc, cam.tiles_x, cam.tiles_y, bounds=ob,
                                                    mode={"global": 0, "binning": 1}[sort_mode], bin_size=8,
                                                    grid=(bx, by))
    assert ctl.duplicates == dup and ctl.coverage_mismatch == 0 and mism == 0 and ctl.overflow == 0
    assert np.array_equal(keys, hi) and np.array_equal(depth, lo) and np.array_equal(vals, ov)
    assert np.array_equal(ranges, orng)
    ref, _ = oracle.render(sc, ocam, bounds=oracle.TIGHT if bounds != "rect" else oracle.RECT,
                           render_splat=render_splat)
    p = oracle.psnr(img[..., :3], ref)
    r16 = gsr.Renderer(gsr_ctx, scene, bounds=bounds, target="rgba32f", sort_mode=sort_mode,
                       render_splat=render_splat)
    r16.render_to_host(cam)
    log(f"  {ctl.entries} entries on {bx}x{by} bins (16-px tiles: {r16.read_control().entries}), PSNR {p:.2f} dB")
    assert p > 60.0
    if bounds == "tight":
        rr = gsr.Renderer(gsr_ctx, scene, bounds="rect", target="rgba32f", sort_mode=sort_mode,
                          render_splat=render_splat, tile_size=8)
        rimg = rr.render_to_host(cam)
        assert rr.read_control().entries > ctl.entries
        assert np.array_equal(img, rimg), "tight 8-px bins changed the image"
        rr.close()
    with pytest.raises(gsr.GsrError, match="tile_size"):
        gsr.Renderer(gsr_ctx, scene, tile_size=4)
    with pytest.raises(gsr.GsrError, match="tile_size 8"):
        gsr.Renderer(gsr_ctx, scene, tile_size=8, sort_window=4)
    r16.close()
    r.close()
    scene.close()


@pytest.mark.parametrize("tile_size,batches", [(16, (128, 512)), (8, (32, 128))])
def test_render_batch_depth_same_image(gsr_ctx, tile_size, batches):
    """render_batch (TileSettings::renderBatch): the number of splats staged per round changes where
    the rounds start and the whole-tile vote falls, not the per-pixel blend order, so every batch
    depth renders the default's image bit for bit; other values are refused."""
    import gsr
    from gsr import camera, scenes

    log(f"case render batch depths, tile size {tile_size}")
    sc = scenes.make_scene(100_000, seed=0, mean_sigma=2.0)
    scene = gsr.Scene(gsr_ctx, sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    cam = _camera(gsr, camera.orbit_view(8.0, 0.0, 0.0), 1280, 720)
    base = gsr.Renderer(gsr_ctx, scene, target="rgba32f", sort_mode="global", tile_size=tile_size)
    want = base.render_to_host(cam).copy()
    for b in batches:
        r = gsr.Renderer(gsr_ctx, scene, target="rgba32f", sort_mode="global", tile_size=tile_size, render_batch=b)
        assert np.array_equal(r.render_to_host(cam), want), f"batch {b} changed the image"
        r.close()
    with pytest.raises(gsr.GsrError, match="render_batch"):
        gsr.Renderer(gsr_ctx, scene, tile_size=tile_size, render_batch=96)
    with pytest.raises(gsr.GsrError, match="render_batch"):
        gsr.Renderer(gsr_ctx, scene, tile_size=16, render_batch=128, sort_window=4)
    base.close()
    scene.close()
