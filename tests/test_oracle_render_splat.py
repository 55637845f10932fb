"""Oracle pins for the packed render splats (render_splat.h:1-69, render_splat.glsl:1-75).

- encodeRenderSplat against an encoder written here from the documented word layout with
  numpy's IEEE fp16 cast (round to nearest even);
- the p24 round trip: the conic and opacity the oracle's preprocess keeps equal numpy's
  decode of the stored halves (conic / 2^12), and the tiles are counted with those;
- what the packing costs in image quality (test_scene_formats.cpp:304): p32 > 75 dB and
  p24 > 70 dB against the full-precision image of the same scene and camera.
"""
import numpy as np


def _scene_and_camera(oracle, n=40_000, W=640, H=360):
    from gsr import camera, scenes

    sc = scenes.make_test_scene(n, seed=99)
    cam = oracle.OracleCamera.from_viewer(camera.look_at([0.3, 0.5, 6.0], [0, 0, 0]), camera.perspective(60.0, W / H),
                                          camera.IDENTITY, W, H)
    return sc, cam


def _h(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)


def _numpy_encode(s, mode):
    cx, cy = s["center"][:, 0], s["center"][:, 1]
    co, col = s["conic_opacity"], s["color"]
    rad = np.minimum(s["radius"], np.float32(65535.0)).astype(np.uint32)
    bits = lambda a: np.ascontiguousarray(a, np.float32).view(np.uint32)
    rg = _h(col[:, 0]) | (_h(col[:, 1]) << 16)
    br = _h(col[:, 2]) | (rad << 16)
    if mode == "p32":
        return np.stack([bits(cx), bits(cy), bits(co[:, 0]), bits(co[:, 1]), bits(co[:, 2]), bits(co[:, 3]), rg, br], 1)
    S = np.float32(4096.0)
    return np.stack([bits(cx), bits(cy), _h(co[:, 0] * S) | (_h(co[:, 1] * S) << 16),
                     _h(co[:, 2] * S) | (_h(co[:, 3]) << 16), rg, br], 1)


def test_encoder_matches_the_documented_layout(oracle):
    sc, cam = _scene_and_camera(oracle)
    sp, counts = oracle.preprocess(sc, cam, bounds=oracle.TIGHT)
    vis = sp[counts > 0]
    assert vis.shape[0] > 1000
    for mode in ("p32", "p24"):
        assert np.array_equal(oracle.encode_render_splats(vis, mode), _numpy_encode(vis, mode)), mode


def test_p24_preprocess_keeps_the_decoded_conic(oracle):
    sc, cam = _scene_and_camera(oracle)
    full, cf = oracle.preprocess(sc, cam, bounds=oracle.TIGHT)
    p24, c24 = oracle.preprocess(sc, cam, bounds=oracle.TIGHT, render_splat="p24")
    vis = full["cull_reason"] == 0
    co = full["conic_opacity"][vis]
    S = np.float32(4096.0)
    want = np.empty_like(co)
    want[:, :3] = (co[:, :3] * S).astype(np.float16).astype(np.float32) * np.float32(1.0 / 4096.0)
    want[:, 3] = co[:, 3].astype(np.float16).astype(np.float32)
    assert np.array_equal(p24["conic_opacity"][vis], want)
    assert np.array_equal(p24["color"][vis], full["color"][vis])  # the 48-byte splat keeps fp32 colour
    assert np.array_equal(oracle.preprocess(sc, cam, bounds=oracle.TIGHT, render_splat="p32")[0], full)
    # the counts follow the decoded ellipse: almost always the same tiles
    assert (c24 != cf).mean() < 0.01


def test_packing_costs_little_image_quality(oracle):
    sc, cam = _scene_and_camera(oracle)
    full, _ = oracle.render(sc, cam, bounds=oracle.TIGHT)
    p32, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, render_splat="p32")
    p24, _ = oracle.render(sc, cam, bounds=oracle.TIGHT, render_splat="p24")
    assert not np.array_equal(full, p32)
    assert oracle.psnr(p32, full) > 75.0
    assert oracle.psnr(p24, full) > 70.0
