"""Oracle pins for the tight tile coverage, the deterministic log and the fp16 scene format.

- Tight coverage (test_tight_bounds.cpp:78-176): on random 2D conics, needles included, the
  tight tiles are a subset of INRIA's rectangle, every pixel whose alpha reaches 1/255 lies in
  a kept tile (alpha evaluated independently here, in numpy float32), and count = the tiles
  enumerated row by row.
- Tight image = rect image, bit for bit (test_tight_bounds.cpp:242-264).
- det_logf = log rounded to float (numpy's float64 log), within 1 ulp, almost always exact.
- fp16: half -> float exhaustively and float -> half (round to nearest even) against numpy.
"""
import math

import numpy as np


def _random_splat(rng, W, H):
    from oracle import SPLAT_DTYPE

    s = np.zeros(1, SPLAT_DTYPE)[0]
    s["center"] = (rng.uniform() * (W + 60) - 30, rng.uniform() * (H + 60) - 30)
    l1 = 0.3 + 10 ** (rng.uniform() * 4)
    l2 = 0.3 + 10 ** (rng.uniform() * 2) * rng.uniform()
    th = rng.uniform() * 2 * math.pi
    c, sn = math.cos(th), math.sin(th)
    cov = np.array([c * c * l1 + sn * sn * l2, c * sn * (l1 - l2), sn * sn * l1 + c * c * l2], np.float32)
    det = cov[0] * cov[2] - cov[1] * cov[1]
    if det == 0:
        return None
    inv = np.float32(1.0) / det
    conic = np.array([cov[2] * inv, -cov[1] * inv, cov[0] * inv], np.float32)
    mid = np.float32(0.5) * (cov[0] + cov[2])
    lam = mid + np.sqrt(max(np.float32(0.1), mid * mid - det))
    radius = np.ceil(np.float32(3.0) * np.sqrt(lam))
    u = rng.uniform()
    o = (1 / 255 + u * 0.01) if rng.uniform() < 0.2 else (0.99 + 0.01 * u if rng.uniform() < 0.15 else u)
    s["conic_opacity"] = (*conic, o)
    s["radius"] = radius
    return s


def _alpha_grid(s, xs, ys):
    """alpha >= 1/255 on a pixel grid (splatAlphaAt restated in numpy float32)."""
    dx = np.float32(s["center"][0]) - xs.astype(np.float32)[None, :]
    dy = np.float32(s["center"][1]) - ys.astype(np.float32)[:, None]
    a, b, c, o = (np.float32(v) for v in s["conic_opacity"])
    power = np.float32(-0.5) * (a * dx * dx + c * dy * dy) - b * dx * dy
    alpha = np.minimum(np.float32(0.99), o * np.exp(power, dtype=np.float32))
    return (power <= 0) & (alpha >= np.float32(1 / 255))


def test_tight_coverage_invariants(oracle):
    rng = np.random.default_rng(1234)
    W, H = 400, 300
    tx, ty = (W + 15) // 16, (H + 15) // 16
    tested = rect_tiles = tight_tiles = 0
    for _ in range(1500):
        s = _random_splat(rng, W, H)
        if s is None:
            continue
        r = oracle.tile_rect(float(s["center"][0]), float(s["center"][1]), float(s["radius"]), tx, ty)
        if r[2] <= r[0] or r[3] <= r[1]:
            continue
        tested += 1
        rect_tiles += (r[2] - r[0]) * (r[3] - r[1])
        tiles, count, _ = oracle.coverage_tiles(s, oracle.TIGHT, tx, ty)
        assert count == len(tiles)
        assert len(set(tiles.tolist())) == len(tiles)
        tight_tiles += count
        kept = np.zeros((ty, tx), bool)
        kept.reshape(-1)[tiles] = True
        # subset of the rectangle
        rmask = np.zeros((ty, tx), bool)
        rmask[r[1]:r[3], r[0]:r[2]] = True
        assert not np.any(kept & ~rmask)
        # every pixel of the rectangle with alpha >= 1/255 lies in a kept tile
        xs = np.arange(r[0] * 16, min(r[2] * 16, W))
        ys = np.arange(r[1] * 16, min(r[3] * 16, H))
        if xs.size and ys.size:
            vis = _alpha_grid(s, xs, ys)
            tile_of = kept[(ys // 16)[:, None], (xs // 16)[None, :]]
            assert not np.any(vis & ~tile_of), "a pixel with alpha >= 1/255 lies in a dropped tile"
        # rect mode enumerates the whole rectangle
        rt, rc, _ = oracle.coverage_tiles(s, oracle.RECT, tx, ty)
        assert rc == (r[2] - r[0]) * (r[3] - r[1]) == len(rt)
    assert tested > 700
    assert tight_tiles < rect_tiles


def test_tight_coverage_invariants_8px_bins(oracle):
    """8-px bins (tileSize 8, test_tight_bounds.cpp:78-176): the tight bins are a subset of the
    16-px rectangle's bins, hold every pixel with alpha >= 1/255, lie inside the 16-px tight tiles
    (checked for conics with |b| / sqrt(a c) < 0.999: for a nearly degenerate needle the two grids'
    row spans are rounded from cancelling terms and may disagree by a bin at the far end, where no
    pixel reaches the cutoff), and rect bounds enumerate the rectangle's four sub-bins per tile,
    clipped to the bin grid."""
    rng = np.random.default_rng(4321)
    W, H = 404, 300  # 51 x 38 bins: the last bin column is clipped
    tx, ty = (W + 15) // 16, (H + 15) // 16
    gx, gy = (W + 7) // 8, (H + 7) // 8
    tested = 0
    for _ in range(1200):
        s = _random_splat(rng, W, H)
        if s is None:
            continue
        r = oracle.tile_rect(float(s["center"][0]), float(s["center"][1]), float(s["radius"]), tx, ty)
        if r[2] <= r[0] or r[3] <= r[1]:
            continue
        tested += 1
        bins, count, _ = oracle.coverage_tiles(s, oracle.TIGHT, tx, ty, S=8, grid=(gx, gy))
        assert count == len(bins) and len(set(bins.tolist())) == count
        kept = np.zeros((gy, gx), bool)
        kept.reshape(-1)[bins] = True
        rmask = np.zeros((gy, gx), bool)
        rmask[2 * r[1]:min(2 * r[3], gy), 2 * r[0]:min(2 * r[2], gx)] = True
        assert not np.any(kept & ~rmask)
        tiles16, _, _ = oracle.coverage_tiles(s, oracle.TIGHT, tx, ty)
        k16 = np.zeros((ty, tx), bool)
        k16.reshape(-1)[tiles16] = True
        a, b, c = (float(v) for v in s["conic_opacity"][:3])
        if abs(b) / np.sqrt(a * c) < 0.999:
            by, bx = np.nonzero(kept)
            assert np.all(k16[by // 2, bx // 2]), "an 8-px bin outside the 16-px tight tiles"
        xs = np.arange(r[0] * 16, min(r[2] * 16, W))
        ys = np.arange(r[1] * 16, min(r[3] * 16, H))
        if xs.size and ys.size:
            vis = _alpha_grid(s, xs, ys)
            assert not np.any(vis & ~kept[(ys // 8)[:, None], (xs // 8)[None, :]]), "a visible pixel was dropped"
        rb, rc, _ = oracle.coverage_tiles(s, oracle.RECT, tx, ty, S=8, grid=(gx, gy))
        assert rc == int(rmask.sum()) == len(rb)
    assert tested > 500


def test_tight_image_bit_identical_to_rect(oracle):
    from gsr import camera, scenes

    sc = scenes.make_test_scene(30_000, seed=77)
    view = camera.look_at([0.3, 0.5, 6.0], [0, 0, 0])
    cam = oracle.OracleCamera.from_viewer(view, camera.perspective(60.0, 320 / 180), camera.IDENTITY, 320, 180)
    rect, st_r = oracle.render(sc, cam, bounds=oracle.RECT)
    tight, st_t = oracle.render(sc, cam, bounds=oracle.TIGHT)
    assert np.array_equal(rect.view(np.uint32), tight.view(np.uint32))
    assert st_t["tileInstances"] < st_r["tileInstances"]
    assert st_t["blendedFragments"] == st_r["blendedFragments"]
    assert st_t["terminatedPixels"] == st_r["terminatedPixels"]


def test_det_logf_is_the_rounded_log(oracle):
    rng = np.random.default_rng(5)
    xs = np.concatenate([
        np.float32(255.0) * rng.uniform(1 / 255, 1.0, 20000).astype(np.float32),  # the tight-bounds domain
        np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32),
        np.array([1.0, 2.0, 255.0, 0.99999994, 1.0000001, 1e-40, 3e38], np.float32),
    ])
    got = np.array([oracle.det_logf(float(x)) for x in xs], np.float32)
    ref = np.log(xs.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert (ulps != 0).mean() < 1e-3
    assert oracle.det_logf(1.0) == 0.0


def test_half_to_float_exhaustive(oracle):
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([oracle.half_to_float(int(v)) for v in h], np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_float_to_half_round_to_nearest_even(oracle):
    rng = np.random.default_rng(9)
    bits = rng.integers(0, 2 ** 32, 60000, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    f = f[np.isfinite(f)]
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 6.1035156e-05, 5.96e-08, 2.98e-08, 1.0 + 2 ** -11,
                        1.0 + 3 * 2 ** -11], np.float32)
    f = np.concatenate([f, special, rng.normal(0, 0.1, 20000).astype(np.float32)])
    got = np.array([oracle.float_to_half(float(v)) for v in f], np.uint16)
    with np.errstate(over="ignore"):
        ref = f.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, ref)


def test_f16_record_round_trip(oracle):
    from gsr import scenes

    sc = scenes.make_scene(2000, seed=4)
    g, c = oracle.encode_f16(sc)
    pa, sc_, ro, sh = oracle.decode_f16(g, c, sc.sh_degree)
    assert np.array_equal(pa[:, :3], sc.pos_alpha[:, :3])  # positions stay fp32
    assert np.array_equal(pa[:, 3], sc.pos_alpha[:, 3].astype(np.float16).astype(np.float32))
    assert np.array_equal(sc_[:, :3], sc.scales[:, :3].astype(np.float16).astype(np.float32))
    assert np.array_equal(sh, sc.sh.astype(np.float16).astype(np.float32))
    assert np.allclose(np.linalg.norm(ro, axis=1), 1.0, atol=1e-6)
    assert np.allclose(ro, sc.rotations, atol=2e-3)
