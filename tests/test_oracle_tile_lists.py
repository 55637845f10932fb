"""Oracle pins for the tile lists of both sort modes (splat_tile_renderer.h:8-36,
test_tile_raster.cpp:332-382).

The oracle's lists are checked against a brute force built here from per-splat coverage
(itself pinned in test_oracle_coverage_fp16.py) and Python's sort:
  - global mode: all (tile, depth bits, index) entries, sorted by (tile, depth bits) with ties in
    index order; past the capacity, the entries of the highest indices are dropped;
  - binning mode: the same list when nothing overflows; past the capacity, the entries are
    dropped in depth order (the farthest splats first), and the rest stay sorted by tile, then
    by (depth, index);
  - ranges = [first, last + 1) of each tile in the list.
"""
import numpy as np
import pytest


def _small_case(oracle, count=3000, seed=21, W=320, H=180):
    from gsr import camera, scenes

    sc = scenes.make_scene(count, seed=seed, mean_sigma=1.5)
    view = camera.orbit_view(6.0, 10.0, 35.0)
    cam = oracle.OracleCamera.from_viewer(view, camera.perspective(60.0, W / H), camera.IDENTITY, W, H)
    splats, counts = oracle.preprocess(sc, cam, bounds=oracle.TIGHT)
    return splats, counts, (W + 15) // 16, (H + 15) // 16


def _entries(oracle, splats, counts, tx, ty, order):
    """(tile, depth bits, index) in duplication order: splats in `order`, tiles row-major."""
    out = []
    for g in order:
        if counts[g] == 0:
            continue
        tiles, c, _ = oracle.coverage_tiles(splats[g], oracle.TIGHT, tx, ty)
        assert c == counts[g]
        d = int(np.float32(splats[g]["depth"]).view(np.uint32))
        out.extend((int(t), d, int(g)) for t in tiles)
    return out


def _ranges(keys, tiles):
    r = np.zeros((tiles, 2), np.uint32)
    for k, t in enumerate(keys):
        if k == 0 or keys[k - 1] != t:
            r[t, 0] = k
        r[t, 1] = k + 1
    return r


@pytest.mark.parametrize("capacity", [None, 2500])
def test_tile_lists_both_modes_vs_brute_force(oracle, capacity):
    splats, counts, tx, ty = _small_case(oracle)
    n = len(counts)
    depth_bits = splats["depth"].view(np.uint32)
    by_index = _entries(oracle, splats, counts, tx, ty, range(n))
    by_depth = _entries(oracle, splats, counts, tx, ty, sorted(range(n), key=lambda g: (int(depth_bits[g]), g)))
    assert len(by_index) == len(by_depth) == int(counts.sum())
    if capacity is not None:
        assert len(by_index) > capacity
    cap = len(by_index) if capacity is None else capacity
    expect = {
        0: sorted(by_index[:cap], key=lambda e: (e[0], e[1], e[2])),  # global
        1: sorted(by_depth[:cap], key=lambda e: e[0]),                # binning: stable by tile
    }
    got = {}
    for mode in (0, 1):
        hi, lo, vals, ranges, dup, mism = oracle.tile_lists(splats, counts, tx, ty, bounds=oracle.TIGHT,
                                                            capacity=capacity, mode=mode)
        assert dup == len(by_index) and mism == 0
        e = expect[mode]
        assert hi.tolist() == [x[0] for x in e]
        assert lo.tolist() == [x[1] for x in e]
        assert vals.tolist() == [x[2] for x in e]
        assert np.array_equal(ranges, _ranges(hi.tolist(), tx * ty))
        got[mode] = (hi, lo, vals)
    same = all(np.array_equal(a, b) for a, b in zip(got[0], got[1]))
    if capacity is None:
        assert same, "without overflow both modes build one list"
    else:
        assert not same, "with overflow the modes drop different entries"
        # binning keeps the nearest entries: every kept depth <= every dropped depth
        kept_max = max(x[1] for x in by_depth[:cap])
        dropped_min = min(x[1] for x in by_depth[cap:])
        assert kept_max <= dropped_min


@pytest.mark.parametrize("bounds", ["rect", "tight", "cutoff"])
def test_tile_lists_8px_bins_vs_brute_force(oracle, bounds):
    """tileSize 8: the preprocess counts are the 8-px coverage counts, and both sort modes' lists
    over the bin grid equal the brute force (bins of each splat row by row, sorted as above)."""
    from gsr import camera, scenes

    b = {"rect": oracle.RECT, "tight": oracle.TIGHT, "cutoff": oracle.CUTOFF}[bounds]
    W, H = 324, 180
    sc = scenes.make_scene(2500, seed=23, mean_sigma=1.5)
    cam = oracle.OracleCamera.from_viewer(camera.orbit_view(6.0, 10.0, 35.0), camera.perspective(60.0, W / H),
                                          camera.IDENTITY, W, H)
    splats, counts = oracle.preprocess(sc, cam, bounds=b, bin_size=8)
    tx, ty, gx, gy = (W + 15) // 16, (H + 15) // 16, (W + 7) // 8, (H + 7) // 8
    depth_bits = splats["depth"].view(np.uint32)
    entries = []
    for g in range(len(counts)):
        if splats["cull_reason"][g] != 0:
            assert counts[g] == 0
            continue
        bins, c, _ = oracle.coverage_tiles(splats[g], b, tx, ty, S=8, grid=(gx, gy))
        assert c == counts[g]
        entries.extend((int(t), int(depth_bits[g]), g) for t in bins)
    expect = sorted(entries, key=lambda e: (e[0], e[1], e[2]))
    for mode in (0, 1):
        hi, lo, vals, ranges, dup, mism = oracle.tile_lists(splats, counts, tx, ty, bounds=b, mode=mode, bin_size=8,
                                                            grid=(gx, gy))
        assert dup == len(entries) and mism == 0
        assert hi.tolist() == [e[0] for e in expect] and vals.tolist() == [e[2] for e in expect]
        assert lo.tolist() == [e[1] for e in expect]
        assert np.array_equal(ranges, _ranges(hi.tolist(), gx * gy))
