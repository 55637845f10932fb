"""The PLY loader (gsr_ply_load, the reference's loadGaussianPly) on files written here, as the
reference's test_ply_loader.cpp does with its PlyBuilder:

- INRIA `save_ply` files of SH degree 0..3 round-trip: raw values bit for bit, positions and
  SH exact (transposed from the file's channel-major f_rest), quaternions normalised exactly as
  q / sqrt(q.q) in float32, alpha and sigma within 1 ulp of 1 / (1 + expf) and expf (expf taken
  as the rounded double exp);
- mixed scalar types, unknown properties (skipped, counted in the stride), an element before
  `vertex`, integer quaternions, a zero quaternion (identity, counted), NaN in a skipped
  property (not counted) and in a used one (counted, warned), comments;
- every rejected file fails with GSR_ERROR_FORMAT and a message naming the problem
  (test_ply_loader.cpp:306-390).
"""
import math
import struct

import numpy as np
import pytest

INRIA_NAMES = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]


def _inria_header(n, degree, fmt="binary_little_endian", extra_lines=""):
    rest = 3 * ((degree + 1) ** 2 - 1)
    names = INRIA_NAMES + [f"f_rest_{i}" for i in range(rest)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                                   "rot_0", "rot_1", "rot_2", "rot_3"]
    lines = ["ply", f"format {fmt} 1.0", "comment synthetic", extra_lines.rstrip("\n")] if extra_lines else \
        ["ply", f"format {fmt} 1.0", "comment synthetic"]
    lines += [f"element vertex {n}"] + [f"property float {p}" for p in names] + ["end_header"]
    return ("\n".join(lines) + "\n").encode(), names


def _raw_scene(n, degree, seed):
    rng = np.random.default_rng(seed)
    K = (degree + 1) ** 2
    raw = {
        "pos": rng.normal(0, 2, (n, 3)).astype(np.float32),
        "dc": rng.normal(0, 0.5, (n, 3)).astype(np.float32),
        "rest": rng.normal(0, 0.1, (n, 3, K - 1)).astype(np.float32),  # [c][k-1]: the file's channel-major order
        "opacity": rng.normal(0, 1.5, n).astype(np.float32),
        "log_scale": rng.normal(-4.5, 0.8, (n, 3)).astype(np.float32),
        "rot": rng.normal(0, 2, (n, 4)).astype(np.float32),
    }
    return raw


def _write_inria(raw, degree):
    n = raw["pos"].shape[0]
    header, names = _inria_header(n, degree)
    cols = [raw["pos"], np.zeros((n, 3), np.float32), raw["dc"], raw["rest"].reshape(n, -1),
            raw["opacity"][:, None], raw["log_scale"], raw["rot"]]
    body = np.concatenate(cols, axis=1).astype("<f4")
    assert body.shape[1] == len(names)
    return header + body.tobytes()


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_inria_round_trip(degree, tmp_path):
    import gsr

    n = 3000
    raw = _raw_scene(n, degree, seed=degree)
    raw["rot"][5] = 0.0  # a zero quaternion -> identity
    path = tmp_path / f"scene_sh{degree}.ply"
    path.write_bytes(_write_inria(raw, degree))
    s = gsr.load_ply(str(path))
    assert s.count == n and s.sh_degree == degree
    assert s.comments == ["synthetic"]
    # raw values, positions, SH: copies
    assert np.array_equal(s.raw_opacity, raw["opacity"])
    assert np.array_equal(s.raw_log_scales, raw["log_scale"])
    assert np.array_equal(s.raw_rotations, raw["rot"])
    assert np.array_equal(s.pos_alpha[:, :3], raw["pos"])
    K = (degree + 1) ** 2
    sh = s.sh.reshape(n, K, 3)
    assert np.array_equal(sh[:, 0, :], raw["dc"])
    assert np.array_equal(sh[:, 1:, :], raw["rest"].transpose(0, 2, 1))
    # activations
    # float32 expf (within one ulp of the rounded double exp), then 1 / (1 + e) in float32
    one = np.float32(1.0)
    e = np.exp(-raw["opacity"].astype(np.float64)).astype(np.float32)
    alpha = one / (one + e)
    assert _ulps(s.pos_alpha[:, 3], alpha).max() <= 1
    assert np.allclose(s.pos_alpha[:, 3], 1.0 / (1.0 + np.exp(-raw["opacity"].astype(np.float64))), rtol=1e-6)
    assert _ulps(s.scales[:, :3], np.exp(raw["log_scale"].astype(np.float64)).astype(np.float32)).max() <= 1
    assert np.all(s.scales[:, 3] == 0)
    q = raw["rot"]
    n2 = ((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2]) + q[:, 3] * q[:, 3]
    with np.errstate(invalid="ignore", divide="ignore"):
        qn = q / np.sqrt(n2)[:, None]
    ok = np.arange(n) != 5
    assert np.array_equal(s.rotations[ok], qn[ok])
    assert np.array_equal(s.rotations[5], np.array([1, 0, 0, 0], np.float32))
    assert s.zero_quaternions == 1 and s.non_finite_values == 0
    assert any("zero-length quaternions" in w for w in s.warnings)
    # the same bytes from memory
    m = gsr.load_ply_bytes(path.read_bytes(), "mem.ply")
    assert np.array_equal(m.sh, s.sh) and np.array_equal(m.rotations, s.rotations)


def test_mixed_types_unknown_properties_and_elements(tmp_path):
    import gsr

    header = (
        "ply\nformat binary_little_endian 1.0\ncomment written by test_ply_loader\n"
        "element extra_info 2\nproperty int a\nproperty uchar b\n"
        "element vertex 3\n"
        "property double x\nproperty double y\nproperty double z\n"
        "property float nxx\nproperty uchar flags\nproperty ushort id\n"
        "property float f_dc_0\nproperty float f_dc_1\nproperty float f_dc_2\n"
        "property float opacity\n"
        "property float scale_0\nproperty float scale_1\nproperty float scale_2\n"
        "property int rot_0\nproperty float rot_1\nproperty float rot_2\nproperty float rot_3\n"
        "property float filter_3D\n"
        "end_header\n").encode()
    body = b""
    for _ in range(2):
        body += struct.pack("<iB", -7, 0xAB)
    for i in range(3):
        body += struct.pack("<ddd", 1.5 + i, -2.25, 1e-3 * i)
        body += struct.pack("<fBH", 99.0, 0xFF, 1000 + i)
        body += struct.pack("<fff", 0.1, 0.2, 0.3)
        body += struct.pack("<f", 0.0)
        body += struct.pack("<fff", 0.0, math.log(2.0), -1.0)
        body += struct.pack("<ifff", 0 if i == 2 else 2, 0.0, 0.0, 0.0)
        body += struct.pack("<f", float("nan") if i == 1 else 0.5)
    path = tmp_path / "mixed.ply"
    path.write_bytes(header + body)
    s = gsr.load_ply(str(path))
    assert s.count == 3 and s.sh_degree == 0
    assert s.pos_alpha[1, 0] == np.float32(2.5) and s.pos_alpha[1, 1] == np.float32(-2.25)
    assert s.pos_alpha[2, 2] == np.float32(np.float64(2e-3))  # double -> float
    assert np.array_equal(s.sh[2], np.array([0.1, 0.2, 0.3], np.float32))
    assert s.pos_alpha[0, 3] == np.float32(0.5)
    assert abs(s.scales[0, 1] - 2.0) < 1e-6 and abs(s.scales[0, 2] - math.exp(-1.0)) < 1e-6
    assert s.raw_rotations[0, 0] == 2.0 and np.array_equal(s.rotations[0], [1, 0, 0, 0])
    assert s.zero_quaternions == 1 and np.array_equal(s.rotations[2], [1, 0, 0, 0])
    assert s.non_finite_values == 0, "NaN in a skipped property must not count"
    assert s.comments == ["written by test_ply_loader"]
    assert any("skipped element 'extra_info'" in w for w in s.warnings)
    assert any("converted to float" in w for w in s.warnings)

    # Inf in a used property is counted and reported.
    raw = _raw_scene(2, 0, seed=1)
    raw["log_scale"][1, 0] = np.inf
    p2 = tmp_path / "inf.ply"
    p2.write_bytes(_write_inria(raw, 0))
    s2 = gsr.load_ply(str(p2))
    assert s2.non_finite_values == 1 and any("NaN/Inf" in w for w in s2.warnings)


def test_crlf_header_and_trailing_bytes(tmp_path):
    import gsr

    raw = _raw_scene(4, 1, seed=2)
    data = _write_inria(raw, 1)
    end = data.index(b"end_header\n") + len(b"end_header\n")
    crlf = data[:end].replace(b"\n", b"\r\n") + data[end:]
    p = tmp_path / "crlf.ply"
    p.write_bytes(crlf + b"xyz")
    s = gsr.load_ply(str(p))
    assert s.count == 4 and s.sh_degree == 1
    assert np.array_equal(s.pos_alpha[:, :3], raw["pos"])
    assert any("3 unexpected trailing bytes" in w for w in s.warnings)


def _deg0(n=1, fmt="binary_little_endian", drop=None, extra_prop="", vertex_list=False, pre=""):
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
             "rot_0", "rot_1", "rot_2", "rot_3"]
    names = [p for p in names if p != drop]
    h = f"ply\nformat {fmt} 1.0\n{pre}element vertex {n}\n" + "".join(f"property float {p}\n" for p in names)
    h += extra_prop
    if vertex_list:
        h += "property list uchar int vertex_indices\n"
    h += "end_header\n"
    return h.encode() + np.full((n, len(names)), 0.5, "<f4").tobytes()


ERROR_CASES = [
    ("ascii", lambda: _deg0(fmt="ascii"), "ASCII"),
    ("big endian", lambda: _deg0(fmt="binary_big_endian"), "big_endian"),
    ("compressed", lambda: _deg0(pre="element chunk 1\nproperty float min_x\n"), "compressed"),
    ("packed property", lambda: _deg0(extra_prop="property uint packed_position\n"), "compressed"),
    ("f_rest 10", lambda: _deg0(extra_prop="".join(f"property float f_rest_{i}\n" for i in range(10))), "f_rest"),
    ("f_rest gap", lambda: _deg0(extra_prop="".join(f"property float f_rest_{i}\n" for i in range(10) if i != 4)),
     "contiguous"),
    ("missing opacity", lambda: _deg0(drop="opacity"), "'opacity'"),
    ("bad magic", lambda: b"plx\n" + _deg0()[4:], "not a PLY"),
    ("list in vertex", lambda: _deg0(vertex_list=True), "list"),
    ("zero gaussians", lambda: _deg0(n=0), "empty"),
    ("unknown type", lambda: _deg0(extra_prop="property float128 q\n"), "unknown property type"),
    ("no vertex", lambda: b"ply\nformat binary_little_endian 1.0\nelement face 0\nproperty int a\nend_header\n",
     "no 'vertex'"),
    ("no end_header", lambda: b"ply\nformat binary_little_endian 1.0\nelement vertex 1\n", "end_header"),
    ("empty file", lambda: b"", "not a PLY"),
    ("one record short", lambda: _deg0(n=3)[:-56], "2 complete records"),
    ("one byte short", lambda: _deg0(n=3)[:-1], "2 complete records"),
]


@pytest.mark.parametrize("name,make,needle", ERROR_CASES, ids=[c[0] for c in ERROR_CASES])
def test_rejected_files(name, make, needle, tmp_path):
    import gsr

    p = tmp_path / "bad.ply"
    p.write_bytes(make())
    with pytest.raises(gsr.PlyError) as e:
        gsr.load_ply(str(p))
    assert needle in str(e.value), str(e.value)
    assert "bad.ply" in str(e.value)


def test_missing_file_is_an_io_error(tmp_path):
    import gsr

    with pytest.raises(gsr.GsrError) as e:
        gsr.load_ply(str(tmp_path / "nope.ply"))
    assert not isinstance(e.value, gsr.PlyError) and "cannot open" in str(e.value)
