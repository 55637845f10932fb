"""CPU checks of the C ABI library: it loads, exports every symbol include/gsr.h declares,
and its host-only entry points (camera set-up, fp16 encoder) agree bit for bit with the
oracle. No compute call needs a GPU here."""
import ctypes
import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    import gsr

    header = open(os.path.join(ROOT, "include", "gsr.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(gsr_[a-z_0-9]+)\(", header, re.M))
    assert len(declared) >= 20
    lib = ctypes.CDLL(gsr.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(gsr.EXPORTED), declared ^ set(gsr.EXPORTED)


def test_camera_make_matches_oracle_bit_exact(oracle):
    import gsr
    from gsr import camera

    for eye, W, H in [((0, 0, 8), 1280, 720), ((3.1, 2.2, -7.5), 1920, 1080), ((0.3, 0.5, 6.0), 640, 360)]:
        view = camera.look_at(eye, [0, 0, 0])
        proj = camera.perspective(60.0, W / H)
        c = gsr.camera_from_viewer(view, proj, camera.IDENTITY, W, H)
        o = oracle.OracleCamera.from_viewer(view, proj, camera.IDENTITY, W, H)
        assert np.array_equal(np.array(c.view_inria[:], np.float32).view(np.uint32), o.view.view(np.uint32))
        assert np.array_equal(np.array(c.position[:], np.float32).view(np.uint32), o.position.view(np.uint32))
        assert np.float32(c.tan_half_fov_x) == np.float32(o.tan_x) and np.float32(c.tan_half_fov_y) == np.float32(o.tan_y)
        assert np.array_equal(np.array([c.focal_x, c.focal_y], np.float32), o.focal)
        assert (c.tiles_x, c.tiles_y) == (o.tiles_x, o.tiles_y)
        assert np.allclose(np.array(c.position[:]), eye, atol=1e-5)
        assert math.isclose(c.tan_half_fov_y, math.tan(math.radians(30)), rel_tol=1e-6)


def test_encode_f16_matches_oracle_and_numpy(oracle):
    import gsr
    from gsr import scenes

    sc = scenes.make_scene(3000, seed=11)
    g, c = gsr.encode_f16(sc.pos_alpha, sc.scales, sc.rotations, sc.sh, sc.sh_degree)
    og, oc = oracle.encode_f16(sc)
    assert np.array_equal(g, og) and np.array_equal(c, oc)
    # numpy's float16 cast is IEEE round-to-nearest-even: an independent check of the encoder.
    assert np.array_equal(c[:, :48], sc.sh.astype(np.float16).view(np.uint16))
    assert np.array_equal((g[:, 3] & 0xFFFF).astype(np.uint16), sc.pos_alpha[:, 3].astype(np.float16).view(np.uint16))


def test_invalid_arguments_report_errors():
    import gsr

    try:
        gsr.camera_make(np.eye(4), 64, 64, 0.0, 1.0)
    except gsr.GsrError as e:
        assert "tan" in str(e)
    else:
        raise AssertionError("expected an error")
    bad = ctypes.c_void_p()
    st = gsr._lib.gsr_scene_create(None, None, ctypes.byref(bad))
    assert st == 1 and b"null" in gsr._lib_last_error()
