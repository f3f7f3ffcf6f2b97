"""C-ABI checks that need no GPU: the library loads, exports every declared entry point, and
its argument validation returns status codes (never exits) before any device work."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def lib():
    from reconflow_amd import _native
    if not _native.lib_path().exists():
        from reconflow_amd.build import build_hip
        build_hip()
    return _native.lib()


def header_symbols():
    text = (ROOT / "include" / "rfhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = header_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    from reconflow_amd import _native
    assert set(header_symbols()) <= set(_native.declared_symbols()) | {"rf_params_default"}


def test_validation_errors_without_device(lib):
    from reconflow_amd._native import RFParams
    ctx = C.c_void_p()
    assert lib.rf_create(C.byref(ctx), 32, 12, 32) == 0
    p = RFParams()
    lib.rf_params_default(C.byref(p), 0, 240)
    assert lib.rf_initialize(ctx, C.byref(p)) == -1  # width 0
    assert b"invalid" in lib.rf_last_error()
    lib.rf_params_default(C.byref(p), 64, 48)
    p.n_levels = 9  # levels run out of pixels: 64 -> 32 -> ... -> 0
    assert lib.rf_initialize(ctx, C.byref(p)) == -1
    assert b"no pixels" in lib.rf_last_error()
    p.n_levels, p.scale = 3, 1.0
    assert lib.rf_initialize(ctx, C.byref(p)) == -1
    p.scale, p.input_type = 2.0, 9
    assert lib.rf_initialize(ctx, C.byref(p)) == -1
    p.input_type, p.path = 0, 7
    assert lib.rf_initialize(ctx, C.byref(p)) == -1
    # sides of 32768 or more: the resampling's 32-bit texel-coordinate products would overflow
    for w, h in ((32768, 8), (64, 32768)):
        lib.rf_params_default(C.byref(p), w, h)
        assert lib.rf_initialize(ctx, C.byref(p)) == -1
        assert b"below 32768" in lib.rf_last_error()
    # below 32768, a scale under 2 leaves a coarse level wide enough for the products to overflow
    lib.rf_params_default(C.byref(p), 30000, 8)
    p.n_levels, p.scale = 2, 1.5
    assert lib.rf_initialize(ctx, C.byref(p)) == -1
    assert b"32-bit texel coordinates" in lib.rf_last_error()
    b = C.c_void_p()
    W, H = (C.c_int * 1)(30000), (C.c_int * 1)(8)
    assert lib.rf_batch_create(C.byref(b), C.byref(p), 1, W, H) == -1
    assert b"32-bit texel coordinates" in lib.rf_last_error()
    lib.rf_params_default(C.byref(p), 64, 48)
    W, H = (C.c_int * 2)(64, 40000), (C.c_int * 2)(48, 16)
    assert lib.rf_batch_create(C.byref(b), C.byref(p), 2, W, H) == -1
    assert b"below 32768" in lib.rf_last_error()
    # the allocation-free solve validates before it looks up its scratch or launches anything
    lib.rf_params_default(C.byref(p), 40000, 8)
    assert lib.rf_solve_flow_noalloc(None, None, None, None, None, None) == -1
    # calls before a successful initialize report a state error
    assert lib.rf_get_stride(ctx) == -2
    assert lib.rf_solve_flow(ctx, 5.0) == -2
    R = (C.c_double * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = (C.c_double * 3)(0.1, 0, 0)
    assert lib.rf_solve_recon(ctx, None, None, R, t, 5.0) == -2
    assert lib.rf_copy_images(ctx, None, None) == -2
    assert lib.rf_destroy(ctx) == 0
    # null context
    assert lib.rf_solve_flow(None, 5.0) == -1


def test_pyramid_level_count_matches_reference_demo(lib):
    # main.cpp:50-56: levels while (int)(h / scale) > minWidth
    assert lib.rf_compute_pyramid_levels(768, 320, 400, 2.0) == 1
    assert lib.rf_compute_pyramid_levels(640, 480, 20, 2.0) == 5
    assert lib.rf_compute_pyramid_levels(1280, 720, 40, 2.0) == 5


def test_kernel_entry_points_reject_bad_geometry(lib):
    # stride smaller than width is rejected before any launch
    assert lib.rf_k_add(None, None, -1, None, None) == -1
    # the uniform division needs a finite positive divisor
    for d in (0.0, -2.0, float("inf"), float("nan")):
        assert lib.rf_k_uniform_divide(None, 4, d, None, None) == -1
    assert lib.rf_k_warp_image(None, 64, 48, 32, None, None, None, None) == -1
    assert lib.rf_k_median_filter(None, None, 8, 8, 8, None, None, 3, None) == -5


def test_synthetic_pairs_are_deterministic(synth_lib):
    import reconflow_amd as rf
    a = rf.make_flow_pair(96, 64, 7)
    b = rf.make_flow_pair(96, 64, 7)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    i0, i1, gu, gv = a
    assert i0.shape == (64, 96) and 0.0 <= i0.min() and i0.max() <= 1.0
    assert np.hypot(gu, gv).max() <= 4.0 + 1e-5
    d = rf.make_recon_pair(160, 120, 3)
    assert d["i0"].shape == (120, 160)
    np.testing.assert_allclose(d["R"] @ d["R"].T, np.eye(3), atol=1e-12)
    assert abs(np.linalg.norm(d["t"]) - 0.1) < 1e-12
    z = d["gt_Z"][:, :160]
    assert z.min() >= 2.0 - 1e-6 and z.max() <= 10.0 + 1e-6


def test_synthetic_sequence_motion_and_poses(synth_lib):
    # the streaming workload's trajectory: deterministic frames, per-frame rotation <= 0.5 deg,
    # camera displacement <= 0.02, proper rotations; chaining the relative poses from frame 0 (the
    # reference pose) gives frame k's absolute rotation, 4.5 deg * |sin(2 pi k / 60)| about one
    # fixed axis
    import reconflow_amd as rf
    n = 40
    a = rf.make_sequence(96, 64, n, 5)
    b = rf.make_sequence(96, 64, n, 5)
    assert np.array_equal(a["frames"], b["frames"])
    assert a["frames"].shape == (n, 64, 96)
    R, t = a["R"], a["t"]
    Rc = np.eye(3)
    axis0 = None
    for k, (Rk, tk) in enumerate(zip(R, t), start=1):
        np.testing.assert_allclose(Rk @ Rk.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Rk) - 1.0) < 1e-12
        ang = np.degrees(np.arccos(np.clip((np.trace(Rk) - 1) / 2, -1, 1)))
        assert ang <= 0.5 + 1e-9
        assert np.linalg.norm(Rk.T @ tk) <= 0.02 + 1e-9  # camera displacement between frames
        Rc = Rk @ Rc
        want = 4.5 * abs(np.sin(2 * np.pi * k / 60.0))
        ax = np.array([Rc[2, 1] - Rc[1, 2], Rc[0, 2] - Rc[2, 0], Rc[1, 0] - Rc[0, 1]])  # 2 sin(angle) axis
        got = np.degrees(np.arctan2(np.linalg.norm(ax) / 2, (np.trace(Rc) - 1) / 2))
        assert abs(got - want) < 1e-9, (k, got, want)
        if want > 0.01:
            ax /= np.linalg.norm(ax)
            axis0 = ax if axis0 is None else axis0
            assert abs(abs(ax @ axis0) - 1.0) < 1e-9, k
    # consecutive frames differ (the camera moves) but stay the same scene
    assert 0.0 < np.abs(a["frames"][1] - a["frames"][0]).mean() < 0.3
