"""Host I/O formats (CPU): Middlebury .flo round trip and reader checks, PLY mask."""
import ctypes as C
import struct
from pathlib import Path

import numpy as np
import pytest

REF_FLO = Path("/root/reference/test_reconflow/data/flo319.flo")


@pytest.fixture(scope="module")
def lib():
    from reconflow_amd import _native
    return _native.lib()


def _read(lib, path, stride_pad=0):
    w, h = C.c_int(), C.c_int()
    assert lib.rf_read_flo(str(path).encode(), C.byref(w), C.byref(h), None, None, 0) == 0
    s = w.value + stride_pad
    u = np.zeros((h.value, s), np.float32)
    v = np.zeros_like(u)
    assert lib.rf_read_flo(str(path).encode(), C.byref(w), C.byref(h), u.ctypes.data, v.ctypes.data, s) == 0
    return u, v, w.value, h.value


def test_flo_round_trip_and_layout(lib, tmp_path):
    rng = np.random.default_rng(0)
    w, h, s = 37, 11, 64
    u = rng.standard_normal((h, s)).astype(np.float32)
    v = rng.standard_normal((h, s)).astype(np.float32)
    p = tmp_path / "x.flo"
    assert lib.rf_write_flo(str(p).encode(), u.ctypes.data, v.ctypes.data, w, h, s) == 0
    raw = p.read_bytes()
    # layout fixed by the Middlebury format: magic, w, h, then (u, v) interleaved row-major
    assert struct.unpack("<fii", raw[:12]) == (202021.25, w, h)
    body = np.frombuffer(raw[12:], np.float32).reshape(h, w, 2)
    np.testing.assert_array_equal(body[..., 0], u[:, :w])
    np.testing.assert_array_equal(body[..., 1], v[:, :w])
    ru, rv, rw, rh = _read(lib, p, stride_pad=s - w)
    assert (rw, rh) == (w, h)
    np.testing.assert_array_equal(ru[:, :w], u[:, :w])
    assert not ru[:, w:].any() and not rv[:, w:].any()  # zero padding to the stride


def test_flo_rejects_bad_magic(lib, tmp_path):
    p = tmp_path / "bad.flo"
    p.write_bytes(struct.pack("<fii", 1.0, 4, 4) + b"\0" * 128)
    w, h = C.c_int(), C.c_int()
    assert lib.rf_read_flo(str(p).encode(), C.byref(w), C.byref(h), None, None, 0) == -1


@pytest.mark.skipif(not REF_FLO.exists(), reason="reference data not mounted")
def test_reads_the_reference_prior_flow_file(lib):
    u, v, w, h = _read(lib, REF_FLO)
    assert (w, h) == (768, 320)
    # value ranges of data/flo319.flo as read independently with numpy
    raw = np.fromfile(REF_FLO, np.float32, offset=12).reshape(h, w, 2)
    np.testing.assert_array_equal(u, raw[..., 0])
    assert -70.2 < u.min() < -70.0 and -4.1 < u.max() < -3.9
    assert -21.1 < v.min() < -20.9 and 25.7 < v.max() < 25.9


def test_ply_mask_matches_demo_rule(lib, tmp_path):
    w, h, s = 4, 2, 8
    X = np.zeros((h, s), np.float32)
    Y = np.zeros_like(X)
    Z = np.zeros_like(X)
    pts = [(0, 0, 20), (80, 0, 20), (-80, 0, 20), (0, -75, 20), (0, 75, 20), (0, 0, 10), (0, 0, 75), (0, 0, 76)]
    keep = [True, False, True, False, True, False, True, False]  # X has no lower bound in the demo
    for i, (x, y, z) in enumerate(pts):
        X.flat[(i // w) * s + i % w], Y.flat[(i // w) * s + i % w], Z.flat[(i // w) * s + i % w] = x, y, z
    p = tmp_path / "c.ply"
    assert lib.rf_write_ply(str(p).encode(), X.ctypes.data, Y.ctypes.data, Z.ctypes.data, None, w, h, s, 10.0, 75.0,
                            75.0) == 0
    lines = p.read_text().splitlines()
    n = int([ln for ln in lines if ln.startswith("element vertex")][0].split()[-1])
    assert n == sum(keep)
    body = [list(map(float, ln.split())) for ln in lines[lines.index("end_header") + 1:]]
    expect = [[x, -y, -z] for (x, y, z), k in zip(pts, keep) if k]
    np.testing.assert_allclose(body, expect)
