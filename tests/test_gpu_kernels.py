"""Per-kernel parity of the HIP kernels (through the rf_k_* C ABI) against the oracle.

Each kernel runs on seeded inputs at two geometries (stride > width, and stride == width so
the triangulation's column-w read wraps into the next row). Bars:
  * bit-exact against the FP32 restatement (same operation order, IEEE division/sqrt);
  * within 1e-5 normwise (max |diff| / max |ref|) of the FP64 oracle, except the
    triangulation, whose 3x3 solve amplifies FP32 rounding (bound 2e-3 here; its FP64
    variant is held to 1e-5).
"""
import numpy as np
import pytest

from tests.gpu_util import call, dev, dzeros, host, normwise, same_bits

pytestmark = pytest.mark.gpu

GEOMS = [(77, 45, 96), (64, 40, 64)]


@pytest.fixture(scope="module")
def lib():
    from reconflow_amd import _native
    return _native.lib()


def flat(rng, w, h, s, lo, hi):
    a = np.zeros((h + 1) * s, np.float32)
    a.reshape(h + 1, s)[:h, :w] = rng.uniform(lo, hi, (h, w)).astype(np.float32)
    return a


def valid(a, w, h, s):
    return np.asarray(a).reshape(-1, s)[:h, :w]


def check(w, h, s, gpu_outs, o32, o64, tol=1e-5):
    for g, a, b in zip(gpu_outs, o32, o64):
        gv, av, bv = valid(g, w, h, s), valid(a, w, h, s), valid(b, w, h, s)
        assert same_bits(gv, av), f"not bit-exact vs FP32 oracle: max diff {np.nanmax(np.abs(gv - av))}"
        assert normwise(gv, bv) <= tol, f"normwise error vs FP64 {normwise(gv, bv)} > {tol}"


@pytest.mark.parametrize("geom", GEOMS)
def test_data_l1(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(11)
    n = (h + 1) * s
    ins = [flat(rng, w, h, s, -1, 1), flat(rng, w, h, s, -1, 1)] + [flat(rng, w, h, s, -0.5, 0.5) for _ in range(4)] + \
          [flat(rng, w, h, s, -0.2, 0.2), flat(rng, w, h, s, -0.2, 0.2), flat(rng, w, h, s, -0.1, 0.1)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("data_l1", kind, *[x.astype(dt) for x in ins], w, h, s, 50.0, 0.3, *o)
        res[kind] = o
    d = [dev(x) for x in ins]
    g = [dzeros(n), dzeros(n)]
    assert call(lib, "solve_data_l1", *d, w, h, s, 50.0, 0.3, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


@pytest.mark.parametrize("geom", GEOMS)
def test_dual_tv(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(12)
    n = (h + 1) * s
    ins = [flat(rng, w, h, s, -2, 2), flat(rng, w, h, s, -2, 2)] + [flat(rng, w, h, s, -0.7, 0.7) for _ in range(4)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(4)]
        orc.call("dual_tv", kind, *[x.astype(dt) for x in ins], w, h, s, 0.25, 0.3, *o)
        res[kind] = o
    d = [dev(x) for x in ins]
    g = [dzeros(n) for _ in range(4)]
    assert call(lib, "solve_smooth_dual_tv_global", *d, w, h, s, 0.25, 0.3, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


@pytest.mark.parametrize("geom", GEOMS)
def test_recon_data_l1_and_prior_variants(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(13)
    n = (h + 1) * s
    ins = [flat(rng, w, h, s, -3, 3) for _ in range(8)] + [flat(rng, w, h, s, -0.5, 0.5) for _ in range(4)] + \
          [flat(rng, w, h, s, -0.2, 0.2) for _ in range(3)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("recon_data_l1", kind, *[x.astype(dt) for x in ins], w, h, s, 50.0, 33.3, 60.0, *o)
        res[kind] = o
    d = [dev(x) for x in ins]
    g = [dzeros(n), dzeros(n)]
    assert call(lib, "solve_recon_data_l1", *d, w, h, s, 50.0, 33.3, 60.0, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    # prior-flow forms (dense: no mask; sparse: mask)
    ufn, vfn = flat(rng, w, h, s, -4, 4), flat(rng, w, h, s, -4, 4)
    mask = (flat(rng, w, h, s, 0, 1) > 0.5).astype(np.float32)
    for m in (None, mask):
        res = {}
        for kind, dt in (("f32", np.float32), ("f64", np.float64)):
            o = [np.zeros(n, dt), np.zeros(n, dt)]
            a = [x.astype(dt) for x in ins]
            orc.call("recon_data_l1_fn", kind, *a[:6], ufn.astype(dt), vfn.astype(dt), *a[6:], w, h, s, 50.0, 33.3,
                     60.0, 100.0, None if m is None else m.astype(dt), *o)
            res[kind] = o
        g = [dzeros(n), dzeros(n)]
        assert call(lib, "solve_recon_data_l1_fn", *d[:6], dev(ufn), dev(vfn), *d[6:], w, h, s, 50.0, 33.3, 60.0,
                    100.0, None if m is None else dev(m), *g) == 0
        check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


@pytest.mark.parametrize("geom", GEOMS)
def test_pointwise_uvproj_sk_uc_cleanup(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(14)
    n = (h + 1) * s
    ins = [flat(rng, w, h, s, -3, 3) for _ in range(8)]
    for name, scal, nout in (("uvproj", (0.1, 60.0), 2), ("update_sk", (), 2)):
        res = {}
        for kind, dt in (("f32", np.float32), ("f64", np.float64)):
            o = [np.zeros(n, dt) for _ in range(nout)]
            orc.call(name, kind, *[x.astype(dt) for x in ins], w, h, s, *scal, *o)
            res[kind] = o
        g = [dzeros(n) for _ in range(nout)]
        cname = "solve_uvproj" if name == "uvproj" else "update_sk"
        assert call(lib, cname, *[dev(x) for x in ins], w, h, s, *scal, *g) == 0
        check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    # reprojection
    R = np.array([[0.9995, -0.0300, 0.0100], [0.0301, 0.9995, -0.0050], [-0.0098, 0.0053, 0.9999]])
    t = np.array([0.08, -0.05, 0.03])
    K = np.array([500.0, 0, 320, 0, 500, 240, 0, 0, 1])
    XYZ = [flat(rng, w, h, s, -2, 2), flat(rng, w, h, s, -2, 2), flat(rng, w, h, s, 2, 10)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("solve_uc", kind, *[x.astype(dt) for x in XYZ], R.ravel().copy(), t.copy(), K.copy(), w, h, s, *o)
        res[kind] = o
    g = [dzeros(n), dzeros(n)]
    assert call(lib, "solve_uc", *[dev(x) for x in XYZ], R.ravel(), t, K, w, h, s, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    # cleanup with non-finite entries
    six = [flat(rng, w, h, s, -10, 10) for _ in range(6)]
    six[0][5] = np.inf
    six[4][s + 3] = np.nan
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(3)]
        orc.call("cleanup3d", kind, *[x.astype(dt) for x in six], w, h, s, *o)
        res[kind] = o
    g = [dzeros(n) for _ in range(3)]
    assert call(lib, "cleanup_3d", *[dev(x) for x in six], w, h, s, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


def _pq(K, R, t):
    P = K @ np.hstack([np.eye(3), np.zeros((3, 1))])
    Q = K @ np.hstack([R, t[:, None]])
    return P.ravel().copy(), Q.ravel().copy()


@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("coarse", [False, True])
def test_solve3d(lib, orc, geom, coarse):
    w, h, s = geom
    rng = np.random.default_rng(15)
    n = (h + 1) * s
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    if coarse:  # the reference's coarse-level intrinsics: K[2][2] = 0
        K = np.array([[125.0, 0, 80], [0, 125, 60], [0, 0, 0]])
    R = np.array([[0.9995, -0.0300, 0.0100], [0.0301, 0.9995, -0.0050], [-0.0098, 0.0053, 0.9999]])
    P, Q = _pq(K, R, np.array([0.1, 0.0, 0.01]))
    up, vp = flat(rng, w, h, s, -20, 5), flat(rng, w, h, s, -3, 3)
    XYZ = [flat(rng, w, h, s, -1, 1), flat(rng, w, h, s, -1, 1), flat(rng, w, h, s, 4, 6)]
    for a in XYZ:  # stale values beyond the valid region are read at column w / row h
        a.reshape(h + 1, s)[:h, w:] = rng.uniform(-1, 1, (h, s - w))
    ins = [up, vp] + XYZ
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(3)]
        orc.call("solve3d", kind, *[x.astype(dt) for x in ins], P, Q, 0.1, 100.0, w, h, s, *o)
        res[kind] = o
    g = [dzeros(n) for _ in range(3)]
    assert call(lib, "solve_3d_minimal_area", *[dev(x) for x in ins], P, Q, 0.1, 100.0, w, h, s, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"], tol=2e-3)


@pytest.mark.parametrize("geom", GEOMS)
def test_medians(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(16)
    n = (h + 1) * s
    # integer-valued fields force ties; negative values make the corner zeros matter
    ins = [np.round(flat(rng, w, h, s, -3, 3)) for _ in range(2)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("median5", kind, *[x.astype(dt) for x in ins], w, h, s, *o)
        res[kind] = o
    g = [dzeros(n), dzeros(n)]
    assert call(lib, "median_filter", *[dev(x) for x in ins], w, h, s, *g, 5) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"], tol=0.0)
    ins3 = [flat(rng, w, h, s, -1, 1) for _ in range(3)]
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(3)]
        orc.call("median3d", kind, *[x.astype(dt) for x in ins3], w, h, s, *o)
        res[kind] = o
    g = [dzeros(n) for _ in range(3)]
    assert call(lib, "median_filter_3d", *[dev(x) for x in ins3], w, h, s, *g, 3) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"], tol=0.0)


@pytest.mark.parametrize("geom", GEOMS)
def test_warp_derivatives_add(lib, orc, geom):
    w, h, s = geom
    rng = np.random.default_rng(17)
    n = (h + 1) * s
    img = flat(rng, w, h, s, 0, 1)
    u, v = flat(rng, w, h, s, -6, 6), flat(rng, w, h, s, -6, 6)
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt)]
        orc.call("warp", kind, img.astype(dt), w, h, s, u.astype(dt), v.astype(dt), *o)
        res[kind] = o
    g = [dzeros(n)]
    assert call(lib, "warp_image", dev(img), w, h, s, dev(u), dev(v), *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    i1 = flat(rng, w, h, s, 0, 1)
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(3)]
        orc.call("derivatives", kind, img.astype(dt), i1.astype(dt), w, h, s, *o)
        res[kind] = o
    g = [dzeros(n) for _ in range(3)]
    assert call(lib, "compute_derivatives", dev(img), dev(i1), w, h, s, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt)]
        orc.call("add", kind, img.astype(dt), i1.astype(dt), h * s, *o)
        res[kind] = o
    g = [dzeros(n)]
    assert call(lib, "add", dev(img), dev(i1), h * s, *g) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


@pytest.mark.parametrize("fine,coarse", [((64, 48, 64), (32, 24, 32)), ((45, 31, 64), (22, 15, 32)),
                                         ((77, 45, 96), (38, 22, 64))])
def test_pyramid_resampling(lib, orc, fine, coarse):
    (W, H, S), (w, h, s) = fine, coarse
    rng = np.random.default_rng(18)
    img = flat(rng, W, H, S, 0, 1)
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros((h + 1) * s, dt)]
        orc.call("downscale", kind, img.astype(dt), W, H, S, w, h, s, *o)
        o2 = [np.zeros((h + 1) * s, dt)]
        orc.call("downscale_scaled", kind, img.astype(dt), W, H, S, w, h, s, 0.5, *o2)
        res[kind] = o + o2
    g = [dzeros((h + 1) * s), dzeros((h + 1) * s)]
    assert call(lib, "downscale", dev(img), W, H, S, w, h, s, g[0]) == 0
    assert call(lib, "downscale_scaled", dev(img), W, H, S, w, h, s, 0.5, g[1]) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    cimg = flat(rng, w, h, s, -3, 3)
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros((H + 1) * S, dt)]
        orc.call("upscale", kind, cimg.astype(dt), w, h, s, W, H, S, 2.0, *o)
        res[kind] = o
    g = [dzeros((H + 1) * S)]
    assert call(lib, "upscale", dev(cimg), w, h, s, W, H, S, 2.0, g[0]) == 0
    check(W, H, S, [host(x) for x in g], res["f32"], res["f64"])


def test_gray_conversions(lib, orc):
    w, h, s = 77, 45, 96
    rng = np.random.default_rng(19)
    rgb = rng.integers(0, 256, (h + 1) * s * 3, dtype=np.uint8)
    res = {}
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros((h + 1) * s, dt)]
        orc.call("gray_rgb", kind, rgb, o[0], w, h, s)
        res[kind] = o
    g = [dzeros((h + 1) * s)]
    assert call(lib, "rgb_to_gray", dev(rgb), g[0], w, h, s) == 0
    check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])
    u8 = rng.integers(0, 256, (h + 1) * s, dtype=np.uint8)
    u16 = rng.integers(0, 65536, (h + 1) * s, dtype=np.uint16)
    for name, arr, cname in (("gray_u8", u8, "cv8u_to_gray"), ("gray_u16", u16, "cv16u_to_gray")):
        res = {}
        for kind, dt in (("f32", np.float32), ("f64", np.float64)):
            o = [np.zeros((h + 1) * s, dt)]
            orc.call(name, kind, arr, o[0], w, h, s)
            res[kind] = o
        g = [dzeros((h + 1) * s)]
        assert call(lib, cname, dev(arr), g[0], w, h, s) == 0
        check(w, h, s, [host(x) for x in g], res["f32"], res["f64"])


def test_uniform_division_equals_ieee(lib):
    """rf_k_uniform_divide (the fused joint kernels' division by a launch-uniform divisor) against
    numpy's IEEE FP32 division, bit for bit: every subnormal and every significand at three
    exponents, both signs, plus infinities and NaN, for the workloads' divisors, divisors with
    exact subnormal ties (98, 6), a power of two, an odd integer and the extremes."""
    sub = np.arange(1 << 23, dtype=np.uint32)
    sets = [sub] + [sub | np.uint32((e + 127) << 23) for e in (-126, 0, 127)]
    special = np.array([np.inf, -np.inf, np.nan, 0.0, -0.0], np.float32).view(np.uint32)
    nums = np.concatenate([np.concatenate([s, s | np.uint32(0x80000000)]) for s in sets] + [special]).view(np.float32)
    g_n = dev(nums)
    g_q = dzeros(nums.size)
    for d in (np.float32(33.3) + np.float32(60.0), np.float32(0.1) + np.float32(60.0), 98.0, 6.0, 1.0, 3.0, 1e-40,
              3.4028235e38):
        d32 = np.float32(d)
        assert call(lib, "uniform_divide", g_n, nums.size, float(d32), g_q) == 0
        got = host(g_q)
        with np.errstate(all="ignore"):
            ref = nums / d32
        both_nan = np.isnan(ref) & np.isnan(got)
        bad = (ref.view(np.uint32) != got.view(np.uint32)) & ~both_nan
        assert not bad.any(), f"d={d32}: {int(bad.sum())} quotients differ, first n={nums[bad][:4]}"
