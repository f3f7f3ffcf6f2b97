"""Pins of the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, golden outputs or expected values for this path, so each
oracle function is checked here against closed forms, library routines (numpy padding,
order statistics, linear algebra) and invariants, chosen so that a plausible transcription
mistake (dropped term, wrong sign or index, transposed operand, wrong border rule) fails.
"""
import numpy as np
import pytest

W, H, S = 37, 23, 64  # odd sizes, stride > width


def field(kind, s=S, h=H, fill=0.0):
    dt = np.float64 if kind == "f64" else np.float32
    return np.full((h + 1) * s, fill, dt)


def put(a2d, kind, s=S):
    """Pack a (h, w) array into a flat stride-s buffer with a zero guard row."""
    h, w = a2d.shape
    f = field(kind, s, h)
    f.reshape(h + 1, s)[:h, :w] = a2d
    return f


def get(flat, w, h, s=S):
    return flat.reshape(-1, s)[:h, :w]


KINDS = ["f64", "f32"]


# ---------------------------------------------------------------- A2: gray conversion
def test_gray_rgb_matches_float32_closed_form_on_all_triples(orc):
    # every (r, g, b) triple in 2^24, laid out as one 4096 x 4096 image
    v = np.arange(1 << 24, dtype=np.uint32)
    rgb = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], -1).astype(np.uint8)
    out = np.zeros(1 << 24, np.float32)
    orc.call("gray_rgb", "f32", rgb, out, 4096, 4096, 4096)
    r, g, b = (rgb[:, i].astype(np.float32) for i in range(3))
    expect = ((r + g + b) / np.float32(3)) / np.float32(256)
    assert np.array_equal(out, expect)


# ---------------------------------------------------------------- A4: warp
@pytest.mark.parametrize("kind", KINDS)
def test_warp_zero_and_integer_flow_is_mirror_shift(orc, kind, rng):
    img = rng.random((H, W))
    src = put(img, kind)
    for du, dv in [(0, 0), (3, -2), (-5, 4), (W + 2, -H - 1)]:
        u, v = put(np.full((H, W), du), kind), put(np.full((H, W), dv), kind)
        out = field(kind)
        orc.call("warp", kind, src, W, H, S, u, v, out)
        # edge-duplicating reflection = numpy "symmetric" padding, periodic with 2n
        big = np.pad(img, ((2 * H, 2 * H), (2 * W, 2 * W)), mode="symmetric")
        yy, xx = np.mgrid[0:H, 0:W]
        expect = big[yy + dv + 2 * H, xx + du + 2 * W]
        np.testing.assert_array_equal(get(out, W, H), expect.astype(out.dtype))


def test_warp_reproduces_linear_images_in_the_interior(orc, rng):
    a, b, c = 0.3, -0.7, 2.0
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = a * xx + b * yy + c
    u = rng.uniform(-3, 3, (H, W))
    v = rng.uniform(-3, 3, (H, W))
    out = field("f64")
    orc.call("warp", "f64", put(img, "f64"), W, H, S, put(u, "f64"), put(v, "f64"), out)
    tx, ty = xx + u, yy + v
    inside = (tx >= 0) & (tx <= W - 1) & (ty >= 0) & (ty <= H - 1)
    np.testing.assert_allclose(get(out, W, H)[inside], (a * tx + b * ty + c)[inside], rtol=0, atol=1e-12)


# ---------------------------------------------------------------- A3 / A10: pyramid
def test_downscale_exact_halving_is_box_mean(orc, rng):
    Wf, Hf, Sf = 64, 46, 64
    w, h, s = 32, 23, 32
    img = rng.random((Hf, Wf))
    out = field("f64", s, h)
    orc.call("downscale", "f64", put(img, "f64", Sf), Wf, Hf, Sf, w, h, s, out)
    box = img.reshape(h, 2, w, 2).mean(axis=(1, 3))
    np.testing.assert_allclose(get(out, w, h, s), box, rtol=0, atol=1e-14)


@pytest.mark.parametrize("Wf,Hf,w,h", [(45, 31, 22, 15), (33, 17, 16, 8), (7, 5, 3, 2)])
def test_downscale_odd_sizes_preserve_linear_at_coarse_centres(orc, Wf, Hf, w, h):
    Sf, s = 64, 32
    yy, xx = np.mgrid[0:Hf, 0:Wf].astype(np.float64)
    img = 0.25 * xx - 0.5 * yy + 1.0
    out = field("f64", s, h)
    orc.call("downscale", "f64", put(img, "f64", Sf), Wf, Hf, Sf, w, h, s, out)
    cy, cx = np.mgrid[0:h, 0:w].astype(np.float64)
    # coarse centre in fine texels: (i + 1/2) * W/w - 1/2; taps straddle it symmetrically
    fx, fy = (cx + 0.5) * Wf / w - 0.5, (cy + 0.5) * Hf / h - 0.5
    rx, ry = 0.25 * Wf / w, 0.25 * Hf / h
    inside = (fx - rx >= 0) & (fx + rx <= Wf - 1) & (fy - ry >= 0) & (fy + ry <= Hf - 1)
    assert inside.sum() > 0
    np.testing.assert_allclose(get(out, w, h, s)[inside], (0.25 * fx - 0.5 * fy + 1.0)[inside], atol=1e-12)
    const = field("f64", s, h)
    orc.call("downscale", "f64", put(np.full((Hf, Wf), 0.7), "f64", Sf), Wf, Hf, Sf, w, h, s, const)
    np.testing.assert_allclose(get(const, w, h, s), 0.7, atol=1e-15)


@pytest.mark.parametrize("w,h,Wf,Hf", [(16, 12, 32, 24), (22, 15, 45, 31)])
def test_upscale_scales_and_preserves_linear(orc, w, h, Wf, Hf):
    s, Sf = 32, 64
    cy, cx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = 0.5 * cx + 0.25 * cy - 3.0
    out = field("f64", Sf, Hf)
    orc.call("upscale", "f64", put(img, "f64", s), w, h, s, Wf, Hf, Sf, 2.0, out)
    fy, fx = np.mgrid[0:Hf, 0:Wf].astype(np.float64)
    tx, ty = (fx + 0.5) * w / Wf - 0.5, (fy + 0.5) * h / Hf - 0.5
    inside = (tx >= 0) & (tx <= w - 1) & (ty >= 0) & (ty <= h - 1)
    np.testing.assert_allclose(get(out, Wf, Hf, Sf)[inside], 2.0 * (0.5 * tx + 0.25 * ty - 3.0)[inside], atol=1e-12)
    # exact 2x: the first fine pixel of each row sits left of texel 0 and clamps to it
    if Wf == 2 * w:
        np.testing.assert_allclose(get(out, Wf, Hf, Sf)[:, 0][(ty[:, 0] >= 0) & (ty[:, 0] <= h - 1)],
                                   2.0 * (0.25 * ty[:, 0] - 3.0)[(ty[:, 0] >= 0) & (ty[:, 0] <= h - 1)], atol=1e-12)


# ---------------------------------------------------------------- A5: derivatives
def test_derivatives_exact_for_quartics_and_mirror_at_border(orc, rng):
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    f0 = 1e-4 * xx**4 - 2e-3 * xx**3 + 0.01 * xx * yy + 1e-4 * yy**4
    f1 = 0.5 * f0 + 0.02 * xx - 0.03 * yy
    Ix, Iy, Iz = field("f64"), field("f64"), field("f64")
    orc.call("derivatives", "f64", put(f0, "f64"), put(f1, "f64"), W, H, S, Ix, Iy, Iz)

    def dx(f):
        return 4e-4 * xx**3 - 6e-3 * xx**2 + 0.01 * yy

    def dy(f):
        return 0.01 * xx + 4e-4 * yy**3

    ex = 0.5 * (dx(f0) + (0.5 * dx(f0) + 0.02))
    ey = 0.5 * (dy(f0) + (0.5 * dy(f0) - 0.03))
    core = (slice(2, H - 2), slice(2, W - 2))
    np.testing.assert_allclose(get(Ix, W, H)[core], ex[core], atol=1e-10)
    np.testing.assert_allclose(get(Iy, W, H)[core], ey[core], atol=1e-10)
    np.testing.assert_allclose(get(Iz, W, H), f1 - f0, atol=1e-15)
    # border: the stencil on the mirror-padded image
    g0, g1 = rng.random((H, W)), rng.random((H, W))
    orc.call("derivatives", "f64", put(g0, "f64"), put(g1, "f64"), W, H, S, Ix, Iy, Iz)
    k = np.array([1, -8, 0, 8, -1]) / 12.0

    def stencil_x(g):
        p = np.pad(g, ((0, 0), (2, 2)), mode="symmetric")
        return sum(k[j] * p[:, j:j + W] for j in range(5))

    def stencil_y(g):
        p = np.pad(g, ((2, 2), (0, 0)), mode="symmetric")
        return sum(k[j] * p[j:j + H, :] for j in range(5))

    np.testing.assert_allclose(get(Ix, W, H), 0.5 * (stencil_x(g0) + stencil_x(g1)), atol=1e-14)
    np.testing.assert_allclose(get(Iy, W, H), 0.5 * (stencil_y(g0) + stencil_y(g1)), atol=1e-14)


# ---------------------------------------------------------------- A6 / A7: TV-L1 steps
def _grad_via_dual(orc, u2d, kind="f64"):
    """Forward-difference gradient read off the dual step with a tiny tau/theta."""
    fac = 1e-9
    u = put(u2d, kind)
    z = field(kind)
    outs = [field(kind) for _ in range(4)]
    orc.call("dual_tv", kind, u, u, z, z, z, z, W, H, S, fac, 1.0, *outs)
    gx = get(outs[0], W, H) / fac
    gy = get(outs[1], W, H) / fac
    return gx, gy


def _div_via_data(orc, p1, p2, kind="f64"):
    """theta * div(p) read off the data step with a zero image gradient and zero residual."""
    z = field(kind)
    o1, o2 = field(kind), field(kind)
    P1, P2 = put(p1, kind), put(p2, kind)
    orc.call("data_l1", kind, z, z, P1, P2, P1, P2, z, z, z, W, H, S, 50.0, 1.0, o1, o2)
    return get(o1, W, H)


def test_divergence_is_minus_adjoint_of_gradient(orc, rng):
    u = rng.standard_normal((H, W))
    p1 = rng.standard_normal((H, W))
    p2 = rng.standard_normal((H, W))
    p1[:, -1] = 0.0  # the dual step keeps these at zero from a zero start
    p2[-1, :] = 0.0
    gx, gy = _grad_via_dual(orc, u)
    # forward differences with zero at the last column / row
    np.testing.assert_allclose(gx[:, :-1], np.diff(u, axis=1), rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(gx[:, -1], 0.0, atol=0)
    np.testing.assert_allclose(gy[:-1], np.diff(u, axis=0), rtol=1e-7, atol=1e-7)
    div = _div_via_data(orc, p1, p2)
    lhs = np.sum(np.diff(u, axis=1) * p1[:, :-1]) + np.sum(np.diff(u, axis=0) * p2[:-1])
    assert abs(lhs + np.sum(u * div)) < 1e-9 * (abs(lhs) + 1)


def test_data_l1_matches_textbook_threshold_branch_by_branch(orc):
    # Zach, Pock & Bischof (2007): v = u + {lt*gI if rho < -lt|gI|^2; -lt*gI if rho > lt|gI|^2;
    # -rho*gI/|gI|^2 otherwise}, with rho = gI.u + It
    lam, theta = 50.0, 0.3
    lt = float(np.float32(lam)) * float(np.float32(theta))  # float parameters, double product
    cases = [  # (Ix, Iy, It, u, v)
        (0.1, -0.2, 5.0, 0.3, -0.1),  # rho > lt g2
        (0.1, -0.2, -5.0, 0.3, -0.1),  # rho < -lt g2
        (0.1, -0.2, 0.01, 0.3, -0.1),  # inside the band
        (0.0, 0.0, 0.0, 0.7, 0.2),  # zero gradient, zero residual: identity
        (0.0, 0.0, 0.3, 0.7, 0.2),  # zero gradient, rho > 0 = upper: step of size 0
    ]
    n = len(cases)
    arr = np.array(cases, np.float64).T
    flat = lambda a: np.concatenate([a, np.zeros(64 * 2 - n)])  # noqa: E731
    Ix, Iy, It, u, v = (flat(a) for a in arr)
    z = np.zeros_like(Ix)
    o1, o2 = np.zeros_like(Ix), np.zeros_like(Ix)
    orc.call("data_l1", "f64", u, v, z, z, z, z, Ix, Iy, It, n, 1, 64, lam, theta, o1, o2)
    for i, (gx, gy, it, uu, vv) in enumerate(cases):
        g2 = gx * gx + gy * gy
        rho = gx * uu + gy * vv + it
        if rho < -lt * g2:
            e = (uu + lt * gx, vv + lt * gy)
        elif rho > lt * g2:
            e = (uu - lt * gx, vv - lt * gy)
        elif g2 > 0:
            e = (uu - rho * gx / g2, vv - rho * gy / g2)
        else:
            e = (uu, vv)
        np.testing.assert_allclose((o1[i], o2[i]), e, rtol=1e-12, atol=1e-15)


def test_data_l1_nan_rho_gives_nan(orc):
    a = np.zeros(128)
    It = a.copy()
    It[0] = np.nan
    o1, o2 = a.copy(), a.copy()
    orc.call("data_l1", "f64", a, a, a, a, a, a, a, a, It, 1, 1, 64, 50.0, 0.3, o1, o2)
    assert np.isnan(o1[0]) and np.isnan(o2[0])


def test_dual_projection_keeps_unit_ball_and_constant_is_fixed(orc, rng):
    u = rng.standard_normal((H, W)) * 3
    ang = rng.uniform(0, 2 * np.pi, (H, W))
    r = rng.uniform(0, 1, (H, W))
    p1, p2 = r * np.cos(ang), r * np.sin(ang)
    outs = [field("f64") for _ in range(4)]
    orc.call("dual_tv", "f64", put(u, "f64"), put(u, "f64"), put(p1, "f64"), put(p2, "f64"), put(p1, "f64"),
             put(p2, "f64"), W, H, S, 0.25, 0.3, *outs)
    mag = np.hypot(get(outs[0], W, H), get(outs[1], W, H))
    assert mag.max() <= 1.0 + 1e-12
    c = np.full((H, W), 1.5)
    orc.call("dual_tv", "f64", put(c, "f64"), put(c, "f64"), put(p1, "f64"), put(p2, "f64"), put(p1, "f64"),
             put(p2, "f64"), W, H, S, 0.25, 0.3, *outs)
    np.testing.assert_allclose(get(outs[0], W, H), p1, atol=0)
    np.testing.assert_allclose(get(outs[1], W, H), p2, atol=0)


# ---------------------------------------------------------------- A8 / A15: medians
def _median_ref5(img):
    # edge replication for taps outside in one direction, zero where outside in both
    h, w = img.shape
    p = np.pad(img, 2, mode="edge")
    for r in (0, 1, h + 2, h + 3):
        for c in (0, 1, w + 2, w + 3):
            p[r, c] = 0.0
    win = np.stack([p[j:j + h, i:i + w] for j in range(5) for i in range(5)], -1)
    return np.partition(win, 12, axis=-1)[..., 12]


def _median_ref3d(img):
    # left / top replicate; the right column and bottom row taps stay zero; corners zero
    h, w = img.shape
    p = np.zeros((h + 2, w + 2))
    p[1:-1, 1:-1] = img
    p[1:-1, 0] = img[:, 0]
    p[0, 1:-1] = img[0, :]
    win = np.stack([p[j:j + h, i:i + w] for j in range(3) for i in range(3)], -1)
    return np.partition(win, 4, axis=-1)[..., 4]


@pytest.mark.parametrize("kind", KINDS)
def test_median5_is_13th_order_statistic_with_reference_border(orc, kind, rng):
    for img in (rng.standard_normal((H, W)), rng.integers(-3, 4, (H, W)).astype(float)):
        ou, ov = field(kind), field(kind)
        f = put(img, kind)
        orc.call("median5", kind, f, f, W, H, S, ou, ov)
        np.testing.assert_array_equal(get(ou, W, H), _median_ref5(img.astype(f.dtype)))
        np.testing.assert_array_equal(get(ov, W, H), get(ou, W, H))


@pytest.mark.parametrize("kind", KINDS)
def test_median3d_is_5th_order_statistic_with_reference_border(orc, kind, rng):
    img = rng.standard_normal((H, W)) + 2.0
    outs = [field(kind) for _ in range(3)]
    f = put(img, kind)
    orc.call("median3d", kind, f, f, f, W, H, S, *outs)
    for o in outs:
        np.testing.assert_array_equal(get(o, W, H), _median_ref3d(img.astype(f.dtype)))


# ---------------------------------------------------------------- A12: joint data step
def test_recon_data_reduces_to_tvl1_when_alpha_proj_is_zero(orc, rng):
    a_tv = 3.0
    theta = 1.0 / a_tv
    f = [put(rng.standard_normal((H, W)) * sc, "f64") for sc in (0.1, 0.1, 0.05, 0.5, 0.5, 0.2, 0.2, 0.2, 0.2)]
    Ix, Iy, It, du, dv, p1, p2, q1, q2 = f
    junk = [put(rng.standard_normal((H, W)), "f64") for _ in range(6)]
    a1, a2, b1, b2 = (field("f64") for _ in range(4))
    orc.call("data_l1", "f64", du, dv, p1, p2, q1, q2, Ix, Iy, It, W, H, S, 50.0, np.float32(theta), a1, a2)
    orc.call("recon_data_l1", "f64", *junk[:2], du, dv, *junk[2:], p1, p2, q1, q2, Ix, Iy, It, W, H, S, 50.0,
             a_tv, 0.0, b1, b2)
    # theta enters as float(1/3) in one call and as 1/3.0 in the other: agreement to ~1e-7
    np.testing.assert_allclose(b1, a1, rtol=0, atol=1e-7)
    np.testing.assert_allclose(b2, a2, rtol=0, atol=1e-7)


def test_recon_data_blend_closed_form(orc, rng):
    # zero image gradient and zero dual: the step returns the blend (aTv*uh + aP*(up - u - sk))/(aTv+aP)
    a_tv, a_p = 33.3, 60.0
    u, v, du, dv, up, vp, sku, skv = (rng.standard_normal((H, W)) for _ in range(8))
    z = field("f64")
    o1, o2 = field("f64"), field("f64")
    args = [put(a, "f64") for a in (u, v, du, dv, up, vp, sku, skv)]
    orc.call("recon_data_l1", "f64", *args, z, z, z, z, z, z, z, W, H, S, 50.0, a_tv, a_p, o1, o2)
    aT, aP = np.float32(a_tv).item(), np.float32(a_p).item()
    np.testing.assert_allclose(get(o1, W, H), (aT * du + aP * (up - u - sku)) / (aT + aP), rtol=1e-12)
    np.testing.assert_allclose(get(o2, W, H), (aT * dv + aP * (vp - v - skv)) / (aT + aP), rtol=1e-12)


# ---------------------------------------------------------------- A13 / A18 / A16 / A17
def test_uvproj_and_sk_closed_forms(orc, rng):
    u, v, du, dv, uc, vc, sku, skv = (rng.standard_normal((H, W)) for _ in range(8))
    up, vp = field("f64"), field("f64")
    orc.call("uvproj", "f64", *[put(a, "f64") for a in (u, v, du, dv, uc, vc, sku, skv)], W, H, S, 0.1, 60.0, up, vp)
    lf, aP = np.float32(0.1).item(), 60.0
    np.testing.assert_allclose(get(up, W, H), (lf * uc + aP * (u + du + sku)) / (lf + aP), rtol=1e-13)
    np.testing.assert_allclose(get(vp, W, H), (lf * vc + aP * (v + dv + skv)) / (lf + aP), rtol=1e-13)
    s1, s2 = field("f64"), field("f64")
    orc.call("update_sk", "f64", *[put(a, "f64") for a in (sku, skv, u, v, du, dv)], up, vp, W, H, S, s1, s2)
    np.testing.assert_allclose(get(s1, W, H), sku + u + du - get(up, W, H), rtol=1e-13, atol=1e-14)


def test_cleanup3d_cases(orc):
    vals = np.array([
        # X, Y, Z, Xm, Ym, Zm
        [1.0, 2.0, 3.0, 1.5, 9.0, -3.0],  # only Y and Z move to the median (|diff| > 5)
        [np.inf, 2.0, 3.0, 1.0, 1.0, 1.0],  # non-finite: all zero
        [1.0, 2.0, 3.0, 1.0, np.nan, 1.0],
        [0.0, 0.0, 0.0, 5.0, -5.0, 5.0000001],  # |diff| == 5 keeps, > 5 replaces
    ])
    n = len(vals)
    cols = [np.concatenate([vals[:, i], np.zeros(64 - n)]) for i in range(6)]
    outs = [np.zeros(64) for _ in range(3)]
    orc.call("cleanup3d", "f64", *cols, n, 1, 64, *outs)
    got = np.stack([o[:n] for o in outs], -1)
    expect = np.array([[1.0, 9.0, -3.0], [0, 0, 0], [0, 0, 0], [0.0, 0.0, 5.0000001]])
    np.testing.assert_array_equal(got, expect)


def _rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def test_solve_uc_is_pinhole_projection(orc, rng):
    R = _rot([0.2, 1.0, -0.3], 0.04)
    t = np.array([0.1, -0.02, 0.01])
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    X, Y = rng.uniform(-2, 2, (H, W)), rng.uniform(-2, 2, (H, W))
    Z = rng.uniform(2, 10, (H, W))
    uc, vc = field("f64"), field("f64")
    orc.call("solve_uc", "f64", put(X, "f64"), put(Y, "f64"), put(Z, "f64"), R.ravel().copy(), t.copy(),
             K.ravel().copy(), W, H, S, uc, vc)
    P = np.stack([X, Y, Z], -1) @ R.T + t
    yy, xx = np.mgrid[0:H, 0:W]
    np.testing.assert_allclose(get(uc, W, H), 500 * P[..., 0] / P[..., 2] + 320 - xx, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(get(vc, W, H), 500 * P[..., 1] / P[..., 2] + 240 - yy, rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------- A14: triangulation
def _PQ(K, R, t):
    P = K @ np.hstack([np.eye(3), np.zeros((3, 1))])
    Q = K @ np.hstack([R, t[:, None]])
    return P, Q


def test_solve3d_without_regulariser_is_linear_triangulation(orc, rng):
    # Hartley & Zisserman DLT: exact correspondences triangulate exactly
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    R = _rot([0.3, -1.0, 0.2], 0.03)
    t = np.array([0.1, 0.01, -0.02])
    P, Q = _PQ(K, R, t)
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    xx += 300
    yy += 200  # pixels near the principal point region of a 640x480 camera
    Z = rng.uniform(2, 10, (H, W))
    Xw = np.stack([(xx - 320) / 500 * Z, (yy - 240) / 500 * Z, Z], -1)
    X1 = Xw @ R.T + t
    x2 = 500 * X1[..., 0] / X1[..., 2] + 320
    y2 = 500 * X1[..., 1] / X1[..., 2] + 240
    # the kernel uses the pixel index as x1, so shift the camera principal point instead
    Ks = K.copy()
    Ks[0, 2] -= 300
    Ks[1, 2] -= 200
    Ps, Qs = _PQ(Ks, R, t)
    up = x2 - 300 - (xx - 300)
    vp = y2 - 200 - (yy - 200)
    outs = [field("f64") for _ in range(3)]
    z = field("f64")
    orc.call("solve3d", "f64", put(up, "f64"), put(vp, "f64"), z, z, z, Ps.ravel().copy(), Qs.ravel().copy(),
             0.1, 0.0, W, H, S, *outs)
    got = np.stack([get(o, W, H) for o in outs], -1)
    np.testing.assert_allclose(got, Xw, rtol=1e-7, atol=1e-9)


def test_solve3d_regulariser_fixes_affine_surfaces(orc, rng):
    # with lambdaf = 0 an affine X field is a fixed point of the minimal-surface step
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    X = 0.01 * xx + 0.002 * yy - 1
    Y = -0.003 * xx + 0.015 * yy + 0.5
    Z = 0.02 * xx - 0.01 * yy + 5
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    P, Q = _PQ(K, np.eye(3), np.array([0.1, 0, 0]))
    outs = [field("f64") for _ in range(3)]
    z = field("f64")
    # extend the affine field one column and one row so the forward differences see it
    Xe, Ye, Ze = (np.pad(a, ((0, 1), (0, 1)), mode="reflect", reflect_type="odd") for a in (X, Y, Z))
    f = [field("f64") for _ in range(3)]
    for dst, src in zip(f, (Xe, Ye, Ze)):
        dst.reshape(H + 1, S)[:, :W + 1] = src
    orc.call("solve3d", "f64", z, z, *f, P.ravel().copy(), Q.ravel().copy(), 0.0, 100.0, W, H, S, *outs)
    core = (slice(1, H), slice(1, W))
    for o, ref in zip(outs, (X, Y, Z)):
        np.testing.assert_allclose(get(o, W, H)[core], ref[core], rtol=1e-9, atol=1e-9)


def test_solve3d_fma_tracks_plain_f64(orc, rng):
    # the FP64-triangulation mode evaluates the same normal equations with FMA chains, a
    # symmetric adjugate and one reciprocal of the determinant: on the same float inputs its
    # FP32 output is the plain FP64 restatement's result rounded to FP32 (the two differ by FP64
    # rounding only). lambdaf > 0 in every case: the minimal-area term alone is rank deficient
    # and its solve is arbitrary in any evaluation order.
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    P, Q = _PQ(K, _rot([0.2, 1, 0.1], 0.02), np.array([0.1, 0.01, 0.0]))
    up, vp = (rng.standard_normal((H, W)) * 0.5 for _ in range(2))
    Xs = [np.full((H, W), c) + 0.05 * rng.standard_normal((H, W)) for c in (0.3, -0.2, 5.0)]
    for lf, lms in ((0.1, 100.0), (0.1, 0.0), (1.0, 10.0), (0.01, 1000.0)):
        a32 = [put(x, "f32") for x in (up, vp, *Xs)]
        a64 = [put(x.astype(np.float32).astype(np.float64), "f64") for x in (up, vp, *Xs)]
        o32 = [field("f32") for _ in range(3)]
        o64 = [field("f64") for _ in range(3)]
        orc.call("solve3d_fma", "f32", *a32, P.ravel().copy(), Q.ravel().copy(), lf, lms, W, H, S, *o32)
        orc.call("solve3d", "f64", *a64, P.ravel().copy(), Q.ravel().copy(), lf, lms, W, H, S, *o64)
        for g, r in zip(o32, o64):
            g, r = get(g, W, H).astype(np.float64), get(r, W, H)
            # within one FP32 ulp of the FP64 value everywhere (half an ulp: correctly rounded)
            ulp = np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)
            d = np.abs(g - r) / ulp
            assert d.max() <= 1.0, (lf, lms, d.max())


# ---------------------------------------------------------------- f32 restatement vs f64
@pytest.mark.parametrize("name", ["data_l1", "dual_tv", "recon_data_l1", "solve3d"])
def test_f32_restatement_tracks_f64(orc, rng, name):
    K = np.array([[500.0, 0, 320], [0, 500, 240], [0, 0, 1]])
    P, Q = _PQ(K, _rot([0, 1, 0], 0.02), np.array([0.1, 0.0, 0.0]))
    ins = [rng.standard_normal((H, W)) * 0.2 for _ in range(15)]
    res = {}
    for kind in KINDS:
        a = [put(x, kind) for x in ins]
        o = [field(kind) for _ in range(4)]
        if name == "data_l1":
            orc.call(name, kind, *a[:9], W, H, S, 50.0, 0.3, o[0], o[1])
        elif name == "dual_tv":
            orc.call(name, kind, *a[:6], W, H, S, 0.25, 0.3, *o)
        elif name == "recon_data_l1":
            orc.call(name, kind, *a[:15], W, H, S, 50.0, 33.3, 60.0, o[0], o[1])
        else:
            Zs = [put(np.full((H, W), 5.0) + x, kind) for x in ins[2:5]]
            orc.call(name, kind, a[0], a[1], *Zs, P.ravel().copy(), Q.ravel().copy(), 0.1, 100.0, W, H, S, *o[:3])
        res[kind] = [get(x, W, H).astype(np.float64) for x in o]
    # The 3x3 Cramer solve of the triangulation is ill-conditioned: FP32 rounding alone moves
    # its result by ~1e-4 of the field scale on these inputs, the other steps stay near 1e-7.
    bound = 2e-3 if name == "solve3d" else 1e-5
    for a64, a32 in zip(res["f64"], res["f32"]):
        scale = max(np.abs(a64).max(), 1e-30)
        assert np.abs(a64 - a32).max() / scale < bound


# ---------------------------------------------------------------- A1: pyramid geometry
@pytest.mark.parametrize("w,h,L,expect", [
    (320, 240, 4, [(320, 240, 320), (160, 120, 160), (80, 60, 96), (40, 30, 64)]),
    (640, 480, 5, [(640, 480, 640), (320, 240, 320), (160, 120, 160), (80, 60, 96), (40, 30, 64)]),
    (1280, 720, 5, [(1280, 720, 1280), (640, 360, 640), (320, 180, 320), (160, 90, 160), (80, 45, 96)]),
    (176, 136, 5, [(176, 136, 192), (88, 68, 96), (44, 34, 64), (22, 17, 32), (11, 8, 32)]),
])
def test_pyramid_geometry(orc, w, h, L, expect):
    W, Hh, S_ = orc.pyramid(orc.flow_params(w, h, L))
    assert list(zip(W, Hh, S_)) == expect
