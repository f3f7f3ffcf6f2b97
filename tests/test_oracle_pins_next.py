"""Pins of the oracle's restatements of the display / post-filter / alternate-data-term kernels
(CPU only): closed forms, textbook IRLS steps and invariants."""
import numpy as np
import pytest

from tests.test_oracle_pins import field, get, put

W, H, S = 37, 23, 64


def middlebury_wheel():
    """The Middlebury colour wheel built from its published recipe (Baker et al., colorcode.cpp):
    hue segments RY=15, YG=6, GC=4, CB=11, BM=13, MR=6, ramps floor(255*i/n)."""
    cols = []
    for n, f in ((15, lambda r: (255, r, 0)), (6, lambda r: (255 - r, 255, 0)), (4, lambda r: (0, 255, r)),
                 (11, lambda r: (0, 255 - r, 255)), (13, lambda r: (r, 0, 255)), (6, lambda r: (255, 0, 255 - r))):
        cols += [f(255 * i // n) for i in range(n)]
    return np.array(cols, float) / 255.0  # (55, 3) RGB


def test_flow_to_hsv_closed_forms(orc):
    fs = 5.0
    wheel = middlebury_wheel()
    # zero flow is white; flow (-fs, 0) has atan2(-0, 1) = 0, fk = 28 exactly: wheel entry 28
    u = np.array([0.0, -fs, 0.0, -3 * fs] + [0.0] * 60)
    v = np.zeros_like(u)
    v[2] = fs  # atan2(-1, -0) = -pi/2: fk just below 14.5, between entries 14 and 15
    rgb = np.zeros(3 * u.size)
    orc.call("flow_to_hsv", "f64", u, v, 4, 1, 64, rgb, fs)
    out = rgb.reshape(-1, 3)[:4]
    np.testing.assert_allclose(out[0], [1, 1, 1])
    np.testing.assert_allclose(out[1], wheel[28][::-1], atol=1e-12)  # stored B, G, R; rad == 1
    a = np.arctan2(-1.0, -0.0) / float(np.float32(3.14159))
    fk = (a + 1) / 2 * 54 + 1
    assert 14 < fk < 15
    f = fk - 14
    np.testing.assert_allclose(out[2], ((1 - f) * wheel[14] + f * wheel[15])[::-1], atol=1e-9)
    # |flow| > 1 keeps the hue at 0.75 brightness
    np.testing.assert_allclose(out[3], 0.75 * wheel[28][::-1], atol=1e-12)


def test_flow_to_hsv_sweeps_the_whole_wheel(orc):
    # flows on the unit circle at the wheel's 54 knot angles hit every entry k = 1..54 exactly
    # (up to the 3.14159 approximation of pi used for the angle normalisation)
    wheel = middlebury_wheel()
    k = np.arange(1, 55)
    a = (k - 1) / 54 * 2 - 1  # fk = k
    ang = a * float(np.float32(3.14159))
    u, v = -np.cos(ang), -np.sin(ang)  # atan2(-v, -u) = ang
    rgb = np.zeros(3 * 64)
    orc.call("flow_to_hsv", "f64", np.r_[u, np.zeros(10)], np.r_[v, np.zeros(10)], 54, 1, 64, rgb, 1.0)
    got = rgb.reshape(-1, 3)[:54, ::-1]
    np.testing.assert_allclose(got, wheel[k % 55], atol=1e-6)


def test_remove_edge_keeps_planes_and_drops_steps(orc):
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    X, Y, Z = 0.01 * xx, 0.01 * yy, 5 + 0.001 * xx
    Z2 = Z.copy()
    Z2[:, 20:] += 3.0  # depth step between columns 19 and 20
    outs = [field("f64") for _ in range(3)]
    for ZZ, expect_drop in ((Z, False), (Z2, True)):
        f = [put(a, "f64") for a in (X, Y, ZZ)]
        for a, src in zip(f, (X, Y, ZZ)):  # extend one column / row so border differences see the field
            a.reshape(H + 1, S)[:H, W] = src[:, -1]
            a.reshape(H + 1, S)[H, :W] = src[-1, :]
        orc.call("remove_edge", "f64", *f, 1, 0.01, W, H, S, *outs)
        kept = get(outs[2], W, H) != 0
        if expect_drop:  # the last row differences against the duplicated guard row: zero area
            assert not kept[:H - 1, 19].any() and kept[:, 21].all() and kept[H - 1].all()
        else:
            assert kept.all()


def test_remove_edge_area_is_the_cross_product(orc, rng):
    # range-2 forward differences (backward within `range` of the far border), parallelogram area
    r, maxarea = 2, 0.05
    fx = [rng.standard_normal((H + 1, S)) * 0.1 for _ in range(3)]
    outs = [field("f64") for _ in range(3)]
    orc.call("remove_edge", "f64", *[a.ravel().copy() for a in fx], r, maxarea, W, H, S, *outs)
    P = np.stack([a[:H, :W] for a in fx], -1)
    full = np.stack(fx, -1)
    dx = np.empty_like(P)
    dy = np.empty_like(P)
    for x in range(W):
        dx[:, x] = full[:H, x + r] - P[:, x] if x + r <= W else P[:, x] - full[:H, x - r]
    for y in range(H):
        dy[y] = full[y + r, :W] - P[y] if y + r <= H else P[y] - full[y - r, :W]
    area = np.linalg.norm(np.cross(dx, dy), axis=-1)
    kept = area <= maxarea
    assert 0 < kept.sum() < kept.size
    for o, a in zip(outs, fx):
        np.testing.assert_array_equal(get(o, W, H), np.where(kept, a[:H, :W], 0.0))


def test_deriv_single_exact_for_quartics(orc):
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    f = 1e-4 * xx**4 + 0.02 * xx * yy - 1e-4 * yy**3
    Ix, Iy = field("f64"), field("f64")
    orc.call("deriv_single", "f64", put(f, "f64"), W, H, S, Ix, Iy)
    core = (slice(2, H - 2), slice(2, W - 2))
    np.testing.assert_allclose(get(Ix, W, H)[core], (4e-4 * xx**3 + 0.02 * yy)[core], atol=1e-10)
    np.testing.assert_allclose(get(Iy, W, H)[core], (0.02 * xx - 3e-4 * yy**2)[core], atol=1e-10)


def test_patch_derivatives_sum_the_patch(orc):
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    f0 = 0.3 * xx - 0.2 * yy
    f1 = f0 + 0.05
    o = [field("f64") for _ in range(3)]
    orc.call("derivatives_patch", "f64", put(f0, "f64"), put(f1, "f64"), W, H, S, *o)
    core = (slice(4, H - 4), slice(4, W - 4))
    np.testing.assert_allclose(get(o[0], W, H)[core], 25 * 0.3, atol=1e-12)
    np.testing.assert_allclose(get(o[1], W, H)[core], 25 * -0.2, atol=1e-12)
    np.testing.assert_allclose(get(o[2], W, H), 25 * 0.05, atol=1e-12)


@pytest.mark.parametrize("grad", [False, True])
def test_charbonnier_step_solves_the_irls_system(orc, rng, grad):
    n = 40
    lam, lgrad, theta = 50.0, 20.0, 0.3
    cols = {k: rng.standard_normal(n) * 0.2 for k in ("du", "dv", "uh", "vh", "Ix", "Iy", "It", "Ixx", "Ixy", "Ixt",
                                                        "Iyx", "Iyy", "Iyt")}
    z = np.zeros(64)

    def pad(a):
        return np.concatenate([a, np.zeros(64 - n)])

    o = [np.zeros(64) for _ in range(4)]
    g = [pad(cols[k]) for k in ("Ixx", "Ixy", "Ixt", "Iyx", "Iyy", "Iyt")] if grad else [None] * 6
    orc.call("data_charb", "f64", *[pad(cols[k]) for k in ("du", "dv", "uh", "vh")], z, z, z, z,
             *[pad(cols[k]) for k in ("Ix", "Iy", "It")], *g, n, 1, 64, lam, lgrad, theta, *o)
    lamf, lgf, thf = (float(np.float32(x)) for x in (lam, lgrad, theta))
    for i in range(n):
        d = np.array([cols["du"][i], cols["dv"][i]])
        gI = np.array([cols["Ix"][i], cols["Iy"][i]])
        terms = [(gI, cols["It"][i], lamf)]
        if grad:
            terms += [(np.array([cols["Ixx"][i], cols["Ixy"][i]]), cols["Ixt"][i], lgf),
                      (np.array([cols["Iyx"][i], cols["Iyy"][i]]), cols["Iyt"][i], lgf)]
        for _ in range(2 if grad else 1):  # IRLS passes: weights from the current estimate
            A = np.eye(2) / thf
            b = np.array([cols["uh"][i], cols["vh"][i]]) / thf
            for gv, gt, lw in terms:
                beta = lw / np.sqrt((gv @ d + gt) ** 2 + 0.001)
                A += beta * np.outer(gv, gv)
                b -= beta * gv * gt
            d = np.linalg.solve(A, b)
        np.testing.assert_allclose([o[0][i], o[1][i]], d, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose([o[2][i], o[3][i]], d, rtol=1e-9, atol=1e-12)  # zero dual: uhat' = d


def test_inpaint_step_freezes_masked_pixels(orc, rng):
    ins = [put(rng.standard_normal((H, W)) * 0.2, "f64") for _ in range(9)]
    mask = (rng.random((H, W)) > 0.5).astype(float)
    a1, a2, b1, b2 = (field("f64") for _ in range(4))
    orc.call("data_l1", "f64", *ins, W, H, S, 50.0, 0.3, a1, a2)
    orc.call("data_l1_inpaint", "f64", ins[0], ins[1], put(mask, "f64"), *ins[2:], W, H, S, 50.0, 0.3, b1, b2)
    on = mask != 0
    np.testing.assert_array_equal(get(b1, W, H)[on], get(a1, W, H)[on])
    # masked pixels keep uhat (plus the dual's divergence, zero here when p is zero)
    z = field("f64")
    orc.call("data_l1_inpaint", "f64", ins[0], ins[1], put(mask, "f64"), z, z, z, z, *ins[6:9], W, H, S, 50.0, 0.3,
             b1, b2)
    np.testing.assert_array_equal(get(b1, W, H)[~on], get(ins[0], W, H)[~on])


def test_scene_flow_back_projection(orc):
    w, h, s = 8, 6, 32
    d0 = np.full((h, s), 1000.0)
    d0[:, :4] = 1500.0
    d0[0, 0] = 300.0  # outside [500, 2500]: zero scene flow
    u = np.zeros((h, s))
    u[:, 3] = 1.0  # column 3 looks at column 4 (depth 1000)
    v = np.zeros((h, s))
    pc0, pc1, sf = (np.zeros(3 * h * s) for _ in range(3))
    orc.call("scene_flow", "f64", u.ravel(), v.ravel(), d0.ravel(), d0.ravel(), 500.0, 500.0, 4.0, 3.0, w, h, s,
             pc0, pc1, sf)
    sf = sf.reshape(h, s, 3)
    np.testing.assert_allclose(sf[2, 3], [((3 - 4) * 1500 - (3 - 4) * 1000) / 500, ((2 - 3) * 1500 - (2 - 3) * 1000) / 500,
                                          500.0])
    np.testing.assert_allclose(sf[0, 0], [0, 0, 0])
    np.testing.assert_allclose(sf[2, 5], [0, 0, 0])


def test_scene_flow_to_hsv_is_the_hsv_sector_map(orc, rng):
    # hue H = atan2(y, x) + 3.1416 in [0, 2pi), sectors of 1.0472; C = |(x, y)|, m = z - C; the
    # textbook HSV-to-RGB sector table (C, X, 0), (X, C, 0), (0, C, X), (0, X, C), (X, 0, C), (C, 0, X)
    n = 64
    ang = rng.uniform(-3.1, 3.1, n)
    mag = rng.uniform(0.1, 2.0, n)
    z = rng.uniform(-1, 1, n)
    sf = np.stack([mag * np.cos(ang), mag * np.sin(ang), z], -1).ravel()
    rgb = np.zeros(3 * n)
    orc.call("scene_flow_to_hsv", "f64", sf, n, 1, n, rgb, 1.0)
    got = rgb.reshape(n, 3)
    for k in range(n):
        x, y, zz = sf[3 * k:3 * k + 3]
        Hh = np.arctan2(y, x) + float(np.float32(3.1416))
        C = np.hypot(x, y)
        hp = Hh / float(np.float32(1.0472))
        X = C * (1 - abs(hp % 2 - 1))
        sector = min(int(Hh // 1.0472), 5)
        table = [(C, X, 0), (X, C, 0), (0, C, X), (0, X, C), (X, 0, C), (C, 0, X)]
        if abs(Hh - round(Hh / 1.0472) * 1.0472) < 1e-3:
            continue  # the reference's sector bounds are rounded literals; skip angles next to them
        np.testing.assert_allclose(got[k], np.array(table[sector]) + (zz - C), atol=1e-12)


def test_flow_methods_dispatch(orc):
    # methods without a data-term branch of their own run the L1 step (Cudaflow.cu:540-557), and
    # the masked solve with an all-ones mask is the plain solve
    import reconflow_amd as rf
    i0, i1, _, _ = rf.make_flow_pair(64, 48, 3)
    base = orc.solve_flow(orc.flow_params(64, 48, 2, n_warp=1, n_iter=5), i0, i1, "f64")
    for m in (4, 5, 6, 8):
        got = orc.solve_flow(orc.flow_params(64, 48, 2, n_warp=1, n_iter=5, method=m), i0, i1, "f64")
        assert all(np.array_equal(a, b) for a, b in zip(got, base)), m
    ones = np.ones_like(i0)
    got = orc.solve_flow_masked(orc.flow_params(64, 48, 2, n_warp=1, n_iter=5, method=8), i0, i1, ones, "f64")
    assert all(np.array_equal(a, b) for a, b in zip(got, base))
    zeros = np.zeros_like(i0)
    got = orc.solve_flow_masked(orc.flow_params(64, 48, 2, n_warp=1, n_iter=5, method=8), i0, i1, zeros, "f64")
    assert not np.any(got[0]) and not np.any(got[1])  # no data term anywhere: the flow stays 0
    # the Charbonnier and patch solves move the flow (they are not the L1 solve)
    for m in (1, 2, 3, 7):
        got = orc.solve_flow(orc.flow_params(64, 48, 2, n_warp=1, n_iter=5, method=m), i0, i1, "f64")
        assert not np.array_equal(got[0], base[0]), m
