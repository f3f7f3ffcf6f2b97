"""Pins of the oracle's warm start (an input this implementation adds; the reference always
starts from zero flow): the rule that decides whether a warm start is used, and what it does.

* With no warps at any level the solve only carries the initial flow up the pyramid, so a
  constant warm start (c, d) comes out as exactly (c, d): the scaling downscale of a constant is
  the constant times w_L / w_0, and each upscale doubles it back (exact halving sizes, powers of
  two throughout).
* A warm start of zeros is the reference's cold start, bit for bit.
* A warm start with a NaN, an infinity or a component beyond max(width, height) is not used:
  the solve equals the cold start; a component of exactly max(width, height) is used.
"""
import numpy as np
import pytest

import reconflow_amd as rf

W, H = 64, 48


def _ws(c, d, s):
    u = np.zeros((H, s), np.float32)
    v = np.zeros_like(u)
    u[:, :W], v[:, :W] = c, d
    return u, v


def _same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_constant_warm_start_without_warps_is_carried_exactly(orc, kind):
    i0, i1, _, _ = rf.make_flow_pair(W, H, 3)
    p = orc.flow_params(W, H, 3, n_warp=0, n_iter=4)
    s = orc.pyramid(p)[2][0]
    u, v = orc.solve_flow(p, i0, i1, kind, ws=_ws(1.75, -0.625, s))
    assert np.all(u[:, :W] == 1.75) and np.all(v[:, :W] == -0.625)
    d = rf.make_recon_pair(W, H, 4)
    pr = orc.recon_params(W, H, 3, n_warp=0, n_iter=4, tri_f64=1)
    out = orc.solve_recon(pr, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], kind,
                          ws=_ws(-3.5, 0.25, s))
    assert np.all(out[0][:, :W] == -3.5) and np.all(out[1][:, :W] == 0.25)


def test_warm_start_rule_and_fallback(orc):
    W, H = 60, 48  # stride 64: padding columns
    i0, i1, gu, gv = rf.make_flow_pair(W, H, 5)
    p = orc.flow_params(W, H, 3, n_warp=2, n_iter=5)
    cold = orc.solve_flow(p, i0, i1, "f32")
    assert _same(orc.solve_flow(p, i0, i1, "f32", ws=(np.zeros_like(gu), np.zeros_like(gv))), cold)
    warm = orc.solve_flow(p, i0, i1, "f32", ws=(gu, gv))
    assert not _same(warm, cold)
    bound = float(max(W, H))
    for bad in (np.nan, np.inf, -np.inf, np.nextafter(np.float32(bound), np.float32(2 * bound))):
        u, v = gu.copy(), gv.copy()
        u[7, 9] = bad
        assert not orc.warm_start_usable(u, v, W, H)
        assert _same(orc.solve_flow(p, i0, i1, "f32", ws=(u, v)), cold), bad
    u, v = gu.copy(), gv.copy()
    v[3, 4] = -bound
    assert orc.warm_start_usable(u, v, W, H)
    # values in the padding columns are not part of the rule
    u[2, W + 1] = np.nan
    assert orc.warm_start_usable(u, v, W, H)
