"""The oracle against the committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py): any change to the restatement that moves a result fails here."""
from pathlib import Path

import numpy as np
import pytest

G = Path(__file__).resolve().parent / "golden" / "golden.npz"
W, H, S = 64, 40, 64


@pytest.fixture(scope="module")
def gold():
    with np.load(G) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name,args,nout", [("data_l1", (50.0, 0.3), 2), ("dual_tv", (0.25, 0.3), 4),
                                            ("median5", (), 2)])
def test_kernel_fixtures(orc, gold, name, args, nout):
    ins = gold[f"{name}_in"]
    n = (H + 1) * S
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(nout)]
        orc.call(name, kind, *[x.astype(dt) for x in ins], W, H, S, *args, *o)
        assert np.array_equal(np.stack(o), gold[f"{name}_out_{kind}"], equal_nan=True), (name, kind)


def test_tvl1_fixture(orc, gold):
    p = orc.flow_params(96, 72, 3, n_warp=2, n_iter=10)
    for kind in ("f32", "f64"):
        u, v = orc.solve_flow(p, gold["tvl1_i0"], gold["tvl1_i1"], kind)
        assert np.array_equal(u, gold[f"tvl1_u_{kind}"]) and np.array_equal(v, gold[f"tvl1_v_{kind}"]), kind


@pytest.mark.parametrize("tri", [0, 1])
def test_joint_fixture(orc, gold, tri):
    p = orc.recon_params(96, 72, 3, n_warp=2, n_iter=10, tri_f64=tri)
    K = gold["joint_K"]
    res = orc.solve_recon(p, K, K, np.eye(3), np.zeros(3), gold["joint_R"], gold["joint_t"], gold["joint_i0"],
                          gold["joint_i1"], "f32")
    for name, a in zip(("u", "v", "X", "Y", "Z"), res):
        assert np.array_equal(a, gold[f"joint_tri{tri}_{name}_f32"], equal_nan=True), name
