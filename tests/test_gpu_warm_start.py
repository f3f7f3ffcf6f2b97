"""The warm-start input on the GPU (rf_copy_warm_start, rf_batch_set_warm_start_device) against
the oracle's warm start: TV-L1 and joint solves on the plain, fused and captured-graph paths,
batches of one size and of uneven sizes, the fallback rule (a warm start with a NaN, an infinity
or a component beyond max(width, height) is not used, and the solve equals the cold start), and
the closed form of a constant warm start with no warps."""
import numpy as np
import pytest

import reconflow_amd as rf
from tests.gpu_util import dev, fetch

pytestmark = pytest.mark.gpu

W, H = 160, 120


def _same(a, b, w):
    return all(np.array_equal(x[:, :w], y[:, :w], equal_nan=True) for x, y in zip(a, b))


def _flow(path, graph, w=W, h=H, L=3, nw=2, ni=6):
    f = rf.CudaFlow()
    f.initialize(w, h, 1, rf.INPUT_32F, L, 2.0, rf.CudaFlow.METHOD_TVL1, 50.0, 0.0, 0.3, 0.25, nw, ni, False,
                 path=path, use_graph=graph)
    return f


def _recon(path, graph, w=W, h=H, L=3, nw=2, ni=6):
    r = rf.ReconFlow()
    r._initializeR(w, h, 1, L, 2.0, 14, 50.0, 0.0, 0.1, 100.0, 33.3, 60.0, 0.125, nw, ni, cvType=rf.INPUT_32F,
                   path=path, use_graph=graph)
    r.params.tri_f64 = 1
    r._init_params(r.params)
    return r


def _flow_solve(f, i0, i1):
    f.copyImagesToDevice(i0, i1)
    f.solveOpticalFlow(5.0)
    return f.copyOpticalFlowToHost()


def _recon_solve(r, d):
    r.setCameraMatrices(d["K"], d["K"])
    r.copyImagesToDevice(d["i0"], d["i1"])
    r.solveR(d["R"], d["t"], 50.0)
    return (*r.copyOpticalFlowToHost(), *r.copy3dToHost())


PATHS = [(rf.PATH_PLAIN, False), (rf.PATH_PLAIN, True), (rf.PATH_FUSED, False), (rf.PATH_FUSED, True)]
IDS = ["plain", "plain-graph", "fused", "fused-graph"]


@pytest.mark.parametrize("path,graph", PATHS, ids=IDS)
def test_flow_warm_start_equals_oracle(orc, path, graph):
    print(f"case: TV-L1 {W}x{H} warm start path={path} graph={graph}", flush=True)
    i0, i1, gu, gv = rf.make_flow_pair(W, H, 41)
    p = orc.flow_params(W, H, 3, n_warp=2, n_iter=6)
    ref_warm = orc.solve_flow(p, i0, i1, "f32", ws=(gu, gv))
    ref_cold = orc.solve_flow(p, i0, i1, "f32")
    f = _flow(path, graph)
    f.copyWarmStartToDevice(gu, gv)
    assert not f.warmStartInUse()  # no solve has applied it yet
    for _ in range(2):  # a captured schedule replays the warm start
        assert _same(_flow_solve(f, i0, i1), ref_warm, W)
        assert f.warmStartInUse()
    bad = gu.copy()
    bad[10, 20] = np.nan
    f.copyWarmStartToDevice(bad, gv)
    assert _same(_flow_solve(f, i0, i1), ref_cold, W)
    assert not f.warmStartInUse()
    f.copyWarmStartToDevice(gu, gv)
    assert _same(_flow_solve(f, i0, i1), ref_warm, W)
    f.copyWarmStartToDevice(None, None)
    assert _same(_flow_solve(f, i0, i1), ref_cold, W)
    assert not f.warmStartInUse()


@pytest.mark.parametrize("path,graph", PATHS, ids=IDS)
def test_recon_warm_start_equals_oracle(orc, path, graph):
    print(f"case: joint {W}x{H} warm start path={path} graph={graph}", flush=True)
    d = rf.make_recon_pair(W, H, 42)
    ws = (d["gt_u"], d["gt_v"])
    pr = orc.recon_params(W, H, 3, n_warp=2, n_iter=6, tri_f64=1)
    K, R, t = d["K"], d["R"], d["t"]
    ref_warm = orc.solve_recon(pr, K, K, np.eye(3), np.zeros(3), R, t, d["i0"], d["i1"], "f32", ws=ws)
    ref_cold = orc.solve_recon(pr, K, K, np.eye(3), np.zeros(3), R, t, d["i0"], d["i1"], "f32")
    r = _recon(path, graph)
    r.copyWarmStartToDevice(*ws)
    assert _same(_recon_solve(r, d), ref_warm, W)
    assert r.warmStartInUse()
    huge = ws[0].copy()
    huge[0, 0] = 1e6  # beyond max(width, height): not used
    r.copyWarmStartToDevice(huge, ws[1])
    assert _same(_recon_solve(r, d), ref_cold, W)
    assert not r.warmStartInUse()


def test_constant_warm_start_without_warps():
    # no warps: the constant warm start is carried up the pyramid unchanged
    print("case: TV-L1 64x48 constant warm start, no warps", flush=True)
    i0, i1, _, _ = rf.make_flow_pair(64, 48, 43)
    for path in (rf.PATH_PLAIN, rf.PATH_FUSED):
        f = _flow(path, False, 64, 48, 3, 0, 4)
        u = np.zeros((48, f.stride), np.float32)
        v = np.zeros_like(u)
        u[:, :64], v[:, :64] = 2.25, -1.5
        f.copyWarmStartToDevice(u, v)
        gu, gv = _flow_solve(f, i0, i1)
        assert np.all(gu[:, :64] == 2.25) and np.all(gv[:, :64] == -1.5), path


@pytest.mark.parametrize("sizes", [[(144, 104)] * 3, [(144, 104), (96, 72), (160, 120)]], ids=["same", "uneven"])
def test_batch_warm_start_device(orc, sizes):
    print(f"case: batch joint warm start sizes={sizes}", flush=True)
    n = len(sizes)
    p = rf.batch_params(14, 3, 2, 5, tau=0.125, use_graph=True, tri_f64=1)
    b = rf.Batch(p, [w for w, _ in sizes], [h for _, h in sizes])
    data, keep = [], []
    for i, (w, h) in enumerate(sizes):
        d = rf.make_recon_pair(w, h, 350 + i)
        d0, d1 = dev(d["i0"].ravel()), dev(d["i1"].ravel())
        keep += [d0, d1]
        b.set_images_device(i, d0.data_ptr(), d1.data_ptr())
        b.set_camera(i, d["K"], d["K"], d["R"], d["t"])
        data.append(d)
    # pairs 0 and 2 warm-started from their ground truth, pair 1 cold; then pair 2's warm start
    # is made unusable
    assert not any(b.warm_start_in_use(i) for i in range(n))
    wsd = {}
    for i in (0, 2):
        wsd[i] = (dev(data[i]["gt_u"].ravel()), dev(data[i]["gt_v"].ravel()))
        b.set_warm_start_device(i, wsd[i][0].data_ptr(), wsd[i][1].data_ptr())
    for rep in range(2):
        if rep == 1:
            wsd[2][0][5] = float("inf")
        b.solve()
        b.synchronize()
        for i, (w, h) in enumerate(sizes):
            s = rf.align_up(w)
            got = [fetch(ptr, h * s).reshape(h, s) for ptr in b.result_ptrs(i)]
            d = data[i]
            warm = i == 0 or (i == 2 and rep == 0)
            pr = orc.recon_params(w, h, 3, n_warp=2, n_iter=5, tri_f64=1)
            ref = orc.solve_recon(pr, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f32",
                                  ws=(d["gt_u"], d["gt_v"]) if warm else None)
            assert _same(got, ref, w), (rep, i)
            assert b.warm_start_in_use(i) == warm, (rep, i)
