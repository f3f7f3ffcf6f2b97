"""The HIP path against the committed golden fixtures (tests/golden/golden.npz): per-kernel
outputs and whole 96x72 solves must equal the oracle's FP32 results bit for bit, and the
per-kernel outputs must be within 1e-5 normwise of its FP64 results."""
from pathlib import Path

import numpy as np
import pytest

import reconflow_amd as rf
from tests.gpu_util import call, dev, dzeros, host, normwise, same_bits

pytestmark = pytest.mark.gpu

G = Path(__file__).resolve().parent / "golden" / "golden.npz"
W, H, S = 64, 40, 64


@pytest.fixture(scope="module")
def gold():
    with np.load(G) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def lib():
    from reconflow_amd import _native
    return _native.lib()


def _valid(a):
    return np.asarray(a).reshape(-1, S)[:H, :W]


@pytest.mark.parametrize("name,entry,args,nout", [("data_l1", "solve_data_l1", (50.0, 0.3), 2),
                                                  ("dual_tv", "solve_smooth_dual_tv_global", (0.25, 0.3), 4),
                                                  ("median5", "median_filter", (5,), 2)])
def test_kernel_fixtures(lib, gold, name, entry, args, nout):
    print(f"case: golden kernel {name}", flush=True)
    ins = [dev(x) for x in gold[f"{name}_in"]]
    n = (H + 1) * S
    outs = [dzeros(n) for _ in range(nout)]
    if name == "median5":  # rf_k_median_filter(u, v, w, h, s, ou, ov, kernelsize)
        assert call(lib, entry, *ins, W, H, S, *outs, *args) == 0
    else:
        assert call(lib, entry, *ins, W, H, S, *args, *outs) == 0
    for g, a32, a64 in zip(outs, gold[f"{name}_out_f32"], gold[f"{name}_out_f64"]):
        gv = _valid(host(g))
        assert same_bits(gv, _valid(a32))
        assert normwise(gv, _valid(a64)) <= 1e-5


@pytest.mark.parametrize("path,graph", [(rf.PATH_PLAIN, False), (rf.PATH_FUSED, True)], ids=["plain", "fused-graph"])
def test_tvl1_fixture(gold, path, graph):
    print(f"case: golden TV-L1 96x72 path={path}", flush=True)
    f = rf.CudaFlow()
    f.initialize(96, 72, 1, rf.INPUT_32F, 3, 2.0, rf.CudaFlow.METHOD_TVL1, 50.0, 0.0, 0.3, 0.25, 2, 10, False,
                 path=path, use_graph=graph)
    f.copyImagesToDevice(gold["tvl1_i0"], gold["tvl1_i1"])
    f.solveOpticalFlow(5.0)
    u, v = f.copyOpticalFlowToHost()
    assert np.array_equal(u[:, :96], gold["tvl1_u_f32"][:, :96]) and np.array_equal(v[:, :96], gold["tvl1_v_f32"][:, :96])


@pytest.mark.parametrize("tri", [0, 1])
@pytest.mark.parametrize("path,graph", [(rf.PATH_PLAIN, False), (rf.PATH_FUSED, True)], ids=["plain", "fused-graph"])
def test_joint_fixture(gold, path, graph, tri):
    print(f"case: golden joint 96x72 path={path} tri={tri}", flush=True)
    r = rf.ReconFlow()
    r._initializeR(96, 72, 1, 3, 2.0, rf.ReconFlow.METHODR_TVL1_MS, 50.0, 0.0, 0.1, 100.0, 33.3, 60.0, 0.125, 2, 10,
                   cvType=rf.INPUT_32F, path=path, use_graph=graph)
    if tri:
        r.params.tri_f64 = 1
        r._init_params(r.params)
    K = gold["joint_K"]
    r.setCameraMatrices(K, K)
    r.copyImagesToDevice(gold["joint_i0"], gold["joint_i1"])
    r.solveR(gold["joint_R"], gold["joint_t"], 50.0)
    got = r.copyOpticalFlowToHost() + r.copy3dToHost()
    for name, a in zip(("u", "v", "X", "Y", "Z"), got):
        assert np.array_equal(a[:, :96], gold[f"joint_tri{tri}_{name}_f32"][:, :96], equal_nan=True), name
