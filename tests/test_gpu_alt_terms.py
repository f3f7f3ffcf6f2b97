"""Parity of the post-filters, colour maps and alternate data terms on the GPU (per-kernel
entries and whole solves through the C ABI) against the oracle.

Bars: bit-exact against the FP32 restatement, except the two colour maps, whose hue comes from
atan2f (the GPU and host libraries may differ in the last ulp): within 1e-5 there."""
import numpy as np
import pytest

import reconflow_amd as rf
from tests.gpu_util import call, dev, dzeros, host, same_bits

pytestmark = pytest.mark.gpu

W, H, S = 77, 45, 96
N = (H + 1) * S


@pytest.fixture(scope="module")
def lib():
    from reconflow_amd import _native
    return _native.lib()


def flat(rng, lo, hi, n3=1):
    a = np.zeros(N * n3, np.float32)
    a.reshape(H + 1, S * n3)[:H, :W * n3] = rng.uniform(lo, hi, (H, W * n3)).astype(np.float32)
    return a


def valid(a, n3=1):
    return np.asarray(a).reshape(-1, S * n3)[:H, :W * n3]


def oracle_f32(orc, name, ins, *scalars, nout, n3=1):
    outs = [np.zeros(N * n3, np.float32) for _ in range(nout)]
    orc.call(name, "f32", *ins, *scalars, *outs)
    return outs


def test_compute_deriv_and_patch_derivatives(lib, orc):
    print("case: ComputeDeriv / ComputeDerivativesPatch 77x45", flush=True)
    rng = np.random.default_rng(31)
    I0, I1 = flat(rng, 0, 1), flat(rng, 0, 1)
    ref = oracle_f32(orc, "deriv_single", [I0], W, H, S, nout=2)
    g = [dzeros(N), dzeros(N)]
    assert call(lib, "compute_deriv", dev(I0), W, H, S, *g) == 0
    assert all(same_bits(valid(host(x)), valid(r)) for x, r in zip(g, ref))
    ref = oracle_f32(orc, "derivatives_patch", [I0, I1], W, H, S, nout=3)
    g = [dzeros(N) for _ in range(3)]
    assert call(lib, "compute_derivatives_patch", dev(I0), dev(I1), W, H, S, *g) == 0
    assert all(same_bits(valid(host(x)), valid(r)) for x, r in zip(g, ref))


@pytest.mark.parametrize("grad", [False, True], ids=["charb", "charbgrad"])
def test_charbonnier_data_steps(lib, orc, grad):
    print(f"case: SolveDataCharbForTV grad={grad}", flush=True)
    rng = np.random.default_rng(32 + grad)
    ins = [flat(rng, -0.5, 0.5) for _ in range(4)] + [flat(rng, -0.3, 0.3) for _ in range(4)] + \
          [flat(rng, -0.2, 0.2), flat(rng, -0.2, 0.2), flat(rng, -0.1, 0.1)]
    gi = [flat(rng, -0.2, 0.2) for _ in range(6)] if grad else [None] * 6
    ref = oracle_f32(orc, "data_charb", ins + gi, W, H, S, 50.0, 20.0, 0.3, nout=4)
    g = [dzeros(N) for _ in range(4)]
    dg = [dev(x) for x in gi] if grad else [None] * 6
    assert call(lib, "solve_data_charb_for_tv", *[dev(x) for x in ins], *dg, W, H, S, 50.0, 20.0, 0.3, *g) == 0
    for x, r in zip(g, ref):
        assert same_bits(valid(host(x)), valid(r))


def test_masked_l1_step(lib, orc):
    print("case: SolveDataL1Inpaint", flush=True)
    rng = np.random.default_rng(34)
    du, dv = flat(rng, -1, 1), flat(rng, -1, 1)
    mask = (flat(rng, 0, 1) > 0.5).astype(np.float32)
    ps = [flat(rng, -0.5, 0.5) for _ in range(4)]
    I = [flat(rng, -0.2, 0.2), flat(rng, -0.2, 0.2), flat(rng, -0.1, 0.1)]
    ref = oracle_f32(orc, "data_l1_inpaint", [du, dv, mask] + ps + I, W, H, S, 50.0, 0.3, nout=2)
    g = [dzeros(N), dzeros(N)]
    assert call(lib, "solve_data_l1_inpaint", dev(du), dev(dv), dev(mask), dev(mask), *[dev(x) for x in ps + I], W, H,
                S, 50.0, 0.3, *g) == 0
    for x, r in zip(g, ref):
        assert same_bits(valid(host(x)), valid(r))


def test_scene_flow_kernels(lib, orc):
    print("case: Cv16uTo32f / SolveSceneFlow / SceneFlowToHSV", flush=True)
    rng = np.random.default_rng(35)
    d16 = np.zeros(N, np.uint16)
    d16.reshape(H + 1, S)[:H, :W] = rng.integers(300, 3000, (H, W))
    g32 = dzeros(N)
    assert call(lib, "cv16u_to_32f", dev(d16.view(np.int16)), g32, W, H, S) == 0
    d32 = host(g32)
    assert np.array_equal(valid(d32), valid(d16).astype(np.float32))
    u, v = flat(rng, -3, 3), flat(rng, -3, 3)
    ref = oracle_f32(orc, "scene_flow", [u, v, d32, d32], 500.0, 510.0, 38.5, 22.5, W, H, S, nout=3, n3=3)
    g = [dzeros(3 * N) for _ in range(3)]
    assert call(lib, "solve_scene_flow", dev(u), dev(v), dev(d32), dev(d32), 500.0, 510.0, 38.5, 22.5, W, H, S,
                *g) == 0
    for x, r in zip(g, ref):
        assert same_bits(valid(host(x), 3), valid(r, 3))
    sf = ref[2]
    grgb = dzeros(3 * N)
    assert call(lib, "scene_flow_to_hsv", dev(sf), W, H, S, grgb, 50.0) == 0
    got = valid(host(grgb), 3)
    rgb_ref = np.zeros(3 * N, np.float32)
    orc.call("scene_flow_to_hsv", "f32", sf, W, H, S, rgb_ref, 50.0)
    np.testing.assert_allclose(got, valid(rgb_ref, 3), atol=1e-5 * max(1.0, np.abs(valid(rgb_ref, 3)).max()))


def test_flow_to_hsv_and_edge_filter(lib, orc):
    print("case: FlowToHSV / RemoveEdgeArtifacts3D", flush=True)
    rng = np.random.default_rng(36)
    u, v = flat(rng, -8, 8), flat(rng, -8, 8)
    ref = [np.zeros(3 * N, np.float32)]
    orc.call("flow_to_hsv", "f32", u, v, W, H, S, ref[0], 5.0)
    g = dzeros(3 * N)
    assert call(lib, "flow_to_hsv", dev(u), dev(v), W, H, S, g, 5.0) == 0
    np.testing.assert_allclose(valid(host(g), 3), valid(ref[0], 3), atol=1e-5)
    X, Y, Z = flat(rng, -1, 1), flat(rng, -1, 1), flat(rng, 2, 6)
    ref = oracle_f32(orc, "remove_edge", [X, Y, Z], 2, 0.05, W, H, S, nout=3)
    g = [dzeros(N) for _ in range(3)]
    assert call(lib, "remove_edge_artifacts_3d", dev(X), dev(Y), dev(Z), 2, 0.05, W, H, S, *g) == 0
    for x, r in zip(g, ref):
        assert same_bits(valid(host(x)), valid(r))


def _flow_solver(method, w, h, L, nw, ni):
    f = rf.CudaFlow()
    f.initialize(w, h, 1, rf.INPUT_32F, L, 2.0, method, 50.0, 20.0, 0.3, 0.25, nw, ni, False, path=rf.PATH_PLAIN)
    return f


@pytest.mark.parametrize("method", [1, 2, 3, 5, 7, 8])
def test_flow_methods_equal_oracle(orc, method):
    print(f"case: CudaFlow method {method} 96x72 plain", flush=True)
    w, h = 96, 72
    i0, i1, _, _ = rf.make_flow_pair(w, h, 40 + method)
    p = orc.flow_params(w, h, 3, n_warp=2, n_iter=5, method=method, lambdagrad=20.0)
    ref = orc.solve_flow(p, i0, i1, "f32")
    f = _flow_solver(method, w, h, 3, 2, 5)
    f.copyImagesToDevice(i0, i1)
    f.solveOpticalFlow(5.0)
    u, v = f.copyOpticalFlowToHost()
    assert np.array_equal(u[:, :w], ref[0][:, :w]) and np.array_equal(v[:, :w], ref[1][:, :w])


def test_inpainting_equals_oracle(orc):
    print("case: CudaFlow solveInpainting 96x72", flush=True)
    w, h = 96, 72
    i0, i1, _, _ = rf.make_flow_pair(w, h, 50)
    mask = (np.random.default_rng(5).random(i0.shape) > 0.3).astype(np.float32)
    p = orc.flow_params(w, h, 3, n_warp=2, n_iter=5, method=8)
    ref = orc.solve_flow_masked(p, i0, i1, mask, "f32")
    f = _flow_solver(8, w, h, 3, 2, 5)
    f.copyMaskToDevice(mask, mask)
    f.copyImagesToDevice(i0, i1)
    f.solveInpainting(5.0)
    u, v = f.copyOpticalFlowToHost()
    assert np.array_equal(u[:, :w], ref[0][:, :w]) and np.array_equal(v[:, :w], ref[1][:, :w])


def test_scene_flow_pipeline(orc):
    print("case: CudaFlow solveSceneFlowAndPointCloud 96x72", flush=True)
    w, h = 96, 72
    i0, i1, _, _ = rf.make_flow_pair(w, h, 51)
    s = i0.shape[1]
    d16 = np.random.default_rng(6).integers(400, 2600, (h, s)).astype(np.uint16)
    f = _flow_solver(4, w, h, 3, 2, 5)
    f.copyImagesToDevice(i0, i1)
    f.copyDepthToDevice(d16, d16)
    f.solveSceneFlowAndPointCloud(520.0, 515.0, 47.5, 35.5, 5.0, 2.0)
    sf, pc0, pc1, rgb = f.copySceneFlowAndPointCloudToHost()
    u, v = orc.solve_flow(orc.flow_params(w, h, 3, n_warp=2, n_iter=5, method=4), i0, i1, "f32")
    d32 = d16.astype(np.float32).ravel()
    n = h * s
    outs = [np.zeros(3 * n, np.float32) for _ in range(3)]
    orc.call("scene_flow", "f32", u.ravel(), v.ravel(), d32, d32, 520.0, 515.0, 47.5, 35.5, w, h, s, *outs)
    for got, ref in zip((pc0, pc1, sf), outs):
        assert same_bits(got[:, :w], ref.reshape(h, s, 3)[:, :w])
    rgb_ref = np.zeros(3 * n, np.float32)
    orc.call("scene_flow_to_hsv", "f32", outs[2], w, h, s, rgb_ref, 2.0)
    np.testing.assert_allclose(rgb[:, :w], rgb_ref.reshape(h, s, 3)[:, :w], atol=1e-3)
