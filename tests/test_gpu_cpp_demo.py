"""The C++ drop-in classes (include/rfhip_cudaflow.hpp) driven by the prebuilt rf_demo binary:
results equal the FP32 restatement of the oracle bit for bit."""
import subprocess

import numpy as np
import pytest

import reconflow_amd as rf
from reconflow_amd import _native

pytestmark = pytest.mark.gpu

DEMO = _native.BUILD / "rf_demo"


def _run(args, tmp_path):
    out = tmp_path / "out.bin"
    print("case: rf_demo", *args, flush=True)
    proc = subprocess.run([str(DEMO), *map(str, args[:7]), str(out), *args[7:]], capture_output=True, text=True,
                          timeout=300)
    assert proc.returncode == 0, proc.stderr
    return out


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_demo_flow(orc, tmp_path, mode):
    w, h, L, nw, ni, seed = 160, 120, 3, 2, 20, 21
    out = _run(["flow", w, h, L, nw, ni, seed] + (["fused"] if mode == "fused" else []), tmp_path)
    s = rf.align_up(w)
    got = np.fromfile(out, np.float32).reshape(2, h, s)
    i0, i1, _, _ = rf.make_flow_pair(w, h, seed)
    ru, rv = orc.solve_flow(orc.flow_params(w, h, L, n_warp=nw, n_iter=ni), i0, i1, "f32")
    assert np.array_equal(got[0][:, :w], ru[:, :w]) and np.array_equal(got[1][:, :w], rv[:, :w])


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_demo_recon(orc, tmp_path, mode):
    w, h, L, nw, ni, seed = 160, 120, 3, 2, 20, 22
    out = _run(["recon", w, h, L, nw, ni, seed] + (["fused"] if mode == "fused" else []) + ["ply"], tmp_path)
    s = rf.align_up(w)
    raw = np.fromfile(out, np.float32)
    got = raw[:5 * h * s].reshape(5, h, s)
    rgb = raw[5 * h * s:].reshape(h, s, 3)
    d = rf.make_recon_pair(w, h, seed)
    p = orc.recon_params(w, h, L, n_warp=nw, n_iter=ni)
    ref = orc.solve_recon(p, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f32")
    for a, b in zip(got, ref):
        assert np.array_equal(a[:, :w], b[:, :w], equal_nan=True)
    assert (tmp_path / "out.bin.ply").read_text().startswith("ply\n")
    # copyOpticalFlowToHost(u, v, uvrgb) after solveR: the FlowToHSV colour map of the returned flow
    # (Reconflow.cu:460, flow scale 50 as the demo passes it), within 1e-5 (atan2f, sqrtf)
    ref_rgb = np.zeros((h + 1) * s * 3, np.float32)
    u = np.zeros((h + 1) * s, np.float32)
    v = np.zeros_like(u)
    u[:h * s], v[:h * s] = got[0].ravel(), got[1].ravel()
    orc.call("flow_to_hsv", "f32", u, v, w, h, s, ref_rgb, 50.0)
    ref_rgb = ref_rgb[:h * s * 3].reshape(h, s, 3)
    assert np.abs(ref_rgb[:, :w]).max() > 0.1
    np.testing.assert_allclose(rgb[:, :w], ref_rgb[:, :w], rtol=0, atol=1e-5)
