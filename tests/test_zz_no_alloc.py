"""rf_solve_flow_noalloc (the allocation-free TV-L1 solve) against the context entry points
(rf_initialize + rf_copy_images_device + rf_solve_flow on the plain path): bit for bit on two
calls in a row, the second smaller than the first, and on a third call on a non-default stream;
solves that do not fit the static scratch are rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import reconflow_amd as rf
from reconflow_amd import _native
from tests.gpu_util import dev, dzeros, host, torch_cuda

pytestmark = pytest.mark.gpu

CASES = [(320, 240, 4, 3, 30, 0), (200, 152, 3, 3, 20, 1), (256, 192, 4, 2, 15, 2)]


def _params(w, h, L, nw, ni):
    p = _native.RFParams()
    _native.lib().rf_params_default(C.byref(p), w, h)
    p.input_type, p.n_levels, p.scale, p.method = rf.INPUT_32F, L, 2.0, 0
    p.lambda_, p.theta, p.tau = 50.0, 0.3, 0.25
    p.n_warp_iters, p.n_solver_iters = nw, ni
    return p


def _context_solve(i0, i1, w, h, L, nw, ni):
    f = rf.CudaFlow()
    f.initialize(w, h, 1, rf.INPUT_32F, L, 2.0, 0, 50.0, 0.0, 0.3, 0.25, nw, ni, False, path=rf.PATH_PLAIN)
    d0, d1 = dev(i0.ravel()), dev(i1.ravel())
    f.copyImagesToDeviceFromDevice(d0.data_ptr(), d1.data_ptr())
    f.solveOpticalFlow(5.0)
    return f.copyOpticalFlowToHost()


def test_noalloc_equals_context_solve(orc):
    torch = torch_cuda()
    lib = _native.lib()
    side = torch.cuda.Stream()
    for k, (w, h, L, nw, ni, seed) in enumerate(CASES):
        print(f"case: rf_solve_flow_noalloc call {k} {w}x{h} L={L} warps={nw} iters={ni}", flush=True)
        i0, i1, _, _ = rf.make_flow_pair(w, h, 70 + seed)
        s = i0.shape[1]
        d0, d1 = dev(i0.ravel()), dev(i1.ravel())
        du, dv = dzeros(h * s), dzeros(h * s)
        du.fill_(np.nan)
        dv.fill_(np.nan)
        torch.cuda.synchronize()
        p = _params(w, h, L, nw, ni)
        stream = side.cuda_stream if k == 2 else None
        rc = lib.rf_solve_flow_noalloc(C.byref(p), d0.data_ptr(), d1.data_ptr(), du.data_ptr(), dv.data_ptr(), stream)
        assert rc == 0, _native.last_error()
        if stream is not None:
            side.synchronize()
        u = host(du).reshape(h, s)
        v = host(dv).reshape(h, s)
        cu, cv = _context_solve(i0, i1, w, h, L, nw, ni)
        assert np.array_equal(u[:, :w], cu[:, :w]) and np.array_equal(v[:, :w], cv[:, :w]), k
        ru, rv = orc.solve_flow(orc.flow_params(w, h, L, n_warp=nw, n_iter=ni), i0, i1, "f32")
        assert np.array_equal(u[:, :w], ru[:, :w]) and np.array_equal(v[:, :w], rv[:, :w]), k


def test_noalloc_rejects_what_does_not_fit():
    lib = _native.lib()
    d = dzeros(64)
    for w, h, L, what in ((1024, 576, 5, "scratch"), (40000, 8, 1, "below 32768")):
        p = _params(w, h, L, 1, 1)
        rc = lib.rf_solve_flow_noalloc(C.byref(p), d.data_ptr(), d.data_ptr(), d.data_ptr(), d.data_ptr(), None)
        assert rc == -1, rc  # RF_ERR_ARG, before any launch
        assert what in _native.last_error(), _native.last_error()
    p = _params(64, 48, 3, 1, 1)
    p.method = 14
    assert lib.rf_solve_flow_noalloc(C.byref(p), d.data_ptr(), d.data_ptr(), d.data_ptr(), d.data_ptr(), None) == -5
    p.method = 0
    assert lib.rf_solve_flow_noalloc(C.byref(p), None, d.data_ptr(), d.data_ptr(), d.data_ptr(), None) == -1
