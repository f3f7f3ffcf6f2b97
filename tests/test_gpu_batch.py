"""Batched solve of uneven-size pairs (rf_batch_*): every pair of a batch equals its own
single-pair oracle solve bit for bit, with and without graph replay."""
import numpy as np
import pytest

import reconflow_amd as rf
from tests.gpu_util import dev, fetch

pytestmark = pytest.mark.gpu

SIZES = [(160, 120), (208, 136), (96, 80), (176, 104)]


def _dev_img(a):
    return dev(a.ravel())


def _fetch(ptr, n):
    return fetch(ptr, n)


@pytest.mark.parametrize("tile", [0, 1, 3, 4, 7], ids=["auto", "64x16-rpt4", "32x16", "32x16-k4", "32x16-k6"])
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_flow(orc, graph, tile):
    print(f"case: batch TV-L1 {len(SIZES)} pairs graph={graph} tile={tile}", flush=True)
    p = rf.batch_params(0, 3, 2, 10, use_graph=graph, flow_tile=tile)
    b = rf.Batch(p, [w for w, _ in SIZES], [h for _, h in SIZES])
    pairs, keep = [], []
    for i, (w, h) in enumerate(SIZES):
        i0, i1, _, _ = rf.make_flow_pair(w, h, 100 + i)
        d0, d1 = _dev_img(i0), _dev_img(i1)
        keep += [d0, d1]
        assert b.stride(i) == rf.align_up(w)
        b.set_images_device(i, d0.data_ptr(), d1.data_ptr())
        pairs.append((i0, i1))
    for rep in range(2):
        b.solve()
        b.synchronize()
        for i, (w, h) in enumerate(SIZES):
            s = rf.align_up(w)
            u_ptr, v_ptr = b.result_ptrs(i)[:2]
            u = _fetch(u_ptr, h * s).reshape(h, s)
            v = _fetch(v_ptr, h * s).reshape(h, s)
            ru, rv = orc.solve_flow(orc.flow_params(w, h, 3, n_warp=2, n_iter=10), *pairs[i], "f32")
            assert np.array_equal(u[:, :w], ru[:, :w]) and np.array_equal(v[:, :w], rv[:, :w]), (rep, i)


@pytest.mark.parametrize("tile", [0, 2, 5], ids=["auto", "32x16", "16x16-k2"])
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_recon(orc, graph, tile):
    print(f"case: batch joint {len(SIZES)} pairs graph={graph} tile={tile}", flush=True)
    p = rf.batch_params(14, 3, 2, 10, tau=0.125, use_graph=graph, joint_tile=tile)
    b = rf.Batch(p, [w for w, _ in SIZES], [h for _, h in SIZES])
    data, keep = [], []
    for i, (w, h) in enumerate(SIZES):
        d = rf.make_recon_pair(w, h, 200 + i)
        d0, d1 = _dev_img(d["i0"]), _dev_img(d["i1"])
        keep += [d0, d1]
        b.set_images_device(i, d0.data_ptr(), d1.data_ptr())
        b.set_camera(i, d["K"], d["K"], d["R"], d["t"])
        data.append(d)
    for rep in range(2):
        b.solve()
        b.synchronize()
        for i, (w, h) in enumerate(SIZES):
            s = rf.align_up(w)
            got = [_fetch(ptr, h * s).reshape(h, s) for ptr in b.result_ptrs(i)]
            d = data[i]
            pr = orc.recon_params(w, h, 3, n_warp=2, n_iter=10)
            ref = orc.solve_recon(pr, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f32")
            for a, r in zip(got, ref):
                assert np.array_equal(a[:, :w], r[:, :w], equal_nan=True), (rep, i)


@pytest.mark.parametrize("tile", [0, 5], ids=["auto", "16x16-k2"])
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_recon_same_size(orc, graph, tile):
    # pairs of one size run the one-pair kernels with the pair taken from the dispatch index
    w, h, n = 144, 104, 3
    print(f"case: batch joint {n} pairs of {w}x{h} graph={graph} tile={tile}", flush=True)
    p = rf.batch_params(14, 3, 2, 7, tau=0.125, use_graph=graph, joint_tile=tile, tri_f64=1)
    b = rf.Batch(p, [w] * n, [h] * n)
    data, keep = [], []
    for i in range(n):
        d = rf.make_recon_pair(w, h, 300 + i)
        d0, d1 = _dev_img(d["i0"]), _dev_img(d["i1"])
        keep += [d0, d1]
        b.set_images_device(i, d0.data_ptr(), d1.data_ptr())
        b.set_camera(i, d["K"], d["K"], d["R"], d["t"])
        data.append(d)
    s = rf.align_up(w)
    for rep in range(2):
        b.solve()
        b.synchronize()
        for i in range(n):
            got = [_fetch(ptr, h * s).reshape(h, s) for ptr in b.result_ptrs(i)]
            d = data[i]
            pr = orc.recon_params(w, h, 3, n_warp=2, n_iter=7, tri_f64=1)
            ref = orc.solve_recon(pr, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f32")
            for a, r in zip(got, ref):
                assert np.array_equal(a[:, :w], r[:, :w], equal_nan=True), (rep, i)


@pytest.mark.parametrize("tile", [0, 4], ids=["auto", "32x16-k4"])
def test_batch_flow_same_size(orc, tile):
    w, h, n = 144, 104, 3
    print(f"case: batch TV-L1 {n} pairs of {w}x{h} tile={tile}", flush=True)
    p = rf.batch_params(0, 3, 2, 9, use_graph=True, flow_tile=tile)
    b = rf.Batch(p, [w] * n, [h] * n)
    pairs, keep = [], []
    for i in range(n):
        i0, i1, _, _ = rf.make_flow_pair(w, h, 400 + i)
        d0, d1 = _dev_img(i0), _dev_img(i1)
        keep += [d0, d1]
        b.set_images_device(i, d0.data_ptr(), d1.data_ptr())
        pairs.append((i0, i1))
    s = rf.align_up(w)
    for rep in range(2):
        b.solve()
        b.synchronize()
        for i in range(n):
            u_ptr, v_ptr = b.result_ptrs(i)[:2]
            u = _fetch(u_ptr, h * s).reshape(h, s)
            v = _fetch(v_ptr, h * s).reshape(h, s)
            ru, rv = orc.solve_flow(orc.flow_params(w, h, 3, n_warp=2, n_iter=9), *pairs[i], "f32")
            assert np.array_equal(u[:, :w], ru[:, :w]) and np.array_equal(v[:, :w], rv[:, :w]), (rep, i)


def _solve_single(p, w, h, setup):
    """One pair alone in a batch of one: the reference result of a pair of a batch."""
    b = rf.Batch(p, [w], [h])
    keep = setup(b, 0)
    b.solve()
    b.synchronize()
    s = rf.align_up(w)
    out = [_fetch(ptr, h * s).reshape(h, s) for ptr in b.result_ptrs(0) if ptr]
    del keep
    return out


@pytest.mark.parametrize("kind", ["joint-f16-derivatives", "joint-f16-derivatives-duals", "tvl1-f16"])
def test_batch_same_size_fp16_pairs_equal_single(kind):
    # 16-bit fields of later pairs sit at the pair offset counted in 16-bit elements: every pair of
    # a same-size FP16 batch equals the same pair solved alone, bit for bit
    w, h, n = 144, 104, 3
    print(f"case: batch {kind} {n} pairs of {w}x{h}", flush=True)
    joint = kind.startswith("joint")
    if joint:
        storage = rf.STORAGE_F16_DERIV_DUALS if kind.endswith("duals") else rf.STORAGE_F16_DERIV
        p = rf.batch_params(14, 3, 2, 7, tau=0.125, use_graph=True, tri_f64=1, storage=storage)
        data = [rf.make_recon_pair(w, h, 500 + i) for i in range(n)]
    else:
        p = rf.batch_params(0, 3, 2, 9, use_graph=True, storage=rf.STORAGE_F16)
        data = [dict(zip(("i0", "i1"), rf.make_flow_pair(w, h, 600 + i)[:2])) for i in range(n)]

    def setup_for(i):
        def setup(b, j):
            d = data[i]
            d0, d1 = _dev_img(d["i0"]), _dev_img(d["i1"])
            b.set_images_device(j, d0.data_ptr(), d1.data_ptr())
            if joint:
                b.set_camera(j, d["K"], d["K"], d["R"], d["t"])
            return d0, d1
        return setup

    b = rf.Batch(p, [w] * n, [h] * n)
    keep = [setup_for(i)(b, i) for i in range(n)]
    b.solve()
    b.synchronize()
    s = rf.align_up(w)
    for i in range(n):
        got = [_fetch(ptr, h * s).reshape(h, s) for ptr in b.result_ptrs(i) if ptr]
        ref = _solve_single(p, w, h, setup_for(i))
        assert len(got) == len(ref)
        for a, r in zip(got, ref):
            assert np.array_equal(a[:, :w], r[:, :w], equal_nan=True), i
    del keep
