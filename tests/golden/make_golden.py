"""Writes tests/golden/golden.npz: small seeded inputs and the oracle's outputs for them.

Run from the repository root after building the oracle and the synthetic-pair library:
``python tests/golden/make_golden.py``. The fixture holds
  * per-kernel cases at 64x40 (stride 64, one guard row): SolveDataL1, SolveSmoothDualTVGlobal and
    the 5x5 median, with their FP32 and FP64 oracle outputs;
  * end-to-end cases at 96x72 (3 levels, 2 warps, 10 iterations): TV-L1 and the joint solver
    (FP32 and FP64-triangulation modes), inputs and the oracle's FP32 (and FP64) results.
The inputs are stored, so the tests that read the fixture do not depend on the generators.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402

W, H, S = 64, 40, 64


def flat(rng, lo, hi):
    a = np.zeros((H + 1) * S, np.float64)
    a.reshape(H + 1, S)[:H, :W] = rng.uniform(lo, hi, (H, W))
    return a.astype(np.float32)


def kernel_cases(out):
    rng = np.random.default_rng(101)
    n = (H + 1) * S
    ins = [flat(rng, -1, 1), flat(rng, -1, 1)] + [flat(rng, -0.5, 0.5) for _ in range(4)] + \
          [flat(rng, -0.2, 0.2), flat(rng, -0.2, 0.2), flat(rng, -0.1, 0.1)]
    out["data_l1_in"] = np.stack(ins)
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("data_l1", kind, *[x.astype(dt) for x in ins], W, H, S, 50.0, 0.3, *o)
        out[f"data_l1_out_{kind}"] = np.stack(o)

    ins = [flat(rng, -2, 2), flat(rng, -2, 2)] + [flat(rng, -0.7, 0.7) for _ in range(4)]
    out["dual_tv_in"] = np.stack(ins)
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt) for _ in range(4)]
        orc.call("dual_tv", kind, *[x.astype(dt) for x in ins], W, H, S, 0.25, 0.3, *o)
        out[f"dual_tv_out_{kind}"] = np.stack(o)

    ins = [flat(rng, -3, 3), flat(rng, -3, 3)]
    ins[0][5 * S + 7] = ins[0][5 * S + 8]  # ties
    out["median5_in"] = np.stack(ins)
    for kind, dt in (("f32", np.float32), ("f64", np.float64)):
        o = [np.zeros(n, dt), np.zeros(n, dt)]
        orc.call("median5", kind, *[x.astype(dt) for x in ins], W, H, S, *o)
        out[f"median5_out_{kind}"] = np.stack(o)


def pair_cases(out):
    import reconflow_amd as rf

    w, h = 96, 72
    i0, i1, _, _ = rf.make_flow_pair(w, h, 5)
    out["tvl1_i0"], out["tvl1_i1"] = i0, i1
    p = orc.flow_params(w, h, 3, n_warp=2, n_iter=10)
    for kind in ("f32", "f64"):
        u, v = orc.solve_flow(p, i0, i1, kind)
        out[f"tvl1_u_{kind}"], out[f"tvl1_v_{kind}"] = u, v

    d = rf.make_recon_pair(w, h, 6)
    for k in ("i0", "i1", "K", "R", "t"):
        out[f"joint_{k}"] = d[k]
    for tri in (0, 1):
        p = orc.recon_params(w, h, 3, n_warp=2, n_iter=10, tri_f64=tri)
        res = orc.solve_recon(p, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f32")
        for name, a in zip(("u", "v", "X", "Y", "Z"), res):
            out[f"joint_tri{tri}_{name}_f32"] = a


def main():
    out = {}
    kernel_cases(out)
    pair_cases(out)
    np.savez_compressed(Path(__file__).with_name("golden.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
