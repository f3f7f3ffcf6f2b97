"""Trains the warm-start network (reconflow_amd.warmnet) on the joint solver's own results.

    python -m reconflow_amd.train_warmnet --pairs 20000 --epochs 8 \\
        --out reconflow_amd/weights/warmnet_640x480.npz

Data: seeded 640x480 joint-mode pairs of the synthetic generator (8 random planes at depths 2-10,
rotation <= 3 deg, translation 0.1), generated on the host by a process pool. Each pair is solved on
the GPU by the batched joint solver (5 levels x 5 warps x 50 iterations, FP64 mode, FP16 derivatives,
K[2][2] = 1 on the coarse levels, eight pairs per launch); the target is the solved flow downscaled to
level 3 by the prior-flow operator (DownscaleScalingKernel, downscale.cu:28-48). The inputs are
levels 3 and 4 of both frames (the solver's downscale). Pairs whose solved flow is not finite are
dropped. Training: Huber loss, Adam, mini-batches of 64, a held-out validation share; every epoch
prints its losses.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import time
from pathlib import Path

import numpy as np

W0, H0, LEVELS = 640, 480, 5


def geometry(w=W0, h=H0, levels=LEVELS):
    W, H, S = [w], [h], [(w + 31) // 32 * 32]
    for _ in range(1, levels):
        W.append(int(np.float32(W[-1]) / np.float32(2.0)))
        H.append(int(np.float32(H[-1]) / np.float32(2.0)))
        S.append((W[-1] + 31) // 32 * 32)
    return W, H, S


def _gen(seed):
    import reconflow_amd as rf
    d = rf.make_recon_pair(W0, H0, seed)
    return seed, d["i0"], d["i1"], d["K"], d["R"], d["t"]


def build_dataset(n_pairs, seed0, workers, log):
    """Solves n_pairs seeded pairs and returns device tensors X3 (N, 2, h3, s3), X4 (N, 2, h4, s4)
    and T (N, h3, w3, 2) of the pairs whose solved flow is finite."""
    import torch

    import reconflow_amd as rf
    from reconflow_amd import _native

    lib = _native.lib()
    W, H, S = geometry()
    B = 8
    p = rf.batch_params(14, LEVELS, 5, 50, lam=50.0, tau=0.125, alpha_tv=33.3, alpha_proj=60.0, lambdaf=0.1,
                        lambdams=100.0, tri_f64=1, storage=rf.STORAGE_F16_DERIV, fix_coarse_k=1)
    batch = rf.Batch(p, [W0] * B, [H0] * B)
    dev = "cuda"
    X3 = torch.zeros((n_pairs, 2, H[3], S[3]), dtype=torch.float32, device=dev)
    X4 = torch.zeros((n_pairs, 2, H[4], S[4]), dtype=torch.float32, device=dev)
    T = torch.zeros((n_pairs, H[3], W[3], 2), dtype=torch.float32, device=dev)
    keep = torch.zeros(n_pairs, dtype=torch.bool)
    lvl = [torch.zeros((2, H[l] + 1, S[l]), dtype=torch.float32, device=dev) for l in range(LEVELS)]
    u3 = torch.zeros((2, H[3] + 1, S[3]), dtype=torch.float32, device=dev)
    t0 = time.time()
    seeds = list(range(seed0, seed0 + n_pairs))
    chunk = 256
    with mp.get_context("spawn").Pool(workers) as pool:
        for c0 in range(0, n_pairs, chunk):
            pairs = pool.map(_gen, seeds[c0:c0 + chunk])
            for g0 in range(0, len(pairs), B):
                grp = pairs[g0:g0 + B]
                imgs = []
                for i, (_, i0, i1, K, R, t) in enumerate(grp):
                    a = torch.from_numpy(i0).to(dev)
                    b = torch.from_numpy(i1).to(dev)
                    imgs.append((a, b))
                    batch.set_images_device(i, a.data_ptr(), b.data_ptr())
                    batch.set_camera(i, K, K, R, t)
                for i in range(len(grp), B):  # a short last group: repeat its first pair
                    batch.set_images_device(i, imgs[0][0].data_ptr(), imgs[0][1].data_ptr())
                    batch.set_camera(i, grp[0][3], grp[0][3], grp[0][4], grp[0][5])
                batch.solve()
                batch.synchronize()
                torch.cuda.synchronize()
                for i in range(len(grp)):
                    k = c0 + g0 + i
                    for f in range(2):
                        lvl[0][f, :H[0]].copy_(imgs[i][f])
                        for l in range(1, LEVELS):
                            assert lib.rf_k_downscale(lvl[l - 1][f].data_ptr(), W[l - 1], H[l - 1], S[l - 1], W[l],
                                                      H[l], S[l], lvl[l][f].data_ptr(), None) == 0
                    u_ptr, v_ptr = batch.result_ptrs(i)[:2]
                    sc = float(np.float32(W[3]) / np.float32(W[0]))
                    for f, ptr in enumerate((u_ptr, v_ptr)):
                        assert lib.rf_k_downscale_scaled(ptr, W[0], H[0], S[0], W[3], H[3], S[3], sc,
                                                         u3[f].data_ptr(), None) == 0
                    torch.cuda.synchronize()
                    X3[k] = lvl[3][:, :H[3]]
                    X4[k] = lvl[4][:, :H[4]]
                    T[k] = u3[:, :H[3], :W[3]].permute(1, 2, 0)
                    keep[k] = bool(torch.isfinite(T[k]).all())
            log(f"dataset: {min(c0 + chunk, n_pairs)}/{n_pairs} pairs solved ({time.time() - t0:.0f} s)")
    idx = torch.nonzero(keep).flatten().to(dev)
    return X3[idx].contiguous(), X4[idx].contiguous(), T[idx].contiguous()


def train(net, X3, X4, T, epochs, batch, lr, val_frac, log, seed=0):
    import torch

    n = X3.shape[0]
    nval = max(1, int(n * val_frac))
    ntr = n - nval
    g = torch.Generator(device="cpu").manual_seed(seed)
    h3, s3, h4, s4 = X3.shape[2], X3.shape[3], X4.shape[2], X4.shape[3]
    hist = []

    def run(ids, rate):
        x3, x4, tg = X3[ids].contiguous(), X4[ids].contiguous(), T[ids].contiguous()
        m = len(ids)
        net.set_input(x3[:, 0].data_ptr(), x3[:, 1].data_ptr(), s3, 2 * h3 * s3, x4[:, 0].data_ptr(),
                      x4[:, 1].data_ptr(), s4, 2 * h4 * s4, m)
        return net.train_step(m, tg.data_ptr(), rate)

    for ep in range(epochs):
        perm = torch.randperm(ntr, generator=g)
        rate = lr * (0.5 ** (ep // max(1, epochs // 3)))  # step decay: thirds of the run
        tl = []
        for b0 in range(0, ntr - batch + 1, batch):
            ids = perm[b0:b0 + batch].to(X3.device)
            tl.append(run(ids, rate))
        vl = []
        for b0 in range(ntr, n, batch):
            ids = torch.arange(b0, min(b0 + batch, n), device=X3.device)
            vl.append(run(ids, 0.0) * len(ids))
        val = float(np.sum(vl) / nval)
        hist.append({"epoch": ep, "lr": rate, "train_loss": float(np.mean(tl)), "val_loss": val})
        log(f"epoch {ep}: lr {rate:.2e} train {np.mean(tl):.4f} val {val:.4f}")
    return hist


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--seed0", type=int, default=100000, help="first generator seed of the training pairs")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--val", type=float, default=0.05)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "weights" / "warmnet_640x480.npz"))
    ap.add_argument("--report", default=None, help="JSON file for the loss history")
    args = ap.parse_args()

    import torch

    from reconflow_amd.warmnet import WarmStartNet, init_params

    assert torch.cuda.is_available(), "training runs on the GPU"

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    X3, X4, T = build_dataset(args.pairs, args.seed0, args.workers, log)
    log(f"dataset: {X3.shape[0]} usable pairs of {args.pairs} in {time.time() - t0:.0f} s")
    W, H, _ = geometry()
    net = WarmStartNet(H[3], W[3], args.batch, init_params(0))
    t1 = time.time()
    hist = train(net, X3, X4, T, args.epochs, args.batch, args.lr, args.val, log)
    log(f"training: {time.time() - t1:.0f} s")
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    net.save(args.out)
    log(f"saved {args.out}")
    if args.report:
        Path(args.report).write_text(json.dumps({"pairs": args.pairs, "usable": int(X3.shape[0]), "epochs": hist,
                                                 "seed0": args.seed0}, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
