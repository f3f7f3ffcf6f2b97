"""Learned warm start for the joint solver (the rf_warmnet_* entry points of include/rfhip.h).

A six-layer 3x3 convolutional encoder-decoder (32 channels, BF16 weights and activations on the
matrix cores, FP32 accumulation) predicts the flow at pyramid level 3 (1/8 of the image) from
levels 3 and 4 of both frames. Its prediction, scaled up to level 0, is handed to the solver as a
warm start (``CudaFlow.copyWarmStartToDevice`` / ``Batch.set_warm_start_device``); the solver's
own rule falls back to a cold start when the prediction is not usable (non-finite or larger than
the image). The network is trained on the joint solver's results on synthetic pairs
(``reconflow_amd.train_warmnet``).

Parameter layout (flat FP32): per layer the weights [9][cout][cin] (tap-major, 3x3 row-major
taps) then the bias [cout].
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native
from ._native import check

# (input channels, output channels, stride) of the six convolutions
LAYERS = [(4, 32, 1), (32, 32, 2), (32, 32, 1), (32, 32, 1), (32, 32, 1), (32, 2, 1)]
WEIGHTS = Path(__file__).resolve().parent / "weights"


def param_count() -> int:
    return sum(9 * co * ci + co for ci, co, _ in LAYERS)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in 9 cin), zero biases; the last layer scaled down so the initial
    prediction is close to zero flow."""
    rng = np.random.default_rng(seed)
    out = []
    for k, (ci, co, _) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (9 * ci)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
        out.append(rng.normal(0.0, std, 9 * co * ci).astype(np.float32))
        out.append(np.zeros(co, np.float32))
    return np.concatenate(out)


def split_params(p: np.ndarray):
    """[(weights [9, cout, cin], bias [cout])] per layer from the flat vector."""
    out, o = [], 0
    for ci, co, _ in LAYERS:
        w = p[o:o + 9 * co * ci].reshape(9, co, ci)
        o += 9 * co * ci
        b = p[o:o + co]
        o += co
        out.append((w, b))
    return out


class WarmStartNet:
    """The network for level-3 images of height x width (both even) and up to max_batch samples."""

    def __init__(self, height: int, width: int, max_batch: int = 1, params: np.ndarray | None = None):
        self._lib = _native.lib()
        self._h = C.c_void_p()
        check(self._lib.rf_warmnet_create(C.byref(self._h), max_batch, height, width), "rf_warmnet_create")
        self.height, self.width, self.max_batch = height, width, max_batch
        n = self._lib.rf_warmnet_param_count(self._h)
        assert n == param_count(), (n, param_count())
        self.set_params(init_params() if params is None else params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.rf_warmnet_destroy(h)
            self._h = None

    # ---- parameters --------------------------------------------------------------------
    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(self._lib.rf_warmnet_set_params(self._h, p.ctypes.data, p.size), "rf_warmnet_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(param_count(), np.float32)
        check(self._lib.rf_warmnet_get_params(self._h, p.ctypes.data, p.size), "rf_warmnet_get_params")
        return p

    def get_grads(self) -> np.ndarray:
        g = np.empty(param_count(), np.float32)
        check(self._lib.rf_warmnet_get_grads(self._h, g.ctypes.data, g.size), "rf_warmnet_get_grads")
        return g

    def save(self, path) -> None:
        np.savez(path, params=self.get_params(), height=self.height, width=self.width)

    @classmethod
    def load(cls, path, max_batch: int = 1) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            return cls(int(z["height"]), int(z["width"]), max_batch, z["params"])

    # ---- compute (device pointers as ints; stream a hipStream_t as int, 0 = null) -------
    def set_input(self, d_i0_l3, d_i1_l3, s3, nstride3, d_i0_l4, d_i1_l4, s4, nstride4, n, stream=0) -> None:
        check(self._lib.rf_warmnet_input(self._h, d_i0_l3, d_i1_l3, s3, nstride3, d_i0_l4, d_i1_l4, s4, nstride4, n,
                                         stream or None), "rf_warmnet_input")

    def forward(self, n: int, d_flow: int = 0, stream=0) -> None:
        check(self._lib.rf_warmnet_forward(self._h, n, d_flow or None, stream or None), "rf_warmnet_forward")

    def train_step(self, n: int, d_target: int, lr: float, stream=0) -> float:
        loss = C.c_float()
        check(self._lib.rf_warmnet_train_step(self._h, n, d_target, float(lr), C.byref(loss), stream or None),
              "rf_warmnet_train_step")
        return float(loss.value)

    def warm_start(self, d_i0: int, d_i1: int, w0: int, h0: int, s0: int, d_u: int, d_v: int, stream=0) -> None:
        """Level-0 warm start (d_u, d_v: stride s0 x h0 floats) of the pair d_i0, d_i1."""
        check(self._lib.rf_warmnet_warm_start(self._h, d_i0, d_i1, w0, h0, s0, d_u, d_v, stream or None),
              "rf_warmnet_warm_start")
