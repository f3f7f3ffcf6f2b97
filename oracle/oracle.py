"""ctypes wrapper of the CPU oracle (oracle.cpp). TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the bench's CPU-baseline leg import this module; it is the
checker, never the thing measured or shipped. ``kind`` selects the arithmetic: ``"f64"`` (the
FP64 oracle) or ``"f32"`` (the same restatement in FP32, reference operation order).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class Params(C.Structure):
    """orc::Params (oracle.cpp)."""

    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("n_levels", C.c_int), ("scale", C.c_float),
        ("method", C.c_int), ("lambda_", C.c_float), ("lambdagrad", C.c_float), ("theta", C.c_float),
        ("tau", C.c_float), ("n_warp_iters", C.c_int), ("n_solver_iters", C.c_int), ("lambdaf", C.c_float),
        ("lambdams", C.c_float), ("alpha_tv", C.c_float), ("alpha_proj", C.c_float), ("alpha_fn", C.c_float),
        ("stride_alignment", C.c_int), ("input_type", C.c_int), ("fix_coarse_k", C.c_int),
        ("tri_f64", C.c_int),
    ]


def build() -> Path:
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True, capture_output=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        assert _lib.orc_params_size() == C.sizeof(Params), "oracle Params layout mismatch"
    return _lib


def _np_t(kind):
    return np.float64 if kind == "f64" else np.float32


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _fn(name, kind):
    f = getattr(lib(), f"orc_{name}_{kind}")
    return f


def _ints(*vals):
    return [C.c_int(int(v)) for v in vals]


def _fl(*vals):
    return [C.c_float(float(v)) for v in vals]


def _field(shape_len, kind):
    return np.zeros(shape_len, _np_t(kind))


# ---- per-kernel calls: every array is a flat buffer in the kind's dtype ----------------
def call(name: str, kind: str, *args):
    """Calls orc_<name>_<kind> with numpy arrays (pointers), ints and floats as given.
    Python floats become C float, Python ints C int, numpy arrays pointers, and
    ``('double*', arr)`` tuples double pointers."""
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(_p(a))
        elif isinstance(a, bool):
            cargs.append(C.c_int(int(a)))
        elif isinstance(a, (int, np.integer)):
            cargs.append(C.c_long(int(a)) if name == "add" and len(cargs) == 2 else C.c_int(int(a)))
        elif isinstance(a, (float, np.floating)):
            cargs.append(C.c_float(float(a)))
        elif a is None:
            cargs.append(C.c_void_p(0))
        else:
            raise TypeError(type(a))
    f = _fn(name, kind)
    f.restype = C.c_int
    return f(*cargs)


def set_threads(n: int) -> None:
    """OpenMP threads of the restatement's row loops."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def pyramid(p: Params):
    n = p.n_levels
    W, H, S = (C.c_int * n)(), (C.c_int * n)(), (C.c_int * n)()
    lib().orc_pyramid(C.byref(p), W, H, S)
    return list(W), list(H), list(S)


def _ws(ws):
    """A warm start (u, v), level-0 (h, stride) float32 planes, as two pointers (or two NULLs)."""
    if ws is None:
        return [None, None], []
    keep = [np.ascontiguousarray(a, np.float32) for a in ws]
    return [_p(a) for a in keep], keep


def warm_start_usable(wsu, wsv, w, h) -> bool:
    """The warm-start rule: every component over the w x h pixels finite and at most max(w, h)."""
    a, b = (np.ascontiguousarray(x, np.float32) for x in (wsu, wsv))
    return bool(lib().orc_warm_start_usable(_p(a), _p(b), int(w), int(h), int(a.shape[1])))


def solve_flow(p: Params, i0: np.ndarray, i1: np.ndarray, kind: str = "f64", ws=None):
    """Whole CudaFlow TV-L1 solve. Images are (h, stride) float32 (or uint8 rgb). ws: an optional
    warm start (u, v) that replaces the zero initial flow at the coarsest level."""
    s0 = pyramid(p)[2][0]
    u = np.zeros((p.height, s0), np.float32)
    v = np.zeros_like(u)
    f = getattr(lib(), f"orc_solve_flow_{kind}")
    f.restype = C.c_int
    wsp, keep = _ws(ws)
    rc = f(C.byref(p), _p(np.ascontiguousarray(i0)), _p(np.ascontiguousarray(i1)), *wsp, _p(u), _p(v))
    assert rc == 0
    del keep
    return u, v


def solve_flow_masked(p: Params, i0: np.ndarray, i1: np.ndarray, mask0: np.ndarray, kind: str = "f64"):
    """CudaFlow::solveInpainting: the TV-L1 solve with the masked L1 step (mask0 (h, stride) float32)."""
    s0 = pyramid(p)[2][0]
    u = np.zeros((p.height, s0), np.float32)
    v = np.zeros_like(u)
    f = getattr(lib(), f"orc_solve_flow_masked_{kind}")
    f.restype = C.c_int
    m = np.ascontiguousarray(mask0, dtype=np.float32)
    rc = f(C.byref(p), _p(np.ascontiguousarray(i0)), _p(np.ascontiguousarray(i1)), _p(m), _p(u), _p(v))
    assert rc == 0
    return u, v


def solve_recon(p: Params, K0, K1, R0, t0, R1, t1, i0, i1, kind="f64", ufn=None, vfn=None, mask=None, ws=None):
    s0 = pyramid(p)[2][0]
    outs = [np.zeros((p.height, s0), np.float32) for _ in range(5)]
    dd = [np.ascontiguousarray(np.asarray(m, np.float64).ravel()) for m in (K0, K1, R0, t0, R1, t1)]
    f = getattr(lib(), f"orc_solve_recon_{kind}")
    f.restype = C.c_int
    optk = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (ufn, vfn, mask)]
    opt = [None if a is None else _p(a) for a in optk]
    wsp, keep = _ws(ws)
    rc = f(C.byref(p), *[_p(d) for d in dd], _p(np.ascontiguousarray(i0)), _p(np.ascontiguousarray(i1)),
           *opt, *wsp, *[_p(o) for o in outs])
    assert rc == 0
    del keep, optk
    return tuple(outs)


def flow_params(w, h, n_levels, scale=2.0, n_warp=3, n_iter=30, lam=50.0, theta=0.3, tau=0.25, input_type=0,
                method=0, lambdagrad=400.0):
    p = Params()
    p.width, p.height, p.n_levels, p.scale, p.method = w, h, n_levels, scale, method
    p.lambda_, p.lambdagrad, p.theta, p.tau = lam, lambdagrad, theta, tau
    p.n_warp_iters, p.n_solver_iters = n_warp, n_iter
    p.stride_alignment, p.input_type = 32, input_type
    return p


def recon_params(w, h, n_levels, scale=2.0, n_warp=5, n_iter=50, lam=50.0, tau=0.125, alpha_tv=33.3,
                 alpha_proj=60.0, lambdaf=0.1, lambdams=100.0, alpha_fn=100.0, method=14, input_type=0,
                 fix_coarse_k=0, tri_f64=0):
    p = Params()
    p.width, p.height, p.n_levels, p.scale, p.method = w, h, n_levels, scale, method
    p.lambda_, p.lambdagrad, p.theta, p.tau = lam, 0.0, float(np.float32(1.0) / np.float32(alpha_tv)), tau
    p.n_warp_iters, p.n_solver_iters = n_warp, n_iter
    p.lambdaf, p.lambdams, p.alpha_tv, p.alpha_proj, p.alpha_fn = lambdaf, lambdams, alpha_tv, alpha_proj, alpha_fn
    p.stride_alignment, p.input_type, p.fix_coarse_k, p.tri_f64 = 32, input_type, fix_coarse_k, tri_f64
    return p
