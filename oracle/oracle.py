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
        ("tri_f64", C.c_int), ("mutation", C.c_int), ("round_mask", C.c_int),
        ("f32_steps", C.c_int), ("proj_f64", C.c_int), ("storage", C.c_int),
    ]


# Planted schedule mistakes (Params.mutation; oracle.cpp MUT_*), for tests that show a check
# detects them. Never reference semantics.
MUT = dict(flow_no_warp_reset=1, flow_no_median=2, flow_no_dual_swap=3, flow_upscale_no_scale=4,
           flow_warp_zero_flow=5, flow_no_uhat_swap=6,
           recon_no_uv_median=11, recon_sk_reset=12, recon_no_uc_upscale=13, recon_no_sk_swap=14,
           recon_no_cleanup=15, recon_no_proj_upscale=16, recon_no_warp_reset=17, recon_xyz_level_reset=18,
           recon_uc_level0_k=19)
# State groups of Params.round_mask (oracle.cpp RND_*): FP64 arithmetic with that group stored in FP32.
RND = dict(img=1, der=2, uhat=4, dual=8, proj=16, xyz=32, uc=64, sk=128, uv=256)
# Steps of Params.f32_steps (oracle.cpp STEP_*): the FP64 joint solve with that step in FP32 arithmetic.
# Precise-projection steps of Params.proj_f64 (oracle.cpp PREC_*), evaluated in binary64 in the FP32 solve.
PREC = dict(proj=1, blend=2, sk=4, data=8)
STEP = dict(warp=1, deriv=2, data=4, dual=8, proj=16, tri=32, clean=64, uc=128, sk=256, med=512, ups=1024)


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
        _lib.orc_params_size.restype = C.c_int
        _lib.orc_params_size.argtypes = []
        assert _lib.orc_params_size() == C.sizeof(Params), "oracle Params layout mismatch"
    return _lib


def _np_t(kind):
    return np.float64 if kind == "f64" else np.float32


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# Signatures of every orc_* symbol (oracle.cpp, extern "C"). Codes: T = pointer to the kind's
# float type (float64 for _f64, float32 for _f32), D = pointer to float64, F = pointer to float32,
# Q = optional T pointer (None allowed), B = pointer to uint8, W = pointer to uint16,
# i = int, l = long, f = float. A call with the wrong arity, a wrong scalar type or an array of
# the wrong dtype or layout raises TypeError before anything reaches the library.
_T9, _T6 = "T" * 9, "T" * 6
SIGS = {
    "gray_rgb": ("BTiii", None), "gray_u8": ("BTiii", None), "gray_u16": ("WTiii", None),
    "downscale": ("Tiiiiii" + "T", None),
    "downscale_scaled": ("Tiiiiiif" + "T", None),
    "upscale": ("Tiiiiiif" + "T", None),
    "warp": ("TiiiTTT", None),
    "derivatives": ("TTiiiTTT", None),
    "data_l1": (_T9 + "iiiffTT", None),
    "dual_tv": (_T6 + "iiiffTTTT", None),
    "median5": ("TTiiiTT", None),
    "median3d": ("TTTiiiTTT", None),
    "add": ("TTlT", None),
    "recon_data_l1": ("T" * 15 + "iiifffTT", None),
    "recon_data_l1_fn": ("T" * 17 + "iiiffffQTT", None),
    "uvproj": ("T" * 8 + "iiiffTT", None),
    "solve3d": ("TTTTTDDffiiiTTT", None),
    "solve3d_fma": ("TTTTTDDffiiiTTT", None),
    "cleanup3d": (_T6 + "iiiTTT", None),
    "solve_uc": ("TTTDDDiiiTT", None),
    "update_sk": ("T" * 8 + "iiiTT", None),
    "scene_flow_to_hsv": ("TiiiTf", None),
    "flow_to_hsv": ("TTiiiTf", None),
    "remove_edge": ("TTTifiiiTTT", None),
    "deriv_single": ("TiiiTT", None),
    "derivatives_patch": ("TTiiiTTT", None),
    "data_charb": ("T" * 11 + "Q" * 6 + "iiifffTTTT", None),
    "data_l1_inpaint": ("T" * 10 + "iiiffTT", None),
    "scene_flow": ("TTTTffffiiiTTT", None),
}
_CT = {"T": C.c_void_p, "D": C.c_void_p, "F": C.c_void_p, "Q": C.c_void_p, "B": C.c_void_p, "W": C.c_void_p,
       "i": C.c_int, "l": C.c_long, "f": C.c_float}
_DT = {"D": np.float64, "F": np.float32, "B": np.uint8, "W": np.uint16}
_bound = {}


def _bind(sym: str, codes: str, restype):
    f = _bound.get(sym)
    if f is None:
        f = getattr(lib(), sym)
        f.argtypes = [_CT[c] for c in codes]
        f.restype = restype
        _bound[sym] = f
    return f


def _conv(code: str, a, kind: str, pos: int, name: str):
    where = f"orc_{name}_{kind} argument {pos}"
    if code in "if":
        if isinstance(a, (bool, np.bool_)) or not isinstance(a, (int, float, np.integer, np.floating)):
            raise TypeError(f"{where}: expected a number, got {type(a).__name__}")
        if code == "i":
            if isinstance(a, (float, np.floating)):
                raise TypeError(f"{where}: expected an int, got a float")
            return int(a)
        return float(a)
    if code == "l":
        if not isinstance(a, (int, np.integer)) or isinstance(a, bool):
            raise TypeError(f"{where}: expected an int count")
        return int(a)
    if a is None:
        if code == "Q":
            return None
        raise TypeError(f"{where}: expected an array, got None")
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{where}: expected a numpy array, got {type(a).__name__}")
    want = _np_t(kind) if code in "TQ" else _DT[code]
    if a.dtype != want:
        raise TypeError(f"{where}: expected {np.dtype(want).name}, got {a.dtype.name}")
    if not a.flags.c_contiguous:
        raise TypeError(f"{where}: array must be C-contiguous")
    return _p(a)


def call(name: str, kind: str, *args):
    """Calls orc_<name>_<kind> (a per-kernel restatement) through its declared signature.
    Arrays are flat buffers: those of the kind's float type, plus float64 matrices (P, Q, R, t,
    K) and uint8 / uint16 inputs where the function takes them."""
    if name not in SIGS:
        raise KeyError(f"unknown oracle function {name!r}")
    if kind not in ("f64", "f32"):
        raise ValueError(kind)
    codes, restype = SIGS[name]
    if len(args) != len(codes):
        raise TypeError(f"orc_{name}_{kind} takes {len(codes)} arguments, got {len(args)}")
    cargs = [_conv(c, a, kind, i, name) for i, (c, a) in enumerate(zip(codes, args))]
    return _bind(f"orc_{name}_{kind}", codes, restype)(*cargs)


_VP, _I, _F = C.c_void_p, C.c_int, C.c_float
_PP = C.POINTER(Params)
# Whole-solve and helper symbols: (argtypes, restype).
_SOLVE_SIGS = {
    "orc_params_size": ([], _I),
    "orc_set_threads": ([_I], None),
    "orc_get_threads": ([], _I),
    "orc_pyramid": ([_PP, _VP, _VP, _VP], None),
    "orc_warm_start_usable": ([_VP, _VP, _I, _I, _I], _I),
    "orc_camera_pyramid": ([_VP, _VP, _VP, _I, _F, _I, _VP, _VP], None),
    "orc_round_half": ([_VP, _VP, C.c_long], None),
    "orc_solve_flow_f64": ([_PP] + [_VP] * 6, _I), "orc_solve_flow_f32": ([_PP] + [_VP] * 6, _I),
    "orc_solve_flow_masked_f64": ([_PP] + [_VP] * 5, _I), "orc_solve_flow_masked_f32": ([_PP] + [_VP] * 5, _I),
    "orc_solve_recon_f64": ([_PP] + [_VP] * 18, _I), "orc_solve_recon_f32": ([_PP] + [_VP] * 18, _I),
}


def _sym(name: str):
    f = _bound.get(name)
    if f is None:
        f = getattr(lib(), name)
        f.argtypes, f.restype = _SOLVE_SIGS[name]
        _bound[name] = f
    return f


def _kind(kind: str) -> str:
    if kind not in ("f64", "f32"):
        raise ValueError(f"kind must be 'f64' or 'f32', got {kind!r}")
    return kind


_IMG = {0: (np.float32, ()), 1: (np.uint8, (3,)), 2: (np.uint8, ()), 3: (np.uint16, ())}


def _image(p: Params, a, s0: int):
    """An input image of p.input_type, (h, stride[, 3]), checked against the solve's size."""
    dt, tail = _IMG[p.input_type]
    a = np.asarray(a)
    if a.dtype != dt or a.shape != (p.height, s0, *tail):
        raise TypeError(f"image must be {np.dtype(dt).name} of shape {(p.height, s0, *tail)}, got "
                        f"{a.dtype.name} {a.shape}")
    return np.ascontiguousarray(a)


def _plane(a, p: Params, s0: int, what: str):
    a = np.ascontiguousarray(a, np.float32)
    if a.shape != (p.height, s0):
        raise TypeError(f"{what} must be float32 of shape {(p.height, s0)}, got {a.shape}")
    return a


def _mat(a, n: int, what: str):
    a = np.ascontiguousarray(np.asarray(a, np.float64).ravel())
    if a.size != n:
        raise TypeError(f"{what} must hold {n} doubles, got {a.size}")
    return a


def set_threads(n: int) -> None:
    """OpenMP threads of the restatement's row loops."""
    _sym("orc_set_threads")(int(n))


def get_threads() -> int:
    return int(_sym("orc_get_threads")())


def pyramid(p: Params):
    n = p.n_levels
    W, H, S = (C.c_int * n)(), (C.c_int * n)(), (C.c_int * n)()
    _sym("orc_pyramid")(C.byref(p), W, H, S)
    return list(W), list(H), list(S)


def camera_pyramid(K, R, t, n_levels: int, scale: float, fix_coarse_k: int = 0):
    """Per-level intrinsics (n_levels, 3, 3) and K_l [R|t] (n_levels, 3, 4) of the joint solve."""
    Kd, Rd, td = _mat(K, 9, "K"), _mat(R, 9, "R"), _mat(t, 3, "t")
    Kl = np.zeros((n_levels, 3, 3))
    Pl = np.zeros((n_levels, 3, 4))
    _sym("orc_camera_pyramid")(_p(Kd), _p(Rd), _p(td), int(n_levels), float(scale), int(fix_coarse_k), _p(Kl),
                               _p(Pl))
    return Kl, Pl


def _ws(ws, p: Params, s0: int):
    """A warm start (u, v), level-0 (h, stride) float32 planes, as two pointers (or two NULLs)."""
    if ws is None:
        return [None, None], []
    keep = [_plane(a, p, s0, "warm start") for a in ws]
    return [_p(a) for a in keep], keep


def round_half(a) -> np.ndarray:
    """The FP16 storage modes' rounding: float32 values to the nearest binary16 (ties to even), as
    float32."""
    x = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(x)
    _sym("orc_round_half")(_p(x), _p(out), x.size)
    return out


def warm_start_usable(wsu, wsv, w, h) -> bool:
    """The warm-start rule: every component over the w x h pixels finite and at most max(w, h)."""
    a, b = (np.ascontiguousarray(x, np.float32) for x in (wsu, wsv))
    if a.shape != b.shape or a.ndim != 2 or a.shape[0] < h or a.shape[1] < w:
        raise TypeError("warm start planes must be (h, stride) float32 of the same shape")
    return bool(_sym("orc_warm_start_usable")(_p(a), _p(b), int(w), int(h), int(a.shape[1])))


def solve_flow(p: Params, i0: np.ndarray, i1: np.ndarray, kind: str = "f64", ws=None):
    """Whole CudaFlow TV-L1 solve. Images are (h, stride) of p.input_type (float32, or uint8 rgb
    (h, stride, 3), uint8, uint16). ws: an optional warm start (u, v) that replaces the zero initial
    flow at the coarsest level."""
    s0 = pyramid(p)[2][0]
    a0, a1 = _image(p, i0, s0), _image(p, i1, s0)
    u = np.zeros((p.height, s0), np.float32)
    v = np.zeros_like(u)
    wsp, keep = _ws(ws, p, s0)
    rc = _sym(f"orc_solve_flow_{_kind(kind)}")(C.byref(p), _p(a0), _p(a1), *wsp, _p(u), _p(v))
    assert rc == 0
    del keep
    return u, v


def solve_flow_masked(p: Params, i0: np.ndarray, i1: np.ndarray, mask0: np.ndarray, kind: str = "f64"):
    """CudaFlow::solveInpainting: the TV-L1 solve with the masked L1 step (mask0 (h, stride) float32)."""
    s0 = pyramid(p)[2][0]
    a0, a1 = _image(p, i0, s0), _image(p, i1, s0)
    u = np.zeros((p.height, s0), np.float32)
    v = np.zeros_like(u)
    m = _plane(mask0, p, s0, "mask0")
    rc = _sym(f"orc_solve_flow_masked_{_kind(kind)}")(C.byref(p), _p(a0), _p(a1), _p(m), _p(u), _p(v))
    assert rc == 0
    return u, v


def solve_recon(p: Params, K0, K1, R0, t0, R1, t1, i0, i1, kind="f64", ufn=None, vfn=None, mask=None, ws=None):
    """Whole ReconFlow joint solve: returns (u, v, X, Y, Z), each (h, stride) float32. ufn, vfn,
    mask: the prior flow of methods 17 / 20; ws: an optional warm start."""
    s0 = pyramid(p)[2][0]
    a0, a1 = _image(p, i0, s0), _image(p, i1, s0)
    outs = [np.zeros((p.height, s0), np.float32) for _ in range(5)]
    dd = [_mat(m, n, w) for m, n, w in ((K0, 9, "K0"), (K1, 9, "K1"), (R0, 9, "R0"), (t0, 3, "t0"), (R1, 9, "R1"),
                                        (t1, 3, "t1"))]
    optk = [None if a is None else _plane(a, p, s0, "prior flow") for a in (ufn, vfn, mask)]
    opt = [None if a is None else _p(a) for a in optk]
    wsp, keep = _ws(ws, p, s0)
    rc = _sym(f"orc_solve_recon_{_kind(kind)}")(C.byref(p), *[_p(d) for d in dd], _p(a0), _p(a1), *opt, *wsp,
                                                *[_p(o) for o in outs])
    assert rc == 0
    del keep, optk
    return tuple(outs)


def flow_params(w, h, n_levels, scale=2.0, n_warp=3, n_iter=30, lam=50.0, theta=0.3, tau=0.25, input_type=0,
                method=0, lambdagrad=400.0, storage=0):
    p = Params()
    p.width, p.height, p.n_levels, p.scale, p.method = w, h, n_levels, scale, method
    p.lambda_, p.lambdagrad, p.theta, p.tau = lam, lambdagrad, theta, tau
    p.n_warp_iters, p.n_solver_iters = n_warp, n_iter
    p.stride_alignment, p.input_type, p.storage = 32, input_type, storage
    return p


def recon_params(w, h, n_levels, scale=2.0, n_warp=5, n_iter=50, lam=50.0, tau=0.125, alpha_tv=33.3,
                 alpha_proj=60.0, lambdaf=0.1, lambdams=100.0, alpha_fn=100.0, method=14, input_type=0,
                 fix_coarse_k=0, tri_f64=0, storage=0):
    p = Params()
    p.width, p.height, p.n_levels, p.scale, p.method = w, h, n_levels, scale, method
    p.lambda_, p.lambdagrad, p.theta, p.tau = lam, 0.0, float(np.float32(1.0) / np.float32(alpha_tv)), tau
    p.n_warp_iters, p.n_solver_iters = n_warp, n_iter
    p.lambdaf, p.lambdams, p.alpha_tv, p.alpha_proj, p.alpha_fn = lambdaf, lambdams, alpha_tv, alpha_proj, alpha_fn
    p.stride_alignment, p.input_type, p.fix_coarse_k, p.tri_f64 = 32, input_type, fix_coarse_k, tri_f64
    p.storage = storage
    return p
