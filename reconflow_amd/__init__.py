"""MI355X-native TV-L1 optical flow and joint flow/depth (Reconflow's CudaFlow / ReconFlow).

The classes here mirror the reference's C++ interface (reconflow/Cudaflow.h:17-361,
reconflow/Reconflow.h:9-159) over the C ABI of ``include/rfhip.h``: initialize with size,
pyramid and solver parameters, copy the image pair (and camera pose) in, solve, copy flow
(and 3D points) out. Host arrays are numpy arrays of ``height x stride`` elements, padded to
the stride as the reference requires (test_reconflow/main.cpp:57-59). Methods that fill
``cv::Mat&`` outputs in the reference return new arrays here.

The solve itself runs only in the HIP library; nothing here computes a result on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from ._native import RFError, RFParams, check

__all__ = ["CudaFlow", "ReconFlow", "RFError", "make_flow_pair", "make_recon_pair", "align_up",
           "INPUT_32F", "INPUT_8UC3", "INPUT_8U", "INPUT_16U", "PATH_PLAIN", "PATH_FUSED",
           "STORAGE_F32", "STORAGE_F16", "STORAGE_F16_DERIV", "STORAGE_F16_DERIV_DUALS"]

INPUT_32F, INPUT_8UC3, INPUT_8U, INPUT_16U = 0, 1, 2, 3
PATH_PLAIN, PATH_FUSED = 0, 1
STORAGE_F32, STORAGE_F16, STORAGE_F16_DERIV, STORAGE_F16_DERIV_DUALS = 0, 1, 2, 3

_DTYPE = {INPUT_32F: np.float32, INPUT_8UC3: np.uint8, INPUT_8U: np.uint8, INPUT_16U: np.uint16}


def align_up(n: int, m: int = 32) -> int:
    """iAlignUp (Cudaflow.cu:985-994)."""
    r = n % m
    return n + m - r if r else n


def _dptr(a: np.ndarray) -> C.POINTER(C.c_double):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(C.POINTER(C.c_double)), a


class CudaFlow:
    """Coarse-to-fine TV-L1 optical flow (Cudaflow.h:17-361)."""

    METHOD_TVL1 = 0
    METHOD_TVCHARB = 1
    METHOD_TVCHARBGRAD = 2
    METHOD_TVL1PATCH = 3
    SCENEFLOW_KINECT_TVL1 = 4
    METHOD_TVL1_FN = 5
    METHOD_TVCHARB_FN = 6
    METHOD_TVCHARBGRAD_FN = 7
    METHOD_TVL1_INPAINTING = 8

    def __init__(self, BlockWidth: int = 32, BlockHeight: int = 12, StrideAlignment: int = 32):
        self._lib = _native.lib()
        self._ctx = C.c_void_p()
        check(self._lib.rf_create(C.byref(self._ctx), BlockWidth, BlockHeight, StrideAlignment), "rf_create")
        self.StrideAlignment = StrideAlignment
        self.params = None

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._lib.rf_destroy(ctx)
            self._ctx = None

    # ---- lifecycle -------------------------------------------------------------------
    def _init_params(self, p: RFParams) -> int:
        check(self._lib.rf_initialize(self._ctx, C.byref(p)), "rf_initialize")
        self.params = p
        self.width, self.height = p.width, p.height
        self.stride = self._lib.rf_get_stride(self._ctx)
        self.nLevels = p.n_levels
        self.inputType = p.input_type
        return 0

    def initialize(self, width, height, channels, cvType, nLevels, scale, method, lambda_=100.0,
                   lambdagrad=400.0, theta=0.33, tau=0.25, nWarpIters=1, nSolverIters=100,
                   withVisualization=True, *, path=PATH_PLAIN, storage=STORAGE_F32, use_graph=False):
        """CudaFlow::initialize (Cudaflow.cu:99-269). ``cvType`` is one of the INPUT_* codes."""
        p = RFParams()
        self._lib.rf_params_default(C.byref(p), width, height)
        p.channels, p.input_type, p.n_levels, p.scale, p.method = channels, cvType, nLevels, scale, method
        p.lambda_, p.lambdagrad, p.theta, p.tau = lambda_, lambdagrad, theta, tau
        p.n_warp_iters, p.n_solver_iters = nWarpIters, nSolverIters
        p.with_visualization = int(bool(withVisualization))
        p.stride_alignment = self.StrideAlignment
        p.path, p.storage, p.use_graph = path, storage, int(bool(use_graph))
        return self._init_params(p)

    def getPyramid(self):
        n = self.nLevels
        W, H, S = (C.c_int * n)(), (C.c_int * n)(), (C.c_int * n)()
        check(self._lib.rf_get_pyramid(self._ctx, W, H, S), "rf_get_pyramid")
        return list(W), list(H), list(S)

    def close(self) -> int:
        """CudaFlow::close (Cudaflow.cu:980-983) is a no-op; buffers live until destruction."""
        return 0

    # ---- inputs ----------------------------------------------------------------------
    def _img(self, a):
        dt = _DTYPE[self.inputType]
        shape = (self.height, self.stride, 3) if self.inputType == INPUT_8UC3 else (self.height, self.stride)
        a = np.ascontiguousarray(a, dtype=dt)
        if a.shape != shape:
            raise ValueError(f"image must be {shape} (padded to the stride), got {a.shape}")
        return a

    def copyWarmStartToDevice(self, u=None, v=None) -> int:
        """Warm-start input (not in the reference): (h, stride) float32 flow that replaces the zero
        initial flow at the coarsest level of the following solves; a non-finite component or one
        larger than max(width, height) makes a solve start from zero; None clears."""
        if u is None or v is None:
            check(self._lib.rf_copy_warm_start(self._ctx, None, None), "rf_copy_warm_start")
            return 0
        u, v = (np.ascontiguousarray(a, np.float32) for a in (u, v))
        check(self._lib.rf_copy_warm_start(self._ctx, u.ctypes.data, v.ctypes.data), "rf_copy_warm_start")
        return 0

    def warmStartInUse(self) -> bool:
        """True when a warm start is set and the last solve used it."""
        rc = self._lib.rf_warm_start_in_use(self._ctx)
        if rc < 0:
            check(rc, "rf_warm_start_in_use")
        return rc == 1

    def copyImagesToDevice(self, i0, i1) -> int:
        """CudaFlow::copyImagesToDevice (Cudaflow.cu:288-311)."""
        a0, a1 = self._img(i0), self._img(i1)
        check(self._lib.rf_copy_images(self._ctx, a0.ctypes.data, a1.ctypes.data), "rf_copy_images")
        return 0

    def copyImagesToDeviceFromDevice(self, d_i0: int, d_i1: int) -> int:
        check(self._lib.rf_copy_images_device(self._ctx, d_i0, d_i1), "rf_copy_images_device")
        return 0

    def _plane(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != (self.height, self.stride):
            raise ValueError(f"plane must be {(self.height, self.stride)} (padded to the stride), got {a.shape}")
        return a

    def copyMaskToDevice(self, mask0, mask1) -> int:
        """CudaFlow::copyMaskToDevice (Inpainting.cu:3-7): float masks of METHOD_TVL1_INPAINTING."""
        m0, m1 = self._plane(mask0, np.float32), self._plane(mask1, np.float32)
        check(self._lib.rf_copy_masks(self._ctx, m0.ctypes.data, m1.ctypes.data), "rf_copy_masks")
        return 0

    def copyDepthToDevice(self, depth0, depth1) -> int:
        """CudaFlow::copyDepthToDevice (Cudaflow.cu:282-286): 16-bit depth of SCENEFLOW_KINECT_TVL1."""
        d0, d1 = self._plane(depth0, np.uint16), self._plane(depth1, np.uint16)
        check(self._lib.rf_copy_depth(self._ctx, d0.ctypes.data, d1.ctypes.data), "rf_copy_depth")
        return 0

    # ---- solve -----------------------------------------------------------------------
    def solveOpticalFlow(self, flowScale: float = 5.0) -> int:
        """CudaFlow::solveOpticalFlow (Cudaflow.cu:395-621); asynchronous."""
        check(self._lib.rf_solve_flow(self._ctx, flowScale), "rf_solve_flow")
        return 0

    def solveInpainting(self, flowScale: float = 5.0) -> int:
        """CudaFlow::solveInpainting (Inpainting.cu:9-152); asynchronous."""
        check(self._lib.rf_solve_inpainting(self._ctx, flowScale), "rf_solve_inpainting")
        return 0

    def solveSceneFlow(self, focalx, focaly, cx, cy, opticalFlowScale=5.0, sceneFlowScale=1.0) -> int:
        """CudaFlow::solveSceneFlow (Cudaflow.cu:816-848); asynchronous."""
        check(self._lib.rf_solve_scene_flow(self._ctx, focalx, focaly, cx, cy, opticalFlowScale, sceneFlowScale, 0),
              "rf_solve_scene_flow")
        return 0

    def solveSceneFlowAndPointCloud(self, focalx, focaly, cx, cy, opticalFlowScale=5.0, sceneFlowScale=1.0) -> int:
        """CudaFlow::solveSceneFlowAndPointCloud (Cudaflow.cu:850-871); asynchronous."""
        check(self._lib.rf_solve_scene_flow(self._ctx, focalx, focaly, cx, cy, opticalFlowScale, sceneFlowScale, 1),
              "rf_solve_scene_flow")
        return 0

    def synchronize(self) -> int:
        check(self._lib.rf_synchronize(self._ctx), "rf_synchronize")
        return 0

    def stream(self) -> int:
        return self._lib.rf_get_stream(self._ctx) or 0

    # ---- outputs ---------------------------------------------------------------------
    def copyOpticalFlowToHost(self, with_rgb: bool = False):
        """CudaFlow::copyOpticalFlowToHost (Cudaflow.cu:313-324): (u, v[, uvrgb])."""
        u = np.empty((self.height, self.stride), np.float32)
        v = np.empty_like(u)
        rgb = np.empty((self.height, self.stride, 3), np.float32) if with_rgb else None
        check(self._lib.rf_copy_flow(self._ctx, u.ctypes.data, v.ctypes.data,
                                     rgb.ctypes.data if with_rgb else None), "rf_copy_flow")
        return (u, v, rgb) if with_rgb else (u, v)

    def copySceneFlowToHost(self, with_rgb: bool = False):
        """CudaFlow::copySceneFlowToHost (Cudaflow.cu:345-354): (h, stride, 3) float (x, y, z)."""
        sf = np.empty((self.height, self.stride, 3), np.float32)
        rgb = np.empty_like(sf) if with_rgb else None
        check(self._lib.rf_copy_scene_flow(self._ctx, sf.ctypes.data, rgb.ctypes.data if with_rgb else None, None,
                                           None), "rf_copy_scene_flow")
        return (sf, rgb) if with_rgb else sf

    def copySceneFlowAndPointCloudToHost(self):
        """CudaFlow::copySceneFlowAndPointCloudToHost (Cudaflow.cu:356-362): (sf, pc0, pc1, rgb)."""
        arrs = [np.empty((self.height, self.stride, 3), np.float32) for _ in range(4)]
        sf, pc0, pc1, rgb = arrs
        check(self._lib.rf_copy_scene_flow(self._ctx, sf.ctypes.data, rgb.ctypes.data, pc0.ctypes.data,
                                           pc1.ctypes.data), "rf_copy_scene_flow")
        return sf, pc0, pc1, rgb

    def resultDevicePointers(self):
        ptrs = [C.c_void_p() for _ in range(5)]
        check(self._lib.rf_get_result_device(self._ctx, *[C.byref(p) for p in ptrs]), "rf_get_result_device")
        return [p.value or 0 for p in ptrs]


class ReconFlow(CudaFlow):
    """Joint optical flow and two-view reconstruction (Reconflow.h:9-159)."""

    METHODR_TVL1_MS = 14
    METHODR_TVL1_MSPLANAR = 15
    METHODR_TVL1_3D = 16
    METHODR_TVL1_MS_FN = 17
    METHODR_TVCHARBGRAD_MS = 18
    METHODR_TVCHARBGRAD_MS_FN = 19
    METHODR_TVL1_MS_FNSPARSE = 20

    def _initializeR(self, width, height, channels, nLevels, scale, method, lambda_=100.0, lambdagrad=400.0,
                     lambdaf=100.0, lambdams=100.0, alphaTv=3.0, alphaProj=1.0, tau=0.25, nWarpIters=1,
                     nSolverIters=100, *, cvType=INPUT_8UC3, path=PATH_PLAIN, storage=STORAGE_F32,
                     use_graph=False, withVisualization=True):
        """ReconFlow::_initializeR (Reconflow.cu:27-63). The reference accepts only 8-bit RGB;
        ``cvType`` extends it to the other CudaFlow input types. Every solve ends with the flow's
        colour map, as solveReconFlow does (Reconflow.cu:460); ``withVisualization=False`` skips it."""
        p = RFParams()
        self._lib.rf_params_default(C.byref(p), width, height)
        p.channels, p.input_type, p.n_levels, p.scale, p.method = channels, cvType, nLevels, scale, method
        p.lambda_, p.lambdagrad, p.tau = lambda_, lambdagrad, tau
        p.theta = float(np.float32(1.0) / np.float32(alphaTv))
        p.lambdaf, p.lambdams, p.alpha_tv, p.alpha_proj = lambdaf, lambdams, alphaTv, alphaProj
        p.n_warp_iters, p.n_solver_iters = nWarpIters, nSolverIters
        p.with_visualization = int(bool(withVisualization))
        p.stride_alignment = self.StrideAlignment
        p.path, p.storage, p.use_graph = path, storage, int(bool(use_graph))
        return self._init_params(p)

    def initializeR(self, width, height, channels, nLevels, scale, method, lambda_=100.0, lambdagrad=400.0,
                    lambdaf=100.0, lambdams=100.0, alphaTv=3.0, alphaProj=1.0, alphaFn=0.5, tau=0.25,
                    nWarpIters=1, nSolverIters=100, **kw):
        """ReconFlow::initializeR (Reconflow.cu:15-25): with the prior-flow buffers."""
        self._initializeR(width, height, channels, nLevels, scale, method, lambda_, lambdagrad, lambdaf,
                          lambdams, alphaTv, alphaProj, tau, nWarpIters, nSolverIters, **kw)
        self.params.alpha_fn = alphaFn
        check(self._lib.rf_initialize(self._ctx, C.byref(self.params)), "rf_initialize")
        check(self._lib.rf_initialize_large_displacement(self._ctx), "rf_initialize_large_displacement")
        return 0

    def setCameraMatrices(self, K0, K1=None) -> int:
        """ReconFlow::setCameraMatrices (Reconflow.cu:72-116); the one-matrix form is a no-op."""
        if K1 is None:
            return 0
        p0, a0 = _dptr(K0)
        p1, a1 = _dptr(K1)
        check(self._lib.rf_set_camera(self._ctx, p0, p1), "rf_set_camera")
        return 0

    def copyInitialOpticalFlowToDevice(self, u, v) -> int:
        u, v = (np.ascontiguousarray(a, np.float32) for a in (u, v))
        check(self._lib.rf_copy_initial_flow(self._ctx, u.ctypes.data, v.ctypes.data), "rf_copy_initial_flow")
        return 0

    def copySparseOpticalFlowToDevice(self, u, v, mask=None) -> int:
        u, v = (np.ascontiguousarray(a, np.float32) for a in (u, v))
        m = None if mask is None else np.ascontiguousarray(mask, np.float32)
        check(self._lib.rf_copy_sparse_flow(self._ctx, u.ctypes.data, v.ctypes.data,
                                            None if m is None else m.ctypes.data), "rf_copy_sparse_flow")
        return 0

    def solveR(self, *args) -> int:
        """solveR(R1, t1, flowScale) or solveR(R0, t0, R1, t1, flowScale) (Reconflow.cu:130-136)."""
        if len(args) == 3:
            R0 = t0 = None
            R1, t1, fs = args
        elif len(args) == 5:
            R0, t0, R1, t1, fs = args
        else:
            raise TypeError("solveR(R1, t1, flowScale) or solveR(R0, t0, R1, t1, flowScale)")
        keep = []

        def dp(a):
            if a is None:
                return None
            p, arr = _dptr(a)
            keep.append(arr)
            return p

        check(self._lib.rf_solve_recon(self._ctx, dp(R0), dp(t0), dp(R1), dp(t1), float(fs)), "rf_solve_recon")
        return 0

    solveReconFlow = solveR

    def copy3dToHost(self):
        """ReconFlow::copy3dToHost (Reconflow.cu:65-70): (X, Y, Z)."""
        X = np.empty((self.height, self.stride), np.float32)
        Y, Z = np.empty_like(X), np.empty_like(X)
        check(self._lib.rf_copy_3d(self._ctx, X.ctypes.data, Y.ctypes.data, Z.ctypes.data), "rf_copy_3d")
        return X, Y, Z

    def removeEdgeArtifacts3D(self, surfaceWidth: int, maxSurfaceArea: float) -> int:
        check(self._lib.rf_remove_edge_artifacts_3d(self._ctx, surfaceWidth, maxSurfaceArea),
              "rf_remove_edge_artifacts_3d")
        return 0


class Batch:
    """Many pairs of uneven sizes solved together (rf_batch_* of rfhip.h): every launch of the
    shared schedule covers the tiles of all pairs. ``params`` is an RFParams whose width and
    height are ignored (taken per pair); images are handed over as device float buffers."""

    def __init__(self, params: RFParams, widths, heights):
        self._lib = _native.lib()
        self._b = C.c_void_p()
        n = len(widths)
        W = (C.c_int * n)(*widths)
        H = (C.c_int * n)(*heights)
        check(self._lib.rf_batch_create(C.byref(self._b), C.byref(params), n, W, H), "rf_batch_create")
        self.n = n
        self.widths, self.heights = list(widths), list(heights)

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            self._lib.rf_batch_destroy(b)
            self._b = None

    def stride(self, i: int) -> int:
        return self._lib.rf_batch_stride(self._b, i)

    def set_images_device(self, i: int, d_i0: int, d_i1: int) -> None:
        check(self._lib.rf_batch_set_images_device(self._b, i, d_i0, d_i1), "rf_batch_set_images_device")

    def set_camera(self, i, K0, K1, R1, t1) -> None:
        ps = [_dptr(a) for a in (K0, K1, R1, t1)]
        check(self._lib.rf_batch_set_camera(self._b, i, *[p for p, _ in ps]), "rf_batch_set_camera")

    def set_warm_start_device(self, i: int, d_u: int, d_v: int) -> None:
        """Warm start of pair i from device flow (stride_i x height_i floats each); 0 clears."""
        check(self._lib.rf_batch_set_warm_start_device(self._b, i, d_u or None, d_v or None),
              "rf_batch_set_warm_start_device")

    def warm_start_in_use(self, i: int) -> bool:
        rc = self._lib.rf_batch_warm_start_in_use(self._b, i)
        if rc < 0:
            check(rc, "rf_batch_warm_start_in_use")
        return rc == 1

    def solve(self) -> None:
        check(self._lib.rf_batch_solve(self._b), "rf_batch_solve")

    def synchronize(self) -> None:
        check(self._lib.rf_batch_synchronize(self._b), "rf_batch_synchronize")

    def stream(self) -> int:
        return self._lib.rf_batch_get_stream(self._b) or 0

    def result_ptrs(self, i: int):
        ptrs = [C.c_void_p() for _ in range(5)]
        check(self._lib.rf_batch_get_result_device(self._b, i, *[C.byref(p) for p in ptrs]),
              "rf_batch_get_result_device")
        return [p.value or 0 for p in ptrs]

    def tiling(self, n_levels: int):
        """Iteration tiling per level (RF_JOINT_TILE_* or RF_FLOW_TILE_* codes)."""
        tl = (C.c_int * n_levels)()
        check(self._lib.rf_batch_tiling(self._b, tl, n_levels), "rf_batch_tiling")
        return list(tl)


def batch_params(method: int, n_levels: int, n_warp: int, n_iter: int, *, lam=50.0, theta=0.3, tau=0.25,
                 alpha_tv=33.3, alpha_proj=60.0, lambdaf=0.1, lambdams=100.0, use_graph=True, tri_f64=0,
                 storage=STORAGE_F32, joint_tile=0, flow_tile=0, with_visualization=False,
                 fix_coarse_k=0) -> RFParams:
    """Parameters for a Batch: TV-L1 flow (method 0) or the joint flow/depth solve (method 14),
    float gray input, scale 2 pyramid."""
    lib = _native.lib()
    p = RFParams()
    lib.rf_params_default(C.byref(p), 1, 1)
    p.input_type, p.n_levels, p.scale, p.method = INPUT_32F, n_levels, 2.0, method
    p.lambda_, p.theta, p.tau = lam, theta, tau
    p.n_warp_iters, p.n_solver_iters = n_warp, n_iter
    p.lambdaf, p.lambdams, p.alpha_tv, p.alpha_proj = lambdaf, lambdams, alpha_tv, alpha_proj
    p.path, p.use_graph, p.tri_f64, p.storage = PATH_FUSED, int(use_graph), tri_f64, storage
    p.joint_tile, p.flow_tile = joint_tile, flow_tile
    p.with_visualization = int(bool(with_visualization))
    p.fix_coarse_k = int(fix_coarse_k)
    return p


# ---- seeded synthetic pairs ---------------------------------------------------------
def make_flow_pair(w: int, h: int, seed: int, stride: int | None = None):
    """Synthetic textured pair moved by a smooth flow field: returns (i0, i1, gt_u, gt_v), each
    (h, stride) float32."""
    s = stride or align_up(w)
    arrs = [np.zeros((h, s), np.float32) for _ in range(4)]
    check(_native.synth().rfs_make_flow_pair(w, h, s, seed, *[a.ctypes.data for a in arrs]), "rfs_make_flow_pair")
    return tuple(arrs)


def make_recon_pair(w: int, h: int, seed: int, stride: int | None = None):
    """Synthetic rigid scene seen by two posed pinhole cameras: returns dict with i0, i1, K, R, t,
    gt_u, gt_v, gt_Z."""
    s = stride or align_up(w)
    i0, i1, gu, gv, gz = (np.zeros((h, s), np.float32) for _ in range(5))
    K, R, t = np.zeros(9), np.zeros(9), np.zeros(3)
    check(_native.synth().rfs_make_recon_pair(w, h, s, seed, K.ctypes.data, R.ctypes.data, t.ctypes.data,
                                              i0.ctypes.data, i1.ctypes.data, gu.ctypes.data, gv.ctypes.data,
                                              gz.ctypes.data), "rfs_make_recon_pair")
    return dict(i0=i0, i1=i1, K=K.reshape(3, 3), R=R.reshape(3, 3), t=t, gt_u=gu, gt_v=gv, gt_Z=gz)


def make_sequence(w: int, h: int, n: int, seed: int, stride: int | None = None):
    """Synthetic camera trajectory over a piecewise-planar scene: returns dict with frames
    (n, h, stride) float32, K (3x3) and the n - 1 relative poses R (n-1, 3, 3), t (n-1, 3) that map
    frame k's camera points into frame k + 1's."""
    s = stride or align_up(w)
    frames = np.zeros((n, h, s), np.float32)
    K, R, t = np.zeros(9), np.zeros((n - 1) * 9), np.zeros((n - 1) * 3)
    check(_native.synth().rfs_make_sequence(w, h, s, seed, n, K.ctypes.data, R.ctypes.data, t.ctypes.data,
                                            frames.ctypes.data), "rfs_make_sequence")
    return dict(frames=frames, K=K.reshape(3, 3), R=R.reshape(n - 1, 3, 3), t=t.reshape(n - 1, 3))
