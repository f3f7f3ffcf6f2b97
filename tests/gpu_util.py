"""Helpers for the GPU parity tests: device buffers (torch as plumbing) and C-ABI calls."""
import ctypes as C

import numpy as np


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return torch


def dev(a):
    """Host numpy array -> device float32/uint8/uint16 tensor (contiguous copy)."""
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def dzeros(n, dtype="float32"):
    torch = torch_cuda()
    return torch.zeros(n, dtype=getattr(torch, dtype), device="cuda")


def host(t):
    torch = torch_cuda()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def ptr(t):
    return C.c_void_p(t.data_ptr())


_hip = None


def fetch(dptr_value, n):
    """n floats at a raw device address -> host numpy (hipMemcpy device-to-host, after a device
    synchronisation)."""
    global _hip
    torch = torch_cuda()
    if _hip is None:
        from reconflow_amd import _native
        _native.lib()  # binds the process's single HIP runtime first
        _hip = C.CDLL("libamdhip64.so.7")  # by soname: the runtime already loaded
        _hip.hipDeviceSynchronize.restype = C.c_int
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    torch.cuda.synchronize()
    assert _hip.hipDeviceSynchronize() == 0
    out = np.empty(n, np.float32)
    assert _hip.hipMemcpy(out.ctypes.data, dptr_value, n * 4, 2) == 0  # 2 = device to host
    return out


def copy_async(dst, src, nbytes, stream):
    """hipMemcpyAsync device to device of nbytes on `stream` (raw addresses)."""
    if _hip is None:
        fetch(0, 0)  # binds the runtime
    _hip.hipMemcpyAsync.restype = C.c_int
    _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip.hipMemcpyAsync(dst, src, nbytes, 3, stream)  # 3 = device to device


def dptr(a):
    a = np.ascontiguousarray(np.asarray(a, np.float64).ravel())
    return a.ctypes.data_as(C.POINTER(C.c_double)), a


def call(lib, name, *args):
    """Calls rf_k_<name> on the null stream; device tensors become pointers."""
    torch = torch_cuda()
    conv = []
    keep = []
    for a in args:
        if hasattr(a, "data_ptr"):
            conv.append(ptr(a))
        elif isinstance(a, np.ndarray) and a.dtype == np.float64:
            p, arr = dptr(a)
            keep.append(arr)
            conv.append(p)
        else:
            conv.append(a)
    rc = getattr(lib, f"rf_k_{name}")(*conv, None)
    torch.cuda.synchronize()
    return rc


def same_bits(a, b):
    """Bit-exact equality treating NaN == NaN (and +0 == -0 only where both are zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a == b) | both_nan))


def normwise(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = max(np.abs(ref[np.isfinite(ref)]).max() if np.isfinite(ref).any() else 0.0, 1e-30)
    return float(np.nanmax(np.abs(a - ref)) / scale)
