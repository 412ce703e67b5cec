"""Vector kernels of the L-BFGS update (pypty/iterative.py:573-932) on HIP devices.

``dots`` evaluates several real inner products Re<a, b> in one launch pair
with FP64 accumulation (``pty_multidot_c64``); ``axpby`` is y <- alpha y +
beta x (``pty_axpby_c64``, ``pty_axpby_f32`` for the real blocks).  Complex64 tensors are reduced as complex,
float32 tensors as real.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def dots(pairs):
    """[sum Re(a * conj(b)) for (a, b) in pairs] as Python floats (one sync)."""
    if not pairs:
        return []
    out_all = []
    for i0 in range(0, len(pairs), 16):
        chunk = pairs[i0:i0 + 16]
        a0 = chunk[0][0]
        is_real = not a0.is_complex()
        n = a0.numel()
        for a, b in chunk:
            _lib.require_device_tensor(a, "a")
            _lib.require_device_tensor(b, "b")
            if a.numel() != n or b.numel() != n or a.is_complex() == is_real or b.is_complex() == is_real:
                raise ValueError("pypty_amd.lbfgs.dots: pairs must share size and kind")
            if a.dtype not in (torch.complex64, torch.float32) or b.dtype != a.dtype:
                raise TypeError("pypty_amd.lbfgs.dots: complex64 or float32 required")
        k = len(chunk)
        dev = a0.device
        out = torch.empty(k, dtype=torch.float64, device=dev)
        partial = torch.empty(k * 512, dtype=torch.float64, device=dev)
        A = (ctypes.c_void_p * k)(*[a.data_ptr() for a, _ in chunk])
        B = (ctypes.c_void_p * k)(*[b.data_ptr() for _, b in chunk])
        rc = _lib.load().pty_multidot_c64(A, B, k, n, 1 if is_real else 0, out.data_ptr(), partial.data_ptr(),
                                          _lib.stream_ptr(dev))
        _lib.check(rc, "multidot")
        out_all.extend(out.tolist())
    return out_all


def axpby(y: torch.Tensor, x, alpha: float, beta: float) -> torch.Tensor:
    """In place y <- alpha * y + beta * x (x None: y <- alpha * y)."""
    _lib.require_device_tensor(y, "y")
    if x is not None:
        _lib.require_device_tensor(x, "x")
        if x.dtype != y.dtype or x.numel() != y.numel():
            raise ValueError("pypty_amd.lbfgs.axpby: x and y must match")
    if y.dtype == torch.complex64:
        fn = _lib.load().pty_axpby_c64
    elif y.dtype == torch.float32:
        fn = _lib.load().pty_axpby_f32
    else:
        raise TypeError("pypty_amd.lbfgs.axpby: complex64 or float32 required")
    rc = fn(y.data_ptr(), None if x is None else x.data_ptr(), y.numel(), float(alpha), float(beta),
            _lib.stream_ptr(y.device))
    _lib.check(rc, "axpby")
    return y


def lincomb(x: torch.Tensor, beta: float) -> torch.Tensor:
    """beta * x as a new tensor."""
    y = torch.empty_like(x)
    return axpby(y, x, 0.0, beta)
