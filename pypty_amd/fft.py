"""Complex 2D FFTs of probe-sized windows (mirror of pypty/fft.py:91-178).

Same call signatures as the reference: ``fft2(arr, axes=(0,1),
overwrite_x=False, xp=None)``.  ``arr`` is a complex64 torch tensor on a HIP
device, a numpy array (copied to the device; the result is a numpy array) or a
device array of another library such as CuPy (viewed through DLPack; the
result is of that library); the two transformed axes are moved innermost, the batch of n x n
tiles is transformed by the LDS-resident gfx950 kernel (pty_fft2_c64) and the
axes are moved back.  numpy convention: forward unnormalised, inverse 1/n^2.
``xp`` is accepted for signature compatibility and ignored.
"""
from __future__ import annotations

import torch

from . import _arrays, _lib


def _run(arr, axes, mode: int, overwrite_x: bool):
    kind = _arrays.kind_of(arr)
    if kind != _arrays.TORCH:
        # numpy (copied to the device) or a device array of another library (viewed through DLPack);
        # the result comes back as the caller's type
        t = _arrays.to_torch(arr)
        if t.dtype == torch.complex128 or not t.is_complex():
            t = t.to(torch.complex64)
        out = _run(t, axes, mode, overwrite_x and kind == _arrays.FOREIGN)
        return _arrays.like(out, kind, arr)
    if arr.device.type != "cuda":
        raise TypeError("pypty_amd.fft: input must be on a HIP device")
    if arr.dtype != torch.complex64:
        raise TypeError(f"pypty_amd.fft: complex64 required, got {arr.dtype}")
    nd = arr.dim()
    a0, a1 = (ax % nd for ax in axes)
    if a0 == a1:
        raise ValueError("pypty_amd.fft: axes must differ")
    n0, n1 = arr.shape[a0], arr.shape[a1]
    lib = _lib.load()
    moved = torch.movedim(arr, (a0, a1), (nd - 2, nd - 1))
    work = moved if (overwrite_x and moved.is_contiguous()) else moved.contiguous()
    if work.data_ptr() == arr.data_ptr() and not overwrite_x:
        work = work.clone()
    batch = work.numel() // (n0 * n1)
    stream = _lib.stream_ptr(arr.device)
    if n0 == n1 and lib.pty_supported_window(n0):
        # LDS-resident window kernel
        _lib.check(lib.pty_fft2_c64(work.data_ptr(), batch, n0, mode, stream), "fft2")
    else:
        # whole-plane kernels (rows pass + columns pass through HBM)
        scratch = torch.empty_like(work)
        _lib.check(lib.pty_plane_fft2_c64(work.data_ptr(), scratch.data_ptr(), batch, n0, n1, mode, stream),
                   "plane fft2")
    return torch.movedim(work, (nd - 2, nd - 1), (a0, a1))


def fft2(arr, axes=(0, 1), overwrite_x=False, xp=None):
    """Unnormalised 2D DFT over ``axes`` (pypty/fft.py:114-134)."""
    return _run(arr, axes, _lib.PTY_FFT_FWD, overwrite_x)


def ifft2(arr, axes=(0, 1), overwrite_x=False, xp=None):
    """Inverse 2D DFT with 1/n^2 (pypty/fft.py:136-156)."""
    return _run(arr, axes, _lib.PTY_FFT_INV, overwrite_x)


def shift_fft2(arr, axes=(0, 1), overwrite_x=False, xp=None):
    """fftshift(fft2(arr)) over ``axes`` (pypty/fft.py:91-112)."""
    return _run(arr, axes, _lib.PTY_FFT_FWD_SHIFT, overwrite_x)


def ifft2_ishift(arr, axes=(0, 1), overwrite_x=False, xp=None):
    """ifft2(ifftshift(arr)) over ``axes`` (pypty/fft.py:158-178)."""
    return _run(arr, axes, _lib.PTY_FFT_INV_ISHIFT, overwrite_x)


def fftfreq(length, sampling=1, xp=None):
    """numpy.fft.fftfreq as a float64 CPU tensor (pypty/fft.py:15-31)."""
    n = int(length)
    k = torch.cat([torch.arange(0, (n - 1) // 2 + 1), torch.arange(-(n // 2), 0)]).to(torch.float64)
    return k * (1.0 / (n * float(sampling)))


def fftshift(array, xp=None):
    """Shift over every axis, like scipy.fft.fftshift (pypty/fft.py:52-68)."""
    return torch.roll(array, shifts=[s // 2 for s in array.shape], dims=list(range(array.dim())))


def ifftshift(array, xp=None):
    """Inverse of fftshift over every axis (pypty/fft.py:70-89)."""
    return torch.roll(array, shifts=[-(s // 2) for s in array.shape], dims=list(range(array.dim())))
