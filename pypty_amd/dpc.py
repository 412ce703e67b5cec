"""Differential phase contrast start (mirror of pypty/dpc.py:9-187).

``fft_based_dpc`` gives the phase map that ``initialize.get_ptycho_obj_from_scan`` turns into a
starting object: the centre of mass of every pattern (the pass over the whole dataset, on the HIP
device through ``pty_center_of_mass``, FP64 sums), the rotation between scan and detector axes
(the angle that minimises the curl of the rotated field, as ``GetPLRotation``), and the integration
of the rotated field through Fourier space on the scan grid.  The last two steps work on arrays of
the scan's size (a few thousand values) and run on the host in float64 with numpy and scipy, in the
reference's operations.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

from . import _arrays, _lib


def get_curl(angle, dpcx, dpcy):
    """Standard deviation of the curl of the DPC field rotated by ``angle`` (dpc.py:9-31)."""
    rotx = dpcx * np.cos(angle) - dpcy * np.sin(angle)
    roty = dpcx * np.sin(angle) + dpcy * np.cos(angle)
    gXY, gXX = np.gradient(rotx)
    gYY, gYX = np.gradient(roty)
    return np.std(gXY - gYX)


def get_curl_derivative(angle, dpcx, dpcy):
    """The reference's derivative of the curl objective with respect to the angle (dpc.py:33-56)."""
    rotx = dpcx * np.cos(angle) - dpcy * np.sin(angle)
    roty = dpcx * np.sin(angle) + dpcy * np.cos(angle)
    gXY, gXX = np.gradient(rotx)
    gYY, gYX = np.gradient(roty)
    return np.sum((gXX + gYY) * (-dpcx * np.sin(angle) - dpcy * np.cos(angle))) / np.sqrt(len(gXY) - 1)


def GetPLRotation(dpcx, dpcy):  # noqa: N802 (reference name)
    """Rotation angle (radians) minimising ``get_curl``: Powell over [-pi, pi] (dpc.py:58-78)."""
    from scipy.optimize import minimize

    sys.stdout.write("\nStarting the DPC rotation angle calculation!")
    sys.stdout.flush()
    result = minimize(get_curl, x0=0, method="Powell", args=(dpcx, dpcy), bounds=[(-np.pi, np.pi)], tol=1e-4,
                      options={"maxiter": 100})
    return result.x[0]


def center_of_mass(patterns, flip_ky=False):
    """Centre of mass of each pattern (P, ny, nx) about its centre, on the device: (comx, comy) float64
    tensors of P values (dpc.py:139-150)."""
    dev = patterns.device if isinstance(patterns, torch.Tensor) and patterns.is_cuda else \
        torch.device("cuda", torch.cuda.current_device())
    if isinstance(patterns, torch.Tensor) and patterns.is_cuda:
        data = patterns.to(torch.float32).contiguous()
    else:
        data = _arrays.stage(patterns, dev, torch.float32).contiguous()
    if data.dim() != 3:
        raise ValueError("pypty_amd.center_of_mass: patterns must be (P, ny, nx)")
    P, ny, nx = data.shape
    comx = torch.empty(P, dtype=torch.float64, device=dev)
    comy = torch.empty(P, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().pty_center_of_mass(data.data_ptr(), P, ny, nx, int(bool(flip_ky)), comx.data_ptr(),
                                              comy.data_ptr(), _lib.stream_ptr(dev)), "center_of_mass")
    return comx, comy


def fft_based_dpc(pypty_params, hpass=0, lpass=0, save=False, comx=None, comy=None, plot=False):
    """FFT-integrated DPC phase (dpc.py:80-187).  Returns (pot, pypty_params) with pot the (scan_y,
    scan_x) float64 potential and pypty_params updated with "comx", "comy" and "PLRotation_deg", as
    there.  The patterns come from pypty_params["dataset"] (an array or device tensor, (P, ny, nx)
    with "scan_size", or 4D) or a .npy "data_path" (HDF5 is not read).  With hpass = lpass = 0 the
    zero frequency divides 0 by 0, so the potential is NaN everywhere, as in the reference; pass
    hpass > 0.  plot is accepted and ignored."""
    save = pypty_params.get("save_preprocessing_files", save)
    scan_size = pypty_params.get("scan_size", None)
    angle = pypty_params.get("PLRotation_deg", None)
    flip_ky = pypty_params.get("flip_ky", False)
    if comx is None:
        comx = pypty_params.get("comx", None)
    if comy is None:
        comy = pypty_params.get("comy", None)
    if comx is None or comy is None:
        data = pypty_params.get("dataset", None)
        if data is None:
            path = pypty_params.get("data_path", "")
            if not path.endswith(".npy"):
                raise NotImplementedError("pypty_amd.fft_based_dpc: data_path must be a .npy file (no h5py)")
            data = np.load(path)
        if len(data.shape) == 4:
            scan_size = [data.shape[0], data.shape[1]]
            data = data.reshape(data.shape[0] * data.shape[1], data.shape[2], data.shape[3])
        cx, cy = center_of_mass(data, flip_ky)
        comx = cx.cpu().numpy()
        comy = cy.cpu().numpy()
        comx -= np.mean(comx)
        comy -= np.mean(comy)
        comx = comx.reshape(scan_size[0], scan_size[1])
        comy = comy.reshape(scan_size[0], scan_size[1])
    if angle is None:
        angle = GetPLRotation(comx, comy)
    else:
        angle = angle * np.pi / 180
    rcomx = comx * np.cos(angle) + comy * np.sin(angle)
    rcomy = -comx * np.sin(angle) + comy * np.cos(angle)
    fCX = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(rcomx)))
    fCY = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(rcomy)))
    kx, ky = np.meshgrid(np.fft.fftshift(np.fft.fftfreq(rcomx.shape[1])),
                         np.fft.fftshift(np.fft.fftfreq(rcomx.shape[0])), indexing="xy")
    fnum = fCX * kx + fCY * ky
    fdenom = np.pi * 2j * (hpass + (kx ** 2 + ky ** 2) + lpass * (kx ** 2 + ky ** 2) ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        fK = np.divide(fnum, fdenom)
    pot = np.real(np.fft.ifftshift(np.fft.ifft2(np.fft.ifftshift(fK))))
    pypty_params["comx"] = comx
    pypty_params["comy"] = comy
    if save:
        folder = os.path.join(pypty_params["output_folder"], "dpc")
        os.makedirs(folder, exist_ok=True)
        np.save(os.path.join(folder, "idpc.npy"), pot)
    pypty_params["PLRotation_deg"] = angle * 180 / np.pi
    return pot, pypty_params
