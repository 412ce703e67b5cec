"""Multislice forward model and its adjoint (mirror of pypty/multislice_core.py).

``multislice``, ``multislice_grads`` and ``scatteradd`` keep the reference
signatures and array layouts and run the gfx950 kernels behind ``pty_ms_forward``,
``pty_ms_adjoint`` and ``pty_scatteradd_c64``.  Supported subset: a single
slice distance (``is_single_dist``), no tilt gradients.  Unlike the
reference these functions do not write through views of their inputs
(the reference's in-place conjugation of the stash and mutation of
``dLoss_dP_out`` are replaced by returned values); ``object_grad`` and
``scatteradd``'s ``full`` are accumulated in place as there.  Arrays may be
torch device tensors, numpy arrays (copied to the device; results come back
as numpy) or device arrays of another library such as CuPy (viewed through
DLPack; results come back as that library's arrays).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _arrays, _lib


def _c64(x):
    return _arrays.to_torch(x, dtype=torch.complex64)


def _problem(n, S, K, M, edge):
    p = _lib.PtyProblem()
    p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = int(n), int(S), int(K), int(M)
    p.compute_batch = 1
    p.edge_multiplier = float(edge)
    return p


def _tables(n, device):
    nbytes = ctypes.c_size_t(0)
    _lib.check(_lib.load().pty_tables_bytes(int(n), ctypes.byref(nbytes)), "tables")
    return torch.empty(int(nbytes.value), dtype=torch.uint8, device=device)


def _edge(edge_slice_thickness_multiplier):
    e = float(edge_slice_thickness_multiplier)
    return e if e > 1 else 1.0


def multislice(full_probe, this_obj_chopped, num_slices, n_obj_modes, n_probe_modes, this_distances,
               edge_slice_thickness_multiplier, this_wavelength, q2, qx, qy, exclude_mask, is_single_dist,
               this_tan_x, this_tan_y, damping_cutoff_multislice, smooth_rolloff, master_propagator_phase_space,
               half_master_propagator_phase_space, mask_clean, waves_multislice, wave, tilt_mode, this_step_tilts,
               default_float, default_complex):
    """Split-step multislice (pypty/multislice_core.py:9-82).

    full_probe (B, N, N, M); this_obj_chopped (B, N, N, S, K) band-limited
    windows; master_propagator_phase_space broadcastable to (N, N).
    Returns (waves_multislice (B, N, N, S, M, K, 1) holding the conjugated
    incident wave of every slice, exit wave (B, N, N, M, K)).
    """
    if not is_single_dist:
        raise NotImplementedError("pypty_amd.multislice: per-slice distances are not on the accelerated path")
    if this_step_tilts:
        raise NotImplementedError("pypty_amd.multislice: tilt gradients are not on the accelerated path")
    kind, template = _arrays.kind_of(full_probe), full_probe
    full_probe, this_obj_chopped = _c64(full_probe), _c64(this_obj_chopped)
    B, N, _, M = full_probe.shape
    S, K = int(num_slices), int(n_obj_modes)
    device = full_probe.device
    probe_b = full_probe.permute(0, 3, 1, 2).contiguous()
    obj_b = this_obj_chopped.permute(0, 3, 4, 1, 2).contiguous()
    H = _c64(master_propagator_phase_space).reshape(N, N).contiguous()
    stash = torch.empty((B, M, K, S, N, N), dtype=torch.complex64, device=device)
    exit_wave = torch.empty((B, M, K, N, N), dtype=torch.complex64, device=device)
    tables = _tables(N, device)
    p = _problem(N, S, K, M, _edge(edge_slice_thickness_multiplier))
    rc = _lib.load().pty_ms_forward(ctypes.byref(p), B, probe_b.data_ptr(), obj_b.data_ptr(), H.data_ptr(),
                                    stash.data_ptr(), exit_wave.data_ptr(), tables.data_ptr(),
                                    _lib.stream_ptr(device))
    _lib.check(rc, "multislice")
    waves = torch.conj(stash.permute(0, 4, 5, 3, 1, 2)).resolve_conj().unsqueeze(-1)
    return (_arrays.like(waves.contiguous(), kind, template),
            _arrays.like(exit_wave.permute(0, 3, 4, 1, 2).contiguous(), kind, template))


def multislice_grads(dLoss_dP_out, waves_multislice, this_obj_chopped, object_grad, tilts_grad, is_single_dist,
                     this_distances, edge_slice_thickness_multiplier, exclude_mask, this_wavelength, q2, qx,
                     this_tan_x, qy, this_tan_y, num_slices, n_obj_modes, tiltind, master_propagator_phase_space,
                     this_step_tilts, damping_cutoff_multislice, smooth_rolloff, tilt_mode, compute_batch,
                     mask_clean, this_step_probe, this_step_obj, this_step_pos_correction, masked_pixels_y,
                     masked_pixels_x, default_float, default_complex, helper_flag_4):
    """Adjoint of ``multislice`` (pypty/multislice_core.py:85-200).

    dLoss_dP_out (B, N, N, M, K) = dL/d(exit)*; waves_multislice as returned
    by ``multislice``; object_grad (Ny, Nx, S, K) accumulates the band-limited
    window gradients at masked_pixels_y/x.  Returns (object_grad,
    interm_probe_grad (B, N, N, M) or None, tilts_grad).
    """
    if not is_single_dist:
        raise NotImplementedError("pypty_amd.multislice_grads: per-slice distances are not on the accelerated path")
    if this_step_tilts:
        raise NotImplementedError("pypty_amd.multislice_grads: tilt gradients are not on the accelerated path")
    kind, template = _arrays.kind_of(dLoss_dP_out), dLoss_dP_out
    caller_grad = object_grad
    dLoss_dP_out, waves_multislice = _c64(dLoss_dP_out), _c64(waves_multislice)
    B, N, _, M, K = dLoss_dP_out.shape
    S = int(num_slices)
    device = dLoss_dP_out.device
    g = dLoss_dP_out.permute(0, 3, 4, 1, 2).contiguous()
    stash = torch.conj(waves_multislice[..., 0].permute(0, 4, 5, 3, 1, 2)).resolve_conj().contiguous()
    obj_b = _c64(this_obj_chopped).permute(0, 3, 4, 1, 2).contiguous()
    H = _c64(master_propagator_phase_space).reshape(N, N).contiguous()
    mask = _arrays.to_torch(mask_clean, dtype=torch.float32).reshape(N, N).contiguous()
    grad_b = torch.empty((B, S, K, N, N), dtype=torch.complex64, device=device)
    tables = _tables(N, device)
    p = _problem(N, S, K, M, _edge(edge_slice_thickness_multiplier))
    rc = _lib.load().pty_ms_adjoint(ctypes.byref(p), B, g.data_ptr(), stash.data_ptr(), obj_b.data_ptr(),
                                    H.data_ptr(), mask.data_ptr(), grad_b.data_ptr(), tables.data_ptr(),
                                    _lib.stream_ptr(device))
    _lib.check(rc, "multislice_grads")
    if this_step_obj > 0:
        og = _c64(object_grad)
        my, mx = _arrays.to_torch(masked_pixels_y), _arrays.to_torch(masked_pixels_x)
        corners = torch.stack((my[:, 0, 0], mx[:, 0, 0]), dim=1).to(torch.int32)
        _scatter_windows(og, grad_b, corners.contiguous())
        object_grad = _writeback(caller_grad, og)
    interm = None
    if helper_flag_4:
        interm = g.sum(dim=2).permute(0, 2, 3, 1).contiguous() if K > 1 else g[:, :, 0].permute(0, 2, 3, 1).contiguous()
        interm = _arrays.like(interm, kind, template)
    return object_grad, interm, tilts_grad


def _writeback(caller, t):
    """The caller's accumulator after an in-place update computed in t: a host array gets t's values
    copied back (in place, as the reference mutates it); a device array viewed by t already has them."""
    if isinstance(caller, np.ndarray):
        np.copyto(caller, t.cpu().numpy())
        return caller
    return caller if not isinstance(caller, torch.Tensor) or caller.data_ptr() == t.data_ptr() else t


def _scatter_windows(full, windows, corners):
    """full (Ny, Nx, S, K) += windows (B, S, K, n, n) placed at corners (B, 2)."""
    ny, nx = full.shape[0], full.shape[1]
    planes = full.shape[2] * full.shape[3]
    B, n = windows.shape[0], windows.shape[-1]
    _lib.require_device_tensor(full, "object_grad")
    if full.dtype != torch.complex64:
        raise TypeError("pypty_amd.scatteradd: complex64 required")
    rc = _lib.load().pty_scatteradd_c64(full.data_ptr(), ny, nx, planes, _lib.PTY_LAYOUT_REFERENCE,
                                        windows.data_ptr(), B, n, corners.data_ptr(), _lib.stream_ptr(full.device))
    _lib.check(rc, "scatteradd")


def scatteradd(full, masky, maskx, chop):
    """full[masky, maskx] += chop with repeated indices accumulating (pypty/multislice_core.py:206-230).

    masky / maskx (B, n, n) are the window index grids the reference builds
    (corner + arange); chop (B, n, n, S, K); full (Ny, Nx, S, K) complex64.
    """
    caller = full
    chop, full = _c64(chop), _c64(full)
    B, n = chop.shape[0], chop.shape[1]
    masky = _arrays.to_torch(masky).broadcast_to((B, n, n))
    maskx = _arrays.to_torch(maskx).broadcast_to((B, n, n))
    corners = torch.stack((masky[:, 0, 0], maskx[:, 0, 0]), dim=1).to(torch.int32).contiguous()
    ar = torch.arange(n, device=chop.device)
    ey = (corners[:, 0, None, None].long() + ar[None, :, None]).expand(B, n, n)
    ex = (corners[:, 1, None, None].long() + ar[None, None, :]).expand(B, n, n)
    if not (torch.equal(masky.long(), ey) and torch.equal(maskx.long(), ex)):
        raise NotImplementedError("pypty_amd.scatteradd: index grids must be contiguous windows")
    windows = chop.reshape(B, n, n, -1).permute(0, 3, 1, 2).contiguous()
    _scatter_windows(full, windows, corners)
    _writeback(caller, full)
