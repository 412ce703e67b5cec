"""Constraint terms of the loss (mirror of pypty/objective.py:1165-1603).

Same names, arguments and results as the reference: each returns (term, gradient), the gradient the
Wirtinger derivative dR/dz* for complex arguments (object, probe) and the plain derivative for real
ones (positions, tilts).  They run on the HIP device: the 2D transforms through this package's FFT
kernels (``pypty_amd.fft``), the rest as device array operations on torch tensors.
``loss_and_direction`` adds them to the loss and gradients in the reference's order
(objective.py:765-921).
"""
from __future__ import annotations

import numpy as np
import torch

from . import fft as pfft


def _roll(t, s, dim):
    return torch.roll(t, shifts=s, dims=dim)


def compute_fast_axis_constraint_on_grid(something, scan_size, tv_reg_weight):
    """Squared second differences along the fast scan axis (objective.py:1165-1198)."""
    g = something.reshape(int(scan_size[0]), int(scan_size[1]), something.shape[-1])
    lap = _roll(g, 1, 1) + _roll(g, -1, 1) - 2 * g
    lap[:, 0] = 0
    lap[:, -1] = 0
    term = tv_reg_weight * torch.sum(lap.double() ** 2)
    lap = lap * (2 * tv_reg_weight)
    grad = -2 * lap + _roll(lap, 1, 1) + _roll(lap, -1, 1)
    return float(term), grad.reshape(something.shape)


def compute_slow_axis_constraint_on_grid(something, scan_size, tv_reg_weight):
    """Squared second differences along the slow axis of the fast-axis sums (objective.py:1201-1238)."""
    g = something.reshape(int(scan_size[0]), int(scan_size[1]), something.shape[-1]).sum(dim=1)
    lap = _roll(g, 1, 0) + _roll(g, -1, 0) - 2 * g
    lap[0, :] = 0
    lap[-1, :] = 0
    term = tv_reg_weight * torch.sum(lap.double() ** 2)
    lap = lap * (2 * tv_reg_weight)
    gs = -2 * lap + _roll(lap, 1, 0) + _roll(lap, -1, 0)
    return float(term), gs[:, None, :].expand(-1, int(scan_size[1]), -1).reshape(something.shape)


def compute_deformation_constraint_on_grid(something, scan_size, reg_weight):
    """Distance of each (y, x) column pair from its least-squares affine fit over the scan grid
    (objective.py:1243-1283); the fit is computed in float64."""
    ny, nx = int(scan_size[0]), int(scan_size[1])
    xp_, yp_ = np.meshgrid(np.arange(nx), np.arange(ny))
    A = np.stack([yp_.ravel(), xp_.ravel(), np.ones(yp_.size)], axis=1)
    m1 = A @ np.linalg.inv(A.T @ A).T
    At = torch.from_numpy(A).to(something.device)
    M1 = torch.from_numpy(m1).to(something.device)
    s = something.double()
    grad = torch.zeros_like(s)
    term = 0.0
    for c in range(something.shape[1] // 2):
        chop = s[:, 2 * c:2 * c + 2]
        r = chop - M1 @ (At.T @ chop)
        term += reg_weight * float(torch.sum(r ** 2))
        r = r * (2 * reg_weight)
        grad[:, 2 * c:2 * c + 2] = r + At @ (M1.T @ (-r))
    return term, grad.to(something.dtype)


def compute_window_constraint(to_reg_probe, current_window, current_window_weight):
    """current_window_weight * sum window |P|^2 (objective.py:1331-1361)."""
    win = torch.as_tensor(np.asarray(current_window) if not isinstance(current_window, torch.Tensor)
                          else current_window, device=to_reg_probe.device, dtype=torch.float32)
    win = (win[:, :, None] if to_reg_probe.dim() == 3 else win[:, :, None, None]) * current_window_weight
    term = torch.sum((to_reg_probe.abs() ** 2 * win).double())
    return float(term), (win * to_reg_probe).to(to_reg_probe.dtype)


def _outside_aperture(shape, aperture, device):
    if isinstance(aperture, (np.ndarray, torch.Tensor)):
        a = torch.as_tensor(np.asarray(aperture.cpu() if isinstance(aperture, torch.Tensor) else aperture),
                            device=device).bool()
        return (~a)[:, :, None].float()
    f = lambda n: torch.fft.fftshift(torch.fft.fftfreq(n, dtype=torch.float64))  # noqa: E731
    fy, fx = torch.meshgrid(f(shape[0]), f(shape[1]), indexing="ij")
    return (((fx ** 2 + fy ** 2) ** 0.5) > 0.5 * float(aperture)).float().to(device)[:, :, None]


def compute_probe_constraint(to_reg_probe, aperture, weight, return_direction=True):
    """weight * sum |shift_fft2(P)| outside the aperture (objective.py:1364-1404): an aperture mask
    (True inside), or a radius ratio r (normalised frequencies beyond r / 2)."""
    outside = _outside_aperture(to_reg_probe.shape, aperture, to_reg_probe.device)
    if to_reg_probe.dim() == 4:
        outside = outside[..., None]
    F = pfft.shift_fft2(to_reg_probe.contiguous(), axes=(0, 1)) * outside
    term = weight * torch.sum(F.abs().double())
    if not return_direction:
        return float(term), None
    G = outside * torch.exp(1j * torch.angle(F)) * (0.5 * weight * F.shape[0] * F.shape[1])
    return float(term), pfft.ifft2_ishift(G.to(torch.complex64).contiguous(), axes=(0, 1))


def _laplace(a):
    return -4 * a + _roll(a, 1, 1) + _roll(a, -1, 1) + _roll(a, 1, 0) + _roll(a, -1, 0)


def probe_modulation_constraint(to_reg_probe, scan_step_A, pixel_size_x_A, pixel_size_y_A, probe_mod_tol=0.01,
                                probe_mod_reg_weight=1, damping_cutoff_multislice=0.63):
    """Squared Laplacian of |shift_fft2(P)| at the spectral points of the scan lattice
    (objective.py:1407-1444)."""
    P = to_reg_probe.contiguous()
    F = pfft.shift_fft2(P, axes=(0, 1))
    Fa = F.abs()
    ffx = np.fft.fftshift(np.fft.fftfreq(P.shape[1], pixel_size_x_A)) * scan_step_A
    ffy = np.fft.fftshift(np.fft.fftfreq(P.shape[0], pixel_size_y_A)) * scan_step_A
    ffx, ffy = np.meshgrid(ffx, ffy, indexing="xy")
    mx = int(np.ceil(0.5 * scan_step_A / pixel_size_x_A))
    t = np.arange(-mx, mx + 1)
    tx, ty = np.meshgrid(t, t, indexing="xy")
    sel = (tx ** 2 + ty ** 2) <= (damping_cutoff_multislice * mx) ** 2
    tx, ty = tx[sel], ty[sel]
    d2 = (tx[:, None, None] - ffx[None]) ** 2 + (ty[:, None, None] - ffy[None]) ** 2
    mask = torch.from_numpy((d2 <= probe_mod_tol ** 2).sum(0).astype(bool)).to(P.device)
    mask = mask[:, :, None] if P.dim() == 3 else mask[:, :, None, None]
    L = mask * _laplace(Fa)
    term = probe_mod_reg_weight * torch.sum(L.double() ** 2)
    L = L * (2 * probe_mod_reg_weight)
    G = (_laplace(L) * 0.5 * torch.exp(1j * torch.angle(F))).to(torch.complex64).contiguous()
    return float(term), pfft.ifft2_ishift(G, axes=(0, 1)) * (P.shape[0] * P.shape[1])


def compute_atv_constraint(obj, atv_weight, atv_q, atv_p, pixel_size_x_A=None, pixel_size_y_A=None,
                           atv_grad_mask=None, return_direction=True, smart_memory=False):
    """Adaptive total variation atv_weight * sum (|dx|^p + |dy|^p + 1e-20)^(q/p) of the periodic
    backward differences (objective.py:1447-1505)."""
    dx = _roll(obj, 1, 1) - obj
    dy = _roll(obj, 1, 0) - obj
    ax, ay = dx.abs(), dy.abs()
    if atv_grad_mask is not None:
        m = torch.as_tensor(atv_grad_mask, device=obj.device, dtype=ax.dtype)[:, :, None, None]
        ax, ay = ax * m, ay * m
    px = ax ** atv_p + ay ** atv_p + 1e-20
    term = atv_weight * torch.sum(px.double() ** (atv_q / atv_p))
    if not return_direction:
        return float(term), torch.zeros_like(obj)
    d = atv_weight * 0.5 * atv_q * px ** (atv_q / atv_p - 1)
    if atv_grad_mask is not None:
        d = d * m
    gx = d * ax ** (atv_p - 1) * torch.exp(1j * torch.angle(dx))
    gy = d * ay ** (atv_p - 1) * torch.exp(1j * torch.angle(dy))
    grad = _roll(gx, -1, 1) - gx + _roll(gy, -1, 0) - gy
    return float(term), grad.to(obj.dtype)


def _dft_axis2(t, inverse):
    """DFT along axis 2 (the slices) of a (Ny, Nx, S, K) array as a product with the S x S matrix."""
    S = t.shape[2]
    k = torch.arange(S, dtype=torch.float64)
    W = torch.exp((2j if inverse else -2j) * np.pi * torch.outer(k, k) / S).to(torch.complex64).to(t.device)
    out = torch.matmul(t.permute(0, 1, 3, 2).contiguous(), W.T.contiguous()).permute(0, 1, 3, 2)
    return (out / S if inverse else out).contiguous()


def compute_missing_wedge_constraint(obj, pixel_size_x_A, pixel_size_y_A, slice_distance, beta_wedge, wedge_mu):
    """sum |fftn(O)|^2 weighted by (2 mu / pi) atan(beta^2 qz^2 / qr^2) (objective.py:1508-1545);
    the 3D transform is the plane FFT of every slice and a DFT along the slices."""
    Ny, Nx, S = obj.shape[:3]
    qx = np.fft.fftfreq(Nx, pixel_size_x_A)
    qy = np.fft.fftfreq(Ny, pixel_size_y_A)
    qz = np.fft.fftfreq(S, slice_distance)
    qx, qy, qz = np.meshgrid(qx, qy, qz)
    qr = qx ** 2 + qy ** 2
    qr[qr == 0] = 1
    w = torch.from_numpy((wedge_mu * 0.63661977236) * np.arctan((beta_wedge ** 2) * (qz ** 2) / qr)).to(obj.device)
    F = _dft_axis2(pfft.fft2(obj.contiguous(), axes=(0, 1)), False)
    term = torch.sum((F.abs() ** 2).double() * w[:, :, :, None])
    G = (F * (w[:, :, :, None] * (Ny * Nx * S)).to(torch.float32)).to(torch.complex64)
    return float(term), pfft.ifft2(_dft_axis2(G, True), axes=(0, 1))


def compute_mixed_object_variance_constraint(this_obj, weight, sigma, return_direction=True, smart_memory=False):
    """Variance over the object modes of the Gaussian-blurred object (objective.py:1548-1603)."""
    Ny, Nx = this_obj.shape[:2]
    f = np.array(np.meshgrid(np.fft.fftfreq(Nx), np.fft.fftfreq(Ny), indexing="xy"))
    mask = torch.from_numpy(np.exp(-np.sum(f ** 2, axis=0) / sigma ** 2).astype(np.float32)).to(this_obj.device)
    blur = pfft.ifft2(pfft.fft2(this_obj.contiguous(), axes=(0, 1)) * mask[:, :, None, None], axes=(0, 1))
    K = this_obj.shape[-1]
    dev = blur - blur.mean(dim=-1, keepdim=True)
    term = weight * torch.sum(dev.abs().double() ** 2) / K
    if not return_direction:
        return float(term), None
    dev = (dev * (weight / K)).contiguous()
    return float(term), pfft.ifft2(pfft.fft2(dev, axes=(0, 1)) * mask[:, :, None, None], axes=(0, 1))
