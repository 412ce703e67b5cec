"""numpy restatement of PyPty's constraint terms (test oracle only).

Each function follows the reference definition it cites (pypty/objective.py)
and returns (term, gradient) in the reference's convention: the Wirtinger
derivative dR/dz* for complex arguments, the plain derivative for real ones.
loss_and_direction adds each term to the loss and each gradient to the
object, probe or position gradient (objective.py:765-921).
"""
from __future__ import annotations

import numpy as np

from . import ms_oracle as orc


def fast_axis_constraint_on_grid(something, scan_size, w):
    """objective.py:1165-1198: second differences along the fast scan axis (interior points)."""
    g = something.reshape(scan_size[0], scan_size[1], something.shape[-1])
    lap = np.roll(g, 1, axis=1) + np.roll(g, -1, axis=1) - 2 * g
    lap[:, 0] = 0
    lap[:, -1] = 0
    term = w * np.sum(lap ** 2)
    lap = lap * (2 * w)
    grad = -2 * lap + np.roll(lap, 1, axis=1) + np.roll(lap, -1, axis=1)
    return term, grad.reshape(something.shape)


def slow_axis_constraint_on_grid(something, scan_size, w):
    """objective.py:1201-1238: second differences along the slow axis of the fast-axis sums."""
    g = something.reshape(scan_size[0], scan_size[1], something.shape[-1]).sum(axis=1)
    lap = np.roll(g, 1, axis=0) + np.roll(g, -1, axis=0) - 2 * g
    lap[0, :] = 0
    lap[-1, :] = 0
    term = w * np.sum(lap ** 2)
    lap = lap * (2 * w)
    gs = -2 * lap + np.roll(lap, 1, axis=0) + np.roll(lap, -1, axis=0)
    return term, np.repeat(gs[:, None, :], scan_size[1], axis=1).reshape(something.shape)


def deformation_constraint_on_grid(something, scan_size, w):
    """objective.py:1243-1283: distance of each (y, x) pair of columns from its least-squares affine
    fit over the scan grid."""
    xp_, yp_ = np.meshgrid(np.arange(scan_size[1]), np.arange(scan_size[0]))
    A = np.stack([yp_.ravel(), xp_.ravel(), np.ones(yp_.size)], axis=1)
    m1 = A @ np.linalg.inv(A.T @ A).T
    term = 0.0
    grad_full = np.zeros_like(something)
    for c in range(something.shape[1] // 2):
        chop = something[:, 2 * c:2 * c + 2]
        fit = m1 @ (A.T @ chop)
        r = chop - fit
        term += w * np.sum(r ** 2)
        r = r * (2 * w)
        grad_full[:, 2 * c:2 * c + 2] = r + A @ (m1.T @ (-r))
    return term, grad_full


def window_constraint(probe, window, w):
    """objective.py:1331-1361: w * sum window |P|^2 over the real-space window."""
    win = window[:, :, None] * w if probe.ndim == 3 else window[:, :, None, None] * w
    return np.sum(np.abs(probe) ** 2 * win), win * probe


def probe_constraint(probe, aperture, w):
    """objective.py:1364-1404: w * sum |shift_fft2(P)| outside the aperture (a mask, or a radius
    ratio r: normalised frequencies beyond r/2)."""
    if isinstance(aperture, np.ndarray):
        outside = 1 - np.expand_dims(aperture.astype(bool), -1)
    else:
        fx = np.fft.fftshift(np.fft.fftfreq(probe.shape[1]))
        fy = np.fft.fftshift(np.fft.fftfreq(probe.shape[0]))
        fx, fy = np.meshgrid(fx, fy, indexing="xy")
        outside = np.expand_dims((fx ** 2 + fy ** 2) ** 0.5, -1) > (0.5 * aperture)
    F = orc.shift_fft2(probe, axes=(0, 1)) * outside
    term = w * np.sum(np.abs(F))
    G = outside * np.exp(1j * np.angle(F)) * (0.5 * w * F.shape[0] * F.shape[1])
    return term, orc.ifft2_ishift(G, axes=(0, 1))


def probe_modulation_constraint(probe, scan_step_A, px, py, tol, w, cutoff):
    """objective.py:1407-1444: squared Laplacian of |shift_fft2(P)| at the spectral points of the
    scan lattice (within tol, inside cutoff)."""
    F = orc.shift_fft2(probe, axes=(0, 1))
    Fa = np.abs(F)
    ffx = np.fft.fftshift(np.fft.fftfreq(probe.shape[1], px)) * scan_step_A
    ffy = np.fft.fftshift(np.fft.fftfreq(probe.shape[0], py)) * scan_step_A
    ffx, ffy = np.meshgrid(ffx, ffy, indexing="xy")
    mx = int(np.ceil(0.5 * scan_step_A / px))
    t = np.arange(-mx, mx + 1)
    tx, ty = np.meshgrid(t, t, indexing="xy")
    sel = (tx ** 2 + ty ** 2) <= (cutoff * mx) ** 2
    tx, ty = tx[sel], ty[sel]
    d2 = (tx[:, None, None] - ffx[None]) ** 2 + (ty[:, None, None] - ffy[None]) ** 2
    mask = (d2 <= tol ** 2).sum(0).astype(bool)[:, :, None]

    def lap(a):
        return -4 * a + np.roll(a, 1, axis=1) + np.roll(a, -1, axis=1) + np.roll(a, 1, axis=0) + np.roll(a, -1, axis=0)

    L = mask * lap(Fa)
    term = w * np.sum(L ** 2)
    L = L * (2 * w)
    G = lap(L) * 0.5 * np.exp(1j * np.angle(F))
    return term, orc.ifft2_ishift(G, axes=(0, 1)) * probe.shape[0] * probe.shape[1]


def atv_constraint(obj, w, q, p):
    """objective.py:1447-1505 (no mask): w * sum (|dx|^p + |dy|^p + 1e-20)^(q/p) of the backward
    differences (periodic)."""
    dx = np.roll(obj, 1, axis=1) - obj
    dy = np.roll(obj, 1, axis=0) - obj
    ax, ay = np.abs(dx), np.abs(dy)
    ex, ey = np.exp(1j * np.angle(dx)), np.exp(1j * np.angle(dy))
    term_px = ax ** p + ay ** p + 1e-20
    term = w * np.sum(term_px ** (q / p))
    d = w * 0.5 * q * term_px ** (q / p - 1)
    gx = d * ax ** (p - 1) * ex
    gy = d * ay ** (p - 1) * ey
    grad = np.roll(gx, -1, axis=1) - gx + np.roll(gy, -1, axis=0) - gy
    return term, grad


def missing_wedge_constraint(obj, px, py, dz, beta, mu):
    """objective.py:1508-1545: sum |fftn(O)|^2 weighted by (2 mu / pi) atan(beta^2 qz^2 / qr^2)."""
    qx = np.fft.fftfreq(obj.shape[1], px)
    qy = np.fft.fftfreq(obj.shape[0], py)
    qz = np.fft.fftfreq(obj.shape[2], dz)
    qx, qy, qz = np.meshgrid(qx, qy, qz)
    qr = qx ** 2 + qy ** 2
    qr[qr == 0] = 1
    wgt = (mu * 0.63661977236) * np.arctan((beta ** 2) * (qz ** 2) / qr)
    F = np.fft.fftn(obj, axes=(0, 1, 2))
    term = np.sum(np.abs(F) ** 2 * wgt[:, :, :, None])
    wgt = wgt * (F.shape[0] * F.shape[1] * F.shape[2])
    return term, np.fft.ifftn(F * wgt[:, :, :, None], axes=(0, 1, 2))


def mixed_object_variance_constraint(obj, w, sigma):
    """objective.py:1548-1603: variance over the object modes of the Gaussian-blurred object."""
    f = np.array(np.meshgrid(np.fft.fftfreq(obj.shape[1]), np.fft.fftfreq(obj.shape[0]), indexing="xy"))
    mask = np.exp(-np.sum(f ** 2, axis=0) / sigma ** 2)
    blur = np.fft.ifft2(np.fft.fft2(obj, axes=(0, 1)) * mask[:, :, None, None], axes=(0, 1))
    K = obj.shape[-1]
    dev = blur - blur.mean(axis=-1)[..., None]
    term = w * np.sum(np.abs(dev) ** 2) / K
    dev = dev * (w / K)
    grad = np.fft.ifft2(np.fft.fft2(dev, axes=(0, 1)) * mask[:, :, None, None], axes=(0, 1))
    return term, grad


def charge_flip(a, delta_phase=0.03, delta_abs=0.14, beta_phase=-0.95, beta_abs=-0.95):
    """objective.py:1014-1110 without fancy_sigma: per slice and mode (axes 2, 3), the phase and the
    absorption -log(|a| + 1e-10) below delta times their largest value are multiplied by beta, and the
    object becomes exp(-absorption + i phase).  Works in a's precision, as the reference."""
    a = np.asarray(a)
    fd = np.float32 if a.dtype == np.complex64 else np.float64
    phase = np.angle(a).astype(fd)
    absorption = (-np.log(np.abs(a) + fd(1e-10))).astype(fd)
    for i in range(phase.shape[-1]):
        for j in range(phase.shape[-2]):
            p = phase[:, :, j, i]
            q = absorption[:, :, j, i]
            p[p < fd(delta_phase) * np.max(p)] *= fd(beta_phase)
            q[q < fd(delta_abs) * np.max(q)] *= fd(beta_abs)
    return np.exp(-absorption + 1j * phase).astype(a.dtype)


def make_states_orthogonal(probe_states):
    """objective.py:1119-1140: Gram-Schmidt over the last axis, mode i minus its projections on the
    modes j < i already orthogonalised, sum(conj(p_j) p_i) / (1e-10 + sum |p_j|^2)."""
    p = np.array(probe_states, copy=True)
    for i in range(p.shape[-1]):
        for j in range(i):
            c = np.sum(np.conj(p[:, :, j]) * p[:, :, i]) / (1e-10 + np.sum(np.abs(p[:, :, j]) ** 2))
            p[:, :, i] = p[:, :, i] - p[:, :, j] * c
    return p
