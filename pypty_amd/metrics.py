"""Reconstruction quality: normalised mean-squared error of the projected phase.

The reference defines no error metric.  NMSE here is
    min_c ||phi_rec - phi_true - c||^2 / ||phi_true - mean(phi_true)||^2
over the scanned field of view shrunk by half a window on every side, where
phi is the phase summed over slices (and object modes), and the constant c
removes the global phase offset that ptychography cannot determine.
"""
from __future__ import annotations

import numpy as np


def projected_phase(obj) -> np.ndarray:
    """Sum over slices and modes of the object's phase, (Ny, Nx) float64."""
    o = obj.detach().cpu().numpy() if hasattr(obj, "detach") else np.asarray(obj)
    return np.angle(o).sum(axis=(2, 3)).astype(np.float64)


def fov_mask(shape, positions_px, window):
    """Boolean mask of pixels inside [min+N/2, max+N/2] of the scanned window corners."""
    pos = np.asarray(positions_px, dtype=np.float64)
    y0, x0 = pos.min(axis=0) + window / 2
    y1, x1 = pos.max(axis=0) + window / 2
    m = np.zeros(shape, bool)
    m[int(np.ceil(y0)):int(np.floor(y1)) + 1, int(np.ceil(x0)):int(np.floor(x1)) + 1] = True
    return m


def nmse(phase_rec, phase_true, mask) -> float:
    d = (phase_rec - phase_true)[mask]
    d = d - d.mean()
    t = phase_true[mask]
    t = t - t.mean()
    return float(np.sum(d * d) / np.sum(t * t))
