"""Starting object from a direct method and restart from checkpoints (mirror of
pypty/initialize.py:629-749 and :1244-1384).

``get_ptycho_obj_from_scan`` interpolates a phase (and amplitude) map sampled at the scan positions
-- the output of ``dpc.fft_based_dpc`` or of another direct method -- onto the object's pixel grid
and spreads it over the slices: the plain start that the learned warm start competes with.
``prepare_restart`` rebuilds the parameter dict of a reconstruction from the files ``run()`` wrote
(``params.pkl``, ``loss.csv``, ``co/cp/cg/cs/cb.npy`` or the per-epoch ``checkpoint_*_epoch_N.npy``).
Both are host set-up code, as in the reference: scipy's Clough-Tocher interpolation and numpy file
reads over arrays of the scan's and the object's size, once per reconstruction.
"""
from __future__ import annotations

import csv
import os
import pickle

import numpy as np


def get_ptycho_obj_from_scan(params, num_slices=None, array_phase=None, array_abs=None, scale_phase=1, scale_abs=1,
                             scan_array_A=None, fill_value_type=None):
    """Object guess (Ny, Nx, S, 1) from phase / amplitude maps at the scan positions
    (initialize.py:629-749); sets params["obj"] and params["slice_distances"] = [total_thickness / S]
    and returns params.

    The window size comes from params["aperture_mask"] (a 2D array) or else params["probe"]: the
    reference's first choice, the dataset's pattern size, reads ``data_pad`` before assigning it
    (:665-681), so that branch always falls through to these two.  The maps are flattened in scan
    order and placed at the positions (rounded, in pixels; or scan_array_A in Angstrom) shifted by
    half a window, then interpolated with scipy's CloughTocher2DInterpolator; pixels outside the
    scan's hull take fill_value_type's value (0, "edge" or "median").  Each slice gets
    |abs|^(1/S) exp(i phase scale_phase / S) times scale_abs; num_slices "auto" takes
    ceil(ptp(phase) / (0.75 pi)) slices."""
    from scipy.interpolate import CloughTocher2DInterpolator

    aperture = params.get("aperture_mask", None)
    if isinstance(aperture, np.ndarray) and aperture.ndim >= 2:
        psx, psy = aperture.shape[1], aperture.shape[0]
    else:
        probe = np.asarray(params.get("probe", None))
        psx, psy = probe.shape[1], probe.shape[0]
    if num_slices is None:
        num_slices = params.get("num_slices", 1)
    if num_slices == "auto":
        num_slices = int(np.ceil((np.max(array_phase) - np.min(array_phase)) / (0.75 * np.pi)))
    positions = np.copy(np.array(params.get("positions", [[0, 0]])))
    sequence = params.get("sequence", None)
    pixel_size_x_A = params.get("pixel_size_x_A", 1)
    total_thickness = params.get("total_thickness", 1)
    use_full_FOV = params.get("use_full_FOV", True)
    positions = np.round(positions).astype(int)
    if scan_array_A is None:
        scan_y, scan_x = positions[:, 0], positions[:, 1]
    else:
        scan_y, scan_x = scan_array_A[:, 0] / pixel_size_x_A, scan_array_A[:, 1] / pixel_size_x_A
    if not use_full_FOV:
        scan_y, scan_x = scan_y - np.min(scan_y), scan_x - np.min(scan_x)
    scan_y, scan_x = (scan_y + psy // 2).astype(int), (scan_x + psx // 2).astype(int)
    max_y_px, max_x_px = np.max(positions[:, 0]) + psy, np.max(positions[:, 1]) + psx
    im_X, im_Y = np.meshgrid(np.arange(max_x_px), np.arange(max_y_px), indexing="xy")
    points = list(zip(scan_x, scan_y))

    def fill(a, edge):
        if fill_value_type == "median":
            return np.median(a)
        if fill_value_type == "edge":
            return edge(a)
        return 0

    if array_phase is None:
        phase_ptycho = 0
    else:
        fv = fill(array_phase, lambda a: np.max([np.min(a[:3, :]), np.min(a[-3:, :]), np.min(a[:, :3]),
                                                 np.min(a[:, -3:])]))
        phase_ptycho = CloughTocher2DInterpolator(points, np.asarray(array_phase).flatten(), fill_value=fv)(im_X, im_Y)
    if array_abs is None:
        abs_ptycho = 1
    else:
        fv = fill(array_abs, lambda a: np.mean([np.mean(a[:1, :]), np.mean(a[-1:, :]), np.mean(a[:, :1]),
                                                np.mean(a[:, -1:])]))
        abs_ptycho = CloughTocher2DInterpolator(points, np.asarray(array_abs).flatten(), fill_value=fv)(im_X, im_Y)
    ptycho = scale_abs * (abs_ptycho ** (1 / num_slices)) * np.exp(1j * phase_ptycho * scale_phase / num_slices)
    if sequence is not None and not use_full_FOV:
        seq = np.asarray(sequence)
        minx, miny = np.min(positions[seq, 1]), np.min(positions[seq, 0])
        maxx, maxy = np.max(positions[seq, 1]), np.max(positions[seq, 0])
        ptycho = ptycho[miny:maxy + psy, minx:maxx + psx]
    ptycho = np.expand_dims(np.tile(np.expand_dims(ptycho, -1), (num_slices)), -1)
    params["slice_distances"] = np.array([total_thickness / num_slices])
    params["obj"] = ptycho
    return params


def _last_true(flag, epoch_prev):
    """The epoch of the newest checkpoint a flag stands for (initialize.py:1277-1300): epoch_prev when
    the flag is set at all, else -1.  (The reference tests truthiness first, so its branch for
    lambda-valued flags, which would look for the last epoch the lambda selects, is never reached;
    a lambda, or its source text from params.pkl, counts as set.)"""
    return epoch_prev if flag else -1


def prepare_restart(output_folder, new_folder=None):
    """Parameter dict for resuming the reconstruction in output_folder (initialize.py:1244-1384).

    epoch_prev is the epoch of loss.csv's last row.  The newest checkpoint is the per-epoch
    checkpoint_*_epoch_N.npy files when save_checkpoints_every_epoch fired later than
    save_inter_checkpoints, else the overwritten co/cp/cg/ct/cs/cb/ca.npy; whichever of them exist
    are loaded into "obj", "probe", "postions" (the reference's key, :1372: the refined positions are
    therefore not taken up by a restart), "tilts", "static_background", "beam_current" and
    "aberrations_array"; epoch_prev becomes that checkpoint's epoch, "skip_preprocessing" is set and
    output_folder changes to new_folder when given.  Raises ValueError when neither flag has fired.

    The reference reads epoch_prev as the csv's text and then compares it with an integer (-1),
    which raises TypeError under Python 3 whenever save_inter_checkpoints is True (its default);
    here it is read as an integer."""
    with open(os.path.join(output_folder, "loss.csv"), "r", newline="") as fh:
        rows = [r for r in csv.reader(fh)]
    epoch_prev = int(float(rows[-1][0]))
    with open(os.path.join(output_folder, "params.pkl"), "rb") as fh:
        params = pickle.load(fh)
    last_true_inter = _last_true(params.get("save_inter_checkpoints", True), epoch_prev)
    last_true_check = _last_true(params.get("save_checkpoints_every_epoch", False), epoch_prev)
    if last_true_check == -1 and last_true_inter == -1:
        raise ValueError("You don't have any saved checkpoints for this reconstruction to restart from!")

    def load(name):
        path = os.path.join(output_folder, name)
        return np.load(path) if os.path.exists(path) else None

    if last_true_check > last_true_inter:
        epoch_prev = last_true_check
        e = last_true_check
        bc, o, p = (load(f"checkpoint_beam_current_epoch_{e}.npy"), load(f"checkpoint_obj_epoch_{e}.npy"),
                    load(f"checkpoint_probe_epoch_{e}.npy"))
        g, t, st = (load(f"checkpoint_positions_epoch_{e}.npy"), load(f"checkpoint_tilts_epoch_{e}.npy"),
                    load(f"checkpoint_stback_epoch_{e}.npy"))
        ab = load(f"checkpoint_aberrations_epcoh_{e}.npy")  # the reference's file name
    else:
        epoch_prev = last_true_inter
        bc, o, p, g, t, st, ab = (load(n) for n in ("cb.npy", "co.npy", "cp.npy", "cg.npy", "ct.npy", "cs.npy",
                                                     "ca.npy"))
    params["epoch_prev"] = epoch_prev
    for key, v in (("obj", o), ("probe", p), ("beam_current", bc), ("postions", g), ("tilts", t),
                   ("static_background", st), ("aberrations_array", ab)):
        if v is not None:
            params[key] = v
    params["skip_preprocessing"] = True
    if new_folder is not None:
        params["output_folder"] = new_folder
    return params
