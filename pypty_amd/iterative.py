"""Object / probe / position update step (mirror of pypty/iterative.py:440-1103).

``bfgs_update`` keeps the reference's 95-argument signature, its result
tuple and its module-global state (``obj``, ``probe``, ``positions``,
``positions_correction``, ``tilts``, ``tilts_correction``,
``static_background``, ``aberrations_array``, ``beam_current``,
``history_bfgs``).  Supported variables: object, probe and sub-pixel
positions (the ones the benchmark configurations refine).  Every large
operation runs on the HIP device through this package's kernels:
loss and gradients (``objective.loss_and_direction``), the band-limits
(``utils.fourier_clean``, plane FFT kernels), the inner products of the
two-loop recursion (``lbfgs.dots``, FP64 accumulation) and the axpy updates
(``lbfgs.axpby``).  Only the Wolfe tests compare host scalars, as in the
reference.
"""
from __future__ import annotations

import sys

import numpy as np
import torch

from . import lbfgs, objective, utils

obj = None
probe = None
positions = None
positions_correction = None
tilts = None
tilts_correction = None
static_background = None
aberrations_array = None
beam_current = None
history_bfgs = None


def reset_bfgs_history():
    """Clear the L-BFGS history (pypty/iterative.py:1055-1103)."""
    global history_bfgs
    history_bfgs = {
        "obj_hist_s": [], "probe_hist_s": [], "positions_hist_s": [], "tilts_hist_s": [], "beam_current_hist_s": [],
        "aberrations_hist_s": [], "static_background_hist_s": [], "obj_hist_y": [], "probe_hist_y": [],
        "positions_hist_y": [], "tilts_hist_y": [], "beam_current_hist_y": [], "aberrations_hist_y": [],
        "static_background_hist_y": [], "rho_hist": [], "prev_loss": None, "prev_sse": None, "obj_grad": None,
        "probe_grad": None, "pos_grad": None, "tilt_grad": None, "beam_current_grad": None, "aberrations_grad": None,
        "static_background_grad": None, "empty_hist": True, "gamma_obj": 1, "gamma_probe": 1, "gamma_pos": 1,
        "gamma_tilts": 1, "gamma_static": 1, "gamma_aberrations": 1, "gamma_beam_current": 1,
    }


reset_bfgs_history()


def _clean(x, cutoff, rolloff):
    return utils.fourier_clean(x, cutoff=cutoff, rolloff=rolloff)


def _neg(x):
    return lbfgs.lincomb(x, -1.0)


def _dot2(a, b):
    """The reference's 2*sum Re(a conj(b)) for complex blocks, sum a*b for real ones."""
    (d,) = lbfgs.dots([(a, b)])
    return 2.0 * d if a.is_complex() else d


def _unsupported(name, cond):
    if cond:
        raise NotImplementedError(f"pypty_amd.bfgs_update: '{name}' is outside the accelerated path")


def bfgs_update(algorithm_type, this_slice_distances, edge_slice_thickness_multiplier, this_step_probe, this_step_obj,
                this_step_pos_correction, this_step_tilts, measured_array, this_wavelength, masks, pixel_size_x_A,
                pixel_size_y_A, phase_norm_weight, abs_norm_weight, stepsize_threshold_low, probe_reg_weight,
                aperture_mask, recon_type, defocus_array, Cs, alpha_near_field, damping_cutoff_multislice,
                smooth_rolloff, update_extra_cut, keep_probe_states_orthogonal, do_charge_flip, cf_delta_phase,
                cf_delta_abs, cf_beta_phase, cf_beta_abs, phase_only_obj, wedge_mu, beta_wedge, wolfe_c1_constant,
                wolfe_c2_constant, atv_weight, atv_q, atv_p, tune_only_probe_phase, propmethod, this_chopped_sequence,
                load_one_by_one, data_multiplier, data_pad, phase_plate_in_h5, print_flag, this_loss_weight,
                max_count, reduce_factor, optimism, mixed_variance_weight, mixed_variance_sigma, data_bin,
                data_shift_vector, smart_memory, default_float, default_complex, default_int, upsample_pattern,
                this_step_static_background, tilt_mode, fancy_sigma, tune_only_probe_abs, aberration_marker,
                this_step_aberrations_array, probe_marker, compute_batch, current_window, current_window_weight,
                dynamically_resize_yx_object, lazy_clean, current_gaussian_filter,
                current_apply_gaussian_filter_amplitude, this_beam_current_step, xp, remove_fft_cache,
                is_first_epoch, hist_length, actual_step, fast_axis_reg_weight_positions, fast_axis_reg_weight_tilts,
                slow_axis_reg_weight_positions, slow_axis_reg_weight_tilts, scan_size, scan_step_A,
                current_deformation_reg_weight_positions, current_deformation_reg_weight_tilts, warnings,
                data_simulation_flag, data_simulation_exitwave, pos_puzzling, probe_mod_reg_weight, probe_mod_tol,
                use_padded_intensity, counter_failed_epochs):
    """One L-BFGS epoch with Wolfe line search (pypty/iterative.py:440-1049).

    Returns (total_loss, sse, constraint_contributions, actual_step, count,
    d_value, new_d_value, warnings, counter_failed_epochs).
    """
    global obj, probe, positions_correction, history_bfgs
    _unsupported("this_step_tilts", bool(this_step_tilts))
    _unsupported("this_step_static_background", bool(this_step_static_background))
    _unsupported("this_step_aberrations_array", bool(this_step_aberrations_array))
    _unsupported("this_beam_current_step", bool(this_beam_current_step))
    _unsupported("phase_only_obj", bool(phase_only_obj))
    _unsupported("tune_only_probe_phase", bool(tune_only_probe_phase))
    _unsupported("tune_only_probe_abs", bool(tune_only_probe_abs))
    _unsupported("do_charge_flip", bool(do_charge_flip))
    _unsupported("keep_probe_states_orthogonal", bool(keep_probe_states_orthogonal) and probe.shape[-1] > 1)
    _unsupported("probe_reg_weight=inf", probe_reg_weight == np.inf)
    _unsupported("gaussian filters", bool(current_gaussian_filter) or bool(current_apply_gaussian_filter_amplitude))
    _unsupported("dynamically_resize_yx_object", bool(dynamically_resize_yx_object) and bool(this_step_pos_correction))
    _unsupported("pos_puzzling", bool(pos_puzzling) and bool(this_step_pos_correction))
    empty_hist = history_bfgs["empty_hist"]
    update_obj = this_step_obj > 0
    update_probe = this_step_probe > 0
    update_pos = this_step_pos_correction > 0
    cut = damping_cutoff_multislice - update_extra_cut
    clean = not lazy_clean

    def evaluate(o, p, pc):
        out = objective.loss_and_direction(
            o, p, positions, pc, tilts, tilts_correction, this_slice_distances, edge_slice_thickness_multiplier,
            measured_array, algorithm_type, this_wavelength, update_probe, update_obj, update_pos, False, masks,
            pixel_size_x_A, pixel_size_y_A, recon_type, Cs, defocus_array, alpha_near_field, damping_cutoff_multislice,
            0, propmethod, this_chopped_sequence, load_one_by_one, data_multiplier, data_pad, phase_plate_in_h5,
            this_loss_weight, data_bin, data_shift_vector, upsample_pattern, None, False, tilt_mode,
            aberration_marker, probe_marker, None, compute_batch, phase_only_obj, None, False, False, default_float,
            default_complex, xp, is_first_epoch, scan_size, scan_step_A, fast_axis_reg_weight_positions,
            slow_axis_reg_weight_positions, slow_axis_reg_weight_tilts, current_deformation_reg_weight_positions,
            current_deformation_reg_weight_tilts, fast_axis_reg_weight_tilts, aperture_mask, 0, current_window_weight,
            current_window, phase_norm_weight, abs_norm_weight, atv_weight, atv_q, atv_p, mixed_variance_weight,
            mixed_variance_sigma, wedge_mu, beta_wedge, probe_mod_reg_weight, probe_mod_tol, smart_memory,
            print_flag, data_simulation_flag, data_simulation_exitwave, use_padded_intensity)
        loss, sse, og, pg, posg = out[0], out[1], out[2], out[3], out[4]
        if update_probe:
            pg = _clean(pg, cut, smooth_rolloff)
        if clean and update_obj:
            og = _clean(og, cut, smooth_rolloff)
        return loss, sse, og, pg, posg, out[9]

    if update_probe:
        probe = _clean(probe, cut, smooth_rolloff)
    if update_obj and clean:
        obj = _clean(obj, cut, smooth_rolloff)

    def direction_dot(uo, up, ux, go, gp, gx):
        d = 0.0
        if update_obj:
            d += _dot2(uo, go)
        if update_probe:
            d += _dot2(up, gp)
        if update_pos:
            d += _dot2(ux, gx)
        return d

    if empty_hist:
        total_loss, this_sse, og, pg, posg, cc = evaluate(obj, probe, positions_correction)
        u_obj = _neg(og) if update_obj else None
        d_val_obj = _dot2(u_obj, u_obj) if update_obj else 0.0
        if d_val_obj == 0:
            d_val_obj = 1.0
        u_probe = lbfgs.lincomb(pg, -d_val_obj / _dot2(pg, pg)) if update_probe else None
        u_pos = lbfgs.lincomb(posg, -d_val_obj / _dot2(posg, posg)) if update_pos else None
        history_bfgs["empty_hist"] = False
    else:
        total_loss, this_sse = history_bfgs["prev_loss"], history_bfgs["prev_sse"]
        og, pg, posg = history_bfgs["obj_grad"], history_bfgs["probe_grad"], history_bfgs["pos_grad"]
        cc = history_bfgs["constraint_contributions"]
        if update_probe:
            pg = _clean(pg, cut, smooth_rolloff)
        rhos = history_bfgs["rho_hist"]
        q_obj = og.clone() if update_obj else None
        q_probe = pg.clone() if update_probe else None
        q_pos = posg.clone() if update_pos else None
        alphas = []
        for i in range(len(rhos) - 1, -1, -1):
            siTq = direction_dot(history_bfgs["obj_hist_s"][i], history_bfgs["probe_hist_s"][i],
                                 history_bfgs["positions_hist_s"][i], q_obj, q_probe, q_pos)
            a = rhos[i] * siTq
            alphas.append(a)
            if update_obj:
                lbfgs.axpby(q_obj, history_bfgs["obj_hist_y"][i], 1.0, -a)
            if update_probe:
                lbfgs.axpby(q_probe, history_bfgs["probe_hist_y"][i], 1.0, -a)
            if update_pos:
                lbfgs.axpby(q_pos, history_bfgs["positions_hist_y"][i], 1.0, -a)
        alphas = alphas[::-1]
        for q, key, on in ((q_obj, "gamma_obj", update_obj), (q_probe, "gamma_probe", update_probe),
                           (q_pos, "gamma_pos", update_pos)):
            g = history_bfgs[key]
            if on and g != 0 and g == g:
                lbfgs.axpby(q, None, g, 0.0)
        for i in range(len(rhos)):
            yiTr = direction_dot(history_bfgs["obj_hist_y"][i], history_bfgs["probe_hist_y"][i],
                                 history_bfgs["positions_hist_y"][i], q_obj, q_probe, q_pos)
            beta = alphas[i] - rhos[i] * yiTr
            if update_obj:
                lbfgs.axpby(q_obj, history_bfgs["obj_hist_s"][i], 1.0, beta)
            if update_probe:
                lbfgs.axpby(q_probe, history_bfgs["probe_hist_s"][i], 1.0, beta)
            if update_pos:
                lbfgs.axpby(q_pos, history_bfgs["positions_hist_s"][i], 1.0, beta)
        u_obj = lbfgs.axpby(q_obj, None, -1.0, 0.0) if update_obj else None
        u_probe = lbfgs.axpby(q_probe, None, -1.0, 0.0) if update_probe else None
        u_pos = lbfgs.axpby(q_pos, None, -1.0, 0.0) if update_pos else None
        if clean and update_obj:
            u_obj = _clean(u_obj, cut, smooth_rolloff)
    d_value = direction_dot(u_obj, u_probe, u_pos, og, pg, posg)
    if d_value >= 0:
        sys.stdout.write("\n\n\nWarning! Positive dir-derivative, using normal gradient-descent!\n\n\n\n\n\n")
        warnings += "\n\\Positive dir-derivative, using normal gradient-descent!"
        u_obj = _neg(og) if update_obj else None
        u_probe = _neg(pg) if update_probe else None
        u_pos = _neg(posg) if update_pos else None
        for u, key, on in ((u_obj, "gamma_obj", update_obj), (u_probe, "gamma_probe", update_probe),
                           (u_pos, "gamma_pos", update_pos)):
            g = history_bfgs[key]
            if on and g > 0 and g == g:
                lbfgs.axpby(u, None, g, 0.0)
        d_value = direction_dot(u_obj, u_probe, u_pos, og, pg, posg)
    count = -1
    new_d_value = d_value
    this_wolfe_1 = False
    while True:
        count += 1
        if update_obj:
            new_obj = lbfgs.axpby(obj.clone(), u_obj, 1.0, actual_step)
            if clean:
                new_obj = _clean(new_obj, cut, smooth_rolloff)
        else:
            new_obj = obj
        new_probe = lbfgs.axpby(probe.clone(), u_probe, 1.0, actual_step) if update_probe else probe
        new_pc = lbfgs.axpby(positions_correction.clone(), u_pos, 1.0, actual_step) if update_pos \
            else positions_correction
        new_total_loss, new_sse, nog, npg, nposg, ncc = evaluate(new_obj, new_probe, new_pc)
        new_d_value = direction_dot(u_obj, u_probe, u_pos, nog, npg, nposg)
        this_wolfe_1 = utils.wolfe_1(total_loss, new_total_loss, d_value, actual_step, wolfe_c1_constant)
        this_wolfe_2 = utils.wolfe_2(d_value, new_d_value, wolfe_c2_constant)
        if print_flag >= 3:
            sys.stdout.write("\nLinesearch iteration %d. This loss is %.3e. Loss change is %.3e. Dir-derivative is "
                             "%.3e. New dir-derivative is %.3e, This step is %.3e." % (
                                 count, total_loss, total_loss - new_total_loss, d_value, new_d_value, actual_step))
            sys.stdout.flush()
        if this_wolfe_1 and this_wolfe_2:
            break
        if not this_wolfe_1 and this_wolfe_2:
            actual_step *= reduce_factor
        if not this_wolfe_2 and this_wolfe_1:
            actual_step *= optimism
        if not (this_wolfe_2 or this_wolfe_1):
            actual_step = type(actual_step)(10 ** (np.random.uniform(0, 20) - 20))
        if max_count is not None and count > max_count:
            break
        if actual_step < stepsize_threshold_low:
            break
    if not this_wolfe_1:
        sys.stdout.write("\nWARNING! The sufficient loss descrese is not achieved, the update is regected, keeping "
                         "the the same step! Resetting the history! Please terminate if this message appears during "
                         "the next epoch!\n")
        warnings += "\nSufficient loss descrese is not achieved, the update is regected"
        counter_failed_epochs += 1
        reset_bfgs_history()
        return total_loss, this_sse, cc, actual_step, count, d_value, new_d_value, warnings, counter_failed_epochs
    counter_failed_epochs = 0
    if print_flag >= 2:
        sys.stdout.write("\n-->Update done with %d linesearch steps! This loss is %.3e. Loss change is %.3e. "
                         "Dir-derivative is %.3e. New dir-derivative is %.3e." % (
                             count, total_loss, total_loss - new_total_loss, d_value, new_d_value))
    h = history_bfgs
    s_obj = lbfgs.lincomb(u_obj, actual_step) if update_obj else 0
    s_probe = lbfgs.lincomb(u_probe, actual_step) if update_probe else 0
    s_pos = lbfgs.lincomb(u_pos, actual_step) if update_pos else 0
    y_obj = lbfgs.axpby(nog.clone(), og, 1.0, -1.0) if update_obj else 0
    y_probe = lbfgs.axpby(npg.clone(), pg, 1.0, -1.0) if update_probe else 0
    y_pos = lbfgs.axpby(nposg.clone(), posg, 1.0, -1.0) if update_pos else 0
    for key, v in (("obj_hist_s", s_obj), ("probe_hist_s", s_probe), ("positions_hist_s", s_pos),
                   ("obj_hist_y", y_obj), ("probe_hist_y", y_probe), ("positions_hist_y", y_pos)):
        h[key].append(v)
    for key in ("tilts_hist_s", "beam_current_hist_s", "aberrations_hist_s", "static_background_hist_s",
                "tilts_hist_y", "beam_current_hist_y", "aberrations_hist_y", "static_background_hist_y"):
        h[key].append(0)
    h["prev_loss"], h["prev_sse"] = new_total_loss, new_sse
    h["obj_grad"], h["probe_grad"], h["pos_grad"] = nog, npg, nposg
    h["constraint_contributions"] = ncc
    yiTsi_tot = 0.0
    for y, s, key, on in ((y_obj, s_obj, "gamma_obj", update_obj), (y_probe, s_probe, "gamma_probe", update_probe),
                          (y_pos, s_pos, "gamma_pos", update_pos)):
        if on:
            (yy,) = lbfgs.dots([(y, y)])
            ys = _dot2(y, s)
            yiTsi_tot += ys
            h[key] = ys / yy
    h["rho_hist"].append(1.0 / yiTsi_tot)
    while len(h["obj_hist_s"]) > hist_length:
        for key in list(h):
            if key.endswith("_hist_s") or key.endswith("_hist_y") or key == "rho_hist":
                h[key].pop(0)
    probe = new_probe
    obj = new_obj
    if update_pos:
        positions_correction = new_pc
    return total_loss, this_sse, cc, actual_step, count, d_value, new_d_value, warnings, counter_failed_epochs
