"""Object / probe / position update step (mirror of pypty/iterative.py:440-1103).

``bfgs_update`` keeps the reference's 95-argument signature, its result
tuple and its module-global state (``obj``, ``probe``, ``positions``,
``positions_correction``, ``tilts``, ``tilts_correction``,
``static_background``, ``aberrations_array``, ``beam_current``,
``history_bfgs``).  Supported variables: object, probe and sub-pixel
positions (the ones the benchmark configurations refine), the static
background held in ``static_background`` (in every loss evaluation; refined
with ``this_step_static_background``), the per-position beam current held
in ``beam_current`` (likewise; refined with ``this_beam_current_step``) and
the aberration rows held in ``aberrations_array`` with ``aberration_marker``
(likewise; refined with ``this_step_aberrations_array``) and the tilts held in
``tilts`` (refined through ``tilts_correction`` with ``this_step_tilts``).
``keep_probe_states_orthogonal`` and ``do_charge_flip`` (without fancy_sigma)
act as in iterative.py:497-505.
Every large
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

from . import _arrays, _lib, lbfgs, objective, utils

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


def _phase_step(o, u, step):
    """|o| exp(i (arg o + step u)): a step of a phase-only object (pypty/iterative.py:737-738)."""
    out = torch.empty_like(o)
    rc = _lib.load().pty_phase_step_c64(o.data_ptr(), u.contiguous().data_ptr(), float(step), out.data_ptr(),
                                        o.numel(), _lib.stream_ptr(o.device))
    _lib.check(rc, "phase_step")
    return out


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
                use_padded_intensity, counter_failed_epochs, *, checkpoint_slices=0):
    """One L-BFGS epoch with Wolfe line search (pypty/iterative.py:440-1049).

    checkpoint_slices (keyword only, not in the reference): passed to every
    loss_and_direction call of the epoch.

    Returns (total_loss, sse, constraint_contributions, actual_step, count,
    d_value, new_d_value, warnings, counter_failed_epochs).
    """
    global obj, probe, positions_correction, tilts, tilts_correction, static_background, beam_current
    global aberrations_array, history_bfgs
    phase_only = bool(phase_only_obj)
    _unsupported("tune_only_probe_phase", bool(tune_only_probe_phase))
    _unsupported("tune_only_probe_abs", bool(tune_only_probe_abs))
    _unsupported("charge flipping with fancy_sigma", bool(do_charge_flip) and fancy_sigma is not None)
    _unsupported("gaussian filters", bool(current_gaussian_filter) or bool(current_apply_gaussian_filter_amplitude))
    _unsupported("dynamically_resize_yx_object", bool(dynamically_resize_yx_object) and bool(this_step_pos_correction))
    _unsupported("pos_puzzling", bool(pos_puzzling) and bool(this_step_pos_correction))
    empty_hist = history_bfgs["empty_hist"]
    update_obj = this_step_obj > 0
    update_probe = this_step_probe > 0
    update_pos = this_step_pos_correction > 0
    # the tilts, a real (rows, 6) block refined through tilts_correction (iterative.py:477, 568, 588-670,
    # 687, 768, 852-868, 892-897, 1027); in every loss evaluation, as there
    update_tilts = this_step_tilts > 0
    if tilts is None:
        tilts = np.zeros((1, 6), np.float32)
    tilts = _arrays.to_torch(tilts, obj.device, torch.float32).reshape(-1, 6).contiguous()
    if tilts_correction is None:
        tilts_correction = torch.zeros_like(tilts)
    tilts_correction = _arrays.to_torch(tilts_correction, obj.device, torch.float32).reshape(-1, 6).contiguous()
    # the static background, a real (N, N) block refined like the positions (iterative.py:480, 569, 604-671,
    # 688, 769, 855-862, 898-903, 1030)
    update_bg = this_step_static_background > 0
    if update_bg and static_background is None:
        raise ValueError("pypty_amd.bfgs_update: refining the static background needs a static_background")
    if static_background is not None and not isinstance(static_background, torch.Tensor) \
            and not np.isscalar(static_background):
        static_background = torch.as_tensor(np.asarray(static_background, np.float32), device=obj.device)
    if update_bg and not isinstance(static_background, torch.Tensor):
        static_background = torch.full((probe.shape[0], probe.shape[1]), float(static_background),
                                       dtype=torch.float32, device=obj.device)
    # the beam current, a real (P,) block (iterative.py:481, 571, 591-673, 690, 771, 853-870, 910-915,
    # 1029); held in every loss evaluation when there is one
    update_bc = this_beam_current_step > 0
    if update_bc and beam_current is None:
        raise ValueError("pypty_amd.bfgs_update: refining the beam current needs a beam_current")
    if beam_current is not None and not isinstance(beam_current, torch.Tensor):
        beam_current = torch.as_tensor(np.asarray(beam_current, np.float32), device=obj.device)
    # the aberrations, a real (rows, coefficients) block (iterative.py:478, 570, 590-672, 689, 770,
    # 854-871, 904-909, 1028); in every loss evaluation when there is a marker, as objective.py:320 uses
    # them only then
    update_ab = this_step_aberrations_array > 0
    if update_ab and (aberration_marker is None or aberrations_array is None):
        raise ValueError("pypty_amd.bfgs_update: refining the aberrations needs aberrations_array and "
                         "aberration_marker")
    if aberrations_array is not None and not isinstance(aberrations_array, torch.Tensor):
        aberrations_array = torch.as_tensor(np.asarray(aberrations_array, np.float32), device=obj.device)
    for key in ("static_background_hist_s", "static_background_hist_y", "beam_current_hist_s",
                "beam_current_hist_y", "aberrations_hist_s", "aberrations_hist_y"):
        history_bfgs.setdefault(key, [0] * len(history_bfgs["rho_hist"]))
    history_bfgs.setdefault("gamma_static", 1)
    history_bfgs.setdefault("gamma_beam_current", 1)
    history_bfgs.setdefault("gamma_aberrations", 1)
    cut = damping_cutoff_multislice - update_extra_cut
    clean = not lazy_clean
    # an infinite probe_reg_weight band-limits the probe and its gradient to the aperture instead of
    # penalising what lies outside it (iterative.py:490-496, 526-530)
    clean_grad_ap = probe_reg_weight is not None and np.isinf(float(probe_reg_weight))

    def _clean_aperture(x):
        if isinstance(aperture_mask, (np.ndarray, torch.Tensor)):
            return utils.fourier_clean(x, mask=aperture_mask)
        return utils.fourier_clean(x, cutoff=aperture_mask)

    if clean_grad_ap:
        probe_reg_weight = 0
        probe = _clean_aperture(probe)

    def evaluate(o, p, pc, b, c, ab, tc):
        out = objective.loss_and_direction(
            o, p, positions, pc, tilts, tc, this_slice_distances, edge_slice_thickness_multiplier,
            measured_array, algorithm_type, this_wavelength, update_probe, update_obj, update_pos, update_tilts, masks,
            pixel_size_x_A, pixel_size_y_A, recon_type, Cs, defocus_array, alpha_near_field, damping_cutoff_multislice,
            0, propmethod, this_chopped_sequence, load_one_by_one, data_multiplier, data_pad, phase_plate_in_h5,
            this_loss_weight, data_bin, data_shift_vector, upsample_pattern, b, update_bg, tilt_mode,
            aberration_marker, probe_marker, ab, compute_batch, phase_only_obj, c, update_bc, update_ab, default_float,
            default_complex, xp, is_first_epoch, scan_size, scan_step_A, fast_axis_reg_weight_positions,
            slow_axis_reg_weight_positions, slow_axis_reg_weight_tilts, current_deformation_reg_weight_positions,
            current_deformation_reg_weight_tilts, fast_axis_reg_weight_tilts, aperture_mask, probe_reg_weight,
            current_window_weight,
            current_window, phase_norm_weight, abs_norm_weight, atv_weight, atv_q, atv_p, mixed_variance_weight,
            mixed_variance_sigma, wedge_mu, beta_wedge, probe_mod_reg_weight, probe_mod_tol, smart_memory,
            print_flag, data_simulation_flag, data_simulation_exitwave, use_padded_intensity,
            checkpoint_slices=checkpoint_slices)
        loss, sse, og, pg, posg = out[0], out[1], out[2], out[3], out[4]
        if clean_grad_ap and update_probe:
            pg = _clean_aperture(pg)
        if update_probe:
            pg = _clean(pg, cut, smooth_rolloff)
        if clean and update_obj:
            og = _clean(og, cut, smooth_rolloff)
        if phase_only and update_obj:
            # a phase-only object is refined in its phase, a real block (iterative.py:549-550, 798-799)
            og = utils.complex_grad_to_phase_grad(og.contiguous(), o.contiguous())
        # an unrefined background has a zero gradient, as objective.py:376-377 returns it, so a history
        # entry stored in such an epoch serves an epoch that starts refining it without a reset
        bgg = out[6] if update_bg else (torch.zeros_like(b) if isinstance(b, torch.Tensor) else 0)
        bcg = out[8] if update_bc else (torch.zeros_like(c) if isinstance(c, torch.Tensor) else 0)
        abg = out[7] if update_ab else (torch.zeros_like(ab) if isinstance(ab, torch.Tensor) else 0)
        tg = out[5] if update_tilts else torch.zeros_like(tc)
        return loss, sse, og, pg, posg, out[9], bgg, bcg, abg, tg

    if update_probe:
        if keep_probe_states_orthogonal and probe.dim() == 3 and probe.shape[-1] > 1:
            probe = objective.make_states_orthogonal(probe)  # iterative.py:498-499
        probe = _clean(probe, cut, smooth_rolloff)
    if update_obj:
        if do_charge_flip:
            # iterative.py:502-505: flip, then start the history afresh
            obj = objective.charge_flip(obj, cf_delta_phase, cf_delta_abs, cf_beta_phase, cf_beta_abs, fancy_sigma)
            reset_bfgs_history()
            empty_hist = True
        if clean:
            obj = _clean(obj, cut, smooth_rolloff)

    # the blocks in the reference's order of its dot products (iterative.py:683-690): object, probe,
    # positions, tilts, then the background, beam current and aberrations
    on = (update_obj, update_probe, update_pos, update_tilts, update_bg, update_bc, update_ab)
    S_KEYS = ("obj_hist_s", "probe_hist_s", "positions_hist_s", "tilts_hist_s", "static_background_hist_s",
              "beam_current_hist_s", "aberrations_hist_s")
    Y_KEYS = ("obj_hist_y", "probe_hist_y", "positions_hist_y", "tilts_hist_y", "static_background_hist_y",
              "beam_current_hist_y", "aberrations_hist_y")
    G_KEYS = ("gamma_obj", "gamma_probe", "gamma_pos", "gamma_tilts", "gamma_static", "gamma_beam_current",
              "gamma_aberrations")
    NB = len(on)

    def direction_dot(u, g):
        # a history entry of 0 (the variable was not refined in that epoch) adds 0, as in the reference
        d = 0.0
        for a, b, f in zip(u, g, on):
            if f and isinstance(a, torch.Tensor) and isinstance(b, torch.Tensor):
                d += _dot2(a, b)
        return d

    if empty_hist:
        total_loss, this_sse, og, pg, posg, cc, bgg, bcg, abg, tg = evaluate(obj, probe, positions_correction,
                                                                             static_background, beam_current,
                                                                             aberrations_array, tilts_correction)
        g = [og, pg, posg, tg, bgg, bcg, abg]
        u_obj = _neg(og) if update_obj else None
        d_val_obj = _dot2(u_obj, u_obj) if update_obj else 0.0
        if d_val_obj == 0:
            d_val_obj = 1.0
        u = [u_obj] + [lbfgs.lincomb(g[j], -d_val_obj / _dot2(g[j], g[j])) if on[j] else None
                       for j in range(1, NB)]
        history_bfgs["empty_hist"] = False
    else:
        total_loss, this_sse = history_bfgs["prev_loss"], history_bfgs["prev_sse"]
        g = [history_bfgs["obj_grad"], history_bfgs["probe_grad"], history_bfgs["pos_grad"],
             history_bfgs.get("tilt_grad"), history_bfgs.get("static_background_grad"),
             history_bfgs.get("beam_current_grad"), history_bfgs.get("aberrations_grad")]
        if update_tilts and not isinstance(g[3], torch.Tensor):
            g[3] = torch.zeros_like(tilts_correction)  # tilts held fixed until now
        if update_bg and not isinstance(g[4], torch.Tensor):
            g[4] = torch.zeros_like(static_background)  # a background held fixed until now
        if update_bc and not isinstance(g[5], torch.Tensor):
            g[5] = torch.zeros_like(beam_current)  # a current held fixed until now
        if update_ab and not isinstance(g[6], torch.Tensor):
            g[6] = torch.zeros_like(aberrations_array)  # aberrations held fixed until now
        cc = history_bfgs["constraint_contributions"]
        if update_probe:
            g[1] = _clean(g[1], cut, smooth_rolloff)
        rhos = history_bfgs["rho_hist"]
        q = [g[j].clone() if on[j] else None for j in range(NB)]
        alphas = []
        for i in range(len(rhos) - 1, -1, -1):
            a = rhos[i] * direction_dot([history_bfgs[k][i] for k in S_KEYS], q)
            alphas.append(a)
            for j, key in enumerate(Y_KEYS):
                if on[j] and isinstance(history_bfgs[key][i], torch.Tensor):
                    lbfgs.axpby(q[j], history_bfgs[key][i], 1.0, -a)
        alphas = alphas[::-1]
        for j, key in enumerate(G_KEYS):
            gm = history_bfgs[key]
            if on[j] and gm != 0 and gm == gm:
                lbfgs.axpby(q[j], None, gm, 0.0)
        for i in range(len(rhos)):
            beta = alphas[i] - rhos[i] * direction_dot([history_bfgs[k][i] for k in Y_KEYS], q)
            for j, key in enumerate(S_KEYS):
                if on[j] and isinstance(history_bfgs[key][i], torch.Tensor):
                    lbfgs.axpby(q[j], history_bfgs[key][i], 1.0, beta)
        u = [lbfgs.axpby(q[j], None, -1.0, 0.0) if on[j] else None for j in range(NB)]
        if clean and update_obj and not phase_only:
            u[0] = _clean(u[0], cut, smooth_rolloff)
    d_value = direction_dot(u, g)
    if d_value >= 0:
        sys.stdout.write("\n\n\nWarning! Positive dir-derivative, using normal gradient-descent!\n\n\n\n\n\n")
        warnings += "\n\\Positive dir-derivative, using normal gradient-descent!"
        u = [_neg(g[j]) if on[j] else None for j in range(NB)]
        for j, key in enumerate(G_KEYS):
            gm = history_bfgs[key]
            if on[j] and gm > 0 and gm == gm:
                lbfgs.axpby(u[j], None, gm, 0.0)
        d_value = direction_dot(u, g)
    u_obj, u_probe, u_pos, u_t, u_bg, u_bc, u_ab = u
    og, pg, posg, tg, bgg, bcg, abg = g
    count = -1
    new_d_value = d_value
    this_wolfe_1 = False
    while True:
        count += 1
        if update_obj:
            if phase_only:
                new_obj = _phase_step(obj, u_obj, actual_step)  # iterative.py:737-738
            else:
                new_obj = lbfgs.axpby(obj.clone(), u_obj, 1.0, actual_step)
            if clean:
                new_obj = _clean(new_obj, cut, smooth_rolloff)
        else:
            new_obj = obj
        new_probe = lbfgs.axpby(probe.clone(), u_probe, 1.0, actual_step) if update_probe else probe
        new_pc = lbfgs.axpby(positions_correction.clone(), u_pos, 1.0, actual_step) if update_pos \
            else positions_correction
        new_bg = lbfgs.axpby(static_background.clone(), u_bg, 1.0, actual_step) if update_bg \
            else static_background
        new_bc = lbfgs.axpby(beam_current.clone(), u_bc, 1.0, actual_step) if update_bc else beam_current
        new_ab = lbfgs.axpby(aberrations_array.clone(), u_ab, 1.0, actual_step) if update_ab else aberrations_array
        new_tc = lbfgs.axpby(tilts_correction.clone(), u_t, 1.0, actual_step) if update_tilts else tilts_correction
        new_total_loss, new_sse, nog, npg, nposg, ncc, nbgg, nbcg, nabg, ntg = evaluate(
            new_obj, new_probe, new_pc, new_bg, new_bc, new_ab, new_tc)
        new_d_value = direction_dot(u, (nog, npg, nposg, ntg, nbgg, nbcg, nabg))
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
    s_t = lbfgs.lincomb(u_t, actual_step) if update_tilts else 0
    s_bg = lbfgs.lincomb(u_bg, actual_step) if update_bg else 0
    s_bc = lbfgs.lincomb(u_bc, actual_step) if update_bc else 0
    s_ab = lbfgs.lincomb(u_ab, actual_step) if update_ab else 0
    y_obj = lbfgs.axpby(nog.clone(), og, 1.0, -1.0) if update_obj else 0
    y_probe = lbfgs.axpby(npg.clone(), pg, 1.0, -1.0) if update_probe else 0
    y_pos = lbfgs.axpby(nposg.clone(), posg, 1.0, -1.0) if update_pos else 0
    y_t = lbfgs.axpby(ntg.clone(), tg, 1.0, -1.0) if update_tilts else 0
    y_bg = lbfgs.axpby(nbgg.clone(), bgg, 1.0, -1.0) if update_bg else 0
    y_bc = lbfgs.axpby(nbcg.clone(), bcg, 1.0, -1.0) if update_bc else 0
    y_ab = lbfgs.axpby(nabg.clone(), abg, 1.0, -1.0) if update_ab else 0
    for key, v in (("obj_hist_s", s_obj), ("probe_hist_s", s_probe), ("positions_hist_s", s_pos),
                   ("tilts_hist_s", s_t), ("static_background_hist_s", s_bg), ("beam_current_hist_s", s_bc),
                   ("aberrations_hist_s", s_ab), ("obj_hist_y", y_obj), ("probe_hist_y", y_probe),
                   ("positions_hist_y", y_pos), ("tilts_hist_y", y_t), ("static_background_hist_y", y_bg),
                   ("beam_current_hist_y", y_bc), ("aberrations_hist_y", y_ab)):
        h[key].append(v)
    h["prev_loss"], h["prev_sse"] = new_total_loss, new_sse
    h["obj_grad"], h["probe_grad"], h["pos_grad"], h["static_background_grad"] = nog, npg, nposg, nbgg
    h["tilt_grad"] = ntg
    h["beam_current_grad"], h["aberrations_grad"] = nbcg, nabg
    h["constraint_contributions"] = ncc
    yiTsi_tot = 0.0
    for y, s, key, flag in ((y_obj, s_obj, "gamma_obj", update_obj), (y_probe, s_probe, "gamma_probe", update_probe),
                            (y_pos, s_pos, "gamma_pos", update_pos), (y_t, s_t, "gamma_tilts", update_tilts),
                            (y_bg, s_bg, "gamma_static", update_bg),
                            (y_bc, s_bc, "gamma_beam_current", update_bc),
                            (y_ab, s_ab, "gamma_aberrations", update_ab)):
        if flag:
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
    if update_tilts:
        tilts_correction = new_tc
    if update_bg:
        static_background = new_bg
    if update_bc:
        beam_current = new_bc
    if update_ab:
        aberrations_array = new_ab
    return total_loss, this_sse, cc, actual_step, count, d_value, new_d_value, warnings, counter_failed_epochs


# ----------------------------------------------------------------- driver
def _per_epoch(v, epoch):
    """PyPty's "lambda-type" parameters (pypty/utils.py:2260-2285, 2366-2425): callables of the epoch,
    or lambda source text (as params.pkl stores them), which ``utils.string_to_lambda`` turns into a
    callable without eval; other values, text included, are used as they are."""
    v = utils.string_to_lambda(v)
    return v(epoch) if callable(v) else v


def _save_epoch_files(folder, epoch, o, p, pos, corr, save_obj, save_probe, save_pos, bg=None, bc=None, ab=None,
                      tilts_in_use=None):
    """checkpoint_{beam_current,obj,probe,positions,tilts,stback}_epoch_N.npy of the variables refined in
    epoch N, as pypty/utils.py:1942-1975 writes them when save_checkpoints_every_epoch selects N, and
    checkpoint_aberrations_epcoh_N.npy (the reference's spelling, which its restart reads) for
    aberrations being refined."""
    import os

    if not folder:
        return
    os.makedirs(folder, exist_ok=True)
    name = lambda what: os.path.join(folder, f"checkpoint_{what}_epoch_{epoch}.npy")  # noqa: E731
    if isinstance(bc, torch.Tensor):
        np.save(name("beam_current"), bc.cpu().numpy())
    if save_obj:
        np.save(name("obj"), o.cpu().numpy())
    if save_probe:
        np.save(name("probe"), p.cpu().numpy())
    if save_pos:
        np.save(name("positions"), (pos.double() + corr.double()).cpu().numpy())
    if isinstance(tilts_in_use, torch.Tensor):
        np.save(name("tilts"), tilts_in_use.cpu().numpy())
    if isinstance(bg, torch.Tensor):
        np.save(name("stback"), bg.cpu().numpy())
    if isinstance(ab, torch.Tensor):
        np.save(os.path.join(folder, f"checkpoint_aberrations_epcoh_{epoch}.npy"), ab.cpu().numpy())


def _save_checkpoint(folder, o, p, pos, corr, save_obj, save_probe, save_pos, bg=None, bc=None, ab=None,
                     tilts_in_use=None):
    """co.npy / cp.npy / cg.npy, cs.npy for a background, cb.npy for a beam current, ca.npy for
    aberrations and ct.npy for the tilts (array plus correction) being refined, as
    pypty/utils.py:2085-2118 writes them."""
    import os

    if not folder:
        return
    os.makedirs(folder, exist_ok=True)
    if save_obj:
        np.save(os.path.join(folder, "co.npy"), o.cpu().numpy())
    if save_probe:
        np.save(os.path.join(folder, "cp.npy"), p.cpu().numpy())
    if save_pos:
        np.save(os.path.join(folder, "cg.npy"), (pos.double() + corr.double()).cpu().numpy())
    if isinstance(tilts_in_use, torch.Tensor):
        np.save(os.path.join(folder, "ct.npy"), tilts_in_use.cpu().numpy())
    if isinstance(bg, torch.Tensor):
        np.save(os.path.join(folder, "cs.npy"), bg.cpu().numpy())
    if isinstance(bc, torch.Tensor):
        np.save(os.path.join(folder, "cb.npy"), bc.cpu().numpy())
    if isinstance(ab, torch.Tensor):
        np.save(os.path.join(folder, "ca.npy"), ab.cpu().numpy())


def lambda_source(f):
    """Source text of a lambda (``lambda e: ...``), the form params.pkl stores for callables
    (pypty/utils.py:2315-2338); other values are returned unchanged."""
    import ast
    import inspect
    import types

    if not isinstance(f, types.LambdaType) or f.__name__ != "<lambda>":
        return f
    try:
        src = inspect.getsource(f)
    except (OSError, TypeError):
        return repr(f)
    start = src.find("lambda")
    # the shortest prefix of the text from "lambda" on that parses as one expression
    for end in range(len(src), start, -1):
        piece = src[start:end].strip()
        try:
            node = ast.parse(piece, mode="eval")
        except SyntaxError:
            continue
        if isinstance(node.body, ast.Lambda):
            return piece
    return src[start:].strip()


def save_params(path, params, strip_dataset=True):
    """params.pkl as pypty/utils.py:2432-2457 writes it: the parameter dict with callables as their
    lambda source and, when strip_dataset, without the 'dataset' entry."""
    import os
    import pickle

    def conv(v):  # lambdas inside [value, lambda] step lists too (the reference converts top-level ones)
        return type(v)(conv(x) for x in v) if isinstance(v, (list, tuple)) else lambda_source(v)

    # epoch_callback is this driver's own hook, not a reconstruction parameter
    out = {k: conv(v) for k, v in params.items()
           if not (strip_dataset and k == "dataset") and k != "epoch_callback"}
    if os.path.exists(path):
        os.remove(path)
    with open(path, "wb") as fh:
        pickle.dump(out, fh)


LOSS_CSV_FIELDS = ["epoch", "time / s", "loss", "sse", "initial step", "matching step", "N linesearch iterations",
                   "dir. derivative", "new dir. derivative", "Constraints contribution", "Free GiB", "Total GiB",
                   "Warnings"]


def _start_loss_csv(folder):
    """A new loss.csv holding the header row (pypty/utils.py:2004-2013)."""
    import csv
    import os

    with open(os.path.join(folder, "loss.csv"), "w+", newline="") as fh:
        csv.DictWriter(fh, fieldnames=LOSS_CSV_FIELDS).writeheader()


def _log_loss(folder, row):
    """Append one row of loss.csv in the reference's 13-column layout (pypty/utils.py:1877-1940)."""
    import csv
    import os

    if not folder:
        return
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "loss.csv"), mode="a", newline="") as fh:
        csv.DictWriter(fh, fieldnames=LOSS_CSV_FIELDS).writerow(row)


def _static_background_from_nothing(value, n, cutoff):
    """pypty/utils.py:694-731 without padding or upsampling: None stays None, an array is used as
    given, a scalar s (at least 1e-7) becomes 0.1 + 0.9 x uniform noise on (n, n) scaled to total s,
    zero where the DC-centred normalised frequency radius exceeds cutoff / 2."""
    if value is None or isinstance(value, (np.ndarray, torch.Tensor)):
        return value
    total = max(float(value), 1e-7)
    bg = 0.1 + 0.9 * np.random.rand(n, n)
    bg *= total / np.sum(bg)
    f = np.fft.fftshift(np.fft.fftfreq(n))
    r = np.sum(np.array(np.meshgrid(f, f, indexing="xy")) ** 2, axis=0) ** 0.5
    bg[r > cutoff * 0.5] = 0
    return bg.astype(np.float32)


def run(pypty_params):
    """Iterative reconstruction (mirror of pypty.iterative.run, pypty/iterative.py:27-436).

    Reads the keys of the far-field multislice path with the reference's
    defaults: dataset (array, a torch tensor used in place on the device, or
    data_path to a .npy), probe (N, N, M),
    positions (P, 2) in pixels, obj, acc_voltage, pixel_size_x_A/y_A,
    slice_distances, num_slices, algorithm, epoch_max, epoch_prev, update_obj,
    update_probe, update_probe_pos (values or callables of the epoch),
    compute_batch, damping_cutoff_multislice, update_extra_cut, smooth_rolloff,
    wolfe_c1_constant, wolfe_c2_constant, max_count, reduce_factor, optimism,
    min_step, hist_length, update_step_bfgs, loss_weight, allow_subPixel_shift,
    extra_space_on_side_px, sequence, output_folder, save_inter_checkpoints,
    save_loss_log, print_flag, static_background and update_static_background,
    beam_current and update_beam_current, aberrations_array, aberration_marker and
    update_aberrations_array, tilts, tilt_mode and update_tilts with the tilts'
    fast_axis_reg_weight_tilts, deformation_reg_weight_tilts (and, as there, the slow-axis
    weight read from slow_axis_reg_weight_positions),
    phase_norm_weight and abs_norm_weight (values or callables of the epoch), and two keys
    not in the reference:
    checkpoint_slices (keep every k-th slice's wave and recompute the rest in
    the adjoint) and epoch_callback (called as f(epoch, obj, probe) with the
    device arrays after every epoch, e.g. to track an error metric).
    Keys selecting features outside the path raise
    NotImplementedError through bfgs_update / loss_and_direction.  Writes
    co.npy, cp.npy, cg.npy and loss.csv like the reference; returns the final
    object, probe, positions and per-epoch losses (the reference returns None).
    """
    import os
    import time

    global obj, probe, positions, positions_correction, tilts, tilts_correction, static_background, beam_current
    global aberrations_array, history_bfgs
    get = pypty_params.get
    dev = torch.device("cuda", torch.cuda.current_device())
    dataset = get("dataset", None)
    if dataset is None:
        path = get("data_path", "")
        if not path.endswith(".npy"):
            raise NotImplementedError("pypty_amd.run: data_path must be a .npy file (HDF5 is not supported)")
        dataset = np.load(path)
    if isinstance(dataset, torch.Tensor):
        # already resident (for example uploaded once for a warm start): used in place on the device
        dataset = dataset.to(dev, torch.float32).contiguous()
    else:
        dataset = np.asarray(dataset, dtype=np.float32)
    if dataset.ndim == 4:
        dataset = dataset.reshape(dataset.shape[0] * dataset.shape[1], dataset.shape[2], dataset.shape[3])
    if get("probe", None) is None:
        raise NotImplementedError("pypty_amd.run: an initial probe is required (probe synthesis is setup code)")
    pr = np.asarray(get("probe"), dtype=np.complex64)
    if pr.ndim == 2:
        pr = pr[:, :, None]
    n = pr.shape[0]
    num_slices = get("num_slices", 1)
    o = np.asarray(get("obj", np.ones((1, 1, num_slices, 1))), dtype=np.complex64)
    acc = np.float32(get("acc_voltage", 60))
    wavelength = np.float32(12.398 / np.sqrt((2 * 511.0 + acc) * acc))  # pypty/utils.py:1721
    px, py = np.float32(get("pixel_size_x_A", 1)), np.float32(get("pixel_size_y_A", 1))
    slice_distances = np.asarray(get("slice_distances", np.array([10])), dtype=np.float32)
    pos = np.asarray(get("positions", np.array([[0.0, 0.0]])), dtype=np.float32).astype(np.float64)
    seq_param = get("sequence", None)
    seq0 = np.arange(dataset.shape[0]) if seq_param is None else np.asarray(_per_epoch(seq_param, 0))
    # positions: shift to the origin, integer corner + sub-pixel correction, pad the object
    # (pypty/utils.py:1722-1764)
    # skip_preprocessing (set by initialize.prepare_restart) keeps the positions as given and adds no
    # margin: the object of a restart carries its margin already (utils.py:1722-1728)
    extra = int(get("extra_space_on_side_px", 0))
    if get("skip_preprocessing", False):
        extra = 0
    else:
        pos[:, 0] -= pos[:, 0].min()
        pos[:, 1] -= pos[:, 1].min()
    pos += extra
    if extra > 0:
        o = np.pad(o, [[extra, 0], [extra, 0], [0, 0], [0, 0]], constant_values=np.mean(o))
    corr = (pos - np.round(pos)) if get("allow_subPixel_shift", True) else np.zeros_like(pos)
    ipos = np.round(pos).astype(np.int32)
    diff_x = int(ipos[:, 1].max() + extra + n - o.shape[1])
    if diff_x > 0:
        o = np.concatenate([o, np.max(np.abs(o)) * np.ones((o.shape[0], diff_x) + o.shape[2:], np.complex64)], 1)
    diff_y = int(ipos[:, 0].max() + extra + n - o.shape[0])
    if diff_y > 0:
        o = np.concatenate([o, np.max(np.abs(o)) * np.ones((diff_y,) + o.shape[1:], np.complex64)], 0)
    cutoff = np.float32(get("damping_cutoff_multislice", 2 / 3))
    rolloff = np.float32(get("smooth_rolloff", 0))
    obj = utils.fourier_clean(torch.from_numpy(np.ascontiguousarray(o)).to(dev), cutoff=cutoff, rolloff=rolloff)
    probe = utils.fourier_clean(torch.from_numpy(pr).to(dev), cutoff=cutoff, rolloff=rolloff)
    positions = torch.from_numpy(ipos).to(dev)
    positions_correction = torch.from_numpy(corr.astype(np.float32)).to(dev)
    # tilts (y, x) before, inside and after the specimen, one row or one per pattern
    # (pypty/iterative.py:134-135), refined through their correction (utils.py:1741)
    tilts = _arrays.to_torch(get("tilts", np.zeros((1, 6), np.float32)), dev, torch.float32).reshape(-1, 6)
    tilts = tilts.contiguous()
    tilts_correction = torch.zeros_like(tilts)
    tilt_mode = int(get("tilt_mode", 0))
    static_background = _static_background_from_nothing(get("static_background", None), n, float(cutoff))
    beam_current = get("beam_current", None)
    if beam_current is not None:
        beam_current = _arrays.to_torch(beam_current, dev, torch.float32).reshape(-1).contiguous()
    # aberration rows and the row of each position (pypty/iterative.py:176-183); without a marker they
    # do not enter the model (objective.py:320)
    aberrations_array = _arrays.to_torch(get("aberrations_array", np.array([[0.0]])), dev, torch.float32)
    aberrations_array = aberrations_array.reshape(aberrations_array.shape[0], -1).contiguous()
    aberration_marker = get("aberration_marker", None)
    measured = dataset if isinstance(dataset, torch.Tensor) else torch.from_numpy(dataset).to(dev)
    folder = get("output_folder", "")
    print_flag = get("print_flag", 3)
    save_inter = get("save_inter_checkpoints", True)
    save_every = get("save_checkpoints_every_epoch", False)
    save_log = get("save_loss_log", 1)
    epoch_prev = int(get("epoch_prev", 0))
    if save_log == 2:
        raise NotImplementedError("pypty_amd.run: save_loss_log=2 (regulariser columns) is not supported")
    if folder:
        os.makedirs(folder, exist_ok=True)
        save_params(os.path.join(folder, "params.pkl"), pypty_params, get("strip_dataset_from_params", True))
        if save_log and epoch_prev == 0:
            _start_loss_csv(folder)  # a restart appends to the log it continues (utils.py:2004-2013)
    _save_checkpoint(folder, obj, probe, positions, positions_correction, True, True, True)
    reset_bfgs_history()
    t0 = time.time()
    counter_failed = 0
    fail_pos, fail_tilts, fail_bg, fail_ab, fail_probe = np.inf, np.inf, np.inf, np.inf, np.inf
    prev_steps = prev_constraints = 0
    losses = []
    is_first = True
    step_param = get("update_step_bfgs", 1)  # first line-search step, a value or a callable of the epoch
    epoch_callback = get("epoch_callback", None)
    for epoch in range(epoch_prev, int(get("epoch_max", 200))):
        warnings = ""
        s_obj = _per_epoch(get("update_obj", 1), epoch)
        s_probe = _per_epoch(get("update_probe", 1), epoch)
        s_pos = _per_epoch(get("update_probe_pos", 0), epoch)
        s_bg = _per_epoch(get("update_static_background", 0), epoch) if static_background is not None else 0
        s_bc = _per_epoch(get("update_beam_current", 0), epoch)
        s_ab = _per_epoch(get("update_aberrations_array", 0), epoch)
        s_t = _per_epoch(get("update_tilts", 0), epoch)
        first_step = _per_epoch(step_param, epoch)  # per epoch, as pypty/iterative.py:325-336 reads it
        # auto-fixer (pypty/iterative.py:338-379): positions after 2 failed epochs, the tilts after 3,
        # the static background after 4, the aberrations after 5, the probe after 6
        if counter_failed >= 2:
            fail_pos = epoch
            warnings += "\nWARNING: Scan grid refinement was manually disabled for the next 10 epochs!"
            if counter_failed >= 3:
                fail_tilts = epoch
                warnings += "\nWARNING: Tilt grid refinement was manually disabled for the next 10 epochs!"
            if counter_failed >= 4:
                fail_bg = epoch
                warnings += "\nWARNING: Static background refinement was manually disabled for the next 10 epochs!"
            if counter_failed >= 5:
                fail_ab = epoch
                warnings += "\nWARNING: Aberrations refinement was manually disabled for the next 10 epochs!"
            if counter_failed >= 6:
                fail_probe = epoch
                warnings += "\nWARNING: Probe refinement was manually disabled for the next 10 epochs!"
        if abs(fail_pos - epoch) < 10:
            s_pos = 0
        if abs(fail_tilts - epoch) < 10:
            s_t = 0
        if abs(fail_bg - epoch) < 10:
            s_bg = 0
        if abs(fail_ab - epoch) < 10:
            s_ab = 0
        if abs(fail_probe - epoch) < 10:
            s_probe = 0
        lw = _per_epoch(get("loss_weight", 1), epoch)
        if lw == "mean":
            lw = 1 / dataset.shape[0]
        w_phase = _per_epoch(get("phase_norm_weight", 0), epoch)
        w_abs = _per_epoch(get("abs_norm_weight", 0), epoch)
        # the constraint weights (pypty/iterative.py:112, 124-126, 188-210), values or callables of the epoch
        cw = {k: _per_epoch(get(k, d), epoch) for k, d in (
            ("fast_axis_reg_weight_positions", 0), ("slow_axis_reg_weight_positions", 0),
            ("deformation_reg_weight_positions", 0), ("probe_reg_constraint_weight", 0), ("window_weight", 0),
            ("atv_weight", 0), ("mixed_variance_weight", 0), ("mixed_variance_sigma", 0.5), ("wedge_mu", 0),
            ("beta_wedge", 0), ("probe_mod_reg_weight", 0), ("probe_mod_tol", 0.05),
            ("fast_axis_reg_weight_tilts", 0), ("deformation_reg_weight_tilts", 0), ("phase_only_obj", False),
            ("keep_probe_states_orthogonal", False), ("do_charge_flip", False))}
        # the reference reads the tilts' slow-axis weight from the positions' key (iterative.py:208)
        cw["slow_axis_reg_weight_tilts"] = _per_epoch(get("slow_axis_reg_weight_positions", 0), epoch)
        window = get("window", None)
        this_window = None
        if cw["window_weight"] > 0 and window is not None:
            # a pair (r0, r_max) is a tapered window, else the window itself (iterative.py:410-416)
            this_window = (utils.get_window(n, window[0], window[1]) if len(window) == 2
                           else np.asarray(window, np.float32))
        # the history is reset when the objective changes: a different set of refined variables or of
        # non-zero constraint weights (pypty/iterative.py:383-389; of its weights run() takes these)
        steps = sum(int(x != 0) for x in (s_bc, s_obj, s_probe, s_pos, s_t, s_bg, s_ab))
        constraints = sum(int(x != 0) for x in (lw, w_phase, w_abs, *[cw[k] for k in cw if k not in (
            "mixed_variance_sigma", "beta_wedge", "probe_mod_tol")], cw["mixed_variance_sigma"]))
        if epoch != epoch_prev and (steps != prev_steps or constraints != prev_constraints):
            reset_bfgs_history()
            warnings += "\nResetting the history!"
        prev_steps, prev_constraints = steps, constraints
        if s_bc:
            # ones for a missing current, a short one padded (pypty/iterative.py:417)
            beam_current = utils.try_to_initialize_beam_current(beam_current, dataset.shape)
        seq = np.sort(np.asarray(_per_epoch(seq_param, epoch))) if seq_param is not None else seq0
        out = bfgs_update(
            algorithm_type=get("algorithm", "lsq_sqrt"), this_slice_distances=slice_distances,
            edge_slice_thickness_multiplier=1.0, this_step_probe=s_probe, this_step_obj=s_obj,
            this_step_pos_correction=s_pos, this_step_tilts=s_t, measured_array=measured, this_wavelength=wavelength,
            masks=None, pixel_size_x_A=px, pixel_size_y_A=py,
            phase_norm_weight=w_phase, abs_norm_weight=w_abs,
            stepsize_threshold_low=get("min_step", 1e-20), probe_reg_weight=cw["probe_reg_constraint_weight"], aperture_mask=get("aperture_mask", None),
            recon_type="far_field", defocus_array=np.array([0.0]), Cs=0.0, alpha_near_field=0.0,
            damping_cutoff_multislice=cutoff, smooth_rolloff=rolloff,
            update_extra_cut=np.float32(get("update_extra_cut", 0.005)),
            keep_probe_states_orthogonal=cw["keep_probe_states_orthogonal"], do_charge_flip=cw["do_charge_flip"],
            cf_delta_phase=get("cf_delta_phase", 0.1), cf_delta_abs=get("cf_delta_abs", 0.01),
            cf_beta_phase=get("cf_beta_phase", -0.95), cf_beta_abs=get("cf_beta_abs", -0.95),
            phase_only_obj=cw["phase_only_obj"], wedge_mu=cw["wedge_mu"], beta_wedge=cw["beta_wedge"],
            wolfe_c1_constant=_per_epoch(get("wolfe_c1_constant", 0.1), epoch),
            wolfe_c2_constant=_per_epoch(get("wolfe_c2_constant", 0.9), epoch), atv_weight=cw["atv_weight"], atv_q=get("atv_q", 1), atv_p=get("atv_p", 2),
            tune_only_probe_phase=False, propmethod="multislice", this_chopped_sequence=seq, load_one_by_one=False,
            data_multiplier=1, data_pad=0, phase_plate_in_h5=None, print_flag=print_flag, this_loss_weight=lw,
            max_count=get("max_count", 20), reduce_factor=np.float32(get("reduce_factor", 0.1)),
            optimism=get("optimism", 3), mixed_variance_weight=cw["mixed_variance_weight"],
            mixed_variance_sigma=cw["mixed_variance_sigma"], data_bin=1,
            data_shift_vector=[0, 0], smart_memory=False, default_float=np.float32, default_complex=np.complex64,
            default_int=np.int32, upsample_pattern=1,
            this_step_static_background=s_bg, tilt_mode=tilt_mode,
            fancy_sigma=get("fancy_sigma", None), tune_only_probe_abs=False, aberration_marker=aberration_marker,
            this_step_aberrations_array=s_ab,
            probe_marker=None, compute_batch=get("compute_batch", 256), current_window=this_window,
            current_window_weight=cw["window_weight"], dynamically_resize_yx_object=False, lazy_clean=False,
            current_gaussian_filter=False, current_apply_gaussian_filter_amplitude=False, this_beam_current_step=s_bc,
            xp=None, remove_fft_cache=False, is_first_epoch=is_first,
            hist_length=_per_epoch(get("hist_length", 10), epoch), actual_step=first_step,
            fast_axis_reg_weight_positions=cw["fast_axis_reg_weight_positions"],
            fast_axis_reg_weight_tilts=cw["fast_axis_reg_weight_tilts"],
            slow_axis_reg_weight_positions=cw["slow_axis_reg_weight_positions"],
            slow_axis_reg_weight_tilts=cw["slow_axis_reg_weight_tilts"], scan_size=get("scan_size", None),
            scan_step_A=get("scan_step_A", 1),
            current_deformation_reg_weight_positions=cw["deformation_reg_weight_positions"],
            current_deformation_reg_weight_tilts=cw["deformation_reg_weight_tilts"], warnings=warnings, data_simulation_flag=False,
            data_simulation_exitwave=False, pos_puzzling=False, probe_mod_reg_weight=cw["probe_mod_reg_weight"], probe_mod_tol=cw["probe_mod_tol"],
            use_padded_intensity=True, counter_failed_epochs=counter_failed,
            checkpoint_slices=int(get("checkpoint_slices", 0)))
        total_loss, sse, cc, actual_step, count, d_value, new_d_value, warnings, counter_failed = out
        is_first = False
        losses.append(total_loss)
        if print_flag:
            print(f"--> epoch {epoch}: loss {total_loss:.6e} sse {sse:.6e} step {actual_step:.3e} "
                  f"linesearch {count + 1} time {time.time() - t0:.2f} s", flush=True)
        if save_inter and epoch % int(save_inter) == 0:
            _save_checkpoint(folder, obj, probe, positions, positions_correction, s_obj, s_probe, s_pos,
                             static_background if s_bg else None, beam_current if s_bc else None,
                             aberrations_array if s_ab else None, (tilts + tilts_correction) if s_t else None)
        if save_every and epoch % int(save_every) == 0:
            _save_epoch_files(folder, epoch, obj, probe, positions, positions_correction, s_obj, s_probe, s_pos,
                              static_background if s_bg else None, beam_current if s_bc else None,
                              aberrations_array if s_ab else None, (tilts + tilts_correction) if s_t else None)
        if save_log:
            free, total = torch.cuda.mem_get_info(dev)
            _log_loss(folder, {"epoch": epoch, "time / s": time.time() - t0, "loss": total_loss, "sse": sse,
                               "initial step": first_step, "matching step": actual_step,
                               "N linesearch iterations": count + 1, "dir. derivative": d_value,
                               "new dir. derivative": new_d_value,
                               "Constraints contribution": float(np.sum(cc)),
                               "Free GiB": free / 1024 ** 3, "Total GiB": total / 1024 ** 3, "Warnings": warnings})
        if epoch_callback is not None:
            epoch_callback(epoch, obj, probe)
    _save_checkpoint(folder, obj, probe, positions, positions_correction, True, True, True)
    result = {"obj": obj, "probe": probe, "positions": positions.double() + positions_correction.double(),
              "beam_current": beam_current, "aberrations_array": aberrations_array,
              "tilts": tilts + tilts_correction, "losses": losses,
              "time_s": time.time() - t0}
    return result
