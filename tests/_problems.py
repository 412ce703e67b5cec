"""Seeded small ptychography problems shared by the CPU and GPU tests."""
from __future__ import annotations

import inspect

import numpy as np

from oracle import ms_oracle as orc


def virtual_detectors(n, J):
    """J virtual detectors in the DC-centred pattern layout: a full-field one and J-1 annular rings
    covering radii up to n/2."""
    y, x = np.mgrid[0:n, 0:n]
    r = np.hypot(y - n // 2, x - n // 2)
    edges = np.linspace(0.0, n / 2, J)
    masks = [np.ones((n, n))]
    for j in range(J - 1):
        masks.append(((r >= edges[j]) & (r < edges[j + 1])).astype(float))
    return np.asarray(masks, np.float32)


def make_problem(n=32, S=3, K=1, M=2, ny=80, nx=80, P=12, shift=True, seed=0, algorithm="lsq_sqrt",
                 dz=2.0, wavelength=0.0196868, px=0.2, step_pos=None, noise=0.05, n_masks=5):
    """Weakly scattering object, aperture-limited probe modes, random scan, measured data
    simulated by the FP64 oracle from a perturbed object plus multiplicative noise (virtual-detector
    coefficients for the compressed losses)."""
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 0.6, (ny, nx, S, K))
    amp = 1.0 - 0.05 * rng.uniform(size=(ny, nx, S, K))
    obj = (amp * np.exp(1j * phase)).astype(np.complex64)
    f = np.fft.fftfreq(n)
    fx, fy = np.meshgrid(f, f, indexing="xy")
    ap = (fx ** 2 + fy ** 2) <= 0.22 ** 2
    probe = np.zeros((n, n, M), np.complex64)
    for m in range(M):
        spec = ap * np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n))) * (1.0 / (1 + m))
        probe[:, :, m] = np.fft.ifft2(spec) * n
    pos = np.stack([rng.integers(0, ny - n + 1, P), rng.integers(0, nx - n + 1, P)], axis=1).astype(np.int32)
    corr = (rng.uniform(-0.5, 0.5, (P, 2)) if shift else np.zeros((P, 2))).astype(np.float32)
    seq = np.arange(P)
    true_obj = (obj * np.exp(1j * 0.1 * rng.standard_normal(obj.shape))).astype(np.complex128)
    _, _, _, _, _, I = orc.loss_and_direction(
        true_obj, probe.astype(np.complex128), pos, corr, np.zeros((P, n, n)), seq, dz=dz,
        wavelength=wavelength, pixel_size_x=px, pixel_size_y=px, algorithm="lsq_sqrt", step_obj=False,
        step_probe=False, step_pos=False, cdtype=np.complex128, return_patterns=True)
    masks = None
    if algorithm in orc.COMPRESSED:
        masks = virtual_detectors(n, n_masks)
        I = np.einsum("pyx,jyx->pj", I, masks.astype(np.float64))
    meas = np.maximum(I * (1 + noise * rng.standard_normal(I.shape)), 0).astype(np.float32)
    return dict(obj=obj, probe=probe, pos=pos, corr=corr, meas=meas, seq=seq, dz=dz, wavelength=wavelength, px=px,
                algorithm=algorithm, step_pos=bool(shift) if step_pos is None else step_pos, masks=masks)


def oracle_pass(pr, cdtype=np.complex64, step_obj=True, step_probe=True, compute_batch=5, loss_weight=1.0):
    return orc.loss_and_direction(
        pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"], dz=pr["dz"],
        wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"], algorithm=pr["algorithm"],
        step_obj=step_obj, step_probe=step_probe, step_pos=pr["step_pos"], cdtype=cdtype,
        compute_batch=compute_batch, loss_weight=loss_weight, masks=pr.get("masks"))


def reference_kwargs(**over):
    """Arguments of loss_and_direction with the reference's defaults for the inactive features."""
    kw = dict(
        this_tilt_array=np.zeros((1, 6), np.float32), this_tilts_correction=np.zeros((1, 6), np.float32),
        edge_slice_thickness_multiplier=1.0, masks=None, recon_type="far_field", Cs=0.0,
        defocus_array=np.array([0.0]), alpha_near_field=0.0, damping_cutoff_multislice=2 / 3, smooth_rolloff=0.0,
        propmethod="multislice", load_one_by_one=False, data_multiplier=1, data_pad=0, phase_plate_in_h5=None,
        this_loss_weight=1.0, data_bin=1, data_shift_vector=[0, 0], upsample_pattern=1, static_background=None,
        this_step_static_background=0, tilt_mode=0, aberration_marker=None, probe_marker=None,
        aberrations_array=np.zeros((1, 1)), compute_batch=5, phase_only_obj=False, beam_current=None,
        this_beam_current_step=0, this_step_aberrations_array=0, default_float=np.float32,
        default_complex=np.complex64, xp=np, is_first_epoch=True, scan_size=None, scan_step_A=1.0,
        fast_axis_reg_weight_positions=0, slow_axis_reg_weight_positions=0, slow_axis_reg_weight_tilts=0,
        current_deformation_reg_weight_positions=0, current_deformation_reg_weight_tilts=0,
        fast_axis_reg_weight_tilts=0, aperture_mask=None, probe_reg_weight=0, current_window_weight=0,
        current_window=None, phase_norm_weight=0, abs_norm_weight=0, atv_weight=0, atv_q=1, atv_p=2,
        mixed_variance_weight=0, mixed_variance_sigma=0.5, wedge_mu=0, beta_wedge=0, probe_mod_reg_weight=0,
        probe_mod_tol=0.05, smart_memory=False, print_flag=0, data_simulation_flag=False,
        data_simulation_exitwave=False, use_padded_intensity=True)
    kw.update(over)
    return kw


def call_loss_and_direction(fn, **kw):
    """Call a loss_and_direction with the reference's positional argument order; keyword-only
    extensions are passed by name when given."""
    params = inspect.signature(fn).parameters.values()
    pos = [kw[p.name] for p in params if p.kind == p.POSITIONAL_OR_KEYWORD]
    extra = {p.name: kw[p.name] for p in params if p.kind == p.KEYWORD_ONLY and p.name in kw}
    return fn(*pos, **extra)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def bfgs_kwargs(pr, **over):
    """Arguments of bfgs_update (reference order via call_loss_and_direction-style lookup)."""
    kw = dict(
        algorithm_type=pr["algorithm"], this_slice_distances=np.array([pr["dz"]], np.float32),
        edge_slice_thickness_multiplier=1.0, this_step_probe=1, this_step_obj=1,
        this_step_pos_correction=int(pr["step_pos"]), this_step_tilts=0, measured_array=pr["meas"],
        this_wavelength=pr["wavelength"], masks=None, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"],
        phase_norm_weight=0, abs_norm_weight=0, stepsize_threshold_low=1e-20, probe_reg_weight=0, aperture_mask=None,
        recon_type="far_field", defocus_array=np.array([0.0]), Cs=0.0, alpha_near_field=0.0,
        damping_cutoff_multislice=np.float32(2 / 3), smooth_rolloff=0.0, update_extra_cut=np.float32(0.005),
        keep_probe_states_orthogonal=False, do_charge_flip=False, cf_delta_phase=0.1, cf_delta_abs=0.01,
        cf_beta_phase=-0.95, cf_beta_abs=-0.95, phase_only_obj=False, wedge_mu=0, beta_wedge=0,
        wolfe_c1_constant=0.1, wolfe_c2_constant=0.9, atv_weight=0, atv_q=1, atv_p=2, tune_only_probe_phase=False,
        propmethod="multislice", this_chopped_sequence=pr["seq"], load_one_by_one=False, data_multiplier=1,
        data_pad=0, phase_plate_in_h5=None, print_flag=0, this_loss_weight=1.0, max_count=20, reduce_factor=0.1,
        optimism=3, mixed_variance_weight=0, mixed_variance_sigma=0.5, data_bin=1, data_shift_vector=[0, 0],
        smart_memory=False, default_float=np.float32, default_complex=np.complex64, default_int=np.int32,
        upsample_pattern=1, this_step_static_background=0, tilt_mode=0, fancy_sigma=None, tune_only_probe_abs=False,
        aberration_marker=None, this_step_aberrations_array=0, probe_marker=None, compute_batch=5,
        current_window=None, current_window_weight=0, dynamically_resize_yx_object=False, lazy_clean=False,
        current_gaussian_filter=False, current_apply_gaussian_filter_amplitude=False, this_beam_current_step=0,
        xp=np, remove_fft_cache=False, is_first_epoch=True, hist_length=10, actual_step=1.0,
        fast_axis_reg_weight_positions=0, fast_axis_reg_weight_tilts=0, slow_axis_reg_weight_positions=0,
        slow_axis_reg_weight_tilts=0, scan_size=None, scan_step_A=1.0, current_deformation_reg_weight_positions=0,
        current_deformation_reg_weight_tilts=0, warnings="", data_simulation_flag=False,
        data_simulation_exitwave=False, pos_puzzling=False, probe_mod_reg_weight=0, probe_mod_tol=0.05,
        use_padded_intensity=True, counter_failed_epochs=0)
    kw.update(over)
    return kw
