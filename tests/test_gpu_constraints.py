"""Constraint terms on the device (pypty_amd.constraints) against the oracle's restatements, alone and
inside loss_and_direction with every weight on (objective.py:765-933): terms and gradients to 1e-5
relative of the FP64 oracle, the whole pass to 1e-4 of the complex64 oracle."""
import numpy as np
import pytest

from oracle import constraints_oracle as co
from oracle import ms_oracle as orc
from tests._problems import call_loss_and_direction, make_problem, oracle_pass, reference_kwargs, rel

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _cmp(label, got, ref, tol=1e-5):
    tg, gg = got
    tr, gr = ref
    et = abs(tg - tr) / max(abs(tr), 1e-300)
    eg = rel(gg.cpu().numpy(), gr)
    print(f"  {label}: term {et:.2e} gradient {eg:.2e}", flush=True)
    assert et < tol and eg < tol, (label, et, eg)


def test_each_constraint_matches_the_oracle():
    from pypty_amd import constraints as C

    print("[constraints] each term against the FP64 oracle", flush=True)
    rng = np.random.default_rng(11)
    pos = (rng.standard_normal((8 * 9, 2)) * 3 + 20).astype(np.float64)
    for name, fn, ofn in (("fast axis", C.compute_fast_axis_constraint_on_grid, co.fast_axis_constraint_on_grid),
                          ("slow axis", C.compute_slow_axis_constraint_on_grid, co.slow_axis_constraint_on_grid),
                          ("deformation", C.compute_deformation_constraint_on_grid,
                           co.deformation_constraint_on_grid)):
        _cmp(name, fn(_t(pos), (8, 9), 0.3), ofn(pos, (8, 9), 0.3), 1e-9)
    P = (rng.standard_normal((96, 96, 2)) + 1j * rng.standard_normal((96, 96, 2))).astype(np.complex64)
    P64 = P.astype(np.complex128)
    win = rng.uniform(0, 1, (96, 96))
    _cmp("window", C.compute_window_constraint(_t(P), win, 0.4), co.window_constraint(P64, win, 0.4))
    _cmp("probe aperture ratio", C.compute_probe_constraint(_t(P), 0.6, 0.2), co.probe_constraint(P64, 0.6, 0.2))
    ap = rng.uniform(size=(96, 96)) > 0.3
    _cmp("probe aperture mask", C.compute_probe_constraint(_t(P), ap, 0.2), co.probe_constraint(P64, ap, 0.2))
    _cmp("probe modulation", C.probe_modulation_constraint(_t(P), 1.2, 0.2, 0.2, 0.3, 0.05, 2 / 3),
         co.probe_modulation_constraint(P64, 1.2, 0.2, 0.2, 0.3, 0.05, 2 / 3))
    O = (np.exp(1j * rng.uniform(-1, 1, (120, 100, 4, 2))) * rng.uniform(0.8, 1.2, (120, 100, 4, 2))).astype(
        np.complex64)
    O64 = O.astype(np.complex128)
    for q, p in ((1, 2), (1, 1), (0.8, 1.5)):
        _cmp(f"atv q={q} p={p}", C.compute_atv_constraint(_t(O), 0.3, q, p), co.atv_constraint(O64, 0.3, q, p))
    _cmp("missing wedge", C.compute_missing_wedge_constraint(_t(O), 0.2, 0.25, 1.5, 2.0, 0.3),
         co.missing_wedge_constraint(O64, 0.2, 0.25, 1.5, 2.0, 0.3))
    _cmp("mixed variance", C.compute_mixed_object_variance_constraint(_t(O), 0.7, 0.3),
         co.mixed_object_variance_constraint(O64, 0.7, 0.3))


def test_loss_and_direction_with_every_constraint():
    import torch

    from pypty_amd import objective

    print("[constraints] loss_and_direction with every constraint weight on", flush=True)
    pr = make_problem(n=64, S=3, K=2, M=2, ny=112, nx=112, P=12, shift=True, seed=71)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    # a probe whose spectrum is nowhere zero: the aperture and modulation terms take |F|, whose
    # gradient F / |F| is round-off where F vanishes (as in the reference)
    pr["probe"] = (pr["probe"] + 0.05 * np.abs(pr["probe"]).mean() * (rng.standard_normal(pr["probe"].shape) + 1j *
                   rng.standard_normal(pr["probe"].shape))).astype(np.complex64)
    window = rng.uniform(0, 1, (64, 64)).astype(np.float32)
    w = dict(fast_axis_reg_weight_positions=0.5, current_deformation_reg_weight_positions=0.3,
             slow_axis_reg_weight_positions=0.2, probe_reg_weight=50.0, aperture_mask=0.35,
             current_window_weight=0.05, current_window=window, phase_norm_weight=3.0, abs_norm_weight=4.0,
             atv_weight=2.0, atv_q=1, atv_p=2, mixed_variance_weight=30.0, mixed_variance_sigma=0.3,
             wedge_mu=0.5, beta_wedge=2.0, probe_mod_reg_weight=1e-4, probe_mod_tol=0.3, scan_size=(3, 4),
             scan_step_A=1.2)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=1, this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], **w)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    # the oracle: the pass, then the constraints in the reference's order
    loss, sse, og, pg, posg = oracle_pass(pr, np.complex64)
    obj, probe = pr["obj"], pr["probe"]
    cc = np.zeros(14)
    pos_abs = pr["pos"].astype(np.float64) + pr["corr"].astype(np.float64)
    for k, fn, wt in ((0, co.fast_axis_constraint_on_grid, 0.5), (1, co.deformation_constraint_on_grid, 0.3),
                      (11, co.slow_axis_constraint_on_grid, 0.2)):
        cc[k], g = fn(pos_abs, (3, 4), wt)
        posg = posg + g
    cc[4], g4 = orc.full_l1_constraint(obj, 0.0, 3.0)
    cc[5], g5 = orc.full_l1_constraint(obj, 4.0, 0.0)
    og = og + g4 + g5
    cc[6], gp6 = co.probe_constraint(probe, 0.35, 50.0)
    cc[7], gp7 = co.window_constraint(probe, window, 0.05)
    cc[8], g8 = co.atv_constraint(obj, 2.0, 1, 2)
    cc[9], g9 = co.mixed_object_variance_constraint(obj, 30.0, 0.3)
    cc[10], g10 = co.missing_wedge_constraint(obj, pr["px"], pr["px"], pr["dz"], 2.0, 0.5)
    cc[13], gp13 = co.probe_modulation_constraint(probe, 1.2, pr["px"], pr["px"], 0.3, 1e-4, 2 / 3)
    og = og + g8 + g9 + g10
    _, _, _, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], 64, 2 / 3)
    pg = pg + orc.ifft2(orc.fft2(gp6 + gp7 + gp13) * mask[:, :, None])
    loss = loss + cc.sum()
    got_cc = np.asarray(out[9])
    print("  contributions gpu   ", np.array2string(got_cc, precision=4), flush=True)
    print("  contributions oracle", np.array2string(cc, precision=4), flush=True)
    assert np.all((got_cc != 0) == (cc != 0))
    assert rel(got_cc, cc) < 1e-4
    errs = dict(loss=abs(out[0] - loss) / abs(loss), obj=rel(out[2].cpu().numpy(), og),
                probe=rel(out[3].cpu().numpy(), pg), pos=rel(out[4].cpu().numpy(), posg))
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


@pytest.mark.parametrize("shape", [(64, 64, 1, 1), (96, 80, 3, 2), (512, 512, 2, 1)])
def test_charge_flip_matches_oracle(shape):
    """objective.charge_flip (pty_charge_flip_c64) against oracle.constraints_oracle.charge_flip
    (objective.py:1014-1110): per plane thresholds, phases and absorptions flipped."""
    import torch

    from oracle import constraints_oracle as con
    from pypty_amd import objective

    print(f"[constraints] charge flip {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    a = (rng.uniform(0.6, 1.0, shape) * np.exp(1j * rng.uniform(-1.5, 2.5, shape))).astype(np.complex64)
    ref = con.charge_flip(a.astype(np.complex128), 0.3, 0.14, -0.95, -0.9)
    got = objective.charge_flip(torch.from_numpy(a).cuda(), 0.3, 0.14, -0.95, -0.9).cpu().numpy()
    e = rel(got, ref)
    print(f"  rel err vs FP64 oracle {e:.2e}", flush=True)
    assert e < 1e-5
    # numpy in, numpy out
    assert isinstance(objective.charge_flip(a[:8, :8].copy()), np.ndarray)


@pytest.mark.parametrize("n,M", [(64, 2), (128, 5), (256, 10)])
def test_orthogonal_probe_states_match_oracle(n, M):
    import torch

    from oracle import constraints_oracle as con
    from pypty_amd import objective

    print(f"[constraints] Gram-Schmidt of {M} modes at {n} px", flush=True)
    rng = np.random.default_rng(n + M)
    p = (rng.standard_normal((n, n, M)) + 1j * rng.standard_normal((n, n, M))).astype(np.complex64)
    p[:, :, 1] += 0.7 * p[:, :, 0]
    ref = con.make_states_orthogonal(p.astype(np.complex128))
    got = objective.make_states_orthogonal(torch.from_numpy(p).cuda()).cpu().numpy()
    e = rel(got, ref)
    print(f"  rel err vs FP64 oracle {e:.2e}", flush=True)
    assert e < 1e-5
