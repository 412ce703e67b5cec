"""Per-position aberrations (pypty/objective.py:320-326, 459-465, 673-681) on the GPU: the drop-in
loss_and_direction with aberrations_array rows and an aberration_marker per position gives the oracle's
loss, object, probe and position gradients and aberration gradient; without the sub-pixel shift (the
plates reach the modelled patterns) and with it (as in the reference, the shift starts from the spectrum
before the plate, so only the gradient sees the plates); one and several probe modes; the register
tile and both slab tiles; with the beam current; several batches over several lanes; patterns
streamed from host memory (accumulating calls); a negative row index (numpy's wrap, as cp.add.at);
and with the aberrations held (zeros of their shape returned)."""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = [
    # n, S, M, algorithm, shift, step, streamed, beam current
    (64, 2, 1, "lsq_sqrt", False, True, False, False),
    (64, 2, 2, "ml", False, True, False, True),
    (32, 2, 2, "lsq", True, True, False, False),
    (96, 2, 2, "lsq_sqrt", False, True, False, False),
    (128, 2, 3, "lsq_sqrt_2", False, True, False, True),
    (128, 2, 2, "lsq_sqrt", True, True, False, True),
    (192, 2, 2, "lsq_sqrt", False, True, False, False),
    (256, 1, 1, "lsq_sqrt", False, True, False, False),
    (64, 2, 2, "lsq_sqrt", False, False, False, False),
    (64, 2, 2, "ml", False, True, True, False),
]

COEF_SCALE = np.array([20.0, 10.0, 10.0, 300.0, 300.0, 300.0, 300.0])


@pytest.mark.parametrize("n,S,M,algorithm,shift,step,streamed,with_bc", CASES)
def test_aberration_pass_matches_oracle(n, S, M, algorithm, shift, step, streamed, with_bc):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print(f"[aberrations] n={n} S={S} M={M} {algorithm} shift={shift} step={step} streamed={streamed} "
          f"bc={with_bc}", flush=True)
    P, R = 11, 3
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=shift, seed=701 + n + M,
                      algorithm=algorithm)
    rng = np.random.default_rng(9)
    ab = (rng.uniform(-1, 1, (R, 7)) * COEF_SCALE).astype(np.float32)
    marker = rng.integers(0, R, P)
    marker[4] = -1  # the last row, as numpy and cp.add.at index
    bc = rng.uniform(0.7, 1.3, P).astype(np.float32) if with_bc else None
    args = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                algorithm=algorithm, step_pos=shift, compute_batch=4, loss_weight=0.7, aberrations=ab,
                aberration_marker=marker, step_aberrations=step, beam_current=bc, step_beam_current=with_bc)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 cdtype=np.complex64, **args)
    ref64 = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                   cdtype=np.complex128, **args)
    k_ab = len(ref) - 1
    checked = (2, 3, k_ab) + ((5,) if with_bc else ())
    assert max(rel(ref[k], ref64[k]) for k in checked) < TOL / 3  # well conditioned
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"] if streamed else torch.from_numpy(pr["meas"]).to(dev),
        load_one_by_one=streamed, algorithm_type=algorithm, this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=int(shift), this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], compute_batch=4,
        this_loss_weight=0.7, aberrations_array=ab, aberration_marker=marker,
        this_step_aberrations_array=int(step), beam_current=bc, this_beam_current_step=int(with_bc))
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    errs = dict(loss=abs(out[0] - ref[0]) / abs(ref[0]), obj=rel(out[2].cpu().numpy(), ref[2]),
                probe=rel(out[3].cpu().numpy(), ref[3]))
    if shift:
        errs["pos"] = rel(out[4].cpu().numpy(), ref[4])
    if with_bc:
        errs["beam current"] = rel(out[8].cpu().numpy(), ref[5])
    got = out[7].cpu().numpy()
    assert got.shape == ab.shape
    if step:
        errs["aberrations"] = rel(got, ref[k_ab])
        assert float(np.abs(ref[k_ab]).max()) > 0
    else:
        assert not got.any()  # zeros of the aberrations' shape, as the reference returns
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (k, v)


def test_plates_reach_the_patterns_only_without_the_shift():
    """The modelled patterns: without the shift the aberrated probe's, equal to the oracle's; with it,
    the same as without aberrations (objective.py:471-474 restarts from the spectrum before the
    plate)."""
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print("[aberrations] modelled patterns with and without the shift", flush=True)
    n, P = 64, 6
    for shift in (False, True):
        pr = make_problem(n=n, S=2, K=1, M=1, ny=n + 32, nx=n + 32, P=P, shift=shift, seed=777)
        ab = np.array([[40.0, 10.0, -5.0]], np.float32)
        marker = np.zeros(P, np.int64)
        dev = torch.device("cuda", 0)
        cfg = objective.PassConfig(n=n, n_slices=2, n_obj_modes=1, n_probe_modes=1, obj_ny=n + 32, obj_nx=n + 32,
                                   n_positions=P, compute_batch=4, step_obj=False, step_probe=False,
                                   step_pos=False, shift_active=shift, slice_distance=pr["dz"],
                                   wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                                   planar=False, n_aberrations=3, n_markers=1)
        eng = objective.PtychoPass(cfg, dev)
        sims = []
        for with_ab in (False, True):
            if not with_ab:
                eng0 = objective.PtychoPass(objective.PassConfig(**{**cfg.__dict__, "n_aberrations": 0,
                                                                    "n_markers": 0}), dev)
            sim = torch.zeros((P, n, n), dtype=torch.float32, device=dev)
            (eng if with_ab else eng0).run(
                torch.from_numpy(pr["obj"]).to(dev), torch.from_numpy(pr["probe"]).permute(2, 0, 1).contiguous().to(dev),
                torch.from_numpy(pr["pos"]).to(dev), torch.from_numpy(pr["corr"]).to(dev),
                torch.from_numpy(pr["meas"]).to(dev), torch.arange(P, dtype=torch.int32, device=dev),
                sim_patterns=sim, aberrations=torch.from_numpy(ab).to(dev) if with_ab else None,
                aberration_marker=torch.from_numpy(marker.astype(np.int32)).to(dev) if with_ab else None)
            sims.append(sim.cpu().numpy())
        ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                     dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"],
                                     pixel_size_y=pr["px"], step_obj=False, step_probe=False, step_pos=False,
                                     cdtype=np.complex128, return_patterns=True, aberrations=ab,
                                     aberration_marker=marker)[5]
        e = rel(sims[1], ref)
        d = rel(sims[1], sims[0])
        print(f"  shift={shift}: aberrated patterns vs FP64 oracle {e:.2e}, vs unaberrated {d:.2e}", flush=True)
        assert e < 1e-4
        if shift:
            assert d < 1e-6
        else:
            assert d > 1e-2
