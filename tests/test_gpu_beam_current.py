"""Beam current (pypty/objective.py:446-450, 688-698) on the GPU: the drop-in loss_and_direction with a
per-position current gives the oracle's loss, object, probe and position gradients and current
gradient, with and without the sub-pixel shift, one and several probe modes, the register tile and the
slab tile, several batches over several lanes, patterns streamed from host memory (accumulating calls),
and with the current held (zeros of its shape returned)."""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = [
    # n, S, M, algorithm, shift, step, streamed
    (64, 2, 1, "lsq_sqrt", True, True, False),
    (64, 3, 2, "ml", True, True, False),
    (32, 2, 2, "lsq", False, True, False),
    (96, 2, 2, "lsq_sqrt", False, True, False),
    (128, 2, 3, "lsq_sqrt_2", True, True, False),
    (192, 2, 2, "lsq_sqrt", True, True, False),
    (192, 2, 2, "ml", False, True, False),
    (64, 2, 2, "lsq_sqrt", True, False, False),
    (64, 2, 2, "ml", True, True, True),
]


@pytest.mark.parametrize("n,S,M,algorithm,shift,step,streamed", CASES)
def test_beam_current_pass_matches_oracle(n, S, M, algorithm, shift, step, streamed):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print(f"[beam current] n={n} S={S} M={M} {algorithm} shift={shift} step={step} streamed={streamed}", flush=True)
    P = 11
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=shift, seed=501 + n + M,
                      algorithm=algorithm)
    rng = np.random.default_rng(6)
    bc = rng.uniform(0.7, 1.3, P).astype(np.float32)
    bc[3] = -bc[3]  # the probe takes |bc|; the reference's factor sign(|bc|) keeps its gradient's sign
    args = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                algorithm=algorithm, step_pos=shift, compute_batch=4, loss_weight=0.7, beam_current=bc,
                step_beam_current=step)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 cdtype=np.complex64, **args)
    ref64 = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                   cdtype=np.complex128, **args)
    assert max(rel(ref[k], ref64[k]) for k in (2, 3, 5)) < TOL / 3  # well conditioned
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"] if streamed else torch.from_numpy(pr["meas"]).to(dev),
        load_one_by_one=streamed, algorithm_type=algorithm, this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=int(shift), this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], compute_batch=4,
        this_loss_weight=0.7, beam_current=bc, this_beam_current_step=int(step))
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    errs = dict(loss=abs(out[0] - ref[0]) / abs(ref[0]), obj=rel(out[2].cpu().numpy(), ref[2]),
                probe=rel(out[3].cpu().numpy(), ref[3]))
    if shift:
        errs["pos"] = rel(out[4].cpu().numpy(), ref[4])
    got_bc = out[8].cpu().numpy()
    assert got_bc.shape == (P,)
    if step:
        errs["beam current"] = rel(got_bc, ref[5])
        assert float(np.abs(ref[5]).max()) > 0
    else:
        assert not got_bc.any()  # zeros of the current's shape, as the reference returns
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (k, v)


def test_beam_current_partial_sequence_leaves_other_rows_zero():
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print("[beam current] a sequence over part of the scan", flush=True)
    pr = make_problem(n=64, S=2, K=1, M=2, ny=104, nx=104, P=10, shift=True, seed=541)
    seq = np.array([7, 2, 9, 4, 0], np.int64)
    bc = np.random.default_rng(8).uniform(0.8, 1.2, 10).astype(np.float32)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], seq, dz=pr["dz"],
                                 wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                                 step_pos=True, compute_batch=2, cdtype=np.complex64, beam_current=bc,
                                 step_beam_current=True)
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=torch.from_numpy(pr["meas"]).to(dev), algorithm_type="lsq_sqrt",
        this_wavelength=pr["wavelength"], this_step_probe=0, this_step_obj=0, this_step_pos_correction=1,
        this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=seq,
        compute_batch=2, beam_current=torch.from_numpy(bc).to(dev), this_beam_current_step=1)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    got = out[8].cpu().numpy()
    assert np.all(got[[1, 3, 5, 6, 8]] == 0)
    e = rel(got, ref[5])
    print(f"  beam current gradient vs complex64 oracle {e:.2e}", flush=True)
    assert e < TOL
