"""Beam tilts and per-slice distances on the GPU (pypty/objective.py:328-350, 428-499, 600-672;
pypty/multislice_core.py:69-77, 163-178): the drop-in loss_and_direction with a tilt array (one row or
one per position), tilt_mode 0 to 4, with and without tilt refinement, with a beam current, and with per-slice
distances, gives the oracle's loss, object, probe, position and tilt gradients; on the register tile
and both slab tiles; one slice (the exit wave band-limited before the after tilt) and several; an edge
multiplier with per-window tilts (each window raises its own tilted propagator); several batches over
several lanes; checkpointed slices; patterns streamed from host memory."""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = [
    # n, S, M, mode, per-position rows, step tilts, shift, distances, edge, extra
    (32, 3, 2, 0, True, True, True, None, 1.0, {}),
    (64, 3, 1, 0, False, True, False, None, 1.0, {}),
    (64, 1, 2, 1, True, True, False, None, 1.0, {}),
    (64, 3, 2, 3, True, True, True, None, 1.0, {}),
    (96, 2, 2, 3, False, True, True, None, 1.0, {}),
    (128, 3, 2, 0, True, True, True, [1.5, 2.5], 1.0, {}),
    (128, 3, 1, 0, True, True, False, None, 1.6, {}),
    (64, 4, 2, 3, True, True, True, [1.0, 3.0, 2.0], 1.4, {}),
    (64, 4, 2, 0, True, False, True, [1.0, 3.0, 2.0], 1.0, {}),
    (32, 4, 1, 0, False, False, False, [0.5, 2.0, 1.0], 1.0, {}),
    (64, 5, 2, 3, True, True, True, None, 1.0, {"checkpoint": 2}),
    (64, 2, 2, 3, True, True, True, None, 1.0, {"streamed": True}),
    (192, 2, 2, 3, True, True, True, None, 1.0, {}),
    (192, 1, 1, 1, True, True, False, None, 1.0, {}),
    (192, 3, 1, 0, True, True, False, [1.0, 2.5], 1.3, {}),
    (256, 2, 1, 3, True, True, True, None, 1.0, {}),
    # tilts before the specimen
    (32, 2, 2, 2, True, True, True, None, 1.0, {}),
    (64, 3, 1, 2, True, True, False, None, 1.0, {}),
    (64, 3, 2, 4, True, True, True, None, 1.0, {"bc": True}),
    (128, 2, 2, 4, False, True, True, None, 1.0, {}),
    (96, 1, 2, 4, True, True, False, None, 1.0, {"bc": True}),
    (64, 2, 2, 2, True, False, True, None, 1.0, {}),
    (64, 3, 2, 4, True, True, True, [1.5, 2.5], 1.0, {"streamed": True}),
    (192, 2, 2, 4, True, True, True, None, 1.0, {"bc": True}),
    (192, 2, 1, 2, True, True, False, None, 1.0, {}),
    (256, 2, 2, 2, False, True, True, None, 1.0, {}),
]


def _tilts(rows, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, (rows, 6)) * np.array([2e-3, 2e-3, 4e-2, 4e-2, 1e-3, 1e-3])).astype(np.float32)


@pytest.mark.parametrize("n,S,M,mode,per_pos,step,shift,dists,edge,extra", CASES)
def test_tilt_pass_matches_oracle(n, S, M, mode, per_pos, step, shift, dists, edge, extra):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print(f"[tilts] n={n} S={S} M={M} mode={mode} rows={'P' if per_pos else 1} step={step} shift={shift} "
          f"dists={dists} edge={edge} {extra}", flush=True)
    P = 11
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=shift, seed=901 + n + S + mode)
    tl = _tilts(P if per_pos else 1, 13 + mode)
    dist = np.asarray(dists if dists is not None else [pr["dz"]], np.float32)
    bc = np.random.default_rng(5).uniform(0.7, 1.3, P).astype(np.float32) if extra.get("bc") else None
    args = dict(dz=pr["dz"], distances=dist.astype(np.float64), wavelength=pr["wavelength"],
                pixel_size_x=pr["px"], pixel_size_y=pr["px"], step_pos=shift, compute_batch=4, loss_weight=0.7,
                edge=edge, tilts=tl, tilt_mode=mode, step_tilts=step, beam_current=bc, step_beam_current=bc is not None)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 cdtype=np.complex64, **args)
    ref64 = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                   cdtype=np.complex128, **args)
    cols = {0: [2, 3], 1: [4, 5], 2: [0, 1], 3: [2, 3, 4, 5], 4: [0, 1, 2, 3, 4, 5]}[mode] if step else []
    if S == 1:
        cols = [c for c in cols if c not in (2, 3)]  # one slice: no propagation for the inside tilt to act on
    assert max(rel(ref[k], ref64[k]) for k in (2, 3)) < TOL / 3  # well conditioned
    for c in cols:
        assert rel(ref[-1][:, c], ref64[-1][:, c]) < TOL / 3, c
    dev = torch.device("cuda", 0)
    streamed = bool(extra.get("streamed"))
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=dist,
        edge_slice_thickness_multiplier=edge,
        measured_array=pr["meas"] if streamed else torch.from_numpy(pr["meas"]).to(dev), load_one_by_one=streamed,
        algorithm_type="lsq_sqrt", this_wavelength=pr["wavelength"], this_step_probe=1, this_step_obj=1,
        this_step_pos_correction=int(shift), this_step_tilts=int(step), pixel_size_x_A=pr["px"],
        pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], compute_batch=4, this_loss_weight=0.7,
        this_tilt_array=tl, this_tilts_correction=np.zeros_like(tl), tilt_mode=mode,
        checkpoint_slices=int(extra.get("checkpoint", 0)), beam_current=bc, this_beam_current_step=int(bc is not None))
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    errs = dict(loss=abs(out[0] - ref[0]) / abs(ref[0]), obj=rel(out[2].cpu().numpy(), ref[2]),
                probe=rel(out[3].cpu().numpy(), ref[3]))
    if shift:
        errs["pos"] = rel(out[4].cpu().numpy(), ref[4])
    if bc is not None:
        errs["beam current"] = rel(out[8].cpu().numpy(), ref[5])
    got = out[5].cpu().numpy()
    assert got.shape == tl.shape
    for c in cols:
        errs[f"tilt col {c}"] = rel(got[:, c], ref[-1][:, c])
        assert float(np.abs(ref[-1][:, c]).max()) > 0
    for c in set(range(6)) - set(cols):
        assert not got[:, c].any(), c
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (k, v)


def test_zero_tilts_and_equal_distances_change_nothing():
    """A tilt array of zeros (mode 0) and a per-slice distance array of equal entries give the pass
    without them bit for bit."""
    import torch

    from pypty_amd import objective

    print("[tilts] zero tilts and equal distances vs the plain pass", flush=True)
    n, S, P = 64, 3, 9
    pr = make_problem(n=n, S=S, K=1, M=2, ny=n + 32, nx=n + 32, P=P, shift=True, seed=77)
    dev = torch.device("cuda", 0)
    base = dict(this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
                this_pos_array=pr["pos"], this_pos_correction=pr["corr"],
                measured_array=torch.from_numpy(pr["meas"]).to(dev), algorithm_type="lsq_sqrt",
                this_wavelength=pr["wavelength"], this_step_probe=1, this_step_obj=1, this_step_pos_correction=1,
                this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"])
    a = call_loss_and_direction(objective.loss_and_direction, **reference_kwargs(
        this_distances=np.array([pr["dz"]], np.float32), **base))
    b = call_loss_and_direction(objective.loss_and_direction, **reference_kwargs(
        this_distances=np.array([pr["dz"]] * (S - 1), np.float32), this_tilt_array=np.zeros((P, 6), np.float32),
        this_tilts_correction=np.zeros((P, 6), np.float32), **base))
    assert a[0] == b[0]
    for k in (2, 3, 4):
        assert torch.equal(a[k], b[k]), k
