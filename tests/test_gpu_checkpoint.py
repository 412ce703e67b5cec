"""Slice checkpointing: a pass that keeps only every k-th slice's wave and recomputes each segment in
the adjoint gives the loss and every gradient of the pass that keeps all slices, bit for bit (the
recompute runs the same instructions on the same stored values, and the wave crossing a segment
boundary is stored exactly).  Covers the register tile with four threads per line (32, 96 px) and
with eight (64, 128 px), the slab tiles (192 px in registers, 256 px in workspace memory), uneven
last segments, object modes, several batches per pass and the degenerate k = 1 (at 192 px its
one-slice segment at slice 0 runs the slab pipeline, the others the register tile)."""
import numpy as np
import pytest

from tests._problems import make_problem

pytestmark = pytest.mark.gpu

CASES = [
    # n, S, K, M, P, batch, k
    (32, 5, 2, 2, 9, 4, 2),
    (64, 4, 1, 3, 7, 7, 1),
    (96, 6, 1, 2, 6, 4, 4),
    (128, 7, 2, 2, 5, 3, 3),
    (192, 5, 1, 2, 4, 4, 2),
    (192, 3, 1, 2, 3, 3, 1),
    (256, 4, 1, 1, 3, 2, 3),
]


def _pass(pr, n, S, K, M, P, batch, ck, dev, lanes=2):
    import torch

    from pypty_amd.objective import PassConfig, PtychoPass

    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=pr["obj"].shape[0],
                     obj_nx=pr["obj"].shape[1], n_positions=P, compute_batch=batch, algorithm="lsq_sqrt",
                     step_obj=True, step_probe=True, step_pos=True, shift_active=True, slice_distance=pr["dz"],
                     wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                     checkpoint_slices=ck, lanes=lanes)
    obj = torch.from_numpy(pr["obj"]).to(dev)
    probe = torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev)
    og, pg = torch.empty_like(obj), torch.empty_like(probe)
    posg = torch.zeros(P, 2, dtype=torch.float32, device=dev)
    eng = PtychoPass(cfg, dev)
    ls = eng.run(obj, probe, torch.from_numpy(pr["pos"]).to(dev), torch.from_numpy(pr["corr"]).to(dev),
                 torch.from_numpy(pr["meas"]).to(dev), torch.arange(P, dtype=torch.int32, device=dev),
                 obj_grad=og, probe_grad=pg, pos_grad=posg, sync=False)
    torch.cuda.synchronize()
    return ls.cpu(), og.cpu(), pg.cpu(), posg.cpu(), eng.workspace_bytes


@pytest.mark.parametrize("n,S,K,M,P,batch,ck", CASES)
def test_checkpointed_pass_is_bit_identical(n, S, K, M, P, batch, ck):
    import torch

    print(f"[checkpoint] n={n} S={S} K={K} M={M} P={P} batch={batch} k={ck}", flush=True)
    pr = make_problem(n=n, S=S, K=K, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=81 + n)
    dev = torch.device("cuda", 0)
    full = _pass(pr, n, S, K, M, P, batch, 0, dev)
    seg = _pass(pr, n, S, K, M, P, batch, ck, dev)
    names = ("loss", "object gradient", "probe gradient", "position gradient")
    for name, a, b in zip(names, full[:4], seg[:4]):
        assert torch.isfinite(a).all(), name
        assert torch.equal(a, b), name
    assert float(full[1].abs().max()) > 0 and float(full[3].abs().max()) > 0
    # wave slots per window and mode: S without checkpoints, else the checkpoints of all segments
    # but the last, one k-slice segment and the wave carried between segments; one set per batch lane
    slots = (-(-S // ck) - 1) + ck + 1
    assert full[4] - seg[4] == 2 * (S - slots) * batch * M * K * n * n * 8


def test_drop_in_checkpoint_keyword():
    """loss_and_direction(..., checkpoint_slices=k) returns what the call without it returns."""
    import torch

    from pypty_amd import objective
    from tests._problems import call_loss_and_direction, reference_kwargs

    print("[checkpoint] drop-in n=64 S=5 k=2", flush=True)
    pr = make_problem(n=64, S=5, K=1, M=2, ny=104, nx=104, P=6, shift=True, seed=95)
    dev = torch.device("cuda", 0)
    outs = []
    for ck in (0, 2):
        kw = reference_kwargs(
            this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
            this_pos_array=pr["pos"], this_pos_correction=pr["corr"],
            this_distances=np.array([pr["dz"]], np.float32), measured_array=pr["meas"], algorithm_type="lsq_sqrt",
            this_wavelength=pr["wavelength"], this_step_probe=1, this_step_obj=1, this_step_pos_correction=1,
            this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"],
            checkpoint_slices=ck)
        out = call_loss_and_direction(objective.loss_and_direction, **kw)
        outs.append((out[0], out[1], out[2].cpu(), out[3].cpu(), out[4].cpu()))
    assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
    for a, b in zip(outs[0][2:], outs[1][2:]):
        assert torch.equal(a, b)
