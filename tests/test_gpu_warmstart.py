"""The learned warm start end to end: train the network on the GPU on patterns the multislice kernels
simulate, stitch its predictions for a held-out multislice dataset (different object and probes),
and show what it gains over the flat start: a lower starting NMSE, and, with a constant small first
line-search step (update_step_bfgs = 1e-3), the plain reconstruction's final NMSE reached in fewer
L-BFGS epochs from the warm start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_warm_start_beats_the_flat_start():
    import torch

    from pypty_amd import iterative, metrics, synth, warmstart

    print("[warm start] 32768 pairs, 10 training epochs; 512x512x10, 128 px x5 modes, 32x32 scan", flush=True)
    dev = torch.device("cuda", 0)
    x, t = warmstart.simulate_pairs(32768, n=128, patch=32, n_slices=10, device=dev, seed=3000)
    net = warmstart.WarmStartNet(128, 32, dev, seed=0)
    hist = net.train(x, t, epochs=10, batch=256, lr=1e-3, seed=1)
    print(f"  training loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.8 * hist[0]
    del x, t
    ds = synth.config1(dev, n_scan=32)
    ny, nx, S, K = ds.obj.shape
    pos = ds.positions.astype(np.float64) + ds.corrections
    truth = metrics.projected_phase(ds.obj)
    warm = warmstart.warm_start_object(net, ds.measured, ds.positions, ds.obj.shape)
    mask0 = metrics.fov_mask(truth.shape, pos, 128)
    e_warm0 = metrics.nmse(metrics.projected_phase(warm), truth, mask0)
    print(f"  start NMSE: flat 1.000, warm {e_warm0:.3f}", flush=True)
    assert e_warm0 < 0.8
    sh = np.round(pos.min(axis=0)).astype(int)
    tr = truth[sh[0]:, sh[1]:]
    mask = metrics.fov_mask(tr.shape, pos - pos.min(axis=0), 128)
    curves = {}
    for label, o0 in (("plain", np.ones_like(ds.obj)), ("warm", warm)):
        o = np.ones_like(ds.obj)
        o[:ny - sh[0], :nx - sh[1]] = o0[sh[0]:, sh[1]:]
        curve = []
        cb = lambda ep, obj, probe, c=curve: c.append(  # noqa: E731
            metrics.nmse(metrics.projected_phase(obj)[:tr.shape[0], :tr.shape[1]], tr, mask))
        np.random.seed(0)
        iterative.run(dict(dataset=ds.measured, probe=ds.probe, positions=pos, obj=o, acc_voltage=300.0,
                           pixel_size_x_A=ds.pixel_A, pixel_size_y_A=ds.pixel_A, slice_distances=np.array([ds.dz_A]),
                           num_slices=S, algorithm="lsq_sqrt", epoch_max=8, update_obj=1, update_probe=0,
                           update_probe_pos=0, compute_batch=256, output_folder="", print_flag=0,
                           update_step_bfgs=1e-3, epoch_callback=cb))
        curves[label] = curve
        print(f"  {label}: " + " ".join(f"{e:.3f}" for e in curve), flush=True)
    target = curves["plain"][-1]
    reached = next(i for i, e in enumerate(curves["warm"] + [0.0]) if e <= target)
    print(f"  warm start reaches the plain final NMSE {target:.3f} after {reached + 1} of 8 epochs", flush=True)
    assert reached + 1 <= 6 and curves["warm"][-1] < target
