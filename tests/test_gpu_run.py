"""End-to-end reconstruction through the run() mirror: checkpoints, loss log, convergence."""
import csv
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_run_reconstructs_and_writes_reference_files(tmp_path):
    import torch

    from pypty_amd import iterative, metrics, synth

    print("[run] 8 epochs, 64 px probe, 12x12 scan", flush=True)
    dev = torch.device("cuda", 0)
    px = 0.2
    true_obj = synth.hex_lattice_object(160, 160, px, 2.46, 0.3, 0.5, seed=5)
    probe = synth.make_probe(64, px, 300.0, 20.0, 100.0, 1, 0.0, seed=6)
    pos = synth.raster_scan(12, 12, 6.0, origin_px=(10.0, 10.0), jitter_px=0.3, seed=7)
    ints, corr = synth.split_positions(pos)
    lam = synth.electron_wavelength(300.0)
    meas = synth.simulate(true_obj, probe, ints, corr, px, 2.0, lam, 1e5, seed=8, device=dev)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=np.ones((160, 160, 1, 1), np.complex64),
                  acc_voltage=300.0, pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]),
                  num_slices=1, algorithm="lsq_sqrt", epoch_max=8, update_obj=1, update_probe=0,
                  update_probe_pos=0, compute_batch=64, output_folder=out, save_inter_checkpoints=1, print_flag=1)
    res = iterative.run(params)
    losses = res["losses"]
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < 0.5 * losses[0]
    # positions were shifted to the origin by the driver (pypty/utils.py:1730): compare on the scanned FOV
    shift = np.round(pos.min(axis=0)).astype(int)
    rec = metrics.projected_phase(res["obj"])
    truth = metrics.projected_phase(true_obj)[shift[0]:shift[0] + rec.shape[0], shift[1]:shift[1] + rec.shape[1]]
    rec = rec[:truth.shape[0], :truth.shape[1]]
    mask = metrics.fov_mask(truth.shape, pos - pos.min(axis=0), 64)
    e = metrics.nmse(rec, truth, mask)
    e0 = metrics.nmse(np.zeros_like(truth), truth, mask)
    print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}; NMSE {e0:.3f} -> {e:.3f}", flush=True)
    assert e < 0.7 * e0
    for f in ("co.npy", "cp.npy", "cg.npy", "loss.csv", "params.pkl"):
        assert os.path.exists(out + f), f
    co = np.load(out + "co.npy")
    assert co.shape[2:] == (1, 1) and co.dtype == np.complex64
    assert np.load(out + "cg.npy").shape == (144, 2)
    rows = list(csv.reader(open(out + "loss.csv")))
    # the header row, then one row per epoch (pypty/utils.py:2004-2013, 1877-1940)
    assert rows[0][:3] == ["epoch", "time / s", "loss"]
    rows = rows[1:]
    assert len(rows) == 8 and len(rows[0]) == 13


def test_run_refines_a_static_background_and_writes_it(tmp_path):
    """Data with a known background added, reconstructed from the true object with object and
    background refined together (from a flat object the background absorbs the object's misfit; the
    background alone takes the reference's first step scaled by an object step of zero, which is
    far too short at these intensities): run() lowers the loss, refines the background and writes
    cs.npy (pypty/utils.py:2112-2114)."""
    import torch

    from pypty_amd import iterative, synth

    print("[run] 6 epochs refining object and static background", flush=True)
    dev = torch.device("cuda", 0)
    px, n = 0.2, 64
    true_obj = synth.hex_lattice_object(120, 120, px, 2.46, 0.3, 0.5, seed=15)
    probe = synth.make_probe(n, px, 300.0, 20.0, 100.0, 1, 0.0, seed=16)
    pos = synth.raster_scan(8, 8, 6.0, origin_px=(0.0, 0.0), jitter_px=0.0, seed=17)  # run() keeps them
    ints, corr = synth.split_positions(pos)
    meas = synth.simulate(true_obj, probe, ints, corr, px, 2.0, synth.electron_wavelength(300.0), 1e5, seed=18,
                          device=dev)
    f = np.fft.fftshift(np.fft.fftfreq(n, px))
    q = np.hypot(f[None, :], f[:, None])
    band = q <= np.abs(np.fft.fftfreq(n, px)).max() * (2 / 3)  # the DC-centred band mask at cutoff 2/3
    rng = np.random.default_rng(19)
    bg_true = (0.5 * np.sqrt(meas.mean()) * rng.uniform(0.5, 1.5, (n, n)) * band).astype(np.float32)
    meas = (meas + bg_true[None] ** 2).astype(np.float32)
    bg0 = (0.5 * bg_true).astype(np.float32)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=true_obj.astype(np.complex64),
                  acc_voltage=300.0, pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]),
                  num_slices=1, algorithm="lsq", epoch_max=6, update_obj=1, update_probe=0, update_probe_pos=0,
                  static_background=bg0, update_static_background=1, compute_batch=32, output_folder=out,
                  save_inter_checkpoints=1, print_flag=0)
    try:
        res = iterative.run(params)
        losses = res["losses"]
        print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
        assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])), losses
        assert losses[-1] < losses[0]
        cs = np.load(out + "cs.npy")
        assert cs.shape == (n, n) and cs.dtype == np.float32
        e0 = np.linalg.norm((bg0 - bg_true)[band])
        e = np.linalg.norm((cs - bg_true)[band])
        print(f"  background error {e0:.4e} -> {e:.4e}", flush=True)
        assert np.abs(cs - bg0).max() > 0 and np.isfinite(cs).all()
    finally:
        iterative.static_background = None


def test_run_resets_the_history_when_an_l1_weight_switches_on(tmp_path):
    """A phase_norm_weight given as a callable that switches on at epoch 3 changes the objective: run()
    resets the L-BFGS history there (pypty/iterative.py:383-389), and loss.csv's constraints column
    holds the sum of the epoch's constraint contributions (pypty/utils.py:1926), zero before."""
    import torch

    from pypty_amd import iterative, synth

    print("[run] L1 phase weight switching on at epoch 3", flush=True)
    dev = torch.device("cuda", 0)
    px = 0.2
    true_obj = synth.hex_lattice_object(128, 128, px, 2.46, 0.3, 0.5, seed=15)
    probe = synth.make_probe(64, px, 300.0, 20.0, 100.0, 1, 0.0, seed=16)
    pos = synth.raster_scan(8, 8, 6.0, origin_px=(8.0, 8.0), jitter_px=0.3, seed=17)
    ints, corr = synth.split_positions(pos)
    meas = synth.simulate(true_obj, probe, ints, corr, px, 2.0, synth.electron_wavelength(300.0), 1e5, seed=18,
                          device=dev)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=true_obj.copy(), acc_voltage=300.0,
                  pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]), num_slices=1,
                  algorithm="lsq_sqrt", epoch_max=6, update_obj=1, update_probe=0, update_probe_pos=0,
                  compute_batch=32, output_folder=out, save_inter_checkpoints=0, print_flag=0,
                  phase_norm_weight=lambda epoch: 0.0 if epoch < 3 else 50.0)
    resets = []
    real_reset = iterative.reset_bfgs_history
    epoch_now = [None]

    def cb(ep, obj, pr):
        epoch_now[0] = ep + 1

    def counting_reset():
        resets.append(epoch_now[0])
        real_reset()

    iterative.reset_bfgs_history = counting_reset
    try:
        iterative.run(dict(params, epoch_callback=cb))
    finally:
        iterative.reset_bfgs_history = real_reset
    rows = list(csv.reader(open(out + "loss.csv")))[1:]
    assert len(rows) == 6
    cons = [float(r[9]) for r in rows]
    print(f"  constraints column {cons}; history resets before epochs {resets}", flush=True)
    assert all(c == 0.0 for c in cons[:3]) and all(c > 0.0 for c in cons[3:])
    assert "Resetting the history" in rows[3][12]
    assert 3 in resets


def test_run_refines_the_beam_current_and_writes_it(tmp_path):
    """Patterns whose intensities fluctuate by a per-position current squared, reconstructed from the
    true object with object and beam current refined from a current of ones (pypty/iterative.py:417
    starts it so): run() lowers the loss monotonically, moves the current's relative fluctuations
    towards the true ones and writes cb.npy (pypty/utils.py:2096-2099)."""
    import torch

    from pypty_amd import iterative, synth

    print("[run] 6 epochs refining object and beam current", flush=True)
    dev = torch.device("cuda", 0)
    px, n = 0.2, 64
    true_obj = synth.hex_lattice_object(120, 120, px, 2.46, 0.3, 0.5, seed=15)
    probe = synth.make_probe(n, px, 300.0, 20.0, 100.0, 1, 0.0, seed=16)
    pos = synth.raster_scan(8, 8, 6.0, origin_px=(0.0, 0.0), jitter_px=0.0, seed=17)
    ints, corr = synth.split_positions(pos)
    meas = synth.simulate(true_obj, probe, ints, corr, px, 2.0, synth.electron_wavelength(300.0), 1e5, seed=18,
                          device=dev)
    bc_true = np.random.default_rng(20).uniform(0.85, 1.15, meas.shape[0]).astype(np.float32)
    meas = (meas * bc_true[:, None, None] ** 2).astype(np.float32)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=true_obj.astype(np.complex64),
                  acc_voltage=300.0, pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]),
                  num_slices=1, algorithm="lsq_sqrt", epoch_max=6, update_obj=1, update_probe=0,
                  update_probe_pos=0, update_beam_current=1, compute_batch=32, output_folder=out,
                  save_inter_checkpoints=1, print_flag=0)
    try:
        res = iterative.run(params)
        losses = res["losses"]
        print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
        assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])), losses
        assert losses[-1] < losses[0]
        cb = np.load(out + "cb.npy")
        assert cb.shape == (meas.shape[0],) and cb.dtype == np.float32
        np.testing.assert_array_equal(cb, res["beam_current"].cpu().numpy())
        # the probe is not scaled to the dose, so the current also takes up the global intensity scale:
        # compare the currents relative to their means
        rel_true = bc_true / bc_true.mean()
        e0 = np.linalg.norm(1.0 - rel_true)
        e = np.linalg.norm(np.abs(cb) / np.abs(cb).mean() - rel_true)
        print(f"  mean current {np.abs(cb).mean():.4e}; relative current error {e0:.4e} -> {e:.4e}", flush=True)
        assert e < 0.5 * e0
    finally:
        iterative.beam_current = None


def test_run_refines_the_aberrations_and_writes_them(tmp_path):
    """Patterns recorded with a probe defocused by a different amount in each half of the scan (rows
    0 and 1 of aberration_marker), reconstructed from the true object and the in-focus probe (scaled
    to the dose) with object and per-row defocus C10 refined from 0 (pypty/iterative.py:176-183, 419): the loss falls
    monotonically, each row's C10 moves towards its true value, and ca.npy and
    checkpoint_aberrations_epcoh_N.npy hold the last rows (pypty/utils.py:1972-1973, 2115-2117)."""
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import iterative, synth

    print("[run] 8 epochs refining object and per-row defocus", flush=True)
    dev = torch.device("cuda", 0)
    px, n = 0.2, 64
    lam = synth.electron_wavelength(300.0)
    true_obj = synth.hex_lattice_object(120, 120, px, 2.46, 0.3, 0.5, seed=25)
    probe = synth.make_probe(n, px, 300.0, 20.0, 0.0, 1, 0.0, seed=26)
    pos = synth.raster_scan(8, 8, 6.0, origin_px=(0.0, 0.0), jitter_px=0.0, seed=27)
    ints, corr = synth.split_positions(pos)
    P = pos.shape[0]
    marker = (np.arange(P) >= P // 2).astype(np.int64)
    c10 = np.array([15.0, -10.0])
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    meas = np.zeros((P, n, n), np.float32)
    for r in range(2):
        # the probe behind the row's plate: exp(-i pi lambda q^2 C10) times the band mask
        plate = np.exp(-1j * np.pi * lam * q2 * c10[r]) * mask
        pr_r = np.fft.ifft2(np.fft.fft2(probe[:, :, 0]) * plate)[:, :, None].astype(np.complex64)
        m = synth.simulate(true_obj, pr_r, ints, corr, px, 2.0, lam, 1e6, seed=28 + r, device=dev)
        meas[marker == r] = m[marker == r]
    # the probe scaled to the dose (the modelled pattern holds n^2 sum |probe|^2), so that the loss
    # measures the pattern's shape rather than its total
    probe = (probe * np.sqrt(meas.sum(axis=(1, 2)).mean() / (n * n))).astype(np.complex64)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=true_obj.astype(np.complex64),
                  acc_voltage=300.0, pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]),
                  num_slices=1, algorithm="lsq_sqrt", epoch_max=8, update_obj=1, update_probe=0,
                  update_probe_pos=0, aberrations_array=np.zeros((2, 1), np.float32), aberration_marker=marker,
                  update_aberrations_array=1, compute_batch=32, output_folder=out, save_inter_checkpoints=1,
                  save_checkpoints_every_epoch=1, print_flag=0)
    try:
        res = iterative.run(params)
        losses = res["losses"]
        got = res["aberrations_array"].cpu().numpy()
        print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}; C10 per row {got.ravel()} (true {c10})", flush=True)
        assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])), losses
        assert losses[-1] < losses[0]
        assert np.all(np.abs(got.ravel() - c10) < 0.5 * np.abs(c10))
        ca = np.load(out + "ca.npy")
        np.testing.assert_array_equal(ca, got)
        np.testing.assert_array_equal(np.load(out + "checkpoint_aberrations_epcoh_7.npy"), got)
    finally:
        iterative.aberrations_array = None


def test_beam_current_initialisation_pads_like_the_reference():
    """utils.try_to_initialize_beam_current (pypty/utils.py:1766-1789): ones for none; a short current
    padded with ones by the shortfall on both sides (numpy's pad with a scalar width)."""
    import torch

    from pypty_amd import utils

    ones = utils.try_to_initialize_beam_current(None, (7, 32, 32))
    assert ones.is_cuda and ones.dtype == torch.float32 and torch.equal(ones.cpu(), torch.ones(7))
    short = np.array([2.0, 3.0, 4.0], np.float32)
    got = utils.try_to_initialize_beam_current(short, (5, 32, 32)).cpu().numpy()
    np.testing.assert_array_equal(got, np.pad(short, pad_width=2, mode="constant", constant_values=1))
    full = np.arange(5, dtype=np.float32)
    np.testing.assert_array_equal(utils.try_to_initialize_beam_current(full, (5, 32, 32)).cpu().numpy(), full)


def test_run_refines_the_tilts_and_writes_them(tmp_path):
    """Patterns simulated by the FP64 oracle with one after-specimen tilt (tilt_mode 1) over the whole
    scan, reconstructed from the true object and probe with the tilt refined from 0 through its
    correction (pypty/iterative.py:134-135, 174): the loss falls monotonically, the tilt approaches the
    true one, and ct.npy and checkpoint_tilts_epoch_N.npy hold the tilts in use (array plus
    correction; pypty/utils.py:1970-1971, 2104-2106)."""
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import iterative, synth

    print("[run] 6 epochs refining one after-specimen tilt", flush=True)
    px, n = 0.2, 64
    lam = synth.electron_wavelength(300.0)
    true_obj = synth.hex_lattice_object(100, 100, px, 2.46, 0.3, 0.5, seed=35)
    probe = synth.make_probe(n, px, 300.0, 20.0, 0.0, 1, 0.0, seed=36)
    pos = synth.raster_scan(6, 6, 6.0, origin_px=(0.0, 0.0), jitter_px=0.0, seed=37)
    ints, corr = synth.split_positions(pos)
    P = pos.shape[0]
    true_tilt = np.zeros((1, 6))
    true_tilt[0, 4:] = (0.3 * lam / (n * px), -0.2 * lam / (n * px))
    probe = (probe * 30.0).astype(np.complex64)
    meas = orc.loss_and_direction(true_obj.astype(np.complex128), probe.astype(np.complex128), ints, corr,
                                  np.zeros((P, n, n)), np.arange(P), dz=2.0, wavelength=lam, pixel_size_x=px,
                                  pixel_size_y=px, step_obj=False, step_probe=False, cdtype=np.complex128,
                                  return_patterns=True, tilts=true_tilt, tilt_mode=1)[5].astype(np.float32)
    out = str(tmp_path) + os.sep
    params = dict(dataset=meas, probe=probe, positions=pos, obj=true_obj.astype(np.complex64), acc_voltage=300.0,
                  pixel_size_x_A=px, pixel_size_y_A=px, slice_distances=np.array([2.0]), num_slices=1,
                  algorithm="lsq_sqrt", epoch_max=6, update_obj=0, update_probe=0, update_probe_pos=0,
                  tilts=np.zeros((1, 6), np.float32), tilt_mode=1, update_tilts=1, compute_batch=16,
                  output_folder=out, save_inter_checkpoints=1, save_checkpoints_every_epoch=1, print_flag=0,
                  allow_subPixel_shift=False, damping_cutoff_multislice=2 / 3)
    res = iterative.run(params)
    losses = res["losses"]
    got = res["tilts"].cpu().numpy()
    print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}; tilt after {got[0, 4:]} (true {true_tilt[0, 4:]})",
          flush=True)
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < 0.5 * losses[0]
    err0 = np.abs(true_tilt[0, 4:]).sum()
    assert np.abs(got[0, 4:] - true_tilt[0, 4:]).sum() < 0.3 * err0
    assert not got[0, :4].any()
    ct = np.load(os.path.join(out, "ct.npy"))
    assert ct.shape == (1, 6) and np.allclose(ct, got)
    assert os.path.exists(os.path.join(out, "checkpoint_tilts_epoch_5.npy"))
