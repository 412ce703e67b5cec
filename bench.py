"""Throughput of the multislice forward + loss + adjoint pass on one MI355X.

One step = one full loss-and-gradient pass (object, probe and position
gradients, probe sub-pixel shift active) over every scan position of every
dataset of the workload, followed by the readback of every dataset's loss and
sse.  patterns/s = positions / median per-step wall time (after --warmup
untimed steps).  The default workload is the ragged one: eight seeded datasets
(windows 96/128/192 px, 4-16 slices, 1-5 probe modes, 39,320 positions), each
with autotuned windows per launch, spread over three streams and replayed as one
captured HIP graph per step.  --also adds further workloads to the same line
(default: the multislice configuration, object 512x512x10, 128 px probe with
5 modes, 4096 positions, four batch lanes); --config thick is the
1536x1536x40 object with a 256 px, 10-mode probe over 16384 positions.

The per-kernel breakdown and the roofline come from a second run of a few
steps with one lane and the datasets one after the other, with HIP events
around every launch on its stream.  The roofline is the FP32 vector roof of
the dominant kernel kind: algorithmic flops of its launches / their summed
duration.

Prints one JSON line: patterns/s, the roofline, the CPU baseline (the
oracle's complex64 numpy path on the first --cpu-sample positions of every
dataset, timed on this host with scipy.fft on one worker, as PyPty calls it,
and on all host cores; the faster is the baseline), "recon" (wall-clock of a
plain L-BFGS reconstruction of the single-slice workload through the run()
driver on the GPU against the oracle's L-BFGS on one CPU core, with the NMSE
each reaches) and "warm_start" (the learned warm start, a CNN trained on the
GPU on patterns simulated by the forward kernels, against the plain flat
start, as the wall-clock each needs to reach the plain reconstruction's final
NMSE).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "patterns/s fwd+adjoint multislice (1 GPU); wall-clock to target NMSE vs plain"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def fft_flops(n):
    import math

    return 5.0 * n * n * math.log2(n * n)


def kernel_flops(kind, n, S, M, K):
    """Algorithmic FP32 flops per window of each kernel (FFT 5 N^2 log2 N^2 plus complex products)."""
    f, n2 = fft_flops(n), n * n
    if kind == "forward":
        return M * K * ((2 * S - 1) * f + 6.0 * (2 * S - 1) * n2)
    if kind == "adjoint":
        return M * K * ((2 * S - 1) * f + 6.0 * (3 * S - 1) * n2)
    if kind in ("patch", "objgrad"):
        return S * K * (2 * f + 2.0 * n2)
    if kind == "probe_shift":
        return M * (f + 6.0 * n2)
    if kind == "probegrad":
        return M * (f + 12.0 * n2)
    return 0.0


def pass_flops(n, S, M, K, shift=True):
    """F_alg per pattern of the whole pass (FFT count 4SK + 3M*sh + 4(S-1)MK + 2MK)."""
    F = 4 * S * K + 3 * M * int(shift) + 4 * (S - 1) * M * K + 2 * M * K
    return F * fft_flops(n) + n * n * (6 * (2 * S - 1) * M * K + 6 * (3 * S - 1) * M * K + 3 * M * K + 10)


def pass_bytes(n, S, K):
    """B_alg per pattern: object window read, object-gradient read-modify-write, measured pattern."""
    return 3 * S * K * n * n * 8 + 4 * n * n


def cpu_baseline(datasets, n_sample, workers):
    """Time the oracle (PyPty's numpy op sequence, complex64, scipy.fft) on the first n_sample
    positions of each dataset; returns (patterns/s, seconds, positions timed)."""
    import numpy as np
    import scipy.fft

    from oracle import ms_oracle as orc

    total_t, total_p = 0.0, 0
    with scipy.fft.set_workers(workers):
        for ds in datasets:
            seq = np.arange(min(n_sample, ds.positions.shape[0]))
            t0 = time.perf_counter()
            orc.loss_and_direction(ds.obj, ds.probe, ds.positions, ds.corrections, ds.measured, seq, dz=ds.dz_A,
                                   wavelength=ds.wavelength_A, pixel_size_x=ds.pixel_A, pixel_size_y=ds.pixel_A,
                                   algorithm="lsq_sqrt", step_obj=True, step_probe=True, step_pos=True,
                                   compute_batch=16, cdtype=np.complex64)
            total_t += time.perf_counter() - t0
            total_p += len(seq)
    return total_p / total_t, total_t, total_p


def cnn_train_flops(n_in, patch, batch):
    """Algorithmic GEMM flops of one training step (forward, weight and input gradients)."""
    from pypty_amd.warmstart import layer_spec

    f = 0.0
    for li, (kind, cin, cout, relu, hw) in enumerate(layer_spec(n_in, patch)):
        P = batch * (hw // 2) ** 2 if kind == "conv" else batch
        Kc = cin * 16 if kind == "conv" else cin
        f += 2.0 * P * Kc * cout * (3 if li > 0 else 2)
    return f


def warm_start_benchmark(dev, pairs, train_epochs, rec_epochs, step0=1e-3):
    """Learned warm start against the plain start on the multislice workload (512x512x10, 128 px x 5
    modes, 64x64 scan): train the network on `pairs` simulated patterns, stitch its predictions
    into a starting object, then run the same L-BFGS reconstruction (object updates) from the flat and
    from the warm start, refining object, probe (5 modes) and positions as the workload specifies,
    both from epoch 5 on (after the object alone, as reconstructions from a flat start are run), with
    two first line-search steps per epoch: `step0` then 1, and `step0` throughout.  The target is the plain run's final NMSE and each run is timed to the first epoch
    that reaches it; the warm wall-clock includes the patterns' upload, the network's inference and
    the stitching (its reconstruction then reads the uploaded patterns; the plain run uploads them
    itself), the one-time simulation and training are reported beside it."""
    import numpy as np
    import torch

    from pypty_amd import iterative, metrics, synth, warmstart

    n, patch, batch = 128, 32, 256
    t0 = time.perf_counter()
    x, t = warmstart.simulate_pairs(pairs, n=n, patch=patch, n_slices=10, device=dev, seed=3000)
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t0
    net = warmstart.WarmStartNet(n, patch, dev, seed=0)
    t0 = time.perf_counter()
    hist = net.train(x, t, epochs=train_epochs, batch=batch, lr=1e-3, seed=1, lr_end=1e-4)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    steps = train_epochs * (pairs // batch)
    del x, t
    ds = synth.config1(dev, n_scan=64)
    ny, nx, S, K = ds.obj.shape
    pos = ds.positions.astype(np.float64) + ds.corrections
    sh = np.round(pos.min(axis=0)).astype(int)
    truth = metrics.projected_phase(ds.obj)
    tr = truth[sh[0]:, sh[1]:]
    mask = metrics.fov_mask(tr.shape, pos - pos.min(axis=0), n)
    base = dict(dataset=ds.measured, probe=ds.probe, positions=pos, acc_voltage=300.0, pixel_size_x_A=ds.pixel_A,
                pixel_size_y_A=ds.pixel_A, slice_distances=np.array([ds.dz_A]), num_slices=S, algorithm="lsq_sqrt",
                update_obj=1, update_probe=lambda e: int(e >= 5), update_probe_pos=lambda e: int(e >= 5),
                compute_batch=256, output_folder="", print_flag=0)
    iterative.run(dict(base, obj=np.ones_like(ds.obj), epoch_max=2, update_step_bfgs=step0))  # untimed warm-up

    def shifted(o0):
        # run() moves the scan to the origin (pypty/utils.py:1730); the object frame moves with it
        o = np.ones_like(ds.obj)
        o[:ny - sh[0], :nx - sh[1]] = o0[sh[0]:, sh[1]:]
        return o

    def recon(o0, t_extra, step, data=None):
        curve, spent = [], [0.0]

        def cb(ep, obj, probe):
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            rec = metrics.projected_phase(obj)[:tr.shape[0], :tr.shape[1]]
            curve.append((c0 - T0 - spent[0] + t_extra, metrics.nmse(rec, tr, mask)))
            spent[0] += time.perf_counter() - c0

        np.random.seed(0)
        torch.cuda.synchronize()
        T0 = time.perf_counter()
        iterative.run(dict(base, obj=o0, epoch_max=rec_epochs, epoch_callback=cb, update_step_bfgs=step,
                           **({} if data is None else {"dataset": data})))
        return curve

    # the warm build uploads the patterns once; its reconstruction then reads them from HBM (the plain
    # run uploads them in its first epoch).  An untimed build from a few patterns first loads the
    # library kernels the build uses, as the untimed run above does for the reconstruction.
    warmstart.warm_start_object(net, ds.measured[:256], ds.positions[:256], ds.obj.shape)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    meas_dev = torch.from_numpy(ds.measured).to(dev)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t0
    warm_obj = warmstart.warm_start_object(net, meas_dev, ds.positions, ds.obj.shape)
    t_build = time.perf_counter() - t0
    mask0 = metrics.fov_mask(truth.shape, pos, n)

    # NMSE need not fall monotonically (the runs end up fitting the noise), so each run is timed to
    # the first epoch at or below a target
    def first(curve, target):
        return next((tt for tt, e in curve if e <= target), None)

    def compare(step):
        plain = recon(np.ones_like(ds.obj), 0.0, step)
        warm = recon(shifted(warm_obj), t_build, step, meas_dev)
        target = plain[-1][1]
        t_plain, t_warm = first(plain, target), first(warm, target)
        # the same comparison with the plain run stopped earlier: target = its NMSE after k epochs
        by_epochs = []
        for k in range(5, rec_epochs + 1, 5):
            ek = plain[k - 1][1]
            tp, tw = first(plain, ek), first(warm, ek)
            by_epochs.append({"plain_epochs": k, "target_nmse": round(ek, 4), "plain_s": round(tp, 3),
                              "warm_s": None if tw is None else round(tw, 3),
                              "warm_over_plain": None if tw is None else round(tw / tp, 3)})
        return {"plain_s_total": round(plain[-1][0], 3), "target_nmse": round(target, 4),
                "plain_s_to_target": round(t_plain, 3), "warm_s_to_target": None if t_warm is None else round(t_warm, 3),
                "warm_over_plain": None if t_warm is None else round(t_warm / t_plain, 3),
                # the warm run's first arrival at plain's final NMSE against plain's whole run (the
                # wall-clock plain takes to produce that final NMSE)
                "warm_over_plain_run": None if t_warm is None else round(t_warm / plain[-1][0], 3),
                "warm_final_nmse": round(warm[-1][1], 4), "by_plain_epochs": by_epochs,
                "curves_s_nmse": {"plain": [[round(t, 3), round(e, 4)] for t, e in plain],
                                  "warm": [[round(t, 3), round(e, 4)] for t, e in warm]}}

    # first line-search step: step0 in the first epoch (the object-norm-scaled first direction overshoots
    # with more) and the reference's default 1 afterwards, or step0 in every epoch
    fast = compare(lambda epoch: step0 if epoch == 0 else 1.0)
    const = compare(step0)
    out = {"workload": "multislice object 512x512x10, probe 128px x5 modes, 64x64 scan, lsq_sqrt, object, probe and "
                       "position updates (probe and positions from epoch 5); each run timed to its first epoch at the target; "
                       f"first L-BFGS step {step0} in epoch 0 then 1 (headline), and {step0} in every epoch",
           "network": "4 conv 4x4/2 (1-16-32-32-64) + dense 4096-1024 + dense 1024-1024, BF16 MFMA GEMMs",
           "train_pairs": pairs, "train_epochs": train_epochs, "sim_s": round(t_sim, 2),
           "train_s": round(t_train, 2), "train_loss": [round(hist[0], 5), round(hist[-1], 5)],
           "train_gemm_tflops": round(cnn_train_flops(n, patch, batch) * steps / t_train / 1e12, 3),
           "recon_epochs": rec_epochs, "warm_build_s": round(t_build, 3),
           "warm_build_upload_s": round(t_upload, 3),
           "start_nmse_flat": round(metrics.nmse(np.zeros_like(truth), truth, mask0), 4),
           "start_nmse_warm": round(metrics.nmse(metrics.projected_phase(warm_obj), truth, mask0), 4)}
    out.update({k: v for k, v in fast.items() if k != "curves_s_nmse"})
    out["curves_s_nmse"] = fast["curves_s_nmse"]
    out["constant_step"] = const
    return out


def thick_recon_benchmark(dev, runs, pairs, train_epochs, rec_epochs, step0=1e-3, n_scan=128):
    """L-BFGS reconstruction of the thick workload at its full size (1536x1536x40 object, 256 px probe
    with 10 modes, 128x128 scan, 512 windows per launch, slice checkpoints every 8): `rec_epochs`
    epochs refining the object, and the probe and positions from epoch 5, with the first line-search
    step `step0` in the epochs that start a fresh L-BFGS history (0, and 5 where the probe and the
    positions join; a first step of 1 there took a trial probe to a NaN loss) and 1 in the others,
    from the flat start ('plain') and from the learned start
    ('warm': a network trained on `pairs` patterns of random 40-slice objects simulated at 256 px, its
    predictions stitched into the starting object; the build's wall-clock counts).  Each run prints
    a progress line per epoch; the result holds each run's (seconds, NMSE) curve."""
    import numpy as np
    import torch

    from pypty_amd import iterative, metrics, synth, warmstart

    n, patch, batch = 256, 48, 256
    t0 = time.perf_counter()
    ds = synth.config4(dev, n_scan=n_scan)
    t_data = time.perf_counter() - t0
    ny, nx, S, K = ds.obj.shape
    pos = ds.positions.astype(np.float64) + ds.corrections
    sh = np.round(pos.min(axis=0)).astype(int)
    truth = metrics.projected_phase(ds.obj)
    tr = truth[sh[0]:, sh[1]:]
    mask = metrics.fov_mask(tr.shape, pos - pos.min(axis=0), n)
    base = dict(dataset=ds.measured, probe=ds.probe, positions=pos, acc_voltage=300.0, pixel_size_x_A=ds.pixel_A,
                pixel_size_y_A=ds.pixel_A, slice_distances=np.array([ds.dz_A]), num_slices=S, algorithm="lsq_sqrt",
                update_obj=1, update_probe=lambda e: int(e >= 5), update_probe_pos=lambda e: int(e >= 5),
                compute_batch=512, checkpoint_slices=8, output_folder="", print_flag=0,
                update_step_bfgs=lambda e: step0 if e in (0, 5) else 1.0)
    print(f"[thick] data {t_data:.1f} s; untimed one-epoch run", file=sys.stderr, flush=True)
    iterative.run(dict(base, obj=np.ones_like(ds.obj), epoch_max=1))

    def recon(name, o0, t_extra, data=None):
        curve, spent = [], [0.0]

        def cb(ep, obj, probe):
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            rec = metrics.projected_phase(obj)[:tr.shape[0], :tr.shape[1]]
            curve.append((c0 - T0 - spent[0] + t_extra, metrics.nmse(rec, tr, mask)))
            print(f"[thick] {name} epoch {ep} {curve[-1][0]:.1f} s NMSE {curve[-1][1]:.4f}", file=sys.stderr,
                  flush=True)
            spent[0] += time.perf_counter() - c0

        np.random.seed(0)
        torch.cuda.synchronize()
        T0 = time.perf_counter()
        iterative.run(dict(base, obj=o0, epoch_max=rec_epochs, epoch_callback=cb,
                           **({} if data is None else {"dataset": data})))
        return [[round(t, 2), round(e, 4)] for t, e in curve]

    out = {"workload": f"thick object 1536x1536x40 dz1A, probe 256px x10 modes, {n_scan}x{n_scan} scan "
                       f"({n_scan * n_scan} positions), "
                       "lsq_sqrt, 512 windows per launch, slice checkpoints every 8; object refined from epoch 0, "
                       f"probe and positions from epoch 5; first L-BFGS step {step0} in epochs 0 and 5, 1 in the others",
           "recon_epochs": rec_epochs,
           "start_nmse_flat": round(metrics.nmse(np.zeros_like(truth), truth, metrics.fov_mask(truth.shape, pos, n)), 4)}
    if runs in ("plain", "both"):
        out["plain"] = recon("plain", np.ones_like(ds.obj), 0.0)
    if runs in ("warm", "both"):
        t0 = time.perf_counter()
        x, t = warmstart.simulate_pairs(pairs, n=n, patch=patch, n_slices=S, pixel_A=ds.pixel_A, dz_A=ds.dz_A,
                                        device=dev, seed=4500)
        torch.cuda.synchronize()
        t_sim = time.perf_counter() - t0
        net = warmstart.WarmStartNet(n, patch, dev, seed=0)
        t0 = time.perf_counter()
        hist = net.train(x, t, epochs=train_epochs, batch=batch, lr=1e-3, seed=1, lr_end=1e-4)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        del x, t
        print(f"[thick] simulated {pairs} pairs in {t_sim:.1f} s, trained in {t_train:.1f} s", file=sys.stderr,
              flush=True)
        warmstart.warm_start_object(net, ds.measured[:256], ds.positions[:256], ds.obj.shape)  # loads kernels
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        meas_dev = torch.from_numpy(ds.measured).to(dev)
        warm_obj = warmstart.warm_start_object(net, meas_dev, ds.positions, ds.obj.shape)
        torch.cuda.synchronize()
        t_build = time.perf_counter() - t0
        o = np.ones_like(ds.obj)
        o[:ny - sh[0], :nx - sh[1]] = warm_obj[sh[0]:, sh[1]:]  # run() moves the scan to the origin
        out.update({"train_pairs": pairs, "train_epochs": train_epochs, "sim_s": round(t_sim, 2),
                    "train_s": round(t_train, 2), "train_loss": [round(hist[0], 5), round(hist[-1], 5)],
                    "warm_build_s": round(t_build, 3),
                    "start_nmse_warm": round(metrics.nmse(metrics.projected_phase(warm_obj), truth,
                                                          metrics.fov_mask(truth.shape, pos, n)), 4)})
        out["warm"] = recon("warm", o, t_build, meas_dev)
    return out


def recon_benchmark(dev, epochs, cpu_epochs):
    """Plain reconstruction of the single-slice workload: run() on the GPU, the oracle's bfgs_epoch on the CPU.

    Both start from a flat object and the true probe and update the object only.  The NMSE is measured over
    the scanned field of view (metrics.nmse: constant phase offset removed).  run() shifts positions to the
    origin (pypty/utils.py:1730), so its object frame is offset by the scan origin."""
    import numpy as np
    import torch

    from oracle import bfgs_oracle
    from oracle import ms_oracle as orc
    from pypty_amd import iterative, metrics, synth

    ds = synth.config0(dev)
    n = ds.probe.shape[0]
    ny, nx = ds.obj.shape[:2]
    P = ds.positions.shape[0]
    pos = ds.positions.astype(np.float64) + ds.corrections
    truth = metrics.projected_phase(ds.obj)
    params = dict(dataset=ds.measured, probe=ds.probe, positions=pos, obj=np.ones((ny, nx, 1, 1), np.complex64),
                  acc_voltage=300.0, pixel_size_x_A=ds.pixel_A, pixel_size_y_A=ds.pixel_A,
                  slice_distances=np.array([ds.dz_A]), num_slices=1, algorithm="lsq_sqrt", epoch_max=epochs,
                  update_obj=1, update_probe=0, update_probe_pos=0, compute_batch=256, output_folder="",
                  print_flag=0)
    # an untimed 2-epoch run first (workspace allocation, tables, library load), as the timed
    # passes of the throughput measurement follow warm-up passes
    iterative.run(dict(params, epoch_max=2))
    sh = np.round(pos.min(axis=0)).astype(int)
    tr = truth[sh[0]:, sh[1]:]
    mask_g = metrics.fov_mask(tr.shape, pos - pos.min(axis=0), n)
    gpu_curve, spent = [], [0.0]

    def cb(ep, obj, probe):
        # NMSE after every epoch; the time spent here is excluded from the wall-clock
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        rec = metrics.projected_phase(obj)[:tr.shape[0], :tr.shape[1]]
        gpu_curve.append((c0 - t0 - spent[0], metrics.nmse(rec, tr[:rec.shape[0], :rec.shape[1]],
                                                           mask_g[:rec.shape[0], :rec.shape[1]])))
        spent[0] += time.perf_counter() - c0

    np.random.seed(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = iterative.run(dict(params, epoch_callback=cb))
    t_gpu, nmse_gpu = gpu_curve[-1]
    mask_c = metrics.fov_mask(truth.shape, pos, n)
    nmse0 = metrics.nmse(np.zeros_like(truth), truth, mask_c)
    out = {"workload": "single slice: object 256x256 (hexagonal lattice, pure phase), probe 64 px x1 mode, "
                       "16x16 raster (256 positions), dose 1e4, lsq_sqrt, object updates, flat start",
           "epochs": epochs, "gpu_s": round(t_gpu, 3), "gpu_nmse": round(nmse_gpu, 5), "start_nmse": round(nmse0, 5),
           "gpu_losses": [float(x) for x in res["losses"][:1] + res["losses"][-1:]]}
    if cpu_epochs > 0:
        state = {"obj": np.ones((ny, nx, 1, 1), np.complex64), "probe": orc.fourier_clean(ds.probe, cutoff=2 / 3)
                 .astype(np.complex64), "positions": ds.positions, "corr": ds.corrections}
        hist = bfgs_oracle.new_history()
        data = {"measured": ds.measured, "sequence": np.arange(P), "dz": ds.dz_A, "wavelength": ds.wavelength_A,
                "px": ds.pixel_A, "algorithm": "lsq_sqrt", "compute_batch": 256}
        step, t_cpu, cpu_curve = 1.0, 0.0, []
        for e in range(cpu_epochs):
            np.random.seed(e)
            t1 = time.perf_counter()
            r = bfgs_oracle.bfgs_epoch(state, hist, data, update_obj=True, update_probe=False, update_pos=False,
                                       actual_step=step)
            t_cpu += time.perf_counter() - t1
            step = r[2]
            cpu_curve.append((t_cpu, metrics.nmse(metrics.projected_phase(state["obj"]), truth, mask_c)))
        # the two runs take slightly different paths; both are timed to the looser of their final NMSEs
        # (from the flat start they can need different numbers of epochs, so both are reported)
        target = max(nmse_gpu, cpu_curve[-1][1])
        ig = next(i for i, (t, e) in enumerate(gpu_curve) if e <= target)
        ic = next(i for i, (t, e) in enumerate(cpu_curve) if e <= target)
        tg, tc = gpu_curve[ig][0], cpu_curve[ic][0]
        out.update({"cpu_epochs": cpu_epochs, "cpu_s": round(t_cpu, 3), "cpu_nmse": round(cpu_curve[-1][1], 5),
                    "cpu_cores": 1, "target_nmse": round(target, 5), "gpu_s_to_target": round(tg, 4),
                    "gpu_epochs_to_target": ig + 1, "cpu_s_to_target": round(tc, 3), "cpu_epochs_to_target": ic + 1,
                    "speedup_to_target": round(tc / tg, 1), "speedup_per_epoch": round((t_cpu / cpu_epochs) /
                                                                                       (t_gpu / epochs), 1)})
    return out


WORKLOADS = {
    "ragged": "8 seeded datasets (seeds 2000-2007), windows {96,128,192} px, 4-16 slices, 1-5 probe modes, "
              "scans 32x32-96x96 (39,320 positions), dose 1e3-1e5, one full pass over every dataset per step, "
              "autotuned windows per launch, HIP-graph replay over --streams streams (default 3)",
    "multislice": "multislice object 512x512x10 dz2A, probe 128px x5 modes, 64x64 scan (4096 positions), "
                  "batch 256, lsq_sqrt, object+probe+position gradients",
    "thick": "thick object 1536x1536x40 dz1A, probe 256px x10 modes, 128x128 scan (16384 positions), batch 512, "
             "object+probe+position gradients, slice waves checkpointed every 8 slices and recomputed",
}
# name of each workload in profiles/ file names
PROFILE_TAG = {"ragged": "ragged_8datasets_n96-192", "multislice": "multislice_512x512x10_n128_m5",
               "thick": "thick_1536x1536x40_n256_m10"}


def build_workload(name, dev, args):
    """The datasets and the RaggedStep of a workload, tuned and (for ragged) graph-captured."""
    from pypty_amd import ragged, synth

    t0 = time.perf_counter()
    if name == "ragged":
        datasets = synth.ragged_datasets(dev)
        kw = dict(batch=256, streams=args.streams, lanes=1, checkpoint=0)
        tune, graph = True, True
    elif name == "multislice":
        datasets = [synth.config1(dev, n_scan=args.scan)]
        kw = dict(batch=256, streams=1, lanes=4, checkpoint=0)
        tune, graph = False, False
    else:
        datasets = [synth.config4(dev, n_scan=args.thick_scan)]
        kw = dict(batch=512, streams=1, lanes=2, checkpoint=8)
        tune, graph = False, False
    t_gen = time.perf_counter() - t0
    if getattr(args, "eager", False):
        graph = False
    step = ragged.RaggedStep(datasets, dev, **kw)
    if args.batches:
        # windows per launch given per dataset (e.g. a tuned choice, to skip the tuner in profiled runs)
        given = [int(b) for b in args.batches.split(",")]
        if len(given) != len(step.jobs):
            raise SystemExit(f"bench.py: --batches needs {len(step.jobs)} values")
        for j, b in zip(step.jobs, given):
            j.set_batch(b)
        step.tuned = given
    elif tune:
        step.tune((64, 128, 256, 512))
    info = {"lanes": kw["lanes"], "streams": step.n_streams, "checkpoint_slices": kw["checkpoint"],
            "hip_graph": graph, "data_gen_s": round(t_gen, 1)}
    return datasets, step, graph, info


def time_steps(step, graph, steps, warmup):
    """Per-step wall times: each step is one full pass over every dataset (the graph replay, or the
    eager launches) followed by the readback of every dataset's loss and sse, which synchronises."""
    import torch

    for _ in range(warmup):
        step.run()
        step.losses().tolist()
    if graph:
        step.capture()
        step.replay()
        step.losses().tolist()
    torch.cuda.synchronize()
    times, losses = [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        if graph:
            step.replay()
        else:
            step.run()
        losses = step.losses().tolist()
        times.append(time.perf_counter() - t0)
    return times, losses


def kernel_breakdown(step, steps):
    """The same steps again, eagerly, datasets one after the other and one batch lane, with HIP events
    around every launch on the stream it is issued on: total ms and launches per kernel kind."""
    import torch

    from pypty_amd import _lib

    lanes = [j.lanes for j in step.jobs]
    for j in step.jobs:
        if j.lanes != 1:
            j.set_lanes(1)
    step.run(concurrent=False)  # the one-lane engines take their workspaces
    torch.cuda.synchronize()
    _lib.timing_reset()
    _lib.timing_enable(True)
    for _ in range(steps):
        step.run(concurrent=False)
    torch.cuda.synchronize()
    kernels = {k: v for k, v in _lib.timing_read().items() if v[1] > 0}
    _lib.timing_enable(False)
    for j, ln in zip(step.jobs, lanes):
        if ln != 1:
            j.set_lanes(ln)
    return kernels


def roofline_of(kernels, step, steps, tag):
    """FP32 roofline of the dominant kernel kind: algorithmic flops of all its launches in a step
    (kernel_flops per window x windows) over the HIP-event time of those launches."""
    dom = max(kernels, key=lambda k: kernels[k][0])
    tot_ms, launches = kernels[dom]
    avg_s = tot_ms / launches / 1e3
    flops_step = sum(kernel_flops(dom, *j.shape) * j.P for j in step.jobs)
    per_launch = flops_step * steps / launches
    achieved = per_launch / avg_s / 1e12
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None, "kernel": dom,
           "avg_launch_ms": round(avg_s * 1e3, 4), "launches_per_step": launches // steps,
           "note": "FP32 vector ALU roof (the FFT-dominated path uses VALU and LDS, not MFMA); "
                   "peak = MI355X FP32 157.3 TFLOP/s"}
    # HBM bytes per launch of this kernel at this workload, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE (scaled by a plain-copy calibration) and committed under profiles/
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_{tag}.txt")
    if os.path.exists(tf):
        with open(tf) as fh:
            for line in fh:
                if line.startswith("hbm_bytes_per_launch:"):
                    out["traffic"] = float(line.split(":", 1)[1])
    return out


def measure(name, dev, args, kernel_timing=True):
    """Throughput of one workload: patterns/s = positions / median step time."""
    import statistics

    datasets, step, graph, info = build_workload(name, dev, args)
    times, losses = time_steps(step, graph, args.steps, args.warmup)
    med = statistics.median(times)
    P = step.positions
    res = {"workload": WORKLOADS[name], "positions": P, "value": round(P / med, 2),
           "ms_per_step": round(med * 1e3, 3), "mean_ms_per_step": round(statistics.fmean(times) * 1e3, 3),
           "min_ms_per_step": round(min(times) * 1e3, 3),
           "compute_batch": step.tuned if step.tuned else [j.cfg.compute_batch for j in step.jobs], **info,
           "datasets": [{"n": j.shape[0], "slices": j.shape[1], "probe_modes": j.shape[2], "obj_modes": j.shape[3],
                         "positions": j.P} for j in step.jobs],
           "loss_sse_last_dataset": losses[-1]}
    flops_step = sum(pass_flops(*j.shape, True) * j.P for j in step.jobs)
    bytes_step = sum(pass_bytes(j.shape[0], j.shape[1], j.shape[3]) * j.P for j in step.jobs)
    res["alg_gflop_per_pattern"] = round(flops_step / P / 1e9, 4)
    res["alg_tflops"] = round(flops_step / med / 1e12, 3)
    res["pass_frac_fp32"] = round(flops_step / med / 1e12 / FP32_PEAK_TFLOPS, 4)
    res["hbm_frac_alg"] = round(bytes_step / med / (HBM_PEAK_GBS * 1e9), 4)
    kernels = {}
    if kernel_timing:
        ksteps = max(1, min(args.steps, args.kernel_steps))
        kernels = kernel_breakdown(step, ksteps)
        res["roofline"] = roofline_of(kernels, step, ksteps, PROFILE_TAG[name])
        res["kernels_ms_per_step"] = {k: round(v[0] / ksteps, 3) for k, v in kernels.items()}
    return datasets, step, res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=sorted(WORKLOADS), default="ragged")
    ap.add_argument("--also", default="multislice",
                    help="comma-separated further workloads measured beside the headline ('none' for none)")
    ap.add_argument("--scan", type=int, default=64, help="scan grid side of the multislice workload")
    ap.add_argument("--thick-scan", type=int, default=128, help="scan grid side of the thick workload")
    ap.add_argument("--cpu-sample", type=int, default=32,
                    help="positions per dataset timed on the CPU (0: skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--kernel-steps", type=int, default=3, help="steps of the per-kernel HIP-event pass")
    ap.add_argument("--batches", default="",
                    help="comma-separated windows per launch of each dataset of the headline workload "
                         "(instead of the autotuner)")
    ap.add_argument("--streams", type=int, default=3,
                    help="streams the ragged workload's datasets are spread over (round robin) inside the graph")
    ap.add_argument("--eager", action="store_true",
                    help="launch the headline steps eagerly instead of replaying a captured HIP graph "
                         "(for counter profiles)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for rocprofv3 counter scaling)")
    ap.add_argument("--recon-epochs", type=int, default=20,
                    help="epochs of the reconstruction wall-clock measurement (0: skip)")
    ap.add_argument("--warm-start-pairs", type=int, default=50000,
                    help="training pairs of the learned warm-start measurement (0: skip)")
    ap.add_argument("--warm-start-epochs", type=int, default=40)
    ap.add_argument("--warm-recon-epochs", type=int, default=20)
    ap.add_argument("--thick-recon", choices=("plain", "warm", "both"), default=None,
                    help="instead of the throughput line: the thick workload's L-BFGS reconstruction at full size "
                         "from the flat start, the learned start or both (one JSON line with their NMSE curves)")
    ap.add_argument("--thick-recon-epochs", type=int, default=50)
    ap.add_argument("--thick-warm-pairs", type=int, default=20000)
    ap.add_argument("--thick-warm-epochs", type=int, default=20)
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the pass is measured on one GPU only (the multi-GPU split is a design, not built); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    dev = torch.device("cuda", 0)
    if args.thick_recon:
        res = thick_recon_benchmark(dev, args.thick_recon, args.thick_warm_pairs, args.thick_warm_epochs,
                                    args.thick_recon_epochs, n_scan=args.thick_scan)
        print(json.dumps({"thick_recon": res}), flush=True)
        return 0
    datasets, step, head = measure(args.config, dev, args, kernel_timing=not args.no_kernel_timing)
    print(f"[bench] {args.config}: {head['value']:.1f} patterns/s, median {head['ms_per_step']:.2f} ms per step",
          file=sys.stderr, flush=True)
    extra = {}
    for name in [x for x in args.also.split(",") if x and x not in (args.config, "none")]:
        _, st, r = measure(name, dev, args, kernel_timing=not args.no_kernel_timing)
        extra[name] = r
        del st
        print(f"[bench] {name}: {r['value']:.1f} patterns/s", file=sys.stderr, flush=True)
    del step

    if args.copy_calibration_mib:
        a = torch.empty(args.copy_calibration_mib << 20, dtype=torch.uint8, device=dev)
        b = a.clone()
        torch.cuda.synchronize()
        del a, b

    cpu = None
    if args.cpu_sample > 0:
        v1, dt1, npos = cpu_baseline(datasets, args.cpu_sample, 1)
        ncores = max(1, min(os.cpu_count() or 1, 16))
        va, dta, _ = cpu_baseline(datasets, args.cpu_sample, ncores)
        faster = max((v1, 1, dt1), (va, ncores, dta))
        cpu = {"value": round(faster[0], 3), "unit": "patterns/s", "cores": faster[1], "kind": "port",
               "sample": f"oracle complex64 numpy/scipy.fft path (PyPty op order), first {args.cpu_sample} "
                         f"positions of each of the {len(datasets)} dataset(s) ({npos} patterns); the faster of "
                         f"scipy.fft workers=1 ({v1:.2f}/s, {dt1:.1f} s) and workers={ncores} "
                         f"({va:.2f}/s, {dta:.1f} s); numpy's elementwise work runs on one core",
               "one_core": {"value": round(v1, 3), "seconds": round(dt1, 1)},
               "all_cores": {"value": round(va, 3), "cores": ncores, "seconds": round(dta, 1)}}
        print(f"[bench] cpu baseline {cpu['value']:.2f} patterns/s", file=sys.stderr, flush=True)

    recon = recon_benchmark(dev, args.recon_epochs, args.recon_epochs if args.cpu_sample > 0 else 0) \
        if args.recon_epochs > 0 else None
    warm = warm_start_benchmark(dev, args.warm_start_pairs, args.warm_start_epochs, args.warm_recon_epochs) \
        if args.warm_start_pairs > 0 else None

    cfg_info = {"workload": head["workload"], "positions": head["positions"], "compute_batch": head["compute_batch"],
                "hip_graph": head["hip_graph"], "streams": head["streams"], "batch_lanes": head["lanes"],
                "checkpoint_slices": head["checkpoint_slices"], "datasets": head["datasets"]}
    out = {
        "metric": METRIC, "value": head["value"], "unit": "patterns/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex64", "data": "synthetic",
        "config": cfg_info, "roofline": head.get("roofline"), "cpu_baseline": cpu,
        "timing": "patterns/s = positions / median of the per-step wall times; a step is one full pass over "
                  "every dataset and the readback of every dataset's loss (synchronising)",
        "pass": {k: head[k] for k in ("mean_ms_per_step", "min_ms_per_step", "alg_gflop_per_pattern", "alg_tflops",
                                      "pass_frac_fp32", "hbm_frac_alg", "loss_sse_last_dataset", "data_gen_s")},
        "gpu_over_cpu": round(head["value"] / cpu["value"], 1) if cpu else None,
        "kernels_ms_per_step": head.get("kernels_ms_per_step", {}),
        "other_workloads": extra,
        "recon": recon,
        "warm_start": warm,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
