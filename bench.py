"""Throughput of the multislice forward + loss + adjoint pass on one MI355X.

One step = one full loss-and-gradient pass (object, probe and position
gradients, probe sub-pixel shift active) over every scan position.  The
default workload is the multislice configuration: object 512x512x10 (dz 2 A,
0.2 A pixels), 128 px probe with 5 incoherent modes, 64x64 raster = 4096
positions, 256 windows per launch, Poisson data at 1e4 e-.  --config ragged
runs eight mixed-shape datasets (windows 96-192 px) one after the other per
step; --config thick the 1536x1536x40 object with a 256 px, 10-mode probe.
The loss/sse readback of every pass is inside the timed region.  Each pass runs
with four batch lanes (--lanes): batch i runs on stream i % 4, so one batch's
memory-bound kernels overlap another's compute.
The per-kernel breakdown and the roofline come from a second run of the same
steps with one lane, where every launch has the GPU to itself.

Prints one JSON line: patterns/s (= positions x steps / seconds), the
roofline of the dominant kernel (FP32 flops per launch / HIP-event average
duration), the CPU baseline (the oracle's complex64 numpy path on a bounded
sample of positions, timed on this host, with scipy.fft on one worker as PyPty
calls it and on all host cores), and "recon": the wall-clock of a plain
L-BFGS reconstruction of the single-slice workload (256x256 hexagonal-lattice
phase object, 64 px probe, 16x16 scan, 20 epochs, object updates) through the
run() driver on the GPU (after an untimed 2-epoch warm-up run) against the
oracle's L-BFGS on one CPU core, with the NMSE each reaches and the CPU time to
reach the GPU's final NMSE.  With the multislice workload, "warm_start" adds the
metric's second half: the learned warm start (a CNN trained on the GPU on
patterns simulated by the forward kernels) against the plain flat start, as
the wall-clock each needs to reach the plain reconstruction's final NMSE.
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
    from the warm start, with two first line-search steps per epoch: `step0` then 1, and `step0`
    throughout.  The target is the plain run's final NMSE and each run is timed to the first epoch
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
                update_obj=1, update_probe=0, update_probe_pos=0, compute_batch=256, output_folder="", print_flag=0)
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
    out = {"workload": "multislice object 512x512x10, probe 128px x5 modes, 64x64 scan, lsq_sqrt, object updates; "
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
    "multislice": "multislice object 512x512x10 dz2A, probe 128px x5 modes, 64x64 scan (4096 positions), "
                  "batch 256, lsq_sqrt, object+probe+position gradients",
    "ragged": "8 seeded datasets, windows {96,128,192} px, 4-16 slices, 1-5 probe modes, scans 32x32-96x96, "
              "dose 1e3-1e5, one pass over every dataset per step",
    "thick": "thick object 1536x1536x40 dz1A, probe 256px x10 modes, 64x64 scan (4096 positions), "
             "object+probe+position gradients, slice waves checkpointed every 8 slices and recomputed",
}


class Job:
    """One dataset resident on the device with its engine and gradient buffers."""

    def __init__(self, ds, batch, dev, checkpoint=0, lanes=4):
        self.ds = ds
        self.checkpoint = checkpoint
        self.lanes = lanes
        self._setup(ds, batch, dev)

    def set_batch(self, batch):
        self._setup(self.ds, batch, self.dev)

    def set_lanes(self, lanes):
        self.lanes = lanes
        self._setup(self.ds, self.cfg.compute_batch, self.dev)

    def _setup(self, ds, batch, dev):
        import numpy as np
        import torch

        from pypty_amd.objective import PassConfig, PtychoPass

        ny, nx, S, K = ds.obj.shape
        n, _, M = ds.probe.shape
        self.P = ds.positions.shape[0]
        self.shape = (n, S, M, K)
        # reconstruction starts from a flat object; data come from the true one
        self.obj = torch.ones((ny, nx, S, K), dtype=torch.complex64, device=dev)
        self.probe = torch.from_numpy(np.ascontiguousarray(ds.probe.transpose(2, 0, 1))).to(dev)
        self.pos = torch.from_numpy(ds.positions).to(dev)
        self.corr = torch.from_numpy(ds.corrections).to(dev)
        self.meas = torch.from_numpy(ds.measured).to(dev)
        self.seq = torch.arange(self.P, dtype=torch.int32, device=dev)
        self.cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx,
                              n_positions=self.P, compute_batch=batch, algorithm="lsq_sqrt", planar=False,
                              step_obj=True, step_probe=True, step_pos=True, shift_active=True,
                              slice_distance=ds.dz_A, wavelength=ds.wavelength_A, pixel_size_x=ds.pixel_A,
                              pixel_size_y=ds.pixel_A, checkpoint_slices=self.checkpoint, lanes=self.lanes)
        self.dev = dev
        self.engine = PtychoPass(self.cfg, dev)
        self.og = torch.empty_like(self.obj)
        self.pg = torch.empty_like(self.probe)
        self.posg = torch.zeros_like(self.corr)

    def step(self, sync=True):
        return self.engine.run(self.obj, self.probe, self.pos, self.corr, self.meas, self.seq, obj_grad=self.og,
                               probe_grad=self.pg, pos_grad=self.posg, sync=sync)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--config", choices=sorted(WORKLOADS), default="multislice")
    ap.add_argument("--scan", type=int, default=64, help="scan grid side of the single-dataset workloads")
    ap.add_argument("--cpu-sample", type=int, default=160,
                    help="positions timed on the CPU (per dataset for the ragged workload; 0: skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for rocprofv3 counter scaling)")
    ap.add_argument("--recon-epochs", type=int, default=20,
                    help="epochs of the reconstruction wall-clock measurement (0: skip)")
    ap.add_argument("--autotune", type=int, default=None,
                    help="choose each dataset's windows per launch by timing (default: on for ragged)")
    ap.add_argument("--graph", type=int, default=None,
                    help="capture one step in a HIP graph and replay it (default: on for ragged)")
    ap.add_argument("--warm-start-pairs", type=int, default=50000,
                    help="training pairs of the learned warm-start measurement (multislice workload; 0: skip)")
    ap.add_argument("--warm-start-epochs", type=int, default=40)
    ap.add_argument("--warm-recon-epochs", type=int, default=20)
    ap.add_argument("--streams", type=int, default=None,
                    help="streams the datasets of a step are spread over (default: 2 for ragged, else 1)")
    ap.add_argument("--lanes", type=int, default=None, choices=(1, 2, 3, 4),
                    help="batch lanes of each pass (batch i runs on stream i %% lanes); "
                         "default 4, and 1 for --config ragged")
    ap.add_argument("--checkpoint", type=int, default=None,
                    help="keep every k-th slice's wave and recompute the rest (default: 8 for thick, else 0)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the pass is measured on one GPU only (the multi-GPU split is a design, not built); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from pypty_amd import _lib, synth

    dev = torch.device("cuda", 0)
    t_gen = time.perf_counter()
    if args.config == "multislice":
        datasets = [synth.config1(dev, n_scan=args.scan)]
    elif args.config == "ragged":
        datasets = synth.ragged_datasets(dev)
    else:
        datasets = [synth.config4(dev, n_scan=args.scan)]
    t_gen = time.perf_counter() - t_gen
    if args.lanes is None:
        args.lanes = 1 if args.config == "ragged" else 4
    if args.config == "ragged" and args.lanes > 1:
        # two lanes under this workload (eight datasets over two streams, autotuned, captured in one
        # graph) ended once in a host segmentation fault whose cause is not known; it runs one lane
        sys.exit("bench.py: --config ragged runs with --lanes 1")
    ck = (8 if args.config == "thick" else 0) if args.checkpoint is None else args.checkpoint
    jobs = [Job(ds, args.batch, dev, ck, args.lanes) for ds in datasets]
    P_total = sum(j.P for j in jobs)
    ragged = args.config == "ragged"
    use_tune = ragged if args.autotune is None else bool(args.autotune)
    use_graph = ragged if args.graph is None else bool(args.graph)
    tuned = None
    if use_tune:
        from pypty_amd import autotune

        tuned = []
        for j in jobs:
            n, S, M, K = j.shape
            key = autotune.shape_key(dev, n, S, K, M, j.P)

            def make_pass(b, j=j):
                j.set_batch(b)
                return lambda: j.step(sync=False)

            best = autotune.choose_batch(key, (64, 128, 256, 512), make_pass)
            j.set_batch(best)
            tuned.append(best)
        from pypty_amd.objective import PtychoPass

        PtychoPass.release_workspaces()  # the candidates' workspaces; the chosen engines take theirs next

    # datasets on several default-priority streams (a workspace per stream and shape), joined at the end
    n_streams = max(1, min(len(jobs), (2 if ragged else 1) if args.streams is None else args.streams))
    lanes = [torch.cuda.Stream(dev) for _ in range(n_streams)] if n_streams > 1 else []

    def step(concurrent=True):
        out = None
        if not lanes or not concurrent:
            for j in jobs:
                out = j.step(sync=False)
            return out
        main = torch.cuda.current_stream(dev)
        for ln in lanes:
            ln.wait_stream(main)
        for i, j in enumerate(jobs):
            with torch.cuda.stream(lanes[i % n_streams]):
                out = j.step(sync=False)
        for ln in lanes:
            main.wait_stream(ln)
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    graph = None
    if use_graph:
        # one whole step (every dataset's pass) captured once; the warm-up above allocated the workspaces
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            step()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            out_t = step()
        graph.replay()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if graph is not None:
            graph.replay()
        else:
            out_t = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    loss, sse = out_t.tolist()
    value = P_total * args.steps / elapsed

    kernels = {}
    roofline = None
    if not args.no_kernel_timing:
        # same steps again with HIP events around every launch, for the per-kernel breakdown
        # (datasets one after the other and one batch lane: events on concurrent streams would time
        # each other's kernels)
        if args.lanes > 1:
            for j in jobs:
                j.set_lanes(1)
        _lib.timing_reset()
        _lib.timing_enable(True)
        for _ in range(args.steps):
            step(concurrent=False)
        torch.cuda.synchronize()
        kernels = {k: v for k, v in _lib.timing_read().items() if v[1] > 0}
        _lib.timing_enable(False)
        dom = max(kernels, key=lambda k: kernels[k][0])
        tot_ms, launches = kernels[dom]
        avg_s = tot_ms / launches / 1e3
        # algorithmic flops of all launches of the dominant kernel in one step, per launch
        flops_step = sum(kernel_flops(dom, *j.shape) * j.P for j in jobs)
        per_launch = flops_step * args.steps / launches
        achieved = per_launch / avg_s / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None, "kernel": dom,
                    "avg_launch_ms": round(avg_s * 1e3, 4), "launches_per_step": launches // args.steps}
        # HBM bytes per launch of this kernel at this workload, measured with rocprofv3
        # FETCH_SIZE / WRITE_SIZE (scaled by a plain-copy calibration) and committed under profiles/
        tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_multislice_512x512x10_n128_m5.txt")
        if args.config == "multislice" and os.path.exists(tf) and args.batch == 256 and args.scan == 64:
            with open(tf) as fh:
                for line in fh:
                    if line.startswith("hbm_bytes_per_launch:"):
                        roofline["traffic"] = float(line.split(":", 1)[1])

    if args.copy_calibration_mib:
        a = torch.empty(args.copy_calibration_mib << 20, dtype=torch.uint8, device=dev)
        b = a.clone()
        torch.cuda.synchronize()
        del a, b

    cpu = None
    if args.cpu_sample > 0:
        n_cpu = args.cpu_sample if args.config != "ragged" else max(1, args.cpu_sample // 10)
        if args.config == "thick":
            n_cpu = max(1, args.cpu_sample // 20)
        v1, dt1, npos = cpu_baseline(datasets, n_cpu, 1)
        cpu = {"value": round(v1, 3), "unit": "patterns/s", "cores": 1, "kind": "port",
               "sample": f"oracle complex64 numpy/scipy.fft path (PyPty op order), first {n_cpu} positions of "
                         f"each of the {len(datasets)} dataset(s) ({npos} patterns), scipy workers=1, {dt1:.1f} s"}
        # the same sample with scipy.fft on all host cores (numpy's elementwise work stays on one core)
        ncores = max(1, min(os.cpu_count() or 1, 16))
        va, dta, _ = cpu_baseline(datasets, n_cpu, ncores)
        cpu["all_cores"] = {"value": round(va, 3), "cores": ncores, "seconds": round(dta, 1)}

    recon = recon_benchmark(dev, args.recon_epochs, args.recon_epochs if args.cpu_sample > 0 else 0) \
        if args.recon_epochs > 0 else None
    warm = warm_start_benchmark(dev, args.warm_start_pairs, args.warm_start_epochs, args.warm_recon_epochs) \
        if args.config == "multislice" and args.warm_start_pairs > 0 else None

    flops_step = sum(pass_flops(*j.shape, True) * j.P for j in jobs)
    bytes_step = sum(pass_bytes(j.shape[0], j.shape[1], j.shape[3]) * j.P for j in jobs)
    cfg_info = {"workload": WORKLOADS[args.config], "positions": P_total,
                "compute_batch": tuned if tuned else args.batch, "hip_graph": bool(use_graph),
                "streams": n_streams, "batch_lanes": args.lanes,
                "checkpoint_slices": ck,
                "datasets": [{"n": j.shape[0], "slices": j.shape[1], "probe_modes": j.shape[2],
                              "obj_modes": j.shape[3], "positions": j.P} for j in jobs]}
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "patterns/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex64", "data": "synthetic",
        "config": cfg_info, "roofline": roofline, "cpu_baseline": cpu,
        "pass": {"loss_last_dataset": loss, "sse_last_dataset": sse,
                 "alg_gflop_per_pattern": round(flops_step / P_total / 1e9, 4),
                 "alg_tflops": round(flops_step * args.steps / elapsed / 1e12, 3),
                 "hbm_frac_alg": round(bytes_step * args.steps / elapsed / (HBM_PEAK_GBS * 1e9), 4),
                 "gpu_over_cpu": round(value / cpu["value"], 1) if cpu else None,
                 "data_gen_s": round(t_gen, 1)},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in kernels.items()},
        "recon": recon,
        "warm_start": warm,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
