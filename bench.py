"""Throughput of the multislice forward + loss + adjoint pass on one MI355X.

One step = one full loss-and-gradient pass (object, probe and position
gradients, probe sub-pixel shift active) over every scan position of the
multislice workload: object 512x512x10 (dz 2 A, 0.2 A pixels), 128 px probe
with 5 incoherent modes, 64x64 raster = 4096 positions, 256 windows per
launch, Poisson data at 1e4 e-.  The loss/sse readback of every step is
inside the timed region.

Prints one JSON line: patterns/s (= positions x steps / seconds), the
roofline of the dominant kernel (FP32 flops per launch / HIP-event average
duration) and the CPU baseline (the oracle's complex64 numpy path on a
bounded sample of positions, timed on this host).
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


def cpu_baseline(ds, n_sample, workers):
    """Time the oracle (PyPty's numpy op sequence, complex64, scipy.fft) on n_sample positions."""
    import numpy as np
    import scipy.fft

    from oracle import ms_oracle as orc

    seq = np.arange(min(n_sample, ds.positions.shape[0]))
    with scipy.fft.set_workers(workers):
        t0 = time.perf_counter()
        orc.loss_and_direction(ds.obj, ds.probe, ds.positions, ds.corrections, ds.measured, seq, dz=ds.dz_A,
                               wavelength=ds.wavelength_A, pixel_size_x=ds.pixel_A, pixel_size_y=ds.pixel_A,
                               algorithm="lsq_sqrt", step_obj=True, step_probe=True, step_pos=True,
                               compute_batch=32, cdtype=np.complex64)
        dt = time.perf_counter() - t0
    return len(seq) / dt, dt


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--scan", type=int, default=64, help="scan grid side (64 -> 4096 positions)")
    ap.add_argument("--cpu-sample", type=int, default=160, help="positions timed on the CPU (0: skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for rocprofv3 counter scaling)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (one-GPU boxes); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pypty_amd import _lib, synth
    from pypty_amd.objective import PassConfig, PtychoPass

    dev = torch.device("cuda", 0)
    t_gen = time.perf_counter()
    ds = synth.config1(dev, n_scan=args.scan)
    t_gen = time.perf_counter() - t_gen
    ny, nx, S, K = ds.obj.shape
    n, _, M = ds.probe.shape
    P = ds.positions.shape[0]
    # reconstruction starts from a flat object; data come from the true one
    obj = torch.ones((ny, nx, S, K), dtype=torch.complex64, device=dev)
    probe = torch.from_numpy(np.ascontiguousarray(ds.probe.transpose(2, 0, 1))).to(dev)
    pos = torch.from_numpy(ds.positions).to(dev)
    corr = torch.from_numpy(ds.corrections).to(dev)
    meas = torch.from_numpy(ds.measured).to(dev)
    seq = torch.arange(P, dtype=torch.int32, device=dev)
    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx, n_positions=P,
                     compute_batch=args.batch, algorithm="lsq_sqrt", planar=False, step_obj=True, step_probe=True,
                     step_pos=True, shift_active=True, slice_distance=ds.dz_A, wavelength=ds.wavelength_A,
                     pixel_size_x=ds.pixel_A, pixel_size_y=ds.pixel_A)
    eng = PtychoPass(cfg, dev)
    og = torch.empty_like(obj)
    pg = torch.empty_like(probe)
    posg = torch.zeros_like(corr)

    def step():
        return eng.run(obj, probe, pos, corr, meas, seq, obj_grad=og, probe_grad=pg, pos_grad=posg)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, sse = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = P * args.steps / elapsed

    kernels = {}
    roofline = None
    if not args.no_kernel_timing:
        # same steps again with HIP events around every launch, for the per-kernel breakdown
        _lib.timing_reset()
        _lib.timing_enable(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        kernels = {k: v for k, v in _lib.timing_read().items() if v[1] > 0}
        _lib.timing_enable(False)
        dom = max(kernels, key=lambda k: kernels[k][0])
        tot_ms, launches = kernels[dom]
        avg_s = tot_ms / launches / 1e3
        per_launch = kernel_flops(dom, n, S, M, K) * args.batch
        achieved = per_launch / avg_s / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None, "kernel": dom,
                    "avg_launch_ms": round(avg_s * 1e3, 4), "windows_per_launch": args.batch}
        # HBM bytes per launch of this kernel at this workload, measured with rocprofv3
        # FETCH_SIZE / WRITE_SIZE (scaled by a plain-copy calibration) and committed under profiles/
        tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_multislice_512x512x10_n128_m5.txt")
        if os.path.exists(tf) and args.batch == 256 and args.scan == 64:
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
        v1, dt1 = cpu_baseline(ds, args.cpu_sample, 1)
        cpu = {"value": round(v1, 3), "unit": "patterns/s", "cores": 1, "kind": "port",
               "sample": f"oracle complex64 numpy/scipy.fft path (PyPty op order), first {min(args.cpu_sample, P)} "
                         f"positions of the same workload, scipy workers=1, {dt1:.1f} s"}

    flops_pat = pass_flops(n, S, M, K, True)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "patterns/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex64", "data": "synthetic",
        "config": {"workload": "multislice object 512x512x10 dz2A, probe 128px x5 modes, 64x64 scan "
                               "(4096 positions), batch 256, lsq_sqrt, object+probe+position gradients",
                   "n": n, "slices": S, "probe_modes": M, "obj_modes": K, "positions": P,
                   "compute_batch": args.batch},
        "roofline": roofline, "cpu_baseline": cpu,
        "pass": {"loss": loss, "sse": sse, "alg_gflop_per_pattern": round(flops_pat / 1e9, 4),
                 "alg_tflops": round(flops_pat * value / 1e12, 3),
                 "hbm_frac_alg": round(pass_bytes(n, S, K) * value / (HBM_PEAK_GBS * 1e9), 4),
                 "data_gen_s": round(t_gen, 1)},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in kernels.items()},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
