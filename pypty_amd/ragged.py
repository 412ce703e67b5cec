"""Ragged multi-dataset steps: several independent datasets of mixed shapes in one replayed step.

The reference reconstructs one dataset per run and evaluates its loss with
``loss_and_direction`` (pypty/objective.py:47-967).  A ragged step holds
several datasets resident on the device at once (different windows, slice
counts, probe modes and scans), each with its own pass engine, gradient
buffers and autotuned windows per launch, spreads them over a few
default-priority streams and captures one whole step (every dataset's pass)
in a HIP graph, so a replay submits all of a step's launches at once.

``DatasetPass`` is one dataset on the device; ``RaggedStep`` is the set of
them with the stream spread, the autotuner and the graph.
"""
from __future__ import annotations

import numpy as np
import torch

from .objective import PassConfig, PtychoPass


class DatasetPass:
    """One dataset resident on the device with its pass engine and gradient buffers.

    ``ds`` carries obj (Ny, Nx, S, K), probe (N, N, M), positions (P, 2) int32, corrections (P, 2)
    float32, measured (P, N, N) float32 and pixel_A / dz_A / wavelength_A (``synth.Dataset``).  The
    reconstruction state starts from a flat object (``start_obj`` overrides it); object, probe and
    position gradients are computed with the sub-pixel shift active."""

    def __init__(self, ds, batch, device, checkpoint=0, lanes=1, start_obj=None, algorithm="lsq_sqrt"):
        self.ds = ds
        self.dev = torch.device(device)
        self.checkpoint = int(checkpoint)
        self.lanes = int(lanes)
        self.algorithm = algorithm
        ny, nx, S, K = ds.obj.shape
        n, _, M = ds.probe.shape
        self.P = int(ds.positions.shape[0])
        self.shape = (n, S, M, K)
        init = np.ones((ny, nx, S, K), np.complex64) if start_obj is None else np.asarray(start_obj, np.complex64)
        self.obj = torch.from_numpy(np.ascontiguousarray(init)).to(self.dev)
        self.probe = torch.from_numpy(np.ascontiguousarray(ds.probe.transpose(2, 0, 1))).to(self.dev)
        self.pos = torch.from_numpy(np.ascontiguousarray(ds.positions, dtype=np.int32)).to(self.dev)
        self.corr = torch.from_numpy(np.ascontiguousarray(ds.corrections, dtype=np.float32)).to(self.dev)
        self.meas = torch.from_numpy(np.ascontiguousarray(ds.measured, dtype=np.float32)).to(self.dev)
        self.seq = torch.arange(self.P, dtype=torch.int32, device=self.dev)
        self.og = torch.empty_like(self.obj)
        self.pg = torch.empty_like(self.probe)
        self.posg = torch.zeros_like(self.corr)
        self._configure(int(batch))

    def _configure(self, batch):
        n, S, M, K = self.shape
        ny, nx = self.obj.shape[:2]
        ds = self.ds
        self.cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx,
                              n_positions=self.P, compute_batch=batch, algorithm=self.algorithm, planar=False,
                              step_obj=True, step_probe=True, step_pos=True, shift_active=True,
                              slice_distance=ds.dz_A, wavelength=ds.wavelength_A, pixel_size_x=ds.pixel_A,
                              pixel_size_y=ds.pixel_A, checkpoint_slices=self.checkpoint, lanes=self.lanes)
        self.engine = PtychoPass(self.cfg, self.dev)

    def set_batch(self, batch):
        """A new engine with ``batch`` windows per launch (same buffers)."""
        self._configure(int(batch))

    def set_lanes(self, lanes):
        self.lanes = int(lanes)
        self._configure(self.cfg.compute_batch)

    @property
    def loss_sse(self) -> torch.Tensor:
        """The engine's device float64 [loss, sse] of its last pass."""
        return self.engine.loss_sse

    def run(self, sync=False, sequence=None):
        """One full pass over the dataset's positions (or ``sequence``) on the current stream."""
        seq = self.seq if sequence is None else sequence
        return self.engine.run(self.obj, self.probe, self.pos, self.corr, self.meas, seq, obj_grad=self.og,
                               probe_grad=self.pg, pos_grad=self.posg, sync=sync)


class RaggedStep:
    """One step = one full pass over every dataset, datasets spread round robin over ``streams``
    default-priority streams that fork from and join the caller's stream.

    ``tune(candidates)`` picks each dataset's windows per launch by timing
    (``autotune.choose_batch``); ``capture()`` records one step in a HIP graph that ``replay()``
    submits; ``losses()`` is the (datasets, 2) float64 [loss, sse] of the last step on the device."""

    def __init__(self, datasets, device, batch=256, streams=2, lanes=1, checkpoint=0):
        self.dev = torch.device(device)
        self.jobs = [DatasetPass(ds, batch, self.dev, checkpoint, lanes) for ds in datasets]
        self.n_streams = max(1, min(len(self.jobs), int(streams)))
        self.streams = [torch.cuda.Stream(self.dev) for _ in range(self.n_streams)] if self.n_streams > 1 else []
        self.graph = None
        self.tuned = None

    @property
    def positions(self) -> int:
        return sum(j.P for j in self.jobs)

    def tune(self, candidates=(64, 128, 256, 512)):
        """Choose every dataset's windows per launch by timing one pass per candidate."""
        from . import autotune

        self.tuned = []
        for j in self.jobs:
            n, S, M, K = j.shape
            key = autotune.shape_key(self.dev, n, S, K, M, j.P)

            def make_pass(b, j=j):
                j.set_batch(b)
                return lambda: j.run(sync=False)

            best = autotune.choose_batch(key, candidates, make_pass)
            j.set_batch(best)
            self.tuned.append(best)
        # the candidates' workspaces; the chosen engines' runs take theirs
        PtychoPass.release_workspaces()
        return self.tuned

    def run(self, concurrent=True):
        """Launch one step (not synchronised)."""
        if not self.streams or not concurrent:
            for j in self.jobs:
                j.run(sync=False)
            return
        main = torch.cuda.current_stream(self.dev)
        for s in self.streams:
            s.wait_stream(main)
        for i, j in enumerate(self.jobs):
            with torch.cuda.stream(self.streams[i % self.n_streams]):
                j.run(sync=False)
        for s in self.streams:
            main.wait_stream(s)

    def losses(self) -> torch.Tensor:
        return torch.stack([j.loss_sse for j in self.jobs])

    def capture(self):
        """Record one step in a HIP graph (the datasets on their streams, whose workspaces the eager
        steps already hold)."""
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self.run()
        torch.cuda.current_stream(self.dev).wait_stream(side)
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        return self.graph

    def replay(self):
        if self.graph is None:
            raise RuntimeError("pypty_amd.ragged: capture() first")
        self.graph.replay()
