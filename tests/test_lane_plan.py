"""The launch plan of a pass, checked on the host (pty_plan_check, no GPU).

The library records the steps pty_loss_and_direction would issue (every
launch with its arguments, every event record and wait, on which lane) and
checks them: each buffer range a launch reads or writes lies in its workspace
buffer (per-batch buffers in the launch's own lane block), launches on
different lanes that touch one range with a write are ordered by the events,
and the batches cover the sequence once.  Cases: the batch-lane case of an
earlier intermittent object-gradient mismatch (32 px, 5 slices, 2 object and
2 probe modes, 9 positions in batches of 4, so the last batch holds one
window) with 2-4 lanes, with and without checkpoints, one and two calls in a
row; every tile family; the data-format options; the ragged workload's
shapes at each autotuner candidate.  Planted faults (waits dropped from the
recorded plan, a stash one window short) must be reported.
"""
import numpy as np
import pytest

from pypty_amd import _lib
from pypty_amd.objective import PassConfig


def _cfg(n, S, K, M, P, batch, ck=0, lanes=2, **kw):
    args = dict(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=n + 48, obj_nx=n + 48, n_positions=P,
                compute_batch=batch, step_obj=True, step_probe=True, step_pos=True, shift_active=True,
                checkpoint_slices=ck, lanes=lanes)
    args.update(kw)
    return PassConfig(**args)


def _check(cfg, n_seq=None, lanes=None, calls=1, accumulate=False):
    n, rep = _lib.plan_check(cfg.to_c(), n_seq or cfg.n_positions, lanes or cfg.lanes, calls, accumulate)
    assert n == 0, rep
    return rep


@pytest.mark.parametrize("lanes", [1, 2, 3, 4])
@pytest.mark.parametrize("ck", [0, 2])
@pytest.mark.parametrize("calls", [1, 2])
def test_batch_lane_mismatch_case_is_ordered_and_in_bounds(lanes, ck, calls):
    cfg = _cfg(32, 5, 2, 2, 9, 4, ck=ck, lanes=lanes)
    rep = _check(cfg, calls=calls)
    assert f"{min(lanes, 3)} lanes, 3 batches" in rep


@pytest.mark.parametrize("n,S,K,M,P,batch,ck", [
    (16, 3, 1, 2, 9, 2, 0), (32, 3, 2, 2, 11, 4, 0), (48, 3, 1, 2, 8, 3, 0), (64, 2, 1, 2, 5, 8, 0),
    (96, 4, 1, 2, 9, 2, 0), (128, 3, 1, 3, 8, 3, 0), (128, 5, 1, 2, 7, 2, 2), (192, 3, 1, 2, 5, 2, 0),
    (192, 16, 1, 5, 7, 2, 4), (256, 2, 1, 1, 4, 2, 0), (256, 40, 1, 10, 5, 2, 8), (128, 10, 1, 5, 1000, 256, 0),
    (64, 7, 2, 3, 13, 4, 3),
])
@pytest.mark.parametrize("lanes", [2, 4])
def test_every_tile_family(n, S, K, M, P, batch, ck, lanes):
    _check(_cfg(n, S, K, M, P, batch, ck=ck, lanes=lanes), calls=2)


@pytest.mark.parametrize("extra", [
    dict(algorithm="lsq_compressed", n_masks=5), dict(algorithm="ml", data_shift_y=2, data_bin=2, raw_n=64,
                                                      data_pad=0, data_multiplier=2.0),
    dict(step_static_background=True), dict(planar=True), dict(algorithm="lsq_sqrt_2"),
    dict(step_beam_current=True), dict(step_beam_current=True, shift_active=False, step_pos=False),
    dict(n_aberrations=7, n_markers=3, step_aberrations=True),
    dict(n_aberrations=12, n_markers=10, step_aberrations=True, shift_active=False, step_pos=False,
         step_beam_current=True),
    dict(n_aberrations=3, n_markers=1, step_aberrations=False, shift_active=False, step_pos=False),
])
def test_data_format_options(extra):
    n = 32 if extra.get("raw_n") != 64 else 32
    _check(_cfg(n, 3, 1, 2, 10, 3, lanes=3, **extra), calls=2)


@pytest.mark.parametrize("n", [128, 192, 256])
def test_aberrations_on_the_large_tiles(n):
    _check(_cfg(n, 3, 1, 2, 9, 2, lanes=4, n_aberrations=5, n_markers=4, step_aberrations=True, shift_active=False,
                step_pos=False), calls=2)


def test_streamed_accumulating_calls():
    # a streamed pass: one batch per call, accumulating into the gradients of the previous call
    _check(_cfg(64, 3, 1, 2, 40, 8, lanes=1), n_seq=8, calls=3, accumulate=True)


def test_ragged_workload_shapes_at_every_candidate():
    for i in range(8):
        rng = np.random.default_rng(2000 + i)
        n = int(rng.choice([96, 128, 192]))
        S = int(rng.integers(4, 17))
        M = int(rng.integers(1, 6))
        scan = int(rng.integers(32, 97))
        for batch in (64, 128, 256, 512):
            for lanes in (1, 2):
                cfg = PassConfig(n=n, n_slices=S, n_obj_modes=1, n_probe_modes=M, obj_ny=6 * scan + n + 8,
                                 obj_nx=6 * scan + n + 8, n_positions=scan * scan, compute_batch=batch,
                                 step_obj=True, step_probe=True, step_pos=True, shift_active=True, lanes=lanes)
                _check(cfg)


@pytest.mark.parametrize("omit,what", [(1, "RACE"), (2, "RACE"), (4, "RACE"), (8, "OUT OF BUFFER")])
def test_planted_faults_are_reported(omit, what):
    # two batches, the last on lane 1: its additions reach lane 0 only through the join
    cfg = _cfg(32, 5, 2, 2, 8, 4, ck=2, lanes=2)
    n, rep = _lib.plan_check(cfg.to_c(), 8, 2, 2, False, omit=omit)
    assert n > 0 and what in rep, rep[-2000:]


@pytest.mark.parametrize("n,S,ck", [(32, 4, 0), (64, 5, 2), (192, 3, 0), (256, 2, 0)])
@pytest.mark.parametrize("tilt", [
    dict(tilt_mode=0, n_tilts=10, step_tilts=True), dict(tilt_mode=3, n_tilts=10, step_tilts=True),
    dict(tilt_mode=1, n_tilts=1, step_tilts=True), dict(tilt_mode=0, n_tilts=1, step_tilts=False),
    dict(tilt_mode=3, n_tilts=10, step_tilts=True, n_distances=4),
    dict(tilt_mode=4, n_tilts=10, step_tilts=True, step_beam_current=True),
    dict(tilt_mode=2, n_tilts=1, step_tilts=True, shift_active=False, step_pos=False),
])
def test_tilts_and_per_slice_distances(n, S, ck, tilt):
    """The tilt buffers (per-lane spectra, exit waves and partials; the tilt gradient, added into by
    every batch with one tilt row) and the per-slice propagator tables."""
    extra = dict(tilt)
    if extra.get("n_distances"):
        extra["n_distances"] = max(S - 1, 2)
    _check(_cfg(n, S, 1, 2, 10, 3, ck=ck, lanes=3, **extra), calls=2)

