"""pty_loss_and_direction_static (the pass without a caller workspace, in a device array inside the
library) against pty_loss_and_direction with a caller workspace and one batch lane: bit-identical
loss and gradients on two calls in a row, the second with a smaller problem (the array keeps the
first call's contents), and a clear error for a problem whose workspace exceeds the array."""
import ctypes

import numpy as np
import pytest

from tests._problems import make_problem

pytestmark = pytest.mark.gpu


def _buffers(torch, pr, dev, workspace=None):
    from pypty_amd import _lib

    ny, nx, S, K = pr["obj"].shape
    n, _, M = pr["probe"].shape
    P = pr["pos"].shape[0]
    t = dict(obj=torch.from_numpy(pr["obj"]).to(dev),
             probe=torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev),
             pos=torch.from_numpy(pr["pos"]).to(dev), corr=torch.from_numpy(pr["corr"]).to(dev),
             meas=torch.from_numpy(pr["meas"]).to(dev), seq=torch.arange(P, dtype=torch.int32, device=dev),
             loss=torch.zeros(2, dtype=torch.float64, device=dev))
    t["og"] = torch.full_like(t["obj"], float("nan"))
    t["pg"] = torch.full_like(t["probe"], float("nan"))
    t["posg"] = torch.zeros_like(t["corr"])
    ptr = lambda x: None if x is None else x.data_ptr()  # noqa: E731
    bufs = _lib.PtyBuffers(ptr(t["obj"]), ptr(t["probe"]), ptr(t["pos"]), ptr(t["corr"]), ptr(t["meas"]),
                           ptr(t["seq"]), P, ptr(t["loss"]), ptr(t["og"]), ptr(t["pg"]), ptr(t["posg"]), None,
                           ptr(workspace), 0 if workspace is None else workspace.numel(), None,
                           _lib.PTY_MEASURED_BY_POSITION, 0, None, None)
    return t, bufs


def _problem(pr, batch):
    from pypty_amd.objective import PassConfig

    ny, nx, S, K = pr["obj"].shape
    n, _, M = pr["probe"].shape
    return PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx,
                      n_positions=pr["pos"].shape[0], compute_batch=batch, step_obj=True, step_probe=True,
                      step_pos=True, shift_active=True, slice_distance=pr["dz"], wavelength=pr["wavelength"],
                      pixel_size_x=pr["px"], pixel_size_y=pr["px"], lanes=1).to_c()


def _compare(torch, pr, batch, dev):
    from pypty_amd import _lib

    lib = _lib.load()
    prob = _problem(pr, batch)
    size = ctypes.c_size_t(0)
    _lib.check(lib.pty_workspace_bytes(ctypes.byref(prob), ctypes.byref(size)))
    ws = torch.empty(int(size.value), dtype=torch.uint8, device=dev)
    stream = _lib.stream_ptr(dev)
    ref, rb = _buffers(torch, pr, dev, ws)
    _lib.check(lib.pty_loss_and_direction(ctypes.byref(prob), ctypes.byref(rb), stream))
    got, gb = _buffers(torch, pr, dev)
    _lib.check(lib.pty_loss_and_direction_static(ctypes.byref(prob), ctypes.byref(gb), stream))
    torch.cuda.synchronize()
    for k in ("loss", "og", "pg", "posg"):
        assert torch.isfinite(got[k]).all(), k
        assert torch.equal(got[k], ref[k]), k
    assert float(got["og"].abs().max()) > 0 and float(got["posg"].abs().max()) > 0


def test_static_workspace_pass_matches_the_caller_workspace_pass():
    import torch

    dev = torch.device("cuda", 0)
    print("[no-alloc] call 1: n=64 S=3 K=1 M=2 P=20 batch 8", flush=True)
    _compare(torch, make_problem(n=64, S=3, K=1, M=2, ny=112, nx=112, P=20, shift=True, seed=901), 8, dev)
    print("[no-alloc] call 2 (smaller): n=32 S=2 K=2 M=1 P=7 batch 4", flush=True)
    _compare(torch, make_problem(n=32, S=2, K=2, M=1, ny=72, nx=72, P=7, shift=True, seed=902), 4, dev)


def test_static_workspace_rejects_a_problem_that_does_not_fit():
    import torch

    from pypty_amd import _lib
    from pypty_amd.objective import PassConfig

    print("[no-alloc] oversized problem", flush=True)
    dev = torch.device("cuda", 0)
    cfg = PassConfig(n=256, n_slices=40, n_obj_modes=1, n_probe_modes=10, obj_ny=400, obj_nx=400, n_positions=64,
                     compute_batch=64, step_obj=True, step_probe=True, step_pos=True, shift_active=True, lanes=1)
    prob = cfg.to_c()
    size = ctypes.c_size_t(0)
    _lib.check(_lib.load().pty_workspace_bytes(ctypes.byref(prob), ctypes.byref(size)))
    assert size.value > _lib.PTY_STATIC_WORKSPACE_BYTES
    x = torch.zeros(16, dtype=torch.float64, device=dev)
    p = x.data_ptr()
    bufs = _lib.PtyBuffers(p, p, p, p, p, p, 64, p, p, p, p, None, None, 0, None, 0, 0, None, None)
    rc = _lib.load().pty_loss_and_direction_static(ctypes.byref(prob), ctypes.byref(bufs), _lib.stream_ptr(dev))
    assert rc == _lib.PTY_ERR_WORKSPACE
