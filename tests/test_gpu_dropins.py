"""GPU parity of the drop-in functions a PyPty caller rebinds one by one.

multislice() (pypty/multislice_core.py:9-82), multislice_grads()
(multislice_core.py:85-200), scatteradd() (multislice_core.py:206-230) and
complex_grad_to_phase_grad() (pypty/utils.py:260-276) run the LDS-tile kernels
behind pty_ms_forward / pty_ms_adjoint / pty_scatteradd_c64 /
pty_phase_grad_c64.  Each is compared with the oracle's restatement of the
same reference function on the same seeded inputs: 1e-4 normwise relative
against the complex64 oracle, the FP64 error printed beside it; the scatter-add
of overlapping windows (repeated corners) against numpy.add.at.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import rel

TOL = 1e-4

pytestmark = pytest.mark.gpu


def _inputs(n, S, K, M, B, edge, seed):
    """Per-window probes (B, N, N, M), band-limited windows (B, N, N, S, K), the master propagator
    (objective.py:337-345) and the band mask (utils.py:153-199)."""
    rng = np.random.default_rng(seed)
    q2, qx, qy, mask_s, mask = orc.create_spatial_frequencies(0.2, 0.2, n, 2 / 3)
    H = orc.propagator(2.0, 0.0196868, q2, qx, qy, mask)
    f = np.fft.fftfreq(n)
    fx, fy = np.meshgrid(f, f, indexing="xy")
    ap = (fx ** 2 + fy ** 2) <= 0.2 ** 2
    spec = ap[None, :, :, None] * np.exp(1j * rng.uniform(0, 2 * np.pi, (B, n, n, M)))
    probe_b = (np.fft.ifft2(spec, axes=(1, 2)) * n).astype(np.complex64)
    phase = rng.uniform(0.0, 0.5, (B, n, n, S, K))
    obj_b = (np.exp(1j * phase) * (1 - 0.05 * rng.uniform(size=phase.shape))).astype(np.complex64)
    obj_b = orc.ifft2(orc.fft2(obj_b, axes=(1, 2)) * mask[None, :, :, None, None], axes=(1, 2)).astype(np.complex64)
    return probe_b, obj_b, H, mask, q2, qx, qy, mask_s


def _forward(torch, probe_b, obj_b, H, mask, q2, qx, qy, mask_s, S, K, M, edge):
    from pypty_amd import multislice_core as mc

    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    B, n = probe_b.shape[0], probe_b.shape[1]
    # the caller's preallocated stash and exit wave (objective.py:389-391); the drop-in returns new arrays
    waves_in = torch.zeros((B, n, n, S, M, K, 1), dtype=torch.complex64, device=dev)
    wave_in = torch.zeros((B, n, n, M, K), dtype=torch.complex64, device=dev)
    Hb = t(H)[None, :, :, None, None]
    return mc.multislice(t(probe_b), t(obj_b), S, K, M, np.array([2.0], np.float32), edge, 0.0196868, t(q2), t(qx),
                         t(qy), t(mask_s), True, 0.0, 0.0, 2 / 3, 0.0, Hb, None, t(mask), waves_in, wave_in, 0, 0,
                         np.float32, np.complex64)


@pytest.mark.parametrize("n,S,K,M,B,edge", [(32, 3, 2, 2, 5, 1.0), (64, 4, 1, 1, 3, 2.0), (128, 3, 1, 3, 2, 1.5),
                                            (16, 5, 1, 2, 7, 1.0), (96, 2, 2, 1, 3, 3.0)])
def test_multislice_exit_wave_and_stash(n, S, K, M, B, edge):
    import torch

    print(f"[multislice] n={n} S={S} K={K} M={M} B={B} edge={edge}", flush=True)
    probe_b, obj_b, H, mask, q2, qx, qy, mask_s = _inputs(n, S, K, M, B, edge, seed=n + S)
    waves, exitw = _forward(torch, probe_b, obj_b, H, mask, q2, qx, qy, mask_s, S, K, M, edge)
    assert tuple(waves.shape) == (B, n, n, S, M, K, 1) and tuple(exitw.shape) == (B, n, n, M, K)
    ref_w, ref_e = orc.multislice(probe_b, obj_b, H, edge)
    ref_w64, ref_e64 = orc.multislice(probe_b.astype(np.complex128), obj_b.astype(np.complex128),
                                      H.astype(np.complex128), edge)
    ew, ee = rel(waves[..., 0].cpu().numpy(), ref_w), rel(exitw.cpu().numpy(), ref_e)
    print(f"  stash {ew:.2e} exit {ee:.2e} | FP64: stash {rel(waves[..., 0].cpu().numpy(), ref_w64):.2e} "
          f"exit {rel(exitw.cpu().numpy(), ref_e64):.2e}", flush=True)
    assert ew < TOL and ee < TOL
    # the stash is stored conjugated (multislice_core.py:81): slice 0 holds conj of the incident probe
    np.testing.assert_allclose(waves[:, :, :, 0, :, 0, 0].cpu().numpy(), np.conj(probe_b), rtol=0, atol=1e-6)


@pytest.mark.parametrize("n,S,K,M,B,edge", [(32, 3, 2, 2, 6, 1.0), (64, 4, 1, 3, 4, 2.0), (128, 2, 1, 2, 3, 1.0),
                                            (48, 3, 1, 1, 5, 1.5)])
def test_multislice_grads_object_and_probe_side(n, S, K, M, B, edge):
    import torch

    from pypty_amd import multislice_core as mc

    print(f"[multislice_grads] n={n} S={S} K={K} M={M} B={B} edge={edge}", flush=True)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    probe_b, obj_b, H, mask, q2, qx, qy, mask_s = _inputs(n, S, K, M, B, edge, seed=3 * n + S)
    rng = np.random.default_rng(n)
    g = (rng.standard_normal((B, n, n, M, K)) + 1j * rng.standard_normal((B, n, n, M, K))).astype(np.complex64)
    ny = nx = n + 20
    # corners that overlap, one repeated (the scatter-add must accumulate both)
    corners = np.stack([rng.integers(0, ny - n + 1, B), rng.integers(0, nx - n + 1, B)], axis=1)
    corners[-1] = corners[0]
    ar = np.arange(n)
    my = corners[:, 0, None, None] + ar[None, :, None] + 0 * ar[None, None, :]
    mx = corners[:, 1, None, None] + 0 * ar[None, :, None] + ar[None, None, :]
    og0 = (0.1 * rng.standard_normal((ny, nx, S, K))).astype(np.complex64)

    waves, _ = _forward(torch, probe_b, obj_b, H, mask, q2, qx, qy, mask_s, S, K, M, edge)
    og = t(og0)
    og_out, interm, tg = mc.multislice_grads(t(g), waves, t(obj_b), og, None, True, np.array([2.0], np.float32),
                                             edge, t(mask_s), 0.0196868, t(q2), t(qx), 0.0, t(qy), 0.0, S, K, 0,
                                             t(H)[None, :, :, None, None], 0, 2 / 3, 0.0, 0, B, t(mask), 1, 1, 0,
                                             t(my), t(mx), np.float32, np.complex64, True)
    assert og_out.data_ptr() == og.data_ptr()  # object_grad is accumulated in place, as there
    ref_w, _ = orc.multislice(probe_b, obj_b, H, edge)
    G, ref_interm = orc.multislice_grads(g, ref_w, obj_b, H, mask, edge)
    ref_og = og0.copy()
    orc.scatteradd(ref_og, my, mx, G)
    ref_w64, _ = orc.multislice(probe_b.astype(np.complex128), obj_b.astype(np.complex128), H.astype(np.complex128),
                                edge)
    G64, interm64 = orc.multislice_grads(g.astype(np.complex128), ref_w64, obj_b.astype(np.complex128),
                                         H.astype(np.complex128), mask.astype(np.float64), edge)
    ref_og64 = og0.astype(np.complex128)
    orc.scatteradd(ref_og64, my, mx, G64)
    e_og, e_pi = rel(og.cpu().numpy(), ref_og), rel(interm.cpu().numpy(), ref_interm)
    print(f"  object_grad {e_og:.2e} probe-side {e_pi:.2e} | FP64: {rel(og.cpu().numpy(), ref_og64):.2e} "
          f"{rel(interm.cpu().numpy(), interm64):.2e}", flush=True)
    assert tuple(interm.shape) == (B, n, n, M)
    assert e_og < TOL and e_pi < TOL


@pytest.mark.parametrize("n,S,K,B,ny,nx", [(32, 2, 1, 9, 70, 90), (64, 1, 2, 5, 100, 64), (16, 3, 1, 12, 40, 40)])
def test_scatteradd_overlapping_windows(n, S, K, B, ny, nx):
    import torch

    from pypty_amd import multislice_core as mc

    print(f"[scatteradd] n={n} S={S} K={K} B={B} object {ny}x{nx}", flush=True)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(B)
    corners = np.stack([rng.integers(0, ny - n + 1, B), rng.integers(0, nx - n + 1, B)], axis=1)
    corners[1] = corners[0]
    corners[2] = corners[0]
    ar = np.arange(n)
    my = corners[:, 0, None, None] + ar[None, :, None] + 0 * ar[None, None, :]
    mx = corners[:, 1, None, None] + 0 * ar[None, :, None] + ar[None, None, :]
    chop = (rng.standard_normal((B, n, n, S, K)) + 1j * rng.standard_normal((B, n, n, S, K))).astype(np.complex64)
    full0 = (rng.standard_normal((ny, nx, S, K)) + 1j * rng.standard_normal((ny, nx, S, K))).astype(np.complex64)
    full = torch.from_numpy(full0).to(dev)
    mc.scatteradd(full, torch.from_numpy(my).to(dev), torch.from_numpy(mx).to(dev), torch.from_numpy(chop).to(dev))
    ref = full0.astype(np.complex128)
    np.add.at(ref, (my, mx), chop.astype(np.complex128))
    e = rel(full.cpu().numpy(), ref)
    print(f"  error {e:.2e}", flush=True)
    assert e < 1e-6
    # the sum is conserved
    assert abs(full.cpu().numpy().astype(np.complex128).sum() - full0.astype(np.complex128).sum()
               - chop.astype(np.complex128).sum()) < 1e-3 * np.abs(chop).sum()


def test_scatteradd_rejects_non_window_grids():
    import torch

    from pypty_amd import multislice_core as mc

    print("[scatteradd] non-window index grids", flush=True)
    dev = torch.device("cuda", 0)
    my = torch.zeros((1, 4, 4), dtype=torch.int64, device=dev)
    mx = torch.zeros((1, 4, 4), dtype=torch.int64, device=dev)
    with pytest.raises(NotImplementedError):
        mc.scatteradd(torch.zeros((8, 8, 1, 1), dtype=torch.complex64, device=dev), my, mx,
                      torch.ones((1, 4, 4, 1, 1), dtype=torch.complex64, device=dev))


@pytest.mark.parametrize("shape", [(64, 48, 3, 1), (1000,), (512, 512, 2, 1)])
def test_complex_grad_to_phase_grad(shape):
    import torch

    from pypty_amd import utils as putils

    print(f"[phase_grad] shape={shape}", flush=True)
    rng = np.random.default_rng(len(shape))
    z = (np.exp(1j * rng.uniform(-3, 3, shape)) * rng.uniform(0.5, 1.5, shape)).astype(np.complex64)
    g = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    dev = torch.device("cuda", 0)
    got = putils.complex_grad_to_phase_grad(torch.from_numpy(g).to(dev), torch.from_numpy(z).to(dev)).cpu().numpy()
    ref = orc.complex_grad_to_phase_grad(g, z)
    ref64 = orc.complex_grad_to_phase_grad(g.astype(np.complex128), z.astype(np.complex128))
    e = rel(got, ref)
    print(f"  error {e:.2e} (FP64 {rel(got, ref64):.2e})", flush=True)
    assert got.dtype == np.float32 and e < 1e-6
