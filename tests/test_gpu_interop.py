"""The drop-ins called the way PyPty's host code calls them: with numpy arrays (its CPU path) and with
another library's device arrays (CuPy's role, played by a DLPack-only array), returning the caller's
array type with the results of the torch-tensor call.

loss_and_direction receives exactly the arguments of pypty/iterative.py:524 (numpy object, probe,
positions, corrections, tilts, distances, measured patterns ...); multislice and multislice_grads
those of pypty/objective.py:486 and :639; the FFT wrappers and fourier_clean those of fft.py and
utils.py.  The measured patterns are copied to the device once across calls.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests import fakedev
from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu


def _kw(pr, obj, probe, **over):
    return reference_kwargs(
        this_obj=obj, full_probe=probe, this_pos_array=pr["pos"], this_pos_correction=pr["corr"],
        this_distances=np.array([pr["dz"]], np.float32), measured_array=pr["meas"],
        algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"], this_step_probe=1, this_step_obj=1,
        this_step_pos_correction=1, this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"],
        this_chopped_sequence=pr["seq"], **over)


def test_loss_and_direction_with_the_reference_numpy_arguments():
    import torch

    from pypty_amd import _arrays, objective

    print("[interop] loss_and_direction with numpy arrays (iterative.py:524)", flush=True)
    pr = make_problem(n=64, S=3, K=1, M=2, ny=112, nx=112, P=10, shift=True, seed=301)
    dev = torch.device("cuda", 0)
    ref = call_loss_and_direction(objective.loss_and_direction,
                                  **_kw(pr, torch.from_numpy(pr["obj"]).to(dev), torch.from_numpy(pr["probe"]).to(dev)))
    _arrays._staged.clear()
    got = call_loss_and_direction(objective.loss_and_direction, **_kw(pr, pr["obj"].copy(), pr["probe"].copy()))
    assert isinstance(got[2], np.ndarray) and isinstance(got[3], np.ndarray) and isinstance(got[4], np.ndarray)
    assert got[2].dtype == np.complex64 and got[2].shape == pr["obj"].shape
    assert got[0] == ref[0] and got[1] == ref[1]
    for a, b in zip(got[2:5], ref[2:5]):
        np.testing.assert_array_equal(a, b.cpu().numpy())
    # the measured patterns were copied to the device once; a second call reuses that copy
    assert len(_arrays._staged) == 1
    staged = next(iter(_arrays._staged.values()))[2]
    again = call_loss_and_direction(objective.loss_and_direction, **_kw(pr, pr["obj"].copy(), pr["probe"].copy()))
    assert next(iter(_arrays._staged.values()))[2] is staged and again[0] == got[0]
    # an in-place change of the host patterns is seen
    meas = pr["meas"]
    meas *= 2
    changed = call_loss_and_direction(objective.loss_and_direction, **_kw(pr, pr["obj"].copy(), pr["probe"].copy()))
    assert changed[0] != got[0]


def test_loss_and_direction_with_foreign_device_arrays():
    import torch

    from pypty_amd import objective

    print("[interop] loss_and_direction with DLPack device arrays", flush=True)
    pr = make_problem(n=32, S=2, K=1, M=2, ny=72, nx=72, P=8, shift=True, seed=302)
    dev = torch.device("cuda", 0)
    ref = call_loss_and_direction(objective.loss_and_direction,
                                  **_kw(pr, torch.from_numpy(pr["obj"]).to(dev), torch.from_numpy(pr["probe"]).to(dev)))
    wrap = lambda a: fakedev.DevArray(torch.from_numpy(np.ascontiguousarray(a)).to(dev))  # noqa: E731
    kw = _kw(pr, wrap(pr["obj"]), wrap(pr["probe"]))
    kw.update(this_pos_array=wrap(pr["pos"]), this_pos_correction=wrap(pr["corr"]), measured_array=wrap(pr["meas"]))
    got = call_loss_and_direction(objective.loss_and_direction, **kw)
    assert all(isinstance(got[i], fakedev.DevArray) for i in (2, 3, 4))
    assert got[0] == ref[0]
    for a, b in zip(got[2:5], ref[2:5]):
        assert torch.equal(a.torch(), b)


@pytest.mark.parametrize("mode", ["fft2", "ifft2", "shift_fft2", "ifft2_ishift"])
def test_fft_wrappers_with_numpy_and_foreign_arrays(mode):
    import torch

    from pypty_amd import fft as pfft

    print(f"[interop] {mode} numpy / DLPack", flush=True)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((4, 64, 64, 2)) + 1j * rng.standard_normal((4, 64, 64, 2))).astype(np.complex64)
    ref = getattr(orc, mode)(x.astype(np.complex128), axes=(1, 2))
    got = getattr(pfft, mode)(x, axes=(1, 2))
    assert isinstance(got, np.ndarray) and rel(got, ref) < 2e-6
    dx = fakedev.DevArray(torch.from_numpy(x).cuda())
    gd = getattr(pfft, mode)(dx, axes=(1, 2))
    assert isinstance(gd, fakedev.DevArray) and rel(gd.torch().cpu().numpy(), ref) < 2e-6
    # plane sizes outside the window kernel (the whole-plane kernels)
    y = (rng.standard_normal((90, 150)) + 1j * rng.standard_normal((90, 150))).astype(np.complex64)
    assert rel(getattr(pfft, mode)(y), getattr(orc, mode)(y.astype(np.complex128))) < 3e-6


def test_multislice_and_grads_with_the_reference_numpy_arguments():
    import torch

    from pypty_amd import multislice_core as mc

    print("[interop] multislice / multislice_grads with numpy arrays (objective.py:486, :639)", flush=True)
    n, S, K, M, B = 64, 3, 1, 2, 4
    rng = np.random.default_rng(6)
    q2, qx, qy, mask_s, mask = orc.create_spatial_frequencies(0.2, 0.2, n, 2 / 3)
    H = orc.propagator(2.0, 0.0196868, q2, qx, qy, mask)
    probe = (rng.standard_normal((B, n, n, M)) + 1j * rng.standard_normal((B, n, n, M))).astype(np.complex64)
    obj = np.exp(1j * rng.uniform(0, 0.5, (B, n, n, S, K))).astype(np.complex64)
    Hb = H[None, :, :, None, None]
    waves_in = np.zeros((B, n, n, S, M, K, 1), np.complex64)
    wave_in = np.zeros((B, n, n, M, K), np.complex64)
    waves, exitw = mc.multislice(probe, obj, S, K, M, np.array([2.0], np.float32), 1.0, 0.0196868, q2, qx, qy,
                                 mask_s, True, 0.0, 0.0, 2 / 3, 0.0, Hb, None, mask, waves_in, wave_in, 0, 0,
                                 np.float32, np.complex64)
    assert isinstance(waves, np.ndarray) and isinstance(exitw, np.ndarray)
    ref_w, ref_e = orc.multislice(probe, obj, H)
    assert rel(waves[..., 0], ref_w) < 1e-4 and rel(exitw, ref_e) < 1e-4
    g = (rng.standard_normal((B, n, n, M, K)) + 1j * rng.standard_normal((B, n, n, M, K))).astype(np.complex64)
    corners = rng.integers(0, 20, (B, 2))
    ar = np.arange(n)
    my = corners[:, 0, None, None] + ar[None, :, None] + 0 * ar[None, None, :]
    mx = corners[:, 1, None, None] + 0 * ar[None, :, None] + ar[None, None, :]
    og = np.zeros((n + 20, n + 20, S, K), np.complex64)
    og_out, interm, _ = mc.multislice_grads(g, waves, obj, og, None, True, np.array([2.0], np.float32), 1.0, mask_s,
                                            0.0196868, q2, qx, 0.0, qy, 0.0, S, K, 0, Hb, 0, 2 / 3, 0.0, 0, B, mask,
                                            1, 1, 0, my, mx, np.float32, np.complex64, True)
    assert og_out is og  # accumulated in place into the caller's array, as there
    assert isinstance(interm, np.ndarray)
    G, ref_i = orc.multislice_grads(g, ref_w, obj, H, mask)
    ref_og = np.zeros_like(og)
    orc.scatteradd(ref_og, my, mx, G)
    assert rel(og, ref_og) < 1e-4 and rel(interm, ref_i) < 1e-4


def test_fourier_clean_and_phase_grad_with_numpy():
    from pypty_amd import utils as putils

    print("[interop] fourier_clean / complex_grad_to_phase_grad with numpy", flush=True)
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((200, 160, 2, 1)) + 1j * rng.standard_normal((200, 160, 2, 1))).astype(np.complex64)
    got = putils.fourier_clean(x, cutoff=0.6)
    assert isinstance(got, np.ndarray) and rel(got, orc.fourier_clean(x.astype(np.complex128), 0.6)) < 3e-6
    z = np.exp(1j * rng.uniform(-3, 3, 500)).astype(np.complex64)
    gr = (rng.standard_normal(500) + 1j * rng.standard_normal(500)).astype(np.complex64)
    pg = putils.complex_grad_to_phase_grad(gr, z)
    assert isinstance(pg, np.ndarray) and rel(pg, orc.complex_grad_to_phase_grad(gr, z)) < 1e-6
