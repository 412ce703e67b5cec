"""The oracle's per-position aberrations (pypty/objective.py:320-326, 459-465, 673-681; the CTF
polynomials of utils.py:1272-1367), pinned by: zero coefficients acting as the aperture mask alone; a defocus-only
set (C10) equal to the free-space propagation of the probe over that distance, as the Fresnel
propagator is exp(-i pi lambda q^2 dz); central finite differences of the FP64 loss in every marker's
coefficients without the sub-pixel shift; and, with the shift, the reference's order of operations
(the shift restarts from the spectrum before the plate, so the loss does not depend on the
aberrations)."""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, rel


def _run(pr, ab, marker, step=True, probe=None, cdtype=np.complex128, step_pos=None, corr=None):
    return orc.loss_and_direction(
        pr["obj"], pr["probe"] if probe is None else probe, pr["pos"], pr["corr"] if corr is None else corr,
        pr["meas"], pr["seq"], dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"],
        pixel_size_y=pr["px"], algorithm=pr["algorithm"], step_obj=True, step_probe=True,
        step_pos=pr["step_pos"] if step_pos is None else step_pos, cdtype=cdtype, compute_batch=3, loss_weight=0.7,
        aberrations=ab, aberration_marker=marker, step_aberrations=step)


def test_nmab_order_and_polynomials():
    n, m, ab = orc.convert_num_to_nmab(7)
    assert list(zip(n, m, ab)) == [(1, 0, ""), (1, 2, "a"), (1, 2, "b"), (2, 1, "a"), (2, 1, "b"), (2, 3, "a"),
                                   (2, 3, "b")]
    kx, ky = np.meshgrid(np.linspace(-0.02, 0.02, 5), np.linspace(-0.03, 0.01, 4))
    A = orc.ctf_matrix(kx, ky, 3, 0.02)
    np.testing.assert_allclose(A[0], 2 * np.pi / 0.02 * (kx ** 2 + ky ** 2) / 2)
    phi = np.arctan2(ky, kx)
    np.testing.assert_allclose(A[2], 2 * np.pi / 0.02 * (kx ** 2 + ky ** 2) / 2 * np.sin(2 * phi))


def test_zero_aberrations_band_limit_the_probe():
    """A plate of zero coefficients is the aperture mask alone (objective.py:324): the same as the
    band-limited probe, whose gradient comes back band-limited."""
    pr = make_problem(n=16, S=2, M=2, ny=40, nx=40, P=5, seed=601, shift=False, step_pos=False)
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], 16, 2 / 3, 0.0, np.float64)
    bl = lambda a: orc.ifft2(orc.fft2(a.astype(np.complex128), axes=(0, 1)) * mask[:, :, None], axes=(0, 1))  # noqa
    base = orc.loss_and_direction(pr["obj"], bl(pr["probe"]), pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                  dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"],
                                  pixel_size_y=pr["px"], step_pos=False, cdtype=np.complex128, compute_batch=3,
                                  loss_weight=0.7)
    out = _run(pr, np.zeros((2, 4)), np.array([0, 1, 0, 1, 1]))
    assert abs(out[0] - base[0]) <= 1e-12 * abs(base[0])
    assert rel(out[2], base[2]) < 1e-12 and rel(out[3], bl(base[3])) < 1e-12


def test_defocus_is_free_propagation_of_the_probe():
    pr = make_problem(n=32, S=1, M=1, ny=64, nx=64, P=3, seed=603, shift=False, step_pos=False)
    dz = 40.0
    # C10 = dz: the plate exp(-i A_0 C10), A_0 = pi lambda q^2, equals the propagator over dz
    out = _run(pr, np.array([[dz]]), np.zeros(3, int), step=False)
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], 32, 2 / 3, 0.0, np.float64)
    H = orc.propagator(dz, pr["wavelength"], q2, qx, qy, mask, cdtype=np.complex128)
    moved = orc.ifft2(orc.fft2(pr["probe"].astype(np.complex128)) * H[:, :, None])
    ref = orc.loss_and_direction(pr["obj"], moved, pr["pos"], pr["corr"], pr["meas"], pr["seq"], dz=pr["dz"],
                                 wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                                 step_pos=False, cdtype=np.complex128, compute_batch=3, loss_weight=0.7)
    assert abs(out[0] - ref[0]) <= 1e-9 * abs(ref[0])


@pytest.mark.parametrize("M", [1, 2])
def test_aberration_gradient_central_fd_without_shift(M):
    pr = make_problem(n=16, S=2, M=M, ny=40, nx=40, P=6, seed=610 + M, shift=False, step_pos=False)
    rng = np.random.default_rng(5)
    ab = rng.uniform(-1, 1, (3, 5)) * np.array([20.0, 10.0, 10.0, 300.0, 300.0])
    marker = np.array([0, 2, 1, 0, 2, 2])
    g = _run(pr, ab, marker)[5]
    assert g.shape == ab.shape and np.abs(g).max() > 0
    for k in range(3):
        for i in range(5):
            eps = 1e-4 * max(abs(ab[k, i]), 1.0)
            up, dn = ab.copy(), ab.copy()
            up[k, i] += eps
            dn[k, i] -= eps
            fd = (_run(pr, up, marker, step=False)[0] - _run(pr, dn, marker, step=False)[0]) / (2 * eps)
            assert abs(g[k, i] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (k, i, g[k, i], fd)


def test_with_the_shift_the_plates_do_not_reach_the_loss():
    pr = make_problem(n=16, S=2, M=2, ny=40, nx=40, P=5, seed=621, shift=True)
    marker = np.array([0, 1, 0, 1, 1])
    a = _run(pr, np.zeros((2, 3)), marker)
    b = _run(pr, np.array([[30.0, 5.0, -5.0], [-20.0, 2.0, 1.0]]), marker)
    assert a[0] == b[0] and rel(b[4], a[4]) < 1e-12  # same loss and position gradient
    assert np.abs(b[5]).max() > 0  # the gradient is still formed (objective.py:673-681)
