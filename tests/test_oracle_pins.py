"""Pins of oracle conventions that a GPU-vs-oracle comparison alone cannot fix.

* fourier_clean's mask radius (pypty/utils.py:129-137): a plane wave survives
  the band limit iff its normalised frequency (ky/Ny, kx/Nx) lies within
  0.5*cutoff, checked on non-square planes so that a transposed axis fails.
* the direction of the sub-pixel probe shift (objective.py:473): an integer
  correction d must equal np.roll of the band-limited probe by +d.
* the edge-slice multiplier (multislice_core.py:73-74 forward, :172-174
  adjoint): the first propagation, and only it, is raised to the power e.
* complex_grad_to_phase_grad (utils.py:260-276): central differences of a
  loss in the phase.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, rel


@pytest.mark.parametrize("shape,cutoff", [((64, 64), 2 / 3 - 0.005), ((48, 80), 0.5), ((90, 36), 0.8)])
def test_fourier_clean_passes_plane_waves_inside_the_radius_only(shape, cutoff):
    ny, nx = shape
    y, x = np.mgrid[0:ny, 0:nx]
    r = 0.5 * cutoff
    inside = outside = 0
    for ky in range(-ny // 2, ny // 2):
        for kx in range(-nx // 2, nx // 2, 3):
            f2 = (ky / ny) ** 2 + (kx / nx) ** 2
            if abs(f2 - r * r) < 1e-9:
                continue  # exactly on the circle: float rounding decides
            w = np.exp(2j * np.pi * (ky * y / ny + kx * x / nx))
            out = orc.fourier_clean(w[:, :, None], cutoff=cutoff, fdtype=np.float64)[:, :, 0]
            if f2 < r * r:
                assert rel(out, w) < 1e-12, (ky, kx)
                inside += 1
            else:
                assert np.abs(out).max() < 1e-12, (ky, kx)
                outside += 1
    assert inside > 20 and outside > 20


@pytest.mark.parametrize("d", [(2, -3), (-1, 4), (0, 1)])
def test_integer_probe_shift_is_a_roll_by_plus_d(d):
    """With a band-limited probe, the loss and gradients with corrections d equal those with no
    correction and the probe rolled by +d (the shift theorem with exp(-2 pi i f d))."""
    pr = make_problem(n=32, S=2, K=1, M=2, ny=72, nx=72, P=5, seed=21, shift=False)
    n = 32
    _, _, _, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], n, 2 / 3, 0.0, np.float64)
    probe = orc.ifft2(orc.fft2(pr["probe"].astype(np.complex128)) * mask[:, :, None])
    corr = np.tile(np.asarray(d, np.float64), (5, 1))
    common = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                  step_obj=True, step_probe=False, step_pos=False, cdtype=np.complex128)
    obj = pr["obj"].astype(np.complex128)
    a = orc.loss_and_direction(obj, probe, pr["pos"], corr, pr["meas"], pr["seq"], **common)
    rolled = np.roll(probe, d, axis=(0, 1))
    b = orc.loss_and_direction(obj, rolled, pr["pos"], np.zeros_like(corr), pr["meas"], pr["seq"], **common)
    # the shift kernel's 2 pi is the float32 literal 6.2831855 (objective.py:305-310), off by 3e-8
    assert abs(a[0] - b[0]) < 1e-6 * abs(b[0])
    assert rel(a[2], b[2]) < 1e-6
    # the opposite direction is a different probe: the check would catch a sign error
    c = orc.loss_and_direction(obj, np.roll(probe, (-d[0], -d[1]), axis=(0, 1)), pr["pos"], np.zeros_like(corr),
                               pr["meas"], pr["seq"], **common)
    assert abs(a[0] - c[0]) > 1e-3 * abs(c[0])


def _edge_setup(S=3, n=32, B=2, M=2):
    rng = np.random.default_rng(41)
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(0.2, 0.2, n, 2 / 3, 0.0, np.float64)
    dz, lam = 2.0, 0.0197

    def H(d):
        return orc.propagator(d, lam, q2, qx, qy, mask, cdtype=np.complex128)

    probe = orc.ifft2(orc.fft2(rng.standard_normal((B, n, n, M)) + 1j * rng.standard_normal((B, n, n, M)),
                               axes=(1, 2)) * mask[None, :, :, None], axes=(1, 2))
    return rng, H, dz, probe, mask

def _prop(wave, h):
    return orc.ifft2(orc.fft2(wave, axes=(1, 2)) * h[None, :, :, None], axes=(1, 2))


@pytest.mark.parametrize("e", [2, 3])
def test_edge_multiplier_free_space_total_distance(e):
    """O = 1: the exit wave is free propagation over (S - 2 + e) dz (integer e, where H^e is exact)."""
    S = 4
    rng, H, dz, probe, mask = _edge_setup(S)
    ones = np.ones(probe.shape[:3] + (S, 1))
    _, exitw = orc.multislice(probe, ones, H(dz), edge=e)
    assert rel(exitw[..., 0], _prop(probe, H((S - 2 + e) * dz))) < 1e-12


def test_edge_multiplier_acts_on_the_first_propagation_only():
    """S = 3 with only slice 1 scattering: exit = H(O1 H^e P) (edge at the first propagation),
    not H^e(O1 H P)."""
    S, e = 3, 2
    rng, H, dz, probe, mask = _edge_setup(S)
    o1 = np.exp(1j * rng.uniform(0, 1.0, probe.shape[:3]))
    obj = np.ones(probe.shape[:3] + (S, 1), np.complex128)
    obj[:, :, :, 1, 0] = o1
    _, exitw = orc.multislice(probe, obj, H(dz), edge=e)
    first = _prop(o1[..., None] * _prop(probe, H(e * dz)), H(dz))
    last = _prop(o1[..., None] * _prop(probe, H(dz)), H(e * dz))
    assert rel(exitw[..., 0], first) < 1e-12
    assert rel(exitw[..., 0], last) > 1e-3


@pytest.mark.parametrize("e", [2.0, 1.5])
def test_edge_multiplier_adjoint_dot_product(e):
    """Re<A x, y> = Re<x, A^H y> for the probe -> exit map with an edge multiplier: the adjoint
    raises conj(H) to e on its propagation back to slice 0 (multislice_core.py:172-174)."""
    S = 4
    rng, H, dz, probe, mask = _edge_setup(S)
    obj = np.exp(1j * rng.uniform(0, 1, probe.shape[:3] + (S, 2)))
    y = rng.standard_normal(probe.shape + (2,)) + 1j * rng.standard_normal(probe.shape + (2,))
    waves, Ax = orc.multislice(probe, obj, H(dz), edge=e)
    _, AHy = orc.multislice_grads(y, waves, obj, H(dz), mask, edge=e, step_obj=False)
    lhs, rhs = np.sum(Ax * np.conj(y)), np.sum(probe * np.conj(AHy))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_edge_multiplier_object_gradient_central_fd():
    pr = make_problem(n=16, S=3, K=1, M=2, ny=40, nx=40, P=5, seed=43)
    obj = pr["obj"].astype(np.complex128)
    probe = pr["probe"].astype(np.complex128)
    kw = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
              step_pos=False, cdtype=np.complex128, edge=2.0)
    g = orc.loss_and_direction(obj, probe, pr["pos"], pr["corr"], pr["meas"], pr["seq"], step_obj=True,
                               step_probe=False, **kw)[2]
    rng = np.random.default_rng(4)
    eps = 1e-6
    for _ in range(4):
        idx = (int(rng.integers(14, 26)), int(rng.integers(14, 26)), int(rng.integers(3)), 0)
        for d in (1.0, 1j):
            def L(s):
                o = obj.copy()
                o[idx] += s * d
                return orc.loss_and_direction(o, probe, pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                              step_obj=False, step_probe=False, **kw)[0]
            fd = (L(eps) - L(-eps)) / (2 * eps)
            an = 2 * np.real(np.conj(g[idx]) * d)
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (idx, d, fd, an)


def test_complex_grad_to_phase_grad_central_fd():
    rng = np.random.default_rng(5)
    a = rng.uniform(0.5, 1.5, 12)
    p = rng.uniform(-3, 3, 12)
    t = rng.standard_normal(12) + 1j * rng.standard_normal(12)

    def L(ph):
        return float(np.sum(np.abs(a * np.exp(1j * ph) - t) ** 2))

    z = a * np.exp(1j * p)
    g = z - t  # dL/dz* of sum |z - t|^2
    an = orc.complex_grad_to_phase_grad(g, z)
    eps = 1e-6
    for j in range(12):
        e = np.zeros(12)
        e[j] = eps
        fd = (L(p + e) - L(p - e)) / (2 * eps)
        assert abs(fd - an[j]) < 1e-6 * max(1.0, abs(fd))
