"""The oracle's constraint terms pinned by central finite differences and closed forms.

Every gradient is checked against central differences of its term: for a real argument
dR = grad . d, for a complex one dR = 2 Re sum(conj(grad) d) (the Wirtinger convention of
pypty/utils.py:260-276), along random directions.  Closed forms pin the terms themselves on small
inputs (a straight scan line has no second differences, an affine grid no deformation, a constant
object no variation, a single object mode no variance).
"""
import numpy as np
import pytest

from oracle import constraints_oracle as co


def _fd_real(f, x, d, eps=1e-6):
    return (f(x + eps * d)[0] - f(x - eps * d)[0]) / (2 * eps)


def _fd_complex(f, z, d, eps=1e-6):
    return (f(z + eps * d)[0] - f(z - eps * d)[0]) / (2 * eps)


def _check_real(f, x, seed=0, rtol=1e-6):
    rng = np.random.default_rng(seed)
    _, g = f(x)
    for _ in range(3):
        d = rng.standard_normal(x.shape)
        fd, an = _fd_real(f, x, d), float(np.sum(g * d))
        assert abs(fd - an) <= rtol * max(abs(an), 1e-8), (fd, an)


def _check_complex(f, z, seed=0, rtol=1e-6):
    rng = np.random.default_rng(seed)
    _, g = f(z)
    for _ in range(3):
        d = rng.standard_normal(z.shape) + 1j * rng.standard_normal(z.shape)
        fd, an = _fd_complex(f, z, d), 2 * float(np.real(np.sum(np.conj(g) * d)))
        assert abs(fd - an) <= rtol * max(abs(an), 1e-8), (fd, an)


@pytest.mark.parametrize("fn", [co.fast_axis_constraint_on_grid, co.slow_axis_constraint_on_grid,
                                co.deformation_constraint_on_grid])
def test_grid_constraints_gradients(fn):
    rng = np.random.default_rng(1)
    pos = rng.standard_normal((6 * 7, 2)) * 3
    _check_real(lambda x: fn(x, (6, 7), 0.7), pos)


def test_grid_constraints_closed_forms():
    y, x = np.meshgrid(np.arange(5.0), np.arange(8.0), indexing="ij")
    affine = np.stack([1.5 * y + 0.2 * x + 3, -0.3 * y + 2.0 * x - 1], axis=-1).reshape(-1, 2)
    assert co.fast_axis_constraint_on_grid(affine, (5, 8), 1.0)[0] < 1e-20
    assert co.slow_axis_constraint_on_grid(affine, (5, 8), 1.0)[0] < 1e-18
    assert co.deformation_constraint_on_grid(affine, (5, 8), 1.0)[0] < 1e-20
    # one kink of height h in the middle of a fast line: second differences -h, 2h, -h (all interior)
    bent = affine.copy().reshape(5, 8, 2)
    bent[2, 4, 1] += 1.0
    assert co.fast_axis_constraint_on_grid(bent.reshape(-1, 2), (5, 8), 1.0)[0] == pytest.approx(6.0)


def test_window_and_probe_constraints_gradients():
    rng = np.random.default_rng(2)
    P = rng.standard_normal((24, 24, 2)) + 1j * rng.standard_normal((24, 24, 2))
    win = rng.uniform(0, 1, (24, 24))
    _check_complex(lambda z: co.window_constraint(z, win, 0.3), P)
    _check_complex(lambda z: co.probe_constraint(z, 0.5, 0.2), P)
    ap = rng.uniform(size=(24, 24)) > 0.5
    _check_complex(lambda z: co.probe_constraint(z, ap, 0.2), P)
    # closed forms: a window of ones is the probe's power; a probe inside the aperture costs nothing
    assert co.window_constraint(P, np.ones((24, 24)), 2.0)[0] == pytest.approx(2.0 * np.sum(np.abs(P) ** 2))
    f = np.fft.fftshift(np.fft.fftfreq(24))
    fx, fy = np.meshgrid(f, f, indexing="xy")
    spec = ((fx ** 2 + fy ** 2) ** 0.5 <= 0.2)[:, :, None] * (1 + 0j)
    inside = np.fft.ifft2(np.fft.ifftshift(spec, axes=(0, 1)), axes=(0, 1))
    assert co.probe_constraint(inside, 0.5, 1.0)[0] < 1e-12


def test_probe_modulation_constraint_gradient():
    rng = np.random.default_rng(3)
    P = rng.standard_normal((32, 32, 2)) + 1j * rng.standard_normal((32, 32, 2))
    _check_complex(lambda z: co.probe_modulation_constraint(z, 1.2, 0.2, 0.2, 0.3, 0.05, 2 / 3), P, rtol=1e-5)


@pytest.mark.parametrize("q,p", [(1, 2), (1, 1), (2, 2), (0.8, 1.5)])
def test_atv_constraint_gradient(q, p):
    rng = np.random.default_rng(4)
    O = np.exp(1j * rng.uniform(-1, 1, (20, 18, 2, 1))) * rng.uniform(0.8, 1.2, (20, 18, 2, 1))
    _check_complex(lambda z: co.atv_constraint(z, 0.4, q, p), O, rtol=1e-5)
    # a constant object has no variation: only the 1e-20 floor of each pixel's term remains
    assert co.atv_constraint(np.full((8, 8, 1, 1), 0.7 + 0.2j), 1.0, q, p)[0] == pytest.approx(64 * 1e-20 ** (q / p))


def test_missing_wedge_and_mixed_variance_gradients():
    rng = np.random.default_rng(5)
    O = np.exp(1j * rng.uniform(-1, 1, (16, 12, 5, 2)))
    _check_complex(lambda z: co.missing_wedge_constraint(z, 0.2, 0.25, 1.5, 2.0, 0.3), O)
    _check_complex(lambda z: co.mixed_object_variance_constraint(z, 0.7, 0.3), O)
    # one object mode has no variance; identical modes neither
    assert co.mixed_object_variance_constraint(O[..., :1], 1.0, 0.3)[0] == 0.0
    same = np.repeat(O[..., :1], 3, axis=-1)
    assert co.mixed_object_variance_constraint(same, 1.0, 0.3)[0] < 1e-20
