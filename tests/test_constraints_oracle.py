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


def test_charge_flip_by_hand():
    """objective.py:1014-1110 on four pixels worked out by hand: the largest phase is 1.0, so phases
    below 0.3 (0.1 and -0.2) turn into -0.095 and 0.19; the largest absorption is -log 0.5, so only the
    two unit amplitudes (absorption ~0) fall below 0.14 of it, and stay ~0 when flipped."""
    from oracle import constraints_oracle as con

    a = np.array([np.exp(0.1j), 0.9 * np.exp(0.5j), 0.5 * np.exp(1.0j), np.exp(-0.2j)]).reshape(2, 2, 1, 1)
    out = con.charge_flip(a, 0.3, 0.14, -0.95, -0.95)
    expect = np.array([np.exp(-0.095j), 0.9 * np.exp(0.5j), 0.5 * np.exp(1.0j), np.exp(0.19j)]).reshape(2, 2, 1, 1)
    assert np.abs(out - expect).max() < 1e-9
    # beta = 1 leaves the object as it is; every plane has its own maxima
    b = np.random.default_rng(1).uniform(0.5, 1.0, (8, 8, 3, 2)) * np.exp(1j * np.random.default_rng(2).uniform(
        -3, 3, (8, 8, 3, 2)))
    assert np.abs(con.charge_flip(b, 0.3, 0.3, 1.0, 1.0) - b).max() < 1e-9
    c = b.copy()
    c[:, :, 1, 1] *= np.exp(0.5j)  # another plane's phases do not move this plane's threshold
    assert np.abs(con.charge_flip(c, 0.2, 0.2, -0.9, -0.9)[:, :, 0, 0] -
                  con.charge_flip(b, 0.2, 0.2, -0.9, -0.9)[:, :, 0, 0]).max() == 0


def test_gram_schmidt_modes_are_orthogonal_and_span_the_same_space():
    """objective.py:1119-1140: the modes come out mutually orthogonal, mode 0 unchanged (no
    normalisation), and mode i differs from the input's by a combination of input modes 0..i-1."""
    from oracle import constraints_oracle as con

    rng = np.random.default_rng(3)
    p = rng.standard_normal((16, 16, 4)) + 1j * rng.standard_normal((16, 16, 4))
    p[:, :, 1] += 0.8 * p[:, :, 0]  # strongly overlapping modes
    q = con.make_states_orthogonal(p)
    assert np.array_equal(q[:, :, 0], p[:, :, 0])
    for i in range(4):
        for j in range(i):
            ip = abs(np.vdot(q[:, :, j], q[:, :, i]))
            assert ip < 1e-10 * np.linalg.norm(q[:, :, j]) * np.linalg.norm(q[:, :, i]), (i, j, ip)
    for i in range(1, 4):
        basis = p[:, :, :i].reshape(-1, i)
        d = (q[:, :, i] - p[:, :, i]).ravel()
        coef, *_ = np.linalg.lstsq(basis, d, rcond=None)
        assert np.linalg.norm(basis @ coef - d) < 1e-10 * np.linalg.norm(p[:, :, i])
