"""The oracle's static background (pypty/objective.py:359-361, 525-526, 588-589), pinned by:
a zero background changing nothing; central finite differences of the FP64 loss at pixels inside
the band (the reference's gradient carries the 1/(M*K) of the mode mean, so with several modes it
is the derivative divided by M*K); and, outside the band, where the model pattern is exactly zero
and the background is masked out of the loss, the closed form that the data alone fix."""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, rel


def _run(pr, bg, step=True, algorithm=None, cdtype=np.complex128, patterns=False):
    return orc.loss_and_direction(
        pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"], dz=pr["dz"],
        wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
        algorithm=algorithm or pr["algorithm"], step_obj=True, step_probe=True, step_pos=pr["step_pos"],
        cdtype=cdtype, compute_batch=3, loss_weight=0.7, static_background=bg, step_static_background=step,
        return_patterns=patterns)


def _background(pr, n, seed):
    rng = np.random.default_rng(seed)
    return (0.3 * np.sqrt(pr["meas"].mean()) * rng.uniform(0.5, 1.5, (n, n))).astype(np.float64)


def _band(n):
    _, _, _, mask_s, _ = orc.create_spatial_frequencies(0.2, 0.2, n, 2 / 3, 0.0, np.float64)
    return mask_s > 0


def test_zero_background_changes_nothing():
    n = 16
    pr = make_problem(n=n, S=2, M=2, ny=40, nx=40, P=5, seed=301)
    base = _run(pr, None, patterns=True)
    with_bg = _run(pr, np.zeros((n, n)), patterns=True)
    assert with_bg[0] == base[0] and with_bg[1] == base[1]
    for a, b in zip(base[2:6], with_bg[2:6]):
        assert np.array_equal(a, b)
    assert not np.any(with_bg[6])


def test_simulated_patterns_leave_the_background_out():
    n = 16
    pr = make_problem(n=n, S=2, M=1, ny=40, nx=40, P=4, seed=305)
    base = _run(pr, None, patterns=True)
    with_bg = _run(pr, _background(pr, n, 1), patterns=True)
    assert np.array_equal(base[5], with_bg[5])
    assert with_bg[0] != base[0]


@pytest.mark.parametrize("algorithm", ["lsq_sqrt", "ml", "lsq", "lsq_sqrt_2"])
@pytest.mark.parametrize("M", [1, 2])
def test_background_gradient_central_fd_inside_the_band(algorithm, M):
    n = 16
    pr = make_problem(n=n, S=2, M=M, ny=40, nx=40, P=4, seed=311 + M, algorithm=algorithm)
    bg = _background(pr, n, 7)
    grad = _run(pr, bg)[5]
    band = np.argwhere(_band(n))
    rng = np.random.default_rng(3)
    for y, x in band[rng.choice(len(band), 5, replace=False)]:
        eps = 1e-4 * bg[y, x]
        up, dn = bg.copy(), bg.copy()
        up[y, x] += eps
        dn[y, x] -= eps
        fd = (_run(pr, up, step=False)[0] - _run(pr, dn, step=False)[0]) / (2 * eps)
        assert abs(grad[y, x] * M - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(grad).max() * M), (y, x, grad[y, x], fd)


@pytest.mark.parametrize("algorithm", ["lsq_sqrt", "lsq"])
def test_background_gradient_outside_the_band_closed_form(algorithm):
    n = 16
    pr = make_problem(n=n, S=2, M=1, ny=40, nx=40, P=5, seed=321, algorithm=algorithm)
    bg = _background(pr, n, 9)
    out = _run(pr, bg)
    grad = out[5]
    outside = ~_band(n)
    assert outside.sum() > 0
    meas = pr["meas"].astype(np.float64)
    # the model pattern is zero there: dL/dI = -sqrt(D) (lsq_sqrt, with the reference's 0 -> 1
    # divisor) or -2 D (lsq)
    dI = -np.sqrt(meas) if algorithm == "lsq_sqrt" else -2 * meas
    want = 2 * bg * 0.7 * dI.sum(0)
    assert rel(grad[outside], want[outside]) < 1e-12
    # and the loss does not depend on the background there
    bg2 = bg.copy()
    bg2[outside] *= 3
    assert _run(pr, bg2, step=False)[0] == out[0]


def test_static_background_start_from_a_scalar():
    """pypty/utils.py:720-731: a scalar total becomes 0.1-1.0 uniform noise scaled to that total, then
    zeroed beyond half the cutoff in DC-centred normalised frequency; arrays and None pass through."""
    from pypty_amd.iterative import _static_background_from_nothing

    assert _static_background_from_nothing(None, 16, 2 / 3) is None
    a = np.ones((16, 16), np.float32)
    assert _static_background_from_nothing(a, 16, 2 / 3) is a
    np.random.seed(0)
    bg = _static_background_from_nothing(5.0, 16, 2 / 3)
    f = np.fft.fftshift(np.fft.fftfreq(16))
    r = np.hypot(f[None, :], f[:, None])
    inside = r <= (2 / 3) * 0.5
    assert bg.dtype == np.float32 and bg.shape == (16, 16)
    assert not np.any(bg[~inside]) and np.all(bg[inside] > 0)
    assert bg[inside].max() / bg[inside].min() <= 10.0 + 1e-5
    assert bg.sum() < 5.0
    np.random.seed(0)
    raw = 0.1 + 0.9 * np.random.rand(16, 16)
    assert np.allclose(bg[inside], (raw * 5.0 / raw.sum())[inside], rtol=1e-6)
    assert _static_background_from_nothing(0.0, 16, 2 / 3).sum() <= 1e-7 * (1 + 1e-6)


def test_oracle_lbfgs_refines_the_background_towards_the_true_one():
    """Data made of the model patterns of the current object plus a known background: refining the
    background alone (iterative.py:480, 569-1030 as the oracle restates them) lowers the loss every
    epoch and moves the background inside the band towards the true one."""
    from oracle import bfgs_oracle

    n = 16
    pr = make_problem(n=n, S=2, M=1, ny=40, nx=40, P=6, seed=331)
    I = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], np.zeros((6, n, n)), pr["seq"],
                               dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                               step_obj=False, step_probe=False, cdtype=np.complex128, return_patterns=True)[5]
    band = _band(n)
    bg_true = _background(pr, n, 11) * band
    meas = I + (bg_true ** 2)[None]
    state = {"obj": pr["obj"].astype(np.complex128), "probe": pr["probe"].astype(np.complex128),
             "positions": pr["pos"], "corr": pr["corr"].astype(np.float64), "bg": 0.5 * bg_true}
    data = {"measured": meas, "sequence": pr["seq"], "dz": pr["dz"], "wavelength": pr["wavelength"], "px": pr["px"],
            "compute_batch": 4, "algorithm": "lsq"}
    hist = bfgs_oracle.new_history()
    err0 = rel(state["bg"][band], bg_true[band])
    losses, step = [], 1.0
    np.random.seed(1)
    for _ in range(15):
        loss, _, step, _, _, _, ok = bfgs_oracle.bfgs_epoch(state, hist, data, update_obj=False, update_probe=False,
                                                            update_bg=True, actual_step=step, cdtype=np.complex128)
        assert ok
        losses.append(loss)
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert rel(state["bg"][band], bg_true[band]) < 0.05 * err0
    assert not np.any(state["bg"][~band])


def test_l1_constraint_closed_form_and_central_fd():
    """pypty/objective.py:1286-1320: on a constant object r e^{i phi} the terms are n (w_p |phi| +
    w_a |log r|); for a generic object the gradient g satisfies dL = 2 Re sum conj(g) dO (the
    convention of every object gradient here), checked by central differences; weights <= 0 add
    nothing."""
    n = 200
    o = np.full(n, 0.8 * np.exp(-0.7j), np.complex128)
    reg, _ = orc.full_l1_constraint(o, 0.3, 0.2)
    assert abs(reg - n * (0.2 * 0.7 + 0.3 * abs(np.log(0.8)))) < 1e-12 * n
    assert orc.full_l1_constraint(o, -1.0, 0.0)[0] == 0.0 and not np.any(orc.full_l1_constraint(o, 0.0, -2.0)[1])
    rng = np.random.default_rng(12)
    o = (rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(-2.5, 2.5, n))).astype(np.complex128)
    reg, g = orc.full_l1_constraint(o, 0.3, 0.2)
    for _ in range(4):
        d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        eps = 1e-7
        fd = (orc.full_l1_constraint(o + eps * d, 0.3, 0.2)[0] - orc.full_l1_constraint(o - eps * d, 0.3, 0.2)[0]) / (2 * eps)
        assert abs(fd - 2 * np.real(np.sum(np.conj(g) * d))) < 1e-6 * abs(fd)
