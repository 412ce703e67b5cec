"""The oracle's beam current (pypty/objective.py:446-450, 688-698), pinned by:
a current of ones changing nothing; the scaling identity (a current c everywhere is the probe c P:
same loss, object and position gradients, probe gradient c times that at c P); central finite
differences of the FP64 loss in each position's current, with and without the sub-pixel shift,
one and two probe modes; central differences of the probe gradient along a random direction with
a ragged current; and the reference's sign(|bc|) factor (a negative current has the gradient of
its magnitude, a zero current none)."""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, rel


def _run(pr, bc, step=True, probe=None, algorithm=None, cdtype=np.complex128, step_pos=None):
    return orc.loss_and_direction(
        pr["obj"], pr["probe"] if probe is None else probe, pr["pos"], pr["corr"], pr["meas"], pr["seq"],
        dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
        algorithm=algorithm or pr["algorithm"], step_obj=True, step_probe=True,
        step_pos=pr["step_pos"] if step_pos is None else step_pos, cdtype=cdtype, compute_batch=3,
        loss_weight=0.7, beam_current=bc, step_beam_current=step)


def _current(P, seed):
    return np.random.default_rng(seed).uniform(0.7, 1.3, P)


def test_unit_current_changes_nothing():
    pr = make_problem(n=16, S=2, M=2, ny=40, nx=40, P=5, seed=401)
    base = _run(pr, None)
    ones = _run(pr, np.ones(5))
    assert ones[0] == base[0] and ones[1] == base[1]
    for a, b in zip(base[2:5], ones[2:5]):
        assert np.array_equal(a, b)
    assert ones[5].shape == (5,)


@pytest.mark.parametrize("shift", [False, True])
def test_uniform_current_is_a_scaled_probe(shift):
    pr = make_problem(n=16, S=2, M=2, ny=40, nx=40, P=6, seed=403, shift=shift)
    c = 1.37
    scaled = _run(pr, np.full(6, c))
    plain = _run(pr, None, probe=pr["probe"].astype(np.complex128) * c)
    assert abs(scaled[0] - plain[0]) <= 1e-12 * abs(plain[0])
    assert rel(scaled[2], plain[2]) < 1e-12
    assert rel(scaled[3], c * plain[3]) < 1e-12
    if shift:
        assert rel(scaled[4], plain[4]) < 1e-10


@pytest.mark.parametrize("algorithm", ["lsq_sqrt", "ml"])
@pytest.mark.parametrize("M,shift", [(1, True), (2, True), (2, False)])
def test_current_gradient_central_fd(algorithm, M, shift):
    P = 5
    pr = make_problem(n=16, S=2, M=M, ny=40, nx=40, P=P, seed=410 + M + 2 * shift, shift=shift,
                      algorithm=algorithm)
    bc = _current(P, 3)
    grad = _run(pr, bc)[5]
    for i in range(P):
        eps = 1e-5
        up, dn = bc.copy(), bc.copy()
        up[i] += eps
        dn[i] -= eps
        fd = (_run(pr, up, step=False)[0] - _run(pr, dn, step=False)[0]) / (2 * eps)
        assert abs(grad[i] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max()), (i, grad[i], fd)


def test_probe_gradient_with_a_current_central_fd():
    pr = make_problem(n=16, S=2, M=2, ny=40, nx=40, P=5, seed=421)
    bc = _current(5, 4)
    g = _run(pr, bc)[3]
    p0 = pr["probe"].astype(np.complex128)
    rng = np.random.default_rng(5)
    # directions inside the band: the reference band-limits the probe gradient
    _, _, _, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], 16, 2 / 3, 0.0, np.float64)
    for _ in range(2):
        d = rng.standard_normal(p0.shape) + 1j * rng.standard_normal(p0.shape)
        d = orc.ifft2(orc.fft2(d) * mask[:, :, None])
        eps = 1e-6 * np.abs(p0).max() / np.abs(d).max()
        fd = (_run(pr, bc, probe=p0 + eps * d)[0] - _run(pr, bc, probe=p0 - eps * d)[0]) / (2 * eps)
        an = 2 * float(np.real(np.sum(np.conj(g) * d)))
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_sign_convention_and_rows_outside_the_sequence():
    pr = make_problem(n=16, S=2, M=1, ny=40, nx=40, P=6, seed=431)
    bc = _current(6, 6)
    pos = _run(pr, bc)
    neg_bc = bc.copy()
    neg_bc[[1, 4]] *= -1
    neg = _run(pr, neg_bc)
    # the probe takes |bc|, and the reference's factor sign(|bc|) is 1: the same loss and gradient
    assert neg[0] == pos[0]
    assert np.array_equal(neg[5], pos[5])
    zero_bc = bc.copy()
    zero_bc[2] = 0.0
    assert _run(pr, zero_bc)[5][2] == 0.0
    # positions outside the sequence keep a zero gradient; without the step every row is zero
    pr_part = dict(pr, seq=np.array([0, 3, 5]))
    part = _run(pr_part, bc)[5]
    assert np.all(part[[1, 2, 4]] == 0) and np.all(part[[0, 3, 5]] != 0)
    assert not np.any(_run(pr, bc, step=False)[5])
