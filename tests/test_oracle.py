"""Pinning of the CPU oracle against mathematics (the reference ships no tests or golden vectors).

Each check is chosen so that a plausible mistake in the oracle (dropped
term, wrong sign or index, transposed operand, sum instead of mean) fails.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, oracle_pass, rel


# ------------------------------------------------------------------ FFTs
@pytest.mark.parametrize("n", [6, 12, 16])
def test_fft_wrappers_match_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, 2)) + 1j * rng.standard_normal((n, n, 2))
    F = orc.dft2_bruteforce(x)
    assert rel(orc.fft2(x), F) < 1e-12
    assert rel(orc.ifft2(F), x) < 1e-12
    assert rel(orc.shift_fft2(x), np.fft.fftshift(F, axes=(0, 1))) < 1e-12
    assert rel(orc.ifft2_ishift(np.fft.fftshift(F, axes=(0, 1))), x) < 1e-12


def test_fft_parseval_and_shift_theorem():
    rng = np.random.default_rng(1)
    n = 32
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    F = orc.fft2(x)
    assert np.isclose(np.sum(np.abs(F) ** 2), n * n * np.sum(np.abs(x) ** 2))
    # circular shift by (3, 5) multiplies the spectrum by exp(-2 pi i (ky*3 + kx*5)/n)
    k = np.fft.fftfreq(n) * n
    ph = np.exp(-2j * np.pi * (k[:, None] * 3 + k[None, :] * 5) / n)
    assert rel(orc.fft2(np.roll(x, (3, 5), axis=(0, 1))), F * ph) < 1e-12


# -------------------------------------------------------- grids, masks
def test_spatial_frequencies_closed_form():
    n, px = 64, 0.2
    q2, qx, qy, mask_s, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    assert qx[0, 1] == pytest.approx(1 / (n * px))
    assert qy[1, 0] == pytest.approx(1 / (n * px))
    assert qx[0, -1] == pytest.approx(-1 / (n * px))
    rmax = (31 / (n * px)) * (2 / 3)
    assert np.array_equal(mask, (q2 <= rmax ** 2).astype(np.float64))
    assert mask_s[n // 2, n // 2] == 1 and mask[0, 0] == 1
    assert np.array_equal(mask_s, np.fft.fftshift(mask))


def test_propagator_composes_and_is_unimodular_in_band():
    n, px = 32, 0.2
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    h1 = orc.propagator(1.5, 0.02, q2, qx, qy, mask, cdtype=np.complex128)
    h2 = orc.propagator(2.5, 0.02, q2, qx, qy, mask, cdtype=np.complex128)
    h3 = orc.propagator(4.0, 0.02, q2, qx, qy, mask, cdtype=np.complex128)
    assert rel(h1 * h2, h3) < 1e-12
    assert np.allclose(np.abs(h3)[mask > 0], 1.0)
    assert np.all(h3[mask == 0] == 0)
    # phase sign: exp(-i pi dz lambda q^2)
    i = np.unravel_index(np.argmax(q2 * mask), q2.shape)
    assert np.isclose(np.angle(h1[i]), np.angle(np.exp(-1j * np.pi * 1.5 * 0.02 * q2[i])))


# ------------------------------------------------------------ multislice
def _ms_setup(n=32, S=4, M=2, K=1, B=2, seed=3):
    rng = np.random.default_rng(seed)
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(0.2, 0.2, n, 2 / 3, 0.0, np.float64)
    H = orc.propagator(2.0, 0.0197, q2, qx, qy, mask, cdtype=np.complex128)
    probe = rng.standard_normal((B, n, n, M)) + 1j * rng.standard_normal((B, n, n, M))
    obj = np.exp(1j * rng.uniform(0, 1, (B, n, n, S, K))) * (1 - 0.1 * rng.uniform(size=(B, n, n, S, K)))
    return q2, qx, qy, mask, H, probe, obj


def test_multislice_free_space_is_propagation():
    q2, qx, qy, mask, H, probe, _ = _ms_setup()
    S = 4
    ones = np.ones(probe.shape[:3] + (S, 1))
    waves, exitw = orc.multislice(probe, ones, H)
    Htot = orc.propagator(2.0 * (S - 1), 0.0197, q2, qx, qy, mask, cdtype=np.complex128)
    expect = orc.ifft2(orc.fft2(probe, axes=(1, 2)) * Htot[None, :, :, None], axes=(1, 2))
    assert rel(exitw[..., 0], expect) < 1e-12
    # stash holds the conjugated incident waves; slice 0 is the probe itself
    assert rel(np.conj(waves[:, :, :, 0, :, 0]), probe) < 1e-14


def test_multislice_constant_phase_slices():
    q2, qx, qy, mask, H, probe, _ = _ms_setup()
    S, phi = 4, 0.3
    const = np.full(probe.shape[:3] + (S, 1), np.exp(1j * phi))
    _, e_const = orc.multislice(probe, const, H)
    _, e_free = orc.multislice(probe, np.ones_like(const), H)
    assert rel(e_const, np.exp(1j * S * phi) * e_free) < 1e-12


def test_multislice_adjoint_dot_product():
    """Re<A x, y> = Re<x, A^H y> for the probe -> exit map with fixed slices."""
    q2, qx, qy, mask, H, probe, obj = _ms_setup(K=2)
    rng = np.random.default_rng(9)
    y = rng.standard_normal(probe.shape + (2,)) + 1j * rng.standard_normal(probe.shape + (2,))
    waves, Ax = orc.multislice(probe, obj, H)
    _, AHy = orc.multislice_grads(y, waves, obj, H, mask, step_obj=False)
    lhs = np.sum(Ax * np.conj(y))
    rhs = np.sum(probe * np.conj(AHy))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_scatteradd_bruteforce_and_conservation():
    rng = np.random.default_rng(4)
    full = rng.standard_normal((20, 22, 2, 1)) + 0j
    chop = rng.standard_normal((5, 6, 6, 2, 1)) + 1j * rng.standard_normal((5, 6, 6, 2, 1))
    corners = np.array([[0, 0], [3, 4], [3, 4], [14, 16], [7, 1]])
    run = np.arange(6)
    my = corners[:, 0, None, None] + run[None, :, None] + 0 * run[None, None, :]
    mx = corners[:, 1, None, None] + run[None, None, :] + 0 * run[None, :, None]
    expect = full.copy()
    for b in range(5):
        expect[corners[b, 0]:corners[b, 0] + 6, corners[b, 1]:corners[b, 1] + 6] += chop[b]
    got = full.copy()
    orc.scatteradd(got, my, mx, chop)
    assert rel(got, expect) < 1e-14
    assert np.isclose(got.sum(), full.sum() + chop.sum())


# ------------------------------------------------------------------ loss
@pytest.mark.parametrize("algorithm", ["lsq_sqrt", "ml", "lsq", "lsq_sqrt_2"])
def test_loss_closed_form_flat_probe(algorithm):
    """Single slice, O = 1, flat probe modes c_m: I is (1/M) sum |c_m N^2|^2 at DC, 0 elsewhere."""
    n, M = 16, 2
    c = np.array([1.0, 2.0])
    probe = np.ones((n, n, M), np.complex128) * c[None, None, :]
    obj = np.ones((n, n, 1, 1), np.complex128)
    meas = np.ones((1, n, n))
    out = orc.loss_and_direction(obj, probe, np.zeros((1, 2), int), np.zeros((1, 2)), meas, [0], dz=1.0,
                                 wavelength=0.02, pixel_size_x=0.2, pixel_size_y=0.2, algorithm=algorithm,
                                 step_obj=False, step_probe=False, cdtype=np.complex128, return_patterns=True)
    I0 = np.mean((c * n * n) ** 2)
    I = out[5][0]
    assert np.isclose(I[n // 2, n // 2], I0)
    assert np.allclose(np.delete(I.ravel(), (n // 2) * n + n // 2), 0.0, atol=1e-9 * I0)
    if algorithm == "lsq_sqrt":
        L = (np.sqrt(I0) - 1) ** 2 + (n * n - 1)
    elif algorithm == "ml":
        L = I0 - np.log(I0)
    elif algorithm == "lsq":
        L = (I0 - 1) ** 2 + (n * n - 1)
    else:  # lsq_sqrt_2 with D = 1 > 0.5 everywhere
        L = 0.5 * (1 - I0) ** 2 + 0.5 * (n * n - 1)
    assert np.isclose(out[0], L, rtol=1e-10)
    assert np.isclose(out[1], (1 - I0) ** 2 + (n * n - 1), rtol=1e-10)


@pytest.mark.parametrize("algorithm", ["lsq_compressed", "gauss_compressed", "poisson_compressed"])
def test_compressed_loss_closed_form_flat_probe(algorithm):
    """Flat probe modes as above (I = I0 at DC only) with three detectors: the full field (c = I0), the
    DC pixel (c = I0) and everything else (c = 0); measured coefficients (1, 2, 3)."""
    n, M = 16, 2
    c = np.array([1.0, 2.0])
    probe = np.ones((n, n, M), np.complex128) * c[None, None, :]
    obj = np.ones((n, n, 1, 1), np.complex128)
    masks = np.zeros((3, n, n))
    masks[0] = 1.0
    masks[1, n // 2, n // 2] = 1.0
    masks[2] = 1.0 - masks[1]
    meas = np.array([[1.0, 2.0, 3.0]])
    out = orc.loss_and_direction(obj, probe, np.zeros((1, 2), int), np.zeros((1, 2)), meas, [0], dz=1.0,
                                 wavelength=0.02, pixel_size_x=0.2, pixel_size_y=0.2, algorithm=algorithm,
                                 step_obj=False, step_probe=False, cdtype=np.complex128, masks=masks)
    I0 = np.mean((c * n * n) ** 2)
    sse = (I0 - 1) ** 2 + (I0 - 2) ** 2 + 9.0
    if algorithm == "lsq_compressed":
        L = sse
    elif algorithm == "gauss_compressed":
        L = (np.sqrt(I0) - 1) ** 2 + (np.sqrt(I0) - np.sqrt(2)) ** 2 + 3.0
    else:  # the zero coefficient is replaced by 1 (objective.py:582)
        L = (I0 - np.log(I0)) + (I0 - 2 * np.log(I0)) + 1.0
    assert np.isclose(out[0], L, rtol=1e-10), (out[0], L)
    assert np.isclose(out[1], sse, rtol=1e-10)


def test_preprocess_dataset_known_values():
    """Hand-computed cases of utils.py:626-651: the roll keeps the wrapped rows and zeroes the band on
    the other side (as the reference writes it), binning sums blocks, padding adds zero borders."""
    x = np.arange(16, dtype=float).reshape(1, 4, 4)
    y = orc.preprocess_dataset(x, (1, 0))
    assert np.array_equal(y[0], np.array([x[0, 3], x[0, 0], x[0, 1], np.zeros(4)]))
    y = orc.preprocess_dataset(x, (0, -1))
    assert np.array_equal(y[0, :, 0], np.zeros(4)) and np.array_equal(y[0, :, 1:], np.roll(x[0], -1, 1)[:, 1:])
    y = orc.preprocess_dataset(x, data_bin=2)
    assert np.array_equal(y[0], np.array([[0 + 1 + 4 + 5, 2 + 3 + 6 + 7], [8 + 9 + 12 + 13, 10 + 11 + 14 + 15]]))
    y = orc.preprocess_dataset(x, data_pad=2, data_multiplier=0.5)
    assert y.shape == (1, 8, 8) and y.sum() == 0.5 * x.sum() and np.array_equal(y[0, 2:6, 2:6], 0.5 * x[0])
    # shift then bin then pad compose in that order
    y = orc.preprocess_dataset(x, (1, 0), data_bin=2, data_pad=1)
    assert y.shape == (1, 4, 4) and y[0, 1, 1] == 12 + 13 + 0 + 1 and y[0, 2, 1] == 4 + 5 + 0 + 0


# -------------------------------------------------------- finite differences
def _fd_loss(pr, what, idx, delta, eps=1e-4):
    def L(shift):
        q = dict(pr)
        q[what] = pr[what].copy()
        q[what][idx] = q[what][idx] + shift
        return oracle_pass(q, np.complex128, step_obj=False, step_probe=False)[0]

    return (L(eps * delta) - L(-eps * delta)) / (2 * eps)


@pytest.mark.parametrize("algorithm", ["lsq_sqrt", "ml", "lsq", "lsq_sqrt_2", "lsq_compressed", "gauss_compressed",
                                       "poisson_compressed"])
def test_object_gradient_central_fd(algorithm):
    pr = make_problem(n=16, S=2, K=2, M=2, ny=40, nx=40, P=6, seed=11, algorithm=algorithm)
    pr["obj"] = pr["obj"].astype(np.complex128)
    pr["probe"] = pr["probe"].astype(np.complex128)
    g = oracle_pass(pr, np.complex128)[2]
    rng = np.random.default_rng(2)
    for _ in range(4):
        idx = tuple(int(rng.integers(12, 28)) for _ in range(2)) + (int(rng.integers(2)), int(rng.integers(2)))
        for d in (1.0, 1j):
            fd = _fd_loss(pr, "obj", idx, d)
            an = 2 * np.real(np.conj(g[idx]) * d)
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (idx, d, fd, an)


def test_probe_gradient_central_fd_with_shift():
    pr = make_problem(n=16, S=2, K=1, M=2, ny=40, nx=40, P=6, seed=12, shift=True)
    pr["obj"] = pr["obj"].astype(np.complex128)
    pr["probe"] = pr["probe"].astype(np.complex128)
    g = oracle_pass(pr, np.complex128)[3]
    rng = np.random.default_rng(5)
    for _ in range(4):
        idx = (int(rng.integers(16)), int(rng.integers(16)), int(rng.integers(2)))
        for d in (1.0, 1j):
            fd = _fd_loss(pr, "probe", idx, d)
            an = 2 * np.real(np.conj(g[idx]) * d)
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (idx, d, fd, an)


def test_position_gradient_central_fd():
    pr = make_problem(n=16, S=2, K=1, M=2, ny=40, nx=40, P=6, seed=13, shift=True)
    pr["obj"] = pr["obj"].astype(np.complex128)
    pr["probe"] = pr["probe"].astype(np.complex128)
    pr["corr"] = pr["corr"].astype(np.float64)
    g = oracle_pass(pr, np.complex128)[4]
    for p in range(3):
        for ax in range(2):
            fd = _fd_loss(pr, "corr", (p, ax), 1.0, eps=1e-5)
            assert abs(fd - g[p, ax]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (p, ax, fd, g[p, ax])


def test_probe_gradient_is_band_limited_projection_without_shift():
    """Without sub-pixel shift the reference band-limits the probe gradient: it equals the FD
    derivative along in-band directions only."""
    pr = make_problem(n=16, S=1, K=1, M=1, ny=40, nx=40, P=5, seed=14, shift=False)
    pr["obj"] = pr["obj"].astype(np.complex128)
    pr["probe"] = pr["probe"].astype(np.complex128)
    g = oracle_pass(pr, np.complex128)[3]
    n = 16
    _, _, _, _, mask = orc.create_spatial_frequencies(pr["px"], pr["px"], n, 2 / 3, 0.0, np.float64)
    rng = np.random.default_rng(6)
    d = np.fft.ifft2(mask * np.fft.fft2(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))))
    eps = 1e-5

    def L(s):
        q = dict(pr)
        q["probe"] = pr["probe"] + s * d[:, :, None]
        return oracle_pass(q, np.complex128, step_obj=False, step_probe=False)[0]

    fd = (L(eps) - L(-eps)) / (2 * eps)
    an = 2 * np.real(np.sum(np.conj(g[:, :, 0]) * d))
    assert abs(fd - an) < 1e-5 * abs(fd)


def test_complex64_oracle_tracks_fp64():
    pr = make_problem(n=32, S=3, K=1, M=2, seed=15)
    a = oracle_pass(pr, np.complex64)
    b = oracle_pass(pr, np.complex128)
    assert abs(a[0] - b[0]) / b[0] < 1e-5
    for i in (2, 3, 4):
        assert rel(a[i], b[i]) < 1e-4


# ---------------------------------------------------------------- L-BFGS
def test_two_loop_matches_textbook_bfgs_inverse_on_quadratic():
    """Nocedal & Wright Alg. 7.4 == explicit inverse-Hessian recursion (single block, gamma = s'y/y'y)."""
    rng = np.random.default_rng(7)
    n = 6
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    S = [rng.standard_normal(n) for _ in range(3)]
    Y = [A @ s for s in S]
    g = rng.standard_normal(n)
    rhos = [1.0 / np.dot(y, s) for s, y in zip(S, Y)]
    gamma = np.dot(S[-1], Y[-1]) / np.dot(Y[-1], Y[-1])
    d = orc.two_loop([g], [[s] for s in S], [[y] for y in Y], rhos, [gamma])[0]
    Hk = gamma * np.eye(n)
    for s, y, r in zip(S, Y, rhos):
        V = np.eye(n) - r * np.outer(y, s)
        Hk = V.T @ Hk @ V + r * np.outer(s, s)
    assert rel(d, -Hk @ g) < 1e-12


def test_complex_inner_product_is_real_inner_product_of_pairs():
    rng = np.random.default_rng(8)
    a = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    b = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    assert np.isclose(orc.cdot(a, b), 2 * (np.dot(a.real, b.real) + np.dot(a.imag, b.imag)))


def test_wolfe_conditions():
    assert orc.wolfe_1(10.0, 9.0, -5.0, 0.1, 0.1)
    assert not orc.wolfe_1(10.0, 9.99, -5.0, 1.0, 0.1)
    assert orc.wolfe_2(-5.0, -1.0, 0.9)
    assert not orc.wolfe_2(-5.0, -4.9, 0.9)


def test_bfgs_oracle_epochs_decrease_loss():
    from oracle import bfgs_oracle

    pr = make_problem(n=16, S=2, K=1, M=1, ny=40, nx=40, P=8, seed=31, shift=True, noise=0.0)
    state = {"obj": np.ones_like(pr["obj"], dtype=np.complex128), "probe": pr["probe"].astype(np.complex128),
             "positions": pr["pos"], "corr": pr["corr"].astype(np.float64)}
    hist = bfgs_oracle.new_history()
    data = {"measured": pr["meas"], "sequence": pr["seq"], "dz": pr["dz"], "wavelength": pr["wavelength"],
            "px": pr["px"], "compute_batch": 4}
    np.random.seed(0)
    losses, accepted, step = [], [], 1.0
    for _ in range(6):
        out = bfgs_oracle.bfgs_epoch(state, hist, data, update_pos=True, actual_step=step, cdtype=np.complex128)
        step = out[2]
        losses.append(out[0])
        accepted.append(out[6])
    # a rejected epoch keeps the iterate (and resets the history); accepted ones satisfy Armijo
    assert all(b <= a * (1 + 1e-12) for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.7 * losses[0]
    assert sum(accepted) >= 4
    assert hist["rho_hist"] and all(r > 0 for r in hist["rho_hist"])
