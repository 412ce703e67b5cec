"""Pins of the oracle's beam tilts and per-slice distances (pypty/objective.py:303-307, 328-350,
428-499, 600-672; pypty/multislice_core.py:69-77, 163-178).

* the tilt gradient of every kind (before, inside, after the specimen), per position and with one
  tilt row, against central differences of the FP64 loss;
* the sign and scale of each tilt by closed forms:
  - inside: with O = 1 a tilt (ty, tx) moves the exit wave by (S-1) d (ty, tx) / p pixels, a roll
    by +that amount when it is a whole number;
  - after (one slice): the real-space ramp moves the far field by k0 = n p tan / lambda bins;
  - before: the same for the probe;
* per-slice distances: O = 1 gives free propagation over their sum, equal distances give the
  single-distance pass, and the inside tilt gradient weighs each propagation by its own distance.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import make_problem, rel

F64 = np.complex128


def _pass(pr, **kw):
    base = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                algorithm=pr["algorithm"], step_obj=True, step_probe=True, step_pos=pr["step_pos"], cdtype=F64,
                compute_batch=4)
    base.update(kw)
    return orc.loss_and_direction(pr["obj"].astype(F64), pr["probe"].astype(F64), pr["pos"], pr["corr"], pr["meas"],
                                  pr["seq"], **base)


def _tilts(rows, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, (rows, 6)) * np.array([2e-3, 2e-3, 4e-2, 4e-2, 2e-3, 2e-3])).astype(np.float64)


@pytest.mark.parametrize("mode,per_position,shift,S,M,extra", [
    (0, True, True, 3, 2, {}),
    (0, False, False, 3, 1, {}),
    (4, True, True, 3, 2, {}),
    (4, True, False, 1, 2, {}),
    (3, False, True, 2, 1, {}),
    (2, True, False, 2, 2, {"ab": True}),
    (1, True, False, 1, 1, {}),
    (4, True, True, 3, 1, {"distances": [1.5, 2.5]}),
    (4, True, True, 3, 2, {"edge": 1.7}),
])
def test_tilt_gradient_matches_central_differences(mode, per_position, shift, S, M, extra):
    pr = make_problem(n=16, S=S, K=1, M=M, ny=40, nx=40, P=6, shift=shift, seed=30 + mode + S)
    T = 6 if per_position else 1
    tl = _tilts(T, 7 + mode)
    kw = dict(tilts=tl, tilt_mode=mode, step_tilts=True, edge=extra.get("edge", 1.0))
    if "distances" in extra:
        kw["distances"] = np.asarray(extra["distances"])
    if extra.get("ab"):
        kw.update(aberrations=np.array([[30.0, 10.0, -5.0]]), aberration_marker=np.zeros(6, int))
    out = _pass(pr, **kw)
    tg = out[-1]
    active = {0: [2, 3], 1: [4, 5], 2: [0, 1, 2, 3], 3: [2, 3, 4, 5], 4: [0, 1, 2, 3, 4, 5]}[mode]
    if mode in (1, 2) and T > 1:
        active = [c for c in active if c not in (2, 3)]  # one master propagator from row 0
    rows = range(T) if T <= 2 else [0, 3, 5]
    checked = 0
    for r in rows:
        for c in active:
            eps = 2e-6 if c in (2, 3) else 1e-7
            tp, tm = tl.copy(), tl.copy()
            tp[r, c] += eps
            tm[r, c] -= eps
            lp = _pass(pr, **{**kw, "tilts": tp, "step_tilts": False})[0]
            lm = _pass(pr, **{**kw, "tilts": tm, "step_tilts": False})[0]
            fd = (lp - lm) / (2 * eps)
            # small components sit in the curvature error of the large ones (the real-space ramps
            # change the loss on a scale of lambda / (n p) in the tilt)
            scale = max(abs(fd), 1e-2 * np.abs(tg[:, active]).max(), 1e-3 * abs(out[0]))
            assert abs(tg[r, c] - fd) <= 2e-5 * scale, (r, c, tg[r, c], fd)
            checked += 1
    # the columns of the kinds this mode does not apply stay zero
    for c in set(range(6)) - set(active) - ({2, 3} if mode in (1, 2) else set()):
        assert np.all(tg[:, c] == 0)
    assert checked >= 2


def test_with_aberrations_and_a_shift_the_before_tilt_does_not_reach_the_loss():
    """The reference's shift starts from the spectrum taken before the aberration plates
    (objective.py:462, 471-474), so the real-space tilt applied between them is lost as well; the
    gradient is still formed (:660-672)."""
    pr = make_problem(n=16, S=2, K=1, M=2, ny=40, nx=40, P=6, shift=True, seed=41)
    kw = dict(tilt_mode=2, step_tilts=True, aberrations=np.array([[30.0, 10.0, -5.0]]),
              aberration_marker=np.zeros(6, int))
    t1 = _tilts(6, 1)
    t2 = t1.copy()
    t2[:, :2] = _tilts(6, 2)[:, :2]
    a = _pass(pr, tilts=t1, **kw)
    b = _pass(pr, tilts=t2, **kw)
    assert a[0] == b[0] and rel(a[2], b[2]) < 1e-13
    assert np.abs(a[-1][:, :2]).max() > 0


def test_inside_tilt_moves_the_exit_wave_by_d_tan():
    n, S, d, px, lam = 32, 3, 2.0, 0.2, 0.0196868
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    rng = np.random.default_rng(3)
    probe = orc.ifft2(mask * np.exp(2j * np.pi * rng.uniform(size=(n, n))))[None, :, :, None]
    ones = np.ones((1, n, n, S, 1), F64)
    # (S-1) d tan / p = 4 * 0.1 / 0.2 = 2 px along x, 4 * -0.05 / 0.2 = -1 px along y
    tx, ty = 0.1, -0.05
    H0 = orc.propagator(d, lam, q2, qx, qy, mask, cdtype=F64)
    Ht = orc.propagator(d, lam, q2, qx, qy, mask, tx=tx, ty=ty, cdtype=F64)
    _, e0 = orc.multislice(probe, ones, H0)
    _, et = orc.multislice(probe, ones, Ht)
    # the propagator's pi is the literal 3.141592654 (objective.py:343), 4e-10 off
    assert rel(et, np.roll(e0, (-1, 2), axis=(1, 2))) < 1e-8
    assert rel(et, np.roll(e0, (1, -2), axis=(1, 2))) > 0.1


@pytest.mark.parametrize("kind", ["after", "before"])
def test_real_space_tilt_moves_the_far_field_by_k0_bins(kind):
    n, px, lam = 32, 0.2, 0.0196868
    pr = make_problem(n=n, S=1, K=1, M=1, ny=48, nx=48, P=3, shift=False, seed=5)
    ones = np.ones_like(pr["obj"])
    k0y, k0x = 2, -3
    tan_y, tan_x = k0y * lam / (n * px), k0x * lam / (n * px)
    tl = np.zeros((1, 6))
    col = {"after": (4, 5), "before": (0, 1)}[kind]
    tl[0, col[0]], tl[0, col[1]] = tan_y, tan_x
    mode = {"after": 1, "before": 2}[kind]
    common = dict(dz=1.0, wavelength=lam, pixel_size_x=px, pixel_size_y=px, step_obj=False, step_probe=False,
                  step_pos=False, cdtype=F64, return_patterns=True)
    probe = pr["probe"].astype(F64)
    I0 = orc.loss_and_direction(ones, probe, pr["pos"], pr["corr"], pr["meas"], pr["seq"], **common)[5]
    It = orc.loss_and_direction(ones, probe, pr["pos"], pr["corr"], pr["meas"], pr["seq"], tilts=tl, tilt_mode=mode,
                                **common)[5]
    # the far field before the detector band mask: the band-limited exit wave's spectrum (after)
    # or the probe's (before, one slice of O = 1), moved by (k0y, k0x); then the mask
    _, _, _, mask_s, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    ex = probe[:, :, 0]
    if kind == "after":
        ex = orc.ifft2(orc.fft2(ex) * mask)
    # the ramp's phase origin is the window centre: exp(2 pi i k0 (j - n/2) / n)
    ramp = np.exp(2j * np.pi * (k0y * (np.arange(n)[:, None] - n // 2) + k0x * (np.arange(n)[None, :] - n // 2)) / n)
    expect = np.abs(orc.shift_fft2(ex * ramp) * mask_s) ** 2
    rolled = np.abs(np.roll(orc.shift_fft2(ex), (k0y, k0x), axis=(0, 1)) * mask_s) ** 2
    assert rel(expect, rolled) < 1e-10
    assert rel(It[0], rolled) < 1e-6
    assert rel(I0[0], rolled) > 0.05


def test_per_slice_distances_free_propagation_and_uniform_case():
    n, S, px, lam = 32, 4, 0.2, 0.0196868
    q2, qx, qy, _, mask = orc.create_spatial_frequencies(px, px, n, 2 / 3, 0.0, np.float64)
    rng = np.random.default_rng(4)
    probe = orc.ifft2(mask * np.exp(2j * np.pi * rng.uniform(size=(n, n))))[None, :, :, None]
    ones = np.ones((1, n, n, S, 1), F64)
    d = [1.0, 2.5, 0.75]
    props = [orc.propagator(x, lam, q2, qx, qy, mask, cdtype=F64) for x in d]
    _, e = orc.multislice(probe, ones, props)
    Hsum = orc.propagator(sum(d), lam, q2, qx, qy, mask, cdtype=F64)
    assert rel(e[:, :, :, :, 0], orc.ifft2(orc.fft2(probe, axes=(1, 2)) * Hsum[None, :, :, None], axes=(1, 2))) < 1e-10
    # equal distances: the per-slice pass is the single-distance pass
    pr = make_problem(n=16, S=3, K=1, M=2, ny=40, nx=40, P=5, seed=9)
    a = _pass(pr)
    b = _pass(pr, distances=np.array([pr["dz"]] * 3))
    assert abs(a[0] - b[0]) < 1e-12 * abs(a[0]) and rel(b[2], a[2]) < 1e-12 and rel(b[3], a[3]) < 1e-12


def test_per_slice_distance_gradients_match_central_differences():
    """The object and probe gradients through slices of different thickness (the propagations use
    their own distances in the adjoint, multislice_core.py:164-168)."""
    pr = make_problem(n=16, S=3, K=1, M=1, ny=40, nx=40, P=4, seed=11, shift=False)
    dist = np.array([1.0, 3.0])
    out = _pass(pr, distances=dist, step_pos=False)
    rng = np.random.default_rng(2)
    for _ in range(4):
        y, x, s = rng.integers(10, 30), rng.integers(10, 30), rng.integers(0, 3)
        eps = 1e-6
        for part, unit in ((0, 1.0), (1, 1j)):
            op = pr["obj"].astype(F64).copy()
            om = op.copy()
            op[y, x, s, 0] += eps * unit
            om[y, x, s, 0] -= eps * unit
            lp = orc.loss_and_direction(op, pr["probe"].astype(F64), pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                        dz=pr["dz"], distances=dist, wavelength=pr["wavelength"],
                                        pixel_size_x=pr["px"], pixel_size_y=pr["px"], step_obj=False,
                                        step_probe=False, cdtype=F64)[0]
            lm = orc.loss_and_direction(om, pr["probe"].astype(F64), pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                        dz=pr["dz"], distances=dist, wavelength=pr["wavelength"],
                                        pixel_size_x=pr["px"], pixel_size_y=pr["px"], step_obj=False,
                                        step_probe=False, cdtype=F64)[0]
            fd = (lp - lm) / (2 * eps)
            # the reference's gradient is dL/dO* = (dL/dRe + i dL/dIm) / 2 in this convention
            g = out[2][y, x, s, 0]
            an = 2 * (g.real if part == 0 else g.imag)
            assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-3), (y, x, s, part, an, fd)
