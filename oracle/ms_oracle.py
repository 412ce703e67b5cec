"""numpy restatement of PyPty's far-field multislice path (test oracle only).

Arrays use the reference's own layouts:
  probe (N, N, M), object (Ny, Nx, S, K), positions / corrections (P, 2) as (y, x),
  measured (P, N, N) with the zero frequency at (N/2, N/2),
  per-window probe (B, N, N, M), windows (B, N, N, S, K),
  wavefield stash (B, N, N, S, M, K), exit wave (B, N, N, M, K).
``cdtype`` selects complex64 (the reference's "single" path, same op order)
or complex128 (the FP64 oracle).  Supported subset: far field, single slice
distance, no tilts, losses lsq_sqrt / ml / lsq / lsq_sqrt_2 and the compressed
lsq / gauss / poisson (virtual detectors), object / probe /
position gradients.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sf

LSQ_SQRT, ML, LSQ, LSQ_SQRT_2 = "lsq_sqrt", "ml", "lsq", "lsq_sqrt_2"
LSQ_COMPRESSED, GAUSS_COMPRESSED, POISSON_COMPRESSED = "lsq_compressed", "gauss_compressed", "poisson_compressed"
COMPRESSED = (LSQ_COMPRESSED, GAUSS_COMPRESSED, POISSON_COMPRESSED)


def _fdt(cdtype):
    return np.float32 if np.dtype(cdtype) == np.complex64 else np.float64


# ----------------------------------------------------------------- FFTs
def fft2(a, axes=(0, 1)):
    """pypty/fft.py:114-134: unnormalised forward 2D DFT."""
    return sf.fft2(a, axes=axes)


def ifft2(a, axes=(0, 1)):
    """pypty/fft.py:136-156: inverse 2D DFT with 1/N^2."""
    return sf.ifft2(a, axes=axes)


def shift_fft2(a, axes=(0, 1)):
    """pypty/fft.py:91-112: fftshift(fft2(a))."""
    return sf.fftshift(sf.fft2(a, axes=axes), axes=axes)


def ifft2_ishift(a, axes=(0, 1)):
    """pypty/fft.py:158-178: ifft2(ifftshift(a))."""
    return sf.ifft2(sf.ifftshift(a, axes=axes), axes=axes)


def dft2_bruteforce(a):
    """Direct O(N^4) evaluation of the 2D DFT over axes (0, 1), in float64/complex128."""
    a = np.asarray(a, dtype=np.complex128)
    ny, nx = a.shape[:2]
    wy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    wx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    return np.einsum("ky,lx,yx...->kl...", wy, wx, a)


# ---------------------------------------------------------- grids, masks
def create_spatial_frequencies(px, py, n, cutoff, rolloff=0.0, fdtype=np.float32):
    """pypty/utils.py:153-199 (rolloff forced to 0 on the loss path, iterative.py:483)."""
    qx, qy = np.meshgrid(np.fft.fftfreq(n, px), np.fft.fftfreq(n, py), indexing="xy")
    qx, qy = qx.astype(fdtype), qy.astype(fdtype)
    q2 = qx ** 2 + qy ** 2
    if cutoff is None:
        cutoff = 2
    if cutoff < 0:
        cutoff = 2 / 3
    max_x = np.max(qx)
    r_max = (max_x * fdtype(cutoff)) ** 2
    mask = (q2 <= r_max).astype(fdtype)
    if rolloff != 0:
        r0 = (max_x * (cutoff - rolloff)) ** 2
        sel = q2 > r0
        mask[sel] *= (0.5 * (1 + np.cos(3.141592654 * (q2 - r0) / (r_max - r0))))[sel]
    mask = mask.astype(fdtype)
    return q2, qx, qy, sf.fftshift(mask), mask


def propagator(dz, wavelength, q2, qx, qy, mask, tx=0.0, ty=0.0, cdtype=np.complex64):
    """Fresnel propagator of one slice, objective.py:337-345 (and multislice_core.py:70-72).

    tx, ty: the in-specimen tilt (tilt columns 3 and 2), scalars or per-window arrays (B, 1, 1),
    which give one propagator per window (objective.py:482-484)."""
    fd = _fdt(cdtype)
    dz, wavelength = fd(dz), fd(wavelength)
    tx, ty = np.asarray(tx, fd), np.asarray(ty, fd)
    h = np.exp(-3.141592654j * dz * (wavelength * q2 + 2 * (qx * tx + qy * ty))) * mask
    return h.astype(cdtype)


def tilt_real_grids(n, px, py, wavelength, cdtype=np.complex64):
    """objective.py:303-307: the real-space tilt grids i 2 pi fftshift(fftfreq(n)) n p / lambda
    along y and x, so that exp(xg tan_x + yg tan_y) tilts a wave by (tan_y, tan_x)."""
    gx, gy = np.meshgrid(sf.fftshift(np.fft.fftfreq(n)), sf.fftshift(np.fft.fftfreq(n)), indexing="xy")
    xg = (gx * 6.2831855j * n * px / wavelength).astype(cdtype)
    yg = (gy * 6.2831855j * n * py / wavelength).astype(cdtype)
    return yg, xg


def shift_masks(n, mask, cdtype=np.complex64):
    """Sub-pixel shift kernels Ky, Kx = -2*pi*i*fftfreq(n)*mask, objective.py:308-313."""
    sx, sy = np.meshgrid(np.fft.fftfreq(n), np.fft.fftfreq(n), indexing="xy")
    sx = (-6.2831855j * sx).astype(cdtype) * mask
    sy = (-6.2831855j * sy).astype(cdtype) * mask
    return sy.astype(cdtype), sx.astype(cdtype)


def fourier_clean(array, cutoff=0.66, rolloff=0.0, fdtype=np.float32):
    """pypty/utils.py:102-150: band-limit over axes (0, 1) in fftshifted normalised frequency."""
    shape = array.shape
    x = sf.fftshift(np.fft.fftfreq(shape[1]))
    y = sf.fftshift(np.fft.fftfreq(shape[0]))
    x, y = np.meshgrid(x, y, indexing="xy")
    r = x ** 2 + y ** 2
    max_r = 0.5 * cutoff
    mask = (r <= max_r ** 2).astype(fdtype)
    if rolloff != 0:
        r0 = (0.5 * (cutoff - rolloff)) ** 2
        sel = r > r0
        mask[sel] *= (0.5 * (1 + np.cos(3.141592654 * (r - r0) / (max_r ** 2 - r0))))[sel]
    ff = shift_fft2(array, axes=(0, 1))
    ff = ff * mask.reshape(mask.shape + (1,) * (len(shape) - 2))
    return ifft2_ishift(ff, axes=(0, 1))


# ------------------------------------------------------------ multislice
def _props(H, S):
    """The propagator after each slice 0..S-2: one array for every slice, or a list of them."""
    return list(H) if isinstance(H, (list, tuple)) else [H] * max(S - 1, 0)


def _per_window(h):
    """(N, N) or (B, N, N) propagator as (1|B, N, N, 1, 1)."""
    return h[None, :, :, None, None] if h.ndim == 2 else h[:, :, :, None, None]


def multislice(probe_b, obj_b, H, edge=1.0, spectral=False):
    """Split-step forward, pypty/multislice_core.py:9-82.

    probe_b (B, N, N, M); obj_b (B, N, N, S, K); H: the propagator (N, N) or one per window
    (B, N, N), or a list of S-1 of them, one per slice (this_distances per slice,
    multislice_core.py:69-71).  The first propagation is raised to edge when edge > 1 (:73-74).
    Returns the conjugated stash (B, N, N, S, M, K) and the exit wave (B, N, N, M, K), and with
    spectral also the conjugated spectra fft2(psi_s O_s) that the forward keeps for the
    in-specimen tilt gradient (:76-77, channel 1 of waves_multislice; slice S-1 stays zero).
    """
    B, N, _, M = probe_b.shape
    S, K = obj_b.shape[3], obj_b.shape[4]
    props = _props(H, S)
    waves = np.zeros((B, N, N, S, M, K), dtype=probe_b.dtype)
    spec = np.zeros_like(waves) if spectral else None
    wave = np.repeat(probe_b[:, :, :, :, None], K, axis=-1)
    for s in range(S):
        waves[:, :, :, s] = wave
        wave = wave * obj_b[:, :, :, s:s + 1, :]
        if s < S - 1:
            Hb = _per_window(props[s])
            prop = Hb ** edge if (edge > 1 and s == 0) else Hb
            wave = fft2(wave, axes=(1, 2))
            if spectral:
                spec[:, :, :, s] = wave
            wave = ifft2(wave * prop, axes=(1, 2))
    if spectral:
        return np.conj(waves), wave, np.conj(spec)
    return np.conj(waves), wave


def multislice_grads(g, waves, obj_b, H, mask, edge=1.0, step_obj=True, spec=None, qx=None, qy=None,
                     distances=None):
    """Reverse-mode adjoint, pypty/multislice_core.py:85-200.

    g (B, N, N, M, K) = dL/d(exit)*; waves = conjugated stash from multislice(); H as there.
    With spec (the conjugated spectra from multislice(..., spectral=True)), qx, qy and the slice
    distances, also the in-specimen tilt gradient (:163-178): per window, for each propagation
    from slice s-1, -4 pi d / N^2 sum Im(spec[s-1] fft2(g) conj(H) q) over pixels and modes, with
    d the distance of that propagation (times edge for the first).
    Returns (band-limited window gradients (B, N, N, S, K) or None,
             probe-side gradient summed over object modes (B, N, N, M)[, tilt gradient (B, 2) as (y, x)]).
    """
    waves = waves.copy()
    objc = np.conj(obj_b)
    B, N = waves.shape[0], waves.shape[1]
    S = waves.shape[3]
    props = _props(H, S)
    tg = None if spec is None else np.zeros((B, 2), dtype=np.float64 if g.dtype == np.complex128 else np.float32)
    dist = None if distances is None else np.asarray(distances).ravel()
    for s in range(S - 1, -1, -1):
        if step_obj:
            waves[:, :, :, s] *= g
        g = g * objc[:, :, :, s:s + 1, :]
        if s > 0:
            Hc = np.conj(_per_window(props[s - 1]))
            prop = Hc ** edge if (edge > 1 and s == 1) else Hc
            g = fft2(g, axes=(1, 2)) * prop
            if tg is not None:
                d = dist[0] if dist.shape[0] == 1 else dist[s - 1]
                if edge > 1 and s == 1:
                    d = d * edge
                sh = -12.566370614359172 * d / (N * N)
                dl = np.sum(spec[:, :, :, s - 1] * g, axis=(3, 4))
                tg[:, 1] += sh * np.sum(np.imag(dl * qx), axis=(1, 2))
                tg[:, 0] += sh * np.sum(np.imag(dl * qy), axis=(1, 2))
            g = ifft2(g, axes=(1, 2))
    grad = None
    if step_obj:
        G = waves[:, :, :, :, 0, :] if waves.shape[4] == 1 else np.sum(waves, axis=4)
        grad = ifft2(fft2(G, axes=(1, 2)) * mask[None, :, :, None, None], axes=(1, 2))
    interm = g[..., 0] if g.shape[-1] == 1 else np.sum(g, axis=-1)
    if tg is not None:
        return grad, interm, tg
    return grad, interm


def scatteradd(full, my, mx, chop):
    """pypty/multislice_core.py:206-230: full[my, mx] += chop with repeated indices accumulating."""
    np.add.at(full, (my, mx), chop)


# ------------------------------------------------------ loss + gradients
def convert_num_to_nmab(num_abs):
    """utils.py:1272-1308: (n, m, a/b) of the first num_abs aberrations in Krivanek order."""
    max_n = int(np.ceil(0.5 * (-5 + np.sqrt(25 + 8 * num_abs))))
    ns, ms, abs_ = [], [], []
    for n in range(1, max_n + 1):
        for m in range((n + 1) % 2, n + 2, 2):
            if m == 0:
                ns.append(n), ms.append(m), abs_.append("")
            else:
                ns.append(n), ms.append(m), abs_.append("a")
                ns.append(n), ms.append(m), abs_.append("b")
    return ns[:num_abs], ms[:num_abs], abs_[:num_abs]


def ctf_matrix(kx, ky, num_abs, wavelength):
    """utils.py:1333-1367: (num_abs, Ny, Nx) phase per unit coefficient, 2 pi / lambda k^(n+1) / (n+1)
    times cos(m phi), sin(m phi) or 1, with k the beam angle (lambda q)."""
    ns, ms, abs_ = convert_num_to_nmab(num_abs)
    kr = np.sqrt(kx ** 2 + ky ** 2)
    phi = np.angle(kx + 1j * ky)
    A = np.zeros((num_abs,) + kr.shape)
    for i in range(num_abs):
        term = kr ** (ns[i] + 1) / (ns[i] + 1)
        if abs_[i] == "a":
            term = term * np.cos(ms[i] * phi)
        elif abs_[i] == "b":
            term = term * np.sin(ms[i] * phi)
        A[i] = term
    return A * 6.283185307179586 / wavelength


def loss_and_direction(obj, probe, positions, corrections, measured, sequence, *, dz, wavelength,
                       pixel_size_x, pixel_size_y, algorithm=LSQ_SQRT, loss_weight=1.0, cutoff=2 / 3,
                       edge=1.0, step_obj=True, step_probe=True, step_pos=False, compute_batch=64,
                       cdtype=np.complex64, return_patterns=False, masks=None, static_background=None,
                       step_static_background=False, beam_current=None, step_beam_current=False, aberrations=None,
                       aberration_marker=None, step_aberrations=False, distances=None, tilts=None, tilt_mode=0,
                       step_tilts=False):
    """objective.loss_and_direction restricted to the far-field multislice path.

    Follows pypty/objective.py:218-361 (grids, propagator, shift kernels),
    :393-703 (chunk loop: window cut and band-limit, probe shift, multislice,
    far field, loss, adjoint, probe shift adjoint, position and probe
    gradients), :742 (loss x loss_weight), :925 (NaN check) and :927-933
    (probe-gradient band-limit).
    Returns loss, sse, obj_grad (Ny, Nx, S, K), probe_grad (N, N, M), pos_grad (P, 2)
    [, modelled intensity (P, N, N) if return_patterns].
    For the compressed losses (objective.py:560-585) measured holds (P, J)
    virtual-detector coefficients and masks the (J, N, N) detectors in the
    DC-centred pattern layout.
    static_background (N, N), DC-centred, follows objective.py:359-361 (b^2
    times the DC-centred band mask), :525-526 (added to the pattern after the
    simulated-pattern copy) and :377, :588-589 (gradient 2 b w sum_b dL/dI,
    with dL/dI after the 1/(M*K) of :586-587, not band-limited); when it is
    given the gradient (zeros unless step_static_background) is appended to
    the result.
    beam_current (P,) follows objective.py:446-450 (each position's probe,
    before its shift, times |beam_current|) and :688-698 (gradient
    2 sign(|bc|) sum Re(dL/dP* conj(P)) of the unscaled probe, then the probe
    gradient times |bc|); when it is given its gradient (P,) (zeros unless
    step_beam_current, rows outside the sequence zero) is appended last.  The
    reference's one-mode branch conjugates the caller's probe in place
    (objective.py:692) and so feeds the conjugate to the next batch; this
    restatement keeps the probe as given, as its many-mode branch does.
    aberrations (n_markers, num_abs) with aberration_marker (P,) follow
    objective.py:320-326, 459-465 (each position's probe spectrum, after the
    beam current, times exp(-i sum_i A_i a_i) mask of its marker's
    coefficients) and :673-681 (the probe-side gradient spectrum times the
    conjugate plate, and per window 2/N^2 sum Re(Fg conj(FP) conj(-i A_i))
    added to its marker's row).  As in the reference, the sub-pixel shift
    (:471-474) starts again from the spectrum before the plate, so with a shift
    the plates leave the loss unchanged while their gradient is still formed.
    When given, the aberration gradient is appended last.
    distances (S-1 or more per-slice distances, or one) replaces dz when given: slice s propagates
    by distances[s] (multislice_core.py:69-71, 164-168).
    tilts (T, 6) with T = 1 or one row per position, columns (y, x) before, inside and after the
    specimen (objective.py:134 of iterative.py, :428-434), and tilt_mode (iterative.py:135):
      * inside (modes 0, 3, 4): the propagator exp(-i pi d (lambda q^2 + 2 (qx tx + qy ty))) with the
        window's inside tilt (objective.py:482-484); one tilt row, or modes 1 and 2, use row 0's
        for every window (:328, :337-345, the master propagator).  The reference keeps that
        master propagator in a module cache across calls (:337), so a refined single tilt would
        not reach later calls; this restatement forms it per call.
      * before (modes 2, 4): the probe times exp(xg tx + yg ty) in real space after the beam
        current and aberrations, before the sub-pixel shift (:465-470); as with the aberration
        plates, a shift starts from the probe spectrum taken before the aberrations when there
        are aberrations, so that then the tilt does not reach the modelled patterns either.
      * after (modes 1, 3, 4): the exit wave times exp(xg tx + yg ty), band-limited first for one
        slice (:490-499).
    step_tilts adds the tilt gradient (T, 6): inside from multislice_grads (:163-178), before
    and after -2 sum Re(g conj(wave before the tilt) (yg, xg)) over pixels and modes
    (:620-636, :660-672).  Rows of the windows are assigned, except inside, which adds; with one
    tilt row every window adds into row 0 (the reference adds a per-window array into that single
    row there, which only broadcasts for one window per batch).  When tilts are given the tilt
    gradient is appended last.
    """
    fd = _fdt(cdtype)
    obj = np.asarray(obj, dtype=cdtype)
    probe = np.asarray(probe, dtype=cdtype)
    corrections = np.asarray(corrections, dtype=fd)
    measured = np.asarray(measured, dtype=fd)
    N, M = probe.shape[0], probe.shape[2]
    S, K = obj.shape[2], obj.shape[3]
    q2, qx, qy, mask_s, mask = create_spatial_frequencies(pixel_size_x, pixel_size_y, N, cutoff, 0.0, fd)
    dists = np.asarray([dz] if distances is None else distances, dtype=np.float64).ravel()
    single_dist = dists.shape[0] == 1
    tl = None if tilts is None else np.asarray(tilts, dtype=fd)
    tilt_mode = int(tilt_mode) if tl is not None else 0
    tilt_before, tilt_after = tilt_mode in (2, 4), tilt_mode in (1, 3, 4)
    tilt_inside_grad = bool(step_tilts) and tl is not None and tilt_mode in (0, 3, 4)
    single_tilt = tl is None or tl.shape[0] == 1
    same_prop = single_tilt or tilt_mode in (1, 2)
    if tl is not None:
        yg, xg = tilt_real_grids(N, pixel_size_x, pixel_size_y, wavelength, cdtype)
        yxg = np.stack((yg, xg), axis=0)  # (2, N, N)
    tilt_grad = None if tl is None else np.zeros_like(tl)
    t0 = np.zeros(6, fd) if tl is None else tl[0]
    H = propagator(dists[0], wavelength, q2, qx, qy, mask, tx=t0[3], ty=t0[2], cdtype=cdtype)
    shy, shx = shift_masks(N, mask, cdtype)
    shift_flag = bool(np.sum(corrections.astype(np.float64) ** 2) != 0.0) or bool(step_pos)
    runx, runy = np.meshgrid(np.arange(N), np.arange(N), indexing="xy")
    coherent = (M == 1 and K == 1)
    obj_grad = np.zeros_like(obj)
    probe_grad = np.zeros((N, N, M), dtype=cdtype)
    pos_grad = np.zeros_like(corrections)
    compressed = algorithm in COMPRESSED
    if compressed:
        masks = np.asarray(masks, dtype=fd)
    patterns = np.zeros((measured.shape[0], N, N), dtype=fd) if return_patterns else None
    bg = None if static_background is None else np.asarray(static_background, dtype=fd)
    bg_square = None if bg is None else (bg[None] ** 2) * mask_s[None]
    bg_grad = None if bg is None else np.zeros((N, N), dtype=fd)
    bc = None if beam_current is None else np.asarray(beam_current, dtype=fd)
    bc_grad = None if bc is None else np.zeros_like(bc)
    ab = None if aberrations is None else np.asarray(aberrations, dtype=fd)
    ab_grad = None if ab is None else np.zeros_like(ab)
    if ab is not None:
        marker = np.asarray(aberration_marker)
        A = ctf_matrix(wavelength * qx, wavelength * qy, ab.shape[1], wavelength)
        polys = (-1j * A).astype(cdtype)
        plates = np.exp(np.sum(polys[None] * ab[:, :, None, None], 1)) * mask[None]
    FP = fft2(probe[None], axes=(1, 2))
    loss, sse = fd(0), fd(0)
    seq = np.asarray(sequence)
    for i in range(0, len(seq), compute_batch):
        tcs = seq[i:i + compute_batch]
        meas = measured[tcs]
        my = positions[tcs, 0][:, None, None] + runy[None]
        mx = positions[tcs, 1][:, None, None] + runx[None]
        chop = obj[my, mx, :, :]
        chop = ifft2(fft2(chop, axes=(1, 2)) * mask[None, :, :, None, None], axes=(1, 2))
        this_probe = probe[None]
        FPb = FP
        if bc is not None:
            thisbc = np.abs(bc[tcs])
            this_probe = this_probe * thisbc[:, None, None, None]
            FPb = FP * thisbc[:, None, None, None]
        if tl is not None:
            tb = tl[[0]] if single_tilt else tl[tcs]
            tcol = [tb[:, c, None, None] for c in range(6)]  # (B|1, 1, 1) each
        FPs = FPb  # the spectrum the shift starts from
        if ab is not None:
            this_probe = ifft2(FPb * plates[marker[tcs]][..., None], axes=(1, 2))
        if tilt_before:
            ramp_b = np.exp(xg[None] * tcol[1] + yg[None] * tcol[0])  # (B|1, N, N)
            probe_bt = this_probe
            this_probe = probe_bt * ramp_b[..., None]
            if ab is None:
                FPs = fft2(this_probe, axes=(1, 2))
        E = None
        if shift_flag:
            corr = corrections[tcs]
            E = np.exp(shx[None] * corr[:, 1, None, None] + shy[None] * corr[:, 0, None, None]) * mask[None]
            this_probe = ifft2(FPs * E[..., None], axes=(1, 2))
        if this_probe.shape[0] == 1 and len(tcs) > 1:
            this_probe = np.repeat(this_probe, len(tcs), axis=0)
        if single_dist:
            props = H if same_prop else propagator(dists[0], wavelength, q2, qx[None], qy[None], mask[None],
                                                   tx=tcol[3], ty=tcol[2], cdtype=cdtype)
        else:
            tx_in, ty_in = (tcol[3], tcol[2]) if tl is not None else (0.0, 0.0)
            props = [propagator(dists[s], wavelength, q2, qx, qy, mask, tx=tx_in, ty=ty_in, cdtype=cdtype)
                     for s in range(S - 1)]
        spec = None
        if tilt_inside_grad:
            waves, exitw, spec = multislice(this_probe, chop, props, edge, spectral=True)
        else:
            waves, exitw = multislice(this_probe, chop, props, edge)
        if tilt_after:
            ramp_a = np.exp(xg[None] * tcol[5] + yg[None] * tcol[4])[:, :, :, None, None]
            if S == 1:
                exitw = ifft2(fft2(exitw, axes=(1, 2)) * mask[None, :, :, None, None], axes=(1, 2))
            exit_bt = exitw
            exitw = exit_bt * ramp_a
        Psi = shift_fft2(exitw, axes=(1, 2)) * mask_s[None, :, :, None, None]
        I = np.abs(Psi) ** 2
        I = I[:, :, :, 0, 0] if coherent else np.mean(I, axis=(-1, -2))
        if return_patterns:
            patterns[tcs] = I
        if bg_square is not None:
            I = I + bg_square
        if algorithm == LSQ_SQRT:
            sse = sse + np.sum((meas - I) ** 2)
            Is, Ds = np.sqrt(I), np.sqrt(meas)
            dI = Is - Ds
            loss = loss + np.sum(dI ** 2)
            Is[Is == 0] = 1
            dI = dI / Is
        elif algorithm == ML:
            dI = I - meas
            sse = sse + np.sum(dI ** 2)
            zero = I == 0
            loss = loss + np.sum(I - meas * np.log(I + zero))
            I = I.copy()
            I[zero] = 1
            dI = dI / I
        elif algorithm == LSQ:
            # objective.py:547-551
            tterm = np.sum((I - meas) ** 2)
            loss = loss + tterm
            sse = sse + tterm
            dI = 2 * (I - meas)
        elif algorithm == LSQ_SQRT_2:
            # objective.py:552-559
            term = (meas - I) ** 2
            sse = sse + np.sum(term)
            nz = meas > 0.5
            loss = loss + 0.5 * np.sum(term[nz] / meas[nz])
            loss = loss + np.sum(I[~nz])
            dI = np.ones_like(I)
            dI[nz] = I[nz] / meas[nz] - 1
        elif compressed:
            # objective.py:560-585: coefficients of the virtual detectors
            c = np.sum(I[:, None, :, :] * masks[None], axis=(2, 3))
            if algorithm == LSQ_COMPRESSED:
                diff = c - meas
                tterm = np.sum(diff ** 2)
                loss = loss + tterm
                sse = sse + tterm
                dc = 2 * diff
            elif algorithm == GAUSS_COMPRESSED:
                ms, cs = np.sqrt(meas), np.sqrt(c)
                sse = sse + np.sum((c - meas) ** 2)
                dc = cs - ms
                loss = loss + np.sum(dc ** 2)
                cs[cs == 0] = 1
                dc = dc / cs
            else:
                dc = c - meas
                sse = sse + np.sum(dc ** 2)
                c[c == 0] = 1
                loss = loss + np.sum(c - meas * np.log(c))
                dc = dc / c
            dI = np.sum(masks[None] * dc[:, :, None, None], axis=1)
        else:
            raise ValueError(f"unsupported loss {algorithm}")
        if not coherent:
            dI = dI / (M * K)
        if step_static_background and bg is not None:
            bg_grad = bg_grad + np.sum(dI, 0) * 2 * bg * loss_weight
        dWave = dI[:, :, :, None, None] * Psi * (loss_weight * N * N)
        g = ifft2_ishift(dWave, axes=(1, 2))
        if tilt_after:
            g = g * np.conj(ramp_a)
            if step_tilts:
                gk = -2 * g * np.conj(exit_bt)
                dl = np.sum(np.real(gk[:, None] * yxg[None, :, :, :, None, None]), axis=(2, 3, 4, 5))
                if single_tilt:
                    tilt_grad[0, 4:] += np.sum(dl, 0)
                else:
                    tilt_grad[tcs, 4:] = dl
            if S == 1:
                g = ifft2(fft2(g, axes=(1, 2)) * mask[None, :, :, None, None], axes=(1, 2))
        if tilt_inside_grad:
            G, interm, tg = multislice_grads(g, waves, chop, props, mask, edge, step_obj, spec, qx[None], qy[None],
                                             dists)
            if single_tilt:
                tilt_grad[0, 2:4] += np.sum(tg, 0)
            else:
                tilt_grad[tcs, 2:4] += tg
        else:
            G, interm = multislice_grads(g, waves, chop, props, mask, edge, step_obj)
        if step_obj:
            scatteradd(obj_grad, my, mx, G)
        Fg = None
        if shift_flag:
            Fg = fft2(interm, axes=(1, 2)) * np.conj(E)[..., None]
            interm = ifft2(Fg, axes=(1, 2))
            if step_pos:
                smg = np.conj(FPs) * Fg
                sh = -2 / (N * N)
                yx = np.stack((shy, shx), axis=0)  # (2, N, N)
                pos_grad[tcs] = sh * np.sum(np.real(yx[None, :, :, :, None] * smg[:, None]), axis=(2, 3, 4))
        if tilt_before:
            interm = interm * np.conj(ramp_b)[..., None]
            if step_tilts:
                gk = -2 * interm * np.conj(probe_bt)
                dl = np.sum(np.real(gk[:, None] * yxg[None, :, :, :, None]), axis=(2, 3, 4))
                if single_tilt:
                    tilt_grad[0, :2] += np.sum(dl, 0)
                else:
                    tilt_grad[tcs, :2] = dl
        if ab is not None:
            Fg = fft2(interm, axes=(1, 2)) if Fg is None else Fg
            Fg = Fg * np.conj(plates[marker[tcs]])[..., None]
            interm = ifft2(Fg, axes=(1, 2))
            if step_aberrations:
                defgr = (2 / (N * N)) * np.sum(np.real((Fg * np.conj(FPb))[:, None] *
                                                       np.conj(polys)[None, :, :, :, None]), axis=(2, 3, 4))
                np.add.at(ab_grad, marker[tcs], defgr)
        if bc is not None:
            if step_beam_current:
                bc_grad[tcs] = 2 * np.sign(thisbc) * np.sum(np.real(interm * np.conj(probe[None])), axis=(1, 2, 3))
            interm = interm * thisbc[:, None, None, None]
        if step_probe:
            probe_grad += np.sum(interm, axis=0)
    loss = loss * loss_weight
    if loss != loss:
        raise ValueError("loss is NaN")
    if step_probe:
        probe_grad = ifft2(fft2(probe_grad, axes=(0, 1)) * mask[:, :, None], axes=(0, 1))
    out = (float(loss), float(sse), obj_grad, probe_grad, pos_grad)
    if return_patterns:
        out = out + (patterns,)
    if bg is not None:
        out = out + (bg_grad,)
    if bc is not None:
        out = out + (bc_grad,)
    if ab is not None:
        out = out + (ab_grad,)
    if tl is not None:
        out = out + (tilt_grad,)
    return out


def full_l1_constraint(obj, abs_norm_weight, phase_norm_weight):
    """pypty/objective.py:1286-1320 (compute_full_l1_constraint with return_direction):
    reg = phase_w sum |angle O| + abs_w sum |log |O||, grad = 0.5 (abs_w sign(log |O|) +
    i phase_w sign(angle O)) / conj(O); a weight <= 0 contributes nothing."""
    reg = 0.0
    grad = np.zeros_like(obj)
    if phase_norm_weight > 0:
        ph = np.angle(obj)
        reg += phase_norm_weight * np.sum(np.abs(ph))
        grad = grad + 1j * phase_norm_weight * np.sign(ph)
    if abs_norm_weight > 0:
        la = np.log(np.abs(obj))
        reg += abs_norm_weight * np.sum(np.abs(la))
        grad = grad + abs_norm_weight * np.sign(la)
    return float(reg), (0.5 * grad / np.conj(obj)).astype(obj.dtype)


def complex_grad_to_phase_grad(grad, array):
    """pypty/utils.py:260-276: dL/dphase = 2 Re(-i conj(z) dL/dz*) for z = |z| exp(i phase)."""
    return 2 * np.real(-1j * np.conj(array) * grad)


def preprocess_dataset(dataset, data_shift_vector=(0, 0), data_bin=1, data_pad=0, data_multiplier=1.0,
                       upsample_pattern=1):
    """pypty/utils.py:626-651 as objective.py:400-405 calls it per batch (not compressed, force_pad):
    np.roll along y then x, each followed by zeroing the edge band the reference zeroes (rows [-sy:]
    for sy > 0, [:-sy] for sy < 0); block sums of data_bin x data_bin pixels (the reference's loop
    indexes an undefined name here, the evident intent is used); times data_multiplier; zero padding
    of data_pad // upsample_pattern on each side."""
    d = np.array(dataset, dtype=np.float64)
    sy, sx = int(data_shift_vector[0]), int(data_shift_vector[1])
    if sy != 0:
        d = np.roll(d, sy, axis=1)
        if sy > 0:
            d[:, -sy:, :] = 0
        else:
            d[:, :-sy, :] = 0
    if sx != 0:
        d = np.roll(d, sx, axis=2)
        if sx > 0:
            d[:, :, -sx:] = 0
        else:
            d[:, :, :-sx] = 0
    if data_bin > 1:
        d2 = np.zeros((d.shape[0], d.shape[1] // data_bin, d.shape[2] // data_bin))
        for i1 in range(data_bin):
            for j1 in range(data_bin):
                d2 += d[:, i1::data_bin, j1::data_bin]
        d = d2
    if data_multiplier != 1:
        d = d * data_multiplier
    pad = int(data_pad) // int(upsample_pattern)
    return np.pad(d, [[0, 0], [pad, pad], [pad, pad]])


# --------------------------------------------------------------- L-BFGS
def wolfe_1(value, new_value, d_value, step, wolfe_c1=0.5):
    """pypty/utils.py:326-348: Armijo condition."""
    return new_value <= value + wolfe_c1 * step * d_value


def wolfe_2(d_value, new_d_value, wolfe_c2=0.9):
    """pypty/utils.py:350-368: curvature condition."""
    return -1 * new_d_value <= -1 * d_value * wolfe_c2


def cdot(a, b):
    """Real inner product used by bfgs_update: 2*Re<a, b> for complex blocks, <a, b> for real ones
    (iterative.py:597, 652, 683-690)."""
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        return 2.0 * float(np.sum(np.real(a * np.conj(b))))
    return float(np.sum(a * b))


def two_loop(grads, hist_s, hist_y, rhos, gammas):
    """L-BFGS two-loop recursion over a list of variable blocks, iterative.py:573-681.

    grads: list of gradient blocks; hist_s / hist_y: lists (oldest first) of lists of blocks;
    gammas: per-block initial scaling.  Returns the search direction blocks (already negated).
    """
    q = [np.array(g, copy=True) for g in grads]
    alphas = []
    for i in range(len(rhos) - 1, -1, -1):
        a = rhos[i] * sum(cdot(s, u) for s, u in zip(hist_s[i], q))
        alphas.append(a)
        q = [u - a * y for u, y in zip(q, hist_y[i])]
    alphas = alphas[::-1]
    q = [u * g if (g != 0 and g == g) else u for u, g in zip(q, gammas)]
    for i in range(len(rhos)):
        beta = alphas[i] - rhos[i] * sum(cdot(y, u) for y, u in zip(hist_y[i], q))
        q = [u + beta * s for u, s in zip(q, hist_s[i])]
    return [-u for u in q]
