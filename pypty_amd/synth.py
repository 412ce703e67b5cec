"""Seeded synthetic ptychography datasets (objects, probes, scans, patterns).

Objects, probes and scan grids are built on the host with numpy (setup, not
per-pattern work); diffraction patterns are simulated on the GPU by the same
forward kernels the reconstruction uses (``sim_patterns`` output of
``pty_loss_and_direction``) and Poisson counts are drawn on the device from a
seeded generator.  Conventions follow PyPty: electron wavelength from
pypty/utils.py:1721, defocus phase exp(-i*pi*df*lambda*q^2) from
pypty/utils.py:1525, positions split into an integer corner and a sub-pixel
correction as in pypty/utils.py:1742-1746, intensity = mean over modes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def electron_wavelength(acc_voltage_kv: float) -> float:
    """Relativistic electron wavelength in Angstrom (pypty/utils.py:1721)."""
    v = float(acc_voltage_kv)
    return float(np.float32(12.398 / np.sqrt((2 * 511.0 + v) * v)))


def _gauss_blur_planes(planes: np.ndarray, sigma_px: float) -> np.ndarray:
    ny, nx = planes.shape[-2:]
    fy = np.fft.fftfreq(ny)[:, None]
    fx = np.fft.fftfreq(nx)[None, :]
    g = np.exp(-2 * (np.pi * sigma_px) ** 2 * (fx ** 2 + fy ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(planes) * g))


def _splat(ny, nx, ys, xs):
    """Bilinear deposit of unit points at fractional pixel positions."""
    img = np.zeros((ny, nx))
    y0 = np.floor(ys).astype(int)
    x0 = np.floor(xs).astype(int)
    wy, wx = ys - y0, xs - x0
    for dy, wyy in ((0, 1 - wy), (1, wy)):
        for dx, wxx in ((0, 1 - wx), (1, wx)):
            np.add.at(img, ((y0 + dy) % ny, (x0 + dx) % nx), wyy * wxx)
    return img


def gaussian_blob_phase(ny, nx, pixel_A, ys_A, xs_A, sigma_A, peak_rad):
    """Sum of Gaussian blobs of height peak_rad at the given positions (Angstrom)."""
    img = _splat(ny, nx, np.asarray(ys_A) / pixel_A, np.asarray(xs_A) / pixel_A)
    s = sigma_A / pixel_A
    img = _gauss_blur_planes(img, s) * (2 * np.pi * s * s)
    return img * peak_rad


def hex_lattice_object(ny, nx, pixel_A=0.2, a_A=2.46, sigma_A=0.3, peak_rad=0.5, seed=1000):
    """Pure-phase single-slice object, blobs on a seeded hexagonal lattice: (ny, nx, 1, 1) complex64."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, np.pi / 3)
    origin = rng.uniform(0, a_A, 2)
    a1 = a_A * np.array([np.cos(theta), np.sin(theta)])
    a2 = a_A * np.array([np.cos(theta + np.pi / 3), np.sin(theta + np.pi / 3)])
    L = max(ny, nx) * pixel_A
    r = int(np.ceil(2 * L / a_A)) + 2
    i, j = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
    pts = origin + i.ravel()[:, None] * a1 + j.ravel()[:, None] * a2
    keep = (pts[:, 0] >= 0) & (pts[:, 0] < ny * pixel_A) & (pts[:, 1] >= 0) & (pts[:, 1] < nx * pixel_A)
    pts = pts[keep]
    phase = gaussian_blob_phase(ny, nx, pixel_A, pts[:, 0], pts[:, 1], sigma_A, peak_rad)
    return np.exp(1j * phase).astype(np.complex64)[:, :, None, None]


def random_atom_object(ny, nx, n_slices, pixel_A=0.2, dz_A=2.0, density=0.1, sigma_A=0.3, peak_rad=0.5,
                       seed=1001):
    """Pure-phase multislice object from seeded atoms at random xyz: (ny, nx, S, 1) complex64."""
    rng = np.random.default_rng(seed)
    vol = ny * pixel_A * nx * pixel_A * n_slices * dz_A
    n_atoms = rng.poisson(density * vol)
    ys = rng.uniform(0, ny * pixel_A, n_atoms)
    xs = rng.uniform(0, nx * pixel_A, n_atoms)
    zs = rng.integers(0, n_slices, n_atoms)
    obj = np.empty((ny, nx, n_slices, 1), np.complex64)
    for s in range(n_slices):
        sel = zs == s
        ph = gaussian_blob_phase(ny, nx, pixel_A, ys[sel], xs[sel], sigma_A, peak_rad)
        obj[:, :, s, 0] = np.exp(1j * ph)
    return obj


def make_probe(n, pixel_A, acc_voltage_kv=300.0, aperture_mrad=20.0, defocus_A=100.0, n_modes=1,
               defocus_spread_A=20.0, seed=1101):
    """Aperture-limited, defocused probe (n, n, M) complex64, centred in the window
    (fftshift of the inverse transform, as pypty/initialize.py:977 builds its probe).

    Mode m carries an extra seeded defocus offset ~ N(0, defocus_spread_A)
    (mode 0 none) and a weight 1/(m+1).  Normalised so that sum |probe|^2 = 1 over all modes.
    """
    rng = np.random.default_rng(seed)
    lam = electron_wavelength(acc_voltage_kv)
    q = np.fft.fftfreq(n, pixel_A)
    qx, qy = np.meshgrid(q, q, indexing="xy")
    q2 = qx ** 2 + qy ** 2
    ap = (np.sqrt(q2) * lam <= aperture_mrad * 1e-3).astype(np.float64)
    probe = np.empty((n, n, n_modes), np.complex64)
    for m in range(n_modes):
        df = defocus_A + (rng.normal(0.0, defocus_spread_A) if m > 0 else 0.0)
        spec = ap * np.exp(-1j * np.pi * df * lam * q2) / (m + 1)
        probe[:, :, m] = np.fft.fftshift(np.fft.ifft2(spec))
    probe /= np.sqrt(np.sum(np.abs(probe) ** 2))
    return probe.astype(np.complex64)


def raster_scan(n_scan_y, n_scan_x, step_px, origin_px=(0.0, 0.0), jitter_px=0.0, seed=1201):
    """Raster scan positions (P, 2) in pixels (y, x) with optional seeded uniform jitter."""
    rng = np.random.default_rng(seed)
    iy, ix = np.meshgrid(np.arange(n_scan_y), np.arange(n_scan_x), indexing="ij")
    pos = np.stack([iy.ravel() * step_px + origin_px[0], ix.ravel() * step_px + origin_px[1]], axis=1)
    if jitter_px:
        pos = pos + rng.uniform(-jitter_px, jitter_px, pos.shape)
    return pos.astype(np.float64)


def split_positions(pos_px):
    """Integer corner + sub-pixel correction (pypty/utils.py:1742-1746)."""
    ints = np.round(pos_px).astype(np.int32)
    corr = (pos_px - ints).astype(np.float32)
    return ints, corr


@dataclass
class Dataset:
    obj: np.ndarray          # (Ny, Nx, S, K) true object
    probe: np.ndarray        # (N, N, M)
    positions: np.ndarray    # (P, 2) int32
    corrections: np.ndarray  # (P, 2) float32
    measured: np.ndarray     # (P, N, N) float32, DC centred
    pixel_A: float
    dz_A: float
    wavelength_A: float
    dose: float


def simulate(obj, probe, positions, corrections, pixel_A, dz_A, wavelength_A, dose, seed, device=None,
             compute_batch=256):
    """Noise-free patterns from the GPU forward model scaled to `dose` e- per pattern, then Poisson."""
    import torch

    from .objective import PassConfig, PtychoPass

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    ny, nx, S, K = obj.shape
    n, _, M = probe.shape
    P = positions.shape[0]
    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx, n_positions=P,
                     compute_batch=min(compute_batch, P), step_obj=False, step_probe=False, step_pos=False,
                     shift_active=bool(np.any(corrections != 0)), slice_distance=dz_A, wavelength=wavelength_A,
                     pixel_size_x=pixel_A, pixel_size_y=pixel_A, lanes=1)  # one stream: profiles see whole launches
    eng = PtychoPass(cfg, dev)
    sim = torch.zeros((P, n, n), dtype=torch.float32, device=dev)
    eng.run(torch.from_numpy(obj).to(dev), torch.from_numpy(np.ascontiguousarray(probe.transpose(2, 0, 1))).to(dev),
            torch.from_numpy(positions).to(dev), torch.from_numpy(corrections).to(dev),
            torch.zeros((P, n, n), dtype=torch.float32, device=dev), torch.arange(P, dtype=torch.int32, device=dev),
            sim_patterns=sim)
    # scale to the dose and draw Poisson counts on the device (seeded generator)
    I = sim.double()
    I *= dose / torch.clamp(I.sum(dim=(1, 2), keepdim=True), min=1e-30)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    return torch.poisson(I, generator=gen).float().cpu().numpy()


def config1(device=None, n_scan=64, seed=1001):
    """Multislice config: 512x512x10 object (dz 2 A, 0.2 A px), 128 px probe with 5 modes,
    64x64 raster (step 6 px, +-0.5 px jitter), 300 kV, 20 mrad, 100 A defocus, 1e4 e- per pattern."""
    px, dz = 0.2, 2.0
    obj = random_atom_object(512, 512, 10, px, dz, 0.1, 0.3, 0.5, seed)
    probe = make_probe(128, px, 300.0, 20.0, 100.0, 5, 20.0, seed + 100)
    pos = raster_scan(n_scan, n_scan, 6.0, origin_px=(1.0, 1.0), jitter_px=0.5, seed=seed + 200)
    ints, corr = split_positions(pos)
    lam = electron_wavelength(300.0)
    meas = simulate(obj, probe, ints, corr, px, dz, lam, 1e4, seed + 300, device)
    return Dataset(obj, probe, ints, corr, meas, px, dz, lam, 1e4)


def config0(device=None, seed=1000):
    """Single-slice config: 256x256 hexagonal-lattice phase object, 64 px probe, 16x16 raster, step 2 px."""
    px, dz = 0.2, 2.0
    obj = hex_lattice_object(256, 256, px, 2.46, 0.3, 0.5, seed)
    probe = make_probe(64, px, 300.0, 20.0, 100.0, 1, 0.0, seed + 100)
    pos = raster_scan(16, 16, 2.0, origin_px=(96.0, 96.0), seed=seed + 200)
    ints, corr = split_positions(pos)
    lam = electron_wavelength(300.0)
    meas = simulate(obj, probe, ints, corr, px, dz, lam, 1e4, seed + 300, device)
    return Dataset(obj, probe, ints, corr, meas, px, dz, lam, 1e4)


def ragged_datasets(device=None, seed=2000, n_datasets=8, max_scan=None):
    """Eight independent seeded datasets of mixed shapes: windows {96, 128, 192} px, 4-16 slices,
    1-5 probe modes, raster scans 32x32 .. 96x96 (step 6 px), dose log-uniform in 1e3 .. 1e5 e-.
    ``max_scan`` caps the scan side (same shapes, fewer positions, smaller objects)."""
    out = []
    px, dz = 0.2, 2.0
    lam = electron_wavelength(300.0)
    for i in range(n_datasets):
        rng = np.random.default_rng(seed + i)
        n = int(rng.choice([96, 128, 192]))
        S = int(rng.integers(4, 17))
        M = int(rng.integers(1, 6))
        scan = int(rng.integers(32, 97))
        if max_scan is not None:
            scan = min(scan, int(max_scan))
        dose = float(10 ** rng.uniform(3, 5))
        step = 6.0
        side = int(np.ceil(((scan - 1) * step + n + 4) / 32.0) * 32)
        obj = random_atom_object(side, side, S, px, dz, 0.1, 0.3, 0.5, seed + 100 + i)
        probe = make_probe(n, px, 300.0, 20.0, 100.0, M, 20.0, seed + 200 + i)
        pos = raster_scan(scan, scan, step, origin_px=(1.0, 1.0), jitter_px=0.5, seed=seed + 300 + i)
        ints, corr = split_positions(pos)
        meas = simulate(obj, probe, ints, corr, px, dz, lam, dose, seed + 400 + i, device)
        out.append(Dataset(obj, probe, ints, corr, meas, px, dz, lam, dose))
    return out


def config4(device=None, seed=4000, n_scan=128):
    """Thick-sample config: 1536x1536x40 object (1 A slices), 256 px probe with 10 modes,
    128x128 raster (step 10 px), 1e4 e- per pattern."""
    px, dz = 0.2, 1.0
    obj = random_atom_object(1536, 1536, 40, px, dz, 0.1, 0.3, 0.5, seed)
    probe = make_probe(256, px, 300.0, 20.0, 100.0, 10, 20.0, seed + 100)
    pos = raster_scan(n_scan, n_scan, 10.0, origin_px=(1.0, 1.0), jitter_px=0.5, seed=seed + 200)
    ints, corr = split_positions(pos)
    lam = electron_wavelength(300.0)
    meas = simulate(obj, probe, ints, corr, px, dz, lam, 1e4, seed + 300, device)
    return Dataset(obj, probe, ints, corr, meas, px, dz, lam, 1e4)
