"""Zero columns of the band mask: wavefronts whose spectrum columns the mask (and so the propagator)
zeroes skip their column transforms.  The pass must give the oracle's loss and gradients whatever
share of the wavefronts skips: many (a small cutoff), the default share, or none (a cutoff that keeps
every column).  The host emulation below counts the skipping wavefronts from the same float32 mask
formula and column order, so each case states which regime it tests."""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _fftfreq_f(f, n):
    k = f if 2 * f < n else f - n
    return np.float32(k * (1.0 / (n * 0.2)))  # pixel 0.2 A as in make_problem


def _zero_columns(n, cutoff):
    f32 = np.float32
    max_x = _fftfreq_f(n // 2 - 1, n)
    rm = f32(max_x * f32(cutoff))
    r_max = f32(rm * rm)
    qs = np.array([_fftfreq_f(f, n) for f in range(n)], np.float32)
    q2 = (qs[None, :] * qs[None, :]).astype(np.float32) + (qs[:, None] * qs[:, None]).astype(np.float32)
    return ~(q2.astype(np.float32) <= r_max).any(axis=0)  # column fx zero for every fy


def _skipping_wavefronts(n, cutoff):
    """(skipping, all) wavefront column groups of the register tile (n <= 128) or wavefront-rounds
    of the slab tile (192, 256)."""
    T = 8 if n % 64 == 0 else 4
    R1, L = n // T, 64 // T
    CPT = R1 // T

    def col_freq(col):
        g = col // T
        return CPT * (col % T) + (g % CPT) + R1 * (g // CPT)

    z = _zero_columns(n, cutoff)
    lines = 64 // 8 if n > 128 else L
    groups = [range(w * lines, (w + 1) * lines) for w in range(n // lines)]
    return sum(all(z[col_freq(c)] for c in g) for g in groups), len(groups)


@pytest.mark.parametrize("n,cutoff,regime", [(128, 0.3, "most"), (128, 1.5, "none"), (64, 0.45, "most"),
                                             (192, 0.3, "most"), (192, 1.5, "none"), (96, 0.3, "some"),
                                             (256, 0.5, "most")])
def test_pass_matches_oracle_whatever_columns_skip(n, cutoff, regime):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    skip, total = _skipping_wavefronts(n, cutoff)
    print(f"[zero columns] n={n} cutoff={cutoff}: {skip} of {total} wavefront column groups skip", flush=True)
    if regime == "none":
        assert skip == 0
    elif regime == "most":
        assert skip >= total // 2
    else:
        assert skip > 0
    pr = make_problem(n=n, S=2, K=1, M=2, ny=n + 40, nx=n + 40, P=5, shift=True, seed=211 + n)
    args = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                algorithm=pr["algorithm"], step_pos=True, cutoff=cutoff, compute_batch=5)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 cdtype=np.complex64, **args)
    ref64 = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                   cdtype=np.complex128, **args)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3  # well conditioned
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=1, this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"],
        damping_cutoff_multislice=cutoff)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    errs = dict(loss=abs(out[0] - ref[0]) / abs(ref[0]), obj=rel(out[2].cpu().numpy(), ref[2]),
                probe=rel(out[3].cpu().numpy(), ref[3]), pos=rel(out[4].cpu().numpy(), ref[4]))
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (k, v)
