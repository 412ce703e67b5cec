"""Static background (pypty/objective.py:359-361, 525-526, 588-589) on the GPU: the drop-in
loss_and_direction with a background array (or a scalar) gives the oracle's loss, object, probe and
position gradients and background gradient, for the four pattern losses, one and several modes, the
register tile and the slab tile, several batches over several lanes, and patterns streamed from host
memory (accumulating calls)."""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, reference_kwargs, rel

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = [
    # n, S, M, algorithm, step, scalar, streamed
    (64, 2, 1, "lsq_sqrt", True, False, False),
    (64, 3, 2, "ml", True, False, False),
    (96, 2, 2, "lsq", True, False, False),
    (128, 2, 3, "lsq_sqrt_2", True, False, False),
    (192, 2, 2, "lsq_sqrt", True, False, False),
    (32, 2, 2, "lsq_sqrt", False, False, False),
    (64, 2, 2, "lsq_sqrt", True, True, False),
    (64, 2, 2, "ml", True, False, True),
]


@pytest.mark.parametrize("n,S,M,algorithm,step,scalar,streamed", CASES)
def test_background_pass_matches_oracle(n, S, M, algorithm, step, scalar, streamed):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print(f"[background] n={n} S={S} M={M} {algorithm} step={step} scalar={scalar} streamed={streamed}", flush=True)
    P = 11
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=401 + n + M,
                      algorithm=algorithm)
    rng = np.random.default_rng(5)
    level = 0.3 * float(np.sqrt(pr["meas"].mean()))
    bg = np.float32(level) if scalar else (level * rng.uniform(0.5, 1.5, (n, n))).astype(np.float32)
    bg_full = np.full((n, n), bg, np.float32) if scalar else bg
    args = dict(dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                algorithm=algorithm, step_pos=True, compute_batch=4, loss_weight=0.7, static_background=bg_full,
                step_static_background=step)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 cdtype=np.complex64, **args)
    ref64 = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                   cdtype=np.complex128, **args)
    assert max(rel(ref[k], ref64[k]) for k in (2, 3, 5)) < TOL / 3  # well conditioned
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"] if streamed else torch.from_numpy(pr["meas"]).to(dev),
        load_one_by_one=streamed, algorithm_type=algorithm, this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=1, this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], compute_batch=4,
        this_loss_weight=0.7, static_background=float(bg) if scalar else bg, this_step_static_background=int(step))
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    errs = dict(loss=abs(out[0] - ref[0]) / abs(ref[0]), obj=rel(out[2].cpu().numpy(), ref[2]),
                probe=rel(out[3].cpu().numpy(), ref[3]), pos=rel(out[4].cpu().numpy(), ref[4]))
    if step:
        errs["background"] = rel(out[6].cpu().numpy(), ref[5])
        assert float(np.abs(ref[5]).max()) > 0
    elif scalar:
        assert out[6] == 0
    else:
        assert not out[6].any()  # zeros of the background's shape, as the reference returns
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (k, v)


def test_background_with_compressed_loss_raises():
    import torch

    from pypty_amd import objective

    pr = make_problem(n=32, S=2, M=1, ny=72, nx=72, P=4, seed=9, algorithm="lsq_compressed")
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type="lsq_compressed", this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=0, this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], masks=pr["masks"],
        static_background=np.ones((32, 32), np.float32), this_step_static_background=1)
    with pytest.raises(NotImplementedError):
        call_loss_and_direction(objective.loss_and_direction, **kw)


@pytest.mark.parametrize("abs_w,phase_w", [(0.3, 0.2), (0.0, 0.5), (0.4, 0.0), (-1.0, -1.0)])
def test_l1_constraint_matches_oracle(abs_w, phase_w):
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print(f"[l1] abs_w={abs_w} phase_w={phase_w}", flush=True)
    rng = np.random.default_rng(31)
    o = (rng.uniform(0.5, 1.5, (70, 90, 3, 2)) * np.exp(1j * rng.uniform(-3, 3, (70, 90, 3, 2)))).astype(np.complex64)
    reg_ref, g_ref = orc.full_l1_constraint(o.astype(np.complex128), abs_w, phase_w)
    reg, g = objective.compute_full_l1_constraint(torch.from_numpy(o).cuda(), abs_w, phase_w, None, True, False)
    if abs_w <= 0 and phase_w <= 0:
        assert reg == 0.0 and not g.any()
        return
    assert abs(reg - reg_ref) <= 1e-5 * abs(reg_ref)
    assert rel(g.cpu().numpy(), g_ref) < 1e-5


def test_drop_in_adds_l1_terms_to_loss_and_object_gradient():
    import torch

    from oracle import ms_oracle as orc
    from pypty_amd import objective

    print("[l1] drop-in with phase and abs weights", flush=True)
    n = 64
    pr = make_problem(n=n, S=2, K=1, M=2, ny=n + 40, nx=n + 40, P=7, shift=True, seed=77)
    ref = orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                 dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"],
                                 pixel_size_y=pr["px"], algorithm="lsq_sqrt", step_pos=True, compute_batch=4,
                                 loss_weight=0.5, cdtype=np.complex64)
    r_ph, g_ph = orc.full_l1_constraint(pr["obj"].astype(np.complex128), 0.0, 2.0)
    r_ab, g_ab = orc.full_l1_constraint(pr["obj"].astype(np.complex128), 3.0, 0.0)
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=torch.from_numpy(pr["meas"]).to(dev), algorithm_type="lsq_sqrt",
        this_wavelength=pr["wavelength"], this_step_probe=1, this_step_obj=1, this_step_pos_correction=1,
        this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"],
        compute_batch=4, this_loss_weight=0.5, phase_norm_weight=2.0, abs_norm_weight=3.0)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    want_loss = ref[0] + r_ph + r_ab
    e_loss = abs(out[0] - want_loss) / abs(want_loss)
    e_obj = rel(out[2].cpu().numpy(), ref[2] + g_ph + g_ab)
    print(f"  loss {e_loss:.2e} obj {e_obj:.2e}", flush=True)
    assert e_loss < 1e-4 and e_obj < 1e-4
    cc = out[9]
    assert len(cc) == 14 and abs(cc[4] - r_ph) <= 1e-5 * r_ph and abs(cc[5] - r_ab) <= 1e-5 * r_ab
    assert not np.any(np.delete(cc, [4, 5]))
