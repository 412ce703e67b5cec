"""The benchmark configurations at their real per-window shapes, against the oracle.

* multislice (128 px window, 10 slices, 5 probe modes, sub-pixel shifts):
  loss and gradients vs the complex64 oracle (1e-4), and the autograd
  gradient vs central finite differences (eps 1e-3) along a direction on 64
  random object pixels, both of the GPU loss and of the FP64 oracle's loss;
* thick sample (256 px, 40 slices, 10 modes): waves checkpointed every 8
  slices give the bit-identical pass of the full stash, and both match the
  oracle on 3 positions;
* ragged step (8 seeded datasets, windows 96/128/192, 4-16 slices, 1-5 modes,
  scans capped to 8x8 positions): the HIP-graph replay of the whole step
  equals each dataset's eager pass bit for bit, and the largest dataset and
  the range's extreme shape (192 px, 16 slices, 5 modes) match the oracle;
* an edge-slice multiplier > 1 (multislice_core.py:73-74, 170-172) in the
  fused pass, and the band limit of whole planes on plane waves at the
  cutoff radius (utils.py:129-137).
Every case prints its shape before it touches the GPU.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import call_loss_and_direction, make_problem, oracle_pass, reference_kwargs, rel

TOL = 1e-4

pytestmark = pytest.mark.gpu


def _dev():
    import torch

    return torch.device("cuda", 0)


def _loss_and_direction(pr, edge=1.0, batch=5, checkpoint_slices=0):
    import torch

    from pypty_amd import objective

    dev = _dev()
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=int(pr["step_pos"]), this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"],
        edge_slice_thickness_multiplier=edge, compute_batch=batch, checkpoint_slices=checkpoint_slices)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    return out[0], out[1], out[2].cpu().numpy(), out[3].cpu().numpy(), out[4].cpu().numpy()


def _oracle(pr, cdtype, edge=1.0):
    return orc.loss_and_direction(pr["obj"], pr["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"], dz=pr["dz"],
                                  wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"],
                                  algorithm=pr["algorithm"], step_obj=True, step_probe=True, step_pos=pr["step_pos"],
                                  cdtype=cdtype, compute_batch=5, edge=edge)


def _check(got, ref, ref64, label):
    errs = dict(loss=abs(got[0] - ref[0]) / abs(ref[0]), obj=rel(got[2], ref[2]), probe=rel(got[3], ref[3]))
    errs64 = dict(loss=abs(got[0] - ref64[0]) / abs(ref64[0]), obj=rel(got[2], ref64[2]), probe=rel(got[3], ref64[3]))
    if np.abs(ref[4]).max() > 0:
        errs["pos"] = rel(got[4], ref[4])
        errs64["pos"] = rel(got[4], ref64[4])
    print(f"  {label} vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    print(f"  {label} vs FP64 oracle:     ", {k: f"{v:.2e}" for k, v in errs64.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, (label, k, v)


def test_multislice_config_shape_matches_oracle():
    print("[config multislice] n=128 S=10 K=1 M=5 P=6 shift", flush=True)
    pr = make_problem(n=128, S=10, K=1, M=5, ny=200, nx=200, P=6, shift=True, seed=1001)
    ref, ref64 = _oracle(pr, np.complex64), _oracle(pr, np.complex128)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3  # well conditioned
    _check(_loss_and_direction(pr, batch=4), ref, ref64, "multislice")


def test_multislice_config_finite_differences_64_pixels():
    """Central differences, eps 1e-3, along a direction supported on 64 random object pixels, against
    the autograd object gradient at the multislice shape: of the FP64 oracle's loss (1e-4), and of
    the GPU loss itself (1e-1: its pixels are float32, and with eps 1e-3 the two evaluations differ
    by about 1e-6 of the loss, where float32 rounding of the 10^6 pixel terms is of the same order;
    measured 4e-2)."""
    import torch

    from pypty_amd.autograd import PassContext, multislice_loss
    from pypty_amd.objective import PassConfig

    print("[config multislice] finite differences n=128 S=10 M=5 P=6", flush=True)
    pr = make_problem(n=128, S=10, K=1, M=5, ny=200, nx=200, P=6, shift=True, seed=1002)
    dev = _dev()
    ny, nx, S, K = pr["obj"].shape
    cfg = PassConfig(n=128, n_slices=S, n_obj_modes=K, n_probe_modes=5, obj_ny=ny, obj_nx=nx, n_positions=6,
                     compute_batch=3, algorithm="lsq_sqrt", slice_distance=pr["dz"], wavelength=pr["wavelength"],
                     pixel_size_x=pr["px"], pixel_size_y=pr["px"])
    pctx = PassContext(cfg, torch.from_numpy(pr["pos"]).to(dev), torch.from_numpy(pr["meas"]).to(dev),
                       torch.from_numpy(pr["seq"].astype(np.int32)).to(dev))
    obj = torch.from_numpy(pr["obj"]).to(dev).requires_grad_(True)
    probe = torch.from_numpy(pr["probe"]).to(dev)
    corr = torch.from_numpy(pr["corr"]).to(dev)
    multislice_loss(obj, probe, corr, pctx).backward()
    grad = obj.grad.cpu().numpy()
    rng = np.random.default_rng(64)
    # 64 pixels inside the scanned windows
    covered = np.zeros((ny, nx), bool)
    for y0, x0 in pr["pos"]:
        covered[y0 + 16:y0 + 112, x0 + 16:x0 + 112] = True
    ys, xs = np.nonzero(covered)
    pick = rng.choice(len(ys), 64, replace=False)
    d = np.zeros(pr["obj"].shape, np.complex128)
    d[ys[pick], xs[pick], rng.integers(0, S, 64), 0] = np.exp(2j * np.pi * rng.uniform(size=64))
    an = float(np.real(np.sum(np.conj(grad) * d)))  # Re<grad, d>, grad = 2 dL/dz*
    eps = 1e-3
    td = torch.from_numpy(d.astype(np.complex64)).to(dev)
    with torch.no_grad():
        o0 = obj.detach()
        fd = (multislice_loss(o0 + eps * td, probe, corr, pctx).item() -
              multislice_loss(o0 - eps * td, probe, corr, pctx).item()) / (2 * eps)

    def L64(s):
        q = dict(pr)
        q["obj"] = pr["obj"].astype(np.complex128) + s * d
        q["probe"] = pr["probe"].astype(np.complex128)
        return orc.loss_and_direction(q["obj"], q["probe"], pr["pos"], pr["corr"], pr["meas"], pr["seq"],
                                      dz=pr["dz"], wavelength=pr["wavelength"], pixel_size_x=pr["px"],
                                      pixel_size_y=pr["px"], step_obj=False, step_probe=False, step_pos=False,
                                      cdtype=np.complex128, compute_batch=6)[0]

    fd64 = (L64(eps) - L64(-eps)) / (2 * eps)
    print(f"  autograd {an:.6e}  fd(GPU loss) {fd:.6e} rel {abs(fd - an) / abs(an):.2e}  "
          f"fd(FP64 oracle) {fd64:.6e} rel {abs(fd64 - an) / abs(an):.2e}", flush=True)
    assert abs(fd64 - an) <= 1e-4 * abs(an)
    assert abs(fd - an) <= 1e-1 * abs(an)


def test_thick_config_checkpoints_bit_identical_and_match_oracle():
    import torch

    from pypty_amd.objective import PassConfig, PtychoPass

    print("[config thick] n=256 S=40 K=1 M=10 P=3 k=8 vs k=0 and the oracle", flush=True)
    pr = make_problem(n=256, S=40, K=1, M=10, ny=300, nx=300, P=3, shift=True, seed=4001, dz=1.0)
    dev = _dev()
    ny, nx, S, K = pr["obj"].shape
    t = dict(obj=torch.from_numpy(pr["obj"]).to(dev),
             probe=torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev),
             pos=torch.from_numpy(pr["pos"]).to(dev), corr=torch.from_numpy(pr["corr"]).to(dev),
             meas=torch.from_numpy(pr["meas"]).to(dev), seq=torch.arange(3, dtype=torch.int32, device=dev))
    outs = {}
    for ck in (0, 8):
        cfg = PassConfig(n=256, n_slices=S, n_obj_modes=K, n_probe_modes=10, obj_ny=ny, obj_nx=nx, n_positions=3,
                         compute_batch=2, algorithm="lsq_sqrt", step_obj=True, step_probe=True, step_pos=True,
                         shift_active=True, slice_distance=pr["dz"], wavelength=pr["wavelength"],
                         pixel_size_x=pr["px"], pixel_size_y=pr["px"], checkpoint_slices=ck, lanes=2)
        og, pg, posg = torch.empty_like(t["obj"]), torch.empty_like(t["probe"]), torch.zeros_like(t["corr"])
        loss, sse = PtychoPass(cfg, dev).run(t["obj"], t["probe"], t["pos"], t["corr"], t["meas"], t["seq"],
                                             obj_grad=og, probe_grad=pg, pos_grad=posg)
        outs[ck] = (loss, sse, og, pg, posg)
    for a, b, name in zip(outs[0], outs[8], ("loss", "sse", "object", "probe", "positions")):
        if isinstance(a, float):
            assert a == b, name
        else:
            assert torch.equal(a, b), name
    got = (outs[8][0], outs[8][1], outs[8][2].cpu().numpy(), outs[8][3].permute(1, 2, 0).cpu().numpy(),
           outs[8][4].cpu().numpy())
    ref, ref64 = _oracle(pr, np.complex64), _oracle(pr, np.complex128)
    print(f"  oracle complex64 vs FP64: obj {rel(ref[2], ref64[2]):.2e} probe {rel(ref[3], ref64[3]):.2e}",
          flush=True)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    _check(got, ref, ref64, "thick k=8")


def test_ragged_step_graph_replay_matches_eager_and_oracle():
    import torch

    from pypty_amd import ragged, synth

    print("[config ragged] 8 datasets, scans capped at 8x8, graph replay vs eager", flush=True)
    dev = _dev()
    datasets = synth.ragged_datasets(dev, max_scan=8)
    step = ragged.RaggedStep(datasets, dev, batch=16, streams=2, lanes=1)
    for j in step.jobs:
        print(f"  dataset n={j.shape[0]} S={j.shape[1]} M={j.shape[2]} P={j.P}", flush=True)
    step.run()
    torch.cuda.synchronize()
    eager = [(j.loss_sse.clone(), j.og.clone(), j.pg.clone(), j.posg.clone()) for j in step.jobs]
    step.capture()
    for rep in range(3):
        for j in step.jobs:
            for t in (j.og, j.pg, j.posg, j.loss_sse):
                t.fill_(float("nan"))  # a replay must rewrite every output
        x = torch.randn(256, 256, device=dev)
        (x @ x).sum().item()  # eager work and a synchronisation between replays
        step.replay()
        torch.cuda.synchronize()
        for i, j in enumerate(step.jobs):
            for a, b, name in zip((j.loss_sse, j.og, j.pg, j.posg), eager[i], ("loss", "obj", "probe", "pos")):
                assert torch.equal(a, b), (rep, i, name)
    # the largest dataset (192 px, 7 slices, 5 modes) against the oracle on its first positions
    big = max(step.jobs, key=lambda j: j.shape[0] * j.shape[1] * j.shape[2])
    seq = np.arange(4)
    print(f"  oracle parity: n={big.shape[0]} S={big.shape[1]} M={big.shape[2]} positions {list(seq)}", flush=True)
    # a scattering object away from the truth: from a flat object the far field vanishes outside the
    # aperture and the gradients there are set by round-off; at the truth they nearly cancel
    ds = big.ds
    start = (np.abs(ds.obj) * np.exp(0.7j * np.angle(ds.obj))).astype(np.complex64)
    big.obj.copy_(torch.from_numpy(start).to(dev))
    big.posg.zero_()  # a pass assigns the gradients of its sequence's positions only (objective.py:658)
    loss, sse = big.run(sync=True, sequence=torch.from_numpy(seq.astype(np.int32)).to(dev))
    kw = dict(dz=ds.dz_A, wavelength=ds.wavelength_A, pixel_size_x=ds.pixel_A, pixel_size_y=ds.pixel_A,
              step_obj=True, step_probe=True, step_pos=True, compute_batch=4)
    ref = orc.loss_and_direction(start, ds.probe, ds.positions, ds.corrections, ds.measured, seq,
                                 cdtype=np.complex64, **kw)
    ref64 = orc.loss_and_direction(start, ds.probe, ds.positions, ds.corrections, ds.measured, seq,
                                   cdtype=np.complex128, **kw)
    print(f"  oracle complex64 vs FP64: obj {rel(ref[2], ref64[2]):.2e} probe {rel(ref[3], ref64[3]):.2e} "
          f"positions {rel(ref[4], ref64[4]):.2e}", flush=True)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    got = (loss, sse, big.og.cpu().numpy(), big.pg.permute(1, 2, 0).cpu().numpy(), big.posg.cpu().numpy())
    _check(got, ref, ref64, "ragged largest")


def test_ragged_extreme_shape_matches_oracle():
    print("[config ragged] extreme shape n=192 S=16 K=1 M=5 P=4", flush=True)
    pr = make_problem(n=192, S=16, K=1, M=5, ny=260, nx=260, P=4, shift=True, seed=2016)
    ref, ref64 = _oracle(pr, np.complex64), _oracle(pr, np.complex128)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    _check(_loss_and_direction(pr, batch=3), ref, ref64, "192/16/5")


@pytest.mark.parametrize("n,S,M,edge", [(64, 3, 2, 2.0), (128, 4, 1, 1.5), (192, 3, 2, 3.0), (32, 5, 3, 1.25)])
def test_edge_multiplier_in_the_pass(n, S, M, edge):
    print(f"[edge multiplier] n={n} S={S} M={M} e={edge}", flush=True)
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 48, nx=n + 48, P=5, shift=True, seed=77 + n)
    ref, ref64 = _oracle(pr, np.complex64, edge), _oracle(pr, np.complex128, edge)
    # the multiplier changes the result (the case would not see it otherwise)
    assert rel(_oracle(pr, np.complex64, 1.0)[2], ref[2]) > 1e-3
    _check(_loss_and_direction(pr, edge=edge), ref, ref64, f"edge {edge}")


@pytest.mark.parametrize("shape,cutoff", [((256, 384, 1, 1), 2 / 3 - 0.005), ((160, 96, 2, 1), 0.5)])
def test_fourier_clean_plane_waves_at_the_radius(shape, cutoff):
    """A plane wave survives the GPU band limit iff its normalised frequency is within 0.5*cutoff."""
    import torch

    from pypty_amd import utils as putils

    print(f"[fourier_clean] plane waves shape={shape} cutoff={cutoff}", flush=True)
    ny, nx = shape[:2]
    r2 = (0.5 * cutoff) ** 2
    y, x = np.mgrid[0:ny, 0:nx]
    rng = np.random.default_rng(ny)
    waves, inside = [], []
    # frequencies just inside and just outside the circle along random directions
    for _ in range(24):
        th = rng.uniform(0, 2 * np.pi)
        for scale in (0.97, 1.03):
            ky = int(round(scale * 0.5 * cutoff * np.sin(th) * ny))
            kx = int(round(scale * 0.5 * cutoff * np.cos(th) * nx))
            f2 = (ky / ny) ** 2 + (kx / nx) ** 2
            if abs(f2 - r2) < 1e-4 * r2:
                continue
            waves.append(np.exp(2j * np.pi * (ky * y / ny + kx * x / nx)))
            inside.append(f2 < r2)
    arr = np.stack(waves, axis=-1).astype(np.complex64).reshape(ny, nx, -1, 1)
    out = putils.fourier_clean(torch.from_numpy(arr).to(_dev()), cutoff=cutoff).cpu().numpy()
    for i, keep in enumerate(inside):
        if keep:
            assert rel(out[:, :, i, 0], arr[:, :, i, 0]) < 1e-5, i
        else:
            assert np.abs(out[:, :, i, 0]).max() < 1e-5, i
    assert 0 < sum(inside) < len(inside)
