"""GPU parity of the gfx950 kernels against the CPU oracle (complex64, 1e-4 relative).

Each case prints its parameters before touching the GPU.  Errors are
normwise relative, ||gpu - oracle|| / ||oracle||; the error against the FP64
oracle is printed beside the complex64 one.
"""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, oracle_pass, reference_kwargs, rel  # noqa: F401

TOL = 1e-4

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128])
@pytest.mark.parametrize("mode", ["fft2", "ifft2", "shift_fft2", "ifft2_ishift"])
def test_fft_modes_match_numpy(n, mode):
    torch = _torch()
    from pypty_amd import fft as pfft
    from oracle import ms_oracle as orc

    print(f"[fft] n={n} mode={mode}", flush=True)
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((3, n, n, 2)) + 1j * rng.standard_normal((3, n, n, 2))).astype(np.complex64)
    ref = getattr(orc, mode)(x.astype(np.complex128), axes=(1, 2))
    got = getattr(pfft, mode)(torch.from_numpy(x).cuda(), axes=(1, 2)).cpu().numpy()
    assert rel(got, ref) < 2e-6


@pytest.mark.parametrize("case", [
    dict(n=32, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt"),
    dict(n=32, S=1, K=1, M=1, shift=False, algorithm="lsq_sqrt"),
    dict(n=32, S=2, K=2, M=3, shift=True, algorithm="ml"),
    dict(n=64, S=4, K=1, M=1, shift=True, algorithm="lsq_sqrt", ny=128, nx=128, P=10),
    dict(n=48, S=2, K=1, M=2, shift=False, algorithm="ml", ny=96, nx=96),
    dict(n=96, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt", ny=160, nx=160, P=6),
    dict(n=128, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt", ny=192, nx=192, P=6),
    dict(n=64, S=2, K=1, M=3, shift=True, algorithm="lsq", ny=96, nx=96, P=7),
    dict(n=32, S=3, K=2, M=1, shift=True, algorithm="lsq_sqrt_2", P=9),
    dict(n=192, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt", ny=256, nx=256, P=5),
    dict(n=256, S=2, K=1, M=2, shift=True, algorithm="ml", ny=320, nx=320, P=4),
    dict(n=128, S=2, K=2, M=1, shift=True, algorithm="ml", ny=160, nx=160, P=5),
    dict(n=64, S=3, K=2, M=2, shift=True, algorithm="lsq_sqrt_2", ny=128, nx=128, P=6),
    dict(n=192, S=4, K=2, M=3, shift=True, algorithm="ml", ny=240, nx=240, P=5),
    dict(n=256, S=3, K=1, M=1, shift=False, algorithm="lsq", ny=300, nx=300, P=4),
    dict(n=96, S=2, K=2, M=3, shift=True, algorithm="ml", ny=144, nx=144, P=6),
    dict(n=48, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt", ny=96, nx=96, P=8),
    dict(n=16, S=3, K=1, M=2, shift=True, algorithm="lsq_sqrt", ny=48, nx=48, P=9),
])
def test_loss_and_direction_matches_oracle(case):
    torch = _torch()
    from pypty_amd import objective

    print(f"[loss_and_direction] {case}", flush=True)
    pr = make_problem(**case)
    ref = oracle_pass(pr, np.complex64)
    ref64 = oracle_pass(pr, np.complex128)
    # the case must be well conditioned: FP32 and FP64 arithmetic of the same
    # algorithm agree far inside the tolerance (with a single slice the ml
    # loss's 1/I makes the gradients depend on round-off by 6e-4)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    dev = torch.device("cuda", 0)
    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=int(pr["step_pos"]), this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"])
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    loss, sse, og, pg, posg = out[0], out[1], out[2].cpu().numpy(), out[3].cpu().numpy(), out[4].cpu().numpy()
    errs = dict(loss=abs(loss - ref[0]) / abs(ref[0]), sse=abs(sse - ref[1]) / abs(ref[1]),
                obj=rel(og, ref[2]), probe=rel(pg, ref[3]))
    errs64 = dict(loss=abs(loss - ref64[0]) / abs(ref64[0]), obj=rel(og, ref64[2]), probe=rel(pg, ref64[3]))
    if pr["step_pos"]:
        errs["pos"] = rel(posg, ref[4])
        errs64["pos"] = rel(posg, ref64[4])
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    print("  vs FP64 oracle:     ", {k: f"{v:.2e}" for k, v in errs64.items()})
    for k, v in errs.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("shape", [(192, 192), (256, 256), (96, 128), (512, 512), (384, 1536),
                                   (160, 160), (350, 210), (143, 242), (2000, 35), (94, 158), (97, 202)])
@pytest.mark.parametrize("mode", ["fft2", "ifft2", "shift_fft2", "ifft2_ishift"])
def test_plane_fft_modes_match_numpy(shape, mode):
    torch = _torch()
    from pypty_amd import fft as pfft
    from oracle import ms_oracle as orc

    print(f"[plane fft] shape={shape} mode={mode}", flush=True)
    rng = np.random.default_rng(shape[0] + shape[1])
    x = (rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)).astype(np.complex64)
    ref = getattr(orc, mode)(x.astype(np.complex128), axes=(1, 2))
    got = getattr(pfft, mode)(torch.from_numpy(x).cuda(), axes=(1, 2)).cpu().numpy()
    assert rel(got, ref) < 3e-6


@pytest.mark.parametrize("shape,cutoff,rolloff", [((512, 512, 3, 1), 2 / 3 - 0.005, 0.0),
                                                  ((256, 384, 2, 2), 0.66, 0.0),
                                                  ((1536, 1536, 1, 1), 2 / 3 - 0.005, 0.0),
                                                  ((192, 192, 1, 1), 0.5, 0.1),
                                                  ((160, 160, 1, 1), 2 / 3 - 0.005, 0.0),
                                                  ((350, 210, 2, 1), 0.6, 0.05),
                                                  ((94, 94, 1, 1), 2 / 3 - 0.005, 0.0),
                                                  ((97, 158, 2, 1), 0.5, 0.0)])
def test_fourier_clean_matches_oracle(shape, cutoff, rolloff):
    torch = _torch()
    from pypty_amd import utils as putils
    from oracle import ms_oracle as orc

    print(f"[fourier_clean] shape={shape} cutoff={cutoff} rolloff={rolloff}", flush=True)
    rng = np.random.default_rng(shape[0])
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = orc.fourier_clean(x.astype(np.complex128), cutoff=cutoff, rolloff=rolloff, fdtype=np.float64)
    got = putils.fourier_clean(torch.from_numpy(x).cuda(), cutoff=cutoff, rolloff=rolloff).cpu().numpy()
    assert rel(got, ref) < 3e-6
    # idempotent
    again = putils.fourier_clean(torch.from_numpy(got).cuda(), cutoff=cutoff, rolloff=0.0).cpu().numpy()
    if rolloff == 0.0:
        assert rel(again, got) < 3e-6


def test_lbfgs_vector_kernels():
    torch = _torch()
    from pypty_amd import lbfgs

    print("[lbfgs vectors]", flush=True)
    rng = np.random.default_rng(5)
    a = [(rng.standard_normal(100003) + 1j * rng.standard_normal(100003)).astype(np.complex64) for _ in range(3)]
    b = [(rng.standard_normal(100003) + 1j * rng.standard_normal(100003)).astype(np.complex64) for _ in range(3)]
    ta = [torch.from_numpy(v).cuda() for v in a]
    tb = [torch.from_numpy(v).cuda() for v in b]
    got = lbfgs.dots(list(zip(ta, tb)))
    ref = [float(np.sum(np.real(x.astype(np.complex128) * np.conj(y.astype(np.complex128))))) for x, y in zip(a, b)]
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-6)
    y = ta[0].clone()
    lbfgs.axpby(y, tb[0], 0.5, -2.0)
    assert rel(y.cpu().numpy(), 0.5 * a[0] - 2.0 * b[0]) < 1e-6
    # real float32 vectors
    r1 = rng.standard_normal(1001).astype(np.float32)
    r2 = rng.standard_normal(1001).astype(np.float32)
    got = lbfgs.dots([(torch.from_numpy(r1).cuda(), torch.from_numpy(r2).cuda())])
    assert np.isclose(got[0], float(np.dot(r1.astype(np.float64), r2.astype(np.float64))), rtol=1e-9)


@pytest.mark.parametrize("shape", [(32, 32, 2), (64, 64, 2, 1), (16, 16), (48, 48, 3)])
def test_fourier_clean_small_planes(shape):
    torch = _torch()
    from pypty_amd import utils as putils
    from oracle import ms_oracle as orc

    print(f"[fourier_clean small] shape={shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    cut = float(np.float32(2 / 3) - np.float32(0.005))
    ref = orc.fourier_clean(x.astype(np.complex128), cutoff=cut, fdtype=np.float64)
    got = putils.fourier_clean(torch.from_numpy(x).cuda(), cutoff=cut).cpu().numpy()
    print(f"  rel err {rel(got, ref):.2e}", flush=True)
    assert got.shape == x.shape
    assert rel(got, ref) < 3e-6


def test_axpby_special_cases():
    torch = _torch()
    from pypty_amd import lbfgs

    print("[axpby special cases]", flush=True)
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((64, 64, 2, 1)) + 1j * rng.standard_normal((64, 64, 2, 1))).astype(np.complex64)
    tx = torch.from_numpy(x).cuda()
    y = torch.full_like(tx, float("nan"))
    lbfgs.axpby(y, tx, 0.0, -1.5)                      # alpha = 0 never reads y
    assert rel(y.cpu().numpy(), -1.5 * x) < 1e-7
    lbfgs.axpby(y, None, 2.0, 0.0)                     # x = NULL
    assert rel(y.cpu().numpy(), -3.0 * x) < 1e-7
    z = lbfgs.lincomb(tx, 0.25)
    assert rel(z.cpu().numpy(), 0.25 * x) < 1e-7
    r = rng.standard_normal((10, 2)).astype(np.float32)
    tr = torch.from_numpy(r).cuda()
    w = lbfgs.lincomb(tr, -2.0)
    assert rel(w.cpu().numpy(), -2.0 * r) < 1e-7
    (d,) = lbfgs.dots([(tx, tx)])
    assert np.isclose(d, float(np.sum(np.abs(x.astype(np.complex128)) ** 2)), rtol=1e-9)


def test_first_bfgs_direction_matches_oracle():
    """Cleaned gradients and the first search direction of an L-BFGS epoch (iterative.py:497-572)."""
    torch = _torch()
    from pypty_amd import objective, utils as putils, lbfgs
    from oracle import ms_oracle as orc

    print("[bfgs first direction]", flush=True)
    pr = make_problem(n=32, S=2, K=1, M=2, ny=64, nx=64, P=10, shift=True, seed=21, step_pos=False, noise=0.02)
    cut = float(np.float32(2 / 3) - np.float32(0.005))
    obj = orc.fourier_clean(pr["obj"], cutoff=cut).astype(np.complex64)
    probe = orc.fourier_clean(pr["probe"], cutoff=cut).astype(np.complex64)
    q = dict(pr, obj=obj, probe=probe)
    ref = oracle_pass(q, np.complex64)
    ref64 = oracle_pass(q, np.complex128)
    assert rel(ref[2], ref64[2]) < 1e-5  # well conditioned (a flat start object is not)
    og_ref = orc.fourier_clean(ref[2], cutoff=cut).astype(np.complex64)
    pg_ref = orc.fourier_clean(ref[3], cutoff=cut).astype(np.complex64)
    dev = torch.device("cuda", 0)
    tobj = putils.fourier_clean(torch.from_numpy(pr["obj"]).to(dev), cutoff=cut)
    tprobe = putils.fourier_clean(torch.from_numpy(pr["probe"]).to(dev), cutoff=cut)
    print(f"  cleaned obj {rel(tobj.cpu().numpy(), obj):.2e} probe {rel(tprobe.cpu().numpy(), probe):.2e}")
    kw = reference_kwargs(
        this_obj=tobj, full_probe=tprobe, this_pos_array=pr["pos"], this_pos_correction=pr["corr"],
        this_distances=np.array([pr["dz"]], np.float32), measured_array=pr["meas"], algorithm_type=pr["algorithm"],
        this_wavelength=pr["wavelength"], this_step_probe=True, this_step_obj=True, this_step_pos_correction=False,
        this_step_tilts=0, pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"])
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    og = putils.fourier_clean(out[2], cutoff=cut)
    pg = putils.fourier_clean(out[3], cutoff=cut)
    e = dict(loss=abs(out[0] - ref[0]) / ref[0], og=rel(og.cpu().numpy(), og_ref), pg=rel(pg.cpu().numpy(), pg_ref))
    u_obj = lbfgs.lincomb(og, -1.0)
    d_val = 2 * lbfgs.dots([(u_obj, u_obj)])[0]
    d_ref = 2 * float(np.sum(np.abs(og_ref.astype(np.complex128)) ** 2))
    e["d_val_obj"] = abs(d_val - d_ref) / d_ref
    print("  ", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    for k, v in e.items():
        assert v < 1e-4, (k, v)
