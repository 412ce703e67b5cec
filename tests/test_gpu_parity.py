"""GPU parity of the gfx950 kernels against the CPU oracle (complex64, 1e-4 relative).

Each case prints its parameters before touching the GPU.  Errors are
normwise relative, ||gpu - oracle|| / ||oracle||; the error against the FP64
oracle is printed beside the complex64 one.
"""
import numpy as np
import pytest

from tests._problems import call_loss_and_direction, make_problem, oracle_pass, reference_kwargs, rel

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
])
def test_loss_and_direction_matches_oracle(case):
    torch = _torch()
    from pypty_amd import objective

    print(f"[loss_and_direction] {case}", flush=True)
    pr = make_problem(**case)
    ref = oracle_pass(pr, np.complex64)
    ref64 = oracle_pass(pr, np.complex128)
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


@pytest.mark.parametrize("shape", [(192, 192), (256, 256), (96, 128), (512, 512), (384, 1536)])
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
                                                  ((192, 192, 1, 1), 0.5, 0.1)])
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
