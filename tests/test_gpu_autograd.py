"""torch.autograd.Function around the fused pass: parity with the oracle and central finite differences."""
import numpy as np
import pytest

from tests._problems import make_problem, oracle_pass, rel

pytestmark = pytest.mark.gpu


def _ctx(pr, dev, batch=5):
    import torch

    from pypty_amd.autograd import PassContext
    from pypty_amd.objective import PassConfig

    ny, nx, S, K = pr["obj"].shape
    n, _, M = pr["probe"].shape
    P = pr["pos"].shape[0]
    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx, n_positions=P,
                     compute_batch=batch, algorithm=pr["algorithm"], slice_distance=pr["dz"],
                     wavelength=pr["wavelength"], pixel_size_x=pr["px"], pixel_size_y=pr["px"])
    return PassContext(cfg, torch.from_numpy(pr["pos"]).to(dev), torch.from_numpy(pr["meas"]).to(dev),
                       torch.from_numpy(pr["seq"].astype(np.int32)).to(dev))


def test_autograd_gradients_match_oracle():
    import torch

    from pypty_amd.autograd import multislice_loss

    print("[autograd] parity", flush=True)
    pr = make_problem(n=32, S=3, K=1, M=2, shift=True, seed=41)
    dev = torch.device("cuda", 0)
    obj = torch.from_numpy(pr["obj"]).to(dev).requires_grad_(True)
    probe = torch.from_numpy(pr["probe"]).to(dev).requires_grad_(True)
    corr = torch.from_numpy(pr["corr"]).to(dev).requires_grad_(True)
    L = multislice_loss(obj, probe, corr, _ctx(pr, dev))
    L.backward()
    ref = oracle_pass(pr, np.complex64)
    e = dict(loss=abs(L.item() - ref[0]) / ref[0], obj=rel(obj.grad.cpu().numpy(), 2 * ref[2]),
             probe=rel(probe.grad.cpu().numpy(), 2 * ref[3]), pos=rel(corr.grad.cpu().numpy(), ref[4]))
    print("  ", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    for k, v in e.items():
        assert v < 1e-4, (k, v)


def test_autograd_central_finite_differences_64_pixels():
    """Central differences (eps 1e-3) along a random direction supported on 64 random object pixels,
    and along a random probe and position direction, against the autograd gradient."""
    import torch

    from pypty_amd.autograd import multislice_loss

    print("[autograd] finite differences", flush=True)
    pr = make_problem(n=64, S=3, K=1, M=2, ny=128, nx=128, P=12, shift=True, seed=42)
    dev = torch.device("cuda", 0)
    pctx = _ctx(pr, dev, batch=6)
    obj = torch.from_numpy(pr["obj"]).to(dev).requires_grad_(True)
    probe = torch.from_numpy(pr["probe"]).to(dev).requires_grad_(True)
    corr = torch.from_numpy(pr["corr"]).to(dev).requires_grad_(True)
    multislice_loss(obj, probe, corr, pctx).backward()
    rng = np.random.default_rng(7)
    eps = 1e-3

    def loss_at(o, p, c):
        with torch.no_grad():
            return multislice_loss(o, p, c, pctx).item()

    # object: 64 random pixels inside the scanned area, random complex directions
    d = np.zeros(pr["obj"].shape, np.complex64)
    ys, xs = rng.integers(16, 112, 64), rng.integers(16, 112, 64)
    ss = rng.integers(0, 3, 64)
    d[ys, xs, ss, 0] = np.exp(2j * np.pi * rng.uniform(size=64))
    td = torch.from_numpy(d).to(dev)
    o0 = obj.detach()
    fd = (loss_at(o0 + eps * td, probe.detach(), corr.detach()) - loss_at(o0 - eps * td, probe.detach(),
                                                                           corr.detach())) / (2 * eps)
    an = float(torch.real(torch.sum(torch.conj(obj.grad) * td)).item())  # Re<grad, d> with grad = 2 dL/dz*
    print(f"  object: fd {fd:.6e} autograd {an:.6e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
    assert abs(fd - an) <= 1e-2 * abs(an)
    # probe
    dp = (rng.standard_normal(pr["probe"].shape) + 1j * rng.standard_normal(pr["probe"].shape)).astype(np.complex64)
    tdp = torch.from_numpy(dp).to(dev) * float(np.abs(pr["probe"]).mean())
    p0 = probe.detach()
    fd = (loss_at(obj.detach(), p0 + eps * tdp, corr.detach()) - loss_at(obj.detach(), p0 - eps * tdp,
                                                                         corr.detach())) / (2 * eps)
    an = float(torch.real(torch.sum(torch.conj(probe.grad) * tdp)).item())
    print(f"  probe: fd {fd:.6e} autograd {an:.6e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
    assert abs(fd - an) <= 1e-2 * abs(an)
    # positions
    dc = torch.from_numpy(rng.uniform(-1, 1, pr["corr"].shape).astype(np.float32)).to(dev)
    c0 = corr.detach()
    e2 = 1e-2
    fd = (loss_at(obj.detach(), probe.detach(), c0 + e2 * dc) - loss_at(obj.detach(), probe.detach(),
                                                                        c0 - e2 * dc)) / (2 * e2)
    an = float(torch.sum(corr.grad * dc).item())
    print(f"  positions: fd {fd:.6e} autograd {an:.6e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
    assert abs(fd - an) <= 1e-2 * abs(an)
