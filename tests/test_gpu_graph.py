"""A full pass captured in a HIP graph and replayed: every replay, with unrelated eager GPU work and
synchronisations between the replays, reproduces the eager pass bit for bit (the kernels are
deterministic and clear their accumulators with fill kernels, not memsets).  Also the batch
autotuner: its choice is one of the candidates and is cached per shape."""
import numpy as np
import pytest

from tests._problems import make_problem

pytestmark = pytest.mark.gpu


def _engine(n, S, M, P, batch, seed):
    import torch

    from pypty_amd.objective import PassConfig, PtychoPass

    dev = torch.device("cuda", 0)
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=seed)
    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=1, n_probe_modes=M, obj_ny=n + 40, obj_nx=n + 40, n_positions=P,
                     compute_batch=batch, algorithm="lsq_sqrt", step_obj=True, step_probe=True, step_pos=True,
                     shift_active=True, slice_distance=pr["dz"], wavelength=pr["wavelength"],
                     pixel_size_x=pr["px"], pixel_size_y=pr["px"])
    t = dict(obj=torch.from_numpy(pr["obj"]).to(dev),
             probe=torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev),
             pos=torch.from_numpy(pr["pos"]).to(dev), corr=torch.from_numpy(pr["corr"]).to(dev),
             meas=torch.from_numpy(pr["meas"]).to(dev), seq=torch.arange(P, dtype=torch.int32, device=dev))
    out = dict(og=torch.empty_like(t["obj"]), pg=torch.empty_like(t["probe"]), posg=torch.zeros_like(t["corr"]))
    eng = PtychoPass(cfg, dev)

    def run():
        return eng.run(t["obj"], t["probe"], t["pos"], t["corr"], t["meas"], t["seq"], obj_grad=out["og"],
                       probe_grad=out["pg"], pos_grad=out["posg"], sync=False)

    return run, out


@pytest.mark.parametrize("n,S,M,P,batch", [(64, 3, 2, 20, 8), (128, 2, 3, 12, 5), (192, 2, 2, 6, 4)])
def test_graph_replays_match_eager(n, S, M, P, batch):
    import torch

    print(f"[graph] n={n} S={S} M={M} P={P} batch={batch}", flush=True)
    run, out = _engine(n, S, M, P, batch, seed=61 + n)
    ls = run()
    torch.cuda.synchronize()
    ref = (ls.clone(), out["og"].clone(), out["pg"].clone(), out["posg"].clone())
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        lt = run()
    for rep in range(4):
        for t in (out["og"], out["pg"], out["posg"]):
            t.fill_(float("nan"))  # replays must rewrite every output
        x = torch.randn(512, 512, device="cuda")
        (x @ x).sum().item()  # unrelated eager work and a synchronisation between replays
        g.replay()
        torch.cuda.synchronize()
        got = (lt, out["og"], out["pg"], out["posg"])
        for a, b in zip(got, ref):
            assert torch.equal(a, b), rep


def test_autotuner_picks_a_candidate_and_caches(tmp_path, monkeypatch):
    import torch

    from pypty_amd import autotune

    monkeypatch.setattr(autotune, "_cache", {})
    monkeypatch.setattr(autotune, "_loaded", False)
    monkeypatch.setenv("PYPTY_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    calls = []

    def make_pass(b):
        run, _ = _engine(64, 2, 2, 24, b, seed=71)
        calls.append(b)
        return run

    key = autotune.shape_key(torch.device("cuda", 0), 64, 2, 1, 2, 24)
    best = autotune.choose_batch(key, (4, 8, 24), make_pass)
    assert best in (4, 8, 24) and calls == [4, 8, 24]
    assert autotune.choose_batch(key, (4, 8, 24), make_pass) == best and len(calls) == 3
    assert (tmp_path / "tune.json").exists()
