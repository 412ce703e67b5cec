"""Batch lanes: a pass whose consecutive batches go round robin over the caller's stream and one to
three library streams gives the loss and every gradient of the one-stream pass bit for bit (the
kernels that add into the shared gradients run in batch order across the lanes).  Covers the
register tile with four and eight threads per line, the slab tile, object modes, checkpointed
passes, an odd number of batches, a single batch (no second lane), and a HIP graph captured from a
two-lane engine on a stream whose lanes were in use eagerly (a capturing call runs one lane),
replayed with eager work between the replays; and two datasets of one shape whose passes run at
the same time on two streams."""
import numpy as np
import pytest

from tests._problems import make_problem
from tests.test_gpu_checkpoint import _pass

pytestmark = pytest.mark.gpu

CASES = [
    # n, S, K, M, P, batch, k
    (32, 3, 2, 2, 11, 4, 0),
    (96, 4, 1, 2, 9, 2, 0),
    (128, 3, 1, 3, 8, 3, 0),
    (128, 5, 1, 2, 7, 2, 2),
    (192, 3, 1, 2, 5, 2, 0),
    (256, 2, 1, 1, 4, 2, 0),
    (64, 2, 1, 2, 5, 8, 0),  # one batch
]


@pytest.mark.parametrize("n,S,K,M,P,batch,ck", CASES)
def test_two_lanes_are_bit_identical_to_one(n, S, K, M, P, batch, ck):
    import torch

    print(f"[lanes] n={n} S={S} K={K} M={M} P={P} batch={batch} k={ck}", flush=True)
    pr = make_problem(n=n, S=S, K=K, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=131 + n + ck)
    dev = torch.device("cuda", 0)
    one = _pass(pr, n, S, K, M, P, batch, ck, dev, lanes=1)
    names = ("loss", "object gradient", "probe gradient", "position gradient")
    assert float(one[1].abs().max()) > 0 and float(one[3].abs().max()) > 0
    for lanes in (2, 3, 4):
        more = _pass(pr, n, S, K, M, P, batch, ck, dev, lanes=lanes)
        for name, a, b in zip(names, one[:4], more[:4]):
            assert torch.isfinite(a).all(), name
            assert torch.equal(a, b), (lanes, name)
        assert more[4] > one[4]  # a set of per-batch buffers per lane


def test_graph_captured_from_a_two_lane_engine_replays_bit_identically():
    import torch

    from pypty_amd.objective import PassConfig, PtychoPass

    n, S, M, P, batch = 128, 2, 2, 10, 3
    print(f"[lanes] graph n={n} S={S} M={M} P={P} batch={batch}", flush=True)
    dev = torch.device("cuda", 0)
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=171)
    t = dict(obj=torch.from_numpy(pr["obj"]).to(dev),
             probe=torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev),
             pos=torch.from_numpy(pr["pos"]).to(dev), corr=torch.from_numpy(pr["corr"]).to(dev),
             meas=torch.from_numpy(pr["meas"]).to(dev), seq=torch.arange(P, dtype=torch.int32, device=dev))

    def engine(lanes):
        cfg = PassConfig(n=n, n_slices=S, n_obj_modes=1, n_probe_modes=M, obj_ny=n + 40, obj_nx=n + 40,
                         n_positions=P, compute_batch=batch, algorithm="lsq_sqrt", step_obj=True, step_probe=True,
                         step_pos=True, shift_active=True, slice_distance=pr["dz"], wavelength=pr["wavelength"],
                         pixel_size_x=pr["px"], pixel_size_y=pr["px"], lanes=lanes)
        eng = PtychoPass(cfg, dev)
        out = dict(og=torch.empty_like(t["obj"]), pg=torch.empty_like(t["probe"]), posg=torch.zeros_like(t["corr"]))

        def run():
            return eng.run(t["obj"], t["probe"], t["pos"], t["corr"], t["meas"], t["seq"], obj_grad=out["og"],
                           probe_grad=out["pg"], pos_grad=out["posg"], sync=False)

        return run, out

    run1, out1 = engine(1)
    ls = run1()
    torch.cuda.synchronize()
    ref = (ls.clone(), out1["og"].clone(), out1["pg"].clone(), out1["posg"].clone())
    run2, out2 = engine(2)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run2()  # eager use of the capture stream, with its second lane
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        lt = run2()
    for rep in range(4):
        for x in (out2["og"], out2["pg"], out2["posg"]):
            x.fill_(float("nan"))  # replays must rewrite every output
        y = torch.randn(512, 512, device="cuda")
        (y @ y).sum().item()  # unrelated eager work and a synchronisation between replays
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((lt, out2["og"], out2["pg"], out2["posg"]), ref):
            assert torch.equal(a, b), rep


def test_engines_of_one_shape_on_two_streams_keep_their_own_workspaces():
    """Two datasets of one shape whose passes run at the same time on two streams (as a step over
    several datasets does) give each dataset's one-stream results bit for bit: engines started on
    different streams take different workspaces, engines started on one stream share one."""
    import torch

    from pypty_amd.objective import PassConfig, PtychoPass

    n, S, M, P, batch = 96, 3, 2, 9, 2
    print(f"[lanes] two streams n={n} S={S} M={M} P={P} batch={batch}", flush=True)
    dev = torch.device("cuda", 0)
    prs = [make_problem(n=n, S=S, K=1, M=M, ny=n + 40, nx=n + 40, P=P, shift=True, seed=191 + k) for k in range(2)]
    cfg = PassConfig(n=n, n_slices=S, n_obj_modes=1, n_probe_modes=M, obj_ny=n + 40, obj_nx=n + 40, n_positions=P,
                     compute_batch=batch, algorithm="lsq_sqrt", step_obj=True, step_probe=True, step_pos=True,
                     shift_active=True, slice_distance=prs[0]["dz"], wavelength=prs[0]["wavelength"],
                     pixel_size_x=prs[0]["px"], pixel_size_y=prs[0]["px"], lanes=2)
    ts = [dict(obj=torch.from_numpy(pr["obj"]).to(dev),
               probe=torch.from_numpy(np.ascontiguousarray(pr["probe"].transpose(2, 0, 1))).to(dev),
               pos=torch.from_numpy(pr["pos"]).to(dev), corr=torch.from_numpy(pr["corr"]).to(dev),
               meas=torch.from_numpy(pr["meas"]).to(dev), seq=torch.arange(P, dtype=torch.int32, device=dev))
          for pr in prs]

    def launch(eng, t):
        og, pg = torch.empty_like(t["obj"]), torch.empty_like(t["probe"])
        posg = torch.zeros_like(t["corr"])
        ls = eng.run(t["obj"], t["probe"], t["pos"], t["corr"], t["meas"], t["seq"], obj_grad=og, probe_grad=pg,
                     pos_grad=posg, sync=False)
        return ls, og, pg, posg

    serial = []
    for t in ts:
        serial.append([x.clone() for x in launch(PtychoPass(cfg, dev), t)])
        torch.cuda.synchronize()
    streams = [torch.cuda.Stream(dev) for _ in ts]
    engines = [PtychoPass(cfg, dev) for _ in ts]
    main = torch.cuda.current_stream(dev)
    for rep in range(3):
        outs = []
        for st, eng, t in zip(streams, engines, ts):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                outs.append(launch(eng, t))
        for st in streams:
            main.wait_stream(st)
        torch.cuda.synchronize()
        assert engines[0].workspace.data_ptr() != engines[1].workspace.data_ptr()
        for k in range(2):
            for a, b in zip(outs[k], serial[k]):
                assert torch.equal(a, b), (rep, k)
    shared = PtychoPass(cfg, dev)
    with torch.cuda.stream(streams[1]):
        launch(shared, ts[0])
    torch.cuda.synchronize()
    assert shared.workspace.data_ptr() == engines[1].workspace.data_ptr()
