"""Measured-data formats on the GPU: compressed (virtual-detector) losses, the per-batch
preprocessing of raw patterns, and patterns streamed from host memory per batch.

Parity is against the CPU oracle (complex64, 1e-4 normwise relative) with the FP64 error
printed beside it; each case is checked for conditioning first (FP32 and FP64 oracles agree
far inside the tolerance).  The streamed pass must reproduce the device-resident pass.
"""
import numpy as np
import pytest

from oracle import ms_oracle as orc
from tests._problems import call_loss_and_direction, make_problem, oracle_pass, reference_kwargs, rel

TOL = 1e-4

pytestmark = pytest.mark.gpu


def _run(pr, dev, **over):
    import torch

    from pypty_amd import objective

    kw = reference_kwargs(
        this_obj=torch.from_numpy(pr["obj"]).to(dev), full_probe=torch.from_numpy(pr["probe"]).to(dev),
        this_pos_array=pr["pos"], this_pos_correction=pr["corr"], this_distances=np.array([pr["dz"]], np.float32),
        measured_array=pr["meas"], algorithm_type=pr["algorithm"], this_wavelength=pr["wavelength"],
        this_step_probe=1, this_step_obj=1, this_step_pos_correction=int(pr["step_pos"]), this_step_tilts=0,
        pixel_size_x_A=pr["px"], pixel_size_y_A=pr["px"], this_chopped_sequence=pr["seq"], masks=pr.get("masks"))
    kw.update(over)
    out = call_loss_and_direction(objective.loss_and_direction, **kw)
    return out[0], out[1], out[2].cpu().numpy(), out[3].cpu().numpy(), out[4].cpu().numpy()


def _compare(got, ref, ref64, step_pos):
    errs = dict(loss=abs(got[0] - ref[0]) / abs(ref[0]), sse=abs(got[1] - ref[1]) / abs(ref[1]),
                obj=rel(got[2], ref[2]), probe=rel(got[3], ref[3]))
    errs64 = dict(loss=abs(got[0] - ref64[0]) / abs(ref64[0]), obj=rel(got[2], ref64[2]), probe=rel(got[3], ref64[3]))
    if step_pos:
        errs["pos"] = rel(got[4], ref[4])
        errs64["pos"] = rel(got[4], ref64[4])
    print("  vs complex64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    print("  vs FP64 oracle:     ", {k: f"{v:.2e}" for k, v in errs64.items()})
    for k, v in errs.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("algorithm", ["lsq_compressed", "gauss_compressed", "poisson_compressed"])
@pytest.mark.parametrize("n,S,M,shift", [(32, 2, 2, True), (64, 3, 1, True), (96, 2, 2, True), (128, 2, 2, True)])
def test_compressed_losses_match_oracle(algorithm, n, S, M, shift):
    import torch

    print(f"[compressed] {algorithm} n={n} S={S} M={M} shift={shift}", flush=True)
    pr = make_problem(n=n, S=S, K=1, M=M, ny=n + 48, nx=n + 48, P=7, shift=shift, seed=31 + n, algorithm=algorithm,
                      n_masks=6)
    ref = oracle_pass(pr, np.complex64)
    ref64 = oracle_pass(pr, np.complex128)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    got = _run(pr, torch.device("cuda", 0))
    _compare(got, ref, ref64, pr["step_pos"])


@pytest.mark.parametrize("n,raw,shift,bin_,pad,mult", [(64, 100, (3, -2), 2, 7, 0.5), (32, 32, (-4, 5), 1, 0, 2.0),
                                                      (128, 248, (0, 6), 2, 2, 1.0)])
def test_preprocessed_raw_patterns_match_oracle(n, raw, shift, bin_, pad, mult):
    """Raw (raw x raw) patterns with shift / bin / pad / multiplier give the pass of the oracle on
    the oracle-preprocessed patterns."""
    import torch

    print(f"[preprocess] n={n} raw={raw} shift={shift} bin={bin_} pad={pad} mult={mult}", flush=True)
    pr = make_problem(n=n, S=2, K=1, M=2, ny=n + 40, nx=n + 40, P=6, shift=True, seed=41 + n)
    rng = np.random.default_rng(7)
    raw_data = (rng.uniform(0.0, 2.0, (6, raw, raw)) * np.mean(pr["meas"]) / max(mult * bin_ * bin_, 1e-9)
                ).astype(np.float32)
    pre = orc.preprocess_dataset(raw_data, shift, bin_, pad, mult).astype(np.float32)
    assert pre.shape == (6, n, n)
    q = dict(pr, meas=pre)
    ref = oracle_pass(q, np.complex64)
    ref64 = oracle_pass(q, np.complex128)
    assert max(rel(ref[2], ref64[2]), rel(ref[3], ref64[3])) < TOL / 3
    got = _run(dict(pr, meas=raw_data), torch.device("cuda", 0), data_shift_vector=list(shift), data_bin=bin_,
               data_pad=pad, data_multiplier=mult)
    _compare(got, ref, ref64, pr["step_pos"])


@pytest.mark.parametrize("n,P,batch,algorithm,pre", [(64, 23, 5, "lsq_sqrt", False), (128, 17, 8, "ml", False),
                                                    (32, 19, 4, "poisson_compressed", False),
                                                    (64, 13, 6, "lsq_sqrt", True)])
def test_streamed_patterns_reproduce_resident_pass(n, P, batch, algorithm, pre):
    """load_one_by_one with numpy patterns streams each batch through pinned buffers; the result equals
    the pass with all patterns on the device (same kernels and summation order; the loss is summed per
    batch in FP64)."""
    import torch

    print(f"[stream] n={n} P={P} batch={batch} {algorithm} preprocess={pre}", flush=True)
    pr = make_problem(n=n, S=2, K=1, M=2, ny=n + 40, nx=n + 40, P=P, shift=True, seed=51 + n, algorithm=algorithm)
    # a shuffled sequence with one position left out
    pr["seq"] = np.random.default_rng(3).permutation(P)[:-1]
    dev = torch.device("cuda", 0)
    over = dict(compute_batch=batch)
    meas = pr["meas"]
    if pre:
        # raw patterns of 2n x 2n whose 2x2 blocks bin back to the problem's patterns exactly
        meas = np.repeat(np.repeat(pr["meas"], 2, axis=1), 2, axis=2) * np.float32(0.25)
        over.update(data_bin=2)
    resident = _run(dict(pr, meas=torch.from_numpy(np.ascontiguousarray(meas)).to(dev)), dev, **over)
    streamed = _run(dict(pr, meas=np.ascontiguousarray(meas)), dev, load_one_by_one=True, **over)
    print(f"  loss resident {resident[0]:.10e} streamed {streamed[0]:.10e}", flush=True)
    assert abs(streamed[0] - resident[0]) <= 1e-12 * abs(resident[0])
    assert abs(streamed[1] - resident[1]) <= 1e-12 * abs(resident[1])
    for k in (2, 3, 4):
        assert np.array_equal(streamed[k], resident[k]), k
    if pre:
        ref = oracle_pass(pr, np.complex64)
        assert abs(resident[0] - ref[0]) / abs(ref[0]) < TOL
