"""Warm-start network kernels on the GPU against the oracle (oracle/cnn_oracle.py): the BF16
matrix-core GEMM in all four operand layouts (ragged sizes, split reductions), im2col / col2im,
and a whole forward + backward pass and Adam steps of the network."""
import numpy as np
import pytest

from oracle import cnn_oracle as cnn

pytestmark = pytest.mark.gpu


def _bits(x):
    """float32 values already on the bfloat16 grid -> int16 tensor of their bf16 bit patterns."""
    import torch

    u = np.ascontiguousarray(x, np.float32).view(np.uint32) >> 16
    return torch.from_numpy(u.astype(np.uint16).view(np.int16).copy()).cuda()


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(37, 45, 29), (64, 64, 64), (130, 70, 200), (16, 16, 70000), (5, 1024, 4096)])
def test_gemm_bf16_matches_numpy(ta, tb, M, N, K):
    import torch

    from pypty_amd import _lib

    print(f"[gemm] ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(M * 7 + N + K)
    a = cnn.bf16_round(rng.standard_normal((M, K)).astype(np.float32))
    b = cnn.bf16_round(rng.standard_normal((K, N)).astype(np.float32))
    A = _bits(a.T.copy() if ta else a)
    B = _bits(b.T.copy() if tb else b)
    lda = M if ta else K
    ldb = K if tb else N
    C = torch.full((M, N + 3), float("nan"), dtype=torch.float32, device="cuda")  # ldc > N
    part = torch.empty(1 << 22, dtype=torch.float32, device="cuda")
    s = _lib.stream_ptr()
    rc = lib.pty_gemm_bf16(ta, tb, M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb, C.data_ptr(), N + 3,
                           part.data_ptr(), part.numel(), s)
    assert rc == 0
    got = C.cpu().numpy()
    ref = a.astype(np.float64) @ b.astype(np.float64)
    assert np.isnan(got[:, N:]).all()  # nothing written past column N
    e = _rel(got[:, :N], ref)
    print(f"  rel err {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("C", [1, 5, 8, 16])
def test_im2col_col2im_match_oracle(C):
    import torch

    from pypty_amd import _lib

    print(f"[im2col] B=3 H=10 W=12 C={C}", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(11 + C)
    X = rng.standard_normal((3, 10, 12, C)).astype(np.float32)
    Kc = 16 * C
    cols = torch.empty((3 * 5 * 6, Kc), dtype=torch.int16, device="cuda")
    s = _lib.stream_ptr()
    assert lib.pty_cnn_im2col(torch.from_numpy(X).cuda().data_ptr(), 3, 10, 12, C, 4, 2, 1, cols.data_ptr(), s) == 0
    got = (cols.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(got, cnn.bf16_round(cnn.im2col(X)))
    D = rng.standard_normal((90, Kc)).astype(np.float32)
    dX = torch.empty((3, 10, 12, C), dtype=torch.float32, device="cuda")
    rc = lib.pty_cnn_col2im(torch.from_numpy(D).cuda().data_ptr(), 3, 10, 12, C, 4, 2, 1, dX.data_ptr(), s)
    if C % 4:
        assert rc == -1  # four channels per thread
        return
    assert rc == 0
    assert np.allclose(dX.cpu().numpy(), cnn.col2im(D, X.shape), rtol=1e-6, atol=1e-6)


def _net_params(net):
    return [(w.cpu().numpy(), b.cpu().numpy()) for w, b in zip(net.W, net.b)]


@pytest.mark.parametrize("n_in,patch,B", [(32, 8, 6), (64, 16, 5)])
def test_network_forward_backward_match_oracle(n_in, patch, B):
    import torch

    from pypty_amd.warmstart import WarmStartNet

    print(f"[cnn] n_in={n_in} patch={patch} B={B}", flush=True)
    net = WarmStartNet(n_in, patch, seed=21)
    rng = np.random.default_rng(22)
    x = rng.uniform(0, 2, (B, n_in, n_in)).astype(np.float32)
    t = rng.standard_normal((B, patch * patch)).astype(np.float32)
    out = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    loss = float(net.backward(torch.from_numpy(t).cuda())[0])
    params = _net_params(net)
    ref_out, _ = cnn.forward(params, x, n_in, patch, bf16=True)
    ref_loss, ref_grads = cnn.loss_and_grads(params, x, t, n_in, patch, bf16=True)
    e_out = _rel(out, ref_out)
    print(f"  out {e_out:.2e} loss {abs(loss - ref_loss) / ref_loss:.2e}", flush=True)
    assert e_out < 1e-4 and abs(loss - ref_loss) < 1e-4 * ref_loss
    for li, ((dW, db), gW, gb) in enumerate(zip(ref_grads, net.gW, net.gb)):
        ew, eb = _rel(gW.cpu().numpy(), dW), _rel(gb.cpu().numpy(), db)
        print(f"  layer {li}: dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew < 2e-3 and eb < 2e-3, li


def test_adam_steps_match_oracle_and_training_fits_a_batch():
    import torch

    from pypty_amd.warmstart import WarmStartNet

    print("[cnn] Adam on one batch, n_in=32 patch=8", flush=True)
    n_in, patch, B = 32, 8, 8
    net = WarmStartNet(n_in, patch, seed=31)
    rng = np.random.default_rng(32)
    x = torch.from_numpy(rng.uniform(0, 2, (B, n_in, n_in)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.standard_normal((B, patch * patch)).astype(np.float32)).cuda()
    params = _net_params(net)
    state = [(np.zeros_like(w), np.zeros_like(w), np.zeros_like(b), np.zeros_like(b)) for w, b in params]
    # one step: GPU gradients and Adam against the oracle's Adam on the oracle's gradients
    net.forward(x)
    net.backward(t)
    net.adam(lr=1e-3)
    _, grads = cnn.loss_and_grads(params, x.cpu().numpy(), t.cpu().numpy(), n_in, patch, bf16=True)
    for li, ((w, b), (dW, db), (mw, vw, mb, vb)) in enumerate(zip(params, grads, state)):
        w1, _, _ = cnn.adam_step(w, dW, mw, vw, 1)
        b1, _, _ = cnn.adam_step(b, db, mb, vb, 1)
        # where a gradient entry is tiny its sign can differ with the summation order; the step there is
        # +-lr either way, so compare the update normwise
        assert _rel(net.W[li].cpu().numpy() - w, w1 - w) < 2e-2, li
        assert _rel(net.b[li].cpu().numpy() - b, b1 - b) < 2e-2, li
    losses = []
    for _ in range(200):
        net.forward(x)
        losses.append(float(net.backward(t)[0]))
        net.adam(lr=1e-3)
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.5f}", flush=True)
    assert losses[-1] < 0.05 * losses[0]
