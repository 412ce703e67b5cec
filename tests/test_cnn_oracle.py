"""Pins the warm-start network's oracle (oracle/cnn_oracle.py) without a GPU: bfloat16 rounding
against torch's conversion and hand-worked ties, im2col against a direct convolution and col2im as
its exact adjoint, the backward pass against central finite differences of the exact (FP64)
network, and one Adam step against its closed form."""
import numpy as np
import pytest

from oracle import cnn_oracle as cnn


def test_bf16_rounding_matches_torch_and_ties_to_even():
    import torch

    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    # exactly half an ulp above 1 and above 1 + ulp: ties go to the even mantissa
    assert cnn.bf16_round(np.array([one + ulp / 2]))[0] == one
    assert cnn.bf16_round(np.array([one + ulp + ulp / 2]))[0] == one + 2 * ulp
    assert cnn.bf16_round(np.array([one + ulp / 2 + np.float32(2.0 ** -20)]))[0] == one + ulp
    x = np.random.default_rng(0).standard_normal(100_000).astype(np.float32) * 10.0 ** np.random.default_rng(
        1).integers(-20, 20, 100_000).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(cnn.bf16_round(x), ref)
    special = np.array([np.inf, -np.inf, 0.0, -0.0], np.float32)
    assert np.array_equal(cnn.bf16_round(special), special)
    assert np.isnan(cnn.bf16_round(np.array([np.nan], np.float32))[0])


def test_im2col_is_a_convolution_and_col2im_its_adjoint():
    rng = np.random.default_rng(2)
    B, H, W, C, Co = 2, 8, 6, 3, 5
    X = rng.standard_normal((B, H, W, C))
    Wt = rng.standard_normal((Co, 16 * C))
    Y = (cnn.im2col(X) @ Wt.T).reshape(B, 4, 3, Co)
    # direct: Y[b, oy, ox, o] = sum over taps of X[b, 2oy - 1 + ky, 2ox - 1 + kx, c] W[o, (ky*4 + kx)*C + c]
    Yd = np.zeros_like(Y)
    for b in range(B):
        for oy in range(4):
            for ox in range(3):
                for ky in range(4):
                    for kx in range(4):
                        iy, ix = 2 * oy - 1 + ky, 2 * ox - 1 + kx
                        if 0 <= iy < H and 0 <= ix < W:
                            Yd[b, oy, ox] += Wt[:, (ky * 4 + kx) * C:(ky * 4 + kx + 1) * C] @ X[b, iy, ix]
    assert np.allclose(Y, Yd, rtol=1e-12, atol=1e-12)
    D = rng.standard_normal((B * 4 * 3, 16 * C))
    lhs = np.sum(cnn.im2col(X) * D)
    rhs = np.sum(X * cnn.col2im(D, X.shape))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def _params(spec, rng, scale=1.0):
    out = []
    for kind, cin, cout, relu, hw in spec:
        fan = cin * 16 if kind == "conv" else cin
        out.append((rng.standard_normal((cout, fan)) * scale / np.sqrt(fan), rng.standard_normal(cout) * 0.1))
    return out


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_matches_finite_differences(seed):
    n_in, patch = 16, 4
    spec = cnn.layer_spec(n_in, patch)
    rng = np.random.default_rng(seed)
    params = _params(spec, rng, 1.5)
    x = rng.standard_normal((3, n_in, n_in))
    t = rng.standard_normal((3, patch * patch))
    loss, grads = cnn.loss_and_grads(params, x, t, n_in, patch, bf16=False)
    h = 1e-6
    for li in range(len(spec)):
        for which in (0, 1):
            arr = params[li][which]
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(3)]:
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = cnn.loss_and_grads(params, x, t, n_in, patch, bf16=False)
                arr[idx] = old - h
                lm, _ = cnn.loss_and_grads(params, x, t, n_in, patch, bf16=False)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                g = grads[li][which][idx]
                assert abs(fd - g) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (li, which, idx, fd, g)


def test_bf16_network_stays_close_to_the_exact_one():
    n_in, patch = 32, 8
    spec = cnn.layer_spec(n_in, patch)
    rng = np.random.default_rng(5)
    params = [(w.astype(np.float32), b.astype(np.float32)) for w, b in _params(spec, rng)]
    x = rng.uniform(0, 2, (4, n_in, n_in)).astype(np.float32)
    t = rng.standard_normal((4, patch * patch)).astype(np.float32)
    l16, g16 = cnn.loss_and_grads(params, x, t, n_in, patch, bf16=True)
    l64, g64 = cnn.loss_and_grads([(w.astype(np.float64), b.astype(np.float64)) for w, b in params],
                                  x.astype(np.float64), t, n_in, patch, bf16=False)
    # bfloat16 keeps 8 significant bits; the first layer's weight gradient sums many terms of both
    # signs, so its relative error is the largest (about 8% here)
    assert abs(l16 - l64) < 2e-2 * l64
    for (a, _), (b, _) in zip(g16, g64):
        assert np.linalg.norm(a - b) < 0.2 * np.linalg.norm(b)


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3], np.float32)
    w, m, v = cnn.adam_step(np.zeros(3, np.float32), g, np.zeros(3, np.float32), np.zeros(3, np.float32), 1,
                            lr=0.01)
    # after one step m_hat = g and v_hat = g^2, so the step is lr * g / (|g| + eps)
    assert np.allclose(w, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-5)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
