"""Learned warm start: a small CNN that maps each diffraction pattern to the
projected phase around its probe position, trained on the GPU on pairs that the
multislice forward model simulates; the stitched predictions replace the flat
starting object of the reconstruction.

The reference has no learned component.  Its reconstruction starts from the
object given in ``pypty_params['obj']`` (pypty/iterative.py:56-250, ones by
default); ``warm_start_object`` builds that array.  The plain start stays the
fallback, and the iterative reconstruction still decides the final error.

Network (activations pixel-major, channels innermost; ``layer_spec``):
four 4x4 stride-2 convolutions 1 -> 16 -> 32 -> 32 -> 64 channels with ReLU,
then a dense layer to patch^2 features with ReLU and a dense layer to the
patch^2 phase values.  Each layer is one BF16 x BF16 -> FP32 GEMM on the
matrix cores (``pty_gemm_bf16``) plus im2col / bias / activation kernels of
``libpypty_amd.so``; weights are FP32 masters with BF16 copies, trained with
Adam on the mean squared error.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, synth

KS, STRIDE, PAD = 4, 2, 1


def layer_spec(n_in: int, patch: int):
    """[(kind, cin, cout, relu, in_hw)] for n_in x n_in inputs and patch x patch outputs."""
    if n_in % 16:
        raise ValueError("pypty_amd.warmstart: the input side must be a multiple of 16")
    spec, hw = [], n_in
    for cin, cout in ((1, 16), (16, 32), (32, 32), (32, 64)):
        spec.append(("conv", cin, cout, True, hw))
        hw //= 2
    spec.append(("dense", hw * hw * 64, patch * patch, True, 1))
    spec.append(("dense", patch * patch, patch * patch, False, 1))
    return spec


def log_scale(patterns: torch.Tensor) -> torch.Tensor:
    """Network input of diffraction patterns: log(1 + I) / 4."""
    return torch.log1p(torch.clamp(patterns, min=0.0)) * 0.25


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class WarmStartNet:
    """The network on one HIP device: parameters, Adam state and per-batch buffers."""

    def __init__(self, n_in: int = 128, patch: int = 32, device=None, seed: int = 0):
        self.lib = _lib.load()
        self.n_in, self.patch = n_in, patch
        self.spec = layer_spec(n_in, patch)
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        rng = np.random.default_rng(seed)
        self.W, self.b, self.Wbf = [], [], []
        for kind, cin, cout, relu, hw in self.spec:
            fan_in = cin * KS * KS if kind == "conv" else cin
            bound = math.sqrt(6.0 / fan_in)  # He-uniform
            w = rng.uniform(-bound, bound, (cout, fan_in)).astype(np.float32)
            self.W.append(torch.from_numpy(w).to(self.dev))
            self.b.append(torch.zeros(cout, dtype=torch.float32, device=self.dev))
            self.Wbf.append(torch.empty((cout, fan_in), dtype=torch.int16, device=self.dev))
            self._bf16(self.W[-1], self.Wbf[-1])
        self.params = [t for pair in zip(self.W, self.b) for t in pair]
        self.m = [torch.zeros_like(t) for t in self.params]
        self.v = [torch.zeros_like(t) for t in self.params]
        self.step_count = 0
        self._buf_B = 0

    # ---------------------------------------------------------------- plumbing
    def _s(self):
        return ctypes.c_void_p(_lib.stream_ptr(self.dev))

    def _bf16(self, x, out):
        _lib.check(self.lib.pty_cnn_to_bf16(_p(x), x.numel(), _p(out), self._s()), "to_bf16")

    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, C, ldc):
        _lib.check(self.lib.pty_gemm_bf16(ta, tb, M, N, K, _p(A), lda, _p(B), ldb, _p(C), ldc, _p(self.partial),
                                          self.partial.numel(), self._s()), "gemm_bf16")

    def _buffers(self, B):
        if B == self._buf_B:
            return
        f32, i16 = torch.float32, torch.int16
        self.cols, self.Y, self.A, self.Abf, self.dY, self.dYbf, self.dcols = [], [], [], [], [], [], []
        for li, (kind, cin, cout, relu, hw) in enumerate(self.spec):
            P = B * (hw // 2) ** 2 if kind == "conv" else B
            Kc = cin * KS * KS if kind == "conv" else cin
            self.cols.append(torch.empty((P, Kc), dtype=i16, device=self.dev) if kind == "conv" else None)
            self.Y.append(torch.empty((P, cout), dtype=f32, device=self.dev))
            self.A.append(torch.empty((P, cout), dtype=f32, device=self.dev))
            nxt = self.spec[li + 1][0] if li + 1 < len(self.spec) else None
            self.Abf.append(torch.empty((P, cout), dtype=i16, device=self.dev) if nxt == "dense" else None)
            self.dY.append(torch.empty((P, cout), dtype=f32, device=self.dev))
            self.dYbf.append(torch.empty((P, cout), dtype=i16, device=self.dev))
            self.dcols.append(torch.empty((P, Kc), dtype=f32, device=self.dev) if li > 0 else None)
        self.dA = [torch.empty_like(y) for y in self.Y]
        self.gW = [torch.empty_like(w) for w in self.W]
        self.gb = [torch.empty_like(b) for b in self.b]
        # split-reduction partials of the long, narrow GEMMs (weight gradients)
        self.partial = torch.empty(1 << 24, dtype=f32, device=self.dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self.loss_part = torch.empty(1024, dtype=torch.float64, device=self.dev)
        self._buf_B = B

    # ----------------------------------------------------------- forward/backward
    def _input_cols(self, li):
        """GEMM operand of layer li: its im2col columns or the previous layer's BF16 activations."""
        return self.cols[li] if self.spec[li][0] == "conv" else self.Abf[li - 1]

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x (B, n, n) float32 (log-scaled) -> (B, patch^2) float32 (a view of an internal buffer)."""
        if x.dtype != torch.float32 or tuple(x.shape[1:]) != (self.n_in, self.n_in):
            raise ValueError(f"pypty_amd.warmstart: input must be float32 (B, {self.n_in}, {self.n_in})")
        x = _lib.require_device_tensor(x.contiguous(), "x")
        B = x.shape[0]
        self._buffers(B)
        s = self._s()
        prev = x
        for li, (kind, cin, cout, relu, hw) in enumerate(self.spec):
            if kind == "conv":
                _lib.check(self.lib.pty_cnn_im2col(_p(prev), B, hw, hw, cin, KS, STRIDE, PAD, _p(self.cols[li]), s),
                           "im2col")
                P, Kc = self.cols[li].shape
            else:
                P, Kc = B, cin
            self._gemm(0, 1, P, cout, Kc, self._input_cols(li), Kc, self.Wbf[li], Kc, self.Y[li], cout)
            _lib.check(self.lib.pty_cnn_bias_act(_p(self.Y[li]), P, cout, _p(self.b[li]), int(relu), _p(self.A[li]),
                                                 _p(self.Abf[li]), s), "bias_act")
            prev = self.A[li]
        return self.A[-1]

    def backward(self, target: torch.Tensor) -> torch.Tensor:
        """Gradients of the mean squared error against target (B, patch^2) after forward(); returns the
        loss as a device float64 tensor."""
        B = self._buf_B
        s = self._s()
        out = self.A[-1]
        _lib.check(self.lib.pty_cnn_mse(_p(out), _p(target.contiguous()), out.numel(), _p(self.dA[-1]),
                                        _p(self.loss), _p(self.loss_part), self.loss_part.numel(), s), "mse")
        for li in range(len(self.spec) - 1, -1, -1):
            kind, cin, cout, relu, hw = self.spec[li]
            P = self.Y[li].shape[0]
            Kc = self.W[li].shape[1]
            _lib.check(self.lib.pty_cnn_act_grad(_p(self.dA[li]), _p(self.A[li]), P * cout, int(relu),
                                                 _p(self.dY[li]), _p(self.dYbf[li]), s), "act_grad")
            _lib.check(self.lib.pty_cnn_bias_grad(_p(self.dY[li]), P, cout, _p(self.gb[li]), _p(self.partial),
                                                  self.partial.numel(), s), "bias_grad")
            self._gemm(1, 0, cout, Kc, P, self.dYbf[li], cout, self._input_cols(li), Kc, self.gW[li], Kc)
            if li > 0:
                self._gemm(0, 0, P, Kc, cout, self.dYbf[li], cout, self.Wbf[li], Kc, self.dcols[li], Kc)
                if kind == "conv":
                    _lib.check(self.lib.pty_cnn_col2im(_p(self.dcols[li]), B, hw, hw, cin, KS, STRIDE, PAD,
                                                       _p(self.dA[li - 1]), s), "col2im")
                else:
                    self.dA[li - 1].view(-1).copy_(self.dcols[li].view(-1))
        return self.loss

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        s = self._s()
        grads = [t for pair in zip(self.gW, self.gb) for t in pair]
        bfs = [t for pair in zip(self.Wbf, [None] * len(self.b)) for t in pair]
        for w, wbf, g, m, v in zip(self.params, bfs, grads, self.m, self.v):
            _lib.check(self.lib.pty_cnn_adam(_p(w), _p(wbf), _p(g), _p(m), _p(v), w.numel(), lr, beta1, beta2, eps,
                                             self.step_count, s), "adam")

    # ------------------------------------------------------------------ use
    def train(self, x_all: torch.Tensor, t_all: torch.Tensor, epochs=5, batch=256, lr=1e-3, seed=0, log=None,
              lr_end=None):
        """Adam on shuffled minibatches (seeded order), the learning rate decaying geometrically from
        lr to lr_end (default: constant) over the epochs; returns the mean loss of each epoch."""
        g = torch.Generator()
        g.manual_seed(int(seed))
        n = x_all.shape[0]
        history = []
        for ep in range(epochs):
            lr_ep = lr if lr_end is None or epochs < 2 else lr * (lr_end / lr) ** (ep / (epochs - 1))
            order = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), dtype=torch.float64, device=self.dev)
            steps = 0
            for i in range(0, n - batch + 1, batch):
                idx = order[i:i + batch]
                self.forward(x_all.index_select(0, idx))
                tot += self.backward(t_all.index_select(0, idx))[0]
                self.adam(lr=lr_ep)
                steps += 1
            history.append(float(tot) / max(steps, 1))
            if log:
                log(f"epoch {ep}: loss {history[-1]:.5f}")
        return history

    def predict(self, x_all: torch.Tensor, batch=256) -> torch.Tensor:
        out = torch.empty((x_all.shape[0], self.patch * self.patch), dtype=torch.float32, device=self.dev)
        for i in range(0, x_all.shape[0], batch):
            xb = x_all[i:i + batch]
            out[i:i + xb.shape[0]] = self.forward(xb)
        return out


def phase_patches(obj, corners, n, patch):
    """Projected phase (sum over slices and modes) of the patch x patch region centred in each
    n x n window: (P, patch^2) float32 (the training targets)."""
    proj = np.angle(obj).sum(axis=(2, 3)).astype(np.float32)
    c0 = (n - patch) // 2
    out = np.empty((len(corners), patch * patch), np.float32)
    for i, (y0, x0) in enumerate(np.asarray(corners)):
        out[i] = proj[y0 + c0:y0 + c0 + patch, x0 + c0:x0 + c0 + patch].ravel()
    return out


def simulate_pairs(n_pairs, n=128, patch=32, n_slices=10, pixel_A=0.2, dz_A=2.0, dose=1e4, per_object=2048,
                   seed=3000, device=None):
    """Training pairs from seeded random-atom objects (n_slices slices, 300 kV) under single-mode
    probes with a random 15-25 mrad aperture and a random defocus in +-200 A, at random positions;
    patterns by the GPU forward model with Poisson noise.  Returns (x (P, n, n), t (P, patch^2)) on
    the device: log-scaled patterns and projected-phase patches."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    side = n + 256
    lam = synth.electron_wavelength(300.0)
    made = 0
    k = 0
    while made < n_pairs:
        P = min(per_object, n_pairs - made)
        obj = synth.random_atom_object(side, side, n_slices, pixel_A, dz_A, 0.1, 0.3, 0.5, seed=seed + 17 * k + 1)
        probe = synth.make_probe(n, pixel_A, 300.0, float(rng.uniform(15, 25)), float(rng.uniform(-200, 200)), 1,
                                 0.0, seed=seed + 17 * k + 2)
        pos = rng.uniform(0, side - n - 1, (P, 2))
        ints, corr = synth.split_positions(pos)
        meas = synth.simulate(obj, probe, ints, corr, pixel_A, dz_A, lam, dose, seed + 17 * k + 3, dev)
        xs.append(log_scale(torch.from_numpy(meas).to(dev)))
        ts.append(torch.from_numpy(phase_patches(obj, ints, n, patch)).to(dev))
        made += P
        k += 1
    return torch.cat(xs), torch.cat(ts)


def warm_start_object(net: WarmStartNet, measured, corners, obj_shape, sigma_px=None, chunk=256):
    """Initial object (Ny, Nx, S, K) complex64 from the network's phase patches: overlapping
    patches are averaged with a Gaussian weight around each patch centre (sigma_px, default an
    eighth of the patch; the network is surest near the probe), the projected phase is spread
    evenly over the slices, and pixels that no patch covers keep the flat start (phase 0).

    Everything runs on the network's device: patterns go up `chunk` at a time and are predicted
    as they arrive; the weighted patches (w * phase, w) are summed as one complex scatter-add
    (``pty_scatteradd_c64``) into a frame padded by a patch on every side, so patches that reach
    past the object edge are clipped."""
    ny, nx, S, K = obj_shape
    n, patch, dev = net.n_in, net.patch, net.dev
    corners = np.asarray(corners)
    P = len(corners)
    pred = torch.empty((P, patch * patch), dtype=torch.float32, device=dev)
    for i in range(0, P, chunk):
        m = measured[i:i + chunk]
        m = m if isinstance(m, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(m, np.float32))
        pred[i:i + m.shape[0]] = net.forward(log_scale(m.to(dev).float()))
    sig = patch / 8.0 if sigma_px is None else float(sigma_px)
    r = torch.arange(patch, device=dev, dtype=torch.float32) - (patch - 1) / 2.0
    w = torch.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / (2 * sig * sig))
    wp = pred.view(P, patch, patch) * w
    patches = torch.complex(wp, w.expand_as(wp)).contiguous()  # (P, 1 plane, patch, patch)
    pad, c0 = patch, (n - patch) // 2
    full = torch.zeros((ny + 2 * pad, nx + 2 * pad), dtype=torch.complex64, device=dev)
    cor = torch.from_numpy((corners.astype(np.int64) + c0 + pad).astype(np.int32)).to(dev).contiguous()
    rc = net.lib.pty_scatteradd_c64(_p(full), ny + 2 * pad, nx + 2 * pad, 1, _lib.PTY_LAYOUT_PLANAR, _p(patches),
                                    P, patch, _p(cor), _lib.stream_ptr(dev))
    _lib.check(rc, "warm_start_object scatter-add")
    full = full[pad:pad + ny, pad:pad + nx]
    acc, cnt = full.real, full.imag
    phase = torch.where(cnt > 1e-6, acc / torch.clamp(cnt, min=1e-6), torch.zeros_like(acc))
    obj = torch.polar(torch.ones_like(phase), phase / (S * K))
    return obj[:, :, None, None].expand(ny, nx, S, K).cpu().numpy()
