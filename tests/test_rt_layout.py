"""CPU emulation of the register-tile layout of pypty_amd/csrc/pty_rt.hpp (N = 64, 128).

The index maps below restate the header's (xslot, col_freq, the row/column phase lane
assignment, the transpose positions).  The emulation checks that the two-phase line split
computes the 2D DFT exactly (against numpy), that the inverse path returns the input, that every
LDS address stays inside the N x (N+8) tile, and that each exchange / transpose instruction puts
the 32 lanes of a half-wavefront on 32 distinct 8-byte bank pairs.
"""
import numpy as np
import pytest


def _maps(N):
    R1 = N // 8
    CPT = R1 // 8
    LS = N + 8
    xslot = lambda j, k1: k1 * 8 + (j ^ (k1 // CPT))  # noqa: E731
    col_freq = lambda col: CPT * (col & 7) + ((col >> 3) % CPT) + R1 * ((col >> 3) // CPT)  # noqa: E731
    return R1, CPT, LS, xslot, col_freq


def _line_fwd(N, xs, T, at):
    R1, CPT, LS, xslot, _ = _maps(N)
    for j in range(8):
        v = np.fft.fft(xs[j]) * np.exp(-2j * np.pi * j * np.arange(R1) / N)
        for k1 in range(R1):
            T[at(xslot(j, k1))] = v[k1]
    out = []
    for j in range(8):
        u = np.zeros(R1, complex)
        for c in range(CPT):
            u[c * 8:(c + 1) * 8] = np.fft.fft([T[at(xslot(jj, CPT * j + c))] for jj in range(8)])
        out.append(u)
    return out


def _line_inv(N, vs, T, at):
    R1, CPT, LS, xslot, _ = _maps(N)
    for j in range(8):
        for c in range(CPT):
            y = np.fft.ifft(vs[j][c * 8:(c + 1) * 8]) * 8
            for jj in range(8):
                T[at(xslot(jj, CPT * j + c))] = y[jj]
    out = []
    for j in range(8):
        z = np.array([T[at(xslot(j, k1))] for k1 in range(R1)]) * np.exp(2j * np.pi * j * np.arange(R1) / N)
        out.append(np.fft.ifft(z) * R1)
    return out


@pytest.mark.parametrize("N", [64, 128])
def test_register_tile_computes_the_2d_dft(N):
    R1, CPT, LS, xslot, col_freq = _maps(N)
    rng = np.random.default_rng(N)
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    T = np.full(N * LS, np.nan + 0j)
    for y in range(N):  # row phase, then the transpose write
        out = _line_fwd(N, [X[y, rj::8] for rj in range(8)], T, lambda p, y=y: y * LS + p)
        for rj in range(8):
            for c in range(CPT):
                for k2 in range(8):
                    T[y * LS + 8 * (CPT * k2 + c) + rj] = out[rj][c * 8 + k2]
    F = np.zeros((N, N), complex)
    for col in range(N):  # column phase
        vs = [np.array([T[(cj + 8 * n) * LS + col] for n in range(R1)]) for cj in range(8)]
        out = _line_fwd(N, vs, T, lambda p, col=col: p * LS + col)
        for cj in range(8):
            for c in range(CPT):
                for k2 in range(8):
                    F[CPT * cj + c + R1 * k2, col_freq(col)] = out[cj][c * 8 + k2]
    assert np.abs(F - np.fft.fft2(X)).max() <= 1e-12 * np.abs(F).max()
    # inverse: column phase, transpose, row phase
    Y = np.zeros((N, N), complex)
    T[:] = 0
    for col in range(N):
        fx = col_freq(col)
        vs = [np.array([F[CPT * cj + c + R1 * k2, fx] for c in range(CPT) for k2 in range(8)]) for cj in range(8)]
        out = _line_inv(N, vs, T, lambda p, col=col: p * LS + col)
        for cj in range(8):
            for n in range(R1):
                T[(cj + 8 * n) * LS + col] = out[cj][n]
    for y in range(N):
        vs = [np.array([T[y * LS + 8 * (CPT * k2 + c) + rj] for c in range(CPT) for k2 in range(8)])
              for rj in range(8)]
        out = _line_inv(N, vs, T, lambda p, y=y: y * LS + p)
        for rj in range(8):
            Y[y, rj::8] = out[rj]
    assert np.abs(Y / N ** 2 - X).max() <= 1e-12 * np.abs(X).max()


@pytest.mark.parametrize("N", [192, 256])
def test_line_split_of_large_windows(N):
    """The 24- and 32-sample line splits (3 and 4 DFT-8 groups per thread) of the 192 / 256 px
    global-slab tiles: one exchange per line through an LDS row of stride N + 8."""
    R1, CPT, LS, xslot, col_freq = _maps(N)
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    T = np.full(LS, np.nan + 0j)
    out = _line_fwd(N, [x[j::8] for j in range(8)], T, lambda p: p)
    X = np.zeros(N, complex)
    for j in range(8):
        for c in range(CPT):
            for k2 in range(8):
                X[CPT * j + c + R1 * k2] = out[j][c * 8 + k2]
    assert np.abs(X - np.fft.fft(x)).max() <= 1e-12 * np.abs(X).max()
    back = _line_inv(N, out, T, lambda p: p)
    y = np.zeros(N, complex)
    for j in range(8):
        y[j::8] = back[j]
    assert np.abs(y / N - x).max() <= 1e-12 * np.abs(x).max()
    assert sorted(col_freq(c) for c in range(N)) == list(range(N))
    # exchanges of 8 lines per wavefront in rows of stride LS: 32 distinct bank pairs per half-wave
    lanes = np.arange(64)
    for k1 in range(R1):
        a = np.array([(l >> 3) * LS + xslot(l & 7, k1) for l in lanes])
        assert a.max() < 8 * LS and all(len(set((h % 32).tolist())) == 32 for h in (a[:32], a[32:]))
    for c in range(CPT):
        for jj in range(8):
            a = np.array([(l >> 3) * LS + xslot(jj, CPT * (l & 7) + c) for l in lanes])
            assert all(len(set((h % 32).tolist())) == 32 for h in (a[:32], a[32:]))


@pytest.mark.parametrize("N", [64, 128])
def test_register_tile_lds_bounds_and_banks(N):
    R1, CPT, LS, xslot, _ = _maps(N)

    def check(addrs):
        a = np.asarray(addrs)
        assert a.min() >= 0 and a.max() < N * LS
        for half in (a[:32], a[32:]):
            assert len(set((half % 32).tolist())) == 32

    lanes = np.arange(64)
    for w in range(N // 8):
        # row phase: lane = j + 8*row_in_wave; column phase: lane = col_in_wave + 8*j
        for k1 in range(R1):
            check([(8 * w + (l >> 3)) * LS + xslot(l & 7, k1) for l in lanes])
            check([xslot(l >> 3, k1) * LS + 8 * w + (l & 7) for l in lanes])
        for c in range(CPT):
            for jj in range(8):
                check([(8 * w + (l >> 3)) * LS + xslot(jj, CPT * (l & 7) + c) for l in lanes])
                check([xslot(jj, CPT * (l >> 3) + c) * LS + 8 * w + (l & 7) for l in lanes])
            for k2 in range(8):
                check([(8 * w + (l >> 3)) * LS + 8 * (CPT * k2 + c) + (l & 7) for l in lanes])
        for n in range(R1):
            check([((l >> 3) + 8 * n) * LS + 8 * w + (l & 7) for l in lanes])
