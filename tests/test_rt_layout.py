"""CPU emulation of the register-tile layout of pypty_amd/csrc/pty_rt.hpp (N = 16 ... 128;
T = 8 threads per line when 64 divides N, else 4).

The index maps below restate the header's (xslot, col_freq, the row/column phase lane
assignment, the transpose positions).  The emulation checks that the two-phase line split
computes the 2D DFT exactly (against numpy), that the inverse path returns the input, that every
LDS address stays inside the N x LS tile, and the bank behaviour of each exchange / transpose
instruction: with T = 8 the 32 lanes of a half-wavefront fall on 32 distinct 8-byte bank pairs;
with T = 4 the row phase does, and the column phase puts at most two lanes on a bank pair.
"""
import numpy as np
import pytest


def _maps(N):
    T = 8 if N % 64 == 0 else 4
    R1 = N // T
    CPT = R1 // T
    LS = N + (8 if T == 8 else 4)
    xslot = lambda j, k1: k1 * T + (j ^ (k1 // CPT))  # noqa: E731
    col_freq = lambda col: CPT * (col % T) + ((col // T) % CPT) + R1 * ((col // T) // CPT)  # noqa: E731
    return T, R1, CPT, LS, xslot, col_freq


def _line_fwd(N, xs, Tl, at):
    T, R1, CPT, LS, xslot, _ = _maps(N)
    for j in range(T):
        v = np.fft.fft(xs[j]) * np.exp(-2j * np.pi * j * np.arange(R1) / N)
        for k1 in range(R1):
            Tl[at(xslot(j, k1))] = v[k1]
    out = []
    for j in range(T):
        u = np.zeros(R1, complex)
        for c in range(CPT):
            u[c * T:(c + 1) * T] = np.fft.fft([Tl[at(xslot(jj, CPT * j + c))] for jj in range(T)])
        out.append(u)
    return out


def _line_inv(N, vs, Tl, at):
    T, R1, CPT, LS, xslot, _ = _maps(N)
    for j in range(T):
        for c in range(CPT):
            y = np.fft.ifft(vs[j][c * T:(c + 1) * T]) * T
            for jj in range(T):
                Tl[at(xslot(jj, CPT * j + c))] = y[jj]
    out = []
    for j in range(T):
        z = np.array([Tl[at(xslot(j, k1))] for k1 in range(R1)]) * np.exp(2j * np.pi * j * np.arange(R1) / N)
        out.append(np.fft.ifft(z) * R1)
    return out


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96, 128])
def test_register_tile_computes_the_2d_dft(N):
    T, R1, CPT, LS, xslot, col_freq = _maps(N)
    rng = np.random.default_rng(N)
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    Tl = np.full(N * LS, np.nan + 0j)
    for y in range(N):  # row phase, then the transpose write
        out = _line_fwd(N, [X[y, rj::T] for rj in range(T)], Tl, lambda p, y=y: y * LS + p)
        for rj in range(T):
            for c in range(CPT):
                for k2 in range(T):
                    Tl[y * LS + T * (CPT * k2 + c) + rj] = out[rj][c * T + k2]
    F = np.zeros((N, N), complex)
    for col in range(N):  # column phase
        vs = [np.array([Tl[(cj + T * n) * LS + col] for n in range(R1)]) for cj in range(T)]
        out = _line_fwd(N, vs, Tl, lambda p, col=col: p * LS + col)
        for cj in range(T):
            for c in range(CPT):
                for k2 in range(T):
                    F[CPT * cj + c + R1 * k2, col_freq(col)] = out[cj][c * T + k2]
    assert np.abs(F - np.fft.fft2(X)).max() <= 1e-12 * np.abs(F).max()
    # inverse: column phase, transpose, row phase
    Y = np.zeros((N, N), complex)
    Tl[:] = 0
    for col in range(N):
        fx = col_freq(col)
        vs = [np.array([F[CPT * cj + c + R1 * k2, fx] for c in range(CPT) for k2 in range(T)]) for cj in range(T)]
        out = _line_inv(N, vs, Tl, lambda p, col=col: p * LS + col)
        for cj in range(T):
            for n in range(R1):
                Tl[(cj + T * n) * LS + col] = out[cj][n]
    for y in range(N):
        vs = [np.array([Tl[y * LS + T * (CPT * k2 + c) + rj] for c in range(CPT) for k2 in range(T)])
              for rj in range(T)]
        out = _line_inv(N, vs, Tl, lambda p, y=y: y * LS + p)
        for rj in range(T):
            Y[y, rj::T] = out[rj]
    assert np.abs(Y / N ** 2 - X).max() <= 1e-12 * np.abs(X).max()


def _rslot(N, y, x):
    T = 8 if N % 64 == 0 else 4
    L, NT = 64 // T, N * T
    return (x // T) * NT + (y // L) * 64 + (y % L) * T + (x % T)


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96, 128])
def test_r_and_c_orders_are_permutations(N):
    """R-order slots of all pixels and C-order frequencies of all slots cover the window once."""
    T = 8 if N % 64 == 0 else 4
    L, NT = 64 // T, N * T
    _, R1, CPT, _, _, col_freq = _maps(N)
    slots = {_rslot(N, y, x) for y in range(N) for x in range(N)}
    assert slots == set(range(N * N))
    freqs = set()
    for e in range(N * N):
        i, t = e // NT, e % NT
        lane = t & 63
        freqs.add((CPT * (lane // L) + i // T + R1 * (i % T), col_freq((t >> 6) * L + lane % L)))
    assert len(freqs) == N * N


@pytest.mark.parametrize("N", [192, 256])
def test_line_split_of_large_windows(N):
    """The 24- and 32-sample line splits (3 and 4 DFT-8 groups per thread) of the 192 / 256 px
    global-slab tiles: one exchange per line through an LDS row of stride N + 8."""
    _, R1, CPT, LS, xslot, col_freq = _maps(N)
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


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96, 128])
def test_register_tile_lds_bounds_and_banks(N):
    T, R1, CPT, LS, xslot, _ = _maps(N)
    L = 64 // T

    def check(addrs, most):
        a = np.asarray(addrs)
        assert a.min() >= 0 and a.max() < N * LS
        for half in (a[:32], a[32:]):
            _, counts = np.unique(half % 32, return_counts=True)
            assert counts.max() <= most

    col_most = 1 if T == 8 else 2  # lanes per 8-byte bank pair in the column phase
    lanes = np.arange(64)
    for w in range(N // L):
        # row phase: lane = j + T*row_in_wave; column phase: lane = col_in_wave + L*j
        for k1 in range(R1):
            check([(L * w + l // T) * LS + xslot(l % T, k1) for l in lanes], 1)
            check([xslot(l // L, k1) * LS + L * w + l % L for l in lanes], col_most)
        for c in range(CPT):
            for jj in range(T):
                check([(L * w + l // T) * LS + xslot(jj, CPT * (l % T) + c) for l in lanes], 1)
                check([xslot(jj, CPT * (l // L) + c) * LS + L * w + l % L for l in lanes], col_most)
            for k2 in range(T):
                check([(L * w + l // T) * LS + T * (CPT * k2 + c) + l % T for l in lanes], 1)
        for n in range(R1):
            check([(l // L + T * n) * LS + L * w + l % L for l in lanes], col_most)


def test_register_slab_192_phase_changes():
    """The 192 px register tile's phase changes (RegSlab in pty_kernels.hip): 768 threads, two rounds
    of 96 lines, LDS of 96 rows x stride 200.  to_cols must turn every thread's row-spectrum registers
    (register c*8 + k2 of round r = position 8*(3*k2 + c) + j of row r*96 + ln) into its column
    registers (register n of round q = row j + 8n of column q*96 + ln), to_rows back; every LDS index
    stays below 96 * 200; writes put a half-wavefront on distinct bank pairs, reads on at most two."""
    N, NT, R1, CPT, LINES, LS, H = 192, 768, 24, 3, 96, 200, 12
    tids = np.arange(NT)
    ln = 8 * (tids >> 6) + ((tids & 63) >> 3)
    j = tids & 7
    tile = np.arange(N * N).reshape(N, N)
    d = np.zeros((NT, 2, R1), np.int64)
    for p in range(2):
        for c in range(CPT):
            for k2 in range(8):
                d[:, p, c * 8 + k2] = tile[p * LINES + ln, 8 * (CPT * k2 + c) + j]
    X = np.full(LINES * LS, -1, np.int64)
    t = np.zeros((NT, 2, 2, H), np.int64)
    top = 0
    for p in range(2):
        X[:] = -1
        for c in range(CPT):
            for k2 in range(8):
                idx = ln * LS + j + 8 * (CPT * k2 + c)
                top = max(top, idx.max())
                X[idx] = d[:, p, c * 8 + k2]
        for q in range(2):
            for n in range(H):
                idx = j * LS + ln + 8 * n * LS + q * LINES
                top = max(top, idx.max())
                t[:, p, q, n] = X[idx]
    assert (t >= 0).all() and top < LINES * LS
    dc = np.zeros_like(d)
    for q in range(2):
        for n in range(R1):
            dc[:, q, n] = t[:, 0, q, n] if n < H else t[:, 1, q, n - H]
    for q in range(2):
        for n in range(R1):
            assert np.array_equal(dc[:, q, n], tile[j + 8 * n, q * LINES + ln]), (q, n)
    back = np.zeros_like(d)
    for p in range(2):
        X[:] = -1
        for q in range(2):
            for n in range(H):
                X[j * LS + ln + 8 * n * LS + q * LINES] = dc[:, q, n + H * p]
        for c in range(CPT):
            for k2 in range(8):
                back[:, p, c * 8 + k2] = X[ln * LS + j + 8 * (CPT * k2 + c)]
    assert np.array_equal(back, d)

    def worst(addr):  # most lanes of one half-wavefront on one 8-byte bank pair (64 banks of 4 bytes)
        w = 0
        for h in range(NT // 32):
            a = addr[32 * h:32 * h + 32]
            w = max(w, np.bincount((a * 2) % 64, minlength=64).max())
        return w

    assert worst(ln * LS + j) == 1 and worst(j * LS + ln) <= 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_register_slab_192_zero_rounds_are_the_zero_groups(seed):
    """The 192 px register tile skips the band mask's zero columns (forward_reg192, adjoint_reg192):
    a column-phase wavefront whose columns are all marked in zcol sets its round to zero in
    registers (RegSlab::zero_round) instead of transforming it, and the next row phase reads the
    tile as is.  After to_rows the zeros must sit exactly in the row-phase registers whose group of
    eight columns gt_row_zero_groups marks, the groups the slab pipeline zeroes on load, for a
    disc-shaped band (the kernels' case) and for arbitrary marks."""
    N, NT, R1, CPT, LINES, LS, H, T = 192, 768, 24, 3, 96, 200, 12, 8
    tids = np.arange(NT)
    ln = 8 * (tids >> 6) + ((tids & 63) >> 3)
    j = tids & 7
    wave = tids >> 6

    def col_freq(col):  # RT<192>::col_freq
        g = col // T
        return CPT * (col % T) + (g % CPT) + R1 * (g // CPT)

    rng = np.random.default_rng(seed)
    f = np.fft.fftfreq(N)
    zsets = [(np.abs(f) > (1.0 / 3.0)).astype(int), (rng.random(N) < 0.5).astype(int)]
    # whole groups marked, so that some wavefronts skip
    g = rng.integers(0, N // 8, 6)
    z = np.zeros(N, int)
    for q in g:
        for c in range(8 * q, 8 * q + 8):
            z[col_freq(c)] = 1
    zsets.append(z)
    for zcol in zsets:
        # column phase: register n of round q = tile[j + 8n, q*96 + ln]; nonzero values everywhere
        dc = np.zeros((NT, 2, R1), np.int64)
        for q in range(2):
            for n in range(R1):
                dc[:, q, n] = 1 + (j + 8 * n) * N + q * LINES + ln
        for q in range(2):
            marked = zcol[col_freq(q * LINES + ln)] != 0
            for w in range(NT // 64):
                lanes = wave == w
                if marked[lanes].all():  # wave_cols_zero: the ballot of unmarked columns is empty
                    dc[lanes, q, :] = 0
        # to_rows (index maps as in the layout test above)
        X = np.zeros(LINES * LS, np.int64)
        back = np.zeros_like(dc)
        for p in range(2):
            X[:] = -1
            for q in range(2):
                for n in range(H):
                    X[j * LS + ln + 8 * n * LS + q * LINES] = dc[:, q, n + H * p]
            for c in range(CPT):
                for k2 in range(8):
                    back[:, p, c * 8 + k2] = X[ln * LS + j + 8 * (CPT * k2 + c)]
        assert (back >= 0).all()
        # gt_row_zero_groups: bit c*8 + k2 when every column 8*(CPT*k2 + c) + j' (j' < 8) is marked
        for c in range(CPT):
            for k2 in range(8):
                grp = CPT * k2 + c
                zero_group = all(zcol[col_freq(8 * grp + jj)] != 0 for jj in range(8))
                for p in range(2):
                    vals = back[:, p, c * 8 + k2]
                    if zero_group:
                        assert (vals == 0).all(), (c, k2, p)
                    else:
                        assert (vals != 0).all(), (c, k2, p)
