// Register-resident window tile for N = 16, 32, 48, 64, 96 and 128 (gfx950).
//
// Each thread keeps R1 = N/T samples of one line in registers; the T threads
// of a line (T = 8 for N a multiple of 64, else 4) are T lanes of the same
// wavefront.  A 1D transform of a line is the split N = R1 * T:
//
//   forward: DFT-R1 over the thread's samples x[j + T*n] (n = 0..R1-1),
//            twiddle W_N^(j*k1), exchange within the line's T lanes through
//            LDS, then R1/T DFT-Ts over j -> X[k1 + R1*k2]
//   inverse: the same steps backwards
//
// The exchange only involves lanes of one wavefront, and LDS instructions of
// a wavefront execute in issue order, so it needs no workgroup barrier.
//
// The window is processed in two phases that alternate (L = 64/T lines per
// wavefront):
//   row phase    : wavefront w owns rows Lw..Lw+L-1 (lane = j + T*row_in_wave)
//   column phase : wavefront w owns columns Lw..Lw+L-1 (lane = col_in_wave + L*j)
// Every step of the multislice recursion that is local to a row (inverse row
// transform, stash write, object product, forward row transform) runs in the
// row phase, every step local to a column (forward column transform,
// propagator, inverse column transform) in the column phase.  Switching phase
// is one transpose through an LDS tile (row stride LS) between two workgroup
// barriers.  With T = 8 (LS = N + 8) every LDS access pattern (exchange
// slots, transposes) puts the 32 lanes of a half-wavefront on 32 distinct
// 8-byte bank pairs; with T = 4 (LS = N + 4) the row phase does, and the
// column phase's accesses meet at most two lanes per bank pair.
//
// Global layouts that go with it (e = i * NT + tid, fully coalesced):
//   R-order: register i of thread tid in the row phase = natural pixel
//            (y, x) = (Lw + rr, j + Ti)
//   C-order: register i of thread tid in the column phase = frequency
//            (fy, fx) after a forward 2D transform (cspec below)
#pragma once
#include "pty_fft.hpp"

namespace pty {

template <int N>
struct RT {
  // 16 to 128 px: the whole window in registers; 192 and 256 px: the line
  // split alone (GT below)
  static_assert(N == 16 || N == 32 || N == 48 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256,
                "register tile line split: N = 16, 32, 48, 64, 96, 128, 192, 256");
  static constexpr int T = (N % 64 == 0) ? 8 : 4;  // threads per line
  static constexpr int L = 64 / T;                 // lines per wavefront
  static constexpr int R1 = N / T;                 // samples per thread and line
  static constexpr int CPT = R1 / T;               // DFT-T groups per thread in stage 2
  static_assert(CPT * T * T == N, "line split");
  static constexpr int NT = N * T;                 // threads per window
  static constexpr int LS = N + (T == 8 ? 8 : 4);  // LDS row stride (float2)
  static constexpr int ELEMS = N * LS;

  // exchange slot of (thread j, stage-1 output k1) within a line
  __host__ __device__ __forceinline__ static int xslot(int j, int k1) { return k1 * T + (j ^ (k1 / CPT)); }
  // column position of frequency f after a row transform, and back
  __host__ __device__ __forceinline__ static int col_freq(int col) {
    const int g = col / T;
    return CPT * (col % T) + (g % CPT) + R1 * (g / CPT);
  }
  // R-order slot of natural pixel (y, x)
  __host__ __device__ __forceinline__ static int rslot(int y, int x) {
    return (x / T) * NT + (y / L) * 64 + (y % L) * T + (x % T);
  }
  // natural pixel of R-order slot e
  __host__ __device__ __forceinline__ static void rpos(int e, int& y, int& x) {
    const int i = e / NT, t = e % NT;
    y = (t >> 6) * L + (t & 63) / T;
    x = (t & 63) % T + T * i;
  }
  // frequency (fy, fx) of C-order slot e
  __host__ __device__ __forceinline__ static void cspec(int e, int& fy, int& fx) {
    const int i = e / NT, t = e % NT, lane = t & 63;
    fx = col_freq((t >> 6) * L + lane % L);
    fy = CPT * (lane / L) + i / T + R1 * (i % T);
  }
};

// Orders the LDS accesses of one wavefront around an exchange: nothing is
// moved across it by the compiler, and the wavefront's own writes are
// complete before its next reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddle factors of the line split, indexed [row][j] (j = 0..T-1) so that a
// lane reads row r at j + r*T: one lane offset per pass and compile-time row
// offsets.  Rows 0..3 hold W_N^(r*j), rows 3+q (q = 1..R1/4-1) W_N^(4*q*j);
// W_N^(j*k1) = W^(j*(k1 & 3)) * W^(4*j*(k1 >> 2)) costs one product.  (A full
// table of every W_N^(j*k1), read through the vector memory path, removed 8% of
// the VALU instructions but made the forward and adjoint 10% slower.)
template <int N>
struct RTTw {
  static constexpr int T = RT<N>::T;
  static constexpr int ROWS = 3 + RT<N>::R1 / 4;
  float v[ROWS * T * 2];
};
template <int N>
constexpr RTTw<N> make_rt_tw() {
  constexpr int T = RT<N>::T;
  RTTw<N> t{};
  for (int row = 0; row < RTTw<N>::ROWS; ++row) {
    const int mult = row < 4 ? row : 4 * (row - 3);
    for (int j = 0; j < T; ++j) {
      int e = (mult * j) % N;
      if (2 * e > N) e -= N;
      const double a = -2.0 * kPi * double(e) / double(N);
      t.v[(row * T + j) * 2] = float(cx_cos(a));
      t.v[(row * T + j) * 2 + 1] = float(cx_sin(a));
    }
  }
  return t;
}
template <int N>
__constant__ __align__(16) RTTw<N> g_rt_tw = make_rt_tw<N>();

// The workgroup's LDS copy of the twiddle table.  Reading the six rows of a
// line transform from LDS instead of through the vector memory path took
// 8-10% off the forward and adjoint kernels: their memory instructions (object,
// propagator, stash) already keep that path busy.
template <int N>
__device__ __forceinline__ float2* rt_tw_lds() {
  __shared__ float2 t[RTTw<N>::ROWS * RT<N>::T];
  return t;
}
// Every kernel that runs line transforms calls this first, with the whole
// workgroup: it copies the twiddle table into LDS and synchronises.
template <int N>
__device__ __forceinline__ void rt_tw_load() {
  constexpr int E = RTTw<N>::ROWS * RT<N>::T;
  float2* t = rt_tw_lds<N>();
  for (int i = threadIdx.x; i < E; i += blockDim.x)
    t[i] = make_float2(g_rt_tw<N>.v[2 * i], g_rt_tw<N>.v[2 * i + 1]);
  __syncthreads();
}

// v[k1] *= W_N^(+-j*k1) for k1 = 1..R1-1
template <int N, bool INV>
__device__ __forceinline__ void rt_twiddle(float2* v, unsigned j) {
  constexpr int R1 = RT<N>::R1;
  const float2* t = rt_tw_lds<N>() + j;
  auto row = [&](int r) { return t[r * RT<N>::T]; };
  float2 a[4], b[R1 / 4];
  a[0] = make_float2(1.f, 0.f);
  b[0] = a[0];
#pragma unroll
  for (int r = 1; r < 4; ++r) a[r] = row(r);
#pragma unroll
  for (int q = 1; q < R1 / 4; ++q) b[q] = row(3 + q);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) {
    const int r = k1 & 3, q = k1 >> 2;
    const float2 w = (q == 0) ? a[r] : (r == 0 ? b[q] : cmul(a[r], b[q]));
    v[k1] = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
  }
}

// Exchange addressing of one line.  Slot s = T*k1 + (j ^ q), q = k1 / CPT,
// lies at line*LS + s (rows) or s*LS + line (columns).  With k1 = CPT*q + c:
//   write of (j, k1) : wbase(q) + c*T*step,  wbase(q) = ((CPT*q)*T + (j ^ q))*step + line0
//   read  of (jj, k1 = CPT*j + c) : rbase(jj) + c*T*step,
//                                   rbase(jj) = ((CPT*j)*T + (jj ^ j))*step + line0
// so one exchange needs T lane addresses and compile-time offsets.  The line
// origin and j pass through an empty asm so that these addresses (and the
// twiddle offsets) are recomputed at every exchange instead of being kept live
// across the kernel.
template <int N, bool ROW>
struct XAddr {
  static constexpr int STEP = ROW ? 1 : RT<N>::LS;
  int line0;
  __device__ __forceinline__ XAddr(int line) {
    line0 = ROW ? line * RT<N>::LS : line;
    asm volatile("" : "+v"(line0));
  }
  __device__ __forceinline__ static int opaque(int j) {
    asm volatile("" : "+v"(j));
    return j;
  }
  __device__ __forceinline__ int w(int j, int q) const {
    return ((RT<N>::CPT * q) * RT<N>::T + (j ^ q)) * STEP + line0;
  }
  __device__ __forceinline__ int r(int j, int jj) const {
    return ((RT<N>::CPT * j) * RT<N>::T + (jj ^ j)) * STEP + line0;
  }
};

// forward line transform: in v[n] = x[j + Tn]; out v[c*T + k2] = X[CPT*j + c + R1*k2]
template <int N, bool ROW>
__device__ __forceinline__ void rt_line_fwd(float2* v, float2* T, int line, int j) {
  using R = RT<N>;
  constexpr int CS = R::T * XAddr<N, ROW>::STEP;  // offset of c
  j = XAddr<N, ROW>::opaque(j);
  dft<R::R1, false>(v);
  rt_twiddle<N, false>(v, unsigned(j));
  const XAddr<N, ROW> X(line);
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < R::T; ++q) {
    float2* p = T + X.w(j, q);
#pragma unroll
    for (int c = 0; c < R::CPT; ++c) p[c * CS] = v[R::CPT * q + c];
  }
  wave_lds_sync();
#pragma unroll
  for (int jj = 0; jj < R::T; ++jj) {
    const float2* p = T + X.r(j, jj);
#pragma unroll
    for (int c = 0; c < R::CPT; ++c) v[c * R::T + jj] = p[c * CS];
  }
#pragma unroll
  for (int c = 0; c < R::CPT; ++c) dft<R::T, false>(&v[c * R::T]);
}

// inverse line transform (unnormalised): exact reverse of rt_line_fwd
template <int N, bool ROW>
__device__ __forceinline__ void rt_line_inv(float2* v, float2* T, int line, int j) {
  using R = RT<N>;
  constexpr int CS = R::T * XAddr<N, ROW>::STEP;
  j = XAddr<N, ROW>::opaque(j);
#pragma unroll
  for (int c = 0; c < R::CPT; ++c) dft<R::T, true>(&v[c * R::T]);
  const XAddr<N, ROW> X(line);
  wave_lds_sync();
#pragma unroll
  for (int jj = 0; jj < R::T; ++jj) {
    float2* p = T + X.r(j, jj);
#pragma unroll
    for (int c = 0; c < R::CPT; ++c) p[c * CS] = v[c * R::T + jj];
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < R::T; ++q) {
    const float2* p = T + X.w(j, q);
#pragma unroll
    for (int c = 0; c < R::CPT; ++c) v[R::CPT * q + c] = p[c * CS];
  }
  rt_twiddle<N, true>(v, unsigned(j));
  dft<R::R1, true>(v);
}

// Thread coordinates of both phases.
template <int N>
struct RTIds {
  int row, rj;  // row phase: line and position within the line
  int col, cj;  // column phase
  __device__ __forceinline__ RTIds() {
    constexpr int T = RT<N>::T, L = RT<N>::L;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    row = L * w + lane / T;
    rj = lane % T;
    col = L * w + lane % L;
    cj = lane / L;
  }
};

// row phase -> column phase: spectrum of the thread's row goes to column
// position T*(CPT*k2 + c) + rj of that row; barrier; read the thread's column
template <int N>
__device__ __forceinline__ void rt_rows_to_cols(float2* v, float2* T, const RTIds<N>& id) {
  using R = RT<N>;
  int rb = id.row * R::LS + id.rj, cb = id.cj * R::LS + id.col;
  asm volatile("" : "+v"(rb), "+v"(cb));
#pragma unroll
  for (int c = 0; c < R::CPT; ++c) {
#pragma unroll
    for (int k2 = 0; k2 < R::T; ++k2) T[rb + R::T * (R::CPT * k2 + c)] = v[c * R::T + k2];
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < R::R1; ++n) v[n] = T[cb + R::T * n * R::LS];
}

// column phase -> row phase: the exact reverse
template <int N>
__device__ __forceinline__ void rt_cols_to_rows(float2* v, float2* T, const RTIds<N>& id) {
  using R = RT<N>;
  int rb = id.row * R::LS + id.rj, cb = id.cj * R::LS + id.col;
  asm volatile("" : "+v"(rb), "+v"(cb));
#pragma unroll
  for (int n = 0; n < R::R1; ++n) T[cb + R::T * n * R::LS] = v[n];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R::CPT; ++c) {
#pragma unroll
    for (int k2 = 0; k2 < R::T; ++k2) v[c * R::T + k2] = T[rb + R::T * (R::CPT * k2 + c)];
  }
}

// Forward 2D transform: row phase natural samples -> column phase spectrum.
template <int N>
__device__ __forceinline__ void rt_fft2_fwd(float2* v, float2* T, const RTIds<N>& id) {
  rt_line_fwd<N, true>(v, T, id.row, id.rj);
  rt_rows_to_cols<N>(v, T, id);
  rt_line_fwd<N, false>(v, T, id.col, id.cj);
}
// Inverse 2D transform (unnormalised): column phase spectrum -> row phase samples.
template <int N>
__device__ __forceinline__ void rt_fft2_inv(float2* v, float2* T, const RTIds<N>& id) {
  rt_line_inv<N, false>(v, T, id.col, id.cj);
  rt_cols_to_rows<N>(v, T, id);
  rt_line_inv<N, true>(v, T, id.row, id.rj);
}

// Windows of 192 and 256 px do not fit in one CU's registers and LDS.  The
// same line split runs in rounds of LINES lines per workgroup, with the
// workgroup's N x N slab in global memory (L2 / MALL resident) as the
// transpose medium between the row phase and the column phase, and LDS only
// for the wave-local exchanges.  Both phases use the row-phase lane mapping
// (lane = j + 8 * line_in_wave); the slab is row-major, so a wavefront's
// accesses are eight 64-byte segments in either phase.  The slab holds
// natural rows after the column phase and row spectra (column positions,
// frequency col_freq) after the row phase.
//   R-order slot (r*R1 + i)*NT + tid: natural pixel (r*LINES + ln, j + 8i)
//   C-order slot (r*R1 + i)*NT + tid: frequency (CPT*j + (i>>3) + R1*(i&7),
//                                                 col_freq(r*LINES + ln))
// with ln = 8*(tid/64) + (tid%64)/8 and j = tid%8.
template <int N, int NT_>
struct GT {
  static constexpr int NT = NT_;
  static_assert(RT<N>::T == 8, "the slab tile splits lines over 8 threads");
  static constexpr int R1 = RT<N>::R1, CPT = RT<N>::CPT, LS = RT<N>::LS;
  static constexpr int LINES = NT / 8;      // lines per round
  static constexpr int ROUNDS = N / LINES;  // rounds per phase
  static_assert(N % LINES == 0, "rounds");
  __host__ __device__ __forceinline__ static void cspec(int e, int& fy, int& fx) {
    const int q = e / NT, t = e % NT, r = q / R1, i = q % R1;
    fx = RT<N>::col_freq(r * LINES + 8 * (t >> 6) + ((t & 63) >> 3));
    fy = CPT * (t & 7) + (i >> 3) + R1 * (i & 7);
  }
};

}  // namespace pty
