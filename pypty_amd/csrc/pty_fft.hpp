// Device-side complex64 FFT building blocks for gfx950 (CDNA4).
//
// A window of N x N complex64 samples lives in LDS as a padded tile (row
// stride N+1 elements, so that both row-wise and column-wise accesses of a
// wavefront fall into distinct banks).  A 2D transform is two line passes; a
// line pass is a two-stage Cooley-Tukey split N = R1 * R2:
//
//   stage 1: R2 threads per line, each holding R1 samples (stride R2) in
//            registers -> radix-R1 DFT -> twiddle W_N^(j*k1) -> back to LDS
//   stage 2: each thread takes R1/R2 groups of R2 samples -> radix-R2 DFT
//
// Both stages work in place, so the forward transform leaves the spectrum in
// a fixed "digit-transposed" order: tile position p = a + R2*b holds
// frequency f = b + R1*a.  The inverse transform runs the stages backwards
// and consumes exactly that order, returning natural spatial order.  Every
// spectral-domain operation on the multislice path is pointwise (propagator,
// band mask, far-field modulus, loss), so the spectrum is never reordered in
// LDS; tables that multiply it are stored pre-permuted ("position order").
//
// Radix-R in-register DFTs are built from radix-2/3/4 butterflies by
// compile-time recursion (textbook mixed-radix Cooley-Tukey); twiddle
// constants of the small radices are evaluated at compile time from their
// Taylor series, tile twiddles W_N^m come from a per-N table in global
// memory that stays L1/L2 resident.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pty {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------- compile-time trigonometry
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 30; ++i) {
    term *= -x * x / double((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 30; ++i) {
    term *= -x * x / double((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}

template <int R>
struct TwTab {
  float re[R];
  float im[R];
};

// W_R^k = exp(-2*pi*i*k/R), angle reduced to [-pi, pi] before the series.
template <int R>
constexpr TwTab<R> make_twtab() {
  TwTab<R> t{};
  for (int k = 0; k < R; ++k) {
    int kk = (2 * k > R) ? k - R : k;
    double a = -2.0 * kPi * double(kk) / double(R);
    t.re[k] = float(cx_cos(a));
    t.im[k] = float(cx_sin(a));
  }
  return t;
}

// multiply v by W_R^e (conjugate when INV); e is a compile-time constant after
// unrolling, so the trivial cases fold away.
template <int R, bool INV>
__device__ __forceinline__ float2 twiddle_const(float2 v, int e) {
  constexpr TwTab<R> T = make_twtab<R>();
  e %= R;
  if (e == 0) return v;
  if (4 * e == R) return mul_mi<INV>(v);
  if (2 * e == R) return make_float2(-v.x, -v.y);
  if (4 * e == 3 * R) return mul_mi<!INV>(v);
  float2 w = make_float2(T.re[e], INV ? -T.im[e] : T.im[e]);
  return cmul(v, w);
}

// ------------------------------------------------------ in-register DFTs
template <int R>
constexpr int first_radix() {
  return (R % 4 == 0 && R != 8) ? 4
                                : (R % 2 == 0 ? 2 : (R % 3 == 0 ? 3 : (R % 5 == 0 ? 5 : (R % 7 == 0 ? 7 : R))));
}

// direct O(R^2) DFT for small odd primes (5, 7): X[k] = sum_n x[n] W_R^(nk)
template <int R, bool INV>
__device__ __forceinline__ void dft_direct(float2* v) {
  float2 out[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float2 acc = v[0];
#pragma unroll
    for (int n = 1; n < R; ++n) acc = cadd(acc, twiddle_const<R, INV>(v[n], (n * k) % R));
    out[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

// Unnormalised DFT of v[0..R) in natural order (sign -1 forward, +1 inverse).
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    // X0 = a+b+c ; X1,2 = a - (b+c)/2 -/+ i*(sqrt3/2)*(b-c)   (forward)
    constexpr float s3 = 0.86602540378443864676f;
    float2 a = v[0], b = v[1], c = v[2];
    float2 t = cadd(b, c);
    float2 d = csub(b, c);
    float2 m = make_float2(a.x - 0.5f * t.x, a.y - 0.5f * t.y);
    // -i*s3*d (forward) ; +i*s3*d (inverse)
    float2 r = INV ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);
    v[0] = cadd(a, t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if constexpr (R == 4) {
    float2 a = v[0], b = v[1], c = v[2], d = v[3];
    float2 s0 = cadd(a, c), d0 = csub(a, c);
    float2 s1 = cadd(b, d), d1 = mul_mi<INV>(csub(b, d));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
  } else if constexpr (R == 5 || R == 7) {
    dft_direct<R, INV>(v);
  } else {
    // R = P*Q, n = Q*n1 + n2, k = k1 + P*k2
    constexpr int P = first_radix<R>();
    constexpr int Q = R / P;
    static_assert(P * Q == R, "radix");
    float2 tmp[R];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      float2 a[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) a[n1] = v[Q * n1 + n2];
      dft<P, INV>(a);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) tmp[k1 * Q + n2] = twiddle_const<R, INV>(a[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      dft<Q, INV>(&tmp[k1 * Q]);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = tmp[k1 * Q + k2];
    }
  }
}

// --------------------------------------------------------- tile geometry
// Upper bound on threads per tile workgroup.  With 1024 threads a 128 x 128
// tile gives each thread 16 samples but caps it at 128 VGPRs; 512 threads
// double the samples per thread and the register budget.
#ifndef PTY_MAX_NT
#define PTY_MAX_NT 1024
#endif
// Split N = R1 * R2 with R2 | R1; R2 threads cooperate on one line.
template <int N>
struct Plan;
template <> struct Plan<16>  { static constexpr int R2 = 4; };
template <> struct Plan<32>  { static constexpr int R2 = 4; };
template <> struct Plan<48>  { static constexpr int R2 = 4; };
template <> struct Plan<64>  { static constexpr int R2 = 8; };
template <> struct Plan<96>  { static constexpr int R2 = 4; };
template <> struct Plan<128> { static constexpr int R2 = 8; };
template <> struct Plan<192> { static constexpr int R2 = 8; };
template <> struct Plan<256> { static constexpr int R2 = 16; };

// Windows up to 128 px live in LDS (128 x 129 x 8 B = 129 KB of the 160 KB).
// Larger windows (192, 256) do not fit: their tile lives in a per-workgroup
// slab of global memory (L2 / MALL resident), with the same line passes.
template <int N>
struct Tile {
  static constexpr int R2 = Plan<N>::R2;
  static constexpr int R1 = N / R2;
  static_assert(R1 % R2 == 0, "R2 must divide R1");
  static constexpr bool GLOBAL = N > 128;
  static constexpr int LS = GLOBAL ? N : N + 1; // row stride (elements), padded in LDS
  static constexpr int ELEMS = N * LS;          // LDS elements per tile
  // threads per tile; global-memory tiles use 512 so each thread may hold
  // 256 VGPRs (their radix-24 / radix-16 stages do not fit in 128)
  static constexpr int NT_CAP = GLOBAL ? 512 : PTY_MAX_NT;
  static constexpr int NT = (N * R2 > NT_CAP) ? NT_CAP : N * R2;
  static constexpr int PER_THREAD = N * N / NT; // pointwise elements per thread
  static_assert((N * N) % NT == 0, "tile/threads");
  __device__ __forceinline__ static int lds(int y, int x) { return y * LS + x; }
  // position index -> frequency index along one axis
  __host__ __device__ __forceinline__ static int pos2freq(int p) { return p / R2 + R1 * (p % R2); }
  // frequency index -> position index along one axis
  __host__ __device__ __forceinline__ static int freq2pos(int f) { return f / R1 + R2 * (f % R1); }
};

// w / N is the same for the 64 lanes of a wavefront when N is a multiple of 64
// (lanes take consecutive lines): make that explicit so the twiddle indices
// live in scalar registers and the table reads become scalar-cache loads.
template <int N>
__device__ __forceinline__ int wave_uniform(int v) {
  if constexpr (N % 64 == 0)
    return __builtin_amdgcn_readfirstlane(v);
  else
    return v;
}

// One line pass over the whole tile.  ALONG_X: lines are rows (transform over
// x); otherwise lines are columns.  tw[m] = W_N^m, m in [0, N).
template <int N, bool INV, bool ALONG_X>
__device__ __forceinline__ void line_pass(float2* __restrict__ t, const float2* __restrict__ tw) {
  using T = Tile<N>;
  constexpr int R1 = T::R1, R2 = T::R2, LS = T::LS, NT = T::NT;
  constexpr int WORK = N * R2;
  // lanes walk across lines (conflict-free LDS, coalesced columns) except for
  // rows of a global-memory tile, where the R2 lanes of a line are adjacent so
  // that a wavefront reads contiguous row segments
  constexpr bool JFAST = T::GLOBAL && ALONG_X;
  const int tid = threadIdx.x;
  auto at = [](int line, int pos) { return ALONG_X ? line * LS + pos : pos * LS + line; };
  auto line_of = [](int w) { return JFAST ? w / R2 : w % N; };
  auto part_of = [](int w) { return JFAST ? w % R2 : wave_uniform<N>(w / N); };
  if constexpr (!INV) {
    for (int w = tid; w < WORK; w += NT) {
      const int line = line_of(w), j = part_of(w);
      float2 v[R1];
#pragma unroll
      for (int n = 0; n < R1; ++n) v[n] = t[at(line, j + R2 * n)];
      dft<R1, false>(v);
#pragma unroll
      for (int k = 1; k < R1; ++k) v[k] = cmul(v[k], tw[j * k]);
#pragma unroll
      for (int k = 0; k < R1; ++k) t[at(line, j + R2 * k)] = v[k];
    }
    __syncthreads();
    for (int w = tid; w < WORK; w += NT) {
      const int line = line_of(w), g = part_of(w);
#pragma unroll
      for (int c = 0; c < R1 / R2; ++c) {
        const int k1 = g * (R1 / R2) + c;
        float2 u[R2];
#pragma unroll
        for (int a = 0; a < R2; ++a) u[a] = t[at(line, a + R2 * k1)];
        dft<R2, false>(u);
#pragma unroll
        for (int a = 0; a < R2; ++a) t[at(line, a + R2 * k1)] = u[a];
      }
    }
    __syncthreads();
  } else {
    for (int w = tid; w < WORK; w += NT) {
      const int line = line_of(w), g = part_of(w);
#pragma unroll
      for (int c = 0; c < R1 / R2; ++c) {
        const int b = g * (R1 / R2) + c;
        float2 u[R2];
#pragma unroll
        for (int a = 0; a < R2; ++a) u[a] = t[at(line, a + R2 * b)];
        dft<R2, true>(u);
#pragma unroll
        for (int j = 1; j < R2; ++j) u[j] = cmulc(u[j], tw[b * j]);
#pragma unroll
        for (int j = 0; j < R2; ++j) t[at(line, j + R2 * b)] = u[j];
      }
    }
    __syncthreads();
    for (int w = tid; w < WORK; w += NT) {
      const int line = line_of(w), j = part_of(w);
      float2 v[R1];
#pragma unroll
      for (int n = 0; n < R1; ++n) v[n] = t[at(line, j + R2 * n)];
      dft<R1, true>(v);
#pragma unroll
      for (int n = 0; n < R1; ++n) t[at(line, j + R2 * n)] = v[n];
    }
    __syncthreads();
  }
}

// Forward 2D transform: natural order in, position order out (unnormalised).
template <int N>
__device__ __forceinline__ void tile_fft_fwd(float2* t, const float2* tw) {
  line_pass<N, false, true>(t, tw);
  line_pass<N, false, false>(t, tw);
}
// Inverse 2D transform: position order in, natural order out (unnormalised;
// the caller folds the 1/N^2 into an adjacent pointwise product).
template <int N>
__device__ __forceinline__ void tile_fft_inv(float2* t, const float2* tw) {
  line_pass<N, true, false>(t, tw);
  line_pass<N, true, true>(t, tw);
}

}  // namespace pty
