// MI355X (gfx950) kernels of the far-field multislice ptychography hot path
// and the C ABI that launches them.
//
// One workgroup owns one n x n window held in LDS (see pty_fft.hpp).  The
// per-batch pipeline of pty_loss_and_direction is
//
//   k_patch      (b, s, k)  cut object window, band-limit           2 FFT
//   k_probe_shift(b, m)     sub-pixel shift of the probe             1 FFT
//   k_forward    (b, m, k)  S slices: stash, transmit, propagate,    2(S-1)+1 FFT
//                           far field (spectrum kept in position order)
//   k_loss       (b)        mode-mean intensity, loss/sse (FP64), dL/dI
//   k_adjoint    (b, m, k)  far-field adjoint, S slices backwards,   1+2(S-1) FFT
//                           per-slice object gradient summed over modes
//   k_objgrad    (b, s, k)  sum the per-mode window gradients, band-  2 FFT
//                           limit them
//   k_gather_grad (tile)    add every overlapping window gradient to
//                           each object tile (deterministic scatter-add)
//   k_probegrad  (g, m)     probe-shift adjoint, position gradient   1 FFT
//                           (FP64 sums), probe gradient in the spectrum
//
// Workgroups of the same window b but different modes are placed on the same
// XCD (blockIdx % 8 equal) so that they share the window's object patch and
// gradient accumulator in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "pty_fft.hpp"
#include "pty_rt.hpp"
#include "pypty_amd.h"

using namespace pty;

// ------------------------------------------------------------ twiddles
template <int N>
struct TwIL {
  float v[2 * N];
};
template <int N>
constexpr TwIL<N> make_twil() {
  TwIL<N> t{};
  for (int k = 0; k < N; ++k) {
    int kk = (2 * k > N) ? k - N : k;
    double a = -2.0 * kPi * double(kk) / double(N);
    t.v[2 * k] = float(cx_cos(a));
    t.v[2 * k + 1] = float(cx_sin(a));
  }
  return t;
}
template <int N>
__constant__ __align__(16) TwIL<N> g_tw = make_twil<N>();

template <int N>
__device__ __forceinline__ const float2* twiddles() {
  return reinterpret_cast<const float2*>(g_tw<N>.v);
}

// ------------------------------------------------------------- helpers
// The window tile of a workgroup: LDS up to 128 px, else its own slab of the
// global scratch buffer (one tile per workgroup of the launch).
#define PTY_TILE(T, t, scratch)                                                 \
  __shared__ float2 t##_lds[T::GLOBAL ? 1 : T::ELEMS];                          \
  float2* t = T::GLOBAL ? (scratch) + size_t(blockIdx.x) * T::ELEMS : t##_lds
// numpy.fft.fftfreq(n, d)[f] rounded to float32 (computed in double as numpy does)
__device__ __forceinline__ float fftfreq_f(int f, int n, float d) {
  int k = (2 * f < n) ? f : f - n;
  double val = 1.0 / (double(n) * double(d));
  return float(double(k) * val);
}

__device__ __forceinline__ void xcd_decode(int bid, int mk_count, int& b, int& mk) {
  const int xcd = bid & 7, slot = bid >> 3;
  b = (slot / mk_count) * 8 + xcd;
  mk = slot % mk_count;
}
static inline int xcd_grid(int B, int mk_count) { return ((B + 7) / 8) * 8 * mk_count; }

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  // wave64 reduction with shuffles, then across waves through LDS
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += red[i];
  }
  return s;  // valid in thread 0
}

__device__ __forceinline__ void atomic_add_c(float2* p, float2 v) {
  unsafeAtomicAdd(&p->x, v.x);
  unsafeAtomicAdd(&p->y, v.y);
}

// ------------------------------------------------------- generic tiles
enum { MODE_FWD = 0, MODE_INV = 1, MODE_FWD_SHIFT = 2, MODE_INV_ISHIFT = 3, MODE_BAND = 4 };

template <int N, int MODE>
__global__ __launch_bounds__(Tile<N>::NT) void k_fft2_tiles(float2* __restrict__ data,
                                                            const float* __restrict__ mask_nat) {
  using T = Tile<N>;
  __shared__ float2 t[T::ELEMS];
  const float2* tw = twiddles<N>();
  float2* g = data + size_t(blockIdx.x) * N * N;
  const int tid = threadIdx.x;
  if (MODE == MODE_INV || MODE == MODE_INV_ISHIFT) {
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT, y = e / N, x = e % N;
      const int fy = (MODE == MODE_INV_ISHIFT) ? (y + N / 2) % N : y;
      const int fx = (MODE == MODE_INV_ISHIFT) ? (x + N / 2) % N : x;
      t[T::lds(T::freq2pos(fy), T::freq2pos(fx))] = g[e];
    }
    __syncthreads();
    tile_fft_inv<N>(t, tw);
    const float s = 1.0f / float(N * N);
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      g[e] = cscale(t[T::lds(e / N, e % N)], s);
    }
    return;
  }
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    t[T::lds(e / N, e % N)] = g[e];
  }
  __syncthreads();
  tile_fft_fwd<N>(t, tw);
  if (MODE == MODE_BAND) {
    const float s = 1.0f / float(N * N);
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT, py = e / N, px = e % N;
      const float m = mask_nat[T::pos2freq(py) * N + T::pos2freq(px)];
      t[T::lds(py, px)] = cscale(t[T::lds(py, px)], m * s);
    }
    __syncthreads();
    tile_fft_inv<N>(t, tw);
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      g[e] = t[T::lds(e / N, e % N)];
    }
    return;
  }
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT, y = e / N, x = e % N;
    const int fy = (MODE == MODE_FWD_SHIFT) ? (y + N / 2) % N : y;
    const int fx = (MODE == MODE_FWD_SHIFT) ? (x + N / 2) % N : x;
    g[e] = t[T::lds(T::freq2pos(fy), T::freq2pos(fx))];
  }
}

// --------------------------------------------------------------- tables
// Propagator H = exp(-i*pi*dz*(lambda*q^2 + 2*(qx*tx + qy*ty))) * mask
// (objective.py:337-345), its first-slice power H^e (multislice_core.py:73)
// and the band mask q^2 <= (max(qx)*cutoff)^2 (utils.py:153-199), maskc,
// stored in the order of the fused forward/adjoint kernels: C-order of the
// register tile (N <= 128) or of the slab tile (192, 256).  H and H^e carry
// the 1/n^2 of the inverse FFT.
struct TableArgs {
  float2* H;   // one table per propagation (blockIdx.y) with per-slice distances
  float2* H0;
  float* maskc;
  int* zcol;  // (N) 1 where column fx of the band mask is zero for every fy
  float dz, lambda, px, py, tx, ty, cutoff, edge;
  const float* dists;  // per-slice distances (blockIdx.y indexes them), or null for dz
  const float* tilt0;  // tilt row 0: its inside tilt (columns 3, 2) replaces tx, ty; or null
};
// float32 products/sums without contraction, in numpy's order
__device__ __forceinline__ float fmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fadd(float a, float b) { return __fadd_rn(a, b); }

// Spectral order of H, H0, maskc, Psi and coef for the fused kernels of a
// window size: register-tile C-order (16 to 128 px) or global-slab C-order
// (192, 256 px).
enum { LAY_RT = 1, LAY_GT = 2 };
// threads of a 192 / 256 px workgroup: 768 at 192 px (two rounds of 96
// lines), 512 at 256 px (four rounds of 64 lines)
template <int N>
constexpr int kGtThreads = (N == 192) ? 768 : 512;
constexpr int kMaxLanes = 4;                 // batch lanes of a pass (pty_problem.lanes)
constexpr int kLayoutOf(int n) {
  return (n <= 128) ? LAY_RT : LAY_GT;
}
template <int N>
constexpr int kLayout = kLayoutOf(N);

template <int N, int LAY>
__host__ __device__ __forceinline__ void spec_of(int e, int& fy, int& fx) {
  if constexpr (LAY == LAY_RT)
    RT<N>::cspec(e, fy, fx);
  else
    GT<N, kGtThreads<N>>::cspec(e, fy, fx);
}

template <int N>
__device__ __forceinline__ float band_mask(const TableArgs& a, int fy, int fx) {
  const float qx = fftfreq_f(fx, N, a.px), qy = fftfreq_f(fy, N, a.py);
  const float q2 = fadd(fmul(qx, qx), fmul(qy, qy));
  const float max_x = fftfreq_f(N / 2 - 1, N, a.px);
  float c = a.cutoff;
  if (c < 0.f) c = float(2.0 / 3.0);
  const float rm = fmul(max_x, c);
  const float r_max = fmul(rm, rm);
  return (q2 <= r_max) ? 1.f : 0.f;
}

template <int N, int LAY>
__global__ void k_tables(TableArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const int t = blockIdx.y;  // the propagation after slice t (one table per distance)
  int fy, fx;
  spec_of<N, LAY>(e, fy, fx);
  const float qx = fftfreq_f(fx, N, a.px), qy = fftfreq_f(fy, N, a.py);
  const float q2 = fadd(fmul(qx, qx), fmul(qy, qy));
  const float m = band_mask<N>(a, fy, fx);
  const float dz = a.dists ? a.dists[t] : a.dz;
  const float tx = a.tilt0 ? a.tilt0[3] : a.tx, ty = a.tilt0 ? a.tilt0[2] : a.ty;
  const float k = fmul(-3.141592654f, dz);
  const float r = fadd(fmul(a.lambda, q2), fmul(2.f, fadd(fmul(qx, tx), fmul(qy, ty))));
  const float ang = fmul(k, r);
  float sn, cs;
  sincosf(ang, &sn, &cs);
  const float s = 1.0f / float(N * N);
  a.H[size_t(t) * N * N + e] = make_float2(cs * m * s, sn * m * s);
  if (t > 0) return;
  if (a.edge > 1.f) {
    // complex power through the principal branch, as numpy's z**e
    const float th = atan2f(sn * m, cs * m);
    const float mag = (m > 0.f) ? 1.f : 0.f;
    float s0, c0;
    sincosf(a.edge * th, &s0, &c0);
    a.H0[e] = make_float2(c0 * mag * m * s, s0 * mag * m * s);
  } else {
    a.H0[e] = a.H[e];
  }
  a.maskc[e] = m;
  if (e < N && a.zcol) {
    unsigned any = 0;
    for (int fy = 0; fy < N; ++fy) any |= unsigned(band_mask<N>(a, fy, e) != 0.f);
    a.zcol[e] = any ? 0 : 1;
  }
}

// (n, m, kind) of aberration j in Krivanek order (utils.py:1272-1308): n =
// 1, 2, ...; m = (n + 1) % 2, ..., n + 1 in steps of 2; kind 0 for m = 0,
// else 1 (a, cos) then 2 (b, sin)
__host__ __device__ inline void nmab_of(int j, int& n, int& m, int& kind) {
  int k = 0;
  for (n = 1;; ++n)
    for (m = (n + 1) % 2; m <= n + 1; m += 2) {
      if (m == 0) {
        if (k == j) { kind = 0; return; }
        k += 1;
      } else {
        if (k == j) { kind = 1; return; }
        if (k + 1 == j) { kind = 2; return; }
        k += 2;
      }
    }
}

// aberration polynomials A_j (utils.py:1333-1367) in the order of maskc:
// 2 pi / lambda k^(n+1) / (n+1) times cos(m phi), sin(m phi) or 1, with
// k = lambda q the beam angle and phi its azimuth; FP64, stored as float
template <int N, int LAY>
__global__ void k_aber_table(float* __restrict__ abt, int NA, float px, float py, float lambda) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  int fy, fx;
  spec_of<N, LAY>(e, fy, fx);
  const double kx = double(fmul(lambda, fftfreq_f(fx, N, px))), ky = double(fmul(lambda, fftfreq_f(fy, N, py)));
  const double kr = sqrt(kx * kx + ky * ky), phi = atan2(ky, kx);
  for (int j = 0; j < NA; ++j) {
    int n, m, kind;
    nmab_of(j, n, m, kind);
    double t = 1.0;
    for (int i = 0; i <= n; ++i) t *= kr;
    t /= double(n + 1);
    if (kind == 1) t *= cos(double(m) * phi);
    if (kind == 2) t *= sin(double(m) * phi);
    abt[size_t(j) * N * N + e] = float(t * 6.283185307179586 / double(lambda));
  }
}

// Position-order tables from a caller-supplied propagator (natural order,
// mask included, as multislice() receives master_propagator_phase_space) and
// band mask; H^e through the principal branch when edge > 1.
template <int N>
__global__ void k_tables_natural(const float2* __restrict__ Hn, const float* __restrict__ mn, float edge,
                                 float2* __restrict__ H, float2* __restrict__ H0, float* __restrict__ mask) {
  using T = Tile<N>;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const int f = T::pos2freq(e / N) * N + T::pos2freq(e % N);
  const float s = 1.0f / float(N * N);
  const float2 h = Hn[f];
  H[e] = cscale(h, s);
  if (edge > 1.f) {
    const float mag = sqrtf(cabs2(h));
    const float th = atan2f(h.y, h.x);
    float s0, c0;
    sincosf(edge * th, &s0, &c0);
    const float me = (mag > 0.f) ? powf(mag, edge) : 0.f;
    H0[e] = make_float2(c0 * me * s, s0 * me * s);
  } else {
    H0[e] = H[e];
  }
  mask[e] = mn ? mn[f] : 1.f;
}

// ------------------------------------------------------------- layouts
// reference (Ny, Nx, P) <-> planar (P, Ny, Nx), P = S*K
__global__ void k_ref_to_planar(const float2* __restrict__ src, float2* __restrict__ dst, int ny, int nx,
                                int planes) {
  const size_t total = size_t(ny) * nx * planes;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    const size_t pix = i / planes;
    const int p = int(i % planes);
    dst[size_t(p) * ny * nx + pix] = src[i];
  }
}
__global__ void k_planar_to_ref(const float2* __restrict__ src, float2* __restrict__ dst, int ny, int nx,
                                int planes) {
  const size_t total = size_t(ny) * nx * planes;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    const size_t pix = i / planes;
    const int p = int(i % planes);
    dst[i] = src[size_t(p) * ny * nx + pix];
  }
}
__global__ void k_zero(float* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.f;
}

// ------------------------------------------------------- batch pipeline
// A window whose corner or sequence index is invalid is skipped (reads as
// zero, contributes nothing): the kernels never touch memory outside the
// object, the data or the position arrays whatever the inputs hold.
__device__ __forceinline__ bool window_inside(int y0, int x0, int n, int ny, int nx) {
  return y0 >= 0 && x0 >= 0 && y0 + n <= ny && x0 + n <= nx;
}
__device__ __forceinline__ int seq_index(const int* seq, int slot, int P) {
  const int idx = seq[slot];
  return (idx >= 0 && idx < P) ? idx : -1;
}

struct BatchArgs {
  // geometry
  int B, S, K, M, ny, nx, b0, P;
  int algorithm, step_obj, step_probe, step_pos, shift, far_field, write_stash, scatter;
  float loss_weight;
  // inputs
  const float2* obj;        // planar (S*K, ny, nx)
  const float2* probe;      // (M, N, N) natural
  const float2* fprobe;     // (M, N, N) spectrum, position order, unnormalised
  const int* pos;           // (P, 2)
  const float* corr;        // (P, 2)
  const float* meas;        // (P, N, N) shifted natural
  const int* seq;           // sequence
  const float2* H;          // position order, /n^2
  const float2* H0;
  const float* mask;        // position order
  const float* maskc;       // order of H (C-order for the register tile)
  const int* zcol;          // (N) columns fx where the band mask, H and H0 are zero, or null
  // per-batch buffers
  float2* Ob;               // (B, S, K, N, N) band-limited windows
  float2* Pp;               // (B, M, N, N) shifted probes
  float2* stash;            // (B, M, K, S, N, N)
  float2* Psi;              // (B, M, K, N, N) far field, position order (or exit wave)
  float* coef;              // (B, N, N) position order
  float2* Gbuf;             // (B, S, K, N, N) window gradients
  float2* gP;               // (B, M, K, N, N)
  float2* probe_part;       // (G, M, N, N)
  double* pos_part;         // (B, M, 2)
  double* loss_part;        // (n_seq, 2)
  float* sim;               // (P, N, N) or null
  float2* scratch;          // tiles of windows larger than 128 px (one per workgroup)
  // outputs
  float2* obj_grad;         // planar
  float* pos_grad;          // (P, 2)
  int G;                    // probe-gradient groups
  // measured data: rows by position index or by sequence slot; raw patterns
  // preprocessed into meas_stage (batch slot order) when prep is set;
  // virtual detectors of the compressed losses
  int meas_order, prep, raw_n, sy, sx, bin, pad, J;
  float mult;
  float* meas_stage;        // (B, N, N)
  const float* masks;       // (J, N, N)
  // static background (N, N), zero frequency at (N/2, N/2), or null; its
  // gradient, summed over the pass, or null when not wanted
  const float* bg;
  float* bg_grad;
  // beam current (P) scaling each position's probe by |bc| (objective.py:446-450), or null;
  // with bc_grad the per-window partials of its gradient (B, M) and the gradient (P)
  const float* bc;
  double* bc_part;
  float* bc_grad;
  // aberrations (objective.py:320-326, 459-465, 673-681): NA coefficients per
  // row (n_markers rows) at ab, the row of each position at marker, and the
  // polynomials A_j (NA, N, N) in the order of maskc at abt, or NA = 0.
  // moved: the sub-pixel shift is active (the shifted probe then starts from
  // the spectrum before the plate, so the plates stay out of the forward).
  // With ab_w, the probe gradient also writes |bc| Im(G conj(F)) per window
  // and mode (B, M, N, N, order of maskc), reduced against A_j into ab_part
  // (B, NA) and added by row into ab_grad (n_markers, NA).  plate: each
  // window's exp(-i sum_j A_j c_j) (B, N, N, order of maskc).
  int NA, n_markers, moved;
  const float* ab;
  const int* marker;
  const float* abt;
  float2* plate;
  float* ab_w;
  double* ab_part;
  float* ab_grad;
  // slice range [s_lo, s_hi) of a forward, adjoint or objgrad launch (a whole
  // pass: 0, S) and the stash blocks of st_slots slots per (window, mode).
  // With r = s - s_lo the forward stores slice s in slot r / st_every when
  // r < st_dense and r % st_every == 0, and in slot
  // st_dense / st_every + r - st_dense when r >= st_dense (stash_slot); the
  // adjoint and objgrad find slice s in slot r
  int s_lo, s_hi, st_every, st_dense, st_slots;
  // forward from a stored wave (segment recompute): the wave entering slice
  // s_lo, in stash order, at fin + (b * M * K + mk) * fin_stride; such a
  // launch writes no exit wave
  const float2* fin;
  size_t fin_stride;
  // adjoint over a segment: the wave of slice s_hi - 1 is read from gstate
  // when s_hi < S, and the wave propagated to slice s_lo - 1 is written there
  // when s_lo > 0 (stash order, (B, M, K, N, N))
  float2* gstate;
  // per-slice distances (multislice_core.py:69-71): the propagation after
  // slice s >= 1 uses H + s * Hstride (H0 after slice 0); dists (S - 1) or
  // null for the single distance dz
  size_t Hstride;
  const float* dists;
  float dz;
  // tilts (objective.py:428-499): n_tilt rows of 6 at tilt (1 or P rows).
  // tilt_win: each window multiplies its propagators by its own inside-tilt
  // factors (the tables hold the untilted propagators); tilt_after: the exit
  // wave is multiplied by the after-tilt ramp.  Gradients: tg_in from the
  // spectra Wsp = fft2(u_s) / N^2 (C-order, slot s of S per window and
  // mode) that the forward keeps; tg_after from the exit wave before the
  // ramp, Xe (R-order); both summed per window and mode into tilt_part
  // (B, M*K, 4: inside y, x; after y, x), then into tilt_grad (n_tilt, 6).
  const float* tilt;
  int n_tilt, tilt_win, tilt_after, tg_in, tg_after;
  float lambda, px, py, edge;
  float2* Wsp;
  float2* Xe;
  double* tilt_part;
  float* tilt_grad;
  // tilts before the specimen (objective.py:465-470, 660-672): each window's
  // probe is |bc| P times the before-tilt ramp (in Pp); with a shift, Fw holds
  // its spectrum (B, M, N, N, C-order) for the shift and the position
  // gradient; tg_before: the gradient parts per window and probe mode in
  // tb_part (B, M, 2: y, x)
  int tilt_before, tg_before;
  float2* Fw;
  double* tb_part;
};

// a launch over every slice with one stash slot per slice
inline void whole_pass(BatchArgs& a) {
  a.s_lo = 0; a.s_hi = a.S; a.st_every = 1; a.st_dense = 0; a.st_slots = a.S;
  a.fin = nullptr; a.fin_stride = 0; a.gstate = nullptr;
}

// forward stash slot of slice s, or -1 when it is not kept
__device__ __forceinline__ int stash_slot(const BatchArgs& a, int s) {
  const int r = s - a.s_lo;
  if (r >= a.st_dense) return a.st_dense / a.st_every + (r - a.st_dense);
  return r % a.st_every == 0 ? r / a.st_every : -1;
}

// |beam current| of position idx, the factor of its probe (1 without a beam current)
__device__ __forceinline__ float bc_scale(const BatchArgs& a, int idx) {
  return (a.bc && idx >= 0) ? fabsf(a.bc[idx]) : 1.f;
}

// 2 sign(|bc|) of position idx (objective.py:694-696): the factor of its
// gradient sum Re(dL/dP* conj(P)); 0 for a zero current
__device__ __forceinline__ double bc_factor(const BatchArgs& a, int idx) {
  return (idx >= 0 && a.bc[idx] != 0.f) ? 2.0 : 0.0;
}

// aberration row of position idx, or null: no aberrations, no position, or a
// row outside [0, n_markers) (the host checks the rows before a pass)
__device__ __forceinline__ const float* ab_row(const BatchArgs& a, int idx) {
  if (a.NA == 0 || idx < 0) return nullptr;
  const int r = a.marker[idx];
  return (r >= 0 && r < a.n_markers) ? a.ab + size_t(r) * a.NA : nullptr;
}

// tilt row of position idx, or null (no tilts, or an invalid window)
__device__ __forceinline__ const float* tilt_row(const BatchArgs& a, int idx) {
  if (!a.tilt || idx < 0) return nullptr;
  return a.tilt + size_t(a.n_tilt == 1 ? 0 : idx) * 6;
}

// distance of the propagation after slice s (multislice_core.py:70, 166)
__device__ __forceinline__ float slice_dist(const BatchArgs& a, int s) { return a.dists ? a.dists[s] : a.dz; }

// Inside-tilt factors of one window's propagator after a slice of thickness
// d (objective.py:482-484): exp(-i pi d 2 (qx tx + qy ty)) = tt[fx] tt[N + fy],
// the untilted exp(-i pi d lambda q^2) mask being in the tables.  The caller
// synchronises.
template <int N>
__device__ __forceinline__ void tilt_tables_inside(float2* tt, float d, float tx, float ty, float px, float py) {
  const float k = fmul(-3.141592654f, d);
  for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) {
    const int f = i % N;
    const float q = i < N ? fftfreq_f(f, N, px) : fftfreq_f(f, N, py);
    float sn, cs;
    sincosf(fmul(k, fmul(2.f, fmul(q, i < N ? tx : ty))), &sn, &cs);
    tt[i] = make_float2(cs, sn);
  }
}

// The real-space tilt grids of objective.py:303-307 (i times these):
// tt[j] = (cx[j], cy[j]) with c = 2 pi (j - N/2) / N * N * p / lambda in
// double rounded to float, as the reference forms its complex64 grids.  The
// caller synchronises.
template <int N>
__device__ __forceinline__ void tilt_tables_grid(float2* tt, float px, float py, float lambda) {
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    const double f = double(j - N / 2) / double(N);
    tt[j] = make_float2(float(f * 6.2831855 * double(N) * double(px) / double(lambda)),
                        float(f * 6.2831855 * double(N) * double(py) / double(lambda)));
  }
}

// exp(xg tx + yg ty) at pixel (y, x) from the grid table (objective.py:491)
__device__ __forceinline__ float2 tilt_ramp(const float2* tt, int y, int x, float tx, float ty) {
  float sn, cs;
  sincosf(fadd(fmul(tt[x].x, tx), fmul(tt[y].y, ty)), &sn, &cs);
  return make_float2(cs, sn);
}

// the propagator value h of the window's propagation after slice s with its
// inside-tilt factors (tt, for frequency (fy, fx)); after slice 0 with an
// edge multiplier e > 1 the whole per-window propagator is raised to e
// through the principal branch, as numpy's z**e (multislice_core.py:73-74)
__device__ __forceinline__ float2 tilted_h(float2 h, const float2* tt, int fy, int fx, int N, bool first, float e) {
  h = cmul(h, cmul(tt[fx], tt[N + fy]));
  if (first && e > 1.f) {
    const float m = (h.x != 0.f || h.y != 0.f) ? 1.f : 0.f;
    const float nn = float(N) * float(N);
    const float th = atan2f(h.y * nn, h.x * nn);
    float s0, c0;
    sincosf(e * th, &s0, &c0);
    h = make_float2(c0 * m / nn, s0 * m / nn);
  }
  return h;
}

// measured row of batch slot b (sequence entry b0 + b, position idx)
__device__ __forceinline__ const float* meas_row(const BatchArgs& a, int b, int idx, size_t row_len) {
  if (a.prep) return a.meas_stage + size_t(b) * row_len;
  return a.meas + size_t(a.meas_order == PTY_MEASURED_BY_SEQUENCE ? a.b0 + b : idx) * row_len;
}

template <int N>
__global__ __launch_bounds__(Tile<N>::NT) void k_forward(BatchArgs a) {
  using T = Tile<N>;
  PTY_TILE(T, t, a.scratch);
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int m = mk / a.K, k = mk % a.K;
  const int tid = threadIdx.x;
  const float2* tw = twiddles<N>();
  const float2* src = a.fin ? a.fin + (size_t(b) * MK + mk) * a.fin_stride
                            : a.shift ? a.Pp + (size_t(b) * a.M + m) * N * N : a.probe + size_t(m) * N * N;
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    t[T::lds(e / N, e % N)] = src[e];
  }
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  for (int s = a.s_lo; s < a.s_hi; ++s) {
    __syncthreads();
    const float2* o = a.Ob + ((size_t(b) * a.S + s) * a.K + k) * N * N;
    const int slot = a.write_stash ? stash_slot(a, s) : -1;
    float2* u_out = slot >= 0 ? st + size_t(slot) * N * N : nullptr;
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      const int l = T::lds(e / N, e % N);
      const float2 u = t[l];
      if (u_out) u_out[e] = u;
      t[l] = cmul(u, o[e]);
    }
    __syncthreads();
    if (s < a.s_hi - 1) {
      tile_fft_fwd<N>(t, tw);
      const float2* H = (s == 0) ? a.H0 : a.H;
#pragma unroll 4
      for (int i = 0; i < T::PER_THREAD; ++i) {
        const int e = tid + i * T::NT;
        const int l = T::lds(e / N, e % N);
        t[l] = cmul(t[l], H[e]);
      }
      __syncthreads();
      tile_fft_inv<N>(t, tw);
    }
  }
  if (a.fin) return;  // segment recompute: the stash is the result
  float2* out = a.Psi + (size_t(b) * MK + mk) * N * N;
  if (!a.far_field) {
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      out[e] = t[T::lds(e / N, e % N)];
    }
    return;
  }
  tile_fft_fwd<N>(t, tw);
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    out[e] = cscale(t[T::lds(e / N, e % N)], a.mask[e]);
  }
}

// loss, sse and dL/dI (objective.py:527-587); a window is split into
// LOSS_CHUNKS pixel ranges, one 256-thread workgroup each, with FP64 partial
// sums per (window, chunk) reduced deterministically by k_loss_final
template <int N>
struct LossSplit {
  static constexpr int CHUNKS = (N * N >= 4096) ? N * N / 1024 : 1;
  static constexpr int PIX = N * N / CHUNKS;
};

template <int N, int LAY>
__global__ __launch_bounds__(256) void k_loss(BatchArgs a) {
  using T = Tile<N>;
  constexpr int CH = LossSplit<N>::CHUNKS, PIX = LossSplit<N>::PIX;
  __shared__ double red[8];
  const int b = blockIdx.x / CH, ch = blockIdx.x % CH;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const int MK = a.M * a.K;
  const size_t slot = (size_t(a.b0 + b) * CH + ch) * 2;
  const int e0 = ch * PIX;
  if (idx < 0) {  // invalid sequence entry: no loss, zero gradient
    for (int e = e0 + threadIdx.x; e < e0 + PIX; e += blockDim.x) a.coef[size_t(b) * N * N + e] = 0.f;
    if (threadIdx.x == 0) {
      a.loss_part[slot] = 0.0;
      a.loss_part[slot + 1] = 0.0;
    }
    return;
  }
  const float* D = meas_row(a, b, idx, size_t(N) * N);
  const float2* psi = a.Psi + size_t(b) * MK * N * N;
  const float scale = a.loss_weight * float(N) * float(N);
  double lsum = 0.0, ssum = 0.0;
  for (int e = e0 + threadIdx.x; e < e0 + PIX; e += blockDim.x) {
    float I = 0.f;
    for (int j = 0; j < MK; ++j) I += cabs2(psi[size_t(j) * N * N + e]);
    if (MK > 1) I = I / float(MK);
    int fy, fx;
    spec_of<N, LAY>(e, fy, fx);
    const int ys = (fy + N / 2) % N, xs = (fx + N / 2) % N;
    const float d = D[ys * N + xs];
    const float I_wave = I;  // the simulated pattern leaves the background out (objective.py:517-526)
    if (a.bg) {
      const float bgv = a.bg[ys * N + xs];
      I += bgv * bgv * a.maskc[e];
    }
    float l;
    if (a.algorithm == PTY_LOSS_ML) {
      const float diff = I - d;
      ssum += double(diff) * double(diff);
      lsum += double(I) - double(d) * double(logf(I + (I == 0.f ? 1.f : 0.f)));
      l = diff / (I == 0.f ? 1.f : I);
    } else if (a.algorithm == PTY_LOSS_LSQ) {
      const float diff = I - d;
      const double t2 = double(diff) * double(diff);
      ssum += t2;
      lsum += t2;
      l = 2.f * diff;
    } else if (a.algorithm == PTY_LOSS_LSQ_SQRT_2) {
      const float r = d - I;
      ssum += double(r) * double(r);
      if (d > 0.5f) {
        lsum += 0.5 * double(r) * double(r) / double(d);
        l = I / d - 1.f;
      } else {
        lsum += double(I);
        l = 1.f;
      }
    } else {
      const float r = d - I;
      ssum += double(r) * double(r);
      float si = sqrtf(I);
      const float dd = si - sqrtf(d);
      lsum += double(dd) * double(dd);
      if (si == 0.f) si = 1.f;
      l = dd / si;
    }
    if (MK > 1) l = l / float(MK);
    a.coef[size_t(b) * N * N + e] = l * scale;
    if (a.sim) a.sim[size_t(idx) * N * N + ys * N + xs] = I_wave;
  }
  const double L = block_sum(lsum, red);
  const double Ssum = block_sum(ssum, red);
  if (threadIdx.x == 0) {
    a.loss_part[slot] = L;
    a.loss_part[slot + 1] = Ssum;
  }
}

// Gradient of the static background (objective.py:588-589): per pixel, the
// batch's dL/dI summed in batch order, times 2 * b * loss_weight, added to the
// pass sum.  coef holds dL/dI * loss_weight * N^2 (zero for invalid entries).
template <int N, int LAY>
__global__ __launch_bounds__(256) void k_bg_grad(BatchArgs a) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= N * N) return;
  float sum = 0.f;
  for (int b = 0; b < a.B; ++b) sum += a.coef[size_t(b) * N * N + e];
  int fy, fx;
  spec_of<N, LAY>(e, fy, fx);
  const int p = ((fy + N / 2) % N) * N + (fx + N / 2) % N;
  a.bg_grad[p] += sum * (2.f / float(N * N)) * a.bg[p];
}

// Per-batch preprocessing of raw patterns (utils.py:587-653 as
// objective.py:400-405 calls it): roll by (sy, sx) with the edge band that the
// reference zeroes after each roll (rows [H-sy, H) for sy > 0, [0, -sy) for
// sy < 0; columns likewise), sum of bin x bin blocks, times mult, zero pad of
// `pad` on each side.  One thread per output pixel of the batch.
template <int N>
__global__ __launch_bounds__(256) void k_preprocess(BatchArgs a) {
  const int b = blockIdx.x / ((N * N + 255) / 256);
  const int e = (blockIdx.x % ((N * N + 255) / 256)) * 256 + threadIdx.x;
  if (e >= N * N) return;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  float v = 0.f;
  const int H = a.raw_n, nb = H / a.bin;
  const int yb = e / N - a.pad, xb = e % N - a.pad;
  if (idx >= 0 && yb >= 0 && yb < nb && xb >= 0 && xb < nb) {
    const float* raw =
        a.meas + size_t(a.meas_order == PTY_MEASURED_BY_SEQUENCE ? a.b0 + b : idx) * size_t(H) * H;
    for (int i1 = 0; i1 < a.bin; ++i1) {
      const int y = yb * a.bin + i1;
      if ((a.sy > 0 && y >= H - a.sy) || (a.sy < 0 && y < -a.sy)) continue;
      const int ys = ((y - a.sy) % H + H) % H;
      for (int j1 = 0; j1 < a.bin; ++j1) {
        const int x = xb * a.bin + j1;
        if ((a.sx > 0 && x >= H - a.sx) || (a.sx < 0 && x < -a.sx)) continue;
        v += raw[size_t(ys) * H + ((x - a.sx) % H + H) % H];
      }
    }
    v *= a.mult;
  }
  a.meas_stage[size_t(b) * N * N + e] = v;
}

// Compressed losses (objective.py:560-585): the model's virtual detector
// coefficients c_j = sum_pix I * mask_j (FP64 sums), the loss and sse of each
// against the measured coefficients, and dL/dI = sum_j mask_j dL/dc_j.  One
// workgroup per window; coef holds I between the passes.
constexpr int kMaxMasks = 256;

template <int N, int LAY>
__global__ __launch_bounds__(256) void k_loss_c(BatchArgs a) {
  using T = Tile<N>;
  constexpr int CH = LossSplit<N>::CHUNKS;
  __shared__ double red[8];
  __shared__ float dcs[kMaxMasks];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const int MK = a.M * a.K;
  const size_t slot = size_t(a.b0 + b) * CH * 2;
  float* coef = a.coef + size_t(b) * N * N;
  if (tid < 2 * CH) a.loss_part[slot + tid] = 0.0;
  if (idx < 0) {
    for (int e = tid; e < N * N; e += blockDim.x) coef[e] = 0.f;
    return;
  }
  auto pix = [](int e) {
    int fy, fx;
    spec_of<N, LAY>(e, fy, fx);
    return ((fy + N / 2) % N) * N + (fx + N / 2) % N;
  };
  const float2* psi = a.Psi + size_t(b) * MK * N * N;
  for (int e = tid; e < N * N; e += blockDim.x) {
    float I = 0.f;
    for (int j = 0; j < MK; ++j) I += cabs2(psi[size_t(j) * N * N + e]);
    if (MK > 1) I = I / float(MK);
    coef[e] = I;
    if (a.sim) a.sim[size_t(idx) * N * N + pix(e)] = I;
  }
  __syncthreads();
  const float* D = meas_row(a, b, idx, size_t(a.J));
  double lsum = 0.0, ssum = 0.0;  // meaningful in thread 0
  constexpr int JC = 8;
  for (int j0 = 0; j0 < a.J; j0 += JC) {
    double acc[JC];
#pragma unroll
    for (int q = 0; q < JC; ++q) acc[q] = 0.0;
    for (int e = tid; e < N * N; e += blockDim.x) {
      const double I = coef[e];
      const int p = pix(e);
#pragma unroll
      for (int q = 0; q < JC; ++q)
        if (j0 + q < a.J) acc[q] += I * double(a.masks[size_t(j0 + q) * N * N + p]);
    }
#pragma unroll
    for (int q = 0; q < JC; ++q) {
      if (j0 + q >= a.J) break;
      const double tot = block_sum(acc[q], red);
      if (tid == 0) {
        float c = float(tot);
        const float d = D[j0 + q];
        float dc;
        if (a.algorithm == PTY_LOSS_LSQ_COMPRESSED) {
          const float diff = c - d;
          const double t2 = double(diff) * double(diff);
          lsum += t2;
          ssum += t2;
          dc = 2.f * diff;
        } else if (a.algorithm == PTY_LOSS_GAUSS_COMPRESSED) {
          float cs = sqrtf(c);
          const float r = c - d, dd = cs - sqrtf(d);
          ssum += double(r) * double(r);
          lsum += double(dd) * double(dd);
          if (cs == 0.f) cs = 1.f;
          dc = dd / cs;
        } else {
          const float r = c - d;
          ssum += double(r) * double(r);
          if (c == 0.f) c = 1.f;
          lsum += double(c) - double(d) * double(logf(c));
          dc = r / c;
        }
        dcs[j0 + q] = dc;
      }
    }
  }
  __syncthreads();
  const float scale = a.loss_weight * float(N) * float(N);
  for (int e = tid; e < N * N; e += blockDim.x) {
    const int p = pix(e);
    float l = 0.f;
    for (int j = 0; j < a.J; ++j) l += a.masks[size_t(j) * N * N + p] * dcs[j];
    if (MK > 1) l = l / float(MK);
    coef[e] = l * scale;
  }
  if (tid == 0) {
    a.loss_part[slot] = lsum;
    a.loss_part[slot + 1] = ssum;
  }
}

template <int N>
__global__ __launch_bounds__(Tile<N>::NT) void k_adjoint(BatchArgs a) {
  using T = Tile<N>;
  PTY_TILE(T, t, a.scratch);
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int k = mk % a.K;
  const int tid = threadIdx.x;
  const float2* tw = twiddles<N>();
  const float2* psi = a.Psi + (size_t(b) * MK + mk) * N * N;
  if (a.s_hi < a.S) {
    // a later segment left the wave of slice s_hi - 1
    const float2* gs = a.gstate + (size_t(b) * MK + mk) * N * N;
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      t[T::lds(e / N, e % N)] = gs[e];
    }
  } else if (a.far_field) {
    const float* c = a.coef + size_t(b) * N * N;
    const float s = 1.0f / float(N * N);
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      t[T::lds(e / N, e % N)] = cscale(psi[e], c[e] * s);
    }
    __syncthreads();
    tile_fft_inv<N>(t, tw);
  } else {
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      t[T::lds(e / N, e % N)] = psi[e];
    }
  }
  // the stash slot of slice s is overwritten by this mode's window gradient
  // conj(u_s) * g (the reference's waves_multislice *= dLoss); k_objgrad sums
  // the modes, so no two workgroups write the same element
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  for (int s = a.s_hi - 1; s >= a.s_lo; --s) {
    __syncthreads();
    const size_t win = ((size_t(b) * a.S + s) * a.K + k) * N * N;
    const float2* o = a.Ob + win;
    float2* u = st + size_t(s - a.s_lo) * N * N;
#pragma unroll 4
    for (int i = 0; i < T::PER_THREAD; ++i) {
      const int e = tid + i * T::NT;
      const int l = T::lds(e / N, e % N);
      const float2 g = t[l];
      if (a.step_obj) u[e] = cmulc(g, u[e]);  // conj(u) * g
      t[l] = cmulc(g, o[e]);
    }
    __syncthreads();
    if (s > 0) {
      tile_fft_fwd<N>(t, tw);
      const float2* H = (s == 1) ? a.H0 : a.H;
#pragma unroll 4
      for (int i = 0; i < T::PER_THREAD; ++i) {
        const int e = tid + i * T::NT;
        const int l = T::lds(e / N, e % N);
        t[l] = cmulc(t[l], H[e]);
      }
      __syncthreads();
      tile_fft_inv<N>(t, tw);
    }
  }
  // the probe-side gradient, or the wave of slice s_lo - 1 for the next segment
  float2* out = (a.s_lo > 0 ? a.gstate : a.gP) + (size_t(b) * MK + mk) * N * N;
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    out[e] = t[T::lds(e / N, e % N)];
  }
}

// workgroup -> (window b, plane = s * K + k) of an objgrad launch over the
// slices [s_lo, s_hi)
__device__ __forceinline__ void objgrad_plane(const BatchArgs& a, int& b, int& plane, int& sl, int& k) {
  const int splanes = (a.s_hi - a.s_lo) * a.K;
  b = blockIdx.x / splanes;
  const int q = blockIdx.x % splanes;
  sl = a.s_lo + q / a.K;
  k = q % a.K;
  plane = sl * a.K + k;
}

// band-limit each window gradient and scatter-add it (multislice_core.py:184-192)
template <int N>
__global__ __launch_bounds__(Tile<N>::NT) void k_objgrad(BatchArgs a) {
  using T = Tile<N>;
  PTY_TILE(T, t, a.scratch);
  const int planes = a.S * a.K;
  int b, plane, sl, k;
  objgrad_plane(a, b, plane, sl, k);
  const int tid = threadIdx.x;
  // sum over probe modes of the per-mode window gradients left in the stash
  const size_t mstride = size_t(a.K) * a.st_slots * N * N;
  const float2* src = a.stash + ((size_t(b) * a.M * a.K + k) * a.st_slots + (sl - a.s_lo)) * N * N;
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    float2 v = src[e];
    for (int m = 1; m < a.M; ++m) v = cadd(v, src[m * mstride + e]);
    t[T::lds(e / N, e % N)] = v;
  }
  __syncthreads();
  const float2* tw = twiddles<N>();
  tile_fft_fwd<N>(t, tw);
  const float s = 1.0f / float(N * N);
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    const int l = T::lds(e / N, e % N);
    t[l] = cscale(t[l], a.mask[e] * s);
  }
  __syncthreads();
  tile_fft_inv<N>(t, tw);
  // band-limited window gradient: to Gbuf (pty_ms_adjoint) or, in a full pass,
  // into the window buffer Ob (free after the adjoint), from where
  // k_gather_grad accumulates the object gradient
  float2* G = (a.scatter ? a.Ob : a.Gbuf) + (size_t(b) * planes + plane) * N * N;
#pragma unroll 4
  for (int i = 0; i < T::PER_THREAD; ++i) {
    const int e = tid + i * T::NT;
    G[e] = t[T::lds(e / N, e % N)];
  }
}

// --------------------------------------------- register-tile pipeline
// The same per-batch steps for N = 16 to 128 over the register tile of
// pty_rt.hpp.  Buffer orders: Ob, stash and gP R-order; Psi, coef, H, H0,
// maskc C-order; Pp and the band-limited window gradient natural order.
//
// Register i of an R/C-order array is element i*NT + tid of one window's
// N*N block.  It is accessed through a buffer resource on the block: the
// lane offset tid*4*W (one vector register for all i) plus the scalar offset
// i*NT*4*W, so no per-register address arithmetic runs on the vector ALU
// (a global load's immediate offset cannot hold i*NT*8 bytes).  The resource
// covers exactly the block, so an access can never leave it.
template <int N, typename E, int NT = RT<N>::NT>
struct RView {
  static_assert(sizeof(E) == 8 || sizeof(E) == 4, "float2 or float");
  __amdgpu_buffer_rsrc_t r;
  unsigned voff;
  __device__ __forceinline__ RView(const E* block, unsigned tid)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<E*>(block), 0, N * N * int(sizeof(E)), 0x00020000)),
        voff(tid * unsigned(sizeof(E))) {}
  __device__ __forceinline__ E operator[](int i) const {
    const int so = i * NT * int(sizeof(E));
    if constexpr (sizeof(E) == 8)
      return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, voff, so, 0));
    else
      return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, voff, so, 0));
  }
  __device__ __forceinline__ void store(int i, E v) const {
    const int so = i * NT * int(sizeof(E));
    if constexpr (sizeof(E) == 8) {
      using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, v), r, voff, so, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, so, 0);
    }
  }
};

// Zero columns.  k_tables marks in zcol[fx] the spectrum columns fx in which
// the band mask, and with it H and H0, is zero for every fy; the default
// mask, a disc of radius 2/3 of the largest frequency, zeroes about a third
// of them.  A spectrum multiplied by these tables is exactly zero there, and
// so is its inverse column transform: a wavefront whose columns are all
// marked skips their transforms.  col is the thread's column position in the
// column phase (frequency col_freq(col)).
template <int N>
__device__ __forceinline__ bool wave_cols_zero(const int* zcol, int col) {
  return zcol != nullptr && __ballot(zcol[RT<N>::col_freq(col)] == 0) == 0ull;
}

// Band limit of a window held in the row phase, in place: forward 2D
// transform, times maskc / N^2, inverse 2D transform.  Wavefronts whose
// columns the mask zeroes skip their column transforms.
template <int N>
__device__ __forceinline__ void rt_band_limit(float2* v, float2* T, const RTIds<N>& id, const float* maskc,
                                              const int* zcol) {
  using R = RT<N>;
  const unsigned tid = threadIdx.x;
  const bool zc = wave_cols_zero<N>(zcol, id.col);
  rt_line_fwd<N, true>(v, T, id.row, id.rj);
  rt_rows_to_cols<N>(v, T, id);
  if (zc) {
#pragma unroll
    for (int i = 0; i < R::R1; ++i) v[i] = make_float2(0.f, 0.f);
  } else {
    rt_line_fwd<N, false>(v, T, id.col, id.cj);
    const float sc = 1.0f / float(N * N);
    const RView<N, float> mask(maskc, tid);
#pragma unroll
    for (int i = 0; i < R::R1; ++i) v[i] = cscale(v[i], mask[i] * sc);
    rt_line_inv<N, false>(v, T, id.col, id.cj);
  }
  rt_cols_to_rows<N>(v, T, id);
  rt_line_inv<N, true>(v, T, id.row, id.rj);
}

// k_patch over the register tile (objective.py:436-441)
template <int N>
__global__ __launch_bounds__(RT<N>::NT) void k_patch_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  const int planes = a.S * a.K;
  const int b = blockIdx.x / planes, plane = blockIdx.x % planes;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const int y0 = idx >= 0 ? a.pos[2 * idx] : -1, x0 = idx >= 0 ? a.pos[2 * idx + 1] : -1;
  const bool inside = window_inside(y0, x0, N, a.ny, a.nx);
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  float2 v[R::R1];
  const float2* src = a.obj + size_t(plane) * a.ny * a.nx + (inside ? size_t(y0 + id.row) * a.nx + x0 + id.rj : 0);
#pragma unroll
  for (int n = 0; n < R::R1; ++n) v[n] = inside ? src[R::T * n] : make_float2(0.f, 0.f);
  rt_band_limit<N>(v, T, id, a.maskc, a.zcol);
  const RView<N, float2> dst(a.Ob + (size_t(b) * planes + plane) * N * N, tid);
#pragma unroll
  for (int n = 0; n < R::R1; ++n) dst.store(n, v[n]);
}

// k_forward over the register tile (multislice_core.py:9-82, objective.py:501-512)
template <int N, bool TILT>
__global__ __launch_bounds__(RT<N>::NT) void k_forward_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  __shared__ float2 tt[TILT ? 2 * N : 1];  // tilt factors or grids
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int m = mk / a.K, k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  const float* trow = TILT ? tilt_row(a, seq_index(a.seq, a.b0 + b, a.P)) : nullptr;
  const float2* src = a.shift ? a.Pp + (size_t(b) * a.M + m) * N * N : a.probe + size_t(m) * N * N;
  // o holds the object window of the current slice; the next slice's is read
  // during the column phase so that its latency hides behind the transforms,
  // and the propagator is read before the column transform that precedes its use
  float2 v[R::R1], o[R::R1];
  if (a.fin) {
    const RView<N, float2> I(a.fin + (size_t(b) * MK + mk) * a.fin_stride, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = I[n];
  } else {
    // the shifted probes carry the beam current already
    const float bcs = a.shift ? 1.f : bc_scale(a, seq_index(a.seq, a.b0 + b, a.P));
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = cscale(src[id.row * N + id.rj + R::T * n], bcs);
  }
  const size_t ostride = size_t(a.K) * N * N;
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  {
    const RView<N, float2> O(ob + size_t(a.s_lo) * ostride, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) o[n] = O[n];
  }
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  float2* wsp = (TILT && a.tg_in) ? a.Wsp + (size_t(b) * MK + mk) * a.S * N * N : nullptr;
  const float tix = trow ? trow[3] : 0.f, tiy = trow ? trow[2] : 0.f;
  if ((TILT && a.tilt_win) && !a.dists) {
    tilt_tables_inside<N>(tt, a.dz, tix, tiy, a.px, a.py);
    __syncthreads();
  }
  // wavefronts whose columns the propagators (the band mask) zero skip the
  // column transforms of the slices (of the far field)
  const bool zh = wave_cols_zero<N>(a.zcol, id.col), zm = zh;
  for (int s = a.s_lo; s < a.s_hi; ++s) {
    const int slot = a.write_stash ? stash_slot(a, s) : -1;
    if (slot >= 0) {
      const RView<N, float2> U(st + size_t(slot) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) U.store(n, v[n]);
    }
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = cmul(v[n], o[n]);
    if (s < a.s_hi - 1) {
      if ((TILT && a.tilt_win) && a.dists) {
        // per-slice distances: this propagation's own tilt factors
        __syncthreads();
        tilt_tables_inside<N>(tt, a.dists[s], tix, tiy, a.px, a.py);
        __syncthreads();
      }
      rt_line_fwd<N, true>(v, T, id.row, id.rj);
      rt_rows_to_cols<N>(v, T, id);
      const RView<N, float2> O(ob + size_t(s + 1) * ostride, tid);
      if (zh) {
#pragma unroll
        for (int n = 0; n < R::R1; ++n) o[n] = O[n];
#pragma unroll
        for (int i = 0; i < R::R1; ++i) v[i] = make_float2(0.f, 0.f);
      } else {
        // with an edge multiplier, the first propagation is H0 = H^e; a window
        // with its own tilt raises its tilted propagator itself
        const float2* hs = (s == 0 && !(TILT && a.tilt_win)) ? a.H0 : a.H + size_t(s) * a.Hstride;
        const RView<N, float2> H(hs, tid);
        float2 h[R::R1];
#pragma unroll
        for (int i = 0; i < R::R1; ++i) h[i] = H[i];
        rt_line_fwd<N, false>(v, T, id.col, id.cj);
        if ((TILT && a.tilt_win)) {
#pragma unroll
          for (int i = 0; i < R::R1; ++i) {
            int fy, fx;
            R::cspec(i * R::NT + tid, fy, fx);
            h[i] = tilted_h(h[i], tt, fy, fx, N, s == 0, a.edge);
          }
        }
#pragma unroll
        for (int i = 0; i < R::R1; ++i) v[i] = cmul(v[i], h[i]);
        if (wsp) {
          // fft2 of the wave entering slice s + 1 (over N^2), for the inside-tilt gradient
          const RView<N, float2> W(wsp + size_t(s + 1) * N * N, tid);
#pragma unroll
          for (int i = 0; i < R::R1; ++i) W.store(i, v[i]);
        }
#pragma unroll
        for (int n = 0; n < R::R1; ++n) o[n] = O[n];
        rt_line_inv<N, false>(v, T, id.col, id.cj);
      }
      rt_cols_to_rows<N>(v, T, id);
      rt_line_inv<N, true>(v, T, id.row, id.rj);
    }
  }
  if (a.fin) return;  // segment recompute: the stash is the result
  if ((TILT && a.tilt_after)) {
    // the exit wave, band-limited first for one slice, times the after-tilt
    // ramp (objective.py:490-499); Xe keeps it before the ramp
    if (a.S == 1) rt_band_limit<N>(v, T, id, a.maskc, a.zcol);
    if ((TILT && a.tg_after)) {
      const RView<N, float2> X(a.Xe + (size_t(b) * MK + mk) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) X.store(n, v[n]);
    }
    __syncthreads();
    tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
    __syncthreads();
    const float tax = trow ? trow[5] : 0.f, tay = trow ? trow[4] : 0.f;
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = cmul(v[n], tilt_ramp(tt, id.row, id.rj + R::T * n, tax, tay));
  }
  rt_line_fwd<N, true>(v, T, id.row, id.rj);
  rt_rows_to_cols<N>(v, T, id);
  const RView<N, float2> out(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
  if (zm) {
#pragma unroll
    for (int i = 0; i < R::R1; ++i) out.store(i, make_float2(0.f, 0.f));
  } else {
    rt_line_fwd<N, false>(v, T, id.col, id.cj);
    const RView<N, float> mask(a.maskc, tid);
#pragma unroll
    for (int i = 0; i < R::R1; ++i) out.store(i, cscale(v[i], mask[i]));
  }
}

// k_adjoint over the register tile (objective.py:595-597, multislice_core.py:85-200)
template <int N, bool TILT>
__global__ __launch_bounds__(RT<N>::NT) void k_adjoint_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  __shared__ float2 tt[TILT ? 2 * N : 1];  // tilt factors or grids
  __shared__ double red[TILT ? 16 : 1];
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  const float sc = 1.0f / float(N * N);
  const float* trow = TILT ? tilt_row(a, seq_index(a.seq, a.b0 + b, a.P)) : nullptr;
  double tg[4] = {0.0, 0.0, 0.0, 0.0};  // tilt gradient: inside y, x; after y, x
  float2 v[R::R1];
  // as in the forward: columns zero in the propagators, and in the band mask
  // (where the far field, and so its gradient, is zero)
  const bool zh = wave_cols_zero<N>(a.zcol, id.col);
  if (a.s_hi < a.S) {
    // a later segment left the wave of slice s_hi - 1
    const RView<N, float2> gs(a.gstate + (size_t(b) * MK + mk) * N * N, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = gs[n];
  } else {
    if (zh) {
#pragma unroll
      for (int i = 0; i < R::R1; ++i) v[i] = make_float2(0.f, 0.f);
    } else {
      const RView<N, float2> psi(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
      const RView<N, float> c(a.coef + size_t(b) * N * N, tid);
#pragma unroll
      for (int i = 0; i < R::R1; ++i) v[i] = cscale(psi[i], c[i] * sc);
      rt_line_inv<N, false>(v, T, id.col, id.cj);
    }
    rt_cols_to_rows<N>(v, T, id);
    rt_line_inv<N, true>(v, T, id.row, id.rj);
    if ((TILT && a.tilt_after)) {
      // the gradient through the after-tilt ramp and of its tilt (objective.py:620-636)
      tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
      __syncthreads();
      const float tax = trow ? trow[5] : 0.f, tay = trow ? trow[4] : 0.f;
      const RView<N, float2> X(a.Xe + (size_t(b) * MK + mk) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) {
        const int x = id.rj + R::T * n;
        v[n] = cmulc(v[n], tilt_ramp(tt, id.row, x, tax, tay));
        if ((TILT && a.tg_after)) {
          const float2 e = X[n];
          const double zi = double(v[n].y) * double(e.x) - double(v[n].x) * double(e.y);  // Im(g conj(e))
          tg[2] += 2.0 * double(tt[id.row].y) * zi;
          tg[3] += 2.0 * double(tt[x].x) * zi;
        }
      }
      if (a.S == 1) rt_band_limit<N>(v, T, id, a.maskc, a.zcol);
      __syncthreads();
    }
  }
  // the stash slot of slice s is overwritten by this mode's window gradient
  // conj(u_s) * g (k_objgrad_rt sums the modes); pu holds slot s, read one
  // column phase ahead like the forward's object window
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  const float2* wsp = (TILT && a.tg_in) ? a.Wsp + (size_t(b) * MK + mk) * a.S * N * N : nullptr;
  const size_t ostride = size_t(a.K) * N * N;
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  const float tix = trow ? trow[3] : 0.f, tiy = trow ? trow[2] : 0.f;
  if ((TILT && a.tilt_win) && !a.dists) {
    tilt_tables_inside<N>(tt, a.dz, tix, tiy, a.px, a.py);
    __syncthreads();
  }
  float2 pu[R::R1];
  if (a.step_obj) {
    const RView<N, float2> U(st + size_t(a.s_hi - 1 - a.s_lo) * N * N, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) pu[n] = U[n];
  }
  for (int s = a.s_hi - 1; s >= a.s_lo; --s) {
    const RView<N, float2> O(ob + size_t(s) * ostride, tid);
    if (a.step_obj) {
      const RView<N, float2> U(st + size_t(s - a.s_lo) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) U.store(n, cmulc(v[n], pu[n]));
    }
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = cmulc(v[n], O[n]);
    if (s > 0) {
      if ((TILT && a.tilt_win) && a.dists) {
        __syncthreads();
        tilt_tables_inside<N>(tt, a.dists[s - 1], tix, tiy, a.px, a.py);
        __syncthreads();
      }
      rt_line_fwd<N, true>(v, T, id.row, id.rj);
      rt_rows_to_cols<N>(v, T, id);
      if (zh) {
#pragma unroll
        for (int i = 0; i < R::R1; ++i) v[i] = make_float2(0.f, 0.f);
      } else {
        const float2* hs = (s == 1 && !(TILT && a.tilt_win)) ? a.H0 : a.H + size_t(s - 1) * a.Hstride;
        const RView<N, float2> H(hs, tid);
        float2 h[R::R1];
        // the propagator is read ahead of the transform, except when the
        // tilt-gradient sum needs the registers first
        if (!wsp) {
#pragma unroll
          for (int i = 0; i < R::R1; ++i) h[i] = H[i];
        }
        rt_line_fwd<N, false>(v, T, id.col, id.cj);
        if (wsp) {
          // -4 pi d / N^2 sum Im(conj(fft2 u_s) fft2(g_s) q), with Wsp = fft2(u_s) / N^2
          // (multislice_core.py:169-178)
          const RView<N, float2> W(wsp + size_t(s) * N * N, tid);
          double sy = 0.0, sx = 0.0;
#pragma unroll
          for (int i = 0; i < R::R1; ++i) {
            int fy, fx;
            R::cspec(i * R::NT + tid, fy, fx);
            const float2 w = W[i];
            const double zi = double(w.x) * double(v[i].y) - double(w.y) * double(v[i].x);
            sx += double(fftfreq_f(fx, N, a.px)) * zi;
            sy += double(fftfreq_f(fy, N, a.py)) * zi;
          }
          double d = double(slice_dist(a, s - 1));
          if (s == 1 && a.edge > 1.f) d *= double(a.edge);
          tg[0] += -12.566370614359172 * d * sy;
          tg[1] += -12.566370614359172 * d * sx;
#pragma unroll
          for (int i = 0; i < R::R1; ++i) h[i] = H[i];
        }
        if ((TILT && a.tilt_win)) {
#pragma unroll
          for (int i = 0; i < R::R1; ++i) {
            int fy, fx;
            R::cspec(i * R::NT + tid, fy, fx);
            h[i] = tilted_h(h[i], tt, fy, fx, N, s == 1, a.edge);
          }
        }
#pragma unroll
        for (int i = 0; i < R::R1; ++i) v[i] = cmulc(v[i], h[i]);
      }
      if (a.step_obj && s > a.s_lo) {
        const RView<N, float2> U(st + size_t(s - 1 - a.s_lo) * N * N, tid);
#pragma unroll
        for (int n = 0; n < R::R1; ++n) pu[n] = U[n];
      }
      if (!zh) rt_line_inv<N, false>(v, T, id.col, id.cj);
      rt_cols_to_rows<N>(v, T, id);
      rt_line_inv<N, true>(v, T, id.row, id.rj);
    }
  }
  // probe-side gradient in R-order (read only by k_probegrad_rt), or the wave
  // of slice s_lo - 1 for the next segment
  const RView<N, float2> out((a.s_lo > 0 ? a.gstate : a.gP) + (size_t(b) * MK + mk) * N * N, tid);
#pragma unroll
  for (int n = 0; n < R::R1; ++n) out.store(n, v[n]);
  if ((TILT && a.tg_in) || (TILT && a.tg_after)) {
    // this segment's part; the segments of a window add in order
    for (int c = 0; c < 4; ++c) {
      const double S = block_sum(tg[c], red);
      if (tid == 0) {
        double* dst = a.tilt_part + (size_t(b) * MK + mk) * 4 + c;
        *dst = (a.s_hi < a.S ? *dst : 0.0) + S;
      }
    }
  }
}

// k_objgrad over the register tile: mode sum from the R-order stash, band
// limit, natural-order window gradient into Ob for k_gather_grad
template <int N>
__global__ __launch_bounds__(RT<N>::NT) void k_objgrad_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  const int planes = a.S * a.K;
  int b, plane, sl, k;
  objgrad_plane(a, b, plane, sl, k);
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  const size_t mstride = size_t(a.K) * a.st_slots * N * N;
  const float2* src = a.stash + ((size_t(b) * a.M * a.K + k) * a.st_slots + (sl - a.s_lo)) * N * N;
  float2 v[R::R1];
  {
    const RView<N, float2> G0(src, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = G0[n];
  }
  for (int m = 1; m < a.M; ++m) {
    const RView<N, float2> Gm(src + m * mstride, tid);
#pragma unroll
    for (int n = 0; n < R::R1; ++n) v[n] = cadd(v[n], Gm[n]);
  }
  rt_band_limit<N>(v, T, id, a.maskc, a.zcol);
  float2* G = (a.scatter ? a.Ob : a.Gbuf) + (size_t(b) * planes + plane) * N * N;
#pragma unroll
  for (int n = 0; n < R::R1; ++n) G[id.row * N + id.rj + R::T * n] = v[n];
}

// --------------------------------- register-tile probe side (16 to 128 px)
// Probe spectra (fprobe) and the spectral probe-gradient partials are in
// C-order here, the probe-side gradient from the adjoint in R-order; the
// shift kernel E = ex[fx] * ey[fy] * mask comes from two axis tables in LDS
// (objective.py:309-313, 473).

// Shift kernel E = exp(Kx*dx + Ky*dy) * mask with K = -2*pi*i*fftfreq * mask
// (objective.py:309-313, 473).  Where the mask is 1 the exponent is a sum of
// an x and a y term, so E = ex[px] * ey[py] * mask with the two axis factors
// exp(-2*pi*i*fftfreq*d) tabulated once per window in LDS (2N sincos instead
// of N^2); where the mask is 0, E is 0 either way.
// exy[f] = exp(i * -2*pi*fftfreq(f) * dx) for f < N, exy[N + f] the same for dy
// (indexed by frequency; the caller synchronises)
template <int N>
__device__ __forceinline__ void shift_tables_freq(float2* exy, float dy, float dx) {
  for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) {
    const int f = i % N;
    const float k = fmul(-6.2831855f, fftfreq_f(f, N, 1.f));
    float sn, cs;
    sincosf(fmul(k, i < N ? dx : dy), &sn, &cs);
    exy[i] = make_float2(cs, sn);
  }
}

// probe spectra, C-order, unnormalised
template <int N>
__global__ __launch_bounds__(RT<N>::NT) void k_fprobe_rt(const float2* __restrict__ probe, float2* __restrict__ fprobe) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  const RTIds<N> id;
  const float2* src = probe + size_t(blockIdx.x) * N * N;
  float2 v[R::R1];
#pragma unroll
  for (int n = 0; n < R::R1; ++n) v[n] = src[id.row * N + id.rj + R::T * n];
  rt_fft2_fwd<N>(v, T, id);
  const RView<N, float2> out(fprobe + size_t(blockIdx.x) * N * N, threadIdx.x);
#pragma unroll
  for (int i = 0; i < R::R1; ++i) out.store(i, v[i]);
}

// shifted probes (objective.py:471-476): ifft2(fprobe * E), natural order;
// without a shift, the aberrated probes ifft2(fprobe * plate) (:459-465)
template <int N, bool AB>
__global__ __launch_bounds__(RT<N>::NT) void k_probe_shift_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  __shared__ float2 exy[2 * N];
  const int b = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const bool mv = a.moved && idx >= 0;
  const float dy = mv ? a.corr[2 * idx] : 0.f, dx = mv ? a.corr[2 * idx + 1] : 0.f;
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  shift_tables_freq<N>(exy, dy, dx);
  __syncthreads();
  // with a before-tilt each window brings its own spectrum, the current in it
  const RView<N, float2> fp(a.Fw ? a.Fw + (size_t(b) * a.M + m) * N * N : a.fprobe + size_t(m) * N * N, tid);
  const RView<N, float> mask(a.maskc, tid);
  const float s = (a.Fw ? 1.f : bc_scale(a, idx)) / float(N * N);
  float2 v[R::R1];
  // E, and so the spectrum, is zero on the columns the band mask zeroes
  if (wave_cols_zero<N>(a.zcol, id.col)) {
#pragma unroll
    for (int i = 0; i < R::R1; ++i) v[i] = make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int i = 0; i < R::R1; ++i) {
      int fy, fx;
      R::cspec(i * R::NT + tid, fy, fx);
      const float2 E = cscale(cmul(exy[fx], exy[N + fy]), mask[i]);
      v[i] = cscale(cmul(fp[i], E), s);
      if constexpr (AB)
        if (!a.moved) v[i] = cmul(v[i], a.plate[size_t(b) * N * N + i * R::NT + tid]);
    }
    rt_line_inv<N, false>(v, T, id.col, id.cj);
  }
  rt_cols_to_rows<N>(v, T, id);
  rt_line_inv<N, true>(v, T, id.row, id.rj);
  float2* dst = a.Pp + (size_t(b) * a.M + m) * N * N;
#pragma unroll
  for (int n = 0; n < R::R1; ++n) dst[id.row * N + id.rj + R::T * n] = v[n];
}

// probe-side adjoint (objective.py:641-658, 697-703): per window and mode
// the gradient summed over object modes is transformed, shifted back and
// accumulated (C-order) into the group's partial, with the position gradient
// -2/n^2 sum Re(K conj(FP) Q) in FP64; without a shift the natural-order
// gradient itself is accumulated (R-order)
template <int N, bool AB, bool SHIFT>
__global__ __launch_bounds__(RT<N>::NT) void k_probegrad_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  __shared__ float2 exy[2 * N];
  __shared__ double red[16];
  const int g = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  // the group's partial sum stays in registers across its windows (one read
  // and one write of the slot per launch; the slot carries earlier batches)
  const RView<N, float2> acc(a.probe_part + (size_t(g) * a.M + m) * N * N, tid);
  const RView<N, float2> fp(a.fprobe + size_t(m) * N * N, tid);
  const RView<N, float> mask(a.maskc, tid);
  float2 sum[R::R1];
#pragma unroll
  for (int i = 0; i < R::R1; ++i) sum[i] = acc[i];
  for (int b = g; b < a.B; b += a.G) {
    const int idx = seq_index(a.seq, a.b0 + b, a.P);
    float2 v[R::R1];
    {
      const RView<N, float2> g0(a.gP + (size_t(b) * a.M + m) * a.K * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) v[n] = g0[n];
    }
    for (int k = 1; k < a.K; ++k) {
      const RView<N, float2> gk(a.gP + ((size_t(b) * a.M + m) * a.K + k) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) v[n] = cadd(v[n], gk[n]);
    }
    const float bcs = bc_scale(a, idx);
    if constexpr (!SHIFT) {
      // natural order: register n of the row phase is R-order slot n*NT + tid,
      // element (row, rj + T*n)
      double sb = 0.0;
#pragma unroll
      for (int n = 0; n < R::R1; ++n) {
        sum[n] = cadd(sum[n], cscale(v[n], bcs));
        if (a.bc_part) {
          const float2 Pn = a.probe[size_t(m) * N * N + id.row * N + id.rj + R::T * n];
          sb += double(v[n].x) * double(Pn.x) + double(v[n].y) * double(Pn.y);
        }
      }
      if (a.bc_part) {
        const double SB = block_sum(sb, red);
        if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) * SB;
      }
      continue;
    }
    const bool mv = a.moved && idx >= 0;
    const float dy = mv ? a.corr[2 * idx] : 0.f, dx = mv ? a.corr[2 * idx + 1] : 0.f;
    float* wrow = AB && a.ab_w ? a.ab_w + (size_t(b) * a.M + m) * N * N : nullptr;
    __syncthreads();  // the previous window's table and transposes are done
    shift_tables_freq<N>(exy, dy, dx);
    rt_fft2_fwd<N>(v, T, id);  // its first barrier also publishes exy
    double sy = 0.0, sx = 0.0, sb = 0.0;
#pragma unroll
    for (int i = 0; i < R::R1; ++i) {
      int fy, fx;
      R::cspec(i * R::NT + tid, fy, fx);
      const float mk = mask[i];
      const float2 E = cscale(cmul(exy[fx], exy[N + fy]), mk);
      const float2 Q = cmulc(v[i], E);
      // after the conjugate aberration plate (objective.py:673-675); the mask is 0 or 1
      float2 Fg = Q;
      if constexpr (AB) Fg = cmulc(Q, a.plate[size_t(b) * N * N + i * R::NT + tid]);
      sum[i] = cadd(sum[i], cscale(Fg, bcs));
      if (a.step_pos || a.bc_part || (AB && wrow)) {
        const float2 F = fp[i];
        // Re(G * conj(FP)): the beam current's sum, Parseval's form of the real-space one
        sb += double(Fg.x) * double(F.x) + double(Fg.y) * double(F.y);
        if constexpr (AB)
          if (wrow) wrow[i * R::NT + tid] = float(double(bcs) * (double(Fg.y) * double(F.x) - double(Fg.x) * double(F.y)));
        // Im(conj(FP) * Q) in FP64: the terms of the sum nearly cancel
        const double zy = double(Q.y) * double(F.x) - double(Q.x) * double(F.y);
        const float kx = fmul(fmul(-6.2831855f, fftfreq_f(fx, N, 1.f)), mk);
        const float ky = fmul(fmul(-6.2831855f, fftfreq_f(fy, N, 1.f)), mk);
        sx += -double(kx) * zy;
        sy += -double(ky) * zy;
      }
    }
    if (a.step_pos) {
      const double SY = block_sum(sy, red);
      const double SX = block_sum(sx, red);
      if (tid == 0) {
        // the shifted probe's spectrum is |bc| FP
        const double sh = -2.0 * double(bcs) / (double(N) * double(N));
        a.pos_part[(size_t(b) * a.M + m) * 2] = sh * SY;
        a.pos_part[(size_t(b) * a.M + m) * 2 + 1] = sh * SX;
      }
    }
    if (a.bc_part) {
      const double SB = block_sum(sb, red);
      if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) / (double(N) * double(N)) * SB;
    }
  }
#pragma unroll
  for (int i = 0; i < R::R1; ++i) acc.store(i, sum[i]);
}

// probe_grad = ifft2(mask * fft2(sum)) (objective.py:927-933) from the reduced
// partials: C-order spectrum (spectral) or R-order natural gradient
template <int N>
__global__ __launch_bounds__(RT<N>::NT) void k_probe_final_rt(const float2* __restrict__ sum, int spectral,
                                                              const float* __restrict__ maskc,
                                                              float2* __restrict__ probe_grad) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  const RTIds<N> id;
  const unsigned tid = threadIdx.x;
  const RView<N, float2> src(sum + size_t(blockIdx.x) * N * N, tid);
  const RView<N, float> mask(maskc, tid);
  float2 v[R::R1];
#pragma unroll
  for (int i = 0; i < R::R1; ++i) v[i] = src[i];
  if (!spectral) rt_fft2_fwd<N>(v, T, id);
  const float s = 1.0f / float(N * N);
#pragma unroll
  for (int i = 0; i < R::R1; ++i) v[i] = cscale(v[i], mask[i] * s);
  rt_fft2_inv<N>(v, T, id);
  float2* dst = probe_grad + size_t(blockIdx.x) * N * N;
#pragma unroll
  for (int n = 0; n < R::R1; ++n) dst[id.row * N + id.rj + R::T * n] = v[n];
}

// ------------------------------------------ tilts before the specimen
// the real-space tilt grid value c_j = 2 pi (j - N/2) / N * N * p / lambda (objective.py:304-306)
__device__ __forceinline__ float tilt_grid(int j, int N, float p, float lambda) {
  return float(double(j - N / 2) / double(N) * 6.2831855 * double(N) * double(p) / double(lambda));
}

// each window's probe before its shift: |bc| P times exp(i (cx tx + cy ty))
// with the window's before-tilt (objective.py:446-450, 465-468), natural
// order into Pp; one thread per element of (B, M, N, N)
template <int N>
__global__ __launch_bounds__(256) void k_tilt_probe(BatchArgs a) {
  const size_t t = blockIdx.x * size_t(256) + threadIdx.x;
  if (t >= size_t(a.B) * a.M * N * N) return;
  const int b = int(t / (size_t(a.M) * N * N)), m = int((t / (N * N)) % a.M), e = int(t % (N * N));
  const int y = e / N, x = e % N;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const float* tr = tilt_row(a, idx);
  const float ty = tr ? tr[0] : 0.f, tx = tr ? tr[1] : 0.f;
  float sn, cs;
  sincosf(fadd(fmul(tilt_grid(x, N, a.px, a.lambda), tx), fmul(tilt_grid(y, N, a.py, a.lambda), ty)), &sn, &cs);
  a.Pp[t] = cmul(cscale(a.probe[size_t(m) * N * N + e], bc_scale(a, idx)), make_float2(cs, sn));
}

// probe-side adjoint with a before-tilt (objective.py:641-672, 688-703) over
// the register tile: per window and mode the gradient summed over object
// modes, with a shift transformed, shifted back (position gradient from the
// window's own spectrum Fw, FP64) and transformed back; then times the
// conjugate ramp, the tilt's gradient 2 sum c Im(g conj(|bc| P)) and the beam
// current's, and added in natural order (R-order) to the group's partial
template <int N>
__global__ __launch_bounds__(RT<N>::NT) void k_probegrad_tilt_rt(BatchArgs a) {
  using R = RT<N>;
  rt_tw_load<N>();
  __shared__ float2 T[R::ELEMS];
  __shared__ float2 exy[2 * N];
  __shared__ float2 tt[N];
  __shared__ double red[16];
  const int g = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const unsigned tid = threadIdx.x;
  const RTIds<N> id;
  tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
  const RView<N, float2> acc(a.probe_part + (size_t(g) * a.M + m) * N * N, tid);
  const RView<N, float> mask(a.maskc, tid);
  float2 sum[R::R1];
#pragma unroll
  for (int i = 0; i < R::R1; ++i) sum[i] = acc[i];
  __syncthreads();
  for (int b = g; b < a.B; b += a.G) {
    const int idx = seq_index(a.seq, a.b0 + b, a.P);
    float2 v[R::R1];
    {
      const RView<N, float2> g0(a.gP + (size_t(b) * a.M + m) * a.K * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) v[n] = g0[n];
    }
    for (int k = 1; k < a.K; ++k) {
      const RView<N, float2> gk(a.gP + ((size_t(b) * a.M + m) * a.K + k) * N * N, tid);
#pragma unroll
      for (int n = 0; n < R::R1; ++n) v[n] = cadd(v[n], gk[n]);
    }
    if (a.moved) {
      const float dy = idx >= 0 ? a.corr[2 * idx] : 0.f, dx = idx >= 0 ? a.corr[2 * idx + 1] : 0.f;
      __syncthreads();  // the previous window's table and transposes are done
      shift_tables_freq<N>(exy, dy, dx);
      rt_fft2_fwd<N>(v, T, id);  // its first barrier also publishes exy
      const RView<N, float2> F(a.Fw + (size_t(b) * a.M + m) * N * N, tid);
      double sy = 0.0, sx = 0.0;
      const float sc = 1.0f / float(N * N);
#pragma unroll
      for (int i = 0; i < R::R1; ++i) {
        int fy, fx;
        R::cspec(i * R::NT + tid, fy, fx);
        const float mk = mask[i];
        const float2 Q = cmulc(v[i], cscale(cmul(exy[fx], exy[N + fy]), mk));
        if (a.step_pos) {
          const float2 f = F[i];
          const double zy = double(Q.y) * double(f.x) - double(Q.x) * double(f.y);
          sx += -double(fmul(fmul(-6.2831855f, fftfreq_f(fx, N, 1.f)), mk)) * zy;
          sy += -double(fmul(fmul(-6.2831855f, fftfreq_f(fy, N, 1.f)), mk)) * zy;
        }
        v[i] = cscale(Q, sc);
      }
      if (a.step_pos) {
        const double SY = block_sum(sy, red);
        const double SX = block_sum(sx, red);
        if (tid == 0) {
          const double sh = -2.0 / (double(N) * double(N));
          a.pos_part[(size_t(b) * a.M + m) * 2] = sh * SY;
          a.pos_part[(size_t(b) * a.M + m) * 2 + 1] = sh * SX;
        }
      }
      rt_fft2_inv<N>(v, T, id);
    }
    const float* tr = tilt_row(a, idx);
    const float ty = tr ? tr[0] : 0.f, tx = tr ? tr[1] : 0.f;
    const float bcs = bc_scale(a, idx);
    double ty_s = 0.0, tx_s = 0.0, sb = 0.0;
#pragma unroll
    for (int n = 0; n < R::R1; ++n) {
      const int y = id.row, x = id.rj + R::T * n;
      v[n] = cmulc(v[n], tilt_ramp(tt, y, x, tx, ty));
      const float2 P = a.probe[size_t(m) * N * N + y * N + x];
      const double zi = double(v[n].y) * double(P.x) - double(v[n].x) * double(P.y);  // Im(g conj(P))
      ty_s += double(tt[y].y) * zi;
      tx_s += double(tt[x].x) * zi;
      sb += double(v[n].x) * double(P.x) + double(v[n].y) * double(P.y);
      sum[n] = cadd(sum[n], cscale(v[n], bcs));
    }
    if (a.tg_before) {
      const double SY = block_sum(ty_s, red);
      const double SX = block_sum(tx_s, red);
      if (tid == 0) {
        a.tb_part[(size_t(b) * a.M + m) * 2] = 2.0 * double(bcs) * SY;
        a.tb_part[(size_t(b) * a.M + m) * 2 + 1] = 2.0 * double(bcs) * SX;
      }
    }
    if (a.bc_part) {
      const double SB = block_sum(sb, red);
      if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) * SB;
    }
  }
#pragma unroll
  for (int i = 0; i < R::R1; ++i) acc.store(i, sum[i]);
}

// ------------------------------------------ global-slab pipeline (192, 256)
// The per-batch steps for windows of 192 and 256 px over GT (pty_rt.hpp):
// each phase runs in rounds of GT::LINES lines, the transposes go through
// the workgroup's slab in `scratch`, LDS holds the wave-local exchanges.
// Buffer orders as for the register tile (GT R-order / C-order).

// the workgroup's slab: element (y, x) at byte (y*N + x)*8 of a buffer
// resource over the slab, lane part in voff, uniform part in soff
template <int N>
struct Slab {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ explicit Slab(float2* p)
      : r(__builtin_amdgcn_make_buffer_rsrc(p, 0, N * N * 8, 0x00020000)) {}
  __device__ __forceinline__ float2 ld(unsigned voff, int soff) const {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
  __device__ __forceinline__ void st(unsigned voff, int soff, float2 v) const {
    using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, v), r, voff, soff, 0);
  }
};

template <int N>
struct GTLane {
  int ln, j;               // line within a round, position within the line
  unsigned vrow, vcol;     // slab lane offsets of the row / column phase
  __device__ __forceinline__ GTLane() {
    const int tid = threadIdx.x;
    ln = 8 * (tid >> 6) + ((tid & 63) >> 3);
    j = tid & 7;
    vrow = unsigned(ln * N + j) * 8u;
    vcol = unsigned(j * N + ln) * 8u;
  }
};

// row spectrum of round r: register c*8 + k2 <-> slab (r*LINES + ln, 8*(CPT*k2 + c) + j)
// column of round r: register n <-> slab (j + 8n, r*LINES + ln)
// The global slab (256 px): the phases meet in workspace memory, so a phase
// change is a workgroup barrier.
template <int N>
struct GSlab {
  using G = GT<N, kGtThreads<N>>;
  Slab<N> sl;
  __device__ __forceinline__ explicit GSlab(float2* p) : sl(p) {}
  __device__ __forceinline__ void load_rowspec(float2* v, const GTLane<N>& L, int r) const {
#pragma unroll
    for (int c = 0; c < G::CPT; ++c)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) v[c * 8 + k2] = sl.ld(L.vrow, (r * G::LINES * N + 8 * (G::CPT * k2 + c)) * 8);
  }
  __device__ __forceinline__ void store_rowspec(const float2* v, const GTLane<N>& L, int r) {
#pragma unroll
    for (int c = 0; c < G::CPT; ++c)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) sl.st(L.vrow, (r * G::LINES * N + 8 * (G::CPT * k2 + c)) * 8, v[c * 8 + k2]);
  }
  __device__ __forceinline__ void load_col(float2* v, const GTLane<N>& L, int r) const {
#pragma unroll
    for (int n = 0; n < G::R1; ++n) v[n] = sl.ld(L.vcol, (8 * n * N + r * G::LINES) * 8);
  }
  __device__ __forceinline__ void store_col(const float2* v, const GTLane<N>& L, int r) {
#pragma unroll
    for (int n = 0; n < G::R1; ++n) sl.st(L.vcol, (8 * n * N + r * G::LINES) * 8, v[n]);
  }
  __device__ __forceinline__ void to_cols(float2*, const GTLane<N>&) { __syncthreads(); }
  __device__ __forceinline__ void to_rows(float2*, const GTLane<N>&) { __syncthreads(); }
  // Columns of the band mask's zero set are exactly zero after a propagation
  // (H and H0 vanish on them).  The slab's columns fall into groups of eight
  // (8g .. 8g+7): one column-phase wavefront of a round, one register of the
  // row phase.  A propagation's column phase does not store a group whose
  // columns are all zero, and the next row phase does not load it (bit q of
  // zg, the same for every lane, for register q) but sets it to zero.
  __device__ __forceinline__ void load_rowspec_z(float2* v, const GTLane<N>& L, int r, unsigned zg) const {
#pragma unroll
    for (int c = 0; c < G::CPT; ++c)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const int q = c * 8 + k2;
        if ((zg >> q) & 1u)
          v[q] = make_float2(0.f, 0.f);
        else
          v[q] = sl.ld(L.vrow, (r * G::LINES * N + 8 * (G::CPT * k2 + c)) * 8);
      }
  }
  // row spectra before a column phase that leaves the zero groups alone
  // (a propagation, the far field): those groups are not stored
  __device__ __forceinline__ void store_rowspec_z(const float2* v, const GTLane<N>& L, int r, unsigned zg) {
#pragma unroll
    for (int c = 0; c < G::CPT; ++c)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const int q = c * 8 + k2;
        if (!((zg >> q) & 1u)) sl.st(L.vrow, (r * G::LINES * N + 8 * (G::CPT * k2 + c)) * 8, v[q]);
      }
  }
};

// bit c*8 + k2: every column of the group that row-phase register c*8 + k2
// reads (8g .. 8g+7, g = CPT*k2 + c) is a zero column of the band mask; the
// same for every lane and round
template <int N>
__device__ __forceinline__ unsigned gt_row_zero_groups(const int* zcol, const GTLane<N>& L) {
  using G = GT<N, kGtThreads<N>>;
  static_assert(G::R1 <= 32, "one bit per register");
  unsigned z = 0;
  if (zcol) {
#pragma unroll
    for (int c = 0; c < G::CPT; ++c)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const bool zero = zcol[RT<N>::col_freq(8 * (G::CPT * k2 + c) + L.j)] != 0;
        if (__ballot(!zero) == 0ull) z |= 1u << (c * 8 + k2);
      }
  }
  return z;
}

// The 192 px tile in registers (768 threads, two rounds of 96 lines, 24
// samples per thread and round: 96 VGPRs).  A phase change moves it through
// LDS in two passes of half the tile (96 rows of stride 200, 150 KB), the LDS
// that the line transforms use for their wave-local exchanges:
//   to_cols, pass p: every thread writes row p*96 + ln (its own line, so no
//   wavefront touches another's exchange rows); barrier; it reads rows
//   j + 8n (n < 12) of its columns ln and 96 + ln, the samples 12p .. 12p+11
//   of both its column rounds; barrier
//   to_rows: barrier (the column exchanges of the other wavefronts are done);
//   then per pass p the reverse
// Registers follow the slab's orders exactly, so every kernel written
// against the slab runs unchanged on it.  The kernels unroll their round
// loops, so d[r] is always a register.
struct RegSlab {
  static constexpr int N = 192;
  using G = GT<N, kGtThreads<N>>;
  static constexpr int R1 = G::R1, CPT = G::CPT, LINES = G::LINES, LS = G::LS, H = R1 / 2;
  static_assert(G::ROUNDS == 2 && 8 * H == LINES, "two rounds of 96 lines");
  float2 d[2][R1];
  __device__ __forceinline__ explicit RegSlab(float2*) {}
  __device__ __forceinline__ void load_rowspec(float2* v, const GTLane<N>&, int r) {
#pragma unroll
    for (int i = 0; i < R1; ++i) v[i] = d[r][i];
  }
  __device__ __forceinline__ void store_rowspec(const float2* v, const GTLane<N>&, int r) {
#pragma unroll
    for (int i = 0; i < R1; ++i) d[r][i] = v[i];
  }
  __device__ __forceinline__ void load_col(float2* v, const GTLane<N>&, int r) {
#pragma unroll
    for (int i = 0; i < R1; ++i) v[i] = d[r][i];
  }
  __device__ __forceinline__ void store_col(const float2* v, const GTLane<N>&, int r) {
#pragma unroll
    for (int i = 0; i < R1; ++i) d[r][i] = v[i];
  }
  // A column phase that skips round r (its columns are zero columns of the
  // band mask) sets it to zero here, so the next row phase reads the zero
  // groups as zeros straight from the tile (GSlab::load_rowspec_z selects them
  // instead; a select per register here cost 20 registers).
  __device__ __forceinline__ void zero_round(int r) {
#pragma unroll
    for (int i = 0; i < R1; ++i) d[r][i] = make_float2(0.f, 0.f);
  }
  __device__ __forceinline__ void to_cols(float2* X, const GTLane<N>& L) {
    float2 t[2][2][H];  // [pass][column round][sample]
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float2* w = X + L.ln * LS + L.j;
#pragma unroll
      for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) w[8 * (CPT * k2 + c)] = d[p][c * 8 + k2];
      __syncthreads();
      const float2* q0 = X + L.j * LS + L.ln;
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int n = 0; n < H; ++n) t[p][q][n] = q0[8 * n * LS + q * LINES];
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int n = 0; n < R1; ++n) d[q][n] = n < H ? t[0][q][n] : t[1][q][n - H];
  }
  __device__ __forceinline__ void to_rows(float2* X, const GTLane<N>& L) {
    float2 t[2][R1];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float2* q0 = X + L.j * LS + L.ln;
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int n = 0; n < H; ++n) q0[8 * n * LS + q * LINES] = d[q][n + H * p];
      __syncthreads();
      const float2* w = X + L.ln * LS + L.j;
#pragma unroll
      for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) t[p][c * 8 + k2] = w[8 * (CPT * k2 + c)];
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int i = 0; i < R1; ++i) d[r][i] = t[r][i];
  }
};

template <int N>
using SlabOf = std::conditional_t<N == 192, RegSlab, GSlab<N>>;

// Between the rounds of a phase on the register tile: the rounds are
// independent, and a schedule that interleaves two rounds' transforms needs
// both rounds' temporaries on top of the tile's 96 registers.
template <class SL>
__device__ __forceinline__ void round_fence() {
  if constexpr (std::is_same_v<SL, RegSlab>) {
    // nor may a round's table or object reads move up into the previous round
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int N, class SL>
__device__ __forceinline__ void gt_load_rowspec(float2* v, SL& sl, const GTLane<N>& L, int r) {
  sl.load_rowspec(v, L, r);
}
template <int N, class SL>
__device__ __forceinline__ void gt_store_rowspec(const float2* v, SL& sl, const GTLane<N>& L, int r) {
  sl.store_rowspec(v, L, r);
}
template <int N, class SL>
__device__ __forceinline__ void gt_load_col(float2* v, SL& sl, const GTLane<N>& L, int r) {
  sl.load_col(v, L, r);
}
template <int N, class SL>
__device__ __forceinline__ void gt_store_col(const float2* v, SL& sl, const GTLane<N>& L, int r) {
  sl.store_col(v, L, r);
}

// bit r set: the wavefront's columns of round r are all marked zero
template <int N>
__device__ __forceinline__ unsigned gt_cols_zero(const int* zcol, const GTLane<N>& L) {
  using G = GT<N, kGtThreads<N>>;
  unsigned z = 0;
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) z |= unsigned(wave_cols_zero<N>(zcol, r * G::LINES + L.ln)) << r;
  return z;
}
template <int N, class SL>
__device__ __forceinline__ void gt_zero_col(float2* v, SL& sl, const GTLane<N>& L, int r) {
#pragma unroll
  for (int i = 0; i < GT<N, kGtThreads<N>>::R1; ++i) v[i] = make_float2(0.f, 0.f);
  gt_store_col<N>(v, sl, L, r);
}

// column phase of a band limit: forward column FFT, times maskc/N^2, inverse;
// zero columns of the mask are written as zeros
template <int N, class SL>
__device__ __forceinline__ void gt_band_columns(float2* v, float2* X, SL& sl, const GTLane<N>& L,
                                                const float* maskc, const int* zcol) {
  using G = GT<N, kGtThreads<N>>;
  const RView<N, float, G::NT> mask(maskc, threadIdx.x);
  const float sc = 1.0f / float(N * N);
  const unsigned zm = gt_cols_zero<N>(zcol, L);
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    if ((zm >> r) & 1u) {
      gt_zero_col<N>(v, sl, L, r);
      continue;
    }
    gt_load_col<N>(v, sl, L, r);
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
    for (int i = 0; i < G::R1; ++i) v[i] = cscale(v[i], mask[r * G::R1 + i] * sc);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    gt_store_col<N>(v, sl, L, r);
  }
}

// k_patch for 192/256 px windows (objective.py:436-441)
template <int N>
__global__ __launch_bounds__(kGtThreads<N>) void k_patch_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  const int planes = a.S * a.K;
  const int b = blockIdx.x / planes, plane = blockIdx.x % planes;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const int y0 = idx >= 0 ? a.pos[2 * idx] : -1, x0 = idx >= 0 ? a.pos[2 * idx + 1] : -1;
  const bool inside = window_inside(y0, x0, N, a.ny, a.nx);
  const GTLane<N> L;
  SlabOf<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  float2 v[G::R1];
  const float2* obj = a.obj + size_t(plane) * a.ny * a.nx;
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    const int row = r * G::LINES + L.ln;
    const float2* src = obj + (inside ? size_t(y0 + row) * a.nx + x0 + L.j : 0);
#pragma unroll
    for (int n = 0; n < G::R1; ++n) v[n] = inside ? src[8 * n] : make_float2(0.f, 0.f);
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
    gt_store_rowspec<N>(v, sl, L, r);
  }
  sl.to_cols(X, L);
  gt_band_columns<N>(v, X, sl, L, a.maskc, a.zcol);
  sl.to_rows(X, L);
  const RView<N, float2, G::NT> dst(a.Ob + (size_t(b) * planes + plane) * N * N, threadIdx.x);
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    gt_load_rowspec<N>(v, sl, L, r);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
#pragma unroll
    for (int n = 0; n < G::R1; ++n) dst.store(r * G::R1 + n, v[n]);
  }
}

// The forward at 192 px with the whole tile in registers (RegSlab, 768
// threads, 168 registers).  The tile leaves about 70 registers for the
// transforms and everything else, so the control flow is kept flat: the
// first slice's row phase (from the probe or a recompute's initial wave, one
// load path for both) is peeled off the slice loop, the far field follows
// the loop, a stash store without a slot goes to an empty buffer resource
// (the hardware drops it), and a recompute segment's last slice runs its
// row transform like any other.  The object and the propagator are read in
// chunks of four registers where they are used.  Same results as the slab
// pipeline (k_forward_gt's general body), which it replaces at 192 px.
__device__ __forceinline__ void forward_reg192(const BatchArgs& a, float2* X) {
  constexpr int N = 192, CH = 12, CHC = 12;
  using G = GT<N, kGtThreads<N>>;
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int m = mk / a.K, k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  RegSlab sl(nullptr);
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  const unsigned zh = gt_cols_zero<N>(a.zcol, L);
  float2 v[G::R1];
  // v holds round r of the wave entering slice s: stash it, times the object
  auto transmit = [&](int s, int r) {
    const int slot = a.write_stash ? stash_slot(a, s) : -1;
    const __amdgpu_buffer_rsrc_t U = __builtin_amdgcn_make_buffer_rsrc(
        st + size_t(slot >= 0 ? slot : 0) * N * N, 0, slot >= 0 ? N * N * 8 : 0, 0x00020000);
    const RView<N, float2, G::NT> O(ob + size_t(s) * a.K * N * N, tid);
    using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(U, 0, 0, 0));
#pragma unroll
    for (int n0 = 0; n0 < G::R1; n0 += CH) {
      __builtin_amdgcn_sched_barrier(0);
      float2 o[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) o[c] = O[r * G::R1 + n0 + c];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, v[n0 + c]), U, tid * 8u,
                                              (r * G::R1 + n0 + c) * G::NT * 8, 0);
        v[n0 + c] = cmul(v[n0 + c], o[c]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  // the wave entering slice s_lo: a recompute's initial wave (R-order slots,
  // element (r*R1 + n)*NT + tid) or the probe (natural rows, element
  // (r*LINES + ln)*N + j + 8n; the shifted probes carry the beam current)
  __amdgpu_buffer_rsrc_t W0;
  unsigned w_voff;
  int w_round, w_step;
  float w_scale;
  if (a.fin) {
    W0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(a.fin + (size_t(b) * MK + mk) * a.fin_stride), 0,
                                           N * N * 8, 0x00020000);
    w_voff = tid * 8u;
    w_round = G::R1 * G::NT * 8;
    w_step = G::NT * 8;
    w_scale = 1.f;
  } else {
    const float2* src = a.shift ? a.Pp + (size_t(b) * a.M + m) * N * N : a.probe + size_t(m) * N * N;
    W0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(src), 0, N * N * 8, 0x00020000);
    w_voff = unsigned(L.ln * N + L.j) * 8u;
    w_round = G::LINES * N * 8;
    w_step = 8 * 8;
    w_scale = a.shift ? 1.f : bc_scale(a, seq_index(a.seq, a.b0 + b, a.P));
  }
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<RegSlab>();
#pragma unroll
    for (int n = 0; n < G::R1; ++n)
      v[n] = cscale(__builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(W0, w_voff, r * w_round + n * w_step, 0)),
                    w_scale);
    transmit(a.s_lo, r);
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
    sl.store_rowspec(v, L, r);
  }
  // a recompute segment wants the stash of its slices, not the far field
  const int s_end = a.fin ? a.s_hi - 1 : a.s_hi;
  for (int s = a.s_lo; s < a.s_hi - 1; ++s) {
    sl.to_cols(X, L);
    // propagation into slice s + 1 (H0 = H^e for the first)
    const RView<N, float2, G::NT> H(s == 0 ? a.H0 : a.H + size_t(s) * a.Hstride, tid);
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<RegSlab>();
      if ((zh >> r) & 1u) {
        sl.zero_round(r);
        continue;
      }
      sl.load_col(v, L, r);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
      for (int i0 = 0; i0 < G::R1; i0 += CHC) {
        __builtin_amdgcn_sched_barrier(0);
        float2 h[CHC];
#pragma unroll
        for (int c = 0; c < CHC; ++c) h[c] = H[r * G::R1 + i0 + c];
#pragma unroll
        for (int c = 0; c < CHC; ++c) v[i0 + c] = cmul(v[i0 + c], h[c]);
      }
      __builtin_amdgcn_sched_barrier(0);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
      sl.store_col(v, L, r);
    }
    sl.to_rows(X, L);
    // row phase of slice s + 1
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<RegSlab>();
      sl.load_rowspec(v, L, r);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
      transmit(s + 1, r);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
      sl.store_rowspec(v, L, r);
    }
  }
  if (s_end < a.s_hi) return;
  // far field: forward column transform times the band mask
  sl.to_cols(X, L);
  const RView<N, float2, G::NT> out(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
  const RView<N, float, G::NT> mask(a.maskc, tid);
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<RegSlab>();
    if ((zh >> r) & 1u) {
#pragma unroll
      for (int i = 0; i < G::R1; ++i) out.store(r * G::R1 + i, make_float2(0.f, 0.f));
      continue;
    }
    sl.load_col(v, L, r);
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
    for (int i = 0; i < G::R1; ++i) out.store(r * G::R1 + i, cscale(v[i], mask[r * G::R1 + i]));
  }
}

// k_forward for 192/256 px windows (multislice_core.py:9-82, objective.py:501-512)
template <int N, bool TILT>
__global__ __launch_bounds__(kGtThreads<N>) void k_forward_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  __shared__ float2 tt[TILT ? 2 * N : 1];  // tilt factors or grids
  if constexpr (N == 192 && !TILT) {
    forward_reg192(a, X);
    return;
  }
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int m = mk / a.K, k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  GSlab<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  const float* trow = TILT ? tilt_row(a, seq_index(a.seq, a.b0 + b, a.P)) : nullptr;
  const float tix = trow ? trow[3] : 0.f, tiy = trow ? trow[2] : 0.f;
  float2* wsp = (TILT && a.tg_in) ? a.Wsp + (size_t(b) * MK + mk) * a.S * N * N : nullptr;
  if ((TILT && a.tilt_win) && !a.dists) {
    tilt_tables_inside<N>(tt, a.dz, tix, tiy, a.px, a.py);
    __syncthreads();
  }
  const float2* src = a.shift ? a.Pp + (size_t(b) * a.M + m) * N * N : a.probe + size_t(m) * N * N;
  // the shifted probes carry the beam current already
  const float bcs = a.shift ? 1.f : bc_scale(a, seq_index(a.seq, a.b0 + b, a.P));
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  const float2* fin = a.fin ? a.fin + (size_t(b) * MK + mk) * a.fin_stride : nullptr;
  // rounds whose columns the propagators (the band mask) zero, and the row
  // phase's groups of zero columns
  const unsigned zh = gt_cols_zero<N>(a.zcol, L), zm = zh;
  const unsigned zrow = gt_row_zero_groups<N>(a.zcol, L);
  // o holds the object registers of the current round; those of the next
  // round (or the next slice's first round) are read while this round's
  // forward row transform runs, and the propagator is read before the column
  // transform that precedes its use
  // at 192 px (768 threads, 168 registers) reading ahead costs spills
  constexpr bool kAhead = N != 192;
  float2 v[G::R1], o[G::R1];
  if constexpr (kAhead) {
    const RView<N, float2, G::NT> O(ob + size_t(a.s_lo) * a.K * N * N, tid);
#pragma unroll
    for (int n = 0; n < G::R1; ++n) o[n] = O[n];
  }
  for (int s = a.s_lo; s < a.s_hi; ++s) {
    // row phase: inverse row transform of the previous slice, stash, object, forward row transform
    const int slot = a.write_stash ? stash_slot(a, s) : -1;
    const bool keep = slot >= 0;
    const RView<N, float2, G::NT> U(st + size_t(keep ? slot : 0) * N * N, tid);
    const bool last_of_recompute = fin && s == a.s_hi - 1;
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      if (s == a.s_lo && fin) {
        const RView<N, float2, G::NT> I(fin, tid);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = I[r * G::R1 + n];
      } else if (s == a.s_lo) {
        const float2* p = src + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = cscale(p[8 * n], bcs);
      } else {
        // after this launch's propagation into slice s: zero columns unread
        sl.load_rowspec_z(v, L, r, zrow);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
      }
      if constexpr (!kAhead) {
        const RView<N, float2, G::NT> O(ob + size_t(s) * a.K * N * N, tid);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) o[n] = O[r * G::R1 + n];
      }
#pragma unroll
      for (int n = 0; n < G::R1; ++n) {
        if (keep) U.store(r * G::R1 + n, v[n]);
        v[n] = cmul(v[n], o[n]);
      }
      if (last_of_recompute) continue;  // only the stash of this slice is wanted
      const int rn = (r + 1 < G::ROUNDS) ? r + 1 : 0, sn = (r + 1 < G::ROUNDS) ? s : s + 1;
      if (kAhead && sn < a.s_hi) {
        const RView<N, float2, G::NT> O(ob + size_t(sn) * a.K * N * N, tid);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) o[n] = O[rn * G::R1 + n];
      }
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
      // an after-tilt reads the last slice's row spectra whole
      if ((TILT && a.tilt_after) && s == a.S - 1)
        gt_store_rowspec<N>(v, sl, L, r);
      else
        sl.store_rowspec_z(v, L, r, zrow);
    }
    if (last_of_recompute) break;
    if ((TILT && a.tilt_after) && s == a.S - 1) {
      // the exit wave, band-limited first for one slice, times the after-tilt
      // ramp (objective.py:490-499); Xe keeps it before the ramp
      sl.to_cols(X, L);
      if (a.S == 1) gt_band_columns<N>(v, X, sl, L, a.maskc, a.zcol);
      sl.to_rows(X, L);
      tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
      __syncthreads();
      const float tax = trow ? trow[5] : 0.f, tay = trow ? trow[4] : 0.f;
      const RView<N, float2, G::NT> XE(a.Xe ? a.Xe + (size_t(b) * MK + mk) * N * N : a.Psi, tid);
      // the row spectra back to rows: the band limit's 1/N^2 leaves the wave
      // itself, else the unnormalised row pair leaves N times it
      const float rs = a.S == 1 ? 1.f : 1.f / float(N);
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        gt_load_rowspec<N>(v, sl, L, r);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) {
          v[n] = cscale(v[n], rs);
          if ((TILT && a.tg_after)) XE.store(r * G::R1 + n, v[n]);
          v[n] = cmul(v[n], tilt_ramp(tt, r * G::LINES + L.ln, L.j + 8 * n, tax, tay));
        }
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        gt_store_rowspec<N>(v, sl, L, r);
      }
    }
    sl.to_cols(X, L);
    // column phase: forward column transform, then the propagator and the
    // inverse column transform, or the far-field band mask after the last slice
    if (s < a.s_hi - 1) {
      if ((TILT && a.tilt_win) && a.dists) {
        // per-slice distances: this propagation's own tilt factors
        __syncthreads();
        tilt_tables_inside<N>(tt, a.dists[s], tix, tiy, a.px, a.py);
        __syncthreads();
      }
      // with an edge multiplier, the first propagation is H0 = H^e; a window
      // with its own tilt raises its tilted propagator itself
      const RView<N, float2, G::NT> H((s == 0 && !(TILT && a.tilt_win)) ? a.H0 : a.H + size_t(s) * a.Hstride, tid);
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        if ((zh >> r) & 1u) continue;  // zero columns: neither computed nor stored
        float2 h[G::R1];
#pragma unroll
        for (int i = 0; i < G::R1; ++i) h[i] = H[r * G::R1 + i];
        gt_load_col<N>(v, sl, L, r);
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        if ((TILT && a.tilt_win)) {
#pragma unroll
          for (int i = 0; i < G::R1; ++i) {
            int fy, fx;
            G::cspec((r * G::R1 + i) * G::NT + tid, fy, fx);
            h[i] = tilted_h(h[i], tt, fy, fx, N, s == 0, a.edge);
          }
        }
#pragma unroll
        for (int i = 0; i < G::R1; ++i) v[i] = cmul(v[i], h[i]);
        if (wsp) {
          // fft2 of the wave entering slice s + 1 (over N^2), for the inside-tilt gradient
          const RView<N, float2, G::NT> W(wsp + size_t(s + 1) * N * N, tid);
#pragma unroll
          for (int i = 0; i < G::R1; ++i) W.store(r * G::R1 + i, v[i]);
        }
        rt_line_inv<N, true>(v, X, L.ln, L.j);
        gt_store_col<N>(v, sl, L, r);
      }
      sl.to_rows(X, L);
    } else {
      const RView<N, float2, G::NT> out(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
      const RView<N, float, G::NT> mask(a.maskc, tid);
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        if ((zm >> r) & 1u) {
#pragma unroll
          for (int i = 0; i < G::R1; ++i) out.store(r * G::R1 + i, make_float2(0.f, 0.f));
          continue;
        }
        gt_load_col<N>(v, sl, L, r);
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
        for (int i = 0; i < G::R1; ++i) out.store(r * G::R1 + i, cscale(v[i], mask[r * G::R1 + i]));
      }
    }
  }
}

// The adjoint at 192 px with the whole tile in registers (RegSlab), laid out
// like forward_reg192: flat control flow around the 168 registers; for
// segments ending above slice 0 (s_hi > 1; k_adjoint_gt runs the others on
// the slab pipeline).  The row phases come in three compile-time forms (from the tile or from a later
// segment's wave; ending in a forward row transform or, at slice 0, in the
// probe-side gradient), chosen outside the round loops; the stash is read
// and the window gradient written in chunks of four registers, through
// buffer resources that are empty when the object is not refined (the
// hardware returns zeros for such reads and drops such writes).  Same
// results as the slab pipeline (k_adjoint_gt's general body).
__device__ __forceinline__ void adjoint_reg192(const BatchArgs& a, float2* X) {
  constexpr int N = 192, CH = 8, CHC = 8;
  using G = GT<N, kGtThreads<N>>;
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  RegSlab sl(nullptr);
  const unsigned zh = gt_cols_zero<N>(a.zcol, L);
  // a later segment left the wave of slice s_hi - 1 (R-order)
  const bool from_state = a.s_hi < a.S;
  const RView<N, float2, G::NT> GS(a.gstate ? a.gstate + (size_t(b) * MK + mk) * N * N : a.Psi, tid);
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  float2 v[G::R1];
  // row phase of slice s: v = the gradient wave at slice s (natural rows);
  // the stash slot becomes this mode's window gradient g conj(u_s), then
  // g conj(O_s) goes on through the forward row transform into the tile, or
  // at slice 0 out as the probe-side gradient (natural order)
  auto row_phase = [&](int s, auto from_gs, auto last) {
    const __amdgpu_buffer_rsrc_t U = __builtin_amdgcn_make_buffer_rsrc(st + size_t(s - a.s_lo) * N * N, 0,
                                                                       a.step_obj ? N * N * 8 : 0, 0x00020000);
    using B64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(U, 0, 0, 0));
    const RView<N, float2, G::NT> O(ob + size_t(s) * a.K * N * N, tid);
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<RegSlab>();
      if constexpr (decltype(from_gs)::value) {
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = GS[r * G::R1 + n];
      } else {
        sl.load_rowspec(v, L, r);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
      }
#pragma unroll
      for (int n0 = 0; n0 < G::R1; n0 += CH) {
        __builtin_amdgcn_sched_barrier(0);
        float2 pu[CH], o[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int so = (r * G::R1 + n0 + c) * G::NT * 8;
          pu[c] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(U, tid * 8u, so, 0));
          o[c] = O[r * G::R1 + n0 + c];
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int so = (r * G::R1 + n0 + c) * G::NT * 8;
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B64, cmulc(v[n0 + c], pu[c])), U, tid * 8u, so, 0);
          v[n0 + c] = cmulc(v[n0 + c], o[c]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (decltype(last)::value) {
        float2* out = a.gP + (size_t(b) * MK + mk) * N * N + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) out[8 * n] = v[n];
      } else {
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        sl.store_rowspec(v, L, r);
      }
    }
  };
  // column phase back through the propagation into slice s (s > 0; H0 for s = 1)
  auto propagate_back = [&](int s) {
    sl.to_cols(X, L);
    const RView<N, float2, G::NT> H(s == 1 ? a.H0 : a.H + size_t(s - 1) * a.Hstride, tid);
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<RegSlab>();
      if ((zh >> r) & 1u) {
        sl.zero_round(r);
        continue;
      }
      sl.load_col(v, L, r);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
      for (int i0 = 0; i0 < G::R1; i0 += CHC) {
        __builtin_amdgcn_sched_barrier(0);
        float2 h[CHC];
#pragma unroll
        for (int c = 0; c < CHC; ++c) h[c] = H[r * G::R1 + i0 + c];
#pragma unroll
        for (int c = 0; c < CHC; ++c) v[i0 + c] = cmulc(v[i0 + c], h[c]);
      }
      __builtin_amdgcn_sched_barrier(0);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
      sl.store_col(v, L, r);
    }
    sl.to_rows(X, L);
  };
  const int s_top = a.s_hi - 1;
  if (!from_state) {
    // far-field adjoint: inverse column transforms of coef * Psi / N^2
    const RView<N, float2, G::NT> psi(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
    const RView<N, float, G::NT> c(a.coef + size_t(b) * N * N, tid);
    const float sc = 1.0f / float(N * N);
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<RegSlab>();
      if ((zh >> r) & 1u) {
        sl.zero_round(r);
        continue;
      }
#pragma unroll
      for (int i = 0; i < G::R1; ++i) v[i] = cscale(psi[r * G::R1 + i], c[r * G::R1 + i] * sc);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
      sl.store_col(v, L, r);
    }
    sl.to_rows(X, L);
  }
  using T_ = std::true_type;
  using F_ = std::false_type;
  if (from_state)
    row_phase(s_top, T_{}, F_{});
  else
    row_phase(s_top, F_{}, F_{});
  propagate_back(s_top);
  const int s_stop = a.s_lo > 1 ? a.s_lo : 1;
  for (int s = s_top - 1; s >= s_stop; --s) {
    row_phase(s, F_{}, F_{});
    propagate_back(s);
  }
  if (a.s_lo == 0) {
    row_phase(0, F_{}, T_{});
    return;
  }
  // the wave propagated to slice s_lo - 1 (R-order) for the next segment
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<RegSlab>();
    sl.load_rowspec(v, L, r);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
#pragma unroll
    for (int n = 0; n < G::R1; ++n) GS.store(r * G::R1 + n, v[n]);
  }
}

// k_adjoint for 192/256 px windows (objective.py:595-597, multislice_core.py:85-200)
template <int N, bool TILT>
__global__ __launch_bounds__(kGtThreads<N>) void k_adjoint_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  __shared__ float2 tt[TILT ? 2 * N : 1];  // tilt factors or grids
  __shared__ double red[TILT ? 16 : 1];
  if constexpr (N == 192 && !TILT) {
    // a first row phase that is also the last (slice 0 alone) takes the
    // slab pipeline below: as a further path beside the register tile's it
    // cost 185-204 spilled registers
    if (a.s_hi > 1) {
      adjoint_reg192(a, X);
      return;
    }
  }
  const int MK = a.M * a.K;
  int b, mk;
  xcd_decode(blockIdx.x, MK, b, mk);
  if (b >= a.B) return;
  const int k = mk % a.K;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  GSlab<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  const float* trow = TILT ? tilt_row(a, seq_index(a.seq, a.b0 + b, a.P)) : nullptr;
  const float tix = trow ? trow[3] : 0.f, tiy = trow ? trow[2] : 0.f;
  const float2* wsp = (TILT && a.tg_in) ? a.Wsp + (size_t(b) * MK + mk) * a.S * N * N : nullptr;
  double tg[4] = {0.0, 0.0, 0.0, 0.0};  // tilt gradient: inside y, x; after y, x
  float2 v[G::R1];
  // a later segment left the wave of slice s_hi - 1 (read in the first row phase)
  const bool from_state = a.s_hi < a.S;
  const RView<N, float2, G::NT> GS(a.gstate ? a.gstate + (size_t(b) * MK + mk) * N * N : a.Psi, tid);
  // as in the forward: columns zero in the propagators, and in the band mask
  // (where the far field, and so its gradient, is zero)
  const unsigned zh = gt_cols_zero<N>(a.zcol, L);
  const unsigned zrow = gt_row_zero_groups<N>(a.zcol, L);
  if (!from_state) {
    // far-field adjoint: inverse column transforms of coef * Psi / N^2
    const RView<N, float2, G::NT> psi(a.Psi + (size_t(b) * MK + mk) * N * N, tid);
    const RView<N, float, G::NT> c(a.coef + size_t(b) * N * N, tid);
    const float sc = 1.0f / float(N * N);
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      if ((zh >> r) & 1u) {
        gt_zero_col<N>(v, sl, L, r);
        continue;
      }
#pragma unroll
      for (int i = 0; i < G::R1; ++i) v[i] = cscale(psi[r * G::R1 + i], c[r * G::R1 + i] * sc);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
      gt_store_col<N>(v, sl, L, r);
    }
    sl.to_rows(X, L);
    if ((TILT && a.tilt_after)) {
      // the gradient through the after-tilt ramp and of its tilt
      // (objective.py:620-636), band-limited for one slice
      tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
      __syncthreads();
      const float tax = trow ? trow[5] : 0.f, tay = trow ? trow[4] : 0.f;
      const RView<N, float2, G::NT> XE(a.Xe ? a.Xe + (size_t(b) * MK + mk) * N * N : a.Psi, tid);
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        gt_load_rowspec<N>(v, sl, L, r);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
        const int y = r * G::LINES + L.ln;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) {
          const int x = L.j + 8 * n;
          v[n] = cmulc(v[n], tilt_ramp(tt, y, x, tax, tay));
          if ((TILT && a.tg_after)) {
            const float2 e = XE[r * G::R1 + n];
            const double zi = double(v[n].y) * double(e.x) - double(v[n].x) * double(e.y);  // Im(g conj(e))
            tg[2] += 2.0 * double(tt[y].y) * zi;
            tg[3] += 2.0 * double(tt[x].x) * zi;
          }
          // the slice loop's row inverse follows this row pair (N times the
          // wave) unless the band limit's 1/N^2 comes between
          if (a.S != 1) v[n] = cscale(v[n], 1.f / float(N));
        }
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        gt_store_rowspec<N>(v, sl, L, r);
      }
      if (a.S == 1) {
        sl.to_cols(X, L);
        gt_band_columns<N>(v, X, sl, L, a.maskc, a.zcol);
        sl.to_rows(X, L);
      } else {
        __syncthreads();  // the row spectra stay in rows; tt may be rewritten below
      }
    }
  }
  if ((TILT && a.tilt_win) && !a.dists) {
    tilt_tables_inside<N>(tt, a.dz, tix, tiy, a.px, a.py);
    __syncthreads();
  }
  // the stash slot of slice s is overwritten by this mode's window gradient
  // conj(u_s) * g; k_objgrad_gt sums the modes
  // pu holds the stash registers of the current round, read one round ahead
  float2* st = a.stash + (size_t(b) * MK + mk) * a.st_slots * N * N;
  const float2* ob = a.Ob + (size_t(b) * a.S * a.K + k) * N * N;
  // at 192 px (768 threads, 168 registers) reading ahead costs spills
  constexpr bool kAhead = N != 192;
  float2 pu[G::R1];
  if (kAhead && a.step_obj) {
    const RView<N, float2, G::NT> U(st + size_t(a.s_hi - 1 - a.s_lo) * N * N, tid);
#pragma unroll
    for (int n = 0; n < G::R1; ++n) pu[n] = U[n];
  }
  for (int s = a.s_hi - 1; s >= a.s_lo; --s) {
    const RView<N, float2, G::NT> U(st + size_t(s - a.s_lo) * N * N, tid), O(ob + size_t(s) * a.K * N * N, tid);
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      if (from_state && s == a.s_hi - 1) {
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = GS[r * G::R1 + n];
      } else {
        // after the far-field adjoint (and an after-tilt) every column was
        // stored; after a propagation the zero columns were not
        if (s == a.s_hi - 1)
          gt_load_rowspec<N>(v, sl, L, r);
        else
          sl.load_rowspec_z(v, L, r, zrow);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
      }
      if (!kAhead && a.step_obj) {
#pragma unroll
        for (int n = 0; n < G::R1; ++n) pu[n] = U[r * G::R1 + n];
      }
#pragma unroll
      for (int n = 0; n < G::R1; ++n) {
        const int q = r * G::R1 + n;
        if (a.step_obj) U.store(q, cmulc(v[n], pu[n]));
        v[n] = cmulc(v[n], O[q]);
      }
      const int rn = (r + 1 < G::ROUNDS) ? r + 1 : 0, sn = (r + 1 < G::ROUNDS) ? s : s - 1;
      if (kAhead && a.step_obj && sn >= a.s_lo) {
        const RView<N, float2, G::NT> Un(st + size_t(sn - a.s_lo) * N * N, tid);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) pu[n] = Un[rn * G::R1 + n];
      }
      if (s > 0) {
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        sl.store_rowspec_z(v, L, r, zrow);  // the propagation skips the zero groups
      } else {
        float2* out = a.gP + (size_t(b) * MK + mk) * N * N + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) out[8 * n] = v[n];
      }
    }
    if (s > 0) {
      sl.to_cols(X, L);
      if ((TILT && a.tilt_win) && a.dists) {
        __syncthreads();
        tilt_tables_inside<N>(tt, a.dists[s - 1], tix, tiy, a.px, a.py);
        __syncthreads();
      }
      const RView<N, float2, G::NT> H((s == 1 && !(TILT && a.tilt_win)) ? a.H0 : a.H + size_t(s - 1) * a.Hstride, tid);
      double sy = 0.0, sx = 0.0;
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        if ((zh >> r) & 1u) continue;  // zero columns: neither computed nor stored
        // the propagator is read before the column transform, except at
        // 192 px (168 registers) and where the tilt sums come first
        constexpr bool kHEarly = N != 192;
        float2 h[G::R1];
        if (kHEarly && !wsp) {
#pragma unroll
          for (int i = 0; i < G::R1; ++i) h[i] = H[r * G::R1 + i];
        }
        gt_load_col<N>(v, sl, L, r);
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        if (!kHEarly && !wsp) {
#pragma unroll
          for (int i = 0; i < G::R1; ++i) h[i] = H[r * G::R1 + i];
        }
        if (wsp) {
          // sum Im(conj(fft2 u_s) fft2(g_s) q) with Wsp = fft2(u_s) / N^2
          // (multislice_core.py:169-178)
          const RView<N, float2, G::NT> W(wsp + size_t(s) * N * N, tid);
#pragma unroll
          for (int i = 0; i < G::R1; ++i) {
            int fy, fx;
            G::cspec((r * G::R1 + i) * G::NT + tid, fy, fx);
            const float2 w = W[r * G::R1 + i];
            const double zi = double(w.x) * double(v[i].y) - double(w.y) * double(v[i].x);
            sx += double(fftfreq_f(fx, N, a.px)) * zi;
            sy += double(fftfreq_f(fy, N, a.py)) * zi;
          }
#pragma unroll
          for (int i = 0; i < G::R1; ++i) h[i] = H[r * G::R1 + i];
        }
        if ((TILT && a.tilt_win)) {
#pragma unroll
          for (int i = 0; i < G::R1; ++i) {
            int fy, fx;
            G::cspec((r * G::R1 + i) * G::NT + tid, fy, fx);
            h[i] = tilted_h(h[i], tt, fy, fx, N, s == 1, a.edge);
          }
        }
#pragma unroll
        for (int i = 0; i < G::R1; ++i) v[i] = cmulc(v[i], h[i]);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
        gt_store_col<N>(v, sl, L, r);
      }
      if (wsp) {
        double d = double(slice_dist(a, s - 1));
        if (s == 1 && a.edge > 1.f) d *= double(a.edge);
        tg[0] += -12.566370614359172 * d * sy;
        tg[1] += -12.566370614359172 * d * sx;
      }
      sl.to_rows(X, L);
    }
  }
  if (a.s_lo > 0) {
    // the wave propagated to slice s_lo - 1, as natural rows, for the next segment
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      sl.load_rowspec_z(v, L, r, zrow);
      rt_line_inv<N, true>(v, X, L.ln, L.j);
#pragma unroll
      for (int n = 0; n < G::R1; ++n) GS.store(r * G::R1 + n, v[n]);
    }
  }
  if ((TILT && a.tg_in) || (TILT && a.tg_after)) {
    // this segment's part; the segments of a window add in order
    for (int c = 0; c < 4; ++c) {
      const double S = block_sum(tg[c], red);
      if (tid == 0) {
        double* dst = a.tilt_part + (size_t(b) * MK + mk) * 4 + c;
        *dst = (a.s_hi < a.S ? *dst : 0.0) + S;
      }
    }
  }
}

// k_objgrad for 192/256 px windows: mode sum from the R-order stash, band
// limit, natural-order window gradient for k_gather_grad
template <int N>
__global__ __launch_bounds__(kGtThreads<N>) void k_objgrad_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  const int planes = a.S * a.K;
  int b, plane, sl_, k;
  objgrad_plane(a, b, plane, sl_, k);
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  SlabOf<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  const size_t mstride = size_t(a.K) * a.st_slots * N * N;
  const float2* src = a.stash + ((size_t(b) * a.M * a.K + k) * a.st_slots + (sl_ - a.s_lo)) * N * N;
  float2 v[G::R1];
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    {
      const RView<N, float2, G::NT> G0(src, tid);
#pragma unroll
      for (int n = 0; n < G::R1; ++n) v[n] = G0[r * G::R1 + n];
    }
    for (int m = 1; m < a.M; ++m) {
      const RView<N, float2, G::NT> Gm(src + m * mstride, tid);
#pragma unroll
      for (int n = 0; n < G::R1; ++n) v[n] = cadd(v[n], Gm[r * G::R1 + n]);
    }
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
    gt_store_rowspec<N>(v, sl, L, r);
  }
  sl.to_cols(X, L);
  gt_band_columns<N>(v, X, sl, L, a.maskc, a.zcol);
  sl.to_rows(X, L);
  float2* Gout = (a.scatter ? a.Ob : a.Gbuf) + (size_t(b) * planes + plane) * N * N;
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    gt_load_rowspec<N>(v, sl, L, r);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    float2* out = Gout + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
    for (int n = 0; n < G::R1; ++n) out[8 * n] = v[n];
  }
}

// ------------------------------- slab-tile probe side (192, 256 px)
// As on the register tile: probe spectra and the spectral probe-gradient
// partials in slab C-order, the shift factor from two axis tables in LDS.
// Without a shift the partials hold the natural-order gradient in slab
// R-order.  Each workgroup transposes through its slab in `scratch`.

// probe spectra, slab C-order, unnormalised
template <int N>
__global__ __launch_bounds__(kGtThreads<N>) void k_fprobe_gt(const float2* __restrict__ probe,
                                                          float2* __restrict__ fprobe, float2* __restrict__ scratch) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  const GTLane<N> L;
  SlabOf<N> sl(scratch + size_t(blockIdx.x) * N * N);
  const float2* src = probe + size_t(blockIdx.x) * N * N;
  float2 v[G::R1];
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    const float2* p = src + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
    for (int n = 0; n < G::R1; ++n) v[n] = p[8 * n];
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
    gt_store_rowspec<N>(v, sl, L, r);
  }
  sl.to_cols(X, L);
  const RView<N, float2, G::NT> out(fprobe + size_t(blockIdx.x) * N * N, threadIdx.x);
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    gt_load_col<N>(v, sl, L, r);
    rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
    for (int i = 0; i < G::R1; ++i) out.store(r * G::R1 + i, v[i]);
  }
}

// shifted probes (objective.py:471-476): ifft2(fprobe * E), natural order
template <int N, bool AB>
__global__ __launch_bounds__(kGtThreads<N>) void k_probe_shift_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  __shared__ float2 exy[2 * N];
  const int b = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  const bool mv = a.moved && idx >= 0;
  const float dy = mv ? a.corr[2 * idx] : 0.f, dx = mv ? a.corr[2 * idx + 1] : 0.f;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  SlabOf<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  shift_tables_freq<N>(exy, dy, dx);
  __syncthreads();
  // with a before-tilt each window brings its own spectrum, the current in it
  const RView<N, float2, G::NT> fp(a.Fw ? a.Fw + (size_t(b) * a.M + m) * N * N : a.fprobe + size_t(m) * N * N,
                                   tid);
  const RView<N, float, G::NT> mask(a.maskc, tid);
  const float s = (a.Fw ? 1.f : bc_scale(a, idx)) / float(N * N);
  float2 v[G::R1];
  // E, and so the spectrum, is zero on the columns the band mask zeroes
  const unsigned zm = gt_cols_zero<N>(a.zcol, L);
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    if ((zm >> r) & 1u) {
      gt_zero_col<N>(v, sl, L, r);
      continue;
    }
#pragma unroll
    for (int i = 0; i < G::R1; ++i) {
      const int q = r * G::R1 + i;
      int fy, fx;
      G::cspec(q * G::NT + tid, fy, fx);
      const float2 E = cscale(cmul(exy[fx], exy[N + fy]), mask[q]);
      v[i] = cscale(cmul(fp[q], E), s);
      if constexpr (AB)
        if (!a.moved) v[i] = cmul(v[i], a.plate[size_t(b) * N * N + q * G::NT + tid]);
    }
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    gt_store_col<N>(v, sl, L, r);
  }
  sl.to_rows(X, L);
  float2* dst = a.Pp + (size_t(b) * a.M + m) * N * N;
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    gt_load_rowspec<N>(v, sl, L, r);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    float2* out = dst + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
    for (int n = 0; n < G::R1; ++n) out[8 * n] = v[n];
  }
}

// probe-side adjoint (objective.py:641-658, 697-703) for 192/256 px windows:
// the gradient summed over object modes (natural order from k_adjoint_gt) is
// transformed, shifted back and added (C-order) to the group's partial, with
// the position gradient -2/n^2 sum Re(K conj(FP) Q) in FP64; without a shift
// the natural-order gradient itself is added (slab R-order)
template <int N, bool AB>
__global__ __launch_bounds__(kGtThreads<N>) void k_probegrad_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  __shared__ float2 exy[2 * N];
  __shared__ double red[16];
  const int g = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  GSlab<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  const RView<N, float2, G::NT> acc(a.probe_part + (size_t(g) * a.M + m) * N * N, tid);
  const RView<N, float2, G::NT> fp(a.fprobe + size_t(m) * N * N, tid);
  const RView<N, float, G::NT> mask(a.maskc, tid);
  float2 v[G::R1];
  // Q = conj(E) * spectrum is zero on the columns the band mask zeroes
  const unsigned zm = gt_cols_zero<N>(a.zcol, L);
  for (int b = g; b < a.B; b += a.G) {
    const int idx = seq_index(a.seq, a.b0 + b, a.P);
    const float bcs = bc_scale(a, idx);
    const float2* gp = a.gP + (size_t(b) * a.M + m) * a.K * N * N;
    if (!a.shift) {
      double sb = 0.0;
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        const size_t row = size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = gp[row + 8 * n];
        for (int k = 1; k < a.K; ++k)
#pragma unroll
          for (int n = 0; n < G::R1; ++n) v[n] = cadd(v[n], gp[size_t(k) * N * N + row + 8 * n]);
#pragma unroll
        for (int n = 0; n < G::R1; ++n) {
          acc.store(r * G::R1 + n, cadd(acc[r * G::R1 + n], cscale(v[n], bcs)));
          if (a.bc_part) {
            const float2 Pn = a.probe[size_t(m) * N * N + row + 8 * n];
            sb += double(v[n].x) * double(Pn.x) + double(v[n].y) * double(Pn.y);
          }
        }
      }
      if (a.bc_part) {
        const double SB = block_sum(sb, red);
        if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) * SB;
      }
      continue;
    }
    const bool mv = a.moved && idx >= 0;
    const float dy = mv ? a.corr[2 * idx] : 0.f, dx = mv ? a.corr[2 * idx + 1] : 0.f;
    float* wrow = AB && a.ab_w ? a.ab_w + (size_t(b) * a.M + m) * N * N : nullptr;
    __syncthreads();  // the previous window's slab and table reads are done
    shift_tables_freq<N>(exy, dy, dx);
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      const size_t row = size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
      for (int n = 0; n < G::R1; ++n) v[n] = gp[row + 8 * n];
      for (int k = 1; k < a.K; ++k)
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = cadd(v[n], gp[size_t(k) * N * N + row + 8 * n]);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
      gt_store_rowspec<N>(v, sl, L, r);
    }
    sl.to_cols(X, L);  // the slab holds the row spectra; the table is published
    double sy = 0.0, sx = 0.0, sb = 0.0;
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      if ((zm >> r) & 1u) {
        if constexpr (AB)
          if (wrow)
#pragma unroll
            for (int i = 0; i < G::R1; ++i) wrow[(r * G::R1 + i) * G::NT + tid] = 0.f;
        continue;
      }
      gt_load_col<N>(v, sl, L, r);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
      for (int i = 0; i < G::R1; ++i) {
        const int q = r * G::R1 + i;
        int fy, fx;
        G::cspec(q * G::NT + tid, fy, fx);
        const float mk = mask[q];
        const float2 E = cscale(cmul(exy[fx], exy[N + fy]), mk);
        const float2 Q = cmulc(v[i], E);
        // after the conjugate aberration plate (objective.py:673-675); the mask is 0 or 1
        float2 Fg = Q;
        if constexpr (AB) Fg = cmulc(Q, a.plate[size_t(b) * N * N + q * G::NT + tid]);
        acc.store(q, cadd(acc[q], cscale(Fg, bcs)));
        if (a.step_pos || a.bc_part || (AB && wrow)) {
          const float2 F = fp[q];
          // Re(G * conj(FP)): the beam current's sum, Parseval's form of the real-space one
          sb += double(Fg.x) * double(F.x) + double(Fg.y) * double(F.y);
          if constexpr (AB)
            if (wrow)
              wrow[q * G::NT + tid] = float(double(bcs) * (double(Fg.y) * double(F.x) - double(Fg.x) * double(F.y)));
          // Im(conj(FP) * Q) in FP64: the terms of the sum nearly cancel
          const double zy = double(Q.y) * double(F.x) - double(Q.x) * double(F.y);
          const float kx = fmul(fmul(-6.2831855f, fftfreq_f(fx, N, 1.f)), mk);
          const float ky = fmul(fmul(-6.2831855f, fftfreq_f(fy, N, 1.f)), mk);
          sx += -double(kx) * zy;
          sy += -double(ky) * zy;
        }
      }
    }
    if (a.step_pos) {
      const double SY = block_sum(sy, red);
      const double SX = block_sum(sx, red);
      if (tid == 0) {
        // the shifted probe's spectrum is |bc| FP
        const double sh = -2.0 * double(bcs) / (double(N) * double(N));
        a.pos_part[(size_t(b) * a.M + m) * 2] = sh * SY;
        a.pos_part[(size_t(b) * a.M + m) * 2 + 1] = sh * SX;
      }
    }
    if (a.bc_part) {
      const double SB = block_sum(sb, red);
      if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) / (double(N) * double(N)) * SB;
    }
  }
}

// k_probegrad_tilt_rt for 192/256 px windows: the shift's transforms through
// the workgroup's slab, the natural-order partial in slab R-order
template <int N>
__global__ __launch_bounds__(kGtThreads<N>) void k_probegrad_tilt_gt(BatchArgs a) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  __shared__ float2 exy[2 * N];
  __shared__ float2 tt[N];
  __shared__ double red[16];
  const int g = blockIdx.x / a.M, m = blockIdx.x % a.M;
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  GSlab<N> sl(a.scratch + size_t(blockIdx.x) * N * N);
  tilt_tables_grid<N>(tt, a.px, a.py, a.lambda);
  const RView<N, float2, G::NT> acc(a.probe_part + (size_t(g) * a.M + m) * N * N, tid);
  const RView<N, float, G::NT> mask(a.maskc, tid);
  const float sc = 1.0f / float(N * N);
  float2 v[G::R1];
  __syncthreads();
  for (int b = g; b < a.B; b += a.G) {
    const int idx = seq_index(a.seq, a.b0 + b, a.P);
    const float2* gp = a.gP + (size_t(b) * a.M + m) * a.K * N * N;
    if (a.moved) {
      const float dy = idx >= 0 ? a.corr[2 * idx] : 0.f, dx = idx >= 0 ? a.corr[2 * idx + 1] : 0.f;
      __syncthreads();  // the previous window's slab and table reads are done
      shift_tables_freq<N>(exy, dy, dx);
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        const size_t row = size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = gp[row + 8 * n];
        for (int k = 1; k < a.K; ++k)
#pragma unroll
          for (int n = 0; n < G::R1; ++n) v[n] = cadd(v[n], gp[size_t(k) * N * N + row + 8 * n]);
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
        gt_store_rowspec<N>(v, sl, L, r);
      }
      sl.to_cols(X, L);  // the slab holds the row spectra; the table is published
      const RView<N, float2, G::NT> F(a.Fw + (size_t(b) * a.M + m) * N * N, tid);
      double sy = 0.0, sx = 0.0;
      for (int r = 0; r < G::ROUNDS; ++r) {
        round_fence<decltype(sl)>();
        gt_load_col<N>(v, sl, L, r);
        rt_line_fwd<N, true>(v, X, L.ln, L.j);
#pragma unroll
        for (int i = 0; i < G::R1; ++i) {
          const int q = r * G::R1 + i;
          int fy, fx;
          G::cspec(q * G::NT + tid, fy, fx);
          const float mk = mask[q];
          const float2 Q = cmulc(v[i], cscale(cmul(exy[fx], exy[N + fy]), mk));
          if (a.step_pos) {
            const float2 f = F[q];
            const double zy = double(Q.y) * double(f.x) - double(Q.x) * double(f.y);
            sx += -double(fmul(fmul(-6.2831855f, fftfreq_f(fx, N, 1.f)), mk)) * zy;
            sy += -double(fmul(fmul(-6.2831855f, fftfreq_f(fy, N, 1.f)), mk)) * zy;
          }
          v[i] = cscale(Q, sc);
        }
        rt_line_inv<N, true>(v, X, L.ln, L.j);
        gt_store_col<N>(v, sl, L, r);
      }
      if (a.step_pos) {
        const double SY = block_sum(sy, red);
        const double SX = block_sum(sx, red);
        if (tid == 0) {
          const double sh = -2.0 / (double(N) * double(N));
          a.pos_part[(size_t(b) * a.M + m) * 2] = sh * SY;
          a.pos_part[(size_t(b) * a.M + m) * 2 + 1] = sh * SX;
        }
      }
      sl.to_rows(X, L);
    }
    const float* tr = tilt_row(a, idx);
    const float ty = tr ? tr[0] : 0.f, tx = tr ? tr[1] : 0.f;
    const float bcs = bc_scale(a, idx);
    double ty_s = 0.0, tx_s = 0.0, sb = 0.0;
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
      const size_t row = size_t(r * G::LINES + L.ln) * N + L.j;
      if (a.moved) {
        gt_load_rowspec<N>(v, sl, L, r);
        rt_line_inv<N, true>(v, X, L.ln, L.j);
      } else {
#pragma unroll
        for (int n = 0; n < G::R1; ++n) v[n] = gp[row + 8 * n];
        for (int k = 1; k < a.K; ++k)
#pragma unroll
          for (int n = 0; n < G::R1; ++n) v[n] = cadd(v[n], gp[size_t(k) * N * N + row + 8 * n]);
      }
      const int y = r * G::LINES + L.ln;
#pragma unroll
      for (int n = 0; n < G::R1; ++n) {
        const int x = L.j + 8 * n;
        v[n] = cmulc(v[n], tilt_ramp(tt, y, x, tx, ty));
        const float2 P = a.probe[size_t(m) * N * N + row + 8 * n];
        const double zi = double(v[n].y) * double(P.x) - double(v[n].x) * double(P.y);  // Im(g conj(P))
        ty_s += double(tt[y].y) * zi;
        tx_s += double(tt[x].x) * zi;
        sb += double(v[n].x) * double(P.x) + double(v[n].y) * double(P.y);
        acc.store(r * G::R1 + n, cadd(acc[r * G::R1 + n], cscale(v[n], bcs)));
      }
    }
    if (a.tg_before) {
      const double SY = block_sum(ty_s, red);
      const double SX = block_sum(tx_s, red);
      if (tid == 0) {
        a.tb_part[(size_t(b) * a.M + m) * 2] = 2.0 * double(bcs) * SY;
        a.tb_part[(size_t(b) * a.M + m) * 2 + 1] = 2.0 * double(bcs) * SX;
      }
    }
    if (a.bc_part) {
      const double SB = block_sum(sb, red);
      if (tid == 0) a.bc_part[size_t(b) * a.M + m] = bc_factor(a, idx) * SB;
    }
  }
}

// probe_grad = ifft2(mask * fft2(sum)) (objective.py:927-933) from the reduced
// partials: slab C-order spectrum (spectral) or slab R-order natural gradient
template <int N>
__global__ __launch_bounds__(kGtThreads<N>) void k_probe_final_gt(const float2* __restrict__ sum, int spectral,
                                                               const float* __restrict__ maskc,
                                                               float2* __restrict__ probe_grad,
                                                               float2* __restrict__ scratch) {
  using G = GT<N, kGtThreads<N>>;
  rt_tw_load<N>();
  __shared__ float2 X[G::LINES * G::LS];
  const unsigned tid = threadIdx.x;
  const GTLane<N> L;
  SlabOf<N> sl(scratch + size_t(blockIdx.x) * N * N);
  const RView<N, float2, G::NT> src(sum + size_t(blockIdx.x) * N * N, tid);
  const RView<N, float, G::NT> mask(maskc, tid);
  const float s = 1.0f / float(N * N);
  float2 v[G::R1];
  if (!spectral) {
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
      round_fence<decltype(sl)>();
#pragma unroll
      for (int n = 0; n < G::R1; ++n) v[n] = src[r * G::R1 + n];
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
      gt_store_rowspec<N>(v, sl, L, r);
    }
    sl.to_cols(X, L);
  }
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    if (spectral) {
#pragma unroll
      for (int i = 0; i < G::R1; ++i) v[i] = src[r * G::R1 + i];
    } else {
      gt_load_col<N>(v, sl, L, r);
      rt_line_fwd<N, true>(v, X, L.ln, L.j);
    }
#pragma unroll
    for (int i = 0; i < G::R1; ++i) v[i] = cscale(v[i], mask[r * G::R1 + i] * s);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    gt_store_col<N>(v, sl, L, r);
  }
  sl.to_rows(X, L);
  float2* dst = probe_grad + size_t(blockIdx.x) * N * N;
#pragma unroll
  for (int r = 0; r < G::ROUNDS; ++r) {
    round_fence<decltype(sl)>();
    gt_load_rowspec<N>(v, sl, L, r);
    rt_line_inv<N, true>(v, X, L.ln, L.j);
    float2* out = dst + size_t(r * G::LINES + L.ln) * N + L.j;
#pragma unroll
    for (int n = 0; n < G::R1; ++n) out[8 * n] = v[n];
  }
}

// Deterministic scatter-add of a batch's window gradients (multislice_core.py:192,
// 206-230): one workgroup per GTILE x GTILE object tile of one plane sums, in window
// order, every window of the batch that overlaps the tile, then adds the sum
// to the object gradient.  Tiles are disjoint, so no atomics are needed and
// the result does not depend on scheduling.
constexpr int GTILE = 32;

template <int N>
__global__ __launch_bounds__(256) void k_gather_grad(BatchArgs a) {
  __shared__ int wy0[256], wx0[256], wb[256];
  const int tiles_x = (a.nx + GTILE - 1) / GTILE, tiles_y = (a.ny + GTILE - 1) / GTILE;
  const int planes = a.S * a.K;
  const int plane = blockIdx.x / (tiles_x * tiles_y);
  const int tile = blockIdx.x % (tiles_x * tiles_y);
  const int ty0 = (tile / tiles_x) * GTILE, tx0 = (tile % tiles_x) * GTILE;
  const int cx = threadIdx.x % GTILE, ry = threadIdx.x / GTILE;  // 8 row groups of 4 rows
  const int x = tx0 + cx;
  float2 acc[GTILE / 8];
#pragma unroll
  for (int r = 0; r < GTILE / 8; ++r) acc[r] = make_float2(0.f, 0.f);
  bool any = false;
  // windows in chunks of 256: each thread tests one window, overlapping ones are
  // compacted into LDS in window order, then all threads sum them
  for (int c0 = 0; c0 < a.B; c0 += 256) {
    const int b = c0 + threadIdx.x;
    bool hit = false;
    int y0 = 0, x0 = 0;
    if (b < a.B) {
      const int idx = seq_index(a.seq, a.b0 + b, a.P);
      if (idx >= 0) {
        y0 = a.pos[2 * idx];
        x0 = a.pos[2 * idx + 1];
        hit = window_inside(y0, x0, N, a.ny, a.nx) && !(y0 >= ty0 + GTILE || y0 + N <= ty0 || x0 >= tx0 + GTILE ||
                                                        x0 + N <= tx0);
      }
    }
    // order-preserving compaction: ballot per wavefront, prefix over the 4 wavefronts
    __shared__ int wcount[4];
    const unsigned long long m = __ballot(hit);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wid] = __popcll(m);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wid; ++w) off += wcount[w];
    const int n = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (hit) {
      wy0[off + before] = y0;
      wx0[off + before] = x0;
      wb[off + before] = b;
    }
    __syncthreads();
    any = any || n > 0;
    // eight windows at a time: all their loads are issued before the sums, which
    // are still taken in window order (a pixel a window does not cover adds 0)
    constexpr int H = 8, RR = GTILE / 8;
    int i = 0;
    for (; i + H <= n; i += H) {
      float2 t[H][RR];
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const int y0s = wy0[i + h], wx = x - wx0[i + h];
        const float2* G = a.Ob + (size_t(wb[i + h]) * planes + plane) * N * N;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
          const int wy = ty0 + ry + 8 * r - y0s;
          t[h][r] = (wx >= 0 && wx < N && wy >= 0 && wy < N) ? G[wy * N + wx] : make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int r = 0; r < RR; ++r) acc[r] = cadd(acc[r], t[h][r]);
    }
    for (; i < n; ++i) {
      const int y0s = wy0[i], x0s = wx0[i];
      const float2* G = a.Ob + (size_t(wb[i]) * planes + plane) * N * N;
      const int wx = x - x0s;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        const int wy = ty0 + ry + 8 * r - y0s;
        if (wx >= 0 && wx < N && wy >= 0 && wy < N) acc[r] = cadd(acc[r], G[wy * N + wx]);
      }
    }
    __syncthreads();
  }
  if (!any || x >= a.nx) return;
  float2* dst = a.obj_grad + size_t(plane) * a.ny * a.nx;
#pragma unroll
  for (int r = 0; r < GTILE / 8; ++r) {
    const int y = ty0 + ry + 8 * r;
    if (y < a.ny) {
      float2* d = &dst[size_t(y) * a.nx + x];
      *d = cadd(*d, acc[r]);
    }
  }
}

// pos_grad[idx] = sum over probe modes (assigned per window, objective.py:658)
__global__ void k_pos_final(BatchArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  if (idx < 0) return;
  double y = 0.0, x = 0.0;
  for (int m = 0; m < a.M; ++m) {
    y += a.pos_part[(size_t(b) * a.M + m) * 2];
    x += a.pos_part[(size_t(b) * a.M + m) * 2 + 1];
  }
  a.pos_grad[2 * idx] = float(y);
  a.pos_grad[2 * idx + 1] = float(x);
}

// beam_current_grad[idx] = sum over probe modes (assigned per window, objective.py:694-696)
__global__ void k_bc_final(BatchArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  if (idx < 0) return;
  double s = 0.0;
  for (int m = 0; m < a.M; ++m) s += a.bc_part[size_t(b) * a.M + m];
  a.bc_grad[idx] = float(s);
}

// tilts_grad (multislice_core.py:176-177, objective.py:618-632): per window the
// sum over its probe and object modes of the inside (columns 2, 3) and after
// (4, 5) parts, into the window's row; with one tilt row every window of the
// batch adds into row 0 in batch order (one thread)
// the window's parts in tilt-column order: before (from tb_part, per probe
// mode), inside and after (from tilt_part, per probe and object mode)
__device__ __forceinline__ void tilt_window_sums(const BatchArgs& a, int b, double* acc) {
  const int MK = a.M * a.K;
  for (int c = 0; c < 6; ++c) acc[c] = 0.0;
  if (a.tg_before)
    for (int m = 0; m < a.M; ++m)
      for (int c = 0; c < 2; ++c) acc[c] += a.tb_part[(size_t(b) * a.M + m) * 2 + c];
  if (a.tg_in || a.tg_after)
    for (int mk = 0; mk < MK; ++mk)
      for (int c = 0; c < 4; ++c) acc[2 + c] += a.tilt_part[(size_t(b) * MK + mk) * 4 + c];
}
__device__ __forceinline__ bool tilt_col_on(const BatchArgs& a, int c) {
  return c < 2 ? a.tg_before != 0 : (c < 4 ? a.tg_in != 0 : a.tg_after != 0);
}
__global__ void k_tilt_final(BatchArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6];
  if (a.n_tilt == 1) {
    if (b != 0) return;
    double tot[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < a.B; ++w) {
      if (seq_index(a.seq, a.b0 + w, a.P) < 0) continue;
      tilt_window_sums(a, w, acc);
      for (int c = 0; c < 6; ++c) tot[c] += acc[c];
    }
    for (int c = 0; c < 6; ++c)
      if (tilt_col_on(a, c)) a.tilt_grad[c] = float(double(a.tilt_grad[c]) + tot[c]);
    return;
  }
  if (b >= a.B) return;
  const int idx = seq_index(a.seq, a.b0 + b, a.P);
  if (idx < 0) return;
  tilt_window_sums(a, b, acc);
  for (int c = 0; c < 6; ++c)
    if (tilt_col_on(a, c)) a.tilt_grad[size_t(idx) * 6 + c] = float(acc[c]);
}

// each window's phase plate exp(-i chi), chi = sum_j A_j c_j with c its
// position's row (objective.py:324; the band mask comes with the shift
// factor), 1 for a window without a row; one thread per element
template <int N>
__global__ __launch_bounds__(256) void k_aber_plate(BatchArgs a) {
  const size_t t = blockIdx.x * size_t(256) + threadIdx.x;
  if (t >= size_t(a.B) * N * N) return;
  const int b = int(t / (N * N)), e = int(t % (N * N));
  const float* c = ab_row(a, seq_index(a.seq, a.b0 + b, a.P));
  float chi = 0.f;
  if (c)
    for (int j = 0; j < a.NA; ++j) chi = fmaf(a.abt[size_t(j) * N * N + e], c[j], chi);
  float sn, cs;
  sincosf(chi, &sn, &cs);
  a.plate[t] = make_float2(cs, -sn);
}

// per window and coefficient: -2/n^2 sum over modes and frequencies of A_j
// times |bc| Im(G conj(F)), the window's term of objective.py:676-679 in FP64
// (Re(Z conj(-i A)) = -A Im(Z))
template <int N>
__global__ __launch_bounds__(256) void k_aber_part(BatchArgs a) {
  __shared__ double red[4];
  const int b = blockIdx.x / a.NA, j = blockIdx.x % a.NA;
  const float* A = a.abt + size_t(j) * N * N;
  double s = 0.0;
  for (int m = 0; m < a.M; ++m) {
    const float* w = a.ab_w + (size_t(b) * a.M + m) * N * N;
    for (int e = threadIdx.x; e < N * N; e += 256) s += double(A[e]) * double(w[e]);
  }
  const double S = block_sum(s, red);
  if (threadIdx.x == 0) a.ab_part[size_t(b) * a.NA + j] = -2.0 / (double(N) * double(N)) * S;
}

// aberrations_grad[r][j] += the batch's windows of row r in batch order
// (objective.py:680, scatteradd_abers), one thread per (row, coefficient)
__global__ void k_aber_final(BatchArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_markers * a.NA) return;
  const int r = t / a.NA, j = t % a.NA;
  double s = 0.0;
  bool any = false;
  for (int b = 0; b < a.B; ++b) {
    const int idx = seq_index(a.seq, a.b0 + b, a.P);
    if (idx < 0 || a.marker[idx] != r) continue;
    s += a.ab_part[size_t(b) * a.NA + j];
    any = true;
  }
  if (any) a.ab_grad[t] = float(double(a.ab_grad[t]) + s);
}

// sum over the G group partials of the probe gradient, one thread per
// element (in group order, so the sum is deterministic)
__global__ __launch_bounds__(256) void k_probe_reduce(const float2* __restrict__ part, int G, size_t per_group,
                                                     float2* __restrict__ sum) {
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (e >= per_group) return;
  float2 v = part[e];
  for (int g = 1; g < G; ++g) v = cadd(v, part[size_t(g) * per_group + e]);
  sum[e] = v;
}

__global__ __launch_bounds__(256) void k_loss_final(const double* __restrict__ part, int n, double w,
                                                   double* __restrict__ out) {
  __shared__ double red[8];
  double l = 0.0, s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    l += part[2 * size_t(i)];
    s += part[2 * size_t(i) + 1];
  }
  const double L = block_sum(l, red);
  const double S = block_sum(s, red);
  if (threadIdx.x == 0) {
    out[0] = L * w;
    out[1] = S;
  }
}

__global__ void k_scatteradd(float2* __restrict__ full, int ny, int nx, int planes, const float2* __restrict__ patches,
                             int n, const int* __restrict__ corners, int layout) {
  // grid: (batch*planes), block 256
  const int b = blockIdx.x / planes, p = blockIdx.x % planes;
  const int y0 = corners[2 * b], x0 = corners[2 * b + 1];
  if (!window_inside(y0, x0, n, ny, nx)) return;
  const float2* src = patches + (size_t(b) * planes + p) * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int y = e / n, x = e % n;
    const size_t pix = size_t(y0 + y) * nx + (x0 + x);
    const size_t at = (layout == PTY_LAYOUT_PLANAR) ? size_t(p) * ny * nx + pix : pix * planes + p;
    atomic_add_c(&full[at], src[e]);
  }
}

// =================================================================== host
namespace {

// ------------------------------------------------------------- timing
struct TimedLaunch {
  int kind;
  hipEvent_t a, b;
};
enum {
  KT_FFT2 = 0, KT_TABLES, KT_LAYOUT, KT_PATCH, KT_FPROBE, KT_PROBE_SHIFT, KT_FORWARD, KT_LOSS, KT_ADJOINT,
  KT_OBJGRAD, KT_PROBEGRAD, KT_POS_FINAL, KT_PROBE_FINAL, KT_LOSS_FINAL, KT_ZERO, KT_SCATTER, KT_GATHER,
  KT_RECOMPUTE, KT_PREPROCESS, KT_BG_GRAD, KT_PROBE_REDUCE, KT_BC_FINAL, KT_AB_TABLE, KT_AB_PLATE, KT_AB_PART,
  KT_AB_FINAL, KT_TILT_FINAL, KT_TILT_PROBE, KT_COUNT
};
const char* kKindNames[KT_COUNT] = {"fft2_tiles", "tables", "layout", "patch", "fprobe", "probe_shift", "forward",
                                    "loss", "adjoint", "objgrad", "probegrad", "pos_final", "probe_final",
                                    "loss_final", "zero", "scatteradd", "gather_grad", "recompute", "preprocess",
                                    "bg_grad", "probe_reduce", "bc_final", "aber_table", "aber_plate",
                                    "aber_part", "aber_final", "tilt_final", "tilt_probe"};
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimedLaunch> g_pending;
double g_total_ms[KT_COUNT];
int64_t g_launches[KT_COUNT];

struct TimeScope {
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  TimeScope(int k, hipStream_t st, bool allowed = true) : kind(k), s(st), on(g_timing && allowed) {
    if (on) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s);
    }
  }
  ~TimeScope() {
    if (on) {
      (void)hipEventRecord(b, s);
      std::lock_guard<std::mutex> lk(g_tmu);
      g_pending.push_back({kind, a, b});
    }
  }
};

inline int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "pypty_amd: HIP launch error: %s\n", hipGetErrorString(e));
    return PTY_ERR_HIP;
  }
  return PTY_OK;
}

#define PTY_DISPATCH_FUSED(n, FN, ...)      \
  switch (n) {                              \
    case 16: return FN<16>(__VA_ARGS__);    \
    case 32: return FN<32>(__VA_ARGS__);    \
    case 48: return FN<48>(__VA_ARGS__);    \
    case 64: return FN<64>(__VA_ARGS__);    \
    case 96: return FN<96>(__VA_ARGS__);    \
    case 128: return FN<128>(__VA_ARGS__);  \
    case 192: return FN<192>(__VA_ARGS__);  \
    case 256: return FN<256>(__VA_ARGS__);  \
    default: return PTY_ERR_SIZE;           \
  }

#define PTY_DISPATCH_N(n, FN, ...)          \
  switch (n) {                              \
    case 16: return FN<16>(__VA_ARGS__);    \
    case 32: return FN<32>(__VA_ARGS__);    \
    case 48: return FN<48>(__VA_ARGS__);    \
    case 64: return FN<64>(__VA_ARGS__);    \
    case 96: return FN<96>(__VA_ARGS__);    \
    case 128: return FN<128>(__VA_ARGS__);  \
    default: return PTY_ERR_SIZE;           \
  }

template <int N>
int launch_fft2(float2* d, int64_t batch, int mode, const float* mask, hipStream_t s) {
  using T = Tile<N>;
  TimeScope ts(KT_FFT2, s);
  const int64_t kMaxGrid = 1LL << 30;
  for (int64_t off = 0; off < batch; off += kMaxGrid) {
    const int g = int(std::min<int64_t>(kMaxGrid, batch - off));
    float2* p = d + off * N * N;
    switch (mode) {
      case PTY_FFT_FWD: k_fft2_tiles<N, MODE_FWD><<<g, T::NT, 0, s>>>(p, mask); break;
      case PTY_FFT_INV: k_fft2_tiles<N, MODE_INV><<<g, T::NT, 0, s>>>(p, mask); break;
      case PTY_FFT_FWD_SHIFT: k_fft2_tiles<N, MODE_FWD_SHIFT><<<g, T::NT, 0, s>>>(p, mask); break;
      case PTY_FFT_INV_ISHIFT: k_fft2_tiles<N, MODE_INV_ISHIFT><<<g, T::NT, 0, s>>>(p, mask); break;
      case MODE_BAND: k_fft2_tiles<N, MODE_BAND><<<g, T::NT, 0, s>>>(p, mask); break;
      default: return PTY_ERR_ARG;
    }
  }
  return check_launch();
}

// Workspace buffers.  Shared buffers first, then one block of per-batch
// buffers per lane: lane l's copy of a per-batch buffer lies lane_bytes * l
// after lane 0's (a buffer of 0 bytes is not needed by the problem).
enum {
  WB_H = 0, WB_H0, WB_MASKC, WB_ZCOL, WB_FPROBE, WB_OBJP, WB_GRADP, WB_PROBE_PART, WB_PROBE_SUM, WB_LOSS_PART,
  WB_ABT, WB_OB, WB_PP, WB_STASH, WB_GSTATE, WB_PSI, WB_COEF, WB_GP, WB_MEAS_STAGE, WB_POS_PART, WB_BC_PART,
  WB_PLATE, WB_AB_W, WB_AB_PART, WB_WSP, WB_XE, WB_TILT_PART, WB_FW, WB_TB_PART, WB_SCRATCH, WB_COUNT
};
constexpr int kFirstLaneBuffer = WB_OB;
const char* kBufNames[WB_COUNT] = {"H", "H0", "maskc", "zcol", "fprobe", "objp", "gradp", "probe_part",
                                   "probe_sum", "loss_part", "abt", "Ob", "Pp", "stash", "gstate", "Psi", "coef",
                                   "gP", "meas_stage", "pos_part", "bc_part", "plate", "ab_w", "ab_part",
                                   "Wsp", "Xe", "tilt_part", "Fw", "tb_part", "scratch"};
struct Layout {
  size_t off[WB_COUNT], bytes[WB_COUNT];  // lane 0's offset for the per-lane buffers
  size_t lane0, lane_bytes, total;
  int lanes;
  int B, G;
  int ck, C;  // slices per checkpoint segment (0: every slice kept) and segments
  int slots;  // stash slots per window and mode
  int scratch_tiles;
};
inline size_t al(size_t x) { return (x + 255) & ~size_t(255); }

int loss_chunks(int n) {
  switch (n) {
    case 16: return LossSplit<16>::CHUNKS;
    case 32: return LossSplit<32>::CHUNKS;
    case 48: return LossSplit<48>::CHUNKS;
    case 64: return LossSplit<64>::CHUNKS;
    case 96: return LossSplit<96>::CHUNKS;
    case 128: return LossSplit<128>::CHUNKS;
    case 192: return LossSplit<192>::CHUNKS;
    case 256: return LossSplit<256>::CHUNKS;
    default: return 1;
  }
}

// windows handled by one launch; windows larger than 128 px keep one tile per
// workgroup in global scratch, so their batch is capped to bound that buffer
// (12 GiB per lane: 512 windows of 256 px with 40 slices fit)
int batch_windows(const pty_problem* p) {
  int B = p->compute_batch;
  if (p->n > 128) {
    const size_t tile = size_t(p->n) * p->n * 8;
    const size_t per_window = size_t(std::max(p->n_slices * p->n_obj_modes, p->n_probe_modes * p->n_obj_modes)) * tile;
    const int cap = int(std::max<size_t>(8, (size_t(12) << 30) / per_window));
    B = std::min(B, cap);
  }
  return B;
}
// the workspace of all lanes stays within this many bytes: a problem whose
// per-lane buffers are larger runs fewer lanes (at least one)
constexpr size_t kLaneBudget = size_t(128) << 30;

// slices per checkpoint segment, or 0 when every slice's wave is kept (the
// stash feeds only the object gradient)
int checkpoint_every(const pty_problem* p) {
  const int c = p->checkpoint_slices;
  return (p->step_obj && c > 0 && c < p->n_slices) ? c : 0;
}

bool compressed(int algorithm) {
  return algorithm >= PTY_LOSS_LSQ_COMPRESSED && algorithm <= PTY_LOSS_POISSON_COMPRESSED;
}
// raw patterns need the per-batch preprocessing pass
bool preprocessing(const pty_problem* p) {
  if (compressed(p->algorithm)) return false;
  return (p->raw_n != 0 && p->raw_n != p->n) || p->data_shift_y != 0 || p->data_shift_x != 0 || p->data_bin > 1 ||
         p->data_pad != 0 || (p->data_multiplier != 0.f && p->data_multiplier != 1.f);
}

// per-slice distances: one propagator table per propagation
bool per_slice(const pty_problem* p) { return p->n_distances > 1; }
int n_prop_tables(const pty_problem* p) { return per_slice(p) ? std::max(1, p->n_slices - 1) : 1; }
// Tilt switches of a pass (objective.py:328-350, 465-499): each window's own
// inside tilt multiplies its propagators when there is a row per position
// and either the distances are per slice (multislice_core.py:71 takes the
// batch's tilts) or the mode tilts inside (0, 3, 4); otherwise the tables
// take row 0's.  The exit wave is tilted in modes 1, 3, 4.
struct TiltPlan {
  bool win, after, before, g_in, g_after, g_before;
};
TiltPlan tilt_plan(const pty_problem* p) {
  TiltPlan t{};
  if (p->n_tilts < 1) return t;
  const int md = p->tilt_mode;
  const bool inside = md == 0 || md == 3 || md == 4;
  t.win = p->n_tilts > 1 && (per_slice(p) || inside);
  t.after = md == 1 || md == 3 || md == 4;
  t.before = md == 2 || md == 4;
  t.g_in = p->step_tilts && inside;
  t.g_after = p->step_tilts && t.after;
  t.g_before = p->step_tilts && t.before;
  return t;
}

Layout plan_layout(const pty_problem* p) {
  Layout L{};
  const size_t n2 = size_t(p->n) * p->n;
  const int B = batch_windows(p);
  const int S = p->n_slices, K = p->n_obj_modes, M = p->n_probe_modes;
  // probe-gradient groups: the register tile keeps a group's sum in
  // registers, so one workgroup per CU (G * M <= 256 CUs) walks several
  // windows; the other tiles add into the slot in memory per window and take
  // enough groups (G * M >= 1024 workgroups) to fill the chip
  const int G = kLayoutOf(p->n) == LAY_RT ? std::min(B, std::max(1, 256 / M))
                                          : std::min(B, std::max(64, (1024 + M - 1) / M));
  const size_t objsz = size_t(p->obj_ny) * p->obj_nx * S * K * 8;
  size_t o = 0;
  auto take = [&](int id, size_t bytes) {
    L.off[id] = bytes ? o : 0;
    L.bytes[id] = bytes;
    o += al(bytes);
  };
  take(WB_H, size_t(n_prop_tables(p)) * n2 * 8);
  take(WB_H0, n2 * 8);
  take(WB_MASKC, n2 * 4);
  take(WB_ZCOL, size_t(p->n) * 4);
  take(WB_FPROBE, size_t(M) * n2 * 8);
  take(WB_OBJP, p->obj_layout == PTY_LAYOUT_REFERENCE ? objsz : 0);
  take(WB_GRADP, p->obj_layout == PTY_LAYOUT_REFERENCE ? objsz : 0);
  take(WB_PROBE_PART, size_t(G) * M * n2 * 8);
  take(WB_PROBE_SUM, size_t(M) * n2 * 8);
  take(WB_LOSS_PART, size_t(std::max(p->n_positions, 1)) * 2 * 8 * loss_chunks(p->n));
  take(WB_ABT, size_t(p->n_aberrations) * n2 * 4);
  L.lane0 = o;
  take(WB_OB, size_t(B) * S * K * n2 * 8);
  take(WB_PP, size_t(B) * M * n2 * 8);
  // every slice, or the checkpoints of all segments but the last followed by
  // one segment of slices (the last segment's, left by the forward, then each
  // recomputed segment's in turn)
  L.ck = checkpoint_every(p);
  L.C = L.ck ? (S + L.ck - 1) / L.ck : 0;
  L.slots = L.ck ? L.C - 1 + L.ck : S;
  take(WB_STASH, size_t(B) * M * K * L.slots * n2 * 8);
  take(WB_GSTATE, L.ck ? size_t(B) * M * K * n2 * 8 : 0);
  take(WB_PSI, size_t(B) * M * K * n2 * 8);
  take(WB_COEF, size_t(B) * n2 * 4);
  take(WB_GP, size_t(B) * M * K * n2 * 8);
  take(WB_MEAS_STAGE, preprocessing(p) ? size_t(B) * n2 * 4 : 0);
  take(WB_POS_PART, size_t(B) * M * 2 * 8);
  take(WB_BC_PART, p->step_beam_current ? size_t(B) * M * 8 : 0);
  take(WB_PLATE, size_t(p->n_aberrations ? B : 0) * n2 * 8);
  take(WB_AB_W, p->step_aberrations ? size_t(B) * M * n2 * 4 : 0);
  take(WB_AB_PART, p->step_aberrations ? size_t(B) * p->n_aberrations * 8 : 0);
  const TiltPlan tp = tilt_plan(p);
  take(WB_WSP, tp.g_in ? size_t(B) * M * K * S * n2 * 8 : 0);
  take(WB_XE, tp.g_after ? size_t(B) * M * K * n2 * 8 : 0);
  take(WB_TILT_PART, (tp.g_in || tp.g_after) ? size_t(B) * M * K * 4 * 8 : 0);
  take(WB_FW, (tp.before && p->shift_active) ? size_t(B) * M * n2 * 8 : 0);
  take(WB_TB_PART, tp.g_before ? size_t(B) * M * 2 * 8 : 0);
  // windows larger than 128 px: one tile per workgroup of the largest launch
  L.scratch_tiles = p->n > 128 ? int(std::max<size_t>({size_t(B) * S * K, size_t(((B + 7) / 8) * 8) * M * K,
                                                       size_t(G) * M, size_t(M)}))
                               : 0;
  take(WB_SCRATCH, size_t(L.scratch_tiles) * n2 * 8);
  L.lane_bytes = o - L.lane0;
  L.lanes = std::max(1, std::min(p->lanes, int(std::max<size_t>(1, kLaneBudget / std::max<size_t>(L.lane_bytes, 1)))));
  L.total = L.lane0 + size_t(L.lanes) * L.lane_bytes;
  L.B = B;
  L.G = G;
  return L;
}

int validate(const pty_problem* p) {
  if (!p) return PTY_ERR_ARG;
  if (!pty_supported_window(p->n) && p->n != 192 && p->n != 256) return PTY_ERR_SIZE;
  if (p->n_slices < 1 || p->n_obj_modes < 1 || p->n_probe_modes < 1 || p->compute_batch < 1) return PTY_ERR_ARG;
  if (p->obj_ny < p->n || p->obj_nx < p->n) return PTY_ERR_ARG;
  if (p->checkpoint_slices < 0) return PTY_ERR_ARG;
  if (p->lanes < 0 || p->lanes > kMaxLanes) return PTY_ERR_ARG;
  if (p->n_aberrations < 0 || p->n_aberrations > PTY_MAX_ABERRATIONS) return PTY_ERR_ARG;
  if (p->n_aberrations > 0 && p->n_markers < 1) return PTY_ERR_ARG;
  if (p->step_aberrations && p->n_aberrations == 0) return PTY_ERR_ARG;
  if (p->algorithm < PTY_LOSS_LSQ_SQRT || p->algorithm > PTY_LOSS_POISSON_COMPRESSED) return PTY_ERR_ARG;
  if (p->n_distances < 0 || (p->n_distances > 1 && p->n_distances < p->n_slices - 1)) return PTY_ERR_ARG;
  if (p->n_tilts < 0 || (p->n_tilts > 1 && p->n_tilts != p->n_positions)) return PTY_ERR_ARG;
  if (p->tilt_mode < 0 || p->tilt_mode > 4) return PTY_ERR_ARG;
  // tilts before the specimen (modes 2, 4) with aberration plates are not on this path
  if (p->n_tilts > 0 && (p->tilt_mode == 2 || p->tilt_mode == 4) && p->n_aberrations > 0) return PTY_ERR_ARG;
  if (p->step_tilts && p->n_tilts < 1) return PTY_ERR_ARG;
  if (compressed(p->algorithm) && (p->n_masks < 1 || p->n_masks > kMaxMasks)) return PTY_ERR_ARG;
  if (preprocessing(p)) {
    const int raw = p->raw_n ? p->raw_n : p->n, bin = std::max(p->data_bin, 1);
    if (raw % bin != 0 || raw / bin + 2 * p->data_pad != p->n || p->data_pad < 0) return PTY_ERR_SIZE;
    if (std::abs(p->data_shift_y) >= raw || std::abs(p->data_shift_x) >= raw) return PTY_ERR_ARG;
  }
  return PTY_OK;
}

// the buffers a pass needs, checked before anything is launched
int validate_buffers(const pty_problem* p, const pty_buffers* b) {
  if (!b || !b->obj || !b->probe || !b->positions || !b->measured || !b->sequence || !b->loss_sse) return PTY_ERR_ARG;
  if (b->n_seq < 1 || b->n_seq > p->n_positions) return PTY_ERR_ARG;
  if ((p->shift_active || p->step_pos) && !b->corrections) return PTY_ERR_ARG;
  if (p->step_pos && !b->pos_grad) return PTY_ERR_ARG;
  if (p->step_obj && !b->obj_grad) return PTY_ERR_ARG;
  if (p->step_probe && !b->probe_grad) return PTY_ERR_ARG;
  if (compressed(p->algorithm) && !b->masks) return PTY_ERR_ARG;
  if (b->static_background && compressed(p->algorithm)) return PTY_ERR_ARG;
  if (p->step_static_background && (!b->static_background || !b->static_background_grad)) return PTY_ERR_ARG;
  if (p->step_beam_current && (!b->beam_current || !b->beam_current_grad)) return PTY_ERR_ARG;
  if (p->n_aberrations > 0 && (!b->aberrations || !b->aberration_marker)) return PTY_ERR_ARG;
  if (p->step_aberrations && !b->aberrations_grad) return PTY_ERR_ARG;
  if (per_slice(p) && !b->distances) return PTY_ERR_ARG;
  if (p->n_tilts > 0 && !b->tilts) return PTY_ERR_ARG;
  if (p->step_tilts && !b->tilts_grad) return PTY_ERR_ARG;
  if (b->measured_order != PTY_MEASURED_BY_POSITION && b->measured_order != PTY_MEASURED_BY_SEQUENCE)
    return PTY_ERR_ARG;
  return PTY_OK;
}

// ------------------------------------------------------------- issuing
// Events of a pass: the fork from the caller's stream, a join per lane and,
// per lane, the event recorded after a batch's additions into the buffers
// every batch shares (the next batch, on the next lane, waits for it).
enum { EV_FORK = 0, EV_JOIN = 1, EV_ACC = 1 + kMaxLanes, EV_COUNT = 1 + 2 * kMaxLanes };
enum { OP_RECORD = 1000, OP_WAIT = 1001 };
// sub-kinds of a fill or layout launch
enum { SUB_NONE = 0, SUB_TO_PLANAR, SUB_TO_REFERENCE, SUB_PER_WINDOW };

// One step of a pass: a launch of kind KT_* on a lane with its arguments, or
// an event record / wait (OP_RECORD / OP_WAIT of event ev).  A fill records
// its buffer (p0, n0).
struct Op {
  int kind, lane, ev, sub;
  BatchArgs a;
  const void* p0;
  size_t n0;
};

// Issues the steps of a pass to the lanes' streams: lane 0 is the caller's
// stream, lanes 1.. the library streams paired with it.
struct Enqueue {
  hipStream_t st[kMaxLanes] = {};
  hipEvent_t ev[EV_COUNT] = {};
  bool timed = true;  // HIP-event timing of the launches when pty_timing_enable is on
  template <class F>
  int launch(int kind, int lane, const BatchArgs&, F&& f, int = SUB_NONE, const void* = nullptr, size_t = 0) {
    TimeScope ts(kind, st[lane], timed);
    f(st[lane]);
    return check_launch();
  }
  int record(int lane, int e) { return hipEventRecord(ev[e], st[lane]) == hipSuccess ? PTY_OK : PTY_ERR_HIP; }
  int wait(int lane, int e) { return hipStreamWaitEvent(st[lane], ev[e], 0) == hipSuccess ? PTY_OK : PTY_ERR_HIP; }
};

// Records the steps instead of issuing them: the plan that plan_check() walks
// on the host (no HIP call is made).
struct Recorder {
  std::vector<Op> ops;
  template <class F>
  int launch(int kind, int lane, const BatchArgs& a, F&&, int sub = SUB_NONE, const void* p0 = nullptr,
             size_t n0 = 0) {
    ops.push_back(Op{kind, lane, -1, sub, a, p0, n0});
    return PTY_OK;
  }
  int record(int lane, int e) {
    ops.push_back(Op{OP_RECORD, lane, e, SUB_NONE, BatchArgs{}, nullptr, 0});
    return PTY_OK;
  }
  int wait(int lane, int e) {
    ops.push_back(Op{OP_WAIT, lane, e, SUB_NONE, BatchArgs{}, nullptr, 0});
    return PTY_OK;
  }
};

#define PTY_TRY(expr)          \
  do {                         \
    const int rc_ = (expr);    \
    if (rc_) return rc_;       \
  } while (0)

template <class X>
int issue_zero(X& x, int lane, const BatchArgs& a, void* ptr, size_t bytes) {
  if (bytes == 0) return PTY_OK;
  return x.launch(KT_ZERO, lane, a, [&](hipStream_t s) {
    const size_t nf = bytes / 4;
    const int grid = int(std::min<size_t>((nf + 255) / 256, 65536));
    k_zero<<<grid, 256, 0, s>>>(static_cast<float*>(ptr), nf);
  }, SUB_NONE, ptr, bytes);
}

// the tilt code paths (per-window propagator factors, after-tilt ramps, tilt
// gradients) are a separate instantiation, so that the kernels without tilts
// keep their registers
inline bool tilted(const BatchArgs& a) { return a.tilt_win || a.tilt_after || a.tg_in || a.tg_after; }

template <int N>
void launch_forward(const BatchArgs& a, hipStream_t s) {
  constexpr int LAY = kLayout<N>;
  const int grid = xcd_grid(a.B, a.M * a.K);
  if constexpr (LAY == LAY_RT) {
    if (tilted(a))
      k_forward_rt<N, true><<<grid, RT<N>::NT, 0, s>>>(a);
    else
      k_forward_rt<N, false><<<grid, RT<N>::NT, 0, s>>>(a);
  } else {
    if (tilted(a))
      k_forward_gt<N, true><<<grid, kGtThreads<N>, 0, s>>>(a);
    else
      k_forward_gt<N, false><<<grid, kGtThreads<N>, 0, s>>>(a);
  }
}

template <int N>
void launch_adjoint(const BatchArgs& a, hipStream_t s) {
  constexpr int LAY = kLayout<N>;
  const int grid = xcd_grid(a.B, a.M * a.K);
  if constexpr (LAY == LAY_RT) {
    if (tilted(a))
      k_adjoint_rt<N, true><<<grid, RT<N>::NT, 0, s>>>(a);
    else
      k_adjoint_rt<N, false><<<grid, RT<N>::NT, 0, s>>>(a);
  } else {
    if (tilted(a))
      k_adjoint_gt<N, true><<<grid, kGtThreads<N>, 0, s>>>(a);
    else
      k_adjoint_gt<N, false><<<grid, kGtThreads<N>, 0, s>>>(a);
  }
}

// one workgroup per (window, slice of [s_lo, s_hi), object mode)
template <int N>
void launch_objgrad(const BatchArgs& a, hipStream_t s) {
  constexpr int LAY = kLayout<N>;
  const int grid = a.B * (a.s_hi - a.s_lo) * a.K;
  if constexpr (LAY == LAY_RT)
    k_objgrad_rt<N><<<grid, RT<N>::NT, 0, s>>>(a);
  else
    k_objgrad_gt<N><<<grid, kGtThreads<N>, 0, s>>>(a);
}

// One batch of windows on one lane.  With checkpoint segments (ck > 0, C
// segments) the forward keeps, per window and mode, the wave entering
// segments 0..C-2 in stash slots 0..C-2 and every wave of the last segment in
// the tail slots C-1.. (a.st_dense).  The object gradient then walks the
// segments from the last: recompute the segment's waves from its checkpoint
// into the tail (except the last segment, already there), run the adjoint
// through the segment (the wave crossing segments goes through a.gstate), and
// band-limit the segment's window gradients.  Every launch computes exactly
// what the whole-pass launch computes for those slices, in the same order.
// The launches that add into buffers shared by all batches (object gradient,
// probe partials, position and background gradients) wait for event wait_ev
// (the previous batch's, on its lane; -1: none) and then record done_ev.
template <int N, class X>
int issue_batch(const BatchArgs& a, const pty_problem* p, int ck, int lane, X& x, int wait_ev, int done_ev) {
  constexpr int LAY = kLayout<N>;
  const int planes = a.S * a.K;
  PTY_TRY(x.launch(KT_PATCH, lane, a, [&](hipStream_t s) {
    if constexpr (LAY == LAY_RT)
      k_patch_rt<N><<<a.B * planes, RT<N>::NT, 0, s>>>(a);
    else
      k_patch_gt<N><<<a.B * planes, kGtThreads<N>, 0, s>>>(a);
  }));
  if (a.NA)
    PTY_TRY(x.launch(KT_AB_PLATE, lane, a, [&](hipStream_t s) {
      k_aber_plate<N><<<int((size_t(a.B) * N * N + 255) / 256), 256, 0, s>>>(a);
    }));
  if (a.tilt_before) {
    // the tilted probes, then with a shift their spectra and the shifted probes
    PTY_TRY(x.launch(KT_TILT_PROBE, lane, a, [&](hipStream_t s) {
      k_tilt_probe<N><<<int((size_t(a.B) * a.M * N * N + 255) / 256), 256, 0, s>>>(a);
    }));
    if (a.moved) {
      PTY_TRY(x.launch(KT_FPROBE, lane, a, [&](hipStream_t s) {
        if constexpr (LAY == LAY_RT)
          k_fprobe_rt<N><<<a.B * a.M, RT<N>::NT, 0, s>>>(a.Pp, a.Fw);
        else
          k_fprobe_gt<N><<<a.B * a.M, kGtThreads<N>, 0, s>>>(a.Pp, a.Fw, a.scratch);
      }, SUB_PER_WINDOW));
      PTY_TRY(x.launch(KT_PROBE_SHIFT, lane, a, [&](hipStream_t s) {
        if constexpr (LAY == LAY_RT)
          k_probe_shift_rt<N, false><<<a.B * a.M, RT<N>::NT, 0, s>>>(a);
        else
          k_probe_shift_gt<N, false><<<a.B * a.M, kGtThreads<N>, 0, s>>>(a);
      }));
    }
  } else if (a.shift)
    PTY_TRY(x.launch(KT_PROBE_SHIFT, lane, a, [&](hipStream_t s) {
      if constexpr (LAY == LAY_RT) {
        if (a.NA)
          k_probe_shift_rt<N, true><<<a.B * a.M, RT<N>::NT, 0, s>>>(a);
        else
          k_probe_shift_rt<N, false><<<a.B * a.M, RT<N>::NT, 0, s>>>(a);
      } else {
        if (a.NA)
          k_probe_shift_gt<N, true><<<a.B * a.M, kGtThreads<N>, 0, s>>>(a);
        else
          k_probe_shift_gt<N, false><<<a.B * a.M, kGtThreads<N>, 0, s>>>(a);
      }
    }));
  PTY_TRY(x.launch(KT_FORWARD, lane, a, [&](hipStream_t s) { launch_forward<N>(a, s); }));
  if (a.prep)
    PTY_TRY(x.launch(KT_PREPROCESS, lane, a, [&](hipStream_t s) {
      k_preprocess<N><<<a.B * ((N * N + 255) / 256), 256, 0, s>>>(a);
    }));
  PTY_TRY(x.launch(KT_LOSS, lane, a, [&](hipStream_t s) {
    if (compressed(a.algorithm))
      k_loss_c<N, LAY><<<a.B, 256, 0, s>>>(a);
    else
      k_loss<N, LAY><<<a.B * LossSplit<N>::CHUNKS, 256, 0, s>>>(a);
  }));
  if (p->step_obj && ck) {
    const int C = (a.S + ck - 1) / ck;
    for (int c = C - 1; c >= 0; --c) {
      BatchArgs g = a;
      g.s_lo = c * ck;
      g.s_hi = std::min(a.S, g.s_lo + ck);
      g.st_every = 1;
      g.st_dense = 0;
      g.stash = a.stash + size_t(C - 1) * N * N;  // the tail slots of each block
      if (c < C - 1) {
        BatchArgs f = g;
        f.fin = a.stash + size_t(c) * N * N;
        f.fin_stride = size_t(a.st_slots) * N * N;
        PTY_TRY(x.launch(KT_RECOMPUTE, lane, f, [&](hipStream_t s) { launch_forward<N>(f, s); }));
      }
      PTY_TRY(x.launch(KT_ADJOINT, lane, g, [&](hipStream_t s) { launch_adjoint<N>(g, s); }));
      PTY_TRY(x.launch(KT_OBJGRAD, lane, g, [&](hipStream_t s) { launch_objgrad<N>(g, s); }));
    }
  } else {
    if (p->step_obj || p->step_probe || p->step_pos || p->step_beam_current || p->step_aberrations || p->step_tilts)
      PTY_TRY(x.launch(KT_ADJOINT, lane, a, [&](hipStream_t s) { launch_adjoint<N>(a, s); }));
    if (p->step_obj) PTY_TRY(x.launch(KT_OBJGRAD, lane, a, [&](hipStream_t s) { launch_objgrad<N>(a, s); }));
  }
  if (wait_ev >= 0) PTY_TRY(x.wait(lane, wait_ev));
  if (p->step_obj)
    PTY_TRY(x.launch(KT_GATHER, lane, a, [&](hipStream_t s) {
      const int tiles = ((a.ny + GTILE - 1) / GTILE) * ((a.nx + GTILE - 1) / GTILE);
      k_gather_grad<N><<<planes * tiles, 256, 0, s>>>(a);
    }));
  if (a.tilt_before && (p->step_probe || (p->step_pos && a.moved) || p->step_beam_current || a.tg_before))
    PTY_TRY(x.launch(KT_PROBEGRAD, lane, a, [&](hipStream_t s) {
      if constexpr (LAY == LAY_RT)
        k_probegrad_tilt_rt<N><<<a.G * a.M, RT<N>::NT, 0, s>>>(a);
      else
        k_probegrad_tilt_gt<N><<<a.G * a.M, kGtThreads<N>, 0, s>>>(a);
    }));
  else if (p->step_probe || (p->step_pos && a.shift) || p->step_beam_current || p->step_aberrations)
    PTY_TRY(x.launch(KT_PROBEGRAD, lane, a, [&](hipStream_t s) {
      if constexpr (LAY == LAY_RT) {
        // the unshifted case (no plates) as its own instantiation: its beam-current
        // sums beside the shifted case's in one kernel spilled
        if (!a.shift)
          k_probegrad_rt<N, false, false><<<a.G * a.M, RT<N>::NT, 0, s>>>(a);
        else if (a.NA)
          k_probegrad_rt<N, true, true><<<a.G * a.M, RT<N>::NT, 0, s>>>(a);
        else
          k_probegrad_rt<N, false, true><<<a.G * a.M, RT<N>::NT, 0, s>>>(a);
      } else {
        if (a.NA)
          k_probegrad_gt<N, true><<<a.G * a.M, kGtThreads<N>, 0, s>>>(a);
        else
          k_probegrad_gt<N, false><<<a.G * a.M, kGtThreads<N>, 0, s>>>(a);
      }
    }));
  if (p->step_pos && a.shift)
    PTY_TRY(x.launch(KT_POS_FINAL, lane, a, [&](hipStream_t s) { k_pos_final<<<(a.B + 63) / 64, 64, 0, s>>>(a); }));
  if (p->step_beam_current)
    PTY_TRY(x.launch(KT_BC_FINAL, lane, a, [&](hipStream_t s) { k_bc_final<<<(a.B + 63) / 64, 64, 0, s>>>(a); }));
  if (a.tg_in || a.tg_after || a.tg_before)
    PTY_TRY(x.launch(KT_TILT_FINAL, lane, a, [&](hipStream_t s) { k_tilt_final<<<(a.B + 63) / 64, 64, 0, s>>>(a); }));
  if (p->step_aberrations) {
    PTY_TRY(x.launch(KT_AB_PART, lane, a, [&](hipStream_t s) { k_aber_part<N><<<a.B * a.NA, 256, 0, s>>>(a); }));
    PTY_TRY(x.launch(KT_AB_FINAL, lane, a, [&](hipStream_t s) {
      k_aber_final<<<(a.n_markers * a.NA + 255) / 256, 256, 0, s>>>(a);
    }));
  }
  if (a.bg_grad)
    PTY_TRY(x.launch(KT_BG_GRAD, lane, a, [&](hipStream_t s) {
      k_bg_grad<N, LAY><<<(N * N + 255) / 256, 256, 0, s>>>(a);
    }));
  if (done_ev >= 0) PTY_TRY(x.record(lane, done_ev));
  return PTY_OK;
}

// The arguments shared by every batch of a pass (lane 0's per-batch buffers).
template <int N>
BatchArgs pass_args(const pty_problem* p, const pty_buffers* bf, const Layout& L, char* ws) {
  BatchArgs a{};
  auto at = [&](int id) { return L.bytes[id] ? ws + L.off[id] : nullptr; };
  a.S = p->n_slices; a.K = p->n_obj_modes; a.M = p->n_probe_modes; a.ny = p->obj_ny; a.nx = p->obj_nx;
  a.P = p->n_positions;
  a.algorithm = p->algorithm; a.step_obj = p->step_obj; a.step_probe = p->step_probe; a.step_pos = p->step_pos;
  // aberrations need the per-window probes and their spectral gradient also without a shift
  a.shift = p->shift_active || p->n_aberrations > 0; a.moved = p->shift_active;
  a.far_field = 1; a.write_stash = p->step_obj; a.scatter = 1;
  a.loss_weight = p->loss_weight;
  const bool ref = p->obj_layout == PTY_LAYOUT_REFERENCE;
  a.obj = ref ? reinterpret_cast<const float2*>(at(WB_OBJP)) : static_cast<const float2*>(bf->obj);
  a.probe = static_cast<const float2*>(bf->probe);
  a.fprobe = reinterpret_cast<float2*>(at(WB_FPROBE));
  a.pos = bf->positions; a.corr = bf->corrections; a.meas = bf->measured; a.seq = bf->sequence;
  a.H = reinterpret_cast<float2*>(at(WB_H)); a.H0 = reinterpret_cast<float2*>(at(WB_H0));
  a.maskc = reinterpret_cast<float*>(at(WB_MASKC)); a.zcol = reinterpret_cast<int*>(at(WB_ZCOL));
  a.Ob = reinterpret_cast<float2*>(at(WB_OB)); a.Pp = reinterpret_cast<float2*>(at(WB_PP));
  a.stash = reinterpret_cast<float2*>(at(WB_STASH)); a.Psi = reinterpret_cast<float2*>(at(WB_PSI));
  a.coef = reinterpret_cast<float*>(at(WB_COEF)); a.Gbuf = nullptr;
  a.gP = reinterpret_cast<float2*>(at(WB_GP)); a.probe_part = reinterpret_cast<float2*>(at(WB_PROBE_PART));
  a.pos_part = reinterpret_cast<double*>(at(WB_POS_PART));
  a.loss_part = reinterpret_cast<double*>(at(WB_LOSS_PART));
  a.sim = bf->sim_patterns;
  a.obj_grad = ref ? reinterpret_cast<float2*>(at(WB_GRADP)) : static_cast<float2*>(bf->obj_grad);
  a.pos_grad = bf->pos_grad; a.G = L.G;
  a.scratch = reinterpret_cast<float2*>(at(WB_SCRATCH));
  a.meas_order = bf->measured_order;
  a.prep = preprocessing(p) ? 1 : 0;
  a.raw_n = p->raw_n ? p->raw_n : N;
  a.sy = p->data_shift_y; a.sx = p->data_shift_x; a.bin = std::max(p->data_bin, 1); a.pad = p->data_pad;
  a.mult = (p->data_multiplier == 0.f) ? 1.f : p->data_multiplier;
  a.meas_stage = reinterpret_cast<float*>(at(WB_MEAS_STAGE));
  a.masks = bf->masks;
  a.J = compressed(p->algorithm) ? p->n_masks : 0;
  a.bg = bf->static_background;
  a.bg_grad = p->step_static_background ? bf->static_background_grad : nullptr;
  a.bc = bf->beam_current;
  a.bc_part = p->step_beam_current ? reinterpret_cast<double*>(at(WB_BC_PART)) : nullptr;
  a.bc_grad = p->step_beam_current ? bf->beam_current_grad : nullptr;
  a.NA = p->n_aberrations; a.n_markers = p->n_markers;
  a.ab = a.NA ? bf->aberrations : nullptr; a.marker = a.NA ? bf->aberration_marker : nullptr;
  a.abt = reinterpret_cast<float*>(at(WB_ABT));
  a.plate = reinterpret_cast<float2*>(at(WB_PLATE));
  a.ab_w = p->step_aberrations ? reinterpret_cast<float*>(at(WB_AB_W)) : nullptr;
  a.ab_part = p->step_aberrations ? reinterpret_cast<double*>(at(WB_AB_PART)) : nullptr;
  a.ab_grad = p->step_aberrations ? bf->aberrations_grad : nullptr;
  a.Hstride = per_slice(p) ? size_t(N) * N : 0;
  a.dists = per_slice(p) ? bf->distances : nullptr;
  a.dz = p->slice_distance;
  const TiltPlan tp = tilt_plan(p);
  a.tilt = p->n_tilts > 0 ? bf->tilts : nullptr;
  a.n_tilt = p->n_tilts;
  a.tilt_win = tp.win; a.tilt_after = tp.after; a.tg_in = tp.g_in; a.tg_after = tp.g_after;
  a.lambda = p->wavelength; a.px = p->pixel_size_x; a.py = p->pixel_size_y; a.edge = p->edge_multiplier;
  a.Wsp = reinterpret_cast<float2*>(at(WB_WSP));
  a.Xe = reinterpret_cast<float2*>(at(WB_XE));
  a.tilt_part = reinterpret_cast<double*>(at(WB_TILT_PART));
  a.tilt_grad = p->step_tilts ? bf->tilts_grad : nullptr;
  a.tilt_before = tp.before; a.tg_before = tp.g_before;
  a.Fw = reinterpret_cast<float2*>(at(WB_FW));
  a.tb_part = reinterpret_cast<double*>(at(WB_TB_PART));
  // a before-tilt gives every window its own probe
  if (tp.before) a.shift = 1;
  whole_pass(a);
  if (L.ck) {
    // the forward keeps the checkpoints and every wave of the last segment
    a.st_every = L.ck;
    a.st_dense = (L.C - 1) * L.ck;
    a.st_slots = L.slots;
    a.gstate = reinterpret_cast<float2*>(at(WB_GSTATE));
  }
  return a;
}

// The whole pass over nl lanes: tables, probe spectra and the planar object
// copy on lane 0; batch i on lane i % nl, each lane with its own per-batch
// buffers; the shared additions of the batches in batch order (events); the
// lanes join lane 0, which reduces the loss, the probe gradient and copies
// the object gradient back.
template <int N, class X>
int issue_pass(const pty_problem* p, const pty_buffers* bf, const Layout& L, int nl, X& x) {
  constexpr int LAY = kLayout<N>;
  const int M = p->n_probe_modes, planes = p->n_slices * p->n_obj_modes;
  const size_t n2 = size_t(N) * N;
  const size_t objelems = size_t(p->obj_ny) * p->obj_nx * planes;
  const int lgrid = int(std::min<size_t>((objelems + 255) / 256, 65536));
  char* ws = static_cast<char*>(bf->workspace);
  const BatchArgs a = pass_args<N>(p, bf, L, ws);
  PTY_TRY(x.launch(KT_TABLES, 0, a, [&](hipStream_t s) {
    // tables without tilt when each window brings its own; else row 0's inside
    // tilt (or the problem's tan_x, tan_y without tilt rows)
    const bool own = a.tilt_win != 0;
    TableArgs t{const_cast<float2*>(a.H), const_cast<float2*>(a.H0), const_cast<float*>(a.maskc),
                const_cast<int*>(a.zcol), p->slice_distance, p->wavelength, p->pixel_size_x, p->pixel_size_y,
                own ? 0.f : p->tan_x, own ? 0.f : p->tan_y, p->damping_cutoff, p->edge_multiplier,
                a.dists, (a.tilt && !own) ? a.tilt : nullptr};
    k_tables<N, LAY><<<dim3((N * N + 255) / 256, n_prop_tables(p)), 256, 0, s>>>(t);
  }));
  if (a.NA)
    PTY_TRY(x.launch(KT_AB_TABLE, 0, a, [&](hipStream_t s) {
      k_aber_table<N, LAY><<<(N * N + 255) / 256, 256, 0, s>>>(const_cast<float*>(a.abt), a.NA, p->pixel_size_x,
                                                                p->pixel_size_y, p->wavelength);
    }));
  PTY_TRY(x.launch(KT_FPROBE, 0, a, [&](hipStream_t s) {
    if constexpr (LAY == LAY_RT)
      k_fprobe_rt<N><<<M, RT<N>::NT, 0, s>>>(a.probe, const_cast<float2*>(a.fprobe));
    else
      k_fprobe_gt<N><<<M, kGtThreads<N>, 0, s>>>(a.probe, const_cast<float2*>(a.fprobe), a.scratch);
  }));
  if (p->obj_layout == PTY_LAYOUT_REFERENCE)
    PTY_TRY(x.launch(KT_LAYOUT, 0, a, [&](hipStream_t s) {
      k_ref_to_planar<<<lgrid, 256, 0, s>>>(static_cast<const float2*>(bf->obj), const_cast<float2*>(a.obj),
                                             p->obj_ny, p->obj_nx, planes);
    }, SUB_TO_PLANAR));
  if (p->step_static_background && !bf->accumulate)
    PTY_TRY(issue_zero(x, 0, a, bf->static_background_grad, n2 * sizeof(float)));
  if (p->step_beam_current && !bf->accumulate)
    PTY_TRY(issue_zero(x, 0, a, bf->beam_current_grad, size_t(p->n_positions) * sizeof(float)));
  if (p->step_aberrations && !bf->accumulate)
    PTY_TRY(issue_zero(x, 0, a, bf->aberrations_grad, size_t(p->n_markers) * p->n_aberrations * sizeof(float)));
  if (p->step_tilts && !bf->accumulate)
    PTY_TRY(issue_zero(x, 0, a, bf->tilts_grad, size_t(p->n_tilts) * 6 * sizeof(float)));
  if (p->step_obj && !bf->accumulate) PTY_TRY(issue_zero(x, 0, a, a.obj_grad, objelems * 8));
  if ((p->step_probe || p->step_pos || p->step_beam_current || p->step_aberrations || a.tg_before) &&
      !bf->accumulate)
    PTY_TRY(issue_zero(x, 0, a, a.probe_part, size_t(L.G) * M * n2 * 8));
  // lane l's arguments: its own copies of the per-batch buffers
  BatchArgs al[kMaxLanes];
  for (int l = 0; l < nl; ++l) {
    BatchArgs& y = al[l];
    y = a;
    const size_t d = size_t(l) * L.lane_bytes;
    auto off = [d](auto* ptr) { return ptr ? reinterpret_cast<decltype(ptr)>(reinterpret_cast<char*>(ptr) + d) : ptr; };
    y.Ob = off(y.Ob); y.Pp = off(y.Pp); y.stash = off(y.stash); y.gstate = off(y.gstate);
    y.Psi = off(y.Psi); y.coef = off(y.coef); y.gP = off(y.gP); y.pos_part = off(y.pos_part);
    y.bc_part = off(y.bc_part); y.plate = off(y.plate); y.ab_w = off(y.ab_w); y.ab_part = off(y.ab_part);
    y.Wsp = off(y.Wsp); y.Xe = off(y.Xe); y.tilt_part = off(y.tilt_part); y.Fw = off(y.Fw);
    y.tb_part = off(y.tb_part);
    y.meas_stage = off(y.meas_stage); y.scratch = off(y.scratch);
  }
  if (nl > 1) {
    PTY_TRY(x.record(0, EV_FORK));
    for (int l = 1; l < nl; ++l) PTY_TRY(x.wait(l, EV_FORK));
  }
  const int n_batches = (bf->n_seq + L.B - 1) / L.B;
  int rc = PTY_OK;
  for (int i = 0; i < n_batches && rc == PTY_OK; ++i) {
    const int l = i % nl;
    BatchArgs y = al[l];
    y.b0 = i * L.B;
    y.B = std::min(L.B, bf->n_seq - y.b0);
    y.G = std::min(L.G, y.B);
    const int wait_ev = (nl > 1 && i > 0) ? EV_ACC + (i - 1) % nl : -1, done_ev = nl > 1 ? EV_ACC + l : -1;
    rc = issue_batch<N>(y, p, L.ck, l, x, wait_ev, done_ev);
  }
  // the lanes' work joins lane 0 also when a launch failed
  for (int l = 1; l < nl; ++l) {
    PTY_TRY(x.record(l, EV_JOIN + l));
    PTY_TRY(x.wait(0, EV_JOIN + l));
  }
  if (rc) return rc;
  PTY_TRY(x.launch(KT_LOSS_FINAL, 0, a, [&](hipStream_t s) {
    k_loss_final<<<1, 256, 0, s>>>(a.loss_part, bf->n_seq * LossSplit<N>::CHUNKS, double(p->loss_weight),
                                   bf->loss_sse);
  }));
  if (p->step_probe) {
    float2* psum = reinterpret_cast<float2*>(ws + L.off[WB_PROBE_SUM]);
    const size_t per_group = size_t(M) * n2;
    PTY_TRY(x.launch(KT_PROBE_REDUCE, 0, a, [&](hipStream_t s) {
      k_probe_reduce<<<int((per_group + 255) / 256), 256, 0, s>>>(a.probe_part, L.G, per_group, psum);
    }));
    PTY_TRY(x.launch(KT_PROBE_FINAL, 0, a, [&](hipStream_t s) {
      if constexpr (LAY == LAY_RT)
        k_probe_final_rt<N><<<M, RT<N>::NT, 0, s>>>(psum, a.shift && !a.tilt_before, a.maskc,
                                                     static_cast<float2*>(bf->probe_grad));
      else
        k_probe_final_gt<N><<<M, kGtThreads<N>, 0, s>>>(psum, a.shift && !a.tilt_before, a.maskc,
                                                     static_cast<float2*>(bf->probe_grad), a.scratch);
    }));
  }
  if (p->step_obj && p->obj_layout == PTY_LAYOUT_REFERENCE)
    PTY_TRY(x.launch(KT_LAYOUT, 0, a, [&](hipStream_t s) {
      k_planar_to_ref<<<lgrid, 256, 0, s>>>(a.obj_grad, static_cast<float2*>(bf->obj_grad), p->obj_ny, p->obj_nx,
                                             planes);
    }, SUB_TO_REFERENCE));
  return PTY_OK;
}

// The library streams that carry lanes 1 .. kMaxLanes-1 of the batches issued
// on a caller stream, one set per (device, caller stream), created on first
// use at the default priority and kept for the life of the process, with the
// events that fork, order and join the lanes.  A call holds the state's mutex
// while it enqueues.  A call runs one lane when its stream is being captured,
// or while the lanes of another caller stream on the device still have work
// in flight (so lanes of several caller streams never run at the same time).
struct LaneState {
  hipStream_t stream[kMaxLanes] = {};  // [0] unused: lane 0 is the caller's stream
  hipEvent_t ev[EV_COUNT] = {};
  int dev = 0;
  std::mutex mu;
};
std::mutex g_lane_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<LaneState>> g_lanes;

// the caller's lane state, or null (one lane) when none can be used now
LaneState* lane_state(hipStream_t caller) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(caller, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
  std::lock_guard<std::mutex> g(g_lane_mu);
  for (auto& kv : g_lanes) {
    if (kv.first.first != dev || kv.first.second == caller) continue;
    for (int l = 1; l < kMaxLanes; ++l)
      if (hipEventQuery(kv.second->ev[EV_JOIN + l]) != hipSuccess) return nullptr;  // busy (or failed)
  }
  auto& slot = g_lanes[{dev, caller}];
  if (!slot) {
    auto st = std::make_unique<LaneState>();
    st->dev = dev;
    bool ok = true;
    for (int l = 1; ok && l < kMaxLanes; ++l)
      ok = hipStreamCreateWithFlags(&st->stream[l], hipStreamNonBlocking) == hipSuccess;
    for (int e = 0; ok && e < EV_COUNT; ++e) ok = hipEventCreateWithFlags(&st->ev[e], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      // release what was made; the call runs one lane
      for (int l = 1; l < kMaxLanes; ++l)
        if (st->stream[l]) (void)hipStreamDestroy(st->stream[l]);
      for (hipEvent_t e : st->ev)
        if (e) (void)hipEventDestroy(e);
      g_lanes.erase({dev, caller});
      return nullptr;
    }
    slot = std::move(st);
  }
  return slot.get();
}

// The workspace of pty_loss_and_direction_static: a file-scope device array,
// one copy per GPU the library runs on, held until the process ends.  Host
// code never passes the array's name to a kernel or copy: it uses the
// device address that hipGetSymbolAddress returns for the current device.
constexpr size_t kStaticWorkspaceBytes = size_t(PTY_STATIC_WORKSPACE_BYTES);
static __device__ __align__(256) char scratch[kStaticWorkspaceBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch`, looked up once per device
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return PTY_ERR_HIP;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PTY_ERR_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PTY_OK;
}

template <int N>
int loss_and_direction_static_n(const pty_problem* p, const pty_buffers* bf, hipStream_t s) {
  Enqueue x;
  x.st[0] = s;
  x.timed = false;  // no timing events are created
  return issue_pass<N>(p, bf, plan_layout(p), 1, x);
}

template <int N>
int loss_and_direction_n(const pty_problem* p, const pty_buffers* bf, hipStream_t s) {
  const Layout L = plan_layout(p);
  if (!bf->workspace || bf->workspace_bytes < L.total) return PTY_ERR_WORKSPACE;
  const int n_batches = (bf->n_seq + L.B - 1) / L.B;
  const int lanes = std::min(L.lanes, n_batches);
  LaneState* ln = lanes >= 2 ? lane_state(s) : nullptr;
  const int nl = ln ? lanes : 1;
  Enqueue x;
  x.st[0] = s;
  std::unique_lock<std::mutex> lane_lock;
  if (ln) {
    lane_lock = std::unique_lock<std::mutex>(ln->mu);
    for (int l = 1; l < nl; ++l) x.st[l] = ln->stream[l];
    for (int e = 0; e < EV_COUNT; ++e) x.ev[e] = ln->ev[e];
  }
  return issue_pass<N>(p, bf, L, nl, x);
}

// ------------------------------------------------------------ plan check
// A pass's plan checked on the host: every buffer range a launch reads or
// writes (from its arguments, as the kernel indexes them) stays inside its
// buffer, per-batch buffers inside the lane's block, every pair of launches
// on different lanes whose ranges overlap with a write is ordered by the
// events, and the batches cover the sequence once.
struct Access {
  uintptr_t lo, hi;  // [lo, hi) bytes
  bool write;
  int buf;           // WB_* or an external buffer (kExt + PB_*)
};
enum { PB_OBJ = 0, PB_PROBE, PB_MEASURED, PB_LOSS_SSE, PB_OBJ_GRAD, PB_PROBE_GRAD, PB_POS_GRAD, PB_SIM,
       PB_BG_GRAD, PB_BC, PB_BC_GRAD, PB_AB, PB_MARKER, PB_AB_GRAD, PB_DIST, PB_TILTS, PB_TILTS_GRAD, PB_COUNT };
constexpr int kExt = 100;

// the bytes of every range an op touches; gt: windows larger than 128 px
void op_accesses(const Op& o, const pty_problem* p, const pty_buffers* bf, const Layout& L, std::vector<Access>& out) {
  out.clear();
  if (o.kind == OP_RECORD || o.kind == OP_WAIT) return;
  const BatchArgs& a = o.a;
  const int n = p->n, S = a.S, K = a.K, M = a.M, MK = M * K;
  const size_t n2 = size_t(n) * n, c8 = n2 * 8, c4 = n2 * 4;
  const bool gt = n > 128;
  const size_t objbytes = size_t(p->obj_ny) * p->obj_nx * S * K * 8;
  auto ws = [&](int id) { return reinterpret_cast<uintptr_t>(bf->workspace) + L.off[id]; };
  auto add = [&](const void* base, size_t bytes, bool w, int buf) {
    if (base && bytes) out.push_back({reinterpret_cast<uintptr_t>(base), reinterpret_cast<uintptr_t>(base) + bytes, w, buf});
  };
  auto shared = [&](int id, bool w) { add(reinterpret_cast<void*>(ws(id)), L.bytes[id], w, id); };
  auto tiles = [&](int count) { if (gt) add(a.scratch, size_t(count) * c8, true, WB_SCRATCH); };
  const int B = a.B;
  // stash slots a launch over [s_lo, s_hi) touches per (window, mode): from
  // slot 0 to the last one it addresses, within blocks of st_slots
  auto stash_range = [&](const float2* base, int last_slot, bool w) {
    if (!base) return;
    add(base, ((size_t(B) * MK - 1) * a.st_slots + last_slot + 1) * c8, w, WB_STASH);
  };
  // each window's aberration plate
  auto plates = [&]() {
    if (a.NA) add(a.plate, size_t(B) * c8, false, WB_PLATE);
  };
  switch (o.kind) {
    case KT_TABLES:
      shared(WB_H, true); shared(WB_H0, true); shared(WB_MASKC, true); shared(WB_ZCOL, true);
      add(a.dists, size_t(p->n_distances) * 4, false, kExt + PB_DIST);
      add(a.tilt, size_t(p->n_tilts) * 24, false, kExt + PB_TILTS);
      break;
    case KT_TILT_FINAL:
      add(a.tilt_part, size_t(B) * MK * 4 * 8, false, WB_TILT_PART);
      add(a.tb_part, size_t(B) * M * 2 * 8, false, WB_TB_PART);
      add(a.tilt_grad, size_t(p->n_tilts) * 24, true, kExt + PB_TILTS_GRAD);
      break;
    case KT_AB_TABLE:
      shared(WB_ABT, true);
      break;
    case KT_AB_PLATE:
      shared(WB_ABT, false);
      add(a.ab, size_t(p->n_markers) * a.NA * 4, false, kExt + PB_AB);
      add(a.marker, size_t(p->n_positions) * 4, false, kExt + PB_MARKER);
      add(a.plate, size_t(B) * c8, true, WB_PLATE);
      break;
    case KT_AB_PART:
      shared(WB_ABT, false);
      add(a.ab_w, size_t(B) * M * c4, false, WB_AB_W);
      add(a.ab_part, size_t(B) * a.NA * 8, true, WB_AB_PART);
      break;
    case KT_AB_FINAL:
      add(a.ab_part, size_t(B) * a.NA * 8, false, WB_AB_PART);
      add(a.marker, size_t(p->n_positions) * 4, false, kExt + PB_MARKER);
      add(a.ab_grad, size_t(p->n_markers) * a.NA * 4, true, kExt + PB_AB_GRAD);
      break;
    case KT_FPROBE:
      if (o.sub == SUB_PER_WINDOW) {
        // the spectra of the windows' tilted probes
        add(a.Pp, size_t(B) * M * c8, false, WB_PP);
        add(a.Fw, size_t(B) * M * c8, true, WB_FW);
        tiles(B * M);
        break;
      }
      add(bf->probe, size_t(M) * c8, false, kExt + PB_PROBE);
      shared(WB_FPROBE, true);
      tiles(M);
      break;
    case KT_TILT_PROBE:
      add(bf->probe, size_t(M) * c8, false, kExt + PB_PROBE);
      add(a.bc, size_t(p->n_positions) * 4, false, kExt + PB_BC);
      add(a.tilt, size_t(p->n_tilts) * 24, false, kExt + PB_TILTS);
      add(a.Pp, size_t(B) * M * c8, true, WB_PP);
      break;
    case KT_LAYOUT:
      if (o.sub == SUB_TO_PLANAR) {
        add(bf->obj, objbytes, false, kExt + PB_OBJ);
        shared(WB_OBJP, true);
      } else {
        shared(WB_GRADP, false);
        add(bf->obj_grad, objbytes, true, kExt + PB_OBJ_GRAD);
      }
      break;
    case KT_ZERO:
      add(o.p0, o.n0, true, -1);
      break;
    case KT_PATCH:
      add(a.obj, objbytes, false, p->obj_layout == PTY_LAYOUT_REFERENCE ? WB_OBJP : kExt + PB_OBJ);
      shared(WB_MASKC, false); shared(WB_ZCOL, false);
      add(a.Ob, size_t(B) * S * K * c8, true, WB_OB);
      tiles(B * S * K);
      break;
    case KT_PROBE_SHIFT:
      if (a.Fw)
        add(a.Fw, size_t(B) * M * c8, false, WB_FW);
      else
        shared(WB_FPROBE, false);
      shared(WB_MASKC, false); shared(WB_ZCOL, false);
      add(a.bc, size_t(p->n_positions) * 4, false, kExt + PB_BC);
      if (!a.moved) plates();
      add(a.Pp, size_t(B) * M * c8, true, WB_PP);
      tiles(B * M);
      break;
    case KT_FORWARD:
    case KT_RECOMPUTE: {
      add(a.Ob, size_t(B) * S * K * c8, false, WB_OB);
      if (!a.shift) add(a.bc, size_t(p->n_positions) * 4, false, kExt + PB_BC);
      if (a.shift) add(a.Pp, size_t(B) * M * c8, false, WB_PP);
      shared(WB_H, false); shared(WB_H0, false); shared(WB_MASKC, false); shared(WB_ZCOL, false);
      int last = -1;  // the last slot the forward stores
      for (int s = a.s_lo; s < a.s_hi; ++s) {
        const int r = s - a.s_lo;
        const int slot = r >= a.st_dense ? a.st_dense / a.st_every + (r - a.st_dense)
                                         : (r % a.st_every == 0 ? r / a.st_every : -1);
        last = std::max(last, slot);
      }
      if (a.write_stash && last >= 0) stash_range(a.stash, last, true);
      add(a.dists, size_t(p->n_distances) * 4, false, kExt + PB_DIST);
      add(a.tilt, size_t(p->n_tilts) * 24, false, kExt + PB_TILTS);
      add(a.Wsp, size_t(B) * MK * S * c8, true, WB_WSP);
      if (!a.fin) add(a.Xe, size_t(B) * MK * c8, true, WB_XE);
      if (a.fin) {
        // the checkpoint of each (window, mode): fin + (b * MK + mk) * fin_stride
        add(a.fin, ((size_t(B) * MK - 1) * a.fin_stride + n2) * 8, false, WB_STASH);
      } else {
        add(a.Psi, size_t(B) * MK * c8, true, WB_PSI);
      }
      tiles(xcd_grid(B, MK));
      break;
    }
    case KT_PREPROCESS:
      add(a.meas_stage, size_t(B) * c4, true, WB_MEAS_STAGE);
      break;
    case KT_LOSS:
      add(a.Psi, size_t(B) * MK * c8, false, WB_PSI);
      if (a.prep) add(a.meas_stage, size_t(B) * c4, false, WB_MEAS_STAGE);
      add(a.coef, size_t(B) * c4, true, WB_COEF);
      add(reinterpret_cast<const char*>(a.loss_part) + size_t(a.b0) * 2 * 8 * loss_chunks(n),
          size_t(B) * 2 * 8 * loss_chunks(n), true, WB_LOSS_PART);
      if (a.sim) add(reinterpret_cast<const char*>(a.sim), size_t(p->n_positions) * c4, true, kExt + PB_SIM);
      break;
    case KT_ADJOINT:
      if (a.s_hi == S) {
        add(a.Psi, size_t(B) * MK * c8, false, WB_PSI);
        add(a.coef, size_t(B) * c4, false, WB_COEF);
      } else {
        add(a.gstate, size_t(B) * MK * c8, false, WB_GSTATE);
      }
      add(a.Ob, size_t(B) * S * K * c8, false, WB_OB);
      shared(WB_H, false); shared(WB_H0, false); shared(WB_ZCOL, false);
      if (a.step_obj) stash_range(a.stash, a.s_hi - 1 - a.s_lo, true);
      add(a.s_lo > 0 ? a.gstate : a.gP, size_t(B) * MK * c8, true, a.s_lo > 0 ? WB_GSTATE : WB_GP);
      add(a.dists, size_t(p->n_distances) * 4, false, kExt + PB_DIST);
      add(a.tilt, size_t(p->n_tilts) * 24, false, kExt + PB_TILTS);
      add(a.Wsp, size_t(B) * MK * S * c8, false, WB_WSP);
      if (a.s_hi == S) add(a.Xe, size_t(B) * MK * c8, false, WB_XE);
      add(a.tilt_part, size_t(B) * MK * 4 * 8, true, WB_TILT_PART);
      tiles(xcd_grid(B, MK));
      break;
    case KT_OBJGRAD:
      stash_range(a.stash, a.s_hi - 1 - a.s_lo, false);
      shared(WB_MASKC, false); shared(WB_ZCOL, false);
      add(a.Ob, size_t(B) * S * K * c8, true, WB_OB);
      tiles(B * (a.s_hi - a.s_lo) * K);
      break;
    case KT_GATHER:
      add(a.Ob, size_t(B) * S * K * c8, false, WB_OB);
      add(a.obj_grad, objbytes, true, p->obj_layout == PTY_LAYOUT_REFERENCE ? WB_GRADP : kExt + PB_OBJ_GRAD);
      break;
    case KT_PROBEGRAD:
      if (a.tilt_before) {
        add(a.gP, size_t(B) * MK * c8, false, WB_GP);
        shared(WB_MASKC, false);
        add(bf->probe, size_t(M) * c8, false, kExt + PB_PROBE);
        add(a.bc, size_t(p->n_positions) * 4, false, kExt + PB_BC);
        add(a.tilt, size_t(p->n_tilts) * 24, false, kExt + PB_TILTS);
        add(a.Fw, size_t(B) * M * c8, false, WB_FW);
        add(a.probe_part, size_t(a.G) * M * c8, true, WB_PROBE_PART);
        if (a.step_pos && a.moved) add(a.pos_part, size_t(B) * M * 2 * 8, true, WB_POS_PART);
        add(a.bc_part, size_t(B) * M * 8, true, WB_BC_PART);
        add(a.tb_part, size_t(B) * M * 2 * 8, true, WB_TB_PART);
        tiles(a.G * M);
        break;
      }
      add(a.gP, size_t(B) * MK * c8, false, WB_GP);
      shared(WB_FPROBE, false); shared(WB_MASKC, false);
      add(a.probe_part, size_t(a.G) * M * c8, true, WB_PROBE_PART);
      if (a.step_pos && a.shift) add(a.pos_part, size_t(B) * M * 2 * 8, true, WB_POS_PART);
      add(a.bc, size_t(p->n_positions) * 4, false, kExt + PB_BC);
      add(a.bc_part, size_t(B) * M * 8, true, WB_BC_PART);
      if (a.bc_part && !a.shift) add(bf->probe, size_t(M) * c8, false, kExt + PB_PROBE);
      plates();
      add(a.ab_w, size_t(B) * M * c4, true, WB_AB_W);
      tiles(a.G * M);
      break;
    case KT_BC_FINAL:
      add(a.bc_part, size_t(B) * M * 8, false, WB_BC_PART);
      add(a.bc_grad, size_t(p->n_positions) * 4, true, kExt + PB_BC_GRAD);
      break;
    case KT_POS_FINAL:
      add(a.pos_part, size_t(B) * M * 2 * 8, false, WB_POS_PART);
      add(a.pos_grad, size_t(p->n_positions) * 2 * 4, true, kExt + PB_POS_GRAD);
      break;
    case KT_BG_GRAD:
      add(a.coef, size_t(B) * c4, false, WB_COEF);
      add(a.bg_grad, c4, true, kExt + PB_BG_GRAD);
      break;
    case KT_LOSS_FINAL:
      add(a.loss_part, size_t(bf->n_seq) * 2 * 8 * loss_chunks(n), false, WB_LOSS_PART);
      add(bf->loss_sse, 16, true, kExt + PB_LOSS_SSE);
      break;
    case KT_PROBE_REDUCE:
      shared(WB_PROBE_PART, false);
      shared(WB_PROBE_SUM, true);
      break;
    case KT_PROBE_FINAL:
      shared(WB_PROBE_SUM, false); shared(WB_MASKC, false);
      add(bf->probe_grad, size_t(M) * c8, true, kExt + PB_PROBE_GRAD);
      tiles(M);
      break;
    default:
      break;
  }
}

const char* op_name(int kind) {
  if (kind == OP_RECORD) return "record";
  if (kind == OP_WAIT) return "wait";
  return (kind >= 0 && kind < KT_COUNT) ? kKindNames[kind] : "?";
}

// Checks `calls` passes of one problem issued one after another on one caller
// stream (lane state and workspace reused), each over n_seq positions on nl
// lanes.  Returns the number of violations; a description of each (and, with
// verbose, the plan) is appended to report.
// omit (for testing the check itself; the issued passes never omit anything):
// bit 0 drops the waits on the batches' ordering events from the recorded
// plan, bit 1 the waits on the fork, bit 2 the waits on the joins, bit 3
// checks the ranges against a stash one window smaller than the workspace's.
template <int N>
int plan_check_n(const pty_problem* p, int n_seq, int nl, int calls, int accumulate, int verbose, int omit,
                 std::string& report) {
  Layout L = plan_layout(p);
  // distinct fake address ranges, 1 TiB apart: the workspace and each external buffer
  constexpr uintptr_t kSpan = uintptr_t(1) << 40;
  pty_buffers bf{};
  auto fake = [&](int i) { return reinterpret_cast<void*>(kSpan * uintptr_t(i + 1)); };
  bf.obj = fake(1 + PB_OBJ); bf.probe = fake(1 + PB_PROBE); bf.positions = static_cast<const int*>(fake(40));
  bf.corrections = static_cast<const float*>(fake(41)); bf.measured = static_cast<const float*>(fake(1 + PB_MEASURED));
  bf.sequence = static_cast<const int*>(fake(42)); bf.n_seq = n_seq;
  bf.loss_sse = static_cast<double*>(fake(1 + PB_LOSS_SSE)); bf.obj_grad = fake(1 + PB_OBJ_GRAD);
  bf.probe_grad = fake(1 + PB_PROBE_GRAD); bf.pos_grad = static_cast<float*>(fake(1 + PB_POS_GRAD));
  bf.workspace = fake(0); bf.workspace_bytes = L.total; bf.masks = static_cast<const float*>(fake(43));
  bf.measured_order = PTY_MEASURED_BY_POSITION; bf.accumulate = accumulate;
  if (p->step_static_background) {
    bf.static_background = static_cast<const float*>(fake(44));
    bf.static_background_grad = static_cast<float*>(fake(1 + PB_BG_GRAD));
  }
  if (p->step_beam_current) {
    bf.beam_current = static_cast<const float*>(fake(1 + PB_BC));
    bf.beam_current_grad = static_cast<float*>(fake(1 + PB_BC_GRAD));
  }
  if (p->n_aberrations) {
    bf.aberrations = static_cast<const float*>(fake(1 + PB_AB));
    bf.aberration_marker = static_cast<const int*>(fake(1 + PB_MARKER));
  }
  if (p->step_aberrations) bf.aberrations_grad = static_cast<float*>(fake(1 + PB_AB_GRAD));
  if (per_slice(p)) bf.distances = static_cast<const float*>(fake(1 + PB_DIST));
  if (p->n_tilts) bf.tilts = static_cast<const float*>(fake(1 + PB_TILTS));
  if (p->step_tilts) bf.tilts_grad = static_cast<float*>(fake(1 + PB_TILTS_GRAD));
  int rc = validate_buffers(p, &bf);
  if (rc) return rc;
  const int n_batches = (n_seq + L.B - 1) / L.B;
  nl = std::max(1, std::min({nl, L.lanes, n_batches}));
  std::vector<Op> ops;
  for (int c = 0; c < calls; ++c) {
    Recorder r;
    PTY_TRY(issue_pass<N>(p, &bf, L, nl, r));
    for (const Op& o : r.ops) {
      const bool drop = o.kind == OP_WAIT && (((omit & 1) && o.ev >= EV_ACC) || ((omit & 2) && o.ev == EV_FORK) ||
                                              ((omit & 4) && o.ev >= EV_JOIN && o.ev < EV_ACC));
      if (!drop) ops.push_back(o);
    }
  }
  if (omit & 8) L.bytes[WB_STASH] -= L.bytes[WB_STASH] / L.B;
  int bad = 0;
  char line[512];
  auto note = [&](const char* s) {
    report += s;
    report += "\n";
  };
  // vector clocks: per lane, per event; an op's stamp is its lane's clock
  std::vector<std::vector<int>> vc(ops.size());
  std::vector<std::vector<int>> lane(kMaxLanes, std::vector<int>(kMaxLanes, 0));
  std::vector<std::vector<int>> evc(EV_COUNT, std::vector<int>(kMaxLanes, 0));
  std::vector<std::vector<Access>> acc(ops.size());
  const uintptr_t ws0 = reinterpret_cast<uintptr_t>(bf.workspace);
  for (size_t i = 0; i < ops.size(); ++i) {
    const Op& o = ops[i];
    auto& lc = lane[o.lane];
    if (o.kind == OP_RECORD) {
      evc[o.ev] = lc;
    } else if (o.kind == OP_WAIT) {
      for (int l = 0; l < kMaxLanes; ++l) lc[l] = std::max(lc[l], evc[o.ev][l]);
    } else {
      lc[o.lane] += 1;
      vc[i] = lc;
      op_accesses(o, p, &bf, L, acc[i]);
    }
    if (verbose) {
      snprintf(line, sizeof line, "op %zu lane %d %s%s b0 %d B %d slices [%d,%d) ev %d", i, o.lane, op_name(o.kind),
               o.kind == KT_LAYOUT ? (o.sub == SUB_TO_PLANAR ? " to-planar" : " to-reference") : "", o.a.b0, o.a.B,
               o.a.s_lo, o.a.s_hi, o.ev);
      note(line);
    }
    // bounds: workspace ranges inside their buffer (per-lane buffers: the op's lane's copy)
    for (const Access& x : acc[i]) {
      if (x.lo < ws0 || x.lo >= ws0 + kSpan) continue;  // an external buffer
      if (x.hi > ws0 + L.total) {
        snprintf(line, sizeof line, "OUT OF WORKSPACE: op %zu %s lane %d touches [%zu, %zu) of %zu bytes", i,
                 op_name(o.kind), o.lane, size_t(x.lo - ws0), size_t(x.hi - ws0), L.total);
        note(line);
        ++bad;
        continue;
      }
      if (x.buf < 0 || x.buf >= WB_COUNT) continue;  // a fill: checked against the buffers below
      const size_t base = L.off[x.buf] + (x.buf >= kFirstLaneBuffer ? size_t(o.lane) * L.lane_bytes : 0);
      if (x.lo - ws0 < base || x.hi - ws0 > base + L.bytes[x.buf]) {
        snprintf(line, sizeof line, "OUT OF BUFFER: op %zu %s lane %d touches %s [%zu, %zu), its buffer is [%zu, %zu)",
                 i, op_name(o.kind), o.lane, kBufNames[x.buf], size_t(x.lo - ws0), size_t(x.hi - ws0), base,
                 base + L.bytes[x.buf]);
        note(line);
        ++bad;
      }
    }
  }
  // races: overlapping ranges on different lanes, one of them written, not ordered by the events
  for (size_t j = 0; j < ops.size(); ++j) {
    if (vc[j].empty()) continue;
    for (size_t i = 0; i < j; ++i) {
      if (vc[i].empty() || ops[i].lane == ops[j].lane) continue;
      const int li = ops[i].lane;
      if (vc[j][li] >= vc[i][li]) continue;  // i happens before j
      for (const Access& x : acc[i])
        for (const Access& y : acc[j])
          if ((x.write || y.write) && x.lo < y.hi && y.lo < x.hi) {
            snprintf(line, sizeof line, "RACE: op %zu %s (lane %d, b0 %d) and op %zu %s (lane %d, b0 %d) on [%zu, %zu)",
                     i, op_name(ops[i].kind), li, ops[i].a.b0, j, op_name(ops[j].kind), ops[j].lane, ops[j].a.b0,
                     size_t(std::max(x.lo, y.lo) - ws0), size_t(std::min(x.hi, y.hi) - ws0));
            note(line);
            ++bad;
          }
    }
  }
  // coverage: the patch launches of each call cover [0, n_seq) once, in order
  int call = 0, at = 0;
  for (const Op& o : ops) {
    if (o.kind != KT_PATCH) continue;
    if (at == n_seq) {
      ++call;
      at = 0;
    }
    if (o.a.b0 != at || o.a.B < 1 || o.a.b0 + o.a.B > n_seq) {
      snprintf(line, sizeof line, "COVERAGE: call %d: batch at %d of %d windows, expected one at %d", call, o.a.b0,
               o.a.B, at);
      note(line);
      ++bad;
    }
    at = o.a.b0 + o.a.B;
  }
  if (call != calls - 1 || at != n_seq) {
    snprintf(line, sizeof line, "COVERAGE: the batches end at call %d, window %d of %d", call, at, n_seq);
    note(line);
    ++bad;
  }
  snprintf(line, sizeof line, "plan: %zu steps, %d lanes, %d batches of up to %d windows per call, %d call(s); %d "
           "violation(s)", ops.size(), nl, n_batches, L.B, calls, bad);
  note(line);
  return bad;
}

template <int N>
int launch_tables_natural(const void* prop, const float* mask_nat, float edge, void* tables, hipStream_t s) {
  const size_t n2 = size_t(N) * N;
  float2* H = static_cast<float2*>(tables);
  float2* H0 = H + n2;
  float* mask = reinterpret_cast<float*>(H0 + n2);
  TimeScope ts(KT_TABLES, s);
  k_tables_natural<N><<<(N * N + 255) / 256, 256, 0, s>>>(static_cast<const float2*>(prop), mask_nat, edge, H, H0,
                                                           mask);
  return check_launch();
}

template <int N>
int ms_forward_n(const pty_problem* p, int batch, const void* probe_b, const void* obj_b, const void* prop,
                 void* stash, void* exit_wave, void* tables, hipStream_t s) {
  using T = Tile<N>;
  const size_t n2 = size_t(N) * N;
  float2* H = static_cast<float2*>(tables);
  float2* H0 = H + n2;
  float* mask = reinterpret_cast<float*>(H0 + n2);
  int rc;
  if ((rc = launch_tables_natural<N>(prop, nullptr, p->edge_multiplier, tables, s))) return rc;
  BatchArgs a{};
  a.B = batch; a.S = p->n_slices; a.K = p->n_obj_modes; a.M = p->n_probe_modes;
  a.shift = 1; a.far_field = 0; a.write_stash = 1;
  a.Pp = const_cast<float2*>(static_cast<const float2*>(probe_b));
  a.Ob = const_cast<float2*>(static_cast<const float2*>(obj_b));
  a.stash = static_cast<float2*>(stash);
  a.Psi = static_cast<float2*>(exit_wave);
  a.H = H; a.H0 = H0; a.mask = mask;
  whole_pass(a);
  TimeScope ts(KT_FORWARD, s);
  k_forward<N><<<xcd_grid(batch, a.M * a.K), T::NT, 0, s>>>(a);
  return check_launch();
}

template <int N>
int ms_adjoint_n(const pty_problem* p, int batch, void* g, void* stash, const void* obj_b, const void* prop,
                 const float* mask_nat, void* grad_b, void* tables, hipStream_t s) {
  using T = Tile<N>;
  const size_t n2 = size_t(N) * N;
  float2* H = static_cast<float2*>(tables);
  float2* H0 = H + n2;
  float* mask = reinterpret_cast<float*>(H0 + n2);
  int rc;
  if ((rc = launch_tables_natural<N>(prop, mask_nat, p->edge_multiplier, tables, s))) return rc;
  BatchArgs a{};
  a.B = batch; a.S = p->n_slices; a.K = p->n_obj_modes; a.M = p->n_probe_modes;
  a.far_field = 0; a.step_obj = 1; a.scatter = 0;
  a.Psi = static_cast<float2*>(g);
  a.gP = static_cast<float2*>(g);  // each workgroup reads its own tile before overwriting it
  a.stash = static_cast<float2*>(stash);
  a.Ob = const_cast<float2*>(static_cast<const float2*>(obj_b));
  a.Gbuf = static_cast<float2*>(grad_b);
  a.H = H; a.H0 = H0; a.mask = mask;
  whole_pass(a);
  {
    TimeScope ts(KT_ADJOINT, s);
    k_adjoint<N><<<xcd_grid(batch, a.M * a.K), T::NT, 0, s>>>(a);
  }
  {
    TimeScope ts(KT_OBJGRAD, s);
    k_objgrad<N><<<batch * a.S * a.K, T::NT, 0, s>>>(a);
  }
  return check_launch();
}

}  // namespace

// ================================================================ C ABI
extern "C" {

const char* pty_status_string(int status) {
  switch (status) {
    case PTY_OK: return "ok";
    case PTY_ERR_ARG: return "invalid argument";
    case PTY_ERR_SIZE: return "unsupported window size";
    case PTY_ERR_HIP: return "HIP runtime error";
    case PTY_ERR_BOUNDS: return "window outside the object";
    case PTY_ERR_WORKSPACE: return "workspace missing or too small";
    default: return "unknown status";
  }
}

int pty_version(void) { return 100; }

int pty_supported_window(int n) {
  return (n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128) ? 1 : 0;
}

int pty_fft2_c64(void* data, int64_t batch, int n, int mode, void* stream) {
  if (batch < 0 || (batch > 0 && !data)) return PTY_ERR_ARG;
  if (mode < PTY_FFT_FWD || mode > PTY_FFT_INV_ISHIFT) return PTY_ERR_ARG;
  if (batch == 0) return pty_supported_window(n) ? PTY_OK : PTY_ERR_SIZE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_N(n, launch_fft2, static_cast<float2*>(data), batch, mode, nullptr, s);
}

int pty_bandlimit_c64(void* data, int64_t batch, int n, const float* mask, void* stream) {
  if (batch < 0 || (batch > 0 && (!data || !mask))) return PTY_ERR_ARG;
  if (batch == 0) return pty_supported_window(n) ? PTY_OK : PTY_ERR_SIZE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_N(n, launch_fft2, static_cast<float2*>(data), batch, (int)MODE_BAND, mask, s);
}

int pty_scatteradd_c64(void* full, int ny, int nx, int planes, int layout, const void* patches, int batch, int n,
                       const int32_t* corners, void* stream) {
  if (!full || !patches || !corners || ny < 1 || nx < 1 || planes < 1 || n < 1 || batch < 0) return PTY_ERR_ARG;
  if (layout != PTY_LAYOUT_PLANAR && layout != PTY_LAYOUT_REFERENCE) return PTY_ERR_ARG;
  if (batch == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TimeScope ts(KT_SCATTER, s);
  k_scatteradd<<<batch * planes, 256, 0, s>>>(static_cast<float2*>(full), ny, nx, planes,
                                               static_cast<const float2*>(patches), n, corners, layout);
  return check_launch();
}

int pty_workspace_bytes(const pty_problem* p, size_t* bytes) {
  int rc = validate(p);
  if (rc) return rc;
  if (!bytes) return PTY_ERR_ARG;
  *bytes = plan_layout(p).total;
  return PTY_OK;
}

int pty_loss_and_direction(const pty_problem* p, const pty_buffers* b, void* stream) {
  int rc = validate(p);
  if (rc) return rc;
  if ((rc = validate_buffers(p, b))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_FUSED(p->n, loss_and_direction_n, p, b, s);
}

int pty_loss_and_direction_static(const pty_problem* p, const pty_buffers* b, void* stream) {
  int rc = validate(p);
  if (rc) return rc;
  if (!b) return PTY_ERR_ARG;
  // one lane: no library streams or events are made
  pty_problem q = *p;
  q.lanes = 1;
  if (plan_layout(&q).total > kStaticWorkspaceBytes) return PTY_ERR_WORKSPACE;
  pty_buffers c = *b;
  c.workspace_bytes = kStaticWorkspaceBytes;
  c.workspace = nullptr;
  if ((rc = scratch_address(&c.workspace))) return rc;
  if ((rc = validate_buffers(&q, &c))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_FUSED(q.n, loss_and_direction_static_n, &q, &c, s);
}

int pty_tables_bytes(int n, size_t* bytes) {
  if (!bytes || !pty_supported_window(n)) return PTY_ERR_SIZE;
  *bytes = size_t(n) * n * (8 + 8 + 4);
  return PTY_OK;
}

int pty_ms_forward(const pty_problem* p, int batch, const void* probe_b, const void* obj_b, const void* propagator,
                   void* stash, void* exit_wave, void* tables, void* stream) {
  if (!p || !pty_supported_window(p->n)) return PTY_ERR_SIZE;
  if (p->n_slices < 1 || p->n_obj_modes < 1 || p->n_probe_modes < 1) return PTY_ERR_ARG;
  if (batch < 1 || !probe_b || !obj_b || !propagator || !stash || !exit_wave || !tables) return PTY_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_N(p->n, ms_forward_n, p, batch, probe_b, obj_b, propagator, stash, exit_wave, tables, s);
}

int pty_ms_adjoint(const pty_problem* p, int batch, void* g, void* stash, const void* obj_b,
                   const void* propagator, const float* mask, void* grad_b, void* tables, void* stream) {
  if (!p || !pty_supported_window(p->n)) return PTY_ERR_SIZE;
  if (p->n_slices < 1 || p->n_obj_modes < 1 || p->n_probe_modes < 1) return PTY_ERR_ARG;
  if (batch < 1 || !g || !stash || !obj_b || !propagator || !mask || !grad_b || !tables) return PTY_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PTY_DISPATCH_N(p->n, ms_adjoint_n, p, batch, g, stash, obj_b, propagator, mask, grad_b, tables, s);
}

int pty_plan_check(const pty_problem* p, int n_seq, int lanes, int calls, int accumulate, int verbose, int omit,
                   char* report, size_t report_bytes) {
  const int rc = validate(p);
  if (rc) return rc;
  if (n_seq < 1 || n_seq > p->n_positions || lanes < 1 || calls < 1) return PTY_ERR_ARG;
  std::string text;
  int bad;
  switch (p->n) {
    case 16: bad = plan_check_n<16>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 32: bad = plan_check_n<32>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 48: bad = plan_check_n<48>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 64: bad = plan_check_n<64>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 96: bad = plan_check_n<96>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 128: bad = plan_check_n<128>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 192: bad = plan_check_n<192>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    case 256: bad = plan_check_n<256>(p, n_seq, lanes, calls, accumulate, verbose, omit, text); break;
    default: return PTY_ERR_SIZE;
  }
  if (report && report_bytes) {
    const size_t k = std::min(text.size(), report_bytes - 1);
    memcpy(report, text.data(), k);
    report[k] = 0;
  }
  return bad;
}

void pty_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
}

int pty_timing_read(const char** names, double* total_ms, int64_t* launches, int cap) {
  std::vector<TimedLaunch> pend;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    pend.swap(g_pending);
  }
  for (auto& t : pend) {
    float ms = 0.f;
    (void)hipEventSynchronize(t.b);
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      g_total_ms[t.kind] += ms;
      g_launches[t.kind] += 1;
    }
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  const int n = std::min<int>(cap, KT_COUNT);
  for (int i = 0; i < n; ++i) {
    if (names) names[i] = kKindNames[i];
    if (total_ms) total_ms[i] = g_total_ms[i];
    if (launches) launches[i] = g_launches[i];
  }
  return KT_COUNT;
}

void pty_timing_reset(void) {
  pty_timing_read(nullptr, nullptr, nullptr, 0);
  for (int i = 0; i < KT_COUNT; ++i) {
    g_total_ms[i] = 0.0;
    g_launches[i] = 0;
  }
}

}  // extern "C"
