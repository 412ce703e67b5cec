// Large 2D FFTs of whole object planes (e.g. 512x512, 1536x1536) and the
// Fourier band-limit of pypty/utils.py:102-150 (fourier_clean), plus the
// vector kernels of the L-BFGS update (pypty/iterative.py:573-932).
//
// A plane is transformed by two kinds of workgroups, each keeping its lines
// in LDS:
//   rows pass : a workgroup owns `rows` full rows of one plane
//   cols pass : a workgroup owns a block of `cols` adjacent columns of one
//               plane (all ny samples of each)
// A line of length L = prod(r_t), r_t in {2, 3, 4, 5, 7, 8} (other primes up to
// 31 through a direct-DFT stage), is transformed in
// place by mixed-radix decimation in frequency: stage t splits blocks of size
// B into r_t interleaved sub-blocks (butterfly radix r_t, stride B/r_t,
// twiddle W_B^(j k)); the spectrum comes out in mixed-radix digit-reversed
// order.  The inverse runs the stages backwards (conjugate twiddle, then
// butterfly) and consumes that order.  The band-limit is pointwise in the
// spectrum, so it is applied in the column pass between the forward and the
// inverse column transforms, with each position mapped to its frequency.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "pty_fft.hpp"
#include "pypty_amd.h"

using namespace pty;

namespace {

constexpr int kMaxStages = 16;

struct LinePlan {
  int L;
  int nst;
  int direct;  // a stage of a prime radix above kMaxGenericRadix (computed through an LDS copy)
  int r[kMaxStages];
};

// factor L into radices 8/4/2/3/5/7 and, for any other prime factor up to
// kMaxGenericRadix, a direct-DFT stage
constexpr int kMaxGenericRadix = 31;

bool make_plan(int L, LinePlan& p) {
  p.L = L;
  p.nst = 0;
  int m = L;
  for (int q = 11; q <= kMaxGenericRadix && m > 1; q += 2) {
    while (m % q == 0 && p.nst < kMaxStages) {
      p.r[p.nst++] = q;
      m /= q;
    }
  }
  while (m % 7 == 0) {
    p.r[p.nst++] = 7;
    m /= 7;
  }
  while (m % 5 == 0) {
    p.r[p.nst++] = 5;
    m /= 5;
  }
  while (m % 3 == 0) {
    p.r[p.nst++] = 3;
    m /= 3;
  }
  while (m % 8 == 0) {
    p.r[p.nst++] = 8;
    m /= 8;
  }
  while (m % 4 == 0) {
    p.r[p.nst++] = 4;
    m /= 4;
  }
  while (m % 2 == 0) {
    p.r[p.nst++] = 2;
    m /= 2;
  }
  // larger primes: a direct DFT stage each
  p.direct = 0;
  for (int q = kMaxGenericRadix + 2; m > 1 && p.nst < kMaxStages; q += 2) {
    if (q * q > m) q = m;  // m is prime
    while (m % q == 0 && p.nst < kMaxStages) {
      p.r[p.nst++] = q;
      m /= q;
      p.direct = 1;
    }
  }
  return m == 1 && p.nst <= kMaxStages && L >= 2;
}

}  // namespace

// frequency held by position p after the forward transform
__device__ __forceinline__ int plan_freq(const LinePlan& pl, int p) {
  int f = 0, mult = 1, B = pl.L;
  for (int t = 0; t < pl.nst; ++t) {
    const int r = pl.r[t], s = B / r;
    f += mult * (p / s);
    p %= s;
    mult *= r;
    B = s;
  }
  return f;
}

__device__ __forceinline__ float2 twiddle_rt(int num, int den, bool inv) {
  // exp(-+2 pi i num/den) with an accurate sincospi
  float s, c;
  sincospif(-2.0f * float(num % den) / float(den), &s, &c);
  return make_float2(c, inv ? -s : s);
}

template <int R, bool INV>
__device__ __forceinline__ void stage_bfly(float2* x, int es, int base, int s, int j, int B) {
  float2 v[R];
  if constexpr (!INV) {
#pragma unroll
    for (int n = 0; n < R; ++n) v[n] = x[(base + s * n) * es];
    dft<R, false>(v);
#pragma unroll
    for (int k = 1; k < R; ++k) v[k] = cmul(v[k], twiddle_rt(j * k, B, false));
#pragma unroll
    for (int k = 0; k < R; ++k) x[(base + s * k) * es] = v[k];
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = x[(base + s * k) * es];
#pragma unroll
    for (int k = 1; k < R; ++k) v[k] = cmul(v[k], twiddle_rt(j * k, B, true));
    dft<R, true>(v);
#pragma unroll
    for (int n = 0; n < R; ++n) x[(base + s * n) * es] = v[n];
  }
}

// butterfly of a prime radix r in (7, 31]: direct DFT with run-time twiddles
template <bool INV>
__device__ __noinline__ void stage_bfly_generic(float2* x, int es, int base, int s, int j, int B, int r) {
  float2 v[kMaxGenericRadix], out[kMaxGenericRadix];
  for (int n = 0; n < r; ++n) v[n] = x[(base + s * n) * es];
  if (INV)
    for (int k = 1; k < r; ++k) v[k] = cmul(v[k], twiddle_rt(j * k, B, true));
  for (int k = 0; k < r; ++k) {
    float2 acc = make_float2(0.f, 0.f);
    for (int n = 0; n < r; ++n) acc = cadd(acc, cmul(v[n], twiddle_rt((n * k) % r, r, INV)));
    out[k] = acc;
  }
  if (!INV)
    for (int k = 1; k < r; ++k) out[k] = cmul(out[k], twiddle_rt(j * k, B, false));
  for (int k = 0; k < r; ++k) x[(base + s * k) * es] = out[k];
}

// butterflies of a prime radix r above kMaxGenericRadix: every thread forms one
// output of one butterfly from the r inputs in LDS (with the stage's twiddles),
// into tmp (the lines' layout), then the lines are copied back.  All threads
// of the block take part.
template <bool INV>
__device__ void stage_direct(float2* buf, float2* tmp, int nlines, int lstride, int es, int L, int r, int s, int B) {
  const int work = nlines * L;
  for (int w = threadIdx.x; w < work; w += blockDim.x) {
    const int line = w / L, q = w % L;
    const int bfly = q / r, k = q % r;
    const int blk = bfly / s, j = bfly % s;
    const int base = blk * B + j;
    const float2* x = buf + line * lstride;
    float2 acc = make_float2(0.f, 0.f);
    for (int n = 0; n < r; ++n) {
      float2 v = x[(base + s * n) * es];
      if (INV && n > 0) v = cmul(v, twiddle_rt(j * n, B, true));
      acc = cadd(acc, cmul(v, twiddle_rt((n * k) % r, r, INV)));
    }
    if (!INV && k > 0) acc = cmul(acc, twiddle_rt(j * k, B, false));
    tmp[line * lstride + (base + s * k) * es] = acc;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < work; w += blockDim.x) {
    const int line = w / L, q = w % L;
    buf[line * lstride + q * es] = tmp[line * lstride + q * es];
  }
}

// Transform `nlines` lines in LDS; line l starts at buf + l*lstride and has
// element stride es.  All threads of the block take part.  tmp: LDS of the
// lines' size, used by the stages of primes above kMaxGenericRadix.
template <bool INV>
__device__ void lines_fft(float2* buf, int nlines, int lstride, int es, const LinePlan& pl, float2* tmp) {
  const int L = pl.L;
  for (int tt = 0; tt < pl.nst; ++tt) {
    const int t = INV ? pl.nst - 1 - tt : tt;
    int B = L;
    for (int u = 0; u < t; ++u) B /= pl.r[u];
    const int r = pl.r[t], s = B / r;
    if (r > kMaxGenericRadix) {
      stage_direct<INV>(buf, tmp, nlines, lstride, es, L, r, s, B);
      __syncthreads();
      continue;
    }
    const int nb = L / r;  // butterflies per line
    const int work = nb * nlines;
    for (int w = threadIdx.x; w < work; w += blockDim.x) {
      const int line = w / nb, q = w % nb;
      const int blk = q / s, j = q % s;
      float2* x = buf + line * lstride;
      const int base = blk * B + j;
      switch (r) {
        case 2: stage_bfly<2, INV>(x, es, base, s, j, B); break;
        case 3: stage_bfly<3, INV>(x, es, base, s, j, B); break;
        case 4: stage_bfly<4, INV>(x, es, base, s, j, B); break;
        case 5: stage_bfly<5, INV>(x, es, base, s, j, B); break;
        case 7: stage_bfly<7, INV>(x, es, base, s, j, B); break;
        case 8: stage_bfly<8, INV>(x, es, base, s, j, B); break;
        default: stage_bfly_generic<INV>(x, es, base, s, j, B, r); break;
      }
    }
    __syncthreads();
  }
}

// rows pass: grid (planes * ceil(ny/rows)); in place on planar (planes, ny, nx)
template <bool INV>
__global__ __launch_bounds__(256) void k_plane_rows(float2* __restrict__ data, int ny, int nx, int rows, LinePlan pl,
                                                    float scale) {
  extern __shared__ float2 lds[];
  const int nblk = (ny + rows - 1) / rows;
  const int plane = blockIdx.x / nblk, y0 = (blockIdx.x % nblk) * rows;
  const int nr = min(rows, ny - y0);
  float2* g = data + (size_t(plane) * ny + y0) * nx;
  for (int i = threadIdx.x; i < nr * nx; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
  lines_fft<INV>(lds, nr, nx, 1, pl, lds + size_t(rows) * nx);
  for (int i = threadIdx.x; i < nr * nx; i += blockDim.x) g[i] = cscale(lds[i], scale);
}

// cols pass: forward column transform, band mask, inverse column transform.
// The mask is the circle nu_x^2 + nu_y^2 <= (cutoff/2)^2 of
// pypty/utils.py:126-131 (normalised frequency, cosine taper of width
// rolloff/2 when rolloff != 0); positions are mapped to frequencies through
// both axes' digit-reversed orders.
__global__ __launch_bounds__(256) void k_plane_cols_band(float2* __restrict__ data, int ny, int nx, int cols,
                                                         LinePlan py, LinePlan px, double cutoff, double rolloff) {
  extern __shared__ float2 lds[];
  const int nblk = (nx + cols - 1) / cols;
  const int plane = blockIdx.x / nblk, x0 = (blockIdx.x % nblk) * cols;
  const int nc = min(cols, nx - x0);
  float2* g = data + size_t(plane) * ny * nx + x0;
  for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
    const int y = i / nc, c = i % nc;
    lds[y * cols + c] = g[size_t(y) * nx + c];
  }
  __syncthreads();
  lines_fft<false>(lds, nc, 1, cols, py, lds + size_t(ny) * cols);
  const double max_r = 0.5 * cutoff;
  const double r0 = (0.5 * (cutoff - rolloff)) * (0.5 * (cutoff - rolloff));
  for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
    const int y = i / nc, c = i % nc;
    const int fy = plan_freq(py, y), fx = plan_freq(px, x0 + c);
    // numpy.fft.fftfreq(n) = k * (1 / n), rounded as numpy does (a pixel on the circle stays on
    // the reference's side of it)
    const double vy = double((2 * fy < ny) ? fy : fy - ny) * (1.0 / double(ny));
    const double vx = double((2 * fx < nx) ? fx : fx - nx) * (1.0 / double(nx));
    const double r = vx * vx + vy * vy;
    float m = (r <= max_r * max_r) ? 1.f : 0.f;
    if (rolloff != 0.0 && r > r0 && m > 0.f)
      m *= float(0.5 * (1.0 + cos(3.141592654 * (r - r0) / (max_r * max_r - r0))));
    lds[y * cols + c] = cscale(lds[y * cols + c], m);
  }
  __syncthreads();
  lines_fft<true>(lds, nc, 1, cols, py, lds + size_t(ny) * cols);
  for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
    const int y = i / nc, c = i % nc;
    g[size_t(y) * nx + c] = lds[y * cols + c];
  }
}

// cols pass of a plain transform, writing natural order (scatter of the
// digit-reversed positions of both axes; shift folds in fftshift)
template <bool INV>
__global__ __launch_bounds__(256) void k_plane_cols(const float2* __restrict__ src, float2* __restrict__ dst, int ny,
                                                    int nx, int cols, LinePlan py, LinePlan px, int shift,
                                                    float scale) {
  extern __shared__ float2 lds[];
  const int nblk = (nx + cols - 1) / cols;
  const int plane = blockIdx.x / nblk, x0 = (blockIdx.x % nblk) * cols;
  const int nc = min(cols, nx - x0);
  const float2* g = src + size_t(plane) * ny * nx;
  float2* o = dst + size_t(plane) * ny * nx;
  if (!INV) {
    for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
      const int y = i / nc, c = i % nc;
      lds[y * cols + c] = g[size_t(y) * nx + x0 + c];
    }
    __syncthreads();
    lines_fft<false>(lds, nc, 1, cols, py, lds + size_t(ny) * cols);
    for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
      const int y = i / nc, c = i % nc;
      int fy = plan_freq(py, y), fx = plan_freq(px, x0 + c);
      if (shift) {
        fy = (fy + ny / 2) % ny;
        fx = (fx + nx / 2) % nx;
      }
      o[size_t(fy) * nx + fx] = cscale(lds[y * cols + c], scale);
    }
  } else {
    // natural (optionally shifted) spectrum in; gather each position's frequency
    for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
      const int y = i / nc, c = i % nc;
      int fy = plan_freq(py, y), fx = plan_freq(px, x0 + c);
      if (shift) {
        fy = (fy + ny / 2) % ny;
        fx = (fx + nx / 2) % nx;
      }
      lds[y * cols + c] = g[size_t(fy) * nx + fx];
    }
    __syncthreads();
    lines_fft<true>(lds, nc, 1, cols, py, lds + size_t(ny) * cols);
    for (int i = threadIdx.x; i < ny * nc; i += blockDim.x) {
      const int y = i / nc, c = i % nc;
      o[size_t(y) * nx + x0 + c] = cscale(lds[y * cols + c], scale);
    }
  }
}

// ----------------------------------------------------------- L-BFGS vectors
// out[p] = sum Re(a_p * conj(b_p)) over n complex elements, FP64 accumulation.
// Pairs are given as arrays of pointers; one block-row of the grid per pair.
struct DotPairs {
  const float2* a[16];
  const float2* b[16];
};

__global__ __launch_bounds__(256) void k_multidot(DotPairs pr, size_t n, int is_real, double* __restrict__ partial) {
  __shared__ double red[4];
  const int p = blockIdx.y;
  double s = 0.0;
  if (is_real) {
    const float* a = reinterpret_cast<const float*>(pr.a[p]);
    const float* b = reinterpret_cast<const float*>(pr.b[p]);
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
      s += double(a[i]) * double(b[i]);
  } else {
    const float2* a = pr.a[p];
    const float2* b = pr.b[p];
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
      const float2 u = a[i], v = b[i];
      s += double(u.x) * double(v.x) + double(u.y) * double(v.y);
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[size_t(p) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_reduce_rows(const double* __restrict__ partial, int cols,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  const int p = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < cols; i += blockDim.x) s += partial[size_t(p) * cols + i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[p] = red[0] + red[1] + red[2] + red[3];
}

// y = alpha * y + beta * x  (float32 complex viewed as 2n floats); alpha == 0
// does not read y, so y may start uninitialised
__global__ __launch_bounds__(256) void k_axpby(float* __restrict__ y, const float* __restrict__ x, size_t n, float alpha,
                                               float beta) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    y[i] = (alpha == 0.f ? 0.f : alpha * y[i]) + (x ? beta * x[i] : 0.f);
}

// L1 norms of the object's phase and log-amplitude (objective.py:1286-1320):
// per block the FP64 sum of phase_w |arg O| + abs_w |log |O||, and, when grad
// is set, grad += 0.5 (abs_w sign(log |O|) + i phase_w sign(arg O)) / conj(O).
// A weight <= 0 contributes nothing.
__global__ __launch_bounds__(256) void k_l1(const float2* __restrict__ obj, size_t n, float abs_w, float phase_w,
                                            float2* __restrict__ grad, double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float2 o = obj[i];
    float cre = 0.f, cim = 0.f;
    if (phase_w > 0.f) {
      const float ph = atan2f(o.y, o.x);
      s += double(phase_w) * double(fabsf(ph));
      cim = 0.5f * phase_w * float((ph > 0.f) - (ph < 0.f));
    }
    if (abs_w > 0.f) {
      const float la = logf(sqrtf(o.x * o.x + o.y * o.y));
      s += double(abs_w) * double(fabsf(la));
      cre = 0.5f * abs_w * float((la > 0.f) - (la < 0.f));
    }
    if (grad) {
      // c / conj(o) = c o / |o|^2
      const float r2 = o.x * o.x + o.y * o.y;
      float2 g = grad[i];
      g.x += (cre * o.x - cim * o.y) / r2;
      g.y += (cre * o.y + cim * o.x) / r2;
      grad[i] = g;
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_phase_grad(const float2* __restrict__ g, const float2* __restrict__ z,
                                                    float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float2 a = z[i], b = g[i];
    out[i] = 2.f * (a.x * b.y - a.y * b.x);
  }
}

// Charge flipping of the object (objective.py:1014-1110 without fancy_sigma):
// per plane the largest phase arg O and absorption -log(|O| + 1e-10); values
// below delta times that largest one are multiplied by beta; the object
// becomes exp(-absorption + i phase).  Planes: element i of plane i % planes
// (the reference layout (ny, nx, S, K)) or i / plane_len (planar).
__device__ __forceinline__ int cf_plane(size_t i, int planes, int layout, size_t plane_len) {
  return layout == PTY_LAYOUT_REFERENCE ? int(i % size_t(planes)) : int(i / plane_len);
}
__device__ __forceinline__ size_t cf_index(size_t k, int plane, int planes, int layout, size_t plane_len) {
  return layout == PTY_LAYOUT_REFERENCE ? k * size_t(planes) + plane : size_t(plane) * plane_len + k;
}
constexpr int kCfBlocks = 64;  // partial maxima per plane
__global__ __launch_bounds__(256) void k_cf_max(const float2* __restrict__ obj, int planes, int layout,
                                                size_t plane_len, float* __restrict__ partial) {
  __shared__ float rp[4], ra[4];
  const int plane = blockIdx.x / kCfBlocks, blk = blockIdx.x % kCfBlocks;
  float mp = -INFINITY, ma = -INFINITY;
  for (size_t k = size_t(blk) * 256 + threadIdx.x; k < plane_len; k += size_t(kCfBlocks) * 256) {
    const float2 o = obj[cf_index(k, plane, planes, layout, plane_len)];
    mp = fmaxf(mp, atan2f(o.y, o.x));
    ma = fmaxf(ma, -logf(hypotf(o.x, o.y) + 1e-10f));
  }
  for (int d = 32; d > 0; d >>= 1) {
    mp = fmaxf(mp, __shfl_down(mp, d, 64));
    ma = fmaxf(ma, __shfl_down(ma, d, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    rp[threadIdx.x >> 6] = mp;
    ra[threadIdx.x >> 6] = ma;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[2 * size_t(blockIdx.x)] = fmaxf(fmaxf(rp[0], rp[1]), fmaxf(rp[2], rp[3]));
    partial[2 * size_t(blockIdx.x) + 1] = fmaxf(fmaxf(ra[0], ra[1]), fmaxf(ra[2], ra[3]));
  }
}
// the per-plane maxima after the partials: out[2 * plane + {0, 1}]
__global__ void k_cf_reduce(const float* __restrict__ partial, int planes, float* __restrict__ out) {
  const int plane = blockIdx.x * blockDim.x + threadIdx.x;
  if (plane >= planes) return;
  float mp = -INFINITY, ma = -INFINITY;
  for (int b = 0; b < kCfBlocks; ++b) {
    mp = fmaxf(mp, partial[2 * (size_t(plane) * kCfBlocks + b)]);
    ma = fmaxf(ma, partial[2 * (size_t(plane) * kCfBlocks + b) + 1]);
  }
  out[2 * plane] = mp;
  out[2 * plane + 1] = ma;
}
__global__ __launch_bounds__(256) void k_cf_apply(float2* __restrict__ obj, size_t n, int planes, int layout,
                                                  size_t plane_len, const float* __restrict__ maxima, float dp, float da,
                                                  float bp, float ba) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const int plane = cf_plane(i, planes, layout, plane_len);
    const float2 o = obj[i];
    float ph = atan2f(o.y, o.x), ab = -logf(hypotf(o.x, o.y) + 1e-10f);
    if (ph < dp * maxima[2 * plane]) ph *= bp;
    if (ab < da * maxima[2 * plane + 1]) ab *= ba;
    const float m = expf(-ab);
    float sn, cs;
    sincosf(ph, &sn, &cs);
    obj[i] = make_float2(m * cs, m * sn);
  }
}

// Gram-Schmidt over probe modes (objective.py:1119-1140), one workgroup: for
// i = 1..M-1 and j < i in the reference's order, p_i -= p_j c with
// c = sum(conj(p_j) p_i) / (1e-10 + sum |p_j|^2) (FP64 sums in a fixed
// order); probe (npix, M) with the mode innermost, as the reference's (N, N, M).
__global__ __launch_bounds__(1024) void k_gram_schmidt(float2* __restrict__ p, int64_t npix, int M) {
  __shared__ double red[3][16];
  __shared__ double coef[2];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  for (int i = 1; i < M; ++i) {
    for (int j = 0; j < i; ++j) {
      double sr = 0.0, si = 0.0, nn = 0.0;
      for (int64_t k = tid; k < npix; k += blockDim.x) {
        const float2 a = p[k * M + j], b = p[k * M + i];
        sr += double(a.x) * double(b.x) + double(a.y) * double(b.y);  // Re conj(a) b
        si += double(a.x) * double(b.y) - double(a.y) * double(b.x);  // Im conj(a) b
        nn += double(a.x) * double(a.x) + double(a.y) * double(a.y);
      }
      for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_down(sr, o, 64);
        si += __shfl_down(si, o, 64);
        nn += __shfl_down(nn, o, 64);
      }
      if (lane == 0) {
        red[0][wid] = sr;
        red[1][wid] = si;
        red[2][wid] = nn;
      }
      __syncthreads();
      if (tid == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int w = 0; w < nw; ++w) {
          a += red[0][w];
          b += red[1][w];
          c += red[2][w];
        }
        coef[0] = a / (1e-10 + c);
        coef[1] = b / (1e-10 + c);
      }
      __syncthreads();
      const float cr = float(coef[0]), ci = float(coef[1]);
      for (int64_t k = tid; k < npix; k += blockDim.x) {
        const float2 a = p[k * M + j];
        float2 b = p[k * M + i];
        b.x -= a.x * cr - a.y * ci;
        b.y -= a.x * ci + a.y * cr;
        p[k * M + i] = b;
      }
      __syncthreads();
    }
  }
}

// A step of the phase of a phase-only object (iterative.py:737-738):
// out = |O| exp(i (arg O + step u)), u the real phase update
__global__ __launch_bounds__(256) void k_phase_step(const float2* __restrict__ obj, const float* __restrict__ u,
                                                    float step, float2* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float2 o = obj[i];
    const float r = hypotf(o.x, o.y), ph = atan2f(o.y, o.x) + step * u[i];
    float sn, cs;
    sincosf(ph, &sn, &cs);
    out[i] = make_float2(r * cs, r * sn);
  }
}

// ================================================================ host
namespace {

inline int plane_launch_check() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "pypty_amd: HIP launch error: %s\n", hipGetErrorString(e));
    return PTY_ERR_HIP;
  }
  return PTY_OK;
}

constexpr size_t kPlaneLds = 64 * 1024;

// lines per workgroup; a plan with a direct stage needs a second copy of the lines in LDS
int rows_per_block(int nx, const LinePlan& p) {
  return int(std::max<size_t>(1, kPlaneLds / (size_t(nx) * 8 * (p.direct ? 2 : 1))));
}
int cols_per_block(int ny, const LinePlan& p) {
  int c = int(kPlaneLds / (size_t(ny) * 8 * (p.direct ? 2 : 1)));
  if (c >= 16) return 16;
  if (c >= 8) return 8;
  return std::max(c, 1);
}
size_t lds_bytes(int lines, int len, const LinePlan& p) { return size_t(lines) * len * 8 * (p.direct ? 2 : 1); }

int plane_checks(void* data, int64_t planes, int ny, int nx, LinePlan& py, LinePlan& px) {
  if (!data || planes < 1 || ny < 2 || nx < 2) return PTY_ERR_ARG;
  if (!make_plan(ny, py) || !make_plan(nx, px)) return PTY_ERR_SIZE;
  if (size_t(nx) * 8 * (px.direct ? 2 : 1) > kPlaneLds || size_t(ny) * 8 * (py.direct ? 2 : 1) > kPlaneLds)
    return PTY_ERR_SIZE;
  if (planes * int64_t((ny + 1)) > (int64_t(1) << 31)) return PTY_ERR_ARG;
  return PTY_OK;
}

// Centre of mass of each pattern about the pattern's centre (pypty/dpc.py:139-150):
// comx = sum_x,y (x - (nx-1)/2) I(y, x), comy likewise in y (negated when the
// pattern's y axis is flipped, dpc.py:133-134), in FP64 as numpy forms them.
// One workgroup per pattern.
__global__ __launch_bounds__(256) void k_center_of_mass(const float* __restrict__ p, int ny, int nx, int flip_y,
                                                        double* __restrict__ comx, double* __restrict__ comy) {
  const float* d = p + size_t(blockIdx.x) * ny * nx;
  const double cx = 0.5 * double(nx - 1), cy = 0.5 * double(ny - 1);
  double sx = 0.0, sy = 0.0;
  const int n = ny * nx;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int y = e / nx, x = e - y * nx;
    const double v = double(d[e]);
    sx += (double(x) - cx) * v;
    sy += (double(y) - cy) * v;
  }
  __shared__ double rx[256], ry[256];
  rx[threadIdx.x] = sx;
  ry[threadIdx.x] = sy;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      rx[threadIdx.x] += rx[threadIdx.x + o];
      ry[threadIdx.x] += ry[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    comx[blockIdx.x] = rx[0];
    comy[blockIdx.x] = flip_y ? -ry[0] : ry[0];
  }
}

}  // namespace

extern "C" {

int pty_fourier_clean_c64(void* data, int64_t planes, int ny, int nx, double cutoff, double rolloff, void* stream) {
  LinePlan py{}, px{};
  int rc = plane_checks(data, planes, ny, nx, py, px);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  float2* d = static_cast<float2*>(data);
  const int rows = rows_per_block(nx, px), cols = cols_per_block(ny, py);
  const int gr = int(planes) * ((ny + rows - 1) / rows), gc = int(planes) * ((nx + cols - 1) / cols);
  k_plane_rows<false><<<gr, 256, lds_bytes(rows, nx, px), s>>>(d, ny, nx, rows, px, 1.0f);
  k_plane_cols_band<<<gc, 256, lds_bytes(cols, ny, py), s>>>(d, ny, nx, cols, py, px, cutoff, rolloff);
  k_plane_rows<true><<<gr, 256, lds_bytes(rows, nx, px), s>>>(d, ny, nx, rows, px, 1.0f / (float(ny) * float(nx)));
  return plane_launch_check();
}

int pty_plane_fft2_c64(void* data, void* scratch, int64_t planes, int ny, int nx, int mode, void* stream) {
  LinePlan py{}, px{};
  int rc = plane_checks(data, planes, ny, nx, py, px);
  if (rc) return rc;
  if (!scratch) return PTY_ERR_WORKSPACE;
  if (mode < PTY_FFT_FWD || mode > PTY_FFT_INV_ISHIFT) return PTY_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  float2* d = static_cast<float2*>(data);
  float2* w = static_cast<float2*>(scratch);
  const int rows = rows_per_block(nx, px), cols = cols_per_block(ny, py);
  const int gr = int(planes) * ((ny + rows - 1) / rows), gc = int(planes) * ((nx + cols - 1) / cols);
  const bool inv = (mode == PTY_FFT_INV || mode == PTY_FFT_INV_ISHIFT);
  const int shift = (mode == PTY_FFT_FWD_SHIFT || mode == PTY_FFT_INV_ISHIFT) ? 1 : 0;
  if (!inv) {
    // rows (digit-reversed x) in place, then columns scattered to natural order into scratch, copy back
    k_plane_rows<false><<<gr, 256, lds_bytes(rows, nx, px), s>>>(d, ny, nx, rows, px, 1.0f);
    k_plane_cols<false><<<gc, 256, lds_bytes(cols, ny, py), s>>>(d, w, ny, nx, cols, py, px, shift, 1.0f);
  } else {
    // gather natural spectrum into digit-reversed columns (scratch), then inverse rows
    k_plane_cols<true><<<gc, 256, lds_bytes(cols, ny, py), s>>>(d, w, ny, nx, cols, py, px, shift, 1.0f);
    k_plane_rows<true><<<gr, 256, lds_bytes(rows, nx, px), s>>>(w, ny, nx, rows, px, 1.0f / (float(ny) * float(nx)));
  }
  const hipError_t e = hipMemcpyAsync(d, w, size_t(planes) * ny * nx * 8, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return PTY_ERR_HIP;
  return plane_launch_check();
}

int pty_multidot_c64(const void* const* a, const void* const* b, int npairs, int64_t n, int is_real, double* out,
                     double* partial, void* stream) {
  if (npairs < 1 || npairs > 16 || n < 0 || !a || !b || !out || !partial) return PTY_ERR_ARG;
  DotPairs pr{};
  for (int i = 0; i < npairs; ++i) {
    if (!a[i] || !b[i]) return PTY_ERR_ARG;
    pr.a[i] = static_cast<const float2*>(a[i]);
    pr.b[i] = static_cast<const float2*>(b[i]);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int cols = 512;
  k_multidot<<<dim3(cols, npairs), 256, 0, s>>>(pr, size_t(n), is_real ? 1 : 0, partial);
  k_reduce_rows<<<npairs, 256, 0, s>>>(partial, cols, out);
  return plane_launch_check();
}

int pty_axpby_c64(void* y, const void* x, int64_t n, float alpha, float beta, void* stream) {
  if (!y || n < 0) return PTY_ERR_ARG;
  if (n == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nf = size_t(n) * 2;
  const int grid = int(std::min<size_t>((nf + 255) / 256, 8192));
  k_axpby<<<grid, 256, 0, s>>>(static_cast<float*>(y), static_cast<const float*>(x), nf, alpha, beta);
  return plane_launch_check();
}

int pty_axpby_f32(float* y, const float* x, int64_t n, float alpha, float beta, void* stream) {
  if (!y || n < 0) return PTY_ERR_ARG;
  if (n == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = int(std::min<size_t>((size_t(n) + 255) / 256, 8192));
  k_axpby<<<grid, 256, 0, s>>>(y, x, size_t(n), alpha, beta);
  return plane_launch_check();
}

int pty_l1_constraint_c64(const void* obj, int64_t n, float abs_weight, float phase_weight, void* grad,
                          double* out, double* partial, void* stream) {
  if (!obj || n < 0 || !out || !partial) return PTY_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int cols = 512;
  k_l1<<<cols, 256, 0, s>>>(static_cast<const float2*>(obj), size_t(n), abs_weight, phase_weight,
                            static_cast<float2*>(grad), partial);
  k_reduce_rows<<<1, 256, 0, s>>>(partial, cols, out);
  return plane_launch_check();
}

int pty_center_of_mass(const float* patterns, int64_t count, int ny, int nx, int flip_y, double* comx, double* comy,
                       void* stream) {
  if (!patterns || !comx || !comy || count < 0 || count > INT32_MAX || ny < 1 || nx < 1) return PTY_ERR_ARG;
  if (count == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_center_of_mass<<<int(count), 256, 0, s>>>(patterns, ny, nx, flip_y ? 1 : 0, comx, comy);
  return plane_launch_check();
}

int pty_charge_flip_c64(void* obj, int64_t n, int planes, int layout, float delta_phase, float delta_abs,
                        float beta_phase, float beta_abs, float* partial, void* stream) {
  if (!obj || !partial || n < 0 || planes < 1 || n % planes != 0) return PTY_ERR_ARG;
  if (layout != PTY_LAYOUT_REFERENCE && layout != PTY_LAYOUT_PLANAR) return PTY_ERR_ARG;
  if (n == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t plane_len = size_t(n) / size_t(planes);
  float2* o = static_cast<float2*>(obj);
  float* maxima = partial + 2 * size_t(planes) * kCfBlocks;
  k_cf_max<<<planes * kCfBlocks, 256, 0, s>>>(o, planes, layout, plane_len, partial);
  k_cf_reduce<<<(planes + 63) / 64, 64, 0, s>>>(partial, planes, maxima);
  const int grid = int(std::min<size_t>((size_t(n) + 255) / 256, 8192));
  k_cf_apply<<<grid, 256, 0, s>>>(o, size_t(n), planes, layout, plane_len, maxima, delta_phase, delta_abs, beta_phase,
                                  beta_abs);
  return plane_launch_check();
}

int pty_orthogonalize_modes_c64(void* probe, int64_t npix, int modes, void* stream) {
  if (!probe || npix < 0 || modes < 1) return PTY_ERR_ARG;
  if (modes == 1 || npix == 0) return PTY_OK;
  k_gram_schmidt<<<1, 1024, 0, static_cast<hipStream_t>(stream)>>>(static_cast<float2*>(probe), npix, modes);
  return plane_launch_check();
}

int pty_phase_step_c64(const void* obj, const float* u, float step, void* out, int64_t n, void* stream) {
  if (!obj || !u || !out || n < 0) return PTY_ERR_ARG;
  if (n == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = int(std::min<size_t>((size_t(n) + 255) / 256, 8192));
  k_phase_step<<<grid, 256, 0, s>>>(static_cast<const float2*>(obj), u, step, static_cast<float2*>(out), size_t(n));
  return plane_launch_check();
}

int pty_phase_grad_c64(const void* g, const void* z, float* out, int64_t n, void* stream) {
  if (!g || !z || !out || n < 0) return PTY_ERR_ARG;
  if (n == 0) return PTY_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = int(std::min<size_t>((size_t(n) + 255) / 256, 8192));
  k_phase_grad<<<grid, 256, 0, s>>>(static_cast<const float2*>(g), static_cast<const float2*>(z), out, size_t(n));
  return plane_launch_check();
}

}  // extern "C"
