// Learned warm start: the kernels of a small convolutional network that maps
// a log-scaled diffraction pattern to the projected phase around the probe.
// The network has no counterpart in the reference; it produces an initial
// object for the reconstruction (the reference's `obj` parameter,
// pypty/iterative.py:56-250).
//
// Every layer is one GEMM on the matrix cores with BF16 operands and FP32
// accumulation (v_mfma_f32_32x32x16_bf16): a convolution as im2col followed by
// cols x W^T, its backward as dY^T x cols (weights) and dY x W then col2im
// (input); a dense layer is the same with one pixel per pattern.  Activations
// are pixel-major with channels innermost (NHWC).  Weights are FP32 masters
// with a BF16 copy for the products.  All reductions run in a fixed order, so
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "pypty_amd.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16 bits, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

int launch_status() {
  return hipGetLastError() == hipSuccess ? PTY_OK : PTY_ERR_HIP;
}

// C[z] (M x N, row stride ldc) = op(A) x op(B) over k in [z*kchunk, (z+1)*kchunk)
//   op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// Workgroup: 4 wavefronts and a 64 x 64 tile of C, K in steps of 32.  Each
// step stages op(A) and op(B)^T as [row][k] tiles in LDS (one 16-byte global
// load per thread and operand, row pitch GLP) while the previous step's
// tiles feed the matrix cores; each wavefront owns a 32 x 32 block and
// accumulates gfx950's v_mfma_f32_32x32x16_bf16 (k 16 per instruction).
// Operand lanes: lane l holds A[m0 + l%32][k + 8*(l/32) + v] and
// B[k + 8*(l/32) + v][n0 + l%32], v = 0..7 (one 16-byte LDS read each);
// accumulator r of lane l is C[m0 + 8*(r/4) + 4*(l/32) + r%4][n0 + l%32].
constexpr int GBK = 32, GLP = GBK + 8;  // k per step; LDS row pitch (bf16), 80 bytes: rows 16-byte aligned

// 8 bf16 of the 64 x 32 tile X[row][k] (rows r0.., k k0..), X[row][k] =
// TRANS ? X[k*ld + row] : X[row*ld + k], zero outside rows < R and k < k_hi.
// Thread t: !TRANS row t/4, k 8*(t%4)..+7; TRANS k t/8, rows 8*(t%8)..+7.
template <bool TRANS>
__device__ __forceinline__ uint4 tile_load(const uint16_t* __restrict__ X, int ld, int R, int r0, int k0, int k_hi,
                                           bool vec) {
  const int t = threadIdx.x;
  uint4 out;
  uint16_t* v = reinterpret_cast<uint16_t*>(&out);
  if (!TRANS) {
    const int r = r0 + (t >> 2), k = k0 + 8 * (t & 3);
    if (vec && r < R && k + 7 < k_hi) return *reinterpret_cast<const uint4*>(X + size_t(r) * ld + k);
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (r < R && k + i < k_hi) ? X[size_t(r) * ld + k + i] : uint16_t(0);
  } else {
    const int k = k0 + (t >> 3), r = r0 + 8 * (t & 7);
    if (vec && k < k_hi && r + 7 < R) return *reinterpret_cast<const uint4*>(X + size_t(k) * ld + r);
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (k < k_hi && r + i < R) ? X[size_t(k) * ld + r + i] : uint16_t(0);
  }
  return out;
}
// the same 8 values into the LDS tile S[row][k] (pitch GLP)
template <bool TRANS>
__device__ __forceinline__ void tile_store(uint16_t* S, uint4 val) {
  const int t = threadIdx.x;
  const uint16_t* v = reinterpret_cast<const uint16_t*>(&val);
  if (!TRANS) {
    uint2* d = reinterpret_cast<uint2*>(S + (t >> 2) * GLP + 8 * (t & 3));
    d[0] = make_uint2(val.x, val.y);
    d[1] = make_uint2(val.z, val.w);
  } else {
    const int k = t >> 3, r = 8 * (t & 7);
#pragma unroll
    for (int i = 0; i < 8; ++i) S[(r + i) * GLP + k] = v[i];
  }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, size_t zstride, int kchunk) {
  __shared__ __align__(16) uint16_t As[64 * GLP];
  __shared__ __align__(16) uint16_t Bs[64 * GLP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  const int li = lane & 31, kq = (lane >> 5) * 8;
  // 16-byte global loads need 8-element aligned rows and base
  const bool avec = (lda % 8 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  const bool bvec = (ldb % 8 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  // op(B)^T[n][k] = TB ? B[n*ldb + k] : B[k*ldb + n]
  uint4 va = tile_load<TA>(A, lda, M, bm, k_lo, k_hi, avec);
  uint4 vb = tile_load<!TB>(B, ldb, N, bn, k_lo, k_hi, bvec);
  for (int k0 = k_lo; k0 < k_hi; k0 += GBK) {
    __syncthreads();  // the previous step's fragments have been read
    tile_store<TA>(As, va);
    tile_store<!TB>(Bs, vb);
    __syncthreads();
    if (k0 + GBK < k_hi) {
      va = tile_load<TA>(A, lda, M, bm, k0 + GBK, k_hi, avec);
      vb = tile_load<!TB>(B, ldb, N, bn, k0 + GBK, k_hi, bvec);
    }
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(As + (wm + li) * GLP + kk + kq);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bs + (wn + li) * GLP + kk + kq);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  }
  float* Cz = C + size_t(blockIdx.z) * zstride;
  const int n = bn + wn + li;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = bm + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (row < M && n < N) Cz[size_t(row) * ldc + n] = acc[r];
  }
}

// C[m][n] = sum over z of part[z][m][n] (z ascending)
__global__ void k_splitk_sum(const float* __restrict__ part, int splits, int M, int N, float* __restrict__ C,
                             int ldc) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(M) * N) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[size_t(z) * M * N + i];
  C[(i / N) * ldc + i % N] = s;
}

// cols[p][(ky*ks + kx)*C + c] = X[b][oy*st - pad + ky][ox*st - pad + kx][c] (0 outside),
// p = (b*Ho + oy)*Wo + ox.  One thread writes 8 consecutive columns of a row
// with one 16-byte store (Kc % 8 == 0); with C % 8 == 0 they are 8 channels
// of one tap, read as two 16-byte loads.
__device__ __forceinline__ float im2col_at(const float* X, int b, int H, int W, int C, int ks, int st, int pad, int oy,
                                           int ox, int q) {
  const int c = q % C, tap = q / C, kx = tap % ks, ky = tap / ks;
  const int iy = oy * st - pad + ky, ix = ox * st - pad + kx;
  return (iy >= 0 && iy < H && ix >= 0 && ix < W) ? X[((size_t(b) * H + iy) * W + ix) * C + c] : 0.f;
}
__global__ void k_im2col(const float* __restrict__ X, int Bn, int H, int W, int C, int ks, int st, int pad, int Ho,
                         int Wo, uint16_t* __restrict__ cols) {
  const int Kc = ks * ks * C, per_row = Kc / 8;
  const size_t total = size_t(Bn) * Ho * Wo * per_row;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    const size_t p = i / per_row;
    const int q0 = int(i % per_row) * 8;
    const int ox = int(p % Wo), oy = int((p / Wo) % Ho), b = int(p / (size_t(Wo) * Ho));
    float v[8];
    if (C % 8 == 0) {
      const int c0 = q0 % C, tap = q0 / C, kx = tap % ks, ky = tap / ks;
      const int iy = oy * st - pad + ky, ix = ox * st - pad + kx;
      if (iy >= 0 && iy < H && ix >= 0 && ix < W) {
        const float4* src = reinterpret_cast<const float4*>(X + ((size_t(b) * H + iy) * W + ix) * C + c0);
        const float4 a = src[0], c = src[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = c.x; v[5] = c.y; v[6] = c.z; v[7] = c.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = 0.f;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = im2col_at(X, b, H, W, C, ks, st, pad, oy, ox, q0 + j);
    }
    uint4 out;
    uint16_t* o = reinterpret_cast<uint16_t*>(&out);
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = to_bf16(v[j]);
    *reinterpret_cast<uint4*>(cols + p * Kc + q0) = out;
  }
}

// dX[b][iy][ix][c] = sum over the taps (ky, kx) that read (iy, ix) of
// dcols[(b, oy, ox)][(ky*ks + kx)*C + c], taps in (ky, kx) order; only the
// taps with ky = (iy + pad) mod st + j*st (likewise kx) can read a pixel.
// One thread per (pixel, 4 channels) with 16-byte loads and stores (C % 4 == 0).
__global__ void k_col2im(const float* __restrict__ dcols, int Bn, int H, int W, int C, int ks, int st, int pad, int Ho,
                         int Wo, float* __restrict__ dX) {
  const int Kc = ks * ks * C, C4 = C / 4;
  const size_t total = size_t(Bn) * H * W * C4;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    const int c = int(i % C4) * 4;
    const size_t pix = i / C4;
    const int ix = int(pix % W), iy = int((pix / W) % H), b = int(pix / (size_t(W) * H));
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int ky = (iy + pad) % st; ky < ks; ky += st) {
      const int oy = (iy + pad - ky) / st;
      if (oy < 0 || oy >= Ho) continue;
      for (int kx = (ix + pad) % st; kx < ks; kx += st) {
        const int ox = (ix + pad - kx) / st;
        if (ox < 0 || ox >= Wo) continue;
        const float4 d =
            *reinterpret_cast<const float4*>(dcols + ((size_t(b) * Ho + oy) * Wo + ox) * Kc + (ky * ks + kx) * C + c);
        s.x += d.x; s.y += d.y; s.z += d.z; s.w += d.w;
      }
    }
    *reinterpret_cast<float4*>(dX + pix * C + c) = s;
  }
}

// A[p][c] = act(Y[p][c] + bias[c]) (act: ReLU or identity); optional FP32
// and BF16 outputs
__global__ void k_bias_act(const float* __restrict__ Y, size_t P, int C, const float* __restrict__ bias, int relu,
                           float* __restrict__ A, uint16_t* __restrict__ Abf) {
  const size_t total = P * C;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    float v = Y[i] + bias[i % C];
    if (relu) v = v > 0.f ? v : 0.f;
    if (A) A[i] = v;
    if (Abf) Abf[i] = to_bf16(v);
  }
}

// dY = dA * relu'(A) (or dA); FP32 and BF16 copies
__global__ void k_act_grad(const float* __restrict__ dA, const float* __restrict__ A, size_t n, int relu,
                           float* __restrict__ dY, uint16_t* __restrict__ dYbf) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float g = (relu && !(A[i] > 0.f)) ? 0.f : dA[i];
    dY[i] = g;
    dYbf[i] = to_bf16(g);
  }
}

// db[c] = sum over p of dY[p][c] in two fixed-order stages.  Stage 1: block g
// sums rows [g*R, (g+1)*R); thread t owns column t % TC (and every TC-th after
// it when C > 256) and rows t / TC + k * (256 / TC), so each pass reads 256
// consecutive floats; row groups are then added in order through LDS.
// Stage 2: part[g][c] summed over g in order.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ dY, size_t P, int C, size_t R,
                                                     float* __restrict__ part) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  const int TC = C < 256 ? C : 256, groups = 256 / TC;
  const int c0 = t % TC, rg = t / TC;
  const size_t r0 = size_t(blockIdx.x) * R, r1 = r0 + R < P ? r0 + R : P;
  for (int cb = 0; cb < C; cb += TC) {
    const int c = cb + c0;
    float s = 0.f;
    if (rg < groups && c < C)
      for (size_t r = r0 + rg; r < r1; r += groups) s += dY[r * C + c];
    red[t] = s;
    __syncthreads();
    if (t < TC) {
      float acc = 0.f;
      for (int g = 0; g < groups; ++g) acc += red[g * TC + t];
      if (cb + t < C) part[size_t(blockIdx.x) * C + cb + t] = acc;
    }
    __syncthreads();
  }
}
// one workgroup per column: thread t sums g = t, t + 256, ... in order, then a fixed tree
__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, int G, int C,
                                                      float* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int g = threadIdx.x; g < G; g += 256) s += part[size_t(g) * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, uint16_t* __restrict__ y) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// mean squared error over n outputs, dY = 2 (y - t) / n; FP64 partial sums
// of (y - t)^2 per block (fixed order: strided per thread, then a tree),
// loss[0] = their ordered sum / n
__global__ __launch_bounds__(256) void k_mse_part(const float* __restrict__ Y, const float* __restrict__ T, size_t n,
                                                  float* __restrict__ dY, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  const float sc = 2.0f / float(n);
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const float d = Y[i] - T[i];
    dY[i] = sc * d;
    s += double(d) * double(d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_mse_final(const double* __restrict__ part, int G, size_t n, double* __restrict__ loss) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part[g];
  loss[0] = s / double(n);
}

// Adam (Kingma & Ba 2015, bias-corrected) on FP32 masters; writes the BF16 copy
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wbf, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = to_bf16(wi);
  }
}

int grid_for(size_t n) { return int(std::min<size_t>((n + 255) / 256, 65536)); }

}  // namespace

extern "C" {

int pty_gemm_bf16(int ta, int tb, int M, int N, int K, const void* A, int lda, const void* B, int ldb, float* C,
                  int ldc, float* partial, size_t partial_elems, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || ldc < N) return PTY_ERR_ARG;
  if (lda < (ta ? M : K) || ldb < (tb ? K : N)) return PTY_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int gx = (N + 63) / 64, gy = (M + 63) / 64;
  // split the reduction when the output has few tiles and the reduction is long
  int splits = 1;
  if (gx * gy < 64 && K >= 4096) splits = std::min(256, std::max(1, K / 2048));
  if (splits > 1) {
    // as many as the partial buffer holds
    const size_t fit = partial ? partial_elems / (size_t(M) * N) : 0;
    splits = int(std::min<size_t>(size_t(splits), fit));
    if (splits < 2) splits = 1;
  }
  const int kchunk = ((K + splits - 1) / splits + 7) / 8 * 8;
  splits = (K + kchunk - 1) / kchunk;
  const dim3 grid(gx, gy, splits);
  const auto* a = static_cast<const uint16_t*>(A);
  const auto* b = static_cast<const uint16_t*>(B);
  float* out = splits > 1 ? partial : C;
  const int ldo = splits > 1 ? N : ldc;
  const size_t zs = splits > 1 ? size_t(M) * N : 0;
  if (ta && tb) k_gemm_bf16<true, true><<<grid, 256, 0, s>>>(M, N, K, a, lda, b, ldb, out, ldo, zs, kchunk);
  else if (ta) k_gemm_bf16<true, false><<<grid, 256, 0, s>>>(M, N, K, a, lda, b, ldb, out, ldo, zs, kchunk);
  else if (tb) k_gemm_bf16<false, true><<<grid, 256, 0, s>>>(M, N, K, a, lda, b, ldb, out, ldo, zs, kchunk);
  else k_gemm_bf16<false, false><<<grid, 256, 0, s>>>(M, N, K, a, lda, b, ldb, out, ldo, zs, kchunk);
  if (splits > 1) k_splitk_sum<<<grid_for(size_t(M) * N), 256, 0, s>>>(partial, splits, M, N, C, ldc);
  return launch_status();
}

int pty_cnn_im2col(const float* X, int batch, int H, int W, int C, int ks, int stride, int pad, void* cols,
                   void* stream) {
  if (batch <= 0 || H <= 0 || W <= 0 || C <= 0 || ks <= 0 || stride <= 0 || pad < 0) return PTY_ERR_ARG;
  const int Ho = (H + 2 * pad - ks) / stride + 1, Wo = (W + 2 * pad - ks) / stride + 1;
  if (Ho <= 0 || Wo <= 0) return PTY_ERR_ARG;
  if ((ks * ks * C) % 8 != 0) return PTY_ERR_ARG;  // 8 columns per thread
  const size_t total = size_t(batch) * Ho * Wo * (ks * ks * C / 8);
  k_im2col<<<grid_for(total), 256, 0, static_cast<hipStream_t>(stream)>>>(X, batch, H, W, C, ks, stride, pad, Ho, Wo,
                                                                          static_cast<uint16_t*>(cols));
  return launch_status();
}

int pty_cnn_col2im(const float* dcols, int batch, int H, int W, int C, int ks, int stride, int pad, float* dX,
                   void* stream) {
  if (batch <= 0 || H <= 0 || W <= 0 || C <= 0 || ks <= 0 || stride <= 0 || pad < 0) return PTY_ERR_ARG;
  const int Ho = (H + 2 * pad - ks) / stride + 1, Wo = (W + 2 * pad - ks) / stride + 1;
  if (Ho <= 0 || Wo <= 0) return PTY_ERR_ARG;
  if (C % 4 != 0) return PTY_ERR_ARG;  // 4 channels per thread
  const size_t total = size_t(batch) * H * W * (C / 4);
  k_col2im<<<grid_for(total), 256, 0, static_cast<hipStream_t>(stream)>>>(dcols, batch, H, W, C, ks, stride, pad, Ho,
                                                                          Wo, dX);
  return launch_status();
}

int pty_cnn_bias_act(const float* Y, int64_t P, int C, const float* bias, int relu, float* A, void* Abf,
                     void* stream) {
  if (P <= 0 || C <= 0 || !Y || !bias) return PTY_ERR_ARG;
  k_bias_act<<<grid_for(size_t(P) * C), 256, 0, static_cast<hipStream_t>(stream)>>>(
      Y, size_t(P), C, bias, relu, A, static_cast<uint16_t*>(Abf));
  return launch_status();
}

int pty_cnn_act_grad(const float* dA, const float* A, int64_t n, int relu, float* dY, void* dYbf, void* stream) {
  if (n <= 0 || !dA || !dY || !dYbf || (relu && !A)) return PTY_ERR_ARG;
  k_act_grad<<<grid_for(size_t(n)), 256, 0, static_cast<hipStream_t>(stream)>>>(dA, A, size_t(n), relu, dY,
                                                                                static_cast<uint16_t*>(dYbf));
  return launch_status();
}

int pty_cnn_bias_grad(const float* dY, int64_t P, int C, float* db, float* partial, size_t partial_elems,
                      void* stream) {
  if (P <= 0 || C <= 0 || !dY || !db || !partial) return PTY_ERR_ARG;
  // about 64 rows per thread-row-group, at most as many blocks as the partials hold
  const int groups = 256 / (C < 256 ? C : 256);
  size_t G = std::min<size_t>((size_t(P) + size_t(groups) * 64 - 1) / (size_t(groups) * 64), 2048);
  G = std::min<size_t>(G, partial_elems / size_t(C));
  if (G < 1) return PTY_ERR_WORKSPACE;
  const size_t R = (size_t(P) + G - 1) / G;
  G = (size_t(P) + R - 1) / R;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_colsum_part<<<int(G), 256, 0, s>>>(dY, size_t(P), C, R, partial);
  k_colsum_final<<<C, 256, 0, s>>>(partial, int(G), C, db);
  return launch_status();
}

int pty_cnn_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  if (n <= 0 || !x || !y) return PTY_ERR_ARG;
  k_to_bf16<<<grid_for(size_t(n)), 256, 0, static_cast<hipStream_t>(stream)>>>(x, size_t(n),
                                                                               static_cast<uint16_t*>(y));
  return launch_status();
}

int pty_cnn_mse(const float* Y, const float* T, int64_t n, float* dY, double* loss, double* partial,
                size_t partial_elems, void* stream) {
  if (n <= 0 || !Y || !T || !dY || !loss || !partial || partial_elems < 1) return PTY_ERR_ARG;
  const int G = int(std::min<size_t>({(size_t(n) + 4095) / 4096, partial_elems, 1024}));
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_mse_part<<<G, 256, 0, s>>>(Y, T, size_t(n), dY, partial);
  k_mse_final<<<1, 64, 0, s>>>(partial, G, size_t(n), loss);
  return launch_status();
}

int pty_cnn_adam(float* w, void* wbf, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                 float beta2, float eps, int step, void* stream) {
  if (n <= 0 || !w || !g || !m || !v || step < 1) return PTY_ERR_ARG;
  const float c1 = 1.f - powf(beta1, float(step)), c2 = 1.f - powf(beta2, float(step));
  k_adam<<<grid_for(size_t(n)), 256, 0, static_cast<hipStream_t>(stream)>>>(
      w, static_cast<uint16_t*>(wbf), g, m, v, size_t(n), lr, beta1, beta2, eps, c1, c2);
  return launch_status();
}

}  // extern "C"
