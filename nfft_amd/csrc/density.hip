// Density-compensation weights for the weighted CGNR: the W that
// solver_init_advanced_complex uses with PRECOMPUTE_WEIGHT (reference
// kernel/solver/solver.c:40-125, 232-292).  The reference computes no weights
// on this path (its MRI applications read them from files made by Matlab
// scripts, applications/mri/mri2d/precompute_weights.m); this file adds two
// ways to get them on the GPU.
//
// 1. Pipe-Menon fixed point (Pipe & Menon, Magn. Reson. Med. 41:179, 1999):
//      w <- w / (C w),   (C w)_j = sum_l w_l sum_g psi(x_j - g) psi(x_l - g),
//    psi the plan's Kaiser-Bessel window on the oversampled grid, so C = B B^T
//    with the plan's own interpolation B and spreading B^T.
//
// 2. A learned model: an MLP (80 -> 64 -> 64 -> 1, ReLU) maps node-density
//    features to log w.  Its two dense layers run on BF16 MFMA
//    (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, forward and
//    backward; training is Adam on FP32 master weights.  The output layer,
//    the loss and the optimiser are FP32 vector code.
//
//    Features of node j (d = 2), 77 values padded with zeros to 80: the
//    radius 2|x_j|, then for the coarse grids G = 32, 64, 128 cells per
//    dimension log(1 + c / cbar) of the node counts c in the 5 x 5 cells
//    around the node's cell (periodic), cbar = M / G^2, and last the
//    node's kernel density estimate log(v_j / mean v) with v = B B^T 1, the
//    first Pipe-Menon denominator (one spreading and one interpolation of
//    the plan).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "engine.hpp"
#include "nfft3_hip.h"

namespace nfft_hip {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kDcfFeat = 80;   // padded feature count (multiple of 16)
constexpr int kDcfRaw = 77;    // 1 + 3 * 25 + 1
constexpr int kDcfHidden = 64;
constexpr int kDcfScales = 3;
__constant__ int c_dcf_grid[kDcfScales] = {32, 64, 128};

// flat FP32 parameter layout; weights [in][out]
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + kDcfFeat * kDcfHidden;
constexpr int P_W2 = P_B1 + kDcfHidden;
constexpr int P_B2 = P_W2 + kDcfHidden * kDcfHidden;
constexpr int P_W3 = P_B2 + kDcfHidden;
constexpr int P_B3 = P_W3 + kDcfHidden;
constexpr int P_COUNT = P_B3 + 1;

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------ Pipe-Menon
__global__ void k_real_to_cplx(const float *__restrict__ w, float2 *__restrict__ out, long M) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) out[j] = float2{w[j], 0.f};
}

__global__ void k_pm_update(float *__restrict__ w, const float2 *__restrict__ v, long M) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const float c = v[j].x;
  w[j] = c > 0.f ? w[j] / c : w[j];
}

static unsigned nblk(long n, int t = 256) { return (unsigned)std::max<long>(1, ceil_div(n, t)); }

// Rows per chunk of learned-weight inference (features and MLP workspace are
// sized for one chunk, not for M).
constexpr long kDcfChunk = 65536;

// The helpers' device scratch, carved from the plan's aux buffer: optional
// features of one chunk (kDcfChunk x 80 floats), coarse-grid counts, two
// complex node vectors and the FP64 partial sums of a mean.
struct DcfScratch {
  float *feat = nullptr;
  int *counts = nullptr;
  float2 *fa = nullptr, *fb = nullptr;
  double *part = nullptr;
};
constexpr int kDcfCells = 32 * 32 + 64 * 64 + 128 * 128;

static DcfScratch dcf_scratch(Engine<float> *e, bool with_feat) {
  const long M = e->M;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_feat = with_feat ? al(sizeof(float) * std::min(M, kDcfChunk) * kDcfFeat) : 0;
  const size_t b_counts = al(sizeof(int) * kDcfCells);
  const size_t b_vec = al(sizeof(float2) * M);
  const size_t b_part = al(sizeof(double) * (ceil_div(M, 256) + 1));
  const size_t need = b_feat + b_counts + 2 * b_vec + b_part;
  if (need > e->aux_cap) {
    NFFT_HIP_CHECK(float, hipStreamSynchronize(e->stream));
    hipFree(e->aux);
    e->aux = nullptr;
    NFFT_HIP_CHECK(float, hipMalloc(&e->aux, need));
    e->aux_cap = need;
  }
  char *p = e->aux;
  DcfScratch sc;
  if (with_feat) sc.feat = reinterpret_cast<float *>(p);
  p += b_feat;
  sc.counts = reinterpret_cast<int *>(p);
  p += b_counts;
  sc.fa = reinterpret_cast<float2 *>(p);
  p += b_vec;
  sc.fb = reinterpret_cast<float2 *>(p);
  p += b_vec;
  sc.part = reinterpret_cast<double *>(p);
  return sc;
}

static void pipe_menon_device(Engine<float> *e, float *w, int iters) {
  const long M = e->M;
  if (M == 0) return;
  hipStream_t s = e->stream;
  const DcfScratch sc = dcf_scratch(e, false);
  float2 *fa = sc.fa, *fb = sc.fb;
  for (int it = 0; it < iters; it++) {
    hipLaunchKernelGGL(k_real_to_cplx, dim3(nblk(M)), dim3(256), 0, s, w, fa, M);
    e->spread_to_grid(fa);
    e->interp_from_grid(fb);
    hipLaunchKernelGGL(k_pm_update, dim3(nblk(M)), dim3(256), 0, s, w, fb, M);
  }
  NFFT_HIP_CHECK(float, hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ features
__device__ __forceinline__ int coarse_cell(float x, int G) {
  int c = (int)floorf((x + 0.5f) * (float)G);
  return c < 0 ? 0 : (c >= G ? G - 1 : c);
}

// Coarse-grid node counts.  Dense trajectories put thousands of nodes in one
// coarse cell, so the two coarse grids (32^2 + 64^2 cells) are counted in a
// per-workgroup LDS histogram first (one global add per touched cell and
// workgroup); the 128^2 grid, whose cells see far fewer nodes, directly.
constexpr int kDcfHistNodes = 4096;  // nodes per workgroup
__global__ __launch_bounds__(256) void k_dcf_hist(const float *__restrict__ x, long M,
                                                  int *__restrict__ counts) {
  constexpr int L = 32 * 32 + 64 * 64;
  __shared__ int h[L];
  for (int q = threadIdx.x; q < L; q += 256) h[q] = 0;
  __syncthreads();
  const long j0 = (long)blockIdx.x * kDcfHistNodes;
  const long j1 = min(M, j0 + kDcfHistNodes);
  for (long j = j0 + threadIdx.x; j < j1; j += 256) {
    const float x0 = x[2 * j], x1 = x[2 * j + 1];
    atomicAdd(&h[coarse_cell(x0, 32) * 32 + coarse_cell(x1, 32)], 1);
    atomicAdd(&h[1024 + coarse_cell(x0, 64) * 64 + coarse_cell(x1, 64)], 1);
    atomicAdd(&counts[L + coarse_cell(x0, 128) * 128 + coarse_cell(x1, 128)], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < L; q += 256)
    if (h[q]) atomicAdd(&counts[q], h[q]);
}

// rows j0 .. j0+n-1 of the features into feat[0 .. n-1], one thread per
// (row, feature) so that the row-major stores are coalesced; the density
// estimate column from v = Re(B B^T 1) and its sum *total
__global__ __launch_bounds__(256) void k_dcf_features_flat(const float *__restrict__ x, long M,
                                                           long j0, long n,
                                                           const int *__restrict__ counts,
                                                           const float2 *__restrict__ v,
                                                           const double *__restrict__ total,
                                                           float *__restrict__ feat) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kDcfFeat) return;
  const long r = i / kDcfFeat;
  const int q = (int)(i - r * kDcfFeat);
  const long j = j0 + r;
  const float x0 = x[2 * j], x1 = x[2 * j + 1];
  float val = 0.f;
  if (q == 0) {
    val = 2.f * sqrtf(x0 * x0 + x1 * x1);
  } else if (q < 1 + kDcfScales * 25) {
    const int s = (q - 1) / 25, ab = (q - 1) - s * 25;
    const int G = c_dcf_grid[s];
    const int off = s == 0 ? 0 : (s == 1 ? 32 * 32 : 32 * 32 + 64 * 64);
    const int ux = (coarse_cell(x0, G) + ab / 5 - 2 + G) % G;
    const int uy = (coarse_cell(x1, G) + ab % 5 - 2 + G) % G;
    val = log1pf((float)counts[off + ux * G + uy] * ((float)G * (float)G / (float)M));
  } else if (q == kDcfRaw - 1) {
    const float mean = (float)(total[0] / (double)M);
    val = logf(fmaxf(v[j].x, 1e-30f) / mean);
  }
  feat[i] = val;
}

__global__ void k_fill_unit(float2 *__restrict__ out, long M) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) out[j] = float2{1.f, 0.f};
}

// FP64 partial sums of Re v over blocks of 256 nodes
__global__ __launch_bounds__(256) void k_sum_re(const float2 *__restrict__ v, long M,
                                                double *__restrict__ part) {
  __shared__ double red[256];
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  red[threadIdx.x] = j < M ? (double)v[j].x : 0.0;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_parts(const double *__restrict__ part, long nb,
                                                   double *__restrict__ total) {
  __shared__ double red[256];
  double a = 0.0;
  for (long i = threadIdx.x; i < nb; i += 256) a += part[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[0] = red[0];
}

// The parts of the features shared by all rows, enqueued on the plan's
// stream: coarse-grid counts, and the density estimate v = Re(B B^T 1) with
// its FP64 sum at part[nb]
static void features_common(Engine<float> *e, const DcfScratch &sc) {
  const long M = e->M;
  hipStream_t s = e->stream;
  const long nb = ceil_div(M, 256);
  NFFT_HIP_CHECK(float, hipMemsetAsync(sc.counts, 0, sizeof(int) * kDcfCells, s));
  hipLaunchKernelGGL(k_dcf_hist, dim3((unsigned)ceil_div(M, kDcfHistNodes)), dim3(256), 0, s,
                     e->xdev, M, sc.counts);
  hipLaunchKernelGGL(k_fill_unit, dim3(nblk(M)), dim3(256), 0, s, sc.fa, M);
  e->spread_to_grid(sc.fa);
  e->interp_from_grid(sc.fb);
  hipLaunchKernelGGL(k_sum_re, dim3((unsigned)nb), dim3(256), 0, s, sc.fb, M, sc.part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, sc.part, nb, sc.part + nb);
}

// feature rows j0 .. j0+n-1 into feat, after features_common
static void features_rows(Engine<float> *e, const DcfScratch &sc, long j0, long n, float *feat) {
  const long M = e->M;
  hipStream_t s = e->stream;
  hipLaunchKernelGGL(k_dcf_features_flat, dim3(nblk(n * kDcfFeat)), dim3(256), 0, s, e->xdev, M,
                     j0, n, sc.counts, sc.fb, sc.part + ceil_div(M, 256), feat);
}

// all M feature rows into the caller's feat
static void features_device(Engine<float> *e, float *feat) {
  if (e->d != 2) family<float>::fatal("nfft: density features are defined for d = 2 plans");
  const long M = e->M;
  if (M == 0) return;
  const DcfScratch sc = dcf_scratch(e, false);
  features_common(e, sc);
  features_rows(e, sc, 0, M, feat);
  NFFT_HIP_CHECK(float, hipStreamSynchronize(e->stream));
}

// ------------------------------------------------------------------ MLP kernels
__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// BF16 copies of the two MFMA layers: forward operands W^T ([out][in]) and
// the backward operand of layer 2 in [in][out] order.
__global__ void k_weights_bf16(const float *__restrict__ P, uint16_t *__restrict__ w1t,
                               uint16_t *__restrict__ w2t, uint16_t *__restrict__ w2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kDcfFeat * kDcfHidden) {
    const int in = i / kDcfHidden, out = i - in * kDcfHidden;
    w1t[out * kDcfFeat + in] = f2bf(P[P_W1 + i]);
  }
  if (i < kDcfHidden * kDcfHidden) {
    const int in = i / kDcfHidden, out = i - in * kDcfHidden;
    const uint16_t b = f2bf(P[P_W2 + i]);
    w2t[out * kDcfHidden + in] = b;
    w2[i] = b;
  }
}

enum { kEpiBiasRelu = 0, kEpiMask = 1 };

// out[r][n] = epi(sum_k A[r][k] Bt[n][k]) for 64 output columns, BF16 in,
// FP32 accumulate, BF16 out.  One wavefront computes 16 rows with four
// 16 x 16 x 16 MFMA tiles per K step; operand lanes: lane l holds
// A[l % 16][4 (l / 16) + 0..3] and B[4 (l / 16) + 0..3][l % 16]; the result
// lane holds D[4 (l / 16) + 0..3][l % 16].  K is a multiple of 16.
template <int EPI>
__global__ __launch_bounds__(256) void k_dense_mfma(const uint16_t *__restrict__ A,
                                                    const uint16_t *__restrict__ Bt,
                                                    const float *__restrict__ bias,
                                                    const uint16_t *__restrict__ mask,
                                                    uint16_t *__restrict__ out, int rows, int K) {
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (row0 >= rows) return;
  const int lr = lane & 15, kq = 4 * (lane >> 4);
  const int ar = row0 + lr;
  floatx4_t acc[4];
#pragma unroll
  for (int n = 0; n < 4; n++) acc[n] = floatx4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a = {0, 0, 0, 0};
    if (ar < rows) a = *reinterpret_cast<const bf16x4_t *>(A + (long)ar * K + k0 + kq);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const bf16x4_t b = *reinterpret_cast<const bf16x4_t *>(Bt + (n * 16 + lr) * K + k0 + kq);
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[n], 0, 0, 0);
    }
  }
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const int col = n * 16 + lr;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = row0 + kq + r;
      if (row >= rows) continue;
      float v = acc[n][r];
      if (EPI == kEpiBiasRelu) {
        v += bias[col];
        v = v > 0.f ? v : 0.f;
      } else {
        v = bf2f(mask[(long)row * kDcfHidden + col]) > 0.f ? v : 0.f;
      }
      out[(long)row * kDcfHidden + col] = f2bf(v);
    }
  }
}

// Output layer, loss and its gradient.  out[r] = H2[r] . w3 + b3; with a
// target: err = out - t, loss += err^2 / rows, dout = 2 err / rows,
// dH2[r][j] = (H2[r][j] > 0) dout w3[j]; per block partial sums of loss,
// dw3[j] = sum dout H2[r][j] and db3 = sum dout go to part[blk][0..65].
__global__ __launch_bounds__(256) void k_head(const uint16_t *__restrict__ H2,
                                              const float *__restrict__ P,
                                              const float *__restrict__ target, int rows,
                                              float *__restrict__ out, uint16_t *__restrict__ dH2,
                                              float *__restrict__ part) {
  __shared__ float s_w3[kDcfHidden];
  __shared__ float s_red[kDcfHidden + 2][4];
  if (threadIdx.x < kDcfHidden) s_w3[threadIdx.x] = P[P_W3 + threadIdx.x];
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  float h[kDcfHidden];
  float o = P[P_B3];
  const bool live = r < rows;
#pragma unroll
  for (int j = 0; j < kDcfHidden; j++) {
    h[j] = live ? bf2f(H2[(long)r * kDcfHidden + j]) : 0.f;
    o += h[j] * s_w3[j];
  }
  if (live && out) out[r] = o;
  if (!target) return;
  const float err = live ? o - target[r] : 0.f;
  const float dout = 2.f * err / (float)rows;
  if (live) {
#pragma unroll
    for (int j = 0; j < kDcfHidden; j++)
      dH2[(long)r * kDcfHidden + j] = f2bf(h[j] > 0.f ? dout * s_w3[j] : 0.f);
  }
  // block sums: loss, dw3[0..63], db3
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  auto wsum = [](float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  };
  float v = wsum(err * err / (float)rows);
  if (lane == 0) s_red[0][wv] = v;
  v = wsum(dout);
  if (lane == 0) s_red[1][wv] = v;
#pragma unroll
  for (int j = 0; j < kDcfHidden; j++) {
    v = wsum(dout * h[j]);
    if (lane == 0) s_red[2 + j][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDcfHidden + 2) {
    const int q = threadIdx.x;
    part[(long)blockIdx.x * (kDcfHidden + 2) + q] =
        s_red[q][0] + s_red[q][1] + s_red[q][2] + s_red[q][3];
  }
}

// Weight gradient of a dense layer: dW[i][j] = sum_r A[r][i] G[r][j] and
// db[j] = sum_r G[r][j] over the block's rows, as partial sums
// part[blk][0 .. KA*64) and part[blk][KA*64 .. KA*64+64).  Rows are staged
// 64 at a time in LDS; the MFMA contracts over rows (A^T as the first
// operand).
constexpr int kWgradRows = 256;
__global__ __launch_bounds__(256) void k_wgrad_mfma(const uint16_t *__restrict__ A,
                                                    const uint16_t *__restrict__ G, int rows,
                                                    int KA, float *__restrict__ part) {
  __shared__ uint16_t s_a[64][kDcfFeat];
  __shared__ uint16_t s_g[64][kDcfHidden];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lr = lane & 15, kq = 4 * (lane >> 4);
  const int mt = KA / 16;            // tiles along i
  const int ntile = mt * 4;          // (i tile, j tile) pairs
  floatx4_t acc[5];
#pragma unroll
  for (int t = 0; t < 5; t++) acc[t] = floatx4_t{0.f, 0.f, 0.f, 0.f};
  float dbias = 0.f;
  const int rbeg = blockIdx.x * kWgradRows;
  const int rend = min(rows, rbeg + kWgradRows);
  for (int r0 = rbeg; r0 < rend; r0 += 64) {
    for (int q = threadIdx.x; q < 64 * KA; q += blockDim.x) {
      const int rr = q / KA, i = q - rr * KA;
      s_a[rr][i] = r0 + rr < rend ? A[(long)(r0 + rr) * KA + i] : 0;
    }
    for (int q = threadIdx.x; q < 64 * kDcfHidden; q += blockDim.x) {
      const int rr = q / kDcfHidden, j = q - rr * kDcfHidden;
      s_g[rr][j] = r0 + rr < rend ? G[(long)(r0 + rr) * kDcfHidden + j] : 0;
    }
    __syncthreads();
    if (threadIdx.x < kDcfHidden)
      for (int rr = 0; rr < 64; rr++) dbias += bf2f(s_g[rr][threadIdx.x]);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
#pragma unroll
      for (int t = 0; t < 5; t++) {
        const int tile = wv + 4 * t;
        if (tile < ntile) {
          const int mi = tile >> 2, nj = tile & 3;
          bf16x4_t a, b;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            a[r] = (short)s_a[kk * 16 + kq + r][mi * 16 + lr];
            b[r] = (short)s_g[kk * 16 + kq + r][nj * 16 + lr];
          }
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const long base = (long)blockIdx.x * (KA * kDcfHidden + kDcfHidden);
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const int tile = wv + 4 * t;
    if (tile < ntile) {
      const int mi = tile >> 2, nj = tile & 3;
#pragma unroll
      for (int r = 0; r < 4; r++)
        part[base + (long)(mi * 16 + kq + r) * kDcfHidden + nj * 16 + lr] = acc[t][r];
    }
  }
  if (threadIdx.x < kDcfHidden) part[base + KA * kDcfHidden + threadIdx.x] = dbias;
}

// out[dst_off + i] = sum over blocks of part[blk][i], i < len
__global__ void k_sum_partials(const float *__restrict__ part, int nb, int len,
                               float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int b = 0; b < nb; b++) s += part[(long)b * len + i];
  out[i] = s;
}

// head sums [loss, db3, dw3[0..63]] -> gradient entries of w3 and b3
__global__ void k_head_grads(const float *__restrict__ sum_head, float *__restrict__ grad) {
  const int j = threadIdx.x;
  if (j < kDcfHidden) grad[P_W3 + j] = sum_head[2 + j];
  if (j == 0) grad[P_B3] = sum_head[1];
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_exp(const float *__restrict__ in, float *__restrict__ out, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = expf(in[i]);
}

// ------------------------------------------------------------------ model
struct DcfModel {
  hipStream_t stream = nullptr;
  float *P = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *w1t = nullptr, *w2t = nullptr, *w2 = nullptr;
  int cap = 0;  // rows the workspace holds
  uint16_t *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dH2 = nullptr, *dH1 = nullptr;
  float *out = nullptr, *part = nullptr, *sum_head = nullptr;
  long step = 0;

  void ensure(int rows) {
    if (rows <= cap) return;
    hipFree(X);
    hipFree(H1);
    hipFree(H2);
    hipFree(dH2);
    hipFree(dH1);
    hipFree(out);
    hipFree(part);
    const long r = rows;
    NFFT_HIP_CHECK(float, hipMalloc(&X, sizeof(uint16_t) * r * kDcfFeat));
    NFFT_HIP_CHECK(float, hipMalloc(&H1, sizeof(uint16_t) * r * kDcfHidden));
    NFFT_HIP_CHECK(float, hipMalloc(&H2, sizeof(uint16_t) * r * kDcfHidden));
    NFFT_HIP_CHECK(float, hipMalloc(&dH2, sizeof(uint16_t) * r * kDcfHidden));
    NFFT_HIP_CHECK(float, hipMalloc(&dH1, sizeof(uint16_t) * r * kDcfHidden));
    NFFT_HIP_CHECK(float, hipMalloc(&out, sizeof(float) * r));
    const long nbw = ceil_div(r, kWgradRows);
    const long nbh = ceil_div(r, 256);
    const long plen = std::max<long>(nbw * (kDcfFeat * kDcfHidden + kDcfHidden),
                                     nbh * (kDcfHidden + 2));
    NFFT_HIP_CHECK(float, hipMalloc(&part, sizeof(float) * plen));
    cap = rows;
  }

  void upload_bf16() {
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(kDcfFeat * kDcfHidden)), dim3(256), 0, stream,
                       P, w1t, w2t, w2);
  }

  // forward of `rows` feature rows (FP32, row-major [rows][80]); with a
  // target also the loss and all gradients (into grad)
  float forward(const float *feat, const float *target, int rows, float *out_dev) {
    ensure(rows);
    hipStream_t s = stream;
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long)rows * kDcfFeat)), dim3(256), 0, s, feat, X,
                       (long)rows * kDcfFeat);
    const unsigned nbd = (unsigned)ceil_div(rows, 64);
    hipLaunchKernelGGL(k_dense_mfma<kEpiBiasRelu>, dim3(nbd), dim3(256), 0, s, X, w1t, P + P_B1,
                       (const uint16_t *)nullptr, H1, rows, kDcfFeat);
    hipLaunchKernelGGL(k_dense_mfma<kEpiBiasRelu>, dim3(nbd), dim3(256), 0, s, H1, w2t, P + P_B2,
                       (const uint16_t *)nullptr, H2, rows, kDcfHidden);
    const unsigned nbh = (unsigned)ceil_div(rows, 256);
    hipLaunchKernelGGL(k_head, dim3(nbh), dim3(256), 0, s, H2, P, target, rows,
                       out_dev ? out_dev : out, dH2, part);
    if (!target) return 0.f;
    // head partials -> loss, db3 and dw3 (grad[P_W3..], grad[P_B3])
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(128), 0, s, part, (int)nbh, kDcfHidden + 2,
                       sum_head);
    // backward through layer 2 (dH1 = (dH2 W2^T) * [H1 > 0]) and the weight
    // gradients of both MFMA layers
    hipLaunchKernelGGL(k_dense_mfma<kEpiMask>, dim3(nbd), dim3(256), 0, s, dH2, w2,
                       (const float *)nullptr, H1, dH1, rows, kDcfHidden);
    const unsigned nbw = (unsigned)ceil_div(rows, kWgradRows);
    const int len2 = kDcfHidden * kDcfHidden + kDcfHidden;
    hipLaunchKernelGGL(k_wgrad_mfma, dim3(nbw), dim3(256), 0, s, H1, dH2, rows, kDcfHidden, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nblk(len2)), dim3(256), 0, s, part, (int)nbw, len2,
                       grad + P_W2);
    const int len1 = kDcfFeat * kDcfHidden + kDcfHidden;
    hipLaunchKernelGGL(k_wgrad_mfma, dim3(nbw), dim3(256), 0, s, X, dH1, rows, kDcfFeat, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nblk(len1)), dim3(256), 0, s, part, (int)nbw, len1,
                       grad + P_W1);
    hipLaunchKernelGGL(k_head_grads, dim3(1), dim3(64), 0, s, sum_head, grad);
    float loss = 0.f;
    NFFT_HIP_CHECK(float, hipMemcpyAsync(&loss, sum_head, sizeof(float), hipMemcpyDeviceToHost, s));
    NFFT_HIP_CHECK(float, hipStreamSynchronize(s));
    return loss;
  }
};

}  // namespace nfft_hip

using nfft_hip::DcfModel;

struct nfft_dcf_model_s {
  DcfModel m;
};

extern "C" {

void nfftf_density_pipe_menon_device(nfftf_plan *ths, float *w_dev, int iters) {
  nfft_hip::clear_launch_error();
  auto *e = nfft_hip::engine_of_mv<float>(ths);
  if (!e) nfft_hip::family<float>::fatal("nfft: density weights need a plan of this library");
  nfft_hip::ensure_nodes_of_mv<float>(ths);
  nfft_hip::pipe_menon_device(e, w_dev, iters);
}

void nfftf_density_pipe_menon(nfftf_plan *ths, float *w, int iters) {
  auto *e = nfft_hip::engine_of_mv<float>(ths);
  if (!e) nfft_hip::family<float>::fatal("nfft: density weights need a plan of this library");
  const long M = ths->M_total;
  if (M == 0) return;
  float *wd = nullptr;
  NFFT_HIP_CHECK(float, hipMalloc(&wd, sizeof(float) * M));
  NFFT_HIP_CHECK(float, hipMemcpy(wd, w, sizeof(float) * M, hipMemcpyHostToDevice));
  nfftf_density_pipe_menon_device(ths, wd, iters);
  NFFT_HIP_CHECK(float, hipMemcpy(w, wd, sizeof(float) * M, hipMemcpyDeviceToHost));
  hipFree(wd);
}

int nfft_dcf_num_features(void) { return nfft_hip::kDcfFeat; }
int nfft_dcf_num_params(void) { return nfft_hip::P_COUNT; }

void nfftf_dcf_features_device(nfftf_plan *ths, float *feat_dev) {
  nfft_hip::clear_launch_error();
  auto *e = nfft_hip::engine_of_mv<float>(ths);
  if (!e) nfft_hip::family<float>::fatal("nfft: density features need a plan of this library");
  nfft_hip::ensure_nodes_of_mv<float>(ths);
  nfft_hip::features_device(e, feat_dev);
}

nfft_dcf_model *nfft_dcf_model_create(unsigned seed) {
  using namespace nfft_hip;
  auto *h = new nfft_dcf_model_s();
  DcfModel &m = h->m;
  NFFT_HIP_CHECK(float, hipStreamCreate(&m.stream));
  NFFT_HIP_CHECK(float, hipMalloc(&m.P, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMalloc(&m.grad, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMalloc(&m.m, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMalloc(&m.v, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMalloc(&m.sum_head, sizeof(float) * (kDcfHidden + 2)));
  NFFT_HIP_CHECK(float, hipMalloc(&m.w1t, sizeof(uint16_t) * kDcfFeat * kDcfHidden));
  NFFT_HIP_CHECK(float, hipMalloc(&m.w2t, sizeof(uint16_t) * kDcfHidden * kDcfHidden));
  NFFT_HIP_CHECK(float, hipMalloc(&m.w2, sizeof(uint16_t) * kDcfHidden * kDcfHidden));
  NFFT_HIP_CHECK(float, hipMemset(m.grad, 0, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMemset(m.m, 0, sizeof(float) * P_COUNT));
  NFFT_HIP_CHECK(float, hipMemset(m.v, 0, sizeof(float) * P_COUNT));
  // workspace of one inference chunk (training batches up to that size reuse
  // it), so that inference allocates nothing after creation
  m.ensure((int)kDcfChunk);
  // He initialisation of the weights, zero biases
  std::vector<float> p(P_COUNT, 0.f);
  std::mt19937 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int i = 0; i < kDcfRaw * kDcfHidden; i++)
    p[P_W1 + i] = nd(rng) * std::sqrt(2.f / kDcfRaw);
  for (int i = 0; i < kDcfHidden * kDcfHidden; i++)
    p[P_W2 + i] = nd(rng) * std::sqrt(2.f / kDcfHidden);
  for (int i = 0; i < kDcfHidden; i++) p[P_W3 + i] = nd(rng) * std::sqrt(1.f / kDcfHidden);
  NFFT_HIP_CHECK(float, hipMemcpy(m.P, p.data(), sizeof(float) * P_COUNT, hipMemcpyHostToDevice));
  m.upload_bf16();
  NFFT_HIP_CHECK(float, hipStreamSynchronize(m.stream));
  return h;
}

void nfft_dcf_model_destroy(nfft_dcf_model *h) {
  if (!h) return;
  DcfModel &m = h->m;
  hipStreamSynchronize(m.stream);
  for (void *q : {(void *)m.P, (void *)m.grad, (void *)m.m, (void *)m.v, (void *)m.w1t,
                  (void *)m.w2t, (void *)m.w2, (void *)m.X, (void *)m.H1, (void *)m.H2,
                  (void *)m.dH2, (void *)m.dH1, (void *)m.out, (void *)m.part,
                  (void *)m.sum_head})
    if (q) hipFree(q);
  hipStreamDestroy(m.stream);
  delete h;
}

void nfft_dcf_model_get_params(const nfft_dcf_model *h, float *params) {
  NFFT_HIP_CHECK(float, hipStreamSynchronize(h->m.stream));
  NFFT_HIP_CHECK(float, hipMemcpy(params, h->m.P, sizeof(float) * nfft_hip::P_COUNT,
                                  hipMemcpyDeviceToHost));
}

void nfft_dcf_model_set_params(nfft_dcf_model *h, const float *params) {
  DcfModel &m = h->m;
  NFFT_HIP_CHECK(float, hipStreamSynchronize(m.stream));
  NFFT_HIP_CHECK(float, hipMemcpy(m.P, params, sizeof(float) * nfft_hip::P_COUNT,
                                  hipMemcpyHostToDevice));
  m.upload_bf16();
  NFFT_HIP_CHECK(float, hipStreamSynchronize(m.stream));
}

float nfft_dcf_model_loss_grad(nfft_dcf_model *h, const float *feat_dev, const float *target_dev,
                               int rows, float *grad_host) {
  nfft_hip::clear_launch_error();
  DcfModel &m = h->m;
  const float loss = m.forward(feat_dev, target_dev, rows, nullptr);
  if (grad_host)
    NFFT_HIP_CHECK(float, hipMemcpy(grad_host, m.grad, sizeof(float) * nfft_hip::P_COUNT,
                                    hipMemcpyDeviceToHost));
  return loss;
}

float nfft_dcf_model_train_step(nfft_dcf_model *h, const float *feat_dev, const float *target_dev,
                                int rows, float lr) {
  nfft_hip::clear_launch_error();
  using namespace nfft_hip;
  DcfModel &m = h->m;
  const float loss = m.forward(feat_dev, target_dev, rows, nullptr);
  m.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)m.step));
  const float c2 = 1.f / (1.f - std::pow(b2, (float)m.step));
  hipLaunchKernelGGL(k_adam, dim3(nblk(P_COUNT)), dim3(256), 0, m.stream, m.P, m.grad, m.m, m.v,
                     (int)P_COUNT, lr, b1, b2, 1e-8f, c1, c2);
  m.upload_bf16();
  NFFT_HIP_CHECK(float, hipGetLastError());
  return loss;
}

void nfft_dcf_model_predict(nfft_dcf_model *h, const float *feat_dev, int rows, float *out_dev) {
  using namespace nfft_hip;
  clear_launch_error();
  DcfModel &m = h->m;
  // in chunks, so the workspace stays at kDcfChunk rows
  for (long r0 = 0; r0 < rows; r0 += kDcfChunk) {
    const int n = (int)std::min<long>(kDcfChunk, rows - r0);
    m.forward(feat_dev + r0 * kDcfFeat, nullptr, n, out_dev + r0);
  }
  NFFT_HIP_CHECK(float, hipStreamSynchronize(m.stream));
}

void nfftf_density_learned_device(nfftf_plan *ths, nfft_dcf_model *h, float *w_dev) {
  using namespace nfft_hip;
  const long M = ths->M_total;
  if (M == 0) return;
  clear_launch_error();
  auto *e = engine_of_mv<float>(ths);
  if (!e) family<float>::fatal("nfft: density weights need a plan of this library");
  ensure_nodes_of_mv<float>(ths);
  if (e->d != 2) family<float>::fatal("nfft: density features are defined for d = 2 plans");
  // features chunk by chunk in the plan's scratch, each chunk's inference
  // right behind it on the plan's stream (the model's kernels follow the
  // plan's work; calls on one model must not overlap)
  const DcfScratch sc = dcf_scratch(e, true);
  features_common(e, sc);
  DcfModel &m = h->m;
  const hipStream_t own = m.stream;
  m.stream = e->stream;
  for (long j0 = 0; j0 < M; j0 += kDcfChunk) {
    const long n = std::min(kDcfChunk, M - j0);
    features_rows(e, sc, j0, n, sc.feat);
    m.forward(sc.feat, nullptr, (int)n, w_dev + j0);
  }
  hipLaunchKernelGGL(k_exp, dim3(nblk(M)), dim3(256), 0, e->stream, w_dev, w_dev, M);
  m.stream = own;
  NFFT_HIP_CHECK(float, hipStreamSynchronize(e->stream));
}

}  // extern "C"
