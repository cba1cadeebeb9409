// B (interpolation, trafo) and B^T (spreading, adjoint) for d = 1, 2, 3.
//
//   B:   f_j = sum_{l} prod_t psi_t(l_t) g[l mod n]
//        (reference kernel/nfft/nfft.c:2131-2153, 2927-3004, 4020-4265)
//   B^T: g[l mod n] += prod_t psi_t(l_t) f_j
//        (reference kernel/nfft/nfft.c:2585-2773, 3583-3805, 5126-5384)
// with taps l_t = c_t - m + w, w = 0 .. 2m+1, c_t = floor(n_t x_t), and
// psi_t(l_t) = phi(n_t x_t - l_t) (kernel/nfft/nfft.c:324-340, 5819-5844).
//
// One workgroup per work item (bin, first node, count).  Tiled variant: the
// bin's grid region, bs_t + 2m + 1 cells per dimension, is staged in LDS.
// Interpolation reads taps from LDS; spreading accumulates into LDS with
// ds_add, then adds the region to the grid with one atomic per touched cell.
// Global variant (region above the LDS budget): taps go straight to HBM
// through the caches.
//
// Two instantiation families: W = 2m+2 as a template parameter for
// m = 1..kMaxStaticM (tiled only; registers hold all window values), and a
// runtime-m family for m = 0..kMaxM (tiled or global) covering the rest.
//
// Roofline: the tap loops are bound by LDS bandwidth and VALU, not HBM.
// Algorithmic HBM bytes per launch: n_total c + M d r + M c (B and B^T alike).
#pragma once

#include "engine.hpp"
#include "window.hpp"

namespace nfft_hip {

constexpr int kMaxStaticM = 8;
constexpr int kMaxM = 16;
constexpr int kMaxW = 2 * kMaxM + 2;

__device__ __forceinline__ int wrap_once(int v, int n) {
  v += (v < 0) ? n : 0;
  v -= (v >= n) ? n : 0;
  return v;
}

// bin -> lower corner (in cells) of the bin
template <int D>
__device__ __forceinline__ void bin_corner(int bin, const BinGeom &gg, int (&lo)[D]) {
#pragma unroll
  for (int t = D - 1; t >= 0; t--) {
    lo[t] = (bin % gg.nb[t]) * gg.bs[t];
    bin /= gg.nb[t];
  }
}

template <int D>
__device__ __forceinline__ long region_to_grid(int q, const int (&lo)[D], const BinGeom &gg) {
  int e[D];
#pragma unroll
  for (int t = D - 1; t >= 0; t--) {
    e[t] = q % gg.E[t];
    q /= gg.E[t];
  }
  // the region may reach past one period when the last bin is partial
  long gi = 0;
#pragma unroll
  for (int t = 0; t < D; t++) {
    int v = (lo[t] - gg.m + e[t]) % gg.n[t];
    v += v < 0 ? gg.n[t] : 0;
    gi = gi * gg.n[t] + v;
  }
  return gi;
}

template <typename R>
__device__ __forceinline__ void lds_add(cplx<R> *a, R vr, R vi) {
  atomicAdd(&a->x, vr);
  atomicAdd(&a->y, vi);
}

template <typename R>
__device__ __forceinline__ void global_add(cplx<R> *a, R vr, R vi) {
  unsafeAtomicAdd(&a->x, vr);
  unsafeAtomicAdd(&a->y, vi);
}

// ===================================================================== static W
template <typename R, int D, int W>
__device__ __forceinline__ void node_window(const int *__restrict__ cell,
                                            const R *__restrict__ frac, long p, long M,
                                            const WinParams<R> &wp, int (&c)[D],
                                            R (&psi)[D][W]) {
  constexpr int m = W / 2 - 1;
#pragma unroll
  for (int t = 0; t < D; t++) {
    c[t] = cell[t * M + p];
    const R fr = frac[t * M + p];
#pragma unroll
    for (int w = 0; w < W; w++)
      psi[t][w] = kb_psi<R>(fr + R(m - w), R(m), wp.b[t], wp.inv_den[t]);
  }
}

template <typename R, int D, int W>
__global__ __launch_bounds__(256) void k_interp_tile(const cplx<R> *__restrict__ g,
                                                     cplx<R> *__restrict__ fout,
                                                     const int *__restrict__ perm,
                                                     const int *__restrict__ cell,
                                                     const R *__restrict__ frac,
                                                     const int4 *__restrict__ subs, BinGeom gg,
                                                     WinParams<R> wp) {
  extern __shared__ unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = g[region_to_grid<D>(q, lo, gg)];
  const int E1 = D >= 2 ? gg.E[1] : 1;
  const int E2 = D == 3 ? gg.E[2] : 1;
  __syncthreads();
  for (int i = threadIdx.x; i < sp.z; i += blockDim.x) {
    const long p = (long)sp.y + i;
    int c[D];
    R psi[D][W];
    node_window<R, D, W>(cell, frac, p, gg.M, wp, c, psi);
    R ar = 0, ai = 0;
    if constexpr (D == 1) {
      const cplx<R> *row = tile + (c[0] - lo[0]);
#pragma unroll
      for (int w = 0; w < W; w++) {
        const cplx<R> v = row[w];
        ar += psi[0][w] * v.x;
        ai += psi[0][w] * v.y;
      }
    } else if constexpr (D == 2) {
      const cplx<R> *base = tile + (c[0] - lo[0]) * E1 + (c[1] - lo[1]);
#pragma unroll
      for (int i0 = 0; i0 < W; i0++) {
        const cplx<R> *row = base + i0 * E1;
        R sr = 0, si = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
          const cplx<R> v = row[w];
          sr += psi[1][w] * v.x;
          si += psi[1][w] * v.y;
        }
        ar += psi[0][i0] * sr;
        ai += psi[0][i0] * si;
      }
    } else {
      const cplx<R> *base = tile + ((c[0] - lo[0]) * E1 + (c[1] - lo[1])) * E2 + (c[2] - lo[2]);
      for (int i0 = 0; i0 < W; i0++) {
        R qr = 0, qi = 0;
#pragma unroll
        for (int i1 = 0; i1 < W; i1++) {
          const cplx<R> *row = base + (i0 * E1 + i1) * E2;
          R sr = 0, si = 0;
#pragma unroll
          for (int w = 0; w < W; w++) {
            const cplx<R> v = row[w];
            sr += psi[2][w] * v.x;
            si += psi[2][w] * v.y;
          }
          qr += psi[1][i1] * sr;
          qi += psi[1][i1] * si;
        }
        ar += psi[0][i0] * qr;
        ai += psi[0][i0] * qi;
      }
    }
    fout[perm[p]] = cplx<R>{ar, ai};
  }
}

template <typename R, int D, int W>
__global__ __launch_bounds__(256) void k_spread_tile(const cplx<R> *__restrict__ fin,
                                                     cplx<R> *__restrict__ g,
                                                     const int *__restrict__ perm,
                                                     const int *__restrict__ cell,
                                                     const R *__restrict__ frac,
                                                     const int4 *__restrict__ subs, BinGeom gg,
                                                     WinParams<R> wp) {
  extern __shared__ unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = cplx<R>{R(0), R(0)};
  const int E1 = D >= 2 ? gg.E[1] : 1;
  const int E2 = D == 3 ? gg.E[2] : 1;
  __syncthreads();
  for (int i = threadIdx.x; i < sp.z; i += blockDim.x) {
    const long p = (long)sp.y + i;
    int c[D];
    R psi[D][W];
    node_window<R, D, W>(cell, frac, p, gg.M, wp, c, psi);
    const cplx<R> fj = fin[perm[p]];
    if constexpr (D == 1) {
      cplx<R> *row = tile + (c[0] - lo[0]);
#pragma unroll
      for (int w = 0; w < W; w++) lds_add<R>(row + w, psi[0][w] * fj.x, psi[0][w] * fj.y);
    } else if constexpr (D == 2) {
      cplx<R> *base = tile + (c[0] - lo[0]) * E1 + (c[1] - lo[1]);
#pragma unroll
      for (int i0 = 0; i0 < W; i0++) {
        cplx<R> *row = base + i0 * E1;
        const R wr = psi[0][i0] * fj.x, wi = psi[0][i0] * fj.y;
#pragma unroll
        for (int w = 0; w < W; w++) lds_add<R>(row + w, psi[1][w] * wr, psi[1][w] * wi);
      }
    } else {
      cplx<R> *base = tile + ((c[0] - lo[0]) * E1 + (c[1] - lo[1])) * E2 + (c[2] - lo[2]);
      for (int i0 = 0; i0 < W; i0++) {
        const R wr0 = psi[0][i0] * fj.x, wi0 = psi[0][i0] * fj.y;
#pragma unroll
        for (int i1 = 0; i1 < W; i1++) {
          cplx<R> *row = base + (i0 * E1 + i1) * E2;
          const R wr = psi[1][i1] * wr0, wi = psi[1][i1] * wi0;
#pragma unroll
          for (int w = 0; w < W; w++) lds_add<R>(row + w, psi[2][w] * wr, psi[2][w] * wi);
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < tot; q += blockDim.x) {
    const cplx<R> v = tile[q];
    if (v.x != R(0) || v.y != R(0)) global_add<R>(g + region_to_grid<D>(q, lo, gg), v.x, v.y);
  }
}

// ==================================================================== runtime m
// Window values and wrapped tap coordinates live in per-thread arrays.
template <typename R, int D>
__device__ __forceinline__ void node_window_rt(const int *__restrict__ cell,
                                               const R *__restrict__ frac, long p,
                                               const BinGeom &gg, const WinParams<R> &wp,
                                               int (&c)[D], R (&psi)[D][kMaxW]) {
  const int W = 2 * gg.m + 2;
  for (int t = 0; t < D; t++) {
    c[t] = cell[t * gg.M + p];
    const R fr = frac[t * gg.M + p];
    for (int w = 0; w < W; w++)
      psi[t][w] = kb_psi<R>(fr + R(gg.m - w), R(gg.m), wp.b[t], wp.inv_den[t]);
  }
}

// TILE: taps in the LDS tile at offset (c - lo); otherwise wrapped into g.
template <typename R, int D, bool TILE>
__global__ __launch_bounds__(256) void k_interp_rt(const cplx<R> *__restrict__ g,
                                                   cplx<R> *__restrict__ fout,
                                                   const int *__restrict__ perm,
                                                   const int *__restrict__ cell,
                                                   const R *__restrict__ frac,
                                                   const int4 *__restrict__ subs, BinGeom gg,
                                                   WinParams<R> wp) {
  extern __shared__ unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int W = 2 * gg.m + 2;
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  if constexpr (TILE) {
    int tot = 1;
    for (int t = 0; t < D; t++) tot *= gg.E[t];
    for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = g[region_to_grid<D>(q, lo, gg)];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < sp.z; i += blockDim.x) {
    const long p = (long)sp.y + i;
    int c[D];
    R psi[D][kMaxW];
    node_window_rt<R, D>(cell, frac, p, gg, wp, c, psi);
    int idx[D][kMaxW];
    long stride[D];
    for (int t = 0; t < D; t++) {
      for (int w = 0; w < W; w++)
        idx[t][w] = TILE ? (c[t] - lo[t] + w) : wrap_once(c[t] - gg.m + w, gg.n[t]);
    }
    // strides of the source (tile or grid), last dimension fastest
    {
      long sdim = 1;
      for (int t = D - 1; t >= 0; t--) {
        stride[t] = sdim;
        sdim *= TILE ? gg.E[t] : gg.n[t];
      }
    }
    const cplx<R> *src = TILE ? tile : g;
    R ar = 0, ai = 0;
    if constexpr (D == 1) {
      for (int w = 0; w < W; w++) {
        const cplx<R> v = src[idx[0][w]];
        ar += psi[0][w] * v.x;
        ai += psi[0][w] * v.y;
      }
    } else if constexpr (D == 2) {
      for (int i0 = 0; i0 < W; i0++) {
        const cplx<R> *row = src + idx[0][i0] * stride[0];
        R sr = 0, si = 0;
        for (int w = 0; w < W; w++) {
          const cplx<R> v = row[idx[1][w]];
          sr += psi[1][w] * v.x;
          si += psi[1][w] * v.y;
        }
        ar += psi[0][i0] * sr;
        ai += psi[0][i0] * si;
      }
    } else {
      for (int i0 = 0; i0 < W; i0++) {
        R qr = 0, qi = 0;
        for (int i1 = 0; i1 < W; i1++) {
          const cplx<R> *row = src + idx[0][i0] * stride[0] + idx[1][i1] * stride[1];
          R sr = 0, si = 0;
          for (int w = 0; w < W; w++) {
            const cplx<R> v = row[idx[2][w]];
            sr += psi[2][w] * v.x;
            si += psi[2][w] * v.y;
          }
          qr += psi[1][i1] * sr;
          qi += psi[1][i1] * si;
        }
        ar += psi[0][i0] * qr;
        ai += psi[0][i0] * qi;
      }
    }
    fout[perm[p]] = cplx<R>{ar, ai};
  }
}

template <typename R, int D, bool TILE>
__global__ __launch_bounds__(256) void k_spread_rt(const cplx<R> *__restrict__ fin,
                                                   cplx<R> *__restrict__ g,
                                                   const int *__restrict__ perm,
                                                   const int *__restrict__ cell,
                                                   const R *__restrict__ frac,
                                                   const int4 *__restrict__ subs, BinGeom gg,
                                                   WinParams<R> wp) {
  extern __shared__ unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int W = 2 * gg.m + 2;
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
  if constexpr (TILE) {
    for (int t = 0; t < D; t++) tot *= gg.E[t];
    for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = cplx<R>{R(0), R(0)};
    __syncthreads();
  }
  for (int i = threadIdx.x; i < sp.z; i += blockDim.x) {
    const long p = (long)sp.y + i;
    int c[D];
    R psi[D][kMaxW];
    node_window_rt<R, D>(cell, frac, p, gg, wp, c, psi);
    const cplx<R> fj = fin[perm[p]];
    int idx[D][kMaxW];
    long stride[D];
    for (int t = 0; t < D; t++) {
      for (int w = 0; w < W; w++)
        idx[t][w] = TILE ? (c[t] - lo[t] + w) : wrap_once(c[t] - gg.m + w, gg.n[t]);
    }
    {
      long sdim = 1;
      for (int t = D - 1; t >= 0; t--) {
        stride[t] = sdim;
        sdim *= TILE ? gg.E[t] : gg.n[t];
      }
    }
    cplx<R> *dst = TILE ? tile : g;
    if constexpr (D == 1) {
      for (int w = 0; w < W; w++) {
        if (TILE) lds_add<R>(dst + idx[0][w], psi[0][w] * fj.x, psi[0][w] * fj.y);
        else global_add<R>(dst + idx[0][w], psi[0][w] * fj.x, psi[0][w] * fj.y);
      }
    } else if constexpr (D == 2) {
      for (int i0 = 0; i0 < W; i0++) {
        cplx<R> *row = dst + idx[0][i0] * stride[0];
        const R wr = psi[0][i0] * fj.x, wi = psi[0][i0] * fj.y;
        for (int w = 0; w < W; w++) {
          if (TILE) lds_add<R>(row + idx[1][w], psi[1][w] * wr, psi[1][w] * wi);
          else global_add<R>(row + idx[1][w], psi[1][w] * wr, psi[1][w] * wi);
        }
      }
    } else {
      for (int i0 = 0; i0 < W; i0++) {
        for (int i1 = 0; i1 < W; i1++) {
          cplx<R> *row = dst + idx[0][i0] * stride[0] + idx[1][i1] * stride[1];
          const R w01 = psi[0][i0] * psi[1][i1];
          const R wr = w01 * fj.x, wi = w01 * fj.y;
          for (int w = 0; w < W; w++) {
            if (TILE) lds_add<R>(row + idx[2][w], psi[2][w] * wr, psi[2][w] * wi);
            else global_add<R>(row + idx[2][w], psi[2][w] * wr, psi[2][w] * wi);
          }
        }
      }
    }
  }
  if constexpr (TILE) {
    __syncthreads();
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
      const cplx<R> v = tile[q];
      if (v.x != R(0) || v.y != R(0)) global_add<R>(g + region_to_grid<D>(q, lo, gg), v.x, v.y);
    }
  }
}

// ==================================================================== dispatch
template <typename R> struct BKernelTable {
  using fn = void (*)(const cplx<R> *, cplx<R> *, const int *, const int *, const R *,
                      const int4 *, BinGeom, WinParams<R>);
  fn interp_static[kMaxDim][kMaxStaticM + 1];
  fn spread_static[kMaxDim][kMaxStaticM + 1];
  fn interp_rt[kMaxDim][2];  // [d-1][tile]
  fn spread_rt[kMaxDim][2];
};

template <typename R, int D, int M_>
void bk_fill_one(BKernelTable<R> &k) {
  k.interp_static[D - 1][M_] = k_interp_tile<R, D, 2 * M_ + 2>;
  k.spread_static[D - 1][M_] = k_spread_tile<R, D, 2 * M_ + 2>;
}

template <typename R, int D, int... Ms>
void bk_fill_dim(BKernelTable<R> &k, std::integer_sequence<int, Ms...>) {
  (bk_fill_one<R, D, Ms + 1>(k), ...);
  k.interp_rt[D - 1][0] = k_interp_rt<R, D, false>;
  k.interp_rt[D - 1][1] = k_interp_rt<R, D, true>;
  k.spread_rt[D - 1][0] = k_spread_rt<R, D, false>;
  k.spread_rt[D - 1][1] = k_spread_rt<R, D, true>;
}

template <typename R>
BKernelTable<R> make_bkernel_table() {
  BKernelTable<R> t{};
  using seq = std::make_integer_sequence<int, kMaxStaticM>;
  bk_fill_dim<R, 1>(t, seq{});
  bk_fill_dim<R, 2>(t, seq{});
  bk_fill_dim<R, 3>(t, seq{});
  return t;
}

template <typename R>
size_t tile_bytes(const BinGeom &gg) {
  size_t tot = 1;
  for (int t = 0; t < gg.d; t++) tot *= (size_t)gg.E[t];
  return tot * sizeof(cplx<R>);
}

// LDS tiles above this size use the global-memory variant
constexpr size_t kTileBudget = 64 * 1024;

template <typename R>
typename BKernelTable<R>::fn pick_bkernel(const BKernelTable<R> &k, const BinGeom &gg,
                                          bool spread, size_t &lds) {
  const size_t bytes = tile_bytes<R>(gg);
  const bool tile = bytes <= kTileBudget;
  lds = tile ? bytes : 0;
  if (tile && gg.m >= 1 && gg.m <= kMaxStaticM)
    return spread ? k.spread_static[gg.d - 1][gg.m] : k.interp_static[gg.d - 1][gg.m];
  return spread ? k.spread_rt[gg.d - 1][tile ? 1 : 0] : k.interp_rt[gg.d - 1][tile ? 1 : 0];
}

template <typename R>
const BKernelTable<R> &bkernels() {
  static BKernelTable<R> k = make_bkernel_table<R>();
  return k;
}

template <typename R>
void launch_interp(const cplx<R> *g, cplx<R> *f, const Engine<R> &e, hipStream_t s) {
  if (e.nsub == 0) return;
  size_t lds;
  auto fn = pick_bkernel<R>(bkernels<R>(), e.geom, false, lds);
  hipLaunchKernelGGL(fn, dim3((unsigned)e.nsub), dim3(256), lds, s, g, f, e.perm, e.cell, e.frac,
                     e.subs, e.geom, e.win);
}

template <typename R>
void launch_spread(const cplx<R> *f, cplx<R> *g, const Engine<R> &e, hipStream_t s) {
  if (e.nsub == 0) return;
  size_t lds;
  auto fn = pick_bkernel<R>(bkernels<R>(), e.geom, true, lds);
  hipLaunchKernelGGL(fn, dim3((unsigned)e.nsub), dim3(256), lds, s, f, g, e.perm, e.cell, e.frac,
                     e.subs, e.geom, e.win);
}

}  // namespace nfft_hip
