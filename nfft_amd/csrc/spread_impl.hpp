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

#include <type_traits>

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

// Spreading tiles in scratch.  Linear order is the LDS order.  Zone order
// (BinGeom::zoned) splits every dimension of the tile into the part that
// belongs to the left neighbour bin's core (m cells), the own core (bs) and
// the right neighbour's core (m + 1), and stores the 3^d zones one after the
// other (row-major over zone indices, each zone row-major inside), so the
// tile reduction reads each contribution as one contiguous block.  The plan
// builds the order once (BinGeom::zmap: tile cell of every scratch position;
// BinGeom::zdesc: where each neighbour's zone lies in its tile and in the
// receiving core).

// Writes the block's LDS tile (values from get(q), q the LDS index) to
// scratch slot `slot` (the work item's place in bin order, which the tile
// reductions rely on; items are launched largest first), or adds it into the
// grid with atomics when there is no scratch.
template <typename R, int D, typename Get>
__device__ __forceinline__ void flush_tile(Get get, cplx<R> *__restrict__ tiles,
                                           cplx<R> *__restrict__ g, const int (&lo)[D],
                                           const BinGeom &gg, int tot, long slot, int tid,
                                           int nthr) {
  if (tiles && gg.zoned) {
    // zone order through the plan's table: coalesced writes, LDS reads
    cplx<R> *out = tiles + slot * tot;
    for (int o = tid; o < tot; o += nthr) out[o] = get(gg.zmap[o]);
    return;
  }
  for (int q = tid; q < tot; q += nthr) {
    const cplx<R> v = get(q);
    if (tiles) tiles[slot * tot + q] = v;
    else if (v.x != R(0) || v.y != R(0)) {
      const long gi = region_to_grid<D>(q, lo, gg);
      unsafeAtomicAdd(&g[gi].x, v.x);
      unsafeAtomicAdd(&g[gi].y, v.y);
    }
  }
}

template <typename R>
__device__ __forceinline__ void global_add(cplx<R> *a, R vr, R vi) {
  unsafeAtomicAdd(&a->x, vr);
  unsafeAtomicAdd(&a->y, vi);
}

// LDS accumulator of B^T.  On this GPU a float atomic add to LDS (ds_add_f32)
// is about 40x slower than a 32-bit integer one and 10x slower than a
// double one, so the float family accumulates in 32-bit fixed point:
// contributions are scaled by 2^30 / S, S = sum over the work item's nodes of
// |Re f_j| + |Im f_j|, which bounds every tile cell because the normalised
// window product never exceeds 1; integer sums are exact and wrap-safe, and
// the rounding error per contribution is below S / 2^31.  The double family
// adds in FP64 (ds_add_f64).
template <typename R> struct LdsAcc;
template <> struct LdsAcc<float> {
  using elem = int2;
  float scale = 1.f, inv_scale = 1.f;
  __device__ void set_bound(float S) {
    scale = S > 0.f ? 1073741824.f / S : 1.f;
    inv_scale = 1.f / scale;
  }
  __device__ __forceinline__ void add(elem *a, float vr, float vi) const {
    atomicAdd(&a->x, __float2int_rn(vr * scale));
    atomicAdd(&a->y, __float2int_rn(vi * scale));
  }
  __device__ __forceinline__ cplx<float> load(const elem *a) const {
    const int2 v = *a;
    return cplx<float>{(float)v.x * inv_scale, (float)v.y * inv_scale};
  }
  __device__ static elem zero() { return make_int2(0, 0); }
};
template <> struct LdsAcc<double> {
  using elem = double2;
  __device__ void set_bound(double) {}
  __device__ __forceinline__ void add(elem *a, double vr, double vi) const {
    atomicAdd(&a->x, vr);
    atomicAdd(&a->y, vi);
  }
  __device__ __forceinline__ cplx<double> load(const elem *a) const { return *a; }
  __device__ static elem zero() { return make_double2(0.0, 0.0); }
};

// sum over the work item's nodes of |Re f| + |Im f|, on every thread
template <typename R>
__device__ R block_abs_sum(const cplx<R> *__restrict__ fin, const int *__restrict__ perm,
                           long first, int count) {
  __shared__ R red[16];
  R s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    const cplx<R> v = fin[perm[first + i]];
    s += fabs(v.x) + fabs(v.y);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  R t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
  __syncthreads();
  return t;
}

// ===================================================================== static W
template <typename R, int D, int W>
__device__ __forceinline__ void node_window(const int *__restrict__ cell,
                                            const R *__restrict__ frac, long p, long M,
                                            const WinParams<R> &wp, int (&c)[D],
                                            R (&psi)[D][W]) {
  static_assert(W % 2 == 0, "W = 2m+2 taps");
  // per-tap polynomials in tt = frac - 1/2, coefficients uniform over the
  // wavefront (scalar loads); the table is mirror-symmetric (tap W-1-w is
  // tap w at -tt), so a tap pair comes from tap w's even and odd parts,
  // E(tt^2) +- tt O(tt^2): PD + 2 FMAs per pair instead of 2 PD
  constexpr int PD = poly_deg_v<R>;
  constexpr int KE = PD - (PD % 2);      // highest even power
  constexpr int KO = PD - 1 + (PD % 2);  // highest odd power
#pragma unroll
  for (int t = 0; t < D; t++) {
    c[t] = cell[t * M + p];
    const R tt = frac[t * M + p] - R(0.5);
    const R t2 = tt * tt;
#pragma unroll
    for (int w = 0; w < W / 2; w++) {
      const R *pc = wp.poly + (t * W + w) * (PD + 1);
      R e = pc[KE];
#pragma unroll
      for (int k = KE - 2; k >= 0; k -= 2) e = fma(e, t2, pc[k]);
      R o = pc[KO];
#pragma unroll
      for (int k = KO - 2; k >= 1; k -= 2) o = fma(o, t2, pc[k]);
      psi[t][w] = fma(tt, o, e);
      psi[t][W - 1 - w] = fma(-tt, o, e);
    }
  }
}

// FP32 window of a node when every dimension shares one mirror-symmetric
// tap table (wp.poly_shared): taps w and W-1-w come from the even part E and
// odd part O of tap w's polynomial, psi = E(t^2) +- t O(t^2).  The W/2
// polynomials of dimension 0 (W/2 * kPolyCoef values) stay in scalar
// registers for the whole kernel instead of the d * W per-tap sets, and a
// tap pair costs kPolyDeg + 1 FMAs instead of 2 kPolyDeg.
template <int D, int W>
__device__ __forceinline__ void node_window_sym(const int *__restrict__ cell,
                                                const float *__restrict__ frac, long p, long M,
                                                const WinParams<float> &wp, int (&c)[D],
                                                float (&psi)[D][W]) {
  constexpr int KE = kPolyDeg - (kPolyDeg % 2);  // highest even power
  constexpr int KO = kPolyDeg - 1 + (kPolyDeg % 2);  // highest odd power
#pragma unroll
  for (int t = 0; t < D; t++) {
    c[t] = cell[t * M + p];
    const float tt = frac[t * M + p] - 0.5f;
    const float t2 = tt * tt;
#pragma unroll
    for (int w = 0; w < W / 2; w++) {
      const float *pc = wp.poly + w * kPolyCoef;
      float e = pc[KE];
#pragma unroll
      for (int k = KE - 2; k >= 0; k -= 2) e = fmaf(e, t2, pc[k]);
      float o = pc[KO];
#pragma unroll
      for (int k = KO - 2; k >= 1; k -= 2) o = fmaf(o, t2, pc[k]);
      psi[t][w] = fmaf(tt, o, e);
      psi[t][W - 1 - w] = fmaf(-tt, o, e);
    }
  }
}

// Stages the grid region of a 3-D work item (E0 x E1 x E2 cells from
// lo - m, periodic) into LDS, rows gg.ip2 and planes gg.ipp cells apart.  Each thread keeps one position along the last
// dimension and walks whole rows, so per element it does one table lookup
// per leading dimension instead of divisions and modulos.
template <typename R>
__device__ __forceinline__ void load_region_3d(const cplx<R> *__restrict__ g, cplx<R> *tile,
                                               const int (&lo)[3], const BinGeom &gg,
                                               long *rowoff0, long *rowoff1) {
  const int E0 = gg.E[0], E1 = gg.E[1], E2 = gg.E[2], P2 = gg.ip2, PP = gg.ipp;
  auto wrap = [](int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
  };
  for (int e = threadIdx.x; e < E0; e += blockDim.x)
    rowoff0[e] = (long)wrap(lo[0] - gg.m + e, gg.n[0]) * gg.n[1] * gg.n[2];
  for (int e = threadIdx.x; e < E1; e += blockDim.x)
    rowoff1[e] = (long)wrap(lo[1] - gg.m + e, gg.n[1]) * gg.n[2];
  __syncthreads();
  const int rows_per_pass = blockDim.x / E2;
  const int e2 = threadIdx.x % E2;
  int r = threadIdx.x / E2;
  if (r < rows_per_pass) {
    const long col = wrap(lo[2] - gg.m + e2, gg.n[2]);
    int e0 = r / E1, e1 = r - (r / E1) * E1;
    // eight loads in flight before their LDS stores
    constexpr int B = 8;
    while (e0 < E0) {
      cplx<R> v[B];
      int dst[B];
#pragma unroll
      for (int u = 0; u < B; u++) {
        dst[u] = -1;
        if (e0 < E0) {
          dst[u] = e0 * PP + e1 * P2 + e2;
          v[u] = g[rowoff0[e0] + rowoff1[e1] + col];
          e1 += rows_per_pass;
          while (e1 >= E1) {
            e1 -= E1;
            e0++;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < B; u++)
        if (dst[u] >= 0) tile[dst[u]] = v[u];
    }
  }
}

// Interpolation of one slot of up to Q nodes of one cell (3-D, FP32, the
// shared symmetric window table): every grid tap of the slot's window is
// read once from LDS for all its nodes.  The dimension-1 and -2 windows of
// the nodes stay in registers; the dimension-0 weights are evaluated row
// pair by row pair (rows i and W-1-i from one polynomial's even and odd
// parts), which keeps the kernel at three wavefronts per SIMD.  Nodes past
// the slot's count repeat its last node and are not written; a lane with
// live = false writes nothing.
template <int W, int Q>
__device__ __forceinline__ void interp_slot_3d(const float2 *tile, float2 *__restrict__ fout,
                                               const int *__restrict__ perm,
                                               const int *__restrict__ cell,
                                               const float *__restrict__ frac, const BinGeom &gg,
                                               const WinParams<float> &wp, const int (&lo)[3],
                                               int PP, int E2, int2 pr, bool live) {
  constexpr int KE = kPolyDeg - (kPolyDeg % 2);      // highest even power
  constexpr int KO = kPolyDeg - 1 + (kPolyDeg % 2);  // highest odd power
  const long M = gg.M;
  long pn[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) pn[q] = (long)pr.x + min(q, pr.y - 1);
  // the symmetric tap pair (w, W-1-w) at tt = frac - 1/2
  auto tap_pair = [&](int w, float tt, float &lo_v, float &hi_v) {
    const float *pc = wp.poly + w * kPolyCoef;
    const float t2 = tt * tt;
    float e = pc[KE];
#pragma unroll
    for (int k = KE - 2; k >= 0; k -= 2) e = fmaf(e, t2, pc[k]);
    float o = pc[KO];
#pragma unroll
    for (int k = KO - 2; k >= 1; k -= 2) o = fmaf(o, t2, pc[k]);
    lo_v = fmaf(tt, o, e);
    hi_v = fmaf(-tt, o, e);
  };
  float p1[Q][W], p2[Q][W], t0[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    t0[q] = frac[pn[q]] - 0.5f;
    const float t1 = frac[M + pn[q]] - 0.5f;
    const float t2 = frac[2 * M + pn[q]] - 0.5f;
#pragma unroll
    for (int w = 0; w < W / 2; w++) {
      tap_pair(w, t1, p1[q][w], p1[q][W - 1 - w]);
      tap_pair(w, t2, p2[q][w], p2[q][W - 1 - w]);
    }
  }
  const float2 *base = tile + (cell[pr.x] - lo[0]) * PP + (cell[M + pr.x] - lo[1]) * E2 +
                       (cell[2 * M + pr.x] - lo[2]);
  float ar[Q] = {}, ai[Q] = {};
  for (int ip = 0; ip < W / 2; ip++) {
    float w0[2][Q];
#pragma unroll
    for (int q = 0; q < Q; q++) tap_pair(ip, t0[q], w0[0][q], w0[1][q]);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int i0 = h == 0 ? ip : W - 1 - ip;
      float qr[Q] = {}, qi[Q] = {};
      for (int i1 = 0; i1 < W; i1++) {
        const float2 *row = base + i0 * PP + i1 * E2;
        // the row's taps are all requested before the first is used
        float2 v[W];
#pragma unroll
        for (int w = 0; w < W; w++) v[w] = row[w];
        float sr[Q] = {}, si[Q] = {};
#pragma unroll
        for (int w = 0; w < W; w++)
#pragma unroll
          for (int q = 0; q < Q; q++) {
            sr[q] = fmaf(p2[q][w], v[w].x, sr[q]);
            si[q] = fmaf(p2[q][w], v[w].y, si[q]);
          }
#pragma unroll
        for (int q = 0; q < Q; q++) {
          qr[q] = fmaf(p1[q][i1], sr[q], qr[q]);
          qi[q] = fmaf(p1[q][i1], si[q], qi[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        ar[q] = fmaf(w0[h][q], qr[q], ar[q]);
        ai[q] = fmaf(w0[h][q], qi[q], ai[q]);
      }
    }
  }
  if (live) {
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (q < pr.y) fout[perm[pr.x + q]] = float2{ar[q], ai[q]};
  }
}

// PAIR (D = 2, 3): threads take the work item's slots (gg.islots / gg.slots)
// instead of single nodes; a slot of two nodes of one cell reads each grid
// tap once from LDS for both, halving the LDS bytes per node of dense runs.
// With SYM in 3-D a slot may hold up to four nodes and the largest count
// among a wavefront's slots picks its code: the pair code for two, the
// slot code (interp_slot_3d) for one, three or four nodes, which reads each
// tap once for all of them.  Precompute orders an item's slots by count, so
// that most wavefronts hold slots of one count.
template <typename R, int D, int W, bool SYM = false, bool PAIR = false>
__global__ __launch_bounds__(256) void k_interp_tile(const cplx<R> *__restrict__ g,
                                                     cplx<R> *__restrict__ fout,
                                                     const int *__restrict__ perm,
                                                     const int *__restrict__ cell,
                                                     const R *__restrict__ frac,
                                                     const int4 *__restrict__ subs, BinGeom gg,
                                                     WinParams<R> wp, cplx<R> *__restrict__ tiles) {
  // right-hand side blockIdx.y of a batched launch
  g += (long)blockIdx.y * gg.bs_g;
  fout += (long)blockIdx.y * gg.bs_f;
  extern __shared__ __align__(16) unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  // next 64-slot chunk of the work item (PAIR).  (The dynamic tile is
  // declared 16-byte aligned: after this 4-byte variable an unaligned
  // declaration started it at a 4 (mod 8) offset, and every 8-byte LDS
  // access of the 2-D kernel ran 4.6x slower.)
  __shared__ int next_chunk;
  if (threadIdx.x == 0) next_chunk = 0;
  if constexpr (D == 3) {
    __shared__ long rowoff[2][64];
    load_region_3d<R>(g, tile, lo, gg, rowoff[0], rowoff[1]);
  } else {
    for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = g[region_to_grid<D>(q, lo, gg)];
  }
  const int E1 = D >= 2 ? gg.E[1] : 1;
  // the tile's row and plane pitches (load_region_3d)
  const int E2 = D == 3 ? gg.ip2 : 1;
  const int PP = D == 3 ? gg.ipp : 0;
  __syncthreads();
  if constexpr (PAIR && (D == 2 || D == 3)) {
    const int2 is = gg.islots[blockIdx.x];
    // wavefronts take the item's 64-slot chunks in turn as they finish
    // their last, so a wavefront with light chunks does not idle while
    // another works through heavy ones
    for (;;) {
      int ch = 0;
      if ((threadIdx.x & 63) == 0) ch = atomicAdd(&next_chunk, 1);
      ch = __shfl(ch, 0, 64);
      if (ch * 64 >= is.y) break;
      const int k = ch * 64 + (threadIdx.x & 63);
      // (first node, count) of this lane's slot; lanes past the item's
      // slots take its first slot and write nothing
      const int2 pr = gg.slots[is.x + (k < is.y ? k : 0)];
      const bool live = k < is.y;
      if constexpr (D == 3 && SYM) {
        // the wavefront's largest slot picks the code path
        int wmax = live ? pr.y : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
        if (wmax == 0) continue;
        if (wmax != 2) {
          // (slots of three or four nodes only for windows of up to 10 taps,
          // where their registers fit; the plan builds pairs otherwise)
          if constexpr (W <= kSlotQuadMaxW) {
            if (wmax == 4)
              interp_slot_3d<W, 4>(tile, fout, perm, cell, frac, gg, wp, lo, PP, E2, pr, live);
            else if (wmax == 3)
              interp_slot_3d<W, 3>(tile, fout, perm, cell, frac, gg, wp, lo, PP, E2, pr, live);
            else
              interp_slot_3d<W, 1>(tile, fout, perm, cell, frac, gg, wp, lo, PP, E2, pr, live);
          } else {
            interp_slot_3d<W, 1>(tile, fout, perm, cell, frac, gg, wp, lo, PP, E2, pr, live);
          }
          continue;
        }
      }
      if (!live) continue;
      const long pa = pr.x;
      const bool two = pr.y >= 2;
      const long pb = two ? pa + 1 : pa;
      int c[D], cb[D];
      R pa_[D][W], pb_[D][W];
      if constexpr (SYM) {
        node_window_sym<D, W>(cell, frac, pa, gg.M, wp, c, pa_);
        node_window_sym<D, W>(cell, frac, pb, gg.M, wp, cb, pb_);
      } else {
        node_window<R, D, W>(cell, frac, pa, gg.M, wp, c, pa_);
        node_window<R, D, W>(cell, frac, pb, gg.M, wp, cb, pb_);
      }
      R ar = 0, ai = 0, br = 0, bi = 0;
      if constexpr (D == 2) {
        const cplx<R> *base = tile + (c[0] - lo[0]) * E1 + (c[1] - lo[1]);
#pragma unroll
        for (int i0 = 0; i0 < W; i0++) {
          const cplx<R> *row = base + i0 * E1;
          R sar = 0, sai = 0, sbr = 0, sbi = 0;
#pragma unroll
          for (int w = 0; w < W; w++) {
            const cplx<R> v = row[w];
            sar += pa_[1][w] * v.x;
            sai += pa_[1][w] * v.y;
            sbr += pb_[1][w] * v.x;
            sbi += pb_[1][w] * v.y;
          }
          ar += pa_[0][i0] * sar;
          ai += pa_[0][i0] * sai;
          br += pb_[0][i0] * sbr;
          bi += pb_[0][i0] * sbi;
        }
      } else {
        const cplx<R> *base = tile + (c[0] - lo[0]) * PP + (c[1] - lo[1]) * E2 + (c[2] - lo[2]);
        for (int i0 = 0; i0 < W; i0++) {
          R qar = 0, qai = 0, qbr = 0, qbi = 0;
#pragma unroll
          for (int i1 = 0; i1 < W; i1++) {
            const cplx<R> *row = base + i0 * PP + i1 * E2;
            // the row's taps are all requested before the first is used, so
            // a wavefront keeps several LDS reads in flight (short windows
            // read in the loop: the compiler would hoist every row)
            cplx<R> v[W];
            if constexpr (W >= 8) {
#pragma unroll
              for (int w = 0; w < W; w++) v[w] = row[w];
            }
            R sar = 0, sai = 0, sbr = 0, sbi = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
              const cplx<R> t = W >= 8 ? v[w] : row[w];
              sar += pa_[2][w] * t.x;
              sai += pa_[2][w] * t.y;
              sbr += pb_[2][w] * t.x;
              sbi += pb_[2][w] * t.y;
            }
            qar += pa_[1][i1] * sar;
            qai += pa_[1][i1] * sai;
            qbr += pb_[1][i1] * sbr;
            qbi += pb_[1][i1] * sbi;
          }
          ar += pa_[0][i0] * qar;
          ai += pa_[0][i0] * qai;
          br += pb_[0][i0] * qbr;
          bi += pb_[0][i0] * qbi;
        }
      }
      fout[perm[pa]] = cplx<R>{ar, ai};
      if (two) fout[perm[pb]] = cplx<R>{br, bi};
    }
    return;
  }
  for (int i = threadIdx.x; i < sp.z; i += blockDim.x) {
    const long p = (long)sp.y + i;
    int c[D];
    R psi[D][W];
    if constexpr (SYM) node_window_sym<D, W>(cell, frac, p, gg.M, wp, c, psi);
    else node_window<R, D, W>(cell, frac, p, gg.M, wp, c, psi);
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
        // the row's taps are all requested before the first is used
        cplx<R> v[W];
        if constexpr (W >= 8) {
#pragma unroll
          for (int w = 0; w < W; w++) v[w] = row[w];
        }
        R sr = 0, si = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
          const cplx<R> t = W >= 8 ? v[w] : row[w];
          sr += psi[1][w] * t.x;
          si += psi[1][w] * t.y;
        }
        ar += psi[0][i0] * sr;
        ai += psi[0][i0] * si;
      }
    } else {
      const cplx<R> *base = tile + (c[0] - lo[0]) * PP + (c[1] - lo[1]) * E2 + (c[2] - lo[2]);
      for (int i0 = 0; i0 < W; i0++) {
        R qr = 0, qi = 0;
#pragma unroll
        for (int i1 = 0; i1 < W; i1++) {
          const cplx<R> *row = base + i0 * PP + i1 * E2;
          cplx<R> v[W];
          if constexpr (W >= 8) {
#pragma unroll
            for (int w = 0; w < W; w++) v[w] = row[w];
          }
          R sr = 0, si = 0;
#pragma unroll
          for (int w = 0; w < W; w++) {
            const cplx<R> t = W >= 8 ? v[w] : row[w];
            sr += psi[2][w] * t.x;
            si += psi[2][w] * t.y;
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
                                                     WinParams<R> wp, cplx<R> *__restrict__ tiles) {
  extern __shared__ __align__(16) unsigned char b_smem[];
  using Acc = LdsAcc<R>;
  typename Acc::elem *tile = reinterpret_cast<typename Acc::elem *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  Acc acc;
  acc.set_bound(block_abs_sum<R>(fin, perm, sp.y, sp.z));
  for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = Acc::zero();
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
      typename Acc::elem *row = tile + (c[0] - lo[0]);
#pragma unroll
      for (int w = 0; w < W; w++) acc.add(row + w, psi[0][w] * fj.x, psi[0][w] * fj.y);
    } else if constexpr (D == 2) {
      typename Acc::elem *base = tile + (c[0] - lo[0]) * E1 + (c[1] - lo[1]);
#pragma unroll
      for (int i0 = 0; i0 < W; i0++) {
        typename Acc::elem *row = base + i0 * E1;
        const R wr = psi[0][i0] * fj.x, wi = psi[0][i0] * fj.y;
#pragma unroll
        for (int w = 0; w < W; w++) acc.add(row + w, psi[1][w] * wr, psi[1][w] * wi);
      }
    } else {
      typename Acc::elem *base = tile + ((c[0] - lo[0]) * E1 + (c[1] - lo[1])) * E2 + (c[2] - lo[2]);
      for (int i0 = 0; i0 < W; i0++) {
        const R wr0 = psi[0][i0] * fj.x, wi0 = psi[0][i0] * fj.y;
#pragma unroll
        for (int i1 = 0; i1 < W; i1++) {
          typename Acc::elem *row = base + (i0 * E1 + i1) * E2;
          const R wr = psi[1][i1] * wr0, wi = psi[1][i1] * wi0;
#pragma unroll
          for (int w = 0; w < W; w++) acc.add(row + w, psi[2][w] * wr, psi[2][w] * wi);
        }
      }
    }
  }
  __syncthreads();
  flush_tile<R, D>([&](int q) { return acc.load(tile + q); }, tiles, g, lo, gg, tot, sp.w, threadIdx.x,
                   blockDim.x);
}

// ============================================================ wave per node
// d = 2, 3: one wavefront handles one node at a time.  Lanes own fixed pairs
// of tap indices over the last two dimensions (W*W pairs in ceil(W*W/64)
// slots); for d = 3 the first dimension is a loop whose weight is read from
// one lane (readlane, uniform).  The d*W window values of the node are
// computed once, one per lane, and fetched with ds_bpermute.  In one wave
// instruction all lanes touch distinct taps, so LDS atomics of a dense
// cluster (many nodes in one cell) never collide inside the wave, and the
// window is evaluated d*W times per node instead of per lane.
template <typename R>
__device__ __forceinline__ R lane_bcast(R v, int src) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
  } else {
    long long u = __builtin_bit_cast(long long, v);
    int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
    return __builtin_bit_cast(R, ((long long)hi << 32) | (unsigned int)lo);
  }
}

template <typename R>
__device__ __forceinline__ R lane_fetch(R v, int src) {
  return __shfl(v, src, 64);
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename R, int D, int W, bool SPREAD>
__global__ __launch_bounds__(256) void k_bwave_tile(const cplx<R> *__restrict__ src,
                                                    cplx<R> *__restrict__ dst,
                                                    const int *__restrict__ perm,
                                                    const int *__restrict__ cell,
                                                    const R *__restrict__ frac,
                                                    const int4 *__restrict__ subs, BinGeom gg,
                                                    WinParams<R> wp, cplx<R> *__restrict__ tiles) {
  static_assert(D == 2 || D == 3, "wave-per-node kernels cover d = 2, 3");
  constexpr int PAIRS = W * W;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  extern __shared__ __align__(16) unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  using Acc = LdsAcc<R>;
  typename Acc::elem *atile = reinterpret_cast<typename Acc::elem *>(b_smem);
  Acc acc;
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  if constexpr (SPREAD) {
    acc.set_bound(block_abs_sum<R>(src, perm, sp.y, sp.z));
    for (int q = threadIdx.x; q < tot; q += blockDim.x) atile[q] = Acc::zero();
  } else {
    for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = src[region_to_grid<D>(q, lo, gg)];
  }
  __syncthreads();
  const int E1 = gg.E[1];
  const int E2 = D == 3 ? gg.E[2] : 1;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  // this lane's tap pairs (pa over dim D-2, pb over dim D-1)
  int pa[SLOTS], pb[SLOTS], off[SLOTS];
  bool pv[SLOTS];
#pragma unroll
  for (int s = 0; s < SLOTS; s++) {
    const int q = lane + 64 * s;
    pv[s] = q < PAIRS;
    pa[s] = pv[s] ? q / W : 0;
    pb[s] = pv[s] ? q - (q / W) * W : 0;
    off[s] = D == 3 ? pa[s] * E2 + pb[s] : pa[s] * E1 + pb[s];
  }
  // window value owned by this lane: dimension lane / W, tap lane % W, its
  // polynomial in registers
  const int wt = lane / W < D ? lane / W : D - 1;
  const int ww = lane - (lane / W) * W;
  TapCoef<R, poly_deg_v<R>> tc;
  tc.load(wp.poly + (wt * W + ww) * (poly_deg_v<R> + 1));
  for (int i = wv; i < sp.z; i += nw) {
    const long p = (long)sp.y + i;
    R mine = R(0);
    if (lane < D * W) mine = tc(frac[wt * gg.M + p]);
    int a[D];
#pragma unroll
    for (int t = 0; t < D; t++) a[t] = cell[t * gg.M + p] - lo[t];
    // product weights of this lane's pairs
    R wpair[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
      const R u = lane_fetch<R>(mine, (D - 2) * W + pa[s]);
      const R v = lane_fetch<R>(mine, (D - 1) * W + pb[s]);
      wpair[s] = pv[s] ? u * v : R(0);
    }
    if constexpr (SPREAD) {
      const cplx<R> fj = src[perm[p]];
      if constexpr (D == 2) {
        typename Acc::elem *base = atile + a[0] * E1 + a[1];
#pragma unroll
        for (int s = 0; s < SLOTS; s++)
          if (pv[s]) acc.add(base + off[s], wpair[s] * fj.x, wpair[s] * fj.y);
      } else {
        for (int i0 = 0; i0 < W; i0++) {
          const R p0 = lane_bcast<R>(mine, i0);
          const R wr = p0 * fj.x, wi = p0 * fj.y;
          typename Acc::elem *base = atile + ((a[0] + i0) * E1 + a[1]) * E2 + a[2];
#pragma unroll
          for (int s = 0; s < SLOTS; s++)
            if (pv[s]) acc.add(base + off[s], wpair[s] * wr, wpair[s] * wi);
        }
      }
    } else {
      R ar = 0, ai = 0;
      if constexpr (D == 2) {
        const cplx<R> *base = tile + a[0] * E1 + a[1];
#pragma unroll
        for (int s = 0; s < SLOTS; s++)
          if (pv[s]) {
            const cplx<R> v = base[off[s]];
            ar += wpair[s] * v.x;
            ai += wpair[s] * v.y;
          }
      } else {
        for (int i0 = 0; i0 < W; i0++) {
          const R p0 = lane_bcast<R>(mine, i0);
          const cplx<R> *base = tile + ((a[0] + i0) * E1 + a[1]) * E2 + a[2];
          R sr = 0, si = 0;
#pragma unroll
          for (int s = 0; s < SLOTS; s++)
            if (pv[s]) {
              const cplx<R> v = base[off[s]];
              sr += wpair[s] * v.x;
              si += wpair[s] * v.y;
            }
          ar += p0 * sr;
          ai += p0 * si;
        }
      }
      ar = wave_sum<R>(ar);
      ai = wave_sum<R>(ai);
      if (lane == 0) dst[perm[p]] = cplx<R>{ar, ai};
    }
  }
  if constexpr (SPREAD) {
    __syncthreads();
    flush_tile<R, D>([&](int q) { return acc.load(atile + q); }, tiles, dst, lo, gg, tot,
                     sp.w, threadIdx.x, blockDim.x);
  }
}

// ======================================================= private-tile B^T
// d = 2, 3, one wavefront per work item and an LDS tile of its own: lanes
// own fixed tap pairs of the last two dimensions (as in k_bwave_tile), so
// within one instruction every lane touches a different cell, and plain
// read-modify-write replaces LDS atomics (exact FP32 sums, no ds_add_f32).
// Per node all W*SLOTS cells of the lane are loaded first, then updated and
// stored, so the LDS latency is paid once per node.  Instructions of one
// wave reach LDS in order, so a later node's loads see earlier stores.
template <typename R, int D, int W>
__global__ __launch_bounds__(64) void k_spread_wave_priv(const cplx<R> *__restrict__ fin,
                                                         cplx<R> *__restrict__ g,
                                                         const int *__restrict__ perm,
                                                         const int *__restrict__ cell,
                                                         const R *__restrict__ frac,
                                                         const int4 *__restrict__ subs,
                                                         BinGeom gg, WinParams<R> wp,
                                                         cplx<R> *__restrict__ tiles) {
  static_assert(D == 2 || D == 3, "private-tile kernels cover d = 2, 3");
  constexpr int PAIRS = W * W;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int ROWS = D == 3 ? W : 1;  // first-dimension taps looped per lane
  extern __shared__ __align__(16) unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  const int lane = threadIdx.x;
  for (int q = lane; q < tot; q += 64) tile[q] = cplx<R>{R(0), R(0)};
  __syncthreads();
  const int E1 = gg.E[1];
  const int E2 = D == 3 ? gg.E[2] : 1;
  int pa[SLOTS], pb[SLOTS], off[SLOTS];
  bool pv[SLOTS];
#pragma unroll
  for (int s = 0; s < SLOTS; s++) {
    const int q = lane + 64 * s;
    pv[s] = q < PAIRS;
    pa[s] = pv[s] ? q / W : 0;
    pb[s] = pv[s] ? q - (q / W) * W : 0;
    off[s] = D == 3 ? pa[s] * E2 + pb[s] : pa[s] * E1 + pb[s];
  }
  const int wt = lane / W < D ? lane / W : D - 1;
  const int ww = lane - (lane / W) * W;
  TapCoef<R, poly_deg_v<R>> tc;
  tc.load(wp.poly + (wt * W + ww) * (poly_deg_v<R> + 1));
  const int rowstride = D == 3 ? E1 * E2 : 0;
  // node data arrive 64 nodes at a time, one node per lane (coalesced), the
  // next chunk loading while the current one is processed; node k of a
  // chunk is broadcast from lane k
  int cc[D], nc[D];
  R cfr[D], nfr[D];
  cplx<R> cf, nf;
  auto load_chunk = [&](int first, int (&c)[D], R (&fr)[D], cplx<R> &fv) {
    const int i = first + lane;
    const bool ok = i < sp.z;
    const long p = (long)sp.y + (ok ? i : 0);
#pragma unroll
    for (int t = 0; t < D; t++) {
      c[t] = cell[t * gg.M + p];
      fr[t] = frac[t * gg.M + p];
    }
    fv = fin[perm[p]];
  };
  load_chunk(0, cc, cfr, cf);
  for (int first = 0; first < sp.z; first += 64) {
    if (first + 64 < sp.z) load_chunk(first + 64, nc, nfr, nf);
    const int cnt = min(64, sp.z - first);
    for (int k = 0; k < cnt; k++) {
    R frk[D];
    int a[D];
#pragma unroll
    for (int t = 0; t < D; t++) {
      frk[t] = lane_bcast<R>(cfr[t], k);
      a[t] = __builtin_amdgcn_readlane(cc[t], k) - lo[t];
    }
    const R frw = D == 3 ? (wt == 0 ? frk[0] : wt == 1 ? frk[1] : frk[D - 1])
                         : (wt == 0 ? frk[0] : frk[D - 1]);
    R mine = R(0);
    if (lane < D * W) mine = tc(frw);
    R wpair[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
      const R u = lane_fetch<R>(mine, (D - 2) * W + pa[s]);
      const R v = lane_fetch<R>(mine, (D - 1) * W + pb[s]);
      wpair[s] = pv[s] ? u * v : R(0);
    }
    const cplx<R> fj = cplx<R>{lane_bcast<R>(cf.x, k), lane_bcast<R>(cf.y, k)};
    cplx<R> *base = D == 3 ? tile + ((a[0] * E1 + a[1]) * E2 + a[2]) : tile + (a[0] * E1 + a[1]);
    // load every cell this lane updates for this node, then update, then store
    cplx<R> v[ROWS][SLOTS];
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int s = 0; s < SLOTS; s++)
        v[r][s] = pv[s] ? base[r * rowstride + off[s]] : cplx<R>{R(0), R(0)};
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const R p0 = D == 3 ? lane_bcast<R>(mine, r) : R(1);
      const R wr = p0 * fj.x, wi = p0 * fj.y;
#pragma unroll
      for (int s = 0; s < SLOTS; s++) {
        v[r][s].x += wpair[s] * wr;
        v[r][s].y += wpair[s] * wi;
      }
    }
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int s = 0; s < SLOTS; s++)
        if (pv[s]) base[r * rowstride + off[s]] = v[r][s];
    }
#pragma unroll
    for (int t = 0; t < D; t++) {
      cc[t] = nc[t];
      cfr[t] = nfr[t];
    }
    cf = nf;
  }
  __syncthreads();
  flush_tile<R, D>([&](int q) { return tile[q]; }, tiles, g, lo, gg, tot, sp.w, lane, 64);
}

// ================================================== row-partitioned B^T, d = 3
// NW wavefronts share one LDS tile and split it by the parity class of the
// first tile coordinate: wavefront w updates only tile rows e0 with
// e0 % NW == w, for every node of the work item.  Rows never overlap between
// wavefronts and pairs never overlap within one instruction, so plain FP32
// read-modify-write is race-free; each wavefront does 1/NW of the taps and
// the full per-node setup, and the tile is shared (NW times the occupancy of
// the private-tile kernel at the same LDS).  Invalid pair slots (lanes past
// W*W) read and rewrite a dummy cell past the tile instead of branching.
template <typename R, int D, int W, int NW>
__global__ __launch_bounds__(64 * NW) void k_spread_rows(const cplx<R> *__restrict__ fin,
                                                         cplx<R> *__restrict__ g,
                                                         const int *__restrict__ perm,
                                                         const int *__restrict__ cell,
                                                         const R *__restrict__ frac,
                                                         const int4 *__restrict__ subs,
                                                         BinGeom gg, WinParams<R> wp,
                                                         cplx<R> *__restrict__ tiles) {
  static_assert(D == 3 && W % NW == 0, "row-partitioned spreading covers d = 3");
  constexpr int PAIRS = W * W;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int RW = W / NW;  // rows per wavefront and node
  extern __shared__ __align__(16) unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[3];
  bin_corner<3>(sp.x, gg, lo);
  const int E1 = gg.E[1], E2 = gg.E[2];
  const int tot = gg.E[0] * E1 * E2;
  for (int q = threadIdx.x; q <= tot; q += blockDim.x) tile[q] = cplx<R>{R(0), R(0)};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int rowstride = E1 * E2;
  // this lane's pairs over dims 1, 2: tile offset within a row; invalid
  // slots point at the dummy cell (index tot) relative to the row base
  int pa[SLOTS], pb[SLOTS], off[SLOTS];
  bool pv[SLOTS];
#pragma unroll
  for (int s = 0; s < SLOTS; s++) {
    const int q = lane + 64 * s;
    pv[s] = q < PAIRS;
    pa[s] = pv[s] ? q / W : 0;
    pb[s] = pv[s] ? q - (q / W) * W : 0;
    off[s] = pa[s] * E2 + pb[s];
  }
  const int wt = lane / W < 3 ? lane / W : 2;
  const int ww = lane - (lane / W) * W;
  TapCoef<R, poly_deg_v<R>> tc;
  tc.load(wp.poly + (wt * W + ww) * (poly_deg_v<R> + 1));
  int cc[3], nc[3];
  R cfr[3], nfr[3];
  cplx<R> cf, nf;
  auto load_chunk = [&](int first, int (&c)[3], R (&fr)[3], cplx<R> &fv) {
    const int i = first + lane;
    const long p = (long)sp.y + (i < sp.z ? i : 0);
#pragma unroll
    for (int t = 0; t < 3; t++) {
      c[t] = cell[t * gg.M + p];
      fr[t] = frac[t * gg.M + p];
    }
    fv = fin[perm[p]];
  };
  load_chunk(0, cc, cfr, cf);
  for (int first = 0; first < sp.z; first += 64) {
    if (first + 64 < sp.z) load_chunk(first + 64, nc, nfr, nf);
    const int cnt = min(64, sp.z - first);
    for (int k = 0; k < cnt; k++) {
      R frk[3];
      int a[3];
#pragma unroll
      for (int t = 0; t < 3; t++) {
        frk[t] = lane_bcast<R>(cfr[t], k);
        a[t] = __builtin_amdgcn_readlane(cc[t], k) - lo[t];
      }
      const R frw = wt == 0 ? frk[0] : (wt == 1 ? frk[1] : frk[2]);
      const R mine = tc(frw);
      R wpair[SLOTS];
      int addr[SLOTS];
      const int row_base = (a[1] * E2 + a[2]);
#pragma unroll
      for (int s = 0; s < SLOTS; s++) {
        const R u = lane_fetch<R>(mine, W + pa[s]);
        const R v = lane_fetch<R>(mine, 2 * W + pb[s]);
        wpair[s] = pv[s] ? u * v : R(0);
        addr[s] = row_base + off[s];
      }
      const cplx<R> fj = cplx<R>{lane_bcast<R>(cf.x, k), lane_bcast<R>(cf.y, k)};
      // first of this wavefront's rows: smallest r >= 0 with (a0 + r) % NW == wv
      const int r0 = ((wv - a[0]) % NW + NW) % NW;
      cplx<R> v[RW][SLOTS];
#pragma unroll
      for (int j = 0; j < RW; j++) {
        const int e0 = a[0] + r0 + j * NW;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          const int idx = pv[s] ? e0 * rowstride + addr[s] : tot;
          v[j][s] = tile[idx];
        }
      }
#pragma unroll
      for (int j = 0; j < RW; j++) {
        // r0 < NW and j < W / NW, so every row r is a valid tap
        const R p0 = lane_bcast<R>(mine, r0 + j * NW);
        const R wr = p0 * fj.x, wi = p0 * fj.y;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          v[j][s].x += wpair[s] * wr;
          v[j][s].y += wpair[s] * wi;
        }
      }
#pragma unroll
      for (int j = 0; j < RW; j++) {
        const int e0 = a[0] + r0 + j * NW;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          const int idx = pv[s] ? e0 * rowstride + addr[s] : tot;
          tile[idx] = v[j][s];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
      cc[t] = nc[t];
      cfr[t] = nfr[t];
    }
    cf = nf;
  }
  __syncthreads();
  flush_tile<R, 3>([&](int q) { return tile[q]; }, tiles, g, lo, gg, tot, sp.w, threadIdx.x, blockDim.x);
}

}  // namespace nfft_hip
#include "spread_rows.hpp"
#include "spread_runs.hpp"
#include "spread_runs2d.hpp"
namespace nfft_hip {

// ==================================================================== runtime m
// Window values and wrapped tap coordinates live in per-thread arrays.
template <typename R, int D>
__device__ __forceinline__ void node_window_rt(const int *__restrict__ cell,
                                               const R *__restrict__ frac, long p,
                                               const BinGeom &gg, const WinParams<R> &wp,
                                               int (&c)[D], R (&psi)[D][kMaxW]) {
  const int W = 2 * gg.m + 2;
  constexpr int PD = poly_deg_v<R>;
  for (int t = 0; t < D; t++) {
    c[t] = cell[t * gg.M + p];
    const R fr = frac[t * gg.M + p];
    for (int w = 0; w < W; w++) psi[t][w] = tap_poly<R, PD>(wp.poly + (t * W + w) * (PD + 1), fr);
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
                                                   WinParams<R> wp, cplx<R> *__restrict__ tiles) {
  extern __shared__ __align__(16) unsigned char b_smem[];
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
    } else if constexpr (D == 3) {
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
    } else {
      // any d: odometer over the taps of the first d-1 dimensions, last
      // dimension innermost (the generic path of the reference,
      // kernel/nfft/nfft.c:630-1100)
      int lt[D];
      for (int t = 0; t < D; t++) lt[t] = 0;
      long nrows = 1;
      for (int t = 0; t < D - 1; t++) nrows *= W;
      for (long q = 0; q < nrows; q++) {
        R wgt = R(1);
        long row = 0;
        for (int t = 0; t < D - 1; t++) {
          wgt *= psi[t][lt[t]];
          row += idx[t][lt[t]] * stride[t];
        }
        R sr = 0, si = 0;
        for (int w = 0; w < W; w++) {
          const cplx<R> v = src[row + idx[D - 1][w]];
          sr += psi[D - 1][w] * v.x;
          si += psi[D - 1][w] * v.y;
        }
        ar += wgt * sr;
        ai += wgt * si;
        for (int t = D - 2; t >= 0; t--) {
          if (++lt[t] < W) break;
          lt[t] = 0;
        }
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
                                                   WinParams<R> wp, cplx<R> *__restrict__ tiles) {
  extern __shared__ __align__(16) unsigned char b_smem[];
  using Acc = LdsAcc<R>;
  typename Acc::elem *tile = reinterpret_cast<typename Acc::elem *>(b_smem);
  Acc acc;
  const int W = 2 * gg.m + 2;
  const int4 sp = subs[blockIdx.x];
  int lo[D];
  bin_corner<D>(sp.x, gg, lo);
  int tot = 1;
  if constexpr (TILE) {
    for (int t = 0; t < D; t++) tot *= gg.E[t];
    acc.set_bound(block_abs_sum<R>(fin, perm, sp.y, sp.z));
    for (int q = threadIdx.x; q < tot; q += blockDim.x) tile[q] = Acc::zero();
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
    // one tap: LDS accumulator (tile) or HBM atomic (global variant)
    auto put = [&](long o, R vr, R vi) {
      if constexpr (TILE) acc.add(tile + o, vr, vi);
      else global_add<R>(g + o, vr, vi);
    };
    if constexpr (D == 1) {
      for (int w = 0; w < W; w++) put(idx[0][w], psi[0][w] * fj.x, psi[0][w] * fj.y);
    } else if constexpr (D == 2) {
      for (int i0 = 0; i0 < W; i0++) {
        const long row = idx[0][i0] * stride[0];
        const R wr = psi[0][i0] * fj.x, wi = psi[0][i0] * fj.y;
        for (int w = 0; w < W; w++) put(row + idx[1][w], psi[1][w] * wr, psi[1][w] * wi);
      }
    } else if constexpr (D == 3) {
      for (int i0 = 0; i0 < W; i0++) {
        for (int i1 = 0; i1 < W; i1++) {
          const long row = idx[0][i0] * stride[0] + idx[1][i1] * stride[1];
          const R w01 = psi[0][i0] * psi[1][i1];
          const R wr = w01 * fj.x, wi = w01 * fj.y;
          for (int w = 0; w < W; w++) put(row + idx[2][w], psi[2][w] * wr, psi[2][w] * wi);
        }
      }
    } else {
      // any d: odometer over the first d-1 dimensions, last innermost
      int lt[D];
      for (int t = 0; t < D; t++) lt[t] = 0;
      long nrows = 1;
      for (int t = 0; t < D - 1; t++) nrows *= W;
      for (long q = 0; q < nrows; q++) {
        R wgt = R(1);
        long row = 0;
        for (int t = 0; t < D - 1; t++) {
          wgt *= psi[t][lt[t]];
          row += idx[t][lt[t]] * stride[t];
        }
        const R wr = wgt * fj.x, wi = wgt * fj.y;
        for (int w = 0; w < W; w++) put(row + idx[D - 1][w], psi[D - 1][w] * wr, psi[D - 1][w] * wi);
        for (int t = D - 2; t >= 0; t--) {
          if (++lt[t] < W) break;
          lt[t] = 0;
        }
      }
    }
  }
  if constexpr (TILE) {
    __syncthreads();
    flush_tile<R, D>([&](int q) { return acc.load(tile + q); }, tiles, g, lo, gg, tot,
                     sp.w, threadIdx.x, blockDim.x);
  }
}

// ============================================================ tile reduction
// Second half of the atomic-free B^T: one workgroup per bin writes every
// grid cell of the bin's core (bs_t cells per dimension, clipped at the grid
// edge) as the sum of the spreading tiles that cover it -- the tiles of this
// bin and of the neighbour bins within 1 + ceil(m / bs_t) bins, all of their
// work items.  Every grid cell is written exactly once, so the grid needs no
// zero fill and no atomics.  Requires E_t <= n_t (a tile never covers a cell
// twice); plans with smaller grids use the atomic flush instead.
constexpr int kMaxSpan = 40;

template <typename R, int D>
__global__ __launch_bounds__(256) void k_gather_tiles(const cplx<R> *__restrict__ tiles,
                                                      const int *__restrict__ sp_first,
                                                      const int *__restrict__ sp_count,
                                                      cplx<R> *__restrict__ g, BinGeom gg) {
  // per dimension: neighbour bin coordinates and the tile coordinate, in
  // that neighbour's tile, of this bin's first core cell
  __shared__ int nb_bt[3][kMaxSpan];
  __shared__ int nb_e[3][kMaxSpan];
  __shared__ int nb_span[3];
  int bc[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, cs[3] = {1, 1, 1}, E[3] = {1, 1, 1}, n[3] = {1, 1, 1},
      nbv[3] = {1, 1, 1};
  {
    int r = blockIdx.x;
#pragma unroll
    for (int t = D - 1; t >= 0; t--) {
      bc[t] = r % gg.nb[t];
      r /= gg.nb[t];
    }
  }
#pragma unroll
  for (int t = 0; t < D; t++) {
    lo[t] = bc[t] * gg.bs[t];
    cs[t] = min(gg.bs[t], gg.n[t] - lo[t]);
    E[t] = gg.E[t];
    n[t] = gg.n[t];
    nbv[t] = gg.nb[t];
  }
  if (threadIdx.x < 3) {
    const int t = threadIdx.x;
    if (t >= D) {
      nb_span[t] = 1;
      nb_bt[t][0] = 0;
      nb_e[t][0] = 0;
    } else {
      const int reach = 1 + (gg.m + gg.bs[t] - 1) / gg.bs[t];
      const bool all = 2 * reach + 1 >= gg.nb[t];
      const int span = all ? gg.nb[t] : 2 * reach + 1;
      nb_span[t] = span;
      for (int j = 0; j < span && j < kMaxSpan; j++) {
        const int bt = all ? j : ((bc[t] + j - reach) % gg.nb[t] + gg.nb[t]) % gg.nb[t];
        int e = (lo[t] - bt * gg.bs[t] + gg.m) % gg.n[t];
        nb_bt[t][j] = bt;
        nb_e[t][j] = e < 0 ? e + gg.n[t] : e;
      }
    }
  }
  __syncthreads();
  const int s0 = nb_span[0], s1 = nb_span[1], s2 = nb_span[2];
  const long tot = (long)E[0] * E[1] * E[2];
  const int core = cs[0] * cs[1] * cs[2];
  for (int q = threadIdx.x; q < core; q += blockDim.x) {
    const int k2 = q % cs[2];
    const int k1 = (q / cs[2]) % cs[1];
    const int k0 = q / (cs[2] * cs[1]);
    R sr = 0, si = 0;
    for (int j0 = 0; j0 < s0; j0++) {
      int e0 = nb_e[0][j0] + k0;
      e0 -= e0 >= n[0] ? n[0] : 0;
      if (e0 >= E[0]) continue;
      for (int j1 = 0; j1 < s1; j1++) {
        int e1 = nb_e[1][j1] + k1;
        e1 -= e1 >= n[1] ? n[1] : 0;
        if (e1 >= E[1]) continue;
        const int b01 = nb_bt[0][j0] * nbv[1] + nb_bt[1][j1];
        const long t01 = (long)e0 * E[1] + e1;
        for (int j2 = 0; j2 < s2; j2++) {
          int e2 = nb_e[2][j2] + k2;
          e2 -= e2 >= n[2] ? n[2] : 0;
          if (e2 >= E[2]) continue;
          const int nbl = b01 * nbv[2] + nb_bt[2][j2];
          const int cnt = sp_count[nbl];
          const cplx<R> *tp = tiles + (long)sp_first[nbl] * tot + t01 * E[2] + e2;
          for (int c = 0; c < cnt; c++) {
            const cplx<R> v = tp[(long)c * tot];
            sr += v.x;
            si += v.y;
          }
        }
      }
    }
    const long gi = ((long)(lo[0] + k0) * n[1] + (lo[1] + k1)) * n[2] + (lo[2] + k2);
    g[gi] = cplx<R>{sr, si};
  }
}

// Tile reduction for zone-ordered tiles: bin b's core receives, per
// dimension, the right zone of its left neighbour (core cells [0, m+1)),
// its own core zone, and the left zone of its right neighbour (core cells
// [bs-m, bs)); every work item of each of the 3^d neighbour bins contributes
// one contiguous zone.  The (zone, work item) pairs are listed in LDS; each
// thread owns CPT core cells, sums in registers and keeps up to 16 loads in
// flight; each grid cell is written once.  P workgroups share a bin, each
// summing CPT * 256 of its core cells, so the heavy neighbourhoods of dense
// regions (hundreds of contributions) are spread over P workgroups.
constexpr int kZonePairs = 512;  // pairs listed per pass
template <typename R, int D, int CPT, int P = 1>
__global__ __launch_bounds__(256) void k_gather_zones(const cplx<R> *__restrict__ tiles,
                                                      const int *__restrict__ sp_first,
                                                      const int *__restrict__ sp_count,
                                                      cplx<R> *__restrict__ g, BinGeom gg,
                                                      const int *__restrict__ order) {
  // right-hand side blockIdx.y of a batched launch
  tiles += (long)blockIdx.y * gg.bs_t;
  g += (long)blockIdx.y * gg.bs_g;
  constexpr int NZ = D == 1 ? 3 : (D == 2 ? 9 : 27);
  constexpr int BATCH = CPT >= 16 ? 1 : 16 / CPT;
  __shared__ long s_base[NZ];   // offset of the zone inside a tile
  __shared__ int s_first[NZ], s_cnt[NZ], s_start[NZ + 1];
  __shared__ int s_off[NZ][3], s_size[NZ][3];
  __shared__ long s_pbase[kZonePairs];
  __shared__ int s_pz[kZonePairs];
  int bc[3] = {0, 0, 0};
  {
    int r = order[blockIdx.x / P];  // heavy neighbourhoods first
#pragma unroll
    for (int t = D - 1; t >= 0; t--) {
      bc[t] = r % gg.nb[t];
      r /= gg.nb[t];
    }
  }
  long tot = 1;
#pragma unroll
  for (int t = 0; t < D; t++) tot *= gg.E[t];
  if (threadIdx.x < NZ) {
    // neighbour offset delta_t = digit - 1; its zone's place in the tile
    // and in this bin's core come from the plan's table
    int dig[3] = {1, 1, 1};
    int r = threadIdx.x;
#pragma unroll
    for (int t = D - 1; t >= 0; t--) {
      dig[t] = r % 3;
      r /= 3;
    }
    int nbl = 0;
#pragma unroll
    for (int t = 0; t < D; t++) {
      int nb = bc[t] + dig[t] - 1;
      nb += nb < 0 ? gg.nb[t] : (nb >= gg.nb[t] ? -gg.nb[t] : 0);
      nbl = nbl * gg.nb[t] + nb;
    }
    const int *zd = gg.zdesc + 8 * threadIdx.x;
    s_base[threadIdx.x] = zd[0];
#pragma unroll
    for (int t = 0; t < D; t++) {
      s_off[threadIdx.x][t] = zd[1 + t];
      s_size[threadIdx.x][t] = zd[4 + t];
    }
    s_first[threadIdx.x] = sp_first[nbl];
    s_cnt[threadIdx.x] = sp_count[nbl];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int z = 0; z < NZ; z++) {
      s_start[z] = acc;
      acc += s_cnt[z];
    }
    s_start[NZ] = acc;
  }
  __syncthreads();
  const int npairs = s_start[NZ];
  int core = 1;
#pragma unroll
  for (int t = 0; t < D; t++) core *= gg.bs[t];
  // this workgroup's slice of the core cells
  const int qbase = (int)(blockIdx.x % P) * CPT * (int)blockDim.x;
  int kc[CPT][3];
  R ar[CPT], ai[CPT];
#pragma unroll
  for (int i = 0; i < CPT; i++) {
    const int q = qbase + threadIdx.x + i * blockDim.x;
    int rem = q < core ? q : 0;
    kc[i][0] = kc[i][1] = kc[i][2] = 0;
#pragma unroll
    for (int t = D - 1; t >= 0; t--) {
      kc[i][t] = rem % gg.bs[t];
      rem /= gg.bs[t];
    }
    ar[i] = ai[i] = R(0);
  }
  for (int w0 = 0; w0 < npairs; w0 += kZonePairs) {
    const int nw = min(kZonePairs, npairs - w0);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
      const int pi = w0 + i;
      int z = 0;
      while (s_start[z + 1] <= pi) z++;
      s_pz[i] = z;
      s_pbase[i] = (long)(s_first[z] + pi - s_start[z]) * tot + s_base[z];
    }
    __syncthreads();
    for (int i0 = 0; i0 < nw; i0 += BATCH) {
      cplx<R> v[BATCH][CPT];
#pragma unroll
      for (int u = 0; u < BATCH; u++) {
        const int i = i0 + u;
        const int z = i < nw ? s_pz[i] : 0;
        const long pb = i < nw ? s_pbase[i] : 0;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
          const int q = qbase + threadIdx.x + c * blockDim.x;
          bool in = i < nw && q < core;
          int idx = 0;
#pragma unroll
          for (int t = 0; t < D; t++) {
            const int uu = kc[c][t] - s_off[z][t];
            in = in && uu >= 0 && uu < s_size[z][t];
            idx = idx * s_size[z][t] + uu;
          }
          v[u][c] = in ? tiles[pb + idx] : cplx<R>{R(0), R(0)};
        }
      }
#pragma unroll
      for (int u = 0; u < BATCH; u++)
#pragma unroll
        for (int c = 0; c < CPT; c++) {
          ar[c] += v[u][c].x;
          ai[c] += v[u][c].y;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < CPT; c++) {
    const int q = qbase + threadIdx.x + c * blockDim.x;
    if (q >= core) break;
    long gi = 0;
#pragma unroll
    for (int t = 0; t < D; t++) gi = gi * gg.n[t] + bc[t] * gg.bs[t] + kc[c][t];
    g[gi] = cplx<R>{ar[c], ai[c]};
  }
}

// zone reduction for d dimensions and core cells per bin; parts: workgroups
// per bin (parts x CPT x 256 >= core)
template <typename R>
void (*gather_zones_kernel(int d, long core, int &parts))(const cplx<R> *, const int *, const int *,
                                                         cplx<R> *, BinGeom, const int *) {
  parts = 1;
  // 3-D keeps one workgroup per bin (two measured slower on the kooshball:
  // 0.25 -> 0.33 ms); 2-D splits 1024-cell cores four ways (clustered 512^2:
  // 0.38 -> 0.32 ms per trafo + adjoint)
  if (d == 3)
    return core <= 512 ? k_gather_zones<R, 3, 2>
                       : (core <= 1024 ? k_gather_zones<R, 3, 4> : k_gather_zones<R, 3, 16>);
  if (d == 2) {
    if (core <= 256) return k_gather_zones<R, 2, 1>;
    if (core <= 1024) {
      parts = 4;
      return k_gather_zones<R, 2, 1, 4>;
    }
    return k_gather_zones<R, 2, 16>;
  }
  return core <= 1024 ? k_gather_zones<R, 1, 4> : k_gather_zones<R, 1, 16>;
}

template <typename R>
void launch_gather_tiles(const Engine<R> &e, cplx<R> *g, hipStream_t s, int nb, const BinGeom *gb,
                         cplx<R> *tiles) {
  long nbins = 1;
  for (int t = 0; t < e.geom.d; t++) nbins *= e.geom.nb[t];
  if (nb > 1) {
    // batched: the zone reduction reads the right-hand side from blockIdx.y
    using F = void (*)(const cplx<R> *, const int *, const int *, cplx<R> *, BinGeom, const int *);
    long core = 1;
    for (int t = 0; t < e.geom.d; t++) core *= e.geom.bs[t];
    int parts = 1;
    F fz = gather_zones_kernel<R>(e.geom.d, core, parts);
    hipLaunchKernelGGL(fz, dim3((unsigned)(nbins * parts), (unsigned)nb), dim3(256), 0, s, tiles,
                       e.sp_first,
                       e.sp_count, g, *gb, e.bin_order);
    return;
  }
  if (e.geom.zoned) {
    long core = 1;
    for (int t = 0; t < e.geom.d; t++) core *= e.geom.bs[t];
    if (core <= 4096) {
      using F = void (*)(const cplx<R> *, const int *, const int *, cplx<R> *, BinGeom,
                         const int *);
      int parts = 1;
      F fz = gather_zones_kernel<R>(e.geom.d, core, parts);
      hipLaunchKernelGGL(fz, dim3((unsigned)(nbins * parts)), dim3(256), 0, s, e.tiles, e.sp_first,
                         e.sp_count, g, e.geom, e.bin_order);
      return;
    }
  }
  auto fn = e.geom.d == 1 ? k_gather_tiles<R, 1> : e.geom.d == 2 ? k_gather_tiles<R, 2>
                                                                 : k_gather_tiles<R, 3>;
  hipLaunchKernelGGL(fn, dim3((unsigned)nbins), dim3(256), 0, s, e.tiles, e.sp_first, e.sp_count,
                     g, e.geom);
}

// ==================================================================== dispatch
template <typename R> struct BKernelTable {
  using fn = void (*)(const cplx<R> *, cplx<R> *, const int *, const int *, const R *,
                      const int4 *, BinGeom, WinParams<R>, cplx<R> *);
  fn interp_static[kMaxDim][kMaxStaticM + 1];
  fn interp_sym[kMaxDim][kMaxStaticM + 1];   // float only, shared tap table
  fn interp_pair[kMaxDim][kMaxStaticM + 1];  // float d = 2, 3: node pairs of one cell
  fn interp_pair_sym[kMaxDim][kMaxStaticM + 1];
  fn spread_static[kMaxDim][kMaxStaticM + 1];
  fn interp_wave[kMaxDim][kMaxStaticM + 1];  // d = 2, 3 only
  fn spread_wave[kMaxDim][kMaxStaticM + 1];
  fn spread_priv[kMaxDim][kMaxStaticM + 1];  // d = 2, 3 only, 64 threads
  fn spread_rows[kMaxStaticM + 1];            // d = 3 only, 128 threads
  // d = 3, float only, 128 threads: run merging, two wavefronts per tile,
  // row weights by DPP row broadcast; two node streams per wavefront (2sbc)
  // or one stream whose runs load and store their tile cells (2rs); runs of
  // kRunCells cells along dimension 2 (a bin edge of at least that many
  // there), else of one (2sbc1)
  fn spread_runs_2sbc[kMaxStaticM + 1];
  fn spread_runs_2sbc1[kMaxStaticM + 1];
  fn spread_runs_2rs[kMaxStaticM + 1];
  fn spread_runs2d[kMaxStaticM + 1];          // d = 2: run-merging, kRunWaves2d wavefronts
  fn spread_runs2d_1w[kMaxStaticM + 1];       // same, one wavefront (large tiles)
  fn spread_runs2d_mc[kMaxStaticM + 1];       // same, kRhsPerGroup right-hand sides each
  fn interp_rt[kMaxDim][2];  // [d-1][tile]
  fn spread_rt[kMaxDim][2];
};

template <typename R, int D, int M_>
void bk_fill_one(BKernelTable<R> &k) {
  k.interp_static[D - 1][M_] = k_interp_tile<R, D, 2 * M_ + 2>;
  if constexpr (std::is_same_v<R, float>) k.interp_sym[D - 1][M_] = k_interp_tile<R, D, 2 * M_ + 2, true>;
  if constexpr (std::is_same_v<R, float> && (D == 2 || D == 3)) {
    k.interp_pair[D - 1][M_] = k_interp_tile<R, D, 2 * M_ + 2, false, true>;
    k.interp_pair_sym[D - 1][M_] = k_interp_tile<R, D, 2 * M_ + 2, true, true>;
  }
  k.spread_static[D - 1][M_] = k_spread_tile<R, D, 2 * M_ + 2>;
  if constexpr (D >= 2 && D * (2 * M_ + 2) <= 64) {
    k.interp_wave[D - 1][M_] = k_bwave_tile<R, D, 2 * M_ + 2, false>;
    k.spread_wave[D - 1][M_] = k_bwave_tile<R, D, 2 * M_ + 2, true>;
    k.spread_priv[D - 1][M_] = k_spread_wave_priv<R, D, 2 * M_ + 2>;
  }
  if constexpr (D == 3 && D * (2 * M_ + 2) <= 64) k.spread_rows[M_] = k_spread_rows2<R, 2 * M_ + 2, 2>;
  if constexpr (D == 2) {
    // FP64 factors take twice the registers: smaller node groups
    constexpr int G2 = std::is_same_v<R, float> ? ((2 * M_ + 2) * (M_ + 1) > 64 ? 4 : 8)
                                               : ((2 * M_ + 2) * (M_ + 1) > 64 ? 2 : 4);
    k.spread_runs2d[M_] = k_spread_runs2d<R, 2 * M_ + 2, G2, 2, kRunWaves2d>;
    k.spread_runs2d_1w[M_] = k_spread_runs2d<R, 2 * M_ + 2, G2, 2, 1>;
    if constexpr (std::is_same_v<R, float>)
      k.spread_runs2d_mc[M_] = k_spread_runs2d_mc<2 * M_ + 2, G2, 2, kRhsPerGroup, kRunWaves2d>;
  }
  // (the row weights of a wavefront's W/2 rows fit one 16-lane DPP row)
  if constexpr (std::is_same_v<R, float> && D == 3 && D * (2 * M_ + 2) <= 64 && M_ + 1 <= 16) {
    // (runs of kRunCells cells for windows of up to 10 taps; wider windows
    // keep one, whose registers already fill the file)
    constexpr int GC = 2 * M_ + 2 <= kSlotQuadMaxW ? kRunCells : 1;
    k.spread_runs_2sbc[M_] = k_spread_runs<2 * M_ + 2, 2, 8, 2, false, GC>;
    k.spread_runs_2sbc1[M_] = k_spread_runs<2 * M_ + 2, 2, 8, 2, false, 1>;
    k.spread_runs_2rs[M_] = k_spread_runs<2 * M_ + 2, 2, 8, 1, true, GC>;
  }
}

template <typename R, int D, int... Ms>
void bk_fill_dim(BKernelTable<R> &k, std::integer_sequence<int, Ms...>) {
  (bk_fill_one<R, D, Ms + 1>(k), ...);
  k.interp_rt[D - 1][0] = k_interp_rt<R, D, false>;
  k.interp_rt[D - 1][1] = k_interp_rt<R, D, true>;
  k.spread_rt[D - 1][0] = k_spread_rt<R, D, false>;
  k.spread_rt[D - 1][1] = k_spread_rt<R, D, true>;
}

template <typename R, int D>
void bk_fill_rt(BKernelTable<R> &t) {
  t.interp_rt[D - 1][0] = k_interp_rt<R, D, false>;
  t.interp_rt[D - 1][1] = k_interp_rt<R, D, true>;
  t.spread_rt[D - 1][0] = k_spread_rt<R, D, false>;
  t.spread_rt[D - 1][1] = k_spread_rt<R, D, true>;
}

template <typename R>
BKernelTable<R> make_bkernel_table() {
  BKernelTable<R> t{};
  using seq = std::make_integer_sequence<int, kMaxStaticM>;
  bk_fill_dim<R, 1>(t, seq{});
  bk_fill_dim<R, 2>(t, seq{});
  bk_fill_dim<R, 3>(t, seq{});
  // d = 4 .. 6: the generic runtime-m kernels only
  bk_fill_rt<R, 4>(t);
  bk_fill_rt<R, 5>(t);
  bk_fill_rt<R, 6>(t);
  static_assert(kMaxDim == 6, "runtime kernels for every d");
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

// variant: 0 lane per node, 1 wave per node, 2 wave per node with a private
// tile (spreading only); falls back to what exists for (d, m)
template <typename R>
typename BKernelTable<R>::fn pick_bkernel(const BKernelTable<R> &k, const BinGeom &gg,
                                          bool spread, int variant, size_t &lds, int &threads,
                                          bool shared_poly = false) {
  size_t bytes = tile_bytes<R>(gg);
  // the lane-per-node 3-D interpolation tile has planes gg.ipp cells apart
  if (!spread && variant == 0 && gg.d == 3) bytes = (size_t)gg.E[0] * gg.ipp * sizeof(cplx<R>);
  const bool tile = bytes <= kTileBudget;
  lds = tile ? bytes : 0;
  threads = 256;
  if (tile && gg.d <= 3 && gg.m >= 1 && gg.m <= kMaxStaticM) {
    const int t = gg.d - 1;
    if (spread && variant == 14 && gg.d == 2 && k.spread_runs2d[gg.m]) {
      // a tile copy per wavefront while they fit 64 KB
      if (lds * kRunWaves2d > 64 * 1024) {
        threads = 64;
        return k.spread_runs2d_1w[gg.m];
      }
      threads = 64 * kRunWaves2d;
      lds *= kRunWaves2d;
      return k.spread_runs2d[gg.m];
    }
    if (spread && (variant == 10 || variant == 13) && gg.d == 3 && k.spread_runs_2sbc[gg.m]) {
      // the tile with its padded plane pitch; runs of one cell for a
      // dimension-2 bin edge below kRunCells (the load-store form exists for
      // kRunCells only)
      threads = 128;
      lds = (size_t)gg.E[0] * gg.sp_pp * sizeof(cplx<R>);
      if (gg.sp_gc == 1) return k.spread_runs_2sbc1[gg.m];
      return variant == 13 ? k.spread_runs_2rs[gg.m] : k.spread_runs_2sbc[gg.m];
    }
    if (spread && variant >= 3 && gg.d == 3 && k.spread_rows[gg.m]) {
      threads = 128;
      lds = bytes + 2 * sizeof(cplx<R>);  // two dummy cells past the tile
      return k.spread_rows[gg.m];
    }
    if (spread && variant >= 2 && k.spread_priv[t][gg.m]) {
      threads = 64;
      return k.spread_priv[t][gg.m];
    }
    if (variant >= 1) {
      auto fn = spread ? k.spread_wave[t][gg.m] : k.interp_wave[t][gg.m];
      if (fn) return fn;
    }
    if (!spread && variant == 0 && gg.slots && k.interp_pair[t][gg.m])
      return shared_poly ? k.interp_pair_sym[t][gg.m] : k.interp_pair[t][gg.m];
    if (!spread && shared_poly && k.interp_sym[t][gg.m]) return k.interp_sym[t][gg.m];
    return spread ? k.spread_static[t][gg.m] : k.interp_static[t][gg.m];
  }
  return spread ? k.spread_rt[gg.d - 1][tile ? 1 : 0] : k.interp_rt[gg.d - 1][tile ? 1 : 0];
}

template <typename R>
const BKernelTable<R> &bkernels() {
  static BKernelTable<R> k = make_bkernel_table<R>();
  return k;
}

template <typename R>
void launch_interp(const cplx<R> *g, cplx<R> *f, const Engine<R> &e, hipStream_t s, int nb,
                   const BinGeom *gb) {
  if (e.nsub == 0) return;
  const BinGeom &gg = gb ? *gb : e.geom;
  size_t lds;
  int threads;
  auto fn = pick_bkernel<R>(bkernels<R>(), gg, false, e.interp_variant, lds, threads,
                            e.win.poly_shared != 0 && e.sym_window);
  hipLaunchKernelGGL(fn, dim3((unsigned)e.nsub, (unsigned)nb), dim3(threads), lds, s, g, f, e.perm,
                     e.cell, e.frac, e.subs, gg, e.win, (cplx<R> *)nullptr);
}

template <typename R>
void launch_spread(const cplx<R> *f, cplx<R> *g, const Engine<R> &e, hipStream_t s, int nb,
                   const BinGeom *gb, cplx<R> *tiles) {
  if (e.sp_nsub == 0) return;
  const BinGeom &gg = gb ? *gb : e.geom;
  size_t lds;
  int threads;
  auto fn = pick_bkernel<R>(bkernels<R>(), gg, true, e.spread_variant, lds, threads);
  // the scratch reduction needs an LDS tile; the global variant adds directly
  if (!tiles) tiles = (lds > 0 && e.tile_scratch) ? e.tiles : nullptr;
  if (nb > 1 && gg.d == 2 && e.spread_variant == 14 && tiles && gg.zoned &&
      fn == bkernels<R>().spread_runs2d[gg.m] && lds * kRhsPerGroup <= 64 * 1024 &&
      bkernels<R>().spread_runs2d_mc[gg.m]) {
    // 2-D batches: kRhsPerGroup right-hand sides per workgroup share the
    // node work (scratch tiles per right-hand side as in the single launch)
    BinGeom gm = gg;
    gm.nrhs = nb;
    hipLaunchKernelGGL(bkernels<R>().spread_runs2d_mc[gg.m],
                       dim3((unsigned)e.sp_nsub, (unsigned)((nb + kRhsPerGroup - 1) / kRhsPerGroup)),
                       dim3(threads), lds * kRhsPerGroup, s, f, g, e.perm, e.cell, e.frac,
                       e.sp_subs, gm, e.win, tiles);
    return;
  }
  hipLaunchKernelGGL(fn, dim3((unsigned)e.sp_nsub, (unsigned)nb), dim3(threads), lds, s, f, g,
                     e.perm, e.cell, e.frac, e.sp_subs, gg, e.win, tiles);
}

// Whether the batched launches above cover this plan: the kernels that read
// their right-hand side from blockIdx.y are the FP32 defaults (lane-per-node
// interpolation, the run-merging spreading kernels with scratch tiles) and
// the zone reduction.
template <typename R>
bool batch_kernels(const Engine<R> &e) {
  if constexpr (!std::is_same_v<R, float>) {
    return false;
  } else {
    const BinGeom &gg = e.geom;
    if (!(e.tile_scratch && gg.zoned && gg.m >= 1 && gg.m <= kMaxStaticM)) return false;
    if (tile_bytes<R>(gg) > kTileBudget || e.interp_variant != 0) return false;
    if (gg.d == 3)
      return (e.spread_variant == 10 && bkernels<R>().spread_runs_2sbc[gg.m]) ||
             (e.spread_variant == 13 && bkernels<R>().spread_runs_2rs[gg.m]);
    if (gg.d == 2) return e.spread_variant == 14 && bkernels<R>().spread_runs2d[gg.m];
    return false;
  }
}

}  // namespace nfft_hip
