// Row-partitioned B^T for d = 3, FP32 read-modify-write without atomics.
//
// NW wavefronts share one LDS tile and split it by the residue class of the
// first tile coordinate: wavefront w updates only tile rows e0 with
// e0 % NW == w, for every node of the work item.  Rows never overlap
// between wavefronts, and within one instruction every lane owns different
// cells, so plain read-modify-write is race-free and FP32 sums are exact
// (no ds_add_f32, which is slow on this GPU).
//
// Lane layout: a lane owns one tap of dimension 1 and a pair of adjacent taps
// of dimension 2 (W * W/2 pair slots over the wavefront), so each update of
// a row is one paired 8-byte LDS read and write.  Invalid slots (lanes past
// the last pair) read and rewrite a dummy cell just past the tile.
//
// Node data arrive 64 nodes at a time, one node per lane (coalesced), the
// next chunk loading while the current one is processed; node k of a chunk
// is broadcast from lane k.  The d*W window values of a node are computed
// once per wavefront, one per lane, and fetched with ds_bpermute / readlane.
#pragma once

#include "engine.hpp"
#include "window.hpp"

namespace nfft_hip {

template <typename R, int W, int NW>
__global__ __launch_bounds__(64 * NW) void k_spread_rows2(const cplx<R> *__restrict__ fin,
                                                          cplx<R> *__restrict__ g,
                                                          const int *__restrict__ perm,
                                                          const int *__restrict__ cell,
                                                          const R *__restrict__ frac,
                                                          const int4 *__restrict__ subs,
                                                          BinGeom gg, WinParams<R> wp,
                                                          cplx<R> *__restrict__ tiles) {
  static_assert(W % 2 == 0 && W % NW == 0, "W = 2m+2 rows split evenly");
  constexpr int HW = W / 2;          // tap pairs along dimension 2
  constexpr int PAIRS = W * HW;      // (dim-1 tap, dim-2 pair) slots
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int RW = W / NW;         // rows per wavefront and node
  extern __shared__ __align__(16) unsigned char b_smem[];
  cplx<R> *tile = reinterpret_cast<cplx<R> *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[3];
  bin_corner<3>(sp.x, gg, lo);
  const int E1 = gg.E[1], E2 = gg.E[2];
  const int tot = gg.E[0] * E1 * E2;
  // the tile plus two dummy cells
  for (int q = threadIdx.x; q < tot + 2; q += blockDim.x) tile[q] = cplx<R>{R(0), R(0)};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int rowstride = E1 * E2;
  int pa[SLOTS], pb[SLOTS], off[SLOTS];
  bool pv[SLOTS];
#pragma unroll
  for (int s = 0; s < SLOTS; s++) {
    const int q = lane + 64 * s;
    pv[s] = q < PAIRS;
    pa[s] = pv[s] ? q / HW : 0;
    pb[s] = pv[s] ? 2 * (q - (q / HW) * HW) : 0;
    off[s] = pa[s] * E2 + pb[s];
  }
  // window value owned by this lane: dimension wt, tap ww
  const int wt = lane / W < 3 ? lane / W : 2;
  const int ww = lane - (lane / W) * W;
  // this lane's window polynomial (tap ww of dimension wt) in registers
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
      const R fr0 = lane_bcast<R>(cfr[0], k);
      const R fr1 = lane_bcast<R>(cfr[1], k);
      const R fr2 = lane_bcast<R>(cfr[2], k);
      const int a0 = __builtin_amdgcn_readlane(cc[0], k) - lo[0];
      const int a1 = __builtin_amdgcn_readlane(cc[1], k) - lo[1];
      const int a2 = __builtin_amdgcn_readlane(cc[2], k) - lo[2];
      const R frw = wt == 0 ? fr0 : (wt == 1 ? fr1 : fr2);
      const R mine = tc(frw);
      // pair weights psi1[pa] * psi2[pb], psi1[pa] * psi2[pb+1]
      R w0[SLOTS], w1[SLOTS];
      int addr[SLOTS];
      const int row_base = a1 * E2 + a2;
#pragma unroll
      for (int s = 0; s < SLOTS; s++) {
        const R u = lane_fetch<R>(mine, W + pa[s]);
        const R v0 = lane_fetch<R>(mine, 2 * W + pb[s]);
        const R v1 = lane_fetch<R>(mine, 2 * W + pb[s] + 1);
        w0[s] = pv[s] ? u * v0 : R(0);
        w1[s] = pv[s] ? u * v1 : R(0);
        addr[s] = row_base + off[s];
      }
      const R fx = lane_bcast<R>(cf.x, k);
      const R fy = lane_bcast<R>(cf.y, k);
      // first of this wavefront's rows: smallest r >= 0 with (a0 + r) % NW == wv
      const int r0 = ((wv - a0) % NW + NW) % NW;
      cplx<R> v[RW][SLOTS][2];
#pragma unroll
      for (int j = 0; j < RW; j++) {
        const int rb = (a0 + r0 + j * NW) * rowstride;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          const int idx = pv[s] ? rb + addr[s] : tot;
          v[j][s][0] = tile[idx];
          v[j][s][1] = tile[idx + 1];
        }
      }
#pragma unroll
      for (int j = 0; j < RW; j++) {
        // r0 < NW and j < W / NW, so every row is a valid tap
        const R p0 = lane_bcast<R>(mine, r0 + j * NW);
        const R wr = p0 * fx, wi = p0 * fy;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          v[j][s][0].x += w0[s] * wr;
          v[j][s][0].y += w0[s] * wi;
          v[j][s][1].x += w1[s] * wr;
          v[j][s][1].y += w1[s] * wi;
        }
      }
#pragma unroll
      for (int j = 0; j < RW; j++) {
        const int rb = (a0 + r0 + j * NW) * rowstride;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
          const int idx = pv[s] ? rb + addr[s] : tot;
          tile[idx] = v[j][s][0];
          tile[idx + 1] = v[j][s][1];
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
