// B^T for d = 3 in FP32 that sums runs of nodes sharing one grid cell in
// registers before touching the LDS tile.
//
// Nodes of a work item arrive sorted by cell (the binning sorts by a
// bin-major cell key), so nodes with the same tap window are adjacent.  For a
// run of such nodes every tap of the window receives sum_j f_j psi_j(tap);
// each lane keeps its taps' partial sums in registers and the tile sees one
// read-modify-write per run instead of one per node.  Dense sampling (the
// centre of radial trajectories) puts several nodes in most occupied cells.
//
// NW wavefronts share one LDS tile and wavefront w updates only tile cells
// whose dimension-1 index e1 has e1 % NW == w (W / NW of a node's window
// rows), so plain FP32 read-modify-write is race-free (no ds_add_f32, which
// is slow on this GPU).  A lane owns one tap of dimension 0 and a pair of
// adjacent taps of dimension 2; lanes past the last pair are masked off.
// (Three wavefronts per tile, the rows split 4/3/3 with the missing row
// masked, fit 150 VGPRs and three waves per SIMD but repeat the window work
// once more per node: B^T 4.03 -> 4.80 ms on the 3-D kooshball.)
//
// Window values: G nodes at a time, each lane evaluates one tap (dimension
// wt, tap ww) of one node as a per-tap polynomial (two nodes per pass when
// 3 W <= 32), with the node's fractional coordinates and value fetched by
// ds_bpermute from the lane that loaded the node.  A lane's pair weights
// psi0 * psi2 are fetched the same way; the row weights psi1 * f_j of each
// node are gathered once by ds_bpermute into lanes 0..RW-1 of every 16-lane
// row and the FMAs read them with a DPP row broadcast.
#pragma once

#include "engine.hpp"
#include "window.hpp"

namespace nfft_hip {

__device__ __forceinline__ float bperm_f(int byte_addr, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
}

// acc += x[lane J of this lane's 16-lane row] * w: the row-broadcast DPP
// form of v_fmac_f32 (row_newbcast, gfx90a and later), so a value that every
// lane needs is read from a register row without a readlane per use.
template <int J>
__device__ __forceinline__ void fmac_row_bcast(float &acc, float x, float w) {
  asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(x), "v"(w), "n"(J));
}

// The RW rows of one node: acc[j] += (sr, si)[row j] * (w0, w1) for the
// lane's two pair cells, the row weights taken from lanes j of every row.
// (A run's first node setting the accumulators with v_mul_f32_dpp instead of
// clearing them at each flush put the FMAs of every node behind a branch and
// measured slower: B^T 4.41 -> 4.83 ms on the 3-D kooshball.)
template <int J, int RW, int SLOTS>
__device__ __forceinline__ void rows_bcast(float2 (&acc)[RW][SLOTS][2], float br, float bi,
                                           const float (&w0)[SLOTS], const float (&w1)[SLOTS]) {
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    fmac_row_bcast<J>(acc[J][sl][0].x, br, w0[sl]);
    fmac_row_bcast<J>(acc[J][sl][0].y, bi, w0[sl]);
    fmac_row_bcast<J>(acc[J][sl][1].x, br, w1[sl]);
    fmac_row_bcast<J>(acc[J][sl][1].y, bi, w1[sl]);
  }
  if constexpr (J + 1 < RW) rows_bcast<J + 1, RW, SLOTS>(acc, br, bi, w0, w1);
}

// S node streams per wavefront: stream s walks its own part of the work
// item's nodes and the updates of the streams interleave node by node, which
// gives the wavefront S independent dependency chains.  Streams of one
// wavefront may touch the same tile cells; a wavefront's LDS operations
// complete in program order, so their read-modify-writes do not race.
// RS (one stream): a run loads its cells of the tile into the accumulators
// when it starts and stores them back when it ends, with no add and no
// clearing; no other stream of the wavefront writes those cells meanwhile
// and the other wavefronts own other rows, so the load-store pair is
// race-free.
// GC: a run is the nodes of GC adjacent cells along dimension 2 (same
// dimension-0 and dimension-1 cell, dimension-2 cells b .. b+GC-1 of the bin
// with b = min(GC floor(c / GC), bs - GC), so that the group's taps stay in
// the tile row), whose windows fit W + GC - 1 dimension-2 taps; a node at
// dimension-2 offset delta in its group takes its taps from lane tap - delta.
// With GC = 3 a run holds the nodes of three cells, so runs are about three
// times as long and flush a third as often, and for even W the (W + 2) / 2
// pairs cover the W + 2 taps exactly (needs a bin edge of 3 or more).
template <int W, int NW, int G, int S, bool RS = false, int GC = 1>
__global__ __launch_bounds__(64 * NW) void k_spread_runs(const float2 *__restrict__ fin,
                                                         float2 *__restrict__ g,
                                                         const int *__restrict__ perm,
                                                         const int *__restrict__ cell,
                                                         const float *__restrict__ frac,
                                                         const int4 *__restrict__ subs,
                                                         BinGeom gg, WinParams<float> wp,
                                                         float2 *__restrict__ tiles) {
  static_assert(W % 2 == 0 && W % NW == 0 && 3 * W <= 64, "W = 2m+2 taps, rows split evenly");
  static_assert(GC >= 1 && GC <= 3, "runs of one to three cells");
  // the dimension-2 pairs cover the run's taps exactly (no cell past them)
  constexpr bool FULL = (W + GC - 1) % 2 == 0;
  constexpr int HW = (W + GC) / 2;  // dimension-2 pairs covering W + GC - 1 taps
  constexpr int PAIRS = W * HW;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int RW = W / NW;
  // (Evaluating one node per pass with its coordinates and value taken by
  // readlane, every 16-lane row holding the dimension-1 taps so that the
  // FMAs read the row weights in place, replaced 4.5 of the 7.5 ds_bpermute
  // per node and wavefront but measured slower: 4.20 -> 5.35 ms.)
  constexpr int NP = 3 * W <= 32 ? 2 : 1;  // nodes per window pass
  constexpr int NPASS = G / NP;
  constexpr int ZL = 3 * W;                 // a window lane that holds 0
  static_assert(G % NP == 0 && 64 % G == 0, "groups tile the 64-node chunk");
  static_assert(RW <= 16, "a wavefront's rows fit one 16-lane DPP row");
  static_assert(!RS || S == 1, "run load-store needs one stream per wavefront");
  static_assert(SLOTS <= 3, "lane table covers three slots");
  // right-hand side blockIdx.y of a batched launch
  fin += (long)blockIdx.y * gg.bs_f;
  g += (long)blockIdx.y * gg.bs_g;
  if (tiles) tiles += (long)blockIdx.y * gg.bs_t;
  extern __shared__ __align__(16) unsigned char b_smem[];
  float2 *tile = reinterpret_cast<float2 *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[3];
  bin_corner<3>(sp.x, gg, lo);
  // LDS layout: cell (e0, e1, e2) at e0 PP + e1 E2 + e2, the plane pitch PP
  // padded so that the lanes of one row update spread over the LDS banks
  const int E2 = gg.E[2];
  const int PP = gg.sp_pp;
  const int plane = gg.E[1] * E2;
  const int tot = gg.E[0] * plane;
  const int ptot = gg.E[0] * PP;
  for (int q = threadIdx.x; q < ptot; q += blockDim.x) tile[q] = float2{0.f, 0.f};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // (dimension-0 tap, dimension-2 pair) slots of this lane from the plan's
  // lane table; c1[sl]: the pair's second cell lies inside the run's taps
  int off[SLOTS];
  bool pv[SLOTS], c1[SLOTS];
  // (adr_t[k]: the lane holding dimension-2 tap pb + 1 - k, so that a node at
  // offset delta in its group reads its pair's taps from adr_t[delta + 1]
  // and adr_t[delta]; separate named arrays, so that choosing between them
  // stays a select and never becomes a dynamically indexed private array)
  int adr_u[NP][SLOTS], adr_p1[NP][SLOTS], adr_0[NP][SLOTS], adr_m1[NP][SLOTS],
      adr_m2[NP][SLOTS];
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = gg.sp_lane[lane + 64 * sl];
    pv[sl] = q < PAIRS;
    const int pa = pv[sl] ? q / HW : 0;
    const int pb = pv[sl] ? 2 * (q - (q / HW) * HW) : 0;
    c1[sl] = pv[sl] && (FULL || pb + 1 < W + GC - 1);
    off[sl] = pa * PP + pb;
#pragma unroll
    for (int h = 0; h < NP; h++) {
      adr_u[h][sl] = 4 * (h * 32 + pa);
      auto tap_adr = [&](int t) { return 4 * (h * 32 + (t >= 0 && t < W ? 2 * W + t : ZL)); };
      adr_p1[h][sl] = tap_adr(pb + 1);
      adr_0[h][sl] = tap_adr(pb);
      adr_m1[h][sl] = tap_adr(pb - 1);
      adr_m2[h][sl] = tap_adr(pb - 2);
    }
  }
  // window tap evaluated by this lane in a pass: node half h, dimension wt,
  // tap ww; lanes without a tap evaluate 0 (all coefficients 0)
  const int h = NP == 2 ? lane >> 5 : 0;
  const int li = NP == 2 ? lane & 31 : lane;
  const bool tap_lane = li < 3 * W;
  const int wt = tap_lane ? li / W : 0;
  const int ww = tap_lane ? li - wt * W : 0;
  float coef[kPolyCoef];
  {
    const float *pc = wp.poly + (wt * W + ww) * kPolyCoef;
#pragma unroll
    for (int k = 0; k < kPolyCoef; k++) coef[k] = tap_lane ? pc[k] : 0.f;
  }

  // streams: node ranges [sb, sb + sl) of the work item
  int sb[S], sn[S];
  {
    const int part = (sp.z + S - 1) / S;
#pragma unroll
    for (int st = 0; st < S; st++) {
      sb[st] = min(sp.z, st * part);
      sn[st] = min(sp.z, (st + 1) * part) - sb[st];
    }
  }
  // run accumulators per stream: [row j][slot][pair cell], complex
  float2 acc[S][RW][SLOTS][2];
#pragma unroll
  for (int st = 0; st < S; st++)
#pragma unroll
    for (int j = 0; j < RW; j++)
#pragma unroll
      for (int sl = 0; sl < SLOTS; sl++) acc[st][j][sl][0] = acc[st][j][sl][1] = float2{0.f, 0.f};
  // current run per stream: packed key (group corner | delta << 20 | r0 <<
  // 21, r0 the first of the node's dimension-1 window rows that this
  // wavefront owns, delta the node's dimension-2 offset in its group; runs
  // compare without delta), from the stream's first node on (a stream
  // without nodes holds the item's first node's corner and only ever adds
  // zeros there)
  constexpr int kDelta = 3 << 20;
  int cur[S];
  auto cell_at = [&](int st, int sl, int j) {
    return tile + (cur[st] & 0xfffff) + ((cur[st] >> 22) + j * NW) * E2 + off[sl];
  };
  // the run's sums into the tile (read-modify-write), accumulators cleared
  auto flush = [&](int st) {
    // all reads, then all writes: the rows are distinct cells
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++) {
      if (pv[sl]) {
        float2 v[RW][2];
#pragma unroll
        for (int j = 0; j < RW; j++) {
          const float2 *t = cell_at(st, sl, j);
          v[j][0] = t[0];
          v[j][1] = FULL || c1[sl] ? t[1] : float2{0.f, 0.f};
        }
#pragma unroll
        for (int j = 0; j < RW; j++) {
          float2 *t = cell_at(st, sl, j);
          t[0] = float2{v[j][0].x + acc[st][j][sl][0].x, v[j][0].y + acc[st][j][sl][0].y};
          if (FULL || c1[sl])
            t[1] = float2{v[j][1].x + acc[st][j][sl][1].x, v[j][1].y + acc[st][j][sl][1].y};
        }
      }
#pragma unroll
      for (int j = 0; j < RW; j++) acc[st][j][sl][0] = acc[st][j][sl][1] = float2{0.f, 0.f};
    }
  };
  // RS: the run's sums are the tile's values; store them, and load the next
  // run's (after the store in program order: the runs' cells overlap)
  auto store = [&](int st) {
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++)
      if (pv[sl]) {
#pragma unroll
        for (int j = 0; j < RW; j++) {
          float2 *t = cell_at(st, sl, j);
          t[0] = acc[st][j][sl][0];
          if (FULL || c1[sl]) t[1] = acc[st][j][sl][1];
        }
      }
  };
  auto load = [&](int st) {
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++)
      if (pv[sl]) {
#pragma unroll
        for (int j = 0; j < RW; j++) {
          const float2 *t = cell_at(st, sl, j);
          acc[st][j][sl][0] = t[0];
          acc[st][j][sl][1] = FULL || c1[sl] ? t[1] : float2{0.f, 0.f};
        }
      }
  };

  // chunk of 64 nodes per stream, one per lane: packed corner, frac - 1/2,
  // value.  Loads run two chunks ahead for the node index (perm) and one
  // chunk ahead for everything else, and are converted only when the chunk
  // becomes current, so no wait for memory falls inside a chunk's work.
  // Lanes past the stream's end repeat its last node with value zero: they
  // stay in the last run and add nothing, so the node loop needs no bounds
  // test per node.
  auto pos = [&](int st, int first) {
    // a valid node of the work item for every lane (an empty stream reads
    // the item's first node)
    const int i = first + lane;
    return (long)sp.y + (sn[st] > 0 ? sb[st] + (i < sn[st] ? i : sn[st] - 1) : 0);
  };
  int ckey[S];
  float cfr[S][3];
  float2 cf[S];
  int nc[S][3];
  float nfr[S][3];
  float2 nf[S];
  int nperm[S], nnperm[S];
  auto load_raw = [&](int st, int first, int pidx) {
    const long p = pos(st, first);
#pragma unroll
    for (int t = 0; t < 3; t++) {
      nc[st][t] = cell[t * gg.M + p];
      nfr[st][t] = frac[t * gg.M + p];
    }
    nf[st] = fin[pidx];
  };
  auto convert = [&](int st, int first) {
    int a[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      a[t] = nc[st][t] - lo[t];
      cfr[st][t] = nfr[st][t] - 0.5f;
    }
    // the dimension-1 cell determines r0, so keys of one wavefront compare
    // as corners
    const int r0 = ((wv - a[1]) % NW + NW) % NW;
    const int g2 = GC == 1 ? a[2] : min(a[2] - a[2] % GC, gg.bs[2] - GC);
    const int dl = a[2] - g2;
    ckey[st] = (a[0] * PP + a[1] * E2 + g2) | (dl << 20) | (r0 << 22);
    cf[st] = first + lane < sn[st] ? nf[st] : float2{0.f, 0.f};
  };
  int longest = 0;
#pragma unroll
  for (int st = 0; st < S; st++) {
    load_raw(st, 0, perm[pos(st, 0)]);
    nperm[st] = perm[pos(st, 64)];
    nnperm[st] = 0;
    convert(st, 0);
    cur[st] = __builtin_amdgcn_readfirstlane(ckey[st]) & ~kDelta;
    longest = max(longest, sn[st]);
  }
  for (int first = 0; first < longest; first += 64) {
    int cnt[S];
    int cmax = 0;
#pragma unroll
    for (int st = 0; st < S; st++) {
      if (first + 64 < sn[st]) {
        load_raw(st, first + 64, nperm[st]);
        nnperm[st] = perm[pos(st, first + 128)];
      }
      cnt[st] = max(0, min(64, sn[st] - first));
      cmax = max(cmax, cnt[st]);
    }
    for (int base = 0; base < cmax; base += G) {
      // window values of nodes base .. base+G-1 of every stream (lanes past
      // the stream's count: unused)
      float xr[S][NPASS], xi[S][NPASS];
      float w0[S][G][SLOTS], w1[S][G][SLOTS];
      int gkey[S][G];
#pragma unroll
      for (int st = 0; st < S; st++) {
#pragma unroll
        for (int k = 0; k < G; k++) gkey[st][k] = __builtin_amdgcn_readlane(ckey[st], base + k);
        float mine[NPASS];
#pragma unroll
        for (int p = 0; p < NPASS; p++) {
          const int src = 4 * (base + NP * p + h);
          const float f0 = bperm_f(src, cfr[st][0]);
          const float f1 = bperm_f(src, cfr[st][1]);
          const float f2 = bperm_f(src, cfr[st][2]);
          const float fx = bperm_f(src, cf[st].x);
          const float fy = bperm_f(src, cf[st].y);
          const float tt = wt == 0 ? f0 : (wt == 1 ? f1 : f2);
          float v = coef[kPolyDeg];
#pragma unroll
          for (int k = kPolyDeg - 1; k >= 0; k--) v = fmaf(v, tt, coef[k]);
          mine[p] = v;
          xr[st][p] = v * fx;
          xi[st][p] = v * fy;
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
          const int p = k / NP, hk = k % NP;
          const int dl = GC == 1 ? 0 : (gkey[st][k] >> 20) & 3;
#pragma unroll
          for (int sl = 0; sl < SLOTS; sl++) {
            const float u = bperm_f(adr_u[hk][sl], mine[p]);
            const int a0 = dl == 0 ? adr_0[hk][sl] : (dl == 1 ? adr_m1[hk][sl] : adr_m2[hk][sl]);
            const int a1 = dl == 0 ? adr_p1[hk][sl] : (dl == 1 ? adr_0[hk][sl] : adr_m1[hk][sl]);
            w0[st][k][sl] = u * bperm_f(a0, mine[p]);
            w1[st][k][sl] = u * bperm_f(a1, mine[p]);
          }
        }
      }
      // row weights of the group's nodes, lane q of every 16-lane row
      // holding dimension-1 row r0 + q NW of the node
      float brr[S][G], bri[S][G];
      {
        const int qoff = 4 * (W + NW * (lane & 15));
#pragma unroll
        for (int st = 0; st < S; st++)
#pragma unroll
          for (int k = 0; k < G; k++) {
            const int p = k / NP, hk = k % NP;
            const int src = ((gkey[st][k] >> 22) << 2) + 4 * hk * 32 + qoff;
            brr[st][k] = bperm_f(src, xr[st][p]);
            bri[st][k] = bperm_f(src, xi[st][p]);
          }
      }
#pragma unroll
      for (int k = 0; k < G; k++) {
#pragma unroll
        for (int st = 0; st < S; st++) {
          const int key = gkey[st][k] & ~kDelta;
          float a0[SLOTS], a1[SLOTS];
#pragma unroll
          for (int sl = 0; sl < SLOTS; sl++) {
            a0[sl] = w0[st][k][sl];
            a1[sl] = w1[st][k][sl];
          }
          // most nodes continue their run: the flush is placed off the
          // straight-line path, so a node without one takes no branch
          if (__builtin_expect(key != cur[st], 0)) {
            if constexpr (RS) {
              store(st);
              cur[st] = key;
              load(st);
            } else {
              flush(st);
              cur[st] = key;
            }
          }
          rows_bcast<0, RW, SLOTS>(acc[st], brr[st][k], bri[st][k], a0, a1);
        }
      }
    }
#pragma unroll
    for (int st = 0; st < S; st++) {
      if (first + 64 < sn[st]) {
        convert(st, first + 64);
        nperm[st] = nnperm[st];
      } else {
        // stream done: zero-valued copies of its last node (its open run)
        ckey[st] = __builtin_amdgcn_readlane(ckey[st], 63);
        cf[st] = float2{0.f, 0.f};
      }
    }
  }
#pragma unroll
  for (int st = 0; st < S; st++) {
    if constexpr (RS) store(st);
    else flush(st);
  }
  __syncthreads();
  // tile cells in the unpadded numbering of the zone table: q -> q + (q /
  // plane) (PP - plane); the quotient in FP32 is exact here ((q + 1/2) /
  // plane stays at least 1 / (2 plane) from an integer)
  const float inv_plane = 1.f / (float)plane;
  const int pad = PP - plane;
  flush_tile<float, 3>(
      [&](int q) { return tile[q + (int)(((float)q + 0.5f) * inv_plane) * pad]; }, tiles, g, lo,
      gg, tot, sp.w, threadIdx.x, blockDim.x);
}

}  // namespace nfft_hip
