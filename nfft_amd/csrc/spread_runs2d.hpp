// B^T for d = 2 (FP32 and FP64) with same-cell runs summed in registers.
//
// NWV wavefronts per work item, each with its own copy of the LDS tile (E0 x
// E1 cells; the copies are summed in a fixed order when the tile is written
// out).  A lane owns one tap of dimension 0 and a pair of adjacent taps of
// dimension 1 of the current node, so its read-modify-writes of the tile
// never collide with another lane's (W (W/2) pairs over SLOTS passes of the
// 64 lanes).  Nodes arrive sorted by cell; consecutive nodes of one cell form
// a run whose contributions the lanes sum in registers and add to the tile
// once.  The item's nodes are cut into NWV S ranges; each wavefront walks S
// of them as node streams interleaved node by node, independent dependency
// chains; a wavefront's LDS operations complete in program order, so its
// streams' updates of shared cells do not race.
//
// Window values: G nodes at a time, each lane evaluates one tap (dimension
// wt, tap ww) of one node, two nodes per pass when 2 W <= 32, as a per-tap
// polynomial in (frac - 1/2).  Lanes of dimension 0 fold the
// node value in (psi0 f_j, complex), and every lane fetches its factors by
// ds_bpermute (two per FP64 value).
#pragma once

#include "engine.hpp"
#include "window.hpp"

namespace nfft_hip {

// lane-to-lane fetch of a real value (ds_bpermute, byte address)
template <typename R>
__device__ __forceinline__ R bperm_r(int byte_addr, R v) {
  if constexpr (sizeof(R) == 4) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
  } else {
    const int lo = __builtin_amdgcn_ds_bpermute(byte_addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(byte_addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
  }
}

template <typename R, int W, int G, int S, int NWV>
__global__ __launch_bounds__(64 * NWV) void k_spread_runs2d(const cplx<R> *__restrict__ fin,
                                                      cplx<R> *__restrict__ g,
                                                      const int *__restrict__ perm,
                                                      const int *__restrict__ cell,
                                                      const R *__restrict__ frac,
                                                      const int4 *__restrict__ subs,
                                                      BinGeom gg, WinParams<R> wp,
                                                      cplx<R> *__restrict__ tiles) {
  using C = cplx<R>;
  static_assert(W % 2 == 0 && 2 * W <= 64, "W = 2m+2 taps");
  constexpr int HW = W / 2;
  constexpr int PAIRS = W * HW;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int NP = 2 * W <= 32 ? 2 : 1;  // nodes per window pass
  constexpr int NPASS = G / NP;
  static_assert(G % NP == 0 && 64 % G == 0, "groups tile the 64-node chunk");
  // right-hand side blockIdx.y of a batched launch
  fin += (long)blockIdx.y * gg.bs_f;
  g += (long)blockIdx.y * gg.bs_g;
  if (tiles) tiles += (long)blockIdx.y * gg.bs_t;
  extern __shared__ __align__(16) unsigned char b_smem[];
  // NWV wavefronts, each with its own copy of the tile (summed when the
  // tile is written out)
  C *tile0 = reinterpret_cast<C *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[2];
  bin_corner<2>(sp.x, gg, lo);
  const int E1 = gg.E[1];
  const int tot = gg.E[0] * E1;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  C *tile = tile0 + wv * tot;
  for (int q = threadIdx.x; q < NWV * tot; q += 64 * NWV) tile0[q] = C{R(0), R(0)};

  // (dimension-0 tap, dimension-1 pair) slots of this lane
  int off[SLOTS];
  bool pv[SLOTS];
  int adr_u[NP][SLOTS], adr_v0[NP][SLOTS], adr_v1[NP][SLOTS];
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = lane + 64 * sl;
    pv[sl] = q < PAIRS;
    const int pa = pv[sl] ? q / HW : 0;
    const int pb = pv[sl] ? 2 * (q - (q / HW) * HW) : 0;
    off[sl] = pa * E1 + pb;
#pragma unroll
    for (int h = 0; h < NP; h++) {
      adr_u[h][sl] = 4 * (h * 32 + pa);
      adr_v0[h][sl] = 4 * (h * 32 + W + pb);
      adr_v1[h][sl] = 4 * (h * 32 + W + pb + 1);
    }
  }
  // window tap evaluated by this lane in a pass: node half h, dimension wt, tap ww
  const int h = NP == 2 ? lane >> 5 : 0;
  const int li = NP == 2 ? lane & 31 : lane;
  const int wt = li / W < 2 ? li / W : 1;
  const int ww = li - (li / W) * W;
  // its tap polynomial in registers
  TapCoef<R, poly_deg_v<R>> tc;
  tc.load(wp.poly + (wt * W + ww) * (poly_deg_v<R> + 1));

  // streams: parts wv S .. wv S + S - 1 of the item's NWV S node ranges
  int sb[S], sn[S];
  {
    const int part = (sp.z + NWV * S - 1) / (NWV * S);
#pragma unroll
    for (int st = 0; st < S; st++) {
      sb[st] = min(sp.z, (wv * S + st) * part);
      sn[st] = min(sp.z, (wv * S + st + 1) * part) - sb[st];
    }
  }
  C acc[S][SLOTS][2];
#pragma unroll
  for (int st = 0; st < S; st++)
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++) acc[st][sl][0] = acc[st][sl][1] = C{R(0), R(0)};
  // current run per stream (tile corner), from the stream's first node on
  int cur[S];
  __syncthreads();
  auto flush = [&](int st) {
    const int corner = cur[st];
    // all reads, then all writes: the slots are distinct cells
    C v[SLOTS][2];
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++)
      if (pv[sl]) {
        v[sl][0] = tile[corner + off[sl]];
        v[sl][1] = tile[corner + off[sl] + 1];
      }
#pragma unroll
    for (int sl = 0; sl < SLOTS; sl++) {
      if (pv[sl]) {
        tile[corner + off[sl]] = C{v[sl][0].x + acc[st][sl][0].x, v[sl][0].y + acc[st][sl][0].y};
        tile[corner + off[sl] + 1] = C{v[sl][1].x + acc[st][sl][1].x, v[sl][1].y + acc[st][sl][1].y};
      }
      acc[st][sl][0] = acc[st][sl][1] = C{R(0), R(0)};
    }
  };

  // chunk of 64 nodes per stream, one per lane (a valid node for every lane:
  // lanes past the stream's end repeat its last node with value zero, so they
  // stay in the last run and add nothing; an empty stream reads the item's
  // first node)
  auto pos = [&](int st, int first) {
    const int i = first + lane;
    return (long)sp.y + (sn[st] > 0 ? sb[st] + (i < sn[st] ? i : sn[st] - 1) : 0);
  };
  int ckey[S];
  R cfr[S][2];
  C cf[S];
  int nc[S][2];
  R nfr[S][2];
  C nf[S];
  int nperm[S], nnperm[S];
  auto load_raw = [&](int st, int first, int pidx) {
    const long p = pos(st, first);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      nc[st][t] = cell[t * gg.M + p];
      nfr[st][t] = frac[t * gg.M + p];
    }
    nf[st] = fin[pidx];
  };
  auto convert = [&](int st, int first) {
    ckey[st] = (nc[st][0] - lo[0]) * E1 + (nc[st][1] - lo[1]);
#pragma unroll
    // frac - 1/2, the polynomials' variable
    for (int t = 0; t < 2; t++) cfr[st][t] = nfr[st][t] - R(0.5);
    cf[st] = first + lane < sn[st] ? nf[st] : C{R(0), R(0)};
  };
  int longest = 0;
#pragma unroll
  for (int st = 0; st < S; st++) {
    load_raw(st, 0, perm[pos(st, 0)]);
    nperm[st] = perm[pos(st, 64)];
    nnperm[st] = 0;
    convert(st, 0);
    cur[st] = __builtin_amdgcn_readfirstlane(ckey[st]);
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
      // factors of nodes base .. base+G-1 of every stream
      R ur[S][G][SLOTS], ui[S][G][SLOTS], v0[S][G][SLOTS], v1[S][G][SLOTS];
#pragma unroll
      for (int st = 0; st < S; st++) {
#pragma unroll
        for (int p = 0; p < NPASS; p++) {
          const int src = 4 * (base + NP * p + h);
          const R f0 = bperm_r<R>(src, cfr[st][0]);
          const R f1 = bperm_r<R>(src, cfr[st][1]);
          const R fx = bperm_r<R>(src, cf[st].x);
          const R fy = bperm_r<R>(src, cf[st].y);
          const R tt = wt == 0 ? f0 : f1;
          const R v = tc.at(tt);
          // dimension-0 lanes carry psi0 f_j, dimension-1 lanes psi1
          const R vr = wt == 0 ? v * fx : v;
          const R vi = wt == 0 ? v * fy : v;
#pragma unroll
          for (int hk = 0; hk < NP; hk++) {
            const int k = NP * p + hk;
#pragma unroll
            for (int sl = 0; sl < SLOTS; sl++) {
              ur[st][k][sl] = bperm_r<R>(adr_u[hk][sl], vr);
              ui[st][k][sl] = bperm_r<R>(adr_u[hk][sl], vi);
              v0[st][k][sl] = bperm_r<R>(adr_v0[hk][sl], v);
              v1[st][k][sl] = bperm_r<R>(adr_v1[hk][sl], v);
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < G; k++) {
#pragma unroll
        for (int st = 0; st < S; st++) {
          const int key = __builtin_amdgcn_readlane(ckey[st], base + k);
          if (__builtin_expect(key != cur[st], 0)) {
            flush(st);
            cur[st] = key;
          }
#pragma unroll
          for (int sl = 0; sl < SLOTS; sl++) {
            acc[st][sl][0].x = fma(ur[st][k][sl], v0[st][k][sl], acc[st][sl][0].x);
            acc[st][sl][0].y = fma(ui[st][k][sl], v0[st][k][sl], acc[st][sl][0].y);
            acc[st][sl][1].x = fma(ur[st][k][sl], v1[st][k][sl], acc[st][sl][1].x);
            acc[st][sl][1].y = fma(ui[st][k][sl], v1[st][k][sl], acc[st][sl][1].y);
          }
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
        cf[st] = C{R(0), R(0)};
      }
    }
  }
#pragma unroll
  for (int st = 0; st < S; st++) flush(st);
  __syncthreads();
  flush_tile<R, 2>(
      [&](int q) {
        C v = tile0[q];
#pragma unroll
        for (int w = 1; w < NWV; w++) v = C{v.x + tile0[w * tot + q].x, v.y + tile0[w * tot + q].y};
        return v;
      },
      tiles, g, lo, gg, tot, sp.w, threadIdx.x, 64 * NWV);
}

// The same B^T for NC right-hand sides that share one node set (a batch of
// coils): workgroup (item, y) spreads right-hand sides NC y .. NC y + NC - 1
// of the launch into NC LDS tiles.  What depends only on the nodes is done
// once for all of them: the coordinate fetches, the window polynomials, the
// pair weights psi0, psi1 and the run keys; per right-hand side a node
// costs its value (two v_readlane from the lanes that loaded it), two
// products psi0 f_j and the 4 FMAs per pair.  Each right-hand side's sums
// are formed exactly as in k_spread_runs2d (psi0 f_j rounded, times psi1,
// added in node order), so the results are bit-identical to it.
template <int W, int G, int S, int NC, int NWV>
__global__ __launch_bounds__(64 * NWV) void k_spread_runs2d_mc(const float2 *__restrict__ fin,
                                                         float2 *__restrict__ g,
                                                         const int *__restrict__ perm,
                                                         const int *__restrict__ cell,
                                                         const float *__restrict__ frac,
                                                         const int4 *__restrict__ subs,
                                                         BinGeom gg, WinParams<float> wp,
                                                         float2 *__restrict__ tiles) {
  static_assert(W % 2 == 0 && 2 * W <= 64, "W = 2m+2 taps");
  constexpr int HW = W / 2;
  constexpr int PAIRS = W * HW;
  constexpr int SLOTS = (PAIRS + 63) / 64;
  constexpr int NP = 2 * W <= 32 ? 2 : 1;  // nodes per window pass
  constexpr int NPASS = G / NP;
  static_assert(G % NP == 0 && 64 % G == 0, "groups tile the 64-node chunk");
  // right-hand sides of this workgroup: c0 .. c0 + nrhs - 1
  const int c0 = blockIdx.y * NC;
  const int nrhs = min(NC, gg.nrhs - c0);
  fin += (long)c0 * gg.bs_f;
  g += (long)c0 * gg.bs_g;
  tiles += (long)c0 * gg.bs_t;
  extern __shared__ __align__(16) unsigned char b_smem[];
  // per wavefront NC tiles of tot cells (the NWV copies summed at the end)
  float2 *tile0 = reinterpret_cast<float2 *>(b_smem);
  const int4 sp = subs[blockIdx.x];
  int lo[2];
  bin_corner<2>(sp.x, gg, lo);
  const int E1 = gg.E[1];
  const int tot = gg.E[0] * E1;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float2 *tile = tile0 + wv * NC * tot;
  for (int q = threadIdx.x; q < NWV * NC * tot; q += 64 * NWV) tile0[q] = float2{0.f, 0.f};

  // (dimension-0 tap, dimension-1 pair) slots of this lane
  int off[SLOTS];
  bool pv[SLOTS];
  int adr_u[NP][SLOTS], adr_v0[NP][SLOTS], adr_v1[NP][SLOTS];
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = lane + 64 * sl;
    pv[sl] = q < PAIRS;
    const int pa = pv[sl] ? q / HW : 0;
    const int pb = pv[sl] ? 2 * (q - (q / HW) * HW) : 0;
    off[sl] = pa * E1 + pb;
#pragma unroll
    for (int h = 0; h < NP; h++) {
      adr_u[h][sl] = 4 * (h * 32 + pa);
      adr_v0[h][sl] = 4 * (h * 32 + W + pb);
      adr_v1[h][sl] = 4 * (h * 32 + W + pb + 1);
    }
  }
  // window tap evaluated by this lane in a pass: node half h, dimension wt,
  // tap ww, its polynomial in registers
  const int h = NP == 2 ? lane >> 5 : 0;
  const int li = NP == 2 ? lane & 31 : lane;
  const int wt = li / W < 2 ? li / W : 1;
  const int ww = li - (li / W) * W;
  TapCoef<float, kPolyDeg> tc;
  tc.load(wp.poly + (wt * W + ww) * kPolyCoef);

  // streams: parts wv S .. wv S + S - 1 of the item's NWV S node ranges
  int sb[S], sn[S];
  {
    const int part = (sp.z + NWV * S - 1) / (NWV * S);
#pragma unroll
    for (int st = 0; st < S; st++) {
      sb[st] = min(sp.z, (wv * S + st) * part);
      sn[st] = min(sp.z, (wv * S + st + 1) * part) - sb[st];
    }
  }
  float2 acc[S][NC][SLOTS][2];
#pragma unroll
  for (int st = 0; st < S; st++)
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int sl = 0; sl < SLOTS; sl++) acc[st][c][sl][0] = acc[st][c][sl][1] = float2{0.f, 0.f};
  int cur[S];
  __syncthreads();
  auto flush = [&](int st) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      if (c < nrhs) {
        float2 *tr = tile + c * tot + cur[st];
        float2 v[SLOTS][2];
#pragma unroll
        for (int sl = 0; sl < SLOTS; sl++)
          if (pv[sl]) {
            v[sl][0] = tr[off[sl]];
            v[sl][1] = tr[off[sl] + 1];
          }
#pragma unroll
        for (int sl = 0; sl < SLOTS; sl++)
          if (pv[sl]) {
            tr[off[sl]] = float2{v[sl][0].x + acc[st][c][sl][0].x, v[sl][0].y + acc[st][c][sl][0].y};
            tr[off[sl] + 1] =
                float2{v[sl][1].x + acc[st][c][sl][1].x, v[sl][1].y + acc[st][c][sl][1].y};
          }
      }
#pragma unroll
      for (int sl = 0; sl < SLOTS; sl++) acc[st][c][sl][0] = acc[st][c][sl][1] = float2{0.f, 0.f};
    }
  };

  // chunk of 64 nodes per stream, one per lane (lanes past the stream's end
  // repeat its last node with value zero; an empty stream reads the item's
  // first node)
  auto pos = [&](int st, int first) {
    const int i = first + lane;
    return (long)sp.y + (sn[st] > 0 ? sb[st] + (i < sn[st] ? i : sn[st] - 1) : 0);
  };
  int ckey[S];
  float cfr[S][2];
  float2 cf[S][NC];
  int nc_[S][2];
  float nfr[S][2];
  float2 nf[S][NC];
  int nperm[S], nnperm[S];
  auto load_raw = [&](int st, int first, int pidx) {
    const long p = pos(st, first);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      nc_[st][t] = cell[t * gg.M + p];
      nfr[st][t] = frac[t * gg.M + p];
    }
#pragma unroll
    for (int c = 0; c < NC; c++)
      nf[st][c] = c < nrhs ? fin[(long)c * gg.bs_f + pidx] : float2{0.f, 0.f};
  };
  auto convert = [&](int st, int first) {
    ckey[st] = (nc_[st][0] - lo[0]) * E1 + (nc_[st][1] - lo[1]);
#pragma unroll
    for (int t = 0; t < 2; t++) cfr[st][t] = nfr[st][t] - 0.5f;
#pragma unroll
    for (int c = 0; c < NC; c++) cf[st][c] = first + lane < sn[st] ? nf[st][c] : float2{0.f, 0.f};
  };
  auto rl = [](float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  };
  int longest = 0;
#pragma unroll
  for (int st = 0; st < S; st++) {
    load_raw(st, 0, perm[pos(st, 0)]);
    nperm[st] = perm[pos(st, 64)];
    nnperm[st] = 0;
    convert(st, 0);
    cur[st] = __builtin_amdgcn_readfirstlane(ckey[st]);
    longest = max(longest, sn[st]);
  }
  for (int first = 0; first < longest; first += 64) {
    int cmax = 0;
#pragma unroll
    for (int st = 0; st < S; st++) {
      if (first + 64 < sn[st]) {
        load_raw(st, first + 64, nperm[st]);
        nnperm[st] = perm[pos(st, first + 128)];
      }
      cmax = max(cmax, max(0, min(64, sn[st] - first)));
    }
    for (int base = 0; base < cmax; base += G) {
      // window factors of nodes base .. base+G-1 of every stream
      float u[S][G][SLOTS], v0[S][G][SLOTS], v1[S][G][SLOTS];
#pragma unroll
      for (int st = 0; st < S; st++) {
#pragma unroll
        for (int p = 0; p < NPASS; p++) {
          const int src = 4 * (base + NP * p + h);
          const float f0 = bperm_r<float>(src, cfr[st][0]);
          const float f1 = bperm_r<float>(src, cfr[st][1]);
          const float v = tc.at(wt == 0 ? f0 : f1);
#pragma unroll
          for (int hk = 0; hk < NP; hk++) {
            const int k = NP * p + hk;
#pragma unroll
            for (int sl = 0; sl < SLOTS; sl++) {
              u[st][k][sl] = bperm_r<float>(adr_u[hk][sl], v);
              v0[st][k][sl] = bperm_r<float>(adr_v0[hk][sl], v);
              v1[st][k][sl] = bperm_r<float>(adr_v1[hk][sl], v);
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < G; k++) {
#pragma unroll
        for (int st = 0; st < S; st++) {
          const int l = base + k;
          const int key = __builtin_amdgcn_readlane(ckey[st], l);
          if (__builtin_expect(key != cur[st], 0)) {
            flush(st);
            cur[st] = key;
          }
#pragma unroll
          for (int c = 0; c < NC; c++) {
            const float fx = rl(cf[st][c].x, l), fy = rl(cf[st][c].y, l);
#pragma unroll
            for (int sl = 0; sl < SLOTS; sl++) {
              const float ur = u[st][k][sl] * fx, ui = u[st][k][sl] * fy;
              acc[st][c][sl][0].x = fmaf(ur, v0[st][k][sl], acc[st][c][sl][0].x);
              acc[st][c][sl][0].y = fmaf(ui, v0[st][k][sl], acc[st][c][sl][0].y);
              acc[st][c][sl][1].x = fmaf(ur, v1[st][k][sl], acc[st][c][sl][1].x);
              acc[st][c][sl][1].y = fmaf(ui, v1[st][k][sl], acc[st][c][sl][1].y);
            }
          }
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
#pragma unroll
        for (int c = 0; c < NC; c++) cf[st][c] = float2{0.f, 0.f};
      }
    }
  }
#pragma unroll
  for (int st = 0; st < S; st++) flush(st);
  __syncthreads();
  for (int c = 0; c < nrhs; c++) {
    const float2 *tr = tile0 + c * tot;
    flush_tile<float, 2>(
        [&](int q) {
          float2 v = tr[q];
#pragma unroll
          for (int w = 1; w < NWV; w++)
            v = float2{v.x + tr[w * NC * tot + q].x, v.y + tr[w * NC * tot + q].y};
          return v;
        },
        tiles + (long)c * gg.bs_t, g + (long)c * gg.bs_g, lo, gg, tot, sp.w, threadIdx.x,
        64 * NWV);
  }
}

}  // namespace nfft_hip
