// Node preprocessing for the B kernels (the job of nfft_precompute_one_psi:
// reference kernel/nfft/nfft.c:5938-5948, 5819-5844, and of its node sort
// sort0 / nfft_sort_node_indices_radix_lsdf, kernel/nfft/nfft.c:75-123,
// kernel/util/sort.c:91-167).
//
// Per node and dimension the coordinate is reduced in FP64:
//   c_t = floor(n_t x_t),  frac_t = n_t x_t - c_t in [0,1)
// The reference's tap window u = c - m .. c + m + 1 (kernel/nfft/nfft.c:324-340)
// follows from c.  Instead of the reference's radix sort of linearised cell
// keys, nodes are counting-sorted by a bin-major cell key
//   key = bin * (bs_0 ... bs_{d-1}) + (cell offset inside the bin, row-major)
// (histogram with atomics and multi-block exclusive scan for the bin and cell
// offsets; the order itself from a stable LSD radix sort of the keys with the
// node indices, rocPRIM's radix_sort_pairs), so each bin is one contiguous
// range and, inside it, nodes of one cell are adjacent (the spreading kernels
// sum such runs in registers) in increasing node index, as the reference's
// stable radix sort leaves them.  The order, and with it every FP32 sum, is
// therefore the same on every run.  Every bin is cut into work items of at
// most max_per_sub nodes for the LDS-tiled B kernels.
#include <cstdio>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "engine.hpp"

namespace nfft_hip {

template <typename R>
__global__ void k_node_prep(const R *__restrict__ x, BinGeom gg, int cap,
                            int *__restrict__ cell_tmp, R *__restrict__ frac_tmp,
                            int *__restrict__ key_of, int *__restrict__ khist) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= gg.M) return;
  int bin = 0, local = 0;
  for (int t = 0; t < gg.d; t++) {
    double y = (double)x[j * gg.d + t] * (double)gg.n[t];
    double c = floor(y);
    double fr = y - c;
    long cw = (long)c % gg.n[t];
    if (cw < 0) cw += gg.n[t];
    cell_tmp[t * gg.M + j] = (int)cw;
    frac_tmp[t * gg.M + j] = (R)fr;
    const int bt = (int)cw / gg.bs[t];
    bin = bin * gg.nb[t] + bt;
    local = local * gg.bs[t] + ((int)cw - bt * gg.bs[t]);
  }
  const int key = bin * cap + local;
  key_of[j] = key;
  atomicAdd(&khist[key], 1);
}

// Exclusive scan of `len` ints by one workgroup of 1024 threads; total in *sum.
// Used for short arrays (bins, block sums of the long scan below).
__global__ __launch_bounds__(1024) void k_scan_exclusive(const int *__restrict__ in,
                                                         int *__restrict__ out, long len,
                                                         int *__restrict__ sum) {
  __shared__ long part[1024];
  const int tid = threadIdx.x;
  const long chunk = (len + 1023) / 1024;
  const long a = tid * chunk;
  const long b = min(len, a + chunk);
  long s = 0;
  for (long i = a; i < b; i++) s += in[i];
  part[tid] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 partial sums
  for (int off = 1; off < 1024; off <<= 1) {
    long v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  long run = tid > 0 ? part[tid - 1] : 0;
  for (long i = a; i < b; i++) {
    int v = in[i];
    out[i] = (int)run;
    run += v;
  }
  if (tid == 1023) *sum = (int)part[1023];
}

// Long exclusive scan in three launches: per-tile sums (k_scan_tile_sums),
// a one-workgroup scan of the tile sums, and per-tile scans with the tile's
// offset (k_scan_tile_apply).  A tile is 256 threads x 16 consecutive ints.
constexpr int kScanPerThread = 16;
constexpr int kScanTile = 256 * kScanPerThread;

__device__ inline int block_exclusive_sum_256(int v, int *sh, int &total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int u = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += u;
    __syncthreads();
  }
  total = sh[255];
  const int incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_tile_sums(const int *__restrict__ in, long len,
                                                        int *__restrict__ tile_sum) {
  __shared__ int sh[256];
  const long q0 = (long)blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) s += q0 + k < len ? in[q0 + k] : 0;
  int total;
  block_exclusive_sum_256(s, sh, total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_tile_apply(const int *__restrict__ in,
                                                         int *__restrict__ out, long len,
                                                         const int *__restrict__ tile_off) {
  __shared__ int sh[256];
  const long q0 = (long)blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  int v[kScanPerThread];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    v[k] = q0 + k < len ? in[q0 + k] : 0;
    s += v[k];
  }
  int total;
  int run = tile_off[blockIdx.x] + block_exclusive_sum_256(s, sh, total);
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    if (q0 + k < len) out[q0 + k] = run;
    run += v[k];
  }
}

// bins from the scanned cell-key histogram: bin b owns keys [b cap, (b+1) cap)
__global__ void k_bins_from_keys(const int *__restrict__ kstart, long nbins, int cap, int total,
                                 int *__restrict__ hist, int *__restrict__ bin_start) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const int first = kstart[b * cap];
  const int next = b + 1 < nbins ? kstart[(b + 1) * cap] : total;
  bin_start[b] = first;
  hist[b] = next - first;
}

// sorted position p holds node order[p]: its index and its cell / frac
template <typename R>
__global__ void k_node_place(BinGeom gg, const int *__restrict__ cell_tmp,
                             const R *__restrict__ frac_tmp, const int *__restrict__ order,
                             int *__restrict__ cell, R *__restrict__ frac, int *__restrict__ perm) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= gg.M) return;
  const long j = order[p];
  perm[p] = (int)j;
  for (int t = 0; t < gg.d; t++) {
    cell[t * gg.M + p] = cell_tmp[t * gg.M + j];
    frac[t * gg.M + p] = frac_tmp[t * gg.M + j];
  }
}

__global__ void k_count_subs(const int *__restrict__ hist, long nbins, int max_per_sub,
                             int max_items, int *__restrict__ nsub_of) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  nsub_of[b] = items_of(hist[b], max_per_sub, max_items);
}

__global__ void k_fill_subs(const int *__restrict__ hist, const int *__restrict__ bin_start,
                            const int *__restrict__ sub_start, long nbins, int max_per_sub,
                            int max_items, int4 *__restrict__ subs) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int cnt = hist[b];
  int s0 = sub_start[b];
  const int k = items_of(cnt, max_per_sub, max_items);
  for (int s = 0; s < k; s++) {
    int first, c;
    split_item(cnt, k, s, first, c);
    // .w: the item's place in bin order (its spreading tile's slot)
    subs[s0 + s] = make_int4((int)b, bin_start[b] + first, c, s0 + s);
  }
}

// Order of the bins for the zone reduction: bins whose 3^d neighbourhoods
// hold the most spreading work items first (16 buckets by log2 of that
// count), so that the few long-running workgroups of dense regions start at
// the beginning of the launch instead of ending it.
constexpr int kOrderBuckets = 16;
__global__ void k_bin_weight(const int *__restrict__ count_of, BinGeom gg, long nbins,
                             int *__restrict__ bucket_of, int *__restrict__ bucket_hist) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int bc[kMaxDim] = {};
  long r = b;
  for (int t = gg.d - 1; t >= 0; t--) {
    bc[t] = (int)(r % gg.nb[t]);
    r /= gg.nb[t];
  }
  int zc = 1;
  for (int t = 0; t < gg.d; t++) zc *= 3;
  long pairs = 0;
  for (int k = 0; k < zc; k++) {
    int rk = k;
    long nbl = 0;
    int dig[kMaxDim] = {};
    for (int t = gg.d - 1; t >= 0; t--) {
      dig[t] = rk % 3;
      rk /= 3;
    }
    for (int t = 0; t < gg.d; t++) {
      int v = bc[t] + dig[t] - 1;
      v += v < 0 ? gg.nb[t] : (v >= gg.nb[t] ? -gg.nb[t] : 0);
      nbl = nbl * gg.nb[t] + v;
    }
    pairs += count_of[nbl];
  }
  int lg = 0;
  while (lg < kOrderBuckets - 1 && (2l << lg) <= pairs + 1) lg++;
  const int bucket = kOrderBuckets - 1 - lg;  // heaviest first
  bucket_of[b] = bucket;
  atomicAdd(&bucket_hist[bucket], 1);
}

__global__ void k_bin_order(const int *__restrict__ bucket_of, long nbins,
                            int *__restrict__ bucket_next, int *__restrict__ order) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  order[atomicAdd(&bucket_next[bucket_of[b]], 1)] = (int)b;
}

// Launch order of the work items: largest node counts first (buckets by
// log2 of the count), so that long items do not end the launch.
__global__ void k_sub_weight(const int4 *__restrict__ subs, long nsub, int *__restrict__ bucket_of,
                             int *__restrict__ bucket_hist) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsub) return;
  const int c = subs[i].z;
  int lg = 0;
  while (lg < kOrderBuckets - 1 && (2 << lg) <= c) lg++;
  const int bucket = kOrderBuckets - 1 - lg;
  bucket_of[i] = bucket;
  atomicAdd(&bucket_hist[bucket], 1);
}

__global__ void k_sub_order(const int4 *__restrict__ in, const int *__restrict__ bucket_of,
                            long nsub, int *__restrict__ bucket_next, int4 *__restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsub) return;
  out[atomicAdd(&bucket_next[bucket_of[i]], 1)] = in[i];
}

// Interpolation slots: up to Q consecutive nodes of one cell run (same
// cell, same work item) whose grid taps the kernel reads once.  Runs start
// where the cell changes or a work item starts; slots start at the run
// positions that are multiples of Q; a slot is (first node, node count).
__global__ void k_mark_item_starts(const int4 *__restrict__ subs, long nsub, int *__restrict__ b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsub) b[subs[i].y] = 1;
}

__global__ void k_run_flags(const int *__restrict__ cell, int d, long M, int *__restrict__ b) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  bool start = j == 0 || b[j] != 0;
  for (int t = 0; t < d && !start; t++) start = cell[t * M + j] != cell[t * M + j - 1];
  b[j] = start ? 1 : 0;
}

// run_id[j] = (exclusive sum of b)[j] + b[j] - 1 in place; run_start[id] = j
__global__ void k_run_starts(const int *__restrict__ b, int *__restrict__ run_id, long M,
                             int *__restrict__ run_start) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int id = run_id[j] + b[j] - 1;
  run_id[j] = id;
  if (b[j]) run_start[id] = (int)j;
}

__global__ void k_slot_flags(const int *__restrict__ run_id, const int *__restrict__ run_start,
                             long M, int Q, int *__restrict__ sflag) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  sflag[j] = ((j - run_start[run_id[j]]) % Q) == 0 ? 1 : 0;
}

__global__ void k_fill_slots(const int *__restrict__ b, const int *__restrict__ sflag,
                             const int *__restrict__ sidx, long M, int Q,
                             int2 *__restrict__ slots) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M || !sflag[j]) return;
  int cnt = 1;
  while (cnt < Q && j + cnt < M && b[j + cnt] == 0) cnt++;
  slots[sidx[j]] = make_int2((int)j, cnt);
}

__global__ void k_item_slots(const int4 *__restrict__ subs, long nsub, const int *__restrict__ sidx,
                             long M, int nslots, int2 *__restrict__ islots) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsub) return;
  const long first = subs[i].y, end = first + subs[i].z;
  const int s0 = sidx[first];
  const int s1 = end < M ? sidx[end] : nslots;
  islots[i] = make_int2(s0, s1 - s0);
}

// Orders each work item's interpolation slots by node count, largest first,
// and deals the slots of one count round-robin over the residue of their
// window's tile index modulo 32 (the LDS bank pair of its first tap, for
// rows P2 and planes PP cells apart), in node order within one residue.
// The interpolation kernel picks its code path per wavefront from the
// largest count among the wavefront's slots, so sorted slots give
// wavefronts of one count and few lanes that compute for absent nodes; the
// dealing makes consecutive lanes read windows that start in different
// banks (count order alone, lanes reading the windows of scattered cells,
// measured slower than node order: more LDS bank conflicts).  One
// workgroup of 256 threads per item; in[] and out[] are the slot arrays
// before and after, islots[item] = (first slot, slot count).
template <int Q>
__global__ __launch_bounds__(256) void k_sort_slots(const int2 *__restrict__ islots,
                                                    const int2 *__restrict__ in,
                                                    int2 *__restrict__ out,
                                                    const int *__restrict__ cell, long M, int P2,
                                                    int PP) {
  constexpr int NR = 32;             // residue buckets
  constexpr int NK = (Q + 1) * NR;   // keys: count * NR + residue
  __shared__ int hist[NK];           // slots per key in the item
  __shared__ int run[NK];            // slots per key placed by earlier chunks
  __shared__ int wcnt[4][NK];        // slots per key in each wavefront's part of a chunk
  const int2 is = islots[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  auto key_of = [&](int2 v) {
    return v.y * NR + ((cell[v.x] * PP + cell[M + v.x] * P2 + cell[2 * M + v.x]) & (NR - 1));
  };
  for (int t = threadIdx.x; t < NK; t += 256) hist[t] = run[t] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < is.y; k += 256) atomicAdd(&hist[key_of(in[is.x + k])], 1);
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int k0 = 0; k0 < is.y; k0 += 256) {
    const int k = k0 + threadIdx.x;
    const bool valid = k < is.y;
    const int2 v = valid ? in[is.x + k] : make_int2(0, 0);
    const int key = valid ? key_of(v) : NK;
    // lanes of this wavefront with the same key (bit by bit)
    unsigned long long same = __ballot(valid);
    if (!valid) same = ~same;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const unsigned long long bal = __ballot((key >> b) & 1);
      same &= ((key >> b) & 1) ? bal : ~bal;
    }
    for (int t = threadIdx.x; t < 4 * NK; t += 256) (&wcnt[0][0])[t] = 0;
    __syncthreads();
    // the key's lowest lane records the wavefront's count
    if (valid && (same & below) == 0) wcnt[wv][key] = __popcll(same);
    __syncthreads();
    if (valid) {
      int j = run[key] + __popcll(same & below);
      for (int w = 0; w < wv; w++) j += wcnt[w][key];
      const int c = v.y, r = key - c * NR;
      int pos = 0;
      for (int cc = Q; cc > c; cc--)
        for (int rr = 0; rr < NR; rr++) pos += hist[cc * NR + rr];
      // round j of the dealing: min(h, j) slots of every residue before
      // it, then those of lower residues that still have a slot
      for (int rr = 0; rr < NR; rr++) {
        const int h = hist[c * NR + rr];
        pos += min(h, j) + (rr < r && h > j ? 1 : 0);
      }
      if (pos < is.y) out[is.x + pos] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NK; t += 256)
      run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
    __syncthreads();
  }
}

__global__ void k_fill_zero_int(int *__restrict__ p, long count) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

static void launch_fill_zero_int(int *p, long count, hipStream_t s) {
  if (count > 0)
    hipLaunchKernelGGL(k_fill_zero_int, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, s, p,
                       count);
}

template <typename R>
void bin_nodes(Engine<R> &e) {
  const BinGeom &gg = e.geom;
  const long M = e.M;
  long nbins = 1;
  int cells_per_bin = 1;
  for (int t = 0; t < gg.d; t++) {
    nbins *= gg.nb[t];
    cells_per_bin *= gg.bs[t];
  }
  const long nkeys = nbins * cells_per_bin;
  if (nkeys > 0x7fffffffL)
    family<R>::fatal("nfft: %ld grid cells exceed the node binning's int32 keys", nkeys);
  const long ntiles = ceil_div(nkeys, kScanTile);
  // scratch layout, each array 256-byte aligned: khist, kstart (nkeys ints),
  // tile_sum, tile_off (ntiles ints), hist, bin_start, nsub_of, sub_start
  // (nbins ints each), key_of, rank_of (M ints each), cell_tmp (d M ints),
  // frac_tmp (d M reals), sums (256 ints: scan totals, ordering buckets)
  auto al = [](long bytes) { return (bytes + 255) / 256 * 256; };
  const long b_keys = al(nkeys * (long)sizeof(int));
  // the tile sums also serve scans of M-long node arrays
  const long b_tiles = al(std::max(ntiles, ceil_div(M + 1, kScanTile)) * (long)sizeof(int));
  const long b_bins = al(nbins * (long)sizeof(int));
  const long b_M = al(M * (long)sizeof(int));
  const long b_cell = al((long)gg.d * M * (long)sizeof(int));
  const long b_frac = al((long)gg.d * M * (long)sizeof(R));
  // the stable node order: keys sorted with the node indices (LSD radix
  // sort over the key's bits); its temporary storage comes from the same
  // scratch
  const unsigned key_bits = nkeys > 1 ? 32u - (unsigned)__builtin_clz((unsigned)(nkeys - 1)) : 1u;
  size_t sort_bytes = 0;
  NFFT_HIP_CHECK(R, rocprim::radix_sort_pairs(nullptr, sort_bytes, (const unsigned *)nullptr,
                                              (unsigned *)nullptr, rocprim::counting_iterator<int>(0),
                                              (int *)nullptr, (size_t)M, 0u, key_bits, e.stream));
  const long b_sort = al((long)sort_bytes + 1);
  const long need = 2 * b_keys + 2 * b_tiles + 4 * b_bins + 4 * b_M + b_cell + b_frac + b_sort + 1024;
  if (need > e.scratch_cap) {
    if (e.scratch_i) NFFT_HIP_CHECK(R, hipFree(e.scratch_i));
    NFFT_HIP_CHECK(R, hipMalloc(&e.scratch_i, need));
    e.scratch_cap = need;
  }
  char *p = reinterpret_cast<char *>(e.scratch_i);
  auto take = [&](long bytes) {
    char *q = p;
    p += bytes;
    return q;
  };
  int *khist = reinterpret_cast<int *>(take(b_keys));
  int *kstart = reinterpret_cast<int *>(take(b_keys));
  int *tile_sum = reinterpret_cast<int *>(take(b_tiles));
  int *tile_off = reinterpret_cast<int *>(take(b_tiles));
  int *hist = reinterpret_cast<int *>(take(b_bins));
  int *bin_start = reinterpret_cast<int *>(take(b_bins));
  int *nsub_of = reinterpret_cast<int *>(take(b_bins));
  int *sub_start = reinterpret_cast<int *>(take(b_bins));
  int *key_of = reinterpret_cast<int *>(take(b_M));
  int *rank_of = reinterpret_cast<int *>(take(b_M));
  unsigned *key_sorted = reinterpret_cast<unsigned *>(take(b_M));
  int *order = reinterpret_cast<int *>(take(b_M));
  void *sort_tmp = take(b_sort);
  int *cell_tmp = reinterpret_cast<int *>(take(b_cell));
  R *frac_tmp = reinterpret_cast<R *>(take(b_frac));
  int *sums = reinterpret_cast<int *>(take(1024));

  hipStream_t s = e.stream;
  launch_fill_zero_int(khist, nkeys, s);
  if (M > 0)
    hipLaunchKernelGGL(k_node_prep<R>, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, e.xdev,
                       gg, cells_per_bin, cell_tmp, frac_tmp, key_of, khist);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)ntiles), dim3(256), 0, s, khist, nkeys,
                     tile_sum);
  hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, tile_sum, tile_off, ntiles,
                     sums);
  hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)ntiles), dim3(256), 0, s, khist, kstart,
                     nkeys, tile_off);
  hipLaunchKernelGGL(k_bins_from_keys, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s,
                     kstart, nbins, cells_per_bin, (int)M, hist, bin_start);
  if (M > 0) {
    NFFT_HIP_CHECK(R, rocprim::radix_sort_pairs(sort_tmp, sort_bytes, (const unsigned *)key_of,
                                                key_sorted, rocprim::counting_iterator<int>(0),
                                                order, (size_t)M, 0u, key_bits, s));
    hipLaunchKernelGGL(k_node_place<R>, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, gg,
                       cell_tmp, frac_tmp, order, e.cell, e.frac, e.perm);
  }
  // work items of interpolation (small, for balance) and of spreading
  // (larger, to bound the number of tiles written to scratch); the spreading
  // list keeps its per-bin first item and count for the tile reduction
  auto make_list = [&](int max_per_sub, int max_items, int *first_of, int *count_of, int4 **subs,
                       long *cap, long *nsub) {
    hipLaunchKernelGGL(k_count_subs, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, hist,
                       nbins, max_per_sub, max_items, count_of);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, count_of, first_of, nbins,
                       sums + 1);
    int total = 0;
    NFFT_HIP_CHECK(R, hipMemcpyAsync(&total, sums + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(s));
    // the list, then the same list in bin order (the ordering's input)
    if (2L * total > *cap) {
      if (*subs) NFFT_HIP_CHECK(R, hipFree(*subs));
      NFFT_HIP_CHECK(R, hipMalloc(subs, (size_t)2 * total * sizeof(int4)));
      *cap = 2L * total;
    }
    *nsub = total;
    if (total > 0) {
      int4 *in_bin_order = *subs + total;
      hipLaunchKernelGGL(k_fill_subs, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, hist,
                         bin_start, first_of, nbins, max_per_sub, max_items, in_bin_order);
      // largest items first; key_of (M ints, total <= M) is free again
      int *bucket_of = key_of;
      int *bhist = sums + 2 + 3 * kOrderBuckets;
      int *boff = bhist + kOrderBuckets;
      launch_fill_zero_int(bhist, kOrderBuckets, s);
      hipLaunchKernelGGL(k_sub_weight, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s,
                         in_bin_order, (long)total, bucket_of, bhist);
      hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, bhist, boff,
                         (long)kOrderBuckets, boff + kOrderBuckets);
      hipLaunchKernelGGL(k_sub_order, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s,
                         in_bin_order, bucket_of, (long)total, boff, *subs);
    }
  };
  make_list(e.max_per_sub, 0, sub_start, nsub_of, &e.subs, &e.subs_cap, &e.nsub);
  // interpolation slots (pairs of same-cell nodes) for the kernels that use them
  e.geom.slots = e.geom.islots = nullptr;
  e.nslots = 0;
  if (e.use_slots() && M > 0 && e.nsub > 0) {
    auto long_scan = [&](const int *in, int *out, long len, int *total) {
      const long nt = ceil_div(len, kScanTile);
      hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nt), dim3(256), 0, s, in, len, tile_sum);
      hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, tile_sum, tile_off, nt,
                         total);
      hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)nt), dim3(256), 0, s, in, out, len,
                         tile_off);
    };
    // node-length temporaries (free after placement): run-start flags,
    // run ids, run starts, slot flags, slot indices
    int *b = key_of;
    int *run_id = rank_of;
    int *run_start = cell_tmp;
    int *sflag = cell_tmp + M;
    int *sidx = reinterpret_cast<int *>(frac_tmp);
    const unsigned gm = (unsigned)ceil_div(M, 256);
    launch_fill_zero_int(b, M, s);
    hipLaunchKernelGGL(k_mark_item_starts, dim3((unsigned)ceil_div(e.nsub, 256)), dim3(256), 0, s,
                       e.subs, e.nsub, b);
    hipLaunchKernelGGL(k_run_flags, dim3(gm), dim3(256), 0, s, e.cell, gg.d, M, b);
    long_scan(b, run_id, M, sums + 100);
    hipLaunchKernelGGL(k_run_starts, dim3(gm), dim3(256), 0, s, b, run_id, M, run_start);
    // quads only where the kernel has their code (the shared symmetric
    // window table); otherwise pairs
    const int Q = e.win.poly_shared && e.sym_window && 2 * gg.m + 2 <= kSlotQuadMaxW ? e.slot_nodes : 2;
    hipLaunchKernelGGL(k_slot_flags, dim3(gm), dim3(256), 0, s, run_id, run_start, M, Q, sflag);
    long_scan(sflag, sidx, M, sums + 101);
    int nslots = 0;
    NFFT_HIP_CHECK(R, hipMemcpyAsync(&nslots, sums + 101, sizeof(int), hipMemcpyDeviceToHost, s));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(s));
    if (nslots > e.slots_cap) {
      if (e.slots) NFFT_HIP_CHECK(R, hipFree(e.slots));
      if (e.slots_raw) NFFT_HIP_CHECK(R, hipFree(e.slots_raw));
      NFFT_HIP_CHECK(R, hipMalloc(&e.slots, sizeof(int2) * (size_t)nslots));
      NFFT_HIP_CHECK(R, hipMalloc(&e.slots_raw, sizeof(int2) * (size_t)nslots));
      e.slots_cap = nslots;
    }
    if (e.nsub > e.islots_cap) {
      if (e.islots) NFFT_HIP_CHECK(R, hipFree(e.islots));
      NFFT_HIP_CHECK(R, hipMalloc(&e.islots, sizeof(int2) * (size_t)e.nsub));
      e.islots_cap = e.nsub;
    }
    // slots in node order, then (e.slot_sort) each item's slots by count
    // and LDS bank
    int2 *filled = e.slot_sort ? e.slots_raw : e.slots;
    hipLaunchKernelGGL(k_fill_slots, dim3(gm), dim3(256), 0, s, b, sflag, sidx, M, Q, filled);
    hipLaunchKernelGGL(k_item_slots, dim3((unsigned)ceil_div(e.nsub, 256)), dim3(256), 0, s,
                       e.subs, e.nsub, sidx, M, nslots, e.islots);
    if (e.slot_sort) {
      if (Q == 4)
        hipLaunchKernelGGL(k_sort_slots<4>, dim3((unsigned)e.nsub), dim3(256), 0, s, e.islots,
                           filled, e.slots, e.cell, M, e.geom.ip2, e.geom.ipp);
      else
        hipLaunchKernelGGL(k_sort_slots<2>, dim3((unsigned)e.nsub), dim3(256), 0, s, e.islots,
                           filled, e.slots, e.cell, M, e.geom.ip2, e.geom.ipp);
    }
    e.nslots = nslots;
    e.geom.slots = e.slots;
    e.geom.islots = e.islots;
  }
  if (nbins > e.bin_arrays_cap) {
    if (e.sp_first) NFFT_HIP_CHECK(R, hipFree(e.sp_first));
    NFFT_HIP_CHECK(R, hipMalloc(&e.sp_first, 3 * nbins * sizeof(int)));
    e.bin_arrays_cap = nbins;
  }
  e.sp_count = e.sp_first + nbins;
  e.bin_order = e.sp_first + 2 * nbins;
  make_list(e.spread_max_per_sub, e.spread_max_items, e.sp_first, e.sp_count, &e.sp_subs,
            &e.sp_subs_cap, &e.sp_nsub);
  // heavy bins first for the zone reduction (the interpolation list's
  // temporaries are free again: bucket per bin, bucket histogram and offsets)
  if (gg.zoned) {
    int *bucket_of = sub_start;
    int *bhist = sums + 2;
    int *boff = sums + 2 + kOrderBuckets;
    launch_fill_zero_int(bhist, kOrderBuckets, s);
    hipLaunchKernelGGL(k_bin_weight, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s,
                       e.sp_count, gg, nbins, bucket_of, bhist);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, bhist, boff,
                       (long)kOrderBuckets, sums + 2 + 2 * kOrderBuckets);
    hipLaunchKernelGGL(k_bin_order, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s,
                       bucket_of, nbins, boff, e.bin_order);
  }
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template void bin_nodes<float>(Engine<float> &);
template void bin_nodes<double>(Engine<double> &);

}  // namespace nfft_hip

// Diagnostics hook (host only): the work-item split of one bin as the
// binning kernels compute it.  Returns the number of items k and writes
// each item's first node (relative to the bin) and node count.
extern "C" int nfft_hip_work_split(int cnt, int max_per_sub, int max_items, int *first,
                                   int *count) {
  const int k = nfft_hip::items_of(cnt, max_per_sub, max_items);
  for (int s = 0; s < k; s++) nfft_hip::split_item(cnt, k, s, first[s], count[s]);
  return k;
}
