// Node preprocessing for the B kernels (the job of nfft_precompute_one_psi:
// reference kernel/nfft/nfft.c:5938-5948, 5819-5844, and of its node sort
// sort0 / nfft_sort_node_indices_radix_lsdf, kernel/nfft/nfft.c:75-123,
// kernel/util/sort.c:91-167).
//
// Per node and dimension the coordinate is reduced in FP64:
//   c_t = floor(n_t x_t),  frac_t = n_t x_t - c_t in [0,1)
// The reference's tap window u = c - m .. c + m + 1 (kernel/nfft/nfft.c:324-340)
// follows from c.  Instead of the reference's radix sort of linearised cell
// keys, nodes are counting-sorted into bins of bs_t cells per dimension
// (histogram with atomics, exclusive scan, placement), and every bin is cut
// into work items of at most max_per_sub nodes for the LDS-tiled B kernels.
// Order inside a bin is not specified; results do not depend on it.
#include <cstdio>

#include "engine.hpp"

namespace nfft_hip {

template <typename R>
__global__ void k_node_prep(const R *__restrict__ x, BinGeom gg, int *__restrict__ cell_tmp,
                            R *__restrict__ frac_tmp, int *__restrict__ bin_of,
                            int *__restrict__ rank_of, int *__restrict__ hist) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= gg.M) return;
  int bin = 0;
  for (int t = 0; t < gg.d; t++) {
    double y = (double)x[j * gg.d + t] * (double)gg.n[t];
    double c = floor(y);
    double fr = y - c;
    long cw = (long)c % gg.n[t];
    if (cw < 0) cw += gg.n[t];
    cell_tmp[t * gg.M + j] = (int)cw;
    frac_tmp[t * gg.M + j] = (R)fr;
    bin = bin * gg.nb[t] + (int)cw / gg.bs[t];
  }
  bin_of[j] = bin;
  rank_of[j] = atomicAdd(&hist[bin], 1);
}

// Exclusive scan of `len` ints by one workgroup of 1024 threads; total in *sum.
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

template <typename R>
__global__ void k_node_place(BinGeom gg, const int *__restrict__ cell_tmp,
                             const R *__restrict__ frac_tmp, const int *__restrict__ bin_of,
                             const int *__restrict__ rank_of, const int *__restrict__ bin_start,
                             int *__restrict__ cell, R *__restrict__ frac, int *__restrict__ perm) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= gg.M) return;
  long p = (long)bin_start[bin_of[j]] + rank_of[j];
  perm[p] = (int)j;
  for (int t = 0; t < gg.d; t++) {
    cell[t * gg.M + p] = cell_tmp[t * gg.M + j];
    frac[t * gg.M + p] = frac_tmp[t * gg.M + j];
  }
}

__global__ void k_count_subs(const int *__restrict__ hist, long nbins, int max_per_sub,
                             int *__restrict__ nsub_of) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  nsub_of[b] = (hist[b] + max_per_sub - 1) / max_per_sub;
}

__global__ void k_fill_subs(const int *__restrict__ hist, const int *__restrict__ bin_start,
                            const int *__restrict__ sub_start, long nbins, int max_per_sub,
                            int4 *__restrict__ subs) {
  long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int cnt = hist[b];
  int s0 = sub_start[b];
  for (int s = 0; s * max_per_sub < cnt; s++) {
    int first = bin_start[b] + s * max_per_sub;
    int c = min(max_per_sub, cnt - s * max_per_sub);
    subs[s0 + s] = make_int4((int)b, first, c, 0);
  }
}

template <typename R>
void bin_nodes(Engine<R> &e) {
  const BinGeom &gg = e.geom;
  const long M = e.M;
  long nbins = 1;
  for (int t = 0; t < gg.d; t++) nbins *= gg.nb[t];
  // scratch layout, each array 256-byte aligned: hist, bin_start, nsub_of,
  // sub_start (nbins ints each), bin_of, rank_of (M ints each), cell_tmp
  // (d M ints), frac_tmp (d M reals), sums (2 ints)
  auto al = [](long bytes) { return (bytes + 255) / 256 * 256; };
  const long b_bins = al(nbins * (long)sizeof(int));
  const long b_M = al(M * (long)sizeof(int));
  const long b_cell = al((long)gg.d * M * (long)sizeof(int));
  const long b_frac = al((long)gg.d * M * (long)sizeof(R));
  const long need = 4 * b_bins + 2 * b_M + b_cell + b_frac + 256;
  if (need > e.scratch_cap) {
    if (e.scratch_i) NFFT_HIP_CHECK(R, hipFree(e.scratch_i));
    NFFT_HIP_CHECK(R, hipMalloc(&e.scratch_i, need));
    e.scratch_cap = need;
  }
  char *base = reinterpret_cast<char *>(e.scratch_i);
  int *hist = reinterpret_cast<int *>(base);
  int *bin_start = reinterpret_cast<int *>(base + b_bins);
  int *nsub_of = reinterpret_cast<int *>(base + 2 * b_bins);
  int *sub_start = reinterpret_cast<int *>(base + 3 * b_bins);
  int *bin_of = reinterpret_cast<int *>(base + 4 * b_bins);
  int *rank_of = reinterpret_cast<int *>(base + 4 * b_bins + b_M);
  int *cell_tmp = reinterpret_cast<int *>(base + 4 * b_bins + 2 * b_M);
  R *frac_tmp = reinterpret_cast<R *>(base + 4 * b_bins + 2 * b_M + b_cell);
  int *sums = reinterpret_cast<int *>(base + 4 * b_bins + 2 * b_M + b_cell + b_frac);

  hipStream_t s = e.stream;
  NFFT_HIP_CHECK(R, hipMemsetAsync(hist, 0, nbins * sizeof(int), s));
  if (M > 0)
    hipLaunchKernelGGL(k_node_prep<R>, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, e.xdev,
                       gg, cell_tmp, frac_tmp, bin_of, rank_of, hist);
  hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, hist, bin_start, nbins, sums);
  if (M > 0)
    hipLaunchKernelGGL(k_node_place<R>, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, gg,
                       cell_tmp, frac_tmp, bin_of, rank_of, bin_start, e.cell, e.frac, e.perm);
  // work items of interpolation (small, for balance) and of spreading
  // (larger, to bound the number of tiles written to scratch); the spreading
  // list keeps its per-bin first item and count for the tile reduction
  auto make_list = [&](int max_per_sub, int *first_of, int *count_of, int4 **subs, long *cap,
                       long *nsub) {
    hipLaunchKernelGGL(k_count_subs, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, hist,
                       nbins, max_per_sub, count_of);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, count_of, first_of, nbins,
                       sums + 1);
    int total = 0;
    NFFT_HIP_CHECK(R, hipMemcpyAsync(&total, sums + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(s));
    if (total > *cap) {
      if (*subs) NFFT_HIP_CHECK(R, hipFree(*subs));
      NFFT_HIP_CHECK(R, hipMalloc(subs, (size_t)total * sizeof(int4)));
      *cap = total;
    }
    *nsub = total;
    if (total > 0)
      hipLaunchKernelGGL(k_fill_subs, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, hist,
                         bin_start, first_of, nbins, max_per_sub, *subs);
  };
  make_list(e.max_per_sub, sub_start, nsub_of, &e.subs, &e.subs_cap, &e.nsub);
  if (nbins > e.bin_arrays_cap) {
    if (e.sp_first) NFFT_HIP_CHECK(R, hipFree(e.sp_first));
    NFFT_HIP_CHECK(R, hipMalloc(&e.sp_first, 2 * nbins * sizeof(int)));
    e.bin_arrays_cap = nbins;
  }
  e.sp_count = e.sp_first + nbins;
  make_list(e.spread_max_per_sub, e.sp_first, e.sp_count, &e.sp_subs, &e.sp_subs_cap, &e.sp_nsub);
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template void bin_nodes<float>(Engine<float> &);
template void bin_nodes<double>(Engine<double> &);

}  // namespace nfft_hip
