// D and D^T (phi_hut deconvolution with zero padding / cropping) and the
// direct NDFT kernels.
//
//   D   (trafo):   g[k mod n] = f_hat[k] * prod_t phi_inv_t[ks_t], zero elsewhere
//                  (reference kernel/nfft/nfft.c:2786-2831, 3833-3897, 5415-5513)
//   D^T (adjoint): f_hat[k]  = g[k mod n] * prod_t phi_inv_t[ks_t]
//                  (reference kernel/nfft/nfft.c:2874-2903, 3941-3997, 5560-5650)
// with k_t = ks_t - N_t/2, ks_t in [0, N_t).  One generic kernel for d <= 3
// replaces the reference's 2^d octant copies.
//
// Roofline: HBM-bound; algorithmic bytes D = N_total c + n_total c (zero fill
// included), D^T = 2 N_total c, with c the bytes of one complex.
#include "engine.hpp"

namespace nfft_hip {

struct DeconvGeom {
  int d;
  int N[kMaxDim];
  int n[kMaxDim];
  int off[kMaxDim];  // offset of dimension t in phi_inv
  long N_total;
};

template <typename R>
__global__ void k_fill_zero(cplx<R> *__restrict__ p, long count) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = cplx<R>{R(0), R(0)};
}

template <typename R>
void launch_fill_zero(cplx<R> *p, long count, hipStream_t s) {
  if (count <= 0) return;
  long blocks = ceil_div(count, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_zero<R>, dim3((unsigned)blocks), dim3(256), 0, s, p, count);
}

// grid index and deconvolution factor of frequency kL
template <typename R>
__device__ __forceinline__ long freq_to_grid(long kL, const DeconvGeom &gg, const R *phi_inv,
                                             R &fac) {
  int ks[kMaxDim];
  for (int t = gg.d - 1; t >= 0; t--) {
    ks[t] = (int)(kL % gg.N[t]);
    kL /= gg.N[t];
  }
  long gi = 0;
  fac = R(1);
  for (int t = 0; t < gg.d; t++) {
    int k = ks[t] - gg.N[t] / 2;
    int l = k < 0 ? k + gg.n[t] : k;
    gi = gi * gg.n[t] + l;
    fac *= phi_inv[gg.off[t] + ks[t]];
  }
  return gi;
}

template <typename R>
__global__ void k_deconv_to_grid(const cplx<R> *__restrict__ fhat, cplx<R> *__restrict__ g,
                                 const R *__restrict__ phi_inv, DeconvGeom gg) {
  long kL = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kL >= gg.N_total) return;
  R fac;
  long gi = freq_to_grid<R>(kL, gg, phi_inv, fac);
  cplx<R> v = fhat[kL];
  g[gi] = cplx<R>{v.x * fac, v.y * fac};
}

template <typename R>
__global__ void k_deconv_from_grid(const cplx<R> *__restrict__ g, cplx<R> *__restrict__ fhat,
                                   const R *__restrict__ phi_inv, DeconvGeom gg) {
  long kL = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kL >= gg.N_total) return;
  R fac;
  long gi = freq_to_grid<R>(kL, gg, phi_inv, fac);
  cplx<R> v = g[gi];
  fhat[kL] = cplx<R>{v.x * fac, v.y * fac};
}

static DeconvGeom make_geom(int d, const long *N, const long *n) {
  DeconvGeom gg{};
  gg.d = d;
  gg.N_total = 1;
  int off = 0;
  for (int t = 0; t < d; t++) {
    gg.N[t] = (int)N[t];
    gg.n[t] = (int)n[t];
    gg.off[t] = off;
    off += (int)N[t];
    gg.N_total *= N[t];
  }
  return gg;
}

template <typename R>
void launch_deconv_to_grid(const cplx<R> *fhat, cplx<R> *g, const R *phi_inv, int d,
                           const long *N, const long *n, hipStream_t s) {
  DeconvGeom gg = make_geom(d, N, n);
  long blocks = ceil_div(gg.N_total, 256);
  hipLaunchKernelGGL(k_deconv_to_grid<R>, dim3((unsigned)blocks), dim3(256), 0, s, fhat, g,
                     phi_inv, gg);
}

template <typename R>
void launch_deconv_from_grid(const cplx<R> *g, cplx<R> *fhat, const R *phi_inv, int d,
                             const long *N, const long *n, hipStream_t s) {
  DeconvGeom gg = make_geom(d, N, n);
  long blocks = ceil_div(gg.N_total, 256);
  hipLaunchKernelGGL(k_deconv_from_grid<R>, dim3((unsigned)blocks), dim3(256), 0, s, g, fhat,
                     phi_inv, gg);
}

// ---------------------------------------------------------------------------
// Direct NDFT (reference kernel/nfft/nfft.c:145-297), O(M N_total):
//   trafo:   f_j   = sum_k f_hat_k exp(-2 pi i k.x_j)   (one thread per node)
//   adjoint: f_hat_k = sum_j f_j exp(+2 pi i k.x_j)     (one thread per k)
// The phase k.x is accumulated in FP64 and reduced mod 1 before sincospi, in
// both precision families.
// ---------------------------------------------------------------------------
struct NdftGeom {
  int d;
  int N[kMaxDim];
  long N_total;
  long M;
};

template <typename R>
__global__ void k_ndft_trafo(const R *__restrict__ x, const cplx<R> *__restrict__ fhat,
                             cplx<R> *__restrict__ f, NdftGeom gg) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= gg.M) return;
  double xj[kMaxDim];
  for (int t = 0; t < gg.d; t++) xj[t] = (double)x[j * gg.d + t];
  double re = 0.0, im = 0.0;
  for (long kL = 0; kL < gg.N_total; kL++) {
    long r = kL;
    double ph = 0.0;
    for (int t = gg.d - 1; t >= 0; t--) {
      int k = (int)(r % gg.N[t]) - gg.N[t] / 2;
      r /= gg.N[t];
      ph += (double)k * xj[t];
    }
    ph -= floor(ph);
    double s, c;
    sincospi(2.0 * ph, &s, &c);
    cplx<R> a = fhat[kL];
    // a * exp(-i 2 pi ph)
    re += (double)a.x * c + (double)a.y * s;
    im += (double)a.y * c - (double)a.x * s;
  }
  f[j] = cplx<R>{(R)re, (R)im};
}

template <typename R>
__global__ void k_ndft_adjoint(const R *__restrict__ x, const cplx<R> *__restrict__ f,
                               cplx<R> *__restrict__ fhat, NdftGeom gg) {
  long kL = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kL >= gg.N_total) return;
  int k[kMaxDim];
  long r = kL;
  for (int t = gg.d - 1; t >= 0; t--) {
    k[t] = (int)(r % gg.N[t]) - gg.N[t] / 2;
    r /= gg.N[t];
  }
  double re = 0.0, im = 0.0;
  for (long j = 0; j < gg.M; j++) {
    double ph = 0.0;
    for (int t = 0; t < gg.d; t++) ph += (double)k[t] * (double)x[j * gg.d + t];
    ph -= floor(ph);
    double s, c;
    sincospi(2.0 * ph, &s, &c);
    cplx<R> a = f[j];
    // a * exp(+i 2 pi ph)
    re += (double)a.x * c - (double)a.y * s;
    im += (double)a.y * c + (double)a.x * s;
  }
  fhat[kL] = cplx<R>{(R)re, (R)im};
}

template <typename R>
void launch_ndft(const R *x, const cplx<R> *in, cplx<R> *out, int d, const long *N, long M,
                 int adjoint, hipStream_t s) {
  NdftGeom gg{};
  gg.d = d;
  gg.N_total = 1;
  for (int t = 0; t < d; t++) {
    gg.N[t] = (int)N[t];
    gg.N_total *= N[t];
  }
  gg.M = M;
  if (adjoint) {
    if (gg.N_total == 0) return;
    hipLaunchKernelGGL(k_ndft_adjoint<R>, dim3((unsigned)ceil_div(gg.N_total, 128)), dim3(128),
                       0, s, x, in, out, gg);
  } else {
    if (M == 0) return;
    hipLaunchKernelGGL(k_ndft_trafo<R>, dim3((unsigned)ceil_div(M, 128)), dim3(128), 0, s, x,
                       in, out, gg);
  }
}

#define NFFT_HIP_INST(R)                                                                    \
  template void launch_fill_zero<R>(cplx<R> *, long, hipStream_t);                          \
  template void launch_deconv_to_grid<R>(const cplx<R> *, cplx<R> *, const R *, int,        \
                                         const long *, const long *, hipStream_t);          \
  template void launch_deconv_from_grid<R>(const cplx<R> *, cplx<R> *, const R *, int,      \
                                           const long *, const long *, hipStream_t);        \
  template void launch_ndft<R>(const R *, const cplx<R> *, cplx<R> *, int, const long *,    \
                               long, int, hipStream_t);
NFFT_HIP_INST(float)
NFFT_HIP_INST(double)
#undef NFFT_HIP_INST

}  // namespace nfft_hip
