// F: the d-dimensional unnormalised DFT of the oversampled grid, done as one
// in-place pass per axis (the role FFTW plays for the reference:
// kernel/nfft/nfft.c:6030-6031 plans, FFTW_FORWARD = sign -1 for trafo,
// FFTW_BACKWARD = sign +1 for adjoint).
//
// Each workgroup stages `lines_per_block` whole lines of one axis in LDS and
// runs a self-sorting Stockham schedule over them (mixed radix 2/3/4/5/7/8 and
// small primes), ping-ponging between two LDS buffers, then writes the lines
// back.  Stage with radix R after sub-transforms of size Ns:
//   v_r = X[j + r n/R] * w^(r (j mod Ns))  with w = exp(sign 2 pi i / (Ns R))
//   v   = DFT_R(v)
//   Y[(j div Ns) Ns R + (j mod Ns) + r Ns] = v_r
// Twiddles come from a table exp(-2 pi i k / n) built in FP64 on the host.
//
// Roofline: HBM-bound; algorithmic bytes per axis pass 2 n_total c.
#include "engine.hpp"

namespace nfft_hip {

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x - b.x, a.y - b.y};
}
// multiply by sign * i
template <typename R, int SIGN>
__device__ __forceinline__ cplx<R> mul_si(cplx<R> a) {
  return SIGN > 0 ? cplx<R>{-a.y, a.x} : cplx<R>{a.y, -a.x};
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft2(cplx<R> *v) {
  cplx<R> a = v[0], b = v[1];
  v[0] = cadd<R>(a, b);
  v[1] = csub<R>(a, b);
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft4(cplx<R> &v0, cplx<R> &v1, cplx<R> &v2, cplx<R> &v3) {
  cplx<R> s02 = cadd<R>(v0, v2), d02 = csub<R>(v0, v2);
  cplx<R> s13 = cadd<R>(v1, v3), d13 = mul_si<R, SIGN>(csub<R>(v1, v3));
  v0 = cadd<R>(s02, s13);
  v2 = csub<R>(s02, s13);
  v1 = cadd<R>(d02, d13);
  v3 = csub<R>(d02, d13);
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft8(cplx<R> *v) {
  // even / odd radix-4 halves, then combine with w8^k, w8 = exp(sign 2 pi i/8)
  cplx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cplx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<R, SIGN>(e0, e1, e2, e3);
  dft4<R, SIGN>(o0, o1, o2, o3);
  const R h = R(0.70710678118654752440084436210484903928);
  // o1 *= (1 + sign i)/sqrt2 ; o2 *= sign i ; o3 *= (-1 + sign i)/sqrt2
  cplx<R> t1 = cadd<R>(o1, mul_si<R, SIGN>(o1));
  t1 = cplx<R>{t1.x * h, t1.y * h};
  cplx<R> t2 = mul_si<R, SIGN>(o2);
  cplx<R> t3 = csub<R>(mul_si<R, SIGN>(o3), o3);
  t3 = cplx<R>{t3.x * h, t3.y * h};
  v[0] = cadd<R>(e0, o0);
  v[4] = csub<R>(e0, o0);
  v[1] = cadd<R>(e1, t1);
  v[5] = csub<R>(e1, t1);
  v[2] = cadd<R>(e2, t2);
  v[6] = csub<R>(e2, t2);
  v[3] = cadd<R>(e3, t3);
  v[7] = csub<R>(e3, t3);
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft3(cplx<R> *v) {
  const R c = R(-0.5);
  const R s = R(SIGN) * R(0.86602540378443864676372317075293618347);
  cplx<R> t = cadd<R>(v[1], v[2]);
  cplx<R> u = csub<R>(v[1], v[2]);
  cplx<R> a = cplx<R>{v[0].x + c * t.x, v[0].y + c * t.y};
  // s * i * u
  cplx<R> iu = cplx<R>{-s * u.y, s * u.x};
  v[0] = cadd<R>(v[0], t);
  v[1] = cadd<R>(a, iu);
  v[2] = csub<R>(a, iu);
}

// generic O(R^2) DFT for the remaining small primes
template <typename R, int RAD, int SIGN>
__device__ __forceinline__ void dft_generic(cplx<R> *v) {
  cplx<R> out[RAD];
#pragma unroll
  for (int k = 0; k < RAD; k++) {
    R sr = 0, si = 0;
#pragma unroll
    for (int r = 0; r < RAD; r++) {
      const int e = (r * k) % RAD;
      double sd, cd;
      sincospi(double(2 * SIGN * e) / double(RAD), &sd, &cd);
      const R s = R(sd), c = R(cd);
      sr += v[r].x * c - v[r].y * s;
      si += v[r].x * s + v[r].y * c;
    }
    out[k] = cplx<R>{sr, si};
  }
#pragma unroll
  for (int k = 0; k < RAD; k++) v[k] = out[k];
}

template <typename R, int RAD, int SIGN>
__device__ __forceinline__ void dft(cplx<R> *v) {
  if constexpr (RAD == 2) dft2<R, SIGN>(v);
  else if constexpr (RAD == 3) dft3<R, SIGN>(v);
  else if constexpr (RAD == 4) dft4<R, SIGN>(v[0], v[1], v[2], v[3]);
  else if constexpr (RAD == 8) dft8<R, SIGN>(v);
  else dft_generic<R, RAD, SIGN>(v);
}

template <typename R, int RAD, int SIGN>
__device__ void stockham_stage(const cplx<R> *__restrict__ src, cplx<R> *__restrict__ dst, int n,
                               int nlines, int Ns, const cplx<R> *__restrict__ tw) {
  const int q = n / RAD;
  const int nbf = nlines * q;
  const int span = Ns * RAD;
  const int twstep = n / span;
  for (int b = threadIdx.x; b < nbf; b += blockDim.x) {
    int ll = b / q;
    int j = b - ll * q;
    int jm = j % Ns;
    const cplx<R> *s = src + ll * n + j;
    cplx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; r++) v[r] = s[r * q];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < RAD; r++) {
        cplx<R> w = tw[jm * r * twstep];
        if (SIGN > 0) w.y = -w.y;
        v[r] = cmul<R>(v[r], w);
      }
    }
    dft<R, RAD, SIGN>(v);
    cplx<R> *o = dst + ll * n + (j / Ns) * span + jm;
#pragma unroll
    for (int r = 0; r < RAD; r++) o[r * Ns] = v[r];
  }
}

template <typename R, int SIGN>
__global__ __launch_bounds__(256) void k_fft_axis(cplx<R> *__restrict__ data, FftAxis<R> ax) {
  extern __shared__ unsigned char fft_smem[];
  const int n = ax.len;
  const int T = ax.lines_per_block;
  cplx<R> *bufA = reinterpret_cast<cplx<R> *>(fft_smem);
  cplx<R> *bufB = bufA + (long)T * n;
  const long L0 = (long)blockIdx.x * T;
  const int TL = (int)min((long)T, ax.nlines - L0);
  const int tot = TL * n;
  const long inner = ax.inner;

  // stage TL lines into LDS: line-major for contiguous lines, otherwise
  // line-minor so that a wave reads TL adjacent lines of one row
  for (int qd = threadIdx.x; qd < tot; qd += blockDim.x) {
    int ll, k;
    if (inner == 1) {
      ll = qd / n;
      k = qd - ll * n;
    } else {
      k = qd / TL;
      ll = qd - k * TL;
    }
    long L = L0 + ll;
    long o = L / inner, i = L - o * inner;
    bufA[ll * n + k] = data[(o * n + k) * inner + i];
  }
  __syncthreads();

  cplx<R> *src = bufA, *dst = bufB;
  int Ns = 1;
  for (int s = 0; s < ax.nstages; s++) {
    const int rad = ax.radix[s];
    switch (rad) {
      case 2: stockham_stage<R, 2, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 3: stockham_stage<R, 3, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 4: stockham_stage<R, 4, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 5: stockham_stage<R, 5, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 7: stockham_stage<R, 7, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 8: stockham_stage<R, 8, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 11: stockham_stage<R, 11, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      case 13: stockham_stage<R, 13, SIGN>(src, dst, n, TL, Ns, ax.twiddle); break;
      default: break;  // excluded at plan time
    }
    __syncthreads();
    cplx<R> *t = src;
    src = dst;
    dst = t;
    Ns *= rad;
  }

  for (int qd = threadIdx.x; qd < tot; qd += blockDim.x) {
    int ll, k;
    if (inner == 1) {
      ll = qd / n;
      k = qd - ll * n;
    } else {
      k = qd / TL;
      ll = qd - k * TL;
    }
    long L = L0 + ll;
    long o = L / inner, i = L - o * inner;
    data[(o * n + k) * inner + i] = src[ll * n + k];
  }
}

template <typename R>
void launch_fft_axis(cplx<R> *data, const FftAxis<R> &ax, int sign, hipStream_t s) {
  long blocks = ceil_div(ax.nlines, ax.lines_per_block);
  size_t lds = 2 * sizeof(cplx<R>) * (size_t)ax.lines_per_block * (size_t)ax.len;
  if (sign < 0)
    hipLaunchKernelGGL((k_fft_axis<R, -1>), dim3((unsigned)blocks), dim3(256), lds, s, data, ax);
  else
    hipLaunchKernelGGL((k_fft_axis<R, +1>), dim3((unsigned)blocks), dim3(256), lds, s, data, ax);
}

template void launch_fft_axis<float>(cplx<float> *, const FftAxis<float> &, int, hipStream_t);
template void launch_fft_axis<double>(cplx<double> *, const FftAxis<double> &, int, hipStream_t);

}  // namespace nfft_hip
