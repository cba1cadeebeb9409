// F: the d-dimensional unnormalised DFT of the oversampled grid (the role
// FFTW plays for the reference: kernel/nfft/nfft.c:6030-6031 plans,
// FFTW_FORWARD = sign -1 for trafo, FFTW_BACKWARD = sign +1 for adjoint),
// done as one pass per dimension with D and D^T fused into the passes at the
// ends and the zero band of the padded grid pruned:
//
//   trafo   passes t = d-1 .. 0.  The first reads f_hat, applies the
//           deconvolution factors and places the values in the band
//           (k - N_t/2) mod n_t of each line (D, reference
//           kernel/nfft/nfft.c:2786-2831); every pass handles only lines whose
//           not-yet-transformed coordinates lie in their band, and reads only
//           the band along its own axis (the rest is known to be zero).
//   adjoint passes t = 0 .. d-1.  Each keeps only the band of its axis (all
//           D^T reads, nfft.c:2874-2903) and handles only lines whose already
//           transformed coordinates lie in the band; the last writes f_hat
//           with the deconvolution factors applied (D^T).
//
// Each workgroup stages `lines_per_block` whole lines in LDS and runs a
// self-sorting Stockham schedule over them in place (mixed radix 2/3/4/5/7/8
// and small primes; every stage reads its butterflies into registers,
// synchronises, then writes).  Stage with radix R after sub-transforms of
// size Ns:
//   v_r = X[j + r n/R] * w^(r (j mod Ns))  with w = exp(sign 2 pi i / (Ns R))
//   v   = DFT_R(v)
//   Y[(j div Ns) Ns R + (j mod Ns) + r Ns] = v_r
// Twiddles come from a table exp(-2 pi i k / n) built in FP64 on the host.
//
// Roofline: HBM-bound.  Algorithmic bytes of the whole F with D fused, for
// d = 3 and c bytes per complex: N_total c + the band-pruned intermediate
// reads and writes (N0 N1 n2 + N0 n1 n2 + n_total) c written, the same minus
// the final n_total read back by the next pass.
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

// Butterflies per thread of one stage: lines * n / RAD over the workgroup.
template <int CAP, int RAD> constexpr int stage_slots() { return (CAP / RAD + 255) / 256; }

// LDS layout of the staged lines: element k of line ll sits at
// ll * ls + k + (k >> 4) with an odd line stride ls = n + n/16 + 1, so that
// neither the strided accesses of the first Stockham stages (stride RAD
// within a line) nor the staging of adjacent lines (one element per line)
// fall into the same LDS bank.
__device__ __forceinline__ int lds_idx(int ll, int k, int ls) { return ll * ls + k + (k >> 4); }

// One in-place Stockham stage over `nlines` lines of length n in LDS.  POW2:
// n and Ns are powers of two (shifts instead of divisions).
template <typename R, int RAD, int SIGN, int CAP, bool POW2>
__device__ __forceinline__ void stage_inplace(cplx<R> *__restrict__ buf, int n, int ls, int nlines,
                                              int Ns, const cplx<R> *tw) {
  constexpr int SL = stage_slots<CAP, RAD>();
  const int q = n / RAD;
  const int nbf = nlines * q;
  const int span = Ns * RAD;
  const int twstep = n / span;
  const int lq = POW2 ? __builtin_ctz(q) : 0;
  const int lns = POW2 ? __builtin_ctz(Ns) : 0;
  cplx<R> v[SL][RAD];
  int dll[SL], dk[SL];
#pragma unroll
  for (int i = 0; i < SL; i++) {
    const int b = threadIdx.x + i * 256;
    dll[i] = -1;
    if (b < nbf) {
      const int ll = POW2 ? b >> lq : b / q;
      const int j = b - ll * q;
      const int jm = POW2 ? j & (Ns - 1) : j % Ns;
      const int jd = POW2 ? j >> lns : j / Ns;
#pragma unroll
      for (int r = 0; r < RAD; r++) v[i][r] = buf[lds_idx(ll, j + r * q, ls)];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < RAD; r++) {
          cplx<R> w = tw[jm * r * twstep];
          if (SIGN > 0) w.y = -w.y;
          v[i][r] = cmul<R>(v[i][r], w);
        }
      }
      dft<R, RAD, SIGN>(v[i]);
      dll[i] = ll;
      dk[i] = jd * span + jm;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SL; i++) {
    if (dll[i] >= 0) {
#pragma unroll
      for (int r = 0; r < RAD; r++) buf[lds_idx(dll[i], dk[i] + r * Ns, ls)] = v[i][r];
    }
  }
  __syncthreads();
}

// A Stockham stage of any radix (a prime factor above 13, known only at run
// time), same math as stage_inplace: each output of a butterfly is its own
// sum over the RAD inputs, read from LDS, with the radix-RAD DFT factor
// exp(sign 2 pi i r k / RAD) taken from the line's twiddle table at
// (r k mod RAD) n / RAD.  All outputs are held in registers until every
// thread has read its inputs, then written.  O(n RAD) work per line.
template <typename R, int SIGN, int CAP>
__device__ __forceinline__ void stage_generic(cplx<R> *__restrict__ buf, int n, int ls, int nlines,
                                              int Ns, int RAD, const cplx<R> *tw) {
  constexpr int PER = CAP / 256;
  const int q = n / RAD;
  const int span = Ns * RAD;
  const int twstep = n / span;
  const int tot = nlines * n;
  cplx<R> out[PER];
  int dll[PER], dpos[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int o = threadIdx.x + u * 256;
    dll[u] = -1;
    if (o < tot) {
      const int ll = o / n;
      const int pos = o - ll * n;
      // output pos = jd span + jm + k Ns of butterfly j = jd Ns + jm
      const int jd = pos / span;
      const int rr = pos - jd * span;
      const int k = rr / Ns;
      const int jm = rr - k * Ns;
      const int j = jd * Ns + jm;
      R sr = 0, si = 0;
      for (int r = 0; r < RAD; r++) {
        cplx<R> x = buf[lds_idx(ll, j + r * q, ls)];
        // stage twiddle w^(r jm) and the DFT factor, one table entry each
        cplx<R> w = tw[jm * r * twstep];
        cplx<R> f = tw[((r * k) % RAD) * q];
        if (SIGN > 0) {
          w.y = -w.y;
          f.y = -f.y;
        }
        x = cmul<R>(cmul<R>(x, w), f);
        sr += x.x;
        si += x.y;
      }
      out[u] = cplx<R>{sr, si};
      dll[u] = ll;
      dpos[u] = pos;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; u++)
    if (dll[u] >= 0) buf[lds_idx(dll[u], dpos[u], ls)] = out[u];
  __syncthreads();
}

// Band position of grid index k along a dimension (n, N): the f_hat
// coordinate ks in [0, N) with k = (ks - N/2) mod n, or -1 outside the band.
__device__ __forceinline__ int band_coord(int k, int n, int N) {
  const int h = N / 2;
  if (k < N - h) return k + h;
  if (k >= n - h) return k - (n - h);
  return -1;
}

constexpr int kMaxLinesPerBlock = 256;

// gsrc / gdst: the grid buffer the pass reads and the one it writes (the
// same buffer except in the second half of a four-step axis)
template <typename R, int SIGN, int CAP, bool POW2>
__global__ __launch_bounds__(256) void k_fft_pass(const cplx<R> *gsrc, cplx<R> *gdst,
                                                  const cplx<R> *__restrict__ fin,
                                                  cplx<R> *__restrict__ fout, FftPass<R> P) {
  // right-hand side blockIdx.y of a batched launch
  gsrc += (long)blockIdx.y * P.bs_g;
  gdst += (long)blockIdx.y * P.bs_g;
  if (fin) fin += (long)blockIdx.y * P.bs_c;
  if (fout) fout += (long)blockIdx.y * P.bs_c;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  cplx<R> *buf = reinterpret_cast<cplx<R> *>(fft_smem);
  const int ls = lds_line_stride(P.len);
  // the line length's twiddle table follows the lines in LDS when it fits
  const cplx<R> *tw = P.twiddle;
  if (P.twiddle_in_lds) {
    cplx<R> *twl = buf + (long)P.lines_per_block * ls;
    for (int k = threadIdx.x; k < P.len; k += blockDim.x) twl[k] = P.twiddle[k];
    tw = twl;
  }
  __shared__ long s_goff[kMaxLinesPerBlock];
  __shared__ long s_foff[kMaxLinesPerBlock];
  __shared__ R s_fac[kMaxLinesPerBlock];
  __shared__ int s_sub[kMaxLinesPerBlock];
  const int n = P.len;
  const long L0 = (long)blockIdx.x * P.lines_per_block;
  const int TL = (int)min((long)P.lines_per_block, P.nlines - L0);
  const int tot = TL * n;
  const bool coeffs = P.in_mode == kFromCoeffs || P.out_mode == kToCoeffs;

  // per line: grid offset, f_hat offset and deconvolution factor of the
  // other dimensions
  for (int ll = threadIdx.x; ll < TL; ll += blockDim.x) {
    long L = L0 + ll;
    // four-step halves: the sub-line index (j1 or k2) runs fastest
    int sub = 0;
    if (P.four) {
      sub = (int)(L % P.nsub);
      L /= P.nsub;
    }
    s_sub[ll] = sub;
    int c[kMaxDim - 1] = {};
    for (int s = P.nod - 1; s >= 0; s--) {
      c[s] = (int)(L % P.cnt[s]);
      L /= P.cnt[s];
    }
    long go = 0, fo = 0;
    R fac = R(1);
    for (int s = 0; s < P.nod; s++) {
      int k = c[s];
      if (P.band[s]) {
        k = c[s] - P.N_od[s] / 2;
        k += k < 0 ? P.n_od[s] : 0;
      }
      go += (long)k * P.gstr[s];
      if (coeffs) {
        fo += (long)c[s] * P.fstr[s];
        fac *= P.phi_od[s][c[s]];
      }
    }
    s_goff[ll] = go;
    s_foff[ll] = fo;
    s_fac[ll] = fac;
  }
  __syncthreads();

  // stage the lines: a wave reads along the line for the contiguous axis and
  // across adjacent lines otherwise.  Each thread issues all its loads
  // (at most CAP / 256) before the first LDS store, so their latencies
  // overlap instead of adding up.
  // element k of line ll: its position along the axis in memory and, for
  // the band tests, its axis index (the two differ only in four-step halves)
  const int n1 = P.fn1, n2 = P.fn2;
  auto in_pos = [&](int ll, int k) {
    return P.four == 0 ? k : (P.four == 1 ? s_sub[ll] + n1 * k : k + n1 * s_sub[ll]);
  };
  auto out_pos = [&](int ll, int k) {
    return P.four == 0 ? k : (P.four == 1 ? s_sub[ll] + n1 * k : s_sub[ll] + n2 * k);
  };
  const int nax = P.four ? P.nfull : n;
  // a wave walks along a line where its elements are adjacent in memory,
  // across lines otherwise
  const bool contiguous = P.stride == 1 && P.four != 1;
  const bool contiguous_out = P.stride == 1 && P.four == 0;
  constexpr int PER = CAP / 256;
  cplx<R> v[PER];
  int dst[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int qd = threadIdx.x + u * 256;
    v[u] = cplx<R>{R(0), R(0)};
    dst[u] = -1;
    if (qd < tot) {
      int ll, k;
      if (contiguous) {
        ll = qd / n;
        k = qd - ll * n;
      } else {
        k = qd / TL;
        ll = qd - k * TL;
      }
      dst[u] = lds_idx(ll, k, ls);
      const int kp = in_pos(ll, k);
      if (P.in_mode == kFromGrid) {
        v[u] = gsrc[s_goff[ll] + (long)kp * P.stride];
      } else {
        const int ks = band_coord(kp, nax, P.Nband);
        if (ks >= 0) {
          if (P.in_mode == kFromGridBand) {
            v[u] = gsrc[s_goff[ll] + (long)kp * P.stride];
          } else {
            const cplx<R> w = fin[s_foff[ll] + (long)ks * P.fstr_ax];
            const R f = s_fac[ll] * P.phi_ax[ks];
            v[u] = cplx<R>{w.x * f, w.y * f};
          }
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < PER; u++)
    if (dst[u] >= 0) buf[dst[u]] = v[u];
  __syncthreads();

  int Ns = 1;
  for (int st = 0; st < P.nstages; st++) {
    const int rad = P.radix[st];
    if constexpr (POW2) {
      switch (rad) {
        case 2: stage_inplace<R, 2, SIGN, CAP, true>(buf, n, ls, TL, Ns, tw); break;
        case 4: stage_inplace<R, 4, SIGN, CAP, true>(buf, n, ls, TL, Ns, tw); break;
        default: stage_inplace<R, 8, SIGN, CAP, true>(buf, n, ls, TL, Ns, tw); break;
      }
    } else {
      switch (rad) {
        case 2: stage_inplace<R, 2, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 3: stage_inplace<R, 3, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 4: stage_inplace<R, 4, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 5: stage_inplace<R, 5, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 7: stage_inplace<R, 7, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 8: stage_inplace<R, 8, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 11: stage_inplace<R, 11, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        case 13: stage_inplace<R, 13, SIGN, CAP, false>(buf, n, ls, TL, Ns, tw); break;
        default: stage_generic<R, SIGN, CAP>(buf, n, ls, TL, Ns, rad, tw); break;
      }
    }
    Ns *= rad;
  }

#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int qd = threadIdx.x + u * 256;
    if (qd >= tot) break;
    int ll, k;
    if (contiguous_out) {
      ll = qd / n;
      k = qd - ll * n;
    } else {
      k = qd / TL;
      ll = qd - k * TL;
    }
    cplx<R> w = buf[lds_idx(ll, k, ls)];
    if (P.four == 1) {
      // w_n^(j1 k2), j1 k2 < n
      cplx<R> t = P.twiddle_full[s_sub[ll] * k];
      if (SIGN > 0) t.y = -t.y;
      w = cmul<R>(w, t);
    }
    const int kp = out_pos(ll, k);
    if (P.out_mode == kToGrid) {
      gdst[s_goff[ll] + (long)kp * P.stride] = w;
    } else {
      const int ks = band_coord(kp, nax, P.Nband);
      if (ks >= 0) {
        if (P.out_mode == kToGridBand) {
          gdst[s_goff[ll] + (long)kp * P.stride] = w;
        } else {
          const R f = s_fac[ll] * P.phi_ax[ks];
          fout[s_foff[ll] + (long)ks * P.fstr_ax] = cplx<R>{w.x * f, w.y * f};
        }
      }
    }
  }
}

// elements of one workgroup's lines
template <typename R> constexpr int fft_cap() { return sizeof(R) == 4 ? 4096 : 2048; }

template <typename R>
void launch_fft_pass(cplx<R> *g, const cplx<R> *fhat_in, cplx<R> *fhat_out, const FftPass<R> &p0,
                     int sign, hipStream_t s, int nb, cplx<R> *g2) {
  if (p0.nlines <= 0) return;
  // Fewer lines per workgroup than fit in LDS when the launch would leave
  // the GPU short of workgroups (small grids: aim for 1024 of them over all
  // right-hand sides), keeping at least 2 lines (4 along a strided axis, so
  // that each load covers 32 contiguous bytes or more).
  FftPass<R> p = p0;
  {
    const long fit = p0.lines_per_block;
    long want = std::max<long>(1, p0.nlines * (long)nb / 1024);
    want = p0.lpb_shift >= 0 ? want << p0.lpb_shift : std::max<long>(1, want >> -p0.lpb_shift);
    const long floor_lines = std::min<long>(fit, p0.stride == 1 ? 2 : 4);
    p.lines_per_block = (int)std::max(floor_lines, std::min(fit, want));
    const size_t line_bytes = sizeof(cplx<R>) * (size_t)p.lines_per_block * lds_line_stride(p.len);
    p.twiddle_in_lds = line_bytes + sizeof(cplx<R>) * (size_t)p.len <= 64 * 1024 ? 1 : 0;
  }
  const long blocks = ceil_div(p.nlines, p.lines_per_block);
  const size_t lds = sizeof(cplx<R>) * ((size_t)p.lines_per_block * (size_t)lds_line_stride(p.len) +
                                        (p.twiddle_in_lds ? (size_t)p.len : 0));
  constexpr int CAP = fft_cap<R>();
  bool pow2 = true;
  for (int i = 0; i < p.nstages; i++) pow2 = pow2 && (p.radix[i] == 2 || p.radix[i] == 4 || p.radix[i] == 8);
  auto fn = sign < 0 ? (pow2 ? k_fft_pass<R, -1, CAP, true> : k_fft_pass<R, -1, CAP, false>)
                     : (pow2 ? k_fft_pass<R, +1, CAP, true> : k_fft_pass<R, +1, CAP, false>);
  cplx<R> *src = p.src ? g2 : g;
  cplx<R> *dst = p.dst ? g2 : g;
  hipLaunchKernelGGL(fn, dim3((unsigned)blocks, (unsigned)nb), dim3(256), lds, s,
                     (const cplx<R> *)src, dst, fhat_in, fhat_out, p);
}

template void launch_fft_pass<float>(cplx<float> *, const cplx<float> *, cplx<float> *,
                                     const FftPass<float> &, int, hipStream_t, int, cplx<float> *);
template void launch_fft_pass<double>(cplx<double> *, const cplx<double> *, cplx<double> *,
                                      const FftPass<double> &, int, hipStream_t, int,
                                      cplx<double> *);

}  // namespace nfft_hip
