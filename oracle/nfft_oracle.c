/*
 * CPU FP64 oracle of the NFFT hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with double / long double arithmetic, what
 * the reference NFFT3 computes on the path nfft_trafo / nfft_adjoint /
 * nfft_trafo_direct / nfft_adjoint_direct.  It is the checker for the GPU
 * library: only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  The product library never links or calls it.
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * The code is written from the formulas, not translated from the reference:
 * the reference specialises 1-D/2-D/3-D with octant pointer arithmetic and
 * wrap-around loop variants; here one generic d-dimensional loop nest does
 * each stage.
 *
 * Pinning (what checks each function against something other than itself):
 *   oracle_bessel_i0      -- 100-entry known-answer table of tests/bessel.c
 *   oracle_ndft_*         -- Mathematica 64-digit golden files tests/data
 *   oracle_fft            -- numpy.fft on random inputs, brute-force DFT
 *   oracle_phi/phi_hut, oracle_D/B/BT/DT, oracle_nfft_*
 *                         -- golden files under the reference's own bound
 *                            (tests/nfft.c:217-284), including m = 6 where the
 *                            untruncated sin branch of the window matters;
 *                            adjointness <A fh, f> = <fh, A^H f>
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ORACLE_PLANT selects a deliberately wrong variant, built only by the test
 * that proves the golden checks catch such mistakes:
 *   1 window truncated at |y| = m (outer taps zeroed)
 *   2 deconvolution misses the factor of the last dimension
 *   3 grid index of every tap off by one cell (weights unchanged)
 *   4 FFT sign flipped in the fast transforms */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

typedef long double ld;

#define ORACLE_PI_L 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------------ */
/* Bessel I0: textbook power series  I0(x) = sum_k ((x/2)^2)^k / (k!)^2,     */
/* summed in long double until the terms stop changing the sum.  The          */
/* reference evaluates I0 by rational Chebyshev fits (kernel/util/            */
/* bessel_i0.c:300-338); the value, not the method, is what matters.          */
/* ------------------------------------------------------------------------ */
/* Threads of the FFT and B / B^T loops (OpenMP).  1 by default, so the test
 * suite's results are bit-reproducible; the benchmark's CPU-baseline leg
 * raises it to time the restatement on the host's cores.  With more than
 * one thread B^T adds into the grid with atomic updates. */
static int oracle_threads = 1;
void oracle_set_threads(int n) { oracle_threads = n > 0 ? n : 1; }
int oracle_get_threads(void) { return oracle_threads; }

double oracle_bessel_i0(double x)
{
  ld q = ((ld)x * (ld)x) / 4.0L;
  ld term = 1.0L, sum = 1.0L;
  for (int k = 1; k < 2000; k++)
  {
    term *= q / ((ld)k * (ld)k);
    sum += term;
    if (term < sum * 1e-22L)
      break;
  }
  return (double)sum;
}

/* Kaiser-Bessel shape parameter b = pi (2 - 1/sigma)  (include/infft.h:216-222) */
double oracle_kb_b(double sigma)
{
  return M_PI * (2.0 - 1.0 / sigma);
}

/* Window phi(x) with y = n x  (include/infft.h:209-215): not truncated; for
 * |y| > m the sin branch applies, at |y| = m the limit b/pi. */
double oracle_phi(long n, double x, long m, double b)
{
  double y = x * (double)n;
  double q = (double)m * (double)m - y * y;
  if (q > 0.0)
  {
    double s = sqrt(q);
    return sinh(b * s) / (M_PI * s);
  }
  if (q < 0.0)
  {
#if ORACLE_PLANT == 1
    /* planted mistake (tests only): truncate the window at |y| = m */
    return 0.0;
#endif
    double s = sqrt(-q);
    return sin(b * s) / (M_PI * s);
  }
  return b / M_PI;
}

/* Fourier coefficient of the window  (include/infft.h:208) */
double oracle_phi_hut(long n, long k, long m, double b)
{
  double w = 2.0 * M_PI * (double)k / (double)n;
  return oracle_bessel_i0((double)m * sqrt(b * b - w * w));
}

/* c_phi_inv[t][ks] = 1 / phi_hut(n_t, ks - N_t/2)  (kernel/nfft/nfft.c:5754-5770).
 * out has sum_t N[t] entries, dimension t starting at offset sum_{s<t} N[s]. */
void oracle_c_phi_inv(int d, const long *N, const long *n, long m, double *out)
{
  long off = 0;
  for (int t = 0; t < d; t++)
  {
    double sigma = (double)n[t] / (double)N[t];
    double b = oracle_kb_b(sigma);
    for (long ks = 0; ks < N[t]; ks++)
      out[off + ks] = 1.0 / oracle_phi_hut(n[t], ks - N[t] / 2, m, b);
    off += N[t];
  }
}

/* ------------------------------------------------------------------------ */
/* NDFT (kernel/nfft/nfft.c:145-297)                                          */
/*   trafo:   f_j   = sum_k f_hat_k exp(-2 pi i k.x_j)                        */
/*   adjoint: f_hat_k = sum_j f_j   exp(+2 pi i k.x_j)                        */
/* k_t = ks_t - N_t/2, f_hat row-major with the last dimension fastest.       */
/* The phase is accumulated in long double and reduced mod 1 before the       */
/* trigonometric call, so the oracle is more accurate than the reference's    */
/* double cexp; both meet the reference's 48 eps bound (tests/nfft.c:211).    */
/* Complex arrays are interleaved (re, im).                                   */
/* ------------------------------------------------------------------------ */
static void unravel(long kL, int d, const long *N, long *k)
{
  for (int t = d - 1; t >= 0; t--)
  {
    k[t] = kL % N[t] - N[t] / 2;
    kL /= N[t];
  }
}

static void phase_exp(ld phase, int sign, ld *c, ld *s)
{
  /* exp(sign * 2 pi i phase) */
  ld fr = phase - floorl(phase);
  ld a = 2.0L * ORACLE_PI_L * fr;
  *c = cosl(a);
  *s = (ld)sign * sinl(a);
}

void oracle_ndft_trafo(int d, const long *N, long M, const double *x,
                       const double *f_hat, double *f)
{
  long Nt = 1;
  for (int t = 0; t < d; t++)
    Nt *= N[t];
  long k[16];
  for (long j = 0; j < M; j++)
  {
    ld re = 0.0L, im = 0.0L;
    for (long kL = 0; kL < Nt; kL++)
    {
      unravel(kL, d, N, k);
      ld ph = 0.0L;
      for (int t = 0; t < d; t++)
        ph += (ld)k[t] * (ld)x[j * d + t];
      ld c, s;
      phase_exp(ph, -1, &c, &s);
      ld ar = f_hat[2 * kL], ai = f_hat[2 * kL + 1];
      re += ar * c - ai * s;
      im += ar * s + ai * c;
    }
    f[2 * j] = (double)re;
    f[2 * j + 1] = (double)im;
  }
}

void oracle_ndft_adjoint(int d, const long *N, long M, const double *x,
                         const double *f, double *f_hat)
{
  long Nt = 1;
  for (int t = 0; t < d; t++)
    Nt *= N[t];
  long k[16];
  for (long kL = 0; kL < Nt; kL++)
  {
    unravel(kL, d, N, k);
    ld re = 0.0L, im = 0.0L;
    for (long j = 0; j < M; j++)
    {
      ld ph = 0.0L;
      for (int t = 0; t < d; t++)
        ph += (ld)k[t] * (ld)x[j * d + t];
      ld c, s;
      phase_exp(ph, +1, &c, &s);
      ld ar = f[2 * j], ai = f[2 * j + 1];
      re += ar * c - ai * s;
      im += ar * s + ai * c;
    }
    f_hat[2 * kL] = (double)re;
    f_hat[2 * kL + 1] = (double)im;
  }
}

/* ------------------------------------------------------------------------ */
/* Multi-dimensional unnormalised DFT of the oversampled grid, the job FFTW   */
/* does for the reference (plans at kernel/nfft/nfft.c:6030-6031):            */
/*   sign = -1 (FFTW_FORWARD, trafo),  sign = +1 (FFTW_BACKWARD, adjoint).    */
/* One 1-D transform per axis: iterative radix-2 for powers of two, a direct  */
/* O(n^2) DFT with long double twiddles otherwise.                            */
/* ------------------------------------------------------------------------ */
static void fft1d_pow2(double *re, double *im, long n, int sign)
{
  /* bit reversal */
  for (long i = 1, j = 0; i < n; i++)
  {
    long bit = n >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j)
    {
      double tr = re[i], ti = im[i];
      re[i] = re[j]; im[i] = im[j];
      re[j] = tr; im[j] = ti;
    }
  }
  for (long len = 2; len <= n; len <<= 1)
  {
    long half = len >> 1;
    for (long k = 0; k < half; k++)
    {
      ld a = (ld)sign * 2.0L * ORACLE_PI_L * (ld)k / (ld)len;
      double wr = (double)cosl(a), wi = (double)sinl(a);
      for (long s = 0; s < n; s += len)
      {
        long p = s + k, q = p + half;
        double xr = re[q] * wr - im[q] * wi;
        double xi = re[q] * wi + im[q] * wr;
        re[q] = re[p] - xr; im[q] = im[p] - xi;
        re[p] += xr; im[p] += xi;
      }
    }
  }
}

static void dft1d_direct(double *re, double *im, long n, int sign)
{
  ld *tr = (ld *)malloc(sizeof(ld) * (size_t)n * 2);
  for (long k = 0; k < n; k++)
  {
    ld sr = 0.0L, si = 0.0L;
    for (long l = 0; l < n; l++)
    {
      long e = (k * l) % n;
      ld a = (ld)sign * 2.0L * ORACLE_PI_L * (ld)e / (ld)n;
      ld c = cosl(a), s = sinl(a);
      sr += re[l] * c - im[l] * s;
      si += re[l] * s + im[l] * c;
    }
    tr[2 * k] = sr;
    tr[2 * k + 1] = si;
  }
  for (long k = 0; k < n; k++)
  {
    re[k] = (double)tr[2 * k];
    im[k] = (double)tr[2 * k + 1];
  }
  free(tr);
}

void oracle_fft(int d, const long *n, double *data, int sign)
{
  long total = 1;
  for (int t = 0; t < d; t++)
    total *= n[t];
  long inner = total;
  for (int t = 0; t < d; t++)
  {
    long len = n[t];
    inner /= len;
    long outer = total / (len * inner);
    int pow2 = (len & (len - 1)) == 0;
#pragma omp parallel num_threads(oracle_threads)
    {
    double *re = (double *)malloc(sizeof(double) * (size_t)len);
    double *im = (double *)malloc(sizeof(double) * (size_t)len);
#pragma omp for collapse(2) schedule(static)
    for (long o = 0; o < outer; o++)
      for (long i = 0; i < inner; i++)
      {
        double *base = data + 2 * (o * len * inner + i);
        for (long l = 0; l < len; l++)
        {
          re[l] = base[2 * l * inner];
          im[l] = base[2 * l * inner + 1];
        }
        if (pow2)
          fft1d_pow2(re, im, len, sign);
        else
          dft1d_direct(re, im, len, sign);
        for (long l = 0; l < len; l++)
        {
          base[2 * l * inner] = re[l];
          base[2 * l * inner + 1] = im[l];
        }
      }
    free(re);
    free(im);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* D: g_hat[k mod n] = f_hat[k] * prod_t c_phi_inv[t][ks_t], zero elsewhere  */
/*    (kernel/nfft/nfft.c:2786-2831, 3833-3897, 5415-5513).                  */
/* DT: f_hat[k] = g_hat[k mod n] * prod_t c_phi_inv[t][ks_t]                 */
/*    (kernel/nfft/nfft.c:2874-2903, 3941-3997, 5560-5650).                  */
/* ------------------------------------------------------------------------ */
static long grid_index_of_freq(int d, const long *N, const long *n, long kL)
{
  long ks[16], g = 0;
  for (int t = d - 1; t >= 0; t--)
  {
    ks[t] = kL % N[t];
    kL /= N[t];
  }
  for (int t = 0; t < d; t++)
  {
    long k = ks[t] - N[t] / 2;
    long l = ((k % n[t]) + n[t]) % n[t];
    g = g * n[t] + l;
  }
  return g;
}

static double deconv_factor(int d, const long *N, const double *cpi, long kL)
{
  long ks[16];
  for (int t = d - 1; t >= 0; t--)
  {
    ks[t] = kL % N[t];
    kL /= N[t];
  }
  double c = 1.0;
  long off = 0;
  for (int t = 0; t < d; t++)
  {
#if ORACLE_PLANT == 2
    if (t == d - 1)
      break;
#endif
    c *= cpi[off + ks[t]];
    off += N[t];
  }
  return c;
}

void oracle_D(int d, const long *N, const long *n, long m, const double *f_hat,
              double *g_hat)
{
  long Nt = 1, nt = 1, sN = 0;
  for (int t = 0; t < d; t++)
  {
    Nt *= N[t];
    nt *= n[t];
    sN += N[t];
  }
  double *cpi = (double *)malloc(sizeof(double) * (size_t)sN);
  oracle_c_phi_inv(d, N, n, m, cpi);
  memset(g_hat, 0, sizeof(double) * 2 * (size_t)nt);
  for (long kL = 0; kL < Nt; kL++)
  {
    long g = grid_index_of_freq(d, N, n, kL);
    double c = deconv_factor(d, N, cpi, kL);
    g_hat[2 * g] = f_hat[2 * kL] * c;
    g_hat[2 * g + 1] = f_hat[2 * kL + 1] * c;
  }
  free(cpi);
}

void oracle_DT(int d, const long *N, const long *n, long m, const double *g_hat,
               double *f_hat)
{
  long Nt = 1, sN = 0;
  for (int t = 0; t < d; t++)
  {
    Nt *= N[t];
    sN += N[t];
  }
  double *cpi = (double *)malloc(sizeof(double) * (size_t)sN);
  oracle_c_phi_inv(d, N, n, m, cpi);
  for (long kL = 0; kL < Nt; kL++)
  {
    long g = grid_index_of_freq(d, N, n, kL);
    double c = deconv_factor(d, N, cpi, kL);
    f_hat[2 * kL] = g_hat[2 * g] * c;
    f_hat[2 * kL + 1] = g_hat[2 * g + 1] * c;
  }
  free(cpi);
}

/* ------------------------------------------------------------------------ */
/* B / B^T.  Per node and dimension: c = floor(n x), taps l = c-m .. c+m+1   */
/* (2m+2 of them, kernel/nfft/nfft.c:324-340), weight                        */
/* psi_t(l) = phi(n_t, x_t - l/n_t) (kernel/nfft/nfft.c:5819-5844), grid     */
/* index l mod n_t, row-major.                                               */
/*   B:   f_j  = sum_l prod_t psi_t(l_t) g[l mod n]                          */
/*        (kernel/nfft/nfft.c:2131-2153, 2927-3004, 4020-4265)               */
/*   B^T: g[l mod n] += prod_t psi_t(l_t) f_j  on a zeroed g                 */
/*        (kernel/nfft/nfft.c:2585-2773, 3583-3805, 5126-5384)               */
/* ------------------------------------------------------------------------ */
typedef struct
{
  int d;
  long m, w; /* w = 2m+2 taps per dimension */
  long n[16];
  double *psi;  /* d * w weights of one node */
  long *idx;    /* d * w wrapped grid coordinates of one node */
} node_taps;

static void node_taps_fill(node_taps *nt, const double *xj, const double *b)
{
  for (int t = 0; t < nt->d; t++)
  {
    long n = nt->n[t];
    long c = (long)floor(xj[t] * (double)n);
    long u = c - nt->m;
    for (long l = 0; l < nt->w; l++)
    {
      long gl = u + l;
      nt->psi[t * nt->w + l] = oracle_phi(n, xj[t] - (double)gl / (double)n,
                                          nt->m, b[t]);
#if ORACLE_PLANT == 3
      gl += 1;
#endif
      nt->idx[t * nt->w + l] = ((gl % n) + n) % n;
    }
  }
}

/* visit all (2m+2)^d taps of every node with an odometer over the tap
 * indices, last dimension fastest */
static void B_core(int d, const long *N, const long *n, long m, long M,
                   const double *x, double *g, double *f, int adjoint)
{
  double b[16];
  for (int t = 0; t < d; t++)
    b[t] = oracle_kb_b((double)n[t] / (double)N[t]);
  const int atomic_add = adjoint && oracle_threads > 1;
#pragma omp parallel num_threads(oracle_threads)
  {
  node_taps nt;
  nt.d = d;
  nt.m = m;
  nt.w = 2 * m + 2;
  for (int t = 0; t < d; t++)
    nt.n[t] = n[t];
  nt.psi = (double *)malloc(sizeof(double) * (size_t)(d * nt.w));
  nt.idx = (long *)malloc(sizeof(long) * (size_t)(d * nt.w));
  long taps = 1;
  for (int t = 0; t < d; t++)
    taps *= nt.w;
  long lt[16];
#pragma omp for schedule(dynamic, 256)
  for (long j = 0; j < M; j++)
  {
    node_taps_fill(&nt, x + j * d, b);
    ld re = 0.0L, im = 0.0L;
    double fr = 0.0, fi = 0.0;
    if (adjoint)
    {
      fr = f[2 * j];
      fi = f[2 * j + 1];
    }
    for (int t = 0; t < d; t++)
      lt[t] = 0;
    for (long q = 0; q < taps; q++)
    {
      double wgt = 1.0;
      long gi = 0;
      for (int t = 0; t < d; t++)
      {
        wgt *= nt.psi[t * nt.w + lt[t]];
        gi = gi * n[t] + nt.idx[t * nt.w + lt[t]];
      }
      if (adjoint && atomic_add)
      {
#pragma omp atomic
        g[2 * gi] += wgt * fr;
#pragma omp atomic
        g[2 * gi + 1] += wgt * fi;
      }
      else if (adjoint)
      {
        g[2 * gi] += wgt * fr;
        g[2 * gi + 1] += wgt * fi;
      }
      else
      {
        re += (ld)wgt * (ld)g[2 * gi];
        im += (ld)wgt * (ld)g[2 * gi + 1];
      }
      /* odometer, last dimension fastest */
      for (int t = d - 1; t >= 0; t--)
      {
        if (++lt[t] < nt.w)
          break;
        lt[t] = 0;
      }
    }
    if (!adjoint)
    {
      f[2 * j] = (double)re;
      f[2 * j + 1] = (double)im;
    }
  }
  free(nt.psi);
  free(nt.idx);
  }
}

void oracle_B(int d, const long *N, const long *n, long m, long M,
              const double *x, const double *g, double *f)
{
  B_core(d, N, n, m, M, x, (double *)g, f, 0);
}

void oracle_BT(int d, const long *N, const long *n, long m, long M,
               const double *x, const double *f, double *g)
{
  long nt = 1;
  for (int t = 0; t < d; t++)
    nt *= n[t];
  memset(g, 0, sizeof(double) * 2 * (size_t)nt);
  B_core(d, N, n, m, M, x, g, (double *)f, 1);
}

/* ------------------------------------------------------------------------ */
/* Full fast transforms (kernel/nfft/nfft.c:5655-5749): the direct-NDFT       */
/* fallback when some N_t <= m or n_t <= 2m+2, otherwise D -> F -> B and      */
/* B^T -> F -> D^T.                                                           */
/* ------------------------------------------------------------------------ */
static int use_direct(int d, const long *N, const long *n, long m)
{
  for (int t = 0; t < d; t++)
    if (N[t] <= m || n[t] <= 2 * m + 2)
      return 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* 3-D B / B^T shaped as the reference's PRE_PSI path, for the CPU baseline  */
/* timing only (results equal oracle_B / oracle_BT to rounding):             */
/*   precompute: psi[j][t][l] = phi(n_t, x_t - (u_t + l)/n_t) and the start  */
/*     u_t = floor(n_t x_t) - m per node (kernel/nfft/nfft.c:5819-5844)      */
/*   trafo:  per node, tensor loops l0, l1 (weights multiplied outside) and  */
/*     a contiguous run over l2 (kernel/nfft/nfft.c:4020-4265, 4687-4914),   */
/*     OpenMP over nodes                                                     */
/*   adjoint: blockwise over the first grid dimension, as                    */
/*     NFFT_OMP_BLOCKWISE_ADJOINT (kernel/nfft/nfft.c:5126-5384, 4289-4685): */
/*     thread k owns grid planes [k n0 / T, (k+1) n0 / T) and adds the taps  */
/*     of every node that fall in its planes, so no atomics are needed      */
/* ------------------------------------------------------------------------ */
void oracle_prepsi_3d(const long *N, const long *n, long m, long M, const double *x,
                      double *psi, long *u)
{
  const long w = 2 * m + 2;
  double b[3];
  for (int t = 0; t < 3; t++)
    b[t] = oracle_kb_b((double)n[t] / (double)N[t]);
#pragma omp parallel for num_threads(oracle_threads) schedule(static)
  for (long j = 0; j < M; j++)
    for (int t = 0; t < 3; t++)
    {
      const long c = (long)floor(x[3 * j + t] * (double)n[t]);
      u[3 * j + t] = c - m;
      for (long l = 0; l < w; l++)
        psi[(j * 3 + t) * w + l] =
            oracle_phi(n[t], x[3 * j + t] - (double)(c - m + l) / (double)n[t], m, b[t]);
    }
}

static inline long wrapl(long v, long n) { return ((v % n) + n) % n; }

void oracle_B_3d_prepsi(const long *n, long m, long M, const double *psi, const long *u,
                        const double *g, double *f)
{
  const long w = 2 * m + 2;
#pragma omp parallel for num_threads(oracle_threads) schedule(static, 1024)
  for (long j = 0; j < M; j++)
  {
    const double *p0 = psi + (j * 3 + 0) * w, *p1 = psi + (j * 3 + 1) * w,
                 *p2 = psi + (j * 3 + 2) * w;
    double re = 0.0, im = 0.0;
    for (long l0 = 0; l0 < w; l0++)
    {
      const long i0 = wrapl(u[3 * j] + l0, n[0]);
      for (long l1 = 0; l1 < w; l1++)
      {
        const double w01 = p0[l0] * p1[l1];
        const long row = (i0 * n[1] + wrapl(u[3 * j + 1] + l1, n[1])) * n[2];
        const long c0 = wrapl(u[3 * j + 2], n[2]);
        double sr = 0.0, si = 0.0;
        for (long l2 = 0; l2 < w; l2++)
        {
          long c = c0 + l2;
          if (c >= n[2]) c -= n[2];
          sr += p2[l2] * g[2 * (row + c)];
          si += p2[l2] * g[2 * (row + c) + 1];
        }
        re += w01 * sr;
        im += w01 * si;
      }
    }
    f[2 * j] = re;
    f[2 * j + 1] = im;
  }
}

void oracle_BT_3d_prepsi(const long *n, long m, long M, const double *psi, const long *u,
                         const double *f, double *g)
{
  const long w = 2 * m + 2;
  memset(g, 0, sizeof(double) * 2 * (size_t)(n[0] * n[1] * n[2]));
#pragma omp parallel num_threads(oracle_threads)
  {
    const int T = omp_get_num_threads(), k = omp_get_thread_num();
    const long p_lo = k * n[0] / T, p_hi = (k + 1) * n[0] / T;
    for (long j = 0; j < M; j++)
    {
      const double *p0 = psi + (j * 3 + 0) * w, *p1 = psi + (j * 3 + 1) * w,
                   *p2 = psi + (j * 3 + 2) * w;
      for (long l0 = 0; l0 < w; l0++)
      {
        const long i0 = wrapl(u[3 * j] + l0, n[0]);
        if (i0 < p_lo || i0 >= p_hi)
          continue;
        const double a0r = p0[l0] * f[2 * j], a0i = p0[l0] * f[2 * j + 1];
        for (long l1 = 0; l1 < w; l1++)
        {
          const double ar = a0r * p1[l1], ai = a0i * p1[l1];
          const long row = (i0 * n[1] + wrapl(u[3 * j + 1] + l1, n[1])) * n[2];
          long c = wrapl(u[3 * j + 2], n[2]);
          for (long l2 = 0; l2 < w; l2++)
          {
            g[2 * (row + c)] += p2[l2] * ar;
            g[2 * (row + c) + 1] += p2[l2] * ai;
            if (++c == n[2])
              c = 0;
          }
        }
      }
    }
  }
}

void oracle_nfft_trafo(int d, const long *N, const long *n, long m, long M,
                       const double *x, const double *f_hat, double *f)
{
  if (use_direct(d, N, n, m))
  {
    oracle_ndft_trafo(d, N, M, x, f_hat, f);
    return;
  }
  long nt = 1;
  for (int t = 0; t < d; t++)
    nt *= n[t];
  double *g = (double *)malloc(sizeof(double) * 2 * (size_t)nt);
  oracle_D(d, N, n, m, f_hat, g);
  oracle_fft(d, n, g, ORACLE_PLANT == 4 ? +1 : -1);
  oracle_B(d, N, n, m, M, x, g, f);
  free(g);
}

void oracle_nfft_adjoint(int d, const long *N, const long *n, long m, long M,
                         const double *x, const double *f, double *f_hat)
{
  if (use_direct(d, N, n, m))
  {
    oracle_ndft_adjoint(d, N, M, x, f, f_hat);
    return;
  }
  long nt = 1;
  for (int t = 0; t < d; t++)
    nt *= n[t];
  double *g = (double *)malloc(sizeof(double) * 2 * (size_t)nt);
  oracle_BT(d, N, n, m, M, x, f, g);
  oracle_fft(d, n, g, ORACLE_PLANT == 4 ? -1 : +1);
  oracle_DT(d, N, n, m, g, f_hat);
  free(g);
}

/* next_power_of_2 with the reference's special cases (kernel/util/int.c:101-132):
 * negative -> -1, 0 -> 1, 1 -> 2, otherwise the smallest power of two >= x. */
long oracle_next_power_of_2(long x)
{
  if (x < 0)
    return -1;
  if (x < 2)
    return x + 1;
  long p = 1;
  while (p < x)
    p <<= 1;
  return p;
}
