// Host utilities of the C API: memory and exit hooks, random numbers,
// printing, norms, BLAS-1 helpers and version queries, for the nfftf_ and
// nfft_ families (reference include/nfft3.h:69-87, 839-894; behaviour of
// kernel/util/{malloc,rand,print,error,int,vector1,vector3,version,window,
// thread,time}.c).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "../../include/nfft3.h"

namespace {

// 64-byte alignment, like FFTW's allocator (which the reference uses).
void *aligned_or_die(size_t n, void (*die)(const char *)) {
  if (n == 0) n = 1;
  void *p = nullptr;
  if (posix_memalign(&p, 64, n) != 0 || !p) die("nfft_malloc: out of memory\n");
  return p;
}

template <typename R> struct Cx { R re, im; };

template <typename R>
R drand() {
  return (R)drand48();
}

template <typename R>
R cabs2(const R *z) {
  return std::sqrt(z[0] * z[0] + z[1] * z[1]);
}

// ||x - y||_inf over complex vectors; y == NULL means ||x||_inf
template <typename R>
R cerr_inf(const R *x, const R *y, NFFT_INT n) {
  R e = 0;
  for (NFFT_INT k = 0; k < n; k++) {
    R d[2] = {x[2 * k] - (y ? y[2 * k] : R(0)), x[2 * k + 1] - (y ? y[2 * k + 1] : R(0))};
    R a = cabs2(d);
    if (a > e) e = a;
  }
  return e;
}

template <typename R>
R cnrm1(const R *x, NFFT_INT n) {
  R s = 0;
  for (NFFT_INT k = 0; k < n; k++) s += cabs2(x + 2 * k);
  return s;
}

// swap halves along every dimension (reference kernel/util/vector3.c fftshift)
template <typename R>
void fftshift(R *x, NFFT_INT d, const NFFT_INT *N) {
  NFFT_INT total = 1;
  for (NFFT_INT t = 0; t < d; t++) total *= N[t];
  R *tmp = (R *)std::malloc(sizeof(R) * 2 * (size_t)(total ? total : 1));
  for (NFFT_INT k = 0; k < total; k++) {
    NFFT_INT r = k, dst = 0, mul = 1;
    for (NFFT_INT t = d - 1; t >= 0; t--) {
      NFFT_INT kt = r % N[t];
      r /= N[t];
      dst += ((kt + N[t] / 2) % N[t]) * mul;
      mul *= N[t];
    }
    tmp[2 * dst] = x[2 * k];
    tmp[2 * dst + 1] = x[2 * k + 1];
  }
  std::memcpy(x, tmp, sizeof(R) * 2 * (size_t)total);
  std::free(tmp);
}

NFFT_INT next_pow2(NFFT_INT x) {
  if (x < 0) return -1;
  if (x < 2) return x + 1;
  NFFT_INT p = 1;
  while (p < x) p <<= 1;
  return p;
}

// ---- damping factors (reference kernel/util/damp.c:25-142) ----
// Each "modified" factor of index kk is the sum of a kernel's Fourier
// coefficients at kk and kk + 1 (the two integer frequencies around the
// half-integer kk + 1/2).

// sum_j a[j] k^j by Horner, degree D
template <typename R, int D>
R horner(const R (&a)[D + 1], R k) {
  R v = a[D];
  for (int j = D - 1; j >= 0; j--) v = v * k + a[j];
  return v;
}

// Fourier coefficient at |k| of the Jackson kernel (sin(n x/2) / sin(x/2))^4,
// normalised to 1 at k = 0: a cubic on [0, n), (2n-k)((2n-k)^2-1) / (2n(2n^2+1))
// on [n, 2n].
template <typename R>
R jackson2_coef(R n, R k) {
  const R den = R(2) * n * (R(2) * n * n + R(1));
  if (k < n) {
    const R a[4] = {0, R(3), R(6) * n, R(-3)};
    return R(1) - horner<R, 3>(a, k) / den;
  }
  const R u = R(2) * n - k;
  return u * (u * u - R(1)) / den;
}

// Fourier coefficient at |k| of the generalised Jackson kernel
// (sin(n x/2) / sin(x/2))^8, normalised to 1 at k = 0: a degree-7 polynomial
// in k on each of [0, n), [n, 2n), [2n, 3n), [3n, 4n), zero beyond.
template <typename R>
R jackson4_coef(R n, R k) {
  const R n2 = n * n, n3 = n2 * n, n4 = n3 * n, n5 = n4 * n, n6 = n5 * n, n7 = n6 * n;
  const R norm = R(2416) * n7 + R(1120) * n5 + R(784) * n3 + R(720) * n;
  if (k < n) {
    const R a[8] = {0,
                    R(1260),
                    R(1680) * n5 + R(2240) * n3 + R(2940) * n,
                    R(-1715),
                    -(R(560) * n3 + R(1400) * n),
                    R(490),
                    R(140) * n,
                    R(-35)};
    return R(1) - horner<R, 7>(a, k) / norm;
  }
  if (k < R(2) * n) {
    const R a[8] = {R(2472) * n7 + R(336) * n5 + R(3528) * n3 - R(1296) * n,
                    -(R(392) * n6 - R(3920) * n4 + R(8232) * n2 - R(756)),
                    -(R(504) * n5 + R(10080) * n3 - R(5292) * n),
                    -(R(1960) * n4 - R(7840) * n2 + R(1029)),
                    R(2520) * n3 - R(2520) * n,
                    -(R(1176) * n2 - R(294)),
                    R(252) * n,
                    R(-21)};
    return horner<R, 7>(a, k) / norm;
  }
  if (k < R(3) * n) {
    const R a[8] = {-(R(1112) * n7 - R(12880) * n5 + R(7448) * n3 - R(720) * n),
                    R(12152) * n6 - R(27440) * n4 + R(8232) * n2 - R(252),
                    -(R(19320) * n5 - R(21280) * n3 + R(2940) * n),
                    R(13720) * n4 - R(7840) * n2 + R(343),
                    -(R(5320) * n3 - R(1400) * n),
                    R(1176) * n2 - R(98),
                    R(-140) * n,
                    R(7)};
    return horner<R, 7>(a, k) / norm;
  }
  if (k < R(4) * n) {
    const R u = R(4) * n - k, u2 = u * u;
    return u * (u2 - R(1)) * (u2 - R(4)) * (u2 - R(9)) / norm;
  }
  return R(0);
}

template <typename R>
R damp_fejer(NFFT_INT N, NFFT_INT kk) {
  const R NN = (R)N;
  return R(2) / (NN * NN) * (R(1) - std::fabs(R(2) * (R)kk + R(1)) / NN);
}

template <typename R>
R damp_jackson2(NFFT_INT N, NFFT_INT kk) {
  const R n = ((R)N / R(2) + R(1)) / R(2);
  R s = 0;
  for (NFFT_INT k = kk; k <= kk + 1; k++) s += jackson2_coef<R>(n, (R)(k < 0 ? -k : k));
  return s;
}

template <typename R>
R damp_jackson4(NFFT_INT N, NFFT_INT kk) {
  const R n = ((R)N / R(2) + R(3)) / R(4);
  R s = 0;
  for (NFFT_INT k = kk; k <= kk + 1; k++) s += jackson4_coef<R>(n, (R)(k < 0 ? -k : k));
  return s;
}

// coefficients |k|^(-2 mu), 1 at k = 0
template <typename R>
R damp_sobolev(R mu, NFFT_INT kk) {
  R s = 0;
  for (NFFT_INT k = kk; k <= kk + 1; k++) {
    const NFFT_INT a = k < 0 ? -k : k;
    s += a == 0 ? R(1) : std::pow((R)a, R(-2) * mu);
  }
  return s;
}

// coefficients (k^2 + c^2)^(-mu)
template <typename R>
R damp_multiquadric(R mu, R c, NFFT_INT kk) {
  R s = 0;
  for (NFFT_INT k = kk; k <= kk + 1; k++) {
    const NFFT_INT a = k < 0 ? -k : k;
    s += std::pow((R)(a * a) + c * c, -mu);
  }
  return s;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ malloc / die
#define NFFT_HIP_MALLOC_API(P)                                                   \
  P##malloc_type_function P##malloc_hook = 0;                                   \
  P##free_type_function P##free_hook = 0;                                       \
  P##die_type_function P##die_hook = 0;                                         \
  void P##die(const char *s) {                                                  \
    if (P##die_hook) P##die_hook(s);                                            \
    std::fprintf(stderr, "%s\n", s);                                            \
    std::exit(EXIT_FAILURE);                                                    \
  }                                                                             \
  void *P##malloc(size_t n) {                                                   \
    if (P##malloc_hook) return P##malloc_hook(n);                               \
    return aligned_or_die(n, P##die);                                           \
  }                                                                             \
  void P##free(void *p) {                                                       \
    if (!p) return;                                                             \
    if (P##free_hook) {                                                         \
      P##free_hook(p);                                                          \
      return;                                                                   \
    }                                                                           \
    std::free(p);                                                               \
  }

NFFT_HIP_MALLOC_API(nfftf_)
NFFT_HIP_MALLOC_API(nfft_)

// ------------------------------------------------------------ util API
#define NFFT_HIP_UTIL_API(P, R, C)                                                     \
  R P##drand48(void) { return drand<R>(); }                                           \
  void P##srand48(long int seed) { ::srand48(seed); }                                 \
  void P##vrand_unit_complex(C *x, const NFFT_INT n) {                                \
    for (NFFT_INT k = 0; k < n; k++) {                                                \
      x[k][0] = drand<R>();                                                           \
      x[k][1] = drand<R>();                                                           \
    }                                                                                 \
  }                                                                                   \
  void P##vrand_shifted_unit_double(R *x, const NFFT_INT n) {                         \
    for (NFFT_INT k = 0; k < n; k++) x[k] = drand<R>() - R(0.5);                      \
  }                                                                                   \
  void P##vrand_real(R *x, const NFFT_INT n, const R a, const R b) {                  \
    for (NFFT_INT k = 0; k < n; k++) x[k] = a + drand<R>() * (b - a);                 \
  }                                                                                   \
  void P##vpr_double(R *x, const NFFT_INT n, const char *text) {                      \
    if (text) std::printf("\n %s, adr=%p\n", text, (void *)x);                        \
    for (NFFT_INT k = 0; k < n; k++) {                                                \
      if (k % 8 == 0) std::printf("%6td.\t", (ptrdiff_t)k);                           \
      std::printf("%+.1e,", (double)x[k]);                                            \
      if (k % 8 == 7) std::printf("\n");                                              \
    }                                                                                 \
    if (n % 8 != 0) std::printf("\n");                                                \
  }                                                                                   \
  void P##vpr_complex(C *x, const NFFT_INT n, const char *text) {                     \
    if (text) std::printf("\n %s, adr=%p\n", text, (void *)x);                        \
    for (NFFT_INT k = 0; k < n; k++) {                                                \
      if (k % 4 == 0) std::printf("%6td.\t", (ptrdiff_t)k);                           \
      std::printf("%+.1e%+.1ei,", (double)x[k][0], (double)x[k][1]);                  \
      if (k % 4 == 3) std::printf("\n");                                              \
    }                                                                                 \
    if (n % 4 != 0) std::printf("\n");                                                \
  }                                                                                   \
  NFFT_INT P##get_num_threads(void) { return 1; }                                     \
  void P##set_num_threads(NFFT_INT nthreads) { (void)nthreads; }                      \
  NFFT_INT P##has_threads_enabled(void) { return 0; }                                 \
  R P##clock_gettime_seconds(void) {                                                  \
    struct timespec ts;                                                               \
    clock_gettime(CLOCK_MONOTONIC, &ts);                                              \
    return (R)ts.tv_sec + (R)ts.tv_nsec * R(1e-9);                                    \
  }                                                                                   \
  R P##error_l_infty_complex(const C *x, const C *y, const NFFT_INT n) {              \
    return cerr_inf<R>((const R *)x, (const R *)y, n) /                               \
           cerr_inf<R>((const R *)x, nullptr, n);                                     \
  }                                                                                   \
  R P##error_l_infty_1_complex(const C *x, const C *y, const NFFT_INT n, const C *z,  \
                               const NFFT_INT m) {                                    \
    return cerr_inf<R>((const R *)x, (const R *)y, n) / cnrm1<R>((const R *)z, m);    \
  }                                                                                   \
  NFFT_INT P##exp2i(const NFFT_INT a) { return (NFFT_INT)1 << a; }                    \
  NFFT_INT P##next_power_of_2(const NFFT_INT N) { return next_pow2(N); }              \
  R P##dot_complex(C *x, NFFT_INT n) {                                                \
    R s = 0;                                                                          \
    for (NFFT_INT k = 0; k < n; k++) s += x[k][0] * x[k][0] + x[k][1] * x[k][1];      \
    return s;                                                                         \
  }                                                                                   \
  void P##upd_axpy_complex(C *x, R a, C *y, NFFT_INT n) {                             \
    for (NFFT_INT k = 0; k < n; k++) {                                                \
      x[k][0] = a * x[k][0] + y[k][0];                                                \
      x[k][1] = a * x[k][1] + y[k][1];                                                \
    }                                                                                 \
  }                                                                                   \
  void P##fftshift_complex(C *x, NFFT_INT d, NFFT_INT *N) { fftshift<R>((R *)x, d, N); } \
  void P##fftshift_complex_int(C *x, int d, int *N) {                                 \
    NFFT_INT NN[16];                                                                  \
    for (int t = 0; t < d && t < 16; t++) NN[t] = N[t];                               \
    fftshift<R>((R *)x, d, NN);                                                       \
  }                                                                                   \
  void P##get_version(unsigned *major, unsigned *minor, unsigned *patch) {            \
    *major = 3;                                                                       \
    *minor = 5;                                                                       \
    *patch = 4;                                                                       \
  }                                                                                   \
  const char *P##get_window_name(void) { return "KAISER_BESSEL"; }                    \
  NFFT_INT P##get_default_window_cut_off(void) { return sizeof(R) == 4 ? 4 : 8; }

NFFT_HIP_UTIL_API(nfftf_, float, fftwf_complex)
NFFT_HIP_UTIL_API(nfft_, double, fftw_complex)

// ------------------------------------------------------------ damping factors
#define NFFT_HIP_DAMP_API(P, R)                                                          \
  R P##modified_fejer(const NFFT_INT N, const NFFT_INT kk) { return damp_fejer<R>(N, kk); } \
  R P##modified_jackson2(const NFFT_INT N, const NFFT_INT kk) {                          \
    return damp_jackson2<R>(N, kk);                                                     \
  }                                                                                     \
  R P##modified_jackson4(const NFFT_INT N, const NFFT_INT kk) {                          \
    return damp_jackson4<R>(N, kk);                                                     \
  }                                                                                     \
  R P##modified_sobolev(const R mu, const NFFT_INT kk) { return damp_sobolev<R>(mu, kk); } \
  R P##modified_multiquadric(const R mu, const R c, const NFFT_INT kk) {                 \
    return damp_multiquadric<R>(mu, c, kk);                                             \
  }

NFFT_HIP_DAMP_API(nfftf_, float)
NFFT_HIP_DAMP_API(nfft_, double)

}  // extern "C"

namespace nfft_hip {

static char g_msg[1024];

[[noreturn]] void fatal_float(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  nfftf_die(g_msg);
  std::exit(EXIT_FAILURE);
}

[[noreturn]] void fatal_double(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  nfft_die(g_msg);
  std::exit(EXIT_FAILURE);
}

}  // namespace nfft_hip
