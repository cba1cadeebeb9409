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
