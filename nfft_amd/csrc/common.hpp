// Shared definitions of the MI355X NFFT library: complex vector types,
// error handling, and the per-precision hooks into the C API.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nfft_hip {

// Complex value types: interleaved (re, im), the memory layout of
// fftw_complex / fftwf_complex.
template <typename R> struct ctype;
template <> struct ctype<float> { using type = float2; };
template <> struct ctype<double> { using type = double2; };
template <typename R> using cplx = typename ctype<R>::type;

// Formats a message and ends the process through the family's die hook
// (nfft_die / nfftf_die: hook, then exit(EXIT_FAILURE)), the reference's
// only fatal-error channel (kernel/util/malloc.c:58-64).
[[noreturn]] void fatal_float(const char *fmt, ...);
[[noreturn]] void fatal_double(const char *fmt, ...);

template <typename R> struct family;
template <> struct family<float> {
  template <typename... A> [[noreturn]] static void fatal(const char *f, A... a) { fatal_float(f, a...); }
};
template <> struct family<double> {
  template <typename... A> [[noreturn]] static void fatal(const char *f, A... a) { fatal_double(f, a...); }
};

#define NFFT_HIP_CHECK(R, expr)                                                   \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      ::nfft_hip::family<R>::fatal("%s failed (%s:%d): %s", #expr, __FILE__,      \
                                   __LINE__, hipGetErrorString(e_));              \
  } while (0)

inline long ceil_div(long a, long b) { return (a + b - 1) / b; }

// Work items of a bin of cnt nodes: ceil(cnt / max_per_sub) of them, at most
// max_items (0: no limit; dense bins then get longer items and fewer tiles).
__host__ __device__ inline int items_of(int cnt, int max_per_sub, int max_items) {
  const int k = (cnt + max_per_sub - 1) / max_per_sub;
  return max_items > 0 ? (k < max_items ? k : max_items) : k;
}

// Item s of k over a bin of cnt nodes: nodes [first, first + count) of the
// bin, split evenly (counts differ by at most one, every count >= 1 when
// k <= cnt, and they sum to cnt).
__host__ __device__ inline void split_item(int cnt, int k, int s, int &first, int &count) {
  const long lo = (long)s * cnt / k, hi = (long)(s + 1) * cnt / k;
  first = (int)lo;
  count = (int)(hi - lo);
}

// The HIP runtime's per-thread last-error state also records failures of
// calls made elsewhere in the process (other libraries, the runtime's own
// fork handling).  Every entry point clears it first, so that its
// hipGetLastError() checks report only this library's launches.
inline void clear_launch_error() { (void)hipGetLastError(); }

}  // namespace nfft_hip
