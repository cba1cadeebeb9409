// Kaiser-Bessel window of the reference (include/infft.h:209-215) evaluated on
// the device, divided by its peak phi(0) = sinh(b m) / (pi m):
//
//   psi(y) = (m/s) sinh(b s) / sinh(b m),   s = sqrt(m^2 - y^2),  |y| < m
//          = (m/s) sin(b s)  / sinh(b m),   s = sqrt(y^2 - m^2),  |y| > m
//          = b m / sinh(b m),                                     |y| = m
//
// with y = n x - l the distance of the node from grid point l in cells.  The
// window is NOT truncated at |y| = m: the two outermost of the 2m+2 taps
// fall in the sin branch, as in the reference.  sinh ratios are written with
// exp/expm1 so that nothing overflows in FP32 for any m.  The 1/phi(0) scale
// is folded back into the deconvolution factors on the host, so B D is the
// reference's B D exactly (up to rounding).
#pragma once

#include "common.hpp"

namespace nfft_hip {

template <typename R>
__device__ __forceinline__ R kb_psi(R y, R m, R b, R inv_den) {
  const R q = m * m - y * y;
  if (q > R(0)) {
    const R s = sqrt(q);
    return (m / s) * exp(b * (s - m)) * (-expm1(R(-2) * b * s)) * inv_den;
  }
  if (q < R(0)) {
    const R s = sqrt(-q);
    return R(2) * (m / s) * sin(b * s) * exp(-b * m) * inv_den;
  }
  return R(2) * b * m * exp(-b * m) * inv_den;
}

// The same window without branches, for FP32 wavefront kernels where every
// lane of a wave evaluates one tap.  The peak branch keeps the accurate exp;
// the factors that only shape the smallest values (1 - exp(-2bs) near |y| = m
// and the sin branch beyond it, both below 1e-6 of the peak) use the fast
// hardware exp / sin.  tail = 2 exp(-b m) / (1 - exp(-2 b m)).
__device__ __forceinline__ float kb_psi_fast(float y, float m, float b, float inv_den,
                                             float tail) {
  const float q = m * m - y * y;
  const float aq = fabsf(q);
  const float rs = __frsqrt_rn(aq);  // 1 / s
  const float s = aq * rs;
  const float pos = (m * rs) * expf(b * (s - m)) * (1.f - __expf(-2.f * b * s)) * inv_den;
  const float neg = (m * rs) * __sinf(b * s) * tail;
  const float zero = b * m * tail;
  return q > 0.f ? pos : (q < 0.f ? neg : zero);
}

}  // namespace nfft_hip
