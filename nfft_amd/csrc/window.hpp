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

}  // namespace nfft_hip
