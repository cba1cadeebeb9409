// Kaiser-Bessel window of the reference (include/infft.h:209-215) evaluated on
// the device, divided by its peak phi(0) = sinh(b m) / (pi m):
//
//   psi(y) = (m/s) sinh(b s) / sinh(b m),   s = sqrt(m^2 - y^2),  |y| < m
//          = (m/s) sin(b s)  / sinh(b m),   s = sqrt(y^2 - m^2),  |y| > m
//          = b m / sinh(b m),                                     |y| = m
//
// with y = n x - l the distance of the node from grid point l in cells.  The
// window is NOT truncated at |y| = m: the two outermost of the 2m+2 taps
// fall in the sin branch, as in the reference.  The kernels evaluate it
// through per-tap polynomials in (frac - 1/2) that the plan fits once in
// long double (WinParams::poly; degree 8 for FP32, 20 for FP64, each at the
// rounding level of its precision), a handful of FMAs per tap instead of a
// square root and exponentials.  The 1/phi(0) scale is folded back into the
// deconvolution factors on the host, so B D is the reference's B D exactly
// (up to rounding).
#pragma once

#include "common.hpp"

namespace nfft_hip {

// The window from the plan's per-tap polynomials (WinParams::poly): pc holds
// the D + 1 coefficients of one tap, lowest first, in t = frac - 1/2, frac
// the node's offset from its cell in [0, 1).
template <typename R, int D>
__device__ __forceinline__ R tap_poly(const R *pc, R frac) {
  const R t = frac - R(0.5);
  R v = pc[D];
#pragma unroll
  for (int k = D - 1; k >= 0; k--) v = fma(v, t, pc[k]);
  return v;
}

// One tap's polynomial held in registers, for kernels whose lanes evaluate
// different taps.
template <typename R, int D>
struct TapCoef {
  R c[D + 1];
  __device__ __forceinline__ void load(const R *pc) {
#pragma unroll
    for (int k = 0; k <= D; k++) c[k] = pc[k];
  }
  __device__ __forceinline__ R operator()(R frac) const { return tap_poly<R, D>(c, frac); }
  // the same at t = frac - 1/2
  __device__ __forceinline__ R at(R t) const {
    R v = c[D];
#pragma unroll
    for (int k = D - 1; k >= 0; k--) v = fma(v, t, c[k]);
    return v;
  }
};

}  // namespace nfft_hip
