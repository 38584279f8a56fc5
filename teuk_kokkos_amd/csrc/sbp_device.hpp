// SBP radial operators on values staged in LDS (or any memory), shared by
// the stage kernels of the first-order path and the second-order kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

// h*D applied at global row `row` to values v(g) = base[(g - lo) * stride].
// INTERIOR: the caller guarantees W <= row < n - W, so only the interior
// stencil is compiled (no branch: the loads of several stencils can then be
// issued together).
template <int SCHEME, bool INTERIOR = false>
__device__ __forceinline__ cdouble apply_d(const SchemeTable& T, const cdouble* base, int stride,
                                           int lo, int row, int n, double inv_h) {
  constexpr int W = SchemeShape<SCHEME>::W;
  constexpr int C = SchemeShape<SCHEME>::C;
  constexpr int B = SchemeShape<SCHEME>::B;
  double re = 0.0, im = 0.0;
  if (!INTERIOR && row < W) {
    const double* c = &T.left[row * kMaxClosureWidth];
    const cdouble* v = base + (0 - lo) * stride;
    // closure rows are rare: keep them compact to save registers
#pragma unroll 2
    for (int k = 0; k < C; ++k) {
      const cdouble x = v[k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else if (!INTERIOR && row + W >= n) {
    const double* c = &T.left[(n - 1 - row) * kMaxClosureWidth];
    const cdouble* v = base + (n - 1 - lo) * stride;
#pragma unroll 2
    for (int k = 0; k < C; ++k) {
      const cdouble x = v[-k * stride];
      re -= c[k] * x.re;
      im -= c[k] * x.im;
    }
  } else {
    const cdouble* v = base + (row - lo) * stride;
#pragma unroll
    for (int k = B; k >= 1; --k) {
      const cdouble a = v[k * stride];
      const cdouble b = v[-k * stride];
      const double d = SchemeShape<SCHEME>::interior(k);
      re += d * (a.re - b.re);
      im += d * (a.im - b.im);
    }
  }
  return cmake(inv_h * re, inv_h * im);
}

// -(eps/h) H^-1 A^T A at global row `row` (radial_discretization.hpp:
// 165-232), with the A^T A rows tabulated on the host.
template <int SCHEME>
__device__ __forceinline__ cdouble apply_diss(const SchemeTable& T, const cdouble* base,
                                              int stride, int lo, int row, int n, double h,
                                              double eps) {
  constexpr int O = SchemeShape<SCHEME>::O;
  constexpr int W = SchemeShape<SCHEME>::W;
  double re = 0.0, im = 0.0;
  if (row < O) {
    const double* c = T.diss_left[row];
    const cdouble* v = base + (0 - lo) * stride;
    for (int k = 0; k <= row + O; ++k) {
      const cdouble x = v[k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else if (row > n - 1 - O) {
    const int d = n - 1 - row;
    const double* c = T.diss_left[d];
    const cdouble* v = base + (n - 1 - lo) * stride;
    for (int k = 0; k <= d + O; ++k) {
      const cdouble x = v[-k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else {
    const cdouble* v = base + (row - lo) * stride;
#pragma unroll
    for (int k = -O; k <= O; ++k) {
      const cdouble x = v[k * stride];
      re += T.diss_interior[k + O] * x.re;
      im += T.diss_interior[k + O] * x.im;
    }
  }
  const int dist = row < n - 1 - row ? row : n - 1 - row;
  const double hw = dist < W ? T.norm[dist] : 1.0;
  const double f = -eps / (h * hw);
  return cmake(f * re, f * im);
}

}  // namespace teuk
