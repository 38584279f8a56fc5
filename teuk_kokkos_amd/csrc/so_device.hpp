// Device helpers of the second-order graph shared by its kernels and the
// diagnostics: complex utilities, the Kerr background, Delta_n, eth / eth'
// and thorn point formulas.
#pragma once

#include <hip/hip_runtime.h>

#include "so_layout.hpp"

namespace teuk {

__device__ __forceinline__ cdouble cconj(cdouble a) { return {a.re, -a.im}; }
__device__ __forceinline__ cdouble cneg(cdouble a) { return {-a.re, -a.im}; }
__device__ __forceinline__ cdouble cmulr(cdouble a, double s) { return {a.re * s, a.im * s}; }

// Complex quotient by Smith's algorithm (Smith 1962, CACM 5:435).
// The branch on the larger component of b is taken with selects on swapped
// operands, so the quotient is branch free (no divergence, and loads can be
// scheduled across it); each branch's terms round exactly as written out:
//   |b.re| >= |b.im|: r = b.im/b.re, d = b.re + b.im r,
//                     ((a.re + a.im r) / d, (a.im - a.re r) / d)
//   otherwise:        r = b.re/b.im, d = b.re r + b.im,
//                     ((a.re r + a.im) / d, (a.im r - a.re) / d)
// (the last term as -((a.re - a.im r) / d), equal under round-to-nearest).
__device__ __forceinline__ cdouble cdiv(cdouble a, cdouble b) {
  const bool big = fabs(b.re) >= fabs(b.im);
  const double p = big ? b.re : b.im, q = big ? b.im : b.re;
  const double x = big ? a.re : a.im, y = big ? a.im : a.re;
  const double r = q / p, d = p + q * r;
  const double e = (x + y * r) / d, f = (y - x * r) / d;
  return {e, big ? f : -f};
}

// Rescaled rotated-Kinnersley background (background.hpp:29-78).
using Bg = SoBackground;

__device__ inline Bg kerr_bg(double M, double a, double L, double R, double ct, double st) {
  const double L2 = L * L, L4 = L2 * L2;
  const double sqrt_two = 1.4142135623730951;
  const cdouble minus_radial = cmake(-L2, a * R * ct);
  const cdouble angular = cmake(L2, -a * R * ct);
  const cdouble angular2 = cmul(angular, angular);
  const double real_pi = L4 + a * a * R * R * ct * ct;
  Bg b;
  b.mu0 = cdiv(cmake(1.0, 0.0), minus_radial);
  // I a st / (sqrt2 angular2)
  b.tau0 = cdiv(cmake(0.0, a * st), cscale(sqrt_two, angular2));
  b.pi0 = cmake(0.0, -(a * st) / (sqrt_two * real_pi));
  const cdouble ingoing = cmake(L2, a * R * ct);
  const cdouble optical = cmul(angular2, ingoing);
  b.rho0 = cdiv(cmake(-0.5 * (L4 - 2.0 * L2 * M * R + a * a * R * R), 0.0), optical);
  b.eps0 = cdiv(cscale(0.5, cmake(L2 * M - a * a * R, -a * (L2 - M * R) * ct)), optical);
  b.psi20 = cdiv(cmake(-M, 0.0), cmul(angular2, angular));
  return b;
}

// Delta_n (ghp.hpp:11-22) and its inverse for the time derivative
// (reconstruction.hpp:108-121).
__device__ __forceinline__ cdouble delta_n(cdouble v, cdouble dt, cdouble dr, int n, double R,
                                           double M, double L) {
  const double L2 = L * L;
  const double tc = 2.0 + 4.0 * M * R / L2;
  const double rc = R / L2;
  return cadd(cscale(tc, dt), cscale(rc, cadd(cscale(R, dr), cscale(static_cast<double>(n), v))));
}

__device__ __forceinline__ cdouble recon_dt(cdouble v, cdouble dr, cdouble rhs, int n, double R,
                                            double M, double L) {
  const double L2 = L * L;
  const double den = 2.0 + 4.0 * M * R / L2;
  const cdouble radial =
      cadd(cscale(R * R / L2, dr), cscale(static_cast<double>(n) * R / L2, v));
  return cdivr(csub(rhs, radial), den);
}

// eth_n / eth'_n point formulas (ghp.hpp:60-116) given the angular part.
__device__ __forceinline__ cdouble eth_point(bool prime, cdouble v, cdouble dt, cdouble ang,
                                             int spin, int boost, double R, double st,
                                             double ct, double a, double L) {
  const double sqrt_two = 1.4142135623730951;
  const double L2 = L * L;
  if (!prime) {
    const cdouble den = cmake(L2, -a * R * ct);
    const cdouble num = cadd(cmul(cmake(0.0, -a * st), dt), ang);
    const cdouble deriv = cdiv(num, cscale(sqrt_two, den));
    const cdouble conn = cdiv(cmul(cmake(0.0, -static_cast<double>(spin + boost) * a * R * st), v),
                              cscale(sqrt_two, cmul(den, den)));
    return cadd(deriv, conn);
  }
  const cdouble den = cmake(L2, a * R * ct);
  const cdouble num = cadd(cmul(cmake(0.0, a * st), dt), ang);
  const cdouble deriv = cdiv(num, cscale(sqrt_two, den));
  const cdouble conn = cdiv(cmul(cmake(0.0, static_cast<double>(-spin + boost) * a * R * st), v),
                            cscale(sqrt_two, cmul(den, den)));
  return cadd(deriv, conn);
}

// thorn_n (ghp.hpp:24-58).
__device__ __forceinline__ cdouble thorn_n(cdouble v, cdouble dt, cdouble dr, int n, int sw,
                                           int bw, int m, double R, double ct, double M,
                                           double a, double L, cdouble eps0) {
  const double L2 = L * L, L4 = L2 * L2;
  const double den = L4 + a * a * R * R * ct * ct;
  const double tc = R * 2.0 * M * (2.0 * M - a * a * R / L2) / den;
  const double rc = -0.5 * (L2 - 2.0 * M * R + a * a * R * R / L2) / den;
  const int pw = sw + bw, qw = -sw + bw;
  cdouble r = cadd(cscale(tc, dt), cscale(rc, cadd(cscale(R, dr), cscale(static_cast<double>(n), v))));
  r = cadd(r, cdivr(cmul(cmake(0.0, R * a * static_cast<double>(m)), v), den));
  const cdouble w = cadd(cscale(static_cast<double>(pw), eps0), cscale(static_cast<double>(qw), cconj(eps0)));
  return csub(r, cscale(R, cmul(w, v)));
}

}  // namespace teuk
