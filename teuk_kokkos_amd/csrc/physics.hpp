// Point physics shared by the host setup code and the HIP stage kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace teuk {

// Complex double as a 16-byte aligned pair so every global or LDS access
// is one 128-bit instruction (an 8-byte aligned pair becomes a
// ds_read2_b64, which moves LDS data at about half the rate of ds_read_b128
// on gfx950).
struct alignas(16) cdouble {
  double re, im;
};

__host__ __device__ inline cdouble cmake(double re, double im) { return {re, im}; }
__host__ __device__ inline cdouble cadd(cdouble a, cdouble b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cdouble csub(cdouble a, cdouble b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cdouble cscale(double s, cdouble a) { return {s * a.re, s * a.im}; }
__host__ __device__ inline cdouble cmul(cdouble a, cdouble b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cdouble cdivr(cdouble a, double d) { return {a.re / d, a.im / d}; }

// Stationary Teukolsky coefficients of the compact first-order reduction
// (P, Q, Psi) at one (R, theta) point, written with the same operation order
// as include/teuk/teukolsky.hpp:38-85 of the reference:
//   C_T = 8M(2M - a^2 R/L^2)(1 + 2MR/L^2) - a^2 sin^2(theta)
//   K   = L^2 - (8M^2 - a^2) R^2/L^2 + 4 a^2 M R^3/L^4
//   H_R = R^2/L^4 (L^4 - 2 L^2 M R + a^2 R^2)
//   G_m = 2 i a m (1 + 4MR/L^2)
//         + 2 (2M(-s + (2+s) 2MR/L^2 - 3a^2R^2/L^4) - a^2 R/L^2 + i s a cos)
//   q   = 2R(1 + s - (3+s) MR/L^2 + 2 a^2 R^2/L^4) - 2 i a m R^2/L^2
//   psi = -2R((1+s) M/L^2 - a^2 R/L^4) - 2 i a m R/L^2
// sin_theta and cos_theta are sin(acos(x_j)) and cos(acos(x_j)) evaluated on
// the host, as the reference evaluates them from theta = acos(x_j).
struct RadialCoefficients {
  double adv;        // K
  double principal;  // H_R
  double time_radial;  // 8M(2M - a^2 R/L^2)(1 + 2MR/L^2)
  double g_real;     // real part of G_m (theta independent)
  double g_imag_m;   // 2 a m (1 + 4MR/L^2) (theta independent part of Im G_m)
  cdouble q, psi;
};

__host__ __device__ inline RadialCoefficients radial_coefficients(
    double M, double a, double L, int spin_weight, int m, double R) {
  const double L2 = L * L;
  const double L4 = L2 * L2;
  const double R2 = R * R;
  const double a2 = a * a;
  const double mrl = M * R / L2;
  const double s = static_cast<double>(spin_weight);
  const double md = static_cast<double>(m);
  RadialCoefficients c;
  c.time_radial = 8.0 * M * (2.0 * M - a2 * R / L2) * (1.0 + 2.0 * mrl);
  c.adv = L2 - (8.0 * M * M - a2) * R2 / L2 + 4.0 * a2 * M * R * R2 / L4;
  c.principal = R2 / L4 * (L4 - 2.0 * L2 * M * R + a2 * R2);
  c.g_real = 2.0 * (2.0 * M * (-s + (2.0 + s) * 2.0 * mrl - 3.0 * a2 * R2 / L4) -
                    a2 * R / L2);
  c.g_imag_m = 2.0 * a * md * (1.0 + 4.0 * mrl);
  c.q = cmake(2.0 * R * (1.0 + s - (3.0 + s) * mrl + 2.0 * a2 * R2 / L4),
              -(2.0 * a * md * R2 / L2));
  c.psi = cmake(-2.0 * R * ((1.0 + s) * M / L2 - a2 * R / L4),
                -(2.0 * a * md * R / L2));
  return c;
}

// Theta-dependent completion: C_T and G_m at one node.
__host__ __device__ inline double time_coefficient(const RadialCoefficients& c,
                                                   double a, double sin_theta) {
  return c.time_radial - a * a * sin_theta * sin_theta;
}

__host__ __device__ inline cdouble definition_coefficient(const RadialCoefficients& c,
                                                          double a, int spin_weight,
                                                          double cos_theta) {
  const double s = static_cast<double>(spin_weight);
  return cmake(c.g_real, c.g_imag_m + 2.0 * (s * a * cos_theta));
}

}  // namespace teuk
