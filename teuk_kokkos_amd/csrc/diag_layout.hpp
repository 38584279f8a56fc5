// Descriptors of the sampled diagnostics (pipeline_diagnostics.hpp:94-429,
// pipeline_independent_reconstruction_diagnostics.hpp,
// pipeline_reconstruction_diagnostics.hpp:22-209).
#pragma once

#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

// Reduced quantities: |x|^2 summed and maximized over every (mode, r, theta)
// point of one instance.
enum DiagQuantity {
  kDqState = 0,         // 13 fields of the state
  kDqRhs = 13,          // 13 fields of its RHS
  kDqFirstConstraint = 26,   // Q - D Psi (first order)
  kDqSecondConstraint = 27,  // the same for the second-order triple
  kDqSource = 28,       // source_over_r3
  kDqForcing = 29,      // latched forcing
  kDqIndependent = 30,  // psi3 Bianchi, psi2 Bianchi, h_ll reality
  kDqTransport = 33,    // transport residuals of G, Lambda, H, B, Pi, C, U
  kDqCount = 40
};

// Non-finite categories (pipeline_diagnostics.hpp:57-68 plus the two
// reconstruction reports).
enum DiagFlag {
  kDfState = 1, kDfRhs = 2, kDfFirstConstraint = 4, kDfSecondConstraint = 8, kDfSource = 16,
  kDfForcing = 32, kDfScri = 64, kDfHorizon = 128, kDfIndependent = 256, kDfTransport = 512
};

// One instance's arrays.  The second-order intermediates of the RHS
// evaluation (per-instance blocks, (mode, slot, r, theta)) are null for
// linear plans, whose reconstruction chain never runs: their contributions
// are then zero, as the reference's zero-initialized views give.
struct DiagArgs {
  const cdouble* state;     // (M, 13, N, T)
  const cdouble* rhs;
  const cdouble* src;       // projected D, D_t, T, eth'3 T
  const cdouble* cons;      // independent constraint residuals
  const cdouble* rdr;       // D of the stage reconstruction fields
  const cdouble* rang;      // eth2 C, eth'2 C#, eth2 G, eth'1 B#
  const cdouble* rangx;     // eth1 F, eth2 Pi
  const int* sharp;         // M: index of the mode -m
  const double* nodes;      // x = cos theta, then sqrt(1 - x^2)
  const SchemeTable* table;
  double mass, spin, length, h, inv_h;
  int scheme, active, M, N, T;
  double* partial;          // per wave: kDqCount sums, then kDqCount maxima
  int* partial_flags;       // per wave
  double* out;              // kDqCount sums, kDqCount maxima, then the flag bits
  int n_waves;
};

// Horizon transverse derivatives (horizon_diagnostics.hpp:10-11): orders
// 0..4 from a 9-point backward stencil.
inline constexpr int kHorizonDerivatives = 5;
inline constexpr int kHorizonStencil = 9;

// Layout of a diagnostics sample (doubles); the C ABI's TEUK_DIAG_* values.
enum DiagOutput {
  kDiagStateRms = 0,            // 13 fields
  kDiagStateMax = 13,
  kDiagRhsRms = 26,
  kDiagRhsMax = 39,
  kDiagFirstConstraint = 52,    // (rms, max) pairs: first and second reduction
  kDiagSecondConstraint = 54,   // constraints, source_over_r3, forcing
  kDiagSource = 56,
  kDiagForcing = 58,
  kDiagIndependent = 60,        // (rms, max): psi3 Bianchi, psi2 Bianchi, h_ll reality, combined
  kDiagTransport = 68,          // (rms, max): G, Lambda, H, B, Pi, C, U, combined
  kDiagHorizonFirst = 84,       // |d^k FirstPsi / dR^k| maxima at R_H, k = 0..4
  kDiagHorizonSecond = 89,      // the same for SecondPsi
  kDiagSourceConstraintMax = 94,
  kDiagSourceActive = 95,
  kDiagActivationTime = 96,
  kDiagScriFinite = 97,
  kDiagHorizonFinite = 98,
  kDiagAllFinite = 99,
  kDiagCount = 100
};

inline constexpr int kDiagThreads = 256;
inline int diag_waves(int M, int N, int T) {
  const long long points = static_cast<long long>(M) * N * T;
  return static_cast<int>((points + kDiagThreads - 1) / kDiagThreads) * (kDiagThreads / 64);
}

hipError_t launch_diagnostics(const DiagArgs& a, hipStream_t s);

}  // namespace teuk
