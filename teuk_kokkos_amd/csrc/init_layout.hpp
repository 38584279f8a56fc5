// Descriptors of the GPU Gaussian initial data with the Galerkin P
// (pipeline_initial_data.hpp:142-340).
#pragma once

#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

// Per line: the seed's modal vector and the radial profile.
struct InitLine {
  int seed_off;   // cdouble offset of the n_ell seed coefficients
  int compact;    // compact-support bump instead of a Gaussian
  int scheme;     // radial scheme of Q = D Psi
  double center, width;
  long long nodal_base;  // nodal output: cdouble offset of the (mode, field 0) block
};

struct InitArgs {
  const LineDesc* lines;
  const InitLine* init;
  const double* mats;
  const double* nodes;
  const SchemeTable* tables;
  const cdouble* seeds;
  cdouble* y;                        // modal state
  // When set, the band-projected fields S A P, S A Q, S A Psi are written
  // instead as nodal rows (field, r, theta) at nodal + nodal_base, and y is
  // not touched (second-order nodal state).
  cdouble* nodal;
  unsigned long long* residual;      // per line: max relative residual (double bits)
  int n_lines;
};

inline constexpr int kInitThreads = 64;
inline constexpr int kInitMaxEll = kMaxBandEll;

int galerkin_init_lds_bytes(int n_ell, int ntheta);
hipError_t launch_galerkin_init(const InitArgs& a, int max_nr, int lds, hipStream_t s);

}  // namespace teuk
