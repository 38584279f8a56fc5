// Device-side descriptors shared by the plan (host) and the stage kernels.
//
// Evolved state layout in HBM (per buffer y, acc, Ya, Yb):
//   for every radial line (instance, active first-order mode m):
//     [radial r = 0..nr-1][field f in {P, Q, Psi}][ell = ell_min..ell_max]
//   complex128 values, one contiguous block per line at LineDesc::base.
// A radial tile of a line is therefore one contiguous byte range, loaded by
// the whole workgroup with coalesced 16-byte accesses.
#pragma once

#include "physics.hpp"

namespace teuk {

inline constexpr int kStageThreads = 256;
inline constexpr int kMaxClosureRows = 11;
inline constexpr int kMaxClosureWidth = 16;
inline constexpr int kMaxDissOrder = 6;

struct LineDesc {
  long long base;   // cdouble offset of this line's block
  int nr, n_ell, ntheta, ell_min;
  int m, constrained, kerr, dissipate;
  int gram;         // apply G = A S (its deviation from I exceeds roundoff)
  int gram_pad;
  // double offset of S (ntheta x n_ell), then A (n_ell x ntheta),
  // G = A S (n_ell x n_ell) and B = G^-1 A (n_ell x ntheta)
  int mat_off;
  int node_off;     // double offset of sin(theta_j)[ntheta] then cos(theta_j)[ntheta]
  double mass, spin, length, h, inv_h, gamma, eps, dt;
};

struct TileDesc {
  int line;
  int t0, t1;    // output rows
  int lo, hi;    // loaded rows
  int mlo, mhi;  // rows of the intermediate psi_t (and D psi when constrained)
  int lds_bytes;
};

// One SBP scheme: closure rows of h*D, interior antisymmetric stencil,
// boundary norm weights and the rows of A^T A for the dissipation.
struct SchemeTable {
  int W, C, B, order;
  double left[kMaxClosureRows * kMaxClosureWidth];
  double interior[6];
  double norm[kMaxClosureRows];
  double diss_left[kMaxDissOrder][2 * kMaxDissOrder + 1];
  double diss_interior[2 * kMaxDissOrder + 1];
};

// Rows of the nodal psi_t chunk processed at once by a workgroup.
__host__ __device__ inline int nodal_chunk_rows(int ntheta) {
  const int r = 512 / ntheta;
  return r < 1 ? 1 : r;
}

__host__ __device__ inline int gram_doubles(int n_ell) {
  return (n_ell * n_ell + 1) / 2 * 2;  // padded to 16 bytes
}

__host__ __device__ inline int tile_lds_bytes(int rows, int mrows, int orows, int n_ell,
                                              int ntheta, int constrained, int kerr, int gram) {
  long long b = 80LL * mrows + 16LL * rows * 3 * n_ell + 16LL * mrows * n_ell;
  if (constrained) b += 16LL * mrows * n_ell;
  if (kerr) b += 16LL * ntheta * n_ell + 16LL * nodal_chunk_rows(ntheta) * ntheta;
  if (gram) b += 8LL * gram_doubles(n_ell) + 16LL * orows * 3 * n_ell;
  return static_cast<int>(b);
}

// Kernel arguments of one stage launch.
struct StageArgs {
  const TileDesc* tiles;
  const LineDesc* lines;
  const double* mats;
  const double* nodes;
  const SchemeTable* table;
  const cdouble* yin;   // RHS input (stage value)
  const cdouble* y;     // state at the start of the step
  cdouble* acc;         // y + dt/6 k1 + dt/3 k2 + dt/3 k3
  cdouble* yout;        // next stage value, or y after stage 4
  cdouble* kout;        // RHS output (stage 0 only)
  double dt_override;   // > 0: use instead of LineDesc::dt
};

// Nodal (mode, field=13, r, theta) <-> modal line conversion of one instance.
struct NodalArgs {
  const LineDesc* lines;  // first line of the instance
  const int* line_mode;   // stored-mode index of each line
  const double* mats;
  int n_lines;
  int nr, ntheta, modes;
};

}  // namespace teuk
