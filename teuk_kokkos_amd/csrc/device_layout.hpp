// Device-side descriptors shared by the plan (host) and the stage kernels.
//
// Evolved state layout in HBM (per buffer y, C, Ya, Yb):
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
  int gram_bf16;    // the correction of the Gram projection on BF16 MFMA (below)
  // double offset of S (ntheta x n_ell), then A (n_ell x ntheta),
  // G = A S (n_ell x n_ell) and B = G^-1 A (n_ell x ntheta)
  int mat_off;
  int node_off;     // double offset of sin(theta_j)[ntheta] then cos(theta_j)[ntheta]
  int frag_off;     // double offset of the MFMA operand fragments of S and A
  int gfrag_off;    // double offset of the FP32 MFMA fragments of G - I and M - Lambda
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

// The scheme table's size in LDS, padded to 16 bytes.
inline constexpr int kSchemeTableDoubles = (sizeof(SchemeTable) + 15) / 16 * 2;
static_assert(sizeof(SchemeTable) % 8 == 0, "SchemeTable is copied as doubles");

// FP64 MFMA (v_mfma_f64_16x16x4_f64) tiling of the Kerr angular coupling.
// Operand layouts of one wave (lane = 0..63, lr = lane % 16, lg = lane / 16),
// measured on gfx950:
//   A operand (16 x 4): lane holds A[lr][lg];  B operand (4 x 16): B[lg][lr];
//   accumulator (16 x 16): element v of lane holds D[lg + 4 v][lr].
// Synthesis (S X^T)[theta][row] uses theta tiles of 16 and ell blocks of 4:
//   S fragment (t, q), lane: S[16 t + lr][4 q + lg].
// The analysis sum over theta runs over blocks (t, v) = {16 t + 4 v + g :
// g = 0..3}, so the synthesis accumulator of a lane is already the A operand
// of the analysis (no data movement between the two products):
//   A fragment (t, v, u), lane: A[16 u + lr][16 t + 4 v + lg].
// Entries outside the band or the theta grid are zero.
__host__ __device__ inline int mfma_theta_tiles(int ntheta) { return (ntheta + 15) / 16; }
__host__ __device__ inline int mfma_ell_blocks(int n_ell) { return (n_ell + 3) / 4; }
__host__ __device__ inline int mfma_ell_tiles(int n_ell) { return (n_ell + 15) / 16; }
__host__ __device__ inline int mfma_s_fragments(int ntheta, int n_ell) {
  return mfma_theta_tiles(ntheta) * mfma_ell_blocks(n_ell);
}
__host__ __device__ inline int mfma_fragment_doubles(int ntheta, int n_ell) {
  return 64 * mfma_theta_tiles(ntheta) * (mfma_ell_blocks(n_ell) + 4 * mfma_ell_tiles(n_ell));
}
// The Gram correction operands of k = G X + M Psi split as G = I + E,
// M = G Lambda G = Lambda + N: E, then N, as FP32 B operands of
// v_mfma_f32_16x16x4_f32, one fragment per (output ell tile uo, contraction
// block kb), lane: E[16 uo + lr][4 kb + lg]; gram_correction_floats floats
// each, gram_fragment_doubles doubles together.  v_mfma_f32_16x16x4_f32 takes
// its operands as the FP64 instruction does, but its accumulator layout
// differs (measured on gfx950): element v of lane holds D[4 lg + v][lr].
__host__ __device__ inline int gram_correction_floats(int n_ell) {
  return 64 * mfma_ell_tiles(n_ell) * mfma_ell_blocks(n_ell);
}
__host__ __device__ inline int gram_fragment_doubles(int n_ell) {
  return gram_correction_floats(n_ell);  // two float matrices
}
// The same two matrices as BF16 operands of v_mfma_f32_16x16x32_bf16 for
// bands of at most 32 ell values (one contraction block covers the band),
// stored after the FP32 fragments: fragment (which = E or N, uo), lane holds
// the eight values E[16 uo + lr][8 lg + j], j = 0..7 (lane layout measured on
// gfx950: A operand A[lr][8 lg + j], B operand B[8 lg + j][lr], accumulator
// element v D[4 lg + v][lr]).  Lines whose E and N are small enough take
// them (LineDesc::gram_bf16).
inline constexpr int kGramBf16MaxEll = 32;
__host__ __device__ inline int gram_bf16_doubles(int n_ell) {
  return n_ell <= kGramBf16MaxEll ? 2 * mfma_ell_tiles(n_ell) * 64 * 2 : 0;
}
// Widest band the MFMA routines hold (4 ell tiles of 16).
inline constexpr int kMaxBandEll = 64;
// ell tiles of 16 a kernel instantiation holds at once: 1, 2 or 4.
__host__ __device__ inline int band_tiles_class(int n_ell) {
  const int lt = mfma_ell_tiles(n_ell);
  return lt <= 1 ? 1 : (lt <= 2 ? 2 : 4);
}

__host__ __device__ inline int gram_doubles(int n_ell) {
  return (n_ell * n_ell + 1) / 2 * 2;  // padded to 16 bytes
}

// LDS of one tile: per-row coefficients, the staged rows, psi_t / D Psi
// buffers where they cannot overwrite P in place, and the node table of the
// Kerr coupling.  The MFMA operand fragments are read from global memory.
__host__ __device__ inline int tile_lds_bytes(int rows, int mrows, int n_ell, int ntheta,
                                              int constrained, int dissipate, int kerr) {
  long long b = 8LL * kSchemeTableDoubles + 80LL * mrows + 16LL * rows * 3 * n_ell;
  if (dissipate) b += 16LL * mrows * n_ell;
  if (constrained) b += 16LL * mrows * n_ell;
  if (kerr) b += 16LL * ntheta;  // sin and cos of the nodes
  return static_cast<int>(b);
}

// Compile-time shape of radial scheme 0 (D4-2), 1 (D8-4), 2 (D10-5): W
// closure rows per boundary, C closure columns, B interior half-width, O the
// dissipation operator's half-width.
template <int SCHEME>
struct SchemeShape {
  static constexpr int W = SCHEME == 2 ? 11 : (SCHEME == 1 ? 8 : 4);
  static constexpr int C = SCHEME == 2 ? 16 : (SCHEME == 1 ? 12 : 6);
  static constexpr int B = SCHEME == 2 ? 5 : (SCHEME == 1 ? 4 : 2);
  static constexpr int O = SCHEME == 2 ? 6 : (SCHEME == 1 ? 5 : 3);
  // Interior antisymmetric stencil d_k (k = 1..B) of h*D: the same rationals,
  // rounded the same way, as the host table (host_math.cpp), so the kernels
  // take them as immediates instead of LDS reads.
  static constexpr double interior(int k) {
    if constexpr (SCHEME == 2) {
      constexpr double d[6] = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0,
                               1.0 / 1260.0};
      return d[k];
    } else if constexpr (SCHEME == 1) {
      constexpr double d[5] = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};
      return d[k];
    } else {
      constexpr double d[3] = {0.0, 2.0 / 3.0, -1.0 / 12.0};
      return d[k];
    }
  }
};

// Stage-kernel variants: line flags fixed at compile time, or read from the
// line descriptor (kStageGeneral: any mix of Schwarzschild / Kerr, Gram,
// dissipation and the stage constraint).
inline constexpr int kStageGeneral = 0;
inline constexpr int kStageKerr = 1;      // Kerr, G = I, no dissipation, free-damped
inline constexpr int kStageKerrGram = 2;  // Kerr, Gram projection, no dissipation, free-damped

__host__ __device__ inline int stage_variant(const LineDesc& L) {
  if (!L.kerr || L.dissipate || L.constrained) return kStageGeneral;
  return L.gram ? kStageKerrGram : kStageKerr;
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
  cdouble* acc;         // C = y + dt/6 (k1 + 2 k2) - y/3 (written by stage 2)
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
