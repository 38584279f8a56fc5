// Descriptors of the second-order path (plans whose instances enable the
// quadratic source): the reference's 13-field nodal state
// (mode, field, radial, theta) evolved with its full stage graph
// (spatial_pipeline.hpp:458-588).
//
// HBM layout, per instance, all complex128:
//   state-like buffers (y, acc, Ya, Yb, k):  (M, 13, N, T) at SoInstance::base
//   scratch arrays with S slots:             (M, S, N, T) at pbase * S
// where pbase counts M*N*T values of the instances before this one.
#pragma once

#include "device_layout.hpp"

namespace teuk {

inline constexpr int kSoMaxModes = 16;
// Dynamic LDS a workgroup can take on gfx950 (160 KiB per CU).
inline constexpr int kMaxLdsBytes = 160 * 1024;
inline constexpr int kSoMaxTargets = 16;
inline constexpr int kSoMaxPairs = 64;
inline constexpr int kSoThreads = 256;
// largest output points (rows x N_theta) of a linear-triple tile per thread
inline constexpr int kSoLinearOut = 4;
// radial rows per source-projection workgroup
inline constexpr int kProjRows = 4;
// radial D workgroups: kDradRows output rows x kDradCols theta columns, with
// the stencil reach (at most one closure width each side) staged in LDS
inline constexpr int kDradRows = 64;
inline constexpr int kDradCols = 16;
inline constexpr int kDradSpan = kDradRows + 2 * kMaxClosureWidth;
inline int so_drad_blocks(int M, int N, int T) {
  return M * 7 * ((N + kDradRows - 1) / kDradRows) * ((T + kDradCols - 1) / kDradCols);
}
// Reconstruction workgroups: kReconThreads threads holding
// kReconThreads / (P * N_theta) radial rows when P * N_theta <= kReconThreads (P:
// the parent modes, the only ones with reconstruction fields),
// else one row of up to kReconMaxThreads threads.
inline constexpr int kReconThreads = 256;
inline constexpr int kReconMaxThreads = 1024;  // largest P * N_theta a plan accepts
inline constexpr int kReconLdsRows = 12;       // theta rows of P * N_theta values per radial row
inline int recon_rows_per_group(int mt) { return mt <= kReconThreads ? kReconThreads / mt : 1; }

// Angular coordinators of the stage graph (spatial_pipeline.hpp:185-218):
// (spin, boost, band) over the parent or target modes.
enum SoCoord {
  kCoFirst = 0,   // spin -2, boost -2, ell_max_first, parents
  kCoSecond,      // spin -2, boost -2, ell_max_second, targets
  kCoG,           // spin -1, boost -1
  kCoB,           // spin -2, boost 0
  kCoPi,          // spin -1, boost 0
  kCoC,           // spin -1, boost 1
  kCoScalar,      // spin 0, boost 0
  kCoBSharp,      // spin 2, boost 0
  kCoCSharp,      // spin 1, boost 1
  kCoT,           // spin -1, boost -2, ell_max_second, targets
  kSoCoords
};
// Per (coordinator, mode) T x T nodal operators (row-major, out = Op * in):
// band projection S A, Laplacian S Lambda A, and the angular parts of eth and
// eth': S' diag(raising or lowering factor) A with the ell alignment of the
// spin-shifted band (ghp_device.hpp:174-196).  The band projections are
// applied as the reference applies them, analysis then synthesis
// (angular_coordinator.hpp:125-146): kOpAnalyze holds A (n_ell x T, its rows
// from n_ell on zero) and kOpSynthesize S (T x n_ell, its columns from n_ell
// on zero), so a projected row is a synthesis of coefficients and stays in
// the band to rounding relative to itself, even where it is a small
// remainder of a much larger input.  kOpProject, the same projection as one
// matrix S A, only names the projection in operator lists.
enum SoOp { kOpProject = 0, kOpLaplacian, kOpRaise, kOpLower, kOpAnalyze, kOpSynthesize, kSoOps };

struct SoInstance {
  long long base;   // offset of the (M, 13, N, T) block in the state buffers
  long long pbase;  // M * N * T summed over the instances before this one
  int M, N, T, scheme;
  int constrained, dissipate;
  int n_targets;
  int node_off;     // x | sqrt(1 - x^2) | cos(acos x) | sin(acos x) | weights, T each
  long long bgoff;  // N * T summed over the instances before this one
  double mass, spin, length, h, inv_h, gamma, eps;
  int modes[kSoMaxModes];
  int sharp[kSoMaxModes];
  // bit f: field f evolves on the mode (fields 0..9 of parent modes, 10..12
  // of target modes); the others are zero for good, see so_rk_kernel
  int field_mask[kSoMaxModes];
  // the parent (first-order) modes, whose reconstruction fields evolve: mode
  // index of each, and the parent index of its sharp partner (-1 when -m is
  // not a parent, whose fields 0..9 are then zero)
  int n_parents;
  int parent_mode[kSoMaxModes];
  int parent_sharp[kSoMaxModes];
  int target_mode[kSoMaxTargets];
  int pair_begin[kSoMaxTargets + 1];
  int pair_m1[kSoMaxPairs], pair_m2[kSoMaxPairs];
  int coord_spin[kSoCoords], coord_boost[kSoCoords];
  int op[kSoCoords][kSoMaxModes][kSoOps];  // double offset, -1 where inactive
  int band_n[kSoCoords][kSoMaxModes];      // n_ell of the coordinator's band on the mode
};

// One radial tile of one (instance, mode) for the linear triple kernel:
// output rows [t0, t1), psi_t rows [mlo, mhi), staged rows [lo, hi).
struct SoTile {
  int inst, mode;
  int t0, t1, lo, hi, mlo, mhi;
};

// Scratch slot counts (per instance, per mode).
enum {
  kSlotsTan = 10,    // tangent RHS of fields 0..9 (first-order triple, reconstruction)
  kSlotsRdr = 7,     // D of the seven reconstruction fields
  kSlotsRang = 4,    // eth2 C, eth'2 C#, eth2 G, eth'1 B# of the stage chain
  kSlotsSrc = 4,     // projected D, projected D_t, projected T, eth'3 T
  kSlotsEx = 4,      // eth1 B, eth'1 F and their tangents
  kSlotsRaw = 4,     // ordered-pair sums D, D_t, T, T_t of the target modes
  kSlotsCons = 3,    // constraint residuals
  kSlotsTerms = 6,   // constraint left/right terms
  kSlotsRangX = 2    // eth1 F, eth2 Pi of the stage chain (constraint passes only)
};
enum { kRangEth2C = 0, kRangEthP2CSharp = 1, kRangEth2G = 2, kRangEthP1BSharp = 3 };

// Kerr background coefficients at one (r, theta) point, rescaled as in the
// reference: mu = R mu0, tau = R^2 tau0, pi = R^2 pi0, Psi2 = R^3 psi20.
struct SoBackground {
  cdouble mu0, tau0, pi0, psi20, rho0, eps0;
};

struct SoArgs {
  const SoInstance* inst;
  SoBackground* bg;           // per instance: storage table then theta table, N * T each
  const double* ops;
  const double* nodes;
  const SchemeTable* tables;  // D4-2, D8-4, D10-5
  const SoTile* tiles;
  const int* active;          // per instance: forcing applied
  const cdouble* in;          // stage, (M, 13, N, T)
  cdouble* out;               // RHS, (M, 13, N, T)
  cdouble* tan;
  cdouble* dpsi;              // D Psi of the stage
  cdouble* dpsi_t;            // D Psi of the RHS
  cdouble* rdr;               // D of the stage reconstruction fields
  cdouble* rdr_t;             // D of the RHS reconstruction fields
  cdouble* rang;
  cdouble* rang_t;            // the same four inputs of the tangent chain
  cdouble* ex;
  cdouble* raw;
  cdouble* src;
  cdouble* cons;
  cdouble* terms;
  cdouble* rangx;             // written with the constraints
  cdouble* pairs;             // single-instance evaluations only, else null: the
                              // value of D and T of every ordered pair, (pair, 2, N, T)
  int want_constraints;
  int first_inst;             // instances [first_inst, first_inst + n_inst)
  int n_inst;
};

// RK4 stage combination over whole state blocks (stage 1..4).
struct SoRkArgs {
  const SoInstance* inst;
  const double* dt;           // per instance
  const cdouble* k;
  const cdouble* y;
  cdouble* acc;
  cdouble* yout;
  long long max_values;       // largest instance block (M * 13 * N * T)
  int stage;
  int n_inst;
};

// Deterministic constraint reduction: per instance 18 sums and maxima plus
// a finiteness flag (source_activation.hpp:120-170).
struct SoReduceArgs {
  const SoInstance* inst;
  const double* nodes;        // Gauss-Legendre weights at node_off + 4 T
  const SchemeTable* tables;
  const cdouble* cons;
  const cdouble* terms;
  double* out;                // per instance: kSoReduceDoubles
  int n_inst;
};
inline constexpr int kSoReduceDoubles = 19;

hipError_t so_launch_linear(int scheme, int kind, const SoArgs& a, int n_tiles, int lds,
                            hipStream_t s);
// `blocks`: largest so_drad_blocks() of the plan's instances; `schemes`: bit
// k set when some instance uses radial scheme k
hipError_t so_launch_drad(int which, const SoArgs& a, int blocks, unsigned schemes,
                          hipStream_t s);
// workgroups of recon_rows_per_group(P * N_theta) radial rows of one
// instance: `groups` = the largest count over the plan's instances,
// max_mt = the largest P * N_theta
hipError_t so_launch_recon(int kind, const SoArgs& a, int groups, int max_mt, int lds,
                           hipStream_t s);
hipError_t so_launch_source(const SoArgs& a, long long max_points, int max_targets,
                            hipStream_t s);
hipError_t so_launch_source_projection(const SoArgs& a, int rows, int lds, hipStream_t s);
hipError_t so_launch_rk(const SoRkArgs& a, hipStream_t s);
hipError_t so_launch_reduce(const SoReduceArgs& a, hipStream_t s);
hipError_t so_set_lds_limit(int bytes);
hipError_t so_launch_background(const SoArgs& a, long long max_points, hipStream_t s);

}  // namespace teuk
