// Host-side numerics used at plan set-up: Gauss-Legendre nodes, spin-
// weighted harmonics and fixed-(s,m) transform matrices, SBP operator
// tables, the dense complex solve of the Galerkin initialiser, and the
// signed-mode registry.
#pragma once

#include <complex>
#include <cstddef>
#include <utility>
#include <vector>

#include "config.hpp"

namespace teuk {

using cplx = std::complex<double>;

// ------------------------------------------------------------ angular
// Gauss-Legendre nodes x_j = cos(theta_j), increasing, and weights
// (angular.hpp:69-115: long-double Newton on P_n).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
// Theta factor of sY_lm with the Condon-Shortley Wigner-d
// (angular.hpp:170-217).
double swsh_theta(int ell, int m, int spin, double theta);
// Teukolsky angular eigenvalue -(l-s)(l+s+1) (angular.hpp:139-165).
inline double laplacian_eigenvalue(int ell, int spin) {
  return -static_cast<double>((ell - spin) * (ell + spin + 1));
}
inline int minimum_ell(int spin, int m) {
  const int as = spin < 0 ? -spin : spin, am = m < 0 ? -m : m;
  return as > am ? as : am;
}

// Fixed-(s,m) band: synthesis S (ntheta x n_ell) and analysis
// A (n_ell x ntheta, A = 2 pi w S^T) (angular.hpp:329-355).
// G = A S equals the identity only to the accuracy of the long-double
// Wigner-d sums: ~1e-14 up to ell ~ 16, 6e-12 at ell = 24 and 2e-9 at
// ell = 32.  The nodal projection S A the reference applies every stage is
// reproduced in modal coordinates with G (`gram` set when |G - I| > 1e-13);
// B = G^-1 A maps a nodal vector S c back to exactly c.
struct BandTransform {
  int spin = 0, m = 0, ell_min = 0, ell_max = 0, nodes = 0;
  std::vector<double> S, A, G, B;
  bool gram = false;
  int modes() const { return ell_max - ell_min + 1; }
  std::vector<cplx> analyze(const cplx* nodal) const;         // A x
  std::vector<cplx> synthesize(const std::vector<cplx>& modal) const;  // S c
  std::vector<cplx> coefficients_of(const cplx* nodal) const;  // c with S c = x on range(S)
  std::vector<cplx> gram_apply(const std::vector<cplx>& c) const;  // G c
};
BandTransform make_band_transform(int spin, int m, int ell_max, int nodes);
bool band_gram_needed(int spin, int m, int ell_max, int nodes);

// ------------------------------------------------------------ SBP
// First-derivative closure description of one scheme.
struct SbpScheme {
  int closure_rows;      // W: boundary rows with a closure stencil
  int closure_width;     // columns used by a closure row
  int half_width;        // interior stencil half width
  int diss_order;        // order of the undivided difference in A^T A
  std::vector<double> left;      // W x closure_width, h*D entries
  std::vector<double> interior;  // half_width+1 entries, c[0]=0, antisymmetric
  std::vector<double> norm;      // W boundary norm weights H/h
  // Dissipation A^T A rows (without the -eps/(h H) factor):
  // diss_left[d] for distance d < diss_order covers columns [0, d+order];
  // diss_interior covers offsets [-order, order].
  std::vector<std::vector<double>> diss_left;
  std::vector<double> diss_interior;
};
const SbpScheme& sbp_scheme(RadialScheme s);
double sbp_derivative_entry(RadialScheme s, int n, int row, int col);
double sbp_norm_weight(RadialScheme s, int n, int index);
// Column range [lo, hi) read by the derivative at `row` of an n-point grid.
std::pair<int, int> derivative_columns(RadialScheme s, int n, int row);
std::pair<int, int> dissipation_columns(RadialScheme s, int n, int row);
// Host derivative of a contiguous complex line (initial data).
void sbp_derivative_line(RadialScheme s, int n, double inv_h, const cplx* in, size_t stride,
                         cplx* out, size_t out_stride);

// ------------------------------------------------------------ dense solve
// Partial-pivot Gaussian elimination; returns false for a singular matrix
// (dense_solve.hpp:20-72).
bool dense_complex_solve(int n, std::vector<cplx> matrix, std::vector<cplx> rhs,
                         std::vector<cplx>& solution);
double dense_complex_residual(int n, const std::vector<cplx>& matrix,
                              const std::vector<cplx>& solution,
                              const std::vector<cplx>& rhs);

// ------------------------------------------------------------ modes
// Signed-mode bookkeeping (modes.hpp:32-213): sorted stored modes, parents
// (first order) and targets (second order), sharp index = index(-m), and
// ordered pairs target-major then m1 ascending.
struct ModePair {
  int m1, m2, target;
  int index1, index2, target_index;
};
class ModeRegistry {
 public:
  ModeRegistry(std::vector<int> modes, std::vector<int> parents, std::vector<int> targets);
  const std::vector<int>& modes() const { return modes_; }
  const std::vector<int>& parents() const { return parents_; }
  const std::vector<int>& targets() const { return targets_; }
  const std::vector<ModePair>& ordered_pairs() const { return pairs_; }
  int size() const { return static_cast<int>(modes_.size()); }
  bool contains(int m) const;
  bool is_parent(int m) const;
  bool is_target(int m) const;
  int index(int m) const;
  int sharp_index(int m) const { return index(-m); }
  bool closed_under_sharp() const;
  std::pair<int, int> pair_range(int target) const;

 private:
  std::vector<int> modes_, parents_, targets_;
  std::vector<ModePair> pairs_;
  std::vector<int> pair_offsets_;
};
ModeRegistry make_run_registry(const RunParameters& p);  // initial_data_factory.hpp:16-24

}  // namespace teuk
