// teuk_oracle.cpp -- CPU FP64 restatement of the reference's linear stage
// operator and RK4 stepper.  TEST INFRASTRUCTURE ONLY: the checker for the
// HIP path, never linked by the product.  See teuk_oracle.h for the scope.
//
// Citations are to the reference (HengruiZhu99/teuk-kokkos), include/teuk/.

#include "teuk_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "sbp_tables.inc"

namespace {

using cplx = std::complex<double>;
constexpr double kPi = std::numbers::pi_v<double>;
constexpr int kFields = 13;

thread_local std::string g_error;

double rat(const OracleRational& r) { return r.num / r.den; }

// ---------------------------------------------------------------- angular
// Gauss-Legendre nodes by long-double Newton iteration (angular.hpp:69-115).
void legendre_pair(int n, long double x, long double& pn, long double& pm1) {
  long double prev = 1.0L;
  if (n == 0) {
    pn = prev;
    pm1 = 0.0L;
    return;
  }
  long double cur = x;
  for (int l = 2; l <= n; ++l) {
    const long double next =
        ((2.0L * l - 1.0L) * x * cur - (l - 1.0L) * prev) / l;
    prev = cur;
    cur = next;
  }
  pn = cur;
  pm1 = prev;
}

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const long double tol = 8.0L * std::numeric_limits<long double>::epsilon();
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double root = std::cos(std::numbers::pi_v<long double> * (i + 0.75L) /
                                (n + 0.5L));
    long double pn, pm1, d;
    for (int it = 0; it < 100; ++it) {
      legendre_pair(n, root, pn, pm1);
      d = n * (root * pn - pm1) / (root * root - 1.0L);
      const long double upd = pn / d;
      root -= upd;
      if (std::abs(upd) <= tol * (1.0L + std::abs(root))) break;
      if (it == 99) throw std::runtime_error("Gauss-Legendre iteration failed");
    }
    legendre_pair(n, root, pn, pm1);
    d = n * (root * pn - pm1) / (root * root - 1.0L);
    const long double weight = 2.0L / ((1.0L - root * root) * d * d);
    x[i] = static_cast<double>(-root);
    x[n - 1 - i] = static_cast<double>(root);
    w[i] = static_cast<double>(weight);
    w[n - 1 - i] = static_cast<double>(weight);
  }
}

long double log_fact(int n) { return std::lgammal(static_cast<long double>(n) + 1.0L); }

long double ipow(long double b, int e) {
  long double r = 1.0L;
  while (e != 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

// Wigner small-d, Condon-Shortley, long-double factorial sum
// (angular.hpp:170-206).
double wigner_d(int l, int mp, int m, double theta) {
  const int kmin = std::max(0, m - mp);
  const int kmax = std::min(l + m, l - mp);
  const long double lognorm = 0.5L * (log_fact(l + m) + log_fact(l - m) +
                                      log_fact(l + mp) + log_fact(l - mp));
  const long double c = std::cos(0.5L * static_cast<long double>(theta));
  const long double s = std::sin(0.5L * static_cast<long double>(theta));
  long double v = 0.0L;
  for (int k = kmin; k <= kmax; ++k) {
    const long double lc = lognorm - log_fact(l + m - k) - log_fact(k) -
                           log_fact(mp - m + k) - log_fact(l - mp - k);
    long double term = std::exp(lc) * ipow(c, 2 * l + m - mp - 2 * k) *
                       ipow(s, mp - m + 2 * k);
    if (((mp - m + k) & 1) != 0) term = -term;
    v += term;
  }
  return static_cast<double>(v);
}

// Theta factor of sY_lm (angular.hpp:211-217).
double swsh(int l, int m, int s, double theta) {
  if (l < 0 || std::abs(s) > l || std::abs(m) > l)
    throw std::invalid_argument("spin-weighted mode requires ell >= |s|,|m|");
  const double phase = (std::abs(s) & 1) ? -1.0 : 1.0;
  return phase * std::sqrt((2.0 * l + 1.0) / (4.0 * kPi)) *
         wigner_d(l, m, -s, theta);
}

struct Transform {
  int spin = 0, m = 0, ell_min = 0, ell_max = 0, nodes = 0;
  std::vector<double> S;  // nodes x modes
  std::vector<double> A;  // modes x nodes
  int modes() const { return ell_max - ell_min + 1; }
};

// SpinWeightedTransform (angular.hpp:330-355).
Transform make_transform(int spin, int m, int ell_max, int nodes) {
  Transform t;
  t.spin = spin;
  t.m = m;
  t.ell_min = std::max(std::abs(spin), std::abs(m));
  t.ell_max = ell_max;
  t.nodes = nodes;
  if (ell_max < t.ell_min) throw std::invalid_argument("ell_max below ell_min");
  if (nodes < t.modes()) throw std::invalid_argument("too few nodes");
  std::vector<double> x, w;
  gauss_legendre(nodes, x, w);
  const int nm = t.modes();
  t.S.assign(static_cast<size_t>(nodes) * nm, 0.0);
  t.A.assign(static_cast<size_t>(nodes) * nm, 0.0);
  for (int j = 0; j < nodes; ++j) {
    const double theta = std::acos(x[j]);
    for (int k = 0; k < nm; ++k) {
      const double b = swsh(t.ell_min + k, m, spin, theta);
      t.S[j * nm + k] = b;
      t.A[k * nodes + j] = 2.0 * kPi * w[j] * b;
    }
  }
  return t;
}

// --------------------------------------------------------------- physics
struct Coeffs {
  double time, adv, principal;
  cplx definition, q, psi;
};

// teukolsky_coefficients (teukolsky.hpp:38-85).
Coeffs teukolsky_coefficients(double M, double a, double L, int sw, int mm,
                              double R, double theta) {
  const double L2 = L * L;
  const double L4 = L2 * L2;
  const double R2 = R * R;
  const double a2 = a * a;
  const double mrl = M * R / L2;
  const cplx I(0.0, 1.0);
  Coeffs c;
  c.time = 8.0 * M * (2.0 * M - a2 * R / L2) * (1.0 + 2.0 * mrl) -
           a2 * std::sin(theta) * std::sin(theta);
  c.adv = L2 - (8.0 * M * M - a2) * R2 / L2 + 4.0 * a2 * M * R * R2 / L4;
  c.principal = R2 / L4 * (L4 - 2.0 * L2 * M * R + a2 * R2);
  const double s = static_cast<double>(sw);
  const double m = static_cast<double>(mm);
  c.definition = 2.0 * I * a * m * (1.0 + 4.0 * mrl) +
                 2.0 * (2.0 * M * (-s + (2.0 + s) * 2.0 * mrl -
                                   3.0 * a2 * R2 / L4) -
                        a2 * R / L2 + I * s * a * std::cos(theta));
  c.q = 2.0 * R * (1.0 + s - (3.0 + s) * mrl + 2.0 * a2 * R2 / L4) -
        2.0 * I * a * m * R2 / L2;
  c.psi = -2.0 * R * ((1.0 + s) * M / L2 - a2 * R / L4) -
          2.0 * I * a * m * R / L2;
  return c;
}

// --------------------------------------------------------------- SBP
int min_points(int scheme) { return scheme == 2 ? 22 : (scheme == 1 ? 16 : 8); }

double norm_weight(int scheme, int n, int i) {
  const int d = std::min(i, n - 1 - i);
  if (scheme == 0) return d < 4 ? rat(kD42Norm[d]) : 1.0;
  if (scheme == 1) return d < 8 ? rat(kD84Norm[d]) : 1.0;
  return d < 11 ? rat(kD105Norm[d]) : 1.0;
}

// D10-5 interior coefficient of column c in row r (sbp105.hpp:44-56).
double d105_interior(int r, int c) {
  static const double k[6] = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0,
                              -5.0 / 504.0, 1.0 / 1260.0};
  const int off = c - r;
  if (off == 0 || off < -5 || off > 5) return 0.0;
  return off > 0 ? k[off] : -k[-off];
}

// Q = H*D left block entry (sbp105.hpp:61-207).
double d105_q(int r, int c) {
  if (c >= 11) return d105_interior(r, c);
  if (r == 0 && c == 0) return -0.5;
  if (r == c) return 0.0;
  if (c > r) return rat(kD105QUpper[r][c - r - 1]);
  return -rat(kD105QUpper[c][r - c - 1]);
}

// Left closure coefficient of h*D (sbp.hpp:33-70, sbp84.hpp:146-160,
// sbp105.hpp:209-213).
double left_coefficient(int scheme, int row, int col) {
  if (scheme == 0) return (col < 6) ? rat(kD42Left[row][col]) : 0.0;
  if (scheme == 1) return (col < 12) ? rat(kD84Left[row][col]) : 0.0;
  if (col >= 16) return 0.0;
  return d105_q(row, col) / rat(kD105Norm[row]);
}

int closure_rows(int scheme) { return scheme == 2 ? 11 : (scheme == 1 ? 8 : 4); }
int closure_width(int scheme) { return scheme == 2 ? 16 : (scheme == 1 ? 12 : 6); }

// Full-matrix entry h*D(row, col).
double derivative_entry(int scheme, int n, int row, int col) {
  const int w = closure_rows(scheme);
  if (row < w) return left_coefficient(scheme, row, col);
  if (row + w >= n) return -left_coefficient(scheme, n - 1 - row, n - 1 - col);
  const int off = col - row;
  if (scheme == 0) {
    static const double k[3] = {0, 2.0 / 3.0, -1.0 / 12.0};
    if (off == 0 || std::abs(off) > 2) return 0.0;
    return off > 0 ? k[off] : -k[-off];
  }
  if (scheme == 1) {
    static const double k[5] = {0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0,
                                -1.0 / 280.0};
    if (off == 0 || std::abs(off) > 4) return 0.0;
    return off > 0 ? k[off] : -k[-off];
  }
  return d105_interior(row, col);
}

// Strided first derivative at one row, with the reference's interior
// expression order (sbp.hpp:125-154, sbp84.hpp:185-217,
// sbp105.hpp:267-305).
cplx derivative_at(int scheme, const cplx* v, int n, int row, double inv_h,
                   size_t stride) {
  const int w = closure_rows(scheme);
  const int cw = closure_width(scheme);
  cplx r = 0.0;
  if (row < w) {
    for (int c = 0; c < cw; ++c) r += left_coefficient(scheme, row, c) * v[c * stride];
    return inv_h * r;
  }
  if (row + w >= n) {
    const int rr = n - 1 - row;
    for (int c = 0; c < cw; ++c)
      r -= left_coefficient(scheme, rr, c) * v[(n - 1 - c) * stride];
    return inv_h * r;
  }
  auto at = [&](int k) { return v[(row + k) * stride]; };
  if (scheme == 0) {
    return inv_h * ((at(-2) - 8.0 * at(-1) + 8.0 * at(1) - at(2)) / 12.0);
  }
  if (scheme == 1) {
    r = (at(-4) - at(4)) / 280.0 + 4.0 * (-at(-3) + at(3)) / 105.0 +
        (at(-2) - at(2)) / 5.0 + 4.0 * (-at(-1) + at(1)) / 5.0;
    return inv_h * r;
  }
  r = (-at(-5) + at(5)) / 1260.0 + 5.0 * (at(-4) - at(4)) / 504.0 +
      5.0 * (-at(-3) + at(3)) / 84.0 + 5.0 * (at(-2) - at(2)) / 21.0 +
      5.0 * (-at(-1) + at(1)) / 6.0;
  return inv_h * r;
}

// Strided compatible dissipation -(eps/h) H^-1 A^T A u at one row
// (radial_discretization.hpp:165-232).
cplx dissipation_at(int scheme, const cplx* v, int n, int index, double h,
                    double eps, size_t stride) {
  static const double s105[7] = {1, -6, 15, -20, 15, -6, 1};
  static const double s84[6] = {-1, 5, -10, 10, -5, 1};
  static const double s42[4] = {-1, 3, -3, 1};
  const int order = scheme == 2 ? 6 : (scheme == 1 ? 5 : 3);
  const double* st = scheme == 2 ? s105 : (scheme == 1 ? s84 : s42);
  const int first = index > order ? index - order : 0;
  const int last = index < n - order ? index : n - order - 1;
  cplx np = 0.0;
  for (int row = first; row <= last; ++row) {
    // Same left-to-right term order as the reference's written stencils.
    cplx diff;
    if (scheme == 2) {
      diff = v[row * stride] - 6.0 * v[(row + 1) * stride] +
             15.0 * v[(row + 2) * stride] - 20.0 * v[(row + 3) * stride] +
             15.0 * v[(row + 4) * stride] - 6.0 * v[(row + 5) * stride] +
             v[(row + 6) * stride];
    } else if (scheme == 1) {
      diff = -v[row * stride] + 5.0 * v[(row + 1) * stride] -
             10.0 * v[(row + 2) * stride] + 10.0 * v[(row + 3) * stride] -
             5.0 * v[(row + 4) * stride] + v[(row + 5) * stride];
    } else {
      diff = -v[row * stride] + 3.0 * v[(row + 1) * stride] -
             3.0 * v[(row + 2) * stride] + v[(row + 3) * stride];
    }
    np += st[index - row] * diff;
  }
  return -eps * np / (h * norm_weight(scheme, n, index));
}

// ------------------------------------------------------------ dense solve
// Partial-pivot Gaussian elimination (dense_solve.hpp:20-72).
bool dense_solve(int n, std::vector<cplx> mat, std::vector<cplx> rhs,
                 std::vector<cplx>& sol) {
  double scale = 0.0;
  for (const cplx& v : mat) scale = std::max(scale, std::abs(v));
  const double floor = 128.0 * std::numeric_limits<double>::epsilon() * scale;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double pm = std::abs(mat[c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      const double cand = std::abs(mat[r * n + c]);
      if (cand > pm) {
        piv = r;
        pm = cand;
      }
    }
    if (!std::isfinite(pm) || pm <= floor) return false;
    if (piv != c) {
      for (int j = c; j < n; ++j) std::swap(mat[c * n + j], mat[piv * n + j]);
      std::swap(rhs[c], rhs[piv]);
    }
    const cplx diag = mat[c * n + c];
    for (int r = c + 1; r < n; ++r) {
      const cplx f = mat[r * n + c] / diag;
      mat[r * n + c] = 0.0;
      for (int j = c + 1; j < n; ++j) mat[r * n + j] -= f * mat[c * n + j];
      rhs[r] -= f * rhs[c];
    }
  }
  sol.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    const int r = n - 1 - k;
    cplx v = rhs[r];
    for (int c = r + 1; c < n; ++c) v -= mat[r * n + c] * sol[c];
    sol[r] = v / mat[r * n + r];
  }
  return true;
}

double dense_residual(int n, const std::vector<cplx>& mat,
                      const std::vector<cplx>& sol,
                      const std::vector<cplx>& rhs) {
  double rs = 0.0, bs = 0.0;
  for (int r = 0; r < n; ++r) {
    cplx res = -rhs[r];
    for (int c = 0; c < n; ++c) res += mat[r * n + c] * sol[c];
    rs += std::abs(res) * std::abs(res);
    bs += std::abs(rhs[r]) * std::abs(rhs[r]);
  }
  return std::sqrt(rs) / std::max(std::sqrt(bs), std::numeric_limits<double>::min());
}

bool contains(const std::vector<int>& v, int x) {
  return std::find(v.begin(), v.end(), x) != v.end();
}

}  // namespace

// ------------------------------------------------------------- pipeline
namespace {

// One angular coordinator (angular_coordinator.hpp:20-108): a (spin, boost,
// ell_max) band over the stored modes it is active for, with the spin +- 1
// transforms its eth / eth' operators synthesize into (ghp_device.hpp:
// 174-196).
struct Coord {
  int spin = 0, boost = 0, ell_max = 0;
  std::vector<bool> active;
  std::vector<Transform> own, raised, lowered;
};

}  // namespace

struct OraclePipeline {
  OracleParams p;
  int M = 0, N = 0, T = 0;
  double h = 0, inv_h = 0, horizon = 0;
  std::vector<int> modes, parents, targets, sharp;
  std::vector<double> theta, cos_t, sin_t;  // acos(x_j), x_j, sqrt(1 - x_j^2)
  std::vector<Transform> first;            // per stored mode (empty if inactive)
  std::vector<bool> parent_active;
  std::vector<cplx> state, k1, k2, k3, k4, stage;
  std::vector<cplx> lap, scratch;          // lap: M*N*T, scratch: M*5*N*T
  // second order (allocated when the source policy is not disabled)
  bool second = false;
  Coord c_first, c_second, c_g, c_b, c_pi, c_c, c_scalar, c_bsharp, c_csharp, c_t;
  std::vector<cplx> tangent, scratch_t, rdr, rdr_t, rang, rang_t, extra, sharpw;
  std::vector<cplx> constraints, terms, inner, inner_t, proj, proj_t, ethprime_t;
  std::vector<cplx> source_over_r3, forcing, zero_plane;
  std::vector<std::pair<int, int>> pair_list;   // (index1, index2) per ordered pair
  std::vector<int> target_slots, pair_offsets;  // target mode index, pair range
  // activation latch (spatial_pipeline.hpp:386-424, 1225-1289)
  OracleActivation act{};
  OracleConstraintNorms norms{};
  int stage_sources[8] = {0};
  int stage_source_count = 0;

  size_t idx(int mode, int field, int r, int t) const {
    return ((static_cast<size_t>(mode) * kFields + field) * N + r) * T + t;
  }
  size_t plane(int mode, int r, int t) const {
    return (static_cast<size_t>(mode) * N + r) * T + t;
  }
};

namespace {

// A (mode, radial, theta) plane inside a larger array: element (mode, r, t)
// at base[mode * ms + r * T + t].
struct PlaneW {
  cplx* b;
  size_t ms;
};
struct Plane {
  const cplx* b;
  size_t ms;
  Plane(const cplx* base, size_t stride) : b(base), ms(stride) {}
  Plane(PlaneW w) : b(w.b), ms(w.ms) {}
};

Plane field_of(const OraclePipeline* p, const cplx* arr, int field) {
  return {arr + static_cast<size_t>(field) * p->N * p->T, static_cast<size_t>(kFields) * p->N * p->T};
}
PlaneW field_of(const OraclePipeline* p, cplx* arr, int field) {
  return {arr + static_cast<size_t>(field) * p->N * p->T, static_cast<size_t>(kFields) * p->N * p->T};
}
Plane slot_of(const OraclePipeline* p, const cplx* arr, int slots, int slot) {
  return {arr + static_cast<size_t>(slot) * p->N * p->T, static_cast<size_t>(slots) * p->N * p->T};
}
PlaneW slot_of(const OraclePipeline* p, cplx* arr, int slots, int slot) {
  return {arr + static_cast<size_t>(slot) * p->N * p->T, static_cast<size_t>(slots) * p->N * p->T};
}
Plane ro(PlaneW w) { return {w.b, w.ms}; }

void set_threads(const OraclePipeline* p) {
#ifdef _OPENMP
  if (p->p.threads > 0) omp_set_num_threads(p->p.threads);
#else
  (void)p;
#endif
}

Coord make_coord(const OraclePipeline* p, int spin, int boost, int E,
                 const std::vector<int>& active_modes, bool ghp) {
  Coord c;
  c.spin = spin;
  c.boost = boost;
  c.ell_max = E;
  c.active.assign(p->M, false);
  c.own.resize(p->M);
  c.raised.resize(p->M);
  c.lowered.resize(p->M);
  for (int i = 0; i < p->M; ++i) {
    if (!contains(active_modes, p->modes[i])) continue;
    c.active[i] = true;
    c.own[i] = make_transform(spin, p->modes[i], E, p->T);
    if (ghp) {
      c.raised[i] = make_transform(spin + 1, p->modes[i], E, p->T);
      c.lowered[i] = make_transform(spin - 1, p->modes[i], E, p->T);
    }
  }
  return c;
}

void analyze(const Transform& tr, const cplx* in, int T, std::vector<cplx>& modal) {
  const int nm = tr.modes();
  modal.assign(nm, 0.0);
  for (int k = 0; k < nm; ++k) {
    cplx v(0.0, 0.0);
    for (int t = 0; t < T; ++t) v += tr.A[k * T + t] * in[t];
    modal[k] = v;
  }
}

void synthesize(const Transform& tr, const std::vector<cplx>& modal, int T, cplx* out) {
  const int nm = tr.modes();
  for (int t = 0; t < T; ++t) {
    cplx v(0.0, 0.0);
    for (int k = 0; k < nm; ++k) v += tr.S[t * nm + k] * modal[k];
    out[t] = v;
  }
}

// Band projection S A (angular_coordinator.hpp:125-146); inactive modes zero.
void project(const OraclePipeline* p, const Coord& c, Plane in, PlaneW out) {
  const int N = p->N, T = p->T;
#pragma omp parallel for collapse(2) schedule(static)
  for (int mode = 0; mode < p->M; ++mode)
    for (int r = 0; r < N; ++r) {
      cplx* o = out.b + mode * out.ms + static_cast<size_t>(r) * T;
      if (!c.active[mode]) {
        for (int t = 0; t < T; ++t) o[t] = 0.0;
        continue;
      }
      std::vector<cplx> modal;
      analyze(c.own[mode], in.b + mode * in.ms + static_cast<size_t>(r) * T, T, modal);
      std::vector<cplx> tmp(T);
      synthesize(c.own[mode], modal, T, tmp.data());
      for (int t = 0; t < T; ++t) o[t] = tmp[t];
    }
}

// S Lambda A (angular_coordinator.hpp:151-174, angular.hpp:139-165).
void laplacian(const OraclePipeline* p, const Coord& c, Plane in, PlaneW out) {
  const int N = p->N, T = p->T;
#pragma omp parallel for collapse(2) schedule(static)
  for (int mode = 0; mode < p->M; ++mode)
    for (int r = 0; r < N; ++r) {
      cplx* o = out.b + mode * out.ms + static_cast<size_t>(r) * T;
      if (!c.active[mode]) {
        for (int t = 0; t < T; ++t) o[t] = 0.0;
        continue;
      }
      const Transform& tr = c.own[mode];
      std::vector<cplx> modal;
      analyze(tr, in.b + mode * in.ms + static_cast<size_t>(r) * T, T, modal);
      for (int k = 0; k < tr.modes(); ++k) {
        const int l = tr.ell_min + k;
        modal[k] *= -static_cast<double>((l - tr.spin) * (l + tr.spin + 1));
      }
      synthesize(tr, modal, T, o);
    }
}

// eth_n X (prime = false) or eth'_n X (prime = true) at one point, from the
// angular part `ang` of X (its spin-raised, or lowered, band synthesis), its
// coordinate-time derivative and the point's (R, cos, sin): the coordinate
// forms of ghp.hpp:60-116 with the connection weight p = s + b (q = -s + b).
cplx eth_point(bool prime, int spin, int boost, double a, double L, double R, double ct,
               double st, cplx x, cplx dt, cplx ang) {
  const double L2 = L * L;
  const double sqrt_two = std::sqrt(2.0);
  const cplx I(0.0, 1.0);
  if (!prime) {
    const int pw = spin + boost;
    const cplx den(L2, -a * R * ct);
    const cplx deriv = (-I * a * st * dt + ang) / (sqrt_two * den);
    const cplx conn = -I * static_cast<double>(pw) * a * R * st * x / (sqrt_two * den * den);
    return deriv + conn;
  }
  const int qw = -spin + boost;
  const cplx den(L2, a * R * ct);
  const cplx deriv = (I * a * st * dt + ang) / (sqrt_two * den);
  const cplx conn = I * static_cast<double>(qw) * a * R * st * x / (sqrt_two * den * den);
  return deriv + conn;
}

// eth_n X and eth'_n X (ghp.hpp:60-116) with the angular part from the
// spin-raised (lowered) band: analyze, multiply by the raising factor
// sqrt((l - s)(l + s + 1)) (lowering factor -sqrt((l + s)(l - s + 1))),
// align ell to the shifted band, synthesize (ghp_device.hpp:174-196,
// angular.hpp:129-137), then the point formula above.
void ghp(const OraclePipeline* p, const Coord& c, bool prime, Plane x, Plane dt, PlaneW out) {
  const int N = p->N, T = p->T;
  const double a = p->p.spin, L = p->p.length;
#pragma omp parallel for collapse(2) schedule(static)
  for (int mode = 0; mode < p->M; ++mode)
    for (int r = 0; r < N; ++r) {
      cplx* o = out.b + mode * out.ms + static_cast<size_t>(r) * T;
      if (!c.active[mode]) {
        for (int t = 0; t < T; ++t) o[t] = 0.0;
        continue;
      }
      const Transform& src = c.own[mode];
      const Transform& dst = prime ? c.lowered[mode] : c.raised[mode];
      const cplx* xv = x.b + mode * x.ms + static_cast<size_t>(r) * T;
      const cplx* dv = dt.b + mode * dt.ms + static_cast<size_t>(r) * T;
      std::vector<cplx> modal;
      analyze(src, xv, T, modal);
      std::vector<cplx> shifted(dst.modes(), cplx(0.0, 0.0));
      for (int j = 0; j < dst.modes(); ++j) {
        const int l = dst.ell_min + j;
        const int k = l - src.ell_min;
        if (k < 0 || k >= src.modes()) continue;
        const double f =
            prime ? -std::sqrt(static_cast<double>((l + c.spin) * (l - c.spin + 1)))
                  : std::sqrt(static_cast<double>((l - c.spin) * (l + c.spin + 1)));
        shifted[j] = f * modal[k];
      }
      std::vector<cplx> ang(T);
      synthesize(dst, shifted, T, ang.data());
      const double R = p->h * r;
      for (int t = 0; t < T; ++t)
        o[t] = eth_point(prime, c.spin, c.boost, a, L, R, p->cos_t[t], p->sin_t[t], xv[t], dv[t],
                         ang[t]);
    }
}

// SBP radial derivative of one plane (radial_discretization.hpp:130-144).
void radial_d(const OraclePipeline* p, Plane in, PlaneW out) {
  const int N = p->N, T = p->T;
#pragma omp parallel for collapse(3) schedule(static)
  for (int mode = 0; mode < p->M; ++mode)
    for (int r = 0; r < N; ++r)
      for (int t = 0; t < T; ++t)
        out.b[mode * out.ms + static_cast<size_t>(r) * T + t] =
            derivative_at(p->p.scheme, in.b + mode * in.ms + t, N, r, p->inv_h, T);
}

// Rescaled rotated-Kinnersley background (background.hpp:29-78).
struct Background {
  cplx mu0, tau0, pi0, psi20, rho0, epsilon0;
};

Background kerr_background(double M, double a, double L, double R, double ct, double st) {
  const double L2 = L * L, L4 = L2 * L2;
  const cplx I(0.0, 1.0);
  const double sqrt_two = std::sqrt(2.0);
  const cplx minus_radial(-L2, a * R * ct);
  const cplx angular(L2, -a * R * ct);
  const cplx angular2 = angular * angular;
  const double real_pi = L4 + a * a * R * R * ct * ct;
  Background b;
  b.mu0 = cplx(1.0, 0.0) / minus_radial;
  b.tau0 = I * a * st / (sqrt_two * angular2);
  b.pi0 = -I * a * st / (sqrt_two * real_pi);
  const cplx ingoing(L2, a * R * ct);
  const cplx optical = angular2 * ingoing;
  b.rho0 = -0.5 * (L4 - 2.0 * L2 * M * R + a * a * R * R) / optical;
  b.epsilon0 = 0.5 * cplx(L2 * M - a * a * R, -a * (L2 - M * R) * ct) / optical;
  b.psi20 = -M / (angular2 * angular);
  return b;
}

Background background_at(const OraclePipeline* p, int r, int t) {
  return kerr_background(p->p.mass, p->p.spin, p->p.length, p->h * r, p->cos_t[t], p->sin_t[t]);
}

// Background at the theta coordinate itself (the reconstruction passes call
// kerr_background_point(parameters, radius, theta), full_spatial.hpp:518).
Background background_theta(const OraclePipeline* p, int r, int t) {
  return kerr_background(p->p.mass, p->p.spin, p->p.length, p->h * r, std::cos(p->theta[t]),
                         std::sin(p->theta[t]));
}

// Delta_n from point values (ghp.hpp:11-22).
cplx delta_n(cplx v, cplx dt, cplx dr, int n, double R, double M, double L) {
  const double L2 = L * L;
  const double tc = 2.0 + 4.0 * M * R / L2;
  const double rc = R / L2;
  return tc * dt + rc * (R * dr + static_cast<double>(n) * v);
}

// Solve Delta_n X = R_X for the time derivative (reconstruction.hpp:108-121).
cplx recon_dt(cplx v, cplx dr, cplx rhs, int n, double R, double M, double L) {
  const double L2 = L * L;
  const double den = 2.0 + 4.0 * M * R / L2;
  const cplx radial = (R * R / L2) * dr + (static_cast<double>(n) * R / L2) * v;
  return (rhs - radial) / den;
}

// thorn_n (ghp.hpp:24-58).
cplx thorn_n(cplx v, cplx dt, cplx dr, int n, int sw, int bw, int m, double R, double ct,
             double M, double a, double L, cplx eps0) {
  const cplx I(0.0, 1.0);
  const double L2 = L * L, L4 = L2 * L2;
  const double den = L4 + a * a * R * R * ct * ct;
  const double tc = R * 2.0 * M * (2.0 * M - a * a * R / L2) / den;
  const double rc = -0.5 * (L2 - 2.0 * M * R + a * a * R * R / L2) / den;
  const int pw = sw + bw, qw = -sw + bw;
  return tc * dt + rc * (R * dr + static_cast<double>(n) * v) +
         R * I * a * static_cast<double>(m) * v / den -
         R * (static_cast<double>(pw) * eps0 + static_cast<double>(qw) * std::conj(eps0)) * v;
}

// Three-pass full-grid Teukolsky RHS of one (P, Q, Psi) triple
// (full_spatial.hpp:99-280, 303-398): stage fields s0..s0+2 -> out fields
// o0..o0+2, Laplacian through `coord`, optional forcing, scratch slots
// (D Psi, Q_eff, D Q_eff, psi_t, D psi_t).  The caller projects.
void linear_rhs(OraclePipeline* p, const cplx* stage, int s0, cplx* out, int o0, const Coord& coord,
                const cplx* forcing, std::vector<cplx>& scratch) {
  const int M = p->M, N = p->N, T = p->T;
  const int scheme = p->p.scheme;
  const size_t sstride = T;
  enum { DPSI = 0, QEFF = 1, DQ = 2, PSIT = 3, DPSIT = 4 };
  auto sc = [&](int mode, int slot, int r, int t) -> cplx& {
    return scratch[((static_cast<size_t>(mode) * 5 + slot) * N + r) * T + t];
  };
  laplacian(p, coord, field_of(p, stage, s0 + 2), PlaneW{p->lap.data(), static_cast<size_t>(N) * T});
  const bool constrained = p->p.reduction == 1;
#pragma omp parallel for collapse(3) schedule(static)
  for (int mode = 0; mode < M; ++mode)
    for (int r = 0; r < N; ++r)
      for (int t = 0; t < T; ++t) {
        const cplx d = derivative_at(scheme, &stage[p->idx(mode, s0 + 2, 0, t)], N, r, p->inv_h,
                                     sstride);
        sc(mode, DPSI, r, t) = d;
        sc(mode, QEFF, r, t) = constrained ? d : stage[p->idx(mode, s0 + 1, r, t)];
      }
#pragma omp parallel for collapse(3) schedule(static)
  for (int mode = 0; mode < M; ++mode)
    for (int r = 0; r < N; ++r)
      for (int t = 0; t < T; ++t) {
        sc(mode, DQ, r, t) = derivative_at(scheme, &sc(mode, QEFF, 0, t), N, r, p->inv_h, sstride);
        const Coeffs c = teukolsky_coefficients(p->p.mass, p->p.spin, p->p.length, -2,
                                                p->modes[mode], p->h * r, p->theta[t]);
        const cplx P = stage[p->idx(mode, s0, r, t)];
        const cplx Q = sc(mode, QEFF, r, t);
        const cplx Psi = stage[p->idx(mode, s0 + 2, r, t)];
        sc(mode, PSIT, r, t) = (P + 2.0 * c.adv * Q - c.definition * Psi) / c.time;
      }
  const double damping = constrained ? 0.0 : p->p.reduction_damping;
#pragma omp parallel for collapse(3) schedule(static)
  for (int mode = 0; mode < M; ++mode)
    for (int r = 0; r < N; ++r)
      for (int t = 0; t < T; ++t) {
        const cplx dpt = derivative_at(scheme, &sc(mode, PSIT, 0, t), N, r, p->inv_h, sstride);
        sc(mode, DPSIT, r, t) = dpt;
        const Coeffs c = teukolsky_coefficients(p->p.mass, p->p.spin, p->p.length, -2,
                                                p->modes[mode], p->h * r, p->theta[t]);
        const cplx Qe = sc(mode, QEFF, r, t);
        const cplx Psi = stage[p->idx(mode, s0 + 2, r, t)];
        const cplx constraint = stage[p->idx(mode, s0 + 1, r, t)] - sc(mode, DPSI, r, t);
        const cplx lapv = p->lap[p->plane(mode, r, t)];
        const cplx frc = forcing ? forcing[p->plane(mode, r, t)] : cplx(0.0, 0.0);
        cplx Pt = c.principal * sc(mode, DQ, r, t) + c.q * Qe + c.psi * Psi + lapv + frc;
        cplx Qt = dpt - damping * constraint;
        cplx St = sc(mode, PSIT, r, t);
        if (p->p.dissipation > 0.0) {
          Pt += dissipation_at(scheme, &stage[p->idx(mode, s0, 0, t)], N, r, p->h,
                               p->p.dissipation, sstride);
          Qt += dissipation_at(scheme, &stage[p->idx(mode, s0 + 1, 0, t)], N, r, p->h,
                               p->p.dissipation, sstride);
          St += dissipation_at(scheme, &stage[p->idx(mode, s0 + 2, 0, t)], N, r, p->h,
                               p->p.dissipation, sstride);
        }
        out[p->idx(mode, o0, r, t)] = Pt;
        out[p->idx(mode, o0 + 1, r, t)] = Qt;
        out[p->idx(mode, o0 + 2, r, t)] = St;
      }
}

void project_triple(OraclePipeline* p, cplx* out, int o0, const Coord& c) {
  for (int f = o0; f < o0 + 3; ++f) project(p, c, field_of(p, out, f), field_of(p, out, f));
}

void evaluate_rhs(OraclePipeline* p, const cplx* stage, cplx* out);

#include "second_order.inc"

// Full RHS (spatial_pipeline.hpp:458-588): the first-order triple; with the
// source policy disabled, fields 3..12 are zero (623-649); otherwise the
// reconstruction chain, its tangent, the quadratic source and the
// second-order triple.
void evaluate_rhs(OraclePipeline* p, const cplx* stage, cplx* out) {
  linear_rhs(p, stage, 0, out, 0, p->c_first, nullptr, p->scratch);
  project_triple(p, out, 0, p->c_first);
  if (!p->second) {
    const int M = p->M, N = p->N, T = p->T;
#pragma omp parallel for collapse(2) schedule(static)
    for (int mode = 0; mode < M; ++mode)
      for (int f = 3; f < kFields; ++f)
        std::fill(&out[p->idx(mode, f, 0, 0)],
                  &out[p->idx(mode, f, 0, 0)] + static_cast<size_t>(N) * T, cplx(0.0, 0.0));
    return;
  }
  second_order_rhs(p, stage, out);
}

// Classical RK4 over the whole 13-field state (device_rk4.hpp:84-128); the
// stage source flags record the latch seen by each stage.
void rk4_step(OraclePipeline* p, double dt) {
  const size_t n = p->state.size();
  cplx* y = p->state.data();
  auto rhs = [&](const cplx* s, cplx* k) {
    if (p->stage_source_count < 8) p->stage_sources[p->stage_source_count++] = p->act.active;
    evaluate_rhs(p, s, k);
  };
  rhs(y, p->k1.data());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) p->stage[i] = y[i] + (0.5 * dt) * p->k1[i];
  rhs(p->stage.data(), p->k2.data());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) p->stage[i] = y[i] + (0.5 * dt) * p->k2[i];
  rhs(p->stage.data(), p->k3.data());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) p->stage[i] = y[i] + dt * p->k3[i];
  rhs(p->stage.data(), p->k4.data());
  const double scale = dt / 6.0;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i)
    y[i] += scale * (p->k1[i] + 2.0 * p->k2[i] + 2.0 * p->k3[i] + p->k4[i]);
}

// SpatialPipeline::step (spatial_pipeline.hpp:590-620): the RK4 step split
// exactly at the source start time, with the latch considered on accepted
// states only.
void pipeline_step(OraclePipeline* p, double time, double dt) {
  p->stage_source_count = 0;
  const double end = time + dt;
  const double start = p->p.source_start_time;
  if (!p->act.active && time >= start) consider_activation(p, time);
  if (!p->act.active && time < start && start < end) {
    if (start - time > 0.0) rk4_step(p, start - time);
    consider_activation(p, start);
    if (end - start > 0.0) rk4_step(p, end - start);
  } else {
    rk4_step(p, dt);
  }
  if (!p->act.active && end >= start) consider_activation(p, end);
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

OraclePipeline* oracle_create(const OracleParams* params) {
  try {
    auto* p = new OraclePipeline();
    p->p = *params;
    const OracleParams& q = *params;
    if (q.n_modes <= 0 || q.n_modes > 64 || q.n_parents <= 0 || q.n_parents > 64 ||
        q.n_targets <= 0 || q.n_targets > 64)
      throw std::invalid_argument("mode lists must be nonempty and <= 64");
    p->modes.assign(q.modes, q.modes + q.n_modes);
    p->parents.assign(q.parents, q.parents + q.n_parents);
    p->targets.assign(q.targets, q.targets + q.n_targets);
    std::sort(p->modes.begin(), p->modes.end());
    if (q.nr < min_points(q.scheme)) throw std::invalid_argument("nr too small");
    if (q.scheme < 0 || q.scheme > 2) throw std::invalid_argument("bad scheme");
    p->M = q.n_modes;
    p->N = q.nr;
    p->T = q.ntheta;
    const double disc = q.mass * q.mass - q.spin * q.spin;
    const double rplus = q.mass + std::sqrt(disc > 0.0 ? disc : 0.0);
    p->horizon = q.length * q.length / rplus;
    p->h = (p->horizon - 0.0) / static_cast<double>(q.nr - 1);
    p->inv_h = 1.0 / p->h;
    std::vector<double> x, w;
    gauss_legendre(q.ntheta, x, w);
    p->theta.resize(q.ntheta);
    for (int t = 0; t < q.ntheta; ++t) p->theta[t] = std::acos(x[t]);
    p->cos_t = x;
    p->sin_t.resize(q.ntheta);
    for (int t = 0; t < q.ntheta; ++t) p->sin_t[t] = std::sqrt(1.0 - x[t] * x[t]);
    p->first.resize(p->M);
    p->parent_active.assign(p->M, false);
    for (int i = 0; i < p->M; ++i) {
      if (contains(p->parents, p->modes[i])) {
        p->parent_active[i] = true;
        p->first[i] = make_transform(-2, p->modes[i], q.ell_max_first, q.ntheta);
      }
    }
    if (q.source_mode < 0 || q.source_mode > 2 || !std::isfinite(q.source_start_time) ||
        q.source_start_time < 0.0 || !std::isfinite(q.constraint_tolerance) ||
        q.constraint_tolerance < 0.0 || q.required_passes < 1)
      throw std::invalid_argument("source policy is invalid");
    p->second = q.source_mode != 0;
    p->c_first = make_coord(p, weight::F.spin, weight::F.boost, q.ell_max_first, p->parents, p->second);
    p->sharp.assign(p->M, -1);
    for (int i = 0; i < p->M; ++i) {
      const auto it = std::find(p->modes.begin(), p->modes.end(), -p->modes[i]);
      if (it != p->modes.end()) p->sharp[i] = static_cast<int>(it - p->modes.begin());
    }
    if (p->second) {
      for (int i = 0; i < p->M; ++i)
        if (p->sharp[i] < 0) throw std::invalid_argument("modes must be closed under m -> -m");
      const int E1 = q.ell_max_first, E2 = q.ell_max_second;
      if (E1 < 3 || E2 < 3 || q.ntheta < std::max({E1 + 1, E2 + 1, (2 * E1 + E2 + 2) / 2}))
        throw std::invalid_argument("angular bands require exact-product padding");
      // Coordinators of spatial_pipeline.hpp:185-218.
      p->c_second = make_coord(p, weight::F.spin, weight::F.boost, E2, p->targets, false);
      p->c_g = make_coord(p, weight::G.spin, weight::G.boost, E1, p->parents, true);
      p->c_b = make_coord(p, weight::B.spin, weight::B.boost, E1, p->parents, true);
      p->c_pi = make_coord(p, weight::Pi.spin, weight::Pi.boost, E1, p->parents, true);
      p->c_c = make_coord(p, weight::C.spin, weight::C.boost, E1, p->parents, true);
      p->c_scalar = make_coord(p, weight::H.spin, weight::H.boost, E1, p->parents, true);
      p->c_bsharp = make_coord(p, weight::sharp(weight::B).spin, weight::sharp(weight::B).boost, E1, p->parents, true);
      p->c_csharp = make_coord(p, weight::sharp(weight::C).spin, weight::sharp(weight::C).boost, E1, p->parents, true);
      p->c_t = make_coord(p, weight::T.spin, weight::T.boost, E2, p->targets, true);
      // Ordered pairs (m1, m2) -> m1 + m2, target-major then m1 ascending,
      // both factors parents (modes.hpp:32-213).
      std::vector<int> sorted_targets = p->targets;
      std::sort(sorted_targets.begin(), sorted_targets.end());
      p->pair_offsets.assign(1, 0);
      for (int m : sorted_targets) {
        const int ti = static_cast<int>(std::find(p->modes.begin(), p->modes.end(), m) - p->modes.begin());
        p->target_slots.push_back(ti);
        std::vector<int> sp = p->parents;
        std::sort(sp.begin(), sp.end());
        for (int m1 : sp) {
          const int m2 = m - m1;
          if (!contains(p->parents, m2)) continue;
          const int i1 = static_cast<int>(std::find(p->modes.begin(), p->modes.end(), m1) - p->modes.begin());
          const int i2 = static_cast<int>(std::find(p->modes.begin(), p->modes.end(), m2) - p->modes.begin());
          p->pair_list.push_back({i1, i2});
        }
        p->pair_offsets.push_back(static_cast<int>(p->pair_list.size()));
      }
      const size_t plane = static_cast<size_t>(p->M) * p->N * p->T;
      const size_t full = plane * kFields;
      p->tangent.assign(full, 0.0);
      p->scratch_t.assign(plane * 5, 0.0);
      p->rdr.assign(plane * 7, 0.0);
      p->rdr_t.assign(plane * 7, 0.0);
      p->rang.assign(plane * 6, 0.0);
      p->rang_t.assign(plane * 6, 0.0);
      p->extra.assign(plane * 4, 0.0);
      p->sharpw.assign(plane * 4, 0.0);
      p->constraints.assign(plane * 3, 0.0);
      p->terms.assign(plane * 6, 0.0);
      p->inner.assign(plane * 2, 0.0);
      p->inner_t.assign(plane * 2, 0.0);
      p->proj.assign(plane * 2, 0.0);
      p->proj_t.assign(plane * 2, 0.0);
      p->ethprime_t.assign(plane, 0.0);
      p->source_over_r3.assign(plane, 0.0);
      p->forcing.assign(plane, 0.0);
    }
    p->act = OracleActivation{0, -1.0, 0, -1.0};
    if (q.source_mode == 2 && q.source_start_time == 0.0)
      p->act = OracleActivation{1, 0.0, q.required_passes, -1.0};
    const size_t count = static_cast<size_t>(p->M) * kFields * p->N * p->T;
    p->state.assign(count, 0.0);
    p->k1.assign(count, 0.0);
    p->k2.assign(count, 0.0);
    p->k3.assign(count, 0.0);
    p->k4.assign(count, 0.0);
    p->stage.assign(count, 0.0);
    p->lap.assign(static_cast<size_t>(p->M) * p->N * p->T, 0.0);
    p->scratch.assign(static_cast<size_t>(p->M) * 5 * p->N * p->T, 0.0);
    return p;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void oracle_destroy(OraclePipeline* p) { delete p; }

long long oracle_value_count(const OraclePipeline* p) {
  return static_cast<long long>(p->state.size());
}

void oracle_get_state(const OraclePipeline* p, double* out) {
  std::memcpy(static_cast<void*>(out), static_cast<const void*>(p->state.data()), p->state.size() * sizeof(cplx));
}

void oracle_set_state(OraclePipeline* p, const double* in) {
  std::memcpy(static_cast<void*>(p->state.data()), static_cast<const void*>(in), p->state.size() * sizeof(cplx));
}

double oracle_init_gaussian(OraclePipeline* p, double center, double width,
                            int compact_support, int n_seeds, const int* ells,
                            const int* ms, const double* amp_re,
                            const double* amp_im) {
  try {
    set_threads(p);
    const int M = p->M, N = p->N, T = p->T;
    const int E = p->p.ell_max_first;
    std::fill(p->state.begin(), p->state.end(), cplx(0.0, 0.0));
    auto profile = [&](double R) {
      const double z = (R - center) / width;
      if (!compact_support) return std::exp(-z * z);
      if (std::abs(R - center) >= width) return 0.0;
      return std::exp(1.0 - 1.0 / (1.0 - z * z));
    };
    // Psi = profile * synthesized seed (pipeline_initial_data.hpp:258-275).
    for (int mode = 0; mode < M; ++mode) {
      const int m = p->modes[mode];
      if (!p->parent_active[mode]) continue;
      const Transform tr = make_transform(-2, m, E, T);
      std::vector<cplx> modal(tr.modes(), cplx(0.0, 0.0));
      for (int s = 0; s < n_seeds; ++s)
        if (ms[s] == m) modal[ells[s] - tr.ell_min] += cplx(amp_re[s], amp_im[s]);
      std::vector<cplx> ang(T, cplx(0.0, 0.0));
      for (int t = 0; t < T; ++t)
        for (int k = 0; k < tr.modes(); ++k) ang[t] += tr.S[t * tr.modes() + k] * modal[k];
      for (int r = 0; r < N; ++r) {
        const double pr = profile(0.0 + p->h * static_cast<double>(r));
        for (int t = 0; t < T; ++t) p->state[p->idx(mode, 2, r, t)] = pr * ang[t];
      }
    }
    // Q = D psi with the selected scheme (pipeline_initial_data.hpp:277-294).
    for (int mode = 0; mode < M; ++mode)
      for (int t = 0; t < T; ++t)
        for (int r = 0; r < N; ++r)
          p->state[p->idx(mode, 1, r, t)] = derivative_at(
              p->p.scheme, &p->state[p->idx(mode, 2, 0, t)], N, r, p->inv_h, T);
    // Galerkin P (pipeline_initial_data.hpp:142-203).
    double max_res = 0.0;
    bool singular = false;
    for (int mode = 0; mode < M; ++mode) {
      if (!p->parent_active[mode]) continue;
      const int m = p->modes[mode];
      const Transform tr = make_transform(-2, m, E, T);
      const int nm = tr.modes();
#pragma omp parallel for schedule(dynamic) reduction(max : max_res)
      for (int r = 0; r < N; ++r) {
        std::vector<cplx> mat(static_cast<size_t>(nm) * nm, 0.0), rhs(nm, 0.0), sol;
        for (int t = 0; t < T; ++t) {
          const Coeffs c = teukolsky_coefficients(p->p.mass, p->p.spin, p->p.length, -2,
                                                  m, 0.0 + p->h * r, p->theta[t]);
          const cplx inv_time = 1.0 / c.time;
          const cplx nodal = inv_time * (-2.0 * c.adv * p->state[p->idx(mode, 1, r, t)] +
                                         c.definition * p->state[p->idx(mode, 2, r, t)]);
          for (int a = 0; a < nm; ++a) {
            rhs[a] += tr.A[a * T + t] * nodal;
            for (int b = 0; b < nm; ++b)
              mat[a * nm + b] += tr.A[a * T + t] * inv_time * tr.S[t * nm + b];
          }
        }
        if (!dense_solve(nm, mat, rhs, sol)) {
#pragma omp critical
          singular = true;
          continue;
        }
        max_res = std::max(max_res, dense_residual(nm, mat, sol, rhs));
        for (int t = 0; t < T; ++t) {
          cplx v(0.0, 0.0);
          for (int k = 0; k < nm; ++k) v += tr.S[t * nm + k] * sol[k];
          p->state[p->idx(mode, 0, r, t)] = v;
        }
      }
    }
    if (singular) throw std::runtime_error("singular Galerkin initialization matrix");
    // Band projection of every field (pipeline_bands.hpp:175-202): fields
    // 0..9 on parents at ell_max_first, 10..12 on targets at ell_max_second,
    // inactive (mode, field) pairs zeroed.
    static const int spins[kFields] = {-2, -2, -2, -1, -2, 0, -2, -1, -1, 0, -2, -2, -2};
    for (int mode = 0; mode < M; ++mode) {
      const int m = p->modes[mode];
      for (int f = 0; f < kFields; ++f) {
        const bool active = f >= 10 ? contains(p->targets, m) : contains(p->parents, m);
        const int emax = f >= 10 ? p->p.ell_max_second : p->p.ell_max_first;
        if (!active) {
          for (int r = 0; r < N; ++r)
            for (int t = 0; t < T; ++t) p->state[p->idx(mode, f, r, t)] = 0.0;
          continue;
        }
        const Transform tr = make_transform(spins[f], m, emax, T);
        const int nm = tr.modes();
        for (int r = 0; r < N; ++r) {
          std::vector<cplx> modal(nm, cplx(0.0, 0.0));
          cplx* v = &p->state[p->idx(mode, f, r, 0)];
          for (int k = 0; k < nm; ++k)
            for (int t = 0; t < T; ++t) modal[k] += tr.A[k * T + t] * v[t];
          for (int t = 0; t < T; ++t) {
            cplx s(0.0, 0.0);
            for (int k = 0; k < nm; ++k) s += tr.S[t * nm + k] * modal[k];
            v[t] = s;
          }
        }
      }
    }
    if (!std::isfinite(max_res) || max_res > 2.0e-12)
      throw std::runtime_error("Galerkin residual check failed");
    return max_res;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1.0;
  }
}

void oracle_evaluate_rhs(OraclePipeline* p, const double* stage, double* out) {
  set_threads(p);
  evaluate_rhs(p, reinterpret_cast<const cplx*>(stage), reinterpret_cast<cplx*>(out));
}

int oracle_step(OraclePipeline* p, double time, double dt) {
  try {
    set_threads(p);
    pipeline_step(p, time, dt);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

int oracle_advance(OraclePipeline* p, int steps, double time0, double dt) {
  try {
    set_threads(p);
    // time accumulates by addition, as the driver's loop does
    // (solver_driver.hpp:289-293)
    double t = time0;
    for (int s = 0; s < steps; ++s) {
      pipeline_step(p, t, dt);
      t += dt;
    }
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

void oracle_get_activation(const OraclePipeline* p, OracleActivation* out) { *out = p->act; }

void oracle_set_activation(OraclePipeline* p, const OracleActivation* in) { p->act = *in; }

void oracle_constraint_norms(const OraclePipeline* p, OracleConstraintNorms* out) {
  *out = p->norms;
}

int oracle_consider_activation(OraclePipeline* p, double time) {
  try {
    set_threads(p);
    return consider_activation(p, time) ? 1 : 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

int oracle_last_stage_sources(const OraclePipeline* p, int* flags) {
  for (int i = 0; i < p->stage_source_count; ++i) flags[i] = p->stage_sources[i];
  return p->stage_source_count;
}

long long oracle_intermediate(const OraclePipeline* p, int which, double* out) {
  if (!p->second) return -1;
  const std::vector<cplx>* v = nullptr;
  switch (which) {
    case 0: v = &p->forcing; break;
    case 1: v = &p->source_over_r3; break;
    case 2: v = &p->proj; break;
    case 3: v = &p->proj_t; break;
    case 4: v = &p->constraints; break;
    case 5: v = &p->terms; break;
    case 6: v = &p->rdr; break;
    case 7: v = &p->rang; break;
    case 8: v = &p->inner; break;
    default: return -1;
  }
  if (out) std::memcpy(static_cast<void*>(out), static_cast<const void*>(v->data()), v->size() * sizeof(cplx));
  return static_cast<long long>(v->size());
}

int oracle_endpoint_modal(const OraclePipeline* p, int mode_index, double* scri_out,
                          double* horizon_out) {
  if (mode_index < 0 || mode_index >= p->M || !p->parent_active[mode_index]) return -1;
  const Transform& tr = p->first[mode_index];
  const int nm = tr.modes();
  for (int side = 0; side < 2; ++side) {
    const int r = side == 0 ? 0 : p->N - 1;
    double* out = side == 0 ? scri_out : horizon_out;
    for (int k = 0; k < nm; ++k) {
      cplx v(0.0, 0.0);
      for (int t = 0; t < p->T; ++t) v += tr.A[k * p->T + t] * p->state[p->idx(mode_index, 2, r, t)];
      out[2 * k] = v.real();
      out[2 * k + 1] = v.imag();
    }
  }
  return nm;
}

void oracle_gauss_legendre(int n, double* x, double* w) {
  std::vector<double> xx, ww;
  gauss_legendre(n, xx, ww);
  std::copy(xx.begin(), xx.end(), x);
  std::copy(ww.begin(), ww.end(), w);
}

double oracle_swsh_theta(int ell, int m, int spin, double theta) {
  return swsh(ell, m, spin, theta);
}

int oracle_transform_matrices(int spin, int m, int ell_max, int ntheta, double* S,
                              double* A) {
  try {
    const Transform t = make_transform(spin, m, ell_max, ntheta);
    std::copy(t.S.begin(), t.S.end(), S);
    std::copy(t.A.begin(), t.A.end(), A);
    return t.modes();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

void oracle_teukolsky_coefficients(double mass, double spin, double length,
                                   int spin_weight, int m, double radius,
                                   double theta, double* out) {
  const Coeffs c = teukolsky_coefficients(mass, spin, length, spin_weight, m, radius, theta);
  out[0] = c.time;
  out[1] = c.adv;
  out[2] = c.principal;
  out[3] = c.definition.real();
  out[4] = c.definition.imag();
  out[5] = c.q.real();
  out[6] = c.q.imag();
  out[7] = c.psi.real();
  out[8] = c.psi.imag();
}

double oracle_sbp_derivative_entry(int scheme, int n, int row, int column) {
  return derivative_entry(scheme, n, row, column);
}

double oracle_sbp_norm_weight(int scheme, int n, int index) {
  return norm_weight(scheme, n, index);
}

int oracle_sbp_minimum_points(int scheme) { return min_points(scheme); }

void oracle_sbp_derivative(int scheme, int n, double h, const double* in, double* out) {
  const cplx* v = reinterpret_cast<const cplx*>(in);
  cplx* o = reinterpret_cast<cplx*>(out);
  for (int r = 0; r < n; ++r) o[r] = derivative_at(scheme, v, n, r, 1.0 / h, 1);
}

void oracle_sbp_dissipation(int scheme, int n, double h, double strength,
                            const double* in, double* out) {
  const cplx* v = reinterpret_cast<const cplx*>(in);
  cplx* o = reinterpret_cast<cplx*>(out);
  for (int r = 0; r < n; ++r) o[r] = dissipation_at(scheme, v, n, r, h, strength, 1);
}

double oracle_dissipation_spectral_radius_bound(int scheme) {
  return scheme == 2 ? 5500.0 : (scheme == 1 ? 1371.0 : 64.0);
}

int oracle_dissipation_step_admissible(int scheme, double strength, double ratio) {
  if (!std::isfinite(strength) || strength < 0.0 || !std::isfinite(ratio) || ratio < 0.0)
    return 0;
  if (strength == 0.0) return 1;
  const double maximum = 0.8 * 2.785293563405282 /
                         (strength * oracle_dissipation_spectral_radius_bound(scheme));
  return ratio <= maximum ? 1 : 0;
}

double oracle_maximum_coordinate_speed(double mass, double spin, double length, int nr,
                                       int ntheta) {
  const double disc = mass * mass - spin * spin;
  const double rplus = mass + std::sqrt(disc > 0.0 ? disc : 0.0);
  const double horizon = length * length / rplus;
  const double h = horizon / static_cast<double>(nr - 1);
  std::vector<double> x, w;
  gauss_legendre(ntheta, x, w);
  double maximum = 0.0;
  for (int r = 0; r < nr; ++r)
    for (int t = 0; t < ntheta; ++t) {
      const Coeffs c = teukolsky_coefficients(mass, spin, length, -2, 0, 0.0 + h * r,
                                              std::acos(x[t]));
      const double root = std::sqrt(c.adv * c.adv + c.time * c.principal);
      const double lm = (c.adv - root) / c.time;
      const double lp = (c.adv + root) / c.time;
      maximum = std::max(maximum, std::max(std::abs(-lm), std::abs(-lp)));
    }
  return maximum;
}

double oracle_dense_solve(int n, const double* matrix, const double* rhs, double* solution) {
  std::vector<cplx> m(reinterpret_cast<const cplx*>(matrix),
                      reinterpret_cast<const cplx*>(matrix) + static_cast<size_t>(n) * n);
  std::vector<cplx> b(reinterpret_cast<const cplx*>(rhs), reinterpret_cast<const cplx*>(rhs) + n);
  std::vector<cplx> x;
  if (!dense_solve(n, m, b, x)) return -1.0;
  std::copy(x.begin(), x.end(), reinterpret_cast<cplx*>(solution));
  return dense_residual(n, m, x, b);
}

void oracle_background(double mass, double spin, double length, double radius,
                       double cos_theta, double sin_theta, double* out) {
  const Background b = kerr_background(mass, spin, length, radius, cos_theta, sin_theta);
  const cplx v[6] = {b.mu0, b.tau0, b.pi0, b.psi20, b.rho0, b.epsilon0};
  for (int k = 0; k < 6; ++k) {
    out[2 * k] = v[k].real();
    out[2 * k + 1] = v[k].imag();
  }
}

void oracle_pair_source(double radius, const double* bg, const double* fields,
                        const double* derivs, double* out) {
  Background b{};
  b.mu0 = cplx(bg[0], bg[1]);
  b.tau0 = cplx(bg[2], bg[3]);
  b.pi0 = cplx(bg[4], bg[5]);
  auto jet = [](const double* v, int k) {
    return Jet{cplx(v[4 * k], v[4 * k + 1]), cplx(v[4 * k + 2], v[4 * k + 3])};
  };
  PairFields f;
  Jet* fp = &f.F1;
  for (int k = 0; k < 19; ++k) fp[k] = jet(fields, k);
  PairDerivs d;
  Jet* dp = &d.d1F2;
  for (int k = 0; k < 11; ++k) dp[k] = jet(derivs, k);
  Jet D, T;
  pair_source(radius, b, f, d, D, T);
  const cplx o[4] = {D.v, D.d, T.v, T.d};
  for (int k = 0; k < 4; ++k) {
    out[2 * k] = o[k].real();
    out[2 * k + 1] = o[k].imag();
  }
}

int oracle_ghp_row(int spin, int boost, int m, int ell_max, int ntheta, int prime,
                   double kerr_spin, double length, double radius, const double* x,
                   const double* dt, double* out) {
  try {
    OraclePipeline p;
    p.p.spin = kerr_spin;
    p.p.length = length;
    p.M = 1;
    p.N = 1;
    p.T = ntheta;
    p.h = radius;
    p.modes = {m};
    std::vector<double> xs, ws;
    gauss_legendre(ntheta, xs, ws);
    p.cos_t = xs;
    p.sin_t.resize(ntheta);
    for (int t = 0; t < ntheta; ++t) p.sin_t[t] = std::sqrt(1.0 - xs[t] * xs[t]);
    const Coord c = make_coord(&p, spin, boost, ell_max, {m}, true);
    // radius enters as h * r with r = 0 on a one-point grid: use h = radius, r = 1
    p.N = 2;
    std::vector<cplx> xv(2 * ntheta, 0.0), dv(2 * ntheta, 0.0), ov(2 * ntheta, 0.0);
    for (int t = 0; t < ntheta; ++t) {
      xv[ntheta + t] = cplx(x[2 * t], x[2 * t + 1]);
      dv[ntheta + t] = cplx(dt[2 * t], dt[2 * t + 1]);
    }
    ghp(&p, c, prime != 0, Plane{xv.data(), 0}, Plane{dv.data(), 0},
        PlaneW{ov.data(), 0});
    for (int t = 0; t < ntheta; ++t) {
      out[2 * t] = ov[ntheta + t].real();
      out[2 * t + 1] = ov[ntheta + t].imag();
    }
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

namespace {
cplx cz(const double* v, int k = 0) { return cplx(v[2 * k], v[2 * k + 1]); }
void put(double* out, int k, cplx v) {
  out[2 * k] = v.real();
  out[2 * k + 1] = v.imag();
}
Background background_from(const double* bg, int count) {
  Background b{};
  cplx* slots[6] = {&b.mu0, &b.tau0, &b.pi0, &b.psi20, &b.rho0, &b.epsilon0};
  for (int k = 0; k < count; ++k) *slots[k] = cz(bg, k);
  return b;
}
}  // namespace

void oracle_eth_point(int prime, int spin, int boost, double kerr_spin, double length,
                      double radius, double cos_theta, double sin_theta, const double* x,
                      const double* dt, const double* ang, double* out) {
  put(out, 0, eth_point(prime != 0, spin, boost, kerr_spin, length, radius, cos_theta, sin_theta,
                        cz(x), cz(dt), cz(ang)));
}

void oracle_recon_rhs(double radius, const double* bg, const double* in, double* out) {
  const Background b = background_from(bg, 3);
  enum { F, G, Lam, H, B, Pi, C, U, Pis, Bs, Cs, e1F, e2G, e2C, e2Pi, ep1Bs, ep2Cs };
  auto v = [&](int k) { return cz(in, k); };
  put(out, 0, recon_rhs_G(radius, b, v(G), v(e1F), v(F)));
  put(out, 1, recon_rhs_Lambda(radius, b, v(Lam), v(F)));
  put(out, 2, recon_rhs_H(radius, b, v(H), v(e2G), v(G)));
  put(out, 3, recon_rhs_B(radius, b, v(B), v(Lam)));
  put(out, 4, recon_rhs_Pi(radius, b, v(G), v(Lam), v(B)));
  put(out, 5, recon_rhs_C(radius, b, v(C), v(Pi), v(B)));
  put(out, 6, recon_rhs_U(radius, b, v(U), v(H), v(Pi), v(C), v(Pis), v(Bs), v(Cs), v(e2C),
                          v(e2Pi), v(ep1Bs), v(ep2Cs)));
}

void oracle_outer_source_point(double radius, const double* bg, const double* in, double* out) {
  put(out, 0, outer_source_point(radius, background_from(bg, 3), cz(in, 0), cz(in, 1), cz(in, 2),
                                 cz(in, 3)));
}

void oracle_coordinate_forcing(double length, double kerr_spin, double radius, double cos_theta,
                               const double* source, double* out) {
  put(out, 0, coordinate_forcing_point(length, kerr_spin, radius, cos_theta, cz(source)));
}

void oracle_constraint_terms(double radius, const double* bg, const double* in, double* out) {
  const Background b = background_from(bg, 6);
  cplx tv[6];
  constraint_terms_point(radius, b, cz(in, 0), cz(in, 1), cz(in, 2), cz(in, 3), cz(in, 4),
                         cz(in, 5), cz(in, 6), cz(in, 7), cz(in, 8), cz(in, 9), cz(in, 10),
                         cz(in, 11), cz(in, 12), tv);
  for (int k = 0; k < 6; ++k) put(out, k, tv[k]);
}

void oracle_thorn(const double* value, const double* dt, const double* dr, int n, int spin,
                  int boost, int m, double radius, double cos_theta, double mass,
                  double kerr_spin, double length, const double* epsilon0, double* out) {
  put(out, 0, thorn_n(cz(value), cz(dt), cz(dr), n, spin, boost, m, radius, cos_theta, mass,
                      kerr_spin, length, cz(epsilon0)));
}

void oracle_field_weights(int* out) {
  const FieldWeight w[10] = {weight::F, weight::G,  weight::H, weight::Lambda, weight::Pi,
                             weight::B, weight::C, weight::U, weight::D,      weight::T};
  for (int k = 0; k < 10; ++k) {
    out[3 * k] = w[k].spin;
    out[3 * k + 1] = w[k].boost;
    out[3 * k + 2] = w[k].falloff;
  }
}

void oracle_delta_n(const double* value, const double* dt, const double* dr, int n,
                    double radius, double mass, double length, double* out) {
  const cplx r = delta_n(cplx(value[0], value[1]), cplx(dt[0], dt[1]), cplx(dr[0], dr[1]), n,
                         radius, mass, length);
  out[0] = r.real();
  out[1] = r.imag();
}

void oracle_recon_dt(const double* value, const double* dr, const double* rhs, int n,
                     double radius, double mass, double length, double* out) {
  const cplx r = recon_dt(cplx(value[0], value[1]), cplx(dr[0], dr[1]), cplx(rhs[0], rhs[1]), n,
                          radius, mass, length);
  out[0] = r.real();
  out[1] = r.imag();
}

}  // extern "C"
