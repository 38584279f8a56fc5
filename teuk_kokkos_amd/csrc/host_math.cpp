#include "host_math.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <string>

namespace teuk {

namespace {
#include "sbp_coefficients.inc"

constexpr double kPi = std::numbers::pi_v<double>;

double rat(const SbpRational& r) { return r.num / r.den; }

void legendre_pair(int n, long double x, long double& pn, long double& pm1) {
  long double a = 1.0L;
  if (n == 0) {
    pn = a;
    pm1 = 0.0L;
    return;
  }
  long double b = x;
  for (int l = 2; l <= n; ++l) {
    const long double c = ((2.0L * l - 1.0L) * x * b - (l - 1.0L) * a) / l;
    a = b;
    b = c;
  }
  pn = b;
  pm1 = a;
}

long double lfact(int n) { return std::lgammal(static_cast<long double>(n) + 1.0L); }

long double power(long double base, int e) {
  long double r = 1.0L;
  while (e) {
    if (e & 1) r *= base;
    base *= base;
    e >>= 1;
  }
  return r;
}

double wigner_small_d(int l, int mp, int m, double theta) {
  const int kmin = std::max(0, m - mp), kmax = std::min(l + m, l - mp);
  const long double norm =
      0.5L * (lfact(l + m) + lfact(l - m) + lfact(l + mp) + lfact(l - mp));
  const long double c = std::cos(0.5L * static_cast<long double>(theta));
  const long double s = std::sin(0.5L * static_cast<long double>(theta));
  long double sum = 0.0L;
  for (int k = kmin; k <= kmax; ++k) {
    long double t = std::exp(norm - lfact(l + m - k) - lfact(k) - lfact(mp - m + k) -
                             lfact(l - mp - k)) *
                    power(c, 2 * l + m - mp - 2 * k) * power(s, mp - m + 2 * k);
    if ((mp - m + k) & 1) t = -t;
    sum += t;
  }
  return static_cast<double>(sum);
}

// Explicit A^T A rows of one scheme's compatible dissipation, using the
// reference's truncation of the difference rows at both ends
// (radial_discretization.hpp:165-232).
void build_dissipation(SbpScheme& s, const std::vector<double>& stencil) {
  const int order = s.diss_order;
  auto coefficient = [&](int n, int index, int col) {
    const int first = index > order ? index - order : 0;
    const int last = index < n - order ? index : n - order - 1;
    double v = 0.0;
    for (int row = first; row <= last; ++row) {
      const int pi = index - row, pc = col - row;
      if (pc >= 0 && pc <= order) v += stencil[pi] * stencil[pc];
    }
    return v;
  };
  const int n = 8 * order + 8;  // large enough to isolate both ends
  s.diss_left.assign(order, {});
  for (int d = 0; d < order; ++d) {
    s.diss_left[d].resize(d + order + 1);
    for (int c = 0; c <= d + order; ++c) s.diss_left[d][c] = coefficient(n, d, c);
  }
  s.diss_interior.resize(2 * order + 1);
  const int mid = n / 2;
  for (int k = -order; k <= order; ++k) s.diss_interior[k + order] = coefficient(n, mid, mid + k);
}

SbpScheme make_d42() {
  SbpScheme s;
  s.closure_rows = 4;
  s.closure_width = 6;
  s.half_width = 2;
  s.diss_order = 3;
  s.left.resize(4 * 6);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 6; ++c) s.left[r * 6 + c] = rat(kD42Left[r][c]);
  s.interior = {0.0, 2.0 / 3.0, -1.0 / 12.0};
  for (int d = 0; d < 4; ++d) s.norm.push_back(rat(kD42Norm[d]));
  build_dissipation(s, {-1, 3, -3, 1});
  return s;
}

SbpScheme make_d84() {
  SbpScheme s;
  s.closure_rows = 8;
  s.closure_width = 12;
  s.half_width = 4;
  s.diss_order = 5;
  s.left.resize(8 * 12);
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 12; ++c) s.left[r * 12 + c] = rat(kD84Left[r][c]);
  s.interior = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};
  for (int d = 0; d < 8; ++d) s.norm.push_back(rat(kD84Norm[d]));
  build_dissipation(s, {-1, 5, -10, 10, -5, 1});
  return s;
}

SbpScheme make_d105() {
  SbpScheme s;
  s.closure_rows = 11;
  s.closure_width = 16;
  s.half_width = 5;
  s.diss_order = 6;
  s.interior = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0};
  for (int d = 0; d < 11; ++d) s.norm.push_back(rat(kD105Norm[d]));
  // Q = H D; the closure row is Q(row, :) / H(row) (sbp105.hpp:209-213).
  auto q = [&](int r, int c) -> double {
    if (c >= 11) {
      const int off = c - r;
      if (off == 0 || off < -5 || off > 5) return 0.0;
      return off > 0 ? s.interior[off] : -s.interior[-off];
    }
    if (r == 0 && c == 0) return -0.5;
    if (r == c) return 0.0;
    if (c > r) return rat(kD105QUpper[r][c - r - 1]);
    return -rat(kD105QUpper[c][r - c - 1]);
  };
  s.left.resize(11 * 16);
  for (int r = 0; r < 11; ++r)
    for (int c = 0; c < 16; ++c) s.left[r * 16 + c] = q(r, c) / s.norm[r];
  build_dissipation(s, {1, -6, 15, -20, 15, -6, 1});
  return s;
}

}  // namespace

// ------------------------------------------------------------ angular
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  if (n <= 0) throw std::invalid_argument("Gauss-Legendre node count must be positive");
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const long double tol = 8.0L * std::numeric_limits<long double>::epsilon();
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = std::cos(std::numbers::pi_v<long double> * (i + 0.75L) / (n + 0.5L));
    long double pn = 0, pm1 = 0, dp = 0;
    for (int it = 0;; ++it) {
      legendre_pair(n, z, pn, pm1);
      dp = n * (z * pn - pm1) / (z * z - 1.0L);
      const long double step = pn / dp;
      z -= step;
      if (std::abs(step) <= tol * (1.0L + std::abs(z))) break;
      if (it == 99) throw std::runtime_error("Gauss-Legendre root iteration failed");
    }
    legendre_pair(n, z, pn, pm1);
    dp = n * (z * pn - pm1) / (z * z - 1.0L);
    const long double weight = 2.0L / ((1.0L - z * z) * dp * dp);
    x[i] = static_cast<double>(-z);
    x[n - 1 - i] = static_cast<double>(z);
    w[i] = w[n - 1 - i] = static_cast<double>(weight);
  }
}

double swsh_theta(int ell, int m, int spin, double theta) {
  if (ell < 0 || std::abs(spin) > ell || std::abs(m) > ell)
    throw std::invalid_argument("spin-weighted mode requires ell >= |s|,|m|");
  const double phase = (std::abs(spin) & 1) ? -1.0 : 1.0;
  return phase * std::sqrt((2.0 * ell + 1.0) / (4.0 * kPi)) *
         wigner_small_d(ell, m, -spin, theta);
}

BandTransform make_band_transform(int spin, int m, int ell_max, int nodes) {
  BandTransform t;
  t.spin = spin;
  t.m = m;
  t.ell_min = minimum_ell(spin, m);
  t.ell_max = ell_max;
  t.nodes = nodes;
  if (ell_max < t.ell_min) throw std::invalid_argument("ell_max is below the minimum allowed ell");
  if (nodes < t.modes())
    throw std::invalid_argument("angular transform needs at least as many nodes as modes");
  std::vector<double> x, w;
  gauss_legendre(nodes, x, w);
  const int nm = t.modes();
  t.S.assign(static_cast<size_t>(nodes) * nm, 0.0);
  t.A.assign(static_cast<size_t>(nodes) * nm, 0.0);
  for (int j = 0; j < nodes; ++j) {
    const double theta = std::acos(x[j]);
    for (int k = 0; k < nm; ++k) {
      const double basis = swsh_theta(t.ell_min + k, m, spin, theta);
      t.S[static_cast<size_t>(j) * nm + k] = basis;
      t.A[static_cast<size_t>(k) * nodes + j] = 2.0 * kPi * w[j] * basis;
    }
  }
  t.G.assign(static_cast<size_t>(nm) * nm, 0.0);
  double dev = 0.0;
  for (int a = 0; a < nm; ++a)
    for (int b = 0; b < nm; ++b) {
      long double g = 0.0L;
      for (int j = 0; j < nodes; ++j)
        g += static_cast<long double>(t.A[static_cast<size_t>(a) * nodes + j]) *
             t.S[static_cast<size_t>(j) * nm + b];
      t.G[a * nm + b] = static_cast<double>(g);
      dev = std::max(dev, std::abs(t.G[a * nm + b] - (a == b ? 1.0 : 0.0)));
    }
  t.gram = dev > 1.0e-13;
  // B = G^-1 A by Gauss-Jordan elimination with partial pivoting.
  std::vector<double> g = t.G;
  t.B = t.A;
  for (int c = 0; c < nm; ++c) {
    int piv = c;
    for (int r = c + 1; r < nm; ++r)
      if (std::abs(g[r * nm + c]) > std::abs(g[piv * nm + c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < nm; ++k) std::swap(g[c * nm + k], g[piv * nm + k]);
      for (int j = 0; j < nodes; ++j) std::swap(t.B[c * nodes + j], t.B[piv * nodes + j]);
    }
    const double d = g[c * nm + c];
    for (int k = 0; k < nm; ++k) g[c * nm + k] /= d;
    for (int j = 0; j < nodes; ++j) t.B[c * nodes + j] /= d;
    for (int r = 0; r < nm; ++r) {
      if (r == c) continue;
      const double f = g[r * nm + c];
      if (f == 0.0) continue;
      for (int k = 0; k < nm; ++k) g[r * nm + k] -= f * g[c * nm + k];
      for (int j = 0; j < nodes; ++j) t.B[r * nodes + j] -= f * t.B[c * nodes + j];
    }
  }
  return t;
}

bool band_gram_needed(int spin, int m, int ell_max, int nodes) {
  return make_band_transform(spin, m, ell_max, nodes).gram;
}

std::vector<cplx> BandTransform::coefficients_of(const cplx* nodal) const {
  const std::vector<double>& M = gram ? B : A;
  std::vector<cplx> out(modes(), cplx(0.0, 0.0));
  for (int k = 0; k < modes(); ++k)
    for (int j = 0; j < nodes; ++j) out[k] += M[static_cast<size_t>(k) * nodes + j] * nodal[j];
  return out;
}

std::vector<cplx> BandTransform::gram_apply(const std::vector<cplx>& c) const {
  if (!gram) return c;
  std::vector<cplx> out(modes(), cplx(0.0, 0.0));
  for (int a = 0; a < modes(); ++a)
    for (int b = 0; b < modes(); ++b) out[a] += G[static_cast<size_t>(a) * modes() + b] * c[b];
  return out;
}

std::vector<cplx> BandTransform::analyze(const cplx* nodal) const {
  std::vector<cplx> out(modes(), cplx(0.0, 0.0));
  for (int k = 0; k < modes(); ++k)
    for (int j = 0; j < nodes; ++j) out[k] += A[static_cast<size_t>(k) * nodes + j] * nodal[j];
  return out;
}

std::vector<cplx> BandTransform::synthesize(const std::vector<cplx>& modal) const {
  std::vector<cplx> out(nodes, cplx(0.0, 0.0));
  for (int j = 0; j < nodes; ++j)
    for (int k = 0; k < modes(); ++k) out[j] += S[static_cast<size_t>(j) * modes() + k] * modal[k];
  return out;
}

// ------------------------------------------------------------ SBP
const SbpScheme& sbp_scheme(RadialScheme s) {
  static const SbpScheme d42 = make_d42();
  static const SbpScheme d84 = make_d84();
  static const SbpScheme d105 = make_d105();
  return s == RadialScheme::D105 ? d105 : (s == RadialScheme::D84 ? d84 : d42);
}

double sbp_derivative_entry(RadialScheme s, int n, int row, int col) {
  const SbpScheme& sc = sbp_scheme(s);
  const int W = sc.closure_rows, C = sc.closure_width;
  if (row < W) return (col >= 0 && col < C) ? sc.left[row * C + col] : 0.0;
  if (row + W >= n) {
    const int rr = n - 1 - row, rc = n - 1 - col;
    return (rc >= 0 && rc < C) ? -sc.left[rr * C + rc] : 0.0;
  }
  const int off = col - row;
  if (off == 0 || std::abs(off) > sc.half_width) return 0.0;
  return off > 0 ? sc.interior[off] : -sc.interior[-off];
}

double sbp_norm_weight(RadialScheme s, int n, int index) {
  const SbpScheme& sc = sbp_scheme(s);
  const int d = std::min(index, n - 1 - index);
  return d < static_cast<int>(sc.norm.size()) ? sc.norm[d] : 1.0;
}

std::pair<int, int> derivative_columns(RadialScheme s, int n, int row) {
  const SbpScheme& sc = sbp_scheme(s);
  if (row < sc.closure_rows) return {0, sc.closure_width};
  if (row + sc.closure_rows >= n) return {n - sc.closure_width, n};
  return {row - sc.half_width, row + sc.half_width + 1};
}

std::pair<int, int> dissipation_columns(RadialScheme s, int n, int row) {
  const int o = sbp_scheme(s).diss_order;
  return {std::max(0, row - o), std::min(n, row + o + 1)};
}

void sbp_derivative_line(RadialScheme s, int n, double inv_h, const cplx* in, size_t stride,
                         cplx* out, size_t out_stride) {
  const SbpScheme& sc = sbp_scheme(s);
  const int W = sc.closure_rows, C = sc.closure_width, B = sc.half_width;
  for (int r = 0; r < n; ++r) {
    cplx acc(0.0, 0.0);
    if (r < W) {
      for (int c = 0; c < C; ++c) acc += sc.left[r * C + c] * in[c * stride];
    } else if (r + W >= n) {
      const int rr = n - 1 - r;
      for (int c = 0; c < C; ++c) acc -= sc.left[rr * C + c] * in[(n - 1 - c) * stride];
    } else {
      for (int k = B; k >= 1; --k)
        acc += sc.interior[k] * (in[(r + k) * stride] - in[(r - k) * stride]);
    }
    out[r * out_stride] = inv_h * acc;
  }
}

// ------------------------------------------------------------ dense solve
bool dense_complex_solve(int n, std::vector<cplx> a, std::vector<cplx> b, std::vector<cplx>& x) {
  double scale = 0.0;
  for (const cplx& v : a) scale = std::max(scale, std::abs(v));
  const double floor = 128.0 * std::numeric_limits<double>::epsilon() * scale;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double best = std::abs(a[c * n + c]);
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[r * n + c]) > best) {
        best = std::abs(a[r * n + c]);
        piv = r;
      }
    if (!std::isfinite(best) || best <= floor) return false;
    if (piv != c) {
      for (int j = c; j < n; ++j) std::swap(a[c * n + j], a[piv * n + j]);
      std::swap(b[c], b[piv]);
    }
    const cplx d = a[c * n + c];
    for (int r = c + 1; r < n; ++r) {
      const cplx f = a[r * n + c] / d;
      a[r * n + c] = 0.0;
      for (int j = c + 1; j < n; ++j) a[r * n + j] -= f * a[c * n + j];
      b[r] -= f * b[c];
    }
  }
  x.assign(n, cplx(0.0, 0.0));
  for (int r = n - 1; r >= 0; --r) {
    cplx v = b[r];
    for (int c = r + 1; c < n; ++c) v -= a[r * n + c] * x[c];
    x[r] = v / a[r * n + r];
  }
  return true;
}

double dense_complex_residual(int n, const std::vector<cplx>& a, const std::vector<cplx>& x,
                              const std::vector<cplx>& b) {
  double rr = 0.0, bb = 0.0;
  for (int r = 0; r < n; ++r) {
    cplx res = -b[r];
    for (int c = 0; c < n; ++c) res += a[r * n + c] * x[c];
    rr += std::norm(res);
    bb += std::norm(b[r]);
  }
  return std::sqrt(rr) / std::max(std::sqrt(bb), std::numeric_limits<double>::min());
}

// ------------------------------------------------------------ modes
namespace {
std::vector<int> sorted_unique(std::vector<int> v, const char* what) {
  std::sort(v.begin(), v.end());
  if (std::adjacent_find(v.begin(), v.end()) != v.end())
    throw std::invalid_argument(std::string("duplicate ") + what);
  return v;
}
}  // namespace

ModeRegistry::ModeRegistry(std::vector<int> modes, std::vector<int> parents,
                           std::vector<int> targets)
    : modes_(sorted_unique(std::move(modes), "mode")),
      parents_(sorted_unique(std::move(parents), "parent mode")),
      targets_(sorted_unique(std::move(targets), "target mode")) {
  if (modes_.empty()) throw std::invalid_argument("ModeRegistry requires at least one mode");
  if (targets_.empty()) throw std::invalid_argument("ModeRegistry requires at least one target mode");
  if (parents_.empty()) throw std::invalid_argument("ModeRegistry requires at least one parent mode");
  for (int p : parents_)
    if (!contains(p)) throw std::invalid_argument("parent mode " + std::to_string(p) + " is not stored");
  for (int t : targets_)
    if (!contains(t)) throw std::invalid_argument("target mode " + std::to_string(t) + " is not stored");
  for (int t : targets_) {
    pair_offsets_.push_back(static_cast<int>(pairs_.size()));
    for (int m1 : parents_) {
      const long long m2w = static_cast<long long>(t) - m1;
      if (m2w < std::numeric_limits<int>::min() || m2w > std::numeric_limits<int>::max()) continue;
      const int m2 = static_cast<int>(m2w);
      if (!is_parent(m2)) continue;
      pairs_.push_back({m1, m2, t, index(m1), index(m2), index(t)});
    }
  }
  pair_offsets_.push_back(static_cast<int>(pairs_.size()));
}

bool ModeRegistry::contains(int m) const { return std::binary_search(modes_.begin(), modes_.end(), m); }
bool ModeRegistry::is_parent(int m) const {
  return std::binary_search(parents_.begin(), parents_.end(), m);
}
bool ModeRegistry::is_target(int m) const {
  return std::binary_search(targets_.begin(), targets_.end(), m);
}
int ModeRegistry::index(int m) const {
  auto it = std::lower_bound(modes_.begin(), modes_.end(), m);
  if (it == modes_.end() || *it != m)
    throw std::out_of_range("mode " + std::to_string(m) + " is not stored");
  return static_cast<int>(it - modes_.begin());
}
bool ModeRegistry::closed_under_sharp() const {
  for (int m : modes_)
    if (m == std::numeric_limits<int>::min() || !contains(-m)) return false;
  return true;
}
std::pair<int, int> ModeRegistry::pair_range(int target) const {
  auto it = std::lower_bound(targets_.begin(), targets_.end(), target);
  if (it == targets_.end() || *it != target)
    throw std::out_of_range("mode " + std::to_string(target) + " is not a target");
  const size_t i = static_cast<size_t>(it - targets_.begin());
  return {pair_offsets_[i], pair_offsets_[i + 1]};
}

ModeRegistry make_run_registry(const RunParameters& p) {
  std::vector<int> stored = p.first_order_modes;
  stored.insert(stored.end(), p.second_order_modes.begin(), p.second_order_modes.end());
  std::sort(stored.begin(), stored.end());
  stored.erase(std::unique(stored.begin(), stored.end()), stored.end());
  return ModeRegistry(stored, p.first_order_modes, p.second_order_modes);
}

}  // namespace teuk
