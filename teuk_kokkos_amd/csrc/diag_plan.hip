// Plan::sample_diagnostics: the reference's sampled diagnostics of one
// instance (solver_driver.hpp:205-242) from one RHS evaluation of its
// current state.
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "diag_layout.hpp"
#include "plan.hpp"
#include "plan_detail.hpp"
#include "so_host.hpp"

namespace teuk {

using detail::check;
using detail::device_alloc;

namespace {

// Backward endpoint stencil for the d-th radial derivative on 9 points
// (horizon_diagnostics.hpp:18-79): the moment equations
//   sum_j w_j (-j)^k = k! delta_{kd},  k = 0..8,
// solved by Gauss-Jordan elimination with partial pivoting in long double,
// then scaled by h^-d.
std::array<double, kHorizonStencil> horizon_weights(int d, double h) {
  constexpr int n = kHorizonStencil;
  long double m[n][n + 1];
  long double fact = 1.0L;
  for (int k = 2; k <= d; ++k) fact *= static_cast<long double>(k);
  for (int k = 0; k < n; ++k) {
    for (int j = 0; j < n; ++j) {
      long double v = 1.0L;
      for (int e = 0; e < k; ++e) v *= -static_cast<long double>(j);
      m[k][j] = v;
    }
    m[k][n] = k == d ? fact : 0.0L;
  }
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    if (std::fabs(m[p][c]) < 1.0e-28L) throw std::runtime_error("singular horizon stencil moment matrix");
    if (p != c)
      for (int e = 0; e <= n; ++e) std::swap(m[p][e], m[c][e]);
    const long double diag = m[c][c];
    for (int e = c; e <= n; ++e) m[c][e] /= diag;
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const long double f = m[r][c];
      for (int e = c; e <= n; ++e) m[r][e] -= f * m[c][e];
    }
  }
  std::array<double, n> w{};
  const long double scale = std::pow(static_cast<long double>(h), -d);
  for (int j = 0; j < n; ++j) w[j] = static_cast<double>(m[j][n] * scale);
  return w;
}

}  // namespace

// The diagnostics tables and reduction buffers of the largest instance, and
// for second-order plans the per-pair source buffer, made at plan creation.
void Plan::allocate_diagnostics() {
  long long waves = 1;
  int m_cap = 1, t_cap = 1;
  long long pair_values = 1;
  for (const Instance& in : inst_) {
    waves = std::max<long long>(waves, diag_waves(in.M, in.N, in.T));
    m_cap = std::max(m_cap, in.M);
    t_cap = std::max(t_cap, in.T);
    pair_values = std::max<long long>(
        pair_values, static_cast<long long>(in.registry.ordered_pairs().size()) * 2 * in.N * in.T);
  }
  d_diag_sharp_ = device_alloc<int>(m_cap);
  d_diag_nodes_ = device_alloc<double>(2 * t_cap);
  d_diag_partial_ = device_alloc<double>(static_cast<size_t>(waves) * 2 * kDqCount);
  d_diag_pflags_ = device_alloc<int>(waves);
  d_diag_out_ = device_alloc<double>(2 * kDqCount + 1);
  diag_waves_cap_ = waves;
  diag_m_cap_ = m_cap;
  diag_t_cap_ = t_cap;
  if (so_mode_) {
    SoHost& H = *so_;
    H.pairs = device_alloc<cdouble>(pair_values);
    H.pairs_cap = pair_values;
  }
}

void Plan::sample_diagnostics(int i, double time, double* out) {
  if (!std::isfinite(time) || time < 0.0)
    throw std::invalid_argument("diagnostics need a finite, nonnegative sample time");
  Instance& in = inst_.at(i);
  const int M = in.M, N = in.N, T = in.T;
  if (N < kHorizonStencil)
    throw std::invalid_argument("horizon diagnostics require nine radial points");
  const long long count = nodal_value_count(i);
  DiagArgs a{};
  std::vector<int> sharp(M, -1);
  for (int mi = 0; mi < M; ++mi) {
    const int m = in.registry.modes()[mi];
    for (int mj = 0; mj < M; ++mj)
      if (in.registry.modes()[mj] == -m) sharp[mi] = mj;
  }
  std::vector<double> nodes(2 * T);
  {
    std::vector<double> x, w;
    gauss_legendre(T, x, w);
    for (int t = 0; t < T; ++t) {
      nodes[t] = x[t];
      nodes[T + t] = std::sqrt(1.0 - x[t] * x[t]);
    }
  }
  const int waves = diag_waves(M, N, T);
  // tables and reduction buffers made at plan creation (allocate_diagnostics)
  if (waves > diag_waves_cap_ || M > diag_m_cap_ || T > diag_t_cap_)
    throw std::logic_error("diagnostics buffers smaller than the instance");
  int* d_sharp = d_diag_sharp_;
  double* d_nodes = d_diag_nodes_;
  double* d_partial = d_diag_partial_;
  int* d_pflags = d_diag_pflags_;
  double* d_out = d_diag_out_;
  std::vector<double> red(2 * kDqCount + 1);
  std::vector<cplx> tail(static_cast<size_t>(M) * 2 * kHorizonStencil * T);
  std::vector<cplx> pairs_host;
  {
    check(hipMemcpy(d_sharp, sharp.data(), M * sizeof(int), hipMemcpyHostToDevice), "sharp");
    check(hipMemcpy(d_nodes, nodes.data(), 2 * T * sizeof(double), hipMemcpyHostToDevice), "nodes");
    const cdouble* state = nullptr;
    const cdouble* rhs = nullptr;
    if (so_mode_) {
      SoHost& H = *so_;
      const SoInstance& I = H.inst.at(i);
      so_upload_active();
      const long long pair_values =
          static_cast<long long>(in.registry.ordered_pairs().size()) * 2 * N * T;
      if (pair_values > H.pairs_cap)
        throw std::logic_error("pair report buffer smaller than the instance");
      H.capture_pairs = true;
      try {
        so_rhs(i, 1, H.y, H.k, true);
      } catch (...) {
        H.capture_pairs = false;
        throw;
      }
      H.capture_pairs = false;
      pairs_host.resize(static_cast<size_t>(pair_values));
      check(hipMemcpyAsync(pairs_host.data(), H.pairs, pair_values * sizeof(cdouble),
                           hipMemcpyDeviceToHost, stream_),
            "read pair sources");
      state = H.y + I.base;
      rhs = H.k + I.base;
      a.src = H.src + I.pbase * kSlotsSrc;
      a.cons = H.cons + I.pbase * kSlotsCons;
      a.rdr = H.rdr + I.pbase * kSlotsRdr;
      a.rang = H.rang + I.pbase * kSlotsRang;
      a.rangx = H.rangx + I.pbase * kSlotsRangX;
      a.active = in.activation.active ? 1 : 0;
      a.table = H.d_tables + static_cast<int>(in.params.scheme);
    } else {
      // the nodal state (export) and its reference-layout RHS, staged in the
      // import buffers
      if (count > imp_cap_) throw std::logic_error("diagnostics staging smaller than the instance");
      std::vector<cplx> host(static_cast<size_t>(count));
      export_state(i, host.data());
      check(hipMemcpy(d_imp_in_, host.data(), count * sizeof(cdouble), hipMemcpyHostToDevice),
            "stage state");
      evaluate_rhs(i, reinterpret_cast<const cplx*>(d_imp_in_), reinterpret_cast<cplx*>(d_imp_res_),
                   time);
      state = d_imp_in_;
      rhs = d_imp_res_;
      a.table = d_tables_ + static_cast<int>(in.params.scheme);
    }
    a.state = state;
    a.rhs = rhs;
    a.sharp = d_sharp;
    a.nodes = d_nodes;
    a.mass = in.params.mass;
    a.spin = in.params.spin;
    a.length = in.params.compactification_length;
    a.h = in.h;
    a.inv_h = 1.0 / in.h;
    a.scheme = static_cast<int>(in.params.scheme);
    a.M = M;
    a.N = N;
    a.T = T;
    a.partial = d_partial;
    a.partial_flags = d_pflags;
    a.out = d_out;
    a.n_waves = waves;
    check(launch_diagnostics(a, stream_), "diagnostics");
    check(hipMemcpyAsync(red.data(), d_out, red.size() * sizeof(double), hipMemcpyDeviceToHost,
                         stream_),
          "read diagnostics");
    // the last nine radial rows of FirstPsi and SecondPsi for the horizon
    // derivatives
    for (int mi = 0; mi < M; ++mi)
      for (int k = 0; k < 2; ++k) {
        const int f = k == 0 ? 2 : 12;
        const long long off = ((static_cast<long long>(mi) * 13 + f) * N + (N - kHorizonStencil)) * T;
        check(hipMemcpyAsync(tail.data() + (static_cast<size_t>(mi) * 2 + k) * kHorizonStencil * T,
                             state + off, static_cast<size_t>(kHorizonStencil) * T * sizeof(cdouble),
                             hipMemcpyDeviceToHost, stream_),
              "read horizon rows");
      }
    check(hipStreamSynchronize(stream_), "diagnostics");
  }

  const int flags = static_cast<int>(red[2 * kDqCount]);
  const bool pipeline_bad = flags & (kDfState | kDfRhs | kDfFirstConstraint | kDfSecondConstraint |
                                     kDfSource | kDfForcing | kDfScri | kDfHorizon);
  if (pipeline_bad) {
    // pipeline_diagnostics.hpp:397-414
    std::string msg = "pipeline diagnostics found nonfinite values in";
    if (flags & kDfState) msg += " state";
    if (flags & kDfRhs) msg += " rhs";
    if (flags & kDfFirstConstraint) msg += " first-constraint";
    if (flags & kDfSecondConstraint) msg += " second-constraint";
    if (flags & kDfSource) msg += " source";
    if (flags & kDfForcing) msg += " forcing";
    if (flags & kDfScri) msg += " scri-endpoint";
    if (flags & kDfHorizon) msg += " horizon-endpoint";
    throw std::runtime_error(msg);
  }
  if (flags & kDfTransport) throw std::runtime_error("nonfinite reconstruction residual");
  if (flags & kDfIndependent)
    throw std::runtime_error("nonfinite independent reconstruction constraint");

  const double points = static_cast<double>(M) * N * T;
  auto rms = [&](int q) { return std::sqrt(red[q] / points); };
  auto mx = [&](int q) { return std::sqrt(red[kDqCount + q]); };
  for (int f = 0; f < 13; ++f) {
    out[kDiagStateRms + f] = rms(kDqState + f);
    out[kDiagStateMax + f] = mx(kDqState + f);
    out[kDiagRhsRms + f] = rms(kDqRhs + f);
    out[kDiagRhsMax + f] = mx(kDqRhs + f);
  }
  const int pairs[4] = {kDqFirstConstraint, kDqSecondConstraint, kDqSource, kDqForcing};
  for (int k = 0; k < 4; ++k) {
    out[kDiagFirstConstraint + 2 * k] = rms(pairs[k]);
    out[kDiagFirstConstraint + 2 * k + 1] = mx(pairs[k]);
  }
  // independent constraints and their combination
  // (pipeline_independent_reconstruction_diagnostics.hpp:80-97)
  double csum = 0.0, cmax = 0.0;
  for (int k = 0; k < 3; ++k) {
    out[kDiagIndependent + 2 * k] = rms(kDqIndependent + k);
    out[kDiagIndependent + 2 * k + 1] = mx(kDqIndependent + k);
    csum += red[kDqIndependent + k];
    cmax = std::max(cmax, red[kDqCount + kDqIndependent + k]);
  }
  out[kDiagIndependent + 6] = std::sqrt(csum / (3.0 * points));
  out[kDiagIndependent + 7] = std::sqrt(cmax);
  // transport residuals (pipeline_reconstruction_diagnostics.hpp:170-190)
  csum = cmax = 0.0;
  for (int k = 0; k < 7; ++k) {
    out[kDiagTransport + 2 * k] = rms(kDqTransport + k);
    out[kDiagTransport + 2 * k + 1] = mx(kDqTransport + k);
    csum += red[kDqTransport + k];
    cmax = std::max(cmax, red[kDqCount + kDqTransport + k]);
  }
  out[kDiagTransport + 14] = std::sqrt(csum / (7.0 * points));
  out[kDiagTransport + 15] = std::sqrt(cmax);
  // horizon transverse derivatives (horizon_diagnostics.hpp:113-160,
  // solver_driver.hpp:174-192): largest |d^k field / dR^k| at R_H
  for (int k = 0; k < 2; ++k) {
    double* o = out + (k == 0 ? kDiagHorizonFirst : kDiagHorizonSecond);
    for (int d = 0; d < kHorizonDerivatives; ++d) {
      const auto w = horizon_weights(d, in.h);
      double best = 0.0;
      for (int mi = 0; mi < M; ++mi) {
        const cplx* rows = tail.data() + (static_cast<size_t>(mi) * 2 + k) * kHorizonStencil * T;
        for (int t = 0; t < T; ++t) {
          cplx v(0.0, 0.0);
          for (int p = 0; p < kHorizonStencil; ++p)
            v += w[p] * rows[static_cast<size_t>(kHorizonStencil - 1 - p) * T + t];
          best = std::max(best, std::abs(v));
        }
      }
      o[d] = best;
    }
  }
  // the independent constraint maximum of the last activation check, the
  // latch (spatial_pipeline.hpp:1264-1276, pipeline_diagnostics.hpp:293-305)
  double norms[14] = {};
  if (so_mode_) constraint_norms(i, norms);
  out[kDiagSourceConstraintMax] = std::max({norms[0], norms[4], norms[8]});
  out[kDiagSourceActive] = in.activation.active ? 1.0 : 0.0;
  out[kDiagActivationTime] = in.activation.activation_time;
  // per ordered pair: RMS and maximum of |D| and |T| over (r, theta)
  // (solver_driver.hpp:251-280)
  in.pair_report.clear();
  if (so_mode_) {
    const auto& pairs = in.registry.ordered_pairs();
    const size_t plane = static_cast<size_t>(N) * T;
    const double inverse_points = 1.0 / static_cast<double>(plane);
    for (size_t k = 0; k < pairs.size(); ++k) {
      double sq[2] = {0.0, 0.0}, mx2[2] = {0.0, 0.0};
      for (int c = 0; c < 2; ++c)
        for (size_t p = 0; p < plane; ++p) {
          const double v = std::abs(pairs_host[(2 * k + c) * plane + p]);
          sq[c] += v * v;
          mx2[c] = std::max(mx2[c], v);
        }
      in.pair_report.insert(in.pair_report.end(),
                            {static_cast<double>(pairs[k].m1), static_cast<double>(pairs[k].m2),
                             static_cast<double>(pairs[k].target), std::sqrt(sq[0] * inverse_points),
                             mx2[0], std::sqrt(sq[1] * inverse_points), mx2[1]});
    }
  }
  out[kDiagScriFinite] = 1.0;
  out[kDiagHorizonFinite] = 1.0;
  out[kDiagAllFinite] = 1.0;
}

}  // namespace teuk
