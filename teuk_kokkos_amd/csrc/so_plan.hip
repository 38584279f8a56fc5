// Plan methods of the second-order graph: operator tables, tiles, device
// buffers, initial data, import/export, the RK4 passes with the exact
// source-event split and the activation latch (spatial_pipeline.hpp:590-620,
// 1225-1310), constraint sampling and waveforms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <tuple>

#include "plan_detail.hpp"
#include "so_host.hpp"

namespace teuk {

hipError_t launch_fill_zero(cdouble* p, long long count, hipStream_t s);

using namespace detail;

namespace {

// Spin and boost of each coordinator (spatial_pipeline.hpp:185-218).
constexpr int kCoordSpin[kSoCoords] = {-2, -2, -1, -2, -1, -1, 0, 2, 1, -1};
constexpr int kCoordBoost[kSoCoords] = {-2, -2, -1, 0, 0, 1, 0, 0, 1, -2};

// T x T nodal operators of one (spin, m, band): S A, S Lambda A, and the
// spin-raised / spin-lowered angular parts S' f A with ell alignment, in long
// double (angular.hpp:129-165, ghp_device.hpp:174-196).
std::array<std::vector<double>, kSoOps> nodal_operators(int spin, int m, int E, int T) {
  const BandTransform own = make_band_transform(spin, m, E, T);
  const int n = own.modes();
  std::array<std::vector<double>, kSoOps> out;
  for (auto& v : out) v.assign(static_cast<size_t>(T) * T, 0.0);
  for (int t = 0; t < T; ++t)
    for (int u = 0; u < T; ++u) {
      long double p = 0.0L, l = 0.0L;
      for (int k = 0; k < n; ++k) {
        const int ell = own.ell_min + k;
        const long double sa = static_cast<long double>(own.S[t * n + k]) *
                               static_cast<long double>(own.A[static_cast<size_t>(k) * T + u]);
        p += sa;
        l += -static_cast<long double>((ell - spin) * (ell + spin + 1)) * sa;
      }
      out[kOpProject][t * T + u] = static_cast<double>(p);
      out[kOpLaplacian][t * T + u] = static_cast<double>(l);
    }
  for (int k = 0; k < n; ++k)
    for (int u = 0; u < T; ++u) {
      out[kOpAnalyze][static_cast<size_t>(k) * T + u] = own.A[static_cast<size_t>(k) * T + u];
      out[kOpSynthesize][static_cast<size_t>(u) * T + k] = own.S[static_cast<size_t>(u) * n + k];
    }
  for (int dir = 0; dir < 2; ++dir) {
    const int s2 = dir == 0 ? spin + 1 : spin - 1;
    const int lmin2 = std::max(std::abs(s2), std::abs(m));
    if (E < lmin2) continue;
    const BandTransform dst = make_band_transform(s2, m, E, T);
    const int n2 = dst.modes();
    std::vector<double>& op = out[dir == 0 ? kOpRaise : kOpLower];
    for (int t = 0; t < T; ++t)
      for (int u = 0; u < T; ++u) {
        long double acc = 0.0L;
        for (int j = 0; j < n2; ++j) {
          const int ell = dst.ell_min + j;
          const int k = ell - own.ell_min;
          if (k < 0 || k >= n) continue;
          const long double f =
              dir == 0 ? std::sqrt(static_cast<long double>((ell - spin) * (ell + spin + 1)))
                       : -std::sqrt(static_cast<long double>((ell + spin) * (ell - spin + 1)));
          acc += static_cast<long double>(dst.S[t * n2 + j]) * f *
                 static_cast<long double>(own.A[static_cast<size_t>(k) * T + u]);
        }
        op[t * T + u] = static_cast<double>(acc);
      }
  }
  return out;
}

// staged rows (reused for the unprojected outputs), Q_eff and psi_t rows
int linear_lds(int rows, int mrows, int T) {
  return 8 * kSchemeTableDoubles + 16 * (rows * 3 * T + 2 * mrows * T);
}

// three linear-triple workgroups per CU (measured best of 40-80 KiB)
constexpr int kSoLdsBudget = 48 * 1024;

}  // namespace

// Size limits of the second-order kernels, checked before any device work:
// the source kernel addresses a radial x theta plane, and the plane's Kerr
// background records, with 32-bit byte offsets; the source projection keeps
// 9 rows of kProjRows x targets x ntheta complex values in LDS.
void Plan::so_check_limits(const RunParameters& p, int n_targets) {
  const long long plane = static_cast<long long>(p.nr) * p.ntheta;
  if (plane * static_cast<long long>(sizeof(SoBackground)) >= (1LL << 32))
    throw std::invalid_argument("second-order plans need nr x ntheta < " +
                                std::to_string((1LL << 32) / sizeof(SoBackground)));
  const long long proj = 16LL * 9 * kProjRows * n_targets * p.ntheta;
  if (proj > kMaxLdsBytes)
    throw std::invalid_argument(
        "second-order source projection needs second-order modes x ntheta <= " +
        std::to_string(kMaxLdsBytes / (16 * 9 * kProjRows)));
}

void Plan::so_build() {
  so_ = std::make_unique<SoHost>();
  SoHost& H = *so_;
  std::map<std::tuple<int, int, int, int>, int> op_index;  // (spin, m, E, T) -> offset
  long long base = 0, pbase = 0, bgoff = 0;
  for (int i = 0; i < instance_count(); ++i) {
    const Instance& in = inst_[i];
    const RunParameters& p = in.params;
    const int M = in.M, N = in.N, T = in.T;
    if (M > kSoMaxModes)
      throw std::invalid_argument("second-order plans support at most 16 stored modes");
    if (p.ell_max_first < 3 || p.ell_max_second < 3 ||
        T < std::max({p.ell_max_first + 1, p.ell_max_second + 1,
                      (2 * p.ell_max_first + p.ell_max_second + 2) / 2}))
      throw std::invalid_argument("full pipeline angular bands require exact-product padding");
    SoInstance I{};
    I.base = base;
    I.pbase = pbase;
    I.bgoff = bgoff;
    I.M = M;
    I.N = N;
    I.T = T;
    I.scheme = static_cast<int>(p.scheme);
    I.constrained = p.reduction == Reduction::StageConstrained ? 1 : 0;
    I.dissipate = p.dissipation > 0.0 ? 1 : 0;
    I.mass = p.mass;
    I.spin = p.spin;
    I.length = p.compactification_length;
    I.h = in.h;
    I.inv_h = 1.0 / in.h;
    I.gamma = p.reduction_damping;
    I.eps = p.dissipation;
    const std::vector<int>& modes = in.registry.modes();
    for (int mi = 0; mi < M; ++mi) {
      I.modes[mi] = modes[mi];
      I.sharp[mi] = in.registry.sharp_index(modes[mi]);
      I.field_mask[mi] = (in.registry.is_parent(modes[mi]) ? 0x3ff : 0) |
                         (in.registry.is_target(modes[mi]) ? 0x1c00 : 0);
    }
    I.n_parents = 0;
    for (int mi = 0; mi < M; ++mi)
      if (in.registry.is_parent(modes[mi])) I.parent_mode[I.n_parents++] = mi;
    for (int pi = 0; pi < I.n_parents; ++pi) {
      I.parent_sharp[pi] = -1;
      for (int pj = 0; pj < I.n_parents; ++pj)
        if (I.parent_mode[pj] == I.sharp[I.parent_mode[pi]]) I.parent_sharp[pi] = pj;
    }
    std::vector<int> targets = in.registry.targets();
    std::sort(targets.begin(), targets.end());
    if (static_cast<int>(targets.size()) > kSoMaxTargets ||
        static_cast<int>(in.registry.ordered_pairs().size()) > kSoMaxPairs)
      throw std::invalid_argument("second-order plans support at most 16 targets and 64 pairs");
    I.n_targets = static_cast<int>(targets.size());
    const auto& pairs = in.registry.ordered_pairs();
    for (int ts = 0; ts < I.n_targets; ++ts) {
      I.target_mode[ts] = in.registry.index(targets[ts]);
      const auto range = in.registry.pair_range(targets[ts]);
      I.pair_begin[ts] = range.first;
      I.pair_begin[ts + 1] = range.second;
    }
    for (size_t k = 0; k < pairs.size(); ++k) {
      I.pair_m1[k] = pairs[k].index1;
      I.pair_m2[k] = pairs[k].index2;
    }
    for (int c = 0; c < kSoCoords; ++c) {
      I.coord_spin[c] = kCoordSpin[c];
      I.coord_boost[c] = kCoordBoost[c];
      const bool second = c == kCoSecond || c == kCoT;
      const int E = second ? p.ell_max_second : p.ell_max_first;
      for (int mi = 0; mi < kSoMaxModes; ++mi) {
        for (int o = 0; o < kSoOps; ++o) I.op[c][mi][o] = -1;
        I.band_n[c][mi] = 0;
      }
      for (int mi = 0; mi < M; ++mi) {
        const int m = modes[mi];
        const bool active = second ? in.registry.is_target(m) : in.registry.is_parent(m);
        if (!active || std::abs(m) > E) continue;
        const auto key = std::make_tuple(kCoordSpin[c], m, E, T);
        auto it = op_index.find(key);
        int off;
        if (it == op_index.end()) {
          const auto ops = nodal_operators(kCoordSpin[c], m, E, T);
          off = static_cast<int>(H.ops.size());
          // stored column-major (entry (t, u) at u * T + t): the lanes of a
          // row operator read consecutive t of one column together
          for (const auto& o : ops)
            for (int u = 0; u < T; ++u)
              for (int t = 0; t < T; ++t) H.ops.push_back(o[static_cast<size_t>(t) * T + u]);
          op_index[key] = off;
        } else {
          off = it->second;
        }
        for (int o = 0; o < kSoOps; ++o) I.op[c][mi][o] = off + o * T * T;
        I.band_n[c][mi] = E - std::max(std::abs(kCoordSpin[c]), std::abs(m)) + 1;
        // a shifted band that does not exist (ell_max below its minimum ell)
        const int s_up = kCoordSpin[c] + 1, s_dn = kCoordSpin[c] - 1;
        if (E < std::max(std::abs(s_up), std::abs(m))) I.op[c][mi][kOpRaise] = -1;
        if (E < std::max(std::abs(s_dn), std::abs(m))) I.op[c][mi][kOpLower] = -1;
      }
    }
    I.node_off = static_cast<int>(H.nodes.size());
    std::vector<double> x, w;
    gauss_legendre(T, x, w);
    for (int t = 0; t < T; ++t) H.nodes.push_back(x[t]);
    for (int t = 0; t < T; ++t) H.nodes.push_back(std::sqrt(1.0 - x[t] * x[t]));
    for (int t = 0; t < T; ++t) H.nodes.push_back(in.cos_theta[t]);
    for (int t = 0; t < T; ++t) H.nodes.push_back(in.sin_theta[t]);
    for (int t = 0; t < T; ++t) H.nodes.push_back(w[t]);
    H.inst.push_back(I);
    // linear-triple tiles: the largest output height whose LDS fits
    const RadialScheme rs = p.scheme;
    const int sc = static_cast<int>(rs);
    std::vector<SoTile> line;
    int rows_out = std::min(64, N);
    for (;;) {
      line.clear();
      int worst = 0;
      const int count = (N + rows_out - 1) / rows_out;
      const int basec = N / count, rem = N % count;
      int t0 = 0;
      for (int k = 0; k < count; ++k) {
        const int len = basec + (k < rem ? 1 : 0);
        SoTile t{};
        t.inst = i;
        t.t0 = t0;
        t.t1 = t0 + len;
        t.mlo = derivative_columns(rs, N, t.t0).first;
        t.mhi = derivative_columns(rs, N, t.t1 - 1).second;
        int lo = std::min(t.mlo, derivative_columns(rs, N, t.t0).first);
        int hi = std::max(t.mhi, derivative_columns(rs, N, t.t1 - 1).second);
        if (p.dissipation > 0.0) {
          lo = std::min(lo, dissipation_columns(rs, N, t.t0).first);
          hi = std::max(hi, dissipation_columns(rs, N, t.t1 - 1).second);
        }
        if (I.constrained) {
          lo = std::min(lo, derivative_columns(rs, N, t.mlo).first);
          hi = std::max(hi, derivative_columns(rs, N, t.mhi - 1).second);
        }
        t.lo = lo;
        t.hi = hi;
        worst = std::max(worst, linear_lds(hi - lo, t.mhi - t.mlo, T));
        // a thread holds at most kSoLinearOut outputs (one row of up to
        // kSoLinearOut * kSoThreads nodes always fits, T <= kReconMaxThreads)
        if (len * T > kSoLinearOut * kSoThreads) worst = std::max(worst, kSoLdsBudget + 1);
        line.push_back(t);
        t0 += len;
      }
      if (worst <= kSoLdsBudget || rows_out == 1) {
        H.lin_lds = std::max(H.lin_lds, worst);
        break;
      }
      rows_out = rows_out > 8 ? rows_out - 4 : rows_out - 1;
    }
    const int first = static_cast<int>(H.tiles[sc].size());
    for (int mi = 0; mi < M; ++mi)
      for (SoTile t : line) {
        t.mode = mi;
        H.tiles[sc].push_back(t);
      }
    H.inst_tiles[sc].resize(instance_count(), {0, 0});
    H.inst_tiles[sc][i] = {first, static_cast<int>(H.tiles[sc].size()) - first};
    H.schemes |= 1u << sc;
    // reconstruction elements of a radial row: parent modes x theta nodes
    const long long MT = static_cast<long long>(I.n_parents) * T;
    if (MT > kReconMaxThreads)
      throw std::invalid_argument("second-order plans need first-order modes x ntheta <= " +
                                  std::to_string(kReconMaxThreads));
    H.max_mt = std::max(H.max_mt, static_cast<int>(MT));
    H.max_drad = std::max(H.max_drad, so_drad_blocks(M, N, T));
    // raw sums and their projections (4 x T per target and row), and the
    // angular part of eth'_3 T (T per target and row)
    H.proj_lds = std::max(H.proj_lds, static_cast<int>(16 * 9LL * kProjRows * I.n_targets * T));
    // sized limits of the kernels (so_check_limits), checked again here
    so_check_limits(p, static_cast<int>(in.registry.targets().size()));
    H.max_source = std::max(H.max_source, static_cast<long long>(N) * T);
    H.max_targets = std::max(H.max_targets, I.n_targets);
    H.max_n = std::max(H.max_n, N);
    const long long block = static_cast<long long>(M) * 13 * N * T;
    H.max_block = std::max(H.max_block, block);
    H.max_plane = std::max(H.max_plane, static_cast<long long>(M) * N * T);
    base += block;
    pbase += static_cast<long long>(M) * N * T;
    bgoff += static_cast<long long>(N) * T;
    H.max_nt = std::max(H.max_nt, static_cast<long long>(N) * T);
  }
  H.bg_points = bgoff;
  // reconstruction workgroups: several radial rows each when every row fits
  // kReconThreads threads, else one row of up to kReconMaxThreads threads
  H.recon_groups = 0;
  H.rec0_lds = 0;
  for (const SoInstance& I : H.inst) {
    const int MT = I.n_parents * I.T;
    const int rpw = H.max_mt <= kReconThreads ? recon_rows_per_group(MT) : 1;
    H.recon_groups = std::max(H.recon_groups, (I.N + rpw - 1) / rpw);
    H.rec0_lds = std::max(H.rec0_lds, 16 * kReconLdsRows * MT * rpw);
  }
  H.rec1_lds = H.rec0_lds;
  for (auto& v : H.inst_tiles) v.resize(instance_count(), {0, 0});
  if (H.rec1_lds > 160 * 1024 || H.rec0_lds > 160 * 1024)
    throw std::invalid_argument(
        "second-order plans need stored modes x theta nodes small enough for one workgroup "
        "(at most about 190)");
  H.values = base;
  H.planes = pbase;
  H.norms.assign(instance_count(), std::array<double, 14>{});
  H.active_host.assign(instance_count(), 0);
  H.d_inst = device_alloc<SoInstance>(H.inst.size());
  check(hipMemcpy(H.d_inst, H.inst.data(), H.inst.size() * sizeof(SoInstance),
                  hipMemcpyHostToDevice),
        "upload instances");
  H.d_ops = device_alloc<double>(H.ops.size());
  check(hipMemcpy(H.d_ops, H.ops.data(), H.ops.size() * sizeof(double), hipMemcpyHostToDevice),
        "upload operators");
  H.d_nodes = device_alloc<double>(H.nodes.size());
  check(hipMemcpy(H.d_nodes, H.nodes.data(), H.nodes.size() * sizeof(double),
                  hipMemcpyHostToDevice),
        "upload nodes");
  const SchemeTable tabs[3] = {make_table(RadialScheme::D42), make_table(RadialScheme::D84),
                               make_table(RadialScheme::D105)};
  H.d_tables = device_alloc<SchemeTable>(3);
  check(hipMemcpy(H.d_tables, tabs, sizeof(tabs), hipMemcpyHostToDevice), "upload tables");
  for (int sc = 0; sc < 3; ++sc) {
    H.d_tiles[sc] = device_alloc<SoTile>(H.tiles[sc].size());
    if (!H.tiles[sc].empty())
      check(hipMemcpy(H.d_tiles[sc], H.tiles[sc].data(), H.tiles[sc].size() * sizeof(SoTile),
                      hipMemcpyHostToDevice),
            "upload tiles");
  }
  H.d_active = device_alloc<int>(instance_count());
  H.d_dt = device_alloc<double>(instance_count());
  H.d_red = device_alloc<double>(static_cast<size_t>(instance_count()) * kSoReduceDoubles);
  {
    std::vector<double> steps(instance_count());
    for (int i = 0; i < instance_count(); ++i) steps[i] = inst_[i].dt;
    H.d_dt_step = device_alloc<double>(steps.size());
    check(hipMemcpy(H.d_dt_step, steps.data(), steps.size() * sizeof(double),
                    hipMemcpyHostToDevice),
          "upload configured steps");
  }
  for (cdouble** b : {&H.y, &H.acc, &H.ya, &H.yb, &H.k}) {
    *b = device_alloc<cdouble>(H.values);
    check(launch_fill_zero(*b, H.values, stream_), "zero state");
  }
  const std::pair<cdouble**, int> scratch[] = {
      {&H.tan, kSlotsTan}, {&H.dpsi, 1},          {&H.dpsi_t, 1},
      {&H.rdr, kSlotsRdr}, {&H.rdr_t, kSlotsRdr}, {&H.rang, kSlotsRang},
      {&H.src, kSlotsSrc}, {&H.cons, kSlotsCons}, {&H.terms, kSlotsTerms},
      {&H.rang_t, kSlotsRang}, {&H.ex, kSlotsEx}, {&H.raw, kSlotsRaw},
      {&H.rangx, kSlotsRangX}};
  for (const auto& [ptr, slots] : scratch) {
    *ptr = device_alloc<cdouble>(H.planes * slots);
    check(launch_fill_zero(*ptr, H.planes * slots, stream_), "zero scratch");
  }
  H.d_bg = device_alloc<SoBackground>(2 * H.bg_points);
  {
    SoArgs a{};
    a.inst = H.d_inst;
    a.nodes = H.d_nodes;
    a.bg = H.d_bg;
    a.n_inst = instance_count();
    check(so_launch_background(a, H.max_nt, stream_), "background tables");
  }
  const int lds = std::max({H.lin_lds, H.rec0_lds, H.rec1_lds});
  if (lds > 64 * 1024) check(so_set_lds_limit(lds), "raise LDS limit");
  for (int i = 0; i < instance_count(); ++i)
    H.active_host[i] = inst_[i].activation.active;
  so_upload_active();
}

void Plan::so_free() {
  if (!so_) return;
  SoHost& H = *so_;
  void* ptrs[] = {H.d_bg, H.d_inst, H.d_ops,  H.d_nodes, H.d_tables, H.d_tiles[0], H.d_tiles[1],
                  H.d_tiles[2], H.d_active, H.d_dt, H.d_dt_step, H.d_red, H.y, H.acc, H.ya,
                  H.yb, H.k,
                  H.tan, H.dpsi, H.dpsi_t, H.rdr, H.rdr_t, H.rang, H.src, H.cons, H.terms,
                  H.rang_t, H.ex, H.raw, H.rangx, H.pairs};
  for (void* p : ptrs)
    if (p) hipFree(p);
  so_.reset();
}

void Plan::so_upload_active() {
  SoHost& H = *so_;
  for (int i = 0; i < instance_count(); ++i) H.active_host[i] = inst_[i].activation.active;
  check(hipMemcpyAsync(H.d_active, H.active_host.data(), H.active_host.size() * sizeof(int),
                       hipMemcpyHostToDevice, stream_),
        "upload source latch");
  check(hipStreamSynchronize(stream_), "latch sync");
}

// One RHS evaluation of instances [first, first + n) (spatial_pipeline.hpp:
// 458-588), stage `in` -> `out`, both in the plan's state layout.
void Plan::so_rhs(int first, int n, const cdouble* in, cdouble* out, bool constraints) {
  SoHost& H = *so_;
  SoArgs a{};
  a.inst = H.d_inst;
  a.bg = H.d_bg;
  a.ops = H.d_ops;
  a.nodes = H.d_nodes;
  a.tables = H.d_tables;
  a.active = H.d_active;
  a.in = in;
  a.out = out;
  a.tan = H.tan;
  a.dpsi = H.dpsi;
  a.dpsi_t = H.dpsi_t;
  a.rdr = H.rdr;
  a.rdr_t = H.rdr_t;
  a.rang = H.rang;
  a.rang_t = H.rang_t;
  a.ex = H.ex;
  a.raw = H.raw;
  a.src = H.src;
  a.cons = H.cons;
  a.terms = H.terms;
  a.rangx = H.rangx;
  a.pairs = H.capture_pairs && n == 1 ? H.pairs : nullptr;
  a.want_constraints = constraints ? 1 : 0;
  a.first_inst = first;
  a.n_inst = n;
  auto linear = [&](int kind) {
    for (int sc = 0; sc < 3; ++sc) {
      int tfirst = 0, tcount = 0;
      if (n == instance_count()) {
        tcount = static_cast<int>(H.tiles[sc].size());
      } else {
        tfirst = H.inst_tiles[sc][first].first;
        tcount = 0;
        for (int i = first; i < first + n; ++i) tcount += H.inst_tiles[sc][i].second;
      }
      if (tcount == 0) continue;
      SoArgs b = a;
      b.tiles = H.d_tiles[sc] + tfirst;
      check(so_launch_linear(sc, kind, b, tcount, H.lin_lds, stream_), "second-order linear");
    }
  };
  linear(0);
  check(so_launch_drad(0, a, H.max_drad, H.schemes, stream_), "reconstruction D");
  check(so_launch_recon(0, a, H.recon_groups, H.max_mt, H.rec0_lds, stream_),
        "reconstruction chain");
  linear(1);
  check(so_launch_drad(1, a, H.max_drad, H.schemes, stream_), "tangent reconstruction D");
  check(so_launch_recon(1, a, H.recon_groups, H.max_mt, H.rec1_lds, stream_),
        "tangent reconstruction chain");
  check(so_launch_source(a, H.max_source, H.max_targets, stream_), "quadratic source");
  check(so_launch_source_projection(a, H.max_n, H.proj_lds, stream_), "source projection");
  linear(2);
}

// One classical RK4 step of every instance with its own dt (0 leaves the
// instance unchanged) under the current latch (device_rk4.hpp:84-128).
void Plan::so_rk_pass(const std::vector<double>& dts) {
  SoHost& H = *so_;
  check(hipMemcpyAsync(H.d_dt, dts.data(), dts.size() * sizeof(double), hipMemcpyHostToDevice,
                       stream_),
        "upload steps");
  so_rk_launches(H.d_dt);
  check(hipStreamSynchronize(stream_), "RK4 pass");
}

// The launches of one RK4 pass with the per-instance steps at d_dt: four RHS
// evaluations, each followed by its RK combination.  Kernel launches only,
// so a stream capture of it is the step graph.
void Plan::so_rk_launches(const double* d_dt) {
  SoHost& H = *so_;
  const cdouble* in_of[5] = {nullptr, H.y, H.ya, H.yb, H.ya};
  cdouble* out_of[5] = {nullptr, H.ya, H.yb, H.ya, H.y};
  for (int s = 1; s <= 4; ++s) {
    so_rhs(0, instance_count(), in_of[s], H.k, false);
    SoRkArgs r{};
    r.inst = H.d_inst;
    r.dt = d_dt;
    r.k = H.k;
    r.y = H.y;
    r.acc = H.acc;
    r.yout = out_of[s];
    r.max_values = H.max_block;
    r.stage = s;
    r.n_inst = instance_count();
    check(so_launch_rk(r, stream_), "RK4 combination");
  }
}

// consider_source_activation (spatial_pipeline.hpp:1242-1289) for one
// instance on its current state.
bool Plan::so_consider(int i, double time) {
  if (!std::isfinite(time) || time < 0.0)
    throw std::invalid_argument("source activation requires a finite accepted-state time");
  Instance& in = inst_.at(i);
  SourceActivation& a = in.activation;
  const RunParameters& p = in.params;
  if (a.active || time < p.source_start_time || time <= a.last_eligibility_time) return a.active;
  if (p.source_mode == SourceMode::Disabled) return false;
  if (p.source_mode == SourceMode::Unrestricted) {
    a.last_eligibility_time = time;
    a.consecutive_passes = p.required_consecutive_passes;
    a.active = 1;
    a.activation_time = time;
    return true;
  }
  SoHost& H = *so_;
  so_rhs(i, 1, H.y, H.k, true);
  SoReduceArgs r{};
  r.inst = H.d_inst + i;
  r.nodes = H.d_nodes;
  r.tables = H.d_tables;
  r.cons = H.cons;
  r.terms = H.terms;
  r.out = H.d_red;
  r.n_inst = 1;
  check(so_launch_reduce(r, stream_), "constraint reduction");
  double red[kSoReduceDoubles];
  check(hipMemcpyAsync(red, H.d_red, sizeof(red), hipMemcpyDeviceToHost, stream_), "read norms");
  check(hipStreamSynchronize(stream_), "constraint sync");
  // SourceConstraintEvaluator::sample's host part (source_activation.hpp:171-212)
  const SoInstance& I = H.inst[i];
  const double weight_sum = static_cast<double>(I.M) * in.horizon * 4.0 * 3.14159265358979323846;
  const double floor = 64.0 * std::numeric_limits<double>::epsilon();
  std::array<double, 14>& out = H.norms[i];
  double controlling = 0.0;
  bool finite = red[18] > 0.5;
  for (int f = 0; f < 3; ++f) {
    out[4 * f + 0] = std::sqrt(red[9 + f]);
    out[4 * f + 1] = std::sqrt(red[15 + f]);
    out[4 * f + 2] = std::sqrt(red[f] / weight_sum);
    const double left = std::sqrt(red[3 + f]), right = std::sqrt(red[6 + f]);
    const double natural = std::max(left, right);
    const double den = left + right + floor * natural;
    out[4 * f + 3] = den > 0.0 ? std::sqrt(red[f]) / den : 0.0;
    controlling = std::max(controlling, std::max(out[4 * f + 1], out[4 * f + 3]));
  }
  if (!finite || !std::isfinite(controlling)) {
    finite = false;
    controlling = std::numeric_limits<double>::infinity();
  }
  out[12] = controlling;
  out[13] = finite ? 1.0 : 0.0;
  const bool passed = finite && controlling <= p.constraint_tolerance;
  a.last_eligibility_time = time;
  a.consecutive_passes =
      passed ? std::min(a.consecutive_passes + 1, p.required_consecutive_passes) : 0;
  if (a.consecutive_passes >= p.required_consecutive_passes) {
    a.active = 1;
    a.activation_time = time;
  }
  return a.active;
}

// SpatialPipeline::step (spatial_pipeline.hpp:590-620) for every instance at
// once: the latch is considered on accepted states, and an instance whose
// source start time falls inside its step is advanced in two exact parts.
void Plan::so_step(const std::vector<double>& t, const std::vector<double>& dt) {
  const int n = instance_count();
  std::vector<double> end(n), start(n);
  std::vector<char> split(n, 0);
  bool any_split = false;
  for (int i = 0; i < n; ++i) {
    start[i] = inst_[i].params.source_start_time;
    end[i] = t[i] + dt[i];
    if (!std::isfinite(end[i])) throw std::invalid_argument("pipeline step end time is not finite");
    if (!inst_[i].activation.active && t[i] >= start[i]) so_consider(i, t[i]);
    split[i] = !inst_[i].activation.active && t[i] < start[i] && start[i] < end[i];
    any_split = any_split || split[i];
  }
  std::vector<double> pass(n);
  for (int i = 0; i < n; ++i) pass[i] = split[i] ? start[i] - t[i] : dt[i];
  so_upload_active();
  so_rk_pass(pass);
  if (any_split) {
    for (int i = 0; i < n; ++i) {
      if (split[i]) so_consider(i, start[i]);
      pass[i] = split[i] ? end[i] - start[i] : 0.0;
    }
    so_upload_active();
    so_rk_pass(pass);
  }
  for (int i = 0; i < n; ++i)
    if (!inst_[i].activation.active && end[i] >= start[i]) so_consider(i, end[i]);
  so_upload_active();
}

// so_step as a replay of the captured step graph, for a step (each instance
// its configured dt) in which no instance splits at its source start and no
// latch decision needs a constraint evaluation: the Unrestricted and
// Disabled policies decide on the host (source_activation.hpp:92-233), and
// the latch reaches the kernels through the device flags, uploaded only when
// one changes.  Returns false, doing nothing, for any other step.
bool Plan::so_graph_step(const std::vector<double>& t, const std::vector<double>& dt) {
  const int n = instance_count();
  for (int i = 0; i < n; ++i) {
    const Instance& in = inst_[i];
    if (in.activation.active) continue;
    const double start = in.params.source_start_time, end = t[i] + dt[i];
    if (!std::isfinite(end)) throw std::invalid_argument("pipeline step end time is not finite");
    if (t[i] < start && start < end) return false;
    if ((t[i] >= start || end >= start) && in.params.source_mode == SourceMode::ConstraintAware)
      return false;
  }
  bool changed = false;
  for (int i = 0; i < n; ++i) {
    const Instance& in = inst_[i];
    if (!in.activation.active && t[i] >= in.params.source_start_time)
      changed = so_consider(i, t[i]) || changed;
  }
  if (changed) so_upload_active();
  hipEvent_t e0 = step_begin_event();
  check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
  step_end_event(e0);
  ++graph_replays_;
  changed = false;
  for (int i = 0; i < n; ++i) {
    const Instance& in = inst_[i];
    if (!in.activation.active && t[i] + dt[i] >= in.params.source_start_time)
      changed = so_consider(i, t[i] + dt[i]) || changed;
  }
  if (changed) so_upload_active();
  return true;
}

// Gaussian data with the Galerkin P (pipeline_initial_data.hpp:142-340) on
// the GPU: one wavefront per (instance, parent mode, radius) forms Psi,
// Q = D Psi and P, and writes the band projections S A of the three fields as
// nodal rows; every other block of the state is zero.
void Plan::so_initialize() {
  SoHost& H = *so_;
  std::vector<LineDesc> ld;
  std::vector<InitLine> il;
  std::vector<double> mats, nodes;
  std::vector<cdouble> seeds;
  int lds = 0, max_nr = 0;
  for (int i = 0; i < instance_count(); ++i) {
    const Instance& in = inst_[i];
    const RunParameters& p = in.params;
    double center, width;
    const auto expanded = gaussian_seeds(in, center, width);
    for (int mi = 0; mi < in.M; ++mi) {
      const int m = in.registry.modes()[mi];
      if (!in.registry.is_parent(m)) continue;
      const BandTransform bt = make_band_transform(-2, m, p.ell_max_first, in.T);
      LineDesc L{};
      L.nr = in.N;
      L.n_ell = bt.modes();
      L.ntheta = in.T;
      L.ell_min = bt.ell_min;
      L.m = m;
      L.mat_off = static_cast<int>(mats.size());
      mats.insert(mats.end(), bt.S.begin(), bt.S.end());
      mats.insert(mats.end(), bt.A.begin(), bt.A.end());
      L.node_off = static_cast<int>(nodes.size());
      nodes.insert(nodes.end(), in.sin_theta.begin(), in.sin_theta.end());
      nodes.insert(nodes.end(), in.cos_theta.begin(), in.cos_theta.end());
      L.mass = p.mass;
      L.spin = p.spin;
      L.length = p.compactification_length;
      L.h = in.h;
      L.inv_h = 1.0 / in.h;
      ld.push_back(L);
      InitLine d{};
      d.seed_off = static_cast<int>(seeds.size());
      d.compact = p.compact_support ? 1 : 0;
      d.scheme = static_cast<int>(p.scheme);
      d.center = center;
      d.width = width;
      d.nodal_base = H.inst[i].base + static_cast<long long>(mi) * 13 * in.N * in.T;
      il.push_back(d);
      seeds.resize(seeds.size() + L.n_ell, cdouble{0.0, 0.0});
      for (const auto& s : expanded)
        if (s.m == m) {
          cdouble& c = seeds[d.seed_off + s.ell - L.ell_min];
          c.re += s.amplitude.real();
          c.im += s.amplitude.imag();
        }
      lds = std::max(lds, galerkin_init_lds_bytes(L.n_ell, L.ntheta));
      max_nr = std::max(max_nr, L.nr);
    }
  }
  const SchemeTable tabs[3] = {make_table(RadialScheme::D42), make_table(RadialScheme::D84),
                               make_table(RadialScheme::D105)};
  LineDesc* d_ld = nullptr;
  double* d_mats = nullptr;
  double* d_nodes = nullptr;
  SchemeTable* d_tabs = nullptr;
  auto release = [&] {
    for (void* q : {static_cast<void*>(d_ld), static_cast<void*>(d_mats),
                    static_cast<void*>(d_nodes), static_cast<void*>(d_tabs)})
      if (q) (void)hipFree(q);
  };
  try {
    d_ld = device_alloc<LineDesc>(ld.size());
    d_mats = device_alloc<double>(mats.size());
    d_nodes = device_alloc<double>(nodes.size());
    d_tabs = device_alloc<SchemeTable>(3);
    check(hipMemcpy(d_ld, ld.data(), ld.size() * sizeof(LineDesc), hipMemcpyHostToDevice),
          "upload initial data lines");
    check(hipMemcpy(d_mats, mats.data(), mats.size() * sizeof(double), hipMemcpyHostToDevice),
          "upload band matrices");
    check(hipMemcpy(d_nodes, nodes.data(), nodes.size() * sizeof(double), hipMemcpyHostToDevice),
          "upload nodes");
    check(hipMemcpy(d_tabs, tabs, sizeof(tabs), hipMemcpyHostToDevice), "upload tables");
    check(launch_fill_zero(H.y, H.values, stream_), "zero state");
    InitArgs a{};
    a.lines = d_ld;
    a.mats = d_mats;
    a.nodes = d_nodes;
    a.tables = d_tabs;
    a.nodal = H.y;
    run_initial_data(a, il, seeds, max_nr, lds);
  } catch (...) {
    release();
    throw;
  }
  release();
  for (int i = 0; i < instance_count(); ++i) inst_[i].activation = initial_activation(inst_[i].params);
  so_upload_active();
}

double Plan::so_import(int i, const cplx* host, double tol) {
  const double combined = import_check(i, host, tol);
  const SoInstance& I = so_->inst[i];
  check(hipMemcpyAsync(so_->y + I.base, d_imp_in_, nodal_value_count(i) * sizeof(cdouble),
                       hipMemcpyDeviceToDevice, stream_),
        "write state");
  check(hipStreamSynchronize(stream_), "import");
  import_release();
  return combined;
}

void Plan::so_export(int i, cplx* host) {
  const SoInstance& I = so_->inst.at(i);
  check(hipStreamSynchronize(stream_), "export sync");
  check(hipMemcpy(host, so_->y + I.base, nodal_value_count(i) * sizeof(cdouble),
                  hipMemcpyDeviceToHost),
        "read state");
}

// Reference-layout RHS of one instance (spatial_pipeline.hpp:458-588) under
// the current latch: the stage is staged into the plan's own buffers.
void Plan::so_evaluate_rhs(int i, const cplx* d_stage, cplx* d_out) {
  SoHost& H = *so_;
  const SoInstance& I = H.inst.at(i);
  const size_t bytes = static_cast<size_t>(nodal_value_count(i)) * sizeof(cdouble);
  so_upload_active();
  check(hipMemcpyAsync(H.ya + I.base, d_stage, bytes, hipMemcpyDeviceToDevice, stream_),
        "stage copy");
  so_rhs(i, 1, H.ya, H.k, false);
  check(hipMemcpyAsync(d_out, H.k + I.base, bytes, hipMemcpyDeviceToDevice, stream_),
        "rhs copy");
  // the RK combination skips the (mode, field) blocks a field does not
  // evolve on and relies on their being zero in the staging buffers: hand
  // the instance's staging block back as a copy of its (zero-padded) state
  check(hipMemcpyAsync(H.ya + I.base, H.y + I.base, bytes, hipMemcpyDeviceToDevice, stream_),
        "restore staging");
}

// EndpointWaveformSampler::sample (waveform_output.hpp:79-160) from the
// nodal FirstPsi and SecondPsi rows.
std::vector<WaveformRecordC> Plan::so_sample_waveform(int i) {
  Instance& in = inst_.at(i);
  const SoInstance& I = so_->inst.at(i);
  check(hipStreamSynchronize(stream_), "waveform sync");
  std::vector<WaveformRecordC> out;
  const double L2 = in.params.compactification_length * in.params.compactification_length;
  std::vector<cplx> row(in.T);
  for (int order = 1; order <= 2; ++order)
    for (int boundary = 0; boundary < 2; ++boundary) {
      const int r = boundary == 0 ? 0 : in.N - 1;
      const double scale = boundary == 0 ? L2 : in.horizon;
      for (int mi = 0; mi < in.M; ++mi) {
        const int m = in.registry.modes()[mi];
        if (order == 1 ? !in.registry.is_parent(m) : !in.registry.is_target(m)) continue;
        const int f = order == 1 ? 2 : 12;
        const int E = order == 1 ? in.params.ell_max_first : in.params.ell_max_second;
        check(hipMemcpy(row.data(), so_->y + I.base + ((static_cast<long long>(mi) * 13 + f) * in.N + r) * in.T,
                        in.T * sizeof(cdouble), hipMemcpyDeviceToHost),
              "read endpoint");
        const BandTransform bt = make_band_transform(-2, m, E, in.T);
        const std::vector<cplx> modal = bt.analyze(row.data());
        for (size_t k = 0; k < modal.size(); ++k) {
          WaveformRecordC rec{};
          rec.step = in.steps_done;
          rec.time = in.time;
          rec.boundary = boundary;
          rec.order = order;
          rec.m = m;
          rec.ell = bt.ell_min + static_cast<int>(k);
          rec.rescaled_re = modal[k].real();
          rec.rescaled_im = modal[k].imag();
          rec.endpoint_re = scale * modal[k].real();
          rec.endpoint_im = scale * modal[k].imag();
          out.push_back(rec);
        }
      }
    }
  return out;
}

void Plan::constraint_norms(int i, double* out14) const {
  if (!so_mode_) throw std::invalid_argument("constraint norms exist only for second-order plans");
  const auto& n = so_->norms.at(i);
  std::copy(n.begin(), n.end(), out14);
}

}  // namespace teuk
