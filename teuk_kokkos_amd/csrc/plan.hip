// Plan: device memory, line/tile tables, initial data, import/export,
// RK4 sequencing, hipGraph capture and waveform sampling.
#include "plan.hpp"
#include "plan_detail.hpp"
#include "so_host.hpp"
#include "init_layout.hpp"
#include "import_layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>

namespace teuk {

hipError_t launch_stage(int scheme, int stage, int ub, int variant, const StageArgs& a,
                        int n_tiles, int lds, hipStream_t s);
hipError_t set_stage_lds_limit(int bytes);
hipError_t launch_nodal_to_modal(const NodalArgs& a, const cdouble* nodal, cdouble* modal,
                                 hipStream_t s);
hipError_t launch_modal_to_nodal(const NodalArgs& a, const cdouble* modal, cdouble* nodal,
                                 hipStream_t s);
hipError_t launch_fill_zero(cdouble* p, long long count, hipStream_t s);

using namespace detail;

Plan::Instance::Instance(RunParameters p)
    : params(std::move(p)), registry(make_run_registry(params)) {
  M = registry.size();
  N = params.nr;
  T = params.ntheta;
  horizon = params.horizon_radius();
  h = (horizon - 0.0) / static_cast<double>(N - 1);
  dt = params.time_step();
  std::vector<double> x, w;
  gauss_legendre(T, x, w);
  sin_theta.resize(T);
  cos_theta.resize(T);
  for (int j = 0; j < T; ++j) {
    const double th = std::acos(x[j]);
    sin_theta[j] = std::sin(th);
    cos_theta[j] = std::cos(th);
  }
  activation = Plan::initial_activation(params);
}

// The latch a pipeline starts with: active at t = 0 only for the
// unrestricted policy with a zero start time (spatial_pipeline.hpp:342-348).
SourceActivation Plan::initial_activation(const RunParameters& p) {
  SourceActivation a;
  if (p.source_start_time == 0.0 && p.second_order_enabled &&
      p.source_mode == SourceMode::Unrestricted) {
    a.active = 1;
    a.activation_time = 0.0;
    a.consecutive_passes = p.required_consecutive_passes;
  }
  return a;
}

Plan::Plan(const std::vector<RunParameters>& params) {
  if (params.empty()) throw std::invalid_argument("a plan needs at least one instance");
  for (const auto& p : params) {
    validate_parameters(p);
    if (p.plus2_enabled)
      throw std::runtime_error("plus2 mode is not production integrated; use plus2.enabled=false "
                               "and plus2.mode=disabled");
    const RunParameters& q = p;
    ModeRegistry reg = make_run_registry(q);
    if (!reg.closed_under_sharp())
      throw std::invalid_argument("spatial pipeline modes must be closed under m -> -m");
    inst_.emplace_back(p);
  }
  // CFL admissibility of the configured step (solver_driver.hpp:36-57,
  // 117-125).
  for (const auto& in : inst_) {
    double vmax = 0.0;
    for (int r = 0; r < in.N; ++r)
      for (int j = 0; j < in.T; ++j) {
        const RadialCoefficients rc = radial_coefficients(
            in.params.mass, in.params.spin, in.params.compactification_length, -2, 0, in.h * r);
        const double ct = time_coefficient(rc, in.params.spin, in.sin_theta[j]);
        const double root = std::sqrt(rc.adv * rc.adv + ct * rc.principal);
        vmax = std::max(vmax, std::max(std::abs((rc.adv - root) / ct),
                                       std::abs((rc.adv + root) / ct)));
      }
    if (in.dt > in.params.cfl * in.h / vmax)
      throw std::invalid_argument("steps are too small for the requested radial CFL limit");
  }
  // Instances with the quadratic source run the 13-field second-order graph;
  // a plan holds either kind, not both.
  so_mode_ = params.front().second_order_enabled;
  for (const auto& p : params)
    if (p.second_order_enabled != so_mode_)
      throw std::invalid_argument(
          "a plan cannot mix first-order-only and second_order.enabled instances");
  if (so_mode_)
    for (const auto& in : inst_)
      so_check_limits(in.params, static_cast<int>(in.registry.targets().size()));
  check(hipStreamCreate(&own_stream_), "hipStreamCreate");
  ++g_allocations;
  stream_ = own_stream_;
  fork_ = create_event(false);
  if (so_mode_) {
    so_build();
    allocate_staging();
    allocate_diagnostics();
    check(hipStreamSynchronize(stream_), "plan setup");
    return;
  }
  build_lines();
  build_tiles();
  upload_tables();
  allocate_staging();
  allocate_diagnostics();
  d_y_ = device_alloc<cdouble>(modal_count_);
  d_acc_ = device_alloc<cdouble>(modal_count_);
  d_ya_ = device_alloc<cdouble>(modal_count_);
  d_yb_ = device_alloc<cdouble>(modal_count_);
  d_k_ = device_alloc<cdouble>(modal_count_);
  for (cdouble* p : {d_y_, d_acc_, d_ya_, d_yb_, d_k_})
    check(launch_fill_zero(p, modal_count_, stream_), "zero state");
  check(hipStreamSynchronize(stream_), "plan setup");
}

Plan::~Plan() {
  if (stream_) hipStreamSynchronize(stream_);
  if (own_stream_ && own_stream_ != stream_) hipStreamSynchronize(own_stream_);
  if (graph_exec_) hipGraphExecDestroy(graph_exec_);
  if (graph_multi_) hipGraphExecDestroy(graph_multi_);
  for (hipEvent_t e : events_) hipEventDestroy(e);
  for (hipEvent_t e : joins_) hipEventDestroy(e);
  if (fork_) hipEventDestroy(fork_);
  for (hipStream_t st : side_) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  for (void* p : {static_cast<void*>(d_lines_), static_cast<void*>(d_tiles_),
                  static_cast<void*>(d_mats_), static_cast<void*>(d_nodes_),
                  static_cast<void*>(d_tables_), static_cast<void*>(d_line_mode_),
                  static_cast<void*>(d_y_), static_cast<void*>(d_acc_), static_cast<void*>(d_ya_),
                  static_cast<void*>(d_yb_), static_cast<void*>(d_k_),
                  static_cast<void*>(d_imp_in_), static_cast<void*>(d_imp_res_),
                  static_cast<void*>(d_imp_flags_), static_cast<void*>(d_imp_norms_),
                  d_imp_rows_, static_cast<void*>(d_imp_mats_),
                  static_cast<void*>(d_diag_sharp_), static_cast<void*>(d_diag_nodes_),
                  static_cast<void*>(d_diag_partial_), static_cast<void*>(d_diag_pflags_),
                  static_cast<void*>(d_diag_out_)})
    if (p) hipFree(p);
  so_free();
  if (own_stream_) hipStreamDestroy(own_stream_);
}

// S and A of one band in the operand order of the Kerr MFMA kernels (the
// fragment layout is described with mfma_fragment_doubles).
// Round a float to BF16, to nearest with ties to even (finite values).
static uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Largest |E| and |N| / max|Lambda| for which the Gram correction runs on
// BF16 operands.  BF16 keeps 8 significant bits, so the correction's own
// relative error is about 4e-3; below these sizes that is under 1e-13 of the
// terms it corrects (ell 24 bands: |E| about 6e-12).
constexpr double kGramBf16Limit = 2e-11;

// Returns whether the band's Gram correction may take the BF16 fragments.
static bool append_mfma_fragments(const BandTransform& bt, int nt, std::vector<double>& out) {
  const int nl = bt.modes();
  const int T = mfma_theta_tiles(nt), K = mfma_ell_blocks(nl), LT = mfma_ell_tiles(nl);
  const size_t first = out.size();
  out.resize(first + mfma_fragment_doubles(nt, nl), 0.0);
  double* s = out.data() + first;
  double* a = s + 64 * T * K;
  for (int t = 0; t < T; ++t)
    for (int lane = 0; lane < 64; ++lane) {
      const int lr = lane % 16, lg = lane / 16;
      for (int q = 0; q < K; ++q) {
        const int th = 16 * t + lr, l = 4 * q + lg;
        s[(t * K + q) * 64 + lane] = th < nt && l < nl ? bt.S[th * nl + l] : 0.0;
      }
      for (int v = 0; v < 4; ++v)
        for (int u = 0; u < LT; ++u) {
          const int l = 16 * u + lr, th = 16 * t + 4 * v + lg;
          a[((t * 4 + v) * LT + u) * 64 + lane] = th < nt && l < nl ? bt.A[l * nt + th] : 0.0;
        }
    }
  // The Gram projection G = A S of k = G X and M = G Lambda G of the
  // Laplacian term, split as G = I + E and M = Lambda + N: E and N (of the
  // size of A S - I) as FP32 B operands, fragment (uo, kb), lane:
  // E[16 uo + lr][4 kb + lg] (gram_correction_floats), both summed in long
  // double from the double G before the final rounding.
  const size_t efirst = out.size();
  out.resize(efirst + gram_fragment_doubles(nl) + gram_bf16_doubles(nl), 0.0);
  float* e = reinterpret_cast<float*>(out.data() + efirst);
  long double emax = 0.0L, nmax = 0.0L, lmax = 0.0L;
  for (int j = 0; j < nl; ++j) {
    const int ell = bt.ell_min + j;
    lmax = std::max(lmax, static_cast<long double>((ell + 2) * (ell - 1)));
  }
  for (int which = 0; which < 2; ++which)
    for (int uo = 0; uo < LT; ++uo)
      for (int kb = 0; kb < K; ++kb)
        for (int lane = 0; lane < 64; ++lane) {
          const int i = 16 * uo + lane % 16, k = 4 * kb + lane / 16;
          float v = 0.0f;
          if (i < nl && k < nl) {
            long double x;
            if (which == 0) {
              x = static_cast<long double>(bt.G[i * nl + k]) - (i == k ? 1.0L : 0.0L);
            } else {
              x = 0.0L;
              for (int j = 0; j < nl; ++j) {
                const int ell = bt.ell_min + j;
                x += static_cast<long double>(bt.G[i * nl + j]) *
                     static_cast<long double>(-(ell + 2) * (ell - 1)) *
                     static_cast<long double>(bt.G[j * nl + k]);
              }
              if (i == k) {
                const int ell = bt.ell_min + i;
                x -= static_cast<long double>(-(ell + 2) * (ell - 1));
              }
            }
            v = static_cast<float>(x);
            (which == 0 ? emax : nmax) = std::max(which == 0 ? emax : nmax, fabsl(x));
          }
          e[which * 64 * LT * K + (uo * K + kb) * 64 + lane] = v;
        }
  if (nl > kGramBf16MaxEll) return false;
  // BF16 fragments: (which, uo), lane, j -> E[16 uo + lr][8 lg + j], the
  // FP32 value rounded once more
  uint16_t* b = reinterpret_cast<uint16_t*>(out.data() + efirst + gram_fragment_doubles(nl));
  for (int which = 0; which < 2; ++which)
    for (int uo = 0; uo < LT; ++uo)
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < 8; ++j) {
          const int i = 16 * uo + lane % 16, k = 8 * (lane / 16) + j;
          float v = 0.0f;
          if (i < nl && k < nl) {
            // the FP32 fragment holding E[i][k]: (uo, kb = k / 4), lane i % 16 + 16 (k % 4)
            v = e[which * 64 * LT * K + (uo * K + k / 4) * 64 + (i % 16) + 16 * (k % 4)];
          }
          b[((which * LT + uo) * 64 + lane) * 8 + j] = bf16_bits(v);
        }
  return emax <= kGramBf16Limit && nmax <= kGramBf16Limit * lmax;
}

void Plan::build_lines() {
  // (m, E, ntheta) -> offsets of the dense matrices and of the MFMA fragments
  std::map<std::tuple<int, int, int>, std::tuple<int, int, int>> mat_index;
  long long base = 0;
  for (int i = 0; i < instance_count(); ++i) {
    Instance& in = inst_[i];
    const RunParameters& p = in.params;
    const int node_off = static_cast<int>(nodes_.size());
    nodes_.insert(nodes_.end(), in.sin_theta.begin(), in.sin_theta.end());
    nodes_.insert(nodes_.end(), in.cos_theta.begin(), in.cos_theta.end());
    in.line_of_mode.assign(in.M, -1);
    for (int mi = 0; mi < in.M; ++mi) {
      const int m = in.registry.modes()[mi];
      if (!in.registry.is_parent(m)) continue;
      const auto key = std::make_tuple(m, p.ell_max_first, in.T);
      int off, frag_off, bf16_ok;
      auto it = mat_index.find(key);
      if (it == mat_index.end()) {
        const BandTransform bt = make_band_transform(-2, m, p.ell_max_first, in.T);
        off = static_cast<int>(mats_.size());
        mats_.insert(mats_.end(), bt.S.begin(), bt.S.end());
        mats_.insert(mats_.end(), bt.A.begin(), bt.A.end());
        mats_.insert(mats_.end(), bt.G.begin(), bt.G.end());
        mats_.resize(mats_.size() + (gram_doubles(bt.modes()) - bt.modes() * bt.modes()), 0.0);
        const std::vector<double>& B = bt.gram ? bt.B : bt.A;
        mats_.insert(mats_.end(), B.begin(), B.end());
        mats_.resize((mats_.size() + 1) / 2 * 2, 0.0);  // fragments start 16-byte aligned
        frag_off = static_cast<int>(mats_.size());
        bf16_ok = append_mfma_fragments(bt, in.T, mats_) ? 1 : 0;
        mat_index[key] = {off, frag_off, bf16_ok};
      } else {
        std::tie(off, frag_off, bf16_ok) = it->second;
      }
      const int n_ell = p.ell_max_first - minimum_ell(-2, m) + 1;
      if (n_ell > kMaxBandEll)
        throw std::runtime_error("bands of more than 64 ell values (ellmax_first - max(2, |m|) "
                                 "+ 1 > 64) are not supported by the stage kernels");
      LineDesc L{};
      L.base = base;
      L.nr = in.N;
      L.n_ell = p.ell_max_first - minimum_ell(-2, m) + 1;
      L.ntheta = in.T;
      L.ell_min = minimum_ell(-2, m);
      L.m = m;
      L.constrained = p.reduction == Reduction::StageConstrained ? 1 : 0;
      L.kerr = p.spin != 0.0 ? 1 : 0;
      L.dissipate = p.dissipation > 0.0 ? 1 : 0;
      L.gram = band_gram_needed(-2, m, p.ell_max_first, in.T) ? 1 : 0;
      L.gram_bf16 = L.gram && bf16_ok && mixed_gram_ ? 1 : 0;
      line_bf16_ok_.push_back(L.gram && bf16_ok ? 1 : 0);
      L.mat_off = off;
      L.frag_off = frag_off;
      L.gfrag_off = frag_off + mfma_fragment_doubles(in.T, n_ell);
      L.node_off = node_off;
      L.mass = p.mass;
      L.spin = p.spin;
      L.length = p.compactification_length;
      L.h = in.h;
      L.inv_h = 1.0 / in.h;
      L.gamma = p.reduction_damping;
      L.eps = p.dissipation;
      L.dt = in.dt;
      in.line_of_mode[mi] = static_cast<int>(lines_.size());
      lines_.push_back(L);
      line_instance_.push_back(i);
      line_mode_.push_back(mi);
      base += static_cast<long long>(in.N) * 3 * L.n_ell;
    }
  }
  modal_count_ = base;
}

void Plan::build_tiles() {
  tiles_.clear();
  groups_.clear();
  max_lds_ = 0;
  long long total_rows = 0;
  for (const auto& L : lines_) total_rows += L.nr;
  // Enough workgroups to fill 256 CUs several times over, but no tile under
  // 64 rows: below that the stencil halo's psi_t rows and re-read rows cost
  // more than the extra parallelism gains (a single nr = 4097 ell-32 run
  // steps 30 % faster with 64-row tiles than with 16).
  const int parallel_rows = static_cast<int>(std::max<long long>(64, total_rows / 1024));
  std::vector<std::vector<TileDesc>> per(instance_count());
  std::vector<int> inst_lds(instance_count(), 0);
  // Execution group of each instance: (scheme, ell-tile class of its widest
  // line, line variant); an instance takes a compile-time variant when all
  // its lines share it (stage_variant).
  std::vector<GroupKey> inst_key(instance_count());
  for (int i = 0; i < instance_count(); ++i) {
    const Instance& in = inst_[i];
    int ub = 1, variant = -1;
    for (int mi = 0; mi < in.M; ++mi) {
      const int li = in.line_of_mode[mi];
      if (li < 0) continue;
      ub = std::max(ub, band_tiles_class(lines_[li].n_ell));
      const int v = specialized_ ? stage_variant(lines_[li]) : kStageGeneral;
      variant = variant < 0 || variant == v ? v : kStageGeneral;
    }
    inst_key[i] = {static_cast<int>(in.params.scheme), ub, std::max(variant, 0)};
  }
  for (int i = 0; i < instance_count(); ++i) {
    Instance& in = inst_[i];
    const RadialScheme rs = in.params.scheme;
    const GroupShape shape = shape_of(inst_key[i]);
    const int budget = shape.lds_budget > 0 ? shape.lds_budget : lds_budget_;
    const int rows_set = shape.tile_rows > 0 ? shape.tile_rows : tile_rows_;
    for (int mi = 0; mi < in.M; ++mi) {
      const int li = in.line_of_mode[mi];
      if (li < 0) continue;
      const LineDesc& L = lines_[li];
      const int n = L.nr;
      auto ranges = [&](int t0, int t1, TileDesc& t) {
        t.t0 = t0;
        t.t1 = t1;
        t.mlo = derivative_columns(rs, n, t0).first;
        t.mhi = derivative_columns(rs, n, t1 - 1).second;
        int lo = t.mlo, hi = t.mhi;
        if (L.dissipate) {
          lo = std::min(lo, dissipation_columns(rs, n, t0).first);
          hi = std::max(hi, dissipation_columns(rs, n, t1 - 1).second);
        }
        if (L.constrained) {
          lo = std::min(lo, derivative_columns(rs, n, t.mlo).first);
          hi = std::max(hi, derivative_columns(rs, n, t.mhi - 1).second);
        }
        t.lo = lo;
        t.hi = hi;
        t.lds_bytes = tile_lds_bytes(hi - lo, t.mhi - t.mlo, L.n_ell, L.ntheta, L.constrained,
                                     L.dissipate, L.kerr);
      };
      // Largest row count whose tiles (boundary-closure tiles included) all
      // fit the LDS budget, capped for parallelism.
      int rows = rows_set > 0 ? rows_set : std::min(256, parallel_rows);
      rows = std::max(1, std::min(rows, n));
      std::vector<TileDesc> line_tiles;
      for (;;) {
        line_tiles.clear();
        const int count = (n + rows - 1) / rows;
        const int basec = n / count, rem = n % count;
        int t0 = 0, worst = 0;
        for (int k = 0; k < count; ++k) {
          const int len = basec + (k < rem ? 1 : 0);
          TileDesc t{};
          t.line = li;
          ranges(t0, t0 + len, t);
          worst = std::max(worst, t.lds_bytes);
          line_tiles.push_back(t);
          t0 += len;
        }
        if (worst <= budget || rows == 1) break;
        rows = rows > 16 ? rows - 4 : rows - 1;
      }
      // Kerr lines: the psi_t rows of a tile (output rows plus the stencil
      // halo) go to the four waves in MFMA groups of 16, so prefer tiles
      // whose interior psi_t rows fill whole rounds of 64.
      if (L.kerr && rows_set <= 0 && n > 2 * rows) {
        TileDesc probe{};
        ranges(n / 2 - rows / 2, n / 2 - rows / 2 + rows, probe);
        const int halo = (probe.mhi - probe.mlo) - rows;
        const int aligned = (rows + halo) / 64 * 64 - halo;
        if (aligned >= 16 && aligned < rows) {
          rows = aligned;
          line_tiles.clear();
          const int count = (n + rows - 1) / rows;
          const int basec = n / count, rem = n % count;
          int t0 = 0;
          for (int k = 0; k < count; ++k) {
            const int len = basec + (k < rem ? 1 : 0);
            TileDesc t{};
            t.line = li;
            ranges(t0, t0 + len, t);
            line_tiles.push_back(t);
            t0 += len;
          }
        }
      }
      for (const TileDesc& t : line_tiles) {
        inst_lds[i] = std::max(inst_lds[i], t.lds_bytes);
        per[i].push_back(t);
      }
    }
  }
  // One launch per execution group; a launch reserves its largest tile's LDS
  // for every workgroup.
  std::map<GroupKey, std::vector<int>> buckets;
  for (int i = 0; i < instance_count(); ++i) buckets[inst_key[i]].push_back(i);
  for (const auto& [key, members] : buckets) {
    Group g{std::get<0>(key), std::get<1>(key), std::get<2>(key),
            static_cast<int>(tiles_.size()), 0, 0};
    for (int i : members) {
      Instance& in = inst_[i];
      in.scheme_group = static_cast<int>(groups_.size());
      in.tile_first = static_cast<int>(tiles_.size());
      in.tile_count = static_cast<int>(per[i].size());
      tiles_.insert(tiles_.end(), per[i].begin(), per[i].end());
      g.lds = std::max(g.lds, inst_lds[i]);
    }
    max_lds_ = std::max(max_lds_, g.lds);
    g.count = static_cast<int>(tiles_.size()) - g.first;
    groups_.push_back(g);
  }
  // a stream and a join event per execution group beyond the first, made
  // here (plan creation and the re-tiling set_* calls) so steps create none
  while (side_.size() + 1 < groups_.size()) {
    side_.push_back(create_stream());
    joins_.push_back(create_event(false));
  }
}

void Plan::upload_tables() {
  if (d_lines_) hipFree(d_lines_);
  if (d_tiles_) hipFree(d_tiles_);
  d_lines_ = device_alloc<LineDesc>(lines_.size());
  d_tiles_ = device_alloc<TileDesc>(tiles_.size());
  check(hipMemcpy(d_lines_, lines_.data(), lines_.size() * sizeof(LineDesc), hipMemcpyHostToDevice),
        "upload lines");
  check(hipMemcpy(d_tiles_, tiles_.data(), tiles_.size() * sizeof(TileDesc), hipMemcpyHostToDevice),
        "upload tiles");
  if (!d_mats_) {
    d_mats_ = device_alloc<double>(mats_.size());
    check(hipMemcpy(d_mats_, mats_.data(), mats_.size() * sizeof(double), hipMemcpyHostToDevice),
          "upload matrices");
    d_nodes_ = device_alloc<double>(nodes_.size());
    check(hipMemcpy(d_nodes_, nodes_.data(), nodes_.size() * sizeof(double),
                    hipMemcpyHostToDevice),
          "upload nodes");
    SchemeTable tabs[3] = {make_table(RadialScheme::D42), make_table(RadialScheme::D84),
                           make_table(RadialScheme::D105)};
    d_tables_ = device_alloc<SchemeTable>(3);
    check(hipMemcpy(d_tables_, tabs, sizeof(tabs), hipMemcpyHostToDevice), "upload tables");
    d_line_mode_ = device_alloc<int>(line_mode_.size());
    check(hipMemcpy(d_line_mode_, line_mode_.data(), line_mode_.size() * sizeof(int),
                    hipMemcpyHostToDevice),
          "upload line modes");
  }
  if (max_lds_ > 64 * 1024) check(set_stage_lds_limit(max_lds_), "raise LDS limit");
}

long long Plan::nodal_value_count(int i) const {
  const Instance& in = inst_.at(i);
  return static_cast<long long>(in.M) * 13 * in.N * in.T;
}

double Plan::grid_points_per_step() const {
  double s = 0.0;
  for (const auto& in : inst_) s += static_cast<double>(in.M) * in.N * in.T;
  return s;
}

double Plan::modal_points_per_step() const {
  double s = 0.0;
  for (const auto& L : lines_) s += static_cast<double>(L.nr) * L.n_ell;
  return s;
}

// Gaussian pulse geometry and expanded seeds of one instance, checked as
// in pipeline_initial_data.hpp:211-257.
std::vector<SeedMode> Plan::gaussian_seeds(const Instance& in, double& center, double& width) {
  const RunParameters& p = in.params;
  center = p.center_fraction * in.horizon;
  width = p.width_fraction * in.horizon;
  if (!(width > 0.0) || center < 0.0 || center > in.horizon)
    throw std::invalid_argument("Gaussian center must be on-grid and width must be positive");
  std::vector<SeedMode> expanded = expanded_seeds(p);
  for (const auto& s : expanded) {
    if (!in.registry.is_parent(s.m))
      throw std::invalid_argument("Gaussian seed m is not a first-order mode");
    if (s.ell < std::max(2, std::abs(s.m)) || s.ell > p.ell_max_first)
      throw std::invalid_argument("Gaussian seed ell is outside the band");
  }
  return expanded;
}

void Plan::run_initial_data(InitArgs a, const std::vector<InitLine>& il,
                            const std::vector<cdouble>& seeds, int max_nr, int lds) {
  InitLine* d_il = device_alloc<InitLine>(il.size());
  cdouble* d_seeds = device_alloc<cdouble>(seeds.size());
  unsigned long long* d_res = device_alloc<unsigned long long>(il.size());
  std::vector<unsigned long long> res(il.size(), 0ULL);
  try {
    check(hipMemcpy(d_il, il.data(), il.size() * sizeof(InitLine), hipMemcpyHostToDevice),
          "upload initial data lines");
    check(hipMemcpy(d_seeds, seeds.data(), seeds.size() * sizeof(cdouble), hipMemcpyHostToDevice),
          "upload seeds");
    check(hipMemcpy(d_res, res.data(), res.size() * sizeof(unsigned long long),
                    hipMemcpyHostToDevice),
          "clear residuals");
    a.init = d_il;
    a.seeds = d_seeds;
    a.residual = d_res;
    a.n_lines = static_cast<int>(il.size());
    check(launch_galerkin_init(a, max_nr, lds, stream_), "Galerkin initial data");
    check(hipMemcpyAsync(res.data(), d_res, res.size() * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, stream_),
          "read residuals");
    check(hipStreamSynchronize(stream_), "initialize");
  } catch (...) {
    (void)hipFree(d_il);
    (void)hipFree(d_seeds);
    (void)hipFree(d_res);
    throw;
  }
  (void)hipFree(d_il);
  (void)hipFree(d_seeds);
  (void)hipFree(d_res);
  // singular matrix, then the relative residual bar (dense_solve.hpp:27-90)
  for (unsigned long long bits : res) {
    double r;
    std::memcpy(&r, &bits, sizeof(r));
    if (std::isinf(r)) throw std::runtime_error("singular Galerkin initialization matrix");
    if (!std::isfinite(r) || r > 2.0e-12)
      throw std::runtime_error("Galerkin initial-time-derivative solve failed its residual check");
  }
}

// initialize_pipeline_state (initial_data_factory.hpp:82-110): Gaussian data
// for every instance, then the checkpoint of each instance whose initial
// data are a checkpoint directory (state, latch and progress).
void Plan::initialize() {
  initialize_gaussian();
  for (int i = 0; i < instance_count(); ++i)
    if (inst_[i].params.initial_type == InitialDataKind::Checkpoint)
      load_checkpoint(i, inst_[i].params.checkpoint_directory);
}

void Plan::initialize_gaussian() {
  if (so_mode_) {
    so_initialize();
  } else {
    // One wavefront per (line, radius) on the GPU: Psi, Q = D Psi, the
    // Galerkin solve for P and the modal coefficients of all three fields.
    std::vector<InitLine> il(lines_.size());
    std::vector<cdouble> seeds;
    int lds = 0, max_nr = 0;
    for (int i = 0; i < instance_count(); ++i) {
      const Instance& in = inst_[i];
      const RunParameters& p = in.params;
      double center, width;
      const auto expanded = gaussian_seeds(in, center, width);
      for (int mi = 0; mi < in.M; ++mi) {
        const int li = in.line_of_mode[mi];
        if (li < 0) continue;
        const LineDesc& L = lines_[li];
        InitLine d{};
        d.seed_off = static_cast<int>(seeds.size());
        d.compact = p.compact_support ? 1 : 0;
        d.scheme = static_cast<int>(p.scheme);
        d.center = center;
        d.width = width;
        il[li] = d;
        seeds.resize(seeds.size() + L.n_ell, cdouble{0.0, 0.0});
        for (const auto& s : expanded)
          if (s.m == L.m) {
            cdouble& c = seeds[d.seed_off + s.ell - L.ell_min];
            c.re += s.amplitude.real();
            c.im += s.amplitude.imag();
          }
        lds = std::max(lds, galerkin_init_lds_bytes(L.n_ell, L.ntheta));
        max_nr = std::max(max_nr, L.nr);
      }
    }
    InitArgs a{};
    a.lines = d_lines_;
    a.mats = d_mats_;
    a.nodes = d_nodes_;
    a.tables = d_tables_;
    a.y = d_y_;
    run_initial_data(a, il, seeds, max_nr, lds);
  }
  for (auto& in : inst_) {
    in.passive.clear();
    in.residual.clear();
    in.time = 0.0;
    in.steps_done = 0;
    in.activation = initial_activation(in.params);
  }
}

// Strict band-validated import (pipeline_bands.hpp:108-232) on the GPU.
// Every (mode, field, r) theta row x is projected onto the field's declared
// band (spin of the field, m, ell_max_first or ell_max_second; nothing for a
// mode the field is inactive on) and |x - S A x| / |x| is reduced to its
// maximum over the instance.
double Plan::import_check(int i, const cplx* host, double tol) {
  if (!std::isfinite(tol) || tol < 0.0)
    throw std::invalid_argument("pipeline band tolerance must be nonnegative");
  import_release();
  Instance& in = inst_.at(i);
  const int M = in.M, N = in.N, T = in.T;
  static const int spins[13] = {-2, -2, -2, -1, -2, 0, -2, -1, -1, 0, -2, -2, -2};
  std::vector<ImportRow> rows(static_cast<size_t>(M) * 13);
  std::vector<double> mats;
  std::map<std::tuple<int, int, int>, ImportRow> pool;
  for (int mi = 0; mi < M; ++mi) {
    const int m = in.registry.modes()[mi];
    for (int f = 0; f < 13; ++f) {
      ImportRow R{-1, 0, 0, 0, 0};
      const bool active = f >= 10 ? in.registry.is_target(m) : in.registry.is_parent(m);
      const int emax = f >= 10 ? in.params.ell_max_second : in.params.ell_max_first;
      if (active) {
        const auto key = std::make_tuple(spins[f], m, emax);
        auto it = pool.find(key);
        if (it == pool.end()) {
          const BandTransform bt = make_band_transform(spins[f], m, emax, T);
          ImportRow P{-1, 0, 0, 0, bt.modes() > 0 ? bt.modes() : 0};
          if (P.n > kMaxImportBand)
            throw std::invalid_argument("imported band is wider than the import supports");
          P.s_off = static_cast<int>(mats.size());
          mats.insert(mats.end(), bt.S.begin(), bt.S.end());
          P.a_off = static_cast<int>(mats.size());
          mats.insert(mats.end(), bt.A.begin(), bt.A.end());
          P.c_off = P.a_off;
          if (bt.gram) {
            P.c_off = static_cast<int>(mats.size());
            mats.insert(mats.end(), bt.B.begin(), bt.B.end());
          }
          it = pool.emplace(key, P).first;
        }
        R = it->second;
      }
      if (!so_mode_ && f < 3 && in.line_of_mode[mi] >= 0)
        R.modal_base = lines_[in.line_of_mode[mi]].base;
      rows[static_cast<size_t>(mi) * 13 + f] = R;
    }
  }
  const long long count = nodal_value_count(i);
  if (count > imp_cap_ || static_cast<long long>(M) * 13 * N > imp_rows_cap_ ||
      static_cast<long long>(rows.size()) > imp_row_tables_cap_ ||
      static_cast<long long>(mats.size()) > imp_mats_cap_)
    throw std::logic_error("import staging smaller than the instance (allocate_staging)");
  check(hipMemcpy(d_imp_rows_, rows.data(), rows.size() * sizeof(ImportRow),
                  hipMemcpyHostToDevice),
        "upload import bands");
  if (!mats.empty())
    check(hipMemcpy(d_imp_mats_, mats.data(), mats.size() * sizeof(double),
                    hipMemcpyHostToDevice),
          "upload import operators");
  static_assert(sizeof(cplx) == sizeof(cdouble), "complex layouts agree");
  check(hipMemcpy(d_imp_in_, host, count * sizeof(cdouble), hipMemcpyHostToDevice),
        "stage imported state");
  check(hipMemsetAsync(d_imp_flags_, 0, 2 * sizeof(unsigned long long), stream_),
        "clear import flags");
  ImportArgs a{};
  a.rows = static_cast<const ImportRow*>(d_imp_rows_);
  a.mats = d_imp_mats_;
  a.in = d_imp_in_;
  a.row_norms = d_imp_norms_;
  a.M = M;
  a.N = N;
  a.T = T;
  check(launch_band_check(a, stream_), "band check");
  std::vector<double> norms(2 * static_cast<size_t>(M) * 13 * N);
  check(hipMemcpyAsync(norms.data(), d_imp_norms_, norms.size() * sizeof(double),
                       hipMemcpyDeviceToHost, stream_),
        "read band check");
  check(hipStreamSynchronize(stream_), "band check");
  // The largest relative residual over the rows, each row against its own
  // norm (pipeline_bands.hpp:127-153): a row that is a rounding remainder
  // must still be a band synthesis, which every state this library writes
  // is (its projections are analysis then synthesis).
  double combined = 0.0;
  for (size_t row = 0; row < norms.size() / 2 && std::isfinite(combined); ++row) {
    const double line = norms[2 * row], res = norms[2 * row + 1];
    if (!std::isfinite(line) || !std::isfinite(res)) {
      combined = std::numeric_limits<double>::quiet_NaN();
      break;
    }
    combined = std::max(combined, res / std::max(line, std::numeric_limits<double>::min()));
  }
  if (!std::isfinite(combined) || combined > tol) {
    import_release();
    throw std::runtime_error("imported pipeline state contains meaningful off-band angular content");
  }
  return combined;
}

// Import and diagnostics staging for the largest instance, allocated once at
// plan creation so that import, checkpoint load and the sampled diagnostics
// allocate nothing: the nodal rows and their residuals (or RHS), per-row
// norms, the flags, and the band tables of one instance (at most four bands
// per mode, spins -2, -1, 0 and two band limits, each S, A and G^-1 A).
void Plan::allocate_staging() {
  long long values = 1, rows = 1, row_tables = 1, mats = 1;
  for (int i = 0; i < instance_count(); ++i) {
    const Instance& in = inst_[i];
    values = std::max(values, nodal_value_count(i));
    rows = std::max(rows, static_cast<long long>(in.M) * 13 * in.N);
    row_tables = std::max(row_tables, static_cast<long long>(in.M) * 13);
    mats = std::max(mats, static_cast<long long>(in.M) * 6 * 3 * in.T * in.T);
  }
  d_imp_in_ = device_alloc<cdouble>(values);
  d_imp_res_ = device_alloc<cdouble>(values);
  imp_cap_ = values;
  d_imp_norms_ = device_alloc<double>(2 * rows);
  imp_rows_cap_ = rows;
  d_imp_flags_ = device_alloc<unsigned long long>(2);
  d_imp_rows_ = device_alloc<ImportRow>(row_tables);
  imp_row_tables_cap_ = row_tables;
  d_imp_mats_ = device_alloc<double>(mats);
  imp_mats_cap_ = mats;
}

// The staged rows of an import stay in the staging buffers until the next
// import; nothing is freed.
void Plan::import_release() {}

double Plan::import_state(int i, const cplx* host, double tol) {
  if (so_mode_) return so_import(i, host, tol);
  const double combined = import_check(i, host, tol);
  // Split: evolved fields (0..2 of first-order modes) go to the modal state
  // plus their off-band residual; everything else is kept as passive data.
  Instance& in = inst_.at(i);
  const int M = in.M, N = in.N, T = in.T;
  const long long count = nodal_value_count(i);
  ImportArgs a{};
  a.rows = static_cast<const ImportRow*>(d_imp_rows_);
  a.mats = d_imp_mats_;
  a.in = d_imp_in_;
  a.residual = d_imp_res_;
  a.modal = d_y_;
  a.any_residual = reinterpret_cast<int*>(d_imp_flags_ + 1);
  a.M = M;
  a.N = N;
  a.T = T;
  try {
    check(hipMemsetAsync(d_imp_res_, 0, count * sizeof(cdouble), stream_), "clear residuals");
    check(launch_import_modal(a, stream_), "import split");
    int any = 0;
    check(hipMemcpyAsync(&any, d_imp_flags_ + 1, sizeof(int), hipMemcpyDeviceToHost, stream_),
          "read residual flag");
    check(hipStreamSynchronize(stream_), "import split");
    in.passive.assign(host, host + count);
    for (int mi = 0; mi < M; ++mi) {
      if (in.line_of_mode[mi] < 0) continue;
      for (int f = 0; f < 3; ++f)
        std::fill_n(in.passive.begin() + ((static_cast<size_t>(mi) * 13 + f) * N) * T,
                    static_cast<size_t>(N) * T, cplx(0.0, 0.0));
    }
    in.residual.clear();
    if (any) {
      in.residual.resize(static_cast<size_t>(count));
      check(hipMemcpy(in.residual.data(), d_imp_res_, count * sizeof(cdouble),
                      hipMemcpyDeviceToHost),
            "read residuals");
    }
  } catch (...) {
    import_release();
    throw;
  }
  import_release();
  return combined;
}

void Plan::export_state(int i, cplx* host) {
  if (so_mode_) return so_export(i, host);
  Instance& in = inst_.at(i);
  const int M = in.M, N = in.N, T = in.T;
  check(hipStreamSynchronize(stream_), "export sync");
  const long long count = nodal_value_count(i);
  if (in.passive.empty()) std::fill(host, host + count, cplx(0.0, 0.0));
  else std::copy(in.passive.begin(), in.passive.end(), host);
  for (int mi = 0; mi < M; ++mi) {
    const int li = in.line_of_mode[mi];
    if (li < 0) continue;
    const LineDesc& L = lines_[li];
    std::vector<cdouble> line(static_cast<size_t>(N) * 3 * L.n_ell);
    check(hipMemcpy(line.data(), d_y_ + L.base, line.size() * sizeof(cdouble),
                    hipMemcpyDeviceToHost),
          "read line");
    const BandTransform bt = make_band_transform(-2, L.m, in.params.ell_max_first, T);
    for (int r = 0; r < N; ++r)
      for (int f = 0; f < 3; ++f) {
        std::vector<cplx> c(L.n_ell);
        for (int l = 0; l < L.n_ell; ++l) {
          const cdouble v = line[static_cast<size_t>(r) * 3 * L.n_ell + f * L.n_ell + l];
          c[l] = cplx(v.re, v.im);
        }
        const std::vector<cplx> nod = bt.synthesize(c);
        const size_t o = ((static_cast<size_t>(mi) * 13 + f) * N + r) * T;
        for (int j = 0; j < T; ++j)
          host[o + j] = in.residual.empty() ? nod[j] : nod[j] + in.residual[o + j];
      }
  }
}

hipEvent_t Plan::timing_event() {
  const size_t k = timing_events_used_++;
  while (events_.size() <= k) {
    hipEvent_t e;
    check(hipEventCreate(&e), "hipEventCreate");
    events_.push_back(e);
  }
  return events_[k];
}

// `steps` RK4 steps of every instance.  Execution groups touch disjoint
// lines, so each runs its stage launches on a stream of its own (forked from
// and joined back into the plan's stream): their tails and their different
// bottlenecks (MFMA-heavy wide bands, memory-bound narrow ones) overlap, and
// over several steps a group that finishes its step early goes on with the
// next instead of waiting at a join for the other.  With `capturing` the
// launches go into a hipGraph and are not timed; otherwise, with kernel
// timing on, HIP events bracket every launch.
void Plan::launch_step(double dt_override, bool capturing, int steps) {
  const bool timed = timing_ && !capturing;
  const int ng = static_cast<int>(groups_.size());
  if (side_.size() + 1 < groups_.size()) throw std::logic_error("group streams missing");
  if (ng > 1) check(hipEventRecord(fork_, stream_), "fork");
  for (int gi = 0; gi < ng; ++gi) {
    const Group& g = groups_[gi];
    hipStream_t st = gi == 0 ? stream_ : side_[gi - 1];
    if (gi > 0) check(hipStreamWaitEvent(st, fork_, 0), "fork wait");
    if (g.count > 0) {
      StageArgs a{};
      a.tiles = d_tiles_ + g.first;
      a.lines = d_lines_;
      a.mats = d_mats_;
      a.nodes = d_nodes_;
      a.table = d_tables_ + g.scheme;
      a.y = d_y_;
      a.acc = d_acc_;
      a.dt_override = dt_override;
      const cdouble* in_of[5] = {nullptr, d_y_, d_ya_, d_yb_, d_ya_};
      cdouble* out_of[5] = {nullptr, d_ya_, d_yb_, d_ya_, d_y_};
      for (int k = 0; k < 4 * steps; ++k) {
        const int s = 1 + k % 4;
        a.yin = in_of[s];
        a.yout = out_of[s];
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
          e0 = timing_event();
          check(hipEventRecord(e0, st), "event");
        }
        check(teuk::launch_stage(g.scheme, s, g.ub, g.variant, a, g.count, g.lds, st),
              "stage kernel");
        if (timed) {
          e1 = timing_event();
          check(hipEventRecord(e1, st), "event");
          launch_events_.push_back({e0, e1, gi});
        }
      }
    }
    if (gi > 0) check(hipEventRecord(joins_[gi - 1], st), "join");
  }
  for (int gi = 1; gi < ng; ++gi) check(hipStreamWaitEvent(stream_, joins_[gi - 1], 0), "join wait");
}

// HIP events around one step on the plan's stream when timing is on.
hipEvent_t Plan::step_begin_event() {
  if (!timing_) return nullptr;
  hipEvent_t e = timing_event();
  check(hipEventRecord(e, stream_), "event");
  return e;
}

void Plan::step_end_event(hipEvent_t begin, int steps) {
  if (!begin) return;
  hipEvent_t e = timing_event();
  check(hipEventRecord(e, stream_), "event");
  step_events_.push_back({begin, e, steps});
}

void Plan::capture_graph() {
  if (graph_exec_) {
    hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
  hipGraph_t graph;
  check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    if (so_mode_)
      so_rk_launches(so_->d_dt_step);
    else
      launch_step(0.0, true);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &graph);
    throw;
  }
  check(hipStreamEndCapture(stream_, &graph), "end capture");
  const hipError_t e = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  check(e, "instantiate graph");
  if (graph_multi_) {
    hipGraphExecDestroy(graph_multi_);
    graph_multi_ = nullptr;
  }
  if (so_mode_) return;
  check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    launch_step(0.0, true, kGraphSteps);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &graph);
    throw;
  }
  check(hipStreamEndCapture(stream_, &graph), "end capture");
  const hipError_t e2 = hipGraphInstantiate(&graph_multi_, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  check(e2, "instantiate graph");
}

void Plan::record_time_step(double dt_override) {
  for (auto& in : inst_) {
    in.time += dt_override > 0.0 ? dt_override : in.dt;
    ++in.steps_done;
  }
}

// SpatialPipeline::step (spatial_pipeline.hpp:590-620) with the Disabled
// source policy: one classical RK4 step of every instance.
void Plan::step(double t, double dt) {
  if (!std::isfinite(t) || !std::isfinite(dt) || t < 0.0 || !(dt > 0.0))
    throw std::invalid_argument("pipeline step requires finite positive time data");
  if (!std::isfinite(t + dt)) throw std::invalid_argument("pipeline step end time is not finite");
  if (so_mode_) {
    const bool timed = timing_;
    hipEvent_t e0 = nullptr;
    if (timed) {
      e0 = timing_event();
      check(hipEventRecord(e0, stream_), "event");
    }
    so_step(std::vector<double>(inst_.size(), t), std::vector<double>(inst_.size(), dt));
    if (timed) {
      hipEvent_t e1 = timing_event();
      check(hipEventRecord(e1, stream_), "event");
      step_events_.push_back({e0, e1});
    }
  } else {
    hipEvent_t e0 = step_begin_event();
    launch_step(dt, false);
    step_end_event(e0);
  }
  for (auto& in : inst_) {
    in.time = t + dt;
    ++in.steps_done;
  }
}

void Plan::advance(int steps) {
  if (steps < 0) throw std::invalid_argument("step count must be nonnegative");
  if (so_mode_) {
    if (use_graph_ && !graph_exec_) capture_graph();
    for (int s = 0; s < steps; ++s) {
      std::vector<double> t, dt;
      for (const auto& in : inst_) {
        t.push_back(in.time);
        dt.push_back(in.dt);
      }
      if (use_graph_ && so_graph_step(t, dt)) {
        record_time_step(0.0);
        continue;
      }
      hipEvent_t e0 = nullptr;
      if (timing_) {
        e0 = timing_event();
        check(hipEventRecord(e0, stream_), "event");
      }
      so_step(t, dt);
      if (timing_) {
        hipEvent_t e1 = timing_event();
        check(hipEventRecord(e1, stream_), "event");
        step_events_.push_back({e0, e1});
      }
      record_time_step(0.0);
    }
    return;
  }
  if (use_graph_) {
    if (!graph_exec_) capture_graph();  // made by set_graph; rebuilt after a failed capture
    for (int s = 0; s < steps;) {
      // kGraphSteps steps per replay while that many remain, then one; with
      // timing on, HIP events around each replay (recorded outside the
      // graph) give the GPU time of its steps
      const bool multi = graph_multi_ && steps - s >= kGraphSteps;
      const int n = multi ? kGraphSteps : 1;
      hipEvent_t e0 = step_begin_event();
      check(hipGraphLaunch(multi ? graph_multi_ : graph_exec_, stream_), "graph launch");
      step_end_event(e0, n);
      graph_replays_ += n;
      for (int k = 0; k < n; ++k) record_time_step(0.0);
      s += n;
    }
    return;
  }
  for (int s = 0; s < steps; ++s) {
    hipEvent_t e0 = step_begin_event();
    launch_step(0.0, false);
    step_end_event(e0);
    record_time_step(0.0);
  }
}

long long Plan::modal_extent(int i) const {
  const Instance& in = inst_.at(i);
  long long end = 0;
  for (int line : in.line_of_mode)
    if (line >= 0)
      end = std::max(end, lines_[line].base + static_cast<long long>(in.N) * 3 * lines_[line].n_ell);
  return end;
}

void Plan::evaluate_rhs(int i, const cplx* d_stage, cplx* d_out, double t) {
  (void)t;  // the stage operator is autonomous; the latch is host state
  if (so_mode_) return so_evaluate_rhs(i, d_stage, d_out);
  evaluate_rhs_staged(i, d_stage, d_out, d_ya_, d_k_);
}

void Plan::evaluate_rhs_staged(int i, const cplx* d_stage, cplx* d_out, cdouble* ya,
                               cdouble* k) {
  if (so_mode_) throw std::invalid_argument("second-order plans stage their RHS in the plan");
  Instance& in = inst_.at(i);
  if (in.tile_count == 0) {
    check(launch_fill_zero(reinterpret_cast<cdouble*>(d_out), nodal_value_count(i), stream_), "zero");
    return;
  }
  // Lines of one instance are contiguous in the line table.
  int first_line = -1, n_lines = 0;
  for (int mi = 0; mi < in.M; ++mi)
    if (in.line_of_mode[mi] >= 0) {
      if (first_line < 0) first_line = in.line_of_mode[mi];
      ++n_lines;
    }
  NodalArgs na{d_lines_ + first_line, d_line_mode_ + first_line, d_mats_, n_lines, in.N, in.T,
               in.M};
  check(launch_nodal_to_modal(na, reinterpret_cast<const cdouble*>(d_stage), ya, stream_),
        "analysis");
  const Group& g = groups_[in.scheme_group];
  StageArgs a{};
  a.tiles = d_tiles_ + in.tile_first;
  a.lines = d_lines_;
  a.mats = d_mats_;
  a.nodes = d_nodes_;
  a.table = d_tables_ + g.scheme;
  a.yin = ya;
  a.kout = k;
  check(teuk::launch_stage(g.scheme, 0, g.ub, g.variant, a, in.tile_count, g.lds, stream_),
        "rhs kernel");
  check(launch_fill_zero(reinterpret_cast<cdouble*>(d_out), nodal_value_count(i), stream_), "zero");
  check(launch_modal_to_nodal(na, k, reinterpret_cast<cdouble*>(d_out), stream_), "synthesis");
}

// EndpointWaveformSampler::sample (waveform_output.hpp:79-160).
std::vector<WaveformRecordC> Plan::sample_waveform(int i) {
  if (so_mode_) return so_sample_waveform(i);
  Instance& in = inst_.at(i);
  check(hipStreamSynchronize(stream_), "waveform sync");
  std::vector<WaveformRecordC> out;
  const double L2 = in.params.compactification_length * in.params.compactification_length;
  for (int order = 1; order <= 2; ++order) {
    for (int boundary = 0; boundary < 2; ++boundary) {
      const int r = boundary == 0 ? 0 : in.N - 1;
      const double scale = boundary == 0 ? L2 : in.horizon;
      for (int mi = 0; mi < in.M; ++mi) {
        const int m = in.registry.modes()[mi];
        std::vector<cplx> modal;
        int ell_min;
        if (order == 1) {
          if (!in.registry.is_parent(m)) continue;
          const LineDesc& L = lines_[in.line_of_mode[mi]];
          std::vector<cdouble> row(L.n_ell);
          check(hipMemcpy(row.data(),
                          d_y_ + L.base + static_cast<long long>(r) * 3 * L.n_ell + 2 * L.n_ell,
                          row.size() * sizeof(cdouble), hipMemcpyDeviceToHost),
                "read endpoint");
          ell_min = L.ell_min;
          modal.resize(L.n_ell);
          for (int l = 0; l < L.n_ell; ++l) modal[l] = cplx(row[l].re, row[l].im);
          // The reference analyses the nodal endpoint S c (+ residual): A S c = G c.
          const BandTransform bt = make_band_transform(-2, m, in.params.ell_max_first, in.T);
          modal = bt.gram_apply(modal);
          if (!in.residual.empty()) {
            const size_t o = ((static_cast<size_t>(mi) * 13 + 2) * in.N + r) * in.T;
            const std::vector<cplx> c = bt.analyze(&in.residual[o]);
            for (int l = 0; l < L.n_ell; ++l) modal[l] += c[l];
          }
        } else {
          if (!in.registry.is_target(m)) continue;
          const BandTransform bt = make_band_transform(-2, m, in.params.ell_max_second, in.T);
          ell_min = bt.ell_min;
          if (in.passive.empty()) {
            modal.assign(bt.modes(), cplx(0.0, 0.0));
          } else {
            const size_t o = ((static_cast<size_t>(mi) * 13 + 12) * in.N + r) * in.T;
            modal = bt.analyze(&in.passive[o]);
          }
        }
        for (size_t k = 0; k < modal.size(); ++k) {
          WaveformRecordC rec{};
          rec.step = in.steps_done;
          rec.time = in.time;
          rec.boundary = boundary;
          rec.order = order;
          rec.m = m;
          rec.ell = ell_min + static_cast<int>(k);
          rec.rescaled_re = modal[k].real();
          rec.rescaled_im = modal[k].imag();
          rec.endpoint_re = scale * modal[k].real();
          rec.endpoint_im = scale * modal[k].imag();
          out.push_back(rec);
        }
      }
    }
  }
  return out;
}

void Plan::sync() { check(hipStreamSynchronize(stream_), "sync"); }

// Graph mode captures one RK4 step here (and again when a set_* call
// re-tiles), so teuk_advance only replays it.  Second-order plans take host
// decisions between passes and keep launching directly.
void Plan::set_graph(bool enable) {
  use_graph_ = enable;
  if (graph_exec_) {
    hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
  if (graph_multi_) {
    hipGraphExecDestroy(graph_multi_);
    graph_multi_ = nullptr;
  }
  if (use_graph_) capture_graph();
}

// The caller's stream for every later enqueue (the reference's execution
// space instance, spatial_pipeline.hpp:458-462); null returns to the plan's
// own stream.  Work already enqueued on the previous stream is waited for.
void Plan::set_stream(hipStream_t s) {
  check(hipStreamSynchronize(stream_), "stream switch");
  stream_ = s ? s : own_stream_;
}

void Plan::set_kernel_timing(bool enable) {
  check(hipStreamSynchronize(stream_), "timing sync");
  timing_ = enable;
  timing_events_used_ = 0;
  launch_events_.clear();
  step_events_.clear();
}

static double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float e = 0.f;
  check(hipEventElapsedTime(&e, a, b), "elapsed");
  return e;
}

void Plan::kernel_timing(double* total_ms, long long* launches) {
  check(hipStreamSynchronize(stream_), "timing sync");
  double ms = 0.0;
  for (const auto& le : launch_events_) ms += elapsed_ms(le.begin, le.end);
  *total_ms = ms;
  *launches = static_cast<long long>(launch_events_.size());
}

void Plan::step_timing(double* total_ms, long long* steps) {
  check(hipStreamSynchronize(stream_), "timing sync");
  double ms = 0.0;
  long long n = 0;
  for (const auto& se : step_events_) {
    ms += elapsed_ms(se.first, se.second);
    n += se.steps;
  }
  *total_ms = ms;
  *steps = n;
}

void Plan::group_info(int g, int* scheme, int* ub, int* tiles, int* lds_bytes,
                      double* kernel_ms) {
  if (g < 0 || g >= group_count()) throw std::invalid_argument("group index out of range");
  const Group& gr = groups_[g];
  *scheme = gr.scheme;
  *ub = gr.ub;
  *tiles = gr.count;
  *lds_bytes = gr.lds;
  check(hipStreamSynchronize(stream_), "timing sync");
  double ms = 0.0;
  for (const auto& le : launch_events_)
    if (le.group == g) ms += elapsed_ms(le.begin, le.end);
  *kernel_ms = ms;
}

void Plan::set_lds_budget(int bytes) {
  if (bytes < 16 * 1024 || bytes > 160 * 1024)
    throw std::invalid_argument("LDS budget must be in [16 KiB, 160 KiB]");
  if (so_mode_) return;  // the second-order kernels size their own tiles
  lds_budget_ = bytes;
  build_tiles();
  upload_tables();
  set_graph(use_graph_);
}


void Plan::set_tile_rows(int rows) {
  if (rows < 0) throw std::invalid_argument("tile rows must be nonnegative");
  if (so_mode_) return;
  tile_rows_ = rows;
  build_tiles();
  upload_tables();
  set_graph(use_graph_);
}

// Mixed-precision Gram correction on or off: the line flags change in place
// (no re-tiling, no allocation); graphs read them at every replay.
void Plan::set_mixed_gram(bool enable) {
  if (so_mode_) return;
  mixed_gram_ = enable;
  for (size_t i = 0; i < lines_.size(); ++i)
    lines_[i].gram_bf16 = enable && line_bf16_ok_[i] ? 1 : 0;
  check(hipStreamSynchronize(stream_), "mixed Gram sync");
  check(hipMemcpy(d_lines_, lines_.data(), lines_.size() * sizeof(LineDesc), hipMemcpyHostToDevice),
        "upload lines");
}

void Plan::set_specialized(bool enable) {
  if (so_mode_) return;
  specialized_ = enable;
  build_tiles();
  upload_tables();
  set_graph(use_graph_);
}

Plan::GroupShape Plan::shape_of(const GroupKey& k) const {
  const auto it = shapes_.find(k);
  return it == shapes_.end() ? GroupShape{} : it->second;
}

void Plan::set_group_shape(int scheme, int ub, int variant, const GroupShape& shape) {
  if (scheme < 0 || scheme > 2 || (ub != 1 && ub != 2 && ub != 4) || variant < kStageGeneral ||
      variant > kStageKerrGram)
    throw std::invalid_argument("execution group key out of range");
  if (shape.lds_budget != 0 && (shape.lds_budget < 16 * 1024 || shape.lds_budget > 160 * 1024))
    throw std::invalid_argument("LDS budget must be 0 or in [16 KiB, 160 KiB]");
  if (shape.tile_rows < 0) throw std::invalid_argument("tile rows must be nonnegative");
  if (so_mode_) return;
  shapes_[{scheme, ub, variant}] = shape;
  build_tiles();
  upload_tables();
  set_graph(use_graph_);
}

Plan::GroupShape Plan::group_shape(int g) const {
  if (g < 0 || g >= group_count()) throw std::invalid_argument("group index out of range");
  const Group& gr = groups_[g];
  return shape_of({gr.scheme, gr.ub, gr.variant});
}

int Plan::group_variant(int g) const {
  if (g < 0 || g >= group_count()) throw std::invalid_argument("group index out of range");
  return groups_[g].variant;
}

void Plan::set_activation(int i, const SourceActivation& a) {
  if (a.consecutive_passes < 0 || !std::isfinite(a.activation_time) ||
      !std::isfinite(a.last_eligibility_time) || (a.active && a.activation_time < 0.0) ||
      (!a.active && a.activation_time != -1.0) || a.last_eligibility_time < -1.0)
    throw std::invalid_argument("invalid second-order source activation state");
  inst_.at(i).activation = a;
  if (so_mode_) so_upload_active();
}

}  // namespace teuk
