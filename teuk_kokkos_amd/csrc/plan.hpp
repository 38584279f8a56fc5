// Plan: the device-resident replacement of the reference's SpatialPipeline
// (include/teuk/spatial_pipeline.hpp:144-620) for the linear first-order
// path, batched over any number of independent instances.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "config.hpp"
#include "device_layout.hpp"
#include "host_math.hpp"
#include "init_layout.hpp"

namespace teuk {

struct WaveformRecordC {
  std::uint64_t step;
  double time;
  int boundary;  // 0 scri, 1 horizon
  int order;     // 1 first (FirstPsi), 2 second (SecondPsi)
  int m;
  int ell;
  double rescaled_re, rescaled_im;
  double endpoint_re, endpoint_im;
};

struct SourceActivation {
  int active = 0;
  int consecutive_passes = 0;
  double activation_time = -1.0;
  double last_eligibility_time = -1.0;
};

class Plan {
 public:
  struct Instance {
    RunParameters params;
    ModeRegistry registry;
    int M = 0, N = 0, T = 0;
    double horizon = 0, h = 0, dt = 0;
    double time = 0;
    std::uint64_t steps_done = 0;
    std::vector<int> line_of_mode;   // stored mode index -> line, -1 if not first order
    std::vector<double> sin_theta, cos_theta;
    // Entries of an imported nodal state that the linear RHS never changes
    // (fields 3..12, inactive modes) and the off-band residual of the
    // evolved fields; empty when the state was generated in band.
    std::vector<cplx> passive;
    std::vector<cplx> residual;
    // per ordered pair of the last diagnostics sample: m1, m2, target, then
    // RMS and maximum of |D| and of |T| over (r, theta)
    std::vector<double> pair_report;
    SourceActivation activation;
    int scheme_group = 0, tile_first = 0, tile_count = 0;
    explicit Instance(RunParameters p);
  };

  explicit Plan(const std::vector<RunParameters>& params);
  ~Plan();
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  int instance_count() const { return static_cast<int>(inst_.size()); }
  const Instance& instance(int i) const { return inst_.at(i); }
  long long nodal_value_count(int i) const;
  long long modal_value_count() const { return modal_count_; }
  // One past the last modal value of instance i's lines (0 without lines).
  long long modal_extent(int i) const;
  double grid_points_per_step() const;    // sum M * N * T
  double modal_points_per_step() const;   // sum over lines N * n_ell
  int tile_count() const { return static_cast<int>(tiles_.size()); }
  int max_lds_bytes() const { return max_lds_; }
  hipStream_t stream() const { return stream_; }

  void initialize();            // every instance from its configuration
  void initialize_gaussian();   // Gaussian data for every instance
  // Checkpoint format v5 of instance i (checkpoint.cpp): write a new
  // directory, or load one whose registry, configuration and geometry match
  // the instance (its state, latch and progress).
  void write_checkpoint(int i, const std::string& directory);
  void load_checkpoint(int i, const std::string& directory);
  double import_state(int i, const cplx* host, double relative_tolerance);
  void export_state(int i, cplx* host);
  void evaluate_rhs(int i, const cplx* d_stage, cplx* d_out, double t);
  // The linear RHS of instance i with its modal staging in `ya` (the stage
  // value) and `k` (the RHS), each modal_value_count() complex values on the
  // plan's device, indexed by line base like the plan's own buffers.
  void evaluate_rhs_staged(int i, const cplx* d_stage, cplx* d_out, cdouble* ya, cdouble* k);
  void step(double t, double dt);  // every instance, explicit dt
  void advance(int steps);         // every instance, its own dt
  std::vector<WaveformRecordC> sample_waveform(int i);
  void sync();
  void set_graph(bool enable);
  // Steps taken as replays of the captured step graph so far.
  long long graph_replays() const { return graph_replays_; }
  void set_stream(hipStream_t s);
  void set_kernel_timing(bool enable);
  void kernel_timing(double* total_ms, long long* launches);
  // GPU time of whole steps (fork of the group streams to their join).
  void step_timing(double* total_ms, long long* steps);
  // Execution groups: one stage launch each (tiles of one radial scheme, ell
  // tile class and occupancy); kernel_ms sums the group's timed launches.
  int group_count() const { return static_cast<int>(groups_.size()); }
  void group_info(int g, int* scheme, int* ub, int* tiles, int* lds_bytes, double* kernel_ms);
  void set_lds_budget(int bytes);   // re-tiles
  void set_tile_rows(int rows);     // 0 = automatic
  // Compile-time line variants for the stage kernels (default on); off runs
  // every line through the general kernel.  Re-tiles.
  void set_specialized(bool enable);
  void set_mixed_gram(bool enable);
  int group_variant(int g) const;
  // Tile shape of one execution group: LDS budget and tile rows (0 = the
  // plan's setting).
  struct GroupShape {
    int lds_budget = 0, tile_rows = 0;
  };
  // Keyed by (scheme, ell-tile class, variant); applies to the group of that
  // key now or after later re-tiling.  Re-tiles.
  void set_group_shape(int scheme, int ub, int variant, const GroupShape& shape);
  GroupShape group_shape(int g) const;
  SourceActivation activation(int i) const { return inst_.at(i).activation; }
  void set_activation(int i, const SourceActivation& a);
  bool second_order() const { return so_mode_; }
  // Constraint norms of the last activation check of instance i
  // (source_activation.hpp:24-37): per family psi3, psi2, h_ll the absolute
  // maximum, normalized maximum, weighted RMS and normalized weighted norm,
  // then the controlling value and the finiteness flag.
  void constraint_norms(int i, double* out14) const;
  // The reference's sampled diagnostics of instance i (solver_driver.hpp:
  // 205-242): one RHS evaluation of the current state at `time`, then the
  // pipeline, independent-constraint, transport-residual and horizon
  // reports, kDiagCount doubles in the DiagOutput layout (diag_plan.hip).
  void sample_diagnostics(int i, double time, double* out);
  static SourceActivation initial_activation(const RunParameters& p);

 private:
  // Validated Gaussian geometry and expanded seeds of one instance.
  static std::vector<SeedMode> gaussian_seeds(const Instance& in, double& center, double& width);
  // Uploads the per-line initial-data descriptors, launches the Galerkin
  // initial data with the line, matrix, node and table pointers in `a` and
  // checks every line's solve residual.
  void run_initial_data(InitArgs a, const std::vector<InitLine>& il,
                        const std::vector<cdouble>& seeds, int max_nr, int lds);
  // second-order graph (so_plan.hip)
  static void so_check_limits(const RunParameters& p, int n_targets);
  void so_build();
  void so_free();
  void so_initialize();
  double so_import(int i, const cplx* host, double tol);
  void so_export(int i, cplx* host);
  void so_rhs(int first, int n, const cdouble* in, cdouble* out, bool constraints);
  void so_rk_pass(const std::vector<double>& dts);
  void so_rk_launches(const double* d_dt);
  bool so_graph_step(const std::vector<double>& t, const std::vector<double>& dt);
  void so_upload_active();
  bool so_consider(int i, double time);
  void so_step(const std::vector<double>& t, const std::vector<double>& dt);
  std::vector<WaveformRecordC> so_sample_waveform(int i);
  void so_evaluate_rhs(int i, const cplx* d_stage, cplx* d_out);
  // Band-validated import on the GPU: stages instance i's host state, checks
  // every theta row against its field's band and returns the largest
  // relative off-band residual (throws above `tol`, the state untouched).
  // On return the staged rows and the band operators of the split stay on
  // the device until import_release().
  double import_check(int i, const cplx* host, double tol);
  void import_release();
  void build_lines();
  void build_tiles();
  void upload_tables();
  void launch_step(double dt_override, bool capturing, int steps = 1);
  void capture_graph();
  hipEvent_t step_begin_event();
  void step_end_event(hipEvent_t begin, int steps = 1);
  void allocate_staging();
  void allocate_diagnostics();
  void record_time_step(double dt_override);

  std::vector<Instance> inst_;
  std::vector<LineDesc> lines_;
  std::vector<int> line_instance_, line_mode_;
  std::vector<double> mats_, nodes_;
  std::vector<TileDesc> tiles_;
  struct Group {
    // ub: ell tiles of 16 per MFMA pass; variant: kStageGeneral, kStageKerr
    // or kStageKerrGram (stage_variant)
    int scheme, ub, variant, first, count, lds;
  };
  std::vector<Group> groups_;
  using GroupKey = std::tuple<int, int, int>;  // (scheme, ub, variant)
  std::map<GroupKey, GroupShape> shapes_;
  GroupShape shape_of(const GroupKey& k) const;
  long long modal_count_ = 0;
  int max_lds_ = 0;
  int lds_budget_ = 80 * 1024;  // two workgroups per CU (160 KiB LDS)
  int tile_rows_ = 0;
  bool specialized_ = true;
  // The Gram correction of eligible bands on BF16 MFMA (LineDesc::gram_bf16);
  // line_bf16_ok_: the band's E and N qualify
  bool mixed_gram_ = false;
  std::vector<char> line_bf16_ok_;

  hipStream_t stream_ = nullptr;      // where work is enqueued (the caller's, or own_stream_)
  hipStream_t own_stream_ = nullptr;
  LineDesc* d_lines_ = nullptr;
  TileDesc* d_tiles_ = nullptr;
  double* d_mats_ = nullptr;
  double* d_nodes_ = nullptr;
  SchemeTable* d_tables_ = nullptr;
  int* d_line_mode_ = nullptr;
  cdouble *d_y_ = nullptr, *d_acc_ = nullptr, *d_ya_ = nullptr, *d_yb_ = nullptr,
          *d_k_ = nullptr;
  // import staging for the largest instance (allocate_staging): its rows,
  // their residuals, per-row norms and the any-residual flag; the band
  // operators of the import in progress
  cdouble *d_imp_in_ = nullptr, *d_imp_res_ = nullptr;
  double* d_imp_norms_ = nullptr;
  unsigned long long* d_imp_flags_ = nullptr;
  long long imp_cap_ = 0, imp_rows_cap_ = 0, imp_row_tables_cap_ = 0, imp_mats_cap_ = 0;
  // sampled-diagnostics tables and reduction buffers (allocate_diagnostics)
  int* d_diag_sharp_ = nullptr;
  double* d_diag_nodes_ = nullptr;
  double* d_diag_partial_ = nullptr;
  int* d_diag_pflags_ = nullptr;
  double* d_diag_out_ = nullptr;
  long long diag_waves_cap_ = 0;
  int diag_m_cap_ = 0, diag_t_cap_ = 0;
  void* d_imp_rows_ = nullptr;
  double* d_imp_mats_ = nullptr;

  bool use_graph_ = false;
  hipGraphExec_t graph_exec_ = nullptr;
  // kGraphSteps RK4 steps in one graph (linear plans): the groups' stage
  // chains run through all of them between one fork and one join
  static constexpr int kGraphSteps = 8;
  hipGraphExec_t graph_multi_ = nullptr;
  long long graph_replays_ = 0;
  bool timing_ = false;
  hipEvent_t timing_event();
  std::vector<hipEvent_t> events_;  // pool, reused after set_kernel_timing
  size_t timing_events_used_ = 0;
  struct LaunchEvents {
    hipEvent_t begin, end;
    int group;
  };
  std::vector<LaunchEvents> launch_events_;
  struct StepEvents {
    hipEvent_t first, second;
    int steps = 1;  // RK4 steps between the two events
  };
  std::vector<StepEvents> step_events_;
  std::vector<hipStream_t> side_;  // streams of execution groups 1..n-1
  std::vector<hipEvent_t> joins_;
  hipEvent_t fork_ = nullptr;

  // second-order graph state
  bool so_mode_ = false;
  struct SoHost;
  std::unique_ptr<SoHost> so_;
};

// Plan::evaluate_rhs of a linear plan with the modal staging in a fixed
// per-device array of two scratch_rhs_capacity() halves (scratch_rhs.hip).
void evaluate_rhs_in_scratch(Plan& plan, int i, const cplx* d_stage, cplx* d_out);
long long scratch_rhs_capacity();

}  // namespace teuk
