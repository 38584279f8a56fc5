// Runtime configuration schema v1: the reference's strict `key = value`
// format (include/teuk/config.hpp, include/teuk/run_parameters.hpp), typed
// parameters, validation and the resolved-provenance text.
#pragma once

#include <complex>
#include <string>
#include <vector>

namespace teuk {

inline constexpr int kConfigSchemaVersion = 1;          // run_parameters.hpp:13
inline constexpr const char* kSolverVersion = "0.2.0";  // run_parameters.hpp:14
inline constexpr unsigned kPlus2NormalizationVersion = 2;  // plus2_source.hpp:16-23
inline constexpr const char* kPlus2NormalizationName =
    "plus2_s0_over_r7_complete_field_factor_v2";        // plus2_source.hpp:24-25

enum class RadialScheme { D42 = 0, D84 = 1, D105 = 2 };
enum class Reduction { FreeDamped = 0, StageConstrained = 1 };
enum class SourceMode { Disabled = 0, ConstraintAware = 1, Unrestricted = 2 };
enum class InitialDataKind { Gaussian = 0, Checkpoint = 1 };
enum class Plus2Mode { Disabled, DiagnosticOnly, Concurrent, Replay };
enum class Plus2Linear { MetricCurvature, Tsi, Both };
enum class Plus2Second { SourcedCompanion };
enum class Plus2Initial { Zero, Checkpoint };

const char* scheme_name(RadialScheme s);
RadialScheme parse_scheme(const std::string& text);
int scheme_minimum_points(RadialScheme s);
double scheme_dissipation_radius_bound(RadialScheme s);
bool dissipation_step_admissible(RadialScheme s, double strength, double ratio);

struct SeedMode {
  int ell = 2;
  int m = 2;
  std::complex<double> amplitude{1.0e-3, 0.0};
};

// Typed parameters with the reference defaults (run_parameters.hpp:20-109).
struct RunParameters {
  int config_version = kConfigSchemaVersion;
  double mass = 1.0, spin = 0.0, compactification_length = 1.0;
  int nr = 65, ntheta = 7, ell_max_first = 4, ell_max_second = 4;
  std::vector<int> first_order_modes{-2, 2};
  std::vector<int> second_order_modes{-2, 2};
  double final_time = 2.0e-3;
  int steps = 20;
  double cfl = 0.1;
  double reduction_damping = 0.1, dissipation = 0.01;
  Reduction reduction = Reduction::FreeDamped;
  RadialScheme scheme = RadialScheme::D42;
  InitialDataKind initial_type = InitialDataKind::Gaussian;
  std::vector<SeedMode> seeds{SeedMode{}};
  bool add_sharp_partner = false, compact_support = false;
  double center_fraction = 0.45, width_fraction = 0.12;
  std::string time_derivative = "zero";
  std::string checkpoint_directory;
  bool second_order_enabled = false;
  SourceMode source_mode = SourceMode::ConstraintAware;
  double source_start_time = 0.0;
  double constraint_tolerance = 1.0e-8;
  int required_consecutive_passes = 1;
  bool allow_truncated_daughter_modes = false;
  bool plus2_enabled = false;
  Plus2Mode plus2_mode = Plus2Mode::Disabled;
  Plus2Linear plus2_linear = Plus2Linear::MetricCurvature;
  bool plus2_linear_evolve_validation = false;
  Plus2Second plus2_second = Plus2Second::SourcedCompanion;
  Plus2Initial plus2_initial = Plus2Initial::Zero;
  std::string plus2_checkpoint;
  int plus2_ell_max_first = 4, plus2_ell_max_second = 4;
  bool plus2_out_regularized = true, plus2_out_physical = true;
  bool plus2_out_families = true, plus2_out_pairs = false;
  std::string output_directory = "teuk-output";
  int diagnostic_interval = 5, checkpoint_interval = 0;

  double horizon_radius() const;          // L^2 / r_+
  double radial_spacing() const;          // R_H / (nr - 1)
  double time_step() const;               // final_time / steps
};

// Parse + merge overrides + resolve + validate (config.hpp:686-713).
// Throws std::invalid_argument on schema violations.
RunParameters parse_configuration_text(const std::string& text,
                                       const std::vector<std::string>& overrides);
void validate_parameters(const RunParameters& p);
std::string resolved_configuration_text(const RunParameters& p);

// Quadratic daughter completeness (modes.hpp:218-264).
std::vector<int> missing_daughter_modes(const std::vector<int>& first,
                                        const std::vector<int>& second);

}  // namespace teuk
