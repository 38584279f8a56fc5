// Strict runtime configuration parser (schema v1).  Behaviour follows the
// reference's include/teuk/config.hpp: `#` comments, exactly one `=` per
// line, duplicate/unknown/malformed keys rejected, command-line overrides
// win, validation before any allocation.
#include "config.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

namespace teuk {

namespace {

using Entries = std::map<std::string, std::string>;

std::string trim(const std::string& s) {
  auto ws = [](unsigned char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; };
  size_t b = 0, e = s.size();
  while (b < e && ws(s[b])) ++b;
  while (e > b && ws(s[e - 1])) --e;
  return s.substr(b, e - b);
}

// config.hpp:36-49
std::pair<std::string, std::string> split_assignment(const std::string& line,
                                                     const std::string& where) {
  const size_t eq = line.find('=');
  if (eq == std::string::npos || line.find('=', eq + 1) != std::string::npos)
    throw std::invalid_argument(where + ": expected exactly one key = value");
  std::string key = trim(line.substr(0, eq));
  std::string value = trim(line.substr(eq + 1));
  if (key.empty() || value.empty())
    throw std::invalid_argument(where + ": empty configuration key or value");
  return {key, value};
}

template <class T>
T number(const std::string& text, const std::string& key) {
  T v{};
  const auto r = std::from_chars(text.data(), text.data() + text.size(), v);
  if (r.ec != std::errc{} || r.ptr != text.data() + text.size())
    throw std::invalid_argument("key '" + key + "' has malformed value '" + text + "'");
  return v;
}

bool boolean(const std::string& text, const std::string& key) {
  if (text == "true" || text == "1") return true;
  if (text == "false" || text == "0") return false;
  throw std::invalid_argument("key '" + key + "' requires true, false, 1, or 0");
}

std::vector<int> int_list(const std::string& text, const std::string& key) {
  std::vector<int> out;
  size_t b = 0;
  while (b < text.size()) {
    const size_t comma = text.find(',', b);
    const size_t e = comma == std::string::npos ? text.size() : comma;
    const std::string item = trim(text.substr(b, e - b));
    if (item.empty()) throw std::invalid_argument("key '" + key + "' contains an empty item");
    out.push_back(number<int>(item, key));
    if (comma == std::string::npos) break;
    b = comma + 1;
  }
  if (out.empty()) throw std::invalid_argument("key '" + key + "' requires a nonempty list");
  return out;
}

Entries parse_stream(const std::string& text, const std::string& source) {
  Entries entries;
  std::istringstream in(text);
  std::string line;
  size_t n = 0;
  while (std::getline(in, line)) {
    ++n;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    if (trim(line).empty()) continue;
    auto kv = split_assignment(line, source + ":" + std::to_string(n));
    if (!entries.emplace(kv.first, kv.second).second)
      throw std::invalid_argument(source + ":" + std::to_string(n) + ": duplicate key '" +
                                  kv.first + "'");
  }
  return entries;
}

Entries parse_overrides(const std::vector<std::string>& overrides) {
  Entries entries;
  for (size_t i = 0; i < overrides.size(); ++i) {
    auto kv = split_assignment(overrides[i], "command-line override " + std::to_string(i + 1));
    if (!entries.emplace(kv.first, kv.second).second)
      throw std::invalid_argument("duplicate command-line override '" + kv.first + "'");
  }
  return entries;
}

void finite(double v, const std::string& key) {
  if (!std::isfinite(v)) throw std::invalid_argument("key '" + key + "' must be finite");
}

const char* reduction_name(Reduction r) {
  return r == Reduction::FreeDamped ? "free_damped" : "stage_constrained";
}
const char* source_name(SourceMode m) {
  switch (m) {
    case SourceMode::Disabled: return "disabled";
    case SourceMode::ConstraintAware: return "constraint_aware";
    case SourceMode::Unrestricted: return "unrestricted";
  }
  return "?";
}
const char* plus2_mode_name(Plus2Mode m) {
  switch (m) {
    case Plus2Mode::Disabled: return "disabled";
    case Plus2Mode::DiagnosticOnly: return "diagnostic_only";
    case Plus2Mode::Concurrent: return "concurrent";
    case Plus2Mode::Replay: return "replay";
  }
  return "?";
}
Plus2Mode parse_plus2_mode(const std::string& t) {
  if (t == "disabled") return Plus2Mode::Disabled;
  if (t == "diagnostic_only") return Plus2Mode::DiagnosticOnly;
  if (t == "concurrent") return Plus2Mode::Concurrent;
  if (t == "replay") return Plus2Mode::Replay;
  throw std::invalid_argument("unknown plus2 run mode: " + t);
}
const char* plus2_linear_name(Plus2Linear m) {
  switch (m) {
    case Plus2Linear::MetricCurvature: return "metric_curvature";
    case Plus2Linear::Tsi: return "tsi";
    case Plus2Linear::Both: return "both";
  }
  return "?";
}
Plus2Linear parse_plus2_linear(const std::string& t) {
  if (t == "metric_curvature") return Plus2Linear::MetricCurvature;
  if (t == "tsi") return Plus2Linear::Tsi;
  if (t == "both") return Plus2Linear::Both;
  throw std::invalid_argument("unknown plus2 linear method: " + t);
}
Plus2Second parse_plus2_second(const std::string& t) {
  if (t == "sourced_companion") return Plus2Second::SourcedCompanion;
  throw std::invalid_argument("unknown plus2 second method: " + t);
}
const char* plus2_initial_name(Plus2Initial p) {
  return p == Plus2Initial::Zero ? "zero" : "checkpoint";
}
Plus2Initial parse_plus2_initial(const std::string& t) {
  if (t == "zero") return Plus2Initial::Zero;
  if (t == "checkpoint") return Plus2Initial::Checkpoint;
  throw std::invalid_argument("unknown plus2 initial policy: " + t);
}

struct SeedBuilder {
  SeedMode mode;
  bool ell = false, m = false, re = false, im = false;
};

// initial_data.mode.<index>.<field> (config.hpp:146-160)
bool seed_key(const std::string& key, size_t& index, std::string& field) {
  static const std::string prefix = "initial_data.mode.";
  if (key.compare(0, prefix.size(), prefix) != 0) return false;
  const size_t dot = key.find('.', prefix.size());
  if (dot == std::string::npos) return false;
  const std::string idx = key.substr(prefix.size(), dot - prefix.size());
  if (idx.empty()) return false;
  index = number<size_t>(idx, key);
  field = key.substr(dot + 1);
  return field == "ell" || field == "m" || field == "amplitude_real" ||
         field == "amplitude_imag";
}

// resolve_run_parameters (config.hpp:194-434)
RunParameters resolve(const Entries& entries) {
  RunParameters p;
  std::set<std::string> seen;
  auto get = [&](const std::string& k) -> const std::string* {
    auto it = entries.find(k);
    if (it == entries.end()) return nullptr;
    seen.insert(k);
    return &it->second;
  };
  auto dbl = [&](const std::string& k, double& d) {
    if (auto* t = get(k)) d = number<double>(*t, k);
  };
  auto integer = [&](const std::string& k, int& d) {
    if (auto* t = get(k)) d = number<int>(*t, k);
  };
  auto flag = [&](const std::string& k, bool& d) {
    if (auto* t = get(k)) d = boolean(*t, k);
  };
  auto str = [&](const std::string& k, std::string& d) {
    if (auto* t = get(k)) d = *t;
  };

  integer("config_version", p.config_version);
  dbl("mass", p.mass);
  dbl("spin", p.spin);
  dbl("compactification_length", p.compactification_length);
  integer("nr", p.nr);
  integer("ntheta", p.ntheta);
  integer("ellmax_first", p.ell_max_first);
  const bool second_explicit = entries.count("ellmax_second") != 0;
  integer("ellmax_second", p.ell_max_second);
  if (!second_explicit) p.ell_max_second = p.ell_max_first;
  if (auto* t = get("first_order_modes")) p.first_order_modes = int_list(*t, "first_order_modes");
  if (auto* t = get("second_order_modes")) p.second_order_modes = int_list(*t, "second_order_modes");
  dbl("final_time", p.final_time);
  integer("steps", p.steps);
  dbl("cfl", p.cfl);
  dbl("reduction_damping", p.reduction_damping);
  dbl("dissipation", p.dissipation);
  if (auto* t = get("radial_discretization")) p.scheme = parse_scheme(*t);
  if (auto* t = get("reduction_mode")) {
    if (*t == "free_damped") p.reduction = Reduction::FreeDamped;
    else if (*t == "stage_constrained") p.reduction = Reduction::StageConstrained;
    else throw std::invalid_argument("key 'reduction_mode' requires free_damped or stage_constrained");
  }
  if (auto* t = get("initial_data.type")) {
    if (*t == "gaussian") p.initial_type = InitialDataKind::Gaussian;
    else if (*t == "checkpoint") p.initial_type = InitialDataKind::Checkpoint;
    else throw std::invalid_argument("key 'initial_data.type' requires gaussian or checkpoint");
  }
  flag("initial_data.add_sharp_partner", p.add_sharp_partner);
  flag("initial_data.compact_support", p.compact_support);
  dbl("initial_data.center", p.center_fraction);
  dbl("initial_data.width", p.width_fraction);
  str("initial_data.time_derivative", p.time_derivative);
  str("initial_data.checkpoint_directory", p.checkpoint_directory);
  if (p.checkpoint_directory == "none") p.checkpoint_directory.clear();

  std::map<size_t, SeedBuilder> builders;
  builders[0].mode = SeedMode{};
  auto& base = builders[0];
  if (auto* t = get("initial_data.seed_ell")) base.mode.ell = number<int>(*t, "initial_data.seed_ell");
  if (auto* t = get("initial_data.seed_m")) base.mode.m = number<int>(*t, "initial_data.seed_m");
  double base_re = base.mode.amplitude.real(), base_im = base.mode.amplitude.imag();
  if (auto* t = get("initial_data.amplitude_real"))
    base_re = number<double>(*t, "initial_data.amplitude_real");
  if (auto* t = get("initial_data.amplitude_imag"))
    base_im = number<double>(*t, "initial_data.amplitude_imag");
  base.mode.amplitude = {base_re, base_im};
  for (const auto& kv : entries) {
    size_t index = 0;
    std::string field;
    if (!seed_key(kv.first, index, field)) continue;
    seen.insert(kv.first);
    auto& b = builders[index];
    if (field == "ell") {
      b.mode.ell = number<int>(kv.second, kv.first);
      b.ell = true;
    } else if (field == "m") {
      b.mode.m = number<int>(kv.second, kv.first);
      b.m = true;
    } else if (field == "amplitude_real") {
      b.mode.amplitude = {number<double>(kv.second, kv.first), b.mode.amplitude.imag()};
      b.re = true;
    } else {
      b.mode.amplitude = {b.mode.amplitude.real(), number<double>(kv.second, kv.first)};
      b.im = true;
    }
  }
  p.seeds.clear();
  for (const auto& kv : builders) {
    const auto& b = kv.second;
    if (kv.first != 0 && !(b.ell && b.m && b.re && b.im))
      throw std::invalid_argument("initial_data.mode." + std::to_string(kv.first) +
                                  " requires ell, m, amplitude_real, and amplitude_imag");
    p.seeds.push_back(b.mode);
  }

  flag("second_order.enabled", p.second_order_enabled);
  if (auto* t = get("second_order.source_mode")) {
    if (*t == "constraint_aware") p.source_mode = SourceMode::ConstraintAware;
    else if (*t == "unrestricted") p.source_mode = SourceMode::Unrestricted;
    else throw std::invalid_argument("key 'second_order.source_mode' requires constraint_aware or unrestricted");
  }
  dbl("second_order.source_start_time", p.source_start_time);
  dbl("second_order.constraint_tolerance", p.constraint_tolerance);
  integer("second_order.required_consecutive_passes", p.required_consecutive_passes);
  flag("second_order.allow_truncated_daughter_modes", p.allow_truncated_daughter_modes);

  flag("plus2.enabled", p.plus2_enabled);
  if (auto* t = get("plus2.mode")) p.plus2_mode = parse_plus2_mode(*t);
  if (auto* t = get("plus2.linear.method")) p.plus2_linear = parse_plus2_linear(*t);
  flag("plus2.linear.evolve_validation", p.plus2_linear_evolve_validation);
  if (auto* t = get("plus2.second.method")) p.plus2_second = parse_plus2_second(*t);
  if (auto* t = get("plus2.second.initial_policy")) p.plus2_initial = parse_plus2_initial(*t);
  str("plus2.second.checkpoint", p.plus2_checkpoint);
  if (p.plus2_checkpoint == "none") p.plus2_checkpoint.clear();
  const bool p2f = entries.count("plus2.ell_max_first") != 0;
  const bool p2s = entries.count("plus2.ell_max_second") != 0;
  integer("plus2.ell_max_first", p.plus2_ell_max_first);
  integer("plus2.ell_max_second", p.plus2_ell_max_second);
  if (!p2f) p.plus2_ell_max_first = p.ell_max_first;
  if (!p2s) p.plus2_ell_max_second = p.ell_max_second;
  flag("plus2.output.regularized", p.plus2_out_regularized);
  flag("plus2.output.physical_tetrad_field", p.plus2_out_physical);
  flag("plus2.output.source_families", p.plus2_out_families);
  flag("plus2.output.ordered_pairs", p.plus2_out_pairs);
  if (auto* t = get("plus2.source_normalization.version")) {
    if (number<int>(*t, "plus2.source_normalization.version") !=
        static_cast<int>(kPlus2NormalizationVersion))
      throw std::invalid_argument("resolved plus2 source-normalization version is incompatible");
  }
  if (auto* t = get("plus2.source_normalization.name")) {
    if (*t != kPlus2NormalizationName)
      throw std::invalid_argument("resolved plus2 source-normalization name is incompatible");
  }
  str("output.directory", p.output_directory);
  integer("output.diagnostic_every", p.diagnostic_interval);
  integer("output.checkpoint_every", p.checkpoint_interval);

  for (const auto& kv : entries)
    if (!seen.count(kv.first))
      throw std::invalid_argument("unknown configuration key '" + kv.first + "'");
  return p;
}

std::string fmt(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

}  // namespace

const char* scheme_name(RadialScheme s) {
  switch (s) {
    case RadialScheme::D42: return "d4-2";
    case RadialScheme::D84: return "d8-4";
    case RadialScheme::D105: return "d10-5";
  }
  return "?";
}

RadialScheme parse_scheme(const std::string& text) {
  if (text == "d4-2") return RadialScheme::D42;
  if (text == "d8-4") return RadialScheme::D84;
  if (text == "d10-5") return RadialScheme::D105;
  throw std::invalid_argument("unknown radial discretization: " + text);
}

int scheme_minimum_points(RadialScheme s) {
  return s == RadialScheme::D105 ? 22 : (s == RadialScheme::D84 ? 16 : 8);
}

// radial_discretization.hpp:23-67
double scheme_dissipation_radius_bound(RadialScheme s) {
  return s == RadialScheme::D105 ? 5500.0 : (s == RadialScheme::D84 ? 1371.0 : 64.0);
}

bool dissipation_step_admissible(RadialScheme s, double strength, double ratio) {
  if (!std::isfinite(strength) || strength < 0.0 || !std::isfinite(ratio) || ratio < 0.0)
    return false;
  if (strength == 0.0) return true;
  const double max_ratio =
      0.8 * 2.785293563405282 / (strength * scheme_dissipation_radius_bound(s));
  return ratio <= max_ratio;
}

double RunParameters::horizon_radius() const {
  const double disc = mass * mass - spin * spin;
  const double rp = mass + std::sqrt(disc > 0.0 ? disc : 0.0);
  return compactification_length * compactification_length / rp;
}

double RunParameters::radial_spacing() const {
  return (horizon_radius() - 0.0) / static_cast<double>(nr > 1 ? nr - 1 : 1);
}

double RunParameters::time_step() const { return final_time / static_cast<double>(steps); }

std::vector<int> missing_daughter_modes(const std::vector<int>& first,
                                        const std::vector<int>& second) {
  if (first.empty()) throw std::invalid_argument("first-order mode set must not be empty");
  std::set<int> daughters;
  for (int a : first)
    for (int b : first) {
      const long long s = static_cast<long long>(a) + b;
      if (s < std::numeric_limits<int>::min() || s > std::numeric_limits<int>::max())
        throw std::invalid_argument("quadratic daughter mode overflows int");
      daughters.insert(static_cast<int>(s));
    }
  std::vector<int> missing;
  for (int d : daughters)
    if (std::find(second.begin(), second.end(), d) == second.end()) missing.push_back(d);
  return missing;
}

// validate_run_parameters (config.hpp:436-684)
void validate_parameters(const RunParameters& p) {
  if (p.config_version != kConfigSchemaVersion)
    throw std::invalid_argument("config_version must be 1");
  finite(p.mass, "mass");
  finite(p.spin, "spin");
  finite(p.compactification_length, "compactification_length");
  if (!(p.mass > 0.0)) throw std::invalid_argument("mass must be positive");
  if (std::abs(p.spin) > p.mass) throw std::invalid_argument("absolute spin must not exceed mass");
  if (!(p.compactification_length > 0.0))
    throw std::invalid_argument("compactification_length must be positive");
  if (p.nr < scheme_minimum_points(p.scheme))
    throw std::invalid_argument(std::string("nr is too small for radial_discretization=") +
                                scheme_name(p.scheme));
  if (p.ell_max_first < 3 || p.ell_max_second < 3)
    throw std::invalid_argument("ellmax_first and ellmax_second must be at least 3");
  const int product_nodes = (2 * p.ell_max_first + p.ell_max_second + 2) / 2;
  if (p.ntheta < std::max({p.ell_max_first + 1, p.ell_max_second + 1, product_nodes}))
    throw std::invalid_argument("ntheta is too small for the selected first/second angular bands");
  auto check_modes = [](const std::vector<int>& modes, int ell_max, const std::string& key) {
    if (modes.empty()) throw std::invalid_argument(key + " must not be empty");
    std::vector<int> s = modes;
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end())
      throw std::invalid_argument(key + " must contain unique signed modes");
    for (int m : s) {
      if (std::abs(m) > ell_max) throw std::invalid_argument(key + " contains |m| above its ellmax");
      if (!std::binary_search(s.begin(), s.end(), -m))
        throw std::invalid_argument(key + " must be closed under m -> -m");
    }
  };
  check_modes(p.first_order_modes, p.ell_max_first, "first_order_modes");
  check_modes(p.second_order_modes, p.ell_max_second, "second_order_modes");
  if (p.second_order_enabled && !p.allow_truncated_daughter_modes) {
    const auto missing = missing_daughter_modes(p.first_order_modes, p.second_order_modes);
    if (!missing.empty()) {
      std::string msg = "second-order mode set omits quadratic daughters:";
      for (int d : missing) msg += " " + std::to_string(d);
      throw std::invalid_argument(msg);
    }
  }
  finite(p.final_time, "final_time");
  finite(p.cfl, "cfl");
  if (!(p.final_time > 0.0) || p.steps <= 0 || !(p.cfl > 0.0))
    throw std::invalid_argument("final_time, steps, and cfl must be positive");
  finite(p.reduction_damping, "reduction_damping");
  finite(p.dissipation, "dissipation");
  if (p.reduction_damping < 0.0 || p.dissipation < 0.0)
    throw std::invalid_argument("reduction_damping and dissipation must be nonnegative");
  const double rp = p.mass + std::sqrt(p.mass * p.mass - p.spin * p.spin);
  const double horizon = p.compactification_length * p.compactification_length / rp;
  const double spacing = horizon / static_cast<double>(p.nr - 1);
  const double dt = p.final_time / static_cast<double>(p.steps);
  if (!dissipation_step_admissible(p.scheme, p.dissipation, dt / spacing))
    throw std::invalid_argument("steps are too small for the selected radial dissipation RK4 bound");
  if (p.initial_type == InitialDataKind::Gaussian) {
    if (p.seeds.empty()) throw std::invalid_argument("Gaussian initial data require a seed mode");
    std::set<std::pair<int, int>> keys;
    for (const auto& s : p.seeds) {
      finite(s.amplitude.real(), "initial_data amplitude_real");
      finite(s.amplitude.imag(), "initial_data amplitude_imag");
      if (s.ell < std::max(2, std::abs(s.m)) || s.ell > p.ell_max_first)
        throw std::invalid_argument("initial-data seed ell exceeds ellmax_first or minimum ell");
      if (std::find(p.first_order_modes.begin(), p.first_order_modes.end(), s.m) ==
          p.first_order_modes.end())
        throw std::invalid_argument("initial-data seed m is absent from first_order_modes");
      if (!keys.emplace(s.ell, s.m).second)
        throw std::invalid_argument("duplicate initial-data (ell,m) seed");
      if (p.add_sharp_partner && s.m != 0 && keys.count({s.ell, -s.m}))
        throw std::invalid_argument("add_sharp_partner would duplicate an explicit seed");
    }
    finite(p.center_fraction, "initial_data.center");
    finite(p.width_fraction, "initial_data.width");
    if (p.center_fraction < 0.0 || p.center_fraction > 1.0 || !(p.width_fraction > 0.0))
      throw std::invalid_argument("initial_data.center must be in [0,1] and width must be positive");
    if (p.time_derivative != "zero")
      throw std::invalid_argument("initial_data.time_derivative currently supports only 'zero'");
  } else if (p.checkpoint_directory.empty()) {
    throw std::invalid_argument("checkpoint initial data require initial_data.checkpoint_directory");
  }
  finite(p.source_start_time, "second_order.source_start_time");
  finite(p.constraint_tolerance, "second_order.constraint_tolerance");
  if (p.source_start_time < 0.0 || p.constraint_tolerance < 0.0 ||
      p.required_consecutive_passes < 1)
    throw std::invalid_argument("invalid second-order source activation policy");

  const bool plus2_mode_enabled = p.plus2_mode != Plus2Mode::Disabled;
  if (p.plus2_enabled != plus2_mode_enabled)
    throw std::invalid_argument("plus2.enabled must be true exactly when plus2.mode is not disabled");
  if (p.plus2_ell_max_first < 2 || p.plus2_ell_max_second < 2)
    throw std::invalid_argument("plus2 ell_max_first and ell_max_second must be at least 2");
  for (int m : p.first_order_modes)
    if (std::abs(m) > p.plus2_ell_max_first)
      throw std::invalid_argument("first_order_modes contains |m| above the plus2 band");
  for (int m : p.second_order_modes)
    if (std::abs(m) > p.plus2_ell_max_second)
      throw std::invalid_argument("second_order_modes contains |m| above the plus2 band");
  const int p2_nodes = (2 * p.plus2_ell_max_first + p.plus2_ell_max_second + 2) / 2;
  if (p.plus2_enabled &&
      p.ntheta < std::max({p.plus2_ell_max_first + 1, p.plus2_ell_max_second + 1, p2_nodes}))
    throw std::invalid_argument("ntheta is too small for the selected plus2 angular bands");
  const bool ckpt = p.plus2_initial == Plus2Initial::Checkpoint;
  if (ckpt != !p.plus2_checkpoint.empty())
    throw std::invalid_argument(
        "plus2.second.checkpoint is required exactly for checkpoint initial policy");
  if (!p.plus2_checkpoint.empty() && p.plus2_checkpoint.find_first_of("\r\n#=") != std::string::npos)
    throw std::invalid_argument("plus2.second.checkpoint must be a config-safe single-line path");
  if ((p.plus2_mode == Plus2Mode::Disabled || p.plus2_mode == Plus2Mode::DiagnosticOnly) && ckpt)
    throw std::invalid_argument(
        "plus2 disabled/diagnostic_only mode cannot load a second-order companion checkpoint");
  if ((p.plus2_mode == Plus2Mode::Concurrent || p.plus2_mode == Plus2Mode::Replay) &&
      !p.second_order_enabled)
    throw std::invalid_argument(
        "plus2 concurrent/replay requires second_order.enabled=true for four-field evolution");
  if (p.plus2_out_pairs && !p.plus2_out_families)
    throw std::invalid_argument("plus2 ordered-pair output requires source-family output");
  if (p.plus2_enabled && !p.plus2_out_regularized && !p.plus2_out_physical)
    throw std::invalid_argument(
        "enabled plus2 mode must output a regularized or physical tetrad field");
  if (p.output_directory.empty() || p.output_directory.find_first_of("\r\n") != std::string::npos)
    throw std::invalid_argument("output.directory must be a nonempty single line");
  if (p.diagnostic_interval <= 0 || p.checkpoint_interval < 0)
    throw std::invalid_argument(
        "diagnostic cadence must be positive and checkpoint cadence nonnegative");
}

RunParameters parse_configuration_text(const std::string& text,
                                       const std::vector<std::string>& overrides) {
  Entries entries = parse_stream(text, "<configuration>");
  for (const auto& kv : parse_overrides(overrides)) entries[kv.first] = kv.second;
  RunParameters p = resolve(entries);
  validate_parameters(p);
  return p;
}

// resolved_configuration_text (config.hpp:715-831)
std::string resolved_configuration_text(const RunParameters& p) {
  validate_parameters(p);
  std::ostringstream o;
  auto b = [](bool v) { return v ? "true" : "false"; };
  auto modes = [&](const char* key, const std::vector<int>& v) {
    o << key << " = ";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << v[i];
    o << '\n';
  };
  o << "config_version = " << p.config_version << '\n'
    << "mass = " << fmt(p.mass) << '\n'
    << "spin = " << fmt(p.spin) << '\n'
    << "compactification_length = " << fmt(p.compactification_length) << '\n'
    << "nr = " << p.nr << '\n'
    << "ntheta = " << p.ntheta << '\n'
    << "ellmax_first = " << p.ell_max_first << '\n'
    << "ellmax_second = " << p.ell_max_second << '\n';
  modes("first_order_modes", p.first_order_modes);
  modes("second_order_modes", p.second_order_modes);
  o << "final_time = " << fmt(p.final_time) << '\n'
    << "steps = " << p.steps << '\n'
    << "cfl = " << fmt(p.cfl) << '\n'
    << "reduction_damping = " << fmt(p.reduction_damping) << '\n'
    << "dissipation = " << fmt(p.dissipation) << '\n'
    << "radial_discretization = " << scheme_name(p.scheme) << '\n'
    << "reduction_mode = " << reduction_name(p.reduction) << '\n'
    << "initial_data.type = "
    << (p.initial_type == InitialDataKind::Gaussian ? "gaussian" : "checkpoint") << '\n'
    << "initial_data.add_sharp_partner = " << b(p.add_sharp_partner) << '\n'
    << "initial_data.compact_support = " << b(p.compact_support) << '\n'
    << "initial_data.center = " << fmt(p.center_fraction) << '\n'
    << "initial_data.width = " << fmt(p.width_fraction) << '\n'
    << "initial_data.time_derivative = " << p.time_derivative << '\n'
    << "initial_data.checkpoint_directory = "
    << (p.checkpoint_directory.empty() ? "none" : p.checkpoint_directory) << '\n';
  for (size_t i = 0; i < p.seeds.size(); ++i) {
    const auto& s = p.seeds[i];
    o << "initial_data.mode." << i << ".ell = " << s.ell << '\n'
      << "initial_data.mode." << i << ".m = " << s.m << '\n'
      << "initial_data.mode." << i << ".amplitude_real = " << fmt(s.amplitude.real()) << '\n'
      << "initial_data.mode." << i << ".amplitude_imag = " << fmt(s.amplitude.imag()) << '\n';
  }
  o << "second_order.enabled = " << b(p.second_order_enabled) << '\n'
    << "second_order.source_mode = " << source_name(p.source_mode) << '\n'
    << "second_order.source_start_time = " << fmt(p.source_start_time) << '\n'
    << "second_order.constraint_tolerance = " << fmt(p.constraint_tolerance) << '\n'
    << "second_order.required_consecutive_passes = " << p.required_consecutive_passes << '\n'
    << "second_order.allow_truncated_daughter_modes = " << b(p.allow_truncated_daughter_modes)
    << '\n'
    << "plus2.enabled = " << b(p.plus2_enabled) << '\n'
    << "plus2.mode = " << plus2_mode_name(p.plus2_mode) << '\n'
    << "plus2.linear.method = " << plus2_linear_name(p.plus2_linear) << '\n'
    << "plus2.linear.evolve_validation = " << b(p.plus2_linear_evolve_validation) << '\n'
    << "plus2.second.method = sourced_companion\n"
    << "plus2.second.initial_policy = " << plus2_initial_name(p.plus2_initial) << '\n'
    << "plus2.second.checkpoint = " << (p.plus2_checkpoint.empty() ? "none" : p.plus2_checkpoint)
    << '\n'
    << "plus2.ell_max_first = " << p.plus2_ell_max_first << '\n'
    << "plus2.ell_max_second = " << p.plus2_ell_max_second << '\n'
    << "plus2.output.regularized = " << b(p.plus2_out_regularized) << '\n'
    << "plus2.output.physical_tetrad_field = " << b(p.plus2_out_physical) << '\n'
    << "plus2.output.source_families = " << b(p.plus2_out_families) << '\n'
    << "plus2.output.ordered_pairs = " << b(p.plus2_out_pairs) << '\n'
    << "plus2.source_normalization.version = " << kPlus2NormalizationVersion << '\n'
    << "plus2.source_normalization.name = " << kPlus2NormalizationName << '\n'
    << "output.directory = " << p.output_directory << '\n'
    << "output.diagnostic_every = " << p.diagnostic_interval << '\n'
    << "output.checkpoint_every = " << p.checkpoint_interval << '\n';
  return o.str();
}

}  // namespace teuk
