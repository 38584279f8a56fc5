// Pipeline checkpoints, format version 5 (pipeline_checkpoint.hpp:40-483,
// 782-943): a directory holding metadata.txt (key=value lines) and state.bin
// (the 13-field state as interleaved little-endian float64 re/im in
// (mode, field, radial, theta) order, io.hpp:96-99).  A checkpoint is written
// under a unique sibling directory and published with one rename; loading
// checks the key set, every value, the registry, the numerical configuration,
// the geometry, the byte count, the FNV-1a checksum and the band content
// before the plan's state changes.  Files are interchangeable with the
// reference's own checkpoints of the same run.
#include <algorithm>
#include <atomic>
#include <bit>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "plan.hpp"

namespace teuk {

namespace {

namespace fs = std::filesystem;

constexpr std::uint64_t kFormatVersion = 5;
constexpr std::uint64_t kLegacyD42Version = 4;  // no radial_discretization key
constexpr const char* kMetadataFile = "metadata.txt";
constexpr const char* kStateFile = "state.bin";

// The parts of a run a checkpoint records (PipelineCheckpointConfiguration,
// pipeline_checkpoint.hpp:49-60).
struct CheckpointConfig {
  double mass = 0, spin = 0, length = 0;
  int ell_max_first = 0, ell_max_second = 0, theta_nodes = 0;
  double reduction_damping = 0, dissipation = 0;
  Reduction reduction = Reduction::FreeDamped;
  RadialScheme scheme = RadialScheme::D42;
  double time_step = 0;
  SourceMode source_mode = SourceMode::Disabled;
  double source_start_time = 0, tolerance = 0;
  int required_passes = 1;
};

struct Metadata {
  std::uint64_t version = kFormatVersion;
  std::uint64_t modes = 0, fields = 13, radial = 0, theta = 0;
  std::vector<int> m_values, parents, targets;
  std::vector<double> radial_coordinates, theta_coordinates;
  CheckpointConfig config;
  double time = 0;
  std::uint64_t step = 0;
  SourceActivation activation;
  std::uint64_t checksum = 0;
};

const char* reduction_text(Reduction r) {
  return r == Reduction::FreeDamped ? "free_damped" : "stage_constrained";
}
Reduction parse_reduction_text(const std::string& t) {
  if (t == "free_damped") return Reduction::FreeDamped;
  if (t == "stage_constrained") return Reduction::StageConstrained;
  throw std::runtime_error("unsupported checkpoint reduction evolution");
}
const char* source_mode_text(SourceMode m) {
  return m == SourceMode::Disabled ? "disabled"
         : m == SourceMode::ConstraintAware ? "constraint_aware" : "unrestricted";
}
SourceMode parse_source_mode_text(const std::string& t) {
  if (t == "disabled") return SourceMode::Disabled;
  if (t == "constraint_aware") return SourceMode::ConstraintAware;
  if (t == "unrestricted") return SourceMode::Unrestricted;
  throw std::runtime_error("unsupported checkpoint second-order source mode");
}

CheckpointConfig config_of(const Plan::Instance& in) {
  const RunParameters& p = in.params;
  CheckpointConfig c;
  c.mass = p.mass;
  c.spin = p.spin;
  c.length = p.compactification_length;
  c.ell_max_first = p.ell_max_first;
  c.ell_max_second = p.ell_max_second;
  c.theta_nodes = p.ntheta;
  c.reduction_damping = p.reduction_damping;
  c.dissipation = p.dissipation;
  c.reduction = p.reduction;
  c.scheme = p.scheme;
  c.time_step = in.dt;
  // make_source_policy: a disabled second order records the disabled policy
  c.source_mode = p.second_order_enabled ? p.source_mode : SourceMode::Disabled;
  c.source_start_time = p.source_start_time;
  c.tolerance = p.constraint_tolerance;
  c.required_passes = p.required_consecutive_passes;
  return c;
}

bool same_config(const CheckpointConfig& a, const CheckpointConfig& b) {
  return a.mass == b.mass && a.spin == b.spin && a.length == b.length &&
         a.ell_max_first == b.ell_max_first && a.ell_max_second == b.ell_max_second &&
         a.theta_nodes == b.theta_nodes && a.reduction_damping == b.reduction_damping &&
         a.dissipation == b.dissipation && a.reduction == b.reduction && a.scheme == b.scheme &&
         a.time_step == b.time_step && a.source_mode == b.source_mode &&
         a.source_start_time == b.source_start_time && a.tolerance == b.tolerance &&
         a.required_passes == b.required_passes;
}

void require_finite(double v, const char* what) {
  if (!std::isfinite(v)) throw std::invalid_argument(std::string(what) + " must be finite");
}

// validate_configuration (pipeline_checkpoint.hpp:134-171)
void validate_config(const CheckpointConfig& c) {
  require_finite(c.mass, "checkpoint mass");
  require_finite(c.spin, "checkpoint spin");
  require_finite(c.length, "checkpoint compactification length");
  require_finite(c.reduction_damping, "checkpoint reduction damping");
  require_finite(c.dissipation, "checkpoint dissipation");
  require_finite(c.time_step, "checkpoint time step");
  require_finite(c.source_start_time, "checkpoint source start time");
  require_finite(c.tolerance, "checkpoint source constraint tolerance");
  if (!(c.mass > 0.0) || std::abs(c.spin) > c.mass || !(c.length > 0.0) ||
      c.ell_max_first < 3 || c.ell_max_second < 3 ||
      c.theta_nodes < std::max({c.ell_max_first + 1, c.ell_max_second + 1,
                                (2 * c.ell_max_first + c.ell_max_second + 2) / 2}) ||
      c.reduction_damping < 0.0 || c.dissipation < 0.0 || !(c.time_step > 0.0) ||
      c.source_start_time < 0.0 || c.tolerance < 0.0 || c.required_passes < 1)
    throw std::invalid_argument("invalid pipeline checkpoint configuration");
}

// FNV-1a over the bytes of every double, real then imaginary
// (pipeline_checkpoint.hpp:283-300).
std::uint64_t fnv1a(const cplx* v, size_t n) {
  std::uint64_t h = 14695981039346656037ULL;
  const std::uint64_t prime = 1099511628211ULL;
  for (size_t i = 0; i < n; ++i) {
    const double parts[2] = {v[i].real(), v[i].imag()};
    for (double d : parts) {
      unsigned char b[8];
      std::memcpy(b, &d, 8);
      for (unsigned char c : b) {
        h ^= static_cast<std::uint64_t>(c);
        h *= prime;
      }
    }
  }
  return h;
}

template <class T>
void write_list(std::ostream& o, const std::vector<T>& v) {
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) o << ',';
    o << v[i];
  }
  o << '\n';
}

// serialize_metadata (pipeline_checkpoint.hpp:419-483): the key order and
// number formatting (max_digits10) of the reference.
std::string serialize(const Metadata& m) {
  std::ostringstream o;
  const CheckpointConfig& c = m.config;
  o << std::setprecision(std::numeric_limits<double>::max_digits10)
    << "format=teuk-kokkos-pipeline-checkpoint\n"
    << "version=" << m.version << '\n'
    << "ordering=mode,field,radial,theta\n"
    << "complex_storage=interleaved_real_imag_float64\n"
    << "byte_order=little\n"
    << "state_file=" << kStateFile << '\n'
    << "modes=" << m.modes << '\n'
    << "fields=" << m.fields << '\n'
    << "radial=" << m.radial << '\n'
    << "theta=" << m.theta << '\n'
    << "m_values=";
  write_list(o, m.m_values);
  o << "target_values=";
  write_list(o, m.targets);
  o << "parent_values=";
  write_list(o, m.parents);
  o << "radial_coordinates=";
  write_list(o, m.radial_coordinates);
  o << "theta_coordinates=";
  write_list(o, m.theta_coordinates);
  o << "mass=" << c.mass << '\n'
    << "spin=" << c.spin << '\n'
    << "compactification_length=" << c.length << '\n'
    << "ell_max_first=" << c.ell_max_first << '\n'
    << "ell_max_second=" << c.ell_max_second << '\n'
    << "theta_nodes=" << c.theta_nodes << '\n'
    << "reduction_damping=" << c.reduction_damping << '\n'
    << "dissipation=" << c.dissipation << '\n'
    << "reduction=" << reduction_text(c.reduction) << '\n'
    << "radial_discretization=" << scheme_name(c.scheme) << '\n'
    << "time_step=" << c.time_step << '\n'
    << "source_mode=" << source_mode_text(c.source_mode) << '\n'
    << "source_start_time=" << c.source_start_time << '\n'
    << "source_normalized_constraint_tolerance=" << c.tolerance << '\n'
    << "source_required_consecutive_passes=" << c.required_passes << '\n'
    << "source_active=" << (m.activation.active ? 1 : 0) << '\n'
    << "source_activation_time=" << m.activation.activation_time << '\n'
    << "source_consecutive_passes=" << m.activation.consecutive_passes << '\n'
    << "source_last_eligibility_time=" << m.activation.last_eligibility_time << '\n'
    << "time=" << m.time << '\n'
    << "step=" << m.step << '\n'
    << "state_checksum_fnv1a64=" << m.checksum << '\n';
  if (!o) throw std::runtime_error("failed serializing checkpoint metadata");
  return o.str();
}

// read_entries (pipeline_checkpoint.hpp:580-625): one key=value per line,
// no duplicates, exactly the format's key set.
std::map<std::string, std::string> read_entries(const std::string& text) {
  std::map<std::string, std::string> e;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    const size_t d = line.find('=');
    if (d == std::string::npos || d == 0) throw std::runtime_error("malformed checkpoint metadata line");
    const std::string key = line.substr(0, d);
    if (!e.emplace(key, line.substr(d + 1)).second)
      throw std::runtime_error("duplicate checkpoint metadata key: " + key);
  }
  std::set<std::string> expected{
      "format", "version", "ordering", "complex_storage", "byte_order", "state_file", "modes",
      "fields", "radial", "theta", "m_values", "parent_values", "target_values",
      "radial_coordinates", "theta_coordinates", "mass", "spin", "compactification_length",
      "ell_max_first", "ell_max_second", "theta_nodes", "reduction_damping", "dissipation",
      "reduction", "radial_discretization", "time_step", "source_mode", "source_start_time",
      "source_normalized_constraint_tolerance", "source_required_consecutive_passes",
      "source_active", "source_activation_time", "source_consecutive_passes",
      "source_last_eligibility_time", "time", "step", "state_checksum_fnv1a64"};
  const auto v = e.find("version");
  if (v != e.end() && v->second == std::to_string(kLegacyD42Version))
    expected.erase("radial_discretization");
  if (e.size() != expected.size())
    throw std::runtime_error("checkpoint metadata key set does not match format");
  for (const auto& kv : e)
    if (!expected.count(kv.first)) throw std::runtime_error("unknown checkpoint metadata key: " + kv.first);
  return e;
}

const std::string& require_text(const std::map<std::string, std::string>& e, const char* key) {
  const auto f = e.find(key);
  if (f == e.end()) throw std::runtime_error(std::string("missing checkpoint key: ") + key);
  return f->second;
}

template <class T>
T parse_value(const std::string& text, const char* key) {
  T v{};
  const auto r = std::from_chars(text.data(), text.data() + text.size(), v);
  if (text.empty() || r.ec != std::errc{} || r.ptr != text.data() + text.size())
    throw std::runtime_error(std::string("invalid checkpoint value for ") + key);
  return v;
}

template <class T>
T number(const std::map<std::string, std::string>& e, const char* key) {
  return parse_value<T>(require_text(e, key), key);
}

template <class T>
std::vector<T> list(const std::map<std::string, std::string>& e, const char* key) {
  const std::string& text = require_text(e, key);
  std::vector<T> out;
  size_t b = 0;
  while (b < text.size()) {
    const size_t d = text.find(',', b);
    const size_t end = d == std::string::npos ? text.size() : d;
    const std::string tok = text.substr(b, end - b);
    if (tok.empty()) throw std::runtime_error(std::string("invalid checkpoint list: ") + key);
    T v{};
    const auto r = std::from_chars(tok.data(), tok.data() + tok.size(), v);
    if (r.ec != std::errc{} || r.ptr != tok.data() + tok.size())
      throw std::runtime_error(std::string("invalid checkpoint list: ") + key);
    out.push_back(v);
    if (d == std::string::npos) break;
    b = d + 1;
  }
  return out;
}

// validate_metadata (pipeline_checkpoint.hpp:341-417)
void validate(const Metadata& m) {
  if ((m.version != kFormatVersion && m.version != kLegacyD42Version) || m.fields != 13 ||
      m.modes != m.m_values.size() || m.radial != m.radial_coordinates.size() ||
      m.theta != m.theta_coordinates.size() ||
      m.config.theta_nodes != static_cast<int>(m.theta) || m.modes == 0 || m.radial == 0 ||
      m.theta == 0)
    throw std::runtime_error("inconsistent pipeline checkpoint metadata");
  auto sorted_unique = [](const std::vector<int>& v) {
    return !v.empty() && std::is_sorted(v.begin(), v.end()) &&
           std::adjacent_find(v.begin(), v.end()) == v.end();
  };
  if (!sorted_unique(m.m_values) || !sorted_unique(m.parents) || !sorted_unique(m.targets))
    throw std::runtime_error("checkpoint modes and targets must be nonempty, sorted, and unique");
  try {
    const ModeRegistry reg(m.m_values, m.parents, m.targets);
    if (!reg.closed_under_sharp()) throw std::runtime_error("checkpoint modes are not sharp closed");
  } catch (const std::invalid_argument&) {
    throw std::runtime_error("invalid checkpoint mode registry");
  }
  try {
    validate_config(m.config);
  } catch (const std::invalid_argument&) {
    throw std::runtime_error("invalid pipeline checkpoint configuration");
  }
  if (!std::isfinite(m.time)) throw std::invalid_argument("checkpoint time must be finite");
  if (m.time < 0.0) throw std::runtime_error("checkpoint time must be nonnegative");
  // validate_source_activation_state (pipeline_checkpoint.hpp:207-220)
  const SourceActivation& a = m.activation;
  if (a.consecutive_passes < 0 || !std::isfinite(a.activation_time) ||
      !std::isfinite(a.last_eligibility_time) ||
      (a.active && (a.activation_time < 0.0 || a.activation_time > m.time)) ||
      (!a.active && a.activation_time != -1.0) || a.last_eligibility_time < -1.0 ||
      a.last_eligibility_time > m.time)
    throw std::runtime_error("invalid checkpoint source activation state");
  for (double r : m.radial_coordinates)
    if (!std::isfinite(r)) throw std::invalid_argument("checkpoint radial coordinate must be finite");
  for (double t : m.theta_coordinates)
    if (!std::isfinite(t)) throw std::invalid_argument("checkpoint theta coordinate must be finite");
  const auto& R = m.radial_coordinates;
  const auto& th = m.theta_coordinates;
  const bool up = std::is_sorted(th.begin(), th.end());
  const bool down = std::is_sorted(th.rbegin(), th.rend());
  if (static_cast<int>(R.size()) < scheme_minimum_points(m.config.scheme) ||
      !std::is_sorted(R.begin(), R.end()) || std::adjacent_find(R.begin(), R.end()) != R.end() ||
      (!up && !down) || std::adjacent_find(th.begin(), th.end()) != th.end())
    throw std::runtime_error("invalid checkpoint spatial geometry");
  // uniform radial grid to two binary64 ulps
  const double spacing = (R.back() - R.front()) / static_cast<double>(R.size() - 1);
  for (size_t i = 0; i < R.size(); ++i) {
    double lo = R.front() + spacing * static_cast<double>(i), hi = lo;
    for (int u = 0; u < 2; ++u) {
      lo = std::nextafter(lo, -std::numeric_limits<double>::infinity());
      hi = std::nextafter(hi, std::numeric_limits<double>::infinity());
    }
    if (R[i] < lo || R[i] > hi) throw std::runtime_error("checkpoint radial grid is not uniform");
  }
}

Metadata parse(const std::map<std::string, std::string>& e) {
  if (require_text(e, "format") != "teuk-kokkos-pipeline-checkpoint" ||
      require_text(e, "ordering") != "mode,field,radial,theta" ||
      require_text(e, "complex_storage") != "interleaved_real_imag_float64" ||
      require_text(e, "byte_order") != "little" || require_text(e, "state_file") != kStateFile)
    throw std::runtime_error("unsupported pipeline checkpoint format");
  Metadata m;
  m.version = number<std::uint64_t>(e, "version");
  m.modes = number<std::uint64_t>(e, "modes");
  m.fields = number<std::uint64_t>(e, "fields");
  m.radial = number<std::uint64_t>(e, "radial");
  m.theta = number<std::uint64_t>(e, "theta");
  m.m_values = list<int>(e, "m_values");
  m.parents = list<int>(e, "parent_values");
  m.targets = list<int>(e, "target_values");
  m.radial_coordinates = list<double>(e, "radial_coordinates");
  m.theta_coordinates = list<double>(e, "theta_coordinates");
  CheckpointConfig& c = m.config;
  c.mass = number<double>(e, "mass");
  c.spin = number<double>(e, "spin");
  c.length = number<double>(e, "compactification_length");
  c.ell_max_first = number<int>(e, "ell_max_first");
  c.ell_max_second = number<int>(e, "ell_max_second");
  c.theta_nodes = number<int>(e, "theta_nodes");
  c.reduction_damping = number<double>(e, "reduction_damping");
  c.dissipation = number<double>(e, "dissipation");
  c.reduction = parse_reduction_text(require_text(e, "reduction"));
  c.scheme = m.version == kLegacyD42Version ? RadialScheme::D42
                                           : parse_scheme(require_text(e, "radial_discretization"));
  c.time_step = number<double>(e, "time_step");
  c.source_mode = parse_source_mode_text(require_text(e, "source_mode"));
  c.source_start_time = number<double>(e, "source_start_time");
  c.tolerance = number<double>(e, "source_normalized_constraint_tolerance");
  c.required_passes = number<int>(e, "source_required_consecutive_passes");
  const int active = number<int>(e, "source_active");
  if (active != 0 && active != 1) throw std::runtime_error("invalid checkpoint source-active flag");
  m.activation.active = active;
  m.activation.activation_time = number<double>(e, "source_activation_time");
  m.activation.consecutive_passes = number<int>(e, "source_consecutive_passes");
  m.activation.last_eligibility_time = number<double>(e, "source_last_eligibility_time");
  m.time = number<double>(e, "time");
  m.step = number<std::uint64_t>(e, "step");
  m.checksum = number<std::uint64_t>(e, "state_checksum_fnv1a64");
  validate(m);
  return m;
}

std::string read_file(const fs::path& path, const char* what) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error(std::string("cannot open checkpoint ") + what);
  const auto end = in.tellg();
  if (end < 0) throw std::runtime_error("cannot size checkpoint content");
  std::string bytes(static_cast<size_t>(end), '\0');
  in.seekg(0);
  if (!bytes.empty()) in.read(bytes.data(), static_cast<std::streamsize>(bytes.size()));
  if (!in) throw std::runtime_error("failed reading checkpoint content");
  return bytes;
}

void write_file(const fs::path& path, const void* data, size_t n) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("cannot open checkpoint byte output");
  out.write(static_cast<const char*>(data), static_cast<std::streamsize>(n));
  if (!out) throw std::runtime_error("failed writing checkpoint bytes");
}

// radial and theta coordinates of an instance as the reference's storage
// holds them: R_i = 0 + h i, theta_t = acos(x_t) (pipeline_storage.hpp:
// 80-100, grid.hpp:50-53, angular.hpp:24-26)
void coordinates(const Plan::Instance& in, std::vector<double>& R, std::vector<double>& theta) {
  R.resize(in.N);
  for (int i = 0; i < in.N; ++i) R[i] = 0.0 + in.h * static_cast<double>(i);
  std::vector<double> x, w;
  gauss_legendre(in.T, x, w);
  theta.resize(in.T);
  for (int t = 0; t < in.T; ++t) theta[t] = std::acos(x[t]);
}

}  // namespace

void Plan::write_checkpoint(int i, const std::string& directory) {
  static_assert(std::endian::native == std::endian::little, "checkpoints are little endian");
  const Instance& in = inst_.at(i);
  const fs::path dir(directory);
  if (directory.empty() || dir.filename().empty() || dir.filename() == "." || dir.filename() == "..")
    throw std::invalid_argument("checkpoint destination must name a directory");
  if (fs::exists(dir)) throw std::runtime_error("checkpoint destination already exists");
  Metadata m;
  m.config = config_of(in);
  validate_config(m.config);
  require_finite(in.time, "checkpoint time");
  if (in.time < 0.0) throw std::invalid_argument("checkpoint time must be nonnegative");
  m.modes = in.M;
  m.radial = in.N;
  m.theta = in.T;
  m.m_values = in.registry.modes();
  m.parents = in.registry.parents();
  m.targets = in.registry.targets();
  coordinates(in, m.radial_coordinates, m.theta_coordinates);
  m.time = in.time;
  m.step = in.steps_done;
  m.activation = in.activation;
  validate(m);
  std::vector<cplx> values(static_cast<size_t>(nodal_value_count(i)));
  export_state(i, values.data());
  m.checksum = fnv1a(values.data(), values.size());
  const std::string text = serialize(m);
  // write under a unique sibling, then publish with one rename
  // (pipeline_checkpoint.hpp:558-578, 840-857)
  static std::atomic<std::uint64_t> sequence{0};
  fs::path parent = dir.parent_path();
  if (parent.empty()) parent = ".";
  fs::create_directories(parent);
  const auto tick = static_cast<std::uint64_t>(
      std::chrono::steady_clock::now().time_since_epoch().count());
  fs::path tmp;
  for (int attempt = 0; attempt < 100 && tmp.empty(); ++attempt) {
    const fs::path cand = parent / (dir.filename().string() + ".tmp." + std::to_string(tick) + "." +
                                    std::to_string(sequence.fetch_add(1)));
    std::error_code ec;
    if (fs::create_directory(cand, ec)) tmp = cand;
    else if (ec && ec != std::errc::file_exists)
      throw std::runtime_error("cannot create temporary checkpoint directory");
  }
  if (tmp.empty()) throw std::runtime_error("cannot reserve temporary checkpoint directory");
  try {
    write_file(tmp / kStateFile, values.data(), values.size() * sizeof(cplx));
    write_file(tmp / kMetadataFile, text.data(), text.size());
    fs::rename(tmp, dir);
  } catch (...) {
    std::error_code ignored;
    fs::remove_all(tmp, ignored);
    throw;
  }
}

void Plan::load_checkpoint(int i, const std::string& directory) {
  Instance& in = inst_.at(i);
  const fs::path dir(directory);
  const Metadata m = parse(read_entries(read_file(dir / kMetadataFile, "metadata")));
  const CheckpointConfig mine = config_of(in);
  validate_config(mine);
  if (m.m_values != in.registry.modes() || m.parents != in.registry.parents() ||
      m.targets != in.registry.targets())
    throw std::runtime_error("checkpoint mode registry does not match caller");
  if (!same_config(m.config, mine))
    throw std::runtime_error("checkpoint numerical configuration does not match caller");
  std::vector<double> R, theta;
  coordinates(in, R, theta);
  if (m.modes != static_cast<std::uint64_t>(in.M) || m.radial != static_cast<std::uint64_t>(in.N) ||
      m.theta != static_cast<std::uint64_t>(in.T))
    throw std::runtime_error("checkpoint shape does not match pipeline");
  if (m.radial_coordinates != R || m.theta_coordinates != theta)
    throw std::runtime_error("checkpoint geometry does not match pipeline");
  const std::string bytes = read_file(dir / kStateFile, "content");
  const size_t count = static_cast<size_t>(nodal_value_count(i));
  if (bytes.size() != count * 2 * sizeof(double))
    throw std::runtime_error("binary snapshot size does not match metadata");
  std::vector<cplx> values(count);
  std::memcpy(static_cast<void*>(values.data()), bytes.data(), bytes.size());
  if (fnv1a(values.data(), values.size()) != m.checksum)
    throw std::runtime_error("pipeline checkpoint state checksum mismatch");
  // band check at the restart tolerance, then the state, the latch and the
  // progress (pipeline_checkpoint.hpp:911-932)
  try {
    import_state(i, values.data(), 5.0e-11);
  } catch (const std::runtime_error& e) {
    if (std::string(e.what()).find("off-band") != std::string::npos)
      throw std::runtime_error("pipeline checkpoint contains meaningful off-band angular content");
    throw;
  }
  set_activation(i, m.activation);
  in.time = m.time;
  in.steps_done = m.step;
}

}  // namespace teuk
