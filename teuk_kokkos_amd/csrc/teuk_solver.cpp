// teuk_solver: the reference's command-line driver (src/main.cpp:26-55,
// include/teuk/solver_driver.hpp:86-353) on the HIP path, written purely
// against the C ABI of libteuk_hip.so.
//
//   teuk_solver --config run.cfg [key=value ...]
//   teuk_solver backend
//   teuk_solver --help
//
// Writes resolved_config.cfg, diagnostics.csv (also echoed to stdout),
// source_pairs.csv and waveforms.csv into output.directory at the start and
// every output.diagnostic_every steps, checkpoint_<step> directories every
// output.checkpoint_every steps (solver_driver.hpp:154-317), then the
// evolution timing (step loop only, fenced at sampling steps) to stdout.  A
// run whose initial data are a checkpoint resumes its time and step count.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/teuk_hip.h"

namespace {

std::string value_of(const std::string& resolved, const std::string& key) {
  std::istringstream in(resolved);
  std::string line;
  const std::string prefix = key + " = ";
  while (std::getline(in, line))
    if (line.compare(0, prefix.size(), prefix) == 0) return line.substr(prefix.size());
  return "";
}

void write_waveforms(std::ofstream& out, teuk_plan* plan) {
  int count = 0;
  teuk_sample_waveform(plan, 0, nullptr, 0, &count);
  std::vector<teuk_waveform_record> recs(count > 0 ? count : 1);
  if (teuk_sample_waveform(plan, 0, recs.data(), count, &count) != TEUK_OK)
    throw std::runtime_error(teuk_last_error(plan));
  char buf[512];
  for (int i = 0; i < count; ++i) {
    const auto& r = recs[i];
    std::snprintf(buf, sizeof(buf), "%llu,%.17g,%s,%d,%d,%d,%.17g,%.17g,%.17g,%.17g\n",
                  static_cast<unsigned long long>(r.step), r.time,
                  r.boundary == 0 ? "scri" : "horizon", r.order, r.m, r.ell, r.rescaled_re,
                  r.rescaled_im, r.endpoint_re, r.endpoint_im);
    out << buf;
  }
  out.flush();
}

// surface_gravity (diagnostics.hpp:18-26)
double surface_gravity(double mass, double spin) {
  if (mass <= 0.0 || std::abs(spin) > mass) return std::nan("");
  const double root = std::sqrt(mass * mass - spin * spin);
  const double rp = mass + root, rm = mass - root;
  return (rp - rm) / (2.0 * (rp * rp + spin * spin));
}

const char* kDiagnosticsHeader =
    "step,time,kappa_time,first_psi_rms,second_psi_rms,"
    "first_constraint_rms,second_constraint_rms,source_rms,forcing_rms,"
    "source_active,source_activation_time,source_constraint_abs_max,"
    "source_constraint_normalized_max,source_constraint_normalized_l2,"
    "transport_residual_rms,psi3_bianchi_rms,psi2_bianchi_rms,"
    "hll_reality_rms,first_horizon_d0,first_horizon_d1,first_horizon_d2,"
    "first_horizon_d3,first_horizon_d4,second_horizon_d0,"
    "second_horizon_d1,second_horizon_d2,second_horizon_d3,"
    "second_horizon_d4";

// One diagnostics.csv line and the source_pairs.csv rows of a sample
// (solver_driver.hpp:205-280).
void write_diagnostics(std::ofstream& diag, std::ofstream& pairs, teuk_plan* plan,
                       unsigned long long step, double time, double kappa, bool second) {
  double d[TEUK_DIAG_COUNT];
  if (teuk_sample_diagnostics(plan, 0, time, d) != TEUK_OK)
    throw std::runtime_error(teuk_last_error(plan));
  double norms[14] = {};
  if (second && teuk_get_source_constraint_norms(plan, 0, norms) != TEUK_OK)
    throw std::runtime_error(teuk_last_error(plan));
  const double normalized_maximum = std::max({norms[1], norms[5], norms[9]});
  const double normalized_l2 = std::max({norms[3], norms[7], norms[11]});
  std::ostringstream line;
  line << std::setprecision(17) << step << ',' << time << ',' << kappa * time << ','
       << d[TEUK_DIAG_STATE_RMS + 2] << ',' << d[TEUK_DIAG_STATE_RMS + 12] << ','
       << d[TEUK_DIAG_FIRST_CONSTRAINT] << ',' << d[TEUK_DIAG_SECOND_CONSTRAINT] << ','
       << d[TEUK_DIAG_SOURCE] << ',' << d[TEUK_DIAG_FORCING] << ','
       << (d[TEUK_DIAG_SOURCE_ACTIVE] > 0.5 ? 1 : 0) << ',' << d[TEUK_DIAG_ACTIVATION_TIME] << ','
       << d[TEUK_DIAG_SOURCE_CONSTRAINT_MAX] << ',' << normalized_maximum << ',' << normalized_l2
       << ',' << d[TEUK_DIAG_TRANSPORT + 14] << ',' << d[TEUK_DIAG_INDEPENDENT] << ','
       << d[TEUK_DIAG_INDEPENDENT + 2] << ',' << d[TEUK_DIAG_INDEPENDENT + 4];
  for (int k = 0; k < 5; ++k) line << ',' << d[TEUK_DIAG_HORIZON_FIRST + k];
  for (int k = 0; k < 5; ++k) line << ',' << d[TEUK_DIAG_HORIZON_SECOND + k];
  std::cout << line.str() << '\n';
  diag << line.str() << '\n';
  diag.flush();
  int n = 0;
  if (teuk_source_pairs(plan, 0, nullptr, 0, &n) != TEUK_OK)
    throw std::runtime_error(teuk_last_error(plan));
  std::vector<double> rec(static_cast<size_t>(7) * (n > 0 ? n : 1));
  if (teuk_source_pairs(plan, 0, rec.data(), n, &n) != TEUK_OK)
    throw std::runtime_error(teuk_last_error(plan));
  for (int k = 0; k < n; ++k) {
    const double* r = rec.data() + 7 * k;
    pairs << std::setprecision(17) << step << ',' << time << ',' << k << ','
          << static_cast<int>(r[0]) << ',' << static_cast<int>(r[1]) << ','
          << static_cast<int>(r[2]) << ',' << r[3] << ',' << r[4] << ',' << r[5] << ',' << r[6]
          << '\n';
  }
  pairs.flush();
}

int run(const std::string& path, const std::vector<std::string>& overrides) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open configuration file: " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  std::vector<const char*> ov;
  for (const auto& o : overrides) ov.push_back(o.c_str());
  size_t need = 0;
  if (teuk_config_resolve(text.c_str(), ov.data(), static_cast<int>(ov.size()), nullptr, 0,
                          &need) != TEUK_OK)
    throw std::invalid_argument(teuk_config_last_error());
  std::string resolved(need, '\0');
  teuk_config_resolve(text.c_str(), ov.data(), static_cast<int>(ov.size()), resolved.data(), need,
                      &need);
  resolved.resize(need - 1);
  if (value_of(resolved, "plus2.enabled") == "true")
    throw std::runtime_error("plus2 mode '" + value_of(resolved, "plus2.mode") +
                             "' is not production integrated; use plus2.enabled=false and "
                             "plus2.mode=disabled");
  const std::filesystem::path outdir = value_of(resolved, "output.directory");
  std::filesystem::create_directories(outdir);
  {
    std::ofstream cfg(outdir / "resolved_config.cfg");
    cfg << "# teuk_solver executable_version = 0.2.0\n"
        << "# backend = hip gfx950 (modal fused RK4 stage kernels)\n"
        << "# build_precision = complex_binary64\n"
        << "# config_schema_version = 1\n"
        << resolved;
  }
  teuk_plan* plan = nullptr;
  if (teuk_plan_create(text.c_str(), ov.data(), static_cast<int>(ov.size()), &plan) != TEUK_OK)
    throw std::invalid_argument(teuk_config_last_error());
  struct Guard {
    teuk_plan* p;
    ~Guard() { teuk_plan_destroy(p); }
  } guard{plan};
  auto ok = [&](int rc) {
    if (rc != TEUK_OK) throw std::runtime_error(teuk_last_error(plan));
  };
  ok(teuk_initialize(plan));
  const int steps = std::stoi(value_of(resolved, "steps"));
  const int every = std::stoi(value_of(resolved, "output.diagnostic_every"));
  const bool second = value_of(resolved, "second_order.enabled") == "true";
  const double kappa =
      surface_gravity(std::stod(value_of(resolved, "mass")), std::stod(value_of(resolved, "spin")));
  std::ofstream diag(outdir / "diagnostics.csv");
  std::ofstream pairs(outdir / "source_pairs.csv");
  std::ofstream wf(outdir / "waveforms.csv");
  if (!diag || !pairs || !wf) throw std::runtime_error("cannot open spatial diagnostic output");
  pairs << "step,time,pair,m1,m2,target,D_rms,D_max,T_rms,T_max\n";
  wf << "step,time,boundary,order,m,ell,rescaled_re,rescaled_im,endpoint_re,endpoint_im\n";
  std::cout << kDiagnosticsHeader << '\n';
  diag << kDiagnosticsHeader << '\n';
  const int checkpoint_every = std::stoi(value_of(resolved, "output.checkpoint_every"));
  // a checkpoint start resumes its time and step count
  double time = 0.0;
  uint64_t completed = 0;
  ok(teuk_plan_progress(plan, 0, &time, &completed));
  write_diagnostics(diag, pairs, plan, static_cast<unsigned long long>(completed), time, kappa,
                    second);
  write_waveforms(wf, plan);
  double evolution = 0.0;
  const auto total_start = std::chrono::steady_clock::now();
  auto start = std::chrono::steady_clock::now();
  for (int s = 1; s <= steps; ++s) {
    ok(teuk_advance(plan, 1));
    const bool diagnostic_step = s % every == 0 || s == steps;
    const bool checkpoint_step = checkpoint_every > 0 && (s % checkpoint_every == 0 || s == steps);
    if (diagnostic_step || checkpoint_step) {
      ok(teuk_sync(plan));
      evolution += std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    }
    uint64_t done = 0;
    ok(teuk_plan_progress(plan, 0, &time, &done));
    if (diagnostic_step) {
      write_diagnostics(diag, pairs, plan, static_cast<unsigned long long>(done), time, kappa,
                        second);
      write_waveforms(wf, plan);
    }
    if (checkpoint_step) {
      // pipeline_checkpoint_path (solver_driver.hpp:59-64)
      std::ostringstream name;
      name << "checkpoint_" << std::setw(8) << std::setfill('0') << done;
      ok(teuk_checkpoint_write(plan, 0, (outdir / name.str()).string().c_str()));
    }
    if (diagnostic_step || checkpoint_step) start = std::chrono::steady_clock::now();
  }
  ok(teuk_sync(plan));
  const double total =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - total_start).count();
  const double points = teuk_plan_grid_points(plan) * steps;
  std::cout << "backend=hip gfx950\n"
            << "boundary_policy=zero-SAT; no incoming propagating modes\n"
            << "source_policy=" << (second ? value_of(resolved, "second_order.source_mode")
                                           : std::string("disabled")) << '\n'
            << "evolution_wall_seconds=" << evolution << '\n'
            << "total_wall_seconds=" << total << '\n'
            << "grid_point_steps_per_second=" << points / evolution << '\n';
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string cmd = argc > 1 ? argv[1] : "--help";
    if (cmd == "backend") {
      std::cout << "teuk-mi355x 0.2.0\nexecution: HIP gfx950\n";
      return 0;
    }
    if (cmd == "--config") {
      if (argc < 3) throw std::invalid_argument("--config requires a configuration path");
      std::vector<std::string> overrides(argv + 3, argv + argc);
      return run(argv[2], overrides);
    }
    if (cmd == "help" || cmd == "--help" || cmd == "-h") {
      std::cout << "Usage:\n  teuk_solver --config run.cfg [key=value ...]\n"
                   "  teuk_solver backend\n  teuk_solver --help\n";
      return 0;
    }
    throw std::invalid_argument("unknown command: " + cmd);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << '\n';
    return 2;
  }
}
