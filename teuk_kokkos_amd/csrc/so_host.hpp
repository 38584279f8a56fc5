// Host-side state of the second-order graph of a Plan.
#pragma once

#include <array>
#include <vector>

#include "plan.hpp"
#include "so_layout.hpp"

namespace teuk {

struct Plan::SoHost {
  std::vector<SoInstance> inst;
  std::vector<double> ops;    // T x T operators
  std::vector<double> nodes;  // per instance: x, sqrt(1-x^2), cos(acos x), sin(acos x), w
  std::vector<SoTile> tiles[3];                      // per radial scheme
  std::vector<std::pair<int, int>> inst_tiles[3];    // per instance: (first, count)
  int lin_lds = 0, rec0_lds = 0, rec1_lds = 0, max_n = 0, max_drad = 0;
  int max_mt = 0;  // largest reconstruction row: parent modes x theta nodes
  int recon_groups = 0;  // reconstruction workgroups per instance (grid.x)
  long long values = 0, planes = 0, max_block = 0, max_plane = 0;
  std::vector<std::array<double, 14>> norms;
  std::vector<int> active_host;

  long long bg_points = 0, max_nt = 0;
  SoBackground* d_bg = nullptr;
  SoInstance* d_inst = nullptr;
  double* d_ops = nullptr;
  double* d_nodes = nullptr;
  SchemeTable* d_tables = nullptr;
  SoTile* d_tiles[3] = {nullptr, nullptr, nullptr};
  int* d_active = nullptr;
  double* d_dt = nullptr;
  double* d_dt_step = nullptr;  // each instance's configured step (the graph's)
  double* d_red = nullptr;
  cdouble *y = nullptr, *acc = nullptr, *ya = nullptr, *yb = nullptr, *k = nullptr;
  cdouble *tan = nullptr, *dpsi = nullptr, *dpsi_t = nullptr, *rdr = nullptr, *rdr_t = nullptr;
  cdouble *rang = nullptr, *src = nullptr, *cons = nullptr, *terms = nullptr;
  cdouble *rang_t = nullptr, *ex = nullptr, *raw = nullptr, *rangx = nullptr;
  int proj_lds = 0;
  long long max_source = 0;  // largest N x ntheta (points of one target)
  int max_targets = 0;
  unsigned schemes = 0;  // bit k: some instance uses radial scheme k
  // per-pair source values of a diagnostics evaluation (one instance)
  cdouble* pairs = nullptr;
  long long pairs_cap = 0;
  bool capture_pairs = false;
};

}  // namespace teuk
