// extern "C" layer of libteuk_hip.so (declarations in include/teuk_hip.h).
// Exceptions never cross the boundary: std::invalid_argument and
// std::out_of_range map to TEUK_EINVAL, everything else to TEUK_ERUNTIME.
#include "../../include/teuk_hip.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "diag_layout.hpp"
#include "host_math.hpp"
#include "plan.hpp"
#include "plan_detail.hpp"

struct teuk_plan {
  std::unique_ptr<teuk::Plan> plan;
  std::string error;
};

namespace {

thread_local std::string g_config_error;

template <class F>
int guarded(std::string& error, F&& f) {
  try {
    f();
    return TEUK_OK;
  } catch (const std::invalid_argument& e) {
    error = e.what();
    return TEUK_EINVAL;
  } catch (const std::out_of_range& e) {
    error = e.what();
    return TEUK_EINVAL;
  } catch (const std::exception& e) {
    error = e.what();
    return TEUK_ERUNTIME;
  } catch (...) {
    error = "unknown error";
    return TEUK_ERUNTIME;
  }
}

std::vector<std::string> to_strings(const char* const* v, int n) {
  std::vector<std::string> out;
  for (int i = 0; i < n; ++i) {
    if (!v || !v[i]) throw std::invalid_argument("null override string");
    out.emplace_back(v[i]);
  }
  return out;
}

teuk::Plan& plan_of(teuk_plan* p) {
  if (!p || !p->plan) throw std::invalid_argument("null plan");
  return *p->plan;
}

void check_instance(const teuk::Plan& p, int i) {
  if (i < 0 || i >= p.instance_count()) throw std::out_of_range("instance index out of range");
}

int create(const std::vector<std::string>& texts, const char* const* overrides, int n_overrides,
           teuk_plan** out) {
  if (!out) {
    g_config_error = "null output pointer";
    return TEUK_EINVAL;
  }
  *out = nullptr;
  auto handle = std::make_unique<teuk_plan>();
  const int rc = guarded(g_config_error, [&] {
    const auto ov = to_strings(overrides, n_overrides);
    std::vector<teuk::RunParameters> params;
    params.reserve(texts.size());
    for (const auto& t : texts) params.push_back(teuk::parse_configuration_text(t, ov));
    handle->plan = std::make_unique<teuk::Plan>(params);
  });
  if (rc == TEUK_OK) *out = handle.release();
  return rc;
}

}  // namespace

extern "C" {

const char* teuk_config_last_error(void) { return g_config_error.c_str(); }

int teuk_config_resolve(const char* cfg_text, const char* const* overrides, int n_overrides,
                        char* out, size_t out_size, size_t* needed) {
  return guarded(g_config_error, [&] {
    if (!cfg_text) throw std::invalid_argument("null configuration text");
    const auto p = teuk::parse_configuration_text(cfg_text, to_strings(overrides, n_overrides));
    const std::string text = teuk::resolved_configuration_text(p);
    if (needed) *needed = text.size() + 1;
    if (out && out_size > 0) {
      const size_t n = std::min(out_size - 1, text.size());
      std::memcpy(out, text.data(), n);
      out[n] = '\0';
    }
  });
}

int teuk_plan_create(const char* cfg_text, const char* const* overrides, int n_overrides,
                     teuk_plan** out) {
  if (!cfg_text) {
    g_config_error = "null configuration text";
    return TEUK_EINVAL;
  }
  return create({cfg_text}, overrides, n_overrides, out);
}

int teuk_plan_create_batch(const char* const* cfg_texts, int n_instances,
                           const char* const* overrides, int n_overrides, teuk_plan** out) {
  if (!cfg_texts || n_instances <= 0) {
    g_config_error = "a batch needs at least one configuration";
    return TEUK_EINVAL;
  }
  std::vector<std::string> texts;
  for (int i = 0; i < n_instances; ++i) {
    if (!cfg_texts[i]) {
      g_config_error = "null configuration text";
      return TEUK_EINVAL;
    }
    texts.emplace_back(cfg_texts[i]);
  }
  return create(texts, overrides, n_overrides, out);
}

void teuk_plan_destroy(teuk_plan* plan) { delete plan; }

const char* teuk_last_error(const teuk_plan* plan) {
  return plan ? plan->error.c_str() : g_config_error.c_str();
}

int teuk_plan_instance_count(const teuk_plan* plan) {
  return plan && plan->plan ? plan->plan->instance_count() : -1;
}

int teuk_plan_shape(const teuk_plan* plan, int instance, int* modes, int* fields, int* nr,
                    int* ntheta) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    check_instance(pl, instance);
    const auto& in = pl.instance(instance);
    if (modes) *modes = in.M;
    if (fields) *fields = 13;
    if (nr) *nr = in.N;
    if (ntheta) *ntheta = in.T;
  });
}

double teuk_plan_time_step(const teuk_plan* plan, int instance) {
  if (!plan || !plan->plan || instance < 0 || instance >= plan->plan->instance_count()) return -1.0;
  return plan->plan->instance(instance).dt;
}

int teuk_plan_progress(const teuk_plan* plan, int instance, double* time, uint64_t* steps) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    check_instance(pl, instance);
    if (time) *time = pl.instance(instance).time;
    if (steps) *steps = pl.instance(instance).steps_done;
  });
}

double teuk_plan_grid_points(const teuk_plan* plan) {
  return plan && plan->plan ? plan->plan->grid_points_per_step() : -1.0;
}

double teuk_plan_modal_points(const teuk_plan* plan) {
  return plan && plan->plan ? plan->plan->modal_points_per_step() : -1.0;
}

int teuk_initialize(teuk_plan* plan) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).initialize(); });
}

int teuk_checkpoint_write(teuk_plan* plan, int instance, const char* directory) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!directory) throw std::invalid_argument("null checkpoint directory");
    pl.write_checkpoint(instance, directory);
  });
}

int teuk_checkpoint_load(teuk_plan* plan, int instance, const char* directory) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!directory) throw std::invalid_argument("null checkpoint directory");
    pl.load_checkpoint(instance, directory);
  });
}

int teuk_state_import(teuk_plan* plan, int instance, const double* host,
                      double relative_tolerance, double* offband) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!host) throw std::invalid_argument("null state buffer");
    const double v =
        pl.import_state(instance, reinterpret_cast<const teuk::cplx*>(host), relative_tolerance);
    if (offband) *offband = v;
  });
}

int teuk_state_export(teuk_plan* plan, int instance, double* host) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!host) throw std::invalid_argument("null state buffer");
    pl.export_state(instance, reinterpret_cast<teuk::cplx*>(host));
  });
}

int teuk_evaluate_rhs(teuk_plan* plan, int instance, const void* d_stage, void* d_out,
                      double time) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!d_stage || !d_out) throw std::invalid_argument("null device buffer");
    if (d_stage == d_out) throw std::invalid_argument("stage and output must not alias");
    pl.evaluate_rhs(instance, static_cast<const teuk::cplx*>(d_stage),
                    static_cast<teuk::cplx*>(d_out), time);
  });
}

int teuk_evaluate_rhs_noalloc(teuk_plan* plan, int instance, const void* d_stage, void* d_out,
                              double time) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!d_stage || !d_out) throw std::invalid_argument("null device buffer");
    if (d_stage == d_out) throw std::invalid_argument("stage and output must not alias");
    (void)time;  // the linear stage operator is autonomous
    teuk::evaluate_rhs_in_scratch(pl, instance, static_cast<const teuk::cplx*>(d_stage),
                                  static_cast<teuk::cplx*>(d_out));
  });
}

int teuk_step(teuk_plan* plan, double time, double dt) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).step(time, dt); });
}

int teuk_advance(teuk_plan* plan, int n_steps) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).advance(n_steps); });
}

int teuk_sync(teuk_plan* plan) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).sync(); });
}

int teuk_sample_waveform(teuk_plan* plan, int instance, teuk_waveform_record* out, int capacity,
                         int* count) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    const auto recs = pl.sample_waveform(instance);
    if (count) *count = static_cast<int>(recs.size());
    if (out) {
      if (capacity < static_cast<int>(recs.size()))
        throw std::invalid_argument("waveform record buffer is too small");
      for (size_t k = 0; k < recs.size(); ++k) {
        const auto& r = recs[k];
        out[k] = {r.step, r.time, r.boundary, r.order, r.m, r.ell,
                  r.rescaled_re, r.rescaled_im, r.endpoint_re, r.endpoint_im};
      }
    }
  });
}

int teuk_get_source_activation(const teuk_plan* plan, int instance,
                               teuk_source_activation* state) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    check_instance(pl, instance);
    if (!state) throw std::invalid_argument("null activation output");
    const auto a = pl.activation(instance);
    *state = {a.active, a.consecutive_passes, a.activation_time, a.last_eligibility_time};
  });
}

int teuk_set_source_activation(teuk_plan* plan, int instance,
                               const teuk_source_activation* state) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!state) throw std::invalid_argument("null activation state");
    pl.set_activation(instance, {state->active, state->consecutive_passes, state->activation_time,
                                 state->last_eligibility_time});
  });
}

int teuk_get_source_constraint_norms(const teuk_plan* plan, int instance, double* norms14) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    check_instance(pl, instance);
    if (!norms14) throw std::invalid_argument("null norms output");
    pl.constraint_norms(instance, norms14);
  });
}

static_assert(TEUK_DIAG_COUNT == teuk::kDiagCount && TEUK_DIAG_TRANSPORT == teuk::kDiagTransport &&
                  TEUK_DIAG_HORIZON_SECOND == teuk::kDiagHorizonSecond &&
                  TEUK_DIAG_ALL_FINITE == teuk::kDiagAllFinite,
              "diagnostics layout of the C ABI and the plan agree");

int teuk_sample_diagnostics(teuk_plan* plan, int instance, double time, double* out) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan& pl = plan_of(plan);
    check_instance(pl, instance);
    if (!out) throw std::invalid_argument("null diagnostics output");
    pl.sample_diagnostics(instance, time, out);
  });
}

int teuk_source_pairs(const teuk_plan* plan, int instance, double* out, int capacity,
                      int* count) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    check_instance(pl, instance);
    if (!count) throw std::invalid_argument("null pair count");
    const auto& rep = pl.instance(instance).pair_report;
    const int n = static_cast<int>(rep.size() / 7);
    *count = n;
    if (!out) return;
    if (capacity < n) throw std::invalid_argument("pair report capacity too small");
    std::copy(rep.begin(), rep.end(), out);
  });
}

int teuk_plan_set_graph(teuk_plan* plan, int enable) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).set_graph(enable != 0); });
}

int teuk_plan_graph_replays(const teuk_plan* plan, long long* replays) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    if (!replays) throw std::invalid_argument("null output pointer");
    *replays = plan_of(p).graph_replays();
  });
}

int teuk_plan_set_stream(teuk_plan* plan, void* stream) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy,
                 [&] { plan_of(plan).set_stream(static_cast<hipStream_t>(stream)); });
}

int teuk_allocation_count(unsigned long long* count) {
  return guarded(g_config_error, [&] {
    if (!count) throw std::invalid_argument("null output pointer");
    *count = teuk::detail::g_allocations.load();
  });
}

int teuk_plan_set_lds_budget(teuk_plan* plan, int bytes) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).set_lds_budget(bytes); });
}

int teuk_plan_set_tile_rows(teuk_plan* plan, int rows) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).set_tile_rows(rows); });
}

int teuk_plan_set_specialized(teuk_plan* plan, int enable) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).set_specialized(enable != 0); });
}

int teuk_plan_set_mixed_gram(teuk_plan* plan, int enable) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] { plan_of(plan).set_mixed_gram(enable != 0); });
}

int teuk_plan_tiles(const teuk_plan* plan, int* tiles, int* max_lds_bytes) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan& pl = plan_of(p);
    if (tiles) *tiles = pl.tile_count();
    if (max_lds_bytes) *max_lds_bytes = pl.max_lds_bytes();
  });
}

int teuk_plan_set_kernel_timing(teuk_plan* plan, int enable) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy,
                 [&] { plan_of(plan).set_kernel_timing(enable != 0); });
}

int teuk_plan_kernel_timing(const teuk_plan* plan, double* total_ms, long long* launches) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    double ms = 0.0;
    long long n = 0;
    plan_of(p).kernel_timing(&ms, &n);
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
  });
}

int teuk_plan_step_timing(const teuk_plan* plan, double* total_ms, long long* steps) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    double ms = 0.0;
    long long n = 0;
    plan_of(p).step_timing(&ms, &n);
    if (total_ms) *total_ms = ms;
    if (steps) *steps = n;
  });
}

int teuk_plan_group_count(const teuk_plan* plan, int* count) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    if (!count) throw std::invalid_argument("null output pointer");
    *count = plan_of(p).group_count();
  });
}

int teuk_plan_group_info(const teuk_plan* plan, int group, int* scheme, int* ell_tiles,
                         int* tiles, int* lds_bytes, double* kernel_ms) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    int s = 0, u = 0, t = 0, l = 0;
    double ms = 0.0;
    plan_of(p).group_info(group, &s, &u, &t, &l, &ms);
    if (scheme) *scheme = s;
    if (ell_tiles) *ell_tiles = u;
    if (tiles) *tiles = t;
    if (lds_bytes) *lds_bytes = l;
    if (kernel_ms) *kernel_ms = ms;
  });
}

int teuk_plan_group_variant(const teuk_plan* plan, int group, int* variant) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    if (!variant) throw std::invalid_argument("null output pointer");
    *variant = plan_of(p).group_variant(group);
  });
}

int teuk_plan_set_group_shape(teuk_plan* plan, int scheme, int ell_tiles, int variant,
                              int lds_budget, int tile_rows) {
  std::string dummy;
  return guarded(plan ? plan->error : dummy, [&] {
    teuk::Plan::GroupShape shape;
    shape.lds_budget = lds_budget;
    shape.tile_rows = tile_rows;
    plan_of(plan).set_group_shape(scheme, ell_tiles, variant, shape);
  });
}

int teuk_plan_group_shape(const teuk_plan* plan, int group, int* lds_budget, int* tile_rows) {
  auto* p = const_cast<teuk_plan*>(plan);
  std::string dummy;
  return guarded(p ? p->error : dummy, [&] {
    const teuk::Plan::GroupShape s = plan_of(p).group_shape(group);
    if (lds_budget) *lds_budget = s.lds_budget;
    if (tile_rows) *tile_rows = s.tile_rows;
  });
}

int teuk_device_alloc(size_t bytes, void** ptr) {
  return guarded(g_config_error, [&] {
    if (!ptr) throw std::invalid_argument("null output pointer");
    *ptr = nullptr;
    const hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    ++teuk::detail::g_allocations;
  });
}

int teuk_device_free(void* ptr) {
  return guarded(g_config_error, [&] {
    const hipError_t e = hipFree(ptr);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipFree: ") + hipGetErrorString(e));
  });
}

int teuk_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  return guarded(g_config_error, [&] {
    if (!dst || !src) throw std::invalid_argument("null copy pointer");
    const hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                            : kind == 1 ? hipMemcpyDeviceToHost
                            : kind == 2 ? hipMemcpyDeviceToDevice
                                        : throw std::invalid_argument("unknown copy kind");
    const hipError_t e = hipMemcpy(dst, src, bytes, k);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
  });
}

double teuk_sbp_derivative_entry(int scheme, int n, int row, int col) {
  if (scheme < 0 || scheme > 2) return 0.0;
  return teuk::sbp_derivative_entry(static_cast<teuk::RadialScheme>(scheme), n, row, col);
}

double teuk_sbp_norm_weight(int scheme, int n, int index) {
  if (scheme < 0 || scheme > 2) return 0.0;
  return teuk::sbp_norm_weight(static_cast<teuk::RadialScheme>(scheme), n, index);
}

int teuk_band_matrices(int spin, int m, int ell_max, int ntheta, double* S, double* A) {
  int n = -1;
  const int rc = guarded(g_config_error, [&] {
    const auto bt = teuk::make_band_transform(spin, m, ell_max, ntheta);
    if (S) std::memcpy(S, bt.S.data(), bt.S.size() * sizeof(double));
    if (A) std::memcpy(A, bt.A.data(), bt.A.size() * sizeof(double));
    n = bt.modes();
  });
  return rc == TEUK_OK ? n : -1;
}

int teuk_config_modes(const char* cfg_text, int* modes, int* sharp, int capacity, int* count,
                      int* pairs, int pair_capacity, int* pair_count) {
  return guarded(g_config_error, [&] {
    if (!cfg_text) throw std::invalid_argument("null configuration text");
    const auto p = teuk::parse_configuration_text(cfg_text, {});
    const auto reg = teuk::make_run_registry(p);
    if (count) *count = reg.size();
    if (pair_count) *pair_count = static_cast<int>(reg.ordered_pairs().size());
    if (reg.size() > capacity && (modes || sharp))
      throw std::invalid_argument("mode buffer is too small");
    for (int i = 0; i < reg.size(); ++i) {
      if (modes) modes[i] = reg.modes()[i];
      if (sharp) sharp[i] = reg.closed_under_sharp() ? reg.sharp_index(reg.modes()[i]) : -1;
    }
    const auto& pr = reg.ordered_pairs();
    if (pairs) {
      if (static_cast<int>(pr.size()) > pair_capacity)
        throw std::invalid_argument("pair buffer is too small");
      for (size_t k = 0; k < pr.size(); ++k) {
        pairs[3 * k] = pr[k].m1;
        pairs[3 * k + 1] = pr[k].m2;
        pairs[3 * k + 2] = pr[k].target;
      }
    }
  });
}

}  // extern "C"
