// Host helpers shared by the plan's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "device_layout.hpp"
#include "host_math.hpp"

namespace teuk {
namespace detail {

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocations, streams and events the library has created (plans
// and teuk_device_alloc): teuk_allocation_count reads it, so a caller or a
// test can see that a call allocates nothing.
inline std::atomic<unsigned long long> g_allocations{0};

template <class T>
inline T* device_alloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
  ++g_allocations;
  return static_cast<T*>(p);
}

inline hipStream_t create_stream() {
  hipStream_t st = nullptr;
  check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  ++g_allocations;
  return st;
}

inline hipEvent_t create_event(bool timing) {
  hipEvent_t e = nullptr;
  check(timing ? hipEventCreate(&e) : hipEventCreateWithFlags(&e, hipEventDisableTiming),
        "hipEventCreate");
  ++g_allocations;
  return e;
}

inline SchemeTable make_table(RadialScheme s) {
  const SbpScheme& sc = sbp_scheme(s);
  SchemeTable t{};
  t.W = sc.closure_rows;
  t.C = sc.closure_width;
  t.B = sc.half_width;
  t.order = sc.diss_order;
  for (int r = 0; r < sc.closure_rows; ++r)
    for (int c = 0; c < sc.closure_width; ++c)
      t.left[r * kMaxClosureWidth + c] = sc.left[r * sc.closure_width + c];
  for (size_t k = 0; k < sc.interior.size(); ++k) t.interior[k] = sc.interior[k];
  for (size_t k = 0; k < sc.norm.size(); ++k) t.norm[k] = sc.norm[k];
  for (int d = 0; d < sc.diss_order; ++d)
    for (size_t c = 0; c < sc.diss_left[d].size(); ++c) t.diss_left[d][c] = sc.diss_left[d][c];
  for (size_t k = 0; k < sc.diss_interior.size(); ++k) t.diss_interior[k] = sc.diss_interior[k];
  return t;
}

// Expanded Gaussian seeds with sharp partners, sorted by (m, ell)
// (initial_data_factory.hpp:45-74).
inline std::vector<SeedMode> expanded_seeds(const RunParameters& p) {
  std::vector<SeedMode> out;
  for (const auto& s : p.seeds) {
    out.push_back(s);
    if (p.add_sharp_partner && s.m != 0) out.push_back({s.ell, -s.m, std::conj(s.amplitude)});
  }
  std::sort(out.begin(), out.end(), [](const SeedMode& a, const SeedMode& b) {
    return a.m != b.m ? a.m < b.m : a.ell < b.ell;
  });
  for (size_t i = 1; i < out.size(); ++i)
    if (out[i].m == out[i - 1].m && out[i].ell == out[i - 1].ell)
      throw std::invalid_argument("expanded Gaussian modes contain duplicate (ell,m) seeds");
  return out;
}

}  // namespace detail
}  // namespace teuk
