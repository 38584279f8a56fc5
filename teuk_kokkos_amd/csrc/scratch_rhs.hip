// The linear RHS evaluation with its modal staging in a fixed device array
// instead of the plan's buffers, for callers that may not allocate device
// memory around the call.  The array is a code-object global: each device
// holds its own copy, placed when the library's code object is loaded there,
// and keeps it until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "plan.hpp"
#include "plan_detail.hpp"

namespace teuk {

// Two halves of kScratchHalf complex doubles: the analysed stage value and
// the modal RHS.  32 MiB per device.
constexpr long long kScratchHalf = 1LL << 20;
static __device__ __align__(256) cdouble scratch[2 * kScratchHalf];

namespace {
using detail::check;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
cdouble* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the host-side symbol is not a device address).
cdouble* scratch_on_current_device() {
  int dev = 0;
  check(hipGetDevice(&dev), "scratch device");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
    g_scratch_ptr[dev] = static_cast<cdouble*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

long long scratch_rhs_capacity() { return kScratchHalf; }

// Plan::evaluate_rhs for a linear plan with the staging in `scratch`.  Every
// staged value the kernels read is written earlier in the same call, so the
// array needs no clearing; calls that overlap on one device race on it.
void evaluate_rhs_in_scratch(Plan& plan, int i, const cplx* d_stage, cplx* d_out) {
  if (plan.second_order())
    throw std::invalid_argument(
        "the fixed-scratch RHS evaluation serves linear plans; second-order plans stage "
        "their RHS in plan buffers");
  const long long need = plan.modal_extent(i);
  if (need > kScratchHalf)
    throw std::invalid_argument("instance " + std::to_string(i) + " needs " +
                                std::to_string(need) +
                                " modal values of staging per buffer; the fixed scratch holds " +
                                std::to_string(kScratchHalf));
  cdouble* base = scratch_on_current_device();
  plan.evaluate_rhs_staged(i, d_stage, d_out, base, base + kScratchHalf);
}

}  // namespace teuk
