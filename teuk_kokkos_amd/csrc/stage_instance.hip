// One (radial scheme, ell-tile class) instantiation of the stage kernels:
// compiled once per pair with -DTEUK_SCHEME=<0|1|2> -DTEUK_UB=<1|2|4>, every
// RK stage (0 = RHS only) and line variant (kStageGeneral, kStageKerr,
// kStageKerrGram) of it.
#include "stage_kernels.hpp"

#if !defined(TEUK_SCHEME) || !defined(TEUK_UB)
#error "stage_instance.hip needs -DTEUK_SCHEME and -DTEUK_UB"
#endif

#define TEUK_CAT5(a, b, c, d, e) a##b##c##d##e
#define TEUK_NAME(prefix, s, u) TEUK_CAT5(prefix, s, _, u, )

namespace teuk {

namespace {
template <int V>
hipError_t launch_variant(int stage, const StageArgs& a, int workgroups, int lds, hipStream_t s) {
  constexpr int S = TEUK_SCHEME, U = TEUK_UB;
  const dim3 grid(workgroups), block(kStageThreads);
  switch (stage) {
    case 0: hipLaunchKernelGGL((teuk_stage_kernel<S, 0, U, V>), grid, block, lds, s, a); break;
    case 1: hipLaunchKernelGGL((teuk_stage_kernel<S, 1, U, V>), grid, block, lds, s, a); break;
    case 2: hipLaunchKernelGGL((teuk_stage_kernel<S, 2, U, V>), grid, block, lds, s, a); break;
    case 3: hipLaunchKernelGGL((teuk_stage_kernel<S, 3, U, V>), grid, block, lds, s, a); break;
    default: hipLaunchKernelGGL((teuk_stage_kernel<S, 4, U, V>), grid, block, lds, s, a); break;
  }
  return hipGetLastError();
}

template <int V>
hipError_t limit_variant(int bytes) {
  constexpr int S = TEUK_SCHEME, U = TEUK_UB;
  const void* fns[] = {reinterpret_cast<const void*>(&teuk_stage_kernel<S, 0, U, V>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 1, U, V>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 2, U, V>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 3, U, V>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 4, U, V>)};
  for (const void* f : fns) {
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace

hipError_t TEUK_NAME(launch_stage_instance_, TEUK_SCHEME, TEUK_UB)(int stage, int variant,
                                                                  const StageArgs& a,
                                                                  int workgroups, int lds,
                                                                  hipStream_t s) {
  switch (variant) {
    case kStageGeneral: return launch_variant<kStageGeneral>(stage, a, workgroups, lds, s);
    case kStageKerr: return launch_variant<kStageKerr>(stage, a, workgroups, lds, s);
    case kStageKerrGram: return launch_variant<kStageKerrGram>(stage, a, workgroups, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t TEUK_NAME(set_lds_limit_instance_, TEUK_SCHEME, TEUK_UB)(int bytes) {
  hipError_t e = limit_variant<kStageGeneral>(bytes);
  if (e == hipSuccess) e = limit_variant<kStageKerr>(bytes);
  if (e == hipSuccess) e = limit_variant<kStageKerrGram>(bytes);
  return e;
}

}  // namespace teuk
