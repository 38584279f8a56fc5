// One (radial scheme, ell-tile class) instantiation of the stage kernels:
// compiled once per pair with -DTEUK_SCHEME=<0|1|2> -DTEUK_UB=<1|2|4>.
#include "stage_kernels.hpp"

#if !defined(TEUK_SCHEME) || !defined(TEUK_UB)
#error "stage_instance.hip needs -DTEUK_SCHEME and -DTEUK_UB"
#endif

#define TEUK_CAT5(a, b, c, d, e) a##b##c##d##e
#define TEUK_NAME(prefix, s, u) TEUK_CAT5(prefix, s, _, u, )

namespace teuk {

// march: the marching kernel (one ell tile only) over segment descriptors.
hipError_t TEUK_NAME(launch_stage_instance_, TEUK_SCHEME, TEUK_UB)(int stage, bool march,
                                                                  const StageArgs& a,
                                                                  int n_tiles, int lds,
                                                                  hipStream_t s) {
  constexpr int S = TEUK_SCHEME, U = TEUK_UB;
  const dim3 grid(n_tiles), block(kStageThreads);
  if (march) {
#if TEUK_UB == 1
    switch (stage) {
      case 0: hipLaunchKernelGGL((teuk_march_kernel<S, 0>), grid, block, lds, s, a); break;
      case 1: hipLaunchKernelGGL((teuk_march_kernel<S, 1>), grid, block, lds, s, a); break;
      case 2: hipLaunchKernelGGL((teuk_march_kernel<S, 2>), grid, block, lds, s, a); break;
      case 3: hipLaunchKernelGGL((teuk_march_kernel<S, 3>), grid, block, lds, s, a); break;
      default: hipLaunchKernelGGL((teuk_march_kernel<S, 4>), grid, block, lds, s, a); break;
    }
    return hipGetLastError();
#else
    return hipErrorInvalidValue;
#endif
  }
  switch (stage) {
    case 0: hipLaunchKernelGGL((teuk_stage_kernel<S, 0, U>), grid, block, lds, s, a); break;
    case 1: hipLaunchKernelGGL((teuk_stage_kernel<S, 1, U>), grid, block, lds, s, a); break;
    case 2: hipLaunchKernelGGL((teuk_stage_kernel<S, 2, U>), grid, block, lds, s, a); break;
    case 3: hipLaunchKernelGGL((teuk_stage_kernel<S, 3, U>), grid, block, lds, s, a); break;
    default: hipLaunchKernelGGL((teuk_stage_kernel<S, 4, U>), grid, block, lds, s, a); break;
  }
  return hipGetLastError();
}

hipError_t TEUK_NAME(set_lds_limit_instance_, TEUK_SCHEME, TEUK_UB)(int bytes) {
  constexpr int S = TEUK_SCHEME, U = TEUK_UB;
  const void* fns[] = {reinterpret_cast<const void*>(&teuk_stage_kernel<S, 0, U>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 1, U>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 2, U>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 3, U>),
                       reinterpret_cast<const void*>(&teuk_stage_kernel<S, 4, U>)};
  for (const void* f : fns) {
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
#if TEUK_UB == 1
  const void* march[] = {reinterpret_cast<const void*>(&teuk_march_kernel<S, 0>),
                         reinterpret_cast<const void*>(&teuk_march_kernel<S, 1>),
                         reinterpret_cast<const void*>(&teuk_march_kernel<S, 2>),
                         reinterpret_cast<const void*>(&teuk_march_kernel<S, 3>),
                         reinterpret_cast<const void*>(&teuk_march_kernel<S, 4>)};
  for (const void* f : march) {
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
#endif
  return hipSuccess;
}

}  // namespace teuk
