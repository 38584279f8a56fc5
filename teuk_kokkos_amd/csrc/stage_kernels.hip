// Band-transform kernels, the zero fill and the stage-kernel dispatch.  The
// stage kernels themselves (stage_kernels.hpp) are instantiated per (radial
// scheme, ell-tile class) in stage_instance.hip, compiled once per pair.
#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

// ------------------------------------------------------------ band transforms
// Nodal (mode, field, r, theta) <-> modal line layout, used by the
// reference-layout evaluate_rhs entry point.
__global__ __launch_bounds__(kStageThreads) void teuk_nodal_to_modal(NodalArgs a,
                                                                     const cdouble* nodal,
                                                                     cdouble* modal) {
  // one workgroup per (line, r): modal[f][l] = sum_j A[l][j] nodal[mode][f][r][j]
  const int line = blockIdx.x / a.nr, r = blockIdx.x % a.nr;
  const LineDesc L = a.lines[line];
  const int mode = a.line_mode[line];
  // B = G^-1 A: the coefficients c of a nodal vector S c.
  const double* A = a.mats + L.mat_off + 2 * L.ntheta * L.n_ell + gram_doubles(L.n_ell);
  for (int i = threadIdx.x; i < 3 * L.n_ell; i += kStageThreads) {
    const int f = i / L.n_ell, l = i % L.n_ell;
    const cdouble* v = nodal + ((static_cast<long long>(mode) * 13 + f) * a.nr + r) * a.ntheta;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < a.ntheta; ++j) {
      re += A[l * a.ntheta + j] * v[j].re;
      im += A[l * a.ntheta + j] * v[j].im;
    }
    modal[L.base + static_cast<long long>(r) * 3 * L.n_ell + i] = cmake(re, im);
  }
}

__global__ __launch_bounds__(kStageThreads) void teuk_modal_to_nodal(NodalArgs a,
                                                                     const cdouble* modal,
                                                                     cdouble* nodal) {
  const int line = blockIdx.x / a.nr, r = blockIdx.x % a.nr;
  const LineDesc L = a.lines[line];
  const int mode = a.line_mode[line];
  const double* S = a.mats + L.mat_off;
  for (int i = threadIdx.x; i < 3 * a.ntheta; i += kStageThreads) {
    const int f = i / a.ntheta, j = i % a.ntheta;
    const cdouble* c = modal + L.base + static_cast<long long>(r) * 3 * L.n_ell + f * L.n_ell;
    double re = 0.0, im = 0.0;
    for (int l = 0; l < L.n_ell; ++l) {
      re += S[j * L.n_ell + l] * c[l].re;
      im += S[j * L.n_ell + l] * c[l].im;
    }
    nodal[((static_cast<long long>(mode) * 13 + f) * a.nr + r) * a.ntheta + j] = cmake(re, im);
  }
}

__global__ void teuk_fill_zero(cdouble* p, long long count) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = cmake(0.0, 0.0);
}

// ------------------------------------------------------------ launchers
#define TEUK_DECLARE_INSTANCE(S, U)                                                        \
  hipError_t launch_stage_instance_##S##_##U(int stage, int variant, const StageArgs& a,   \
                                             int workgroups, int lds, hipStream_t s);      \
  hipError_t set_lds_limit_instance_##S##_##U(int bytes);
TEUK_DECLARE_INSTANCE(0, 1)
TEUK_DECLARE_INSTANCE(0, 2)
TEUK_DECLARE_INSTANCE(0, 4)
TEUK_DECLARE_INSTANCE(1, 1)
TEUK_DECLARE_INSTANCE(1, 2)
TEUK_DECLARE_INSTANCE(1, 4)
TEUK_DECLARE_INSTANCE(2, 1)
TEUK_DECLARE_INSTANCE(2, 2)
TEUK_DECLARE_INSTANCE(2, 4)
#undef TEUK_DECLARE_INSTANCE

namespace {
using LaunchFn = hipError_t (*)(int, int, const StageArgs&, int, int, hipStream_t);
using LimitFn = hipError_t (*)(int);
// [scheme][ub class 0, 1, 2 for ub 1, 2, 4]
constexpr LaunchFn kLaunch[3][3] = {
    {launch_stage_instance_0_1, launch_stage_instance_0_2, launch_stage_instance_0_4},
    {launch_stage_instance_1_1, launch_stage_instance_1_2, launch_stage_instance_1_4},
    {launch_stage_instance_2_1, launch_stage_instance_2_2, launch_stage_instance_2_4}};
constexpr LimitFn kLimit[3][3] = {
    {set_lds_limit_instance_0_1, set_lds_limit_instance_0_2, set_lds_limit_instance_0_4},
    {set_lds_limit_instance_1_1, set_lds_limit_instance_1_2, set_lds_limit_instance_1_4},
    {set_lds_limit_instance_2_1, set_lds_limit_instance_2_2, set_lds_limit_instance_2_4}};
}  // namespace

// ub = ell tiles of the tile kernel (1, 2, 4); variant: kStageGeneral,
// kStageKerr or kStageKerrGram; one workgroup per tile.
hipError_t launch_stage(int scheme, int stage, int ub, int variant, const StageArgs& a,
                        int n_tiles, int lds, hipStream_t s) {
  const int uc = ub == 1 ? 0 : (ub == 2 ? 1 : (ub == 4 ? 2 : -1));
  if (scheme < 0 || scheme > 2 || uc < 0 || stage < 0 || stage > 4 || n_tiles <= 0)
    return hipErrorInvalidValue;
  return kLaunch[scheme][uc](stage, variant, a, n_tiles, lds, s);
}

hipError_t set_stage_lds_limit(int bytes) {
  for (const auto& row : kLimit)
    for (LimitFn f : row) {
      const hipError_t e = f(bytes);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

hipError_t launch_nodal_to_modal(const NodalArgs& a, const cdouble* nodal, cdouble* modal,
                                 hipStream_t s) {
  hipLaunchKernelGGL(teuk_nodal_to_modal, dim3(a.n_lines * a.nr), dim3(kStageThreads), 0, s, a,
                     nodal, modal);
  return hipGetLastError();
}

hipError_t launch_modal_to_nodal(const NodalArgs& a, const cdouble* modal, cdouble* nodal,
                                 hipStream_t s) {
  hipLaunchKernelGGL(teuk_modal_to_nodal, dim3(a.n_lines * a.nr), dim3(kStageThreads), 0, s, a,
                     modal, nodal);
  return hipGetLastError();
}

hipError_t launch_fill_zero(cdouble* p, long long count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const long long blocks = (count + kStageThreads - 1) / kStageThreads;
  hipLaunchKernelGGL(teuk_fill_zero, dim3(static_cast<unsigned>(blocks)), dim3(kStageThreads), 0,
                     s, p, count);
  return hipGetLastError();
}

}  // namespace teuk
