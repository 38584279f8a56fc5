// Band-validated import on the GPU (pipeline_bands.hpp:108-232): one 64-lane
// workgroup per (mode, field, r) theta row of the imported instance.  The
// check projects the row onto its field's declared band (S A x, with the
// spin, m and ell_max of the field) and records |x| and |x - S A x|; the host
// reduces them.  The split writes the modal coefficients of the evolved rows
// into the plan's modal state and keeps their nodal residual x - S c for
// export.
#include <hip/hip_runtime.h>

#include "import_layout.hpp"

namespace teuk {

namespace {

// theta row `row` of the staged instance into LDS: x (T values), then the
// coefficients c = C x (n values) of the band with C at `c_off`
__device__ __forceinline__ void row_coefficients(const ImportArgs& a, const ImportRow& R,
                                                 long long row, int c_off, cdouble* x,
                                                 cdouble* c) {
  const int T = a.T, n = R.n, lane = threadIdx.x;
  const cdouble* src = a.in + row * T;
  for (int j = lane; j < T; j += kImportThreads) x[j] = src[j];
  __syncthreads();
  const double* C = a.mats + c_off;
  for (int k = lane; k < n; k += kImportThreads) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < T; ++j) {
      re += C[k * T + j] * x[j].re;
      im += C[k * T + j] * x[j].im;
    }
    c[k] = cmake(re, im);
  }
  __syncthreads();
}

// S c at node j
__device__ __forceinline__ cdouble synthesize_at(const ImportArgs& a, const ImportRow& R,
                                                 const cdouble* c, int j) {
  const double* S = a.mats + R.s_off + j * R.n;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < R.n; ++k) {
    re += S[k] * c[k].re;
    im += S[k] * c[k].im;
  }
  return cmake(re, im);
}

}  // namespace

__global__ __launch_bounds__(kImportThreads) void teuk_band_check_kernel(ImportArgs a) {
  extern __shared__ __align__(16) double smem[];
  const long long row = blockIdx.x;  // (mode * 13 + field) * N + r
  const ImportRow& R = a.rows[row / a.N];
  cdouble* x = reinterpret_cast<cdouble*>(smem);
  cdouble* c = x + a.T;
  row_coefficients(a, R, row, R.a_off, x, c);
  double line2 = 0.0, res2 = 0.0;
  for (int j = threadIdx.x; j < a.T; j += kImportThreads) {
    const cdouble p = R.n > 0 ? synthesize_at(a, R, c, j) : cmake(0.0, 0.0);
    const double dre = x[j].re - p.re, dim = x[j].im - p.im;
    line2 += x[j].re * x[j].re + x[j].im * x[j].im;
    res2 += dre * dre + dim * dim;
  }
  for (int off = kImportThreads / 2; off > 0; off >>= 1) {
    line2 += __shfl_xor(line2, off);
    res2 += __shfl_xor(res2, off);
  }
  if (threadIdx.x == 0) {
    a.row_norms[2 * row] = sqrt(line2);
    a.row_norms[2 * row + 1] = sqrt(res2);
  }
}

__global__ __launch_bounds__(kImportThreads) void teuk_import_modal_kernel(ImportArgs a) {
  extern __shared__ __align__(16) double smem[];
  // grid: (r, mode * 3 + field) over the evolved fields
  const int r = blockIdx.x, mf3 = blockIdx.y, mode = mf3 / 3, f = mf3 - mode * 3;
  const ImportRow& R = a.rows[mode * 13 + f];
  if (R.modal_base < 0 || R.n <= 0) return;
  const long long row = static_cast<long long>(mode * 13 + f) * a.N + r;
  cdouble* x = reinterpret_cast<cdouble*>(smem);
  cdouble* c = x + a.T;
  row_coefficients(a, R, row, R.c_off, x, c);
  cdouble* out = a.modal + R.modal_base + static_cast<long long>(r) * 3 * R.n + f * R.n;
  for (int k = threadIdx.x; k < R.n; k += kImportThreads) out[k] = c[k];
  bool nonzero = false;
  for (int j = threadIdx.x; j < a.T; j += kImportThreads) {
    const cdouble p = synthesize_at(a, R, c, j);
    const cdouble d = cmake(x[j].re - p.re, x[j].im - p.im);
    a.residual[row * a.T + j] = d;
    nonzero = nonzero || d.re != 0.0 || d.im != 0.0;
  }
  if (nonzero) atomicOr(a.any_residual, 1);
}

int import_lds_bytes(int ntheta, int max_band) { return 16 * (ntheta + max_band); }

hipError_t launch_band_check(const ImportArgs& a, hipStream_t s) {
  const long long rows = static_cast<long long>(a.M) * 13 * a.N;
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(teuk_band_check_kernel, dim3(static_cast<unsigned>(rows)),
                     dim3(kImportThreads), import_lds_bytes(a.T, kMaxImportBand), s, a);
  return hipGetLastError();
}

hipError_t launch_import_modal(const ImportArgs& a, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(teuk_import_modal_kernel, dim3(a.N, a.M * 3), dim3(kImportThreads),
                     import_lds_bytes(a.T, kMaxImportBand), s, a);
  return hipGetLastError();
}

}  // namespace teuk
