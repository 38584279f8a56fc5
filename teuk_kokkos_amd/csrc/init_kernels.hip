// Gaussian initial data with the Galerkin zero-time-derivative P on the GPU
// (pipeline_initial_data.hpp:142-340, dense_solve.hpp:20-72): one wavefront
// per (line, radius) assembles the n_ell x n_ell system
//   (A diag(1/C_T) S) p = A (C_T^-1 (-2K Q + G_m Psi))
// in LDS, solves it by Gaussian elimination with partial pivoting, checks
// the relative residual and writes the modal coefficients A x of P, Q and Psi,
// or for a nodal state their band projections S A x
// (pipeline_initial_data.hpp:334-335).
#include <hip/hip_runtime.h>

#include "init_layout.hpp"

namespace teuk {

namespace {

__device__ __forceinline__ double profile(const InitLine& il, double R) {
  const double z = (R - il.center) / il.width;
  if (!il.compact) return exp(-z * z);
  if (fabs(R - il.center) >= il.width) return 0.0;
  return exp(1.0 - 1.0 / (1.0 - z * z));
}

}  // namespace

// grid: (radius, line).  LDS: M (n x n), its copy, rhs and its copy, the
// nodal rows and the solution.
__global__ __launch_bounds__(kInitThreads) void teuk_galerkin_init(InitArgs a) {
  extern __shared__ __align__(16) double smem[];
  const int line = blockIdx.y, r = blockIdx.x;
  const LineDesc& L = a.lines[line];
  if (r >= L.nr) return;
  const InitLine& il = a.init[line];
  const SchemeTable& TB = a.tables[il.scheme];
  const int n = L.n_ell, T = L.ntheta, N = L.nr, tid = threadIdx.x;
  const double* S = a.mats + L.mat_off;       // T x n
  const double* A = S + T * n;                // n x T
  const double* sin_t = a.nodes + L.node_off;
  const double* cos_t = sin_t + T;
  double* Mx = smem;                          // n x n, eliminated in place
  double* M0 = Mx + n * n;                    // n x n
  cdouble* rhs = reinterpret_cast<cdouble*>(M0 + n * n);  // n
  cdouble* b0 = rhs + n;                       // n
  cdouble* x = b0 + n;                         // n
  cdouble* ang = x + n;                        // T: S seed
  cdouble* psi = ang + T;                      // T
  cdouble* q = psi + T;                        // T
  cdouble* pn = q + T;                         // T
  double* w = reinterpret_cast<double*>(pn + T);  // T: 1 / C_T
  __shared__ double red[kInitThreads];
  __shared__ int redi[kInitThreads];
  __shared__ int singular;

  const cdouble* seed = a.seeds + il.seed_off;
  const double R = L.h * r;
  for (int t = tid; t < T; t += kInitThreads) {
    double re = 0.0, im = 0.0;
    for (int l = 0; l < n; ++l) {
      re += S[t * n + l] * seed[l].re;
      im += S[t * n + l] * seed[l].im;
    }
    ang[t] = cmake(re, im);
  }
  if (tid == 0) singular = 0;
  __syncthreads();
  // Psi = profile * angular seed; Q = D Psi with the scheme's first
  // derivative on the nodal column (pipeline_initial_data.hpp:258-294).
  const int W = TB.W, C = TB.C, B = TB.B;
  for (int t = tid; t < T; t += kInitThreads) {
    const double pr = profile(il, R);
    psi[t] = cmake(pr * ang[t].re, pr * ang[t].im);
    double re = 0.0, im = 0.0;
    auto col = [&](int c) {
      const double p = profile(il, L.h * c);
      return cmake(p * ang[t].re, p * ang[t].im);
    };
    if (r < W) {
      for (int c = 0; c < C; ++c) {
        const cdouble v = col(c);
        re += TB.left[r * kMaxClosureWidth + c] * v.re;
        im += TB.left[r * kMaxClosureWidth + c] * v.im;
      }
    } else if (r + W >= N) {
      const int rr = N - 1 - r;
      for (int c = 0; c < C; ++c) {
        const cdouble v = col(N - 1 - c);
        re -= TB.left[rr * kMaxClosureWidth + c] * v.re;
        im -= TB.left[rr * kMaxClosureWidth + c] * v.im;
      }
    } else {
      for (int k = B; k >= 1; --k) {
        const cdouble hi = col(r + k), lo = col(r - k);
        re += TB.interior[k] * (hi.re - lo.re);
        im += TB.interior[k] * (hi.im - lo.im);
      }
    }
    q[t] = cmake(L.inv_h * re, L.inv_h * im);
  }
  __syncthreads();
  // nodal right-hand side and 1 / C_T (pipeline_initial_data.hpp:160-186)
  const RadialCoefficients rc = radial_coefficients(L.mass, L.spin, L.length, -2, L.m, R);
  for (int t = tid; t < T; t += kInitThreads) {
    const double ct = time_coefficient(rc, L.spin, sin_t[t]);
    w[t] = 1.0 / ct;
    const cdouble g = definition_coefficient(rc, L.spin, -2, cos_t[t]);
    const cdouble num = cadd(cscale(-2.0 * rc.adv, q[t]), cmul(g, psi[t]));
    pn[t] = cscale(w[t], num);  // nodal C_T^-1 (-2K Q + G Psi), reused below for P
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += kInitThreads) {
    const int i = e / n, j = e - i * n;
    double v = 0.0;
    for (int t = 0; t < T; ++t) v += A[i * T + t] * w[t] * S[t * n + j];
    Mx[e] = v;
    M0[e] = v;
  }
  for (int i = tid; i < n; i += kInitThreads) {
    double re = 0.0, im = 0.0;
    for (int t = 0; t < T; ++t) {
      re += A[i * T + t] * pn[t].re;
      im += A[i * T + t] * pn[t].im;
    }
    rhs[i] = cmake(re, im);
    b0[i] = rhs[i];
  }
  __syncthreads();
  // scale and singularity floor (dense_solve.hpp:27-31)
  double local = 0.0;
  for (int e = tid; e < n * n; e += kInitThreads) local = fmax(local, fabs(Mx[e]));
  red[tid] = local;
  __syncthreads();
  for (int s = kInitThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  const double floor = 128.0 * 2.220446049250313e-16 * red[0];
  __syncthreads();
  // elimination with partial pivoting: the first row of largest |a| wins
  for (int c = 0; c < n; ++c) {
    double v = -1.0;
    int idx = n;
    for (int rr = c + tid; rr < n; rr += kInitThreads) {
      const double m = fabs(Mx[rr * n + c]);
      if (m > v) {
        v = m;
        idx = rr;
      }
    }
    red[tid] = v;
    redi[tid] = idx;
    __syncthreads();
    for (int s = kInitThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const double o = red[tid + s];
        const int oi = redi[tid + s];
        if (o > red[tid] || (o == red[tid] && oi < redi[tid])) {
          red[tid] = o;
          redi[tid] = oi;
        }
      }
      __syncthreads();
    }
    const double best = red[0];
    const int piv = redi[0];
    __syncthreads();
    if (!isfinite(best) || best <= floor) {
      if (tid == 0) singular = 1;
      break;
    }
    if (piv != c) {
      for (int j = c + tid; j < n; j += kInitThreads) {
        const double tmp = Mx[c * n + j];
        Mx[c * n + j] = Mx[piv * n + j];
        Mx[piv * n + j] = tmp;
      }
      if (tid == 0) {
        const cdouble tmp = rhs[c];
        rhs[c] = rhs[piv];
        rhs[piv] = tmp;
      }
    }
    __syncthreads();
    const double d = Mx[c * n + c];
    for (int rr = c + 1 + tid; rr < n; rr += kInitThreads) {
      const double f = Mx[rr * n + c] / d;
      Mx[rr * n + c] = 0.0;
      for (int j = c + 1; j < n; ++j) Mx[rr * n + j] -= f * Mx[c * n + j];
      rhs[rr] = csub(rhs[rr], cscale(f, rhs[c]));
    }
    __syncthreads();
  }
  __syncthreads();
  if (singular) {
    if (tid == 0) atomicMax(a.residual + line, 0x7ff0000000000000ULL);  // +inf
    return;
  }
  // back substitution, the row of x[k] folded into the rows above it
  for (int k = n - 1; k >= 0; --k) {
    if (tid == 0) x[k] = cdivr(rhs[k], Mx[k * n + k]);
    __syncthreads();
    for (int rr = tid; rr < k; rr += kInitThreads)
      rhs[rr] = csub(rhs[rr], cscale(Mx[rr * n + k], x[k]));
    __syncthreads();
  }
  // relative residual |M x - b| / |b| (dense_solve.hpp:74-90)
  double rr2 = 0.0, bb2 = 0.0;
  for (int i = tid; i < n; i += kInitThreads) {
    double re = -b0[i].re, im = -b0[i].im;
    for (int j = 0; j < n; ++j) {
      re += M0[i * n + j] * x[j].re;
      im += M0[i * n + j] * x[j].im;
    }
    rr2 += re * re + im * im;
    bb2 += b0[i].re * b0[i].re + b0[i].im * b0[i].im;
  }
  red[tid] = rr2;
  __syncthreads();
  for (int s = kInitThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double rr_sum = red[0];
  __syncthreads();
  red[tid] = bb2;
  __syncthreads();
  for (int s = kInitThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    const double rel = sqrt(rr_sum) / fmax(sqrt(red[0]), 2.2250738585072014e-308);
    const double v = isfinite(rel) ? rel : __builtin_inf();
    atomicMax(a.residual + line, static_cast<unsigned long long>(__double_as_longlong(v)));
  }
  // P nodal = S x, then the modal coefficients A x of every field
  for (int t = tid; t < T; t += kInitThreads) {
    double re = 0.0, im = 0.0;
    for (int l = 0; l < n; ++l) {
      re += S[t * n + l] * x[l].re;
      im += S[t * n + l] * x[l].im;
    }
    pn[t] = cmake(re, im);
  }
  __syncthreads();
  if (a.nodal) {
    // rhs, b0 and x are free now and contiguous: the 3 n coefficients A v
    cdouble* coef = rhs;
    for (int e = tid; e < 3 * n; e += kInitThreads) {
      const int f = e / n, l = e - f * n;
      const cdouble* v = f == 0 ? pn : (f == 1 ? q : psi);
      double re = 0.0, im = 0.0;
      for (int t = 0; t < T; ++t) {
        re += A[l * T + t] * v[t].re;
        im += A[l * T + t] * v[t].im;
      }
      coef[e] = cmake(re, im);
    }
    __syncthreads();
    cdouble* out = a.nodal + il.nodal_base + static_cast<long long>(r) * T;
    const long long field_stride = static_cast<long long>(N) * T;
    for (int e = tid; e < 3 * T; e += kInitThreads) {
      const int f = e / T, t = e - f * T;
      double re = 0.0, im = 0.0;
      for (int l = 0; l < n; ++l) {
        re += S[t * n + l] * coef[f * n + l].re;
        im += S[t * n + l] * coef[f * n + l].im;
      }
      out[f * field_stride + t] = cmake(re, im);
    }
    return;
  }
  cdouble* out = a.y + L.base + static_cast<long long>(r) * 3 * n;
  for (int e = tid; e < 3 * n; e += kInitThreads) {
    const int f = e / n, l = e - f * n;
    const cdouble* v = f == 0 ? pn : (f == 1 ? q : psi);
    double re = 0.0, im = 0.0;
    for (int t = 0; t < T; ++t) {
      re += A[l * T + t] * v[t].re;
      im += A[l * T + t] * v[t].im;
    }
    out[e] = cmake(re, im);
  }
}

int galerkin_init_lds_bytes(int n_ell, int ntheta) {
  return 16 * n_ell * n_ell + 16 * 3 * n_ell + 16 * 4 * ntheta + 8 * ntheta;
}

hipError_t launch_galerkin_init(const InitArgs& a, int max_nr, int lds, hipStream_t s) {
  if (a.n_lines <= 0) return hipSuccess;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&teuk_galerkin_init),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(teuk_galerkin_init, dim3(max_nr, a.n_lines), dim3(kInitThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace teuk
