// Fused per-RK-stage Teukolsky kernels for gfx950.
//
// One workgroup owns one radial tile of one (instance, m) line and does the
// whole linear stage for it:
//   1. stage the tile (+ stencil halo) of (P, Q, Psi) modal coefficients
//      into LDS with coalesced 16-byte loads;
//   2. psi_t = A [(S(P + 2K Q_eff) - G_m S Psi) / C_T] on the halo rows
//      (the theta-dependent Kerr coupling as a small synthesis / pointwise /
//      analysis contraction in LDS; diagonal when a = 0);
//   3. SBP radial derivatives D Q_eff, D Psi, D psi_t and the compatible
//      dissipation from LDS, the point RHS
//        P_t = H_R D Q_eff + q Q_eff + psi Psi + Lambda_l Psi
//        Q_t = D psi_t - gamma (Q - D Psi)
//        Psi_t = psi_t
//      (reference: full_spatial.hpp:99-280, teukolsky.hpp:117-146), and
//   4. the classical RK4 stage combination written straight to HBM.
// The band projection of every stage derivative (spatial_pipeline.hpp:
// 499-500, 756-766) is implicit: the state lives in the retained band.
#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

namespace {

// h*D applied at global row `row` to values v(g) = base[(g - lo) * stride].
template <int SCHEME>
__device__ __forceinline__ cdouble apply_d(const SchemeTable& T, const cdouble* base, int stride,
                                           int lo, int row, int n, double inv_h) {
  constexpr int W = SCHEME == 2 ? 11 : (SCHEME == 1 ? 8 : 4);
  constexpr int C = SCHEME == 2 ? 16 : (SCHEME == 1 ? 12 : 6);
  constexpr int B = SCHEME == 2 ? 5 : (SCHEME == 1 ? 4 : 2);
  double re = 0.0, im = 0.0;
  if (row < W) {
    const double* c = &T.left[row * kMaxClosureWidth];
    const cdouble* v = base + (0 - lo) * stride;
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const cdouble x = v[k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else if (row + W >= n) {
    const double* c = &T.left[(n - 1 - row) * kMaxClosureWidth];
    const cdouble* v = base + (n - 1 - lo) * stride;
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const cdouble x = v[-k * stride];
      re -= c[k] * x.re;
      im -= c[k] * x.im;
    }
  } else {
    const cdouble* v = base + (row - lo) * stride;
#pragma unroll
    for (int k = B; k >= 1; --k) {
      const cdouble a = v[k * stride];
      const cdouble b = v[-k * stride];
      re += T.interior[k] * (a.re - b.re);
      im += T.interior[k] * (a.im - b.im);
    }
  }
  return cmake(inv_h * re, inv_h * im);
}

// -(eps/h) H^-1 A^T A at global row `row` (radial_discretization.hpp:
// 165-232), with the A^T A rows tabulated on the host.
template <int SCHEME>
__device__ __forceinline__ cdouble apply_diss(const SchemeTable& T, const cdouble* base,
                                              int stride, int lo, int row, int n, double h,
                                              double eps) {
  constexpr int O = SCHEME == 2 ? 6 : (SCHEME == 1 ? 5 : 3);
  constexpr int W = SCHEME == 2 ? 11 : (SCHEME == 1 ? 8 : 4);
  double re = 0.0, im = 0.0;
  if (row < O) {
    const double* c = T.diss_left[row];
    const cdouble* v = base + (0 - lo) * stride;
    for (int k = 0; k <= row + O; ++k) {
      const cdouble x = v[k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else if (row > n - 1 - O) {
    const int d = n - 1 - row;
    const double* c = T.diss_left[d];
    const cdouble* v = base + (n - 1 - lo) * stride;
    for (int k = 0; k <= d + O; ++k) {
      const cdouble x = v[-k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else {
    const cdouble* v = base + (row - lo) * stride;
#pragma unroll
    for (int k = -O; k <= O; ++k) {
      const cdouble x = v[k * stride];
      re += T.diss_interior[k + O] * x.re;
      im += T.diss_interior[k + O] * x.im;
    }
  }
  const int dist = row < n - 1 - row ? row : n - 1 - row;
  const double hw = dist < W ? T.norm[dist] : 1.0;
  const double f = -eps / (h * hw);
  return cmake(f * re, f * im);
}

}  // namespace

// STAGE 0: k -> kout.  STAGE 1..4: classical RK4 stage combination with the
// accumulator carrying y (16 state vectors of HBM traffic per step).
template <int SCHEME, int STAGE>
__global__ __launch_bounds__(kStageThreads) void teuk_stage_kernel(StageArgs a) {
  extern __shared__ double smem[];
  const TileDesc tile = a.tiles[blockIdx.x];
  const LineDesc L = a.lines[tile.line];
  const SchemeTable& T = *a.table;
  const int tid = threadIdx.x;
  const int nl = L.n_ell;
  const int nt = L.ntheta;
  const int n = L.nr;
  const int rows = tile.hi - tile.lo;
  const int mrows = tile.mhi - tile.mlo;
  const int us = 3 * nl;  // LDS row stride of the staged tile (cdouble)

  cdouble* U = reinterpret_cast<cdouble*>(smem);
  cdouble* PT = U + rows * us;
  cdouble* DP = PT + mrows * nl;
  cdouble* after_dp = L.constrained ? DP + mrows * nl : DP;
  double* Smat = reinterpret_cast<double*>(after_dp);
  double* Amat = Smat + nt * nl;
  cdouble* NOD = reinterpret_cast<cdouble*>(Amat + nt * nl);

  // 1. stage the tile rows [lo, hi) of all three fields.
  {
    const cdouble* src = a.yin + L.base + static_cast<long long>(tile.lo) * us;
    const int total = rows * us;
    for (int i = tid; i < total; i += kStageThreads) U[i] = src[i];
    if (L.kerr) {
      const double* m = a.mats + L.mat_off;
      for (int i = tid; i < 2 * nt * nl; i += kStageThreads) Smat[i] = m[i];
    }
  }
  __syncthreads();

  const double M = L.mass, spin = L.spin, len = L.length;
  // D Psi on the intermediate rows (stage-constrained Q_eff = D Psi).
  if (L.constrained) {
    for (int i = tid; i < mrows * nl; i += kStageThreads) {
      const int j = tile.mlo + i / nl, l = i % nl;
      DP[i] = apply_d<SCHEME>(T, U + 2 * nl + l, us, tile.lo, j, n, L.inv_h);
    }
    __syncthreads();
  }

  // 2. psi_t on rows [mlo, mhi).
  if (!L.kerr) {
    for (int i = tid; i < mrows * nl; i += kStageThreads) {
      const int j = tile.mlo + i / nl, l = i % nl;
      const RadialCoefficients rc = radial_coefficients(M, spin, len, -2, L.m, L.h * j);
      const double ct = time_coefficient(rc, spin, 0.0);
      const cdouble g = definition_coefficient(rc, spin, -2, 0.0);
      const cdouble* u = U + (j - tile.lo) * us;
      const cdouble P = u[l];
      const cdouble Q = L.constrained ? DP[i] : u[nl + l];
      const cdouble S = u[2 * nl + l];
      const double k2 = 2.0 * rc.adv;
      const cdouble num = csub(cadd(P, cscale(k2, Q)), cmul(g, S));
      PT[i] = cdivr(num, ct);
    }
    __syncthreads();
  } else {
    const double* sinv = a.nodes + L.node_off;
    const double* cosv = sinv + nt;
    const int ch = nodal_chunk_rows(nt);
    for (int c0 = tile.mlo; c0 < tile.mhi; c0 += ch) {
      const int cn = min(ch, tile.mhi - c0);
      for (int i = tid; i < cn * nt; i += kStageThreads) {
        const int j = c0 + i / nt, th = i % nt;
        const RadialCoefficients rc = radial_coefficients(M, spin, len, -2, L.m, L.h * j);
        const double k2 = 2.0 * rc.adv;
        const cdouble* u = U + (j - tile.lo) * us;
        const cdouble* dq = DP + (j - tile.mlo) * nl;
        const double* srow = Smat + th * nl;
        double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
        for (int l = 0; l < nl; ++l) {
          const double s = srow[l];
          const cdouble P = u[l];
          const cdouble Q = L.constrained ? dq[l] : u[nl + l];
          const cdouble S = u[2 * nl + l];
          xr += s * (P.re + k2 * Q.re);
          xi += s * (P.im + k2 * Q.im);
          yr += s * S.re;
          yi += s * S.im;
        }
        const double ct = time_coefficient(rc, spin, sinv[th]);
        const cdouble g = definition_coefficient(rc, spin, -2, cosv[th]);
        const cdouble num = csub(cmake(xr, xi), cmul(g, cmake(yr, yi)));
        NOD[i] = cdivr(num, ct);
      }
      __syncthreads();
      for (int i = tid; i < cn * nl; i += kStageThreads) {
        const int jj = i / nl, l = i % nl;
        const double* arow = Amat + l * nt;
        const cdouble* nod = NOD + jj * nt;
        double re = 0.0, im = 0.0;
        for (int th = 0; th < nt; ++th) {
          re += arow[th] * nod[th].re;
          im += arow[th] * nod[th].im;
        }
        PT[(c0 - tile.mlo + jj) * nl + l] = cmake(re, im);
      }
      __syncthreads();
    }
  }

  // 3-4. point RHS on the output rows and the RK combination.
  const double dt = a.dt_override > 0.0 ? a.dt_override : L.dt;
  const double gamma = L.constrained ? 0.0 : L.gamma;
  const int orows = tile.t1 - tile.t0;
  // Terms of the RHS that are linear in the stage with theta-independent
  // coefficients: the reference forms them on the nodal stage S c and
  // projects with S A, i.e. they carry G = A S in modal coordinates.
  auto linear_terms = [&](int r, int l, const cdouble* lap_in_ptr, cdouble& xp, cdouble& xq,
                          cdouble& xs) {
    const cdouble* u = U + (r - tile.lo) * us;
    const cdouble Q = u[nl + l], S = u[2 * nl + l];
    const cdouble dpsi = apply_d<SCHEME>(T, U + 2 * nl + l, us, tile.lo, r, n, L.inv_h);
    cdouble qeff, dq;
    if (L.constrained) {
      qeff = DP[(r - tile.mlo) * nl + l];
      dq = apply_d<SCHEME>(T, DP + l, nl, tile.mlo, r, n, L.inv_h);
    } else {
      qeff = Q;
      dq = apply_d<SCHEME>(T, U + nl + l, us, tile.lo, r, n, L.inv_h);
    }
    const RadialCoefficients rc = radial_coefficients(M, spin, len, -2, L.m, L.h * r);
    const int ell = L.ell_min + l;
    const double lam = -static_cast<double>((ell + 2) * (ell - 1));
    const cdouble lap_in = lap_in_ptr ? *lap_in_ptr : S;
    xp = cadd(cadd(cscale(rc.principal, dq), cmul(rc.q, qeff)),
              cadd(cmul(rc.psi, S), cscale(lam, lap_in)));
    xq = cscale(-gamma, csub(Q, dpsi));
    xs = cmake(0.0, 0.0);
    if (L.dissipate) {
      xp = cadd(xp, apply_diss<SCHEME>(T, U + l, us, tile.lo, r, n, L.h, L.eps));
      xq = cadd(xq, apply_diss<SCHEME>(T, U + nl + l, us, tile.lo, r, n, L.h, L.eps));
      xs = apply_diss<SCHEME>(T, U + 2 * nl + l, us, tile.lo, r, n, L.h, L.eps);
    }
  };
  auto store = [&](int r, int l, cdouble kp, cdouble kq, cdouble ks) {
    const long long g = L.base + static_cast<long long>(r) * us + l;
    if constexpr (STAGE == 0) {
      a.kout[g] = kp;
      a.kout[g + nl] = kq;
      a.kout[g + 2 * nl] = ks;
    } else {
      const cdouble* u = U + (r - tile.lo) * us;
      const cdouble k[3] = {kp, kq, ks};
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const long long gi = g + f * nl;
        if constexpr (STAGE == 1) {
          const cdouble y0 = u[f * nl + l];  // the stage-1 input is y, already in LDS
          a.acc[gi] = cadd(y0, cscale(dt / 6.0, k[f]));
          a.yout[gi] = cadd(y0, cscale(0.5 * dt, k[f]));
        } else if constexpr (STAGE == 2 || STAGE == 3) {
          const cdouble y0 = a.y[gi];
          a.acc[gi] = cadd(a.acc[gi], cscale(dt / 3.0, k[f]));
          a.yout[gi] = cadd(y0, cscale(STAGE == 2 ? 0.5 * dt : dt, k[f]));
        } else {
          a.yout[gi] = cadd(a.acc[gi], cscale(dt / 6.0, k[f]));
        }
      }
    }
  };

  if (!L.gram) {
    for (int i = tid; i < orows * nl; i += kStageThreads) {
      const int r = tile.t0 + i / nl, l = i % nl;
      cdouble xp, xq, xs;
      linear_terms(r, l, nullptr, xp, xq, xs);
      const cdouble dpt = apply_d<SCHEME>(T, PT + l, nl, tile.mlo, r, n, L.inv_h);
      const cdouble pt = PT[(r - tile.mlo) * nl + l];
      store(r, l, xp, cadd(dpt, xq), cadd(pt, xs));
    }
    return;
  }

  // Gram path: G in LDS, pre-projection terms X in LDS, then k = G X.
  double* Gm = reinterpret_cast<double*>(L.kerr ? NOD + nodal_chunk_rows(nt) * nt
                                               : reinterpret_cast<cdouble*>(Smat));
  cdouble* X = reinterpret_cast<cdouble*>(Gm + gram_doubles(nl));
  {
    const double* g = a.mats + L.mat_off + 2 * nt * nl;
    for (int i = tid; i < nl * nl; i += kStageThreads) Gm[i] = g[i];
  }
  __syncthreads();
  // X = pre-projection terms; the Laplacian acts on G Psi (the reference's
  // S Lambda A applied to the nodal stage S c).
  for (int i = tid; i < orows * nl; i += kStageThreads) {
    const int r = tile.t0 + i / nl, l = i % nl;
    const cdouble* u = U + (r - tile.lo) * us + 2 * nl;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nl; ++k) {
      re += Gm[l * nl + k] * u[k].re;
      im += Gm[l * nl + k] * u[k].im;
    }
    const cdouble gpsi = cmake(re, im);
    cdouble xp, xq, xs;
    linear_terms(r, l, &gpsi, xp, xq, xs);
    cdouble* x = X + (r - tile.t0) * us;
    x[l] = xp;
    x[nl + l] = xq;
    x[2 * nl + l] = xs;
  }
  __syncthreads();
  for (int i = tid; i < orows * nl; i += kStageThreads) {
    const int r = tile.t0 + i / nl, l = i % nl;
    const cdouble* x = X + (r - tile.t0) * us;
    const double* grow = Gm + l * nl;
    double pr = 0, pi = 0, qr = 0, qi = 0, sr = 0, si = 0;
    for (int k = 0; k < nl; ++k) {
      const double gk = grow[k];
      pr += gk * x[k].re;
      pi += gk * x[k].im;
      qr += gk * x[nl + k].re;
      qi += gk * x[nl + k].im;
      sr += gk * x[2 * nl + k].re;
      si += gk * x[2 * nl + k].im;
    }
    const cdouble dpt = apply_d<SCHEME>(T, PT + l, nl, tile.mlo, r, n, L.inv_h);
    const cdouble pt = PT[(r - tile.mlo) * nl + l];
    store(r, l, cmake(pr, pi), cadd(dpt, cmake(qr, qi)), cadd(pt, cmake(sr, si)));
  }
}

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
template <int SCHEME>
static hipError_t launch_stage_scheme(int stage, const StageArgs& a, int n_tiles, int lds,
                                      hipStream_t s) {
  const dim3 grid(n_tiles), block(kStageThreads);
  switch (stage) {
    case 0: hipLaunchKernelGGL((teuk_stage_kernel<SCHEME, 0>), grid, block, lds, s, a); break;
    case 1: hipLaunchKernelGGL((teuk_stage_kernel<SCHEME, 1>), grid, block, lds, s, a); break;
    case 2: hipLaunchKernelGGL((teuk_stage_kernel<SCHEME, 2>), grid, block, lds, s, a); break;
    case 3: hipLaunchKernelGGL((teuk_stage_kernel<SCHEME, 3>), grid, block, lds, s, a); break;
    default: hipLaunchKernelGGL((teuk_stage_kernel<SCHEME, 4>), grid, block, lds, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_stage(int scheme, int stage, const StageArgs& a, int n_tiles, int lds,
                        hipStream_t s) {
  if (scheme == 2) return launch_stage_scheme<2>(stage, a, n_tiles, lds, s);
  if (scheme == 1) return launch_stage_scheme<1>(stage, a, n_tiles, lds, s);
  return launch_stage_scheme<0>(stage, a, n_tiles, lds, s);
}

hipError_t set_stage_lds_limit(int bytes) {
  const void* fns[] = {
      reinterpret_cast<const void*>(&teuk_stage_kernel<0, 0>), reinterpret_cast<const void*>(&teuk_stage_kernel<0, 1>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<0, 2>), reinterpret_cast<const void*>(&teuk_stage_kernel<0, 3>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<0, 4>), reinterpret_cast<const void*>(&teuk_stage_kernel<1, 0>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<1, 1>), reinterpret_cast<const void*>(&teuk_stage_kernel<1, 2>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<1, 3>), reinterpret_cast<const void*>(&teuk_stage_kernel<1, 4>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<2, 0>), reinterpret_cast<const void*>(&teuk_stage_kernel<2, 1>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<2, 2>), reinterpret_cast<const void*>(&teuk_stage_kernel<2, 3>),
      reinterpret_cast<const void*>(&teuk_stage_kernel<2, 4>)};
  for (const void* f : fns) {
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
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
