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
    // closure rows are rare: keep them compact to save registers
#pragma unroll 2
    for (int k = 0; k < C; ++k) {
      const cdouble x = v[k * stride];
      re += c[k] * x.re;
      im += c[k] * x.im;
    }
  } else if (row + W >= n) {
    const double* c = &T.left[(n - 1 - row) * kMaxClosureWidth];
    const cdouble* v = base + (n - 1 - lo) * stride;
#pragma unroll 2
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

// Global -> LDS copy with kCopyBatch loads in flight per thread before the
// first LDS write (a plain loop would wait for every load in turn).
constexpr int kCopyBatch = 8;
template <class V>
__device__ __forceinline__ void stage_copy(V* __restrict__ dst, const V* __restrict__ src,
                                           int count, int tid) {
  for (int base = tid; base < count; base += kCopyBatch * kStageThreads) {
    V tmp[kCopyBatch];
    // Clamped indices keep every load and store unconditional, so all loads
    // are in flight before the first wait (out-of-range lanes rewrite the
    // last element with its own value).
#pragma unroll
    for (int k = 0; k < kCopyBatch; ++k) tmp[k] = src[min(base + k * kStageThreads, count - 1)];
#pragma unroll
    for (int k = 0; k < kCopyBatch; ++k) dst[min(base + k * kStageThreads, count - 1)] = tmp[k];
  }
}

}  // namespace

// Per-row radial coefficients, tabulated once per tile row in LDS.
struct RowCoeffs {
  double adv2;         // 2 K
  double principal;    // H_R
  double time_radial;  // theta-independent part of C_T
  double g_real, g_imag_m;
  double q_re, q_im, psi_re, psi_im;
  double pad;
};
static_assert(sizeof(RowCoeffs) == 80, "RowCoeffs must stay 16-byte aligned");

// STAGE 0: k -> kout.  STAGE 1..4: classical RK4 stage combination with the
// accumulator carrying y (16 state vectors of HBM traffic per step).
template <int SCHEME, int STAGE>
__global__ __launch_bounds__(kStageThreads, 2) void teuk_stage_kernel(StageArgs a) {
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
  const int orows = tile.t1 - tile.t0;
  const int us = 3 * nl;  // LDS row stride of the staged tile (cdouble)

  // (row, ell) and (row, theta) thread decompositions: one division each.
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  const bool on_l = row_l < step_l;
  const int lane_t = tid % nt, row_t = tid / nt, step_t = kStageThreads / nt;
  const bool on_t = row_t < step_t;

  RowCoeffs* RC = reinterpret_cast<RowCoeffs*>(smem);
  cdouble* U = reinterpret_cast<cdouble*>(RC + mrows);
  cdouble* PT = U + rows * us;
  cdouble* DP = PT + mrows * nl;
  cdouble* after_dp = L.constrained ? DP + mrows * nl : DP;
  double* Smat = reinterpret_cast<double*>(after_dp);
  double* Amat = Smat + nt * nl;
  cdouble* NOD = reinterpret_cast<cdouble*>(Amat + nt * nl);

  const double M = L.mass, spin = L.spin, len = L.length;
  // 1. stage the tile rows [lo, hi) of all three fields, the band matrices
  //    and the per-row coefficients.
  {
    stage_copy(U, a.yin + L.base + static_cast<long long>(tile.lo) * us, rows * us, tid);
    if (L.kerr) stage_copy(Smat, a.mats + L.mat_off, 2 * nt * nl, tid);
    for (int j = tid; j < mrows; j += kStageThreads) {
      const RadialCoefficients rc =
          radial_coefficients(M, spin, len, -2, L.m, L.h * (tile.mlo + j));
      RC[j] = {2.0 * rc.adv, rc.principal, rc.time_radial, rc.g_real, rc.g_imag_m,
               rc.q.re,      rc.q.im,      rc.psi.re,      rc.psi.im, 0.0};
    }
  }
  __syncthreads();

  // D Psi on the intermediate rows (stage-constrained Q_eff = D Psi).
  if (L.constrained) {
    if (on_l)
      for (int j = tile.mlo + row_l; j < tile.mhi; j += step_l)
        DP[(j - tile.mlo) * nl + lane_l] =
            apply_d<SCHEME>(T, U + 2 * nl + lane_l, us, tile.lo, j, n, L.inv_h);
    __syncthreads();
  }

  // 2. psi_t on rows [mlo, mhi).  First X = P + 2K Q_eff (into PT's slots).
  if (on_l)
    for (int j = tile.mlo + row_l; j < tile.mhi; j += step_l) {
      const int jj = j - tile.mlo;
      const cdouble* u = U + (j - tile.lo) * us;
      const cdouble Q = L.constrained ? DP[jj * nl + lane_l] : u[nl + lane_l];
      const double k2 = RC[jj].adv2;
      const cdouble P = u[lane_l];
      if (!L.kerr) {
        // a = 0: C_T and G_m are theta independent, the coupling is diagonal.
        const cdouble g = cmake(RC[jj].g_real, RC[jj].g_imag_m);
        const cdouble num = csub(cadd(P, cscale(k2, Q)), cmul(g, u[2 * nl + lane_l]));
        PT[jj * nl + lane_l] = cdivr(num, RC[jj].time_radial);
      } else {
        PT[jj * nl + lane_l] = cadd(P, cscale(k2, Q));
      }
    }
  __syncthreads();
  if (L.kerr) {
    const double* sinv = a.nodes + L.node_off;
    const double* cosv = sinv + nt;
    const double sn = on_t ? sinv[lane_t] : 0.0, cs = on_t ? cosv[lane_t] : 0.0;
    const int ch = nodal_chunk_rows(nt);
    for (int c0 = tile.mlo; c0 < tile.mhi; c0 += ch) {
      const int c1 = min(c0 + ch, tile.mhi);
      // synthesis and the pointwise (X - G_m Psi) / C_T at (j, theta)
      if (on_t)
        for (int j = c0 + row_t; j < c1; j += step_t) {
          const int jj = j - tile.mlo;
          const cdouble* x = PT + jj * nl;
          const cdouble* ps = U + (j - tile.lo) * us + 2 * nl;
          const double* srow = Smat + lane_t * nl;
          double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
#pragma unroll 4
          for (int l = 0; l < nl; ++l) {
            const double w = srow[l];
            const cdouble xv = x[l], pv = ps[l];
            xr += w * xv.re;
            xi += w * xv.im;
            yr += w * pv.re;
            yi += w * pv.im;
          }
          const double ct = RC[jj].time_radial - spin * spin * sn * sn;
          const cdouble g = cmake(RC[jj].g_real, RC[jj].g_imag_m + 2.0 * (-2.0 * spin * cs));
          const cdouble num = csub(cmake(xr, xi), cmul(g, cmake(yr, yi)));
          NOD[(j - c0) * nt + lane_t] = cdivr(num, ct);
        }
      __syncthreads();
      // analysis A (nodal psi_t) -> PT rows of this chunk
      if (on_l)
        for (int j = c0 + row_l; j < c1; j += step_l) {
          const double* arow = Amat + lane_l * nt;
          const cdouble* nod = NOD + (j - c0) * nt;
          double re = 0.0, im = 0.0;
#pragma unroll 4
          for (int th = 0; th < nt; ++th) {
            re += arow[th] * nod[th].re;
            im += arow[th] * nod[th].im;
          }
          PT[(j - tile.mlo) * nl + lane_l] = cmake(re, im);
        }
      __syncthreads();
    }
  }

  // 3-4. point RHS on the output rows and the RK combination.
  const double dt = a.dt_override > 0.0 ? a.dt_override : L.dt;
  const double gamma = L.constrained ? 0.0 : L.gamma;
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
    const RowCoeffs& rc = RC[r - tile.mlo];
    const int ell = L.ell_min + l;
    const double lam = -static_cast<double>((ell + 2) * (ell - 1));
    const cdouble lap_in = lap_in_ptr ? *lap_in_ptr : S;
    xp = cadd(cadd(cscale(rc.principal, dq), cmul(cmake(rc.q_re, rc.q_im), qeff)),
              cadd(cmul(cmake(rc.psi_re, rc.psi_im), S), cscale(lam, lap_in)));
    xq = cscale(-gamma, csub(Q, dpsi));
    xs = cmake(0.0, 0.0);
    if (L.dissipate) {
      xp = cadd(xp, apply_diss<SCHEME>(T, U + l, us, tile.lo, r, n, L.h, L.eps));
      xq = cadd(xq, apply_diss<SCHEME>(T, U + nl + l, us, tile.lo, r, n, L.h, L.eps));
      xs = apply_diss<SCHEME>(T, U + 2 * nl + l, us, tile.lo, r, n, L.h, L.eps);
    }
  };
  // RK registers of one output point, loaded before the stencil work so the
  // global latency overlaps it (the stores below could alias these loads, so
  // the compiler would not hoist them by itself).
  struct RkIn {
    cdouble y0[3], acc0[3];
  };
  auto rk_load = [&](int r, int l) {
    RkIn v{};
    const long long g = L.base + static_cast<long long>(r) * us + l;
    if constexpr (STAGE == 2 || STAGE == 3) {
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        v.y0[f] = a.y[g + f * nl];
        v.acc0[f] = a.acc[g + f * nl];
      }
    } else if constexpr (STAGE == 4) {
#pragma unroll
      for (int f = 0; f < 3; ++f) v.acc0[f] = a.acc[g + f * nl];
    }
    return v;
  };
  auto store = [&](int r, int l, const RkIn& in, cdouble kp, cdouble kq, cdouble ks) {
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
          a.acc[gi] = cadd(in.acc0[f], cscale(dt / 3.0, k[f]));
          a.yout[gi] = cadd(in.y0[f], cscale(STAGE == 2 ? 0.5 * dt : dt, k[f]));
        } else {
          a.yout[gi] = cadd(in.acc0[f], cscale(dt / 6.0, k[f]));
        }
      }
    }
  };

  if (!L.gram) {
    if (on_l)
      for (int r = tile.t0 + row_l; r < tile.t1; r += step_l) {
        const RkIn in = rk_load(r, lane_l);
        cdouble xp, xq, xs;
        linear_terms(r, lane_l, nullptr, xp, xq, xs);
        const cdouble dpt = apply_d<SCHEME>(T, PT + lane_l, nl, tile.mlo, r, n, L.inv_h);
        const cdouble pt = PT[(r - tile.mlo) * nl + lane_l];
        store(r, lane_l, in, xp, cadd(dpt, xq), cadd(pt, xs));
      }
    return;
  }

  // Gram path: G in LDS, pre-projection terms X in LDS, then k = G X.
  double* Gm = reinterpret_cast<double*>(L.kerr ? NOD + nodal_chunk_rows(nt) * nt
                                               : reinterpret_cast<cdouble*>(Smat));
  cdouble* X = reinterpret_cast<cdouble*>(Gm + gram_doubles(nl));
  {
    stage_copy(Gm, a.mats + L.mat_off + 2 * nt * nl, nl * nl, tid);
  }
  __syncthreads();
  // X = pre-projection terms; the Laplacian acts on G Psi (the reference's
  // S Lambda A applied to the nodal stage S c).
  if (on_l)
    for (int r = tile.t0 + row_l; r < tile.t1; r += step_l) {
      const cdouble* u = U + (r - tile.lo) * us + 2 * nl;
      double re = 0.0, im = 0.0;
#pragma unroll 4
      for (int k = 0; k < nl; ++k) {
        re += Gm[lane_l * nl + k] * u[k].re;
        im += Gm[lane_l * nl + k] * u[k].im;
      }
      const cdouble gpsi = cmake(re, im);
      cdouble xp, xq, xs;
      linear_terms(r, lane_l, &gpsi, xp, xq, xs);
      cdouble* x = X + (r - tile.t0) * us;
      x[lane_l] = xp;
      x[nl + lane_l] = xq;
      x[2 * nl + lane_l] = xs;
    }
  __syncthreads();
  if (on_l)
    for (int r = tile.t0 + row_l; r < tile.t1; r += step_l) {
      const RkIn in = rk_load(r, lane_l);
      const cdouble* x = X + (r - tile.t0) * us;
      const double* grow = Gm + lane_l * nl;
      double pr = 0, pi = 0, qr = 0, qi = 0, sr = 0, si = 0;
#pragma unroll 4
      for (int k = 0; k < nl; ++k) {
        const double gk = grow[k];
        pr += gk * x[k].re;
        pi += gk * x[k].im;
        qr += gk * x[nl + k].re;
        qi += gk * x[nl + k].im;
        sr += gk * x[2 * nl + k].re;
        si += gk * x[2 * nl + k].im;
      }
      const cdouble dpt = apply_d<SCHEME>(T, PT + lane_l, nl, tile.mlo, r, n, L.inv_h);
      const cdouble pt = PT[(r - tile.mlo) * nl + lane_l];
      store(r, lane_l, in, cmake(pr, pi), cadd(dpt, cmake(qr, qi)), cadd(pt, cmake(sr, si)));
    }
  (void)orows;
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
