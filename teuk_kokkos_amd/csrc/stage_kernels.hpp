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
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_layout.hpp"
#include "sbp_device.hpp"

namespace teuk {

namespace {

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

typedef double mfma_acc __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_acc mfma_f64(double a, double b, mfma_acc c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Kerr coupling psi_t = A[(S X - G_m S Psi) / C_T] of the 16 rows
// [g0, g0 + 16) (tile-relative, clipped to mrows) by one wave, with
// X = P + 2K Q_eff read from PT (row stride pts) and overwritten there by
// psi_t:
//   synthesis   (S X^T)[theta][row]     16 x 16 tiles, FP64 MFMA
//   pointwise   ((S X) - G_m (S Psi)) / C_T on the accumulator registers
//   analysis    psi[row][ell] += N[row][theta] A^T[theta][ell]
// The accumulator of the synthesis is the A operand of the analysis, so the
// nodal values never leave registers.  The S and A fragments come from
// global memory (cache resident, shared by every tile of the band), one
// theta tile ahead; sin and cos of the nodes from LDS.  UB = ell tiles of 16
// held at once (K <= 4 UB blocks).
// S and A fragments of theta tile t, lane `lane` (KB = 4 UB S blocks, 4 x UB
// A blocks).
template <int UB>
__device__ __forceinline__ void kerr_load_tile(const double* Sf, const double* Af, int K, int LT,
                                               int t, int lane, double (&sf)[4 * UB],
                                               double (&af)[4][UB]) {
  const double* s = Sf + t * K * 64 + lane;
#pragma unroll
  for (int q = 0; q < 4 * UB; ++q) sf[q] = q < K ? s[q * 64] : 0.0;
  const double* a = Af + t * 4 * LT * 64 + lane;
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int u = 0; u < UB; ++u) af[v][u] = u < LT ? a[(v * LT + u) * 64] : 0.0;
}

// s0, a0: theta tile 0's fragments when `have0` (requested before X was
// formed, so their L2 latency is already spent).
template <int UB>
__device__ __forceinline__ void kerr_psi_group(int nl, int nt, double spin, const RowCoeffs* RC,
                                               cdouble* PT, int pts, const cdouble* psi, int us,
                                               int g0, int mrows, const double* Sf,
                                               const double* Af, const double* nodes,
                                               int lane, const double (&s0)[4 * UB],
                                               const double (&a0)[4][UB], bool have0) {
  constexpr int KB = 4 * UB;
  const int T = mfma_theta_tiles(nt), K = mfma_ell_blocks(nl), LT = mfma_ell_tiles(nl);
  const int lr = lane & 15, lg = lane >> 4;
  const int jj = g0 + lr;
  const bool row_ok = jj < mrows;
  const int jc = row_ok ? jj : mrows - 1;
  const cdouble* x = PT + jc * pts;
  const cdouble* ps = psi + jc * us;
  const double tau = RC[jc].time_radial, gre = RC[jc].g_real, gim = RC[jc].g_imag_m;
  const double a2 = spin * spin, c4 = -4.0 * spin;
  const double* sinv = nodes;
  const double* cosv = nodes + nt;
  mfma_acc ar[UB], ai[UB];
#pragma unroll
  for (int u = 0; u < UB; ++u) ar[u] = ai[u] = mfma_acc{0.0, 0.0, 0.0, 0.0};
  // S and A fragments of one theta tile; the next tile's are in flight while
  // the current one is processed.
  double s_cur[KB], s_nxt[KB], a_cur[4][UB], a_nxt[4][UB];
  auto load_tile = [&](int t, double (&sf)[KB], double (&af)[4][UB]) {
    kerr_load_tile<UB>(Sf, Af, K, LT, t, lane, sf, af);
  };
  if (have0) {
#pragma unroll
    for (int q = 0; q < KB; ++q) s_cur[q] = s0[q];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int u = 0; u < UB; ++u) a_cur[v][u] = a0[v][u];
  } else {
    load_tile(0, s_cur, a_cur);
  }
  // Wait here for every load issued so far (tile 0's fragments among them).
  // Otherwise the compiler's wait at the loop head, which must cover them on
  // the first pass, also covers the next tile's fragments on every later
  // pass, and their cache latency is paid in full once per theta tile.
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt open
  for (int t = 0; t < T; ++t) {
    if (t + 1 < T) load_tile(t + 1, s_nxt, a_nxt);
    mfma_acc sxr = {0.0, 0.0, 0.0, 0.0}, sxi = sxr, spr = sxr, spi = sxr;
#pragma unroll
    for (int q = 0; q < KB; ++q) {
      if (q >= K) break;
      // B operands X, Psi of (row jj, ell 4 q + lg) from LDS.  (Reading
      // block q + 1's while block q's products run, and keeping block 0's
      // for every theta tile, measured slower: more registers.)
      const int l = 4 * q + lg;
      const bool ok = row_ok && l < nl;
      const int lc = ok ? l : 0;
      const cdouble xv = x[lc], pv = ps[lc];
      sxr = mfma_f64(s_cur[q], ok ? xv.re : 0.0, sxr);
      sxi = mfma_f64(s_cur[q], ok ? xv.im : 0.0, sxi);
      spr = mfma_f64(s_cur[q], ok ? pv.re : 0.0, spr);
      spi = mfma_f64(s_cur[q], ok ? pv.im : 0.0, spi);
    }
    // element v of this lane: theta = 16 t + lg + 4 v, radial row jj
    double nre[4], nim[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int th = 16 * t + lg + 4 * v;
      const bool tok = th < nt;
      const int tc = tok ? th : 0;
      const double sn = sinv[tc], cs = cosv[tc];
      const double ct = tau - a2 * sn * sn;
      const double gi = gim + c4 * cs;
      const double pr = spr[v], pi = spi[v];
      const double num_re = sxr[v] - (gre * pr - gi * pi);
      const double num_im = sxi[v] - (gre * pi + gi * pr);
      // one reciprocal and two products instead of two divisions (1.3 % of
      // the sweep's stage time), the reciprocal from the hardware estimate
      // and two Newton steps (C_T is positive and normal on the grid; the
      // result differs from the quotient by about an ulp)
      double rct = __builtin_amdgcn_rcp(ct);
      rct = __builtin_fma(__builtin_fma(-ct, rct, 1.0), rct, rct);
      rct = __builtin_fma(__builtin_fma(-ct, rct, 1.0), rct, rct);
      nre[v] = tok ? num_re * rct : 0.0;
      nim[v] = tok ? num_im * rct : 0.0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      // theta block 16 t + 4 v .. + 3: past the last node it is all
      // padding (zero operands), so its products are skipped
      if (16 * t + 4 * v >= nt) break;
#pragma unroll
      for (int u = 0; u < UB; ++u)
        if (u < LT) {
          ar[u] = mfma_f64(nre[v], a_cur[v][u], ar[u]);
          ai[u] = mfma_f64(nim[v], a_cur[v][u], ai[u]);
        }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) s_cur[q] = s_nxt[q];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int u = 0; u < UB; ++u) a_cur[v][u] = a_nxt[v][u];
  }
  // element w of this lane: psi_t at row g0 + lg + 4 w, ell 16 u + lr.  The
  // wave's LDS reads of X were issued before these writes, and no other wave
  // touches rows of this group, so X can be overwritten in place.
#pragma unroll
  for (int u = 0; u < UB; ++u)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int row = g0 + lg + 4 * w, l = 16 * u + lr;
      if (u < LT && row < mrows && l < nl) PT[row * pts + l] = cmake(ar[u][w], ai[u][w]);
    }
}

// Stage outputs written with the nontemporal hint (one 128-bit store each):
// measured faster for the Gram bands' group-of-16-ell store pattern and
// slower for the narrow bands' row-major one, so rk_store takes it as a
// parameter.  Loads stay plain (the hint slowed every class).
typedef double double2v __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ void store_out(cdouble* p, cdouble v) {
  if constexpr (NT)
    __builtin_nontemporal_store(double2v{v.re, v.im}, reinterpret_cast<double2v*>(p));
  else
    *p = v;
}

// Classical RK4 stage combination of one value (STAGE 0: k -> kout), in a
// form that moves 12 state vectors per step through HBM.  With Y_s the
// input of stage s (Y_1 = y) and the partial sum C = y + dt/6 (k1 + 2 k2) - y/3:
//   S1: Y2 = y + dt/2 k1
//   S2: Y3 = y + dt/2 k2,  C = (y + Y2)/3 + dt/3 k2   (dt/6 k1 = (Y2 - y)/3)
//   S3: Y4 = y + dt k3
//   S4: y  = C + Y4/3 + dt/6 k4                        (dt/3 k3 = (Y4 - y)/3)
// which is y + dt/6 (k1 + 2 k2 + 2 k3 + k4).  Reads per stage: Y_s (staged)
// plus y in S2 and S3, C in S4; writes: Y_{s+1}, plus C in S2.  Every term is
// of the size of the state, so the rewriting costs no accuracy.
struct RkIn {
  cdouble v;   // y (stages 2, 3) or C (stage 4)
  cdouble ys;  // Y_s at the point, when re-read from HBM
};

// need_ys: the stage input at the point is read from HBM (an L2 hit: the
// tile staged it this launch) because psi_t overwrote the staged copy of P.
template <int STAGE>
__device__ __forceinline__ RkIn rk_load(const StageArgs& a, long long gi, bool need_ys = false) {
  RkIn v{};
  if constexpr (STAGE == 2 || STAGE == 3) v.v = a.y[gi];
  if constexpr (STAGE == 4) v.v = a.acc[gi];
  if constexpr (STAGE == 1 || STAGE == 2 || STAGE == 4)
    if (need_ys) v.ys = a.yin[gi];
  return v;
}

// c k + b with explicit fused multiply-adds, so that every instantiation of
// the stage kernels rounds the combination identically.
__device__ __forceinline__ cdouble cfma(double c, cdouble k, cdouble b) {
  return cmake(__builtin_fma(c, k.re, b.re), __builtin_fma(c, k.im, b.im));
}

// ys: the stage input Y_s at the point.
template <int STAGE, bool NT = false>
__device__ __forceinline__ void rk_store(const StageArgs& a, long long gi, const RkIn& in,
                                         cdouble ys, cdouble k, double dt) {
  constexpr double third = 1.0 / 3.0;
  if constexpr (STAGE == 0) {
    store_out<NT>(a.kout + gi, k);
  } else if constexpr (STAGE == 1) {
    store_out<NT>(a.yout + gi, cfma(0.5 * dt, k, ys));
  } else if constexpr (STAGE == 2) {
    const cdouble mean = cmake(third * (in.v.re + ys.re), third * (in.v.im + ys.im));
    store_out<NT>(a.acc + gi, cfma(dt * third, k, mean));
    store_out<NT>(a.yout + gi, cfma(0.5 * dt, k, in.v));
  } else if constexpr (STAGE == 3) {
    store_out<NT>(a.yout + gi, cfma(dt, k, in.v));
  } else {
    store_out<NT>(a.yout + gi, cfma(dt / 6.0, k, cfma(third, ys, in.v)));
  }
}

// Per-tile context shared by the output routines.
struct TileCtx {
  const LineDesc* L;
  const TileDesc* tile;
  const SchemeTable* T;
  const RowCoeffs* RC;  // coefficients of row r at RC[r - rc_lo]
  int rc_lo;
  const cdouble* U;   // staged rows [lo, hi): P | Q | Psi per row
  const cdouble* PT;  // psi_t on rows [mlo, mhi), row stride pts
  const cdouble* DP;  // D Psi on rows [mlo, mhi) (stage-constrained only)
  int pts;
  bool inplace;       // psi_t overwrote P in U (P is then read from HBM)
  double dt, gamma;
};

// Pre-projection (theta-independent, linear) terms of the RHS at (r, l):
// the reference forms them on the nodal stage S c and projects with S A,
// i.e. they carry G = A S in modal coordinates (full_spatial.hpp:99-280).
//   P: H_R D Q_eff + q Q_eff + psi Psi [+ Lambda_l lap_in] (+ dissipation)
//   Q: -gamma (Q - D Psi) (+ dissipation)        Psi: (dissipation)
template <int SCHEME, int FIELD, bool WITH_LAP = true, bool IN = false>
__device__ __forceinline__ cdouble pre_projection(const TileCtx& c, int r, int l, cdouble lap_in) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl, n = L.nr;
  const cdouble* u = c.U + (r - tile.lo) * us;
  cdouble x;
  if constexpr (FIELD == 0) {
    cdouble qeff, dq;
    if (L.constrained) {
      qeff = c.DP[(r - tile.mlo) * nl + l];
      dq = apply_d<SCHEME, IN>(*c.T, c.DP + l, nl, tile.mlo, r, n, L.inv_h);
    } else {
      qeff = u[nl + l];
      dq = apply_d<SCHEME, IN>(*c.T, c.U + nl + l, us, tile.lo, r, n, L.inv_h);
    }
    const RowCoeffs& rc = c.RC[r - c.rc_lo];
    const double lam = -static_cast<double>((L.ell_min + l + 2) * (L.ell_min + l - 1));
    const cdouble radial = cadd(cscale(rc.principal, dq), cmul(cmake(rc.q_re, rc.q_im), qeff));
    const cdouble coupling = cmul(cmake(rc.psi_re, rc.psi_im), u[2 * nl + l]);
    if constexpr (WITH_LAP)
      x = cadd(radial, cadd(coupling, cscale(lam, lap_in)));
    else
      x = cadd(radial, coupling);
  } else if constexpr (FIELD == 1) {
    const cdouble dpsi = apply_d<SCHEME, IN>(*c.T, c.U + 2 * nl + l, us, tile.lo, r, n, L.inv_h);
    x = cscale(-c.gamma, csub(u[nl + l], dpsi));
  } else {
    x = cmake(0.0, 0.0);
  }
  if (L.dissipate)
    x = cadd(x, apply_diss<SCHEME>(*c.T, c.U + FIELD * nl + l, us, tile.lo, r, n, L.h, L.eps));
  return x;
}

// Output rows of the tile, one (row, ell) point per thread, for bands whose
// A S is the identity to roundoff.  A thread walks rows row_l, row_l + step,
// ...; the RK inputs of its first kPlainDepth rows are requested when the
// tile starts (plain_prefetch, before the staging copy, so they arrive while
// the tile is staged and psi_t is formed) and each consumed slot is refilled
// with the row kPlainDepth steps ahead.
// Rows in flight per thread: three for narrow bands (measured best of 2 to
// 6 with the present staging order; four was best before it); wide bands
// (UB > 1) keep two, within their register budget.
template <int UB>
inline constexpr int kPlainDepth = UB == 1 ? 3 : 2;
template <int STAGE>
__device__ __forceinline__ void plain_fetch(const StageArgs& a, const LineDesc& L, bool inplace,
                                            int r, int lane_l, RkIn (&v)[3]) {
  const int nl = L.n_ell;
  const long long g = L.base + static_cast<long long>(r) * 3 * nl + lane_l;
#pragma unroll
  for (int f = 0; f < 3; ++f) v[f] = rk_load<STAGE>(a, g + f * nl, f == 0 && inplace);
}

template <int STAGE, int D>
__device__ __forceinline__ void plain_prefetch(const StageArgs& a, const LineDesc& L,
                                               const TileDesc& tile, bool inplace, int tid,
                                               RkIn (&in)[D][3]) {
  const int nl = L.n_ell;
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  if (row_l >= step_l) return;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const int r = tile.t0 + row_l + j * step_l;
    if (r < tile.t1) plain_fetch<STAGE>(a, L, inplace, r, lane_l, in[j]);
  }
}

template <int SCHEME, int STAGE, int D, bool IN>
__device__ __forceinline__ void plain_output(const StageArgs& a, const TileCtx& c, int tid,
                                             RkIn (&in)[D][3]) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl, n = L.nr;
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  if (row_l >= step_l) return;
  for (int r0 = tile.t0 + row_l; r0 < tile.t1; r0 += D * step_l) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int r = r0 + j * step_l;
      if (r >= tile.t1) break;
      const long long g = L.base + static_cast<long long>(r) * us + lane_l;
      const cdouble* u = c.U + (r - tile.lo) * us;
      const cdouble kp = pre_projection<SCHEME, 0, true, IN>(c, r, lane_l, u[2 * nl + lane_l]);
      const cdouble xq = pre_projection<SCHEME, 1, true, IN>(c, r, lane_l, u[2 * nl + lane_l]);
      const cdouble xs = pre_projection<SCHEME, 2, true, IN>(c, r, lane_l, u[2 * nl + lane_l]);
      const cdouble dpt = apply_d<SCHEME, IN>(*c.T, c.PT + lane_l, c.pts, tile.mlo, r, n, L.inv_h);
      const cdouble pt = c.PT[(r - tile.mlo) * c.pts + lane_l];
      const RkIn (&v)[3] = in[j];
      rk_store<STAGE>(a, g, v[0], c.inplace ? v[0].ys : u[lane_l], kp, c.dt);
      rk_store<STAGE>(a, g + nl, v[1], u[nl + lane_l], cadd(dpt, xq), c.dt);
      rk_store<STAGE>(a, g + 2 * nl, v[2], u[2 * nl + lane_l], cadd(pt, xs), c.dt);
      const int rf = r + D * step_l;
      if (rf < tile.t1) plain_fetch<STAGE>(a, L, c.inplace, rf, lane_l, in[j]);
    }
  }
}

typedef float mfma_accf __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mfma_accf mfma_f32(float a, float b, mfma_accf c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Gram bands: the projection G = A S of the pre-projection terms X and the
// Laplacian M Psi (M = G Lambda G) are applied split as G = I + E and
// M = Lambda + N.  E and N are of the size of A S - I (6e-12 at ell 24,
// 2e-9 at ell 32), so
//   k_P = [X_P + Lambda Psi] + [E X_P + N Psi],  k_Q = [D psi_t + X_Q] + E X_Q,
//   k_Psi = [psi_t + X_Psi] + E X_Psi  (X_Psi: dissipation only)
// with the first bracket (the identity part, gram_identity) in FP64 at each
// point and the correction on FP32 MFMA (operands, E, N and sums in FP32):
// its relative error, about 1e-6, is 1e-17 of k, below FP64 rounding.
// Both output routines below form these two parts with the same functions in
// the same order, so they agree bit for bit.
template <int SCHEME, int F, bool IN>
__device__ __forceinline__ cdouble gram_identity(const TileCtx& c, int r, int l, cdouble x) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell;
  if constexpr (F == 0) {
    const double lam = -static_cast<double>((L.ell_min + l + 2) * (L.ell_min + l - 1));
    return cadd(x, cscale(lam, c.U[(r - tile.lo) * 3 * nl + 2 * nl + l]));
  } else if constexpr (F == 1) {
    return cadd(apply_d<SCHEME, IN>(*c.T, c.PT + l, c.pts, tile.mlo, r, L.nr, L.inv_h), x);
  } else {
    return cadd(c.PT[(r - tile.mlo) * c.pts + l], x);
  }
}

// One contraction block kb of the correction of field F into the FP32
// accumulators (uo = output ell tile): E X (+ N Psi for P).  Fragment (uo,
// kb), lane: E[16 uo + lr][4 kb + lg] (gram_correction_floats).
template <int F, int UB>
__device__ __forceinline__ void gram_correction_step(int LT, int K, int kb, int lane,
                                                     const float* Ef, const float* Nf, float xr,
                                                     float xi, float pr, float pi,
                                                     mfma_accf (&cr)[UB], mfma_accf (&ci)[UB]) {
#pragma unroll
  for (int uo = 0; uo < UB; ++uo) {
    if (uo >= LT) break;
    const float e = Ef[(uo * K + kb) * 64 + lane];
    cr[uo] = mfma_f32(xr, e, cr[uo]);
    ci[uo] = mfma_f32(xi, e, ci[uo]);
    if constexpr (F == 0) {
      const float nn = Nf[(uo * K + kb) * 64 + lane];
      cr[uo] = mfma_f32(pr, nn, cr[uo]);
      ci[uo] = mfma_f32(pi, nn, ci[uo]);
    }
  }
}

// BF16 form of the correction for lines with LineDesc::gram_bf16 (E and N
// below 2e-11 of the terms they correct): one v_mfma_f32_16x16x32_bf16 per
// (output ell tile, matrix, real or imaginary part) covers the whole band
// (n_ell <= 32).  Operand lane (lr, lg): row lr, ell 8 lg + j, j = 0..7; the
// FP32 operand values rounded once more to BF16.  The correction's own
// relative error, about 4e-3, is below 1e-13 of k.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ mfma_accf mfma_bf16(bf16x8 a, bf16x8 b, mfma_accf c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <int UB>
struct GramFragmentsB {
  bf16x8 e[UB], n[UB];
};
template <int UB>
__device__ __forceinline__ void gram_load_fragments_bf16(const double* mats, const LineDesc& L,
                                                         int lane, GramFragmentsB<UB>& f) {
  const int LT = mfma_ell_tiles(L.n_ell);
  const bf16x8* p =
      reinterpret_cast<const bf16x8*>(mats + L.gfrag_off + gram_fragment_doubles(L.n_ell));
#pragma unroll
  for (int uo = 0; uo < UB; ++uo) {
    // unconditional loads (clamped to the last tile): the correction then
    // waits for these alone
    const int u = min(uo, LT - 1);
    const bf16x8 e = p[u * 64 + lane], n = p[(LT + u) * 64 + lane];
    const bf16x8 z = {};
    f.e[uo] = uo < LT ? e : z;
    f.n[uo] = uo < LT ? n : z;
  }
}
// Eight FP32 values to one BF16 operand.
__device__ __forceinline__ bf16x8 to_bf16x8(const float (&v)[8]) {
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = static_cast<__bf16>(v[j]);
  return r;
}

// Output rows [g0, g0 + 16) of the tile (relative to t0) by one wave for
// Gram bands (the general kernel, and tiles with more rows than the split
// routine below holds).  Per field: the correction from the pre-projection
// terms at the operand positions (lane: row g0 + lr, ell 4 kb + lg), then at
// each accumulator position (row g0 + 4 lg + v, ell 16 uo + lr) the identity
// part from the stencils there, the sum and the RK combination; every RK
// load or store instruction covers 16 consecutive ell of four rows.
template <int SCHEME, int STAGE, int UB, bool NT, bool IN, bool BF>
__device__ __forceinline__ void gram_output_group(const StageArgs& a, const TileCtx& c,
                                                  const float* Ef, int g0, int lane) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl;
  const int LT = mfma_ell_tiles(nl), K = mfma_ell_blocks(nl);
  const int lr = lane & 15, lg = lane >> 4;
  const int orows = tile.t1 - tile.t0;
  // operand row of this lane
  const int r_in = tile.t0 + min(g0 + lr, orows - 1);
  const cdouble* u_in = c.U + (r_in - tile.lo) * us;
  const float* Nf = Ef + gram_correction_floats(nl);
  const cdouble zero = cmake(0.0, 0.0);
  // accumulator position (uo, v): output row and ell, clamped for addresses
  auto out_row = [&](int v) { return tile.t0 + min(g0 + 4 * lg + v, orows - 1); };
  auto out_ok = [&](int uo, int v) { return g0 + 4 * lg + v < orows && 16 * uo + lr < nl; };
  auto out_index = [&](int F, int uo, int v) {
    return L.base + static_cast<long long>(out_row(v)) * us + F * nl + min(16 * uo + lr, nl - 1);
  };
  constexpr int NR = UB <= 2 ? UB : 1;  // ell tiles of RK inputs requested early
  auto field = [&](auto field_tag) {
    constexpr int F = decltype(field_tag)::value;
    const bool need_ys = F == 0 && c.inplace;
    RkIn in[NR][4];
    if constexpr (UB <= 2) {
#pragma unroll
      for (int uo = 0; uo < UB; ++uo)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (uo < LT) in[uo][v] = rk_load<STAGE>(a, out_index(F, uo, v), need_ys);
    }
    mfma_accf cr[UB], ci[UB];
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) cr[uo] = ci[uo] = mfma_accf{0.0f, 0.0f, 0.0f, 0.0f};
    // without dissipation the Psi pre-projection terms vanish
    if constexpr (BF) {
      if (F != 2 || L.dissipate) {
        // fragments per field (kept across all three they spill)
        GramFragmentsB<UB> fb;
        gram_load_fragments_bf16<UB>(a.mats, L, lane, fb);
        // the operands X at (r_in, ell 8 lg + j), one stencil evaluation at a
        // time, then E X; then Psi and N Psi (gram_split_correction's order)
        // two passes of four operands (a loop, not unrolled: eight stencil
        // evaluations inline spill)
        bf16x8 bxr = {}, bxi = {};
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
          float xr[4], xi[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int l = 8 * lg + 4 * h + j;
            const cdouble x =
                l < nl ? pre_projection<SCHEME, F, false, IN>(c, r_in, l, zero) : zero;
            xr[j] = static_cast<float>(x.re);
            xi[j] = static_cast<float>(x.im);
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if (h == 0) {
              bxr[j] = static_cast<__bf16>(xr[j]);
              bxi[j] = static_cast<__bf16>(xi[j]);
            } else {
              bxr[4 + j] = static_cast<__bf16>(xr[j]);
              bxi[4 + j] = static_cast<__bf16>(xi[j]);
            }
          }
        }
#pragma unroll
        for (int uo = 0; uo < UB; ++uo) {
          if (uo >= LT) break;
          cr[uo] = mfma_bf16(bxr, fb.e[uo], cr[uo]);
          ci[uo] = mfma_bf16(bxi, fb.e[uo], ci[uo]);
        }
        if constexpr (F == 0) {
          float sr[8], si[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int l = 8 * lg + j;
            const cdouble ps = u_in[2 * nl + (l < nl ? l : 0)];
            sr[j] = l < nl ? static_cast<float>(ps.re) : 0.0f;
            si[j] = l < nl ? static_cast<float>(ps.im) : 0.0f;
          }
          const bf16x8 bsr = to_bf16x8(sr), bsi = to_bf16x8(si);
#pragma unroll
          for (int uo = 0; uo < UB; ++uo) {
            if (uo >= LT) break;
            cr[uo] = mfma_bf16(bsr, fb.n[uo], cr[uo]);
            ci[uo] = mfma_bf16(bsi, fb.n[uo], ci[uo]);
          }
        }
      }
    } else if (F != 2 || L.dissipate) {
      for (int kb = 0; kb < K; ++kb) {
        const int l = 4 * kb + lg;
        cdouble x = zero, ps = zero;
        if (l < nl) {
          x = pre_projection<SCHEME, F, false, IN>(c, r_in, l, zero);
          if constexpr (F == 0) ps = u_in[2 * nl + l];
        }
        gram_correction_step<F, UB>(LT, K, kb, lane, Ef, Nf, static_cast<float>(x.re),
                                    static_cast<float>(x.im), static_cast<float>(ps.re),
                                    static_cast<float>(ps.im), cr, ci);
      }
    }
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      const int slot = UB <= 2 ? uo : 0;
      if constexpr (UB > 2) {
#pragma unroll
        for (int v = 0; v < 4; ++v) in[0][v] = rk_load<STAGE>(a, out_index(F, uo, v), need_ys);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (!out_ok(uo, v)) continue;
        const int r = out_row(v), l = 16 * uo + lr;
        const cdouble x = pre_projection<SCHEME, F, false, IN>(c, r, l, zero);
        const cdouble k = cadd(gram_identity<SCHEME, F, IN>(c, r, l, x),
                               cmake(static_cast<double>(cr[uo][v]),
                                     static_cast<double>(ci[uo][v])));
        const cdouble ys = need_ys ? in[slot][v].ys : c.U[(r - tile.lo) * us + F * nl + l];
        rk_store<STAGE, NT>(a, out_index(F, uo, v), in[slot][v], ys, k, c.dt);
      }
    }
  };
  field(std::integral_constant<int, 0>{});
  field(std::integral_constant<int, 1>{});
  field(std::integral_constant<int, 2>{});
}

// The same output for Gram bands without dissipation in free-damped mode
// (the kStageKerrGram variant), one (row, ell) point per thread like
// plain_output, for tiles of at most kGramPoints rows per thread:
//   1. per point: the pre-projection terms, the identity parts of P and Q
//      (kept in registers) and Psi's whole output (its RK inputs were
//      requested while psi_t was formed);
//   2. the FP32 operands X_P, X_Q written over the point's psi_t slot (a
//      row's psi_t slot holds exactly n_ell complex doubles = 2 n_ell complex
//      floats), once every thread has read the psi_t stencils;
//   3. per wave and 16 rows the corrections on FP32 MFMA, written over their
//      operands;
//   4. per point: identity part + correction and the RK combination.
// The stencils run once per point with every thread busy, where the
// accumulator-position routine above runs them at the operand and at the
// output positions.
inline constexpr int kGramPoints = 5;
template <int STAGE>
__device__ __forceinline__ void gram_split_prefetch(const StageArgs& a, const LineDesc& L,
                                                    const TileDesc& tile, int tid,
                                                    RkIn (&in)[kGramPoints]) {
  const int nl = L.n_ell;
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  // Unconditional loads (rows past the tile clamped to its last row, their
  // values unused): with every load issued on every path, the compiler's
  // waits can count exactly which loads a use needs, instead of waiting for
  // every load in flight.
#pragma unroll
  for (int j = 0; j < kGramPoints; ++j) {
    const int r = min(tile.t0 + row_l + j * step_l, tile.t1 - 1);
    in[j] = rk_load<STAGE>(a, L.base + static_cast<long long>(r) * 3 * nl + 2 * nl + lane_l);
  }
}

// The split output serves the Gram variant up to two ell tiles (its E and N
// fragments live in registers); wider bands take gram_output_group.
template <int VARIANT, int UB>
constexpr bool split_gram() {
  return VARIANT == kStageKerrGram && UB <= 2;
}

__device__ __forceinline__ bool gram_split_fits(const LineDesc& L, const TileDesc& tile) {
  return tile.t1 - tile.t0 <= kGramPoints * (kStageThreads / L.n_ell);
}

// This lane's E and N fragments of every (output ell tile, contraction
// block): requested once per tile, ahead of the correction, so it finds them
// in registers instead of waiting on a cache load per block.
template <int UB>
struct GramFragments {
  float e[UB][4 * UB], n[UB][4 * UB];
};
template <int UB>
__device__ __forceinline__ void gram_load_fragments(const float* Ef, int nl, int lane,
                                                    GramFragments<UB>& f) {
  const int LT = mfma_ell_tiles(nl), K = mfma_ell_blocks(nl);
  const float* Nf = Ef + gram_correction_floats(nl);
  // unconditional loads (clamped to the last fragment), as in
  // gram_split_prefetch, so the correction waits for these alone
#pragma unroll
  for (int uo = 0; uo < UB; ++uo)
#pragma unroll
    for (int kb = 0; kb < 4 * UB; ++kb) {
      const bool ok = uo < LT && kb < K;
      const int fi = min(uo * K + kb, LT * K - 1) * 64 + lane;
      const float e = Ef[fi], n = Nf[fi];
      f.e[uo][kb] = ok ? e : 0.0f;
      f.n[uo][kb] = ok ? n : 0.0f;
    }
}

// Corrections E X_P + N Psi and E X_Q of rows [g0, g0 + 16) on FP32 MFMA,
// written over their operands.  Every accumulator takes its terms in the
// order of gram_correction_step (per block: E X, then N Psi), so the result
// equals gram_output_group's bit for bit.  The LDS operands of block kb + 1
// are read while block kb's products run.
template <int UB>
__device__ __forceinline__ void gram_split_correction(const TileCtx& c, const GramFragments<UB>& f,
                                                      int g0, int lane) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl;
  const int LT = mfma_ell_tiles(nl), K = mfma_ell_blocks(nl);
  const int lr = lane & 15, lg = lane >> 4;
  const int orows = tile.t1 - tile.t0;
  const int r_in = tile.t0 + min(g0 + lr, orows - 1);
  cdouble* urow = const_cast<cdouble*>(c.U) + (r_in - tile.lo) * us;
  const float* x = reinterpret_cast<const float*>(urow);  // X_P | X_Q as complex floats
  mfma_accf pr[UB], pi[UB], qr[UB], qi[UB];
#pragma unroll
  for (int uo = 0; uo < UB; ++uo) pr[uo] = pi[uo] = qr[uo] = qi[uo] = mfma_accf{0.0f, 0.0f, 0.0f, 0.0f};
  // operands of block kb at (row r_in, ell 4 kb + lg): X_P, Psi, X_Q
  auto operands = [&](int kb, float (&o)[6]) {
    const int l = 4 * kb + lg;
    const bool ok = l < nl;
    const int lc = ok ? l : 0;
    const cdouble ps = urow[2 * nl + lc];
    o[0] = ok ? x[2 * lc] : 0.0f;
    o[1] = ok ? x[2 * lc + 1] : 0.0f;
    o[2] = ok ? static_cast<float>(ps.re) : 0.0f;
    o[3] = ok ? static_cast<float>(ps.im) : 0.0f;
    o[4] = ok ? x[2 * nl + 2 * lc] : 0.0f;
    o[5] = ok ? x[2 * nl + 2 * lc + 1] : 0.0f;
  };
  float cur[6], nxt[6];
  operands(0, cur);
#pragma unroll
  for (int kb = 0; kb < 4 * UB; ++kb) {
    if (kb >= K) break;
    if (kb + 1 < K) operands(kb + 1, nxt);
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      pr[uo] = mfma_f32(cur[0], f.e[uo][kb], pr[uo]);
      pi[uo] = mfma_f32(cur[1], f.e[uo][kb], pi[uo]);
      pr[uo] = mfma_f32(cur[2], f.n[uo][kb], pr[uo]);
      pi[uo] = mfma_f32(cur[3], f.n[uo][kb], pi[uo]);
    }
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      qr[uo] = mfma_f32(cur[4], f.e[uo][kb], qr[uo]);
      qi[uo] = mfma_f32(cur[5], f.e[uo][kb], qi[uo]);
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) cur[i] = nxt[i];
  }
  // element v: row g0 + 4 lg + v, ell 16 uo + lr.  This wave alone reads
  // and writes these rows, and its operand reads were issued first.
#pragma unroll
  for (int uo = 0; uo < UB; ++uo)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = g0 + 4 * lg + v, l = 16 * uo + lr;
      if (uo < LT && row < orows && l < nl) {
        float* o = reinterpret_cast<float*>(const_cast<cdouble*>(c.U) +
                                            (tile.t0 + row - tile.lo) * us);
        o[2 * l] = pr[uo][v];
        o[2 * l + 1] = pi[uo][v];
        o[2 * nl + 2 * l] = qr[uo][v];
        o[2 * nl + 2 * l + 1] = qi[uo][v];
      }
    }
}

// The same corrections from the BF16 fragments (gram_output_group's BF16
// form, term for term).
template <int UB>
__device__ __forceinline__ void gram_split_correction(const TileCtx& c,
                                                      const GramFragmentsB<UB>& f, int g0,
                                                      int lane) {
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl;
  const int LT = mfma_ell_tiles(nl);
  const int lr = lane & 15, lg = lane >> 4;
  const int orows = tile.t1 - tile.t0;
  const int r_in = tile.t0 + min(g0 + lr, orows - 1);
  cdouble* urow = const_cast<cdouble*>(c.U) + (r_in - tile.lo) * us;
  const float* x = reinterpret_cast<const float*>(urow);  // X_P | X_Q as complex floats
  // one operand at a time (all six gathered at once spill): X_P with E,
  // Psi with N, X_Q with E, each accumulator taking its terms in
  // gram_output_group's order
  auto gather = [&](int which) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int l = 8 * lg + j;
      const bool ok = l < nl;
      const int lc = ok ? l : 0;
      float t;
      if (which < 2) t = x[2 * lc + which];
      else if (which < 4) t = static_cast<float>(which == 2 ? urow[2 * nl + lc].re
                                                            : urow[2 * nl + lc].im);
      else t = x[2 * nl + 2 * lc + (which - 4)];
      v[j] = ok ? t : 0.0f;
    }
    const bf16x8 r = to_bf16x8(v);
    __builtin_amdgcn_sched_barrier(0);
    return r;
  };
  mfma_accf pr[UB], pi[UB], qr[UB], qi[UB];
#pragma unroll
  for (int uo = 0; uo < UB; ++uo) pr[uo] = pi[uo] = qr[uo] = qi[uo] = mfma_accf{0.0f, 0.0f, 0.0f, 0.0f};
  {
    const bf16x8 xr = gather(0), xi = gather(1);
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      pr[uo] = mfma_bf16(xr, f.e[uo], pr[uo]);
      pi[uo] = mfma_bf16(xi, f.e[uo], pi[uo]);
    }
  }
  {
    const bf16x8 sr = gather(2), si = gather(3);
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      pr[uo] = mfma_bf16(sr, f.n[uo], pr[uo]);
      pi[uo] = mfma_bf16(si, f.n[uo], pi[uo]);
    }
  }
  {
    const bf16x8 xr = gather(4), xi = gather(5);
#pragma unroll
    for (int uo = 0; uo < UB; ++uo) {
      if (uo >= LT) break;
      qr[uo] = mfma_bf16(xr, f.e[uo], qr[uo]);
      qi[uo] = mfma_bf16(xi, f.e[uo], qi[uo]);
    }
  }
  // element v: row g0 + 4 lg + v, ell 16 uo + lr (as the FP32 form)
#pragma unroll
  for (int uo = 0; uo < UB; ++uo)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = g0 + 4 * lg + v, l = 16 * uo + lr;
      if (uo < LT && row < orows && l < nl) {
        float* w = reinterpret_cast<float*>(const_cast<cdouble*>(c.U) +
                                            (tile.t0 + row - tile.lo) * us);
        w[2 * l] = pr[uo][v];
        w[2 * l + 1] = pi[uo][v];
        w[2 * nl + 2 * l] = qr[uo][v];
        w[2 * nl + 2 * l + 1] = qi[uo][v];
      }
    }
}

template <int SCHEME, int STAGE, int UB, bool IN, bool BF>
__device__ __forceinline__ void gram_split_output(const StageArgs& a, const TileCtx& c,
                                                  const float* Ef, int tid,
                                                  RkIn (&psi_in)[kGramPoints]) {
  constexpr int NP = kGramPoints;
  const LineDesc& L = *c.L;
  const TileDesc& tile = *c.tile;
  const int nl = L.n_ell, us = 3 * nl;
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  const bool on = row_l < step_l;
  const cdouble zero = cmake(0.0, 0.0);
  cdouble kP[NP], kQ[NP];
  float xf[NP][4];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r = tile.t0 + row_l + j * step_l;
    if (on && r < tile.t1) {
      const cdouble xp = pre_projection<SCHEME, 0, false, IN>(c, r, lane_l, zero);
      const cdouble xq = pre_projection<SCHEME, 1, false, IN>(c, r, lane_l, zero);
      const cdouble xs = pre_projection<SCHEME, 2, false, IN>(c, r, lane_l, zero);
      kP[j] = gram_identity<SCHEME, 0, IN>(c, r, lane_l, xp);
      kQ[j] = gram_identity<SCHEME, 1, IN>(c, r, lane_l, xq);
      const cdouble ks = cadd(gram_identity<SCHEME, 2, IN>(c, r, lane_l, xs), cmake(0.0, 0.0));
      xf[j][0] = static_cast<float>(xp.re);
      xf[j][1] = static_cast<float>(xp.im);
      xf[j][2] = static_cast<float>(xq.re);
      xf[j][3] = static_cast<float>(xq.im);
      const long long g = L.base + static_cast<long long>(r) * us + lane_l;
      rk_store<STAGE>(a, g + 2 * nl, psi_in[j], c.U[(r - tile.lo) * us + 2 * nl + lane_l], ks,
                      c.dt);
    }
  }
  // E and N fragments for the correction, in flight across the barriers
  std::conditional_t<BF, GramFragmentsB<UB>, GramFragments<UB>> frag;
  if constexpr (BF)
    gram_load_fragments_bf16<UB>(a.mats, L, tid & 63, frag);
  else
    gram_load_fragments<UB>(Ef, nl, tid & 63, frag);
  __syncthreads();  // every psi_t stencil has been read
  RkIn inP[NP], inQ[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r = tile.t0 + row_l + j * step_l;
    if (on && r < tile.t1) {
      float* o = reinterpret_cast<float*>(const_cast<cdouble*>(c.U) + (r - tile.lo) * us);
      o[2 * lane_l] = xf[j][0];
      o[2 * lane_l + 1] = xf[j][1];
      o[2 * nl + 2 * lane_l] = xf[j][2];
      o[2 * nl + 2 * lane_l + 1] = xf[j][3];
    }
    // unconditional loads (clamped rows), as in gram_split_prefetch
    const long long g = L.base + static_cast<long long>(min(r, tile.t1 - 1)) * us + lane_l;
    inP[j] = rk_load<STAGE>(a, g, c.inplace);
    inQ[j] = rk_load<STAGE>(a, g + nl);
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int orows = tile.t1 - tile.t0;
  for (int g0 = 16 * wave; g0 < orows; g0 += 16 * (kStageThreads / 64))
    gram_split_correction<UB>(c, frag, g0, lane);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r = tile.t0 + row_l + j * step_l;
    if (on && r < tile.t1) {
      const float* o = reinterpret_cast<const float*>(c.U + (r - tile.lo) * us);
      const cdouble kp = cadd(kP[j], cmake(static_cast<double>(o[2 * lane_l]),
                                           static_cast<double>(o[2 * lane_l + 1])));
      const cdouble kq = cadd(kQ[j], cmake(static_cast<double>(o[2 * nl + 2 * lane_l]),
                                           static_cast<double>(o[2 * nl + 2 * lane_l + 1])));
      const long long g = L.base + static_cast<long long>(r) * us + lane_l;
      rk_store<STAGE>(a, g, inP[j], c.inplace ? inP[j].ys : c.U[(r - tile.lo) * us + lane_l], kp,
                      c.dt);
      rk_store<STAGE>(a, g + nl, inQ[j], c.U[(r - tile.lo) * us + nl + lane_l], kq, c.dt);
    }
  }
}

// LDS pointers of one pass of the stage kernels.  A pass covers the rows of
// a TileDesc: [lo, hi) staged (U, RC; buffer row = row - lo), [mlo, mhi) whose
// psi_t (and D Psi when stage-constrained) it computes, [t0, t1) output rows.
struct StageLds {
  const SchemeTable* T;
  RowCoeffs* RC;   // coefficients of row j at RC[j - rc_lo]
  int rc_lo;
  cdouble* U;      // rows [lo, hi): P | Q | Psi
  cdouble* PT;     // psi_t of row j at PT[(j - mlo) * pts]
  int pts;
  cdouble* DP;     // D Psi on rows [mlo, mhi) (stage-constrained only)
  double* NODES;   // sin | cos of the nodes (Kerr only)
  bool inplace;    // psi_t overwrites P in U
};

// Per-row coefficients of rows [r0, r1) into RC (row j at RC[j - lo]).
__device__ __forceinline__ void row_coefficients(const LineDesc& L, RowCoeffs* RC, int lo, int r0,
                                                 int r1, int tid) {
  for (int j = r0 + tid; j < r1; j += kStageThreads) {
    const RadialCoefficients rc =
        radial_coefficients(L.mass, L.spin, L.length, -2, L.m, L.h * j);
    RC[j - lo] = {2.0 * rc.adv, rc.principal, rc.time_radial, rc.g_real, rc.g_imag_m,
                  rc.q.re,      rc.q.im,      rc.psi.re,      rc.psi.im, 0.0};
  }
}

// psi_t on rows [mlo, mhi): D Psi (stage-constrained), X = P + 2K Q_eff,
// then the Kerr coupling (or the diagonal a = 0 form).  Ends with a barrier.
// CONSTRAINED = false compiles the stage-constrained form out.
template <int SCHEME, int UB, bool CONSTRAINED = true>
__device__ __forceinline__ void psi_t_phase(const StageArgs& a, const LineDesc& L,
                                            const TileDesc& tile, const StageLds& s, int tid) {
  const int nl = L.n_ell, nt = L.ntheta, n = L.nr, us = 3 * nl;
  const int mrows = tile.mhi - tile.mlo;
  const int lane_l = tid % nl, row_l = tid / nl, step_l = kStageThreads / nl;
  const bool on_l = row_l < step_l;
  const int wave = tid >> 6, lane = tid & 63;
  const RowCoeffs* RC = s.RC;
  const cdouble* U = s.U;
  const bool constrained = CONSTRAINED && L.constrained;
  // theta tile 0's S and A fragments, in flight while X is formed
  const double* Sf = a.mats + L.frag_off;
  const double* Af = Sf + 64 * mfma_s_fragments(nt, nl);
  double s0[4 * UB], a0[4][UB];
  if (L.kerr) kerr_load_tile<UB>(Sf, Af, mfma_ell_blocks(nl), mfma_ell_tiles(nl), 0, lane, s0, a0);
  if (constrained) {
    if (on_l)
      for (int j = tile.mlo + row_l; j < tile.mhi; j += step_l)
        s.DP[(j - tile.mlo) * nl + lane_l] =
            apply_d<SCHEME>(*s.T, U + 2 * nl + lane_l, us, tile.lo, j, n, L.inv_h);
    __syncthreads();
  }
  if (on_l)
    for (int j = tile.mlo + row_l; j < tile.mhi; j += step_l) {
      const int jj = j - tile.mlo;
      const RowCoeffs& rc = RC[j - s.rc_lo];
      const cdouble* u = U + (j - tile.lo) * us;
      const cdouble Q = constrained ? s.DP[jj * nl + lane_l] : u[nl + lane_l];
      const double k2 = rc.adv2;
      const cdouble P = u[lane_l];
      if (!L.kerr) {
        // a = 0: C_T and G_m are theta independent, the coupling is diagonal.
        const cdouble g = cmake(rc.g_real, rc.g_imag_m);
        const cdouble num = csub(cadd(P, cscale(k2, Q)), cmul(g, u[2 * nl + lane_l]));
        s.PT[jj * s.pts + lane_l] = cdivr(num, rc.time_radial);
      } else {
        s.PT[jj * s.pts + lane_l] = cadd(P, cscale(k2, Q));
      }
    }
  __syncthreads();
  if (L.kerr) {
    const cdouble* psi = U + (tile.mlo - tile.lo) * us + 2 * nl;
    for (int g0 = 16 * wave; g0 < mrows; g0 += 16 * (kStageThreads / 64))
      kerr_psi_group<UB>(nl, nt, L.spin, RC + (tile.mlo - s.rc_lo), s.PT, s.pts, psi, us, g0,
                         mrows, Sf, Af, s.NODES, lane, s0, a0, g0 == 16 * wave);
    __syncthreads();
  }
}

// Point RHS and RK combination on the output rows [t0, t1).  IN: every
// output row is an interior row of the radial operator (no closure rows).
template <int SCHEME, int STAGE, int UB, int VARIANT, bool IN>
__device__ __forceinline__ void output_phase(const StageArgs& a, const LineDesc& L,
                                             const TileDesc& tile, const StageLds& s, int tid,
                                             RkIn (&pre)[kPlainDepth<UB>][3],
                                             RkIn (&psi_in)[kGramPoints]) {
  const int wave = tid >> 6, lane = tid & 63;
  const int orows = tile.t1 - tile.t0;
  TileCtx c{&L,     &tile, s.T,       s.RC,      s.rc_lo, s.U, s.PT, s.DP, s.pts, s.inplace,
            a.dt_override > 0.0 ? a.dt_override : L.dt, L.constrained ? 0.0 : L.gamma};
  if (!L.gram) {
    plain_output<SCHEME, STAGE, kPlainDepth<UB>, IN>(a, c, tid, pre);
    return;
  }
  const float* Ef = reinterpret_cast<const float*>(a.mats + L.gfrag_off);
  // The BF16 correction serves the Kerr-Gram variant's bands of at most two
  // ell tiles; the general kernel keeps the FP32 form for every line (its
  // routines with BF16 inlined spill).
  constexpr bool kBf = UB <= 2 && VARIANT == kStageKerrGram;
  const bool bf = kBf && L.gram_bf16;
  if constexpr (split_gram<VARIANT, UB>()) {
    if (gram_split_fits(L, tile)) {
      if (bf)
        gram_split_output<SCHEME, STAGE, UB, IN, true>(a, c, Ef, tid, psi_in);
      else
        gram_split_output<SCHEME, STAGE, UB, IN, false>(a, c, Ef, tid, psi_in);
      return;
    }
  }
  constexpr bool NT = VARIANT == kStageKerrGram;
  if constexpr (kBf) {
    if (bf) {
      for (int g0 = 16 * wave; g0 < orows; g0 += 16 * (kStageThreads / 64))
        gram_output_group<SCHEME, STAGE, UB, NT, IN, true>(a, c, Ef, g0, lane);
      return;
    }
  }
  for (int g0 = 16 * wave; g0 < orows; g0 += 16 * (kStageThreads / 64))
    gram_output_group<SCHEME, STAGE, UB, NT, IN, false>(a, c, Ef, g0, lane);
}

// One radial tile of one line; UB = ell tiles of 16 the MFMA routines hold
// at once (1, 2 or 4).  VARIANT fixes the line flags at compile time for the
// common case (kStageGeneral keeps them at run time): kStageKerr and
// kStageKerrGram serve Kerr lines in free-damped mode without dissipation,
// with G = I and with the Gram projection respectively, and compile every
// other path out.
template <int SCHEME, int STAGE, int UB, int VARIANT>
__device__ __forceinline__ void stage_tile(const StageArgs& a, const TileDesc tile,
                                           double* smem) {
  LineDesc L = a.lines[tile.line];
  if constexpr (VARIANT != kStageGeneral) {
    L.kerr = 1;
    L.dissipate = 0;
    L.constrained = 0;
    L.gram = VARIANT == kStageKerrGram ? 1 : 0;
  }
  const int tid = threadIdx.x;
  const int nl = L.n_ell;
  const int nt = L.ntheta;
  const int rows = tile.hi - tile.lo;
  const int mrows = tile.mhi - tile.mlo;
  const int us = 3 * nl;  // LDS row stride of the staged tile (cdouble)

  // LDS: SBP table | RC[mrows] | U[rows][P|Q|Psi] | PT[mrows] (dissipation
  // only) | DP[mrows] (stage-constrained only) | sin, cos of the nodes (Kerr
  // only).  Without dissipation psi_t replaces P in U (P's only other use,
  // the stage-1 input, is read from HBM).  The SBP coefficient table lives in
  // LDS: closure rows index it per lane, which from global memory costs a
  // full load latency per tap.
  StageLds s{};
  s.inplace = !L.dissipate;
  s.T = reinterpret_cast<const SchemeTable*>(smem);
  s.RC = reinterpret_cast<RowCoeffs*>(smem + kSchemeTableDoubles);
  s.rc_lo = tile.mlo;
  s.U = reinterpret_cast<cdouble*>(s.RC + mrows);
  s.PT = s.inplace ? s.U + (tile.mlo - tile.lo) * us : s.U + rows * us;
  s.pts = s.inplace ? us : nl;
  s.DP = s.inplace ? s.U + rows * us : s.PT + mrows * nl;
  s.NODES = reinterpret_cast<double*>(L.constrained ? s.DP + mrows * nl : s.DP);

  // 1. stage the tile rows [lo, hi) of all three fields, the tables and the
  //    per-row coefficients; the RK inputs of the first output rows (G = I
  //    bands) are requested once the staging loads are issued.
  // (Requested before the staging loads they delay the staging, which the
  // rest of the tile waits for: slower.)
  // The SBP table and the sines and cosines of the nodes, a few doubles per
  // thread, are requested ahead of the staging loads and written to LDS after
  // them: their latency hides behind the staging instead of following it, and
  // no wait for them also waits for the RK inputs requested below.
  constexpr int kTab = static_cast<int>(sizeof(SchemeTable) / 8);
  constexpr int kTabPer = (kTab + kStageThreads - 1) / kStageThreads;
  const double* tab = reinterpret_cast<const double*>(a.table);
  double tabv[kTabPer];
#pragma unroll
  for (int k = 0; k < kTabPer; ++k) tabv[k] = tab[min(tid + k * kStageThreads, kTab - 1)];
  const int nn = L.kerr ? 2 * nt : 0;
  double nodev = 0.0;
  if (tid < nn) nodev = a.nodes[L.node_off + tid];
  RkIn pre[kPlainDepth<UB>][3];
  RkIn psi_in[kGramPoints];
  stage_copy(s.U, a.yin + L.base + static_cast<long long>(tile.lo) * us, rows * us, tid);
  // (written before the RK inputs are requested: the compiler's wait for
  // these values would otherwise also cover those loads)
#pragma unroll
  for (int k = 0; k < kTabPer; ++k)
    if (tid + k * kStageThreads < kTab) smem[tid + k * kStageThreads] = tabv[k];
  if (tid < nn) s.NODES[tid] = nodev;
  if (!L.gram) plain_prefetch<STAGE>(a, L, tile, s.inplace, tid, pre);
  if constexpr (split_gram<VARIANT, UB>())
    if (gram_split_fits(L, tile)) gram_split_prefetch<STAGE>(a, L, tile, tid, psi_in);
  for (int j = tid + kStageThreads; j < nn; j += kStageThreads)
    s.NODES[j] = a.nodes[L.node_off + j];
  row_coefficients(L, s.RC, tile.mlo, tile.mlo, tile.mhi, tid);
  __syncthreads();

  // 2. psi_t on rows [mlo, mhi); 3-4. the point RHS and the RK combination.
  psi_t_phase<SCHEME, UB>(a, L, tile, s, tid);
  constexpr int W = SchemeShape<SCHEME>::W;
  if (tile.t0 >= W && tile.t1 + W <= L.nr)
    output_phase<SCHEME, STAGE, UB, VARIANT, true>(a, L, tile, s, tid, pre, psi_in);
  else
    output_phase<SCHEME, STAGE, UB, VARIANT, false>(a, L, tile, s, tid, pre, psi_in);
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD b mod
// 8), each XCD with its own L2.  Within every run of 8 * kXcdChunk
// workgroups, XCD x takes kXcdChunk consecutive tiles, so neighbouring tiles
// of a line, which stage each other's halo rows, mostly share an L2; runs
// stay short so every XCD still gets an even mix of the group's tiles (HBM
// reads of the ell-24 class: -10 %).
inline constexpr int kXcdChunk = 8;
__device__ __forceinline__ int xcd_tile_order(int b, int n) {
  constexpr int span = 8 * kXcdChunk;
  if constexpr (kXcdChunk <= 1) return b;
  if (b >= n / span * span) return b;  // the last partial run keeps its order
  const int w = b % span;
  return b - w + (w % 8) * kXcdChunk + w / 8;
}

// One workgroup per tile.  (A loop over several tiles per workgroup keeps
// more values live across tiles and measured slower at every count.)  The
// LDS budget allows two workgroups per CU, so registers are sized for two.
template <int SCHEME, int STAGE, int UB, int VARIANT>
__global__ __launch_bounds__(kStageThreads, 2) void teuk_stage_kernel(
    StageArgs a) {
  extern __shared__ __align__(16) double smem[];
  stage_tile<SCHEME, STAGE, UB, VARIANT>(a, a.tiles[xcd_tile_order(blockIdx.x, gridDim.x)],
                                         smem);
}

}  // namespace teuk
