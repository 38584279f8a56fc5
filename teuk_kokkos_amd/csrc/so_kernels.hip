// Second-order stage graph on gfx950: the reference's 13-field nodal state
// advanced with its reconstruction chain, stage tangent, quadratic
// ordered-pair source and second-order Teukolsky triple
// (spatial_pipeline.hpp:458-588).
//
// One RHS evaluation is nine launches:
//   so_linear_kernel<0>   first-order triple of the stage -> k[0..2], D Psi
//   so_drad_kernel        D of the stage reconstruction fields
//   so_recon_kernel<0>    reconstruction chain (stage, k) -> k[3..9], and the
//                         independent constraints when asked
//   so_linear_kernel<1>   first-order triple of k -> tangent, D Psi_t
//   so_drad_kernel        D of k's reconstruction fields
//   so_recon_kernel<1>    tangent chain (k, tangent) -> tangent[3..9], its
//                         angular inputs, eth1 B and eth'1 F with tangents
//   so_source_kernel      ordered-pair source with Jet1 tangents, one thread
//                         per (target, r, theta)
//   so_source_projection_kernel  band projection of D and T, eth'_3 T
//   so_linear_kernel<2>   second-order triple of the stage with the forcing
//                         (outer operator evaluated in the kernel) -> k[10..12]
// Angular operators are T x T nodal matrices built on the host (projection,
// Laplacian, eth and eth' angular parts), applied to rows staged in LDS.
#include <hip/hip_runtime.h>

#include "sbp_device.hpp"
#include "so_device.hpp"
#include "so_layout.hpp"

namespace teuk {

namespace {

// Jet1 (jet.hpp:9-136): value and coordinate-time derivative.
struct Jet {
  cdouble v, d;
};
__device__ __forceinline__ Jet jadd(Jet a, Jet b) { return {cadd(a.v, b.v), cadd(a.d, b.d)}; }
__device__ __forceinline__ Jet jneg(Jet a) { return {cneg(a.v), cneg(a.d)}; }
__device__ __forceinline__ Jet jmul(Jet a, Jet b) {
  return {cmul(a.v, b.v), cadd(cmul(a.d, b.v), cmul(a.v, b.d))};
}
__device__ __forceinline__ Jet jscale(Jet a, cdouble s) { return {cmul(a.v, s), cmul(a.d, s)}; }
__device__ __forceinline__ Jet jscaler(Jet a, double s) { return {cmulr(a.v, s), cmulr(a.d, s)}; }
__device__ __forceinline__ Jet jdiv(Jet a, cdouble s) { return {cdiv(a.v, s), cdiv(a.d, s)}; }
__device__ __forceinline__ Jet jconj(Jet a) { return {cconj(a.v), cconj(a.d)}; }

// corrected_ordered_pair_source (second_order.hpp:74-142) on jets.
// f: F1 G1 Lambda1 Pi1 B1 C1 U1 F2 G2 H2 B2 C2 U2 U2# C2# C1# B1# Pi2# B2#
// d: d1F2 d3U2 eth2C2 eth'2C2# eth1B2 d2C2 d2G2 eth2G2 eth'1F2 d2C2# eth'1B2#
// f(i) and d(i) return the field and derivative jets on demand, so each is
// loaded next to the terms that use it rather than all at once.
template <class FJ, class DJ>
__device__ __forceinline__ void pair_source(double R, const Bg& bg, FJ f, DJ d, Jet& D, Jet& T) {
  enum { F1, G1, La1, Pi1, B1, C1, U1, F2, G2, H2, B2, C2, U2, U2s, C2s, C1s, B1s, Pi2s, B2s };
  enum { d1F2, d3U2, e2C2, ep2C2s, e1B2, d2C2, d2G2, e2G2, ep1F2, d2C2s, ep1B2s };
  const cdouble mu0b = cconj(bg.mu0), pi0b = cconj(bg.pi0), tau0b = cconj(bg.tau0);
  const double R2 = R * R;
  Jet acc;
  // sd01-sd02
  acc = jmul(jscaler(f(U1), 0.5), jadd(jdiv(d(d1F2), bg.mu0), jscaler(f(F2), R)));
  // sd03
  acc = jadd(acc, jmul(f(F1), jscaler(d(e2C2), 0.5 * R)));
  // sd04
  acc = jadd(acc, jmul(f(F1), jscale(f(C2), cscale(0.5 * R2, cadd(pi0b, cscale(2.0, bg.tau0))))));
  // sd05
  acc = jadd(acc, jmul(f(F1), jadd(jdiv(d(d3U2), bg.mu0), jscaler(f(U2s), R))));
  // sd06
  acc = jadd(acc, jmul(f(F1), jscaler(d(ep2C2s), -0.5 * R)));
  // sd07
  acc = jadd(acc, jmul(f(F1), jscale(f(C2s), cscale(0.5 * R2, cadd(cscale(5.0, bg.pi0),
                                                                  cscale(4.0, tau0b))))));
  // sd08 - sd11: -0.5 G1 * (...)
  const Jet g1 = jscaler(f(G1), -0.5);
  acc = jadd(acc, jmul(g1, jscaler(d(e1B2), R)));
  acc = jadd(acc, jmul(g1, jscale(f(B2), cscale(R2, cadd(pi0b, bg.tau0)))));
  acc = jadd(acc, jmul(g1, jscaler(d(d2C2), R)));
  acc = jadd(acc, jmul(g1, jscale(f(C2), cscale(R2, cadd(cscale(-2.0, bg.mu0), mu0b)))));
  // sd12 - sd15
  acc = jadd(acc, jmul(jscaler(f(C1), -R), d(d2G2)));
  acc = jadd(acc, jmul(jscaler(f(B1), 0.5 * R), d(e2G2)));
  acc = jadd(acc, jmul(jscaler(f(Pi1), 4.0 * R), f(G2)));
  acc = jadd(acc, jmul(jscaler(f(La1), -3.0 * R), f(H2)));
  D = acc;
  // st01-st02
  Jet t = jmul(jneg(f(C1s)),
               jadd(d(d1F2), jscale(f(F2), cscale(R, cadd(bg.mu0, cscale(2.0, mu0b))))));
  t = jadd(t, jmul(jscaler(f(B1s), 0.5), d(ep1F2)));
  t = jadd(t, jmul(f(F1), f(Pi2s)));
  t = jadd(t, jmul(jneg(f(F1)), d(d2C2s)));
  t = jadd(t, jmul(f(F1), d(ep1B2s)));
  t = jadd(t, jmul(f(F1), jscale(f(B2s), cscale(-0.5 * R, cadd(bg.pi0, tau0b)))));
  T = t;
}

// A value every lane of the wave holds, moved to a scalar register.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uniform64(long long v) {
  const unsigned long long u = static_cast<unsigned long long>(v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u)));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
  return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

__device__ __forceinline__ const double* node_ptr(const SoArgs& a, const SoInstance& I) {
  return a.nodes + I.node_off;
}

// element (mode, field, r, t) of a 13-field buffer
__device__ __forceinline__ long long fidx(const SoInstance& I, int mode, int f, int r, int t) {
  return I.base + ((static_cast<long long>(mode) * 13 + f) * I.N + r) * I.T + t;
}
// element (mode, slot, r, t) of a scratch array with S slots
__device__ __forceinline__ long long sidx(const SoInstance& I, int S, int mode, int s, int r,
                                          int t) {
  return I.pbase * S + ((static_cast<long long>(mode) * S + s) * I.N + r) * I.T + t;
}

// out[t] = sum_t' Op[t][t'] in[t'] (in: one theta row), zero where inactive.
// The operator column entries are read in chunks of CH independent loads, so
// a chunk costs one memory latency rather than one per coefficient.
template <int CH = 16>
__device__ __forceinline__ cdouble row_op(const double* ops, int off, const cdouble* in, int t,
                                          int T) {
  if (off < 0) return cmake(0.0, 0.0);
  const double* o = ops + off + t;  // column-major: entry (t, u) at u * T + t
  double re = 0.0, im = 0.0;
  for (int k0 = 0; k0 < T; k0 += CH) {
    double c[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) c[j] = k0 + j < T ? o[(k0 + j) * T] : 0.0;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (k0 + j < T) {
        re += c[j] * in[k0 + j].re;
        im += c[j] * in[k0 + j].im;
      }
    }
  }
  return cmake(re, im);
}

// Background coefficients at (r, t) from the per-instance tables: `theta`
// selects the theta-coordinate evaluation (cos, sin of acos x) used by the
// reconstruction passes, otherwise the storage evaluation (x, sqrt(1-x^2)).
__device__ __forceinline__ Bg bg_at(const SoArgs& a, const SoInstance& I, bool theta, int r,
                                    int t) {
  const long long i = I.bgoff * 2 + (theta ? static_cast<long long>(I.N) * I.T : 0) +
                      static_cast<long long>(r) * I.T + t;
  return a.bg[i];
}

}  // namespace

typedef double mfma4 __attribute__((ext_vector_type(4)));

// Out[m][t] = sum_u A[m][u] Op[t][u] for one 16-row block on FP64 MFMA
// (v_mfma_f64_16x16x4_f64): the rows are the real and imaginary parts of up
// to eight complex theta rows (row m: complex row m >> 1, part m & 1; the
// operators are real, so each part is contracted on its own).  Operand
// lanes: A row lr, u = 4 kb + lg; B column t = 16 tile + lr, u = 4 kb + lg
// (Op column-major: entry (t, u) at u T + t); accumulator element w: row
// lg + 4 w, column t = 16 tile + lr.
template <class AF>
__device__ __forceinline__ mfma4 row_block_op(const double* op, int T, int tile, int lane,
                                              AF a_of) {
  const int lr = lane & 15, lg = lane >> 4;
  const int tcol = 16 * tile + lr;
  mfma4 acc = {0.0, 0.0, 0.0, 0.0};
  // operands of four k blocks (16 u) requested before the products
  for (int k0 = 0; k0 < T; k0 += 16) {
    double av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int u = k0 + 4 * q + lg;
      bv[q] = (u < T && tcol < T) ? op[u * T + tcol] : 0.0;
      av[q] = u < T ? a_of(lr, u) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (k0 + 4 * q < T) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc, 0, 0, 0);
  }
  return acc;
}

// Band projection of one 16-row block as analysis then synthesis
// (angular_coordinator.hpp:125-146), both on FP64 MFMA:
//   C^T[ell][m] = sum_u A[ell][u] X[m][u]     (ell tiles of 16, up to four)
//   Out[m][t]   = sum_ell C[m][ell] S[t][ell]
// Accumulator element w of the first product holds C^T[16 lt + lg + 4 w][lr],
// which is the A operand of ell block (lt, w) of the second, so the
// coefficients never leave registers.  `an` and `sn` are the kOpAnalyze and
// kOpSynthesize operators (column-major like every operator).  Rows m and the
// output layout are those of row_block_op.
template <class AF>
__device__ __forceinline__ mfma4 row_block_project(const double* an, const double* sn, int T,
                                                   int n_ell, int tile, int lane, AF a_of) {
  const int lr = lane & 15, lg = lane >> 4;
  const int LT = (n_ell + 15) / 16;
  mfma4 c[4];
#pragma unroll
  for (int lt = 0; lt < 4; ++lt) c[lt] = mfma4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < T; k0 += 16) {
    double bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int u = k0 + 4 * q + lg;
      bv[q] = u < T ? a_of(lr, u) : 0.0;
    }
#pragma unroll
    for (int lt = 0; lt < 4; ++lt) {
      if (lt >= LT) break;
      const int ell = 16 * lt + lr;
      double av[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int u = k0 + 4 * q + lg;
        av[q] = (u < T && ell < n_ell) ? an[u * T + ell] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (k0 + 4 * q < T) c[lt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], c[lt], 0, 0, 0);
    }
  }
  const int tcol = 16 * tile + lr;
  mfma4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int lt = 0; lt < 4; ++lt) {
    if (lt >= LT) break;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (16 * lt + 4 * w >= n_ell) break;
      const int ell = 16 * lt + 4 * w + lg;
      const double b = (ell < n_ell && tcol < T) ? sn[ell * T + tcol] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(c[lt][w], b, acc, 0, 0, 0);
    }
  }
  return acc;
}

// The same projection of one theta row at node t by one thread (rows wider
// than an MFMA tile): the n_ell coefficients, then their synthesis at t.
__device__ __forceinline__ cdouble row_project(const double* ops, int an_off, int sn_off,
                                               int n_ell, const cdouble* in, int t, int T) {
  if (an_off < 0) return cmake(0.0, 0.0);
  const double* an = ops + an_off;
  const double* sn = ops + sn_off;
  double re = 0.0, im = 0.0;
  for (int ell = 0; ell < n_ell; ++ell) {
    double cr = 0.0, ci = 0.0;
    for (int u = 0; u < T; ++u) {
      const double w = an[u * T + ell];
      cr += w * in[u].re;
      ci += w * in[u].im;
    }
    const double s = sn[ell * T + t];
    re += s * cr;
    im += s * ci;
  }
  return cmake(re, im);
}

// ------------------------------------------------------------ linear triple
// KIND 0: stage fields 0..2 -> out 0..2, D Psi -> dpsi      (first order)
// KIND 1: out fields 0..2 -> tan 0..2, D Psi -> dpsi_t      (its tangent)
// KIND 2: stage fields 10..12 -> out 10..12 with the forcing (second order)
// full_spatial.hpp:99-398 with the band projection of spatial_pipeline.hpp:
// 756-766 and, for KIND 2, the outer operator and coordinate forcing of
// source_spatial.hpp:357-425, second_order.hpp:152-182.
template <int SCHEME, int KIND>
// Four waves per SIMD (at most 128 registers): measured 5 % faster per
// second-order step than the three the unconstrained allocation allows.
__global__ __launch_bounds__(kSoThreads, 4) void so_linear_kernel(SoArgs a) {
  extern __shared__ __align__(16) double smem[];
  const SoTile tile = a.tiles[blockIdx.x];
  const SoInstance& I = a.inst[tile.inst];
  const int N = I.N, T = I.T, mode = tile.mode, tid = threadIdx.x;
  // a triple that does not evolve on the mode has a zero RHS (its band
  // projection is the zero operator): the zero-initialized outputs stand
  if (!((I.field_mask[mode] >> (KIND == 2 ? 10 : 0)) & 1)) return;
  const int rows = tile.hi - tile.lo, mrows = tile.mhi - tile.mlo, orows = tile.t1 - tile.t0;
  const int us = 3 * T;
  const int coord = KIND == 2 ? kCoSecond : kCoFirst;
  SchemeTable* TS = reinterpret_cast<SchemeTable*>(smem);
  cdouble* U = reinterpret_cast<cdouble*>(smem + kSchemeTableDoubles);
  cdouble* QE = U + rows * us;
  cdouble* PT = QE + mrows * T;
  // the unprojected outputs reuse the staged rows once every thread is done
  // reading them (orows <= rows)
  cdouble* OUT = U;
  const double* nodes = node_ptr(a, I);
  const double* cos_th = nodes + 2 * T;
  const double* sin_th = nodes + 3 * T;

  for (int j = tid; j < static_cast<int>(sizeof(SchemeTable) / 8); j += kSoThreads)
    smem[j] = reinterpret_cast<const double*>(a.tables + I.scheme)[j];
  auto in_at = [&](int f, int r, int t) -> cdouble {
    if constexpr (KIND == 0) return a.in[fidx(I, mode, f, r, t)];
    if constexpr (KIND == 1) return a.out[fidx(I, mode, f, r, t)];
    return a.in[fidx(I, mode, 10 + f, r, t)];
  };
  for (int e = tid; e < rows * us; e += kSoThreads) {
    const int row = e / us, rem = e - row * us, f = rem / T, t = rem - f * T;
    U[e] = in_at(f, tile.lo + row, t);
  }
  __syncthreads();
  const SchemeTable& TB = *TS;
  const bool constrained = I.constrained != 0;
  const int m = I.modes[mode];
  for (int e = tid; e < mrows * T; e += kSoThreads) {
    const int jj = e / T, t = e - jj * T, j = tile.mlo + jj;
    const cdouble* u = U + (j - tile.lo) * us;
    const cdouble qe = constrained ? apply_d<SCHEME>(TB, U + 2 * T + t, us, tile.lo, j, N, I.inv_h)
                                   : u[T + t];
    QE[e] = qe;
    const RadialCoefficients rc = radial_coefficients(I.mass, I.spin, I.length, -2, m, I.h * j);
    const double ct = time_coefficient(rc, I.spin, sin_th[t]);
    const cdouble g = definition_coefficient(rc, I.spin, -2, cos_th[t]);
    const cdouble num = csub(cadd(u[t], cscale(2.0 * rc.adv, qe)), cmul(g, u[2 * T + t]));
    PT[e] = cdivr(num, ct);
  }
  __syncthreads();
  const int lap_off = I.op[coord][mode][kOpLaplacian];
  const double damping = constrained ? 0.0 : I.gamma;
  const bool forced = KIND == 2 && I.op[kCoSecond][mode][kOpProject] >= 0;
  const double a_spin = I.spin, L = I.length, L2 = L * L, L4 = L2 * L2;
  const double* cos_x = nodes;
  // each thread keeps its (at most kSoLinearOut) outputs in registers
  cdouble out_p[kSoLinearOut], out_q[kSoLinearOut], out_s[kSoLinearOut];
#pragma unroll
  for (int j = 0; j < kSoLinearOut; ++j) {
    const int e = tid + j * kSoThreads;
    if (e >= orows * T) break;
    const int rr = e / T, t = e - rr * T, r = tile.t0 + rr;
    const cdouble* u = U + (r - tile.lo) * us;
    const cdouble dpsi = apply_d<SCHEME>(TB, U + 2 * T + t, us, tile.lo, r, N, I.inv_h);
    const cdouble dq = apply_d<SCHEME>(TB, QE + t, T, tile.mlo, r, N, I.inv_h);
    const cdouble dpt = apply_d<SCHEME>(TB, PT + t, T, tile.mlo, r, N, I.inv_h);
    const cdouble qe = QE[(r - tile.mlo) * T + t];
    const cdouble lap = row_op(a.ops, lap_off, u + 2 * T, t, T);
    const RadialCoefficients rc = radial_coefficients(I.mass, I.spin, I.length, -2, m, I.h * r);
    cdouble pt = cadd(cadd(cadd(cscale(rc.principal, dq), cmul(rc.q, qe)), cmul(rc.psi, u[2 * T + t])),
                      lap);
    if constexpr (KIND == 2) {
      cdouble frc = cmake(0.0, 0.0);
      if (forced && a.active[tile.inst]) {
        const double R = I.h * r;
        const cdouble* dcol = a.src + sidx(I, kSlotsSrc, mode, 0, 0, t);
        const cdouble Dv = dcol[static_cast<long long>(r) * T];
        const cdouble Dt = a.src[sidx(I, kSlotsSrc, mode, 1, r, t)];
        const cdouble Tv = a.src[sidx(I, kSlotsSrc, mode, 2, r, t)];
        const cdouble ethT = a.src[sidx(I, kSlotsSrc, mode, 3, r, t)];
        const cdouble Dr = apply_d<SCHEME>(TB, dcol, T, 0, r, N, I.inv_h);
        const cdouble d3D = delta_n(Dv, Dt, Dr, 3, R, I.mass, L);
        const Bg bg = bg_at(a, I, false, r, t);
        cdouble s = d3D;
        s = cadd(s, cmul(cscale(R, cadd(cscale(4.0, bg.mu0), cconj(bg.mu0))), Dv));
        s = cadd(s, cscale(R, ethT));
        s = cadd(s, cmul(cscale(R * R, csub(cscale(4.0, bg.pi0), cconj(bg.tau0))), Tv));
        const double norm = 2.0 * (L4 + a_spin * a_spin * R * R * cos_x[t] * cos_x[t]);
        frc = cscale(norm, s);
      }
      pt = cadd(pt, frc);
    }
    cdouble qt = csub(dpt, cscale(damping, csub(u[T + t], dpsi)));
    cdouble st = PT[(r - tile.mlo) * T + t];
    if (I.dissipate) {
      pt = cadd(pt, apply_diss<SCHEME>(TB, U + t, us, tile.lo, r, N, I.h, I.eps));
      qt = cadd(qt, apply_diss<SCHEME>(TB, U + T + t, us, tile.lo, r, N, I.h, I.eps));
      st = cadd(st, apply_diss<SCHEME>(TB, U + 2 * T + t, us, tile.lo, r, N, I.h, I.eps));
    }
    out_p[j] = pt;
    out_q[j] = qt;
    out_s[j] = st;
    if constexpr (KIND == 0) a.dpsi[sidx(I, 1, mode, 0, r, t)] = dpsi;
    if constexpr (KIND == 1) a.dpsi_t[sidx(I, 1, mode, 0, r, t)] = dpsi;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSoLinearOut; ++j) {
    const int e = tid + j * kSoThreads;
    if (e >= orows * T) break;
    const int rr = e / T, t = e - rr * T;
    cdouble* o = OUT + rr * us;
    o[t] = out_p[j];
    o[T + t] = out_q[j];
    o[2 * T + t] = out_s[j];
  }
  __syncthreads();
  // band projection of the outputs on FP64 MFMA: the rows (output row,
  // field) are contiguous theta rows of OUT; a block takes eight of them,
  // real and imaginary parts, and the blocks go round the waves
  const int an_off = I.op[coord][mode][kOpAnalyze], sn_off = I.op[coord][mode][kOpSynthesize];
  const int n_band = I.band_n[coord][mode];
  const int lane = tid & 63, lg = lane >> 4, lr = lane & 15;
  const int crows = orows * 3, ttiles = (T + 15) / 16;
  for (int blk = tid >> 6; blk < ((crows + 7) / 8) * ttiles; blk += kSoThreads / 64) {
    const int b8 = 8 * (blk / ttiles), tile_t = blk % ttiles;
    auto a_of = [&](int mm, int u) {
      const int cr = b8 + (mm >> 1);
      if (cr >= crows) return 0.0;
      const cdouble v = OUT[cr * T + u];
      return (mm & 1) ? v.im : v.re;
    };
    const mfma4 acc = an_off < 0 ? mfma4{0.0, 0.0, 0.0, 0.0}
                                 : row_block_project(a.ops + an_off, a.ops + sn_off, T, n_band,
                                                     tile_t, lane, a_of);
    const int t = 16 * tile_t + lr;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int mm = lg + 4 * w, cr = b8 + (mm >> 1);
      if (cr >= crows || t >= T) continue;
      const int rr = cr / 3, f = cr - 3 * rr, r = tile.t0 + rr;
      cdouble* dst;
      if constexpr (KIND == 0) dst = a.out + fidx(I, mode, f, r, t);
      if constexpr (KIND == 1) dst = a.tan + sidx(I, kSlotsTan, mode, f, r, t);
      if constexpr (KIND == 2) dst = a.out + fidx(I, mode, 10 + f, r, t);
      reinterpret_cast<double*>(dst)[mm & 1] = acc[w];
    }
  }
}

// ------------------------------------------------------------ radial D
// D of the seven reconstruction fields of the stage (which 0) or of the RHS
// (which 1) (full_spatial.hpp:400-455).  One workgroup per (mode, field,
// kDradRows radial rows, kDradCols theta columns) of an instance: the rows
// the stencils reach are staged in LDS once, so HBM sees each value about
// once instead of once per stencil tap.  grid.y = instance.
template <int SCHEME>
__global__ __launch_bounds__(kSoThreads) void so_drad_kernel(SoArgs a, int which) {
  __shared__ __align__(16) cdouble tile[kDradSpan * kDradCols];
  const SoInstance& I = a.inst[a.first_inst + blockIdx.y];
  if (I.scheme != SCHEME) return;
  constexpr int W = SchemeShape<SCHEME>::W, C = SchemeShape<SCHEME>::C;
  constexpr int B = SchemeShape<SCHEME>::B;
  const int N = I.N, T = I.T;
  const int n_rt = (N + kDradRows - 1) / kDradRows, n_tc = (T + kDradCols - 1) / kDradCols;
  int q = blockIdx.x;
  if (q >= I.M * 7 * n_rt * n_tc) return;
  const int tc = q % n_tc;
  q /= n_tc;
  const int rt = q % n_rt;
  q /= n_rt;
  const int c = q % 7, mode = q / 7;
  // reconstruction fields of a non-parent mode are zero for good (their D too)
  if (!((I.field_mask[mode] >> (3 + c)) & 1)) return;
  const int r0 = rt * kDradRows, r1 = min(N, r0 + kDradRows);
  const int t0 = tc * kDradCols, nt = min(kDradCols, T - t0);
  // rows [lo, hi) cover every stencil of rows [r0, r1): interior taps reach
  // B rows, the closures the first or last C rows
  const bool left = r0 < W, right = r1 - 1 + W >= N;
  int lo = left ? 0 : r0 - B;
  int hi = right ? N : r1 + B;
  if (right) lo = min(lo, N - C);
  if (left) hi = max(hi, C);
  lo = max(lo, 0);
  hi = min(hi, N);
  const cdouble* src = (which == 0 ? a.in : a.out) + fidx(I, mode, 3 + c, 0, 0);
  for (int e = threadIdx.x; e < (hi - lo) * nt; e += kSoThreads) {
    const int rr = e / nt, tt = e - rr * nt;
    tile[rr * kDradCols + tt] = src[static_cast<long long>(lo + rr) * T + t0 + tt];
  }
  __syncthreads();
  cdouble* dst = (which == 0 ? a.rdr : a.rdr_t) + sidx(I, kSlotsRdr, mode, c, 0, 0);
  const SchemeTable& TB = a.tables[SCHEME];
  for (int e = threadIdx.x; e < (r1 - r0) * nt; e += kSoThreads) {
    const int rr = e / nt, tt = e - rr * nt, r = r0 + rr;
    dst[static_cast<long long>(r) * T + t0 + tt] =
        apply_d<SCHEME>(TB, tile + tt, kDradCols, lo, r, N, I.inv_h);
  }
}

// ------------------------------------------------------------ reconstruction
// KIND 0: chain of (stage, k) -> k[3..9], the stage angular inputs of the
//         source and its eth1 B, eth'1 F (with k as their time derivatives),
//         plus the independent constraints when asked.
// KIND 1: chain of (k, tangent) -> tangent[3..9], the tangent angular inputs
//         and eth1 B, eth'1 F of k (with the tangent as time derivatives).
// Reconstruction passes: full_spatial.hpp:486-693, reconstruction.hpp:48-105;
// chain order: spatial_pipeline.hpp:664-754; constraints: 768-893,
// reconstruction_constraints.hpp:24-56; source inputs: 934-1181,
// source_tangent_spatial.hpp:65-224; projection: 1183-1209.
//
// One thread per (parent mode, theta) element of a radial row, several
// radial rows per workgroup, so every value a thread only needs at its own
// element stays in registers across the phases.  LDS holds just the rows
// other threads read: the theta rows the angular operators contract over
// and the sharp partners conj(X[-m]).  Rows: F, G, B, Pi, C of the input,
// the time derivatives of B and C, four rows of reconstructed time
// derivatives awaiting their band projection, and one for an angular part:
// 12 * P * N_theta * 16 bytes per radial row (P parent modes).  The angular
// operators run on FP64 MFMA over the workgroup's rows (apply_rows) and
// leave their results in these rows.
template <int KIND, int NT>
// The 256-thread form at three waves per SIMD (at most 168 registers, a few
// spilled): measured 5 % faster per second-order step than at two (at four,
// with about 100 spilled, 7 % slower).
__global__ __launch_bounds__(NT, NT == 256 ? 3 : 1) void so_recon_kernel(SoArgs a) {
  constexpr int CH = 8;  // operator loads in flight per row_op chunk
  extern __shared__ __align__(16) double smem[];
  const int inst = a.first_inst + blockIdx.y;
  const SoInstance& I = a.inst[inst];
  // Only the parent modes carry reconstruction fields (the others are zero
  // for good), so a radial row has MT = n_parents * N_theta elements:
  // element e is node t of parent pi, its LDS rows are indexed by e and the
  // sharp partner's by its parent index.
  const int T = I.T, MT = I.n_parents * T, tid = threadIdx.x;
  // radial rows of this workgroup: thread tid works on element e of row
  // r0 + sub; rows past the grid and threads past the rows only take part
  // in the barriers
  const int rpw = NT == kReconMaxThreads ? 1 : NT / MT;
  const int r0 = blockIdx.x * rpw;
  if (r0 >= I.N) return;
  const int sub = tid / MT;
  const int r = r0 + (sub < rpw ? sub : 0);
  const bool live = sub < rpw && r < I.N;
  const int e = live ? tid - sub * MT : 0;
  const int pi = e / T, t = e - pi * T;
  const int mode = I.parent_mode[pi];
  const int spi = I.parent_sharp[pi];  // -1: the partner's fields are zero
  const int sm = I.sharp[mode], se = (spi < 0 ? 0 : spi) * T + t;
  const double R = I.h * r, R2 = R * R, R3 = R2 * R;
  const double mass = I.mass, aspin = I.spin, L = I.length;
  const double* nodes = node_ptr(a, I);
  const double sx = nodes[T + t], cx = nodes[t];
  cdouble* LX = reinterpret_cast<cdouble*>(smem) + (live ? sub : 0) * kReconLdsRows * MT;
  enum { lF = 0, lG, lB, lPi, lC, ltB, ltC, lRaw, lAng = lRaw + 4 };
  auto row = [&](int k) { return LX + k * MT; };
  auto op = [&](int coord, int kind) { return I.op[coord][mode][kind]; };
  const cdouble czero = cmake(0.0, 0.0);
  // the sharp partner's value of LDS row k
  auto sharp_of = [&](int k) { return spi < 0 ? czero : cconj(row(k)[se]); };
  auto eth = [&](int coord, bool prime, cdouble xo, cdouble dto, cdouble ang) {
    return eth_point(prime, xo, dto, ang, I.coord_spin[coord], I.coord_boost[coord], R, sx, cx,
                     aspin, L);
  };
  // eth / eth' of coordinator `coord` at this element: theta rows `xr` of the
  // field (every parent mode), its own value `xo` and time derivative `dto`
  auto ghp = [&](int coord, bool prime, const cdouble* xr, cdouble xo, cdouble dto) {
    const int off = op(coord, prime ? kOpLower : kOpRaise);
    if (off < 0) return cmake(0.0, 0.0);
    return eth(coord, prime, xo, dto, row_op<CH>(a.ops, off, xr + pi * T, t, T));
  };
  // eth' of the sharp field conj(X[-m]) from the rows of X and its time
  // derivative: the operators are real, so the angular part is the
  // conjugate of the operator applied to the partner's row (bit for bit)
  auto ghp_sharp = [&](int coord, const cdouble* xr, const cdouble* dr_row) {
    const int off = op(coord, kOpLower);
    if (off < 0) return cmake(0.0, 0.0);
    if (spi < 0) return czero;  // the partner's rows are zero
    const cdouble ang = cconj(row_op<CH>(a.ops, off, xr + spi * T, t, T));
    return eth(coord, true, cconj(xr[se]), cconj(dr_row[se]), ang);
  };
  auto project = [&](int coord, const cdouble* in) {
    return row_project(a.ops, op(coord, kOpAnalyze), op(coord, kOpSynthesize),
                       I.band_n[coord][mode], in + pi * T, t, T);
  };
  // Angular operators applied to whole LDS rows on FP64 MFMA when a theta
  // row fits one 16-wide tile: a block is (operator j, parent mode, eight
  // radial rows), its sixteen MFMA rows the real and imaginary parts.
  // Operator j (coordinator coords[j], kind kinds[j], of the block's mode)
  // reads row srow[j] of that mode, or of its sharp partner when sharp[j],
  // and writes row drow[j] of the mode; a block reads and writes only its
  // own rows (in place when srow == drow), so the waves take blocks
  // independently, and the caller syncs before and after.  Wider rows keep
  // the per-thread contractions (`ghp`, `ghp_sharp`, `project`).
  const bool mfma_rows = T <= 16;
  const int rows_here = min(rpw, I.N - r0);
  auto apply_rows = [&](int nops, const int* coords, const int* kinds, const int* srow,
                        const int* drow, const bool* sharp) {
    const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
    const int np = I.n_parents, nsb = (rows_here + 7) / 8;
    const int waves = static_cast<int>(blockDim.x) / 64;
    for (int blk = tid >> 6; blk < nops * np * nsb; blk += waves) {
      const int j = blk / (np * nsb), rem = blk - j * (np * nsb), p = rem / nsb;
      const int sb = rem - p * nsb;
      const int pm = I.parent_mode[p];
      const bool two_step = kinds[j] == kOpProject;
      const int off = I.op[coords[j]][pm][two_step ? kOpAnalyze : kinds[j]];
      const int ps = sharp[j] ? I.parent_sharp[p] : p;
      const cdouble* in = reinterpret_cast<const cdouble*>(smem) + srow[j] * MT + (ps < 0 ? 0 : ps) * T;
      cdouble* out = reinterpret_cast<cdouble*>(smem) + drow[j] * MT + p * T;
      auto a_of = [&](int m, int u) {
        const int s = 8 * sb + (m >> 1);
        if (s >= rows_here) return 0.0;
        const cdouble v = in[s * kReconLdsRows * MT + u];
        return (m & 1) ? v.im : v.re;
      };
      mfma4 acc = {0.0, 0.0, 0.0, 0.0};
      if (off >= 0 && ps >= 0)
        acc = two_step ? row_block_project(a.ops + off, a.ops + I.op[coords[j]][pm][kOpSynthesize],
                                           T, I.band_n[coords[j]][pm], 0, lane, a_of)
                       : row_block_op(a.ops + off, T, 0, lane, a_of);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int m = lg + 4 * w, s = 8 * sb + (m >> 1);
        if (s < rows_here && lr < T)
          reinterpret_cast<double*>(out + s * kReconLdsRows * MT + lr)[m & 1] = acc[w];
      }
    }
  };
  // eth / eth' of this element from an angular part computed by apply_rows
  auto ghp_from = [&](int coord, bool prime, cdouble xo, cdouble dto, cdouble ang) {
    if (op(coord, prime ? kOpLower : kOpRaise) < 0) return czero;
    return eth(coord, prime, xo, dto, ang);
  };
  auto ghp_sharp_from = [&](int coord, const cdouble* xr, const cdouble* dr_row, cdouble ang) {
    if (op(coord, kOpLower) < 0 || spi < 0) return czero;
    return eth(coord, true, cconj(xr[se]), cconj(dr_row[se]), cconj(ang));
  };

  const cdouble* src = KIND == 0 ? a.in : a.out;
  cdouble* dst = KIND == 0 ? a.out : a.tan;
  const cdouble* rdr = KIND == 0 ? a.rdr : a.rdr_t;
  // state field of chain slot k: F is field 2, G..U are fields 3..9
  auto load = [&](int k) { return src[fidx(I, mode, 2 + k, r, t)]; };
  auto store = [&](int k, cdouble v) {
    if (KIND == 0) dst[fidx(I, mode, 2 + k, r, t)] = v;
    else dst[sidx(I, kSlotsTan, mode, 2 + k, r, t)] = v;
  };
  auto dr = [&](int c) { return rdr[sidx(I, kSlotsRdr, mode, c, r, t)]; };
  // eth1 B and eth'1 F of `src` with the chain's time derivatives: slots
  // 0, 1 from the stage chain, 2, 3 from the tangent chain
  auto ex = [&](int k, cdouble v) { a.ex[sidx(I, kSlotsEx, mode, 2 * KIND + k, r, t)] = v; };

  cdouble xF{}, xG{}, xLa{}, xH{}, xB{}, xPi{}, xC{}, xU{}, tF{}, tG{}, tH{}, tPi{}, tB{}, tC{};
  cdouble e1F{}, e2G{}, e2C{}, e2Pi{};
  Bg bg{};
  // D of the seven fields, each requested one phase before it is needed
  cdouble d0{}, d1{}, d2{}, d3{}, d4{}, d5{}, d6{};
  // stage the row
  if (live) {
    xF = load(0);
    xG = load(1);
    xLa = load(2);
    xH = load(3);
    xB = load(4);
    xPi = load(5);
    xC = load(6);
    xU = load(7);
    tF = KIND == 0 ? a.out[fidx(I, mode, 2, r, t)] : a.tan[sidx(I, kSlotsTan, mode, 2, r, t)];
    bg = bg_at(a, I, true, r, t);
    d0 = dr(0);
    d1 = dr(1);
    row(lF)[e] = xF;
    row(lG)[e] = xG;
    row(lB)[e] = xB;
    row(lPi)[e] = xPi;
    row(lC)[e] = xC;
  }
  __syncthreads();
  if (mfma_rows) {
    const int co[2] = {kCoFirst, kCoFirst}, ki[2] = {kOpRaise, kOpLower};
    const int sr[2] = {lF, lF}, dr_[2] = {lRaw + 2, lRaw + 3};
    const bool sh[2] = {false, false};
    apply_rows(2, co, ki, sr, dr_, sh);
    __syncthreads();
  }
  // pass 1: G, Lambda
  if (live) {
    e1F = mfma_rows ? ghp_from(kCoFirst, false, xF, tF, row(lRaw + 2)[e])
                    : ghp(kCoFirst, false, row(lF), xF, tF);
    ex(1, mfma_rows ? ghp_from(kCoFirst, true, xF, tF, row(lRaw + 3)[e])
                    : ghp(kCoFirst, true, row(lF), xF, tF));
    const cdouble dg = csub(cadd(cmul(cscale(-4.0 * R, bg.mu0), xG), e1F),
                            cmul(cscale(R, bg.tau0), xF));
    const cdouble dl = csub(cmul(cscale(-R, cadd(bg.mu0, cconj(bg.mu0))), xLa), xF);
    row(lRaw)[e] = recon_dt(xG, d0, dg, 2, R, mass, L);
    row(lRaw + 1)[e] = recon_dt(xLa, d1, dl, 1, R, mass, L);
  }
  __syncthreads();
  if (mfma_rows) {
    const int co[3] = {kCoG, kCoFirst, kCoG}, ki[3] = {kOpProject, kOpProject, kOpRaise};
    const int sr[3] = {lRaw, lRaw + 1, lG}, dr_[3] = {lRaw, lRaw + 1, lRaw + 2};
    const bool sh[3] = {false, false, false};
    apply_rows(3, co, ki, sr, dr_, sh);
    __syncthreads();
  }
  if (live) {
    d2 = dr(2);
    d3 = dr(3);
    d4 = dr(4);
    d5 = dr(5);
    tG = mfma_rows ? row(lRaw)[e] : project(kCoG, row(lRaw));
    store(1, tG);
    store(2, mfma_rows ? row(lRaw + 1)[e] : project(kCoFirst, row(lRaw + 1)));
    e2G = mfma_rows ? ghp_from(kCoG, false, xG, tG, row(lRaw + 2)[e])
                    : ghp(kCoG, false, row(lG), xG, tG);
  }
  __syncthreads();
  // pass 2: H, B, Pi, C
  if (live) {
    const cdouble mub = cconj(bg.mu0), pib = cconj(bg.pi0);
    const cdouble dh = csub(cadd(cmul(cscale(-3.0 * R, bg.mu0), xH), e2G),
                            cmul(cscale(2.0 * R, bg.tau0), xG));
    const cdouble db = csub(cmul(cscale(R, csub(bg.mu0, mub)), xB), cscale(2.0, xLa));
    const cdouble pt = cadd(pib, bg.tau0);
    const cdouble dpi = cadd(csub(cneg(xG), cmul(cscale(R, pt), xLa)),
                             cmul(cmul(cscale(0.5 * R2, bg.mu0), pt), xB));
    const cdouble dc = csub(csub(cmul(cscale(-R, mub), xC), cscale(2.0, xPi)),
                            cmul(cscale(R, bg.tau0), xB));
    row(lRaw)[e] = recon_dt(xH, d2, dh, 3, R, mass, L);
    row(lRaw + 1)[e] = recon_dt(xB, d3, db, 1, R, mass, L);
    row(lRaw + 2)[e] = recon_dt(xPi, d4, dpi, 2, R, mass, L);
    row(lRaw + 3)[e] = recon_dt(xC, d5, dc, 2, R, mass, L);
  }
  __syncthreads();
  if (mfma_rows) {
    // the angular parts of eth C, eth B go to the rows that take the
    // projected C_t, B_t afterwards (each element is read by its thread
    // before it writes there), eth Pi's to the spare row
    const int co[7] = {kCoScalar, kCoB, kCoPi, kCoC, kCoC, kCoPi, kCoB};
    const int ki[7] = {kOpProject, kOpProject, kOpProject, kOpProject,
                       kOpRaise,   kOpRaise,   kOpRaise};
    const int sr[7] = {lRaw, lRaw + 1, lRaw + 2, lRaw + 3, lC, lPi, lB};
    const int dr_[7] = {lRaw, lRaw + 1, lRaw + 2, lRaw + 3, ltC, lAng, ltB};
    const bool sh[7] = {false, false, false, false, false, false, false};
    apply_rows(7, co, ki, sr, dr_, sh);
    __syncthreads();
  }
  if (live) {
    d6 = dr(6);
    if (mfma_rows) {
      tH = row(lRaw)[e];
      tB = row(lRaw + 1)[e];
      tPi = row(lRaw + 2)[e];
      tC = row(lRaw + 3)[e];
      e2C = ghp_from(kCoC, false, xC, tC, row(ltC)[e]);
      e2Pi = ghp_from(kCoPi, false, xPi, tPi, row(lAng)[e]);
      ex(0, ghp_from(kCoB, false, xB, tB, row(ltB)[e]));
    } else {
      tH = project(kCoScalar, row(lRaw));
      tB = project(kCoB, row(lRaw + 1));
      tPi = project(kCoPi, row(lRaw + 2));
      tC = project(kCoC, row(lRaw + 3));
      e2C = ghp(kCoC, false, row(lC), xC, tC);
      e2Pi = ghp(kCoPi, false, row(lPi), xPi, tPi);
      ex(0, ghp(kCoB, false, row(lB), xB, tB));
    }
    store(3, tH);
    store(4, tB);
    store(5, tPi);
    store(6, tC);
    row(ltB)[e] = tB;
    row(ltC)[e] = tC;
  }
  __syncthreads();
  if (mfma_rows) {
    const int co[2] = {kCoBSharp, kCoCSharp}, ki[2] = {kOpLower, kOpLower};
    const int sr[2] = {lB, lC}, dr_[2] = {lRaw + 2, lRaw + 3};
    const bool sh[2] = {true, true};
    apply_rows(2, co, ki, sr, dr_, sh);
    __syncthreads();
  }
  // pass 3: U, with the sharp fields conj(X[-m]) (spatial_pipeline.hpp:
  // 895-932) read from their partners' rows.  The raw rows are free from
  // here on: the first takes the raw U derivative, the second the H row
  // (for the constraint eth').
  if (live) {
    const cdouble ep1BS = mfma_rows
                              ? ghp_sharp_from(kCoBSharp, row(lB), row(ltB), row(lRaw + 2)[e])
                              : ghp_sharp(kCoBSharp, row(lB), row(ltB));
    const cdouble ep2CS = mfma_rows
                              ? ghp_sharp_from(kCoCSharp, row(lC), row(ltC), row(lRaw + 3)[e])
                              : ghp_sharp(kCoCSharp, row(lC), row(ltC));
    const cdouble mu0b = cconj(bg.mu0), pi0b = cconj(bg.pi0), tau0b = cconj(bg.tau0);
    const cdouble pi_s = sharp_of(lPi), b_s = sharp_of(lB), c_s = sharp_of(lC);
    cdouble du = cmul(cscale(-R, mu0b), xU);
    du = cadd(du, cmul(cscale(-R, bg.mu0), e2C));
    du = cadd(du, cmul(cmul(cscale(-R2, bg.mu0), cadd(pi0b, cscale(2.0, bg.tau0))), xC));
    du = cadd(du, cscale(-2.0, e2Pi));
    du = cadd(du, cmul(cscale(-2.0 * R, pi0b), xPi));
    du = cadd(du, cscale(-2.0, xH));
    du = cadd(du, cmul(cscale(-2.0 * R, bg.pi0), pi_s));
    du = cadd(du, cmul(cscale(-R, bg.pi0), ep1BS));
    du = cadd(du, cmul(cmul(cscale(R2, bg.pi0), bg.pi0), b_s));
    du = cadd(du, cmul(cscale(R, bg.mu0), ep2CS));
    const cdouble sc = cadd(csub(cscale(-3.0, cmul(bg.mu0, bg.pi0)), cscale(-2.0, cmul(mu0b, bg.pi0))),
                            cscale(-2.0, cmul(bg.mu0, tau0b)));
    du = cadd(du, cmul(cscale(R2, sc), c_s));
    row(lRaw)[e] = recon_dt(xU, d6, du, 3, R, mass, L);
    row(lRaw + 1)[e] = xH;
    cdouble* ro = KIND == 0 ? a.rang : a.rang_t;
    ro[sidx(I, kSlotsRang, mode, kRangEth2C, r, t)] = e2C;
    ro[sidx(I, kSlotsRang, mode, kRangEthP2CSharp, r, t)] = ep2CS;
    ro[sidx(I, kSlotsRang, mode, kRangEth2G, r, t)] = e2G;
    ro[sidx(I, kSlotsRang, mode, kRangEthP1BSharp, r, t)] = ep1BS;
    if (KIND == 0 && a.want_constraints) {
      a.rangx[sidx(I, kSlotsRangX, mode, 0, r, t)] = e1F;
      a.rangx[sidx(I, kSlotsRangX, mode, 1, r, t)] = e2Pi;
    }
  }
  __syncthreads();
  if (mfma_rows) {
    const int co[1] = {kCoScalar}, ki[1] = {kOpProject}, sr[1] = {lRaw}, dr_[1] = {lRaw};
    const bool sh[1] = {false};
    apply_rows(1, co, ki, sr, dr_, sh);
    __syncthreads();
  }
  if (!live) return;
  store(7, mfma_rows ? row(lRaw)[e] : project(kCoScalar, row(lRaw)));
  if constexpr (KIND == 0) {
    if (!a.want_constraints) return;
    // reloaded rather than held in registers through the chain
    xF = load(0);
    xG = load(1);
    xLa = load(2);
    xH = load(3);
    xB = load(4);
    xPi = load(5);
    xC = load(6);
    xU = load(7);
    tF = a.out[fidx(I, mode, 2, r, t)];
    tG = a.out[fidx(I, mode, 3, r, t)];
    tH = a.out[fidx(I, mode, 5, r, t)];
    const cdouble ethpG = ghp(kCoG, true, row(lG), xG, tG);
    const cdouble ethpH = ghp(kCoScalar, true, row(lRaw + 1), xH, tH);
    const Bg bs = bg_at(a, I, false, r, t);
    const cdouble Us = cconj(a.in[fidx(I, sm, 9, r, t)]);
    const int m = I.modes[mode];
    const cdouble th1 = thorn_n(xF, tF, a.dpsi[sidx(I, 1, mode, 0, r, t)], 1, -2, -2, m, R, cx,
                                mass, aspin, L, bs.eps0);
    const cdouble th2 = thorn_n(xG, tG, a.rdr[sidx(I, kSlotsRdr, mode, 0, r, t)], 2, -1, -1, m, R,
                                cx, mass, aspin, L, bs.eps0);
    const cdouble l3 = cadd(cadd(cscale(R, ethpG), cmul(cscale(4.0 * R2, bs.pi0), xG)),
                            cmul(bs.rho0, xF));
    const cdouble r3 = cadd(th1, cmul(cscale(3.0 * R2, bs.psi20), xLa));
    const cdouble bgt = cmul(bs.psi20, csub(csub(cmul(cscale(-3.0 * R3, bs.mu0), xC),
                                                 cmul(cscale(1.5 * R3, bs.tau0), xB)),
                                            cscale(3.0 * R2, xPi)));
    const cdouble l2 = csub(th2, cmul(cscale(2.0, bs.rho0), xG));
    const cdouble r2 = csub(cadd(cscale(R, ethpH), cmul(cscale(3.0 * R2, bs.pi0), xH)), bgt);
    const cdouble lh = cdiv(xU, bs.mu0);
    const cdouble rh = cdiv(Us, cconj(bs.mu0));
    const cdouble tv[6] = {l3, r3, l2, r2, lh, rh};
    for (int k = 0; k < 3; ++k)
      a.cons[sidx(I, kSlotsCons, mode, k, r, t)] = csub(tv[2 * k], tv[2 * k + 1]);
    for (int k = 0; k < 6; ++k) a.terms[sidx(I, kSlotsTerms, mode, k, r, t)] = tv[k];
  }
}

// ------------------------------------------------------------ quadratic source
// One thread per (instance, target, r, theta): the ordered-pair sum with Jet1
// tangents (source_tangent_spatial.hpp:137-219, second_order.hpp:74-142) from
// the stage, its RHS and tangent and the derivative inputs in HBM
// (spatial_pipeline.hpp:934-1181).  grid.y = instance.
// Workgroup order: runs of eight point blocks per target, the targets of
// the same eight blocks next to each other (block x -> target (x / 8) %
// targets, point block 8 (x / (8 targets)) + x % 8).  Workgroups x and x + 8
// go to the same XCD, so the targets of one point block, which read the same
// parent values, meet in one L2 at about the same time.
// (Two waves per SIMD spill about 108 registers here and measured 17 %
// slower per second-order step; the kernel keeps one.)
__global__ __launch_bounds__(kSoThreads) void so_source_kernel(SoArgs a, int max_targets) {
  const SoInstance& I = a.inst[a.first_inst + blockIdx.y];
  // The instance record is read through a pointer the compiler cannot prove
  // unclobbered, so its fields would stay in vector registers; taking them
  // as wave-uniform makes every plane start a scalar address.
  const int N = uniform(I.N), T = uniform(I.T);
  const int run = 8 * max_targets;
  const int within = static_cast<int>(blockIdx.x) % run;
  const int ts = within / 8;
  const long long p0 = (static_cast<long long>(blockIdx.x / run) * 8 + within % 8) * kSoThreads;
  if (ts >= uniform(I.n_targets) || p0 >= static_cast<long long>(N) * T) return;
  const double mass = I.mass, L = I.length;
  const long long plane = static_cast<long long>(N) * T;
  const long long base = uniform64(I.base), pbase = uniform64(I.pbase);
  auto slot_of = [](int k) { return 2 + k; };
  enum { xF = 0, xG, xLa, xH, xB, xPi, xC, xU };
  const int pt = static_cast<int>(p0) + static_cast<int>(threadIdx.x);
  if (pt >= N * T) return;
  const int r = pt / T, t = pt - r * T;
  const double R = I.h * r;
  // Every access is a plane start plus the one in-plane byte offset, a
  // 32-bit value (a plane and its background records stay below 4 GiB,
  // Plan::so_check_limits), so the per-thread part of every address is one
  // register shared by all the loads.
  const unsigned off = static_cast<unsigned>(pt) * static_cast<unsigned>(sizeof(cdouble));
  auto at = [&](const cdouble* p) {
    return *reinterpret_cast<const cdouble*>(reinterpret_cast<const char*>(p) + off);
  };
  auto put = [&](cdouble* p, cdouble v) {
    *reinterpret_cast<cdouble*>(reinterpret_cast<char*>(p) + off) = v;
  };
  auto fplane = [&](int mode, int f) { return base + (mode * 13LL + f) * plane; };
  auto splane = [&](int slots, int mode, int sl) {
    return pbase * slots + (static_cast<long long>(mode) * slots + sl) * plane;
  };
  auto S = [&](int mode, int k) { return at(a.in + fplane(mode, slot_of(k))); };
  auto X = [&](int mode, int k) { return at(a.out + fplane(mode, slot_of(k))); };
  auto XT = [&](int mode, int k) { return at(a.tan + splane(kSlotsTan, mode, slot_of(k))); };
  auto rd = [&](int mode, int c) { return at(a.rdr + splane(kSlotsRdr, mode, c)); };
  auto rdt = [&](int mode, int c) { return at(a.rdr_t + splane(kSlotsRdr, mode, c)); };
  auto rg = [&](int mode, int k) { return at(a.rang + splane(kSlotsRang, mode, k)); };
  auto rgt = [&](int mode, int k) { return at(a.rang_t + splane(kSlotsRang, mode, k)); };
  auto ex = [&](int mode, int k) { return at(a.ex + splane(kSlotsEx, mode, k)); };
  const Bg bg = *reinterpret_cast<const Bg*>(reinterpret_cast<const char*>(
      a.bg + uniform64(I.bgoff) * 2) + static_cast<unsigned>(pt) * static_cast<unsigned>(sizeof(Bg)));
  Jet totD{cmake(0.0, 0.0), cmake(0.0, 0.0)}, totT = totD;
  const int k_end = uniform(I.pair_begin[ts + 1]);
  for (int k = uniform(I.pair_begin[ts]); k < k_end; ++k) {
    const int m1 = uniform(I.pair_m1[k]), m2 = uniform(I.pair_m2[k]);
    const int s1 = uniform(I.sharp[m1]), s2 = uniform(I.sharp[m2]);
  auto fj = [&](int mode, int k2) { return Jet{S(mode, k2), X(mode, k2)}; };
    // field jets: value from the stage, tangent from its RHS
    auto f = [&](int i) -> Jet {
      switch (i) {
        case 0: return fj(m1, xF);
        case 1: return fj(m1, xG);
        case 2: return fj(m1, xLa);
        case 3: return fj(m1, xPi);
        case 4: return fj(m1, xB);
        case 5: return fj(m1, xC);
        case 6: return fj(m1, xU);
        case 7: return fj(m2, xF);
        case 8: return fj(m2, xG);
        case 9: return fj(m2, xH);
        case 10: return fj(m2, xB);
        case 11: return fj(m2, xC);
        case 12: return fj(m2, xU);
        case 13: return jconj(fj(s2, xU));
        case 14: return jconj(fj(s2, xC));
        case 15: return jconj(fj(s1, xC));
        case 16: return jconj(fj(s1, xB));
        case 17: return jconj(fj(s2, xPi));
        default: return jconj(fj(s2, xB));
      }
    };
    // the eleven derivative jets of m2 (spatial_pipeline.hpp:1019-1180)
    auto dl = [&](int mode, int k2, int n, int c) {
      const cdouble x = X(mode, k2);
      return Jet{delta_n(S(mode, k2), x, rd(mode, c), n, R, mass, L),
                 delta_n(x, XT(mode, k2), rdt(mode, c), n, R, mass, L)};
    };
    auto d = [&](int i) -> Jet {
      switch (i) {
        case 0: {
          const cdouble xf = X(m2, xF);
          return {delta_n(S(m2, xF), xf, at(a.dpsi + splane(1, m2, 0)), 1, R, mass, L),
                  delta_n(xf, XT(m2, xF), at(a.dpsi_t + splane(1, m2, 0)), 1, R, mass,
                          L)};
        }
        case 1: return dl(m2, xU, 3, 6);
        case 2: return {rg(m2, kRangEth2C), rgt(m2, kRangEth2C)};
        case 3: return {rg(m2, kRangEthP2CSharp), rgt(m2, kRangEthP2CSharp)};
        case 4: return {ex(m2, 0), ex(m2, 2)};
        case 5: return dl(m2, xC, 2, 5);
        case 6: return dl(m2, xG, 2, 0);
        case 7: return {rg(m2, kRangEth2G), rgt(m2, kRangEth2G)};
        case 8: return {ex(m2, 1), ex(m2, 3)};
        case 9: {
          const cdouble xcs = cconj(X(s2, xC));
          return {delta_n(cconj(S(s2, xC)), xcs, cconj(rd(s2, 5)), 2, R, mass, L),
                  delta_n(xcs, cconj(XT(s2, xC)), cconj(rdt(s2, 5)), 2, R, mass, L)};
        }
        default: return {rg(m2, kRangEthP1BSharp), rgt(m2, kRangEthP1BSharp)};
      }
    };
    Jet Dp, Tp;
    pair_source(R, bg, f, d, Dp, Tp);
    if (a.pairs) {
      put(a.pairs + 2 * k * plane, Dp.v);
      put(a.pairs + (2 * k + 1) * plane, Tp.v);
    }
    totD = jadd(totD, Dp);
    totT = jadd(totT, Tp);
  }
  const int mode = uniform(I.target_mode[ts]);
  put(a.raw + splane(kSlotsRaw, mode, 0), totD.v);
  put(a.raw + splane(kSlotsRaw, mode, 1), totD.d);
  put(a.raw + splane(kSlotsRaw, mode, 2), totT.v);
  put(a.raw + splane(kSlotsRaw, mode, 3), totT.d);
}

// ------------------------------------------------------------ source projection
// Band projection of D (spin -2) and T (spin -1) with their tangents and
// eth'_3 T (spatial_pipeline.hpp:1183-1209): one workgroup per kProjRows (4)
// radial rows of an instance.  Each operator application is a 16-row MFMA
// block: the projection of target ts and coordinator (D pair or T pair)
// takes rows (radial row, component of the pair, real/imag), eth'_3 T rows
// (radial row, real/imag); the blocks go round the four waves.
// Six waves per SIMD (at most 80 registers, a few spilled; the LDS of a
// workgroup allows six per CU): measured 2 % faster per second-order step than
// the three its unconstrained 96 registers gave.
__global__ __launch_bounds__(kSoThreads, 6) void so_source_projection_kernel(SoArgs a) {
  static_assert(kProjRows == 4, "a projection block is 4 rows x 2 components x 2 parts");
  extern __shared__ __align__(16) double smem[];
  const SoInstance& I = a.inst[a.first_inst + blockIdx.y];
  const int r0 = blockIdx.x * kProjRows;
  if (r0 >= I.N) return;
  const int T = I.T, NT = I.n_targets, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, lg = lane >> 4, lr = lane & 15;
  constexpr int kWaves = kSoThreads / 64;
  const int per_row = NT * 4 * T;  // raw sums of one row: [targets][4][T]
  const int rows = min(kProjRows, I.N - r0);
  const int ttiles = (T + 15) / 16;
  const double* nodes = node_ptr(a, I);
  cdouble* RW = reinterpret_cast<cdouble*>(smem);    // [rows][targets][4][T] raw sums
  cdouble* PRJ = RW + kProjRows * per_row;            // the same, projected
  cdouble* ANG = PRJ + kProjRows * per_row;           // [rows][targets][T]: angular part of eth'_3 T
  for (int e = tid; e < rows * per_row; e += kSoThreads) {
    const int rr = e / per_row, q = e - rr * per_row;
    const int ts = q / (4 * T), rem = q - ts * 4 * T, c = rem / T, t = rem - c * T;
    RW[e] = a.raw[sidx(I, kSlotsRaw, I.target_mode[ts], c, r0 + rr, t)];
  }
  __syncthreads();
  // projections: block (ts, pair, tile), pair 0 = (D, D_t), 1 = (T, T_t)
  for (int blk = wave; blk < NT * 2 * ttiles; blk += kWaves) {
    const int ts = blk / (2 * ttiles), pr = (blk / ttiles) % 2, tile = blk % ttiles;
    const int pc = pr == 0 ? kCoSecond : kCoT, pm = I.target_mode[ts];
    const int off = I.op[pc][pm][kOpAnalyze];
    auto a_of = [&](int m, int u) {
      const int rr = m >> 2, c = 2 * pr + ((m >> 1) & 1);
      if (rr >= rows) return 0.0;
      const cdouble v = RW[rr * per_row + (ts * 4 + c) * T + u];
      return (m & 1) ? v.im : v.re;
    };
    const mfma4 acc = off < 0 ? mfma4{0.0, 0.0, 0.0, 0.0}
                              : row_block_project(a.ops + off, a.ops + I.op[pc][pm][kOpSynthesize],
                                                  T, I.band_n[pc][pm], tile, lane, a_of);
    const int t = 16 * tile + lr;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int m = lg + 4 * w, rr = m >> 2, c = 2 * pr + ((m >> 1) & 1);
      if (rr < rows && t < T)
        reinterpret_cast<double*>(PRJ + rr * per_row + (ts * 4 + c) * T + t)[m & 1] = acc[w];
    }
  }
  __syncthreads();
  // eth'_3 T: block (ts, tile) over rows (radial row, real/imag), 8 of 16 used
  for (int blk = wave; blk < NT * ttiles; blk += kWaves) {
    const int ts = blk / ttiles, tile = blk % ttiles;
    const int off = I.op[kCoT][I.target_mode[ts]][kOpLower];
    if (off < 0) continue;
    auto a_of = [&](int m, int u) {
      const int rr = m >> 1;
      if (rr >= rows) return 0.0;
      const cdouble v = PRJ[rr * per_row + (ts * 4 + 2) * T + u];
      return (m & 1) ? v.im : v.re;
    };
    const mfma4 acc = row_block_op(a.ops + off, T, tile, lane, a_of);
    const int t = 16 * tile + lr;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int m = lg + 4 * w, rr = m >> 1;
      if (rr < rows && t < T)
        reinterpret_cast<double*>(ANG + (rr * NT + ts) * T + t)[m & 1] = acc[w];
    }
  }
  __syncthreads();
  for (int e = tid; e < rows * NT * T; e += kSoThreads) {
    const int rr = e / (NT * T), q = e - rr * NT * T;
    const int ts = q / T, t = q - ts * T, r = r0 + rr;
    const int mode = I.target_mode[ts];
    const int off = I.op[kCoT][mode][kOpLower];
    const cdouble* pr = PRJ + rr * per_row + ts * 4 * T;
    const cdouble* tv = pr + 2 * T;
    const cdouble* tt = pr + 3 * T;
    cdouble ethT = cmake(0.0, 0.0);
    if (off >= 0)
      ethT = eth_point(true, tv[t], tt[t], ANG[(rr * NT + ts) * T + t], I.coord_spin[kCoT],
                       I.coord_boost[kCoT], I.h * r, nodes[T + t], nodes[t], I.spin, I.length);
    a.src[sidx(I, kSlotsSrc, mode, 0, r, t)] = pr[t];
    a.src[sidx(I, kSlotsSrc, mode, 1, r, t)] = pr[T + t];
    a.src[sidx(I, kSlotsSrc, mode, 2, r, t)] = tv[t];
    a.src[sidx(I, kSlotsSrc, mode, 3, r, t)] = ethT;
  }
}

// ------------------------------------------------------------ background tables
// Kerr background (background.hpp:29-78) at every (r, theta node) of every
// instance, once per plan: storage evaluation then theta evaluation.
__global__ __launch_bounds__(kSoThreads) void so_background_kernel(SoArgs a) {
  const SoInstance& I = a.inst[blockIdx.y];
  const double* nodes = a.nodes + I.node_off;
  const long long n = static_cast<long long>(I.N) * I.T;
  for (long long e = static_cast<long long>(blockIdx.x) * kSoThreads + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * kSoThreads) {
    const int r = static_cast<int>(e / I.T), t = static_cast<int>(e % I.T);
    const double R = I.h * r;
    a.bg[I.bgoff * 2 + e] = kerr_bg(I.mass, I.spin, I.length, R, nodes[t], nodes[I.T + t]);
    a.bg[I.bgoff * 2 + n + e] =
        kerr_bg(I.mass, I.spin, I.length, R, nodes[2 * I.T + t], nodes[3 * I.T + t]);
  }
}

hipError_t so_launch_background(const SoArgs& a, long long max_points, hipStream_t s) {
  if (a.n_inst <= 0) return hipSuccess;
  const long long want = (max_points + kSoThreads - 1) / kSoThreads;
  const dim3 g(static_cast<unsigned>(want < 1 ? 1 : (want > 4096 ? 4096 : want)), a.n_inst);
  hipLaunchKernelGGL(so_background_kernel, g, dim3(kSoThreads), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------------ RK4 combination
// Classical RK4 (device_rk4.hpp:84-128) on whole instance blocks:
//   stage 1: acc = y + dt/6 k, Y2 = y + dt/2 k;   stage 2: acc += dt/3 k,
//   Y3 = y + dt/2 k;   stage 3: acc += dt/3 k, Y4 = y + dt k;
//   stage 4: y = acc + dt/6 k.   grid.y = instance.
// (mode, field) blocks outside the field's mode set are skipped: their
// state is zero in every buffer (initial data, band-validated imports) and
// their RHS is exactly zero (the projections of inactive modes are zero
// operators), so the combination would leave them zero bit for bit.
__global__ __launch_bounds__(kSoThreads) void so_rk_kernel(SoRkArgs a) {
  const SoInstance& I = a.inst[blockIdx.y];
  const double dt = a.dt[blockIdx.y];
  const long long plane = static_cast<long long>(I.N) * I.T;
  const long long n = static_cast<long long>(I.M) * 13 * plane;
  for (long long e = static_cast<long long>(blockIdx.x) * kSoThreads + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * kSoThreads) {
    const int mf = static_cast<int>(e / plane), mode = mf / 13, f = mf - mode * 13;
    if (!((I.field_mask[mode] >> f) & 1)) continue;
    const long long g = I.base + e;
    const cdouble k = a.k[g];
    if (a.stage == 1) {
      const cdouble y = a.y[g];
      a.acc[g] = cadd(y, cscale(dt / 6.0, k));
      a.yout[g] = cadd(y, cscale(0.5 * dt, k));
    } else if (a.stage == 2) {
      a.acc[g] = cadd(a.acc[g], cscale(dt / 3.0, k));
      a.yout[g] = cadd(a.y[g], cscale(0.5 * dt, k));
    } else if (a.stage == 3) {
      a.acc[g] = cadd(a.acc[g], cscale(dt / 3.0, k));
      a.yout[g] = cadd(a.y[g], cscale(dt, k));
    } else {
      a.yout[g] = cadd(a.acc[g], cscale(dt / 6.0, k));
    }
  }
}

// ------------------------------------------------------------ constraint norms
// One workgroup per instance; fixed-order tree reductions, so the sums do not
// depend on scheduling (the reference reduces with atomics,
// source_activation.hpp:120-170).  Output per instance: weighted sums of
// c^2, l^2, r^2 (3 each), max c^2 (3), max natural (3), max normalized^2 (3),
// finiteness flag.
__global__ __launch_bounds__(kSoThreads) void so_reduce_kernel(SoReduceArgs a) {
  __shared__ double red[kSoThreads];
  const SoInstance& I = a.inst[blockIdx.x];
  const int tid = threadIdx.x;
  const SchemeTable& TB = a.tables[I.scheme];
  const int W = TB.W;
  const double* wts = a.nodes + I.node_off + 4 * I.T;
  const long long plane = static_cast<long long>(I.M) * I.N * I.T;
  double acc[15];
  for (double& v : acc) v = 0.0;
  int finite = 1;
  const double two_pi = 6.283185307179586;
  for (long long e = tid; e < plane; e += kSoThreads) {
    const int t = static_cast<int>(e % I.T);
    const int r = static_cast<int>((e / I.T) % I.N);
    const int mode = static_cast<int>(e / (static_cast<long long>(I.T) * I.N));
    const int dist = r < I.N - 1 - r ? r : I.N - 1 - r;
    const double hw = dist < W ? TB.norm[dist] : 1.0;
    const double weight = I.h * hw * two_pi * wts[t];
    for (int f = 0; f < 3; ++f) {
      const cdouble cv = a.cons[sidx(I, kSlotsCons, mode, f, r, t)];
      const cdouble lv = a.terms[sidx(I, kSlotsTerms, mode, 2 * f, r, t)];
      const cdouble rv = a.terms[sidx(I, kSlotsTerms, mode, 2 * f + 1, r, t)];
      const double c = hypot(cv.re, cv.im), l = hypot(lv.re, lv.im), rr = hypot(rv.re, rv.im);
      if (!isfinite(c) || !isfinite(l) || !isfinite(rr)) {
        finite = 0;
        continue;
      }
      acc[f] += weight * c * c;
      acc[3 + f] += weight * l * l;
      acc[6 + f] += weight * rr * rr;
      acc[9 + f] = fmax(acc[9 + f], c * c);
      acc[12 + f] = fmax(acc[12 + f], fmax(l * l, rr * rr));
    }
  }
  double* out = a.out + static_cast<long long>(blockIdx.x) * kSoReduceDoubles;
  auto reduce = [&](double v, bool is_max) {
    red[tid] = v;
    __syncthreads();
    for (int s = kSoThreads / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] = is_max ? fmax(red[tid], red[tid + s]) : red[tid] + red[tid + s];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  double total[15];
  for (int k = 0; k < 15; ++k) total[k] = reduce(acc[k], k >= 9);
  const double fin = reduce(static_cast<double>(1 - finite), true) < 0.5 ? 1.0 : 0.0;
  const double floor = 64.0 * 2.220446049250313e-16;
  double mnorm[3] = {0.0, 0.0, 0.0};
  int finite2 = 1;
  for (long long e = tid; e < plane; e += kSoThreads) {
    const int t = static_cast<int>(e % I.T);
    const int r = static_cast<int>((e / I.T) % I.N);
    const int mode = static_cast<int>(e / (static_cast<long long>(I.T) * I.N));
    for (int f = 0; f < 3; ++f) {
      const cdouble cv = a.cons[sidx(I, kSlotsCons, mode, f, r, t)];
      const cdouble lv = a.terms[sidx(I, kSlotsTerms, mode, 2 * f, r, t)];
      const cdouble rv = a.terms[sidx(I, kSlotsTerms, mode, 2 * f + 1, r, t)];
      const double c = hypot(cv.re, cv.im), l = hypot(lv.re, lv.im), rr = hypot(rv.re, rv.im);
      const double den = l + rr + floor * sqrt(total[12 + f]);
      const double nv = den > 0.0 ? c / den : 0.0;
      if (isfinite(nv)) mnorm[f] = fmax(mnorm[f], nv * nv);
      else finite2 = 0;
    }
  }
  double m2[3];
  for (int f = 0; f < 3; ++f) m2[f] = reduce(mnorm[f], true);
  const double fin2 = reduce(static_cast<double>(1 - finite2), true) < 0.5 ? 1.0 : 0.0;
  if (tid == 0) {
    for (int k = 0; k < 15; ++k) out[k] = total[k];
    for (int f = 0; f < 3; ++f) out[15 + f] = m2[f];
    out[18] = fin * fin2;
  }
}

// ------------------------------------------------------------ launchers
namespace {
template <int K>
hipError_t launch_linear_k(int scheme, const SoArgs& a, int n, int lds, hipStream_t s) {
  const dim3 g(n), b(kSoThreads);
  switch (scheme) {
    case 0: hipLaunchKernelGGL((so_linear_kernel<0, K>), g, b, lds, s, a); break;
    case 1: hipLaunchKernelGGL((so_linear_kernel<1, K>), g, b, lds, s, a); break;
    default: hipLaunchKernelGGL((so_linear_kernel<2, K>), g, b, lds, s, a); break;
  }
  return hipGetLastError();
}
}  // namespace

hipError_t so_launch_linear(int scheme, int kind, const SoArgs& a, int n_tiles, int lds,
                            hipStream_t s) {
  if (n_tiles <= 0) return hipSuccess;
  if (kind == 0) return launch_linear_k<0>(scheme, a, n_tiles, lds, s);
  if (kind == 1) return launch_linear_k<1>(scheme, a, n_tiles, lds, s);
  return launch_linear_k<2>(scheme, a, n_tiles, lds, s);
}

hipError_t so_launch_drad(int which, const SoArgs& a, int blocks, unsigned schemes,
                          hipStream_t s) {
  if (a.n_inst <= 0 || blocks <= 0) return hipSuccess;
  const dim3 g(blocks, a.n_inst);
  if (schemes & 1u) hipLaunchKernelGGL(so_drad_kernel<0>, g, dim3(kSoThreads), 0, s, a, which);
  if (schemes & 2u) hipLaunchKernelGGL(so_drad_kernel<1>, g, dim3(kSoThreads), 0, s, a, which);
  if (schemes & 4u) hipLaunchKernelGGL(so_drad_kernel<2>, g, dim3(kSoThreads), 0, s, a, which);
  return hipGetLastError();
}

hipError_t so_launch_recon(int kind, const SoArgs& a, int groups, int max_mt, int lds,
                           hipStream_t s) {
  if (a.n_inst <= 0) return hipSuccess;
  if (max_mt > kReconMaxThreads) return hipErrorInvalidValue;
  const dim3 g(groups, a.n_inst);
  if (max_mt <= kReconThreads) {
    const dim3 b(kReconThreads);
    if (kind == 0) hipLaunchKernelGGL((so_recon_kernel<0, kReconThreads>), g, b, lds, s, a);
    else hipLaunchKernelGGL((so_recon_kernel<1, kReconThreads>), g, b, lds, s, a);
  } else {
    const dim3 b((max_mt + 63) / 64 * 64);
    if (kind == 0) hipLaunchKernelGGL((so_recon_kernel<0, kReconMaxThreads>), g, b, lds, s, a);
    else hipLaunchKernelGGL((so_recon_kernel<1, kReconMaxThreads>), g, b, lds, s, a);
  }
  return hipGetLastError();
}

hipError_t so_launch_source(const SoArgs& a, long long max_points, int max_targets,
                            hipStream_t s) {
  if (a.n_inst <= 0 || max_targets <= 0 || max_points <= 0) return hipSuccess;
  const long long blocks = (max_points + kSoThreads - 1) / kSoThreads;
  const long long gx = (blocks + 7) / 8 * 8 * max_targets;
  if (gx > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  const dim3 g(static_cast<unsigned>(gx), a.n_inst);
  hipLaunchKernelGGL(so_source_kernel, g, dim3(kSoThreads), 0, s, a, max_targets);
  return hipGetLastError();
}

hipError_t so_launch_source_projection(const SoArgs& a, int rows, int lds, hipStream_t s) {
  if (a.n_inst <= 0) return hipSuccess;
  if (lds > 64 * 1024) {
    const hipError_t e =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&so_source_projection_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  const dim3 g((rows + kProjRows - 1) / kProjRows, a.n_inst);
  hipLaunchKernelGGL(so_source_projection_kernel, g, dim3(kSoThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t so_launch_rk(const SoRkArgs& a, hipStream_t s) {
  if (a.n_inst <= 0) return hipSuccess;
  const long long want = (a.max_values + kSoThreads - 1) / kSoThreads;
  const dim3 g(static_cast<unsigned>(want < 1 ? 1 : (want > 8192 ? 8192 : want)), a.n_inst);
  hipLaunchKernelGGL(so_rk_kernel, g, dim3(kSoThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t so_launch_reduce(const SoReduceArgs& a, hipStream_t s) {
  if (a.n_inst <= 0) return hipSuccess;
  hipLaunchKernelGGL(so_reduce_kernel, dim3(a.n_inst), dim3(kSoThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t so_set_lds_limit(int bytes) {
  const void* fns[] = {
      reinterpret_cast<const void*>(&so_linear_kernel<0, 0>),
      reinterpret_cast<const void*>(&so_linear_kernel<0, 1>),
      reinterpret_cast<const void*>(&so_linear_kernel<0, 2>),
      reinterpret_cast<const void*>(&so_linear_kernel<1, 0>),
      reinterpret_cast<const void*>(&so_linear_kernel<1, 1>),
      reinterpret_cast<const void*>(&so_linear_kernel<1, 2>),
      reinterpret_cast<const void*>(&so_linear_kernel<2, 0>),
      reinterpret_cast<const void*>(&so_linear_kernel<2, 1>),
      reinterpret_cast<const void*>(&so_linear_kernel<2, 2>),
      reinterpret_cast<const void*>(&so_recon_kernel<0, kReconThreads>),
      reinterpret_cast<const void*>(&so_recon_kernel<1, kReconThreads>),
      reinterpret_cast<const void*>(&so_recon_kernel<0, kReconMaxThreads>),
      reinterpret_cast<const void*>(&so_recon_kernel<1, kReconMaxThreads>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace teuk
