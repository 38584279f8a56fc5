// Sampled diagnostics of one instance on the GPU: the pipeline report
// (pipeline_diagnostics.hpp:145-271: RMS and maximum of every state and RHS
// field, of the two reduction constraints Q - D Psi, of source_over_r3 and of
// the latched forcing), the independent reconstruction constraints
// (pipeline_independent_reconstruction_diagnostics.hpp:33-98) and the
// transport-equation residuals (pipeline_reconstruction_diagnostics.hpp:
// 43-190, reconstruction_diagnostics.hpp:30-140).  The reference accumulates
// with atomics; here every wave reduces its 64 points with a fixed butterfly
// and a second pass adds the wave partials in wave order, so a sample does
// not depend on scheduling.
#include <hip/hip_runtime.h>

#include "diag_layout.hpp"
#include "sbp_device.hpp"
#include "so_device.hpp"

namespace teuk {

namespace {

__device__ __forceinline__ bool finite_c(cdouble v) { return isfinite(v.re) && isfinite(v.im); }
__device__ __forceinline__ double norm2(cdouble v) { return v.re * v.re + v.im * v.im; }

// (mode, slot, r, t) of a per-instance block with S slots
__device__ __forceinline__ long long bidx(const DiagArgs& a, int S, int mode, int s, int r, int t) {
  return ((static_cast<long long>(mode) * S + s) * a.N + r) * a.T + t;
}

// LHS of Delta_n X = R_X for the transport residual (reconstruction_diagnostics.hpp:30-40)
__device__ __forceinline__ cdouble delta_lhs(cdouble v, cdouble dt, cdouble dr, int n, double R,
                                             double M, double L) {
  const double L2 = L * L;
  return cadd(cscale(2.0 + 4.0 * M * R / L2, dt),
              cscale(R / L2, cadd(cscale(R, dr), cscale(static_cast<double>(n), v))));
}

}  // namespace

// One thread per (mode, r, theta) point; each wave writes its partials.
template <int SCHEME>
__global__ __launch_bounds__(kDiagThreads) void teuk_diag_kernel(DiagArgs a) {
  const long long total = static_cast<long long>(a.M) * a.N * a.T;
  const long long e = static_cast<long long>(blockIdx.x) * kDiagThreads + threadIdx.x;
  const bool live = e < total;
  const int t = live ? static_cast<int>(e % a.T) : 0;
  const int r = live ? static_cast<int>((e / a.T) % a.N) : 0;
  const int mode = live ? static_cast<int>(e / (static_cast<long long>(a.T) * a.N)) : 0;
  const int lane = threadIdx.x & 63;
  const long long wave = (static_cast<long long>(blockIdx.x) * kDiagThreads + threadIdx.x) >> 6;
  double* psum = a.partial + wave * 2 * kDqCount;
  double* pmax = psum + kDqCount;
  int flags = 0;
  const bool edge_scri = r == 0, edge_hor = r == a.N - 1;
  auto fail = [&](int bit, bool endpoint) {
    flags |= bit;
    if (endpoint && edge_scri) flags |= kDfScri;
    if (endpoint && edge_hor) flags |= kDfHorizon;
  };
  // one quantity: finite |v|^2 of live points, reduced over the wave
  auto reduce = [&](int q, cdouble v, int bit, bool endpoint) {
    double x = 0.0;
    if (live) {
      if (finite_c(v)) x = norm2(v);
      else fail(bit, endpoint);
    }
    double s = x, m = x;
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_xor(s, off);
      m = fmax(m, __shfl_xor(m, off));
    }
    if (lane == 0) {
      psum[q] = s;
      pmax[q] = m;
    }
  };
  auto st = [&](int f, int rr) { return a.state[bidx(a, 13, mode, f, rr, t)]; };
  const cdouble zero = cmake(0.0, 0.0);
  for (int f = 0; f < 13; ++f) reduce(kDqState + f, live ? st(f, r) : zero, kDfState, true);
  for (int f = 0; f < 13; ++f)
    reduce(kDqRhs + f, live ? a.rhs[bidx(a, 13, mode, f, r, t)] : zero, kDfRhs, true);
  const SchemeTable& TB = *a.table;
  auto dcol = [&](int f) {
    return apply_d<SCHEME>(TB, a.state + bidx(a, 13, mode, f, 0, t), a.T, 0, r, a.N, a.inv_h);
  };
  reduce(kDqFirstConstraint, live ? csub(st(1, r), dcol(2)) : zero, kDfFirstConstraint, false);
  reduce(kDqSecondConstraint, live ? csub(st(11, r), dcol(12)) : zero, kDfSecondConstraint,
         false);
  const double R = a.h * r, R2 = R * R, L = a.length;
  const double x = a.nodes[t], sx = a.nodes[a.T + t];
  const Bg bg = kerr_bg(a.mass, a.spin, L, R, x, sx);
  // source_over_r3 and the forcing (source_spatial.hpp:357-425,
  // second_order.hpp:144-182)
  cdouble src = zero, frc = zero;
  if (live && a.src) {
    const cdouble* dc = a.src + bidx(a, kSlotsSrc, mode, 0, 0, t);
    const cdouble Dv = dc[static_cast<long long>(r) * a.T];
    const cdouble Dt = a.src[bidx(a, kSlotsSrc, mode, 1, r, t)];
    const cdouble Tv = a.src[bidx(a, kSlotsSrc, mode, 2, r, t)];
    const cdouble ethT = a.src[bidx(a, kSlotsSrc, mode, 3, r, t)];
    const cdouble Dr = apply_d<SCHEME>(TB, dc, a.T, 0, r, a.N, a.inv_h);
    src = delta_n(Dv, Dt, Dr, 3, R, a.mass, L);
    src = cadd(src, cmul(cscale(R, cadd(cscale(4.0, bg.mu0), cconj(bg.mu0))), Dv));
    src = cadd(src, cscale(R, ethT));
    src = cadd(src, cmul(cscale(R2, csub(cscale(4.0, bg.pi0), cconj(bg.tau0))), Tv));
    const double L2 = L * L;
    const double norm = 2.0 * (L2 * L2 + a.spin * a.spin * R * R * x * x);
    if (a.active) frc = cscale(norm, src);
  }
  reduce(kDqSource, src, kDfSource, true);
  reduce(kDqForcing, frc, kDfForcing, true);
  for (int k = 0; k < 3; ++k)
    reduce(kDqIndependent + k, live && a.cons ? a.cons[bidx(a, kSlotsCons, mode, k, r, t)] : zero,
           kDfIndependent, false);
  // transport residuals L_X - R_X (reconstruction_diagnostics.hpp:46-140)
  {
    const int sm = a.sharp[mode] >= 0 ? a.sharp[mode] : mode;
    auto sv = [&](int f) { return live ? cconj(a.state[bidx(a, 13, sm, f, r, t)]) : zero; };
    auto fv = [&](int f) { return live ? st(f, r) : zero; };
    auto rh = [&](int f) { return live ? a.rhs[bidx(a, 13, mode, f, r, t)] : zero; };
    auto dr = [&](int c) {
      return live && a.rdr ? a.rdr[bidx(a, kSlotsRdr, mode, c, r, t)] : zero;
    };
    auto ang = [&](int k) {
      return live && a.rang ? a.rang[bidx(a, kSlotsRang, mode, k, r, t)] : zero;
    };
    auto angx = [&](int k) {
      return live && a.rangx ? a.rangx[bidx(a, kSlotsRangX, mode, k, r, t)] : zero;
    };
    const cdouble F = fv(2), G = fv(3), La = fv(4), H = fv(5), B = fv(6), Pi = fv(7), C = fv(8),
                  U = fv(9);
    const cdouble Pis = sv(7), Bs = sv(6), Cs = sv(8);
    const cdouble mub = cconj(bg.mu0), pib = cconj(bg.pi0), taub = cconj(bg.tau0);
    const cdouble rG = csub(cadd(cmul(cscale(-4.0 * R, bg.mu0), G), angx(0)),
                            cmul(cscale(R, bg.tau0), F));
    const cdouble rLa = csub(cmul(cscale(-R, cadd(bg.mu0, mub)), La), F);
    const cdouble rH = csub(cadd(cmul(cscale(-3.0 * R, bg.mu0), H), ang(kRangEth2G)),
                            cmul(cscale(2.0 * R, bg.tau0), G));
    const cdouble rB = csub(cmul(cscale(R, csub(bg.mu0, mub)), B), cscale(2.0, La));
    const cdouble rPi = cadd(csub(cneg(G), cmul(cscale(R, cadd(pib, bg.tau0)), La)),
                             cmul(cmul(cscale(0.5 * R2, bg.mu0), cadd(pib, bg.tau0)), B));
    const cdouble rC = csub(csub(cmul(cscale(-R, mub), C), cscale(2.0, Pi)),
                            cmul(cscale(R, bg.tau0), B));
    cdouble rU = cmul(cscale(-R, mub), U);
    rU = cadd(rU, cmul(cscale(-R, bg.mu0), ang(kRangEth2C)));
    rU = cadd(rU, cmul(cmul(cscale(-R2, bg.mu0), cadd(pib, cscale(2.0, bg.tau0))), C));
    rU = cadd(rU, cscale(-2.0, angx(1)));
    rU = cadd(rU, cmul(cscale(-2.0 * R, pib), Pi));
    rU = cadd(rU, cscale(-2.0, H));
    rU = cadd(rU, cmul(cscale(-2.0 * R, bg.pi0), Pis));
    rU = cadd(rU, cmul(cscale(-R, bg.pi0), ang(kRangEthP1BSharp)));
    rU = cadd(rU, cmul(cmul(cscale(R2, bg.pi0), bg.pi0), Bs));
    rU = cadd(rU, cmul(cscale(R, bg.mu0), ang(kRangEthP2CSharp)));
    const cdouble sc = cadd(csub(cscale(-3.0, cmul(bg.mu0, bg.pi0)), cscale(-2.0, cmul(mub, bg.pi0))),
                            cscale(-2.0, cmul(bg.mu0, taub)));
    rU = cadd(rU, cmul(cscale(R2, sc), Cs));
    const double M = a.mass;
    const cdouble res[7] = {
        csub(delta_lhs(G, rh(3), dr(0), 2, R, M, L), rG),
        csub(delta_lhs(La, rh(4), dr(1), 1, R, M, L), rLa),
        csub(delta_lhs(H, rh(5), dr(2), 3, R, M, L), rH),
        csub(delta_lhs(B, rh(6), dr(3), 1, R, M, L), rB),
        csub(delta_lhs(Pi, rh(7), dr(4), 2, R, M, L), rPi),
        csub(delta_lhs(C, rh(8), dr(5), 2, R, M, L), rC),
        csub(delta_lhs(U, rh(9), dr(6), 3, R, M, L), rU)};
    for (int k = 0; k < 7; ++k) reduce(kDqTransport + k, res[k], kDfTransport, false);
  }
  for (int off = 32; off > 0; off >>= 1) flags |= __shfl_xor(flags, off);
  if (lane == 0) a.partial_flags[wave] = flags;
}

// Wave partials added in wave order: thread q < 2 kDqCount owns one sum or
// maximum; thread 2 kDqCount ORs the flags.
__global__ __launch_bounds__(128) void teuk_diag_finish_kernel(DiagArgs a) {
  const int q = threadIdx.x;
  if (q < 2 * kDqCount) {
    const bool is_max = q >= kDqCount;
    double acc = 0.0;
    for (int w = 0; w < a.n_waves; ++w) {
      const double v = a.partial[static_cast<long long>(w) * 2 * kDqCount + q];
      acc = is_max ? fmax(acc, v) : acc + v;
    }
    a.out[q] = acc;
  } else if (q == 2 * kDqCount) {
    int f = 0;
    for (int w = 0; w < a.n_waves; ++w) f |= a.partial_flags[w];
    a.out[q] = static_cast<double>(f);
  }
}

hipError_t launch_diagnostics(const DiagArgs& a, hipStream_t s) {
  const long long points = static_cast<long long>(a.M) * a.N * a.T;
  if (points <= 0) return hipSuccess;
  const dim3 g(static_cast<unsigned>((points + kDiagThreads - 1) / kDiagThreads));
  if (a.scheme == 0) hipLaunchKernelGGL(teuk_diag_kernel<0>, g, dim3(kDiagThreads), 0, s, a);
  else if (a.scheme == 1) hipLaunchKernelGGL(teuk_diag_kernel<1>, g, dim3(kDiagThreads), 0, s, a);
  else hipLaunchKernelGGL(teuk_diag_kernel<2>, g, dim3(kDiagThreads), 0, s, a);
  hipLaunchKernelGGL(teuk_diag_finish_kernel, dim3(1), dim3(128), 0, s, a);
  return hipGetLastError();
}

}  // namespace teuk
