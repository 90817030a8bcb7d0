// dg_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the Navier-Stokes residual for the
// matrix-core element sizes, FP64.
//
// Forward residual of one element (the MFMA stage kernels):
//   u_q = V U, traces u_s = Vf U;  A_s = +-(u_L + u_R)/2 (x) n |J_f| w,  J_s = (u_R - u_L)/2 (x) n |J_f| w
//   R_w = A Phi_f - Q Phi_grad (Q = u_q (x) (J^T)^-1 |J| w),  R_t = R_w + J Phi_f
//   grad u_q = R_t (Mref^-1 V^T) / |J|,  G_s = R_w (Mref^-1 Vf^T) / |J| + J_s Lf_s / |J| (BR2)
//   F_q = (F_c - F_v)(u_q, grad u_q) (J^T)^-1 |J| w,  t = (H_c - <F_v> . n) |J_f| w -> slots -t | +t
//   R = F Phi_grad + H Phi_f
// Its transpose, given mu (the adjoint of R), runs as four kernels:
//   nsAdjVolumeMfma  recompute the lifting (and Gf), H-bar = Vf mu, F-bar = Phi_grad mu, the
//                    pointwise volume-flux VJP -> (u-bar_q, grad-bar_q), R-bar_t(vol) =
//                    (Mref^-1 V^T)^T grad-bar / |J|, U-bar += V^T u-bar_q
//   nsAdjFaceKernel  per face point: t-bar = H-bar_R - H-bar_L, face-flux VJP -> trace and
//                    face-gradient adjoints of both parents' slots
//   nsAdjLiftMfma    R-bar_w, R-bar_t from the face-gradient adjoint, A-bar, J-bar, Q-bar, and
//                    their images on u_q and on both parents' traces
//   nsAdjGatherMfma  U-bar += Vf^T (trace adjoints of own slots and of the neighbours' lifting terms)
// Pointwise Jacobians come from forward-mode dual numbers (dual.hpp) for the state dependence and
// from the linearity of the viscous flux in the gradient for the gradient dependence.
#ifndef SDG_DG_ADJOINT_HPP_
#define SDG_DG_ADJOINT_HPP_

#include <hip/hip_runtime.h>

#include "dg_mfma.hpp"
#include "dual.hpp"
#include "physics.hpp"
#include "physics_adjoint.hpp"

namespace sdg {

// transposed operator tables of the adjoint (padded T[r4(K)][r16(M)] like MOps)
struct AdjOps {
  const double* gradT;     // K = N_b,   M = N_q D : Phi_grad^T
  const double* interpT;   // K = N_q,   M = N_b   : V^T
  const double* pinterpT;  // K = N_q,   M = N_b   : (Mref^-1 V^T)^T
  const double* ptraceT;   // K = N_aq,  M = N_b   : (Mref^-1 Vf^T)^T
  const double* liftT;     // K = N_aq,  M = N_aq  : block-diagonal Lf_s^T
};

struct AdjArgs {
  AdjOps op;
  double* lam;        // E x NV x N_b  adjoint seed (updated in place when rk == 1)
  int apply_minv;     // 1: lam is the adjoint of M^-1 R (mu = Mref^-1 lam / |J|); 0: of R (mu = lam)
  double* Hb;         // E x NV x N_aq  mu traced to the face slots
  double* RtV;        // E x NG x N_b   volume part of the adjoint of R_t
  double* Ua;         // E x NV x N_b   partial result
  double* UFf;        // E x NV x N_aq  face-flux adjoint with respect to the own traces
  double* GFb;        // E x NG x N_aq  adjoint of the lifted gradient traces
  double* Pown;       // E x NV x N_aq  lifting-face adjoint to the own trace
  double* Qnb;        // E x NV x N_aq  lifting-face adjoint to the neighbour's trace of the same point
  double* out;        // E x NV x N_b   (dR/dU)^T mu (rk == 0)
  // rk == 1: instead of `out`, apply the transpose of the RK stage U_{s+1} = c0 U_0 + c1 U_s +
  // c2 dt L(U_s) in place: acc (+)= c0 lam, lam = c1 lam + c2dt VJP, and on the last transposed
  // stage lam += acc
  int rk, rk_first, rk_last;
  double* acc;
  double c0, c1, c2dt;
};

// ---- pointwise VJPs --------------------------------------------------------------------------
// raw volume flux F_c - F_v (D x NV column-major) in the FP64 operation order of the stage kernels
template <int D, typename T, bool VISC>
__device__ __forceinline__ void rawVolumeFlux(const Phys& ph, const T* u, const T* g, T* f) {
  constexpr int NV = D + 2, NG = NV * D;
  T w[D + 3];
  computational<D>(ph, u, w);
  const T rho = w[0], p = w[D + 2];
  const T et = w[D + 1] + velocitySquared<D>(w) / T(2.0);
  for (int aa = 0; aa < D; aa++) {
    f[aa] = rho * w[1 + aa];
    for (int b = 0; b < D; b++) f[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? p * T(1.0) : p * T(0.0));
    f[aa + D * (D + 1)] = (rho * et + p) * w[1 + aa];
  }
  if constexpr (VISC) {
    T pg[NG], fv[D * NV];
    primitiveGradient<D>(ph, w, u, g, pg);
    viscousRawFlux<D>(ph, w, pg, fv);
    for (int i = 0; i < D * NV; i++) f[i] = f[i] - fv[i];
  }
}

// gb[j] = sum_{a,v} Fb[a + D v] dF_v[a][v] / dg_j for every gradient component j: the transpose of
// the linear map g -> F_v (primitiveGradient then viscousRawFlux at the fixed state u), applied
// in reverse order of those two functions' operations.
template <int D>
__device__ __forceinline__ void viscousGradientVjp(const Phys& ph, const double* u, const double* Fb, double* gb) {
  constexpr int NV = D + 2, NG = NV * D;
  double w[D + 3];
  computational<D>(ph, u, w);
  const double rho = w[0];
  const double t = w[D + 1] / ph.cv;
  const double mu = ph.transport == 2 ? ph.mu * sutherlandRatio(t) : ph.mu;
  const double kap = ph.transport == 2 ? ph.kappa * sutherlandRatio(t) : ph.kappa;
  // fv[a + D (1 + b)] = tau[a + D b];  fv[a + D (D + 1)] = sum_b tau[a + D b] u_b + kap tg_a
  double taub[D * D], deb[D];
  for (int a = 0; a < D; a++) {
    const double fe = Fb[a + D * (D + 1)];
    for (int b = 0; b < D; b++) taub[a + D * b] = Fb[a + D * (1 + b)] + fe * w[1 + b];
    deb[a] = kap * fe / ph.cv;  // tg_a = de_a / cv
  }
  // tau = mu (vg + vg^T) - 2/3 mu tr(vg) I
  double trb = 0.0;
  for (int a = 0; a < D; a++) trb += taub[a + D * a];
  double vgb[D * D];
  for (int b = 0; b < D; b++)
    for (int a = 0; a < D; a++)
      vgb[a + D * b] = mu * (taub[a + D * b] + taub[b + D * a]) - (a == b ? 2.0 / 3.0 * mu * trb : 0.0);
  for (int r = 0; r < NG; r++) gb[r] = 0.0;
  // de_a = (g_E,a - g_rho,a etot) / rho - sum_b vg[a + D b] u_b
  const double etot = u[D + 1] / rho;
  for (int a = 0; a < D; a++) {
    gb[(D + 1) * D + a] += deb[a] / rho;
    gb[a] -= etot * deb[a] / rho;
    for (int b = 0; b < D; b++) vgb[a + D * b] -= deb[a] * w[1 + b];
  }
  // vg[a + D b] = (g_m_b,a - g_rho,a u_b) / rho
  for (int b = 0; b < D; b++)
    for (int a = 0; a < D; a++) {
      gb[(1 + b) * D + a] += vgb[a + D * b] / rho;
      gb[a] -= w[1 + b] * vgb[a + D * b] / rho;
    }
}

// Fb[a + D v] = sum_b fbar[b NV + v] G[a + D b]: the adjoint of the raw flux from that of the
// contracted volume flux out[b NV + v] = sum_a F[a + D v] G[a + D b]
template <int D>
__device__ __forceinline__ void volumeFluxBar(const double* fbar, const double* G, double* Fb) {
  constexpr int NV = D + 2;
  for (int aa = 0; aa < D; aa++)
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int b = 0; b < D; b++) s += fbar[b * NV + v] * G[aa + D * b];
      Fb[aa + D * v] = s;
    }
}

// One part of the volume-flux VJP: dir < NV returns ub[dir] = sum_i Fb_i dF_i/du_dir (one dual
// seed); dir == NV writes gb = -(dF/dg)^T Fb (the viscous flux is linear in g) and returns 0.
template <int D>
__device__ __forceinline__ double volumeFluxVjpDir(const Phys& ph, const double* u, const double* g, const double* Fb,
                                                   int dir, double* gb) {
  constexpr int NV = D + 2, NG = NV * D;
  if (dir < NV) {
    using DT = Dual<1>;
    DT uD[NV], gD[NG], fD[D * NV];
    for (int k = 0; k < NV; k++) uD[k] = k == dir ? seed<1>(u[k], 0) : DT(u[k]);
    for (int r = 0; r < NG; r++) gD[r] = DT(g[r]);
    rawVolumeFlux<D, DT, true>(ph, uD, gD, fD);
    double s = 0.0;
    for (int i = 0; i < D * NV; i++) s += Fb[i] * fD[i].d[0];
    return s;
  }
  viscousGradientVjp<D>(ph, u, Fb, gb);
  for (int r = 0; r < NG; r++) gb[r] = -gb[r];
  return 0.0;
}

// One component of the state part of one side of the face-flux VJP (a dual number seeded on
// component k of side SIDE): sum_v tb[v] d t_v / du_SIDE[k] for t = H_c(ul, ur; n) -
// (F_v(ul, gl) + F_v(ur, gr)) . n / 2. One seed per evaluation keeps the flux's registers those of
// a single-direction dual number (the NV-direction form spilled to scratch in three dimensions).
template <int D, int SIDE>
__device__ __forceinline__ double faceStateVjpDir(const Phys& ph, const double* n, const double* tb, const double* ul,
                                                  const double* ur, const double* gs, int k) {
  constexpr int NV = D + 2, NG = NV * D;
  using DT = Dual<1>;
  DT nD[D], uL[NV], uR[NV], wL[D + 3], wR[D + 3], h[NV];
  for (int d = 0; d < D; d++) nD[d] = DT(n[d]);
  for (int c = 0; c < NV; c++) {
    uL[c] = (SIDE == 0 && c == k) ? seed<1>(ul[c], 0) : DT(ul[c]);
    uR[c] = (SIDE == 1 && c == k) ? seed<1>(ur[c], 0) : DT(ur[c]);
  }
  computational<D>(ph, uL, wL);
  computational<D>(ph, uR, wR);
  convectiveFlux<D>(ph, nD, uL, wL, uR, wR, h);
  DT gD[NG], pg[NG], fv[D * NV];
  for (int r = 0; r < NG; r++) gD[r] = DT(gs[r]);
  if (SIDE == 0) {
    primitiveGradient<D>(ph, wL, uL, gD, pg);
    viscousRawFlux<D>(ph, wL, pg, fv);
  } else {
    primitiveGradient<D>(ph, wR, uR, gD, pg);
    viscousRawFlux<D>(ph, wR, pg, fv);
  }
  double s = 0.0;
  for (int v = 0; v < NV; v++) {
    double dv = 0.0;
    for (int d = 0; d < D; d++) dv += fv[d + D * v].d[0] * n[d];
    s += tb[v] * (h[v].d[0] - dv / 2.0);
  }
  return s;
}

// ---- kernel 1: volume part ----------------------------------------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ, int VISC>
__global__ __launch_bounds__(kThreads) void nsAdjVolumeMfma(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  using LD = GradLds<D, NB, NQ, NF, NFQ>;
  constexpr int NV = LD::NV, NG = LD::NG, EPB = LD::EPB, NAQ = LD::NAQ;
  constexpr int NBk = LD::NBk, NBm = LD::NBm, NQm = LD::NQm, NQk = r4(NQ), NAQm = LD::NAQm;
  constexpr int NBr = LD::NBr, NDQm = r16(NQ * D);
  constexpr int AJV = LD::AJ > 2 * NQm * 16 ? LD::AJ : 2 * NQm * 16;  // grad u_q and grad-bar_q
  static_assert(2 * NBr >= NQm, "u-bar_q reuses the lifting work area");
  __shared__ double sUq[NQm * 16];
  __shared__ double sAJ[AJV];
  __shared__ double sRR[2 * NBr * 16];
  __shared__ double sFb[NDQm * 16];
  const MOps& op = g.op;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int e0 = blockIdx.x * EPB;
  gradLift<double, D, NB, NQ, NF, NFQ, VISC>(g, ph, a, e0, sUq, sAJ, sRR);
  double* sGq = sAJ;
  double* sLam = sRR;
  double* sMu = sRR + NBr * 16;
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sLam[idx] = (b < NB && cc < EPB * NV && e < g.E) ? j.lam[(static_cast<int64_t>(e) * NV + v) * NB + b] : 0.0;
  }
  __syncthreads();
  const int le_c = c / NV, v_c = c % NV, e_c = e0 + le_c;
  const bool col_ok = c < EPB * NV && e_c < g.E;
  // mu = Mref^-1 lam / |J| (adjoint of M^-1 R) or lam
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    if (j.apply_minv) {
      tileAcc<double, NBm>(acc, op.minv, m0, sLam, 0, NBk, lane);
      const double idj = col_ok ? 1.0 / g.detj[e_c] : 0.0;
#pragma unroll
      for (int i = 0; i < 4; i++) acc[i] *= idj;
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) acc[i] = sLam[(m0 + Mfma<double>::row(rr, i)) * 16 + c];
    }
    tileStore<double>(acc, sMu, m0, lane);
  }
  __syncthreads();
  // H-bar = Vf mu at the face slots -> HBM
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NAQm>(acc, op.trace, m0, sMu, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + Mfma<double>::row(rr, i);
      if (col_ok && sl < NAQ) j.Hb[(static_cast<int64_t>(e_c) * NV + v_c) * NAQ + sl] = acc[i];
    }
  }
  // F-bar = Phi_grad mu at the (point, direction) rows
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NDQm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NDQm>(acc, j.op.gradT, m0, sMu, 0, NBk, lane);
    tileStore<double>(acc, sFb, m0, lane);
  }
  __syncthreads();
  // pointwise VJP of the volume flux, one thread per (part, element, point) -- part 0 the state
  // (one reverse-mode pass), part 1 the gradient: u-bar_q into sUb (the free sRR area) and
  // grad-bar_q into sGb (behind grad u_q in sAJ)
  double* sUb = sRR;
  double* sGb = sAJ + NQm * 16;
  for (int idx = tid; idx < 2 * EPB * NQ; idx += kThreads) {
    const int part = idx / (EPB * NQ), rem = idx % (EPB * NQ), le = rem / NQ, q = rem % NQ, e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], gq[NG], G[D * D], fb[D * NV], Fb[D * NV];
    for (int v = 0; v < NV; v++) u[v] = sUq[q * 16 + le * NV + v];
    for (int r = 0; r < NG; r++) gq[r] = sGq[q * 16 + le * NG + r];
    const double jw = g.detj[e] * g.wq[q];
    for (int i = 0; i < D * D; i++) G[i] = g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
    for (int b = 0; b < D; b++)
      for (int v = 0; v < NV; v++) fb[b * NV + v] = sFb[(q * D + b) * 16 + le * NV + v];
    volumeFluxBar<D>(fb, G, Fb);
    if (part == 0) {  // the state (reverse mode)
      double ub[NV];
      volumeFluxStateVjp<D>(ph, u, gq, Fb, ub);
      for (int v = 0; v < NV; v++) sUb[q * 16 + le * NV + v] = ub[v];
    } else {  // the gradient (linear)
      double gbv[NG];
      volumeFluxVjpDir<D>(ph, u, gq, Fb, NV, gbv);
      for (int r = 0; r < NG; r++) sGb[q * 16 + le * NG + r] = gbv[r];
    }
  }
  __syncthreads();
  // R-bar_t (volume part) = (Mref^-1 V^T)^T grad-bar / |J| -> HBM
  const int gle = c / NG, ge = e0 + gle;
  const bool gok = c < EPB * NG && ge < g.E;
  const double idj = gok ? 1.0 / g.detj[ge] : 0.0;
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NBm>(acc, j.op.pinterpT, m0, sGb, 0, NQk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int b = m0 + Mfma<double>::row(rr, i);
      if (gok && b < NB) j.RtV[(static_cast<int64_t>(ge) * NG + c % NG) * NB + b] = acc[i] * idj;
    }
  }
  // U-bar (partial) = V^T u-bar_q -> HBM
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NBm>(acc, j.op.interpT, m0, sUb, 0, NQk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int b = m0 + Mfma<double>::row(rr, i);
      if (col_ok && b < NB) j.Ua[(static_cast<int64_t>(e_c) * NV + v_c) * NB + b] = acc[i];
    }
  }
}

// ---- kernel 2: face points --------------------------------------------------------------------
template <int D, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsAdjFaceKernel(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  // one thread per (part, side, face point): parts 0 .. NV-1 are the adjoints of one component of
  // that side's state (a dual number seeded on it), part NV that of its gradient (the viscous flux
  // is linear in the gradient); part and side are uniform over consecutive face points
  const int64_t npts = static_cast<int64_t>(g.F) * NFQ;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= npts * 2 * (NV + 1)) return;
  const int part = static_cast<int>(idx / (2 * npts));
  const int side = static_cast<int>((idx / npts) & 1);
  const int64_t pt = idx % npts;
  const int64_t f = pt / NFQ;
  const int jq = static_cast<int>(pt % NFQ);
  const int k = g.fseq[f * NFQ + jq];
  const int64_t L = g.fpar[2 * f], R = g.fpar[2 * f + 1];
  const int slL = g.floc[2 * f] * NFQ + jq, slR = g.floc[2 * f + 1] * NFQ + k;
  double ul[NV], ur[NV], gs[NG], tb[NV], n[D];
  const double jw = g.fjw[f * NFQ + jq];
  for (int v = 0; v < NV; v++) {
    ul[v] = a.Uf_in[(L * NV + v) * NAQ + slL];
    ur[v] = a.Uf_in[(R * NV + v) * NAQ + slR];
    tb[v] = (j.Hb[(R * NV + v) * NAQ + slR] - j.Hb[(L * NV + v) * NAQ + slL]) * jw;
  }
  const int64_t E = side == 0 ? L : R;
  const int sl = side == 0 ? slL : slR;
  for (int d = 0; d < D; d++) n[d] = g.fnormal[(f * NFQ + jq) * D + d];
  if (part < NV) {
    for (int r = 0; r < NG; r++) gs[r] = a.Gf[(E * NG + r) * NAQ + sl];
    const double ub = side == 0 ? faceStateVjpDir<D, 0>(ph, n, tb, ul, ur, gs, part)
                                : faceStateVjpDir<D, 1>(ph, n, tb, ul, ur, gs, part);
    j.UFf[(E * NV + part) * NAQ + sl] = ub;
  } else {
    // -(1/2) sum_{v,a} tb[v] n_a dF_v[a][v]/dg of this side
    double Fb[D * NV], gb[NG];
    for (int v = 0; v < NV; v++)
      for (int d = 0; d < D; d++) Fb[d + D * v] = -0.5 * tb[v] * n[d];
    viscousGradientVjp<D>(ph, side == 0 ? ul : ur, Fb, gb);
    for (int r = 0; r < NG; r++) j.GFb[(E * NG + r) * NAQ + sl] = gb[r];
  }
}

// ---- kernel 2, reverse mode (Central, Lax-Friedrichs, HLLC, Roe): one thread per face point ----
// The numerical flux and both sides' viscous fluxes are differentiated in one reverse pass
// (physics_adjoint.hpp) instead of one dual-number pass per state component and side.
template <int D, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsAdjFaceRevKernel(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  const int64_t pt = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (pt >= static_cast<int64_t>(g.F) * NFQ) return;
  const int64_t f = pt / NFQ;
  const int jq = static_cast<int>(pt % NFQ);
  const int k = g.fseq[f * NFQ + jq];
  const int64_t L = g.fpar[2 * f], R = g.fpar[2 * f + 1];
  const int slL = g.floc[2 * f] * NFQ + jq, slR = g.floc[2 * f + 1] * NFQ + k;
  double ul[NV], ur[NV], tb[NV], n[D];
  const double jw = g.fjw[f * NFQ + jq];
  for (int v = 0; v < NV; v++) {
    ul[v] = a.Uf_in[(L * NV + v) * NAQ + slL];
    ur[v] = a.Uf_in[(R * NV + v) * NAQ + slR];
    tb[v] = (j.Hb[(R * NV + v) * NAQ + slR] - j.Hb[(L * NV + v) * NAQ + slL]) * jw;
  }
  for (int d = 0; d < D; d++) n[d] = g.fnormal[(f * NFQ + jq) * D + d];
  double wl[D + 3], wr[D + 3], wbl[D + 3], wbr[D + 3], ubl[NV], ubr[NV];
  computational<D>(ph, ul, wl);
  computational<D>(ph, ur, wr);
  for (int i = 0; i < D + 3; i++) wbl[i] = wbr[i] = 0.0;
  for (int v = 0; v < NV; v++) ubl[v] = ubr[v] = 0.0;
  convectiveFluxVjp<D>(ph, n, ul, wl, ur, wr, tb, wbl, wbr, ubl, ubr);
  // t = H_c - (F_v(u_L, g_L) + F_v(u_R, g_R)) . n / 2: each side's F_v has the adjoint -tb n / 2
  double Fvb[D * NV];
  for (int v = 0; v < NV; v++)
    for (int d = 0; d < D; d++) Fvb[d + D * v] = -0.5 * tb[v] * n[d];
#pragma unroll 1
  for (int sd = 0; sd < 2; sd++) {
    const int64_t E = sd ? R : L;
    const int sl = sd ? slR : slL;
    double gs[NG], gb[NG], x[NV];
    for (int r = 0; r < NG; r++) gs[r] = a.Gf[(E * NG + r) * NAQ + sl];
    if (sd == 0) {
      viscousStateVjp<D>(ph, ul, wl, gs, Fvb, wbl);
      viscousGradientVjp<D>(ph, ul, Fvb, gb);
      computationalVjp<D>(ph, ul, wl, wbl, x);
      for (int v = 0; v < NV; v++) x[v] += ubl[v];
    } else {
      viscousStateVjp<D>(ph, ur, wr, gs, Fvb, wbr);
      viscousGradientVjp<D>(ph, ur, Fvb, gb);
      computationalVjp<D>(ph, ur, wr, wbr, x);
      for (int v = 0; v < NV; v++) x[v] += ubr[v];
    }
    for (int r = 0; r < NG; r++) j.GFb[(E * NG + r) * NAQ + sl] = gb[r];
    for (int v = 0; v < NV; v++) j.UFf[(E * NV + v) * NAQ + sl] = x[v];
  }
}

// ---- kernel 3: lifting terms --------------------------------------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ, int VISC>
__global__ __launch_bounds__(kThreads) void nsAdjLiftMfma(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  using LD = GradLds<D, NB, NQ, NF, NFQ>;
  constexpr int NV = LD::NV, NG = LD::NG, EPB = LD::EPB, NAQ = LD::NAQ;
  constexpr int NBk = LD::NBk, NBm = LD::NBm, NQm = LD::NQm, NQk = r4(NQ), NAQk = LD::NAQk, NAQm = LD::NAQm;
  constexpr int NBr = LD::NBr, NDQm = r16(NQ * D);
  __shared__ double sRt[NBr * 16];
  __shared__ double sRw[NBr * 16];
  __shared__ double sGb[NAQm * 16];
  __shared__ double sAb[NAQm * 16];
  __shared__ double sJb[NAQm * 16];
  __shared__ double sQb[NDQm * 16];
  __shared__ double sUQ[NQm * 16];
  const MOps& op = g.op;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int e0 = blockIdx.x * EPB;
  zeroLds(sUQ, NQm * 16, tid);
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NG, r = cc % NG, e = e0 + le;
    sRt[idx] = (b < NB && cc < EPB * NG && e < g.E) ? j.RtV[(static_cast<int64_t>(e) * NG + r) * NB + b] : 0.0;
  }
  for (int idx = tid; idx < NAQm * 16; idx += kThreads) {
    const int sl = idx / 16, cc = idx % 16, le = cc / NG, r = cc % NG, e = e0 + le;
    sGb[idx] = (sl < NAQ && cc < EPB * NG && e < g.E) ? j.GFb[(static_cast<int64_t>(e) * NG + r) * NAQ + sl] : 0.0;
  }
  __syncthreads();
  const int gle = c / NG, ge = e0 + gle;
  const bool gok = c < EPB * NG && ge < g.E;
  const double idj = gok ? 1.0 / g.detj[ge] : 0.0;
  // face-gradient adjoint -> R-bar_w (BR2) or R-bar_t (BR1); each lane updates its own entries
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NBm>(acc, j.op.ptraceT, m0, sGb, 0, NAQk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int b = m0 + Mfma<double>::row(rr, i);
      const double t = sRt[b * 16 + c];
      if (VISC == 2) {
        sRw[b * 16 + c] = t + acc[i] * idj;
      } else {
        sRt[b * 16 + c] = t + acc[i] * idj;
        sRw[b * 16 + c] = t + acc[i] * idj;
      }
    }
  }
  __syncthreads();
  // J-bar = Vf R-bar_t (+ Lf^T G-bar / |J|), A-bar = Vf R-bar_w
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    acc_t<double> aj = zeroAcc<double>(), aa = zeroAcc<double>();
    tileAcc<double, NAQm>(aj, op.trace, m0, sRt, 0, NBk, lane);
    if (VISC == 2) {
      acc_t<double> al = zeroAcc<double>();
      const int s_lo = m0 / NFQ;
      const int s_hi = (m0 + 15 < NAQ ? m0 + 15 : NAQ - 1) / NFQ;
      const int k0 = (s_lo * NFQ) / 4 * 4;
      const int k1 = r4((s_hi + 1) * NFQ);
      tileAcc<double, NAQm>(al, j.op.liftT, m0, sGb, k0, k1 < NAQk ? k1 : NAQk, lane);
#pragma unroll
      for (int i = 0; i < 4; i++) aj[i] += al[i] * idj;
    }
    tileStore<double>(aj, sJb, m0, lane);
    tileAcc<double, NAQm>(aa, op.trace, m0, sRw, 0, NBk, lane);
    tileStore<double>(aa, sAb, m0, lane);
  }
  // Q-bar = -Phi_grad R-bar_w
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NDQm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NDQm>(acc, j.op.gradT, m0, sRw, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = -acc[i];
    tileStore<double>(acc, sQb, m0, lane);
  }
  __syncthreads();
  // Q = u_q (x) (J^T)^-1 |J| w: u-bar_q[k] = sum_{a,b} Q-bar[(q,b)][(k,a)] G[a][b]
  for (int idx = tid; idx < EPB * NQ * NV; idx += kThreads) {
    const int le = idx / (NQ * NV), rem = idx % (NQ * NV), q = rem / NV, kv = rem % NV, e = e0 + le;
    if (e >= g.E) continue;
    const double jw = g.detj[e] * g.wq[q];
    double s = 0.0;
    for (int b = 0; b < D; b++)
      for (int aa = 0; aa < D; aa++)
        s += sQb[(q * D + b) * 16 + le * NG + kv * D + aa] * (g.ginv[static_cast<int64_t>(e) * D * D + aa + D * b] * jw);
    sUQ[q * 16 + le * NV + kv] = s;
  }
  // A and J face terms -> adjoints of the own trace (Pown) and of the neighbour's trace (Qnb)
  for (int idx = tid; idx < EPB * NAQ * NV; idx += kThreads) {
    const int le = idx / (NAQ * NV), rem = idx % (NAQ * NV), sl = rem / NV, kv = rem % NV, e = e0 + le;
    if (e >= g.E) continue;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const double jw = g.fjw[fp];
    double ca = 0.0, cj = 0.0;
    for (int d = 0; d < D; d++) {
      const double n = g.fnormal[fp * D + d];
      ca += sAb[sl * 16 + le * NG + kv * D + d] * n;
      cj += sJb[sl * 16 + le * NG + kv * D + d] * n;
    }
    ca = (side == 0 ? ca : -ca) * jw / 2.0;
    cj = cj * jw / 2.0;
    const double sg = side == 0 ? -1.0 : 1.0;
    const int64_t o = (static_cast<int64_t>(e) * NV + kv) * NAQ + sl;
    j.Pown[o] = ca + sg * cj;
    j.Qnb[o] = ca - sg * cj;
  }
  __syncthreads();
  // U-bar (partial) += V^T u-bar_q
  const int le_c = c / NV, v_c = c % NV, e_c = e0 + le_c;
  const bool col_ok = c < EPB * NV && e_c < g.E;
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NBm>(acc, j.op.interpT, m0, sUQ, 0, NQk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int b = m0 + Mfma<double>::row(rr, i);
      if (col_ok && b < NB) j.Ua[(static_cast<int64_t>(e_c) * NV + v_c) * NB + b] += acc[i];
    }
  }
}

// ---- kernel 4: gather the trace adjoints and apply Vf^T ---------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsAdjGatherMfma(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  constexpr int NV = D + 2, NAQ = NF * NFQ, EPB = 16 / NV;
  constexpr int NBm = r16(NB), NAQk = r4(NAQ);
  __shared__ double sUf[NAQk * 16];
  const MOps& op = g.op;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < NAQk * 16; idx += kThreads) {
    const int sl = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    double x = 0.0;
    if (sl < NAQ && cc < EPB * NV && e < g.E) {
      const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
      const int nbs = g.snb[es];
      const int64_t o = (static_cast<int64_t>(e) * NV + v) * NAQ + sl;
      x = j.UFf[o] + j.Pown[o] + j.Qnb[(static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ];
    }
    sUf[idx] = x;
  }
  __syncthreads();
  const int le_c = c / NV, v_c = c % NV, e_c = e0 + le_c;
  const bool col_ok = c < EPB * NV && e_c < g.E;
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    tileAcc<double, NBm>(acc, op.face, m0, sUf, 0, NAQk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int b = m0 + Mfma<double>::row(rr, i);
      if (col_ok && b < NB) {
        const int64_t o = (static_cast<int64_t>(e_c) * NV + v_c) * NB + b;
        const double vjp = j.Ua[o] + acc[i];
        if (!j.rk) {
          j.out[o] = vjp;
        } else {
          const double l = j.lam[o];
          const double ac = j.rk_first ? j.c0 * l : j.c0 * l + j.acc[o];
          const double ln = j.c2dt * vjp + j.c1 * l;
          if (j.rk_last) {
            j.lam[o] = ac + ln;
          } else {
            j.acc[o] = ac;
            j.lam[o] = ln;
          }
        }
      }
    }
  }
}

}  // namespace sdg

#endif  // SDG_DG_ADJOINT_HPP_
