// dg_mfma.hpp -- MFMA (matrix-core) versions of the stage kernels for elements whose basis is large
// enough to fill 16x16 tiles (hexahedra P >= 2, quadrangles P >= 3).
//
// Every dense contraction of the stage is cast as Out[m][n] = sum_k Op[m][k] X[k][n] where
//   Op  is a shared operator (basis values, basis gradients, face traces, mass inverse, face lift),
//       stored zero-padded and transposed as T[k][m] so that the A fragment of
//       v_mfma_f64_16x16x4f64 (lane l: A[l%16][l/16]) is one coalesced 4 x 16 read;
//   X   is per-element data staged in LDS as [k][16]: the 16 columns are (element, variable) or
//       (element, gradient component) pairs, so one 16x16 tile covers a whole element's panel;
//   Out accumulates in registers (lane l holds rows l/16 + 4 i, column l%16) and is either
//       written to LDS as the next contraction's X, or straight to HBM.
// Pointwise physics (fluxes, conversions) stays on the VALU.
#ifndef SDG_DG_MFMA_HPP_
#define SDG_DG_MFMA_HPP_

#include <hip/hip_runtime.h>

#include "dg_kernels.hpp"
#include "physics.hpp"

namespace sdg {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// acc += T[k][m0 + 0..16) * X[k][0..16) for k in [k0, k1), with X in LDS ([k][16])
template <int MP>
__device__ __forceinline__ void tileAcc(d4& acc, const double* __restrict__ T, int m0, const double* X, int k0, int k1,
                                        int lane) {
  const int c = lane & 15, kr = lane >> 4;
#pragma unroll 8
  for (int k = k0; k < k1; k += 4) acc = mfma64(T[(k + kr) * MP + m0 + c], X[(k + kr) * 16 + c], acc);
}

// Y[m0 + r][c] = acc (Y in LDS, [M][16])
__device__ __forceinline__ void tileStore(const d4& acc, double* Y, int m0, int lane) {
  const int c = lane & 15, r = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; i++) Y[(m0 + r + 4 * i) * 16 + c] = acc[i];
}

__device__ __forceinline__ void zeroLds(double* p, int n, int tid) {
  for (int i = tid; i < n; i += kThreads) p[i] = 0.0;
}

// ---- shared tail: residual -> (monitor) -> update -> next traces --------------------------------
// sR: [NBm][16] residual panel (row basis, column le*NV+v), sU: [NBk][16] current state panel.
template <int D, int NB, int NQ, int NF, int NFQ, int EPB>
__device__ __forceinline__ void mfmaStageTail(const Geo& g, const StageArgs& a, int e0, double* sR,
                                              double* sU, double* sRed) {
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQm = r16(NAQ);
  const MOps& op = g.op;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int le = c / NV, v = c % NV;
  const int e = e0 + le;
  const bool col_ok = c < EPB * NV && e < g.E;
  if (a.last_stage) {  // residual monitor: mean_q |R V^T| (TimeIntegration.cpp:279-298)
    double part = 0.0;
    for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      tileAcc<NQm>(acc, op.interp, m0, sR, 0, NBk, lane);
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (m0 + rr + 4 * i < NQ) part += fabs(acc[i]);
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    if (lane < 16) sRed[wave * 16 + lane] = part;
    __syncthreads();
    if (tid < EPB * NV && e0 + tid / NV < g.E) {
      const double s = sRed[tid] + sRed[16 + tid] + sRed[32 + tid] + sRed[48 + tid];
      a.mon[static_cast<int64_t>(e0 + tid / NV) * NV + tid % NV] = s / NQ;
    }
  }
  // update: U <- c1 U + c0 Un + c2dt (R Mref^-1) / |J|
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NBm>(acc, op.minv, m0, sR, 0, NBk, lane);
    const double idj = col_ok ? g.detj[e] : 1.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int bi = m0 + rr + 4 * i;
      if (col_ok && bi < NB) {
        const int64_t gi = (static_cast<int64_t>(e) * NV + v) * NB + bi;
        double u = sU[bi * 16 + c] * a.c1;
        u += a.c0 * a.Un[gi];
        u += a.c2dt * (acc[i] / idj);
        a.U[gi] = u;
        if (a.last_stage) a.Un[gi] = u;
        acc[i] = u;
      }
    }
    // rows of this tile are read and written by this wave only (the contraction above reads sR),
    // so no barrier is needed here; the traces below read sU after the barrier that follows
    if (m0 < NBk) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int bi = m0 + rr + 4 * i;
        if (bi < NBk) sU[bi * 16 + c] = (col_ok && bi < NB) ? acc[i] : 0.0;
      }
    }
  }
  __syncthreads();
  // traces of the new state
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NAQm>(acc, op.trace, m0, sU, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + rr + 4 * i;
      if (col_ok && sl < NAQ) a.Uf_out[(static_cast<int64_t>(e) * NV + v) * NAQ + sl] = acc[i];
    }
  }
}

// ---- Euler stage (MFMA) -------------------------------------------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void eulerStageMfma(Geo g, Phys ph, StageArgs a) {
  const MOps& op = g.op;
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int EPB = 16 / NV;
  constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQk = r4(NAQ), NDQk = r4(NQ * D);
  constexpr int NBr = NBk > NBm ? NBk : NBm;
  __shared__ double sU[NBr * 16];
  __shared__ double sUq[NQm * 16];
  __shared__ double sF[NDQk * 16];
  __shared__ double sH[NAQk * 16];
  __shared__ double sR[NBr * 16];
  __shared__ double sRed[64];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int e0 = blockIdx.x * EPB;
  zeroLds(sF, NDQk * 16, tid);
  zeroLds(sH, NAQk * 16, tid);
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, c = idx % 16, le = c / NV, v = c % NV, e = e0 + le;
    sU[idx] = (b < NB && c < EPB * NV && e < g.E) ? a.U[(static_cast<int64_t>(e) * NV + v) * NB + b] : 0.0;
  }
  __syncthreads();
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NQm>(acc, op.interp, m0, sU, 0, NBk, lane);
    tileStore(acc, sUq, m0, lane);
  }
  __syncthreads();
  // volume flux at the points: (F^T (J^T)^-1 |J| w) -> sF[(q D + b)][le NV + v]
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ, e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) u[v] = sUq[q * 16 + le * NV + v];
    computational<D>(ph, u, w);
    const double rho = w[0], p = w[D + 2];
    const double et = w[D + 1] + velocitySquared<D>(w) / 2.0;
    double fc[D * NV];
    for (int aa = 0; aa < D; aa++) {
      fc[aa] = rho * w[1 + aa];
      for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? p * 1.0 : p * 0.0);
      fc[aa + D * (D + 1)] = (rho * et + p) * w[1 + aa];
    }
    const double jw = g.detj[e] * g.wq[q];
    double G[D * D];
    for (int i = 0; i < D * D; i++) G[i] = g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
    for (int v = 0; v < NV; v++)
      for (int b = 0; b < D; b++) {
        double s = 0.0;
        for (int aa = 0; aa < D; aa++) s += fc[aa + D * v] * G[aa + D * b];
        sF[(q * D + b) * 16 + le * NV + v] = s;
      }
  }
  // face fluxes -> sH[slot][le NV + v] (negated: the face term enters the residual with -)
  for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
    const int le = idx / NAQ, sl = idx % NAQ, e = e0 + le;
    if (e >= g.E) continue;
    const int s = sl / NFQ, j = sl % NFQ;
    int f, p, side;
    int64_t nbslot;
    slotRoles<NF, NFQ>(g, e, s, j, f, p, nbslot, side);
    double uo[NV], un[NV];
    const int64_t own = static_cast<int64_t>(e) * NV * NAQ + sl;
    const int64_t nbe = (nbslot / NAQ) * NV * NAQ + nbslot % NAQ;
    for (int v = 0; v < NV; v++) {
      uo[v] = a.Uf_in[own + v * NAQ];
      un[v] = a.Uf_in[nbe + v * NAQ];
    }
    const double* ul = side == 0 ? uo : un;
    const double* ur = side == 0 ? un : uo;
    double wl[D + 3], wr[D + 3], n[D], h[NV];
    computational<D>(ph, ul, wl);
    computational<D>(ph, ur, wr);
    for (int d = 0; d < D; d++) n[d] = g.fnormal[(static_cast<int64_t>(f) * NFQ + p) * D + d];
    convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
    const double jw = g.fjw[static_cast<int64_t>(f) * NFQ + p];
    for (int v = 0; v < NV; v++) {
      const double t = h[v] * jw;
      sH[sl * 16 + le * NV + v] = side == 0 ? -t : t;
    }
  }
  __syncthreads();
  // residual R = F Phi_grad - H Phi_face
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NBm>(acc, op.grad, m0, sF, 0, NDQk, lane);
    tileAcc<NBm>(acc, op.face, m0, sH, 0, NAQk, lane);
    tileStore(acc, sR, m0, lane);
  }
  for (int idx = NBm * 16 + tid; idx < NBr * 16; idx += kThreads) sR[idx] = 0.0;
  __syncthreads();
  if (a.mode == 1) {
    for (int idx = tid; idx < NB * 16; idx += kThreads) {
      const int b = idx / 16, c = idx % 16, le = c / NV, v = c % NV, e = e0 + le;
      if (c < EPB * NV && e < g.E) a.Rout[(static_cast<int64_t>(e) * NV + v) * NB + b] = sR[idx];
    }
    return;
  }
  mfmaStageTail<D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sR, sU, sRed);
}

// ---- Navier-Stokes kernel 1 (MFMA): gradient lifting + volume flux -------------------------------
// One element per workgroup in 3-D (15 gradient columns), two in 2-D (2 x 8 columns).
// The lifted coefficients w = R Mref^-1 / |J| are only ever evaluated at points, so Mref^-1 is folded
// into the evaluation operators: grad U_q = R_t (Mref^-1 V^T) / |J| and the BR2 face trace
// R_w (Mref^-1 Vf^T) / |J| + J Lf / |J| (BR1: R_t (Mref^-1 Vf^T) / |J|).
template <int D, int NB, int NQ, int NF, int NFQ, int VISC>
__global__ __launch_bounds__(kThreads) void nsGradMfma(Geo g, Phys ph, StageArgs a) {
  const MOps& op = g.op;
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  constexpr int EPB = 16 / NG;
  constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQk = r4(NAQ), NAQm = r16(NAQ), NDQk = r4(NQ * D);
  constexpr int NBr = NBk > NBm ? NBk : NBm;
  constexpr int AJ = 2 * NAQk * 16 > NQm * 16 ? 2 * NAQk * 16 : NQm * 16;
  __shared__ double sUq[NQm * 16];    // state at volume points
  __shared__ double sAJ[AJ];          // face terms A | J ([slot][le*NG + r]); later the gradient at points
  __shared__ double sRR[2 * NBr * 16];  // state panel; then R_w | R_t
  double* sU = sRR;
  double* sA = sAJ;
  double* sJ = sAJ + NAQk * 16;
  double* sRw = sRR;
  double* sRt = sRR + NBr * 16;
  double* sGq = sAJ;
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int e0 = blockIdx.x * EPB;
  zeroLds(sAJ, AJ, tid);
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sU[idx] = (b < NB && cc < EPB * NV && e < g.E) ? a.U[(static_cast<int64_t>(e) * NV + v) * NB + b] : 0.0;
  }
  __syncthreads();
  // face terms: (u- + u+)/2 (x) n |J_f| w (signed by side) and (u+ - u-)/2 (x) n |J_f| w
  for (int idx = tid; idx < EPB * NAQ * NV; idx += kThreads) {
    const int le = idx / (NAQ * NV), rem = idx % (NAQ * NV), sl = rem / NV, v = rem % NV, e = e0 + le;
    if (e >= g.E) continue;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es], fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const double uo = a.Uf_in[(static_cast<int64_t>(e) * NV + v) * NAQ + sl];
    const double un = a.Uf_in[(static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ];
    const double ul = side == 0 ? uo : un, ur = side == 0 ? un : uo;
    const double jw = g.fjw[fp];
    for (int d = 0; d < D; d++) {
      const double n = g.fnormal[fp * D + d];
      const double avg = (n * (ul + ur) / 2.0) * jw;
      sA[sl * 16 + le * NG + v * D + d] = side == 0 ? avg : -avg;
      sJ[sl * 16 + le * NG + v * D + d] = (n * (ur - ul) / 2.0) * jw;
    }
  }
  // state at the volume points
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NQm>(acc, op.interp, m0, sU, 0, NBk, lane);
    tileStore(acc, sUq, m0, lane);
  }
  __syncthreads();
  // gradient residuals: R_w = A Phi_f - Q Phi_grad, R_t = R_w + J Phi_f, with
  // Q[(le,k,a)][(q,b)] = u_k(q) (J^T)^-1(a,b) |J| w_q generated in the B fragment
  const int gle = c / NG, ge = e0 + gle;
  const bool gok = c < EPB * NG && ge < g.E;
  const int gec = gok ? ge : 0;
  const double dj = g.detj[gec];
  {
    const int gk = (c % NG) / D, ga = (c % NG) % D;
    double gi[D];
#pragma unroll
    for (int b = 0; b < D; b++) gi[b] = g.ginv[static_cast<int64_t>(gec) * D * D + ga + D * b];
    for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
      d4 aw = {0.0, 0.0, 0.0, 0.0}, aj = {0.0, 0.0, 0.0, 0.0};
      tileAcc<NBm>(aw, op.face, m0, sA, 0, NAQk, lane);
      tileAcc<NBm>(aj, op.face, m0, sJ, 0, NAQk, lane);
#pragma unroll 4
      for (int k = 0; k < NDQk; k += 4) {
        const int kk = k + rr;
        double bq = 0.0;
        if (gok && kk < NQ * D) {
          const int q = kk / D, b = kk % D;
          const double G = gi[b] * (dj * g.wq[q]);
          bq = -(sUq[q * 16 + gle * NV + gk] * G);
        }
        aw = mfma64(op.grad[kk * NBm + m0 + c], bq, aw);
      }
      // the state panel (aliased by R) was last read before the barrier above
      tileStore(aw, sRw, m0, lane);
      d4 at = aw + aj;
      tileStore(at, sRt, m0, lane);
    }
  }
  for (int idx = NBm * 16 + tid; idx < NBr * 16; idx += kThreads) sRw[idx] = sRt[idx] = 0.0;
  __syncthreads();
  const double idj = 1.0 / dj;
  // optional output of the lifted total-gradient coefficients w = R_t Mref^-1 / |J|
  if (a.Wout) {
    for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
      d4 wt = {0.0, 0.0, 0.0, 0.0};
      tileAcc<NBm>(wt, op.minv, m0, sRt, 0, NBk, lane);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int bi = m0 + rr + 4 * i;
        if (gok && bi < NB) a.Wout[(static_cast<int64_t>(ge) * NB + bi) * NG + (c % NG)] = wt[i] * idj;
      }
    }
  }
  // lifted gradient traces at the face points -> HBM
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    if (VISC == 2) {
      tileAcc<NAQm>(acc, op.ptrace, m0, sRw, 0, NBk, lane);
      // block-diagonal BR2 lift: only the faces this row tile touches
      const int s_lo = m0 / NFQ;
      const int s_hi = (m0 + 15 < NAQ ? m0 + 15 : NAQ - 1) / NFQ;
      const int k0 = (s_lo * NFQ) / 4 * 4;
      const int k1 = r4((s_hi + 1) * NFQ);
      tileAcc<NAQm>(acc, op.lift, m0, sJ, k0, k1 < NAQk ? k1 : NAQk, lane);
    } else {
      tileAcc<NAQm>(acc, op.ptrace, m0, sRt, 0, NBk, lane);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + rr + 4 * i;
      if (gok && sl < NAQ) a.Gf[(static_cast<int64_t>(ge) * NG + (c % NG)) * NAQ + sl] = acc[i] * idj;
    }
  }
  __syncthreads();
  // total gradient at the volume points (into the face-term buffer, no longer needed)
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NQm>(acc, op.pinterp, m0, sRt, 0, NBk, lane);
    acc *= idj;
    tileStore(acc, sGq, m0, lane);
  }
  __syncthreads();
  // volume flux (F_c - F_v)^T (J^T)^-1 |J| w at the points -> HBM Fv[e][(q D + b)][v]; the weak
  // divergence F Phi_grad is taken in nsResidualMfma, which packs three elements per 16 columns
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ, e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3], gq[NG], pg[NG], fv[D * NV];
    for (int v = 0; v < NV; v++) u[v] = sUq[q * 16 + le * NV + v];
    for (int r = 0; r < NG; r++) gq[r] = sGq[q * 16 + le * NG + r];
    computational<D>(ph, u, w);
    primitiveGradient<D>(ph, w, u, gq, pg);
    viscousRawFlux<D>(ph, w, pg, fv);
    const double rho = w[0], p = w[D + 2];
    const double et = w[D + 1] + velocitySquared<D>(w) / 2.0;
    double fc[D * NV];
    for (int aa = 0; aa < D; aa++) {
      fc[aa] = rho * w[1 + aa];
      for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? p * 1.0 : p * 0.0);
      fc[aa + D * (D + 1)] = (rho * et + p) * w[1 + aa];
    }
    const double jw = g.detj[e] * g.wq[q];
    double G[D * D];
    for (int i = 0; i < D * D; i++) G[i] = g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
    double* out = a.Fv + (static_cast<int64_t>(e) * NQ * D + q * D) * NV;
    for (int b = 0; b < D; b++)
      for (int v = 0; v < NV; v++) {
        double s = 0.0;
        for (int aa = 0; aa < D; aa++) s += (fc[aa + D * v] - fv[aa + D * v]) * G[aa + D * b];
        out[b * NV + v] = s;
      }
  }
}

// ---- Navier-Stokes face fluxes, one thread per face point ---------------------------------------
// calculateInteriorAdjacencyElementQuadrature (SpatialDiscrete.cpp:633-748): each face point is
// evaluated once with its left/right roles; t = (H_c - H_v) |J_f| w goes to the left parent's slot
// as -t and to the right parent's slot (point seq[j]) as +t, the signs with which the face term
// enters R = R_vol + sum_slots Phi_f s.
template <int D, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsFaceFluxKernel(Geo g, Phys ph, StageArgs a) {
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  if (a.status && a.status[0]) return;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<int64_t>(g.F) * NFQ) return;
  const int64_t f = idx / NFQ;
  const int j = static_cast<int>(idx % NFQ);
  const int k = g.fseq[f * NFQ + j];
  const int64_t L = g.fpar[2 * f], R = g.fpar[2 * f + 1];
  const int slL = g.floc[2 * f] * NFQ + j, slR = g.floc[2 * f + 1] * NFQ + k;
  double ul[NV], ur[NV], gl[NG], gr[NG];
  for (int v = 0; v < NV; v++) {
    ul[v] = a.Uf_in[(L * NV + v) * NAQ + slL];
    ur[v] = a.Uf_in[(R * NV + v) * NAQ + slR];
  }
  for (int r = 0; r < NG; r++) {
    gl[r] = a.Gf[(L * NG + r) * NAQ + slL];
    gr[r] = a.Gf[(R * NG + r) * NAQ + slR];
  }
  double wl[D + 3], wr[D + 3], n[D], h[NV], pgl[NG], pgr[NG], fvl[D * NV], fvr[D * NV];
  computational<D>(ph, ul, wl);
  computational<D>(ph, ur, wr);
  for (int d = 0; d < D; d++) n[d] = g.fnormal[(f * NFQ + j) * D + d];
  convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
  primitiveGradient<D>(ph, wl, ul, gl, pgl);
  primitiveGradient<D>(ph, wr, ur, gr, pgr);
  viscousRawFlux<D>(ph, wl, pgl, fvl);
  viscousRawFlux<D>(ph, wr, pgr, fvr);
  const double jw = g.fjw[f * NFQ + j];
  for (int v = 0; v < NV; v++) {
    double vl = 0.0, vr = 0.0;
    for (int d = 0; d < D; d++) {
      vl += fvl[d + D * v] * n[d];
      vr += fvr[d + D * v] * n[d];
    }
    const double t = (h[v] - (vl + vr) / 2.0) * jw;
    a.Hs[(L * NV + v) * NAQ + slL] = -t;
    a.Hs[(R * NV + v) * NAQ + slR] = t;
  }
}

// ---- Navier-Stokes residual + update + next traces (MFMA), after nsFaceFluxKernel ---------------
// calculateResidual + updateBasisFunctionCoefficient (SpatialDiscrete.cpp:1016-1032,
// TimeIntegration.cpp:181-198)
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsResidualMfma(Geo g, Phys ph, StageArgs a) {
  const MOps& op = g.op;
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int EPB = 16 / NV;
  constexpr int NBk = r4(NB), NBm = r16(NB), NAQk = r4(NAQ);
  constexpr int NBr = NBk > NBm ? NBk : NBm;
  constexpr int NDQ = NQ * D, NDQk = r4(NDQ);
  __shared__ double sU[NBr * 16];
  __shared__ double sH[NAQk * 16];
  __shared__ double sF[NDQk * 16];
  __shared__ double sR[NBr * 16];
  __shared__ double sRed[64];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sU[idx] = (b < NB && cc < EPB * NV && e < g.E) ? a.U[(static_cast<int64_t>(e) * NV + v) * NB + b] : 0.0;
  }
  for (int idx = tid; idx < NAQk * 16; idx += kThreads) {
    const int sl = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sH[idx] = (sl < NAQ && cc < EPB * NV && e < g.E) ? a.Hs[(static_cast<int64_t>(e) * NV + v) * NAQ + sl] : 0.0;
  }
  for (int idx = tid; idx < NDQk * 16; idx += kThreads) {
    const int k = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sF[idx] = (k < NDQ && cc < EPB * NV && e < g.E) ? a.Fv[(static_cast<int64_t>(e) * NDQ + k) * NV + v] : 0.0;
  }
  __syncthreads();
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    tileAcc<NBm>(acc, op.grad, m0, sF, 0, NDQk, lane);
    tileAcc<NBm>(acc, op.face, m0, sH, 0, NAQk, lane);
    tileStore(acc, sR, m0, lane);
  }
  __syncthreads();
  if (a.mode == 1) {
    for (int idx = tid; idx < NB * 16; idx += kThreads) {
      const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
      if (cc < EPB * NV && e < g.E) a.Rout[(static_cast<int64_t>(e) * NV + v) * NB + b] = sR[idx];
    }
    return;
  }
  mfmaStageTail<D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sR, sU, sRed);
}

}  // namespace sdg

#endif  // SDG_DG_MFMA_HPP_
