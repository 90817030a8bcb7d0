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
typedef float f4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step per contraction type. Inputs: lane l holds A[l%16][l/16] and
// B[l/16][l%16] for both types; the accumulator of lane l holds column l%16 and rows
// row(l/16, i), i = 0..3, which differ between the FP64 and FP32 instructions.
template <typename TC>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = d4;
  static __device__ __forceinline__ d4 mma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int rr, int i) { return rr + 4 * i; }
};
template <>
struct Mfma<float> {
  using acc_t = f4;
  static __device__ __forceinline__ f4 mma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int rr, int i) { return 4 * rr + i; }
};
template <typename TC>
using acc_t = typename Mfma<TC>::acc_t;

template <typename TC>
__device__ __forceinline__ acc_t<TC> zeroAcc() {
  acc_t<TC> z = {TC(0), TC(0), TC(0), TC(0)};
  return z;
}

// acc += T[k][m0 + 0..16) * X[k][0..16) for k in [k0, k1), with X in LDS ([k][16])
template <typename TC, int MP>
__device__ __forceinline__ void tileAcc(acc_t<TC>& acc, const TC* __restrict__ T, int m0, const TC* X, int k0, int k1,
                                        int lane) {
  const int c = lane & 15, kr = lane >> 4;
#pragma unroll 8
  for (int k = k0; k < k1; k += 4) acc = Mfma<TC>::mma(T[(k + kr) * MP + m0 + c], X[(k + kr) * 16 + c], acc);
}

// Y[m0 + row][c] = acc (Y in LDS, [M][16])
template <typename TC>
__device__ __forceinline__ void tileStore(const acc_t<TC>& acc, TC* Y, int m0, int lane) {
  const int c = lane & 15, r = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; i++) Y[(m0 + Mfma<TC>::row(r, i)) * 16 + c] = acc[i];
}

template <typename TC>
__device__ __forceinline__ void zeroLds(TC* p, int n, int tid) {
  for (int i = tid; i < n; i += kThreads) p[i] = TC(0);
}

// FP32 deviation of a coefficient from the reference state's (one expression for every producer
// of contraction operands, so traces computed after a stage and from a stored state agree bitwise)
__device__ __forceinline__ float deviationOperand(double u, double uref, double one) {
  return static_cast<float>(fma(-uref, one, u));
}

// Coefficient b of variable v of element e as a contraction operand: the FP64 state itself, or in
// the mixed-precision mode its deviation from the reference state, rounded to FP32.
template <typename TC>
__device__ __forceinline__ TC stateOperand(const StageArgs& a, int64_t e, int v, int b, int NV, int NB) {
  const double u = a.U[(e * NV + v) * NB + b];
  if constexpr (sizeof(TC) == 8)
    return u;
  else
    return deviationOperand(u, a.mix.uref[v], a.mix.one[b]);
}

// Conserved state at a point from a contraction result (FP64 in both modes).
template <typename TC>
__device__ __forceinline__ double pointState(const StageArgs& a, TC x, int v) {
  if constexpr (sizeof(TC) == 8)
    return x;
  else
    return a.mix.uref[v] + static_cast<double>(x);
}

// Pressure entering the momentum flux: p itself, or p - pref in the mixed-precision mode.
template <typename TC>
__device__ __forceinline__ double fluxPressure(const StageArgs& a, double p) {
  if constexpr (sizeof(TC) == 8)
    return p;
  else
    return p - a.mix.pref;
}

// Mixed-precision mode: the numerical flux minus pref n in the momentum and h0 H_rho in the energy.
template <int D>
__device__ __forceinline__ void reduceFaceFlux(const StageArgs& a, const double* n, double* h) {
  for (int d = 0; d < D; d++) h[1 + d] -= a.mix.pref * n[d];
  h[D + 1] -= a.mix.h0 * h[0];
}

// Mixed-precision mode: value x of column (element, variable v) of a residual-derived tile, with
// h0 times the density column of the same element added back to the energy column. All lanes
// must call it (cross-lane exchange).
template <int NV>
__device__ __forceinline__ double restoreEnergy(const StageArgs& a, double x, int v, int lane) {
  const int src = v == NV - 1 ? lane - (NV - 1) : lane;
  const double xr = __shfl(x, src);
  return v == NV - 1 ? x + a.mix.h0 * xr : x;
}

// ---- shared tail: residual -> (monitor) -> update -> next traces --------------------------------
// sR: [NBm][16] residual panel (row basis, column le*NV+v), sU: [NBk][16] state panel (FP64 mode:
// the current state; mixed mode: scratch, the update reads the FP64 state from HBM).
// + the modal source residual of the last momentum component (Rsrc), before the update or the
// residual output; the FP64 kernels only (the mixed mode refuses a source term)
template <typename TC, int D, int NB, int EPB>
__device__ __forceinline__ void addSource(const Geo& g, const StageArgs& a, int e0, TC* sR) {
  if (!a.Rsrc) return;
  constexpr int NV = D + 2;
  for (int idx = threadIdx.x; idx < NB * EPB; idx += kThreads) {
    const int b = idx / EPB, le = idx % EPB, e = e0 + le;
    if (e < g.E) sR[b * 16 + le * NV + D] += static_cast<TC>(a.Rsrc[static_cast<int64_t>(e) * NB + b]);
  }
  __syncthreads();
}

template <typename TC, int D, int NB, int NQ, int NF, int NFQ, int EPB>
__device__ __forceinline__ void mfmaStageTail(const Geo& g, const StageArgs& a, int e0, TC* sR, TC* sU,
                                              double* sRed) {
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQm = r16(NAQ);
  constexpr bool MIX = sizeof(TC) == 4;
  const MOpsT<TC>& op = opsOf<TC>(g);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int le = c / NV, v = c % NV;
  const int e = e0 + le;
  const bool col_ok = c < EPB * NV && e < g.E;
  if (a.last_stage) {  // residual monitor: mean_q |R V^T| (TimeIntegration.cpp:279-298)
    double part = 0.0;
    #pragma unroll 1
    for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
      acc_t<TC> acc = zeroAcc<TC>();
      tileAcc<TC, NQm>(acc, op.rinterp, m0, sR, 0, NBk, lane);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        double x = static_cast<double>(acc[i]);
        if constexpr (MIX) x = restoreEnergy<NV>(a, x, v, lane);
        if (m0 + Mfma<TC>::row(rr, i) < NQ) part += fabs(x);
      }
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
  // update: U <- c1 U + c0 Un + c2dt (R Mref^-1) / |J|, in FP64 in both modes; in the mixed mode
  // the FP32 residual panel is mapped to the hierarchic coefficients by FP64 matrix cores
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<double> acc = zeroAcc<double>();
    if constexpr (MIX) {
      const int kr = lane >> 4;
#pragma unroll 4
      for (int k = 0; k < NBk; k += 4)
        acc = Mfma<double>::mma(g.minv_mixed64[(k + kr) * NBm + m0 + c], static_cast<double>(sR[(k + kr) * 16 + c]), acc);
    } else {
      tileAcc<double, NBm>(acc, op.minv, m0, sR, 0, NBk, lane);
    }
    const double idj = col_ok ? g.detj[e] : 1.0;
    double un[4], rm[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      rm[i] = acc[i];
      if constexpr (MIX) rm[i] = restoreEnergy<NV>(a, rm[i], v, lane);
    }
    // U_n (and, in the mixed mode, U) of the tile's four rows, loaded before any store: the
    // compiler must assume the stores to U may alias them, which would serialise the rows
    double pn[4], pu[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int bi = m0 + Mfma<double>::row(rr, i);
      const int64_t gi = (static_cast<int64_t>(col_ok ? e : 0) * NV + v) * NB + (bi < NB ? bi : 0);
      pn[i] = a.Un[gi];
      pu[i] = MIX ? a.U[gi] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int bi = m0 + Mfma<double>::row(rr, i);
      un[i] = 0.0;
      if (col_ok && bi < NB) {
        const int64_t gi = (static_cast<int64_t>(e) * NV + v) * NB + bi;
        double u;
        if constexpr (MIX)
          u = pu[i] * a.c1;
        else
          u = sU[bi * 16 + c] * a.c1;
        u += a.c0 * pn[i];
        u += a.c2dt * (rm[i] / idj);
        a.U[gi] = u;
        if (a.last_stage) a.Un[gi] = u;
        if (a.rec) {
          const int step = a.status[1] + (a.last_stage ? 1 : 0);
          const int stage = a.last_stage ? 0 : a.rec_stage + 1;
          if (step < a.rec_nsteps) a.rec[(static_cast<int64_t>(step) * a.rec_S + stage) * a.rec_n + gi] = u;
        }
        un[i] = u;
      }
    }
    // rows of this tile are read and written by this wave only (the contraction above reads sR),
    // so no barrier is needed here; the traces below read sU after the barrier that follows
    if (m0 < NBk) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int bi = m0 + Mfma<double>::row(rr, i);
        if (bi < NBk) {
          TC x = TC(0);
          if (col_ok && bi < NB) {
            if constexpr (MIX)
              x = deviationOperand(un[i], a.mix.uref[v], a.mix.one[bi]);
            else
              x = un[i];
          }
          sU[bi * 16 + c] = x;
        }
      }
    }
  }
  __syncthreads();
  // traces of the new state
  TC* Uf_out = pick<TC>(a.Uf_out, a.Uf_out32);
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NAQm>(acc, op.trace, m0, sU, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + Mfma<TC>::row(rr, i);
      if (col_ok && sl < NAQ) {
        const int64_t ti = (static_cast<int64_t>(e) * NV + v) * NAQ + sl;
        Uf_out[ti] = acc[i];
        if (a.rec_t) {
          const int step = a.status[1] + (a.last_stage ? 1 : 0);
          const int stage = a.last_stage ? 0 : a.rec_stage + 1;
          if (step < a.rec_nsteps) a.rec_t[(static_cast<int64_t>(step) * a.rec_S + stage) * a.rec_nt + ti] = acc[i];
        }
      }
    }
  }
}

// volume convective flux minus viscous flux, contracted with (J^T)^-1 |J| w:
// out[b * NV + v] = sum_a (F_c - F_v)[a][v] G[a][b] (ConvectiveFlux.cpp:27-57, ViscousFlux.cpp:58-103)
template <typename TC, int D, bool VISC>
__device__ __forceinline__ void volumeFlux(const StageArgs& a, const double* w, const double* fv, const double* G,
                                           double* out) {
  constexpr int NV = D + 2;
  const double rho = w[0], p = w[D + 2];
  const double et = w[D + 1] + velocitySquared<D>(w) / 2.0;
  double fc[D * NV];
  for (int aa = 0; aa < D; aa++) {
    fc[aa] = rho * w[1 + aa];
    if constexpr (sizeof(TC) == 8) {
      for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? p * 1.0 : p * 0.0);
    } else {
      const double pf = fluxPressure<TC>(a, p);
      for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? pf : 0.0);
    }
    if constexpr (sizeof(TC) == 8)
      fc[aa + D * (D + 1)] = (rho * et + p) * w[1 + aa];
    else
      fc[aa + D * (D + 1)] = (rho * et + p - a.mix.h0 * rho) * w[1 + aa];
  }
  for (int b = 0; b < D; b++)
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int aa = 0; aa < D; aa++) {
        const double f = VISC ? fc[aa + D * v] - fv[aa + D * v] : fc[aa + D * v];
        s += f * G[aa + D * b];
      }
      out[b * NV + v] = s;
    }
}

// ---- Euler stage (MFMA) -------------------------------------------------------------------------
template <typename TC, int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void eulerStageMfma(Geo g, Phys ph, StageArgs a) {
  const MOpsT<TC>& op = opsOf<TC>(g);
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int EPB = 16 / NV;
  constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQk = r4(NAQ), NDQk = r4(NQ * D);
  constexpr int NBr = NBk > NBm ? NBk : NBm;
  __shared__ TC sU[NBr * 16];
  __shared__ TC sUq[NQm * 16];
  __shared__ TC sF[NDQk * 16];
  __shared__ TC sH[NAQk * 16];
  __shared__ TC sR[NBr * 16];
  __shared__ double sRed[64];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int e0 = blockIdx.x * EPB;
  const TC* Uf_in = pick<TC>(a.Uf_in, a.Uf_in32);
  zeroLds(sF, NDQk * 16, tid);
  zeroLds(sH, NAQk * 16, tid);
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, c = idx % 16, le = c / NV, v = c % NV, e = e0 + le;
    sU[idx] = (b < NB && c < EPB * NV && e < g.E) ? stateOperand<TC>(a, e, v, b, NV, NB) : TC(0);
  }
  __syncthreads();
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NQm>(acc, op.interp, m0, sU, 0, NBk, lane);
    tileStore<TC>(acc, sUq, m0, lane);
  }
  __syncthreads();
  // volume flux at the points: (F^T (J^T)^-1 |J| w) -> sF[(q D + b)][le NV + v]
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ, e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) u[v] = pointState<TC>(a, sUq[q * 16 + le * NV + v], v);
    computational<D>(ph, u, w);
    const double jw = g.detj[e] * g.wq[q];
    double G[D * D], out[D * NV];
    for (int i = 0; i < D * D; i++) G[i] = g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
    volumeFlux<TC, D, false>(a, w, nullptr, G, out);
    for (int b = 0; b < D; b++)
      for (int v = 0; v < NV; v++) sF[(q * D + b) * 16 + le * NV + v] = static_cast<TC>(out[b * NV + v]);
  }
  // face fluxes -> sH[slot][le NV + v] (negated: the face term enters the residual with -)
  for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
    const int le = idx / NAQ, sl = idx % NAQ, e = e0 + le;
    if (e >= g.E) continue;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es], fps = g.sfp[es];
    if (nbs < 0) {  // boundary face point
      double ub[NV], t[NV];
      for (int v = 0; v < NV; v++) ub[v] = pointState<TC>(a, Uf_in[static_cast<int64_t>(e) * NV * NAQ + sl + v * NAQ], v);
      boundaryFlux<D, false>(ph, g, -1 - nbs, ub, nullptr, sizeof(TC) == 4, a.mix.pref, a.mix.h0, t);
      for (int v = 0; v < NV; v++) sH[sl * 16 + le * NV + v] = static_cast<TC>(-t[v]);
      continue;
    }
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    double uo[NV], un[NV];
    const int64_t own = static_cast<int64_t>(e) * NV * NAQ + sl;
    const int64_t nbe = static_cast<int64_t>(nbs / NAQ) * NV * NAQ + nbs % NAQ;
    for (int v = 0; v < NV; v++) {
      uo[v] = pointState<TC>(a, Uf_in[own + v * NAQ], v);
      un[v] = pointState<TC>(a, Uf_in[nbe + v * NAQ], v);
    }
    double ul[NV], ur[NV];  // values, not pointers: keeps the arrays in registers
    for (int v = 0; v < NV; v++) {
      ul[v] = side == 0 ? uo[v] : un[v];
      ur[v] = side == 0 ? un[v] : uo[v];
    }
    double wl[D + 3], wr[D + 3], n[D], h[NV];
    computational<D>(ph, ul, wl);
    computational<D>(ph, ur, wr);
    for (int d = 0; d < D; d++) n[d] = g.fnormal[fp * D + d];
    convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
    if constexpr (sizeof(TC) == 4) reduceFaceFlux<D>(a, n, h);
    const double jw = g.fjw[fp];
    for (int v = 0; v < NV; v++) {
      const double t = h[v] * jw;
      sH[sl * 16 + le * NV + v] = static_cast<TC>(side == 0 ? -t : t);
    }
  }
  __syncthreads();
  // residual R = F Phi_grad - H Phi_face
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NBm>(acc, op.grad, m0, sF, 0, NDQk, lane);
    tileAcc<TC, NBm>(acc, op.face, m0, sH, 0, NAQk, lane);
    tileStore<TC>(acc, sR, m0, lane);
  }
  for (int idx = NBm * 16 + tid; idx < NBr * 16; idx += kThreads) sR[idx] = TC(0);
  __syncthreads();
  addSource<TC, D, NB, EPB>(g, a, e0, sR);
  if (a.mode == 1) {
    for (int idx = tid; idx < NB * 16; idx += kThreads) {
      const int b = idx / 16, c = idx % 16, le = c / NV, v = c % NV, e = e0 + le;
      if (c < EPB * NV && e < g.E) {
        double r = sR[idx];
        if (sizeof(TC) == 4 && v == NV - 1) r += a.mix.h0 * static_cast<double>(sR[b * 16 + le * NV]);
        a.Rout[(static_cast<int64_t>(e) * NV + v) * NB + b] = r;
      }
    }
    return;
  }
  mfmaStageTail<TC, D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sR, sU, sRed);
}

// ---- Navier-Stokes kernel 1 (MFMA): gradient lifting + volume flux -------------------------------
// One element per workgroup in 3-D (15 gradient columns), two in 2-D (2 x 8 columns).
// The lifted coefficients w = R Mref^-1 / |J| are only ever evaluated at points, so Mref^-1 is folded
// into the evaluation operators: grad U_q = R_t (Mref^-1 V^T) / |J| and the BR2 face trace
// R_w (Mref^-1 Vf^T) / |J| + J Lf / |J| (BR1: R_t (Mref^-1 Vf^T) / |J|).
// In the mixed-precision mode the lifting acts on the deviation traces (the lifted gradient of the
// constant reference state is zero).
// LDS layout shared by the gradient-lifting kernels (forward and adjoint)
template <int D, int NB, int NQ, int NF, int NFQ>
struct GradLds {
  static constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  static constexpr int EPB = 16 / NG;
  static constexpr int NBk = r4(NB), NBm = r16(NB), NQm = r16(NQ), NAQk = r4(NAQ), NAQm = r16(NAQ), NDQk = r4(NQ * D);
  static constexpr int NBr = NBk > NBm ? NBk : NBm;
  static constexpr int AJ = 2 * NAQk * 16 > NQm * 16 ? 2 * NAQk * 16 : NQm * 16;
};

// Gradient lifting (BR1 / BR2) of the workgroup's EPB elements starting at e0. On return:
//   sUq[q][le*NV + v]   state at the volume points,
//   sAJ[q][le*NG + r]   total (lifted) gradient at the volume points (first NQm*16 entries),
//   Gf (HBM)            lifted gradient traces at the element's face points, and, if a.Wout is
//                       set, the lifted total-gradient coefficients;
// sRR is free again. Ends with a barrier.
template <typename TC, int D, int NB, int NQ, int NF, int NFQ, int VISC>
__device__ __forceinline__ void gradLift(const Geo& g, const Phys& ph, const StageArgs& a, int e0, TC* sUq, TC* sAJ,
                                         TC* sRR) {
  using LD = GradLds<D, NB, NQ, NF, NFQ>;
  const MOpsT<TC>& op = opsOf<TC>(g);
  constexpr int NV = LD::NV, NG = LD::NG, NAQ = LD::NAQ, EPB = LD::EPB;
  constexpr int NBk = LD::NBk, NBm = LD::NBm, NQm = LD::NQm, NAQk = LD::NAQk, NAQm = LD::NAQm, NDQk = LD::NDQk;
  constexpr int NBr = LD::NBr, AJ = LD::AJ;
  TC* sU = sRR;
  TC* sA = sAJ;
  TC* sJ = sAJ + NAQk * 16;
  TC* sRw = sRR;
  TC* sRt = sRR + NBr * 16;
  TC* sGq = sAJ;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const TC* Uf_in = pick<TC>(a.Uf_in, a.Uf_in32);
  zeroLds(sAJ, AJ, tid);
  for (int idx = tid; idx < NBr * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sU[idx] = (b < NB && cc < EPB * NV && e < g.E) ? stateOperand<TC>(a, e, v, b, NV, NB) : TC(0);
  }
  __syncthreads();
  // face terms: (u- + u+)/2 (x) n |J_f| w (signed by side) and (u+ - u-)/2 (x) n |J_f| w
  for (int idx = tid; idx < EPB * NAQ * NV; idx += kThreads) {
    const int le = idx / (NAQ * NV), rem = idx % (NAQ * NV), sl = rem / NV, v = rem % NV, e = e0 + le;
    if (e >= g.E) continue;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es], fps = g.sfp[es];
    if (nbs < 0) {  // boundary face point: terms from the boundary-condition states
      double uo[NV], Ab[D], Jb[D];
      for (int k = 0; k < NV; k++) uo[k] = pointState<TC>(a, Uf_in[(static_cast<int64_t>(e) * NV + k) * NAQ + sl], k);
      boundaryGradientTerms<D>(ph, g, -1 - nbs, uo, v, sizeof(TC) == 4 ? a.mix.uref[v] : 0.0, Ab, Jb);
      for (int d = 0; d < D; d++) {
        sA[sl * 16 + le * NG + v * D + d] = static_cast<TC>(Ab[d]);
        sJ[sl * 16 + le * NG + v * D + d] = static_cast<TC>(Jb[d]);
      }
      continue;
    }
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const double uo = Uf_in[(static_cast<int64_t>(e) * NV + v) * NAQ + sl];
    const double un = Uf_in[(static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ];
    const double ul = side == 0 ? uo : un, ur = side == 0 ? un : uo;
    const double jw = g.fjw[fp];
    for (int d = 0; d < D; d++) {
      const double n = g.fnormal[fp * D + d];
      const double avg = (n * (ul + ur) / 2.0) * jw;
      sA[sl * 16 + le * NG + v * D + d] = static_cast<TC>(side == 0 ? avg : -avg);
      sJ[sl * 16 + le * NG + v * D + d] = static_cast<TC>((n * (ur - ul) / 2.0) * jw);
    }
  }
  // state at the volume points
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NQm>(acc, op.interp, m0, sU, 0, NBk, lane);
    tileStore<TC>(acc, sUq, m0, lane);
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
    #pragma unroll 1
    for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
      acc_t<TC> aw = zeroAcc<TC>(), aj = zeroAcc<TC>();
      tileAcc<TC, NBm>(aw, op.face, m0, sA, 0, NAQk, lane);
      tileAcc<TC, NBm>(aj, op.face, m0, sJ, 0, NAQk, lane);
#pragma unroll 4
      for (int k = 0; k < NDQk; k += 4) {
        const int kk = k + rr;
        TC bq = TC(0);
        if (gok && kk < NQ * D) {
          const int q = kk / D, b = kk % D;
          const double G = gi[b] * (dj * g.wq[q]);
          bq = static_cast<TC>(-(static_cast<double>(sUq[q * 16 + gle * NV + gk]) * G));
        }
        aw = Mfma<TC>::mma(op.grad[kk * NBm + m0 + c], bq, aw);
      }
      // the state panel (aliased by R) was last read before the barrier above
      tileStore<TC>(aw, sRw, m0, lane);
      acc_t<TC> at = aw + aj;
      tileStore<TC>(at, sRt, m0, lane);
    }
  }
  for (int idx = NBm * 16 + tid; idx < NBr * 16; idx += kThreads) sRw[idx] = sRt[idx] = TC(0);
  __syncthreads();
  const double idj = 1.0 / dj;
  // optional output of the lifted total-gradient coefficients w = R_t Mref^-1 / |J|
  if (a.Wout) {
    #pragma unroll 1
    for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
      acc_t<TC> wt = zeroAcc<TC>();
      tileAcc<TC, NBm>(wt, op.minv, m0, sRt, 0, NBk, lane);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int bi = m0 + Mfma<TC>::row(rr, i);
        if (gok && bi < NB) a.Wout[(static_cast<int64_t>(ge) * NB + bi) * NG + (c % NG)] = static_cast<double>(wt[i]) * idj;
      }
    }
  }
  // lifted gradient traces at the face points -> HBM
  TC* Gf = pick<TC>(a.Gf, a.Gf32);
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    if (VISC == 2) {
      tileAcc<TC, NAQm>(acc, op.ptrace, m0, sRw, 0, NBk, lane);
      // block-diagonal BR2 lift: only the faces this row tile touches
      const int s_lo = m0 / NFQ;
      const int s_hi = (m0 + 15 < NAQ ? m0 + 15 : NAQ - 1) / NFQ;
      const int k0 = (s_lo * NFQ) / 4 * 4;
      const int k1 = r4((s_hi + 1) * NFQ);
      tileAcc<TC, NAQm>(acc, op.lift, m0, sJ, k0, k1 < NAQk ? k1 : NAQk, lane);
    } else {  // BR1: the total gradient; VISC == 0 (artificial viscosity): the volume gradient
      tileAcc<TC, NAQm>(acc, op.ptrace, m0, VISC == 0 ? sRw : sRt, 0, NBk, lane);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + Mfma<TC>::row(rr, i);
      if (gok && sl < NAQ)
        Gf[(static_cast<int64_t>(ge) * NG + (c % NG)) * NAQ + sl] = static_cast<TC>(static_cast<double>(acc[i]) * idj);
    }
  }
  __syncthreads();
  // total gradient at the volume points (into the face-term buffer, no longer needed)
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NQm>(acc, op.pinterp, m0, VISC == 0 ? sRw : sRt, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = static_cast<TC>(static_cast<double>(acc[i]) * idj);
    tileStore<TC>(acc, sGq, m0, lane);
  }
  __syncthreads();
}

template <typename TC, int D, int NB, int NQ, int NF, int NFQ, int VISC>
__global__ __launch_bounds__(kThreads) void nsGradMfma(Geo g, Phys ph, StageArgs a) {
  using LD = GradLds<D, NB, NQ, NF, NFQ>;
  constexpr int NV = LD::NV, NG = LD::NG, EPB = LD::EPB;
  __shared__ TC sUq[LD::NQm * 16];      // state at volume points
  __shared__ TC sAJ[LD::AJ];            // face terms A | J ([slot][le*NG + r]); later the gradient at points
  __shared__ TC sRR[2 * LD::NBr * 16];  // state panel; then R_w | R_t
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  gradLift<TC, D, NB, NQ, NF, NFQ, VISC>(g, ph, a, e0, sUq, sAJ, sRR);
  const TC* sGq = sAJ;
  // volume flux (F_c - F_v)^T (J^T)^-1 |J| w at the points -> HBM Fv[e][(q D + b)][v]; the weak
  // divergence F Phi_grad is taken in nsResidualMfma, which packs three elements per 16 columns
  TC* Fv = pick<TC>(a.Fv, a.Fv32);
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ, e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3], gq[NG], pg[NG], fv[D * NV];
    for (int v = 0; v < NV; v++) u[v] = pointState<TC>(a, sUq[q * 16 + le * NV + v], v);
    for (int r = 0; r < NG; r++) gq[r] = sGq[q * 16 + le * NG + r];
    computational<D>(ph, u, w);
    if constexpr (VISC == 0) {  // Euler with artificial viscosity: F_v = eps grad U (volume gradient)
      const double eps = avAt(g, g.nodq, q, e);
      for (int v = 0; v < NV; v++)
        for (int aa = 0; aa < D; aa++) fv[aa + D * v] = eps * gq[v * D + aa];
      if constexpr (sizeof(TC) == 4)  // mixed: the energy row reduced by h0 times the density row
        for (int aa = 0; aa < D; aa++) fv[aa + D * (NV - 1)] -= a.mix.h0 * fv[aa];
      (void)pg;
    } else {
      primitiveGradient<D>(ph, w, u, gq, pg);
      viscousRawFlux<D>(ph, w, pg, fv);
    }
    const double jw = g.detj[e] * g.wq[q];
    double G[D * D], out[D * NV];
    for (int i = 0; i < D * D; i++) G[i] = g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
    volumeFlux<TC, D, true>(a, w, fv, G, out);
    TC* o = Fv + (static_cast<int64_t>(e) * NQ * D + q * D) * NV;
    for (int i = 0; i < D * NV; i++) o[i] = static_cast<TC>(out[i]);
  }
}

// ---- Navier-Stokes face fluxes, one thread per face point ---------------------------------------
// calculateInteriorAdjacencyElementQuadrature (SpatialDiscrete.cpp:633-748): each face point is
// evaluated once with its left/right roles; t = (H_c - H_v) |J_f| w goes to the left parent's slot
// as -t and to the right parent's slot (point seq[j]) as +t, the signs with which the face term
// enters R = R_vol + sum_slots Phi_f s.
// VN: each side's normal viscous flux F_v(u, grad u) . n_out was formed by the side's gradient
// kernel (a.Vn holding TC values, with the same arithmetic as below); the face kernel then reads
// 2 NV values per point instead of both sides' 2 NV D gradient traces (boundary points still
// read Gf).
// AV: Euler with artificial viscosity: Gf holds the volume-gradient traces and the viscous part is
// the central average of eps grad U . n (SpatialDiscrete.cpp:714-735,812-830); in the mixed mode
// its energy row is reduced by h0 times its density row like the convective part. VN with AV
// (Navier-Stokes, sum-factorised kernels): each side's Vn already holds (F_v + eps grad U) . n_out;
// boundary points take eps grad U from the volume-gradient traces GfV.
template <typename TC, int D, int NF, int NFQ, bool VN = false, bool AV = false>
__global__ __launch_bounds__(kThreads) void nsFaceFluxKernel(Geo g, Phys ph, StageArgs a) {
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  if (a.status && a.status[0]) return;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<int64_t>(g.F + g.Fb) * NFQ) return;
  const TC* Uf_in = pick<TC>(a.Uf_in, a.Uf_in32);
  const TC* Gf = pick<TC>(static_cast<const double*>(a.Gf), static_cast<const float*>(a.Gf32));
  TC* Hs = pick<TC>(a.Hs, a.Hs32);
  if (idx >= static_cast<int64_t>(g.F) * NFQ) {  // boundary face point
    const int bp = static_cast<int>(idx - static_cast<int64_t>(g.F) * NFQ);
    const int64_t os = g.bslot[bp];
    const int64_t e = os / NAQ;
    const int sl = static_cast<int>(os % NAQ);
    double ub[NV], gb[NG], t[NV];
    for (int v = 0; v < NV; v++) ub[v] = pointState<TC>(a, Uf_in[(e * NV + v) * NAQ + sl], v);
    for (int r = 0; r < NG; r++) gb[r] = Gf[(e * NG + r) * NAQ + sl];
    // Euler with AV: no viscous flux; Navier-Stokes (VN) with AV: the viscous flux from Gf and the
    // artificial term from the volume-gradient traces GfV
    boundaryFlux<D, VN || !AV>(ph, g, bp, ub, gb, sizeof(TC) == 4, a.mix.pref, a.mix.h0, t);
    if constexpr (AV) {  // minus eps grad U . n of the element's own volume gradient
      const double eps = avAt(g, g.nodf, sl, static_cast<int>(e)), bjw = g.bjw[bp];
      double f[NV];
      for (int v = 0; v < NV; v++) {
        f[v] = 0.0;
        for (int d = 0; d < D; d++) {
          const double gv = VN ? a.GfV[(e * NG + v * D + d) * NAQ + sl] : gb[v * D + d];
          f[v] += (eps * gv) * g.bnormal[static_cast<int64_t>(bp) * D + d];
        }
      }
      if constexpr (sizeof(TC) == 4) f[NV - 1] -= a.mix.h0 * f[0];  // mixed: energy reduced by h0 mass
      for (int v = 0; v < NV; v++) t[v] -= f[v] * bjw;
    }
    for (int v = 0; v < NV; v++) Hs[(e * NV + v) * NAQ + sl] = static_cast<TC>(-t[v]);
    return;
  }
  const int64_t f = idx / NFQ;
  const int j = static_cast<int>(idx % NFQ);
  const int k = g.fseq[f * NFQ + j];
  const int64_t L = g.fpar[2 * f], R = g.fpar[2 * f + 1];
  const int slL = g.floc[2 * f] * NFQ + j, slR = g.floc[2 * f + 1] * NFQ + k;
  double ul[NV], ur[NV];
  for (int v = 0; v < NV; v++) {
    ul[v] = pointState<TC>(a, Uf_in[(L * NV + v) * NAQ + slL], v);
    ur[v] = pointState<TC>(a, Uf_in[(R * NV + v) * NAQ + slR], v);
  }
  double wl[D + 3], wr[D + 3], n[D], h[NV];
  computational<D>(ph, ul, wl);
  computational<D>(ph, ur, wr);
  for (int d = 0; d < D; d++) n[d] = g.fnormal[(f * NFQ + j) * D + d];
  convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
  if constexpr (sizeof(TC) == 4) reduceFaceFlux<D>(a, n, h);
  // normal viscous flux of each side, one side at a time (keeps one side's gradient live)
  double vl[NV], vr[NV];
  if constexpr (AV && !VN) {  // Euler: eps grad U . n of each side with the left normal
    const double el = avAt(g, g.nodf, slL, static_cast<int>(L)), er = avAt(g, g.nodf, slR, static_cast<int>(R));
    for (int v = 0; v < NV; v++) {
      double al = 0.0, ar = 0.0;
      for (int d = 0; d < D; d++) al += (el * static_cast<double>(Gf[(L * NG + v * D + d) * NAQ + slL])) * n[d];
      for (int d = 0; d < D; d++) ar += (er * static_cast<double>(Gf[(R * NG + v * D + d) * NAQ + slR])) * n[d];
      vl[v] = al;
      vr[v] = ar;
    }
    if constexpr (sizeof(TC) == 4) {  // mixed: the energy rows reduced by h0 times the density rows
      vl[NV - 1] -= a.mix.h0 * vl[0];
      vr[NV - 1] -= a.mix.h0 * vr[0];
    }
  } else if constexpr (VN) {  // the right side's outward normal is -n
    const TC* Vn = reinterpret_cast<const TC*>(a.Vn);
    for (int v = 0; v < NV; v++) {
      vl[v] = static_cast<double>(Vn[(L * NV + v) * NAQ + slL]);
      vr[v] = -static_cast<double>(Vn[(R * NV + v) * NAQ + slR]);
    }
  } else {
#pragma unroll 1
    for (int sd = 0; sd < 2; sd++) {
      const int64_t es = sd ? R : L;
      const int ss = sd ? slR : slL;
      double us[NV], ws[D + 3], gs[NG], pg[NG], fv[D * NV];
      for (int v = 0; v < NV; v++) us[v] = sd ? ur[v] : ul[v];
      for (int i = 0; i < D + 3; i++) ws[i] = sd ? wr[i] : wl[i];
      for (int r = 0; r < NG; r++) gs[r] = Gf[(es * NG + r) * NAQ + ss];
      primitiveGradient<D>(ph, ws, us, gs, pg);
      viscousRawFlux<D>(ph, ws, pg, fv);
      for (int v = 0; v < NV; v++) {
        double s = 0.0;
        for (int d = 0; d < D; d++) s += fv[d + D * v] * n[d];
        if (sd)
          vr[v] = s;
        else
          vl[v] = s;
      }
    }
  }
  const double jw = g.fjw[f * NFQ + j];
  for (int v = 0; v < NV; v++) {
    const double t = (h[v] - (vl[v] + vr[v]) / 2.0) * jw;
    Hs[(L * NV + v) * NAQ + slL] = static_cast<TC>(-t);
    Hs[(R * NV + v) * NAQ + slR] = static_cast<TC>(t);
  }
}

// ---- Navier-Stokes residual + update + next traces (MFMA), after nsFaceFluxKernel ---------------
// calculateResidual + updateBasisFunctionCoefficient (SpatialDiscrete.cpp:1016-1032,
// TimeIntegration.cpp:181-198)
template <typename TC, int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void nsResidualMfma(Geo g, Phys ph, StageArgs a) {
  const MOpsT<TC>& op = opsOf<TC>(g);
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  constexpr int EPB = 16 / NV;
  constexpr int NBk = r4(NB), NBm = r16(NB), NAQk = r4(NAQ);
  constexpr int NBr = NBk > NBm ? NBk : NBm;
  constexpr int NDQ = NQ * D, NDQk = r4(NDQ);
  __shared__ TC sU[NBr * 16];
  __shared__ TC sH[NAQk * 16];
  __shared__ TC sF[NDQk * 16];
  __shared__ TC sR[NBr * 16];
  __shared__ double sRed[64];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int e0 = blockIdx.x * EPB;
  const TC* Hs = pick<TC>(static_cast<const double*>(a.Hs), static_cast<const float*>(a.Hs32));
  const TC* Fv = pick<TC>(static_cast<const double*>(a.Fv), static_cast<const float*>(a.Fv32));
  if constexpr (sizeof(TC) == 8) {  // the mixed-precision tail reads the FP64 state from HBM
    for (int idx = tid; idx < NBr * 16; idx += kThreads) {
      const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
      sU[idx] = (b < NB && cc < EPB * NV && e < g.E) ? a.U[(static_cast<int64_t>(e) * NV + v) * NB + b] : 0.0;
    }
  }
  for (int idx = tid; idx < NAQk * 16; idx += kThreads) {
    const int sl = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sH[idx] = (sl < NAQ && cc < EPB * NV && e < g.E) ? Hs[(static_cast<int64_t>(e) * NV + v) * NAQ + sl] : TC(0);
  }
  for (int idx = tid; idx < NDQk * 16; idx += kThreads) {
    const int k = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    sF[idx] = (k < NDQ && cc < EPB * NV && e < g.E) ? Fv[(static_cast<int64_t>(e) * NDQ + k) * NV + v] : TC(0);
  }
  __syncthreads();
  #pragma unroll 1
  for (int m0 = wave * 16; m0 < NBm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NBm>(acc, op.grad, m0, sF, 0, NDQk, lane);
    tileAcc<TC, NBm>(acc, op.face, m0, sH, 0, NAQk, lane);
    tileStore<TC>(acc, sR, m0, lane);
  }
  __syncthreads();
  addSource<TC, D, NB, EPB>(g, a, e0, sR);
  if (a.mode == 1) {
    for (int idx = tid; idx < NB * 16; idx += kThreads) {
      const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
      if (cc < EPB * NV && e < g.E) {
        double r = sR[idx];
        if (sizeof(TC) == 4 && v == NV - 1) r += a.mix.h0 * static_cast<double>(sR[b * 16 + le * NV]);
        a.Rout[(static_cast<int64_t>(e) * NV + v) * NB + b] = r;
      }
    }
    return;
  }
  mfmaStageTail<TC, D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sR, sU, sRed);
}

// Traces at the element's own face points from the state in HBM, with the same operands and
// matrix-core accumulation order as the stage tail that produces them after every stage, so a state
// set from the host (a restart, a checkpoint) continues bit-identically. EPB elements per block.
template <typename TC, int D, int NB, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void mfmaTraceKernel(Geo g, const double* U, void* Uf_out, MixRef mix) {
  constexpr int NV = D + 2, NAQ = NF * NFQ, EPB = 16 / NV;
  constexpr int NBk = r4(NB), NAQm = r16(NAQ);
  __shared__ TC sU[NBk * 16];
  const MOpsT<TC>& op = opsOf<TC>(g);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rr = lane >> 4;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < NBk * 16; idx += kThreads) {
    const int b = idx / 16, cc = idx % 16, le = cc / NV, v = cc % NV, e = e0 + le;
    TC x = TC(0);
    if (b < NB && cc < EPB * NV && e < g.E) {
      const double u = U[(static_cast<int64_t>(e) * NV + v) * NB + b];
      if constexpr (sizeof(TC) == 8)
        x = u;
      else
        x = deviationOperand(u, mix.uref[v], mix.one[b]);
    }
    sU[idx] = x;
  }
  __syncthreads();
  const int le = c / NV, v = c % NV, e = e0 + le;
  const bool col_ok = c < EPB * NV && e < g.E;
  TC* out = static_cast<TC*>(Uf_out);
#pragma unroll 1
  for (int m0 = wave * 16; m0 < NAQm; m0 += 64) {
    acc_t<TC> acc = zeroAcc<TC>();
    tileAcc<TC, NAQm>(acc, op.trace, m0, sU, 0, NBk, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sl = m0 + Mfma<TC>::row(rr, i);
      if (col_ok && sl < NAQ) out[(static_cast<int64_t>(e) * NV + v) * NAQ + sl] = acc[i];
    }
  }
}

}  // namespace sdg

#endif  // SDG_DG_MFMA_HPP_
