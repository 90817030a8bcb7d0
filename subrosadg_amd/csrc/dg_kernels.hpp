// dg_kernels.hpp -- gfx950 kernels of one RK stage of the DG residual + update.
//
// Device layout (structure of arrays, element-major panels):
//   U, Un      [E][NV][N_b]        modal coefficients (row (e,v) contiguous over basis)
//   Uf[2]      [E][NV][N_aq]       traces of U at the element's face points (ping-pong)
//   Gf         [E][NV*D][N_aq]     lifted-gradient traces at face points (Navier-Stokes)
//   Rvol       [E][NV][N_b]        volume part of the residual (Navier-Stokes)
// Operand tables are small and shared by all elements; geometry is per element (affine).
//
// Each kernel processes EPB elements per workgroup with the element panels staged in LDS.
// Faces are processed element-centrically: every element evaluates the numerical flux of each of
// its face points itself, using the face's own left/right roles and the left parent's normal,
// so both parents obtain bit-identical +H / -H values without a separate face pass or atomics
// (the reference writes +H to the left slot and -H to the right slot of a single face
// evaluation, SpatialDiscrete.cpp:738-744).
#ifndef SDG_DG_KERNELS_HPP_
#define SDG_DG_KERNELS_HPP_

#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace sdg {

constexpr int kThreads = 256;

constexpr int r16(int x) { return (x + 15) / 16 * 16; }
constexpr int r4(int x) { return (x + 3) / 4 * 4; }

// Padded, transposed operator tables T[K4][M16] for the matrix-core kernels (zero outside the
// operator): Out[m][n] = sum_k T[k][m] X[k][n].
struct MOps {
  const double* interp;  // K = N_b,   M = N_q   : basis values          (U -> U_q, w -> grad U_q, R -> R_q)
  const double* grad;    // K = N_q D, M = N_b   : basis gradients       (weak divergence, volume gradient term)
  const double* face;    // K = N_aq,  M = N_b   : face values           (face terms of residuals)
  const double* minv;    // K = N_b,   M = N_b   : reference mass inverse
  const double* trace;   // K = N_b,   M = N_aq  : face values           (coefficients -> face traces)
  const double* lift;    // K = N_aq,  M = N_aq  : block-diagonal Vf_s Mref^-1 Vf_s^T (BR2 face lift)
  const double* pinterp; // K = N_b,   M = N_q   : Mref^-1 V^T   (R -> gradient at volume points)
  const double* ptrace;  // K = N_b,   M = N_aq  : Mref^-1 Vf^T  (R -> gradient traces at face points)
};

struct Geo {
  const double* Vt;    // N_b x N_q        basis at volume points, transposed
  const double* dV;    // (N_q D) x N_b    reference basis gradient, row q*D+d
  const double* Vft;   // N_b x N_aq       basis at face points, transposed
  const double* Vf;    // N_aq x N_b
  const double* Minv;  // N_b x N_b        reference-element mass inverse
  const double* Lf;    // N_f x N_fq x N_fq  face lifting blocks Vf_s Minv Vf_s^T (BR2)
  const double* wq;    // N_q               volume weights
  const double* ginv;  // E x D*D           (J^T)^-1 column-major
  const double* detj;  // E
  const double* fnormal;  // F x N_fq x D   outward from the left parent
  const double* fjw;      // F x N_fq       |J_f| w
  const int* fseq;        // F x N_fq       right point of left point j
  const int* fseqinv;     // F x N_fq       left point of right point k
  const int* efid;        // E x N_f        face id
  const int* enb;         // E x N_f        neighbour element
  const int* enbf;        // E x N_f        neighbour's local face
  const int* eside;       // E x N_f        0: this element is the face's left parent
  const int* snb;         // E x N_aq       neighbour slot: neighbour element * N_aq + its slot
  const int* sfp;         // E x N_aq       2 (f N_fq + p) + side: face point of the slot and side
  const int* fpar;        // F x 2          left, right parent element
  const int* floc;        // F x 2          local face index in left, right parent
  int E;
  int F;
  MOps op;
};

struct StageArgs {
  double* U;
  double* Un;
  const double* Uf_in;
  double* Uf_out;
  double* Gf;
  double* Rvol;
  double* Rout;   // residual output (mode 1)
  double* Wout;   // lifted total gradient output (optional)
  double* Hs;     // E x NV x N_aq signed face-flux slots (Navier-Stokes, matrix-core path)
  double* Fv;     // E x (N_q D) x NV contravariant volume flux (Navier-Stokes, matrix-core path)
  double* mon;    // E x NV monitor partials (last stage)
  int* status;    // [0] NaN stop flag, [1] steps taken; nullptr: no check
  double c0, c1, c2dt;  // U <- c1 U + c0 Un + c2dt R Minv (TimeIntegration.cpp:181-198)
  int last_stage;
  int mode;  // 0: stage update, 1: residual only
};

template <int D, int NB, int NQ, int NF, int NFQ>
struct Sizes {
  static constexpr int NV = D + 2;
  static constexpr int NG = NV * D;
  static constexpr int NAQ = NF * NFQ;
  static constexpr int MAXN = (NB > NQ ? (NB > NAQ ? NB : NAQ) : (NQ > NAQ ? NQ : NAQ));
  // elements per workgroup: enough (element, variable, point) work items for 256 threads
  static constexpr int EPB_RAW = kThreads / (NV * MAXN);
  static constexpr int EPB = EPB_RAW < 1 ? 1 : (EPB_RAW > 16 ? 16 : EPB_RAW);
};

// Face point roles of slot (s, j) of element e: face id, face point p (left ordering), neighbour
// slot, and side (0 left, 1 right).
template <int NF, int NFQ>
__device__ __forceinline__ void slotRoles(const Geo& g, int e, int s, int j, int& f, int& p, int64_t& nbslot,
                                          int& side) {
  constexpr int NAQ = NF * NFQ;
  f = g.efid[e * NF + s];
  side = g.eside[e * NF + s];
  const int nb = g.enb[e * NF + s];
  const int nbs = g.enbf[e * NF + s];
  int k;
  if (side == 0) {
    p = j;
    k = g.fseq[f * NFQ + j];
  } else {
    p = g.fseqinv[f * NFQ + j];
    k = p;
  }
  nbslot = static_cast<int64_t>(nb) * NAQ + nbs * NFQ + k;
}

// ---- monitor, update and next-stage traces (shared tail of the last kernel of a stage) ----
template <int D, int NB, int NQ, int NF, int NFQ, int EPB>
__device__ __forceinline__ void stageTail(const Geo& g, const StageArgs& a, int e0, double (*sU)[D + 2][NB],
                                          double (*sR)[D + 2][NB]) {
  using S = Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int NV = S::NV, NAQ = S::NAQ;
  const int tid = threadIdx.x;
  if (a.mode == 1) {
    for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
      const int le = idx / (NV * NB), rem = idx % (NV * NB);
      const int e = e0 + le;
      if (e < g.E) a.Rout[static_cast<int64_t>(e) * NV * NB + rem] = sR[le][rem / NB][rem % NB];
    }
    return;
  }
  if (a.last_stage) {  // residual monitor: mean_q |R V^T| per variable (TimeIntegration.cpp:279-298)
    for (int idx = tid; idx < EPB * NV; idx += kThreads) {
      const int le = idx / NV, v = idx % NV;
      const int e = e0 + le;
      if (e >= g.E) continue;
      double s = 0.0;
      for (int q = 0; q < NQ; q++) {
        double r = 0.0;
        for (int b = 0; b < NB; b++) r += sR[le][v][b] * g.Vt[b * NQ + q];
        s += fabs(r);
      }
      a.mon[static_cast<int64_t>(e) * NV + v] = s / NQ;
    }
  }
  // update: U <- c1 U + c0 Un + c2dt (R Minv) / |J|
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int v = rem / NB, i = rem % NB;
    const int e = e0 + le;
    if (e >= g.E) continue;
    double rm = 0.0;
    for (int b = 0; b < NB; b++) rm += sR[le][v][b] * g.Minv[b * NB + i];
    rm = rm / g.detj[e];
    const int64_t gi = static_cast<int64_t>(e) * NV * NB + rem;
    double u = sU[le][v][i] * a.c1;
    u += a.c0 * a.Un[gi];
    u += a.c2dt * rm;
    sU[le][v][i] = u;
    a.U[gi] = u;
    if (a.last_stage) a.Un[gi] = u;
  }
  __syncthreads();
  // traces of the new state for the next stage
  for (int idx = tid; idx < EPB * NV * NAQ; idx += kThreads) {
    const int le = idx / (NV * NAQ), rem = idx % (NV * NAQ);
    const int v = rem / NAQ, sl = rem % NAQ;
    const int e = e0 + le;
    if (e >= g.E) continue;
    double t = 0.0;
    for (int b = 0; b < NB; b++) t += sU[le][v][b] * g.Vft[b * NAQ + sl];
    a.Uf_out[static_cast<int64_t>(e) * NV * NAQ + rem] = t;
  }
}

// ---- Euler: the whole stage in one kernel ----------------------------------------------------
// passes calculateQuadrature, calculateAdjacencyQuadrature, calculateResidual and
// updateBasisFunctionCoefficient (SpatialDiscrete.cpp:194-266,633-748,1016-1032;
// TimeIntegration.cpp:181-198). The reference also runs the gradient passes for Euler; they do
// not influence U without artificial viscosity (SpatialDiscrete.cpp:211-215) and are skipped.
template <int D, int NB, int NQ, int NF, int NFQ, int EPB>
__global__ __launch_bounds__(kThreads) void eulerStageKernel(Geo g, Phys ph, StageArgs a) {
  using S = Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int NV = S::NV, NAQ = S::NAQ;
  __shared__ double sU[EPB][NV][NB];
  __shared__ double sUq[EPB][NV][NQ];
  __shared__ double sF[EPB][NV][D * NQ];
  __shared__ double sH[EPB][NV][NAQ];
  __shared__ double sR[EPB][NV][NB];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int e = e0 + le;
    sU[le][rem / NB][rem % NB] = e < g.E ? a.U[static_cast<int64_t>(e) * NV * NB + rem] : 1.0;
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NV * NQ; idx += kThreads) {
    const int le = idx / (NV * NQ), rem = idx % (NV * NQ);
    const int v = rem / NQ, q = rem % NQ;
    double s = 0.0;
    for (int b = 0; b < NB; b++) s += sU[le][v][b] * g.Vt[b * NQ + q];
    sUq[le][v][q] = s;
  }
  __syncthreads();
  // volume flux at the quadrature points, contravariant: (F^T (J^T)^-1 |J| w)
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ;
    const int e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) u[v] = sUq[le][v][q];
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
        sF[le][v][q * D + b] = s;
      }
  }
  // face fluxes
  for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
    const int le = idx / NAQ, sl = idx % NAQ;
    const int e = e0 + le;
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
      sH[le][v][sl] = side == 0 ? t : -t;
    }
  }
  __syncthreads();
  // residual R = F Phi_grad - H Phi_face (SpatialDiscrete.cpp:1016-1032)
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int v = rem / NB, i = rem % NB;
    double r = 0.0;
    for (int k = 0; k < D * NQ; k++) r += sF[le][v][k] * g.dV[k * NB + i];
    double h = 0.0;
    for (int k = 0; k < NAQ; k++) h += sH[le][v][k] * g.Vf[k * NB + i];
    sR[le][v][i] = r - h;
  }
  __syncthreads();
  stageTail<D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sU, sR);
}

// ---- Navier-Stokes, kernel 1: gradient lifting + volume flux --------------------------------
// passes calculateGardientQuadrature .. updateGardientBasisFunctionCoefficient and
// calculateQuadrature (SpatialDiscrete.cpp:294-346,844-911,1034-1068; TimeIntegration.cpp:200-228;
// SpatialDiscrete.cpp:194-266). Lifting, for affine elements (M^-1 = Mref^-1/|J|):
//   w_vol = (A Phi_f - Q Phi_grad) M^-1 ,  w = (A Phi_f + J Phi_f - Q Phi_grad) M^-1
//   BR2 face trace on face s: w_vol Phi_f,s^T + J_s (Phi_f,s Mref^-1 Phi_f,s^T) / |J|
//   BR1 face trace: w Phi_f^T
// which equals the reference's per-face lifting r_s = R_w,s M^-1 (BR2, eta = 1,
// VariableConvertor.cpp:674-688) up to rounding.
template <int D, int NB, int NQ, int NF, int NFQ, int EPB, int VISC>
__global__ __launch_bounds__(kThreads) void nsGradKernel(Geo g, Phys ph, StageArgs a) {
  using S = Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int NV = S::NV, NG = S::NG, NAQ = S::NAQ;
  constexpr int NBQ = NB > NQ ? NB : NQ;
  __shared__ double sU[EPB][NV][NB];
  __shared__ double sUq[EPB][NV][NQ];
  __shared__ double sA[EPB][NG][NAQ];   // averaged face term, later the face gradient traces
  __shared__ double sJ[EPB][NG][NAQ];   // jump face term
  __shared__ double sR1[EPB][NG][NBQ];  // R_w (vol), later gradient at volume points
  __shared__ double sR2[EPB][NG][NBQ];  // R_w (total), later the contravariant volume flux
  __shared__ double sWv[EPB][NG][NB];
  __shared__ double sWt[EPB][NG][NB];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int e = e0 + le;
    sU[le][rem / NB][rem % NB] = e < g.E ? a.U[static_cast<int64_t>(e) * NV * NB + rem] : 1.0;
  }
  // face terms of the gradient: (u- + u+)/2 (x) n |J_f| w and (u+ - u-)/2 (x) n |J_f| w
  // (ViscousFlux.cpp:32-56, SpatialDiscrete.cpp:875-907)
  for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
    const int le = idx / NAQ, sl = idx % NAQ;
    const int e = e0 + le;
    if (e >= g.E) {
      for (int r = 0; r < NG; r++) sA[le][r][sl] = sJ[le][r][sl] = 0.0;
      continue;
    }
    const int s = sl / NFQ, j = sl % NFQ;
    int f, p, side;
    int64_t nbslot;
    slotRoles<NF, NFQ>(g, e, s, j, f, p, nbslot, side);
    const int64_t own = static_cast<int64_t>(e) * NV * NAQ + sl;
    const int64_t nbe = (nbslot / NAQ) * NV * NAQ + nbslot % NAQ;
    double n[D];
    for (int d = 0; d < D; d++) n[d] = g.fnormal[(static_cast<int64_t>(f) * NFQ + p) * D + d];
    const double jw = g.fjw[static_cast<int64_t>(f) * NFQ + p];
    for (int v = 0; v < NV; v++) {
      const double uo = a.Uf_in[own + v * NAQ], un = a.Uf_in[nbe + v * NAQ];
      const double ul = side == 0 ? uo : un, ur = side == 0 ? un : uo;
      for (int d = 0; d < D; d++) {
        const double avg = (n[d] * (ul + ur) / 2.0) * jw;
        sA[le][v * D + d][sl] = side == 0 ? avg : -avg;
        sJ[le][v * D + d][sl] = (n[d] * (ur - ul) / 2.0) * jw;
      }
    }
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NV * NQ; idx += kThreads) {
    const int le = idx / (NV * NQ), rem = idx % (NV * NQ);
    const int v = rem / NQ, q = rem % NQ;
    double s = 0.0;
    for (int b = 0; b < NB; b++) s += sU[le][v][b] * g.Vt[b * NQ + q];
    sUq[le][v][q] = s;
  }
  __syncthreads();
  // gradient residuals (SpatialDiscrete.cpp:1034-1068)
  for (int idx = tid; idx < EPB * NG * NB; idx += kThreads) {
    const int le = idx / (NG * NB), rem = idx % (NG * NB);
    const int r = rem / NB, i = rem % NB;
    const int k = r / D, aa = r % D;
    const int e = e0 + le;
    const int ec = e < g.E ? e : g.E - 1;
    double fa = 0.0, fj = 0.0;
    for (int sl = 0; sl < NAQ; sl++) {
      const double vf = g.Vf[sl * NB + i];
      fa += sA[le][r][sl] * vf;
      fj += sJ[le][r][sl] * vf;
    }
    double vol = 0.0;
    const double dj = g.detj[ec];
    for (int q = 0; q < NQ; q++) {
      const double uk = sUq[le][k][q];
      const double jw = dj * g.wq[q];
      for (int b = 0; b < D; b++) {
        const double G = g.ginv[static_cast<int64_t>(ec) * D * D + aa + D * b] * jw;
        vol += (uk * G) * g.dV[(q * D + b) * NB + i];
      }
    }
    sR1[le][r][i] = fa - vol;
    sR2[le][r][i] = (fa - vol) + fj;
  }
  __syncthreads();
  // lifted coefficients (TimeIntegration.cpp:200-228)
  for (int idx = tid; idx < EPB * NG * NB; idx += kThreads) {
    const int le = idx / (NG * NB), rem = idx % (NG * NB);
    const int r = rem / NB, i = rem % NB;
    const int e = e0 + le;
    const double idj = 1.0 / g.detj[e < g.E ? e : g.E - 1];
    double wv = 0.0, wt = 0.0;
    for (int b = 0; b < NB; b++) {
      const double m = g.Minv[b * NB + i];
      wv += sR1[le][r][b] * m;
      wt += sR2[le][r][b] * m;
    }
    sWv[le][r][i] = wv * idj;
    sWt[le][r][i] = wt * idj;
    if (a.Wout && e < g.E) a.Wout[static_cast<int64_t>(e) * NG * NB + static_cast<int64_t>(i) * NG + r] = wt * idj;
  }
  __syncthreads();
  // total gradient at volume points and lifted gradient traces at face points
  for (int idx = tid; idx < EPB * NG * NQ; idx += kThreads) {
    const int le = idx / (NG * NQ), rem = idx % (NG * NQ);
    const int r = rem / NQ, q = rem % NQ;
    double s = 0.0;
    for (int b = 0; b < NB; b++) s += sWt[le][r][b] * g.Vt[b * NQ + q];
    sR1[le][r][q] = s;
  }
  for (int idx = tid; idx < EPB * NG * NAQ; idx += kThreads) {
    const int le = idx / (NG * NAQ), rem = idx % (NG * NAQ);
    const int r = rem / NAQ, sl = rem % NAQ;
    const int e = e0 + le;
    double t = 0.0;
    if (VISC == 2) {
      for (int b = 0; b < NB; b++) t += sWv[le][r][b] * g.Vft[b * NAQ + sl];
      const int s = sl / NFQ, j = sl % NFQ;
      double l = 0.0;
      for (int pp = 0; pp < NFQ; pp++) l += sJ[le][r][s * NFQ + pp] * g.Lf[(s * NFQ + pp) * NFQ + j];
      t += l / g.detj[e < g.E ? e : g.E - 1];
    } else {
      for (int b = 0; b < NB; b++) t += sWt[le][r][b] * g.Vft[b * NAQ + sl];
    }
    if (e < g.E) a.Gf[static_cast<int64_t>(e) * NG * NAQ + rem] = t;
  }
  __syncthreads();
  // volume flux: (F_c - F_v)^T (J^T)^-1 |J| w (SpatialDiscrete.cpp:216-253)
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ;
    const int e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3], gq[NG], pg[NG], fv[D * NV];
    for (int v = 0; v < NV; v++) u[v] = sUq[le][v][q];
    for (int r = 0; r < NG; r++) gq[r] = sR1[le][r][q];
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
    double* out = &sR2[le][0][0];
    for (int v = 0; v < NV; v++)
      for (int b = 0; b < D; b++) {
        double s = 0.0;
        for (int aa = 0; aa < D; aa++) s += (fc[aa + D * v] - fv[aa + D * v]) * G[aa + D * b];
        out[v * (D * NQ) + q * D + b] = s;
      }
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int v = rem / NB, i = rem % NB;
    const int e = e0 + le;
    const double* F = &sR2[le][0][0];
    double r = 0.0;
    for (int k = 0; k < D * NQ; k++) r += F[v * (D * NQ) + k] * g.dV[k * NB + i];
    if (e < g.E) a.Rvol[static_cast<int64_t>(e) * NV * NB + rem] = r;
  }
}

// ---- Navier-Stokes, kernel 2: face fluxes + residual + update + next traces -------------------
// calculateAdjacencyQuadrature, calculateResidual, updateBasisFunctionCoefficient
// (SpatialDiscrete.cpp:633-748,1016-1032; TimeIntegration.cpp:181-198)
template <int D, int NB, int NQ, int NF, int NFQ, int EPB>
__global__ __launch_bounds__(kThreads) void nsFluxKernel(Geo g, Phys ph, StageArgs a) {
  using S = Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int NV = S::NV, NG = S::NG, NAQ = S::NAQ;
  __shared__ double sU[EPB][NV][NB];
  __shared__ double sH[EPB][NV][NAQ];
  __shared__ double sR[EPB][NV][NB];
  if (a.status && a.status[0]) return;
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int e = e0 + le;
    sU[le][rem / NB][rem % NB] = e < g.E ? a.U[static_cast<int64_t>(e) * NV * NB + rem] : 1.0;
  }
  for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
    const int le = idx / NAQ, sl = idx % NAQ;
    const int e = e0 + le;
    if (e >= g.E) {
      for (int v = 0; v < NV; v++) sH[le][v][sl] = 0.0;
      continue;
    }
    const int s = sl / NFQ, j = sl % NFQ;
    int f, p, side;
    int64_t nbslot;
    slotRoles<NF, NFQ>(g, e, s, j, f, p, nbslot, side);
    const int64_t nbel = nbslot / NAQ, nbsl = nbslot % NAQ;
    double uo[NV], un[NV], go[NG], gn[NG];
    for (int v = 0; v < NV; v++) {
      uo[v] = a.Uf_in[static_cast<int64_t>(e) * NV * NAQ + v * NAQ + sl];
      un[v] = a.Uf_in[nbel * NV * NAQ + v * NAQ + nbsl];
    }
    for (int r = 0; r < NG; r++) {
      go[r] = a.Gf[static_cast<int64_t>(e) * NG * NAQ + r * NAQ + sl];
      gn[r] = a.Gf[nbel * NG * NAQ + r * NAQ + nbsl];
    }
    const double* ul = side == 0 ? uo : un;
    const double* ur = side == 0 ? un : uo;
    const double* gl = side == 0 ? go : gn;
    const double* gr = side == 0 ? gn : go;
    double wl[D + 3], wr[D + 3], n[D], h[NV], pgl[NG], pgr[NG], fvl[D * NV], fvr[D * NV];
    computational<D>(ph, ul, wl);
    computational<D>(ph, ur, wr);
    for (int d = 0; d < D; d++) n[d] = g.fnormal[(static_cast<int64_t>(f) * NFQ + p) * D + d];
    convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
    primitiveGradient<D>(ph, wl, ul, gl, pgl);
    primitiveGradient<D>(ph, wr, ur, gr, pgr);
    viscousRawFlux<D>(ph, wl, pgl, fvl);
    viscousRawFlux<D>(ph, wr, pgr, fvr);
    const double jw = g.fjw[static_cast<int64_t>(f) * NFQ + p];
    for (int v = 0; v < NV; v++) {
      double vl = 0.0, vr = 0.0;
      for (int d = 0; d < D; d++) {
        vl += fvl[d + D * v] * n[d];
        vr += fvr[d + D * v] * n[d];
      }
      const double t = (h[v] - (vl + vr) / 2.0) * jw;
      sH[le][v][sl] = side == 0 ? t : -t;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), rem = idx % (NV * NB);
    const int v = rem / NB, i = rem % NB;
    const int e = e0 + le;
    double h = 0.0;
    for (int k = 0; k < NAQ; k++) h += sH[le][v][k] * g.Vf[k * NB + i];
    const double r = e < g.E ? a.Rvol[static_cast<int64_t>(e) * NV * NB + rem] : 0.0;
    sR[le][v][i] = r - h;
  }
  __syncthreads();
  stageTail<D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sU, sR);
}

// ---- auxiliary kernels ---------------------------------------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void traceKernel(Geo g, const double* U, double* Uf) {
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<int64_t>(g.E) * NV * NAQ) return;
  const int64_t e = idx / (NV * NAQ);
  const int rem = static_cast<int>(idx % (NV * NAQ));
  const int v = rem / NAQ, sl = rem % NAQ;
  double t = 0.0;
  for (int b = 0; b < NB; b++) t += U[e * NV * NB + v * NB + b] * g.Vft[b * NAQ + sl];
  Uf[idx] = t;
}

// CFL time step per element: min_q CFL minEdge / ((|u| + c)(P+1)^2) (TimeIntegration.cpp:104-131)
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void dtKernel(Geo g, Phys ph, const double* U, const double* minedge, double cfl,
                                                     int order, double* out) {
  constexpr int NV = D + 2;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= g.E) return;
  double best = 1.0e308;
  for (int q = 0; q < NQ; q++) {
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int b = 0; b < NB; b++) s += U[static_cast<int64_t>(e) * NV * NB + v * NB + b] * g.Vt[b * NQ + q];
      u[v] = s;
    }
    computational<D>(ph, u, w);
    const double c = soundSpeedOf(ph, w[0], w[D + 2]);
    const double sr = sqrt(velocitySquared<D>(w)) + c;
    const double t = cfl * minedge[e] / (sr * (order + 1.0) * (order + 1.0));
    best = fmin(best, t);
  }
  out[e] = best;
}

// Fixed-order reduction of the monitor partials; raises the NaN stop flag when every variable is
// NaN (SystemControl.cpp:185-191).
template <int NV>
__global__ __launch_bounds__(kThreads) void monitorKernel(const double* mon, int E, double* rel, int* status) {
  __shared__ double red[kThreads][NV];
  if (status[0]) return;
  const int tid = threadIdx.x;
  double acc[NV];
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  const int chunk = (E + kThreads - 1) / kThreads;
  const int lo = tid * chunk, hi = lo + chunk < E ? lo + chunk : E;
  for (int e = lo; e < hi; e++)
    for (int v = 0; v < NV; v++) acc[v] += mon[static_cast<int64_t>(e) * NV + v];
  for (int v = 0; v < NV; v++) red[tid][v] = acc[v];
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (tid < st)
      for (int v = 0; v < NV; v++) red[tid][v] += red[tid + st][v];
    __syncthreads();
  }
  if (tid == 0) {
    bool all_nan = true;
    for (int v = 0; v < NV; v++) {
      const double r = red[0][v] / static_cast<double>(E);
      rel[v] = r;
      all_nan = all_nan && isnan(r);
    }
    status[1] += 1;
    if (all_nan) status[0] = 1;
  }
}

__global__ void resetStatusKernel(int* status) {
  if (threadIdx.x == 0) {
    status[0] = 0;
    status[1] = 0;
  }
}

}  // namespace sdg

#endif  // SDG_DG_KERNELS_HPP_
