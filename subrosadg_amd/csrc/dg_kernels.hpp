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
template <typename TC>
struct MOpsT {
  const TC* interp;  // K = N_b,   M = N_q   : basis values          (U -> U_q, w -> grad U_q, R -> R_q)
  const TC* grad;    // K = N_q D, M = N_b   : basis gradients       (weak divergence, volume gradient term)
  const TC* face;    // K = N_aq,  M = N_b   : face values           (face terms of residuals)
  const TC* minv;    // K = N_b,   M = N_b   : reference mass inverse
  const TC* trace;   // K = N_b,   M = N_aq  : face values           (coefficients -> face traces)
  const TC* lift;    // K = N_aq,  M = N_aq  : block-diagonal Vf_s Mref^-1 Vf_s^T (BR2 face lift)
  const TC* pinterp; // K = N_b,   M = N_q   : Mref^-1 V^T   (R -> gradient at volume points)
  const TC* ptrace;  // K = N_b,   M = N_aq  : Mref^-1 Vf^T  (R -> gradient traces at face points)
  const TC* rinterp; // K = N_b,   M = N_q   : residual -> residual at volume points (monitor)
};
using MOps = MOpsT<double>;

// Sum-factorisation tables of a hexahedron order (see dg_tensor.hpp).
constexpr int kTensorMaxN = 6;

// 1-D operators of the hexahedron of one order (by value in the kernel arguments)
struct TensorOps {
  double B[kTensorMaxN * kTensorMaxN];     // B[i N + k] = l_k(x_i)
  double Binv[kTensorMaxN * kTensorMaxN];  // (B^-1)[k N + i]
  double Dn[kTensorMaxN * kTensorMaxN];    // Dn[i N + m] = d/dxi Lagrange_m (x_i)
  double lf[2 * kTensorMaxN];              // lf[s N + i] = Lagrange_i(-1) (s = 0), (+1) (s = 1)
  double w[kTensorMaxN];                   // Gauss weights
  double kappa[2];                         // sum_i lf[s][i]^2 / w_i
  int axis[6], side[6];                    // local face -> normal axis (x=0, y=1, z=2), side (0: -1, 1: +1)
  const short* perm;                       // modal index -> tensor index kx + N ky + N^2 kz
  const short* iperm;                      // tensor index -> modal index
  const short* slot_of;                    // [face][node] -> face point of the node's tangential position
  const short* base;                       // [face][face point] -> node with normal index 0
  const double* fwt;                       // [face][face point] face Gauss weight (tangential w w)
  const double* nodt;                      // artificial viscosity: [node][vertex] P1 nodal functions at the
                                           // tensor-ordered Gauss nodes (else null)
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
  const int* snb;         // E x N_aq       neighbour slot: neighbour element * N_aq + its slot
  const int* sfp;         // E x N_aq       2 (f N_fq + p) + side: face point of the slot and side
  // sum-factorised kernels only (else null): per element slot, the outward normal of the element
  // at the slot's face point and the point's |J_f| w, so that they are one load away
  const double* snrm;     // E x N_aq x D
  const double* sjw;      // E x N_aq
  const int* fpar;        // F x 2          left, right parent element
  const int* floc;        // F x 2          local face index in left, right parent
  // boundary face points bp = b N_fq + p (a slot on a boundary face has snb = -1 - bp)
  const double* bnormal;  // F_b N_fq x D    outward normal
  const double* bjw;      // F_b N_fq        |J_f| w
  const int* btype;       // F_b N_fq        sdg_boundary type
  const int* bslot;       // F_b N_fq        own slot: element * N_aq + local slot
  const double* bdata;    // F_b N_fq x (D+3) boundary data, computational variables
  int E;
  int F;
  int Fb;
  MOps op;                // FP64 operator tables
  MOpsT<float> opf;       // the same tables rounded to FP32 (mixed-precision mode)
  TensorOps tens;         // hexahedra: 1-D operators of the sum-factorised kernels
  // mixed-precision mode: the residual-to-update map L^-T of opf.minv kept in FP64, so the update's
  // hierarchic coefficients are formed without FP32 rounding of their (cancelling) terms
  const double* minv_mixed64;
  // artificial viscosity: per element the node of each vertex and the inner radius, the P1 nodal
  // functions at the volume and face points, and the node values of the current step
  const int* ntag;        // E x nvert
  const double* inrad;    // E
  const double* nodq;     // N_q x nvert
  const double* nodf;     // N_aq x nvert
  const double* nodeav;   // nodes
  int nvert;
  int order;              // P
  // curved (non-affine) elements, else null (vector kernels only): per element and volume point
  // (J^T)^-1 |J| w (column-major D x D) and |J| w, the element mass inverse (Geometry.cpp:88-100)
  // and the BR2 face blocks Vf_s M_e^-1 Vf_s^T
  const double* qG;       // E x N_q x D*D
  const double* qjw;      // E x N_q
  const double* minvE;    // E x N_b x N_b
  const double* LfE;      // E x N_f x N_fq x N_fq
};

// |J| w at volume point q of element e, and the (a, b) entry (i = a + D b) of (J^T)^-1 |J| w given
// that value: from the element's constant Jacobian, or from the per-point tables of curved elements
template <int NQ>
__device__ __forceinline__ double pointJw(const Geo& g, int e, int q) {
  return g.qjw ? g.qjw[static_cast<int64_t>(e) * NQ + q] : g.detj[e] * g.wq[q];
}
template <int D, int NQ>
__device__ __forceinline__ double pointG(const Geo& g, int e, int q, int i, double jw) {
  return g.qG ? g.qG[(static_cast<int64_t>(e) * NQ + q) * D * D + i] : g.ginv[static_cast<int64_t>(e) * D * D + i] * jw;
}

// P1 interpolant of the node values of element e at row `row` of a nodal table
__device__ __forceinline__ double avAt(const Geo& g, const double* table, int row, int e) {
  double s = 0.0;
  for (int j = 0; j < g.nvert; j++) s += table[row * g.nvert + j] * g.nodeav[g.ntag[e * g.nvert + j]];
  return s;
}

template <typename TC>
__device__ __forceinline__ const MOpsT<TC>& opsOf(const Geo& g) {
  if constexpr (sizeof(TC) == 8)
    return g.op;
  else
    return g.opf;
}

// Mixed-precision mode: the contractions work on the deviation dU = U - uref[v] one[b] from a
// constant reference state (one[] = modal coefficients of the constant function 1), which the DG
// operators annihilate exactly (the lifted gradient of a constant is zero, and the weak divergence
// of a constant flux is zero on affine elements with exact quadrature). For the same reason the
// reference pressure pref is taken out of the momentum fluxes, and the energy fluxes carry
// F_E - h0 F_rho with the reference total enthalpy h0: the residual is linear in the fluxes, so
// R_E = R'_E + h0 R_rho is restored in FP64 after the contractions. Without this, FP32 would round
// the O(p0) = O(1/Ma^2) parts of the momentum and energy fluxes, whose divergence nearly cancels
// at low Mach number. Unused (on = 0) in FP64 mode.
struct MixRef {
  double uref[5];
  double pref;
  double h0;
  const double* one;
  int on;
};

struct StageArgs {
  double* U;
  double* Un;
  const double* Uf_in;
  double* Uf_out;
  double* Gf;
  double* GfV;    // Navier-Stokes with artificial viscosity: E x (NV D) x N_aq volume-gradient traces
  double* Rvol;   // (sum-factorised path: the weak divergence of the volume flux, FP32 values in
                  // the mixed mode)
  double* Rout;   // residual output (mode 1)
  double* Wout;   // lifted total gradient output (optional)
  double* Hs;     // E x NV x N_aq signed face-flux slots (Navier-Stokes, matrix-core path)
  double* Fv;     // E x (N_q D) x NV contravariant volume flux (Navier-Stokes, matrix-core path)
  const double* Rsrc;  // E x N_b modal source residual of the last momentum component (null: none;
                       // the sum-factorised kernels form the source at their nodes instead)
  double* Vn;     // E x NV x N_aq the element's own normal viscous flux F_v . n_out at its face
                  // points (sum-factorised path, else null; holds FP32 values in the mixed mode)
  // FP32 versions of the work buffers above (mixed-precision mode; traces hold the deviation)
  const float* Uf_in32;
  float* Uf_out32;
  float* Gf32;
  float* Hs32;
  float* Fv32;
  double* mon;    // E x NV monitor partials (last stage)
  int* status;    // [0] NaN stop flag, [1] steps taken; nullptr: no check
  double c0, c1, c2dt;  // U <- c1 U + c0 Un + c2dt R Minv (TimeIntegration.cpp:181-198)
  int last_stage;
  int mode;  // 0: stage update, 1: residual only
  // recording (matrix-core stage kernels): the updated state is also written to pool slot
  // (step, stage + 1), or (step + 1, 0) after the last stage, with step = status[1]; slots of
  // rec_S * state-size doubles per step, rec_nsteps steps
  double* rec;
  double* rec_t;  // optional: the traces of the updated state, slots of rec_S * trace-size doubles
  int rec_stage, rec_S, rec_nsteps;
  int64_t rec_n, rec_nt;
  MixRef mix;
};

// the work buffer of contraction type TC
template <typename TC>
__device__ __forceinline__ TC* pick(double* d, float* f) {
  if constexpr (sizeof(TC) == 8)
    return d;
  else
    return f;
}
template <typename TC>
__device__ __forceinline__ const TC* pick(const double* d, const float* f) {
  if constexpr (sizeof(TC) == 8)
    return d;
  else
    return f;
}

// Navier-Stokes / Euler flux t at boundary point bp from the interior trace ul (conserved) and, for
// Navier-Stokes, its lifted gradient gl: calculateBoundaryAdjacencyElementQuadrature
// (SpatialDiscrete.cpp:750-842): the physical flux of the boundary state (no Riemann solver), and
// the central viscous flux of the (wall-modified) interior state and the boundary state, whose
// primitive gradient is the interior one with zero temperature gradient on adiabatic walls. With
// mixed = true the numerical flux is reduced as in the mixed-precision interior faces.
template <int D, bool VISC>
__device__ __forceinline__ void boundaryFlux(const Phys& ph, const Geo& g, int bp, const double* ul, const double* gl,
                                             bool mixed, double pref, double h0, double* t) {
  constexpr int NV = D + 2, NG = NV * D;
  const int bc = g.btype[bp];
  double n[D], wl[D + 3], wb[D + 3], wd[D + 3], h[NV];
  for (int d = 0; d < D; d++) n[d] = g.bnormal[static_cast<int64_t>(bp) * D + d];
  for (int i = 0; i < D + 3; i++) wd[i] = g.bdata[static_cast<int64_t>(bp) * (D + 3) + i];
  computational<D>(ph, ul, wl);
  boundaryState<D>(ph, bc, n, wl, wd, wb);
  normalFlux<D>(n, wb, h);
  if (mixed) {
    for (int d = 0; d < D; d++) h[1 + d] -= pref * n[d];
    h[D + 1] -= h0 * h[0];
  }
  const double jw = g.bjw[bp];
  if constexpr (VISC) {
    double pgl[NG], pgb[NG], wm[D + 3], fvl[D * NV], fvb[D * NV];
    primitiveGradient<D>(ph, wl, ul, gl, pgl);
    for (int r = 0; r < NG; r++) pgb[r] = pgl[r];
    if (bc == 4 || bc == 5)
      for (int aa = 0; aa < D; aa++) pgb[(D + 1) * D + aa] = 0.0;
    for (int i = 0; i < D + 3; i++) wm[i] = bc >= 3 ? wb[i] : wl[i];  // modifyBoundaryVariable
    viscousRawFlux<D>(ph, wm, pgl, fvl);
    viscousRawFlux<D>(ph, wb, pgb, fvb);
    for (int v = 0; v < NV; v++) {
      double vl = 0.0, vr = 0.0;
      for (int d = 0; d < D; d++) {
        vl += fvl[d + D * v] * n[d];
        vr += fvb[d + D * v] * n[d];
      }
      t[v] = (h[v] - (vl + vr) / 2.0) * jw;
    }
  } else {
    for (int v = 0; v < NV; v++) t[v] = h[v] * jw;
  }
}

// Gradient-pass face terms at boundary point bp (calculateBoundaryAdjacencyElementGardientQuadrature,
// SpatialDiscrete.cpp:913-968): A = n (x) (u_vg - shift) |J_f| w, J = n (x) u_ig |J_f| w, component v
// (shift: the reference state of the mixed-precision deviation form, else 0).
template <int D>
__device__ __forceinline__ void boundaryGradientTerms(const Phys& ph, const Geo& g, int bp, const double* ul, int v,
                                                      double shift, double* A, double* J) {
  constexpr int NV = D + 2;
  double n[D], wd[D + 3], uvg[NV], uig[NV];
  for (int d = 0; d < D; d++) n[d] = g.bnormal[static_cast<int64_t>(bp) * D + d];
  for (int i = 0; i < D + 3; i++) wd[i] = g.bdata[static_cast<int64_t>(bp) * (D + 3) + i];
  boundaryGradientStates<D>(ph, g.btype[bp], n, ul, wd, uvg, uig);
  const double jw = g.bjw[bp];
  for (int d = 0; d < D; d++) {
    A[d] = (n[d] * (uvg[v] - shift)) * jw;
    J[d] = (n[d] * uig[v]) * jw;
  }
}

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
    if (g.minvE) {
      const double* M = g.minvE + static_cast<int64_t>(e) * NB * NB;
      for (int b = 0; b < NB; b++) rm += sR[le][v][b] * M[b * NB + i];
    } else {
      for (int b = 0; b < NB; b++) rm += sR[le][v][b] * g.Minv[b * NB + i];
      rm = rm / g.detj[e];
    }
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
    const double jw = pointJw<NQ>(g, e, q);
    double G[D * D];
    for (int i = 0; i < D * D; i++) G[i] = pointG<D, NQ>(g, e, q, i, jw);
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
    const int64_t own = static_cast<int64_t>(e) * NV * NAQ + sl;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es];
    if (nbs < 0) {  // boundary face point
      double uo[NV], t[NV];
      for (int v = 0; v < NV; v++) uo[v] = a.Uf_in[own + v * NAQ];
      boundaryFlux<D, false>(ph, g, -1 - nbs, uo, nullptr, false, 0.0, 0.0, t);
      for (int v = 0; v < NV; v++) sH[le][v][sl] = t[v];
      continue;
    }
    const int fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    // select the left/right index rather than a pointer to a private array (which would put the
    // arrays in scratch memory)
    const int64_t nbe = static_cast<int64_t>(nbs / NAQ) * NV * NAQ + nbs % NAQ;
    const int64_t li = side == 0 ? own : nbe, ri = side == 0 ? nbe : own;
    double ul[NV], ur[NV];
    for (int v = 0; v < NV; v++) {
      ul[v] = a.Uf_in[li + v * NAQ];
      ur[v] = a.Uf_in[ri + v * NAQ];
    }
    double wl[D + 3], wr[D + 3], n[D], h[NV];
    computational<D>(ph, ul, wl);
    computational<D>(ph, ur, wr);
    for (int d = 0; d < D; d++) n[d] = g.fnormal[fp * D + d];
    convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
    const double jw = g.fjw[fp];
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
    const int e = e0 + le;
    sR[le][v][i] = (v == D && a.Rsrc && e < g.E) ? r - h + a.Rsrc[static_cast<int64_t>(e) * NB + i] : r - h;
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
// VISC = 0: Euler with artificial viscosity (the traces and the flux use the volume gradient);
// VISC = 1, 2 with AV: Navier-Stokes with artificial viscosity (the volume-gradient traces go to
// GfV, the volume flux also subtracts eps grad U with the volume gradient).
template <int D, int NB, int NQ, int NF, int NFQ, int EPB, int VISC, bool AV = false>
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
    const int64_t own = static_cast<int64_t>(e) * NV * NAQ + sl;
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es];
    if (nbs < 0) {  // boundary face point
      double uo[NV];
      for (int v = 0; v < NV; v++) uo[v] = a.Uf_in[own + v * NAQ];
      for (int v = 0; v < NV; v++) {
        double Ab[D], Jb[D];
        boundaryGradientTerms<D>(ph, g, -1 - nbs, uo, v, 0.0, Ab, Jb);
        for (int d = 0; d < D; d++) {
          sA[le][v * D + d][sl] = Ab[d];
          sJ[le][v * D + d][sl] = Jb[d];
        }
      }
      continue;
    }
    const int fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const int64_t nbe = static_cast<int64_t>(nbs / NAQ) * NV * NAQ + nbs % NAQ;
    double n[D];
    for (int d = 0; d < D; d++) n[d] = g.fnormal[fp * D + d];
    const double jw = g.fjw[fp];
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
    for (int q = 0; q < NQ; q++) {
      const double uk = sUq[le][k][q];
      const double jw = pointJw<NQ>(g, ec, q);
      for (int b = 0; b < D; b++) {
        const double G = pointG<D, NQ>(g, ec, q, aa + D * b, jw);
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
    const int ec = e < g.E ? e : g.E - 1;
    double wv = 0.0, wt = 0.0;
    if (g.minvE) {  // curved: the element's own mass inverse
      const double* M = g.minvE + static_cast<int64_t>(ec) * NB * NB;
      for (int b = 0; b < NB; b++) {
        wv += sR1[le][r][b] * M[b * NB + i];
        wt += sR2[le][r][b] * M[b * NB + i];
      }
    } else {
      const double idj = 1.0 / g.detj[ec];
      for (int b = 0; b < NB; b++) {
        const double m = g.Minv[b * NB + i];
        wv += sR1[le][r][b] * m;
        wt += sR2[le][r][b] * m;
      }
      wv *= idj;
      wt *= idj;
    }
    sWv[le][r][i] = wv;
    sWt[le][r][i] = wt;
    if (a.Wout && e < g.E) a.Wout[static_cast<int64_t>(e) * NG * NB + static_cast<int64_t>(i) * NG + r] = wt;
  }
  __syncthreads();
  // total gradient at volume points and lifted gradient traces at face points (VISC == 0,
  // artificial viscosity: the volume gradient, get<ViscousFluxEnum::None>)
  for (int idx = tid; idx < EPB * NG * NQ; idx += kThreads) {
    const int le = idx / (NG * NQ), rem = idx % (NG * NQ);
    const int r = rem / NQ, q = rem % NQ;
    double s = 0.0;
    for (int b = 0; b < NB; b++) s += (VISC == 0 ? sWv[le][r][b] : sWt[le][r][b]) * g.Vt[b * NQ + q];
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
      const int ec = e < g.E ? e : g.E - 1;
      double l = 0.0;
      if (g.LfE) {  // curved: the element's own face blocks
        const double* Lb = g.LfE + static_cast<int64_t>(ec) * NAQ * NFQ;
        for (int pp = 0; pp < NFQ; pp++) l += sJ[le][r][s * NFQ + pp] * Lb[(s * NFQ + pp) * NFQ + j];
        t += l;
      } else {
        for (int pp = 0; pp < NFQ; pp++) l += sJ[le][r][s * NFQ + pp] * g.Lf[(s * NFQ + pp) * NFQ + j];
        t += l / g.detj[ec];
      }
    } else if (VISC == 0) {
      for (int b = 0; b < NB; b++) t += sWv[le][r][b] * g.Vft[b * NAQ + sl];
    } else {
      for (int b = 0; b < NB; b++) t += sWt[le][r][b] * g.Vft[b * NAQ + sl];
    }
    if (e < g.E) a.Gf[static_cast<int64_t>(e) * NG * NAQ + rem] = t;
    if constexpr (AV && VISC != 0) {
      double tv = 0.0;
      for (int b = 0; b < NB; b++) tv += sWv[le][r][b] * g.Vft[b * NAQ + sl];
      if (e < g.E) a.GfV[static_cast<int64_t>(e) * NG * NAQ + rem] = tv;
    }
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
    if constexpr (VISC != 0) {
      primitiveGradient<D>(ph, w, u, gq, pg);
      viscousRawFlux<D>(ph, w, pg, fv);
    }
    const double rho = w[0], p = w[D + 2];
    const double et = w[D + 1] + velocitySquared<D>(w) / 2.0;
    double fc[D * NV];
    for (int aa = 0; aa < D; aa++) {
      fc[aa] = rho * w[1 + aa];
      for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + (aa == b ? p * 1.0 : p * 0.0);
      fc[aa + D * (D + 1)] = (rho * et + p) * w[1 + aa];
    }
    const double jw = pointJw<NQ>(g, e, q);
    double G[D * D];
    for (int i = 0; i < D * D; i++) G[i] = pointG<D, NQ>(g, e, q, i, jw);
    double* out = &sR2[le][0][0];
    if constexpr (VISC == 0) {  // F_c^T G - (eps grad U)^T G (SpatialDiscrete.cpp:211-249)
      const double eps = avAt(g, g.nodq, q, e);
      for (int v = 0; v < NV; v++)
        for (int b = 0; b < D; b++) {
          double s = 0.0, t = 0.0;
          for (int aa = 0; aa < D; aa++) s += fc[aa + D * v] * G[aa + D * b];
          for (int aa = 0; aa < D; aa++) t += (eps * gq[v * D + aa]) * G[aa + D * b];
          out[v * (D * NQ) + q * D + b] = s - t;
        }
    } else if constexpr (AV) {  // (F_c - F_v)^T G - (eps grad U)^T G with the volume gradient
      const double eps = avAt(g, g.nodq, q, e);
      double gv[NG];
      for (int r = 0; r < NG; r++) {
        double sv = 0.0;
        for (int b = 0; b < NB; b++) sv += sWv[le][r][b] * g.Vt[b * NQ + q];
        gv[r] = sv;
      }
      for (int v = 0; v < NV; v++)
        for (int b = 0; b < D; b++) {
          double s = 0.0, t = 0.0;
          for (int aa = 0; aa < D; aa++) s += (fc[aa + D * v] - fv[aa + D * v]) * G[aa + D * b];
          for (int aa = 0; aa < D; aa++) t += (eps * gv[v * D + aa]) * G[aa + D * b];
          out[v * (D * NQ) + q * D + b] = s - t;
        }
    } else {
      for (int v = 0; v < NV; v++)
        for (int b = 0; b < D; b++) {
          double s = 0.0;
          for (int aa = 0; aa < D; aa++) s += (fc[aa + D * v] - fv[aa + D * v]) * G[aa + D * b];
          out[v * (D * NQ) + q * D + b] = s;
        }
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

// ---- raw files of Navier-Stokes BR2 runs: the gradient blocks of boundary faces -------------
// RawBinary.cpp:93-110 writes, for every boundary face, its parent's volume-gradient coefficients
// plus the BR2 lift of that face, (NV D) x N_b column-major (row var*D+dim). One workgroup per
// boundary face, FP64, from the state U: the parent's own and its neighbours' traces are formed
// here from U, the face terms and the lifting follow nsGradKernel:
//   w_vol = (A Phi_f - Q Phi_grad) Mref^-1 / |J|,   r_s = J_s Phi_f,s Mref^-1 / |J|.
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void br2BoundaryBlockKernel(Geo g, Phys ph, const double* U, double* out) {
  constexpr int NV = D + 2, NG = NV * D, NAQ = NF * NFQ;
  __shared__ double sU[NV][NB];
  __shared__ double sUq[NV][NQ];
  __shared__ double sA[NG][NAQ];
  __shared__ double sJ[NG][NFQ];
  __shared__ double sRv[NG][NB];
  __shared__ double sRl[NG][NB];
  const int tid = threadIdx.x;
  const int bf = blockIdx.x;
  const int own0 = g.bslot[static_cast<int64_t>(bf) * NFQ];  // parent * N_aq + slot of the face's point 0
  const int e = own0 / NAQ, s = (own0 % NAQ) / NFQ;
  for (int idx = tid; idx < NV * NB; idx += kThreads) sU[idx / NB][idx % NB] = U[static_cast<int64_t>(e) * NV * NB + idx];
  __syncthreads();
  for (int idx = tid; idx < NV * NQ; idx += kThreads) {
    const int v = idx / NQ, q = idx % NQ;
    double acc = 0.0;
    for (int b = 0; b < NB; b++) acc += sU[v][b] * g.Vt[b * NQ + q];
    sUq[v][q] = acc;
  }
  for (int sl = tid; sl < NAQ; sl += kThreads) {
    double uo[NV];
    for (int v = 0; v < NV; v++) {
      double acc = 0.0;
      for (int b = 0; b < NB; b++) acc += sU[v][b] * g.Vf[sl * NB + b];
      uo[v] = acc;
    }
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es];
    const bool lifted = sl / NFQ == s;
    if (nbs < 0) {
      for (int v = 0; v < NV; v++) {
        double Ab[D], Jb[D];
        boundaryGradientTerms<D>(ph, g, -1 - nbs, uo, v, 0.0, Ab, Jb);
        for (int d = 0; d < D; d++) {
          sA[v * D + d][sl] = Ab[d];
          if (lifted) sJ[v * D + d][sl % NFQ] = Jb[d];
        }
      }
      continue;
    }
    const int fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const int ne = nbs / NAQ, nsl = nbs % NAQ;
    double n[D];
    for (int d = 0; d < D; d++) n[d] = g.fnormal[fp * D + d];
    const double jw = g.fjw[fp];
    for (int v = 0; v < NV; v++) {
      double un = 0.0;
      for (int b = 0; b < NB; b++) un += U[(static_cast<int64_t>(ne) * NV + v) * NB + b] * g.Vf[nsl * NB + b];
      const double ul = side == 0 ? uo[v] : un, ur = side == 0 ? un : uo[v];
      for (int d = 0; d < D; d++) {
        const double avg = (n[d] * (ul + ur) / 2.0) * jw;
        sA[v * D + d][sl] = side == 0 ? avg : -avg;
        if (lifted) sJ[v * D + d][sl % NFQ] = (n[d] * (ur - ul) / 2.0) * jw;
      }
    }
  }
  __syncthreads();
  for (int idx = tid; idx < NG * NB; idx += kThreads) {
    const int r = idx / NB, i = idx % NB;
    const int k = r / D, aa = r % D;
    double fa = 0.0, fj = 0.0;
    for (int sl = 0; sl < NAQ; sl++) fa += sA[r][sl] * g.Vf[sl * NB + i];
    for (int j = 0; j < NFQ; j++) fj += sJ[r][j] * g.Vf[(s * NFQ + j) * NB + i];
    double vol = 0.0;
    for (int q = 0; q < NQ; q++) {
      const double uk = sUq[k][q];
      const double jw = pointJw<NQ>(g, e, q);
      for (int b = 0; b < D; b++) {
        const double G = pointG<D, NQ>(g, e, q, aa + D * b, jw);
        vol += (uk * G) * g.dV[(q * D + b) * NB + i];
      }
    }
    sRv[r][i] = fa - vol;
    sRl[r][i] = fj;
  }
  __syncthreads();
  for (int idx = tid; idx < NG * NB; idx += kThreads) {
    const int r = idx / NB, i = idx % NB;
    double wv = 0.0, wl = 0.0;
    if (g.minvE) {
      const double* M = g.minvE + static_cast<int64_t>(e) * NB * NB;
      for (int b = 0; b < NB; b++) {
        wv += sRv[r][b] * M[b * NB + i];
        wl += sRl[r][b] * M[b * NB + i];
      }
    } else {
      const double idj = 1.0 / g.detj[e];
      for (int b = 0; b < NB; b++) {
        const double m = g.Minv[b * NB + i];
        wv += sRv[r][b] * m;
        wl += sRl[r][b] * m;
      }
      wv *= idj;
      wl *= idj;
    }
    out[static_cast<int64_t>(bf) * NG * NB + static_cast<int64_t>(i) * NG + r] = wv + wl;
  }
}

// ---- Navier-Stokes, kernel 2: face fluxes + residual + update + next traces -------------------
// calculateAdjacencyQuadrature, calculateResidual, updateBasisFunctionCoefficient
// (SpatialDiscrete.cpp:633-748,1016-1032; TimeIntegration.cpp:181-198)
// AV = 1: Euler with artificial viscosity, where Gf holds the volume gradient; the face flux is
// the convective one minus the central average of eps grad U . n (boundary faces: the element's
// own value), SpatialDiscrete.cpp:714-735,812-830. AV = 2: Navier-Stokes with artificial
// viscosity, the same term from the volume-gradient traces GfV after the viscous flux.
template <int D, int NB, int NQ, int NF, int NFQ, int EPB, int AV = 0>
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
    const int64_t es = static_cast<int64_t>(e) * NAQ + sl;
    const int nbs = g.snb[es];
    if (nbs < 0) {  // boundary face point
      double uo[NV], go[NG], t[NV];
      for (int v = 0; v < NV; v++) uo[v] = a.Uf_in[static_cast<int64_t>(e) * NV * NAQ + v * NAQ + sl];
      for (int r = 0; r < NG; r++) go[r] = a.Gf[static_cast<int64_t>(e) * NG * NAQ + r * NAQ + sl];
      const int bp = -1 - nbs;
      boundaryFlux<D, AV != 1>(ph, g, bp, uo, go, false, 0.0, 0.0, t);
      if constexpr (AV != 0) {
        const double eps = avAt(g, g.nodf, sl, e), jw = g.bjw[bp];
        const double* gsrc = AV == 2 ? a.GfV : a.Gf;
        for (int v = 0; v < NV; v++) {
          double f = 0.0;
          for (int d = 0; d < D; d++)
            f += (eps * gsrc[static_cast<int64_t>(e) * NG * NAQ + (v * D + d) * NAQ + sl]) * g.bnormal[static_cast<int64_t>(bp) * D + d];
          t[v] -= f * jw;
        }
      }
      for (int v = 0; v < NV; v++) sH[le][v][sl] = t[v];
      continue;
    }
    const int fps = g.sfp[es];
    const int side = fps & 1;
    const int64_t fp = fps >> 1;
    const int64_t nbel = nbs / NAQ, nbsl = nbs % NAQ;
    const int64_t lo = side == 0 ? static_cast<int64_t>(e) : nbel, ro = side == 0 ? nbel : static_cast<int64_t>(e);
    const int lsl = side == 0 ? sl : nbsl, rsl = side == 0 ? nbsl : sl;
    double ul[NV], ur[NV], gl[NG], gr[NG];
    for (int v = 0; v < NV; v++) {
      ul[v] = a.Uf_in[lo * NV * NAQ + v * NAQ + lsl];
      ur[v] = a.Uf_in[ro * NV * NAQ + v * NAQ + rsl];
    }
    for (int r = 0; r < NG; r++) {
      gl[r] = a.Gf[lo * NG * NAQ + r * NAQ + lsl];
      gr[r] = a.Gf[ro * NG * NAQ + r * NAQ + rsl];
    }
    double wl[D + 3], wr[D + 3], n[D], h[NV], pgl[NG], pgr[NG], fvl[D * NV], fvr[D * NV];
    computational<D>(ph, ul, wl);
    computational<D>(ph, ur, wr);
    for (int d = 0; d < D; d++) n[d] = g.fnormal[fp * D + d];
    convectiveFlux<D>(ph, n, ul, wl, ur, wr, h);
    if constexpr (AV == 1) {
      const double el = avAt(g, g.nodf, lsl, static_cast<int>(lo)), er = avAt(g, g.nodf, rsl, static_cast<int>(ro));
      const double jw = g.fjw[fp];
      for (int v = 0; v < NV; v++) {
        double al = 0.0, ar = 0.0;
        for (int d = 0; d < D; d++) al += (el * gl[v * D + d]) * n[d];
        for (int d = 0; d < D; d++) ar += (er * gr[v * D + d]) * n[d];
        double t = h[v] * jw;
        t -= ((al + ar) / 2.0) * jw;
        sH[le][v][sl] = side == 0 ? t : -t;
      }
      continue;
    }
    primitiveGradient<D>(ph, wl, ul, gl, pgl);
    primitiveGradient<D>(ph, wr, ur, gr, pgr);
    viscousRawFlux<D>(ph, wl, pgl, fvl);
    viscousRawFlux<D>(ph, wr, pgr, fvr);
    const double jw = g.fjw[fp];
    double el = 0.0, er = 0.0;
    if constexpr (AV == 2) {
      el = avAt(g, g.nodf, lsl, static_cast<int>(lo));
      er = avAt(g, g.nodf, rsl, static_cast<int>(ro));
    }
    for (int v = 0; v < NV; v++) {
      double vl = 0.0, vr = 0.0;
      for (int d = 0; d < D; d++) {
        vl += fvl[d + D * v] * n[d];
        vr += fvr[d + D * v] * n[d];
      }
      double t = (h[v] - (vl + vr) / 2.0) * jw;
      if constexpr (AV == 2) {  // minus the central average of eps grad U . n (volume gradient)
        double al = 0.0, ar = 0.0;
        for (int d = 0; d < D; d++) al += (el * a.GfV[lo * NG * NAQ + (v * D + d) * NAQ + lsl]) * n[d];
        for (int d = 0; d < D; d++) ar += (er * a.GfV[ro * NG * NAQ + (v * D + d) * NAQ + rsl]) * n[d];
        t -= ((al + ar) / 2.0) * jw;
      }
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
    sR[le][v][i] = (v == D && a.Rsrc && e < g.E) ? r - h + a.Rsrc[static_cast<int64_t>(e) * NB + i] : r - h;
  }
  __syncthreads();
  stageTail<D, NB, NQ, NF, NFQ, EPB>(g, a, e0, sU, sR);
}

// ---- artificial viscosity (SpatialDiscrete.cpp:37-192) ----------------------------------------
__global__ __launch_bounds__(kThreads) void avClearKernel(double* node, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    node[i] = 0.0;
}

// One thread per element: the density sensor s = log10(sum_q rho_h^2 |J|w / sum_q rho^2 |J|w) with
// rho_h the part of the modes above order P-1, the element value eps_e, and its maximum into the
// element's nodes (non-negative doubles order as their bit patterns, so the maximum is an integer
// atomic on the bits).
template <int D, int NB, int NQ>
__global__ __launch_bounds__(kThreads) void avSensorKernel(Geo g, Phys ph, const double* U, double* node) {
  constexpr int NV = D + 2;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= g.E) return;
  const double* u = U + static_cast<int64_t>(e) * NV * NB;  // density row: the first N_b entries
  const double dj = g.detj[e];
  double num = 0.0, den = 0.0;
  for (int q = 0; q < NQ; q++) {
    double all = 0.0, high = 0.0;
    for (int b = 0; b < NB; b++) all += g.Vt[b * NQ + q] * u[b];
    for (int b = ph.av_low; b < NB; b++) high += g.Vt[b * NQ + q] * u[b];
    const double jw = dj * g.wq[q];
    num += high * (high * jw);
    den += all * (all * jw);
  }
  const double s = log10(num / den);
  const double top = ph.av_factor * (g.inrad[e] / static_cast<double>(g.order));
  double eps;
  if (s < ph.av_s0 - ph.av_tol)
    eps = 0.0;
  else if (s > ph.av_s0 + ph.av_tol)
    eps = top;
  else
    eps = top * (1.0 + sin(3.14159265358979323846 * (s - ph.av_s0) / (2.0 * ph.av_tol))) / 2.0;
  const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(eps));
  for (int j = 0; j < g.nvert; j++)
    atomicMax(reinterpret_cast<unsigned long long*>(node + g.ntag[e * g.nvert + j]), bits);
}

// ---- auxiliary kernels ---------------------------------------------------------------------
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void traceKernel(Geo g, const double* U, void* Uf_out, MixRef mix) {
  constexpr int NV = D + 2, NAQ = NF * NFQ;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<int64_t>(g.E) * NV * NAQ) return;
  const int64_t e = idx / (NV * NAQ);
  const int rem = static_cast<int>(idx % (NV * NAQ));
  const int v = rem / NAQ, sl = rem % NAQ;
  double t = 0.0;
  if (mix.on) {
    for (int b = 0; b < NB; b++)
      t += (U[e * NV * NB + v * NB + b] - mix.uref[v] * mix.one[b]) * g.Vft[b * NAQ + sl];
    static_cast<float*>(Uf_out)[idx] = static_cast<float>(t);
  } else {
    for (int b = 0; b < NB; b++) t += U[e * NV * NB + v * NB + b] * g.Vft[b * NAQ + sl];
    static_cast<double*>(Uf_out)[idx] = t;
  }
}

// CFL time step per element: min_q CFL minEdge / ((|u| + c)(P+1)^2) (TimeIntegration.cpp:104-131)
// CFL time step per element (TimeIntegration.cpp:104-131): min over its volume points of
// cfl minEdge / ((|u| + c) (P+1)^2). EPB elements per workgroup: the coefficient blocks are staged
// in LDS with coalesced loads, one thread per point, then a sequential minimum per element.
template <int D, int NB, int NQ>
struct DtShape {
  static constexpr int EPB = kThreads / NQ < 1 ? 1 : (kThreads / NQ > 16 ? 16 : kThreads / NQ);
};
template <int D, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void dtKernel(Geo g, Phys ph, const double* U, const double* minedge, double cfl,
                                                     int order, double* out) {
  constexpr int NV = D + 2, EPB = DtShape<D, NB, NQ>::EPB;
  __shared__ double sU[EPB][NV * NB];
  __shared__ double sT[EPB][NQ];
  const int tid = threadIdx.x;
  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * EPB;
  for (int idx = tid; idx < EPB * NV * NB; idx += kThreads) {
    const int le = idx / (NV * NB), k = idx % (NV * NB);
    const int64_t e = e0 + le;
    sU[le][k] = e < g.E ? U[e * NV * NB + k] : 1.0;
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ;
    const int64_t e = e0 + le;
    if (e >= g.E) continue;
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int b = 0; b < NB; b++) s += sU[le][v * NB + b] * g.Vt[b * NQ + q];
      u[v] = s;
    }
    computational<D>(ph, u, w);
    const double c = soundSpeedOf(ph, w[0], w[D + 2]);
    const double sr = sqrt(velocitySquared<D>(w)) + c;
    sT[le][q] = cfl * minedge[e] / (sr * (order + 1.0) * (order + 1.0));
  }
  __syncthreads();
  for (int le = tid; le < EPB; le += kThreads) {
    const int64_t e = e0 + le;
    if (e >= g.E) continue;
    double best = 1.0e308;
    for (int q = 0; q < NQ; q++) best = fmin(best, sT[le][q]);
    out[e] = best;
  }
}

// Fixed-order reduction of the monitor partials; raises the NaN stop flag when every variable is
// NaN (SystemControl.cpp:185-191).
// Boussinesq source, modal: Rsrc[e][b] = sum_q rho beta (T - T_ref) g |J| w_q V[q][b] for the last
// momentum component (SpatialDiscrete.cpp:254-262, 1026-1029), from the state U of the stage;
// point values in LDS, one thread per point, then one per mode. Li (mixed mode, else null): the
// residuals are formed against the orthonormal test functions L^-1 phi, so Rsrc <- L^-1 Rsrc.
template <int D, int NB, int NQ>
__global__ __launch_bounds__(kThreads) void sourceKernel(Geo g, Phys ph, const double* U, double* Rsrc,
                                                         const double* Li) {
  constexpr int NV = D + 2, NM = NQ > NB ? NQ : NB, EPB = kThreads / NM > 0 ? kThreads / NM : 1;
  __shared__ double sS[EPB][NQ];
  __shared__ double sP[EPB][NB];
  const int tid = threadIdx.x, e0 = blockIdx.x * EPB;
  for (int idx = tid; idx < EPB * NQ; idx += kThreads) {
    const int le = idx / NQ, q = idx % NQ, e = e0 + le;
    double s = 0.0;
    if (e < g.E) {
      double u[NV], w[D + 3];
      for (int v = 0; v < NV; v++) {
        double x = 0.0;
        for (int b = 0; b < NB; b++) x += U[(static_cast<int64_t>(e) * NV + v) * NB + b] * g.Vt[b * NQ + q];
        u[v] = x;
      }
      computational<D>(ph, u, w);
      s = boussinesqSource<D>(ph, w) * pointJw<NQ>(g, e, q);
    }
    sS[le][q] = s;
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * NB; idx += kThreads) {
    const int le = idx / NB, b = idx % NB, e = e0 + le;
    double r = 0.0;
    for (int q = 0; q < NQ; q++) r += sS[le][q] * g.Vt[b * NQ + q];
    if (!Li) {
      if (e < g.E) Rsrc[static_cast<int64_t>(e) * NB + b] = r;
    } else {
      sP[le][b] = r;
    }
  }
  if (!Li) return;
  __syncthreads();
  for (int idx = tid; idx < EPB * NB; idx += kThreads) {
    const int le = idx / NB, j = idx % NB, e = e0 + le;
    if (e >= g.E) continue;
    double r = 0.0;
    for (int i = 0; i <= j; i++) r += Li[j * NB + i] * sP[le][i];  // L^-1 is lower triangular
    Rsrc[static_cast<int64_t>(e) * NB + j] = r;
  }
}

constexpr int kMonitorBlocks = 256;

// sum of a block's LANES-strided values, one variable per lane (LANES a multiple of NV), -> out[v]
template <int NV>
__device__ __forceinline__ void blockVariableSums(const double* x, int64_t n, double* red, double* out) {
  constexpr int LANES = (kThreads / NV) * NV;
  const int tid = threadIdx.x;
  double acc = 0.0;
  if (tid < LANES) {
#pragma unroll 4
    for (int64_t i = tid; i < n; i += LANES) acc += x[i];
  }
  red[tid] = acc;
  __syncthreads();
  if (tid < NV) {
    double s = 0.0;
    for (int k = tid; k < LANES; k += NV) s += red[k];
    out[tid] = s;
  }
}

// first pass: block b sums the monitor values of its contiguous element range -> part[b][v]
template <int NV>
__global__ __launch_bounds__(kThreads) void monitorPartialKernel(const double* mon, int E, double* part,
                                                                 const int* status) {
  __shared__ double red[kThreads];
  if (status[0]) return;
  const int chunk = (E + kMonitorBlocks - 1) / kMonitorBlocks;
  const int lo = blockIdx.x * chunk, hi = lo + chunk < E ? lo + chunk : E;
  const int64_t n = hi > lo ? static_cast<int64_t>(hi - lo) * NV : 0;
  blockVariableSums<NV>(mon + static_cast<int64_t>(lo) * NV, n, red, part + static_cast<int64_t>(blockIdx.x) * NV);
}

// second pass: the domain mean of each variable's monitor; all-NaN stops the run
// (SystemControl.cpp:185-191)
template <int NV>
__global__ __launch_bounds__(kThreads) void monitorKernel(const double* part, int E, double* rel, int* status) {
  __shared__ double red[kThreads];
  __shared__ double sum[NV];
  if (status[0]) return;
  blockVariableSums<NV>(part, static_cast<int64_t>(kMonitorBlocks) * NV, red, sum);
  __syncthreads();
  if (threadIdx.x == 0) {
    bool all_nan = true;
    for (int v = 0; v < NV; v++) {
      const double r = sum[v] / static_cast<double>(E);
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
