// dg_tensor_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the sum-factorised
// Navier-Stokes stage of dg_tensor.hpp (hexahedra P2-P4, FP64, periodic meshes).
//
// The forward stage, at the element's Gauss nodes (one thread per node):
//   Y = (B x B x B) c                      nodal state from the modal coefficients c
//   u_s = sum_i lf_s(i) Y(line of s)       own traces (written by the previous stage)
//   alpha_s = ((u_o + u_n)/2 - u_o) fw_s,  beta_s = ((u_n - u_o)/2) fw_s   per slot and variable
//   Dr_a = Dn_a Y,  g_tot = J^-T Dr + sum_s lf_s(i) (alpha_s + beta_s) n_s / (w |J|)
//   g_face = J^-T Dr + sum_s lf_s(i) alpha_s n_s / (w |J|)   (BR2; BR1: g_face = g_tot)
//   gs = sum_i lf_s(i) g_face(line) + kappa_side beta_s n_s / (|J| fw_s)   (BR2 own-face lift)
//   Rvol = sum_a Dn_a^T (F(Y, g_tot) G),  F = F_c - F_v, G = J^-T w |J|
//   H_s = -+ (H_c(u_L, u_R) - (F_v(u_L, gs_L) + F_v(u_R, gs_R)).n / 2) |J_f| w   (face kernel)
//   R = Rvol + sum_s lf_s(i) H_s;   increment = (B^-1 x B^-1 x B^-1) (R / (w |J|))
// Its transpose, given lam (the adjoint of the increment, or of the reference residual
// (B^T x B^T x B^T) R), runs as four kernels that fill the buffers of AdjArgs with nodal
// (tensor-ordered) fields where the dense kernels hold modal ones:
//   nsAdjVolumeTensor  recompute Y, alpha, beta, g_tot, g_face and gs (-> Gf); mu = the adjoint of
//                      R at the nodes; H-bar = lf-traces of mu (-> Hb); the pointwise volume-flux
//                      VJP of W-bar_a = Dn_a mu -> u-bar (-> Ua) and g_tot-bar (-> RtV)
//   nsAdjFaceRevKernel (dg_adjoint.hpp, or nsAdjFaceKernel for the Exact flux) face-flux
//                      VJP -> UFf, GFb
//   nsAdjLiftTensor    gs-bar -> g_face-bar (lf transposes) and beta-bar (own-face lift); the
//                      transpose of the gradient: Y-bar += Dn^T J^-1 (g_tot-bar + g_face-bar)
//                      (-> Ua), alpha-bar + beta-bar per slot -> own and neighbour trace adjoints
//                      (-> Pown, Qnb)
//   nsAdjGatherTensor  trace adjoints of the own slots (UFf + Pown + the neighbours' Qnb) through
//                      the lf transposes into Y-bar, then (B^T x B^T x B^T) -> modal, fused with the
//                      transpose of the RK stage as in nsAdjGatherMfma
#ifndef SDG_DG_TENSOR_ADJOINT_HPP_
#define SDG_DG_TENSOR_ADJOINT_HPP_

#include <hip/hip_runtime.h>

#include "dg_adjoint.hpp"
#include "dg_tensor.hpp"

namespace sdg {

// out[c][q] = sum_k M[k N + idx_axis(q)] in[c][q with axis index k]: the transpose of linePass
template <int N, int C>
__device__ __forceinline__ void linePassT(const double* M, int axis, const double* in, double* out, int q, bool ok) {
  constexpr int NPT = N * N * N;
  if (!ok) return;
  const int i = nodeIndex(q, axis, N), st = axisStride(axis, N);
  const int q0 = q - i * st;
  double m[N];
#pragma unroll
  for (int k = 0; k < N; k++) m[k] = M[k * N + i];
#pragma unroll
  for (int c = 0; c < C; c++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) s += m[k] * in[c * NPT + q0 + k * st];
    out[c * NPT + q] = s;
  }
}

// ---- kernel 1: forward recomputation, mu, H-bar, volume-flux VJP --------------------------------
// LDS per element: sW [NG NPT] = Y + alpha (+ pass scratch), then the face-gradient field;
// sB [NV NAQ] = the basis-pass panel, then beta; sM [NV NPT] = lam, then mu. The volume-flux VJP
// follows the gradient directly, so that the node's state and total gradient die before the
// face-gradient phase.
template <int N, int VISC>
__global__ __launch_bounds__(TensorShape<N>::TPB, 2) void nsAdjVolumeTensor(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV, NG = S::NG, NFQ = S::NFQ, NAQ = S::NAQ, D = 3;
  static_assert(NV * NPT + NV * NAQ <= NG * NPT && NV * NPT <= NV * NAQ, "LDS regions do not nest");
  __shared__ double sW[EPB][NG * NPT];
  __shared__ double sB[EPB][NV * NAQ];
  __shared__ double sM[EPB][NV * NPT];
  __shared__ TensorLds<N> sT;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  double* W = sW[lec];
  double* Y = W;
  double* Al = W + NV * NPT;
  double* Bt = sB[lec];
  double* Mu = sM[lec];
  double dj = 1.0;
  if (ok) dj = g.detj[e];
  constexpr int ITEMS = (NV * NAQ + NPT - 1) / NPT;
  double fo[ITEMS], fn[ITEMS], fw[ITEMS];
  if (ok) {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int idx = q + k * NPT;
      const bool in = idx < NV * NAQ;
      const int v = in ? idx / NAQ : 0, sl = in ? idx % NAQ : 0;
      const int nbs = g.snb[e * NAQ + sl];
      const bool inner = in && nbs >= 0;
      const int64_t own = (e * NV + v) * NAQ + sl;
      const int64_t nbr = inner ? (static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ : own;
      fo[k] = a.Uf_in[own];
      fn[k] = a.Uf_in[nbr];
      fw[k] = inner ? g.sjw[e * NAQ + sl] : 0.0;
    }
  }
  loadModal<double, N>(t, a, a.U, e, Bt, q, ok);
  if (ok) {
    const int tq = t.perm[q];
    for (int v = 0; v < NV; v++) Mu[v * NPT + tq] = j.lam[(e * NV + v) * NPT + q];
  }
  __syncthreads();
  linePass<double, N, NV>(sT.B, 0, Bt, Y, q, ok);
  elemSync<NPT>();
  linePass<double, N, NV>(sT.B, 1, Y, Bt, q, ok);
  elemSync<NPT>();
  linePass<double, N, NV>(sT.B, 2, Bt, Y, q, ok);
  elemSync<NPT>();
  // mu: the adjoint of the nodal residual R (Bt and the alpha region as pass scratch)
  if (j.apply_minv) {  // increment = (B^-1)^3 (R / (w |J|)): mu = ((B^-1)^T)^3 lam / (w |J|)
    linePassT<N, NV>(sT.Binv, 0, Mu, Bt, q, ok);
    elemSync<NPT>();
    linePassT<N, NV>(sT.Binv, 1, Bt, Al, q, ok);
    elemSync<NPT>();
    linePassT<N, NV>(sT.Binv, 2, Al, Mu, q, ok);
    if (ok) {
      const double ijw = 1.0 / (sT.wn[q] * dj);
      for (int v = 0; v < NV; v++) Mu[v * NPT + q] *= ijw;
    }
  } else {  // reference residual (B^T)^3 R: mu = B^3 lam
    linePass<double, N, NV>(sT.B, 0, Mu, Bt, q, ok);
    elemSync<NPT>();
    linePass<double, N, NV>(sT.B, 1, Bt, Al, q, ok);
    elemSync<NPT>();
    linePass<double, N, NV>(sT.B, 2, Al, Mu, q, ok);
  }
  elemSync<NPT>();
  if (ok) {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int idx = q + k * NPT;
      if (idx < NV * NAQ) {
        Al[idx] = ((fo[k] + fn[k]) / 2.0 - fo[k]) * fw[k];
        Bt[idx] = ((fn[k] - fo[k]) / 2.0) * fw[k];
      }
    }
  }
  elemSync<NPT>();
  // gradients at the node (the forward kernel's arithmetic)
  double gi[D * D], u[NV], gtot[NG], gface[NG];
  for (int i = 0; i < D * D; i++) gi[i] = ok ? g.ginv[e * D * D + i] : 0.0;
  for (int v = 0; v < NV; v++) u[v] = 0.0;
  for (int r = 0; r < NG; r++) gtot[r] = gface[r] = 0.0;
  if (ok) {
    {
      double dr[NV][D];
      for (int ax = 0; ax < D; ax++) {
        const int i = nodeIndex(q, ax, N), st = axisStride(ax, N), q0 = q - i * st;
        for (int v = 0; v < NV; v++) {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < N; m++) s += sT.Dn[i * N + m] * Y[v * NPT + q0 + m * st];
          dr[v][ax] = s;
        }
      }
      for (int v = 0; v < NV; v++) {
        u[v] = Y[v * NPT + q];
        for (int d = 0; d < D; d++) {
          double s = 0.0;
          for (int ax = 0; ax < D; ax++) s += gi[d + D * ax] * dr[v][ax];
          gtot[v * D + d] = s;
          gface[v * D + d] = s;
        }
      }
    }
    const double inw = 1.0 / (sT.wn[q] * dj);
#pragma unroll 1
    for (int s = 0; s < 6; s++) {
      const int ax = t.axis[s], i = nodeIndex(q, ax, N);
      const int sl = s * NFQ + sT.slot_of[s * NPT + q];
      double n[D];
      for (int d = 0; d < D; d++) n[d] = g.snrm[(e * NAQ + sl) * D + d];
      const double coef = sT.lf[t.side[s] * N + i] * inw;
      for (int v = 0; v < NV; v++) {
        const double al = Al[v * NAQ + sl] * coef, be = Bt[v * NAQ + sl] * coef;
        for (int d = 0; d < D; d++) {
          gtot[v * D + d] += (al + be) * n[d];
          gface[v * D + d] += (VISC == 2 ? al : al + be) * n[d];
        }
      }
    }
  }
  elemSync<NPT>();  // Y and alpha are read: sW takes the face-gradient field (BR2: without the jump
                    // lifts; BR1: the total gradient)
  if (ok)
    for (int r = 0; r < NG; r++) W[r * NPT + q] = gface[r];
  // volume flux: W-bar_a = Dn_a mu (transpose of Rvol = sum_a Dn_a^T W_a), then the pointwise VJP
  if (ok) {
    double fbar[D * NV];
    for (int ax = 0; ax < D; ax++) {
      const int i = nodeIndex(q, ax, N), st = axisStride(ax, N), q0 = q - i * st;
      for (int v = 0; v < NV; v++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++) s += sT.Dn[i * N + k] * Mu[v * NPT + q0 + k * st];
        fbar[ax * NV + v] = s;
      }
    }
    const double jw = dj * sT.wn[q];
    double G[D * D], Fb[D * NV], gb[NG], ub[NV];
    for (int i = 0; i < D * D; i++) G[i] = gi[i] * jw;
    volumeFluxBar<D>(fbar, G, Fb);
    volumeFluxStateVjp<D>(ph, u, gtot, Fb, ub);
    for (int v = 0; v < NV; v++) j.Ua[(e * NV + v) * NPT + q] = ub[v];
    volumeFluxVjpDir<D>(ph, u, gtot, Fb, NV, gb);
    for (int r = 0; r < NG; r++) j.RtV[(e * NG + r) * NPT + q] = gb[r];
  }
  // H-bar: R += lf_s(i) H_s at the nodes of the slot's line
  if (ok)
    for (int idx = q; idx < NV * NAQ; idx += NPT) {
      const int sl = idx % NAQ, v = idx / NAQ, s = sl / NFQ;
      const int b = sT.base[sl], st = axisStride(sT.axis[s], N);
      const double* l = sT.lf + sT.side[s] * N;
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < N; i++) acc += l[i] * Mu[v * NPT + b + i * st];
      j.Hb[(e * NV + v) * NAQ + sl] = acc;
    }
  elemSync<NPT>();  // the face-gradient field is complete
  // gs per own slot: the face-gradient trace plus (BR2) the own face's lift -> Gf (FP64), the
  // gradient the face-flux VJP differentiates the viscous flux at
  if (ok) {
    const double idj = 1.0 / dj;
    for (int sl = q; sl < NAQ; sl += NPT) {
      const int s = sl / NFQ;
      const int b = sT.base[sl], st = axisStride(sT.axis[s], N);
      const double* l = sT.lf + sT.side[s] * N;
      const double kf = sT.kappa[sT.side[s]] * idj / sT.fwt[sl];
      double n[D];
      for (int d = 0; d < D; d++) n[d] = g.snrm[(e * NAQ + sl) * D + d];
      for (int r = 0; r < NG; r++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) acc += l[i] * W[r * NPT + b + i * st];
        if constexpr (VISC == 2) acc += kf * Bt[(r / D) * NAQ + sl] * n[r % D];
        a.Gf[(e * NG + r) * NAQ + sl] = acc;
      }
    }
  }
}

// ---- kernel 3 (after nsAdjFaceKernel): transpose of the lifted gradient -------------------------
template <int N, int VISC>
__global__ __launch_bounds__(TensorShape<N>::TPB) void nsAdjLiftTensor(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV, NG = S::NG, NFQ = S::NFQ, NAQ = S::NAQ, D = 3;
  __shared__ double sZ[EPB][NG * NPT];
  __shared__ TensorLds<N> sT;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  double* Z = sZ[lec];
  double dj = 1.0, gi[D * D];
  for (int i = 0; i < D * D; i++) gi[i] = 0.0;
  double gt[NG], gf[NG];
  for (int r = 0; r < NG; r++) gt[r] = gf[r] = 0.0;
  if (ok) {
    dj = g.detj[e];
    for (int i = 0; i < D * D; i++) gi[i] = g.ginv[e * D * D + i];
    for (int r = 0; r < NG; r++) gt[r] = j.RtV[(e * NG + r) * NPT + q];
  }
  __syncthreads();  // tables
  if (ok) {
    // gs = sum_i lf_s(i) g_face(line): g_face-bar(node) = sum_s lf_s(i) gs-bar(slot of the node)
#pragma unroll 1
    for (int s = 0; s < 6; s++) {
      const int i = nodeIndex(q, t.axis[s], N);
      const int sl = s * NFQ + sT.slot_of[s * NPT + q];
      const double l = sT.lf[t.side[s] * N + i];
      for (int r = 0; r < NG; r++) gf[r] += l * j.GFb[(e * NG + r) * NAQ + sl];
    }
    if constexpr (VISC == 1)
      for (int r = 0; r < NG; r++) {
        gt[r] += gf[r];
        gf[r] = 0.0;
      }
    // Dr-bar[v][a] = sum_d J^-T[d][a] (g_tot-bar + g_face-bar)[v][d]
    for (int v = 0; v < NV; v++)
      for (int ax = 0; ax < D; ax++) {
        double s = 0.0;
        for (int d = 0; d < D; d++) s += gi[d + D * ax] * (gt[v * D + d] + gf[v * D + d]);
        Z[(v * D + ax) * NPT + q] = s;
      }
  }
  elemSync<NPT>();
  if (ok) {  // Y-bar += sum_a Dn_a^T Dr-bar_a (+ the volume part already in Ua)
    double yb[NV];
    for (int v = 0; v < NV; v++) yb[v] = j.Ua[(e * NV + v) * NPT + q];
    for (int ax = 0; ax < D; ax++) {
      const int i = nodeIndex(q, ax, N), st = axisStride(ax, N), q0 = q - i * st;
      for (int v = 0; v < NV; v++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; m++) s += sT.Dn[m * N + i] * Z[(v * D + ax) * NPT + q0 + m * st];
        yb[v] += s;
      }
    }
    for (int v = 0; v < NV; v++) j.Ua[(e * NV + v) * NPT + q] = yb[v];
  }
  elemSync<NPT>();
  // alpha enters g_tot and g_face, beta g_tot only, both through lf_s(i) n_s / (w |J|)
  if (ok) {
    const double inw = 1.0 / (sT.wn[q] * dj);
    for (int r = 0; r < NG; r++) Z[r * NPT + q] = inw * (2.0 * gt[r] + gf[r]);
  }
  elemSync<NPT>();
  if (ok) {
    const double idj = 1.0 / dj;
    for (int idx = q; idx < NV * NAQ; idx += NPT) {
      const int sl = idx % NAQ, v = idx / NAQ, s = sl / NFQ;
      const int b = sT.base[sl], st = axisStride(sT.axis[s], N);
      const double* l = sT.lf + sT.side[s] * N;
      const int64_t es = e * NAQ + sl;
      double ab = 0.0;
      for (int d = 0; d < D; d++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) acc += l[i] * Z[(v * D + d) * NPT + b + i * st];
        ab += acc * g.snrm[es * D + d];
      }
      if constexpr (VISC == 2) {  // own-face lift: gs += kappa beta n / (|J| fw)
        const double kf = sT.kappa[sT.side[s]] * idj / sT.fwt[sl];
        for (int d = 0; d < D; d++) ab += kf * j.GFb[(e * NG + v * D + d) * NAQ + sl] * g.snrm[es * D + d];
      }
      // alpha = beta = (u_n - u_o)/2 fw in their derivatives
      const double sb = ab * (g.snb[es] >= 0 ? g.sjw[es] : 0.0) / 2.0;
      j.Pown[(e * NV + v) * NAQ + sl] = -sb;
      j.Qnb[(e * NV + v) * NAQ + sl] = sb;
    }
  }
  (void)ph;
  (void)a;
}

// ---- kernel 4: trace adjoints -> modal, transpose of the RK stage -------------------------------
template <int N>
__global__ __launch_bounds__(TensorShape<N>::TPB) void nsAdjGatherTensor(Geo g, Phys ph, StageArgs a, AdjArgs j) {
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV, NFQ = S::NFQ, NAQ = S::NAQ;
  __shared__ double sX[EPB][NV * NAQ];
  __shared__ double sP[EPB][NV * NPT];
  __shared__ double sQ[EPB][NV * NPT];
  __shared__ TensorLds<N> sT;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  double* X = sX[lec];
  double* P = sP[lec];
  double* Q = sQ[lec];
  if (ok)
    for (int idx = q; idx < NV * NAQ; idx += NPT) {
      const int sl = idx % NAQ, v = idx / NAQ;
      const int nbs = g.snb[e * NAQ + sl];
      const int64_t o = (e * NV + v) * NAQ + sl;
      double x = j.UFf[o] + j.Pown[o];
      if (nbs >= 0) x += j.Qnb[(static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ];
      X[v * NAQ + sl] = x;
    }
  __syncthreads();
  if (ok) {  // u_s = sum_i lf_s(i) Y(line): Y-bar(node) += lf_s(i) u_s-bar(slot of the node)
    double yb[NV];
    for (int v = 0; v < NV; v++) yb[v] = j.Ua[(e * NV + v) * NPT + q];
#pragma unroll 1
    for (int s = 0; s < 6; s++) {
      const int i = nodeIndex(q, t.axis[s], N);
      const int sl = s * NFQ + sT.slot_of[s * NPT + q];
      const double l = sT.lf[t.side[s] * N + i];
      for (int v = 0; v < NV; v++) yb[v] += l * X[v * NAQ + sl];
    }
    for (int v = 0; v < NV; v++) P[v * NPT + q] = yb[v];
  }
  elemSync<NPT>();
  // Y = B^3 c: c-bar = (B^T)^3 Y-bar
  linePassT<N, NV>(sT.B, 0, P, Q, q, ok);
  elemSync<NPT>();
  linePassT<N, NV>(sT.B, 1, Q, P, q, ok);
  elemSync<NPT>();
  linePassT<N, NV>(sT.B, 2, P, Q, q, ok);
  if (ok) {
    const int b = sT.iperm[q];
    for (int v = 0; v < NV; v++) {
      const int64_t o = (e * NV + v) * NPT + b;
      const double vjp = Q[v * NPT + q];
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
  (void)ph;
  (void)a;
}

}  // namespace sdg

#endif  // SDG_DG_TENSOR_ADJOINT_HPP_
