// dg_tensor.hpp -- sum-factorised Navier-Stokes stage kernels for hexahedra.
//
// On a hexahedron the modal basis is a tensor product of 1-D Lobatto functions and the volume and
// face rules are tensor Gauss rules with the same N = P+1 nodes per direction, so every operator
// of the stage factors into 1-D operators along grid lines of N nodes:
//   u(nodes) = (B x B x B) c  with B[i][k] = l_k(x_i);  c = (B^-1 x B^-1 x B^-1) u(nodes);
//   a reference derivative along one axis is the collocation derivative Dn of the Lagrange
//   polynomials at the Gauss nodes; the trace of a nodal field on the face at xi_a = -1/+1 is the
//   contraction with the Lagrange values l(-1) / l(+1) along axis a;
//   the mass matrix of the nodal (Lagrange) basis is diag(w |J|) (Gauss is exact for it).
// The stage is then the reference's pass chain (SpatialDiscrete.cpp:194-1116,
// TimeIntegration.cpp:181-253) in exact arithmetic, with O(N^4) instead of O(N^6) work per element:
//   lifted gradient (BR1/BR2, TimeIntegration.cpp:200-228): integrating the weak gradient term by
//     parts (exact for these polynomials), M^-1 [ sum_faces (A - u n |J_f| w) phi - int u grad phi ]
//     = grad u + sum_faces l_s(i) alpha / (w_i |J|) at node (i, tangential point), alpha the face
//     term minus the own trace; the BR2 face lift r_s likewise with the jump term beta;
//   residual R_psi = sum_axes Dn^T (w |J| J^-1 F) + sum_faces l_s(i) Hs, the increment R_psi/(w|J|)
//   taken back to modal coefficients with B^-1 along each axis; the reference residual (before
//   M^-1) is (B^T x B^T x B^T) R_psi.
// Each thread owns one node of one element. FP64 mode: FP64 throughout. Mixed-precision mode:
// contractions and work buffers in TC = FP32 on deviations from the reference state, the
// conversion to computational variables (the low-Mach pressure recovery) and the flux deviations
// p - pref, rho H - h0 rho in FP64, the primitive gradient, viscous and volume fluxes in FP32
// (volumeFlux32, normalViscous32), FP64 state and RK accumulation.
#ifndef SDG_DG_TENSOR_HPP_
#define SDG_DG_TENSOR_HPP_

#include <hip/hip_runtime.h>

#include "dg_mfma.hpp"

namespace sdg {

template <int N>
struct TensorShape {
  static constexpr int NPT = N * N * N;
  // threads per workgroup (one thread per node, as many elements as fit; one-element workgroups
  // at N = 4 measured slower: the order tables are then staged once per element)
  static constexpr int TPB = kThreads;
  static constexpr int EPB = TPB / NPT;
  static constexpr int NV = 5, NG = 15, NFQ = N * N, NAQ = 6 * N * N;
};

// the order's 1-D operators and node / face-point maps, staged into LDS once per workgroup: the
// kernels index them per lane, which from the kernel-argument segment or HBM would be vector
// memory loads
template <int N>
struct TensorLds {
  static constexpr int NPT = N * N * N, NAQ = 6 * N * N;
  double B[N * N], Binv[N * N], Dn[N * N], lf[2 * N], wn[NPT], fwt[NAQ], kappa[2];
  float Bf[N * N], Dnf[N * N], lff[2 * N];  // FP32 copies for the mixed mode's contractions
  short iperm[NPT], slot_of[6 * NPT], base[NAQ], axis[6], side[6];
};

// the contraction-type copy of a staged table
template <typename TC, int N>
__device__ __forceinline__ const TC* tabB(const TensorLds<N>& L) {
  if constexpr (sizeof(TC) == 8) return L.B; else return L.Bf;
}
template <typename TC, int N>
__device__ __forceinline__ const TC* tabDn(const TensorLds<N>& L) {
  if constexpr (sizeof(TC) == 8) return L.Dn; else return L.Dnf;
}
template <typename TC, int N>
__device__ __forceinline__ const TC* tabLf(const TensorLds<N>& L) {
  if constexpr (sizeof(TC) == 8) return L.lf; else return L.lff;
}

// all threads of the workgroup; the caller's next barrier publishes the tables. Every thread
// issues its loads unconditionally (indices clamped into range) before any store, so the copies
// cost one memory round trip.
template <int N>
__device__ __forceinline__ void stageTables(const TensorOps& t, TensorLds<N>& L) {
  constexpr int NPT = N * N * N, NAQ = 6 * N * N, TPB = TensorShape<N>::TPB;
  constexpr int NS = (6 * NPT + TPB - 1) / TPB;
  static_assert(NPT <= TPB && NAQ <= TPB, "one table entry per thread");
  const int tid = threadIdx.x;
  const int i2 = tid < N * N ? tid : 0, i3 = tid < NPT ? tid : 0, ia = tid < NAQ ? tid : 0;
  const double b = t.B[i2], bi = t.Binv[i2], dn = t.Dn[i2];
  const double w0 = t.w[i3 % N], w1 = t.w[(i3 / N) % N], w2 = t.w[i3 / (N * N)];
  const short ip = t.iperm[i3], bs = t.base[ia];
  const double fw = t.fwt[ia], lf = t.lf[tid < 2 * N ? tid : 0], kp = t.kappa[tid < 2 ? tid : 0];
  const int ax = t.axis[tid < 6 ? tid : 0], sd = t.side[tid < 6 ? tid : 0];
  short so[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const int j = tid + k * TPB;
    so[k] = t.slot_of[j < 6 * NPT ? j : 0];
  }
  if (tid < N * N) {
    L.B[tid] = b;
    L.Binv[tid] = bi;
    L.Dn[tid] = dn;
    L.Bf[tid] = static_cast<float>(b);
    L.Dnf[tid] = static_cast<float>(dn);
  }
  if (tid < NPT) {
    L.wn[tid] = w0 * w1 * w2;
    L.iperm[tid] = ip;
  }
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (tid + k * TPB < 6 * NPT) L.slot_of[tid + k * TPB] = so[k];
  if (tid < NAQ) {
    L.fwt[tid] = fw;
    L.base[tid] = bs;
  }
  if (tid < 2 * N) {
    L.lf[tid] = lf;
    L.lff[tid] = static_cast<float>(lf);
  }
  if (tid < 6) {
    L.axis[tid] = static_cast<short>(ax);
    L.side[tid] = static_cast<short>(sd);
  }
  if (tid < 2) L.kappa[tid] = kp;
}

// barrier between the phases of one element's work: at NPT = 64 an element is exactly one
// wavefront that touches only its own LDS region (the shared tables are read-only by then), so
// the lanes of the wavefront need only see each other's LDS writes, which wavefront-scope fences
// give without waiting for the other wavefronts of the workgroup; otherwise elements straddle
// wavefronts and the workgroup barrier is needed
template <int NPT>
__device__ __forceinline__ void elemSync() {
  if constexpr (NPT == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

__device__ __forceinline__ int axisStride(int axis, int N) { return axis == 0 ? 1 : (axis == 1 ? N : N * N); }
__device__ __forceinline__ int nodeIndex(int q, int axis, int N) {
  return axis == 0 ? q % N : (axis == 1 ? (q / N) % N : q / (N * N));
}

// out[c][q] = sum_m M[idx_axis(q) N + m] in[c][q with axis index m], for C components of NPT
// values each, one thread per node (all threads of the element call it; no barrier inside)
template <typename TC, int N, int C>
__device__ __forceinline__ void linePass(const TC* M, int axis, const TC* in, TC* out, int q, bool ok) {
  constexpr int NPT = N * N * N;
  if (!ok) return;
  const int i = nodeIndex(q, axis, N), st = axisStride(axis, N);
  const int q0 = q - i * st;
  TC m[N];
#pragma unroll
  for (int k = 0; k < N; k++) m[k] = M[i * N + k];
#pragma unroll
  for (int c = 0; c < C; c++) {
    TC s = TC(0);
#pragma unroll
    for (int k = 0; k < N; k++) s += m[k] * in[c * NPT + q0 + k * st];
    out[c * NPT + q] = s;
  }
}

// the element's state as TC operands in tensor order (FP64 state, or its FP32 deviation)
template <typename TC, int N>
__device__ __forceinline__ void loadModal(const TensorOps& t, const StageArgs& a, const double* U, int64_t e, TC* sA,
                                          int q, bool ok) {
  constexpr int NPT = N * N * N, NV = 5;
  if (!ok) return;
  const int tq = t.perm[q];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double u = U[(e * NV + v) * NPT + q];
    if constexpr (sizeof(TC) == 8)
      sA[v * NPT + tq] = u;
    else
      sA[v * NPT + tq] = deviationOperand(u, a.mix.uref[v], a.mix.one[q]);
  }
}

// traces of the nodal state sA at the element's own face points -> Uf_out (TC)
// (rec: optional FP64 copy of the traces, same layout, for the adjoint's recorded run)
template <typename TC, int N>
__device__ __forceinline__ void writeTraces(const TensorLds<N>& L, const TC* sA, TC* Uf_out, int64_t e, int le, int tid,
                                            double* rec = nullptr) {
  constexpr int NPT = N * N * N, NV = 5, NFQ = N * N, NAQ = 6 * NFQ, EPB = TensorShape<N>::EPB;
  for (int idx = tid - le * NPT; idx < NAQ * NV; idx += NPT) {
    if (idx < 0) break;
    const int sl = idx % NAQ, v = idx / NAQ, s = sl / NFQ;
    const int b = L.base[sl], st = axisStride(L.axis[s], N);
    const TC* l = tabLf<TC, N>(L) + L.side[s] * N;
    TC acc = TC(0);
#pragma unroll
    for (int i = 0; i < N; i++) acc += l[i] * sA[v * NPT + b + i * st];
    Uf_out[(e * NV + v) * NAQ + sl] = acc;
    if (rec) rec[(e * NV + v) * NAQ + sl] = static_cast<double>(acc);
  }
  (void)EPB;
}

// element slot of a thread (tid / NPT); at NPT = 64 a wavefront is one element, and the index is
// made wave-uniform so that the element's geometry is read through scalar loads into SGPRs
template <int NPT>
__device__ __forceinline__ int localElement(int tid) {
  int le = tid / NPT;
  if constexpr (NPT == 64) le = __builtin_amdgcn_readfirstlane(le);
  return le;
}

// Mixed-precision mode, pointwise physics in FP32 after the FP64 pressure recovery: w =
// computational(u) stays FP64 (p = (gamma - 1)(rho E - rho |v|^2 / 2) cancels two digits at
// Ma 0.1), so do the deviations p - pref and rho H - h0 rho of the momentum and energy fluxes;
// the primitive gradient, the viscous flux and the contraction with (J^T)^-1 |J| w are FP32.
// out[b NV + v] = sum_a (F_c - F_v)[a][v] G[a][b] as volumeFlux.
// avg (optional): the artificial term eps grad U per gradient component, added to F_v (its energy
// row reduced by h0 times its density row like the convective one)
template <int D>
__device__ __forceinline__ void volumeFlux32(const Phys& ph, const StageArgs& a, const double* u, const double* w,
                                             const float* g, const double* G, float* out,
                                             const float* avg = nullptr) {
  constexpr int NV = D + 2, NG = NV * D;
  float wf[D + 3], uf[NV], pg[NG], fv[D * NV], fc[D * NV];
#pragma unroll
  for (int i = 0; i < D + 2; i++) wf[i] = static_cast<float>(w[i]);
  wf[D + 2] = 0.0f;  // the viscous flux does not read p
#pragma unroll
  for (int v = 0; v < NV; v++) uf[v] = static_cast<float>(u[v]);
  primitiveGradient<D>(ph, wf, uf, g, pg);
  viscousRawFlux<D>(ph, wf, pg, fv);
  if (avg)
#pragma unroll
    for (int r = 0; r < NG; r++) fv[(r % D) + D * (r / D)] += avg[r];
  const double rho = w[0], et = w[D + 1] + velocitySquared<D>(w) / 2.0;
  const float pdev = static_cast<float>(w[D + 2] - a.mix.pref);
  const float hdev = static_cast<float>(rho * et + w[D + 2] - a.mix.h0 * rho);
  const float rf = wf[0];
#pragma unroll
  for (int aa = 0; aa < D; aa++) {
    fc[aa] = rf * wf[1 + aa];
#pragma unroll
    for (int b = 0; b < D; b++) fc[aa + D * (1 + b)] = rf * wf[1 + aa] * wf[1 + b] + (aa == b ? pdev : 0.0f);
    fc[aa + D * (D + 1)] = hdev * wf[1 + aa];
  }
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int v = 0; v < NV; v++) {
      float s = 0.0f;
#pragma unroll
      for (int aa = 0; aa < D; aa++) s += (fc[aa + D * v] - fv[aa + D * v]) * static_cast<float>(G[aa + D * b]);
      out[b * NV + v] = s;
    }
}

// Mixed-precision mode: the normal viscous flux F_v(u, g) . n in FP32 (w = computational(u) FP64)
template <int D>
__device__ __forceinline__ void normalViscous32(const Phys& ph, const double* u, const double* w, const float* g,
                                                const double* n, float* out) {
  constexpr int NV = D + 2, NG = NV * D;
  float wf[D + 3], uf[NV], pg[NG], fv[D * NV];
#pragma unroll
  for (int i = 0; i < D + 2; i++) wf[i] = static_cast<float>(w[i]);
  wf[D + 2] = 0.0f;
#pragma unroll
  for (int v = 0; v < NV; v++) uf[v] = static_cast<float>(u[v]);
  primitiveGradient<D>(ph, wf, uf, g, pg);
  viscousRawFlux<D>(ph, wf, pg, fv);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < D; d++) s += fv[d + D * v] * static_cast<float>(n[d]);
    out[v] = s;
  }
}

// ---- kernel 1: lifted gradient, volume flux, gradient traces ------------------------------------
// LDS per element (TC values), two regions reused across the phases:
//   sW [NG NPT]:  nodal state Y [v][node] + alpha [v][slot]  ->  the face-gradient field [r][node]
//   sB [NV NAQ]:  the basis-pass panel (modal, then axis-0/1 partial sums)  ->  beta [v][slot]
// AV (FP64): artificial viscosity eps grad U with grad U the volume gradient (the gradient with the
// faces' average traces, SpatialDiscrete.cpp:844-911: BR2's face-gradient field, the alpha lifts
// alone under BR1) added to the volume flux and to each side's normal viscous flux
// (ViscousFlux.cpp:105-190); boundary points keep its trace in GfV for the boundary flux
template <typename TC, int N, int VISC, bool AV = false>
__global__ __launch_bounds__(TensorShape<N>::TPB) void nsGradTensor(Geo g, Phys ph, StageArgs a) {
  static_assert(!AV || VISC == 1 || VISC == 2, "artificial viscosity: BR1 / BR2");
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV, NG = S::NG, NFQ = S::NFQ, NAQ = S::NAQ, D = 3;
  static_assert(NV * NPT + NV * NAQ <= NG * NPT && NV * NPT <= NV * NAQ, "LDS regions do not nest");
  __shared__ TC sW[EPB][NG * NPT];
  __shared__ TC sB[EPB][NV * NAQ];
  __shared__ TensorLds<N> sT;
  if (a.status && a.status[0]) return;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  TC* W = sW[lec];
  TC* Y = W;              // nodal state [v][node]
  TC* Al = W + NV * NPT;  // alpha [v][slot]: face term minus the own trace, scalar (times n)
  TC* Bt = sB[lec];       // basis-pass panel, then beta [v][slot]: jump term, scalar (times n)
  // element geometry (affine hexahedra: constant J^-1 and |J|), read before any store
  double dj = 0.0, gi[D * D];
  if (ok) {
    dj = g.detj[e];
    for (int i = 0; i < D * D; i++) gi[i] = g.ginv[e * D * D + i];
  }
  // face terms per (variable, slot): alpha = A - u_own n |J_f| w and beta = J (scalars of n_out).
  // The interior faces' operands are gathered first, so that their latency overlaps the basis
  // passes; boundary slots get jw = 0 here and are filled in after the passes.
  const TC* Uf_in = pick<TC>(a.Uf_in, a.Uf_in32);
  constexpr int ITEMS = (NV * NAQ + NPT - 1) / NPT;
  TC fo[ITEMS], fn[ITEMS];
  double fw[ITEMS];
  if (ok) {
    // branch-free: the neighbour slots first, then every trace and weight load at once (a slot
    // past the end or on a boundary reads its own trace and gets weight 0)
    int nbk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int idx = q + k * NPT;
      nbk[k] = g.snb[e * NAQ + (idx < NV * NAQ ? idx % NAQ : 0)];
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int idx = q + k * NPT;
      const bool in = idx < NV * NAQ;
      const int v = in ? idx / NAQ : 0, sl = in ? idx % NAQ : 0, nbs = nbk[k];
      const bool inner = in && nbs >= 0;
      const int64_t own = (e * NV + v) * NAQ + sl;
      const int64_t nbr = inner ? (static_cast<int64_t>(nbs / NAQ) * NV + v) * NAQ + nbs % NAQ : own;
      fo[k] = Uf_in[own];
      fn[k] = Uf_in[nbr];
      const double jw = g.sjw[e * NAQ + sl];
      fw[k] = inner ? jw : 0.0;
    }
  }
  loadModal<TC, N>(t, a, a.U, e, Bt, q, ok);
  __syncthreads();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 0, Bt, Y, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 1, Y, Bt, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 2, Bt, Y, q, ok);  // Y: nodal state
  elemSync<NPT>();
  if (ok) {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int idx = q + k * NPT;
      if (idx < NV * NAQ) {
        // in the contraction type (the FP64 mode's arithmetic is FP64; the mixed mode's traces
        // are FP32 deviations already)
        const TC uo = fo[k], un = fn[k], w = static_cast<TC>(fw[k]);
        Al[idx] = ((uo + un) / TC(2) - uo) * w;
        Bt[idx] = ((un - uo) / TC(2)) * w;
      }
    }
    if (g.Fb > 0)
      for (int idx = q; idx < NV * NAQ; idx += NPT) {
        const int v = idx / NAQ, sl = idx % NAQ;
        const int nbs = g.snb[e * NAQ + sl];
        if (nbs >= 0) continue;
        const double uo = Uf_in[(e * NV + v) * NAQ + sl];
        double ub[NV], Ab[D], Jb[D];
        for (int k = 0; k < NV; k++) ub[k] = pointState<TC>(a, Uf_in[(e * NV + k) * NAQ + sl], k);
        boundaryGradientTerms<D>(ph, g, -1 - nbs, ub, v, sizeof(TC) == 4 ? a.mix.uref[v] : 0.0, Ab, Jb);
        // A and J are (scalar) n |J_f| w; recover the scalars along the dominant normal component
        const int bp = -1 - nbs;
        int dm = 0;
        for (int d = 1; d < D; d++)
          if (fabs(g.bnormal[static_cast<int64_t>(bp) * D + d]) > fabs(g.bnormal[static_cast<int64_t>(bp) * D + dm])) dm = d;
        const double nd = g.bnormal[static_cast<int64_t>(bp) * D + dm];
        Al[idx] = static_cast<TC>(Ab[dm] / nd - uo * g.bjw[bp]);
        Bt[idx] = static_cast<TC>(Jb[dm] / nd);
      }
  }
  elemSync<NPT>();
  // gradient at the node: J^-T (Dn u along each axis) + the face corrections, assembled in the
  // contraction type (the mixed mode's lifted gradient is an FP32 quantity throughout)
  double u[NV];
  TC wtot[NG], wface[VISC == 2 || AV ? NG : 1];  // wface: the volume gradient (alpha lifts only)
  if (ok) {
    TC dr[NV][D];
    for (int ax = 0; ax < D; ax++) {
      const int i = nodeIndex(q, ax, N), st = axisStride(ax, N), q0 = q - i * st;
      for (int v = 0; v < NV; v++) {
        TC s = TC(0);
#pragma unroll
        for (int m = 0; m < N; m++) s += tabDn<TC, N>(sT)[i * N + m] * Y[v * NPT + q0 + m * st];
        dr[v][ax] = s;
      }
    }
    TC git[D * D];
    for (int i = 0; i < D * D; i++) git[i] = static_cast<TC>(gi[i]);
    for (int v = 0; v < NV; v++) {
      u[v] = pointState<TC>(a, Y[v * NPT + q], v);
      for (int d = 0; d < D; d++) {
        TC s = TC(0);
        for (int ax = 0; ax < D; ax++) s += git[d + D * ax] * dr[v][ax];
        wtot[v * D + d] = s;
        if constexpr (VISC == 2 || AV) wface[v * D + d] = s;
      }
    }
    // the outward normal at the node's face point, loaded one face ahead of its use
    double nx[D];
    {
      const int64_t es = e * NAQ + sT.slot_of[q];
      for (int d = 0; d < D; d++) nx[d] = g.snrm[es * D + d];
    }
    const double inw = 1.0 / (sT.wn[q] * dj);  // M^-1 at the node, one reciprocal for the six faces
#pragma unroll 1  // unrolled, the six faces' LDS loads are hoisted together and double the VGPRs
    for (int s = 0; s < 6; s++) {
      const int ax = t.axis[s], i = nodeIndex(q, ax, N);
      const int sl = s * NFQ + sT.slot_of[s * NPT + q];
      TC n[D];
      for (int d = 0; d < D; d++) n[d] = static_cast<TC>(nx[d]);
      if (s < 5) {
        const int64_t es = e * NAQ + (s + 1) * NFQ + sT.slot_of[(s + 1) * NPT + q];
        for (int d = 0; d < D; d++) nx[d] = g.snrm[es * D + d];
      }
      // alpha, beta carry |J_f| w_t1 w_t2; M^-1 = 1 / (w_i w_t1 w_t2 |J|) at the node
      const TC coef = static_cast<TC>(sT.lf[t.side[s] * N + i] * inw);
      for (int v = 0; v < NV; v++) {
        const TC al = Al[v * NAQ + sl] * coef;
        const TC be = Bt[v * NAQ + sl] * coef;
        for (int d = 0; d < D; d++) {
          wtot[v * D + d] += (al + be) * n[d];
          if constexpr (VISC == 2 || AV) wface[v * D + d] += al * n[d];
        }
      }
    }
  }
  elemSync<NPT>();  // Y and alpha are read; sW takes the gradient fields from here on
  // optional: modal coefficients of the total lifted gradient (sdg_get_gradient), B^-1 per axis
  if (a.Wout) {
    if (ok)
      for (int r = 0; r < NG; r++) W[r * NPT + q] = static_cast<TC>(wtot[r]);
    elemSync<NPT>();
    if (ok) {
      // modal c[k] = sum_i Binv[kx][i] Binv[ky][j] Binv[kz][l] w(i, j, l): direct triple sum (output path)
      const int kx = q % N, ky = (q / N) % N, kz = q / (N * N);
      const int b = sT.iperm[q];
#pragma unroll 1
      for (int r = 0; r < NG; r++) {
        double s = 0.0;
#pragma unroll 1
        for (int l = 0; l < N; l++)
          for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++)
              s += sT.Binv[kx * N + i] * sT.Binv[ky * N + j] * sT.Binv[kz * N + l] *
                   static_cast<double>(W[r * NPT + i + N * j + N * N * l]);
        a.Wout[(e * NPT + b) * NG + r] = s;
      }
    }
    elemSync<NPT>();
  }
  // BR1 with artificial viscosity: the AV term of each own slot's normal flux takes the volume
  // gradient's trace, which is not the field BR1's viscous flux traces, so it goes first: eps times
  // the trace . n per slot (NV values kept in registers) and, at boundary slots, the trace itself
  constexpr int SI0 = (NAQ + NPT - 1) / NPT;
  double avn[AV && VISC == 1 ? SI0 : 1][NV];
  if constexpr (AV && VISC == 1) {
    if (ok)
      for (int r = 0; r < NG; r++) W[r * NPT + q] = wface[r];
    elemSync<NPT>();
    if (ok) {
#pragma unroll
      for (int k = 0; k < SI0; k++) {
        const int sl = q + k * NPT;
        if (sl >= NAQ) break;
        const int s = sl / NFQ;
        const int b = sT.base[sl], st = axisStride(sT.axis[s], N);
        const int64_t es = e * NAQ + sl;
        const int nbs = g.snb[es];
        double n[D];
        for (int d = 0; d < D; d++) n[d] = g.snrm[es * D + d];
        const TC* lt = tabLf<TC, N>(sT) + sT.side[s] * N;
        const double eps = avAt(g, g.nodf, sl, static_cast<int>(e));
        for (int v = 0; v < NV; v++) avn[k][v] = 0.0;
        for (int r = 0; r < NG; r++) {
          double tr = 0.0;
#pragma unroll
          for (int i = 0; i < N; i++) tr += lt[i] * W[r * NPT + b + i * st];
          avn[k][r / D] += eps * tr * n[r % D];
          if (nbs < 0) a.GfV[(e * NG + r) * NAQ + sl] = tr;
        }
      }
    }
    elemSync<NPT>();
  }
  // the face-gradient field (BR2: without the jump lifts; BR1: the total) -> sW for the traces
  if (ok)
    for (int r = 0; r < NG; r++) W[r * NPT + q] = static_cast<TC>(VISC == 2 ? wface[VISC == 2 ? r : 0] : wtot[r]);
  elemSync<NPT>();
  // per own face point: the trace of the face-gradient field (+ the own face's BR2 lift, the
  // gradient the face flux of the reference uses for this side), then this side's normal viscous
  // flux F_v(u, grad u) . n_out (FP64 physics) -> Vn in TC, so the face kernel reads NV values per
  // side instead of NV D; a boundary point also keeps its gradient trace in Gf for the boundary flux
  if (ok) {
    TC* Gf = pick<TC>(a.Gf, a.Gf32);
    const double idj = 1.0 / dj;
    constexpr int SI = (NAQ + NPT - 1) / NPT;
    // the slots' neighbour flags, normals and own traces, all gathered before the loop (a slot
    // past the end reads slot 0)
    int nbk[SI];
    double nk[SI][D];
    TC uk[SI][NV];
#pragma unroll
    for (int k = 0; k < SI; k++) {
      const int sl = q + k * NPT < NAQ ? q + k * NPT : 0;
      const int64_t es = e * NAQ + sl;
      nbk[k] = g.snb[es];
#pragma unroll
      for (int d = 0; d < D; d++) nk[k][d] = g.snrm[es * D + d];
#pragma unroll
      for (int v = 0; v < NV; v++) uk[k][v] = Uf_in[(e * NV + v) * NAQ + sl];
    }
#pragma unroll
    for (int k = 0; k < SI; k++) {
      const int sl = q + k * NPT;
      if (sl >= NAQ) break;
      const int s = sl / NFQ;
      const int b = sT.base[sl], st = axisStride(sT.axis[s], N);
      const int nbs = nbk[k];
      double n[D], us[NV], gs[NG];
      for (int d = 0; d < D; d++) n[d] = nk[k][d];
      for (int v = 0; v < NV; v++) us[v] = pointState<TC>(a, uk[k][v], v);
      // BR2 own-face lift factor of the slot, times the normal (in the contraction type)
      const double kf = sT.kappa[sT.side[s]] * idj / sT.fwt[sl];
      TC kn[D];
      for (int d = 0; d < D; d++) kn[d] = static_cast<TC>(kf * n[d]);
      const TC* lt = tabLf<TC, N>(sT) + sT.side[s] * N;
      TC gt[NG];
      for (int r = 0; r < NG; r++) {
        TC acc = TC(0);
#pragma unroll
        for (int i = 0; i < N; i++) acc += lt[i] * W[r * NPT + b + i * st];
        if constexpr (VISC == 2) {  // r_s at its own face: kappa_side beta n_d / |J|
          if constexpr (sizeof(TC) == 8)
            acc += kf * Bt[(r / D) * NAQ + sl] * n[r % D];
          else
            acc += Bt[(r / D) * NAQ + sl] * kn[r % D];
        }
        gt[r] = acc;
        gs[r] = static_cast<double>(gt[r]);
        if (nbs < 0) Gf[(e * NG + r) * NAQ + sl] = gt[r];
      }
      double ws[D + 3];
      computational<D>(ph, us, ws);
      TC* Vn = reinterpret_cast<TC*>(a.Vn);  // TC values in the FP64-sized buffer
      if constexpr (sizeof(TC) == 4) {
        float vn[NV];
        normalViscous32<D>(ph, us, ws, gt, n, vn);
        if constexpr (AV) {  // + eps grad U . n_out of the volume gradient, energy reduced by h0 mass
          double av[NV];
          if constexpr (VISC == 2) {
            const double eps = avAt(g, g.nodf, sl, static_cast<int>(e));
            for (int v = 0; v < NV; v++) av[v] = 0.0;
            for (int r = 0; r < NG; r++) {
              double tr = 0.0;
#pragma unroll
              for (int i = 0; i < N; i++) tr += static_cast<double>(lt[i] * W[r * NPT + b + i * st]);
              av[r / D] += eps * tr * n[r % D];
              if (nbs < 0) a.GfV[(e * NG + r) * NAQ + sl] = tr;
            }
          } else {
            for (int v = 0; v < NV; v++) av[v] = avn[k][v];
          }
          av[NV - 1] -= a.mix.h0 * av[0];
          for (int v = 0; v < NV; v++) vn[v] = static_cast<float>(static_cast<double>(vn[v]) + av[v]);
        }
        for (int v = 0; v < NV; v++) Vn[(e * NV + v) * NAQ + sl] = vn[v];
      } else {
        double pg[NG], fv[D * NV];
        primitiveGradient<D>(ph, ws, us, gs, pg);
        viscousRawFlux<D>(ph, ws, pg, fv);
        if constexpr (AV && VISC == 2) {  // + eps grad U with the volume-gradient trace at the slot (without the lift)
          const double eps = avAt(g, g.nodf, sl, static_cast<int>(e));
          for (int r = 0; r < NG; r++) {
            double tr = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) tr += lt[i] * W[r * NPT + b + i * st];
            fv[(r % D) + D * (r / D)] += eps * tr;
            if (nbs < 0) a.GfV[(e * NG + r) * NAQ + sl] = tr;
          }
        }
        for (int v = 0; v < NV; v++) {
          double sum = 0.0;
          for (int d = 0; d < D; d++) sum += fv[d + D * v] * n[d];
          if constexpr (AV && VISC == 1) sum += avn[k][v];
          Vn[(e * NV + v) * NAQ + sl] = static_cast<TC>(sum);
        }
      }
    }
  }
  elemSync<NPT>();  // the face-gradient field is read; sW takes the volume flux
  // volume flux (F_c - F_v)^T (J^T)^-1 |J| w at the node -> sW [b NV + v][node], then its weak
  // divergence sum_axes Dn^T (.) -> HBM Rvol[e][v][node] in TC: the residual kernel continues
  // from there with the face lift
  double src = 0.0;  // Boussinesq source at the node times w |J| (the test function is the node's)
  if (ok) {
    double w[D + 3];
    computational<D>(ph, u, w);
    const double jw = dj * sT.wn[q];
    if (ph.source == kSourceBoussinesq) src = boussinesqSource<D>(ph, w) * jw;
    double G[D * D];
    for (int i = 0; i < D * D; i++) G[i] = gi[i] * jw;
    if constexpr (sizeof(TC) == 4) {
      float out[D * NV];
      if constexpr (AV) {  // eps grad U of the volume gradient, energy row reduced by h0 mass
        double eps = 0.0;
        for (int j = 0; j < g.nvert; j++) eps += t.nodt[q * g.nvert + j] * g.nodeav[g.ntag[e * g.nvert + j]];
        float avg[NG];
        for (int r = 0; r < NG; r++) {
          double gv = static_cast<double>(wface[r]);
          if (r / D == NV - 1) gv -= a.mix.h0 * static_cast<double>(wface[r % D]);
          avg[r] = static_cast<float>(eps * gv);
        }
        volumeFlux32<D>(ph, a, u, w, wtot, G, out, avg);
      } else {
        volumeFlux32<D>(ph, a, u, w, wtot, G, out);
      }
      for (int i = 0; i < D * NV; i++) W[i * NPT + q] = out[i];
    } else {
      double pg[NG], fv[D * NV], out[D * NV];
      primitiveGradient<D>(ph, w, u, wtot, pg);
      viscousRawFlux<D>(ph, w, pg, fv);
      if constexpr (AV) {  // + eps grad U, the volume gradient (alpha lifts only) at the node
        double eps = 0.0;
        for (int j = 0; j < g.nvert; j++) eps += t.nodt[q * g.nvert + j] * g.nodeav[g.ntag[e * g.nvert + j]];
        for (int r = 0; r < NG; r++) fv[(r % D) + D * (r / D)] += eps * wface[r];
      }
      volumeFlux<TC, D, true>(a, w, fv, G, out);
      for (int i = 0; i < D * NV; i++) W[i * NPT + q] = static_cast<TC>(out[i]);
    }
  }
  elemSync<NPT>();
  if (ok) {
    TC R[NV];
    for (int v = 0; v < NV; v++) R[v] = TC(0);
    for (int ax = 0; ax < D; ax++) {
      const int i = nodeIndex(q, ax, N), st = axisStride(ax, N), q0 = q - i * st;
      for (int v = 0; v < NV; v++) {
        TC s = TC(0);
#pragma unroll
        for (int m = 0; m < N; m++) s += tabDn<TC, N>(sT)[m * N + i] * W[(ax * NV + v) * NPT + q0 + m * st];
        R[v] += s;
      }
    }
    R[D] += static_cast<TC>(src);
    TC* Rv = reinterpret_cast<TC*>(a.Rvol);  // TC values in the FP64-sized buffer
    for (int v = 0; v < NV; v++) Rv[(e * NV + v) * NPT + q] = R[v];
  }
}

// ---- kernel 2 (after nsFaceFluxKernel): residual, update, next-stage traces, monitor -----------
template <typename TC, int N>
__global__ __launch_bounds__(TensorShape<N>::TPB) void nsResidualTensor(Geo g, Phys ph, StageArgs a) {
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV, NFQ = S::NFQ, NAQ = S::NAQ, D = 3;
  constexpr bool MIX = sizeof(TC) == 4;
  // face-flux slots (TC), then the nodal increment / residual (FP64), in one region
  constexpr int HB = (NAQ * NV * static_cast<int>(sizeof(TC)) > NV * NPT * 8 ? NAQ * NV * static_cast<int>(sizeof(TC))
                                                                              : NV * NPT * 8) / 8;
  __shared__ double sF[EPB][NV * NPT];  // FP64 scratch of the basis passes
  __shared__ double sHR[EPB][HB];
  __shared__ double sM[EPB][NV];        // monitor partials
  __shared__ TensorLds<N> sT;
  if (a.status && a.status[0]) return;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  TC* sH = reinterpret_cast<TC*>(sHR[lec]);
  double* SR = sHR[lec];
  const TC* Hs = pick<TC>(static_cast<const double*>(a.Hs), static_cast<const float*>(a.Hs32));
  // the state and step-start state for the update, gathered early so that their latency overlaps
  // the residual (modal index of this thread's tensor position)
  double pu[NV], pn[NV], rv[NV];
  if (ok) {
    const int b = t.iperm[q];
    for (int v = 0; v < NV; v++) {
      pu[v] = a.U[(e * NV + v) * NPT + b];
      pn[v] = a.Un[(e * NV + v) * NPT + b];
    }
    // compile-time trip counts, so that all the loads are issued before the first wait
#pragma unroll  // the weak divergence of the volume flux at the node (gradient kernel)
    for (int v = 0; v < NV; v++) rv[v] = reinterpret_cast<const TC*>(a.Rvol)[(e * NV + v) * NPT + q];
    constexpr int HI = (NAQ * NV + NPT - 1) / NPT;
    TC h[HI];
#pragma unroll
    for (int k = 0; k < HI; k++) {
      const int idx = q + k * NPT;
      if (idx < NAQ * NV) h[k] = Hs[(e * NV + idx / NAQ) * NAQ + idx % NAQ];
    }
#pragma unroll
    for (int k = 0; k < HI; k++) {  // HBM [v][slot] -> LDS [slot][v]
      const int idx = q + k * NPT;
      if (idx < NAQ * NV) sH[(idx % NAQ) * NV + idx / NAQ] = h[k];
    }
  }
  __syncthreads();
  double R[NV];
  if (ok) {
    for (int v = 0; v < NV; v++) R[v] = rv[v];
#pragma unroll 1
    for (int s = 0; s < 6; s++) {
      const int i = nodeIndex(q, t.axis[s], N);
      const int sl = s * NFQ + sT.slot_of[s * NPT + q];
      const double l = sT.lf[t.side[s] * N + i];
      for (int v = 0; v < NV; v++) R[v] += l * static_cast<double>(sH[sl * NV + v]);
    }
    if constexpr (MIX) R[NV - 1] += a.mix.h0 * R[0];  // energy: add back h0 R_rho
  }
  elemSync<NPT>();  // the face-flux slots are read; their region takes the nodal residual
  const bool monitor = a.mode == 1 || a.last_stage;
  const double jw = ok ? g.detj[e] * sT.wn[q] : 1.0;
  if (ok) {  // residual R_psi for the monitor stages, else directly the nodal increment R_psi / (w |J|)
    const double ijw = 1.0 / jw;
    for (int v = 0; v < NV; v++) SR[v * NPT + q] = monitor ? R[v] : R[v] * ijw;
  }
  elemSync<NPT>();
  // the reference residual (before M^-1) is (B^T x B^T x B^T) R_psi (modal); its monitor is
  // mean_q |R V^T| = mean over nodes of |(B x B x B) of it| (TimeIntegration.cpp:279-298)
  double* T0 = sF[lec];
  auto pass64 = [&](const double* M, bool transpose, int axis, const double* in, double* out) {
    if (!ok) return;
    const int i = nodeIndex(q, axis, N), st = axisStride(axis, N), q0 = q - i * st;
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int k = 0; k < N; k++) s += (transpose ? M[k * N + i] : M[i * N + k]) * in[v * NPT + q0 + k * st];
      out[v * NPT + q] = s;
    }
  };
  if (monitor) {
    pass64(sT.B, true, 0, SR, T0);
    elemSync<NPT>();
    pass64(sT.B, true, 1, T0, SR);
    elemSync<NPT>();
    pass64(sT.B, true, 2, SR, T0);  // T0: modal residual, tensor order
    elemSync<NPT>();
    if (a.mode == 1) {
      if (ok)
        for (int v = 0; v < NV; v++) a.Rout[(e * NV + v) * NPT + sT.iperm[q]] = T0[v * NPT + q];
      return;
    }
    pass64(sT.B, false, 0, T0, SR);
    elemSync<NPT>();
    pass64(sT.B, false, 1, SR, T0);
    elemSync<NPT>();
    pass64(sT.B, false, 2, T0, SR);  // SR: R V^T at the nodes
    elemSync<NPT>();
    if (le < EPB && q < NV) {
      double s = 0.0;
      for (int k = 0; k < NPT; k++) s += fabs(SR[q * NPT + k]);
      sM[le][q] = s;
    }
    elemSync<NPT>();
    if (ok && q < NV) a.mon[e * NV + q] = sM[le][q] / NPT;
    // the nodal increment R_psi / (w |J|) from the registers (SR was overwritten)
    elemSync<NPT>();
    if (ok) {
      const double ijw = 1.0 / jw;
      for (int v = 0; v < NV; v++) SR[v * NPT + q] = R[v] * ijw;
    }
    elemSync<NPT>();
  }
  // nodal increment -> modal with B^-1 along each axis (FP64)
  pass64(sT.Binv, false, 0, SR, T0);
  elemSync<NPT>();
  pass64(sT.Binv, false, 1, T0, SR);
  elemSync<NPT>();
  pass64(sT.Binv, false, 2, SR, T0);  // T0: modal increment, tensor order
  elemSync<NPT>();
  // RK update (FP64) in the modal order; the new coefficients (TC operands, tensor order) -> X
  TC* X = reinterpret_cast<TC*>(SR);  // SR is free again
  if (ok) {
    const int b = sT.iperm[q];
    for (int v = 0; v < NV; v++) {
      const int64_t gi = (e * NV + v) * NPT + b;
      double u = pu[v] * a.c1;
      u += a.c0 * pn[v];
      u += a.c2dt * T0[v * NPT + q];
      a.U[gi] = u;
      if (a.last_stage) a.Un[gi] = u;
      if (a.rec) {
        const int step = a.status[1] + (a.last_stage ? 1 : 0);
        const int stage = a.last_stage ? 0 : a.rec_stage + 1;
        if (step < a.rec_nsteps) a.rec[(static_cast<int64_t>(step) * a.rec_S + stage) * a.rec_n + gi] = u;
      }
      if constexpr (MIX)
        X[v * NPT + q] = deviationOperand(u, a.mix.uref[v], a.mix.one[b]);
      else
        X[v * NPT + q] = u;
    }
  }
  elemSync<NPT>();
  TC* Y = reinterpret_cast<TC*>(T0);
  linePass<TC, N, NV>(tabB<TC, N>(sT), 0, X, Y, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 1, Y, X, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 2, X, Y, q, ok);  // Y: new nodal state
  elemSync<NPT>();
  double* rec_t = nullptr;
  if (a.rec_t) {
    const int step = a.status[1] + (a.last_stage ? 1 : 0);
    const int stage = a.last_stage ? 0 : a.rec_stage + 1;
    if (step < a.rec_nsteps) rec_t = a.rec_t + (static_cast<int64_t>(step) * a.rec_S + stage) * a.rec_nt;
  }
  if (le < EPB && e < g.E) writeTraces<TC, N>(sT, Y, pick<TC>(a.Uf_out, a.Uf_out32), e, le, tid, rec_t);
  (void)ph;
}

// traces from the state in HBM with the residual kernel's own arithmetic (restarts, checkpoints)
template <typename TC, int N>
__global__ __launch_bounds__(TensorShape<N>::TPB) void tensorTraceKernel(Geo g, const double* U, void* Uf_out, MixRef mix) {
  const TensorOps& t = g.tens;
  using S = TensorShape<N>;
  constexpr int NPT = S::NPT, EPB = S::EPB, NV = S::NV;
  __shared__ TC sX[EPB][NV * NPT];
  __shared__ TC sY[EPB][NV * NPT];
  __shared__ TensorLds<N> sT;
  stageTables<N>(t, sT);
  const int tid = threadIdx.x, le = localElement<NPT>(tid), q = tid % NPT;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPB + le;
  const bool ok = le < EPB && e < g.E;
  const int lec = le < EPB ? le : 0;
  TC* X = sX[lec];
  TC* Y = sY[lec];
  if (ok) {
    const int tq = t.perm[q];
    for (int v = 0; v < NV; v++) {
      const double u = U[(e * NV + v) * NPT + q];
      if constexpr (sizeof(TC) == 8)
        X[v * NPT + tq] = u;
      else
        X[v * NPT + tq] = deviationOperand(u, mix.uref[v], mix.one[q]);
    }
  }
  __syncthreads();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 0, X, Y, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 1, Y, X, q, ok);
  elemSync<NPT>();
  linePass<TC, N, NV>(tabB<TC, N>(sT), 2, X, Y, q, ok);
  elemSync<NPT>();
  if (ok) writeTraces<TC, N>(sT, Y, static_cast<TC*>(Uf_out), e, le, tid);
}

}  // namespace sdg

#endif  // SDG_DG_TENSOR_HPP_
