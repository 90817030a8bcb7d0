// sdg_batch.hip -- batched solver: many independent 2-D compressible Euler problems on meshes of
// mixed triangles and quadrangles, orders 1..4, advanced together. An RK stage of the whole batch
// is one persistent launch per element variant present (element type and order): each variant's
// kernel has its tile shape and numerical flux fixed at compile time and strides over that
// variant's work items (first element, count); the variants' launches run side by side on forked
// streams, which a captured step turns into parallel branches of one hipGraph. The work-item size
// and the grid of each variant are picked by a measuring autotuner (sdg_batch_tune).
//
// Per stage and element this is the reference's stepSolver pass chain for the Euler model
// (TimeIntegration.cpp:327-350, SpatialDiscrete.cpp:194-266,633-748,1016-1032): volume flux,
// numerical face flux (evaluated element-centrically with the face's own left/right roles, so both
// parents get bit-identical +-H), residual, RK update; the traces of the new state are written for
// the next stage. In the mixed-precision mode the traces and contractions are FP32 on deviations
// from each problem's reference state (as in the single-problem engine); state, geometry,
// pointwise physics and the RK accumulation stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "kit.hpp"
#include "physics.hpp"
#include "sdg_internal.hpp"

namespace sdg {
namespace batch {

constexpr int kThreads = 256;
constexpr int kVariants = 8;  // code = (quadrangle ? 4 : 0) + P - 1
constexpr int NV = 4, D = 2;

// Per variant (element type and order) the operator tables, all in the orthonormal basis
// psi = L^-1 phi of the reference element (Mref = L L^T): with it the mass matrix of an affine
// element is |J| I and the RK update of a mode needs only that mode's residual. The state is kept
// in psi coefficients on the device (sdg_batch_set_state / get_state convert), which is the same
// scheme as the reference's hierarchic basis in exact arithmetic.
struct VariantDev {
  const double* Vt;    // NB x NQ   psi at volume points
  const double* dV;    // (NQ D) x NB reference gradients of psi (test functions)
  const double* Vft;   // NB x NAQ  psi at face points
  const double* Vf;    // NAQ x NB  the same, transposed (test functions)
  const double* one;   // NB psi coefficients of the constant 1
  const double* wq;    // NQ
  int64_t E, u_base, t_base, f_base, e_base;
};

struct WorkItem {
  int code;
  int count;
  int64_t first;
};

// Per element face (index f_base + e N_F + s): the neighbour's trace index of its point 0 and
// variable 0, the face (normal and |J_f| w per point, shared by both parents) and flags: bit 0 the
// element is the face's right parent, bit 1 the neighbour's points run in reverse order.
template <typename TC>
struct Args {
  double* U;
  double* Un;
  const TC* Tin;
  TC* Tout;
  const double* ginv;   // per element: (J^T)^-1 column-major 2 x 2
  const double* detj;   // per element
  const int* pid;       // per element: problem
  const int64_t* fnb;   // per element face: neighbour trace index of its point 0, variable 0
  const int* fid;       // per element face: face index (rows of fnrm, fjw)
  const int8_t* fflag;  // per element face: side | reversed << 1
  const double* fnrm;   // per face: normal (outward from the left parent), constant on the face
  const double* fjw;    // per face: |J_f| w at the left parent's points, stride kMaxNfq
  const double* dt;     // per problem
  const double* mix;    // per problem: uref[4], pref, h0
  const VariantDev* var;
  const WorkItem* items;
  int nitems;
  int last_stage;
  double c0, c1, c2;
  Phys ph;
};

constexpr int kMaxNfq = 5;
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Elements per panel: enough that the point phase (one thread per volume point or face point)
// covers the workgroup, within an LDS budget for the panel (state, volume flux, face flux) so that
// several workgroups share a CU.
template <typename TC, int NB, int NQ, int NF, int NFQ>
struct Shape {
  static constexpr int NAQ = NF * NFQ;
  static constexpr int PANEL = NV * (NB + NQ * D + NAQ);  // LDS values per element
  static constexpr int BUDGET = 16384;                     // bytes of panel per workgroup
  static constexpr int EPB_T = cmax(1, kThreads / (NQ + NAQ) + 1);
  static constexpr int EPB_L = cmax(1, BUDGET / (PANEL * static_cast<int>(sizeof(TC))));
  static constexpr int EPB = cmin(cmin(EPB_T, EPB_L), 64);
  static constexpr int LDS_ELEMS = EPB * PANEL;
  static constexpr int TAB_ELEMS = NB * NQ + NQ * D * NB + 2 * NB * NAQ;
};

// numerical flux at one face point (FLUX: the convective flux, fixed at compile time)
template <int FLUX>
__device__ __forceinline__ void faceFluxPoint(const Phys& ph, const double* n, const double* ul, const double* ur, double* h) {
  double wl[D + 3], wr[D + 3];
  computational<D>(ph, ul, wl);
  computational<D>(ph, ur, wr);
  convectiveFlux<D, double, FLUX>(ph, n, ul, wl, ur, wr, h);
}

template <typename TC>
__device__ __forceinline__ double pointValue(const Args<TC>& a, int p, TC x, int v) {
  if constexpr (sizeof(TC) == 8)
    return x;
  else
    return a.mix[p * 6 + v] + static_cast<double>(x);
}

// The contraction operand of a state coefficient: the coefficient itself (FP64), or in the mixed
// mode its deviation from the problem's reference state rounded to FP32. Every site that forms it
// (panel load, stage update, trace kernel) uses this one expression, so the traces an element reads
// for itself and the ones its neighbours read are bitwise the same.
template <typename TC>
__device__ __forceinline__ TC stateOperand(double u, double uref, double one) {
  if constexpr (sizeof(TC) == 8)
    return u;
  else
    return static_cast<TC>(__fma_rn(-uref, one, u));
}

// value of the panel state of one variable at face slot sl (same loop order everywhere)
template <typename TC, int NB, int NAQ>
__device__ __forceinline__ TC traceAt(const TC* su, const TC* tVft, int sl) {
  TC t = TC(0);
#pragma unroll 5
  for (int b = 0; b < NB; b++) t += su[b] * tVft[b * NAQ + sl];
  return t;
}

// One variant's elements [first, first + count), EPB at a time (out of line: each variant keeps its
// own register allocation). Per panel, with a barrier after each:
//   load   state panel (coalesced; all loads issued before the first wait)
//   A      one thread per volume point: point values (FP32 / FP64 contraction), FP64 flux
//          F^T (J^T)^-1 |J| w; one thread per face slot: the own trace from the panel, the
//          neighbour's from HBM, the numerical flux with the face's left/right roles
//   B      one thread per mode: R = F . grad psi - H psi, RK update in FP64 (U, U_n), new state
//          operand into the panel
//   C      traces of the new state for the neighbours (slot-major, the variables of a slot adjacent)
template <typename TC, int FLUX, int NB, int NQ, int NF, int NFQ>
__device__ __forceinline__ void eulerTile(const Args<TC>& a, const VariantDev& vd, int64_t first, int count, TC* smem,
                                          TC* tab, bool reload) {
  using S = Shape<TC, NB, NQ, NF, NFQ>;
  constexpr int NAQ = S::NAQ, EPB = S::EPB;
  constexpr bool MIX = sizeof(TC) == 4;
  TC(*sU)[NV][NB] = reinterpret_cast<TC(*)[NV][NB]>(smem);
  TC(*sF)[NV][NQ * D] = reinterpret_cast<TC(*)[NV][NQ * D]>(smem + EPB * NV * NB);
  TC(*sH)[NV][NAQ] = reinterpret_cast<TC(*)[NV][NAQ]>(smem + EPB * NV * (NB + NQ * D));
  TC* tVt = tab;                  // [NB][NQ]
  TC* tdV = tVt + NB * NQ;        // [NQ D][NB]
  TC* tVft = tdV + NQ * D * NB;   // [NB][NAQ]
  TC* tVf = tVft + NB * NAQ;      // [NAQ][NB]
  const int tid = threadIdx.x;
  if (reload) {
    for (int i = tid; i < NB * NQ; i += kThreads) tVt[i] = static_cast<TC>(vd.Vt[i]);
    for (int i = tid; i < NQ * D * NB; i += kThreads) tdV[i] = static_cast<TC>(vd.dV[i]);
    for (int i = tid; i < NB * NAQ; i += kThreads) tVft[i] = static_cast<TC>(vd.Vft[i]);
    for (int i = tid; i < NAQ * NB; i += kThreads) tVf[i] = static_cast<TC>(vd.Vf[i]);
  }
#pragma unroll 1
  for (int64_t c0 = first; c0 < first + count; c0 += EPB) {
    const int ne = static_cast<int>(first + count - c0 < EPB ? first + count - c0 : EPB);
    // state panel: EPB consecutive elements are one contiguous stretch of U
    {
      constexpr int IT = (EPB * NV * NB + kThreads - 1) / kThreads;
      double u[IT];
#pragma unroll
      for (int k = 0; k < IT; k++) {
        const int idx = tid + k * kThreads;
        const int le = idx / (NV * NB);
        const int64_t e = c0 + (le < ne ? le : 0);
        u[k] = a.U[vd.u_base + e * NV * NB + idx % (NV * NB)];
      }
      double ur[IT], on[IT];
#pragma unroll
      for (int k = 0; k < IT; k++) {
        const int idx = tid + k * kThreads;
        const int le = idx / (NV * NB), rem = idx % (NV * NB);
        if constexpr (MIX) {
          ur[k] = a.mix[a.pid[vd.e_base + c0 + (le < ne ? le : 0)] * 6 + rem / NB];
          on[k] = vd.one[rem % NB];
        } else {
          ur[k] = on[k] = 0.0;
        }
      }
#pragma unroll
      for (int k = 0; k < IT; k++) {
        const int idx = tid + k * kThreads;
        if (idx < EPB * NV * NB) {
          const int le = idx / (NV * NB), rem = idx % (NV * NB);
          sU[le][rem / NB][rem % NB] = le < ne ? stateOperand<TC>(u[k], ur[k], on[k]) : TC(0);
        }
      }
    }
    __syncthreads();
    // A: volume points and face slots in one index space
#pragma unroll 1
    for (int idx = tid; idx < EPB * (NQ + NAQ); idx += kThreads) {
      if (idx < EPB * NQ) {
        const int le = idx / NQ, q = idx % NQ;
        if (le >= ne) {
          for (int v = 0; v < NV; v++)
            for (int b = 0; b < D; b++) sF[le][v][q * D + b] = TC(0);
          continue;
        }
        const int64_t eg = vd.e_base + c0 + le;
        const int p = a.pid[eg];
        double G[D * D];
        const double jw = a.detj[eg] * vd.wq[q];
        for (int i = 0; i < D * D; i++) G[i] = a.ginv[eg * D * D + i] * jw;
        TC s[NV] = {};
#pragma unroll 5
        for (int b = 0; b < NB; b++) {
          const TC t = tVt[b * NQ + q];
#pragma unroll
          for (int v = 0; v < NV; v++) s[v] += sU[le][v][b] * t;
        }
        double u[NV], w[D + 3];
        for (int v = 0; v < NV; v++) u[v] = pointValue<TC>(a, p, s[v], v);
        computational<D>(a.ph, u, w);
        const double rho = w[0], pr = w[D + 2];
        const double et = w[D + 1] + velocitySquared<D>(w) / 2.0;
        double fc[D * NV];
        for (int aa = 0; aa < D; aa++) {
          fc[aa] = rho * w[1 + aa];
          for (int b = 0; b < D; b++) {
            double pd;
            if constexpr (MIX)
              pd = aa == b ? pr - a.mix[p * 6 + 4] : 0.0;
            else
              pd = aa == b ? pr * 1.0 : pr * 0.0;
            fc[aa + D * (1 + b)] = rho * w[1 + aa] * w[1 + b] + pd;
          }
          if constexpr (MIX)
            fc[aa + D * (D + 1)] = (rho * et + pr - a.mix[p * 6 + 5] * rho) * w[1 + aa];
          else
            fc[aa + D * (D + 1)] = (rho * et + pr) * w[1 + aa];
        }
        for (int v = 0; v < NV; v++)
          for (int b = 0; b < D; b++) {
            double t = 0.0;
            for (int aa = 0; aa < D; aa++) t += fc[aa + D * v] * G[aa + D * b];
            sF[le][v][q * D + b] = static_cast<TC>(t);
          }
      } else {
        const int j2 = idx - EPB * NQ;
        const int le = j2 / NAQ, sl = j2 % NAQ;
        if (le >= ne) {
          for (int v = 0; v < NV; v++) sH[le][v][sl] = TC(0);
          continue;
        }
        const int64_t e = c0 + le, eg = vd.e_base + e;
        const int s = sl / NFQ, j = sl % NFQ;
        const int64_t ef = vd.f_base + e * NF + s;
        const int64_t nb0 = a.fnb[ef];
        const int fl = a.fflag[ef], face = a.fid[ef];
        const int side = fl & 1, kpt = (fl & 2) ? NFQ - 1 - j : j;
        const int p = a.pid[eg];
        double ul[NV], ur[NV];
        for (int v = 0; v < NV; v++) {
          const double uo = pointValue<TC>(a, p, traceAt<TC, NB, NAQ>(sU[le][v], tVft, sl), v);
          const double un = pointValue<TC>(a, p, a.Tin[nb0 + kpt * NV + v], v);
          ul[v] = side == 0 ? uo : un;
          ur[v] = side == 0 ? un : uo;
        }
        double n[D], h[NV];
        n[0] = a.fnrm[static_cast<int64_t>(face) * D];
        n[1] = a.fnrm[static_cast<int64_t>(face) * D + 1];
        faceFluxPoint<FLUX>(a.ph, n, ul, ur, h);
        if constexpr (MIX) {
          h[1] -= a.mix[p * 6 + 4] * n[0];
          h[2] -= a.mix[p * 6 + 4] * n[1];
          h[3] -= a.mix[p * 6 + 5] * h[0];
        }
        const double jw = a.fjw[static_cast<int64_t>(face) * kMaxNfq + (side == 0 ? j : kpt)];
        for (int v = 0; v < NV; v++) {
          const double t = h[v] * jw;
          sH[le][v][sl] = static_cast<TC>(side == 0 ? t : -t);
        }
      }
    }
    __syncthreads();
    // B: residual against psi_i and the RK update in FP64 (TimeIntegration.cpp:181-198):
    // U <- c1 U + c0 U_n + c2 dt R / |J|
#pragma unroll 1
    for (int idx = tid; idx < EPB * NB; idx += kThreads) {
      const int le = idx / NB, i = idx % NB;
      if (le >= ne) continue;
      const int64_t e = c0 + le, eg = vd.e_base + e;
      const int p = a.pid[eg];
      double uo[NV], uno[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int64_t gi = vd.u_base + (e * NV + v) * NB + i;
        uo[v] = a.U[gi];
        uno[v] = a.c0 != 0.0 ? a.Un[gi] : 0.0;
      }
      TC r[NV] = {}, hs[NV] = {};
#pragma unroll 4
      for (int k = 0; k < D * NQ; k++) {
        const TC t = tdV[k * NB + i];
#pragma unroll
        for (int v = 0; v < NV; v++) r[v] += sF[le][v][k] * t;
      }
#pragma unroll 4
      for (int k = 0; k < NAQ; k++) {
        const TC t = tVf[k * NB + i];
#pragma unroll
        for (int v = 0; v < NV; v++) hs[v] += sH[le][v][k] * t;
      }
      double rm[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) rm[v] = static_cast<double>(r[v] - hs[v]);
      if constexpr (MIX) rm[NV - 1] += a.mix[p * 6 + 5] * rm[0];  // energy: add back h0 R_rho
      const double scale = a.c2 * a.dt[p] / a.detj[eg];
      const double onei = MIX ? vd.one[i] : 0.0;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int64_t gi = vd.u_base + (e * NV + v) * NB + i;
        double u = uo[v] * a.c1;
        u += a.c0 * uno[v];
        u += scale * rm[v];
        a.U[gi] = u;
        if (a.last_stage) a.Un[gi] = u;
        sU[le][v][i] = stateOperand<TC>(u, MIX ? a.mix[p * 6 + v] : 0.0, onei);
      }
    }
    __syncthreads();
    // C: traces of the new state, per slot the NV variables adjacent
#pragma unroll 1
    for (int idx = tid; idx < EPB * NAQ; idx += kThreads) {
      const int le = idx / NAQ, sl = idx % NAQ;
      if (le >= ne) continue;
      TC* out = a.Tout + vd.t_base + ((c0 + le) * NAQ + sl) * NV;
#pragma unroll
      for (int v = 0; v < NV; v++) out[v] = traceAt<TC, NB, NAQ>(sU[le][v], tVft, sl);
    }
    __syncthreads();
  }
}

// variant table: (NB, NQ, NF, NFQ) of triangles and quadrangles P1..P4
#define SDG_BATCH_VARIANTS(X) \
  X(0, 3, 3, 3, 2)            \
  X(1, 6, 6, 3, 3)            \
  X(2, 10, 12, 3, 4)          \
  X(3, 15, 16, 3, 5)          \
  X(4, 4, 4, 4, 2)            \
  X(5, 9, 9, 4, 3)            \
  X(6, 16, 16, 4, 4)          \
  X(7, 25, 25, 4, 5)

// elements per panel of variant `code` in the precision of TC
template <typename TC>
constexpr int epbOf(int code) {
#define SDG_BATCH_EPB(c, nb, nq, nf, nfq) code == c ? Shape<TC, nb, nq, nf, nfq>::EPB :
  return SDG_BATCH_VARIANTS(SDG_BATCH_EPB) 1;
#undef SDG_BATCH_EPB
}

// One RK stage of one variant's elements (all problems): a persistent grid striding over that
// variant's work items [begin, end). Each variant has its own kernel, so that its registers and LDS
// are sized for its element type and order alone; the variants of a stage run concurrently (graph
// branches).
template <typename TC, int FLUX, int CODE, int NB, int NQ, int NF, int NFQ>
__global__ __launch_bounds__(kThreads) void batchVariantKernel(Args<TC> a, int begin, int end) {
  using S = Shape<TC, NB, NQ, NF, NFQ>;
  __shared__ TC smem[S::LDS_ELEMS];
  __shared__ TC tab[S::TAB_ELEMS];
  const VariantDev& vd = a.var[CODE];
  bool reload = true;
  for (int it = begin + static_cast<int>(blockIdx.x); it < end; it += static_cast<int>(gridDim.x)) {
    const WorkItem w = a.items[it];
    eulerTile<TC, FLUX, NB, NQ, NF, NFQ>(a, vd, w.first, w.count, smem, tab, reload);
    reload = false;
  }
}

template <typename TC>
using VariantKernel = void (*)(Args<TC>, int, int);

// the kernel of (flux, variant)
template <typename TC, int FLUX>
VariantKernel<TC> variantKernelOf(int code) {
  switch (code) {
#define SDG_BATCH_KERNEL(c, nb, nq, nf, nfq) \
  case c:                                    \
    return batchVariantKernel<TC, FLUX, c, nb, nq, nf, nfq>;
    SDG_BATCH_VARIANTS(SDG_BATCH_KERNEL)
#undef SDG_BATCH_KERNEL
    default:
      return nullptr;
  }
}
template <typename TC>
VariantKernel<TC> variantKernel(int flux, int code) {
  switch (flux) {
    case 0: return variantKernelOf<TC, 0>(code);
    case 1: return variantKernelOf<TC, 1>(code);
    case 2: return variantKernelOf<TC, 2>(code);
    case 3: return variantKernelOf<TC, 3>(code);
    default: return variantKernelOf<TC, 4>(code);
  }
}

// traces of the state (initial state, or after host writes), slot-major like the stage kernel's
template <typename TC>
__global__ __launch_bounds__(kThreads) void batchTraceKernel(Args<TC> a, int code, int nb_, int naq) {
  const VariantDev& vd = a.var[code];
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= vd.E * naq) return;
  const int64_t e = idx / naq;
  const int sl = static_cast<int>(idx % naq);
  const int p = a.pid[vd.e_base + e];
  for (int v = 0; v < NV; v++) {
    TC t = TC(0);
    for (int b = 0; b < nb_; b++) {
      const double u = a.U[vd.u_base + (e * NV + v) * nb_ + b];
      t += stateOperand<TC>(u, sizeof(TC) == 4 ? a.mix[p * 6 + v] : 0.0, sizeof(TC) == 4 ? vd.one[b] : 0.0) *
           static_cast<TC>(vd.Vft[b * naq + sl]);
    }
    a.Tout[vd.t_base + idx * NV + v] = t;
  }
}

// CFL time step per element (TimeIntegration.cpp:104-131)
__global__ __launch_bounds__(kThreads) void batchDtKernel(const double* U, const VariantDev* var, int code, int nb_,
                                                          int nq, int order, Phys ph, const double* min_edge, double cfl,
                                                          double* out) {
  const VariantDev& vd = var[code];
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= vd.E) return;
  double best = 1.0e308;
  for (int q = 0; q < nq; q++) {
    double u[NV], w[D + 3];
    for (int v = 0; v < NV; v++) {
      double s = 0.0;
      for (int b = 0; b < nb_; b++) s += U[vd.u_base + (e * NV + v) * nb_ + b] * vd.Vt[b * nq + q];
      u[v] = s;
    }
    computational<D>(ph, u, w);
    const double c = soundSpeedOf(ph, w[0], w[D + 2]);
    const double sr = sqrt(velocitySquared<D>(w)) + c;
    best = fmin(best, cfl * min_edge[vd.e_base + e] / (sr * (order + 1.0) * (order + 1.0)));
  }
  out[vd.e_base + e] = best;
}

}  // namespace batch
}  // namespace sdg

// ---- host side ------------------------------------------------------------------------------------
struct sdg_batch {
  std::string err;
  sdg_problem_desc prob{};
  sdg::Phys phys{};
  bool mixed = false;
  int nprob = 0;
  std::vector<const sdg::MixedMesh*> meshes;
  struct Var {
    bool used = false;
    int type = 0, order = 0, nb = 0, nq = 0, nf = 0, nfq = 0, naq = 0;
    int64_t E = 0, u_base = 0, t_base = 0, f_base = 0, e_base = 0;
    sdg::RefElement ref;
    std::vector<double> one, mean;
  } var[sdg::batch::kVariants];
  std::vector<std::vector<std::pair<int, int64_t>>> emap;  // problem, element -> (variant, index)
  std::vector<std::vector<int64_t>> coeff_off;             // problem, element -> offset in reference layout
  int64_t totalE = 0, totalU = 0, totalT = 0, totalF = 0;
  double *U = nullptr, *Un = nullptr, *ginv = nullptr, *detj = nullptr, *fnrm = nullptr, *fjw = nullptr, *dt = nullptr,
         *mix = nullptr, *min_edge = nullptr, *dtbuf = nullptr;
  int* pid = nullptr;
  int64_t* fnb = nullptr;
  int* fid = nullptr;
  int8_t* fflag = nullptr;
  // per variant the reference element's Cholesky factor L of Mref (row-major) and L^-1: the
  // device state is in psi = L^-1 phi coefficients, c_psi = L^T c_phi
  std::vector<double> chol[sdg::batch::kVariants], chol_inv[sdg::batch::kVariants];
  void* T[2] = {nullptr, nullptr};
  sdg::batch::VariantDev* dvar = nullptr;
  sdg::batch::WorkItem* ditems = nullptr;
  int ditems_cap = 0;
  std::vector<sdg::batch::WorkItem> items;
  int chunk[sdg::batch::kVariants] = {1, 1, 1, 1, 1, 1, 1, 1};  // work-item size in tiles
  int vgrid[sdg::batch::kVariants] = {};                         // persistent grid per variant kernel
  int ibeg[sdg::batch::kVariants] = {}, iend[sdg::batch::kVariants] = {};  // work items per variant
  // the variants of a stage run on their own streams, forked from and joined back into `stream`
  hipStream_t aux[sdg::batch::kVariants] = {};
  hipEvent_t fork_ev = nullptr, join_ev[sdg::batch::kVariants] = {};
  std::vector<double> mixh;
  std::vector<double> host_dt;
  std::vector<char> has_state;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int cur = 0;
  float last_ms = 0.0f;
  std::vector<void*> allocs;
};

namespace {

using sdg::batch::Args;
using sdg::batch::VariantDev;
using sdg::batch::WorkItem;

int bfail(sdg_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  sdg::setGlobalError(msg);
  return code;
}

#define BHIP(b, call)                                                                                     \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return bfail((b), SDG_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int balloc(sdg_batch* b, size_t count, T** out) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return bfail(b, SDG_E_NOMEM, "hipMalloc failed");
  b->allocs.push_back(p);
  *out = static_cast<T*>(p);
  return SDG_OK;
}
template <typename T>
int bupload(sdg_batch* b, const std::vector<T>& h, T** out) {
  int rc = balloc(b, h.size(), out);
  if (rc) return rc;
  if (!h.empty() && hipMemcpy(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return bfail(b, SDG_E_DEVICE, "hipMemcpy failed");
  return SDG_OK;
}

int variantCode(int type, int order) { return (type == SDG_ELEM_QUADRANGLE ? 4 : 0) + order - 1; }

template <typename TC>
Args<TC> makeArgs(sdg_batch* b, int parity) {
  Args<TC> a{};
  a.U = b->U;
  a.Un = b->Un;
  a.Tin = static_cast<const TC*>(b->T[parity]);
  a.Tout = static_cast<TC*>(b->T[parity ^ 1]);
  a.ginv = b->ginv;
  a.detj = b->detj;
  a.pid = b->pid;
  a.fnb = b->fnb;
  a.fid = b->fid;
  a.fflag = b->fflag;
  a.fnrm = b->fnrm;
  a.fjw = b->fjw;
  a.dt = b->dt;
  a.mix = b->mix;
  a.var = b->dvar;
  a.items = b->ditems;
  a.nitems = static_cast<int>(b->items.size());
  a.ph = b->phys;
  return a;
}

void rkCoeff(const sdg_batch* b, int stage, double c[3]) {  // TimeIntegration.cpp:45-65
  const int rk = b->prob.time_integration;
  if (rk == SDG_RK_FORWARD_EULER) {
    c[0] = 1.0; c[1] = 0.0; c[2] = 1.0;
  } else if (rk == SDG_RK_HEUN2) {
    const double t[2][3] = {{1.0, 0.0, 1.0}, {0.5, 0.5, 0.5}};
    for (int i = 0; i < 3; i++) c[i] = t[stage][i];
  } else {
    const double t[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    for (int i = 0; i < 3; i++) c[i] = t[stage][i];
  }
}
int stagesOf(const sdg_batch* b) {
  return b->prob.time_integration == SDG_RK_FORWARD_EULER ? 1 : (b->prob.time_integration == SDG_RK_HEUN2 ? 2 : 3);
}

// work items: per variant, chunks of chunk[v] tiles
int buildItems(sdg_batch* b, const int* only = nullptr) {
  b->items.clear();
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    const auto& V = b->var[v];
    b->ibeg[v] = b->iend[v] = static_cast<int>(b->items.size());
    if (!V.used || (only && v != *only)) continue;
    const int epb = b->mixed ? sdg::batch::epbOf<float>(v) : sdg::batch::epbOf<double>(v);
    const int64_t step = static_cast<int64_t>(b->chunk[v]) * epb;
    for (int64_t f = 0; f < V.E; f += step) b->items.push_back({v, static_cast<int>(std::min<int64_t>(step, V.E - f)), f});
    b->iend[v] = static_cast<int>(b->items.size());
  }
  if (static_cast<int>(b->items.size()) > b->ditems_cap) {
    sdg::batch::WorkItem* p = nullptr;
    if (balloc(b, b->items.size(), &p)) return SDG_E_NOMEM;
    b->ditems = p;
    b->ditems_cap = static_cast<int>(b->items.size());
  }
  if (!b->items.empty())
    BHIP(b, hipMemcpy(b->ditems, b->items.data(), b->items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  for (int i = 0; i < 2; i++) {
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
    b->graph[i] = nullptr;
  }
  return SDG_OK;
}

template <typename TC>
void launchVariant(sdg_batch* b, const Args<TC>& a, int v, hipStream_t st) {
  const int n = b->iend[v] - b->ibeg[v];
  if (n <= 0) return;
  const int grid = std::max(1, std::min(b->vgrid[v], n));
  hipLaunchKernelGGL(sdg::batch::variantKernel<TC>(b->prob.convective_flux, v), dim3(grid), dim3(sdg::batch::kThreads), 0,
                     st, a, b->ibeg[v], b->iend[v]);
}

// One RK stage: every used variant's kernel, on its own stream forked from b->stream and joined
// back (concurrent branches when captured in a graph); only >= 0: that variant alone on b->stream.
template <typename TC>
void launchStageT(sdg_batch* b, int stage, int parity, int only) {
  double c[3];
  rkCoeff(b, stage, c);
  Args<TC> a = makeArgs<TC>(b, parity);
  a.c0 = c[0];
  a.c1 = c[1];
  a.c2 = c[2];
  a.last_stage = stage == stagesOf(b) - 1;
  if (only >= 0) {
    launchVariant<TC>(b, a, only, b->stream);
    return;
  }
  (void)hipEventRecord(b->fork_ev, b->stream);
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    if (b->iend[v] <= b->ibeg[v]) continue;
    (void)hipStreamWaitEvent(b->aux[v], b->fork_ev, 0);
    launchVariant<TC>(b, a, v, b->aux[v]);
    (void)hipEventRecord(b->join_ev[v], b->aux[v]);
    (void)hipStreamWaitEvent(b->stream, b->join_ev[v], 0);
  }
}
void launchStage(sdg_batch* b, int stage, int parity, int only = -1) {
  if (b->mixed)
    launchStageT<float>(b, stage, parity, only);
  else
    launchStageT<double>(b, stage, parity, only);
}

void launchTraces(sdg_batch* b, int parity) {
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    const auto& V = b->var[v];
    if (!V.used) continue;
    const int64_t n = V.E * V.naq;
    const dim3 grid(static_cast<unsigned>((n + sdg::batch::kThreads - 1) / sdg::batch::kThreads));
    if (b->mixed) {
      Args<float> a = makeArgs<float>(b, parity ^ 1);  // writes T[parity]
      hipLaunchKernelGGL(sdg::batch::batchTraceKernel<float>, grid, dim3(sdg::batch::kThreads), 0, b->stream, a, v, V.nb, V.naq);
    } else {
      Args<double> a = makeArgs<double>(b, parity ^ 1);
      hipLaunchKernelGGL(sdg::batch::batchTraceKernel<double>, grid, dim3(sdg::batch::kThreads), 0, b->stream, a, v, V.nb, V.naq);
    }
  }
}

// mixed-precision reference state of problem p from its coefficients (reference layout)
void updateMix(sdg_batch* b, int p, const double* coeff) {
  const sdg::MixedMesh& m = *b->meshes[p];
  double u[4] = {0, 0, 0, 0};
  for (int64_t e = 0; e < m.elements; e++) {
    const auto& V = b->var[b->emap[p][e].first];
    const double* c = coeff + b->coeff_off[p][e];
    for (int v = 0; v < 4; v++)
      for (int k = 0; k < V.nb; k++) u[v] += c[k * 4 + v] * V.mean[k];
  }
  for (int v = 0; v < 4; v++) u[v] /= static_cast<double>(m.elements);
  double* o = &b->mixh[static_cast<size_t>(p) * 6];
  for (int v = 0; v < 4; v++) o[v] = u[v];
  const double rho = u[0], vx = u[1] / rho, vy = u[2] / rho;
  const double e = u[3] / rho - (vx * vx + vy * vy) / 2.0;
  const sdg::Phys& ph = b->phys;
  o[4] = ph.eos == SDG_EOS_IDEAL_GAS ? (sdg::kGamma - 1.0) * rho * e : ph.c0 * ph.c0 * (rho - ph.rho0) + ph.padd;
  o[5] = (u[3] + o[4]) / rho;
}

}  // namespace

extern "C" {

const char* sdg_batch_last_error(const sdg_batch* b) { return b ? b->err.c_str() : sdg::globalError(); }

int sdg_batch_create(const sdg_problem_desc* problem, const sdg_mesh* const* meshes, int count, sdg_batch** out) {
  if (!problem || !meshes || !out || count < 1) return bfail(nullptr, SDG_E_INVALID, "sdg_batch_create: bad argument");
  *out = nullptr;
  if (problem->abi_version != SDG_ABI_VERSION) return bfail(nullptr, SDG_E_INVALID, "sdg_batch_create: abi_version mismatch");
  if (problem->source_term != SDG_SOURCE_NONE)
    return bfail(nullptr, SDG_E_UNSUPPORTED, "sdg_batch_create: source terms are not supported by the batched solver");
  if (problem->shock_capturing != SDG_SHOCK_NONE)
    return bfail(nullptr, SDG_E_UNSUPPORTED, "sdg_batch_create: shock capturing is not supported by the batched solver");
  if (problem->equation != SDG_EQ_COMPRESSIBLE_EULER)
    return bfail(nullptr, SDG_E_UNSUPPORTED, "sdg_batch_create: the batched solver runs the compressible Euler model");
  if (problem->convective_flux < SDG_FLUX_CENTRAL || problem->convective_flux > SDG_FLUX_EXACT ||
      problem->time_integration < SDG_RK_FORWARD_EULER || problem->time_integration > SDG_RK_SSPRK3 ||
      !(problem->cv > 0.0) || !(problem->cp > 0.0))
    return bfail(nullptr, SDG_E_INVALID, "sdg_batch_create: bad problem descriptor");
  sdg_batch* b = new (std::nothrow) sdg_batch();
  if (!b) return bfail(nullptr, SDG_E_NOMEM, "sdg_batch_create: out of host memory");
  auto bail = [&](int code) {
    std::string m = b->err;
    sdg_batch_destroy(b);
    sdg::setGlobalError(m);
    return code;
  };
  b->prob = *problem;
  b->mixed = problem->precision == SDG_PREC_MIXED;
  b->nprob = count;
  sdg::Phys& ph = b->phys;
  ph.eos = problem->equation_of_state;
  ph.transport = SDG_TRANSPORT_NONE;
  ph.flux = problem->convective_flux;
  ph.cp = problem->cp;
  ph.cv = problem->cv;
  ph.c0 = problem->reference_sound_speed;
  ph.rho0 = problem->reference_density;
  ph.padd = 0.01 * problem->reference_density * problem->reference_sound_speed * problem->reference_sound_speed;
  for (int p = 0; p < count; p++) {
    const sdg::MixedMesh* m = sdg::mixedOf(meshes[p]);
    if (!m) {
      b->err = "sdg_batch_create: meshes must come from sdg_mesh_create_split";
      return bail(SDG_E_UNSUPPORTED);
    }
    if (m->order < 1 || m->order > 4) {
      b->err = "sdg_batch_create: orders 1..4 are supported";
      return bail(SDG_E_UNSUPPORTED);
    }
    b->meshes.push_back(m);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
    b->err = "sdg_batch_create: no HIP device / stream";
    return bail(SDG_E_DEVICE);
  }
  for (int v = 0; v < sdg::batch::kVariants; v++)
    if (hipStreamCreateWithFlags(&b->aux[v], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join_ev[v], hipEventDisableTiming) != hipSuccess) {
      b->err = "sdg_batch_create: stream creation failed";
      return bail(SDG_E_DEVICE);
    }
  // variants present and their element ranges (problem-major within a variant)
  b->emap.resize(count);
  b->coeff_off.resize(count);
  for (int p = 0; p < count; p++) {
    const sdg::MixedMesh& m = *b->meshes[p];
    b->emap[p].resize(m.elements);
    b->coeff_off[p].resize(m.elements);
    int64_t off = 0;
    for (int64_t e = 0; e < m.elements; e++) {
      const sdg::RefElement& r = m.refOf(e);
      const int v = variantCode(r.type, r.order);
      auto& V = b->var[v];
      if (!V.used) {
        V.used = true;
        V.type = r.type;
        V.order = r.order;
        V.ref = r;
        V.nb = r.nb;
        V.nq = r.nq;
        V.nf = r.nf;
        V.nfq = r.nfq;
        V.naq = r.nf * r.nfq;
      }
      b->emap[p][e] = {v, V.E++};
      b->coeff_off[p][e] = off;
      off += static_cast<int64_t>(r.nb) * 4;
    }
  }
  // re-index variant elements problem-major: index = (running count at creation) already is
  int64_t eb = 0, ub = 0, tb = 0, fb = 0;
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    auto& V = b->var[v];
    if (!V.used) continue;
    V.e_base = eb;
    V.u_base = ub;
    V.t_base = tb;
    V.f_base = fb;
    eb += V.E;
    ub += V.E * 4 * V.nb;
    tb += V.E * 4 * V.naq;
    fb += V.E * V.nf;
    // constant-function coefficients (psi basis) and the reference-element means of the phi basis
    // functions (the mixed mode's reference state is formed from phi coefficients on the host)
    const sdg::RefElement& r = V.ref;
    std::vector<double> colsum(r.nb, 0.0);
    double wsum = 0.0;
    V.mean.assign(r.nb, 0.0);
    for (int q = 0; q < r.nq; q++) {
      wsum += r.qweight[q];
      for (int k = 0; k < r.nb; k++) {
        colsum[k] += r.V[static_cast<size_t>(q) * r.nb + k];
        V.mean[k] += r.qweight[q] * r.V[static_cast<size_t>(q) * r.nb + k];
      }
    }
    for (int k = 0; k < r.nb; k++) V.mean[k] /= wsum;
    std::vector<double> one_phi(r.nb, 0.0);
    for (int k = 0; k < r.nb; k++) {
      double x = 0.0;
      for (int c = 0; c < r.nb; c++) x += colsum[c] * r.lsq_inverse[static_cast<size_t>(c) * r.nb + k];
      if (std::fabs(x) < 1e-12) x = 0.0;
      if (std::fabs(x - 1.0) < 1e-12) x = 1.0;
      one_phi[k] = x;
    }
    b->chol[v] = r.mass_chol;
    b->chol_inv[v] = r.mass_chol_inverse;
    V.one.assign(r.nb, 0.0);  // c_psi = L^T c_phi
    for (int j2 = 0; j2 < r.nb; j2++) {
      double x = 0.0;
      for (int i2 = 0; i2 < r.nb; i2++) x += r.mass_chol[static_cast<size_t>(i2) * r.nb + j2] * one_phi[i2];
      V.one[j2] = x;
    }
  }
  b->totalE = eb;
  b->totalU = ub;
  b->totalT = tb;
  b->totalF = fb;
  // per element: geometry and problem; per element face: neighbour trace base, face, flags;
  // per face (all problems' faces in one list): normal and |J_f| w at the left parent's points
  std::vector<int64_t> face_off(count + 1, 0);
  for (int p = 0; p < count; p++) face_off[p + 1] = face_off[p] + static_cast<int64_t>(b->meshes[p]->faces.size());
  if (face_off[count] > INT_MAX) {
    b->err = "sdg_batch_create: too many faces";
    return bail(SDG_E_UNSUPPORTED);
  }
  std::vector<double> ginv(static_cast<size_t>(eb) * 4), detj(eb), me(eb);
  std::vector<double> fnrm(static_cast<size_t>(face_off[count]) * 2), fjw(static_cast<size_t>(face_off[count]) * sdg::batch::kMaxNfq, 0.0);
  std::vector<int> pid(eb), fid(fb);
  std::vector<int64_t> fnb(fb);
  std::vector<int8_t> fflag(fb);
  for (int p = 0; p < count; p++) {
    const sdg::MixedMesh& m = *b->meshes[p];
    for (size_t f = 0; f < m.faces.size(); f++) {
      const sdg::Face& fc = m.faces[f];
      const int64_t gf = face_off[p] + static_cast<int64_t>(f);
      const int nfq = static_cast<int>(fc.jw.size());
      if (nfq > sdg::batch::kMaxNfq) {
        b->err = "sdg_batch_create: face rule larger than the batch supports";
        return bail(SDG_E_UNSUPPORTED);
      }
      for (int d = 0; d < 2; d++) {
        fnrm[static_cast<size_t>(gf) * 2 + d] = fc.normal[d];
        for (int j2 = 1; j2 < nfq; j2++)  // straight (affine) faces: one normal per face
          if (std::fabs(fc.normal[static_cast<size_t>(j2) * 2 + d] - fc.normal[d]) > 1e-13) {
            b->err = "sdg_batch_create: the batched solver needs straight faces";
            return bail(SDG_E_UNSUPPORTED);
          }
      }
      for (int j2 = 0; j2 < nfq; j2++) fjw[static_cast<size_t>(gf) * sdg::batch::kMaxNfq + j2] = fc.jw[j2];
    }
    for (int64_t e = 0; e < m.elements; e++) {
      const auto [v, k] = b->emap[p][e];
      const auto& V = b->var[v];
      const int64_t ge = V.e_base + k;
      const double* J = &m.jac[static_cast<size_t>(e) * 4];
      const double jt[4] = {J[0], J[2], J[1], J[3]};
      double jti[4];
      sdg::invertMatrix(2, jt, jti);
      for (int aa = 0; aa < 2; aa++)
        for (int bb = 0; bb < 2; bb++) ginv[static_cast<size_t>(ge) * 4 + aa + 2 * bb] = jti[aa * 2 + bb];
      detj[ge] = m.detj[e];
      me[ge] = m.min_edge[e];
      pid[ge] = p;
      for (int s = 0; s < V.nf; s++) {
        const int f = m.elem_face[static_cast<size_t>(e) * 4 + s];
        const int side = m.elem_side[static_cast<size_t>(e) * 4 + s];
        const sdg::Face& fc = m.faces[f];
        const int64_t r = fc.parent[side ^ 1];
        const int rs = fc.local[side ^ 1];
        const auto [rv, rk] = b->emap[p][r];
        const auto& RV = b->var[rv];
        // the neighbour's point of own point j: the left parent's point j has the right's seq[j]
        bool ident = true, rev = true;
        for (int j2 = 0; j2 < V.nfq; j2++) {
          int kpt = fc.seq[j2];
          if (side == 1)
            for (int q = 0; q < V.nfq; q++)
              if (fc.seq[q] == j2) kpt = q;
          ident = ident && kpt == j2;
          rev = rev && kpt == V.nfq - 1 - j2;
        }
        if (!ident && !rev) {
          b->err = "sdg_batch_create: face point order neither identical nor reversed";
          return bail(SDG_E_UNSUPPORTED);
        }
        const int64_t ef = V.f_base + k * V.nf + s;
        fnb[ef] = RV.t_base + (rk * RV.naq + static_cast<int64_t>(rs) * RV.nfq) * 4;
        fid[ef] = static_cast<int>(face_off[p] + f);
        fflag[ef] = static_cast<int8_t>(side | ((ident ? 0 : 1) << 1));
      }
    }
  }
  std::vector<VariantDev> dv(sdg::batch::kVariants);
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    auto& V = b->var[v];
    VariantDev& d = dv[v];
    d = VariantDev{};
    if (!V.used) continue;
    const sdg::RefElement& r = V.ref;
    const int nb = r.nb;
    // psi_j = sum_i (L^-1)_{j i} phi_i at the volume and face points, and the reference gradients
    auto li = [&](int j2, int i2) { return r.mass_chol_inverse[static_cast<size_t>(j2) * nb + i2]; };
    auto psi = [&](const std::vector<double>& phi, int rows) {  // rows x nb (phi) -> rows x nb (psi)
      std::vector<double> out(static_cast<size_t>(rows) * nb, 0.0);
      for (int k = 0; k < rows; k++)
        for (int j2 = 0; j2 < nb; j2++) {
          double acc = 0.0;
          for (int i2 = 0; i2 <= j2; i2++) acc += phi[static_cast<size_t>(k) * nb + i2] * li(j2, i2);
          out[static_cast<size_t>(k) * nb + j2] = acc;
        }
      return out;
    };
    const std::vector<double> Vq = psi(r.V, r.nq), dVq = psi(r.dV, r.nq * 2), Vfq = psi(r.Vf, V.naq);
    std::vector<double> Vt(static_cast<size_t>(nb) * r.nq), Vft(static_cast<size_t>(nb) * V.naq);
    for (int q = 0; q < r.nq; q++)
      for (int k = 0; k < nb; k++) Vt[static_cast<size_t>(k) * r.nq + q] = Vq[static_cast<size_t>(q) * nb + k];
    for (int sl = 0; sl < V.naq; sl++)
      for (int k = 0; k < nb; k++) Vft[static_cast<size_t>(k) * V.naq + sl] = Vfq[static_cast<size_t>(sl) * nb + k];
    double *pVt, *pdV, *pVft, *pVf, *pone, *pw;
    if (bupload(b, Vt, &pVt) || bupload(b, dVq, &pdV) || bupload(b, Vft, &pVft) || bupload(b, Vfq, &pVf) ||
        bupload(b, V.one, &pone) || bupload(b, r.qweight, &pw))
      return bail(SDG_E_NOMEM);
    d = VariantDev{pVt, pdV, pVft, pVf, pone, pw, V.E, V.u_base, V.t_base, V.f_base, V.e_base};
  }
  const size_t tbytes = b->mixed ? sizeof(float) : sizeof(double);
  void *t0 = nullptr, *t1 = nullptr;
  if (bupload(b, dv, &b->dvar) || bupload(b, ginv, &b->ginv) || bupload(b, detj, &b->detj) || bupload(b, me, &b->min_edge) ||
      bupload(b, pid, &b->pid) || bupload(b, fnb, &b->fnb) || bupload(b, fid, &b->fid) || bupload(b, fflag, &b->fflag) ||
      bupload(b, fnrm, &b->fnrm) || bupload(b, fjw, &b->fjw) || balloc(b, static_cast<size_t>(ub), &b->U) ||
      balloc(b, static_cast<size_t>(ub), &b->Un) || balloc(b, static_cast<size_t>(count), &b->dt) ||
      balloc(b, static_cast<size_t>(count) * 6, &b->mix) || balloc(b, static_cast<size_t>(eb), &b->dtbuf) ||
      balloc(b, static_cast<size_t>(tb) * tbytes, reinterpret_cast<char**>(&t0)) ||
      balloc(b, static_cast<size_t>(tb) * tbytes, reinterpret_cast<char**>(&t1)))
    return bail(SDG_E_NOMEM);
  b->T[0] = t0;
  b->T[1] = t1;
  if (hipMemset(b->dt, 0, sizeof(double) * count) != hipSuccess || hipMemset(b->mix, 0, sizeof(double) * 6 * count) != hipSuccess) {
    b->err = "sdg_batch_create: device initialisation failed";
    return bail(SDG_E_DEVICE);
  }
  b->mixh.assign(static_cast<size_t>(count) * 6, 0.0);
  b->host_dt.assign(count, 0.0);
  b->has_state.assign(count, 0);
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    b->chunk[v] = 4;
    b->vgrid[v] = cus * 4;
  }
  if (buildItems(b)) return bail(SDG_E_DEVICE);
  *out = b;
  return SDG_OK;
}

void sdg_batch_destroy(sdg_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (int i = 0; i < 2; i++)
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    if (b->aux[v]) (void)hipStreamDestroy(b->aux[v]);
    if (b->join_ev[v]) (void)hipEventDestroy(b->join_ev[v]);
  }
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int sdg_batch_set_state(sdg_batch* b, int p, const double* coeff) {
  if (!b || !coeff || p < 0 || p >= b->nprob) return bfail(b, SDG_E_INVALID, "sdg_batch_set_state: bad argument");
  const sdg::MixedMesh& m = *b->meshes[p];
  updateMix(b, p, coeff);
  // scatter into the device layout of each variant: the problem's elements of one variant are
  // contiguous there, so one copy per variant
  int64_t lo[sdg::batch::kVariants], hi[sdg::batch::kVariants];
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    lo[v] = INT64_MAX;
    hi[v] = -1;
  }
  for (int64_t e = 0; e < m.elements; e++) {
    const auto [v, k] = b->emap[p][e];
    lo[v] = std::min(lo[v], k);
    hi[v] = std::max(hi[v], k);
  }
  std::vector<std::vector<double>> host(sdg::batch::kVariants);
  for (int v = 0; v < sdg::batch::kVariants; v++)
    if (hi[v] >= 0) host[v].assign(static_cast<size_t>(hi[v] - lo[v] + 1) * 4 * b->var[v].nb, 0.0);
  for (int64_t e = 0; e < m.elements; e++) {
    const auto [v, k] = b->emap[p][e];
    const auto& V = b->var[v];
    double* panel = &host[v][static_cast<size_t>(k - lo[v]) * 4 * V.nb];
    const double* c = coeff + b->coeff_off[p][e];
    const std::vector<double>& Lc = b->chol[v];  // c_psi[j] = sum_{i >= j} L[i][j] c_phi[i]
    for (int vv = 0; vv < 4; vv++)
      for (int j = 0; j < V.nb; j++) {
        double acc = 0.0;
        for (int i = j; i < V.nb; i++) acc += Lc[static_cast<size_t>(i) * V.nb + j] * c[i * 4 + vv];
        panel[static_cast<size_t>(vv) * V.nb + j] = acc;
      }
  }
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    if (hi[v] < 0) continue;
    const size_t off = static_cast<size_t>(b->var[v].u_base + lo[v] * 4 * b->var[v].nb);
    BHIP(b, hipMemcpyAsync(b->U + off, host[v].data(), host[v].size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->Un + off, host[v].data(), host[v].size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  }
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipMemcpyAsync(b->mix + static_cast<size_t>(p) * 6, &b->mixh[static_cast<size_t>(p) * 6], 6 * sizeof(double),
                         hipMemcpyHostToDevice, b->stream));
  b->has_state[p] = 1;
  b->cur = 0;
  launchTraces(b, 0);
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < 2; i++) {
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
    b->graph[i] = nullptr;
  }
  return SDG_OK;
}

int sdg_batch_get_state(sdg_batch* b, int p, double* coeff) {
  if (!b || !coeff || p < 0 || p >= b->nprob) return bfail(b, SDG_E_INVALID, "sdg_batch_get_state: bad argument");
  const sdg::MixedMesh& m = *b->meshes[p];
  std::vector<double> host(static_cast<size_t>(b->totalU));
  BHIP(b, hipMemcpyAsync(host.data(), b->U, host.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int64_t e = 0; e < m.elements; e++) {
    const auto [v, k] = b->emap[p][e];
    const auto& V = b->var[v];
    const double* panel = &host[static_cast<size_t>(V.u_base + k * 4 * V.nb)];
    double* c = coeff + b->coeff_off[p][e];
    const std::vector<double>& Li = b->chol_inv[v];  // c_phi[i] = sum_{j >= i} (L^-1)[j][i] c_psi[j]
    for (int vv = 0; vv < 4; vv++)
      for (int i = 0; i < V.nb; i++) {
        double acc = 0.0;
        for (int j = i; j < V.nb; j++) acc += Li[static_cast<size_t>(j) * V.nb + i] * panel[static_cast<size_t>(vv) * V.nb + j];
        c[i * 4 + vv] = acc;
      }
  }
  return SDG_OK;
}

int sdg_batch_set_initial_primitive(sdg_batch* b, int p, const double* prim) {
  if (!b || !prim || p < 0 || p >= b->nprob) return bfail(b, SDG_E_INVALID, "sdg_batch_set_initial_primitive: bad argument");
  const sdg::MixedMesh& m = *b->meshes[p];
  std::vector<double> coeff(b->coeff_off[p].empty() ? 0 : static_cast<size_t>(b->coeff_off[p].back()) +
                                                              4 * static_cast<size_t>(m.refOf(m.elements - 1).nb));
  int64_t qoff = 0;
  for (int64_t e = 0; e < m.elements; e++) {  // InitialCondition.cpp:85-104
    const sdg::RefElement& r = m.refOf(e);
    std::vector<double> uq(static_cast<size_t>(r.nq) * 4), t(static_cast<size_t>(4) * r.nb);
    for (int q = 0; q < r.nq; q++) {
      const double* pr = prim + (qoff + q) * 4;
      const double rho = pr[0];
      uq[q * 4] = rho;
      uq[q * 4 + 1] = rho * pr[1];
      uq[q * 4 + 2] = rho * pr[2];
      uq[q * 4 + 3] = rho * (b->phys.cv * pr[3] + (pr[1] * pr[1] + pr[2] * pr[2]) / 2.0);
    }
    qoff += r.nq;
    for (int v = 0; v < 4; v++)
      for (int k = 0; k < r.nb; k++) {
        double s = 0.0;
        for (int q = 0; q < r.nq; q++) s += uq[q * 4 + v] * r.V[static_cast<size_t>(q) * r.nb + k];
        t[static_cast<size_t>(v) * r.nb + k] = s;
      }
    double* c = &coeff[static_cast<size_t>(b->coeff_off[p][e])];
    for (int v = 0; v < 4; v++)
      for (int k = 0; k < r.nb; k++) {
        double s = 0.0;
        for (int cc = 0; cc < r.nb; cc++) s += t[static_cast<size_t>(v) * r.nb + cc] * r.lsq_inverse[static_cast<size_t>(cc) * r.nb + k];
        c[k * 4 + v] = s;
      }
  }
  return sdg_batch_set_state(b, p, coeff.data());
}

int sdg_batch_compute_dt(sdg_batch* b, double cfl, double* dt) {
  if (!b || !dt) return bfail(b, SDG_E_INVALID, "sdg_batch_compute_dt: bad argument");
  for (int p = 0; p < b->nprob; p++)
    if (!b->has_state[p]) return bfail(b, SDG_E_STATE, "sdg_batch_compute_dt: a problem has no state");
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    const auto& V = b->var[v];
    if (!V.used) continue;
    hipLaunchKernelGGL(sdg::batch::batchDtKernel, dim3(static_cast<unsigned>((V.E + sdg::batch::kThreads - 1) / sdg::batch::kThreads)),
                       dim3(sdg::batch::kThreads), 0, b->stream, b->U, b->dvar, v, V.nb, V.nq, V.order, b->phys,
                       b->min_edge, cfl, b->dtbuf);
  }
  BHIP(b, hipGetLastError());
  std::vector<double> h(static_cast<size_t>(b->totalE));
  std::vector<int> pid(static_cast<size_t>(b->totalE));
  BHIP(b, hipMemcpyAsync(h.data(), b->dtbuf, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(pid.data(), b->pid, pid.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int p = 0; p < b->nprob; p++) dt[p] = 1.0e308;
  for (int64_t e = 0; e < b->totalE; e++) dt[pid[e]] = std::min(dt[pid[e]], h[e]);
  return SDG_OK;
}

int sdg_batch_step(sdg_batch* b, const double* dt, int nsteps) {
  if (!b || !dt || nsteps < 0) return bfail(b, SDG_E_INVALID, "sdg_batch_step: bad argument");
  for (int p = 0; p < b->nprob; p++)
    if (!b->has_state[p]) return bfail(b, SDG_E_STATE, "sdg_batch_step: a problem has no state");
  if (std::memcmp(b->host_dt.data(), dt, sizeof(double) * b->nprob) != 0) {
    std::memcpy(b->host_dt.data(), dt, sizeof(double) * b->nprob);
    BHIP(b, hipMemcpyAsync(b->dt, b->host_dt.data(), sizeof(double) * b->nprob, hipMemcpyHostToDevice, b->stream));
  }
  const int S = stagesOf(b);
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  for (int i = 0; i < nsteps; i++) {
    const int c = b->cur;
    if (!b->graph[c]) {
      hipGraph_t g;
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < S; s++) launchStage(b, s, (c + s) & 1);
      BHIP(b, hipStreamEndCapture(b->stream, &g));
      const hipError_t e = hipGraphInstantiate(&b->graph[c], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return bfail(b, SDG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    BHIP(b, hipGraphLaunch(b->graph[c], b->stream));
    b->cur = (c + S) & 1;
  }
  BHIP(b, hipGetLastError());
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  return SDG_OK;
}

int sdg_batch_last_step_ms(const sdg_batch* b, double* ms) {
  if (!b || !ms) return SDG_E_INVALID;
  *ms = b->last_ms;
  return SDG_OK;
}

// Autotuner: for every variant, the work-item size (tiles per item) that makes a stage over that
// variant's elements fastest, then the persistent grid size; the state is saved and restored.
int sdg_batch_tune(sdg_batch* b, double* stage_ms) {
  if (!b) return bfail(b, SDG_E_INVALID, "sdg_batch_tune: null batch");
  for (int p = 0; p < b->nprob; p++)
    if (!b->has_state[p]) return bfail(b, SDG_E_STATE, "sdg_batch_tune: a problem has no state");
  const size_t tbytes = (b->mixed ? sizeof(float) : sizeof(double)) * static_cast<size_t>(b->totalT);
  std::vector<char> saveU(static_cast<size_t>(b->totalU) * sizeof(double)), saveUn(saveU.size()), saveT(tbytes);
  BHIP(b, hipMemcpy(saveU.data(), b->U, saveU.size(), hipMemcpyDeviceToHost));
  BHIP(b, hipMemcpy(saveUn.data(), b->Un, saveUn.size(), hipMemcpyDeviceToHost));
  BHIP(b, hipMemcpy(saveT.data(), b->T[b->cur], tbytes, hipMemcpyDeviceToHost));
  auto timeStage = [&](int reps, int only) -> float {
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
      (void)hipEventRecord(b->ev0, b->stream);
      launchStage(b, 0, b->cur, only);
      (void)hipEventRecord(b->ev1, b->stream);
      (void)hipStreamSynchronize(b->stream);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
      best = std::min(best, ms);
    }
    return best;
  };
  // per variant, alone: the work-item size in tiles and its kernel's persistent grid
  static const int kChunks[] = {1, 2, 4, 8, 16};
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    if (!b->var[v].used) continue;
    float best = 1e30f;
    int bestc = b->chunk[v], bestg = b->vgrid[v];
    for (int c : kChunks) {
      b->chunk[v] = c;
      if (buildItems(b, &v)) return SDG_E_DEVICE;
      for (int mult : {1, 2, 4, 8}) {
        b->vgrid[v] = cus * mult;
        const float t = timeStage(3, v);
        if (t < best) {
          best = t;
          bestc = c;
          bestg = b->vgrid[v];
        }
      }
    }
    b->chunk[v] = bestc;
    b->vgrid[v] = bestg;
  }
  if (buildItems(b)) return SDG_E_DEVICE;
  const float best = timeStage(3, -1);
  BHIP(b, hipGetLastError());
  BHIP(b, hipMemcpy(b->U, saveU.data(), saveU.size(), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->Un, saveUn.data(), saveUn.size(), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->T[b->cur], saveT.data(), tbytes, hipMemcpyHostToDevice));
  for (int i = 0; i < 2; i++) {
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
    b->graph[i] = nullptr;
  }
  if (stage_ms) *stage_ms = best;
  return SDG_OK;
}

int sdg_batch_get_config(const sdg_batch* b, int* chunk_tiles, int* epb, int* grid) {
  if (!b) return SDG_E_INVALID;
  for (int v = 0; v < sdg::batch::kVariants; v++) {
    if (chunk_tiles) chunk_tiles[v] = b->var[v].used ? b->chunk[v] : 0;
    if (epb) epb[v] = b->mixed ? sdg::batch::epbOf<float>(v) : sdg::batch::epbOf<double>(v);
  }
  if (grid) {
    *grid = 0;
    for (int v = 0; v < sdg::batch::kVariants; v++) *grid += b->var[v].used ? b->vgrid[v] : 0;
  }
  return SDG_OK;
}

}  // extern "C"
