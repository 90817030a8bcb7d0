// sdg_solver.hip -- solver context and the device part of the C-ABI (include/sdg.h).
//
// One context owns one HIP stream, the operand tables and the state in HBM. A time step is the
// reference's stepSolver (TimeIntegration.cpp:327-350): per RK stage one kernel for Euler, two for
// Navier-Stokes, then a fixed-order monitor reduction; whole steps are captured in hipGraphs and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "dg_adjoint.hpp"
#include "dg_tensor_adjoint.hpp"
#include "dg_tensor.hpp"
#include "dg_kernels.hpp"
#include "dg_mfma.hpp"
#include "kit.hpp"
#include "sdg_internal.hpp"

using sdg::Geo;
using sdg::Phys;
using sdg::StageArgs;

namespace {

using StageFn = void (*)(hipStream_t, const Geo&, const Phys&, const StageArgs&);
using TraceFn = void (*)(hipStream_t, const Geo&, const double*, void*, const sdg::MixRef&);
using DtFn = void (*)(hipStream_t, const Geo&, const Phys&, const double*, const double*, double, int, double*);
using MonFn = void (*)(hipStream_t, const double*, int, double*, int*);
using SrcFn = void (*)(hipStream_t, const Geo&, const Phys&, const double*, double*, const double*);
using AdjFn = void (*)(hipStream_t, const Geo&, const Phys&, const StageArgs&, const sdg::AdjArgs&);
using AvFn = void (*)(hipStream_t, const Geo&, const Phys&, const double*, double*);
using BlockFn = void (*)(hipStream_t, const Geo&, const Phys&, const double*, double*);

struct Launchers {
  bool mfma = false;
  bool tensor = false;  // hexahedra, Navier-Stokes: sum-factorised stage kernels
  StageFn euler = nullptr;
  StageFn ns_grad_br1 = nullptr;
  StageFn ns_grad_br2 = nullptr;
  StageFn ns_flux = nullptr;
  StageFn ns_face = nullptr;  // matrix-core path: face fluxes before ns_flux
  TraceFn trace = nullptr;
  TraceFn trace64 = nullptr;  // FP64 traces of the FP64 state (the adjoint of mixed-precision contexts)
  SrcFn source = nullptr;  // modal Boussinesq source (all kernel sets but the sum-factorised one)
  DtFn dt = nullptr;
  MonFn monitor = nullptr;
  // Euler with artificial viscosity (vector kernels): gradient pass + volume flux, face fluxes +
  // residual + update, and the per-step sensor
  StageFn av_grad = nullptr;
  StageFn av_flux = nullptr;
  StageFn av_ns_grad[2] = {nullptr, nullptr};  // Navier-Stokes with artificial viscosity: BR1, BR2
  StageFn av_ns_flux = nullptr;
  AvFn av_sensor = nullptr;
  // Navier-Stokes BR2 raw files: per boundary face the parent's volume gradient plus that face's lift
  BlockFn br2_blocks = nullptr;
  // discrete adjoint (Navier-Stokes, FP64: the dense matrix-core set, or the sum-factorised set on hexahedra)
  AdjFn adj_volume[2] = {nullptr, nullptr};  // BR1, BR2
  AdjFn adj_face = nullptr;
  AdjFn adj_lift[2] = {nullptr, nullptr};
  AdjFn adj_gather = nullptr;
};

inline int blocksFor(int E, int epb) { return (E + epb - 1) / epb; }

// VALU kernels: any size; MFMA kernels: bases that fill 16x16 tiles. TC is the contraction type of
// the matrix-core kernels: double (FP64 mode) or float (mixed-precision mode).
template <int D, int NB, int NQ, int NF, int NFQ, bool MFMA, typename TC = double>
Launchers makeLaunchers() {
  using S = sdg::Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int EPB = S::EPB;
  constexpr int NV = D + 2;
  Launchers l;
  l.mfma = MFMA;
  if constexpr (MFMA) {
    constexpr int EPB_S = 16 / NV, EPB_G = 16 / (NV * D);
    l.euler = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::eulerStageMfma<TC, D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_grad_br1 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradMfma<TC, D, NB, NQ, NF, NFQ, 1>), dim3(blocksFor(g.E, EPB_G)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    l.ns_grad_br2 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradMfma<TC, D, NB, NQ, NF, NFQ, 2>), dim3(blocksFor(g.E, EPB_G)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    l.ns_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsFaceFluxKernel<TC, D, NF, NFQ>), dim3(blocksFor((g.F + g.Fb) * NFQ, sdg::kThreads)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsResidualMfma<TC, D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    {  // Euler with artificial viscosity: volume gradient, AV fluxes (both precision modes)
      l.av_grad = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
        hipLaunchKernelGGL((sdg::nsGradMfma<TC, D, NB, NQ, NF, NFQ, 0>), dim3(blocksFor(g.E, EPB_G)),
                           dim3(sdg::kThreads), 0, st, g, ph, a);
      };
      l.av_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
        hipLaunchKernelGGL((sdg::nsFaceFluxKernel<TC, D, NF, NFQ, false, true>),
                           dim3(blocksFor((g.F + g.Fb) * NFQ, sdg::kThreads)), dim3(sdg::kThreads), 0, st, g, ph, a);
        hipLaunchKernelGGL((sdg::nsResidualMfma<TC, D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)),
                           dim3(sdg::kThreads), 0, st, g, ph, a);
      };
      l.av_sensor = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, double* node) {
        hipLaunchKernelGGL((sdg::avSensorKernel<D, NB, NQ>), dim3(blocksFor(g.E, sdg::kThreads)), dim3(sdg::kThreads),
                           0, st, g, ph, U, node);
      };
    }
    if constexpr (sizeof(TC) == 8) {
      l.adj_volume[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        hipLaunchKernelGGL((sdg::nsAdjVolumeMfma<D, NB, NQ, NF, NFQ, 1>), dim3(blocksFor(g.E, EPB_G)),
                           dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
      l.adj_volume[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        hipLaunchKernelGGL((sdg::nsAdjVolumeMfma<D, NB, NQ, NF, NFQ, 2>), dim3(blocksFor(g.E, EPB_G)),
                           dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
      l.adj_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        if (ph.flux <= SDG_FLUX_ROE)  // reverse mode, one thread per face point
          hipLaunchKernelGGL((sdg::nsAdjFaceRevKernel<D, NF, NFQ>), dim3(blocksFor(g.F * NFQ, sdg::kThreads)),
                             dim3(sdg::kThreads), 0, st, g, ph, a, j);
        else  // dual numbers, one thread per (state component or gradient, side, face point)
          hipLaunchKernelGGL((sdg::nsAdjFaceKernel<D, NF, NFQ>),
                             dim3(blocksFor(g.F * NFQ * 2 * (D + 3), sdg::kThreads)),
                             dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
      l.adj_lift[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        hipLaunchKernelGGL((sdg::nsAdjLiftMfma<D, NB, NQ, NF, NFQ, 1>), dim3(blocksFor(g.E, EPB_G)),
                           dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
      l.adj_lift[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        hipLaunchKernelGGL((sdg::nsAdjLiftMfma<D, NB, NQ, NF, NFQ, 2>), dim3(blocksFor(g.E, EPB_G)),
                           dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
      l.adj_gather = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
        hipLaunchKernelGGL((sdg::nsAdjGatherMfma<D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)),
                           dim3(sdg::kThreads), 0, st, g, ph, a, j);
      };
    }
  } else {
    l.euler = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::eulerStageKernel<D, NB, NQ, NF, NFQ, EPB>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_grad_br1 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradKernel<D, NB, NQ, NF, NFQ, EPB, 1>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_grad_br2 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradKernel<D, NB, NQ, NF, NFQ, EPB, 2>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsFluxKernel<D, NB, NQ, NF, NFQ, EPB>), dim3(blocksFor(g.E, EPB)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    l.av_grad = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradKernel<D, NB, NQ, NF, NFQ, EPB, 0>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.av_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsFluxKernel<D, NB, NQ, NF, NFQ, EPB, 1>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.av_ns_grad[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradKernel<D, NB, NQ, NF, NFQ, EPB, 1, true>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.av_ns_grad[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradKernel<D, NB, NQ, NF, NFQ, EPB, 2, true>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.av_ns_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsFluxKernel<D, NB, NQ, NF, NFQ, EPB, 2>), dim3(blocksFor(g.E, EPB)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.av_sensor = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, double* node) {
      hipLaunchKernelGGL((sdg::avSensorKernel<D, NB, NQ>), dim3(blocksFor(g.E, sdg::kThreads)), dim3(sdg::kThreads), 0,
                         st, g, ph, U, node);
    };
  }
  if constexpr (MFMA) {
    // the stage tail's own contraction (bit-identical traces after a host-side state change)
    l.trace = [](hipStream_t st, const Geo& g, const double* U, void* Uf, const sdg::MixRef& mix) {
      hipLaunchKernelGGL((sdg::mfmaTraceKernel<TC, D, NB, NF, NFQ>), dim3(blocksFor(g.E, 16 / (D + 2))),
                         dim3(sdg::kThreads), 0, st, g, U, Uf, mix);
    };
  } else {
    l.trace = [](hipStream_t st, const Geo& g, const double* U, void* Uf, const sdg::MixRef& mix) {
      const int64_t n = static_cast<int64_t>(g.E) * (D + 2) * NF * NFQ;
      hipLaunchKernelGGL((sdg::traceKernel<D, NB, NQ, NF, NFQ>), dim3((n + sdg::kThreads - 1) / sdg::kThreads),
                         dim3(sdg::kThreads), 0, st, g, U, Uf, mix);
    };
  }
  l.source = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, double* out, const double* Li) {
    constexpr int NM = NQ > NB ? NQ : NB, EPBS = sdg::kThreads / NM > 0 ? sdg::kThreads / NM : 1;
    hipLaunchKernelGGL((sdg::sourceKernel<D, NB, NQ>), dim3(blocksFor(g.E, EPBS)), dim3(sdg::kThreads), 0, st, g, ph,
                       U, out, Li);
  };
  l.br2_blocks = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, double* out) {
    hipLaunchKernelGGL((sdg::br2BoundaryBlockKernel<D, NB, NQ, NF, NFQ>), dim3(g.Fb), dim3(sdg::kThreads), 0, st, g, ph,
                       U, out);
  };
  l.dt = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, const double* me, double cfl, int order,
            double* out) {
    hipLaunchKernelGGL((sdg::dtKernel<D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, sdg::DtShape<D, NB, NQ>::EPB)),
                       dim3(sdg::kThreads), 0, st, g, ph, U, me, cfl, order, out);
  };
  l.monitor = [](hipStream_t st, const double* mon, int E, double* rel, int* status) {
    // mon holds E x NV per-element values followed by kMonitorBlocks x NV partial sums
    double* part = const_cast<double*>(mon) + static_cast<int64_t>(E) * (D + 2);
    hipLaunchKernelGGL((sdg::monitorPartialKernel<D + 2>), dim3(sdg::kMonitorBlocks), dim3(sdg::kThreads), 0, st, mon,
                       E, part, status);
    hipLaunchKernelGGL((sdg::monitorKernel<D + 2>), dim3(1), dim3(sdg::kThreads), 0, st, part, E, rel, status);
  };
  return l;
}

// Kernel set for (element type, order). Mixed precision runs on the matrix-core kernels only.
template <int D, int NB, int NQ, int NF, int NFQ>
bool pickMfma(bool force_valu, bool mixed, Launchers& l) {
  if (mixed)
    l = makeLaunchers<D, NB, NQ, NF, NFQ, true, float>();
  else
    l = force_valu ? makeLaunchers<D, NB, NQ, NF, NFQ, false>() : makeLaunchers<D, NB, NQ, NF, NFQ, true>();
  return true;
}
template <int D, int NB, int NQ, int NF, int NFQ>
bool pickValu(bool mixed, Launchers& l) {
  if (mixed) return false;
  l = makeLaunchers<D, NB, NQ, NF, NFQ, false>();
  return true;
}

// Navier-Stokes on hexahedra P2..P4: the sum-factorised stage kernels replace the matrix-core
// gradient and residual kernels (the face-flux kernel, the adjoint kernels and Euler stay).
template <typename TC, int N>
void useTensorKernels(Launchers& l) {
  constexpr int EPB = sdg::TensorShape<N>::EPB, TPB = sdg::TensorShape<N>::TPB;
  l.ns_grad_br1 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    hipLaunchKernelGGL((sdg::nsGradTensor<TC, N, 1>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a);
  };
  l.ns_grad_br2 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    hipLaunchKernelGGL((sdg::nsGradTensor<TC, N, 2>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a);
  };
  // face fluxes from the elements' own normal viscous fluxes (written by the gradient kernel)
  l.ns_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    constexpr int NFQ = N * N;
    hipLaunchKernelGGL((sdg::nsFaceFluxKernel<TC, 3, 6, NFQ, true>), dim3(blocksFor((g.F + g.Fb) * NFQ, sdg::kThreads)),
                       dim3(sdg::kThreads), 0, st, g, ph, a);
  };
  l.ns_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    hipLaunchKernelGGL((sdg::nsResidualTensor<TC, N>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph,
                       a);
  };
  l.trace = [](hipStream_t st, const Geo& g, const double* U, void* Uf, const sdg::MixRef& mix) {
    hipLaunchKernelGGL((sdg::tensorTraceKernel<TC, N>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, U,
                       Uf, mix);
  };
  // the transpose of these stage kernels, FP64 in both precision modes: a mixed-precision context's
  // adjoint differentiates the FP64 residual at the states of its run, on FP64 traces (trace64)
  l.trace64 = [](hipStream_t st, const Geo& g, const double* U, void* Uf, const sdg::MixRef& mix) {
    hipLaunchKernelGGL((sdg::tensorTraceKernel<double, N>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, U, Uf, mix);
  };
  l.adj_volume[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    hipLaunchKernelGGL((sdg::nsAdjVolumeTensor<N, 1>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a, j);
  };
  l.adj_volume[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    hipLaunchKernelGGL((sdg::nsAdjVolumeTensor<N, 2>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a, j);
  };
  l.adj_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    constexpr int NFQ = N * N;
    if (ph.flux <= SDG_FLUX_ROE)
      hipLaunchKernelGGL((sdg::nsAdjFaceRevKernel<3, 6, NFQ>), dim3(blocksFor(g.F * NFQ, sdg::kThreads)),
                         dim3(sdg::kThreads), 0, st, g, ph, a, j);
    else
      hipLaunchKernelGGL((sdg::nsAdjFaceKernel<3, 6, NFQ>), dim3(blocksFor(g.F * NFQ * 2 * 6, sdg::kThreads)),
                         dim3(sdg::kThreads), 0, st, g, ph, a, j);
  };
  l.adj_lift[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    hipLaunchKernelGGL((sdg::nsAdjLiftTensor<N, 1>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a, j);
  };
  l.adj_lift[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    hipLaunchKernelGGL((sdg::nsAdjLiftTensor<N, 2>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a, j);
  };
  l.adj_gather = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a, const sdg::AdjArgs& j) {
    hipLaunchKernelGGL((sdg::nsAdjGatherTensor<N>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a, j);
  };
  l.tensor = true;
}
// Navier-Stokes (BR1 or BR2, FP64 or mixed) with artificial viscosity on the sum-factorised kernels: the
// gradient kernel adds eps grad U (the volume gradient) to the volume flux and to each side's normal
// viscous flux, boundary points keep the volume-gradient trace (GfV) for the boundary flux
template <typename TC, int N>
void useTensorAvKernels(Launchers& l) {
  constexpr int EPB = sdg::TensorShape<N>::EPB, TPB = sdg::TensorShape<N>::TPB;
  l.av_ns_grad[0] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    hipLaunchKernelGGL((sdg::nsGradTensor<TC, N, 1, true>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a);
  };
  l.av_ns_grad[1] = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    hipLaunchKernelGGL((sdg::nsGradTensor<TC, N, 2, true>), dim3(blocksFor(g.E, EPB)), dim3(TPB), 0, st, g, ph, a);
  };
  l.ns_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
    constexpr int NFQ = N * N;
    hipLaunchKernelGGL((sdg::nsFaceFluxKernel<TC, 3, 6, NFQ, true, true>),
                       dim3(blocksFor((g.F + g.Fb) * NFQ, sdg::kThreads)), dim3(sdg::kThreads), 0, st, g, ph, a);
  };
  l.av_ns_flux = l.ns_flux;
}
template <typename TC>
bool selectTensorAvT(int order, Launchers& l) {
  switch (order) {
    case 2: useTensorAvKernels<TC, 3>(l); return true;
    case 3: useTensorAvKernels<TC, 4>(l); return true;
    case 4: useTensorAvKernels<TC, 5>(l); return true;
    default: return false;
  }
}
bool selectTensorAv(int order, bool mixed, Launchers& l) {
  return mixed ? selectTensorAvT<float>(order, l) : selectTensorAvT<double>(order, l);
}

template <int N>
bool pickTensor(bool mixed, Launchers& l) {
  if (mixed)
    useTensorKernels<float, N>(l);
  else
    useTensorKernels<double, N>(l);
  return true;
}

bool selectLaunchers(int type, int order, bool force_valu, bool mixed, Launchers& l, bool tensor = false) {
  if (type == SDG_ELEM_HEXAHEDRON && tensor && !force_valu && order >= 2 && order <= 4) {
    if (!selectLaunchers(type, order, false, mixed, l)) return false;
    switch (order) {
      case 2: return pickTensor<3>(mixed, l);
      case 3: return pickTensor<4>(mixed, l);
      default: return pickTensor<5>(mixed, l);
    }
  }
  if (type == SDG_ELEM_LINE) {  // 1-D: vector kernels (P+1 modes, P+1 Gauss points, two end-point faces)
    switch (order) {
      case 1: return pickValu<1, 2, 2, 2, 1>(mixed, l);
      case 2: return pickValu<1, 3, 3, 2, 1>(mixed, l);
      case 3: return pickValu<1, 4, 4, 2, 1>(mixed, l);
      case 4: return pickValu<1, 5, 5, 2, 1>(mixed, l);
      case 5: return pickValu<1, 6, 6, 2, 1>(mixed, l);
      default: return false;
    }
  }
  if (type == SDG_ELEM_TRIANGLE) {  // bases of 3..15 functions: vector kernels
    switch (order) {
      case 1: return pickValu<2, 3, 3, 3, 2>(mixed, l);
      case 2: return pickValu<2, 6, 6, 3, 3>(mixed, l);
      case 3: return pickValu<2, 10, 12, 3, 4>(mixed, l);
      case 4: return pickValu<2, 15, 16, 3, 5>(mixed, l);
      default: return false;
    }
  }
  if (type == SDG_ELEM_QUADRANGLE) {
    switch (order) {
      case 1: return pickValu<2, 4, 4, 4, 2>(mixed, l);
      case 2: return pickValu<2, 9, 9, 4, 3>(mixed, l);
      case 3: return pickMfma<2, 16, 16, 4, 4>(force_valu, mixed, l);
      case 4: return pickMfma<2, 25, 25, 4, 5>(force_valu, mixed, l);
      case 5: return pickMfma<2, 36, 36, 4, 6>(force_valu, mixed, l);
      default: return false;
    }
  }
  if (type == SDG_ELEM_TETRAHEDRON) {  // bases of 4..35 functions, Grundmann-Moeller volume rules: vector kernels
    switch (order) {
      case 1: return pickValu<3, 4, 5, 4, 4>(mixed, l);
      case 2: return pickValu<3, 10, 15, 4, 7>(mixed, l);
      case 3: return pickValu<3, 20, 35, 4, 13>(mixed, l);
      case 4: return pickValu<3, 35, 70, 4, 19>(mixed, l);
      default: return false;
    }
  }
  if (type == SDG_ELEM_PYRAMID) {  // (P+1)^3 collapsed points, five faces of (P+1)^2 slots: vector kernels
    switch (order) {
      case 1: return pickValu<3, 5, 8, 5, 4>(mixed, l);
      case 2: return pickValu<3, 14, 27, 5, 9>(mixed, l);
      case 3: return pickValu<3, 30, 64, 5, 16>(mixed, l);
      case 4: return pickValu<3, 55, 125, 5, 25>(mixed, l);
      default: return false;
    }
  }
  if (type == SDG_ELEM_HEXAHEDRON) {
    switch (order) {
      case 1: return pickValu<3, 8, 8, 6, 4>(mixed, l);
      case 2: return pickMfma<3, 27, 27, 6, 9>(force_valu, mixed, l);
      case 3: return pickMfma<3, 64, 64, 6, 16>(force_valu, mixed, l);
      case 4: return pickMfma<3, 125, 125, 6, 25>(force_valu, mixed, l);
      default: return false;
    }
  }
  return false;
}

// 1-D tables and node/face maps of the sum-factorised hexahedron kernels (see dg_tensor.hpp).
bool buildTensorTables(const sdg::RefElement& r, sdg::TensorOps& t, std::vector<short>& perm, std::vector<short>& iperm,
                       std::vector<short>& slot_of, std::vector<short>& base, std::vector<double>& fwt) {
  const int N = r.order + 1, NPT = N * N * N, NFQ = r.nfq;
  if (r.type != SDG_ELEM_HEXAHEDRON || N > sdg::kTensorMaxN || r.nb != NPT || r.nq != NPT || NFQ != N * N) return false;
  std::vector<double> gx, gw;
  sdg::gaussLegendre(N, gx, gw);
  t = sdg::TensorOps{};
  std::vector<double> B(static_cast<size_t>(N) * N), Bi(B.size());
  for (int i = 0; i < N; i++)
    for (int k = 0; k < N; k++) {
      double v, d;
      sdg::lobatto(k, gx[i], v, d);
      B[static_cast<size_t>(i) * N + k] = v;
    }
  if (!sdg::invertMatrix(N, B.data(), Bi.data())) return false;
  std::vector<double> c(N, 1.0);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      if (j != i) c[i] *= gx[i] - gx[j];
  for (int i = 0; i < N; i++) {
    t.w[i] = gw[i];
    for (int k = 0; k < N; k++) {
      t.B[i * N + k] = B[static_cast<size_t>(i) * N + k];
      t.Binv[i * N + k] = Bi[static_cast<size_t>(i) * N + k];
      double dnm = 0.0;
      if (k != i) {
        dnm = (c[i] / c[k]) / (gx[i] - gx[k]);
      } else {
        for (int j = 0; j < N; j++)
          if (j != i) dnm += 1.0 / (gx[i] - gx[j]);
      }
      t.Dn[i * N + k] = dnm;  // derivative of Lagrange k at node i
    }
    for (int s = 0; s < 2; s++) {
      const double x0 = s == 0 ? -1.0 : 1.0;
      double l = 1.0;
      for (int j = 0; j < N; j++)
        if (j != i) l *= (x0 - gx[j]) / (gx[i] - gx[j]);
      t.lf[s * N + i] = l;
    }
  }
  for (int s = 0; s < 2; s++) {
    t.kappa[s] = 0.0;
    for (int i = 0; i < N; i++) t.kappa[s] += t.lf[s * N + i] * t.lf[s * N + i] / t.w[i];
  }
  auto nodeOf = [&](double x) {
    for (int i = 0; i < N; i++)
      if (std::fabs(x - gx[i]) < 1e-12) return i;
    return -1;
  };
  perm.assign(NPT, 0);
  iperm.assign(NPT, 0);
  for (int b = 0; b < NPT; b++) {
    const int tq = r.modes[b][0] + N * r.modes[b][1] + N * N * r.modes[b][2];
    perm[b] = static_cast<short>(tq);
    iperm[tq] = static_cast<short>(b);
  }
  slot_of.assign(static_cast<size_t>(6) * NPT, -1);
  base.assign(static_cast<size_t>(6) * NFQ, -1);
  fwt.assign(static_cast<size_t>(6) * NFQ, 0.0);
  const int stride[3] = {1, N, N * N};
  for (int s = 0; s < 6; s++) {
    const double* fp = &r.face_param[static_cast<size_t>(s) * NFQ * 3];
    int ax = -1;
    for (int d = 0; d < 3 && ax < 0; d++) {
      bool on = true;
      for (int p = 0; p < NFQ; p++) on = on && std::fabs(std::fabs(fp[p * 3 + d]) - 1.0) < 1e-12 && fp[p * 3 + d] * fp[d] > 0.0;
      if (on) ax = d;
    }
    if (ax < 0) return false;
    t.axis[s] = ax;
    t.side[s] = fp[ax] > 0.0 ? 1 : 0;
    for (int p = 0; p < NFQ; p++) {
      int bq = 0;
      for (int d = 0; d < 3; d++) {
        if (d == ax) continue;
        const int i = nodeOf(fp[p * 3 + d]);
        if (i < 0) return false;
        bq += i * stride[d];
      }
      base[static_cast<size_t>(s) * NFQ + p] = static_cast<short>(bq);
      fwt[static_cast<size_t>(s) * NFQ + p] = r.fweight[p];
      for (int i = 0; i < N; i++) slot_of[static_cast<size_t>(s) * NPT + bq + i * stride[ax]] = static_cast<short>(p);
    }
  }
  for (short v : slot_of)
    if (v < 0) return false;
  return true;
}

// Zero-padded transposed operator T[r4(K)][r16(M)] with T[k][m] = src(m, k).
template <typename F>
std::vector<double> paddedOperator(int K, int M, F src) {
  const int Kp = sdg::r4(K), Mp = sdg::r16(M);
  std::vector<double> t(static_cast<size_t>(Kp) * Mp, 0.0);
  for (int k = 0; k < K; k++)
    for (int m = 0; m < M; m++) t[static_cast<size_t>(k) * Mp + m] = src(m, k);
  return t;
}

}  // namespace

// ---- scratch of the allocation-free entry point (sdg_compute_dt_noalloc) ----------------------
// The per-element CFL minima and, after them, the minimum over the mesh. Every device the library
// runs on holds one copy until the process ends; calls on one device share it, so calls that
// overlap on one device race. In host code the name is only a stand-in: kernels and copies take
// the device address that hipGetSymbolAddress returns for the current device (noallocScratch).
constexpr int64_t kNoAllocMaxElements = int64_t(1) << 21;  // 16 MiB of minima
static __device__ __align__(256) double scratch[kNoAllocMaxElements + 32];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device; resolved once per
// device (first context created there) and kept. Null with msg set on failure.
double* noallocScratch(int device, std::string& msg) {
  if (device < 0 || device >= kMaxDevices) {
    msg = "device index beyond the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[device];
}

// state-shaped arrays between the reference layout [E][N_b][NV] and the device layout [E][NV][N_b]
// (one thread per entry, writes coalesced)
__global__ __launch_bounds__(256) void referenceToDeviceKernel(const double* ref, double* dev, int E, int nv, int nb) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= static_cast<int64_t>(E) * nv * nb) return;
  const int64_t e = i / (nv * nb);
  const int r = static_cast<int>(i % (nv * nb)), v = r / nb, b = r % nb;
  dev[i] = ref[(e * nb + b) * nv + v];
}
__global__ __launch_bounds__(256) void deviceToReferenceKernel(const double* dev, double* ref, int E, int nv, int nb) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= static_cast<int64_t>(E) * nv * nb) return;
  const int64_t e = i / (nv * nb);
  const int r = static_cast<int>(i % (nv * nb)), b = r / nv, v = r % nv;
  ref[i] = dev[(e * nv + v) * nb + b];
}

// out = min over v[0..n) in one workgroup (the minimum is exact, so the order does not matter).
__global__ __launch_bounds__(256) void minReduceKernel(const double* v, int64_t n, double* out) {
  __shared__ double part[256];
  const int tid = threadIdx.x;
  double best = 1.0e308;
  for (int64_t i = tid; i < n; i += 256) best = fmin(best, v[i]);
  part[tid] = best;
  __syncthreads();
  for (int width = 128; width > 0; width >>= 1) {
    if (tid < width) part[tid] = fmin(part[tid], part[tid + width]);
    __syncthreads();
  }
  if (tid == 0) *out = part[0];
}

}  // namespace

struct sdg_ctx {
  std::string err;
  sdg_problem_desc prob{};
  Phys phys{};
  Launchers launch;
  int dim = 0, nv = 0, ng = 0, nb = 0, nq = 0, nf = 0, nfq = 0, naq = 0, order = 0, type = 0;
  int E = 0;
  int64_t F = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> allocations;
  Geo geo{};
  double* minedge = nullptr;
  double* lsq = nullptr;  // host copy
  std::vector<double> h_lsq, h_V, h_chol;
  std::vector<int> h_bparent;  // parent element of each boundary face
  double* GfV = nullptr;  // Navier-Stokes with artificial viscosity: volume-gradient traces
  double *U = nullptr, *Un = nullptr, *Uf[2] = {nullptr, nullptr}, *Gf = nullptr, *Rvol = nullptr, *Hs = nullptr, *Fv = nullptr,
         *Vn = nullptr, *Rsrc = nullptr, *src_li = nullptr;
  double *Rout = nullptr, *Wout = nullptr, *mon = nullptr, *rel = nullptr, *dtbuf = nullptr;
  int* status = nullptr;
  double* Wbnd = nullptr;  // BR2 raw files: per boundary face the gradient block (allocated on first use)
  // device the context was created on, and that device's address of the file-scope scratch array
  // of sdg_compute_dt_noalloc (null with scratch_err set if it could not be resolved)
  int device = 0;
  double* scratch = nullptr;
  std::string scratch_err;
  // artificial viscosity: node values (written once per step by the sensor kernel)
  bool av = false;
  double* nodeav = nullptr;
  int64_t nnode = 0;
  // boundary faces: count and boundary data (computational variables per boundary face point)
  int64_t Fb = 0;
  double* bdata = nullptr;
  bool has_bdata = false;
  // mixed-precision mode: FP32 work buffers and the reference state of the deviation form
  bool mixed = false;
  float *Uf32[2] = {nullptr, nullptr}, *Gf32 = nullptr, *Hs32 = nullptr, *Fv32 = nullptr;
  double* one = nullptr;            // device: modal coefficients of the constant function 1
  std::vector<double> h_one, h_mean;  // host: the same; reference-element mean of each basis function
  sdg::MixRef mix{};
  int cur = 0;
  bool has_state = false;
  // discrete adjoint: transposed operator tables and lazily allocated work buffers
  sdg::AdjOps adj_ops{};
  bool adj_ready = false;
  double *adj_lam = nullptr, *adj_Hb = nullptr, *adj_RtV = nullptr, *adj_Ua = nullptr, *adj_UFf = nullptr;
  double *adj_GFb = nullptr, *adj_Pown = nullptr, *adj_Qnb = nullptr, *adj_out = nullptr;
  double *adj_S[3] = {nullptr, nullptr, nullptr}, *adj_acc = nullptr, *adj_tmp = nullptr;
  double *adj_Uf = nullptr, *adj_Gf = nullptr;  // mixed-precision contexts: FP64 traces and gradient traces
  std::vector<double*> snaps;   // checkpoint states (device layout)
  std::vector<int> snap_step;   // step index held by each snapshot, -1 when free
  int adj_nsteps = -1, adj_nslots = 0;
  double adj_dt = 0.0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  double graph_dt[2] = {0.0, 0.0};
  hipGraphExec_t adj_graph[2] = {nullptr, nullptr};  // one adjoint step, by trace-buffer parity
  double adj_graph_dt[2] = {0.0, 0.0};
  // recorded run that keeps every stage state (when the snapshot budget covers every step): the
  // pool holds steps x stages states, written by the recording graph; the reverse sweep then
  // needs no recomputation
  double* adj_pool = nullptr;
  size_t adj_pool_cap = 0;  // in doubles
  double* adj_tpool = nullptr;  // the stage states' traces, steps x stages x trace size
  size_t adj_tpool_cap = 0;
  double* adj_T = nullptr;      // the current step's stage traces (stages x trace size)
  bool adj_stored = false;
  bool recording = false;  // stage launches also write the recording pool (graph capture only)
  int rec_nsteps = 0;
  hipGraphExec_t rec_graph[2] = {nullptr, nullptr};
  double rec_graph_dt[2] = {0.0, 0.0};
  int rec_graph_n[2] = {0, 0};
  hipGraphExec_t adjst_graph[2] = {nullptr, nullptr};
  double adjst_graph_dt[2] = {0.0, 0.0};
  bool use_graph = true;
  float last_ms = 0.0f;

  bool ns() const { return prob.equation == SDG_EQ_COMPRESSIBLE_NS; }
  void* traces(int i) const { return mixed ? static_cast<void*>(Uf32[i]) : static_cast<void*>(Uf[i]); }
  int stages() const {
    return prob.time_integration == SDG_RK_FORWARD_EULER ? 1 : (prob.time_integration == SDG_RK_HEUN2 ? 2 : 3);
  }
  void rkCoefficients(int stage, double c[3]) const {  // TimeIntegration.cpp:45-65
    if (prob.time_integration == SDG_RK_FORWARD_EULER) {
      c[0] = 1.0; c[1] = 0.0; c[2] = 1.0;
    } else if (prob.time_integration == SDG_RK_HEUN2) {
      const double t[2][3] = {{1.0, 0.0, 1.0}, {0.5, 0.5, 0.5}};
      for (int i = 0; i < 3; i++) c[i] = t[stage][i];
    } else {
      const double t[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
      for (int i = 0; i < 3; i++) c[i] = t[stage][i];
    }
  }
};

namespace {

int fail(sdg_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  sdg::setGlobalError(msg);
  return code;
}

#define HIP_TRY(ctx, call)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail((ctx), SDG_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int upload(sdg_ctx* ctx, const std::vector<T>& host, T** dev) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(host.size(), 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(ctx, SDG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocations.push_back(p);
  if (!host.empty()) {
    e = hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  *dev = static_cast<T*>(p);
  return SDG_OK;
}

template <typename T>
int allocate(sdg_ctx* ctx, size_t count, T** dev) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return fail(ctx, SDG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocations.push_back(p);
  *dev = static_cast<T*>(p);
  return SDG_OK;
}

void destroyGraphs(sdg_ctx* ctx) {
  for (int i = 0; i < 2; i++) {
    if (ctx->graph[i]) (void)hipGraphExecDestroy(ctx->graph[i]);
    ctx->graph[i] = nullptr;
    if (ctx->adj_graph[i]) (void)hipGraphExecDestroy(ctx->adj_graph[i]);
    ctx->adj_graph[i] = nullptr;
    if (ctx->rec_graph[i]) (void)hipGraphExecDestroy(ctx->rec_graph[i]);
    ctx->rec_graph[i] = nullptr;
    if (ctx->adjst_graph[i]) (void)hipGraphExecDestroy(ctx->adjst_graph[i]);
    ctx->adjst_graph[i] = nullptr;
  }
}

StageArgs baseArgs(sdg_ctx* ctx) {
  StageArgs a{};
  a.U = ctx->U;
  a.Un = ctx->Un;
  a.Uf_in = ctx->Uf[ctx->cur];
  a.Uf_out = ctx->Uf[ctx->cur ^ 1];
  a.Gf = ctx->Gf;
  a.GfV = ctx->GfV;
  a.Rvol = ctx->Rvol;
  a.Rout = ctx->Rout;
  a.Wout = nullptr;
  a.Hs = ctx->Hs;
  a.Vn = ctx->Vn;
  a.Rsrc = ctx->Rsrc;
  a.Fv = ctx->Fv;
  a.Uf_in32 = ctx->Uf32[ctx->cur];
  a.Uf_out32 = ctx->Uf32[ctx->cur ^ 1];
  a.Gf32 = ctx->Gf32;
  a.Hs32 = ctx->Hs32;
  a.Fv32 = ctx->Fv32;
  a.mix = ctx->mix;
  a.mon = ctx->mon;
  a.status = ctx->status;
  return a;
}

// Enqueue one stage (mode 0: update, 1: residual only). Advances ctx->cur for mode 0. With ev
// (six events), records ev[0]/ev[1] around the first kernel, ev[2]/ev[3] around the face-flux
// kernel and ev[4]/ev[5] around the residual kernel (empty intervals where a path has none).
void enqueueStage(sdg_ctx* ctx, int stage, double dt, int mode, double* wout, hipEvent_t* ev = nullptr) {
  StageArgs a = baseArgs(ctx);
  double c[3];
  ctx->rkCoefficients(stage < 0 ? 0 : stage, c);
  a.c0 = c[0];
  a.c1 = c[1];
  a.c2dt = c[2] * dt;
  a.last_stage = (stage == ctx->stages() - 1) ? 1 : 0;
  a.mode = mode;
  if (ctx->recording && mode == 0) {
    a.rec = ctx->adj_pool;
    a.rec_t = ctx->mixed ? nullptr : ctx->adj_tpool;  // a mixed run's traces are FP32 deviations
    a.rec_stage = stage;
    a.rec_S = ctx->stages();
    a.rec_nsteps = ctx->rec_nsteps;
    a.rec_n = static_cast<int64_t>(ctx->E) * ctx->nv * ctx->nb;
    a.rec_nt = static_cast<int64_t>(ctx->E) * ctx->nv * ctx->naq;
  }
  a.Wout = wout;
  if (mode == 1) a.status = nullptr;
  if (ctx->Rsrc) ctx->launch.source(ctx->stream, ctx->geo, ctx->phys, ctx->U, ctx->Rsrc, ctx->src_li);
  if (ev) (void)hipEventRecord(ev[0], ctx->stream);
  if (ctx->ns()) {
    if (ctx->av)
      ctx->launch.av_ns_grad[ctx->prob.viscous_flux == SDG_VISC_BR1 ? 0 : 1](ctx->stream, ctx->geo, ctx->phys, a);
    else if (ctx->prob.viscous_flux == SDG_VISC_BR1)
      ctx->launch.ns_grad_br1(ctx->stream, ctx->geo, ctx->phys, a);
    else
      ctx->launch.ns_grad_br2(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) {
      (void)hipEventRecord(ev[1], ctx->stream);
      (void)hipEventRecord(ev[2], ctx->stream);
    }
    if (ctx->launch.ns_face) ctx->launch.ns_face(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) {
      (void)hipEventRecord(ev[3], ctx->stream);
      (void)hipEventRecord(ev[4], ctx->stream);
    }
    if (ctx->av)
      ctx->launch.av_ns_flux(ctx->stream, ctx->geo, ctx->phys, a);
    else
      ctx->launch.ns_flux(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) (void)hipEventRecord(ev[5], ctx->stream);
  } else if (ctx->av) {
    ctx->launch.av_grad(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) {
      (void)hipEventRecord(ev[1], ctx->stream);
      (void)hipEventRecord(ev[2], ctx->stream);
      (void)hipEventRecord(ev[3], ctx->stream);
      (void)hipEventRecord(ev[4], ctx->stream);
    }
    ctx->launch.av_flux(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) (void)hipEventRecord(ev[5], ctx->stream);
  } else {
    ctx->launch.euler(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev)
      for (int k = 1; k < 6; k++) (void)hipEventRecord(ev[k], ctx->stream);
  }
  if (mode == 0) ctx->cur ^= 1;
}

// Node artificial viscosity from the current state (calculateArtificialViscosity,
// SpatialDiscrete.cpp:123-192): clear with a kernel (the launch may be captured in a graph), then
// the per-element sensor with its maximum into the nodes.
void enqueueArtificialViscosity(sdg_ctx* ctx) {
  if (!ctx->av) return;
  hipLaunchKernelGGL(sdg::avClearKernel, dim3(static_cast<unsigned>(std::min<int64_t>((ctx->nnode + sdg::kThreads - 1) / sdg::kThreads, 1024))),
                     dim3(sdg::kThreads), 0, ctx->stream, ctx->nodeav, ctx->nnode);
  ctx->launch.av_sensor(ctx->stream, ctx->geo, ctx->phys, ctx->U, ctx->nodeav);
}

void enqueueStep(sdg_ctx* ctx, double dt) {
  enqueueArtificialViscosity(ctx);  // once per step, before the stages (TimeIntegration.cpp:336-338)
  for (int s = 0; s < ctx->stages(); s++) enqueueStage(ctx, s, dt, 0, nullptr);
  ctx->launch.monitor(ctx->stream, ctx->mon, ctx->E, ctx->rel, ctx->status);
}

// Enqueue nsteps whole steps (each replayed from a captured graph) without waiting for them.
int launchSteps(sdg_ctx* ctx, double dt, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    if (ctx->use_graph) {
      const int c = ctx->cur;
      if (!ctx->graph[c] || ctx->graph_dt[c] != dt) {
        if (ctx->graph[c]) (void)hipGraphExecDestroy(ctx->graph[c]);
        ctx->graph[c] = nullptr;
        hipGraph_t graph;
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueueStep(ctx, dt);
        HIP_TRY(ctx, hipStreamEndCapture(ctx->stream, &graph));
        hipError_t e = hipGraphInstantiate(&ctx->graph[c], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        ctx->graph_dt[c] = dt;
        ctx->cur = c;  // capture advanced cur; replay below advances it for real
      }
      HIP_TRY(ctx, hipGraphLaunch(ctx->graph[c], ctx->stream));
      ctx->cur = c ^ (ctx->stages() & 1);
    } else {
      enqueueStep(ctx, dt);
    }
  }
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// reference layout (E x N_b x NV, NV fastest) <-> device layout (E x NV x N_b)
void toDevice(const sdg_ctx* ctx, const double* ref, std::vector<double>& dev) {
  const int nv = ctx->nv, nb = ctx->nb;
  dev.resize(static_cast<size_t>(ctx->E) * nv * nb);
  for (int64_t e = 0; e < ctx->E; e++)
    for (int b = 0; b < nb; b++)
      for (int v = 0; v < nv; v++) dev[(e * nv + v) * nb + b] = ref[(e * nb + b) * nv + v];
}
void toReference(const sdg_ctx* ctx, const std::vector<double>& dev, double* ref) {
  const int nv = ctx->nv, nb = ctx->nb;
  for (int64_t e = 0; e < ctx->E; e++)
    for (int b = 0; b < nb; b++)
      for (int v = 0; v < nv; v++) ref[(e * nb + b) * nv + v] = dev[(e * nv + v) * nb + b];
}

// Mixed-precision reference state: the domain mean of the conserved variables (mean over elements
// of the element means) and its pressure.
void updateReference(sdg_ctx* ctx, const double* ref) {
  sdg::MixRef& m = ctx->mix;
  m = sdg::MixRef{};
  m.on = ctx->mixed ? 1 : 0;
  m.one = ctx->one;
  if (!ctx->mixed) return;
  const int nv = ctx->nv, nb = ctx->nb, D = ctx->dim;
  for (int v = 0; v < nv; v++) {
    double s = 0.0;
    for (int64_t e = 0; e < ctx->E; e++)
      for (int b = 0; b < nb; b++) s += ref[(e * nb + b) * nv + v] * ctx->h_mean[b];
    m.uref[v] = s / static_cast<double>(ctx->E);
  }
  const double rho = m.uref[0];
  double v2 = 0.0;
  for (int d = 0; d < D; d++) v2 += (m.uref[1 + d] / rho) * (m.uref[1 + d] / rho);
  const double e = m.uref[D + 1] / rho - v2 / 2.0;
  const sdg::Phys& ph = ctx->phys;
  m.pref = ph.eos == SDG_EOS_IDEAL_GAS ? (sdg::kGamma - 1.0) * rho * e : ph.c0 * ph.c0 * (rho - ph.rho0) + ph.padd;
  m.h0 = (m.uref[D + 1] + m.pref) / rho;
}

int validateProblem(const sdg_problem_desc* p, std::string& msg) {
  if (p->abi_version != SDG_ABI_VERSION) {
    msg = "abi_version mismatch";
    return SDG_E_INVALID;
  }
  if (p->equation != SDG_EQ_COMPRESSIBLE_EULER && p->equation != SDG_EQ_COMPRESSIBLE_NS) {
    msg = "unknown equation model";
    return SDG_E_INVALID;
  }
  if (p->convective_flux < SDG_FLUX_CENTRAL || p->convective_flux > SDG_FLUX_EXACT) {
    msg = "unknown convective flux";
    return SDG_E_INVALID;
  }
  if (p->equation == SDG_EQ_COMPRESSIBLE_NS && p->viscous_flux != SDG_VISC_BR1 && p->viscous_flux != SDG_VISC_BR2) {
    msg = "Navier-Stokes needs the BR1 or BR2 viscous flux";
    return SDG_E_INVALID;
  }
  if (p->time_integration < SDG_RK_FORWARD_EULER || p->time_integration > SDG_RK_SSPRK3) {
    msg = "unknown time integration";
    return SDG_E_INVALID;
  }
  if (p->equation_of_state != SDG_EOS_IDEAL_GAS && p->equation_of_state != SDG_EOS_WEAK_COMPRESSIBLE) {
    msg = "unknown equation of state";
    return SDG_E_INVALID;
  }
  if (p->kernels < SDG_KERNELS_AUTO || p->kernels > SDG_KERNELS_VECTOR) {
    msg = "unknown kernel family";
    return SDG_E_INVALID;
  }
  if (p->launch != SDG_LAUNCH_GRAPH && p->launch != SDG_LAUNCH_EAGER) {
    msg = "unknown launch mode";
    return SDG_E_INVALID;
  }
  if (p->equation == SDG_EQ_COMPRESSIBLE_NS && p->transport != SDG_TRANSPORT_CONSTANT &&
      p->transport != SDG_TRANSPORT_SUTHERLAND) {
    msg = "Navier-Stokes needs the constant or Sutherland transport model";
    return SDG_E_INVALID;
  }
  if (!(p->cv > 0.0) || !(p->cp > 0.0)) {
    msg = "specific heats must be positive";
    return SDG_E_INVALID;
  }
  if (p->precision != SDG_PREC_FP64 && p->precision != SDG_PREC_MIXED) {
    msg = "unknown precision mode";
    return SDG_E_INVALID;
  }
  if (p->source_term != SDG_SOURCE_NONE && p->source_term != SDG_SOURCE_BOUSSINESQ) {
    msg = "unknown source term";
    return SDG_E_INVALID;
  }
  if (p->shock_capturing != SDG_SHOCK_NONE && p->shock_capturing != SDG_SHOCK_ARTIFICIAL_VISCOSITY) {
    msg = "unknown shock capturing";
    return SDG_E_INVALID;
  }
  if (p->shock_capturing == SDG_SHOCK_ARTIFICIAL_VISCOSITY &&
      (!(p->empirical_tolerance > 0.0) || !(p->artificial_viscosity_factor >= 0.0))) {
    msg = "artificial viscosity needs a positive empirical tolerance and a non-negative factor";
    return SDG_E_INVALID;
  }
  return SDG_OK;
}

}  // namespace

extern "C" {

const char* sdg_last_error(const sdg_ctx* ctx) { return ctx ? ctx->err.c_str() : sdg::globalError(); }

int sdg_create(const sdg_problem_desc* problem, const sdg_mesh* mesh_handle, sdg_ctx** out) {
  if (!problem || !mesh_handle || !out) return fail(nullptr, SDG_E_INVALID, "sdg_create: null argument");
  *out = nullptr;
  std::string msg;
  int rc = validateProblem(problem, msg);
  if (rc != SDG_OK) return fail(nullptr, rc, "sdg_create: " + msg);
  if (!sdg::meshOf(mesh_handle))
    return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: meshes of mixed element types run through sdg_batch_create");
  const sdg::Mesh& mesh = *sdg::meshOf(mesh_handle);
  const bool curved = !mesh.affine;
  if (curved && (problem->precision == SDG_PREC_MIXED || problem->kernels == SDG_KERNELS_DENSE ||
                 problem->shock_capturing == SDG_SHOCK_ARTIFICIAL_VISCOSITY))
    return fail(nullptr, SDG_E_UNSUPPORTED,
                "sdg_create: curved elements run on the FP64 vector kernels (no mixed precision, dense kernels or "
                "artificial viscosity)");
  if (mesh.elements > (int64_t(1) << 30)) return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: too many elements");
  sdg_ctx* ctx = new (std::nothrow) sdg_ctx();
  if (!ctx) return fail(nullptr, SDG_E_NOMEM, "sdg_create: out of host memory");
  ctx->prob = *problem;
  const sdg::RefElement& r = mesh.ref;
  ctx->dim = r.dim;
  ctx->nv = r.dim + 2;
  ctx->ng = ctx->nv * r.dim;
  ctx->nb = r.nb;
  ctx->nq = r.nq;
  ctx->nf = r.nf;
  ctx->nfq = r.nfq;
  ctx->naq = r.nf * r.nfq;
  ctx->order = r.order;
  ctx->type = r.type;
  ctx->E = static_cast<int>(mesh.elements);
  ctx->F = static_cast<int64_t>(mesh.faces.size());
  const bool force_vector = problem->kernels == SDG_KERNELS_VECTOR || curved;
  const bool mixed = problem->precision == SDG_PREC_MIXED;
  ctx->mixed = mixed;
  const bool tensor = problem->equation == SDG_EQ_COMPRESSIBLE_NS && problem->kernels == SDG_KERNELS_AUTO && !curved;
  ctx->av = problem->shock_capturing == SDG_SHOCK_ARTIFICIAL_VISCOSITY;
  const bool tensor_size = r.type == SDG_ELEM_HEXAHEDRON && r.order >= 2 && r.order <= 4 &&
                           (problem->viscous_flux == SDG_VISC_BR1 || problem->viscous_flux == SDG_VISC_BR2);
  if (ctx->av && mixed && problem->equation != SDG_EQ_COMPRESSIBLE_EULER && !(tensor && tensor_size)) {
    delete ctx;
    return fail(nullptr, SDG_E_UNSUPPORTED,
                "sdg_create: artificial viscosity with Navier-Stokes in mixed precision runs on the sum-factorised "
                "kernels (hexahedra P2-P4, BR1 / BR2); elsewhere it runs in FP64");
  }
  if (force_vector && mixed) {
    delete ctx;
    return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: the vector kernels run in FP64");
  }
  // artificial viscosity: the vector kernels, or for Euler in FP64 the matrix-core kernels where the
  // element size has them (the Navier-Stokes terms need the vector kernels' second gradient trace)
  // Navier-Stokes with artificial viscosity: the sum-factorised kernels for BR1 and BR2 on hexahedra
  // P2-P4 (FP64 or mixed), else the FP64 vector kernels
  const bool tensor_av = ctx->av && tensor && tensor_size;
  const bool av_vector = ctx->av && problem->equation == SDG_EQ_COMPRESSIBLE_NS && !tensor_av;
  if (!selectLaunchers(r.type, r.order, force_vector || av_vector, mixed, ctx->launch,
                       tensor && (!ctx->av || tensor_av))) {
    delete ctx;
    return fail(nullptr, SDG_E_UNSUPPORTED,
                mixed ? "sdg_create: the mixed-precision mode needs a matrix-core kernel set (quadrangle P3-P5, "
                             "hexahedron P2-P4)"
                           : "sdg_create: no kernel instantiation for this element type and order");
  }
  if (tensor_av && !selectTensorAv(r.order, mixed, ctx->launch)) {
    delete ctx;
    return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: no sum-factorised artificial-viscosity kernels for this order");
  }
  Phys& ph = ctx->phys;
  ph.eos = problem->equation_of_state;
  ph.transport = problem->transport;
  ph.flux = problem->convective_flux;
  ph.cp = problem->cp;
  ph.cv = problem->cv;
  ph.mu = problem->dynamic_viscosity;
  ph.kappa = problem->cp * problem->dynamic_viscosity / sdg::kPrandtl;  // PhysicalModel.cpp:149-153
  ph.c0 = problem->reference_sound_speed;
  ph.rho0 = problem->reference_density;
  ph.padd = 0.01 * problem->reference_density * problem->reference_sound_speed * problem->reference_sound_speed;
  ph.source = problem->source_term;
  ph.beta = problem->thermal_expansion_coefficient;
  ph.tref = problem->reference_temperature;
  ph.shock = problem->shock_capturing;
  ph.av_tol = problem->empirical_tolerance;
  ph.av_factor = problem->artificial_viscosity_factor;
  {
    // kArtificialViscosityTolerance (SimulationControl.cpp:892-893): log10(1/P^4) to 12 digits
    static const double kThreshold[5] = {0.0, -1.20411998266, -1.90848501888, -2.40823996531, -2.79588001734};
    ph.av_s0 = kThreshold[r.order - 1];
    const int p = r.order - 1;  // the hierarchic basis of order P-1 is a prefix of order P's
    ph.av_low = p == 0 ? 0 : sdg::basisSize(r.type, r.dim, p);
  }
  ctx->use_graph = problem->launch == SDG_LAUNCH_GRAPH;

  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) {
    delete ctx;
    return fail(nullptr, SDG_E_DEVICE, "sdg_create: no HIP device available");
  }
  auto bail = [&](int code) {
    std::string m = ctx->err;
    sdg_destroy(ctx);
    sdg::setGlobalError(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    ctx->err = "sdg_create: stream/event creation failed";
    return bail(SDG_E_DEVICE);
  }
  if (hipGetDevice(&ctx->device) != hipSuccess) {
    ctx->err = "sdg_create: hipGetDevice failed";
    return bail(SDG_E_DEVICE);
  }
  ctx->scratch = noallocScratch(ctx->device, ctx->scratch_err);
  const int D = r.dim, nb = r.nb, nq = r.nq, nf = r.nf, nfq = r.nfq, naq = nf * nfq;
  const int64_t E = mesh.elements, F = ctx->F;
  // operand tables
  std::vector<double> Vt(static_cast<size_t>(nb) * nq), Vft(static_cast<size_t>(nb) * naq), Lf(static_cast<size_t>(nf) * nfq * nfq);
  for (int q = 0; q < nq; q++)
    for (int b = 0; b < nb; b++) Vt[static_cast<size_t>(b) * nq + q] = r.V[static_cast<size_t>(q) * nb + b];
  for (int s = 0; s < naq; s++)
    for (int b = 0; b < nb; b++) Vft[static_cast<size_t>(b) * naq + s] = r.Vf[static_cast<size_t>(s) * nb + b];
  for (int s = 0; s < nf; s++)
    for (int p = 0; p < nfq; p++)
      for (int j = 0; j < nfq; j++) {
        double acc = 0.0;
        for (int b = 0; b < nb; b++) {
          double t = 0.0;
          for (int c = 0; c < nb; c++) t += r.mass_inverse[static_cast<size_t>(b) * nb + c] * r.Vf[static_cast<size_t>(s * nfq + j) * nb + c];
          acc += r.Vf[static_cast<size_t>(s * nfq + p) * nb + b] * t;
        }
        Lf[(static_cast<size_t>(s) * nfq + p) * nfq + j] = acc;
      }
  // per element: (J^T)^-1 column-major and |J|
  std::vector<double> ginv(static_cast<size_t>(E) * D * D), detj(E);
  for (int64_t e = 0; e < E; e++) {
    const double* J = &mesh.jac[static_cast<size_t>(e) * D * D];
    double jt[9], jti[9];
    for (int k = 0; k < D; k++)
      for (int l = 0; l < D; l++) jt[k * D + l] = J[l * D + k];
    sdg::invertMatrix(D, jt, jti);
    for (int a = 0; a < D; a++)
      for (int b = 0; b < D; b++) ginv[static_cast<size_t>(e) * D * D + a + D * b] = jti[a * D + b];
    detj[e] = mesh.detj[e];
  }
  std::vector<double> fnormal(static_cast<size_t>(F) * nfq * D), fjw(static_cast<size_t>(F) * nfq);
  std::vector<int> fseq(static_cast<size_t>(F) * nfq), fseqinv(static_cast<size_t>(F) * nfq);
  for (int64_t f = 0; f < F; f++) {
    const sdg::Face& fc = mesh.faces[f];
    for (int j = 0; j < nfq; j++) {
      for (int d = 0; d < D; d++) fnormal[(static_cast<size_t>(f) * nfq + j) * D + d] = fc.normal[static_cast<size_t>(j) * D + d];
      fjw[static_cast<size_t>(f) * nfq + j] = fc.jw[j];
      fseq[static_cast<size_t>(f) * nfq + j] = fc.seq[j];
      fseqinv[static_cast<size_t>(f) * nfq + fc.seq[j]] = j;
    }
  }
  std::vector<int> fpar(static_cast<size_t>(F) * 2), floc(static_cast<size_t>(F) * 2);
  for (int64_t f = 0; f < F; f++) {
    fpar[2 * f] = mesh.faces[f].parent[0];
    fpar[2 * f + 1] = mesh.faces[f].parent[1];
    floc[2 * f] = mesh.faces[f].local[0];
    floc[2 * f + 1] = mesh.faces[f].local[1];
  }
  // per element slot: neighbour slot and (face point, side), so a kernel reaches the neighbour's
  // trace with one indexed load (no face -> permutation -> slot chain); a slot on boundary face b,
  // point j has snb = -1 - (b N_fq + j)
  const int64_t Fb = static_cast<int64_t>(mesh.bfaces.size());
  if (static_cast<int64_t>(E) * naq >= (int64_t(1) << 31) || (F + Fb) * nfq >= (int64_t(1) << 30)) {
    ctx->err = "sdg_create: mesh too large for 32-bit slot indices";
    return bail(SDG_E_UNSUPPORTED);
  }
  std::vector<int> snb(static_cast<size_t>(E) * naq), sfp(static_cast<size_t>(E) * naq);
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      const int fid = mesh.elem_face[static_cast<size_t>(e) * nf + s];
      const int side = mesh.elem_side[static_cast<size_t>(e) * nf + s];
      if (fid == -1) {
        ctx->err = "sdg_create: element face without a neighbour or boundary condition";
        return bail(SDG_E_INVALID);
      }
      for (int j = 0; j < nfq; j++) {
        const size_t o = static_cast<size_t>(e) * naq + s * nfq + j;
        if (fid <= -2) {
          snb[o] = -1 - ((-2 - fid) * nfq + j);
          sfp[o] = 0;
          continue;
        }
        const sdg::Face& fc = mesh.faces[fid];
        const int p = side == 0 ? j : fseqinv[static_cast<size_t>(fid) * nfq + j];
        const int k = side == 0 ? fseq[static_cast<size_t>(fid) * nfq + j] : p;
        snb[o] = fc.parent[side ^ 1] * naq + fc.local[side ^ 1] * nfq + k;
        sfp[o] = (fid * nfq + p) * 2 + side;
      }
    }
  // boundary face points
  std::vector<double> bnormal(static_cast<size_t>(std::max<int64_t>(Fb, 1)) * nfq * D), bjw(static_cast<size_t>(std::max<int64_t>(Fb, 1)) * nfq);
  std::vector<int> btype(bjw.size()), bslot(bjw.size());
  for (int64_t b = 0; b < Fb; b++) {
    const sdg::Face& fc = mesh.bfaces[b];
    for (int j = 0; j < nfq; j++) {
      const size_t bp = static_cast<size_t>(b) * nfq + j;
      for (int d = 0; d < D; d++) bnormal[bp * D + d] = fc.normal[static_cast<size_t>(j) * D + d];
      bjw[bp] = fc.jw[j];
      btype[bp] = fc.bc;
      bslot[bp] = fc.parent[0] * naq + fc.local[0] * nfq + j;
    }
    ctx->h_bparent.push_back(fc.parent[0]);
  }
  ctx->Fb = Fb;
  Geo& g = ctx->geo;
  if (ctx->launch.tensor) {  // per slot normal and face weight for the sum-factorised kernels
    std::vector<double> snrm(static_cast<size_t>(E) * naq * D), sjw(static_cast<size_t>(E) * naq);
    for (size_t o = 0; o < sjw.size(); o++) {
      if (snb[o] < 0) {
        const size_t bp = static_cast<size_t>(-1 - snb[o]);
        for (int d = 0; d < D; d++) snrm[o * D + d] = bnormal[bp * D + d];
        sjw[o] = bjw[bp];
      } else {
        const size_t fp = static_cast<size_t>(sfp[o] >> 1);
        const double sg = (sfp[o] & 1) ? -1.0 : 1.0;
        for (int d = 0; d < D; d++) snrm[o * D + d] = sg * fnormal[fp * D + d];
        sjw[o] = fjw[fp];
      }
    }
    double *dsn, *dsj;
    if (upload(ctx, snrm, &dsn) || upload(ctx, sjw, &dsj)) return bail(SDG_E_NOMEM);
    g.snrm = dsn;
    g.sjw = dsj;
  }
  double *dVt, *ddV, *dVft, *dVf, *dMinv, *dLf, *dwq, *dginv, *ddetj, *dnormal, *djw, *dme;
  int *dseq, *dseqinv, *dfpar, *dfloc, *dsnb, *dsfp, *dbtype, *dbslot;
  double *dbnormal, *dbjw;
  if (upload(ctx, Vt, &dVt) || upload(ctx, r.dV, &ddV) || upload(ctx, Vft, &dVft) || upload(ctx, r.Vf, &dVf) ||
      upload(ctx, r.mass_inverse, &dMinv) || upload(ctx, Lf, &dLf) || upload(ctx, r.qweight, &dwq) ||
      upload(ctx, ginv, &dginv) || upload(ctx, detj, &ddetj) || upload(ctx, fnormal, &dnormal) ||
      upload(ctx, fjw, &djw) || upload(ctx, fseq, &dseq) || upload(ctx, fseqinv, &dseqinv) ||
      upload(ctx, mesh.min_edge, &dme) || upload(ctx, fpar, &dfpar) || upload(ctx, floc, &dfloc) ||
      upload(ctx, snb, &dsnb) || upload(ctx, sfp, &dsfp) || upload(ctx, bnormal, &dbnormal) || upload(ctx, bjw, &dbjw) ||
      upload(ctx, btype, &dbtype) || upload(ctx, bslot, &dbslot) ||
      allocate(ctx, bjw.size() * (D + 3), &ctx->bdata))
    return bail(SDG_E_NOMEM);
  g.Vt = dVt;
  g.dV = ddV;
  g.Vft = dVft;
  g.Vf = dVf;
  g.Minv = dMinv;
  g.Lf = dLf;
  g.wq = dwq;
  g.ginv = dginv;
  g.detj = ddetj;
  g.fnormal = dnormal;
  g.fjw = djw;
  g.fseq = dseq;
  g.fseqinv = dseqinv;
  g.snb = dsnb;
  g.bnormal = dbnormal;
  g.bjw = dbjw;
  g.btype = dbtype;
  g.bslot = dbslot;
  g.bdata = ctx->bdata;
  g.Fb = static_cast<int>(Fb);
  g.sfp = dsfp;
  g.fpar = dfpar;
  g.floc = dfloc;
  g.E = ctx->E;
  g.F = static_cast<int>(F);
  g.nvert = r.nvert;
  g.order = r.order;
  if (curved) {  // per-point geometry, element mass inverses and BR2 face blocks (Geometry.cpp:44-100)
    std::vector<double> qG(static_cast<size_t>(E) * nq * D * D), qjw(static_cast<size_t>(E) * nq),
        minvE(static_cast<size_t>(E) * nb * nb), LfE(static_cast<size_t>(E) * naq * nfq);
    std::vector<double> t(nb);
    for (int64_t e = 0; e < E; e++) {
      sdg::elementGeometry(mesh, e, &qjw[static_cast<size_t>(e) * nq], &qG[static_cast<size_t>(e) * nq * D * D], nullptr);
      double* M = &minvE[static_cast<size_t>(e) * nb * nb];
      sdg::elementMassInverse(mesh, e, M);
      for (int s = 0; s < nf; s++)
        for (int j = 0; j < nfq; j++) {
          for (int b = 0; b < nb; b++) {
            double acc = 0.0;
            for (int c = 0; c < nb; c++) acc += M[static_cast<size_t>(b) * nb + c] * r.Vf[static_cast<size_t>(s * nfq + j) * nb + c];
            t[b] = acc;
          }
          for (int p = 0; p < nfq; p++) {
            double acc = 0.0;
            for (int b = 0; b < nb; b++) acc += r.Vf[static_cast<size_t>(s * nfq + p) * nb + b] * t[b];
            LfE[((static_cast<size_t>(e) * nf + s) * nfq + p) * nfq + j] = acc;
          }
        }
    }
    double *dqG, *dqjw, *dminvE, *dLfE;
    if (upload(ctx, qG, &dqG) || upload(ctx, qjw, &dqjw) || upload(ctx, minvE, &dminvE) || upload(ctx, LfE, &dLfE))
      return bail(SDG_E_NOMEM);
    g.qG = dqG;
    g.qjw = dqjw;
    g.minvE = dminvE;
    g.LfE = dLfE;
  }
  if (ctx->av) {
    double *dnq, *dnf, *drad;
    int* dtag;
    std::vector<int> tag(mesh.node_tag.begin(), mesh.node_tag.end());
    ctx->nnode = mesh.nodes;
    if (upload(ctx, r.nodal_value, &dnq) || upload(ctx, r.nodal_face_value, &dnf) || upload(ctx, mesh.inner_radius, &drad) ||
        upload(ctx, tag, &dtag) || allocate(ctx, static_cast<size_t>(ctx->nnode), &ctx->nodeav))
      return bail(SDG_E_NOMEM);
    // zero until the first step forms them, as the reference initialises them (InitialCondition.cpp:156-157)
    if (hipMemsetAsync(ctx->nodeav, 0, static_cast<size_t>(ctx->nnode) * sizeof(double), ctx->stream) != hipSuccess) {
      ctx->err = "sdg_create: device initialisation failed";
      return bail(SDG_E_DEVICE);
    }
    g.nodq = dnq;
    g.nodf = dnf;
    g.inrad = drad;
    g.ntag = dtag;
    g.nodeav = ctx->nodeav;
  }
  {
    const int nfq_ = nfq;
    auto V = [&](int q, int b) { return r.V[static_cast<size_t>(q) * nb + b]; };
    auto Vfv = [&](int sl, int b) { return r.Vf[static_cast<size_t>(sl) * nb + b]; };
    std::vector<double> t_interp = paddedOperator(nb, nq, [&](int m, int k) { return V(m, k); });
    std::vector<double> t_grad = paddedOperator(nq * D, nb, [&](int m, int k) { return r.dV[static_cast<size_t>(k) * nb + m]; });
    std::vector<double> t_face = paddedOperator(naq, nb, [&](int m, int k) { return Vfv(k, m); });
    std::vector<double> t_minv = paddedOperator(nb, nb, [&](int m, int k) { return r.mass_inverse[static_cast<size_t>(k) * nb + m]; });
    std::vector<double> t_trace = paddedOperator(nb, naq, [&](int m, int k) { return Vfv(m, k); });
    std::vector<double> t_lift = paddedOperator(naq, naq, [&](int m, int k) {
      const int sm = m / nfq_, sk = k / nfq_;
      return sm == sk ? Lf[(static_cast<size_t>(sm) * nfq_ + k % nfq_) * nfq_ + m % nfq_] : 0.0;
    });
    // Mref^-1 folded into the evaluation operators: (Mref^-1 V^T)(b, q), (Mref^-1 Vf^T)(b, sl)
    std::vector<double> pinterp = paddedOperator(nb, nq, [&](int m, int k) {
      double acc = 0.0;
      for (int i = 0; i < nb; i++) acc += r.mass_inverse[static_cast<size_t>(k) * nb + i] * V(m, i);
      return acc;
    });
    std::vector<double> ptrace = paddedOperator(nb, naq, [&](int m, int k) {
      double acc = 0.0;
      for (int i = 0; i < nb; i++) acc += r.mass_inverse[static_cast<size_t>(k) * nb + i] * Vfv(m, i);
      return acc;
    });
    double *pi, *pg, *pf, *pm, *pt, *pl, *ppi, *ppt;
    if (upload(ctx, t_interp, &pi) || upload(ctx, t_grad, &pg) || upload(ctx, t_face, &pf) || upload(ctx, t_minv, &pm) ||
        upload(ctx, t_trace, &pt) || upload(ctx, t_lift, &pl) || upload(ctx, pinterp, &ppi) || upload(ctx, ptrace, &ppt))
      return bail(SDG_E_NOMEM);
    if (ctx->mixed) {
      // FP32 tables. Residuals are formed against the orthonormal test functions psi = L^-1 phi
      // (mass = L L^T): R_psi = L^-1 R, so the update (Mref^-1 R) = L^-T R_psi, the lifted
      // gradient at points is V L^-T R_psi, and the monitor's R V^T is V L R_psi. In exact
      // arithmetic this is the same scheme; in FP32 the residual's rounding then reaches the
      // solution unamplified by the conditioning of the hierarchic basis's mass matrix.
      const std::vector<double>& Li = r.mass_chol_inverse;
      const std::vector<double>& Lc = r.mass_chol;
      auto li = [&](int j, int i) { return Li[static_cast<size_t>(j) * nb + i]; };  // (L^-1)_{j i}
      std::vector<double> m_grad = paddedOperator(nq * D, nb, [&](int j, int k) {
        double acc = 0.0;
        for (int i = 0; i <= j; i++) acc += r.dV[static_cast<size_t>(k) * nb + i] * li(j, i);
        return acc;
      });
      std::vector<double> m_face = paddedOperator(naq, nb, [&](int j, int k) {
        double acc = 0.0;
        for (int i = 0; i <= j; i++) acc += Vfv(k, i) * li(j, i);
        return acc;
      });
      std::vector<double> m_minv = paddedOperator(nb, nb, [&](int m, int k) { return li(k, m); });
      std::vector<double> m_pinterp = paddedOperator(nb, nq, [&](int q, int j) {
        double acc = 0.0;
        for (int i = 0; i <= j; i++) acc += V(q, i) * li(j, i);
        return acc;
      });
      std::vector<double> m_ptrace = paddedOperator(nb, naq, [&](int sl, int j) {
        double acc = 0.0;
        for (int i = 0; i <= j; i++) acc += Vfv(sl, i) * li(j, i);
        return acc;
      });
      std::vector<double> m_rinterp = paddedOperator(nb, nq, [&](int q, int j) {
        double acc = 0.0;
        for (int i = j; i < nb; i++) acc += V(q, i) * Lc[static_cast<size_t>(i) * nb + j];
        return acc;
      });
      auto f32 = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
      float *fi, *fg, *ff, *fm, *ft, *fl, *fpi, *fpt, *fri;
      if (upload(ctx, f32(t_interp), &fi) || upload(ctx, f32(m_grad), &fg) || upload(ctx, f32(m_face), &ff) ||
          upload(ctx, f32(m_minv), &fm) || upload(ctx, f32(t_trace), &ft) || upload(ctx, f32(t_lift), &fl) ||
          upload(ctx, f32(m_pinterp), &fpi) || upload(ctx, f32(m_ptrace), &fpt) || upload(ctx, f32(m_rinterp), &fri))
        return bail(SDG_E_NOMEM);
      g.opf = {fi, fg, ff, fm, ft, fl, fpi, fpt, fri};
      double* fm64;
      if (upload(ctx, m_minv, &fm64)) return bail(SDG_E_NOMEM);
      g.minv_mixed64 = fm64;
    }
    if (ctx->launch.adj_volume[0]) {
      // transposes for the discrete adjoint
      std::vector<double> gradT = paddedOperator(nb, nq * D, [&](int m, int k) { return r.dV[static_cast<size_t>(m) * nb + k]; });
      std::vector<double> interpT = paddedOperator(nq, nb, [&](int m, int k) { return V(k, m); });
      std::vector<double> pinterpT = paddedOperator(nq, nb, [&](int m, int k) {
        double acc = 0.0;
        for (int i = 0; i < nb; i++) acc += r.mass_inverse[static_cast<size_t>(m) * nb + i] * V(k, i);
        return acc;
      });
      std::vector<double> ptraceT = paddedOperator(naq, nb, [&](int m, int k) {
        double acc = 0.0;
        for (int i = 0; i < nb; i++) acc += r.mass_inverse[static_cast<size_t>(m) * nb + i] * Vfv(k, i);
        return acc;
      });
      std::vector<double> liftT = paddedOperator(naq, naq, [&](int m, int k) {
        const int sm = m / nfq_, sk = k / nfq_;
        return sm == sk ? Lf[(static_cast<size_t>(sm) * nfq_ + m % nfq_) * nfq_ + k % nfq_] : 0.0;
      });
      double *ag, *ai, *api, *apt, *al;
      if (upload(ctx, gradT, &ag) || upload(ctx, interpT, &ai) || upload(ctx, pinterpT, &api) ||
          upload(ctx, ptraceT, &apt) || upload(ctx, liftT, &al))
        return bail(SDG_E_NOMEM);
      ctx->adj_ops = {ag, ai, api, apt, al};
    }
    g.op.pinterp = ppi;
    g.op.ptrace = ppt;
    g.op.interp = pi;
    g.op.grad = pg;
    g.op.face = pf;
    g.op.minv = pm;
    g.op.trace = pt;
    g.op.lift = pl;
    g.op.rinterp = pi;
  }
  if (ctx->launch.tensor) {
    std::vector<short> perm, iperm, slot_of, base;
    std::vector<double> fwt;
    if (!buildTensorTables(r, g.tens, perm, iperm, slot_of, base, fwt))
      return bail(fail(ctx, SDG_E_INVALID, "sdg_create: inconsistent hexahedron tables for the sum-factorised kernels"));
    short *dp, *dip, *ds, *db;
    double* dw;
    if (upload(ctx, perm, &dp) || upload(ctx, iperm, &dip) || upload(ctx, slot_of, &ds) || upload(ctx, base, &db) ||
        upload(ctx, fwt, &dw))
      return bail(SDG_E_NOMEM);
    g.tens.perm = dp;
    g.tens.iperm = dip;
    g.tens.slot_of = ds;
    g.tens.base = db;
    g.tens.fwt = dw;
    if (ctx->av) {  // P1 nodal functions at the tensor-ordered nodes (node artificial viscosity)
      const int N = r.order + 1;
      std::vector<double> gx, gw, nodt(static_cast<size_t>(N) * N * N * r.nvert);
      sdg::gaussLegendre(N, gx, gw);
      for (int q = 0; q < N * N * N; q++) {
        const double xi[3] = {gx[q % N], gx[(q / N) % N], gx[q / (N * N)]};
        sdg::nodalValues(r, xi, &nodt[static_cast<size_t>(q) * r.nvert]);
      }
      double* dn;
      if (upload(ctx, nodt, &dn)) return bail(SDG_E_NOMEM);
      g.tens.nodt = dn;
    }
  }
  ctx->minedge = dme;
  ctx->h_lsq = r.lsq_inverse;
  ctx->h_chol = r.mass_chol;
  ctx->h_V = r.V;
  {
    // coefficients of the constant 1 by the initial-condition projection (exactly 1 on the vertex
    // modes and 0 elsewhere for the Lobatto basis; rounding residue snapped), and the mean of each
    // basis function over the reference element
    std::vector<double> colsum(nb, 0.0);
    double wsum = 0.0;
    ctx->h_mean.assign(nb, 0.0);
    for (int q = 0; q < nq; q++) {
      wsum += r.qweight[q];
      for (int b = 0; b < nb; b++) {
        colsum[b] += r.V[static_cast<size_t>(q) * nb + b];
        ctx->h_mean[b] += r.qweight[q] * r.V[static_cast<size_t>(q) * nb + b];
      }
    }
    for (int b = 0; b < nb; b++) ctx->h_mean[b] /= wsum;
    ctx->h_one.assign(nb, 0.0);
    for (int b = 0; b < nb; b++) {
      double x = 0.0;
      for (int c = 0; c < nb; c++) x += colsum[c] * r.lsq_inverse[static_cast<size_t>(c) * nb + b];
      if (std::fabs(x) < 1e-12) x = 0.0;
      if (std::fabs(x - 1.0) < 1e-12) x = 1.0;
      ctx->h_one[b] = x;
    }
    if (upload(ctx, ctx->h_one, &ctx->one)) return bail(SDG_E_NOMEM);
  }
  const size_t nU = static_cast<size_t>(E) * ctx->nv * nb, nUf = static_cast<size_t>(E) * ctx->nv * naq;
  if (ctx->mixed ? (allocate(ctx, nUf, &ctx->Uf32[0]) || allocate(ctx, nUf, &ctx->Uf32[1]))
                 : (allocate(ctx, nUf, &ctx->Uf[0]) || allocate(ctx, nUf, &ctx->Uf[1])))
    return bail(SDG_E_NOMEM);
  if (allocate(ctx, nU, &ctx->U) || allocate(ctx, nU, &ctx->Un) || allocate(ctx, nU, &ctx->Rout) ||
      allocate(ctx, (static_cast<size_t>(E) + sdg::kMonitorBlocks) * ctx->nv, &ctx->mon) || allocate(ctx, static_cast<size_t>(ctx->nv), &ctx->rel) ||
      allocate(ctx, static_cast<size_t>(E), &ctx->dtbuf) || allocate(ctx, size_t(2), &ctx->status))
    return bail(SDG_E_NOMEM);
  if (ctx->ns()) {
    const size_t nGf = static_cast<size_t>(E) * ctx->ng * naq, nFv = static_cast<size_t>(E) * nq * D * ctx->nv;
    const bool tensor = ctx->launch.tensor;  // sum-factorised: Vn and the weak divergence Rvol, no Fv
    if (ctx->mixed) {
      if (allocate(ctx, nGf, &ctx->Gf32) || allocate(ctx, nUf, &ctx->Hs32) || (!tensor && allocate(ctx, nFv, &ctx->Fv32)))
        return bail(SDG_E_NOMEM);
    } else {
      if (allocate(ctx, nGf, &ctx->Gf) || allocate(ctx, nUf, &ctx->Hs)) return bail(SDG_E_NOMEM);
      if (!tensor && (ctx->launch.mfma ? allocate(ctx, nFv, &ctx->Fv) : allocate(ctx, nU, &ctx->Rvol)))
        return bail(SDG_E_NOMEM);
    }
    if (tensor && (allocate(ctx, nUf, &ctx->Vn) || allocate(ctx, nU, &ctx->Rvol))) return bail(SDG_E_NOMEM);
  }
  if (ctx->av && !ctx->ns() && !ctx->mixed &&
      (allocate(ctx, static_cast<size_t>(E) * ctx->ng * naq, &ctx->Gf) ||
       (ctx->launch.mfma ? (allocate(ctx, nUf, &ctx->Hs) || allocate(ctx, static_cast<size_t>(E) * nq * D * ctx->nv, &ctx->Fv))
                         : allocate(ctx, nU, &ctx->Rvol))))
    return bail(SDG_E_NOMEM);
  // mixed-precision Euler with artificial viscosity (matrix-core kernels): the FP32 work buffers
  if (ctx->av && !ctx->ns() && ctx->mixed &&
      (allocate(ctx, static_cast<size_t>(E) * ctx->ng * naq, &ctx->Gf32) || allocate(ctx, nUf, &ctx->Hs32) ||
       allocate(ctx, static_cast<size_t>(E) * nq * D * ctx->nv, &ctx->Fv32)))
    return bail(SDG_E_NOMEM);
  if (ctx->av && ctx->ns() && allocate(ctx, static_cast<size_t>(E) * ctx->ng * naq, &ctx->GfV)) return bail(SDG_E_NOMEM);
  // modal Boussinesq source for the kernel sets that take it from a buffer (the sum-factorised
  // kernels form it at their nodes)
  if (ctx->phys.source != SDG_SOURCE_NONE && !ctx->launch.tensor) {
    if (allocate(ctx, static_cast<size_t>(E) * ctx->nb, &ctx->Rsrc)) return bail(SDG_E_NOMEM);
    if (ctx->mixed && upload(ctx, r.mass_chol_inverse, &ctx->src_li)) return bail(SDG_E_NOMEM);
  }
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ctx->err = "sdg_create: device initialisation failed";
    return bail(SDG_E_DEVICE);
  }
  *out = ctx;
  return SDG_OK;
}

void sdg_destroy(sdg_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  destroyGraphs(ctx);
  for (void* p : ctx->allocations) (void)hipFree(p);
  if (ctx->adj_pool) (void)hipFree(ctx->adj_pool);
  if (ctx->adj_tpool) (void)hipFree(ctx->adj_tpool);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int sdg_set_state(sdg_ctx* ctx, const double* coefficients) {
  if (!ctx || !coefficients) return fail(ctx, SDG_E_INVALID, "sdg_set_state: null argument");
  std::vector<double> dev;
  toDevice(ctx, coefficients, dev);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->U, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->Un, ctx->U, dev.size() * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  ctx->cur = 0;
  updateReference(ctx, coefficients);
  ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->traces(0), ctx->mix);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  destroyGraphs(ctx);
  ctx->has_state = true;
  return SDG_OK;
}

int sdg_get_state(sdg_ctx* ctx, double* coefficients) {
  if (!ctx || !coefficients) return fail(ctx, SDG_E_INVALID, "sdg_get_state: null argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_get_state: no state set");
  std::vector<double> dev(static_cast<size_t>(ctx->E) * ctx->nv * ctx->nb);
  HIP_TRY(ctx, hipMemcpyAsync(dev.data(), ctx->U, dev.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  toReference(ctx, dev, coefficients);
  return SDG_OK;
}

int sdg_set_initial_primitive(sdg_ctx* ctx, const double* prim) {
  if (!ctx || !prim) return fail(ctx, SDG_E_INVALID, "sdg_set_initial_primitive: null argument");
  const int nv = ctx->nv, nq = ctx->nq, nb = ctx->nb, D = ctx->dim;
  std::vector<double> coeff(static_cast<size_t>(ctx->E) * nb * nv);
  std::vector<double> uq(static_cast<size_t>(nv) * nq), t(static_cast<size_t>(nv) * nb);
  for (int64_t e = 0; e < ctx->E; e++) {
    // VariableConvertor.cpp:341-366 then InitialCondition.cpp:99-101
    for (int q = 0; q < nq; q++) {
      const double* pr = &prim[(e * nq + q) * nv];
      const double rho = pr[0];
      double v2 = 0.0;
      uq[static_cast<size_t>(q) * nv] = rho;
      for (int d = 0; d < D; d++) {
        uq[static_cast<size_t>(q) * nv + 1 + d] = rho * pr[1 + d];
        v2 += pr[1 + d] * pr[1 + d];
      }
      uq[static_cast<size_t>(q) * nv + D + 1] = rho * (ctx->phys.cv * pr[D + 1] + v2 / 2.0);
    }
    for (int v = 0; v < nv; v++)
      for (int b = 0; b < nb; b++) {
        double s = 0.0;
        for (int q = 0; q < nq; q++) s += uq[static_cast<size_t>(q) * nv + v] * ctx->h_V[static_cast<size_t>(q) * nb + b];
        t[static_cast<size_t>(v) * nb + b] = s;
      }
    for (int v = 0; v < nv; v++)
      for (int b = 0; b < nb; b++) {
        double s = 0.0;
        for (int c = 0; c < nb; c++) s += t[static_cast<size_t>(v) * nb + c] * ctx->h_lsq[static_cast<size_t>(c) * nb + b];
        coeff[(e * nb + b) * nv + v] = s;
      }
  }
  return sdg_set_state(ctx, coeff.data());
}

int sdg_set_boundary_primitive(sdg_ctx* ctx, const double* prim) {
  if (!ctx || (!prim && ctx->Fb > 0)) return fail(ctx, SDG_E_INVALID, "sdg_set_boundary_primitive: null argument");
  const int D = ctx->dim, nv = ctx->nv, nfq = ctx->nfq;
  const size_t npts = static_cast<size_t>(ctx->Fb) * nfq;
  std::vector<double> w(npts * (D + 3));
  const sdg::Phys& ph = ctx->phys;
  for (size_t i = 0; i < npts; i++) {  // primitive -> computational (VariableConvertor.cpp:118-139,341-366)
    const double* pr = prim + i * nv;
    double* o = &w[i * (D + 3)];
    o[0] = pr[0];
    for (int d = 0; d < D; d++) o[1 + d] = pr[1 + d];
    o[D + 1] = ph.cv * pr[D + 1];
    o[D + 2] = ph.eos == SDG_EOS_IDEAL_GAS ? (sdg::kGamma - 1.0) * pr[0] * o[D + 1]
                                           : ph.c0 * ph.c0 * (pr[0] - ph.rho0) + ph.padd;
  }
  if (npts) {
    HIP_TRY(ctx, hipMemcpyAsync(ctx->bdata, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  ctx->has_bdata = true;
  return SDG_OK;
}

int sdg_compute_dt(sdg_ctx* ctx, double cfl, double* dt) {
  if (!ctx || !dt) return fail(ctx, SDG_E_INVALID, "sdg_compute_dt: null argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_compute_dt: no state set");
  ctx->launch.dt(ctx->stream, ctx->geo, ctx->phys, ctx->U, ctx->minedge, cfl, ctx->order, ctx->dtbuf);
  HIP_TRY(ctx, hipGetLastError());
  std::vector<double> h(ctx->E);
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), ctx->dtbuf, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  double best = h.empty() ? 0.0 : h[0];
  for (double v : h) best = std::min(best, v);
  *dt = best;
  return SDG_OK;
}

int sdg_compute_dt_noalloc(sdg_ctx* ctx, double cfl, double* dt) {
  if (!ctx || !dt) return fail(ctx, SDG_E_INVALID, "sdg_compute_dt_noalloc: null argument");
  if (ctx->E > kNoAllocMaxElements)
    return fail(ctx, SDG_E_INVALID, "sdg_compute_dt_noalloc: more elements than the scratch holds (2^21)");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_compute_dt_noalloc: no state set");
  if (!ctx->scratch) return fail(ctx, SDG_E_DEVICE, "sdg_compute_dt_noalloc: " + ctx->scratch_err);
  double* ptr = ctx->scratch;
  ctx->launch.dt(ctx->stream, ctx->geo, ctx->phys, ctx->U, ctx->minedge, cfl, ctx->order, ptr);
  hipLaunchKernelGGL(minReduceKernel, dim3(1), dim3(256), 0, ctx->stream, ptr, static_cast<int64_t>(ctx->E),
                     ptr + kNoAllocMaxElements);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(dt, ptr + kNoAllocMaxElements, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SDG_OK;
}

int sdg_residual(sdg_ctx* ctx, double* residual) {
  if (!ctx || !residual) return fail(ctx, SDG_E_INVALID, "sdg_residual: null argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_residual: no state set");
  if (ctx->Fb > 0 && !ctx->has_bdata)
    return fail(ctx, SDG_E_STATE, "sdg_residual: boundary data not set (sdg_set_boundary_primitive)");
  enqueueArtificialViscosity(ctx);
  enqueueStage(ctx, -1, 0.0, 1, nullptr);
  HIP_TRY(ctx, hipGetLastError());
  std::vector<double> dev(static_cast<size_t>(ctx->E) * ctx->nv * ctx->nb);
  HIP_TRY(ctx, hipMemcpyAsync(dev.data(), ctx->Rout, dev.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  // the dense mixed-precision kernels form R_psi = L^-1 R; return R = L R_psi (the sum-factorised
  // kernels write R itself)
  if (ctx->mixed && !ctx->launch.tensor) {
    const int nb = ctx->nb;
    std::vector<double> t(nb);
    for (size_t o = 0; o < dev.size(); o += nb) {
      for (int i = 0; i < nb; i++) {
        double acc = 0.0;
        for (int j = 0; j <= i; j++) acc += ctx->h_chol[static_cast<size_t>(i) * nb + j] * dev[o + j];
        t[i] = acc;
      }
      std::copy(t.begin(), t.end(), dev.begin() + static_cast<std::ptrdiff_t>(o));
    }
  }
  toReference(ctx, dev, residual);
  return SDG_OK;
}

int sdg_get_gradient(sdg_ctx* ctx, double* gradient) {
  if (!ctx || !gradient) return fail(ctx, SDG_E_INVALID, "sdg_get_gradient: null argument");
  if (!ctx->ns()) return fail(ctx, SDG_E_UNSUPPORTED, "sdg_get_gradient: only defined for Navier-Stokes");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_get_gradient: no state set");
  if (ctx->Fb > 0 && !ctx->has_bdata)
    return fail(ctx, SDG_E_STATE, "sdg_get_gradient: boundary data not set (sdg_set_boundary_primitive)");
  if (!ctx->Wout && allocate(ctx, static_cast<size_t>(ctx->E) * ctx->ng * ctx->nb, &ctx->Wout)) return SDG_E_NOMEM;
  StageArgs a = baseArgs(ctx);
  a.mode = 1;
  a.status = nullptr;
  a.Wout = ctx->Wout;
  if (ctx->prob.viscous_flux == SDG_VISC_BR1)
    ctx->launch.ns_grad_br1(ctx->stream, ctx->geo, ctx->phys, a);
  else
    ctx->launch.ns_grad_br2(ctx->stream, ctx->geo, ctx->phys, a);
  HIP_TRY(ctx, hipGetLastError());
  const size_t n = static_cast<size_t>(ctx->E) * ctx->ng * ctx->nb;
  HIP_TRY(ctx, hipMemcpyAsync(gradient, ctx->Wout, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SDG_OK;
}

int sdg_step(sdg_ctx* ctx, double dt, int nsteps, double* relative_error, int* nan_flag, int* steps_taken) {
  if (!ctx) return fail(ctx, SDG_E_INVALID, "sdg_step: null context");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_step: no state set");
  if (ctx->Fb > 0 && !ctx->has_bdata)
    return fail(ctx, SDG_E_STATE, "sdg_step: boundary data not set (sdg_set_boundary_primitive)");
  if (nsteps < 0 || !(dt >= 0.0)) return fail(ctx, SDG_E_INVALID, "sdg_step: bad dt or step count");
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  int rc = launchSteps(ctx, dt, nsteps);
  if (rc) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  int st[2] = {0, 0};
  std::vector<double> rel(ctx->nv, 0.0);
  HIP_TRY(ctx, hipMemcpyAsync(st, ctx->status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(rel.data(), ctx->rel, rel.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
  // A NaN stop leaves the later launches as no-ops; the trace buffer parity then no longer follows
  // the host counter, so recompute the traces from the stopped state.
  if (st[0] && st[1] < nsteps) {
    ctx->cur = 0;
    ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->traces(0), ctx->mix);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (relative_error) std::memcpy(relative_error, rel.data(), rel.size() * sizeof(double));
  if (nan_flag) *nan_flag = st[0];
  if (steps_taken) *steps_taken = st[1];
  return SDG_OK;
}

int sdg_last_step_ms(const sdg_ctx* ctx, double* ms) {
  if (!ctx || !ms) return SDG_E_INVALID;
  *ms = ctx->last_ms;
  return SDG_OK;
}

int sdg_synchronize(sdg_ctx* ctx) {
  if (!ctx) return SDG_E_INVALID;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SDG_OK;
}

int sdg_time_kernels(sdg_ctx* ctx, double dt, int nsteps, double* avg_ms, int n_out) {
  if (!ctx || !avg_ms || n_out < 1) return fail(ctx, SDG_E_INVALID, "sdg_time_kernels: bad argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_time_kernels: no state set");
  if (ctx->Fb > 0 && !ctx->has_bdata)
    return fail(ctx, SDG_E_STATE, "sdg_time_kernels: boundary data not set (sdg_set_boundary_primitive)");
  if (nsteps < 1 || !(dt >= 0.0)) return fail(ctx, SDG_E_INVALID, "sdg_time_kernels: bad dt or step count");
  const int st = ctx->stages();
  const size_t per_step = static_cast<size_t>(6 * st + 2);
  std::vector<hipEvent_t> ev(per_step * nsteps, nullptr);
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
      return fail(ctx, SDG_E_DEVICE, "sdg_time_kernels: event creation failed");
    }
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  for (int i = 0; i < nsteps; i++) {
    hipEvent_t* e = &ev[per_step * i];
    enqueueArtificialViscosity(ctx);
    for (int s = 0; s < st; s++) enqueueStage(ctx, s, dt, 0, nullptr, e + 6 * s);
    (void)hipEventRecord(e[6 * st], ctx->stream);
    ctx->launch.monitor(ctx->stream, ctx->mon, ctx->E, ctx->rel, ctx->status);
    (void)hipEventRecord(e[6 * st + 1], ctx->stream);
  }
  hipError_t err = hipStreamSynchronize(ctx->stream);
  double sum[SDG_KERNEL_CLASSES] = {0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < nsteps && err == hipSuccess; i++) {
    hipEvent_t* e = &ev[per_step * i];
    for (int s = 0; s < st; s++)
      for (int k = 0; k < 3; k++) {
        float x = 0.0f;
        (void)hipEventElapsedTime(&x, e[6 * s + 2 * k], e[6 * s + 2 * k + 1]);
        sum[k] += x;
      }
    float m = 0.0f;
    (void)hipEventElapsedTime(&m, e[6 * st], e[6 * st + 1]);
    sum[3] += m;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  if (err != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("sdg_time_kernels: ") + hipGetErrorString(err));
  const double launches[SDG_KERNEL_CLASSES] = {double(nsteps) * st, double(nsteps) * st, double(nsteps) * st,
                                               double(nsteps)};
  for (int k = 0; k < n_out; k++) avg_ms[k] = k < SDG_KERNEL_CLASSES ? sum[k] / launches[k] : 0.0;
  if (!ctx->ns() && !ctx->av)
    for (int k = 1; k < n_out && k < 3; k++) avg_ms[k] = 0.0;
  return SDG_OK;
}

int sdg_get_artificial_viscosity(sdg_ctx* ctx, double* node_values) {
  if (!ctx || !node_values) return fail(ctx, SDG_E_INVALID, "sdg_get_artificial_viscosity: null argument");
  if (!ctx->av) return fail(ctx, SDG_E_UNSUPPORTED, "sdg_get_artificial_viscosity: shock capturing is off");
  HIP_TRY(ctx, hipMemcpyAsync(node_values, ctx->nodeav, static_cast<size_t>(ctx->nnode) * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SDG_OK;
}

}  // extern "C"

// ---- raw state files ------------------------------------------------------------------------------
extern "C" {

int sdg_raw_binary_write(sdg_ctx* ctx, const char* path, int64_t node_count) {
  if (!ctx || !path || node_count < 0) return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_write: bad argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_raw_binary_write: no state set");
  if (ctx->av && node_count != ctx->nnode)
    return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_write: with artificial viscosity node_count must be the mesh's node count");
  const bool ns = ctx->ns();
  const size_t E = static_cast<size_t>(ctx->E), nu = static_cast<size_t>(ctx->nv) * ctx->nb,
               ng = ns ? static_cast<size_t>(ctx->ng) * ctx->nb : 0;
  std::vector<double> U(E * nu), G(E * ng);
  int rc = sdg_get_state(ctx, U.data());
  if (rc) return rc;
  if (ns && (rc = sdg_get_gradient(ctx, G.data()))) return rc;
  // BR2 with boundary faces: each boundary face's block is its parent's volume gradient plus that
  // face's lift (RawBinary.cpp:99-107), formed by a kernel of its own
  const bool br2_faces = ns && ctx->prob.viscous_flux == SDG_VISC_BR2 && ctx->Fb > 0;
  std::vector<double> GB;
  if (br2_faces) {
    const size_t nbd = static_cast<size_t>(ctx->Fb) * ng;
    if (!ctx->Wbnd && allocate(ctx, nbd, &ctx->Wbnd)) return SDG_E_NOMEM;
    ctx->launch.br2_blocks(ctx->stream, ctx->geo, ctx->phys, ctx->U, ctx->Wbnd);
    HIP_TRY(ctx, hipGetLastError());
    GB.resize(nbd);
    HIP_TRY(ctx, hipMemcpyAsync(GB.data(), ctx->Wbnd, nbd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  std::vector<double> payload;
  payload.reserve((E + ctx->h_bparent.size()) * (nu + ng) + static_cast<size_t>(node_count));
  auto block = [&](size_t e) {
    payload.insert(payload.end(), U.begin() + static_cast<std::ptrdiff_t>(e * nu),
                   U.begin() + static_cast<std::ptrdiff_t>((e + 1) * nu));
    if (ns)
      payload.insert(payload.end(), G.begin() + static_cast<std::ptrdiff_t>(e * ng),
                     G.begin() + static_cast<std::ptrdiff_t>((e + 1) * ng));
  };
  for (size_t e = 0; e < E; e++) block(e);  // RawBinary.cpp:76-90
  for (size_t bf = 0; bf < ctx->h_bparent.size(); bf++) {  // RawBinary.cpp:95-123
    const size_t p = static_cast<size_t>(ctx->h_bparent[bf]);
    if (!br2_faces) {
      block(p);
      continue;
    }
    payload.insert(payload.end(), U.begin() + static_cast<std::ptrdiff_t>(p * nu),
                   U.begin() + static_cast<std::ptrdiff_t>((p + 1) * nu));
    payload.insert(payload.end(), GB.begin() + static_cast<std::ptrdiff_t>(bf * ng),
                   GB.begin() + static_cast<std::ptrdiff_t>((bf + 1) * ng));
  }
  payload.resize(payload.size() + static_cast<size_t>(node_count), 0.0);  // node artificial viscosity
  if (ctx->av && node_count > 0 &&
      (rc = sdg_get_artificial_viscosity(ctx, payload.data() + payload.size() - static_cast<size_t>(node_count))))
    return rc;
  rc = sdg::writeRawFile(path, payload.data(), payload.size() * sizeof(double));
  if (rc) return fail(ctx, rc, sdg::globalError());
  return SDG_OK;
}

int sdg_raw_binary_read(sdg_ctx* ctx, const char* path, int initial_condition) {
  if (!ctx || !path) return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_read: null argument");
  if (initial_condition != SDG_IC_LAST_STEP && initial_condition != SDG_IC_SPECIFIC_FILE)
    return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_read: initial_condition must be LAST_STEP or SPECIFIC_FILE");
  int nb_src = ctx->nb;
  if (initial_condition == SDG_IC_SPECIFIC_FILE) {
    if (ctx->order < 2) return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_read: SPECIFIC_FILE needs P >= 2");
    nb_src = sdg::basisSize(ctx->type, ctx->dim, ctx->order - 1);  // basis of order P - 1
  }
  std::vector<char> raw;
  int rc = sdg::readRawFile(path, raw);
  if (rc) return fail(ctx, rc, sdg::globalError());
  const size_t E = static_cast<size_t>(ctx->E), nv = static_cast<size_t>(ctx->nv);
  const size_t blk = nv * nb_src + (ctx->ns() ? static_cast<size_t>(ctx->ng) * nb_src : 0);
  if (raw.size() < E * blk * sizeof(double))
    return fail(ctx, SDG_E_INVALID, "sdg_raw_binary_read: the file holds fewer element blocks than the mesh has elements");
  const double* src = reinterpret_cast<const double*>(raw.data());
  std::vector<double> coeff(E * nv * ctx->nb, 0.0);
  for (size_t e = 0; e < E; e++)
    std::copy(src + e * blk, src + e * blk + nv * nb_src, coeff.begin() + static_cast<std::ptrdiff_t>(e * nv * ctx->nb));
  return sdg_set_state(ctx, coeff.data());
}

}  // extern "C"

// ---- discrete adjoint ---------------------------------------------------------------------------
namespace {

int ensureAdjoint(sdg_ctx* ctx) {
  if (ctx->adj_ready) return SDG_OK;
  if (!ctx->ns() || (ctx->mixed && !ctx->launch.tensor) || !ctx->launch.adj_volume[0] || ctx->Fb > 0 ||
      ctx->phys.source != SDG_SOURCE_NONE || ctx->av)
    return fail(ctx, SDG_E_UNSUPPORTED,
                "adjoint: implemented for Navier-Stokes on periodic meshes of the matrix-core element sizes "
                "(quadrangle P3-P5, hexahedron P2-P4; mixed precision on hexahedra P2-P4), without a source term or "
                "artificial viscosity");
  const size_t E = static_cast<size_t>(ctx->E);
  const size_t nU = E * ctx->nv * ctx->nb, nS = E * ctx->nv * ctx->naq, nG = E * ctx->ng * ctx->nb,
               nGS = E * ctx->ng * ctx->naq;
  if (allocate(ctx, nU, &ctx->adj_lam) || allocate(ctx, nS, &ctx->adj_Hb) || allocate(ctx, nG, &ctx->adj_RtV) ||
      allocate(ctx, nU, &ctx->adj_Ua) || allocate(ctx, nS, &ctx->adj_UFf) || allocate(ctx, nGS, &ctx->adj_GFb) ||
      allocate(ctx, nS, &ctx->adj_Pown) || allocate(ctx, nS, &ctx->adj_Qnb) || allocate(ctx, nU, &ctx->adj_out) ||
      allocate(ctx, 3 * nU, &ctx->adj_S[0]) || allocate(ctx, nU, &ctx->adj_acc) ||
      allocate(ctx, nU, &ctx->adj_tmp))
    return SDG_E_NOMEM;
  if (ctx->mixed && (allocate(ctx, nS, &ctx->adj_Uf) || allocate(ctx, nGS, &ctx->adj_Gf))) return SDG_E_NOMEM;
  ctx->adj_S[1] = ctx->adj_S[0] + nU;  // stage states, one contiguous block
  ctx->adj_S[2] = ctx->adj_S[0] + 2 * nU;
  ctx->adj_ready = true;
  return SDG_OK;
}

size_t stateCount(const sdg_ctx* ctx) { return static_cast<size_t>(ctx->E) * ctx->nv * ctx->nb; }

// out = (dR/dU)^T mu at state U (device layout), mu = lam, or Mref^-1 lam / |J| with apply_minv
// With stage >= 0 the gather kernel applies the transpose of RK stage `stage` to adj_lam in place
// (accumulating into adj_acc) instead of writing `out`.
int vjpAt(sdg_ctx* ctx, const double* U, double* lam, int apply_minv, double* out, int stage = -1,
          const double* traces = nullptr) {
  double* Uf = const_cast<double*>(traces);
  if (ctx->mixed) {  // FP64 traces of the FP64 state (the run's own traces are FP32 deviations)
    Uf = ctx->adj_Uf;
    ctx->launch.trace64(ctx->stream, ctx->geo, U, Uf, ctx->mix);
  } else if (!Uf) {
    Uf = ctx->Uf[ctx->cur ^ 1];  // scratch: the next stage rewrites this trace buffer
    ctx->launch.trace(ctx->stream, ctx->geo, U, Uf, ctx->mix);
  }
  StageArgs a = baseArgs(ctx);
  a.U = const_cast<double*>(U);
  a.Uf_in = Uf;
  if (ctx->mixed) a.Gf = ctx->adj_Gf;
  a.Wout = nullptr;
  a.status = nullptr;
  a.mode = 1;
  sdg::AdjArgs j{};
  j.op = ctx->adj_ops;
  j.lam = lam;
  j.apply_minv = apply_minv;
  j.Hb = ctx->adj_Hb;
  j.RtV = ctx->adj_RtV;
  j.Ua = ctx->adj_Ua;
  j.UFf = ctx->adj_UFf;
  j.GFb = ctx->adj_GFb;
  j.Pown = ctx->adj_Pown;
  j.Qnb = ctx->adj_Qnb;
  j.out = out;
  if (stage >= 0) {
    double c[3];
    ctx->rkCoefficients(stage, c);
    j.rk = 1;
    j.rk_first = stage == ctx->stages() - 1;
    j.rk_last = stage == 0;
    j.acc = ctx->adj_acc;
    j.c0 = c[0];
    j.c1 = c[1];
    j.c2dt = c[2] * ctx->adj_dt;
  }
  const int vi = ctx->prob.viscous_flux == SDG_VISC_BR1 ? 0 : 1;
  ctx->launch.adj_volume[vi](ctx->stream, ctx->geo, ctx->phys, a, j);
  ctx->launch.adj_face(ctx->stream, ctx->geo, ctx->phys, a, j);
  ctx->launch.adj_lift[vi](ctx->stream, ctx->geo, ctx->phys, a, j);
  ctx->launch.adj_gather(ctx->stream, ctx->geo, ctx->phys, a, j);
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// state <- src (device layout), with its traces
int loadState(sdg_ctx* ctx, const double* src) {
  const size_t bytes = stateCount(ctx) * sizeof(double);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->U, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->Un, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->traces(ctx->cur), ctx->mix);
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// k forward steps, enqueued without waiting; the NaN/step-count status is read by the caller once a
// whole run is enqueued (recomputations repeat steps the recorded run already took).
int advance(sdg_ctx* ctx, int k) {
  if (k <= 0) return SDG_OK;
  return launchSteps(ctx, ctx->adj_dt, k);
}

// forward steps replayed inside the reverse sweep (from a checkpoint the recorded run passed)
int recompute(sdg_ctx* ctx, int k) {
  if (k <= 0) return SDG_OK;
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  return launchSteps(ctx, ctx->adj_dt, k);
}

int saveState(sdg_ctx* ctx, double* dst) {
  HIP_TRY(ctx, hipMemcpyAsync(dst, ctx->U, stateCount(ctx) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return SDG_OK;
}

// adj_lam: adjoint of U_{n+1} -> adjoint of U_n, for the step that starts from U0 (device layout).
// Stage s maps U_s to U_{s+1} = c0 U_0 + c1 U_s + c2 dt L(U_s) (TimeIntegration.cpp:45-65,181-198),
// so its transpose is lam_s = c1 lam_{s+1} + c2 dt L'(U_s)^T lam_{s+1}, plus c0 lam_{s+1} to U_0.
// The body of one adjoint step, starting from the state in adj_S[0]: recompute the stage states,
// then transpose the stages in reverse order. The state buffers and trace parity it leaves behind
// are those of S - 1 forward stages. No memset (the body is captured in a graph).
int enqueueAdjointStep(sdg_ctx* ctx) {
  const int S = ctx->stages();
  int rc = loadState(ctx, ctx->adj_S[0]);
  if (rc) return rc;
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  for (int s = 0; s + 1 < S; s++) {
    enqueueStage(ctx, s, ctx->adj_dt, 0, nullptr);
    if ((rc = saveState(ctx, ctx->adj_S[s + 1]))) return rc;
  }
  for (int s = S - 1; s >= 0; s--)
    if ((rc = vjpAt(ctx, ctx->adj_S[s], ctx->adj_lam, 1, nullptr, s))) return rc;
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// Capture (once per trace parity and dt) and replay a graph built by `body`; the body leaves the
// trace parity flipped by `flips`.
template <typename Body>
int replayGraph(sdg_ctx* ctx, hipGraphExec_t* graphs, double* graph_dt, double dt, int flips, Body body) {
  const int c = ctx->cur;
  if (!graphs[c] || graph_dt[c] != dt) {
    if (graphs[c]) (void)hipGraphExecDestroy(graphs[c]);
    graphs[c] = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t captured = nullptr;
    const hipError_t ee = hipStreamEndCapture(ctx->stream, &captured);
    if (rc) {
      if (captured) (void)hipGraphDestroy(captured);
      ctx->cur = c;
      return rc;
    }
    if (ee != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
    const hipError_t e = hipGraphInstantiate(&graphs[c], captured, nullptr, nullptr, 0);
    (void)hipGraphDestroy(captured);
    if (e != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    graph_dt[c] = dt;
    ctx->cur = c;  // capture advanced cur; the replay below advances it for real
  }
  HIP_TRY(ctx, hipGraphLaunch(graphs[c], ctx->stream));
  ctx->cur = c ^ flips;
  return SDG_OK;
}

int stepAdjoint(sdg_ctx* ctx, const double* U0) {
  HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_S[0], U0, stateCount(ctx) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  if (!ctx->use_graph) return enqueueAdjointStep(ctx);
  return replayGraph(ctx, ctx->adj_graph, ctx->adj_graph_dt, ctx->adj_dt, (ctx->stages() - 1) & 1,
                     [&]() { return enqueueAdjointStep(ctx); });
}

// One adjoint step from the stage states in adj_S[0 .. S-1] (no recomputation).
int enqueueAdjointStepStored(sdg_ctx* ctx) {
  const int S = ctx->stages();
  const size_t nt = static_cast<size_t>(ctx->E) * ctx->nv * ctx->naq;
  int rc;
  for (int s = S - 1; s >= 0; s--)
    if ((rc = vjpAt(ctx, ctx->adj_S[s], ctx->adj_lam, 1, nullptr, s, ctx->mixed ? nullptr : ctx->adj_T + s * nt)))
      return rc;
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// One recorded forward step: every stage's updated state also goes to the pool slot of the next
// stage input, the step index being read on the device from the status counter.
int enqueueRecordStep(sdg_ctx* ctx, double dt, int nsteps) {
  ctx->recording = true;
  ctx->rec_nsteps = nsteps;
  for (int s = 0; s < ctx->stages(); s++) enqueueStage(ctx, s, dt, 0, nullptr);
  ctx->recording = false;
  ctx->launch.monitor(ctx->stream, ctx->mon, ctx->E, ctx->rel, ctx->status);
  HIP_TRY(ctx, hipGetLastError());
  return SDG_OK;
}

// first checkpoint position of a binomial schedule for n steps with s free snapshots: with r the
// least repetition count such that C(s + r, s) >= n, the right part n - m must be reversible with
// s - 1 snapshots (n - m <= C(s - 1 + r, s - 1)) and the left part with r - 1 repetitions
double binom(int n, int k) {
  double b = 1.0;
  for (int i = 1; i <= k; i++) b = b * (n - k + i) / i;
  return b;
}
int splitBinomial(int n, int s) {
  int r = 0;
  while (binom(s + r, s) < n) r++;
  int m = n - static_cast<int>(binom(s - 1 + r, s - 1));
  if (m < 1) m = 1;
  if (m > n - 1) m = n - 1;
  return m;
}

// Reverse steps [i0, i0 + n): snapshot `slot` holds U_{i0}, slots slot+1 .. slot+s are free (or
// hold the states the forward run placed there); adj_lam holds the adjoint of U_{i0+n}.
int reverseSweep(sdg_ctx* ctx, int i0, int n, int s, int slot) {
  int rc;
  if (n <= 0) return SDG_OK;
  if (n == 1) return stepAdjoint(ctx, ctx->snaps[slot]);
  if (s == 0) {
    for (int jj = n - 1; jj >= 0; jj--) {
      if ((rc = loadState(ctx, ctx->snaps[slot])) || (rc = recompute(ctx, jj)) || (rc = saveState(ctx, ctx->adj_tmp)) ||
          (rc = stepAdjoint(ctx, ctx->adj_tmp)))
        return rc;
    }
    return SDG_OK;
  }
  const int m = splitBinomial(n, s);
  if (ctx->snap_step[slot + 1] != i0 + m) {
    if ((rc = loadState(ctx, ctx->snaps[slot])) || (rc = recompute(ctx, m)) || (rc = saveState(ctx, ctx->snaps[slot + 1])))
      return rc;
    ctx->snap_step[slot + 1] = i0 + m;
  }
  if ((rc = reverseSweep(ctx, i0 + m, n - m, s - 1, slot + 1))) return rc;
  ctx->snap_step[slot + 1] = -1;
  return reverseSweep(ctx, i0, m, s, slot);
}

int reverseRecorded(sdg_ctx* ctx);

// state size (doubles) from which the stored reverse sweep launches its kernels on the recording
// pools directly instead of replaying a graph over copies of each step's states
constexpr size_t kDirectAdjointStates = size_t(1) << 22;

}  // namespace

extern "C" {

int sdg_vjp_residual(sdg_ctx* ctx, const double* lambda, double* out) {
  if (!ctx || !lambda || !out) return fail(ctx, SDG_E_INVALID, "sdg_vjp_residual: null argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_vjp_residual: no state set");
  int rc = ensureAdjoint(ctx);
  if (rc) return rc;
  std::vector<double> dev;
  toDevice(ctx, lambda, dev);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_lam, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = vjpAt(ctx, ctx->U, ctx->adj_lam, 0, ctx->adj_out))) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(dev.data(), ctx->adj_out, dev.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  toReference(ctx, dev, out);
  return SDG_OK;
}

int sdg_adjoint_forward(sdg_ctx* ctx, double dt, int nsteps, int snapshots) {
  if (!ctx) return fail(ctx, SDG_E_INVALID, "sdg_adjoint_forward: null context");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_adjoint_forward: no state set");
  if (nsteps < 0 || snapshots < 1 || !(dt >= 0.0)) return fail(ctx, SDG_E_INVALID, "sdg_adjoint_forward: bad argument");
  int rc = ensureAdjoint(ctx);
  if (rc) return rc;
  ctx->adj_stored = false;
  if (snapshots >= nsteps && nsteps > 0 && ctx->use_graph) {
    // every step fits: record every stage state instead of recomputing them in the reverse sweep
    const size_t need = static_cast<size_t>(nsteps) * ctx->stages() * stateCount(ctx);
    if (need > ctx->adj_pool_cap) {
      if (ctx->adj_pool) (void)hipFree(ctx->adj_pool);
      ctx->adj_pool = nullptr;
      ctx->adj_pool_cap = 0;
      void* p = nullptr;
      if (hipMalloc(&p, need * sizeof(double)) == hipSuccess) {
        ctx->adj_pool = static_cast<double*>(p);
        ctx->adj_pool_cap = need;
      } else {
        (void)hipGetLastError();  // out of device memory: fall back to the checkpointed schedule
      }
    }
    const size_t nt = static_cast<size_t>(ctx->E) * ctx->nv * ctx->naq;
    // FP64 mode: the stage states' traces are kept too; a mixed-precision run's are FP32 deviations,
    // and its adjoint recomputes FP64 traces from the states
    const bool keep_t = !ctx->mixed;
    const size_t need_t = keep_t ? static_cast<size_t>(nsteps) * ctx->stages() * nt : 0;
    if (keep_t && ctx->adj_pool && need_t > ctx->adj_tpool_cap) {
      if (ctx->adj_tpool) (void)hipFree(ctx->adj_tpool);
      ctx->adj_tpool = nullptr;
      ctx->adj_tpool_cap = 0;
      void* p = nullptr;
      if (hipMalloc(&p, need_t * sizeof(double)) == hipSuccess) {
        ctx->adj_tpool = static_cast<double*>(p);
        ctx->adj_tpool_cap = need_t;
      } else {
        (void)hipGetLastError();
      }
    }
    if (keep_t && ctx->adj_pool && ctx->adj_tpool && !ctx->adj_T && allocate(ctx, 3 * nt, &ctx->adj_T))
      return SDG_E_NOMEM;
    if (ctx->adj_pool && need <= ctx->adj_pool_cap && (!keep_t || (ctx->adj_tpool && need_t <= ctx->adj_tpool_cap))) {
      ctx->adj_dt = dt;
      ctx->adj_nsteps = -1;
      hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
      HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_pool, ctx->U, stateCount(ctx) * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));  // slot (0, 0): the initial state and its traces
      if (keep_t)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_tpool, ctx->traces(ctx->cur), nt * sizeof(double),
                                    hipMemcpyDeviceToDevice, ctx->stream));
      const int flips = ctx->stages() & 1;
      for (int i = 0; i < nsteps; i++) {
        // the graph bakes in the pool's step count; rebuild it when that changes
        if (ctx->rec_graph_n[ctx->cur] != nsteps) ctx->rec_graph_dt[ctx->cur] = -1.0;
        if ((rc = replayGraph(ctx, ctx->rec_graph, ctx->rec_graph_dt, dt, flips,
                              [&]() { return enqueueRecordStep(ctx, dt, nsteps); })))
          return rc;
        ctx->rec_graph_n[ctx->cur ^ flips] = nsteps;
      }
      int st[2] = {0, 0};
      HIP_TRY(ctx, hipMemcpyAsync(st, ctx->status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      if (st[0] || st[1] != nsteps) {
        ctx->cur = 0;
        ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->traces(0), ctx->mix);
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return fail(ctx, SDG_E_STATE, "sdg_adjoint_forward: the residual monitor turned NaN during the recorded run");
      }
      ctx->adj_stored = true;
      ctx->adj_nsteps = nsteps;
      ctx->adj_nslots = nsteps;
      return SDG_OK;
    }
  }
  const int nslots = std::max(1, std::min(snapshots, nsteps));
  while (static_cast<int>(ctx->snaps.size()) < nslots) {
    double* p = nullptr;
    if (allocate(ctx, stateCount(ctx), &p)) return SDG_E_NOMEM;
    ctx->snaps.push_back(p);
  }
  ctx->snap_step.assign(ctx->snaps.size(), -1);
  ctx->adj_dt = dt;
  ctx->adj_nsteps = -1;
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  if ((rc = saveState(ctx, ctx->snaps[0]))) return rc;
  ctx->snap_step[0] = 0;
  // the forward run places the checkpoints of the schedule's first descent
  int i0 = 0, n = nsteps, s = nslots - 1, slot = 0;
  while (n > 1 && s > 0) {
    const int m = splitBinomial(n, s);
    if ((rc = advance(ctx, m))) return rc;
    slot++;
    if ((rc = saveState(ctx, ctx->snaps[slot]))) return rc;
    ctx->snap_step[slot] = i0 + m;
    i0 += m;
    n -= m;
    s--;
  }
  if ((rc = advance(ctx, n))) return rc;
  int st[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(st, ctx->status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (st[0] || st[1] != nsteps) {
    // a NaN stop leaves the later launches as no-ops; rebuild the traces of the stopped state
    ctx->cur = 0;
    ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->traces(0), ctx->mix);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return fail(ctx, SDG_E_STATE, "sdg_adjoint_forward: the residual monitor turned NaN during the recorded run");
  }
  ctx->adj_nsteps = nsteps;
  ctx->adj_nslots = nslots;
  return SDG_OK;
}

int sdg_adjoint_backward(sdg_ctx* ctx, const double* lambda_final, double* lambda_initial) {
  if (!ctx || !lambda_final || !lambda_initial) return fail(ctx, SDG_E_INVALID, "sdg_adjoint_backward: null argument");
  if (ctx->adj_nsteps < 0) return fail(ctx, SDG_E_STATE, "sdg_adjoint_backward: no recorded sdg_adjoint_forward run");
  std::vector<double> dev;
  toDevice(ctx, lambda_final, dev);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_lam, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  int rc = reverseRecorded(ctx);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(dev.data(), ctx->adj_lam, dev.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if ((rc = loadState(ctx, ctx->adj_stored ? ctx->adj_pool : ctx->snaps[0]))) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  toReference(ctx, dev, lambda_initial);
  return SDG_OK;
}

int sdg_adjoint_backward_device(sdg_ctx* ctx, const double* lambda_final, double* lambda_initial) {
  if (!ctx || !lambda_final || !lambda_initial)
    return fail(ctx, SDG_E_INVALID, "sdg_adjoint_backward_device: null argument");
  if (ctx->adj_nsteps < 0)
    return fail(ctx, SDG_E_STATE, "sdg_adjoint_backward_device: no recorded sdg_adjoint_forward run");
  hipPointerAttribute_t at[2];
  const void* ptrs[2] = {lambda_final, lambda_initial};
  for (int k = 0; k < 2; k++) {
    if (hipPointerGetAttributes(&at[k], ptrs[k]) != hipSuccess || at[k].type != hipMemoryTypeDevice ||
        at[k].device != ctx->device) {
      (void)hipGetLastError();
      return fail(ctx, SDG_E_INVALID, "sdg_adjoint_backward_device: the adjoints must be device memory of the "
                                      "context's device");
    }
  }
  const int64_t n = static_cast<int64_t>(stateCount(ctx));
  const int grid = static_cast<int>((n + 255) / 256);
  hipLaunchKernelGGL(referenceToDeviceKernel, dim3(grid), dim3(256), 0, ctx->stream, lambda_final, ctx->adj_lam,
                     ctx->E, ctx->nv, ctx->nb);
  HIP_TRY(ctx, hipGetLastError());
  int rc = reverseRecorded(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(deviceToReferenceKernel, dim3(grid), dim3(256), 0, ctx->stream, ctx->adj_lam, lambda_initial,
                     ctx->E, ctx->nv, ctx->nb);
  HIP_TRY(ctx, hipGetLastError());
  if ((rc = loadState(ctx, ctx->adj_stored ? ctx->adj_pool : ctx->snaps[0]))) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SDG_OK;
}

}  // extern "C"

namespace {

// the reverse sweep of the recorded run, from the adjoint of the final state in adj_lam
int reverseRecorded(sdg_ctx* ctx) {
  int rc;
  if (ctx->adj_stored) {
    const int S = ctx->stages();
    const size_t n = stateCount(ctx);
    const size_t nt = static_cast<size_t>(ctx->E) * ctx->nv * ctx->naq;
    if (n >= kDirectAdjointStates) {
      // large meshes: the stage kernels read the recorded states and traces in the pools directly
      // (no per-step copies into the graph's fixed buffers; a launch is short against the kernels)
      for (int i = ctx->adj_nsteps - 1; i >= 0; i--)
        for (int st = S - 1; st >= 0; st--) {
          const size_t slot = static_cast<size_t>(i) * S + st;
          const double* tr = ctx->mixed ? nullptr : ctx->adj_tpool + slot * nt;
          if ((rc = vjpAt(ctx, ctx->adj_pool + slot * n, ctx->adj_lam, 1, nullptr, st, tr))) return rc;
        }
      return SDG_OK;
    }
    for (int i = ctx->adj_nsteps - 1; i >= 0; i--) {
      HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_S[0], ctx->adj_pool + static_cast<size_t>(i) * S * n, S * n * sizeof(double),
                                  hipMemcpyDeviceToDevice, ctx->stream));
      if (!ctx->mixed)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->adj_T, ctx->adj_tpool + static_cast<size_t>(i) * S * nt,
                                    S * nt * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      if ((rc = replayGraph(ctx, ctx->adjst_graph, ctx->adjst_graph_dt, ctx->adj_dt, 0,
                            [&]() { return enqueueAdjointStepStored(ctx); })))
        return rc;
    }
  } else if ((rc = reverseSweep(ctx, 0, ctx->adj_nsteps, ctx->adj_nslots - 1, 0))) {
    return rc;
  }
  return SDG_OK;
}

}  // namespace
