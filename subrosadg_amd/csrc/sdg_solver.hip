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
#include <new>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "dg_kernels.hpp"
#include "dg_mfma.hpp"
#include "kit.hpp"
#include "sdg_internal.hpp"

using sdg::Geo;
using sdg::Phys;
using sdg::StageArgs;

namespace {

using StageFn = void (*)(hipStream_t, const Geo&, const Phys&, const StageArgs&);
using TraceFn = void (*)(hipStream_t, const Geo&, const double*, double*);
using DtFn = void (*)(hipStream_t, const Geo&, const Phys&, const double*, const double*, double, int, double*);
using MonFn = void (*)(hipStream_t, const double*, int, double*, int*);

struct Launchers {
  bool mfma = false;
  StageFn euler = nullptr;
  StageFn ns_grad_br1 = nullptr;
  StageFn ns_grad_br2 = nullptr;
  StageFn ns_flux = nullptr;
  StageFn ns_face = nullptr;  // matrix-core path: face fluxes before ns_flux
  TraceFn trace = nullptr;
  DtFn dt = nullptr;
  MonFn monitor = nullptr;
};

inline int blocksFor(int E, int epb) { return (E + epb - 1) / epb; }

// VALU kernels: any size; MFMA kernels: bases that fill 16x16 tiles
template <int D, int NB, int NQ, int NF, int NFQ, bool MFMA>
Launchers makeLaunchers() {
  using S = sdg::Sizes<D, NB, NQ, NF, NFQ>;
  constexpr int EPB = S::EPB;
  constexpr int NV = D + 2;
  Launchers l;
  l.mfma = MFMA;
  if constexpr (MFMA) {
    constexpr int EPB_S = 16 / NV, EPB_G = 16 / (NV * D);
    l.euler = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::eulerStageMfma<D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_grad_br1 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradMfma<D, NB, NQ, NF, NFQ, 1>), dim3(blocksFor(g.E, EPB_G)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    l.ns_grad_br2 = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsGradMfma<D, NB, NQ, NF, NFQ, 2>), dim3(blocksFor(g.E, EPB_G)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
    l.ns_face = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsFaceFluxKernel<D, NF, NFQ>), dim3(blocksFor(g.F * NFQ, sdg::kThreads)),
                         dim3(sdg::kThreads), 0, st, g, ph, a);
    };
    l.ns_flux = [](hipStream_t st, const Geo& g, const Phys& ph, const StageArgs& a) {
      hipLaunchKernelGGL((sdg::nsResidualMfma<D, NB, NQ, NF, NFQ>), dim3(blocksFor(g.E, EPB_S)), dim3(sdg::kThreads),
                         0, st, g, ph, a);
    };
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
  }
  l.trace = [](hipStream_t st, const Geo& g, const double* U, double* Uf) {
    const int64_t n = static_cast<int64_t>(g.E) * (D + 2) * NF * NFQ;
    hipLaunchKernelGGL((sdg::traceKernel<D, NB, NQ, NF, NFQ>), dim3((n + sdg::kThreads - 1) / sdg::kThreads),
                       dim3(sdg::kThreads), 0, st, g, U, Uf);
  };
  l.dt = [](hipStream_t st, const Geo& g, const Phys& ph, const double* U, const double* me, double cfl, int order,
            double* out) {
    hipLaunchKernelGGL((sdg::dtKernel<D, NB, NQ, NF, NFQ>), dim3((g.E + sdg::kThreads - 1) / sdg::kThreads),
                       dim3(sdg::kThreads), 0, st, g, ph, U, me, cfl, order, out);
  };
  l.monitor = [](hipStream_t st, const double* mon, int E, double* rel, int* status) {
    hipLaunchKernelGGL((sdg::monitorKernel<D + 2>), dim3(1), dim3(sdg::kThreads), 0, st, mon, E, rel, status);
  };
  return l;
}

bool selectLaunchers(int type, int order, bool force_valu, Launchers& l) {
  if (type == SDG_ELEM_QUADRANGLE) {
    switch (order) {
      case 1: l = makeLaunchers<2, 4, 4, 4, 2, false>(); return true;
      case 2: l = makeLaunchers<2, 9, 9, 4, 3, false>(); return true;
      case 3:
        l = force_valu ? makeLaunchers<2, 16, 16, 4, 4, false>() : makeLaunchers<2, 16, 16, 4, 4, true>();
        return true;
      case 4:
        l = force_valu ? makeLaunchers<2, 25, 25, 4, 5, false>() : makeLaunchers<2, 25, 25, 4, 5, true>();
        return true;
      case 5:
        l = force_valu ? makeLaunchers<2, 36, 36, 4, 6, false>() : makeLaunchers<2, 36, 36, 4, 6, true>();
        return true;
      default: return false;
    }
  }
  if (type == SDG_ELEM_HEXAHEDRON) {
    switch (order) {
      case 1: l = makeLaunchers<3, 8, 8, 6, 4, false>(); return true;
      case 2:
        l = force_valu ? makeLaunchers<3, 27, 27, 6, 9, false>() : makeLaunchers<3, 27, 27, 6, 9, true>();
        return true;
      case 3:
        l = force_valu ? makeLaunchers<3, 64, 64, 6, 16, false>() : makeLaunchers<3, 64, 64, 6, 16, true>();
        return true;
      case 4:
        l = force_valu ? makeLaunchers<3, 125, 125, 6, 25, false>() : makeLaunchers<3, 125, 125, 6, 25, true>();
        return true;
      default: return false;
    }
  }
  return false;
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
  std::vector<double> h_lsq, h_V;
  double *U = nullptr, *Un = nullptr, *Uf[2] = {nullptr, nullptr}, *Gf = nullptr, *Rvol = nullptr, *Hs = nullptr, *Fv = nullptr;
  double *Rout = nullptr, *Wout = nullptr, *mon = nullptr, *rel = nullptr, *dtbuf = nullptr;
  int* status = nullptr;
  int cur = 0;
  bool has_state = false;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  double graph_dt[2] = {0.0, 0.0};
  bool use_graph = true;
  float last_ms = 0.0f;

  bool ns() const { return prob.equation == SDG_EQ_COMPRESSIBLE_NS; }
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
  }
}

StageArgs baseArgs(sdg_ctx* ctx) {
  StageArgs a{};
  a.U = ctx->U;
  a.Un = ctx->Un;
  a.Uf_in = ctx->Uf[ctx->cur];
  a.Uf_out = ctx->Uf[ctx->cur ^ 1];
  a.Gf = ctx->Gf;
  a.Rvol = ctx->Rvol;
  a.Rout = ctx->Rout;
  a.Wout = nullptr;
  a.Hs = ctx->Hs;
  a.Fv = ctx->Fv;
  a.mon = ctx->mon;
  a.status = ctx->status;
  return a;
}

// Enqueue one stage (mode 0: update, 1: residual only). Advances ctx->cur for mode 0. With ev
// (four events), records ev[0]/ev[1] around the first kernel and ev[2]/ev[3] around the second.
void enqueueStage(sdg_ctx* ctx, int stage, double dt, int mode, double* wout, hipEvent_t* ev = nullptr) {
  StageArgs a = baseArgs(ctx);
  double c[3];
  ctx->rkCoefficients(stage < 0 ? 0 : stage, c);
  a.c0 = c[0];
  a.c1 = c[1];
  a.c2dt = c[2] * dt;
  a.last_stage = (stage == ctx->stages() - 1) ? 1 : 0;
  a.mode = mode;
  a.Wout = wout;
  if (mode == 1) a.status = nullptr;
  if (ev) (void)hipEventRecord(ev[0], ctx->stream);
  if (ctx->ns()) {
    if (ctx->prob.viscous_flux == SDG_VISC_BR1)
      ctx->launch.ns_grad_br1(ctx->stream, ctx->geo, ctx->phys, a);
    else
      ctx->launch.ns_grad_br2(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) {
      (void)hipEventRecord(ev[1], ctx->stream);
      (void)hipEventRecord(ev[2], ctx->stream);
    }
    if (ctx->launch.ns_face) ctx->launch.ns_face(ctx->stream, ctx->geo, ctx->phys, a);
    ctx->launch.ns_flux(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) (void)hipEventRecord(ev[3], ctx->stream);
  } else {
    ctx->launch.euler(ctx->stream, ctx->geo, ctx->phys, a);
    if (ev) {
      (void)hipEventRecord(ev[1], ctx->stream);
      (void)hipEventRecord(ev[2], ctx->stream);
      (void)hipEventRecord(ev[3], ctx->stream);
    }
  }
  if (mode == 0) ctx->cur ^= 1;
}

void enqueueStep(sdg_ctx* ctx, double dt) {
  for (int s = 0; s < ctx->stages(); s++) enqueueStage(ctx, s, dt, 0, nullptr);
  ctx->launch.monitor(ctx->stream, ctx->mon, ctx->E, ctx->rel, ctx->status);
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
  if (p->equation == SDG_EQ_COMPRESSIBLE_NS && p->transport != SDG_TRANSPORT_CONSTANT &&
      p->transport != SDG_TRANSPORT_SUTHERLAND) {
    msg = "Navier-Stokes needs the constant or Sutherland transport model";
    return SDG_E_INVALID;
  }
  if (!(p->cv > 0.0) || !(p->cp > 0.0)) {
    msg = "specific heats must be positive";
    return SDG_E_INVALID;
  }
  if (p->precision != SDG_PREC_FP64) {
    msg = "only the FP64 precision mode is implemented in this build";
    return SDG_E_UNSUPPORTED;
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
  const sdg::Mesh& mesh = *sdg::meshOf(mesh_handle);
  if (!mesh.affine) return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: curved (non-affine) elements are not supported");
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
  const char* fv = std::getenv("SDG_FORCE_VALU");
  if (!selectLaunchers(r.type, r.order, fv && fv[0] == '1', ctx->launch)) {
    delete ctx;
    return fail(nullptr, SDG_E_UNSUPPORTED, "sdg_create: no kernel instantiation for this element type and order");
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
  const char* ng = std::getenv("SDG_NO_GRAPH");
  ctx->use_graph = !(ng && ng[0] == '1');

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
  std::vector<int> efid(static_cast<size_t>(E) * nf), enb(static_cast<size_t>(E) * nf), enbf(static_cast<size_t>(E) * nf),
      eside(static_cast<size_t>(E) * nf);
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      const int fid = mesh.elem_face[static_cast<size_t>(e) * nf + s];
      const int side = mesh.elem_side[static_cast<size_t>(e) * nf + s];
      if (fid < 0) {
        ctx->err = "sdg_create: element face without a neighbour (boundary faces are not supported)";
        return bail(SDG_E_UNSUPPORTED);
      }
      const sdg::Face& fc = mesh.faces[fid];
      efid[static_cast<size_t>(e) * nf + s] = fid;
      eside[static_cast<size_t>(e) * nf + s] = side;
      enb[static_cast<size_t>(e) * nf + s] = fc.parent[side ^ 1];
      enbf[static_cast<size_t>(e) * nf + s] = fc.local[side ^ 1];
    }
  // per element slot: neighbour slot and (face point, side), so a kernel reaches the neighbour's
  // trace with one indexed load (no face -> permutation -> slot chain)
  if (static_cast<int64_t>(E) * naq >= (int64_t(1) << 31) || F * nfq >= (int64_t(1) << 30)) {
    ctx->err = "sdg_create: mesh too large for 32-bit slot indices";
    return bail(SDG_E_UNSUPPORTED);
  }
  std::vector<int> snb(static_cast<size_t>(E) * naq), sfp(static_cast<size_t>(E) * naq);
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      const int fid = efid[static_cast<size_t>(e) * nf + s];
      const int side = eside[static_cast<size_t>(e) * nf + s];
      for (int j = 0; j < nfq; j++) {
        const int p = side == 0 ? j : fseqinv[static_cast<size_t>(fid) * nfq + j];
        const int k = side == 0 ? fseq[static_cast<size_t>(fid) * nfq + j] : p;
        snb[static_cast<size_t>(e) * naq + s * nfq + j] = enb[static_cast<size_t>(e) * nf + s] * naq + enbf[static_cast<size_t>(e) * nf + s] * nfq + k;
        sfp[static_cast<size_t>(e) * naq + s * nfq + j] = (fid * nfq + p) * 2 + side;
      }
    }
  Geo& g = ctx->geo;
  double *dVt, *ddV, *dVft, *dVf, *dMinv, *dLf, *dwq, *dginv, *ddetj, *dnormal, *djw, *dme;
  int *dseq, *dseqinv, *defid, *denb, *denbf, *deside, *dfpar, *dfloc, *dsnb, *dsfp;
  if (upload(ctx, Vt, &dVt) || upload(ctx, r.dV, &ddV) || upload(ctx, Vft, &dVft) || upload(ctx, r.Vf, &dVf) ||
      upload(ctx, r.mass_inverse, &dMinv) || upload(ctx, Lf, &dLf) || upload(ctx, r.qweight, &dwq) ||
      upload(ctx, ginv, &dginv) || upload(ctx, detj, &ddetj) || upload(ctx, fnormal, &dnormal) ||
      upload(ctx, fjw, &djw) || upload(ctx, fseq, &dseq) || upload(ctx, fseqinv, &dseqinv) ||
      upload(ctx, efid, &defid) || upload(ctx, enb, &denb) || upload(ctx, enbf, &denbf) ||
      upload(ctx, eside, &deside) || upload(ctx, mesh.min_edge, &dme) || upload(ctx, fpar, &dfpar) ||
      upload(ctx, floc, &dfloc) || upload(ctx, snb, &dsnb) || upload(ctx, sfp, &dsfp))
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
  g.efid = defid;
  g.enb = denb;
  g.enbf = denbf;
  g.eside = deside;
  g.snb = dsnb;
  g.sfp = dsfp;
  g.fpar = dfpar;
  g.floc = dfloc;
  g.E = ctx->E;
  g.F = static_cast<int>(F);
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
    g.op.pinterp = ppi;
    g.op.ptrace = ppt;
    g.op.interp = pi;
    g.op.grad = pg;
    g.op.face = pf;
    g.op.minv = pm;
    g.op.trace = pt;
    g.op.lift = pl;
  }
  ctx->minedge = dme;
  ctx->h_lsq = r.lsq_inverse;
  ctx->h_V = r.V;
  const size_t nU = static_cast<size_t>(E) * ctx->nv * nb, nUf = static_cast<size_t>(E) * ctx->nv * naq;
  if (allocate(ctx, nU, &ctx->U) || allocate(ctx, nU, &ctx->Un) || allocate(ctx, nUf, &ctx->Uf[0]) ||
      allocate(ctx, nUf, &ctx->Uf[1]) || allocate(ctx, nU, &ctx->Rout) ||
      allocate(ctx, static_cast<size_t>(E) * ctx->nv, &ctx->mon) || allocate(ctx, static_cast<size_t>(ctx->nv), &ctx->rel) ||
      allocate(ctx, static_cast<size_t>(E), &ctx->dtbuf) || allocate(ctx, size_t(2), &ctx->status))
    return bail(SDG_E_NOMEM);
  if (ctx->ns()) {
    if (allocate(ctx, static_cast<size_t>(E) * ctx->ng * naq, &ctx->Gf) || allocate(ctx, nU, &ctx->Rvol) ||
        allocate(ctx, nUf, &ctx->Hs))
      return bail(SDG_E_NOMEM);
    if (ctx->launch.mfma && allocate(ctx, static_cast<size_t>(E) * nq * D * ctx->nv, &ctx->Fv)) return bail(SDG_E_NOMEM);
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
  ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->Uf[0]);
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

int sdg_residual(sdg_ctx* ctx, double* residual) {
  if (!ctx || !residual) return fail(ctx, SDG_E_INVALID, "sdg_residual: null argument");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_residual: no state set");
  enqueueStage(ctx, -1, 0.0, 1, nullptr);
  HIP_TRY(ctx, hipGetLastError());
  std::vector<double> dev(static_cast<size_t>(ctx->E) * ctx->nv * ctx->nb);
  HIP_TRY(ctx, hipMemcpyAsync(dev.data(), ctx->Rout, dev.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  toReference(ctx, dev, residual);
  return SDG_OK;
}

int sdg_get_gradient(sdg_ctx* ctx, double* gradient) {
  if (!ctx || !gradient) return fail(ctx, SDG_E_INVALID, "sdg_get_gradient: null argument");
  if (!ctx->ns()) return fail(ctx, SDG_E_UNSUPPORTED, "sdg_get_gradient: only defined for Navier-Stokes");
  if (!ctx->has_state) return fail(ctx, SDG_E_STATE, "sdg_get_gradient: no state set");
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
  if (nsteps < 0 || !(dt >= 0.0)) return fail(ctx, SDG_E_INVALID, "sdg_step: bad dt or step count");
  hipLaunchKernelGGL(sdg::resetStatusKernel, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
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
    ctx->launch.trace(ctx->stream, ctx->geo, ctx->U, ctx->Uf[0]);
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
  if (nsteps < 1 || !(dt >= 0.0)) return fail(ctx, SDG_E_INVALID, "sdg_time_kernels: bad dt or step count");
  const int st = ctx->stages();
  const size_t per_step = static_cast<size_t>(4 * st + 2);
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
    for (int s = 0; s < st; s++) enqueueStage(ctx, s, dt, 0, nullptr, e + 4 * s);
    (void)hipEventRecord(e[4 * st], ctx->stream);
    ctx->launch.monitor(ctx->stream, ctx->mon, ctx->E, ctx->rel, ctx->status);
    (void)hipEventRecord(e[4 * st + 1], ctx->stream);
  }
  hipError_t err = hipStreamSynchronize(ctx->stream);
  double sum[SDG_KERNEL_CLASSES] = {0.0, 0.0, 0.0};
  for (int i = 0; i < nsteps && err == hipSuccess; i++) {
    hipEvent_t* e = &ev[per_step * i];
    for (int s = 0; s < st; s++) {
      float a = 0.0f, b = 0.0f;
      (void)hipEventElapsedTime(&a, e[4 * s], e[4 * s + 1]);
      (void)hipEventElapsedTime(&b, e[4 * s + 2], e[4 * s + 3]);
      sum[0] += a;
      sum[1] += b;
    }
    float m = 0.0f;
    (void)hipEventElapsedTime(&m, e[4 * st], e[4 * st + 1]);
    sum[2] += m;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  if (err != hipSuccess) return fail(ctx, SDG_E_DEVICE, std::string("sdg_time_kernels: ") + hipGetErrorString(err));
  const double launches[SDG_KERNEL_CLASSES] = {double(nsteps) * st, double(nsteps) * st, double(nsteps)};
  for (int k = 0; k < n_out; k++) avg_ms[k] = k < SDG_KERNEL_CLASSES ? sum[k] / launches[k] : 0.0;
  if (!ctx->ns() && n_out > 1) avg_ms[1] = 0.0;
  return SDG_OK;
}

}  // extern "C"
