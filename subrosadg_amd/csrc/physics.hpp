// physics.hpp -- pointwise compressible-flow physics for the gfx950 kernels: variable conversions,
// equation of state, transport, convective Riemann fluxes and the viscous flux.
//
// Each function restates one reference routine (cited) with the same operation order, so the
// FP64 kernels agree with the reference arithmetic to rounding. Arrays are tiny fixed-size
// register arrays; D is a template parameter.
#ifndef SDG_PHYSICS_HPP_
#define SDG_PHYSICS_HPP_

#include <hip/hip_runtime.h>

namespace sdg {

constexpr double kGamma = 1.4;                 // PhysicalModel.cpp:45
constexpr double kPrandtl = 0.71;              // PhysicalModel.cpp:91,107
constexpr double kSutherland = 110.4 / 273.15; // PhysicalModel.cpp:109

// runtime physical model of one problem (no process-global state)
struct Phys {
  int eos;        // sdg_eos
  int transport;  // sdg_transport
  int flux;       // sdg_convective_flux
  int pad_;
  double cp, cv, mu, kappa, c0, rho0, padd;
  int source;  // sdg_source_term
  int pad2_;
  double beta, tref;  // Boussinesq thermal expansion coefficient and reference temperature
  // artificial viscosity: on/off, number of modes of order P-1, empirical tolerance k, factor f
  // and the sensor threshold s0 (kArtificialViscosityTolerance, SimulationControl.cpp:892-897)
  int shock;
  int av_low;
  double av_tol, av_factor, av_s0;
};

constexpr int kSourceBoussinesq = 1;  // sdg_source_term value of the Boussinesq source

// Boussinesq source on the last momentum component at a point (SourceTerm.cpp:41-55): rho beta
// (T - T_ref) g with g = 1, T = e / c_v; w: computational variables (rho, u.., e, p)
template <int D>
__device__ __forceinline__ double boussinesqSource(const Phys& ph, const double* w) {
  return w[0] * ph.beta * (w[D + 1] / ph.cv - ph.tref) * 1.0;
}

template <typename T>
__device__ __forceinline__ T pressureOf(const Phys& ph, T rho, T e) {  // PhysicalModel.cpp:45-75
  if (ph.eos == 0) return (T(kGamma) - T(1.0)) * rho * e;
  return T(ph.c0) * T(ph.c0) * (rho - T(ph.rho0)) + T(ph.padd);
}
template <typename T>
__device__ __forceinline__ T soundSpeedOf(const Phys& ph, T rho, T p) {
  if (ph.eos == 0) return sqrt(T(kGamma) * p / rho);
  return T(ph.c0);
}
template <typename T>
__device__ __forceinline__ T sutherlandRatio(T t) {  // PhysicalModel.cpp:111-114
  return sqrt(t * t * t) * (T(1.0) + T(kSutherland)) / (t + T(kSutherland));
}

// conserved -> computational (rho, u.., e, p): VariableConvertor.cpp:315-339. The divisions by rho
// are products with one reciprocal (an FP64 division is a multi-instruction sequence around a
// quarter-rate reciprocal; this differs from the reference's quotients by rounding only).
template <int D, typename T>
__device__ __forceinline__ void computational(const Phys& ph, const T* u, T* w) {
  const T rho = u[0];
  const T ir = T(1.0) / rho;
  w[0] = rho;
  T v2 = T(0);
#pragma unroll
  for (int d = 0; d < D; d++) {
    w[1 + d] = u[1 + d] * ir;
    v2 += w[1 + d] * w[1 + d];
  }
  const T e = u[D + 1] * ir - v2 / T(2.0);
  w[D + 1] = e;
  w[D + 2] = pressureOf(ph, rho, e);
}
template <int D, typename T>
__device__ __forceinline__ T velocitySquared(const T* w) {
  T s = T(0);
#pragma unroll
  for (int d = 0; d < D; d++) s += w[1 + d] * w[1 + d];
  return s;
}

// conserved gradient (row var*D+dim) -> primitive gradient: VariableConvertor.cpp:574-620.
// pg layout: [0,D) grad rho; [D, D+D*D) velocity gradient column-major (a,b) = du_b/dx_a;
// [(D+1)D, (D+2)D) grad T
// (divisions by rho and c_v as products with reciprocals, as in computational)
template <int D, typename T>
__device__ __forceinline__ void primitiveGradient(const Phys& ph, const T* w, const T* u, const T* g, T* pg) {
  const T rho = w[0];
  const T ir = T(1.0) / rho, icv = T(1.0) / T(ph.cv);
#pragma unroll
  for (int a = 0; a < D; a++) pg[a] = g[a];
  T vg[D * D];
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int a = 0; a < D; a++) {
      vg[a + D * b] = (g[(1 + b) * D + a] - g[a] * w[1 + b]) * ir;
      pg[(1 + b) * D + a] = vg[a + D * b];
    }
  const T etot = u[D + 1] * ir;
#pragma unroll
  for (int a = 0; a < D; a++) {
    T vgu = T(0);
#pragma unroll
    for (int b = 0; b < D; b++) vgu += vg[a + D * b] * w[1 + b];
    const T de = (g[(D + 1) * D + a] - g[a] * etot) * ir - vgu;
    pg[(D + 1) * D + a] = de * icv;
  }
}

// normal convective flux F(u).n: ConvectiveFlux.cpp:59-91
template <int D, typename T>
__device__ __forceinline__ void normalFlux(const T* n, const T* w, T* f) {
  const T rho = w[0];
  T un = T(0);
#pragma unroll
  for (int d = 0; d < D; d++) un += w[1 + d] * n[d];
  const T p = w[D + 2];
  f[0] = rho * un;
#pragma unroll
  for (int d = 0; d < D; d++) f[1 + d] = rho * un * w[1 + d] + p * n[d];
  const T et = w[D + 1] + velocitySquared<D>(w) / T(2.0);
  f[D + 1] = (rho * et + p) * un;
}

// numerical convective flux H(uL, uR; n): ConvectiveFlux.cpp:93-439
// FLUX >= 0 fixes the flux at compile time (the other cases are not compiled in); FLUX < 0 takes
// it from ph.flux.
template <int D, typename T, int FLUX = -1>
__device__ void convectiveFlux(const Phys& ph, const T* n, const T* ul, const T* wl, const T* ur, const T* wr, T* h) {
  constexpr int NV = D + 2;
  T fl[NV], fr[NV];
  const int flux = FLUX < 0 ? ph.flux : FLUX;
  if (flux == 0) {  // Central :93-104
    normalFlux<D>(n, wl, fl);
    normalFlux<D>(n, wr, fr);
#pragma unroll
    for (int v = 0; v < NV; v++) h[v] = (fl[v] + fr[v]) / T(2.0);
  } else if (flux == 1) {  // Lax-Friedrichs :106-134
    normalFlux<D>(n, wl, fl);
    normalFlux<D>(n, wr, fr);
    T unl = T(0), unr = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      unl += wl[1 + d] * n[d];
      unr += wr[1 + d] * n[d];
    }
    const T cl = soundSpeedOf(ph, wl[0], wl[D + 2]);
    const T cr = soundSpeedOf(ph, wr[0], wr[D + 2]);
    const T lam = fmax(fabs(unl) + cl, fabs(unr) + cr);
#pragma unroll
    for (int v = 0; v < NV; v++) h[v] = ((fl[v] + fr[v]) - lam * (ur[v] - ul[v])) / T(2.0);
  } else if (flux == 2) {  // HLLC :136-238
    const T rl = wl[0], rr = wr[0], pl = wl[D + 2], pr = wr[D + 2];
    T unl = T(0), unr = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      unl += wl[1 + d] * n[d];
      unr += wr[1 + d] * n[d];
    }
    const T cl = soundSpeedOf(ph, rl, pl), cr = soundSpeedOf(ph, rr, pr);
    const T rbar = (rl + rr) / T(2.0), cbar = (cl + cr) / T(2.0);
    const T ps = fmax(T(0), (pl + pr) / T(2.0) - (unr - unl) * rbar * cbar);
    const T g = T(kGamma);
    const T sl = unl - cl * (ps <= pl ? T(1.0) : sqrt(T(1.0) + (g + T(1.0)) * (ps / pl - T(1.0)) / T(2.0) / g));
    if (sl >= T(0)) {
      normalFlux<D>(n, wl, h);
      return;
    }
    const T sr = unr + cr * (ps <= pr ? T(1.0) : sqrt(T(1.0) + (g + T(1.0)) * (ps / pr - T(1.0)) / T(2.0) / g));
    if (sr <= T(0)) {
      normalFlux<D>(n, wr, h);
      return;
    }
    const T sm = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / (rl * (sl - unl) - rr * (sr - unr));
    T us[NV];
    if (sm >= T(0)) {  // the star state's divisions by (s - s_m) as products with one reciprocal
      normalFlux<D>(n, wl, fl);
      const T is = T(1.0) / (sl - sm);
      us[0] = rl * (sl - unl) * is;
#pragma unroll
      for (int d = 0; d < D; d++) us[1 + d] = ((sl - unl) * rl * wl[1 + d] + (ps - pl) * n[d]) * is;
      us[D + 1] = ((sl - unl) * rl * (wl[D + 1] + velocitySquared<D>(wl) / T(2.0)) - pl * unl + ps * sm) * is;
#pragma unroll
      for (int v = 0; v < NV; v++) h[v] = fl[v] + sl * (us[v] - ul[v]);
    } else {
      normalFlux<D>(n, wr, fr);
      const T is = T(1.0) / (sr - sm);
      us[0] = rr * (sr - unr) * is;
#pragma unroll
      for (int d = 0; d < D; d++) us[1 + d] = ((sr - unr) * rr * wr[1 + d] + (ps - pr) * n[d]) * is;
      us[D + 1] = ((sr - unr) * rr * (wr[D + 1] + velocitySquared<D>(wr) / T(2.0)) - pr * unr + ps * sm) * is;
#pragma unroll
      for (int v = 0; v < NV; v++) h[v] = fr[v] + sr * (us[v] - ur[v]);
    }
  } else if (flux == 3) {  // Roe with Harten entropy fix :240-350
    normalFlux<D>(n, wl, fl);
    normalFlux<D>(n, wr, fr);
    const T sql = sqrt(wl[0]), sqr = sqrt(wr[0]);
    const T sqs = sql + sqr;
    const T isqs = T(1.0) / sqs;  // Roe averages: products with one reciprocal of the weight sum
    const T rho = sqrt(wl[0] * wr[0]);
    T vel[D];
#pragma unroll
    for (int d = 0; d < D; d++) vel[d] = (sql * wl[1 + d] + sqr * wr[1 + d]) * isqs;
    const T g = T(kGamma);
    const T hl = wl[D + 1] * g + velocitySquared<D>(wl) / T(2.0);
    const T hr = wr[D + 1] * g + velocitySquared<D>(wr) / T(2.0);
    const T htot = (sql * hl + sqr * hr) * isqs;
    T v2 = T(0), un = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      v2 += vel[d] * vel[d];
      un += vel[d] * n[d];
    }
    const T e = (htot - v2 / T(2.0)) / g;
    const T p = pressureOf(ph, rho, e);
    const T c = soundSpeedOf(ph, rho, p);
    T dw[D + 3];
#pragma unroll
    for (int i = 0; i < D + 3; i++) dw[i] = wr[i] - wl[i];
    T dun = T(0), vdv = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      dun += dw[1 + d] * n[d];
      vdv += vel[d] * dw[1 + d];
    }
    const T delta = c / T(20.0);
    const T a1 = fabs(un - c) > delta ? fabs(un - c) : ((un - c) * (un - c) + delta * delta) / (T(2.0) * delta);
    const T a3 = fabs(un + c) > delta ? fabs(un + c) : ((un + c) * (un + c) + delta * delta) / (T(2.0) * delta);
    const T dp = dw[D + 2], drho = dw[0];
    // dissipation sum over the wave columns, accumulated column by column (the same per-variable
    // addition order as summing the NV x NV column matrix row-wise, without materialising it)
    T sum[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) sum[v] = T(0);
    const T ic2 = T(1.0) / (c * c);
    const T s0 = a1 * (dp - rho * c * dun) * (T(0.5) * ic2);
    sum[0] += T(1.0) * s0;
#pragma unroll
    for (int d = 0; d < D; d++) sum[1 + d] += (vel[d] - c * n[d]) * s0;
    sum[D + 1] += (htot - c * un) * s0;
    const T s1 = fabs(un) * (drho - dp * ic2);
    sum[0] += T(1.0) * s1;
#pragma unroll
    for (int d = 0; d < D; d++) sum[1 + d] += vel[d] * s1;
    sum[D + 1] += v2 / T(2.0) * s1;
    if (D >= 2) {
      const T s2 = fabs(un) * rho;
      sum[0] += T(0.0) * s2;
#pragma unroll
      for (int d = 0; d < D; d++) sum[1 + d] += (dw[1 + d] - dun * n[d]) * s2;
      sum[D + 1] += (vdv - un * dun) * s2;
    }
    // in 3-D the fourth column stays zero (ConvectiveFlux.cpp:240-350); adding it changes nothing
    const T s3 = a3 * (dp + rho * c * dun) * (T(0.5) * ic2);
    sum[0] += T(1.0) * s3;
#pragma unroll
    for (int d = 0; d < D; d++) sum[1 + d] += (vel[d] + c * n[d]) * s3;
    sum[D + 1] += (htot + c * un) * s3;
#pragma unroll
    for (int v = 0; v < NV; v++) h[v] = ((fl[v] + fr[v]) - sum[v]) / T(2.0);
  } else {  // Exact flux of the weakly-compressible model :352-414
    const T c = soundSpeedOf(ph, T(0), T(0));
    T unl = T(0), unr = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      unl += wl[1 + d] * n[d];
      unr += wr[1 + d] * n[d];
    }
    const T rho = sqrt(wl[0] * wr[0] * exp((unl - unr) / c));
    const T un = (unl + unr) / T(2.0) + log(wl[0] / wr[0]) * c / T(2.0);
    const T* ws = un < T(0) ? wr : wl;
    const T uns = un < T(0) ? unr : unl;
    T we[D + 3];
    const T e = ws[D + 1] * ws[0] / rho;
    we[0] = rho;
#pragma unroll
    for (int d = 0; d < D; d++) we[1 + d] = ws[1 + d] + (un - uns) * n[d];
    we[D + 1] = e;
    we[D + 2] = pressureOf(ph, rho, e);
    normalFlux<D>(n, we, h);
  }
}

// viscous raw flux F_v, D x NV column-major: ViscousFlux.cpp:58-103
template <int D, typename T>
__device__ __forceinline__ void viscousRawFlux(const Phys& ph, const T* w, const T* pg, T* fv) {
  T mu = T(ph.mu), kap = T(ph.kappa);
  if (ph.transport == 2) {
    const T r = sutherlandRatio(w[D + 1] / T(ph.cv));
    mu = mu * r;
    kap = kap * r;
  }
  const T* vg = pg + D;
  T tr = T(0);
#pragma unroll
  for (int a = 0; a < D; a++) tr += vg[a + D * a];
#pragma unroll
  for (int a = 0; a < D; a++) fv[a] = T(0);
  T tau[D * D];
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int a = 0; a < D; a++)
      tau[a + D * b] = mu * (vg[a + D * b] + vg[b + D * a]) - T(2.0) / T(3.0) * mu * tr * (a == b ? T(1.0) : T(0.0));
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int a = 0; a < D; a++) fv[a + D * (1 + b)] = tau[a + D * b];
  const T* tg = pg + (D + 1) * D;
#pragma unroll
  for (int a = 0; a < D; a++) {
    T s = T(0);
#pragma unroll
    for (int b = 0; b < D; b++) s += tau[a + D * b] * w[1 + b];
    fv[a + D * (D + 1)] = s + kap * tg[a];
  }
}

// ---- boundary conditions (BoundaryCondition.cpp:79-547) ---------------------------------------
// Computational variables w = (rho, u.., e, p). wl: interior state at the boundary point, wd: the
// boundary data (boundary_dummy_variable_), n: outward normal. Returns the boundary state wb of
// calculateBoundaryVariable for type bc (sdg_boundary order).
template <int D>
__device__ __forceinline__ void boundaryState(const Phys& ph, int bc, const double* n, const double* wl,
                                              const double* wd, double* wb) {
  constexpr int NW = D + 3;
  double unl = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) unl += wl[1 + d] * n[d];
  const double cl = soundSpeedOf(ph, wl[0], wl[D + 2]);
  const double mn = unl / cl;
  if (bc == 0) {  // Riemann-invariant farfield :79-247 (compressible branch)
    if (fabs(mn) > 1.0) {  // supersonic: all characteristics from one side
      const bool in = mn < 0.0;
#pragma unroll
      for (int i = 0; i < NW; i++) wb[i] = in ? wd[i] : wl[i];
      return;
    }
    const double g = kGamma;
    double und = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) und += wd[1 + d] * n[d];
    const double cd = soundSpeedOf(ph, wd[0], wd[D + 2]);
    const double rm = und - 2.0 * cd / (g - 1.0);  // invariant carried in from outside
    const double rp = unl + 2.0 * cl / (g - 1.0);  // invariant carried out from inside
    const double unb = (rm + rp) / 2.0;
    const bool in = mn < 0.0;  // subsonic inflow: tangential velocity and entropy from outside
    double uns = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) uns += (in ? wd[1 + d] : wl[1 + d]) * n[d];
#pragma unroll
    for (int d = 0; d < D; d++) wb[1 + d] = (in ? wd[1 + d] : wl[1 + d]) + (unb - uns) * n[d];
    const double cb = (g - 1.0) * (rp - rm) / 4.0;
    const double rs = in ? wd[0] : wl[0], ps = in ? wd[D + 2] : wl[D + 2];
    const double sb = ps / pow(rs, g);  // calculateEntropyFromDensityPressure (PhysicalModel.cpp:148-150)
    const double rb = pow(cb * cb / (g * sb), 1.0 / (g - 1.0));
    const double pb = rb * cb * cb / g;
    wb[0] = rb;
    wb[D + 1] = pb / ((g - 1.0) * rb);
    wb[D + 2] = pb;
  } else if (bc == 1) {  // velocity inflow :309-333
#pragma unroll
    for (int i = 0; i < NW; i++) wb[i] = wd[i];
    if (mn > -1.0) wb[D + 2] = wl[D + 2];
  } else if (bc == 2) {  // pressure outflow :362-386
#pragma unroll
    for (int i = 0; i < NW; i++) wb[i] = wl[i];
    if (mn < 1.0) wb[D + 2] = wd[D + 2];
  } else if (bc == 3) {  // isothermal no-slip wall :416-434
    wb[0] = wl[0];
#pragma unroll
    for (int d = 0; d < D; d++) wb[1 + d] = wd[1 + d];
    wb[D + 1] = wd[D + 1];
    wb[D + 2] = pressureOf(ph, wl[0], wd[D + 1]);
  } else if (bc == 4) {  // adiabatic slip wall :474-488
#pragma unroll
    for (int i = 0; i < NW; i++) wb[i] = wl[i];
#pragma unroll
    for (int d = 0; d < D; d++) wb[1 + d] = wl[1 + d] - unl * n[d];
  } else {  // adiabatic no-slip wall :521-533
#pragma unroll
    for (int i = 0; i < NW; i++) wb[i] = wl[i];
#pragma unroll
    for (int d = 0; d < D; d++) wb[1 + d] = wd[1 + d];
  }
}

// computational -> conserved (VariableConvertor.cpp:341-366 on computational input)
template <int D>
__device__ __forceinline__ void conservedOf(const double* w, double* u) {
  u[0] = w[0];
#pragma unroll
  for (int d = 0; d < D; d++) u[1 + d] = w[0] * w[1 + d];
  u[D + 1] = w[0] * (w[D + 1] + velocitySquared<D>(w) / 2.0);
}

// States of the gradient pass at a boundary point (calculateBoundaryGradientVariable): the value
// entering the volume-gradient face term and the jump entering the interface term. Farfield,
// inflow and outflow: (u-, 0); walls: (u_b, u_b - u-) with u_b the conserved boundary state.
template <int D>
__device__ __forceinline__ void boundaryGradientStates(const Phys& ph, int bc, const double* n, const double* ul,
                                                       const double* wd, double* uvg, double* uig) {
  constexpr int NV = D + 2;
  if (bc <= 2) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      uvg[v] = ul[v];
      uig[v] = 0.0;
    }
    return;
  }
  double wl[D + 3], wb[D + 3], ub[NV];
  computational<D>(ph, ul, wl);
  boundaryState<D>(ph, bc, n, wl, wd, wb);
  conservedOf<D>(wb, ub);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    uvg[v] = ub[v];
    uig[v] = ub[v] - ul[v];
  }
}

}  // namespace sdg

#endif  // SDG_PHYSICS_HPP_
