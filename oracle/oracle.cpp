// oracle.cpp -- CPU ORACLE (test infrastructure only).
//
// A plain C++/OpenMP restatement of the reference's DG residual and explicit time stepping
// (SubrosaDG src/Solver/*), kept in the reference's own pass structure and per-element
// array-of-structures layout so that it can serve as the parity checker of the HIP engine and as
// the timed CPU baseline. Nothing in the product links or calls this file: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Per-element matrices are stored column-major, as the reference's Eigen members are
// (SolveControl.cpp:45-174); operand tables arrive through oracle_create in the row-major layout
// sdg_mesh_export writes and are transposed here.
//
// Pinning (what makes this restatement trustworthy; the checks live in tests/):
//   quadrature, basis, geometry    -> closed forms and numpy (tests/test_kit.py)
//   pointwise physics and fluxes   -> consistency, conservation, rotation invariance, Roe vs an
//                                     eigen-decomposition |A| built with numpy, supersonic
//                                     upwinding, stationary contact (tests/test_oracle_flux.py)
//   whole residual / step          -> free-stream preservation, discrete conservation, exact
//                                     lifted gradient of continuous polynomials, convergence
//                                     order on a smooth Euler solution, viscous decay rate,
//                                     SSP-RK3 on a linear ODE (tests/test_oracle_scheme.py)
//   closed forms                   -> BR1/BR2 face lifts of a single jump (Legendre kernel at
//                                     the end points), CFL dt of a uniform state, the monitor
//                                     of element-constant and linear residuals, the Exact flux's
//                                     star state on the two isothermal rarefaction invariants,
//                                     HLLC at symmetric pairs (zero mass and energy flux, normal
//                                     momentum flux = p*, vacuum), Sutherland's law through the
//                                     viscous flux (tests/test_oracle_pins.py)
//   parity unpinned                -> the value of HLLC's star-pressure estimate away from
//                                     symmetric pairs is the reference's formula
//                                     (ConvectiveFlux.cpp:162-165), which no closed form fixes.
//
// The arithmetic type is ORACLE_REAL (default double). The same source is also built with
// ORACLE_REAL=long double (x87 extended, 64-bit mantissa) into liboracle_ld.so: that build takes
// the same FP64 operand tables and states and evaluates the same passes with 11 more mantissa
// bits, which measures the FP64 rounding floor of a case -- how far the FP64 oracle itself is
// from the exact-arithmetic value of the same discrete operator. Its interface takes and returns
// arrays of ORACLE_REAL.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "../include/sdg.h"

#ifdef ORACLE_TANGENT
// Tangent-linear build (liboracle_tl.so): every Real carries a value and a directional derivative
// (forward mode), so the same passes that evaluate R(U) also evaluate J v = dR/dU v when the state
// is seeded with derivative v. Branches (flux cases, max / min, upwinding) follow the value, as the
// reference's FP64 code does; the derivative is that of the branch taken. The interface arrays are
// pairs (value, derivative) of doubles, i.e. complex128 in numpy.
struct Tangent {
  double v = 0.0, d = 0.0;
  constexpr Tangent() = default;
  constexpr Tangent(double value) : v(value) {}  // NOLINT: constants and tables enter as values
  constexpr Tangent(double value, double der) : v(value), d(der) {}
  Tangent& operator+=(Tangent b) { d += b.d; v += b.v; return *this; }
  Tangent& operator-=(Tangent b) { d -= b.d; v -= b.v; return *this; }
  Tangent& operator*=(Tangent b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
  Tangent& operator/=(Tangent b) { const double q = v / b.v; d = (d - q * b.d) / b.v; v = q; return *this; }
};
inline Tangent operator+(Tangent a, Tangent b) { return {a.v + b.v, a.d + b.d}; }
inline Tangent operator-(Tangent a, Tangent b) { return {a.v - b.v, a.d - b.d}; }
inline Tangent operator-(Tangent a) { return {-a.v, -a.d}; }
inline Tangent operator*(Tangent a, Tangent b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Tangent operator/(Tangent a, Tangent b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline bool operator<(Tangent a, Tangent b) { return a.v < b.v; }
inline bool operator>(Tangent a, Tangent b) { return a.v > b.v; }
inline bool operator<=(Tangent a, Tangent b) { return a.v <= b.v; }
inline bool operator>=(Tangent a, Tangent b) { return a.v >= b.v; }
inline bool operator==(Tangent a, Tangent b) { return a.v == b.v; }
inline bool operator!=(Tangent a, Tangent b) { return a.v != b.v; }
inline Tangent sqrt(Tangent a) {
  const double s = std::sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
inline Tangent fabs(Tangent a) { return a.v < 0.0 ? -a : a; }
inline Tangent exp(Tangent a) {
  const double e = std::exp(a.v);
  return {e, e * a.d};
}
inline Tangent log(Tangent a) { return {std::log(a.v), a.d / a.v}; }
inline Tangent log10(Tangent a) { return {std::log10(a.v), a.d / (a.v * 2.302585092994045684)}; }
inline Tangent sin(Tangent a) { return {std::sin(a.v), std::cos(a.v) * a.d}; }
inline Tangent pow(Tangent a, Tangent b) {  // a > 0
  const double p = std::pow(a.v, b.v);
  return {p, p * (b.v * a.d / a.v + std::log(a.v) * b.d)};
}
inline bool isnan(Tangent a) { return std::isnan(a.v); }
#pragma omp declare reduction(min : Tangent : omp_out = omp_in < omp_out ? omp_in : omp_out) \
    initializer(omp_priv = omp_orig)
#define ORACLE_REAL Tangent
namespace std {
template <>
struct numeric_limits<Tangent> {
  static constexpr Tangent max() noexcept { return Tangent(numeric_limits<double>::max()); }
};
}  // namespace std
#endif

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif

// loops over fewer items than this run on the calling thread: for small meshes the per-pass
// fork/join of the thread team costs more than the work (results do not depend on it: the loops
// are over independent elements or faces, and the reductions are exact minima)
constexpr int64_t kOmpMinItems = 256;

using std::exp;
using std::fabs;
using std::isnan;
using std::log;
using std::log10;
using std::pow;
using std::sin;
using std::sqrt;

namespace {

using Real = ORACLE_REAL;

constexpr Real kGamma = 1.4;        // PhysicalModel.cpp:45
constexpr Real kPrandtl = 0.71;     // PhysicalModel.cpp:91,107
constexpr Real kSutherland = 110.4 / 273.15;  // PhysicalModel.cpp:109

struct Physics {
  int eos = SDG_EOS_IDEAL_GAS;
  int transport = SDG_TRANSPORT_NONE;
  Real cp = 0, cv = 0, mu = 0, kappa = 0, c0 = 0, rho0 = 0, padd = 0;
  int source = SDG_SOURCE_NONE;  // SourceTerm.cpp:29-58
  Real beta = 0, tref = 0;

  // ThermodynamicModel<Constant> (PhysicalModel.cpp:24-38)
  Real internalEnergyFromTemperature(Real t) const { return cv * t; }
  Real temperatureFromInternalEnergy(Real e) const { return e / cv; }
  // EquationOfState (PhysicalModel.cpp:40-75)
  Real pressure(Real rho, Real e) const {
    if (eos == SDG_EOS_IDEAL_GAS) return (kGamma - 1.0) * rho * e;
    return c0 * c0 * (rho - rho0) + padd;
  }
  Real soundSpeed(Real rho, Real p) const {
    if (eos == SDG_EOS_IDEAL_GAS) return sqrt(kGamma * p / rho);
    return c0;
  }
  // TransportModel (PhysicalModel.cpp:77-122)
  Real sutherlandRatio(Real t) const {
    return sqrt(t * t * t) * (1.0 + kSutherland) / (t + kSutherland);
  }
  Real viscosity(Real t) const { return transport == SDG_TRANSPORT_SUTHERLAND ? mu * sutherlandRatio(t) : mu; }
  Real conductivity(Real t) const {
    return transport == SDG_TRANSPORT_SUTHERLAND ? kappa * sutherlandRatio(t) : kappa;
  }
};

// column-major helpers: C (m x n) (+)= A (m x k) * B (k x n)
inline void gemm(Real* c, const Real* a, const Real* b, int m, int k, int n, bool accumulate, Real alpha = 1.0) {
  for (int j = 0; j < n; j++) {
    Real* cj = c + static_cast<size_t>(j) * m;
    if (!accumulate)
      for (int i = 0; i < m; i++) cj[i] = 0.0;
    for (int l = 0; l < k; l++) {
      const Real blj = alpha * b[l + static_cast<size_t>(j) * k];
      const Real* al = a + static_cast<size_t>(l) * m;
      for (int i = 0; i < m; i++) cj[i] += al[i] * blj;
    }
  }
}
// C (m x n) (+)= A (m x k) * B^T where B is (n x k)
inline void gemmBt(Real* c, const Real* a, const Real* b, int m, int k, int n, bool accumulate) {
  for (int j = 0; j < n; j++) {
    Real* cj = c + static_cast<size_t>(j) * m;
    if (!accumulate)
      for (int i = 0; i < m; i++) cj[i] = 0.0;
    for (int l = 0; l < k; l++) {
      const Real blj = b[j + static_cast<size_t>(l) * n];
      const Real* al = a + static_cast<size_t>(l) * m;
      for (int i = 0; i < m; i++) cj[i] += al[i] * blj;
    }
  }
}

struct Oracle {
  // sizes
  int dim = 0, nv = 0, nb = 0, nq = 0, nf = 0, nfq = 0, naq = 0, order = 0;
  int64_t ne = 0, nface = 0;
  int equation = 0, flux = 0, viscous = 0, rk = 0;
  Physics phys;
  // operands (column-major like Eigen)
  std::vector<Real> V;    // N_q x N_b
  std::vector<Real> dV;   // (N_q D) x N_b
  std::vector<Real> Vf;   // N_aq x N_b
  std::vector<Real> lsq;  // N_b x N_b
  std::vector<Real> detjw, tiw, minv, min_edge;
  std::vector<int32_t> fparent, flocal, fseq;
  std::vector<Real> fnormal, fjw;
  // per-element solver data (SolveControl.cpp:45-174), flattened element-major
  std::vector<Real> coeff, coeff_last, vol_quad, adj_quad, residual, src_quad;
  std::vector<Real> vg_coeff, vg_quad, vg_adj, vg_res;           // volume gradient
  std::vector<Real> ig_adj, ig_res, ig_coeff, g_coeff;           // interface / total gradient (NS)
  int nig = 0;  // interface-gradient residual blocks per element: 1 (BR1) or N_f (BR2)
  std::vector<Real> rel_err;
  bool has_state = false;
  // boundary faces (the element is the left parent): element, local face, type; per point the
  // outward normal, |J_f| w and the boundary data (boundary_dummy_variable_, computational)
  int64_t nbface = 0;
  std::vector<int32_t> bparent, blocal, btype;
  std::vector<Real> bnormal, bjw, bdummy;
  // artificial viscosity (ShockCapturingEnum::ArtificialViscosity): node tags (0-based), inner
  // radii, the P1 nodal functions at volume and face points (row-major N x nvert), the tolerance
  // k and factor f of setArtificialViscosity, node values and per-element vertex values
  bool av = false;
  int nvert = 0;
  int64_t nnode = 0;
  std::vector<int32_t> node_tag;
  std::vector<Real> inner_radius, nodal_q, nodal_f, node_av, elem_av;
  Real av_tol = 0.0, av_factor = 1.0;

  size_t szC() const { return static_cast<size_t>(nv) * nb; }
  size_t szG() const { return static_cast<size_t>(nv) * dim * nb; }
  bool ns() const { return equation == SDG_EQ_COMPRESSIBLE_NS; }

  // --- variable conversions (VariableConvertor.cpp:315-339, 574-620) ---
  // computational: (rho, u.., e, p)
  void computational(const Real* u, Real* w) const {
    const Real rho = u[0];
    w[0] = rho;
    Real v2 = 0.0;
    for (int d = 0; d < dim; d++) {
      w[1 + d] = u[1 + d] / rho;
      v2 += w[1 + d] * w[1 + d];
    }
    const Real e = u[dim + 1] / rho - v2 / 2.0;
    w[dim + 1] = e;
    w[dim + 2] = phys.pressure(rho, e);
  }
  Real velocitySquared(const Real* w) const {
    Real s = 0.0;
    for (int d = 0; d < dim; d++) s += w[1 + d] * w[1 + d];
    return s;
  }
  // primitive gradient from conserved gradient g (row var*D+dim); out: (NV*D): rho, u (D x D
  // column-major, (a,b) = du_b/dx_a), T
  void primitiveGradient(const Real* w, const Real* u, const Real* g, Real* pg) const {
    const Real rho = w[0];
    const Real* drho = g;
    for (int a = 0; a < dim; a++) pg[a] = drho[a];
    Real vg[9];
    for (int b = 0; b < dim; b++)
      for (int a = 0; a < dim; a++) {
        vg[a + dim * b] = (g[(1 + b) * dim + a] - drho[a] * w[1 + b]) / rho;
        pg[(1 + b) * dim + a] = vg[a + dim * b];
      }
    const Real etot = u[dim + 1] / rho;
    for (int a = 0; a < dim; a++) {
      Real vgu = 0.0;
      for (int b = 0; b < dim; b++) vgu += vg[a + dim * b] * w[1 + b];
      const Real de = (g[(dim + 1) * dim + a] - drho[a] * etot) / rho - vgu;
      pg[(dim + 1) * dim + a] = phys.temperatureFromInternalEnergy(de);
    }
  }

  // --- fluxes (ConvectiveFlux.cpp:27-439, ViscousFlux.cpp:58-170) ---
  void normalFlux(const Real* n, const Real* w, Real* f) const {  // :59-91
    const Real rho = w[0];
    Real un = 0.0;
    for (int d = 0; d < dim; d++) un += w[1 + d] * n[d];
    const Real p = w[dim + 2];
    f[0] = rho * un;
    for (int d = 0; d < dim; d++) f[1 + d] = rho * un * w[1 + d] + p * n[d];
    const Real et = w[dim + 1] + velocitySquared(w) / 2.0;
    f[dim + 1] = (rho * et + p) * un;
  }
  void convectiveFlux(const Real* n, const Real* ul, const Real* wl, const Real* ur, const Real* wr,
                      Real* h) const {
    Real fl[5], fr[5];
    if (flux == SDG_FLUX_CENTRAL) {  // :93-104
      normalFlux(n, wl, fl);
      normalFlux(n, wr, fr);
      for (int v = 0; v < nv; v++) h[v] = (fl[v] + fr[v]) / 2.0;
    } else if (flux == SDG_FLUX_LAX_FRIEDRICHS) {  // :106-134
      normalFlux(n, wl, fl);
      normalFlux(n, wr, fr);
      Real unl = 0, unr = 0;
      for (int d = 0; d < dim; d++) {
        unl += wl[1 + d] * n[d];
        unr += wr[1 + d] * n[d];
      }
      const Real cl = phys.soundSpeed(wl[0], wl[dim + 2]);
      const Real cr = phys.soundSpeed(wr[0], wr[dim + 2]);
      const Real lam = std::max(fabs(unl) + cl, fabs(unr) + cr);
      for (int v = 0; v < nv; v++) h[v] = ((fl[v] + fr[v]) - lam * (ur[v] - ul[v])) / 2.0;
    } else if (flux == SDG_FLUX_HLLC) {  // :136-238
      const Real rl = wl[0], rr = wr[0], pl = wl[dim + 2], pr = wr[dim + 2];
      Real unl = 0, unr = 0;
      for (int d = 0; d < dim; d++) {
        unl += wl[1 + d] * n[d];
        unr += wr[1 + d] * n[d];
      }
      const Real cl = phys.soundSpeed(rl, pl), cr = phys.soundSpeed(rr, pr);
      const Real rbar = (rl + rr) / 2.0, cbar = (cl + cr) / 2.0;
      const Real ps = std::max(Real(0), (pl + pr) / 2.0 - (unr - unl) * rbar * cbar);
      const Real sl =
          unl - cl * (ps <= pl ? 1.0 : sqrt(1.0 + (kGamma + 1.0) * (ps / pl - 1.0) / 2.0 / kGamma));
      if (sl >= 0.0) {
        normalFlux(n, wl, h);
        return;
      }
      const Real sr =
          unr + cr * (ps <= pr ? 1.0 : sqrt(1.0 + (kGamma + 1.0) * (ps / pr - 1.0) / 2.0 / kGamma));
      if (sr <= 0.0) {
        normalFlux(n, wr, h);
        return;
      }
      const Real sm = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / (rl * (sl - unl) - rr * (sr - unr));
      Real us[5];
      if (sm >= 0.0) {
        normalFlux(n, wl, fl);
        us[0] = rl * (sl - unl) / (sl - sm);
        for (int d = 0; d < dim; d++) us[1 + d] = ((sl - unl) * rl * wl[1 + d] + (ps - pl) * n[d]) / (sl - sm);
        us[dim + 1] = ((sl - unl) * rl * (wl[dim + 1] + velocitySquared(wl) / 2.0) - pl * unl + ps * sm) / (sl - sm);
        for (int v = 0; v < nv; v++) h[v] = fl[v] + sl * (us[v] - ul[v]);
      } else {
        normalFlux(n, wr, fr);
        us[0] = rr * (sr - unr) / (sr - sm);
        for (int d = 0; d < dim; d++) us[1 + d] = ((sr - unr) * rr * wr[1 + d] + (ps - pr) * n[d]) / (sr - sm);
        us[dim + 1] = ((sr - unr) * rr * (wr[dim + 1] + velocitySquared(wr) / 2.0) - pr * unr + ps * sm) / (sr - sm);
        for (int v = 0; v < nv; v++) h[v] = fr[v] + sr * (us[v] - ur[v]);
      }
    } else if (flux == SDG_FLUX_ROE) {  // :240-350
      normalFlux(n, wl, fl);
      normalFlux(n, wr, fr);
      const Real sql = sqrt(wl[0]), sqr = sqrt(wr[0]);
      const Real sqs = sql + sqr;
      const Real rho = sqrt(wl[0] * wr[0]);
      Real vel[3];
      for (int d = 0; d < dim; d++) vel[d] = (sql * wl[1 + d] + sqr * wr[1 + d]) / sqs;
      const Real hl = wl[dim + 1] * kGamma + velocitySquared(wl) / 2.0;
      const Real hr = wr[dim + 1] * kGamma + velocitySquared(wr) / 2.0;
      const Real htot = (sql * hl + sqr * hr) / sqs;
      Real v2 = 0.0, un = 0.0;
      for (int d = 0; d < dim; d++) {
        v2 += vel[d] * vel[d];
        un += vel[d] * n[d];
      }
      const Real e = (htot - v2 / 2.0) / kGamma;
      const Real p = phys.pressure(rho, e);
      const Real c = phys.soundSpeed(rho, p);
      Real dw[8];
      for (int i = 0; i < dim + 3; i++) dw[i] = wr[i] - wl[i];
      Real dun = 0.0, vdv = 0.0;
      for (int d = 0; d < dim; d++) {
        dun += dw[1 + d] * n[d];
        vdv += vel[d] * dw[1 + d];
      }
      const Real delta = c / 20.0;
      const Real a1 = fabs(un - c) > delta ? fabs(un - c) : ((un - c) * (un - c) + delta * delta) / (2.0 * delta);
      const Real a3 = fabs(un + c) > delta ? fabs(un + c) : ((un + c) * (un + c) + delta * delta) / (2.0 * delta);
      const Real dp = dw[dim + 2], drho = dw[0];
      Real col[5][5] = {};
      const Real s0 = a1 * (dp - rho * c * dun) / (2.0 * c * c);
      col[0][0] = 1.0 * s0;
      for (int d = 0; d < dim; d++) col[0][1 + d] = (vel[d] - c * n[d]) * s0;
      col[0][dim + 1] = (htot - c * un) * s0;
      const Real s1 = fabs(un) * (drho - dp / (c * c));
      col[1][0] = 1.0 * s1;
      for (int d = 0; d < dim; d++) col[1][1 + d] = vel[d] * s1;
      col[1][dim + 1] = v2 / 2.0 * s1;
      if (dim >= 2) {
        const Real s2 = fabs(un) * rho;
        col[2][0] = 0.0 * s2;
        for (int d = 0; d < dim; d++) col[2][1 + d] = (dw[1 + d] - dun * n[d]) * s2;
        col[2][dim + 1] = (vdv - un * dun) * s2;
      }
      const Real s3 = a3 * (dp + rho * c * dun) / (2.0 * c * c);
      col[dim + 1][0] = 1.0 * s3;
      for (int d = 0; d < dim; d++) col[dim + 1][1 + d] = (vel[d] + c * n[d]) * s3;
      col[dim + 1][dim + 1] = (htot + c * un) * s3;
      for (int v = 0; v < nv; v++) {
        Real s = 0.0;
        for (int k = 0; k < nv; k++) s += col[k][v];  // rowwise().sum()
        h[v] = ((fl[v] + fr[v]) - s) / 2.0;
      }
    } else {  // Exact (weakly compressible), :352-414
      const Real c = phys.soundSpeed(0.0, 0.0);
      Real unl = 0, unr = 0;
      for (int d = 0; d < dim; d++) {
        unl += wl[1 + d] * n[d];
        unr += wr[1 + d] * n[d];
      }
      const Real rho = sqrt(wl[0] * wr[0] * exp((unl - unr) / c));
      const Real un = (unl + unr) / 2.0 + log(wl[0] / wr[0]) * c / 2.0;
      Real we[8];
      const Real* ws = un < 0.0 ? wr : wl;
      const Real uns = un < 0.0 ? unr : unl;
      const Real e = ws[dim + 1] * ws[0] / rho;
      we[0] = rho;
      for (int d = 0; d < dim; d++) we[1 + d] = ws[1 + d] + (un - uns) * n[d];
      we[dim + 1] = e;
      we[dim + 2] = phys.pressure(rho, e);
      normalFlux(n, we, h);
    }
  }
  // viscous raw flux F_v (D x NV, column-major), ViscousFlux.cpp:58-103
  void viscousRawFlux(const Real* w, const Real* pg, Real* fv) const {
    const Real t = phys.temperatureFromInternalEnergy(w[dim + 1]);
    const Real mu = phys.viscosity(t);
    const Real* vg = pg + dim;  // D x D column-major, (a,b) = du_b/dx_a
    Real tr = 0.0;
    for (int a = 0; a < dim; a++) tr += vg[a + dim * a];
    for (int a = 0; a < dim; a++) fv[a] = 0.0;
    Real tau[9];
    for (int b = 0; b < dim; b++)
      for (int a = 0; a < dim; a++)
        tau[a + dim * b] = mu * (vg[a + dim * b] + vg[b + dim * a]) - (a == b ? 2.0 / 3.0 * mu * tr * 1.0 : 2.0 / 3.0 * mu * tr * 0.0);
    for (int b = 0; b < dim; b++)
      for (int a = 0; a < dim; a++) fv[a + dim * (1 + b)] = tau[a + dim * b];
    const Real k = phys.conductivity(t);
    const Real* tg = pg + (dim + 1) * dim;
    for (int a = 0; a < dim; a++) {
      Real s = 0.0;
      for (int b = 0; b < dim; b++) s += tau[a + dim * b] * w[1 + b];
      fv[a + dim * (dim + 1)] = s + k * tg[a];
    }
  }
  void viscousNormalFlux(const Real* n, const Real* w, const Real* pg, Real* f) const {  // :115-124
    Real fv[15];
    viscousRawFlux(w, pg, fv);
    for (int v = 0; v < nv; v++) {
      Real s = 0.0;
      for (int a = 0; a < dim; a++) s += fv[a + dim * v] * n[a];
      f[v] = s;
    }
  }

  int accumulate(int s) const { return s * nfq; }  // getElementAccumulateAdjacencyQuadratureNumber

  // --- boundary conditions (BoundaryCondition.cpp:79-547), computational variables ---
  // calculateBoundaryVariable: boundary state wb from the interior state wl and boundary data wd
  void boundaryVariable(int type, const Real* n, const Real* wl, const Real* wd, Real* wb) const {
    const int nw = dim + 3;
    Real unl = 0.0;
    for (int d = 0; d < dim; d++) unl += wl[1 + d] * n[d];
    const Real mach = unl / phys.soundSpeed(wl[0], wl[dim + 2]);
    switch (type) {
      case 0: {  // RiemannFarfield :79-247
        if (fabs(mach) > 1.0) {
          const Real* src = mach < 0.0 ? wd : wl;  // supersonic inflow / outflow
          for (int i = 0; i < nw; i++) wb[i] = src[i];
          break;
        }
        const Real gm1 = kGamma - 1.0;
        Real und = 0.0;
        for (int d = 0; d < dim; d++) und += wd[1 + d] * n[d];
        const Real left_toward = und - 2.0 * phys.soundSpeed(wd[0], wd[dim + 2]) / gm1;
        const Real right_toward = unl + 2.0 * phys.soundSpeed(wl[0], wl[dim + 2]) / gm1;
        const Real unb = (left_toward + right_toward) / 2.0;
        const Real* ref = mach < 0.0 ? wd : wl;  // subsonic inflow: outside, outflow: inside
        Real unr = 0.0;
        for (int d = 0; d < dim; d++) unr += ref[1 + d] * n[d];
        for (int d = 0; d < dim; d++) wb[1 + d] = ref[1 + d] + (unb - unr) * n[d];
        const Real c = gm1 * (right_toward - left_toward) / 4.0;
        const Real entropy = ref[dim + 2] / pow(ref[0], kGamma);
        const Real rho = pow(c * c / (kGamma * entropy), 1.0 / gm1);
        const Real p = rho * c * c / kGamma;
        wb[0] = rho;
        wb[dim + 1] = p / (gm1 * rho);
        wb[dim + 2] = p;
        break;
      }
      case 1:  // VelocityInflow :309-333
        for (int i = 0; i < nw; i++) wb[i] = wd[i];
        if (mach > -1.0) wb[dim + 2] = wl[dim + 2];
        break;
      case 2:  // PressureOutflow :362-386
        for (int i = 0; i < nw; i++) wb[i] = wl[i];
        if (mach < 1.0) wb[dim + 2] = wd[dim + 2];
        break;
      case 3:  // IsoThermalNonSlipWall :416-434
        wb[0] = wl[0];
        for (int d = 0; d < dim; d++) wb[1 + d] = wd[1 + d];
        wb[dim + 1] = wd[dim + 1];
        wb[dim + 2] = phys.pressure(wl[0], wd[dim + 1]);
        break;
      case 4:  // AdiabaticSlipWall :474-488
        for (int i = 0; i < nw; i++) wb[i] = wl[i];
        for (int d = 0; d < dim; d++) wb[1 + d] = wl[1 + d] - unl * n[d];
        break;
      default:  // AdiabaticNonSlipWall :521-533
        for (int i = 0; i < nw; i++) wb[i] = wl[i];
        for (int d = 0; d < dim; d++) wb[1 + d] = wd[1 + d];
        break;
    }
  }
  void conservedFromComputational(const Real* w, Real* u) const {
    u[0] = w[0];
    for (int d = 0; d < dim; d++) u[1 + d] = w[0] * w[1 + d];
    u[dim + 1] = w[0] * (w[dim + 1] + velocitySquared(w) / 2.0);
  }
  void boundaryTraces(int64_t b, Real* ul) const {  // NV x N_fq of the boundary face's element
    const int64_t e = bparent[b];
    const int s = blocal[b];
    for (int j = 0; j < nfq; j++)
      for (int v = 0; v < nv; v++) {
        Real a = 0.0;
        for (int i = 0; i < nb; i++) a += coeff[e * szC() + v + nv * i] * Vf[(accumulate(s) + j) + static_cast<size_t>(naq) * i];
        ul[v + nv * j] = a;
      }
  }
  void boundaryGradientQuadrature() {  // SpatialDiscrete.cpp:913-968 with calculateBoundaryGradientVariable
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(nbface >= kOmpMinItems)
    for (int64_t b = 0; b < nbface; b++) {
      std::vector<Real> ul(static_cast<size_t>(nv) * nfq);
      boundaryTraces(b, ul.data());
      const int64_t e = bparent[b];
      const int s = blocal[b];
      for (int j = 0; j < nfq; j++) {
        const size_t bp = static_cast<size_t>(b) * nfq + j;
        const Real* n = &bnormal[bp * dim];
        const Real jw = bjw[bp];
        const Real* u = &ul[nv * j];
        Real vg[5], ig[5];
        if (btype[b] <= 2) {  // farfield, inflow, outflow: interior state, no interface term
          for (int v = 0; v < nv; v++) {
            vg[v] = u[v];
            ig[v] = 0.0;
          }
        } else {  // walls: conserved boundary state and its jump
          Real wl[8], wb[8];
          computational(u, wl);
          boundaryVariable(btype[b], n, wl, &bdummy[bp * (dim + 3)], wb);
          conservedFromComputational(wb, vg);
          for (int v = 0; v < nv; v++) ig[v] = vg[v] - u[v];
        }
        Real* cl = &vg_adj[static_cast<size_t>(e) * rows * naq + static_cast<size_t>(rows) * (accumulate(s) + j)];
        for (int v = 0; v < nv; v++)
          for (int a = 0; a < dim; a++) cl[v * dim + a] = (n[a] * vg[v]) * jw;
        if (ns()) {
          Real* il = &ig_adj[static_cast<size_t>(e) * rows * naq + static_cast<size_t>(rows) * (accumulate(s) + j)];
          for (int v = 0; v < nv; v++)
            for (int a = 0; a < dim; a++) il[v * dim + a] = (n[a] * ig[v]) * jw;
        }
      }
    }
  }
  void boundaryQuadrature() {  // SpatialDiscrete.cpp:750-842
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(nbface >= kOmpMinItems)
    for (int64_t b = 0; b < nbface; b++) {
      std::vector<Real> ul(static_cast<size_t>(nv) * nfq), gl;
      boundaryTraces(b, ul.data());
      const int64_t e = bparent[b];
      const int s = blocal[b];
      if (ns()) {
        gl.resize(static_cast<size_t>(rows) * nfq);
        faceGradientTrace(e, s, gl.data());
      }
      std::vector<Real> vl;
      if (av) {
        vl.resize(static_cast<size_t>(rows) * nfq);
        faceVolumeGradientTrace(e, s, vl.data());
      }
      for (int j = 0; j < nfq; j++) {
        const size_t bp = static_cast<size_t>(b) * nfq + j;
        const Real* n = &bnormal[bp * dim];
        const Real* u = &ul[nv * j];
        Real wl[8], wb[8], h[5], tmp[5];
        computational(u, wl);
        boundaryVariable(btype[b], n, wl, &bdummy[bp * (dim + 3)], wb);
        normalFlux(n, wb, h);  // calculateConvectiveNormalFlux of the boundary state
        if (ns()) {
          Real pgl[15], pgb[15], fl[5], fb[5], wm[8];
          primitiveGradient(wl, u, &gl[static_cast<size_t>(rows) * j], pgl);
          for (int r = 0; r < rows; r++) pgb[r] = pgl[r];
          if (btype[b] == 4 || btype[b] == 5)  // adiabatic walls: no temperature gradient
            for (int a = 0; a < dim; a++) pgb[(dim + 1) * dim + a] = 0.0;
          for (int i = 0; i < dim + 3; i++) wm[i] = btype[b] >= 3 ? wb[i] : wl[i];  // modifyBoundaryVariable
          viscousNormalFlux(n, wm, pgl, fl);
          viscousNormalFlux(n, wb, pgb, fb);
          for (int v = 0; v < nv; v++) tmp[v] = (h[v] - (fl[v] + fb[v]) / 2.0) * bjw[bp];
        } else {
          for (int v = 0; v < nv; v++) tmp[v] = h[v] * bjw[bp];
        }
        if (av) {  // the element's own value (SpatialDiscrete.cpp:812-830)
          Real al[5];
          avNormalFlux(n, avAt(nodal_f, accumulate(s) + j, e), &vl[static_cast<size_t>(rows) * j], al);
          for (int v = 0; v < nv; v++) tmp[v] -= al[v] * bjw[bp];
        }
        Real* cl = &adj_quad[static_cast<size_t>(e) * nv * naq + static_cast<size_t>(nv) * (accumulate(s) + j)];
        for (int v = 0; v < nv; v++) cl[v] = tmp[v];
      }
    }
  }

  // --- artificial viscosity (SpatialDiscrete.cpp:37-192) ---
  // kArtificialViscosityTolerance (SimulationControl.cpp:892-893): log10(1/P^4) to 12 digits, 0 for P = 1
  static Real sensorThreshold(int p) {
    static const Real t[5] = {0.0, -1.20411998266, -1.90848501888, -2.40823996531, -2.79588001734};
    return t[p - 1];
  }
  // number of basis functions of order P-1 (the hierarchic basis of order P-1 is a prefix)
  int lowModes() const {
    if (order == 1) return 0;
    const int p = order - 1;
    if (dim == 2 && nvert == 3) return (p + 1) * (p + 2) / 2;
    if (dim == 3 && nvert == 4) return (p + 1) * (p + 2) * (p + 3) / 6;  // tetrahedra
    if (dim == 3 && nvert == 5) return (p + 1) * (p + 2) * (2 * p + 3) / 6;  // pyramids
    if (dim == 1) return p + 1;
    return dim == 2 ? (p + 1) * (p + 1) : (p + 1) * (p + 1) * (p + 1);
  }
  void artificialViscosity() {
    const Real s0 = sensorThreshold(order);
    const int lo = lowModes();
    std::vector<Real> eps(static_cast<size_t>(ne));
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {  // calculateElementArtificialViscosity, :37-88
      Real num = 0.0, den = 0.0;
      for (int q = 0; q < nq; q++) {
        Real all = 0.0, high = 0.0;
        for (int b = 0; b < nb; b++) all += V[q + static_cast<size_t>(nq) * b] * coeff[e * szC() + nv * b];
        for (int b = lo; b < nb; b++) high += V[q + static_cast<size_t>(nq) * b] * coeff[e * szC() + nv * b];
        const Real jw = detjw[static_cast<size_t>(e) * nq + q];
        num += high * (high * jw);
        den += all * (all * jw);
      }
      const Real s = log10(num / den);
      const Real top = av_factor * (inner_radius[e] / static_cast<Real>(order));
      if (s < s0 - av_tol)
        eps[e] = 0.0;
      else if (s > s0 + av_tol)
        eps[e] = top;
      else
        eps[e] = top * (1.0 + sin(Real(3.14159265358979323846) * (s - s0) / (2.0 * av_tol))) / 2.0;
    }
    std::fill(node_av.begin(), node_av.end(), Real(0.0));  // maxElementArtificialViscosity, :90-107
    for (int64_t e = 0; e < ne; e++)
      for (int j = 0; j < nvert; j++) {
        Real& n = node_av[node_tag[static_cast<size_t>(e) * nvert + j]];
        n = std::max(n, eps[e]);
      }
    for (int64_t e = 0; e < ne; e++)  // storeElementArtificialViscosity, :109-121
      for (int j = 0; j < nvert; j++)
        elem_av[static_cast<size_t>(e) * nvert + j] = node_av[node_tag[static_cast<size_t>(e) * nvert + j]];
  }
  // P1 interpolant of the element's vertex values at row `row` of a nodal table
  Real avAt(const std::vector<Real>& table, int row, int64_t e) const {
    Real s = 0.0;
    for (int j = 0; j < nvert; j++)
      s += table[static_cast<size_t>(row) * nvert + j] * elem_av[static_cast<size_t>(e) * nvert + j];
    return s;
  }
  // volume-gradient coefficients at the face points of side s (get<ViscousFluxEnum::None>)
  void faceVolumeGradientTrace(int64_t e, int s, Real* gt) const {
    const int rows = nv * dim;
    for (int j = 0; j < nfq; j++)
      for (int r = 0; r < rows; r++) {
        Real acc = 0.0;
        for (int i = 0; i < nb; i++)
          acc += vg_coeff[e * szG() + r + static_cast<size_t>(rows) * i] * Vf[(accumulate(s) + j) + static_cast<size_t>(naq) * i];
        gt[r + static_cast<size_t>(rows) * j] = acc;
      }
  }
  // artificial viscous normal flux eps g^T n per variable (ViscousFlux.cpp:105-134)
  void avNormalFlux(const Real* n, Real eps, const Real* g, Real* out) const {
    for (int v = 0; v < nv; v++) {
      Real s = 0.0;
      for (int a = 0; a < dim; a++) s += (eps * g[v * dim + a]) * n[a];
      out[v] = s;
    }
  }

  // --- passes of stepSolver ---
  void gradientQuadrature() {  // SpatialDiscrete.cpp:294-346
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      std::vector<Real> uq(static_cast<size_t>(nv) * nq);
      gemmBt(uq.data(), &coeff[e * szC()], V.data(), nv, nb, nq, false);
      Real* out = &vg_quad[static_cast<size_t>(e) * nv * dim * dim * nq];
      const int rows = nv * dim;
      for (int q = 0; q < nq; q++) {
        const Real* G = &tiw[(static_cast<size_t>(e) * nq + q) * dim * dim];
        for (int k = 0; k < nv; k++)
          for (int b = 0; b < dim; b++)
            for (int a = 0; a < dim; a++)
              out[(k * dim + a) + static_cast<size_t>(rows) * (q * dim + b)] = uq[k + static_cast<size_t>(nv) * q] * G[a + dim * b];
      }
    }
  }
  void faceTraces(int64_t f, Real* ul, Real* ur) const {
    const int32_t L = fparent[2 * f], R = fparent[2 * f + 1];
    const int sl = flocal[2 * f], sr = flocal[2 * f + 1];
    // NV x N_fq = C * Vf[face rows]^T
    for (int j = 0; j < nfq; j++)
      for (int v = 0; v < nv; v++) {
        Real a = 0.0, b = 0.0;
        for (int i = 0; i < nb; i++) {
          a += coeff[L * szC() + v + nv * i] * Vf[(accumulate(sl) + j) + static_cast<size_t>(naq) * i];
          b += coeff[R * szC() + v + nv * i] * Vf[(accumulate(sr) + j) + static_cast<size_t>(naq) * i];
        }
        ul[v + nv * j] = a;
        ur[v + nv * j] = b;
      }
  }
  void adjacencyGradientQuadrature() {  // SpatialDiscrete.cpp:844-911
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(nface >= kOmpMinItems)
    for (int64_t f = 0; f < nface; f++) {
      std::vector<Real> ul(static_cast<size_t>(nv) * nfq), ur(static_cast<size_t>(nv) * nfq);
      faceTraces(f, ul.data(), ur.data());
      const int32_t L = fparent[2 * f], R = fparent[2 * f + 1];
      const int sl = flocal[2 * f], sr = flocal[2 * f + 1];
      for (int j = 0; j < nfq; j++) {
        const int k = fseq[f * nfq + j];
        const Real* n = &fnormal[(static_cast<size_t>(f) * nfq + j) * dim];
        const Real jw = fjw[f * nfq + j];
        Real tmp[15];
        for (int v = 0; v < nv; v++)
          for (int a = 0; a < dim; a++) tmp[v * dim + a] = (n[a] * (ul[v + nv * j] + ur[v + nv * k]) / 2.0) * jw;
        Real* cl = &vg_adj[static_cast<size_t>(L) * rows * naq + static_cast<size_t>(rows) * (accumulate(sl) + j)];
        Real* cr = &vg_adj[static_cast<size_t>(R) * rows * naq + static_cast<size_t>(rows) * (accumulate(sr) + k)];
        for (int r = 0; r < rows; r++) {
          cl[r] = tmp[r];
          cr[r] = -tmp[r];
        }
        if (ns()) {
          for (int v = 0; v < nv; v++)
            for (int a = 0; a < dim; a++) tmp[v * dim + a] = (n[a] * (ur[v + nv * k] - ul[v + nv * j]) / 2.0) * jw;
          Real* il = &ig_adj[static_cast<size_t>(L) * rows * naq + static_cast<size_t>(rows) * (accumulate(sl) + j)];
          Real* ir = &ig_adj[static_cast<size_t>(R) * rows * naq + static_cast<size_t>(rows) * (accumulate(sr) + k)];
          for (int r = 0; r < rows; r++) {
            il[r] = tmp[r];
            ir[r] = tmp[r];
          }
        }
      }
    }
  }
  void gradientResidual() {  // SpatialDiscrete.cpp:1034-1068
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      Real* res = &vg_res[e * szG()];
      gemm(res, &vg_adj[static_cast<size_t>(e) * rows * naq], Vf.data(), rows, naq, nb, false);
      gemm(res, &vg_quad[static_cast<size_t>(e) * rows * dim * nq], dV.data(), rows, dim * nq, nb, true, -1.0);
      if (ns()) {
        if (viscous == SDG_VISC_BR1) {
          gemm(&ig_res[e * szG()], &ig_adj[static_cast<size_t>(e) * rows * naq], Vf.data(), rows, naq, nb, false);
        } else {
          for (int s = 0; s < nf; s++) {
            Real* out = &ig_res[(static_cast<size_t>(e) * nf + s) * szG()];
            for (int j = 0; j < nb; j++)
              for (int r = 0; r < rows; r++) {
                Real acc = 0.0;
                for (int p = 0; p < nfq; p++)
                  acc += ig_adj[static_cast<size_t>(e) * rows * naq + r + static_cast<size_t>(rows) * (accumulate(s) + p)] *
                         Vf[(accumulate(s) + p) + static_cast<size_t>(naq) * j];
                out[r + static_cast<size_t>(rows) * j] = acc;
              }
          }
        }
      }
    }
  }
  void updateGradientCoefficient() {  // TimeIntegration.cpp:200-228
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      const Real* mi = &minv[static_cast<size_t>(e) * nb * nb];
      gemm(&vg_coeff[e * szG()], &vg_res[e * szG()], mi, rows, nb, nb, false);
      if (ns()) {
        Real* g = &g_coeff[e * szG()];
        std::memcpy(g, &vg_coeff[e * szG()], szG() * sizeof(Real));
        for (int s = 0; s < nig; s++) {
          Real* ic = &ig_coeff[(static_cast<size_t>(e) * nig + s) * szG()];
          gemm(ic, &ig_res[(static_cast<size_t>(e) * nig + s) * szG()], mi, rows, nb, nb, false);
          for (size_t i = 0; i < szG(); i++) g[i] += ic[i];
        }
      }
    }
  }
  void volumeQuadrature() {  // SpatialDiscrete.cpp:194-266
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      std::vector<Real> uq(static_cast<size_t>(nv) * nq), gq;
      gemmBt(uq.data(), &coeff[e * szC()], V.data(), nv, nb, nq, false);
      if (ns()) {
        gq.resize(static_cast<size_t>(rows) * nq);
        gemmBt(gq.data(), &g_coeff[e * szG()], V.data(), rows, nb, nq, false);
      }
      std::vector<Real> gv;
      if (av) {
        gv.resize(static_cast<size_t>(rows) * nq);
        gemmBt(gv.data(), &vg_coeff[e * szG()], V.data(), rows, nb, nq, false);
      }
      Real* out = &vol_quad[static_cast<size_t>(e) * nv * dim * nq];
      for (int q = 0; q < nq; q++) {
        Real w[8], fc[15], fv[15], pg[15];
        const Real* u = &uq[static_cast<size_t>(nv) * q];
        computational(u, w);
        if (phys.source == SDG_SOURCE_BOUSSINESQ)  // rho beta (T - T_ref) g |J| w, g = 1 (SourceTerm.cpp:41-55)
          src_quad[static_cast<size_t>(e) * nq + q] =
              w[0] * phys.beta * (phys.temperatureFromInternalEnergy(w[dim + 1]) - phys.tref) * 1.0 *
              detjw[static_cast<size_t>(e) * nq + q];
        // F_c (D x NV column-major), ConvectiveFlux.cpp:27-44
        const Real rho = w[0], p = w[dim + 2];
        const Real et = w[dim + 1] + velocitySquared(w) / 2.0;
        for (int a = 0; a < dim; a++) {
          fc[a] = rho * w[1 + a];
          for (int b = 0; b < dim; b++) fc[a + dim * (1 + b)] = rho * w[1 + a] * w[1 + b] + (a == b ? p * 1.0 : p * 0.0);
          fc[a + dim * (dim + 1)] = (rho * et + p) * w[1 + a];
        }
        if (ns()) {
          primitiveGradient(w, u, &gq[static_cast<size_t>(rows) * q], pg);
          viscousRawFlux(w, pg, fv);
        }
        const Real* G = &tiw[(static_cast<size_t>(e) * nq + q) * dim * dim];
        for (int v = 0; v < nv; v++)
          for (int b = 0; b < dim; b++) {
            Real s = 0.0;
            for (int a = 0; a < dim; a++) {
              const Real f = ns() ? (fc[a + dim * v] - fv[a + dim * v]) : fc[a + dim * v];
              s += f * G[a + dim * b];
            }
            out[v + static_cast<size_t>(nv) * (q * dim + b)] = s;
          }
        if (av) {  // - (eps grad U)^T G (SpatialDiscrete.cpp:211-215,226-228,245-249)
          const Real eps = avAt(nodal_q, q, e);
          const Real* g = &gv[static_cast<size_t>(rows) * q];
          for (int v = 0; v < nv; v++)
            for (int b = 0; b < dim; b++) {
              Real s = 0.0;
              for (int a = 0; a < dim; a++) s += (eps * g[v * dim + a]) * G[a + dim * b];
              out[v + static_cast<size_t>(nv) * (q * dim + b)] -= s;
            }
        }
      }
    }
  }
  void faceGradientTrace(int64_t e, int s, Real* gt) const {  // VariableConvertor.cpp:640-689
    const int rows = nv * dim;
    for (int j = 0; j < nfq; j++)
      for (int r = 0; r < rows; r++) {
        Real acc = 0.0;
        for (int i = 0; i < nb; i++) {
          Real c;
          if (viscous == SDG_VISC_BR2)
            c = vg_coeff[e * szG() + r + static_cast<size_t>(rows) * i] +
                ig_coeff[(static_cast<size_t>(e) * nig + s) * szG() + r + static_cast<size_t>(rows) * i];
          else
            c = g_coeff[e * szG() + r + static_cast<size_t>(rows) * i];
          acc += c * Vf[(accumulate(s) + j) + static_cast<size_t>(naq) * i];
        }
        gt[r + static_cast<size_t>(rows) * j] = acc;
      }
  }
  void adjacencyQuadrature() {  // SpatialDiscrete.cpp:633-748
    const int rows = nv * dim;
#pragma omp parallel for schedule(static) if(nface >= kOmpMinItems)
    for (int64_t f = 0; f < nface; f++) {
      std::vector<Real> ul(static_cast<size_t>(nv) * nfq), ur(static_cast<size_t>(nv) * nfq);
      std::vector<Real> wl(static_cast<size_t>(dim + 3) * nfq), wr(static_cast<size_t>(dim + 3) * nfq);
      std::vector<Real> gl, gr;
      faceTraces(f, ul.data(), ur.data());
      const int32_t L = fparent[2 * f], R = fparent[2 * f + 1];
      const int sl = flocal[2 * f], sr = flocal[2 * f + 1];
      for (int j = 0; j < nfq; j++) {
        computational(&ul[nv * j], &wl[(dim + 3) * j]);
        computational(&ur[nv * j], &wr[(dim + 3) * j]);
      }
      if (ns()) {
        gl.resize(static_cast<size_t>(rows) * nfq);
        gr.resize(static_cast<size_t>(rows) * nfq);
        faceGradientTrace(L, sl, gl.data());
        faceGradientTrace(R, sr, gr.data());
      }
      std::vector<Real> vl, vr;
      if (av) {
        vl.resize(static_cast<size_t>(rows) * nfq);
        vr.resize(static_cast<size_t>(rows) * nfq);
        faceVolumeGradientTrace(L, sl, vl.data());
        faceVolumeGradientTrace(R, sr, vr.data());
      }
      for (int j = 0; j < nfq; j++) {
        const int k = fseq[f * nfq + j];
        const Real* n = &fnormal[(static_cast<size_t>(f) * nfq + j) * dim];
        Real h[5], tmp[5];
        convectiveFlux(n, &ul[nv * j], &wl[(dim + 3) * j], &ur[nv * k], &wr[(dim + 3) * k], h);
        if (ns()) {
          Real pgl[15], pgr[15], fl[5], fr[5];
          primitiveGradient(&wl[(dim + 3) * j], &ul[nv * j], &gl[static_cast<size_t>(rows) * j], pgl);
          primitiveGradient(&wr[(dim + 3) * k], &ur[nv * k], &gr[static_cast<size_t>(rows) * k], pgr);
          viscousNormalFlux(n, &wl[(dim + 3) * j], pgl, fl);
          viscousNormalFlux(n, &wr[(dim + 3) * k], pgr, fr);
          for (int v = 0; v < nv; v++) tmp[v] = (h[v] - (fl[v] + fr[v]) / 2.0) * fjw[f * nfq + j];
        } else {
          for (int v = 0; v < nv; v++) tmp[v] = h[v] * fjw[f * nfq + j];
        }
        if (av) {  // central average of the two sides (ViscousFlux.cpp:172-187, SpatialDiscrete.cpp:714-735)
          Real al[5], ar[5];
          avNormalFlux(n, avAt(nodal_f, accumulate(sl) + j, L), &vl[static_cast<size_t>(rows) * j], al);
          avNormalFlux(n, avAt(nodal_f, accumulate(sr) + k, R), &vr[static_cast<size_t>(rows) * k], ar);
          for (int v = 0; v < nv; v++) tmp[v] -= ((al[v] + ar[v]) / 2.0) * fjw[f * nfq + j];
        }
        Real* cl = &adj_quad[static_cast<size_t>(L) * nv * naq + static_cast<size_t>(nv) * (accumulate(sl) + j)];
        Real* cr = &adj_quad[static_cast<size_t>(R) * nv * naq + static_cast<size_t>(nv) * (accumulate(sr) + k)];
        for (int v = 0; v < nv; v++) {
          cl[v] = tmp[v];
          cr[v] = -tmp[v];
        }
      }
    }
  }
  void calcResidual() {  // SpatialDiscrete.cpp:1016-1032
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      Real* res = &residual[e * szC()];
      gemm(res, &vol_quad[static_cast<size_t>(e) * nv * dim * nq], dV.data(), nv, dim * nq, nb, false);
      gemm(res, &adj_quad[static_cast<size_t>(e) * nv * naq], Vf.data(), nv, naq, nb, true, -1.0);
      if (phys.source == SDG_SOURCE_BOUSSINESQ)  // + S V on the last momentum row (SpatialDiscrete.cpp:1026-1029)
        for (int b = 0; b < nb; b++) {
          Real s = 0.0;
          for (int q = 0; q < nq; q++) s += src_quad[static_cast<size_t>(e) * nq + q] * V[q + static_cast<size_t>(nq) * b];
          res[dim + static_cast<size_t>(nv) * b] += s;
        }
    }
  }
  void rkCoefficients(int stage, Real c[3]) const {  // TimeIntegration.cpp:45-65
    if (rk == SDG_RK_FORWARD_EULER) {
      c[0] = 1.0; c[1] = 0.0; c[2] = 1.0;
    } else if (rk == SDG_RK_HEUN2) {
      const Real t[2][3] = {{1.0, 0.0, 1.0}, {0.5, 0.5, 0.5}};
      for (int i = 0; i < 3; i++) c[i] = t[stage][i];
    } else {
      const Real t[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
      for (int i = 0; i < 3; i++) c[i] = t[stage][i];
    }
  }
  int stages() const { return rk == SDG_RK_FORWARD_EULER ? 1 : (rk == SDG_RK_HEUN2 ? 2 : 3); }
  void updateCoefficient(int stage, Real dt) {  // TimeIntegration.cpp:181-198
    Real c[3];
    rkCoefficients(stage, c);
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      Real* u = &coeff[e * szC()];
      const Real* un = &coeff_last[e * szC()];
      std::vector<Real> rm(szC());
      gemm(rm.data(), &residual[e * szC()], &minv[static_cast<size_t>(e) * nb * nb], nv, nb, nb, false);
      for (size_t i = 0; i < szC(); i++) u[i] *= c[1];
      for (size_t i = 0; i < szC(); i++) u[i] += c[0] * un[i];
      for (size_t i = 0; i < szC(); i++) u[i] += c[2] * dt * rm[i];
    }
  }
  void relativeError() {  // TimeIntegration.cpp:279-324
    std::vector<Real> acc(nv, 0.0);
    std::vector<Real> per(static_cast<size_t>(ne) * nv);
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      std::vector<Real> rq(static_cast<size_t>(nv) * nq);
      gemmBt(rq.data(), &residual[e * szC()], V.data(), nv, nb, nq, false);
      for (int v = 0; v < nv; v++) {
        Real s = 0.0;
        for (int q = 0; q < nq; q++) s += fabs(rq[v + static_cast<size_t>(nv) * q]);
        per[e * nv + v] = s / nq;
      }
    }
    for (int64_t e = 0; e < ne; e++)
      for (int v = 0; v < nv; v++) acc[v] += per[e * nv + v];
    for (int v = 0; v < nv; v++) rel_err[v] = acc[v] / static_cast<Real>(ne);
  }
  void spatial() {
    gradientQuadrature();
    adjacencyGradientQuadrature();
    boundaryGradientQuadrature();
    gradientResidual();
    updateGradientCoefficient();
    volumeQuadrature();
    adjacencyQuadrature();
    boundaryQuadrature();
    calcResidual();
  }
  void step(Real dt) {  // TimeIntegration.cpp:327-350
    coeff_last = coeff;
    if (av) artificialViscosity();
    for (int s = 0; s < stages(); s++) {
      spatial();
      updateCoefficient(s, dt);
    }
    relativeError();
  }
};

// Euler solver on a 2-D mesh of triangles and quadrangles: the reference holds one ElementSolver
// per element type (Solver::triangle_, quadrangle_) and runs each pass over both, the faces
// coupling elements of either type (TimeIntegration.cpp:279-350, SpatialDiscrete.cpp:194-266,
// 633-748, 1016-1032). Same passes and arithmetic as Oracle's Euler path; per-type operand
// tables, per-element blocks of each element's own size. Periodic meshes (no boundary faces).
struct MixedOracle {
  Oracle pt;  // pointwise physics, flux and RK tables (dim 2, NV 4)
  struct Part {
    int nb = 0, nq = 0, nf = 0, nfq = 0, naq = 0;
    std::vector<Real> V, dV, Vf, lsq;  // column-major like Oracle's
  } part[2];                           // [0] triangle, [1] quadrangle
  int64_t ne = 0, nface = 0;
  std::vector<int8_t> etype;
  std::vector<int64_t> coff, qoff, aoff, moff;  // per element: coefficient, point, slot, mass offsets
  std::vector<Real> detjw, tiw, minv, min_edge;
  std::vector<int32_t> fparent, flocal, fseq;
  std::vector<Real> fnormal, fjw;
  std::vector<Real> coeff, coeff_last, vol_quad, adj_quad, residual, rel_err;
  static constexpr int kNv = 4, kDim = 2;

  const Part& of(int64_t e) const { return part[etype[e]]; }
  void faceTrace(int64_t e, int s, Real* u) const {  // NV x N_fq = C Vf[face rows]^T
    const Part& p = of(e);
    for (int j = 0; j < p.nfq; j++)
      for (int v = 0; v < kNv; v++) {
        Real a = 0.0;
        for (int i = 0; i < p.nb; i++) a += coeff[coff[e] + v + kNv * i] * p.Vf[(s * p.nfq + j) + static_cast<size_t>(p.naq) * i];
        u[v + kNv * j] = a;
      }
  }
  void volumeQuadrature() {
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      const Part& p = of(e);
      std::vector<Real> uq(static_cast<size_t>(kNv) * p.nq);
      gemmBt(uq.data(), &coeff[coff[e]], p.V.data(), kNv, p.nb, p.nq, false);
      Real* out = &vol_quad[qoff[e] * kNv * kDim];
      for (int q = 0; q < p.nq; q++) {
        Real w[8], fc[15];
        pt.computational(&uq[static_cast<size_t>(kNv) * q], w);
        const Real rho = w[0], pr = w[kDim + 2];
        const Real et = w[kDim + 1] + pt.velocitySquared(w) / 2.0;
        for (int a = 0; a < kDim; a++) {
          fc[a] = rho * w[1 + a];
          for (int b = 0; b < kDim; b++) fc[a + kDim * (1 + b)] = rho * w[1 + a] * w[1 + b] + (a == b ? pr * 1.0 : pr * 0.0);
          fc[a + kDim * (kDim + 1)] = (rho * et + pr) * w[1 + a];
        }
        const Real* G = &tiw[(qoff[e] + q) * kDim * kDim];
        for (int v = 0; v < kNv; v++)
          for (int b = 0; b < kDim; b++) {
            Real acc = 0.0;
            for (int a = 0; a < kDim; a++) acc += fc[a + kDim * v] * G[a + kDim * b];
            out[v + static_cast<size_t>(kNv) * (q * kDim + b)] = acc;
          }
      }
    }
  }
  void adjacencyQuadrature() {
#pragma omp parallel for schedule(static) if(nface >= kOmpMinItems)
    for (int64_t f = 0; f < nface; f++) {
      const int32_t L = fparent[2 * f], R = fparent[2 * f + 1];
      const int sl = flocal[2 * f], sr = flocal[2 * f + 1];
      const int nfq = of(L).nfq;
      std::vector<Real> ul(static_cast<size_t>(kNv) * nfq), ur(static_cast<size_t>(kNv) * nfq);
      faceTrace(L, sl, ul.data());
      faceTrace(R, sr, ur.data());
      for (int j = 0; j < nfq; j++) {
        const int k = fseq[f * nfq + j];
        Real wl[8], wr[8], h[5];
        pt.computational(&ul[kNv * j], wl);
        pt.computational(&ur[kNv * k], wr);
        pt.convectiveFlux(&fnormal[(static_cast<size_t>(f) * nfq + j) * kDim], &ul[kNv * j], wl, &ur[kNv * k], wr, h);
        Real* cl = &adj_quad[(aoff[L] + sl * nfq + j) * kNv];
        Real* cr = &adj_quad[(aoff[R] + sr * nfq + k) * kNv];
        for (int v = 0; v < kNv; v++) {
          const Real t = h[v] * fjw[f * nfq + j];
          cl[v] = t;
          cr[v] = -t;
        }
      }
    }
  }
  void calcResidual() {
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      const Part& p = of(e);
      Real* res = &residual[coff[e]];
      gemm(res, &vol_quad[qoff[e] * kNv * kDim], p.dV.data(), kNv, kDim * p.nq, p.nb, false);
      gemm(res, &adj_quad[aoff[e] * kNv], p.Vf.data(), kNv, p.naq, p.nb, true, -1.0);
    }
  }
  void updateCoefficient(int stage, Real dt) {
    Real c[3];
    pt.rkCoefficients(stage, c);
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      const int nb = of(e).nb;
      const size_t sz = static_cast<size_t>(kNv) * nb;
      std::vector<Real> rm(sz);
      gemm(rm.data(), &residual[coff[e]], &minv[moff[e]], kNv, nb, nb, false);
      Real* u = &coeff[coff[e]];
      const Real* un = &coeff_last[coff[e]];
      for (size_t i = 0; i < sz; i++) u[i] *= c[1];
      for (size_t i = 0; i < sz; i++) u[i] += c[0] * un[i];
      for (size_t i = 0; i < sz; i++) u[i] += c[2] * dt * rm[i];
    }
  }
  void relativeError() {  // sum over both element types, divided by the element count
    std::vector<Real> per(static_cast<size_t>(ne) * kNv);
#pragma omp parallel for schedule(static) if(ne >= kOmpMinItems)
    for (int64_t e = 0; e < ne; e++) {
      const Part& p = of(e);
      std::vector<Real> rq(static_cast<size_t>(kNv) * p.nq);
      gemmBt(rq.data(), &residual[coff[e]], p.V.data(), kNv, p.nb, p.nq, false);
      for (int v = 0; v < kNv; v++) {
        Real acc = 0.0;
        for (int q = 0; q < p.nq; q++) acc += fabs(rq[v + static_cast<size_t>(kNv) * q]);
        per[e * kNv + v] = acc / p.nq;
      }
    }
    std::vector<Real> acc(kNv, 0.0);
    for (int64_t e = 0; e < ne; e++)
      for (int v = 0; v < kNv; v++) acc[v] += per[e * kNv + v];
    for (int v = 0; v < kNv; v++) rel_err[v] = acc[v] / static_cast<Real>(ne);
  }
  void spatial() {
    volumeQuadrature();
    adjacencyQuadrature();
    calcResidual();
  }
  void step(Real dt) {
    coeff_last = coeff;
    for (int s = 0; s < pt.stages(); s++) {
      spatial();
      updateCoefficient(s, dt);
    }
    relativeError();
  }
};

template <typename T>
void colMajorFromRowMajor(std::vector<T>& dst, const T* src, int rows, int cols) {
  dst.resize(static_cast<size_t>(rows) * cols);
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) dst[r + static_cast<size_t>(rows) * c] = src[static_cast<size_t>(r) * cols + c];
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int oracle_get_threads(void) { return omp_get_max_threads(); }

// Sizes: dim, nb, nq, nf, nfq, order, ne, nface. Row-major tables as sdg_mesh_export writes them.
void* oracle_create(const sdg_problem_desc* p, int dim, int nb, int nq, int nf, int nfq, int order, int64_t ne,
                    int64_t nface, const Real* V, const Real* dV, const Real* Vf, const Real* lsq,
                    const Real* detjw, const Real* tiw, const Real* minv, const Real* min_edge,
                    const int32_t* fparent, const int32_t* flocal, const int32_t* fseq, const Real* fnormal,
                    const Real* fjw) {
  if (!p || dim < 1 || dim > 3) return nullptr;
  Oracle* o = new (std::nothrow) Oracle();
  if (!o) return nullptr;
  o->dim = dim;
  o->nv = dim + 2;
  o->nb = nb;
  o->nq = nq;
  o->nf = nf;
  o->nfq = nfq;
  o->naq = nf * nfq;
  o->order = order;
  o->ne = ne;
  o->nface = nface;
  o->equation = p->equation;
  o->flux = p->convective_flux;
  o->viscous = p->equation == SDG_EQ_COMPRESSIBLE_NS ? p->viscous_flux : SDG_VISC_NONE;
  o->rk = p->time_integration;
  o->phys.eos = p->equation_of_state;
  o->phys.transport = p->transport;
  o->phys.cp = p->cp;
  o->phys.cv = p->cv;
  o->phys.mu = p->dynamic_viscosity;
  o->phys.kappa = p->cp * p->dynamic_viscosity / kPrandtl;  // PhysicalModel.cpp:149-153
  o->phys.c0 = p->reference_sound_speed;
  o->phys.rho0 = p->reference_density;
  o->phys.padd = 0.01 * p->reference_density * p->reference_sound_speed * p->reference_sound_speed;
  o->phys.source = p->source_term;
  o->phys.beta = p->thermal_expansion_coefficient;
  o->phys.tref = p->reference_temperature;
  colMajorFromRowMajor(o->V, V, nq, nb);
  colMajorFromRowMajor(o->dV, dV, nq * dim, nb);
  colMajorFromRowMajor(o->Vf, Vf, nf * nfq, nb);
  colMajorFromRowMajor(o->lsq, lsq, nb, nb);
  o->detjw.assign(detjw, detjw + ne * nq);
  o->tiw.assign(tiw, tiw + ne * nq * dim * dim);
  o->minv.resize(static_cast<size_t>(ne) * nb * nb);
  for (int64_t e = 0; e < ne; e++)
    for (int r = 0; r < nb; r++)
      for (int c = 0; c < nb; c++)
        o->minv[e * nb * nb + r + static_cast<size_t>(nb) * c] = minv[e * nb * nb + static_cast<size_t>(r) * nb + c];
  o->min_edge.assign(min_edge, min_edge + ne);
  o->fparent.assign(fparent, fparent + 2 * nface);
  o->flocal.assign(flocal, flocal + 2 * nface);
  o->fseq.assign(fseq, fseq + nface * nfq);
  o->fnormal.assign(fnormal, fnormal + nface * nfq * dim);
  o->fjw.assign(fjw, fjw + nface * nfq);
  const size_t nv = o->nv;
  o->coeff.assign(ne * nv * nb, 0.0);
  o->coeff_last.assign(ne * nv * nb, 0.0);
  o->vol_quad.assign(ne * nv * dim * nq, 0.0);
  o->src_quad.assign(ne * nq, 0.0);
  o->adj_quad.assign(ne * nv * o->naq, 0.0);
  o->residual.assign(ne * nv * nb, 0.0);
  o->vg_coeff.assign(ne * o->szG(), 0.0);
  o->vg_quad.assign(ne * nv * dim * dim * nq, 0.0);
  o->vg_adj.assign(ne * nv * dim * o->naq, 0.0);
  o->vg_res.assign(ne * o->szG(), 0.0);
  if (o->ns()) {
    o->nig = o->viscous == SDG_VISC_BR2 ? nf : 1;
    o->ig_adj.assign(ne * nv * dim * o->naq, 0.0);
    o->ig_res.assign(ne * o->nig * o->szG(), 0.0);
    o->ig_coeff.assign(ne * o->nig * o->szG(), 0.0);
    o->g_coeff.assign(ne * o->szG(), 0.0);
  }
  o->rel_err.assign(nv, 0.0);
  return o;
}

void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }

// Boundary faces: parent_local (F_b x 2: element, local face), type (F_b), normal
// (F_b x N_fq x D), jw (F_b x N_fq).
int oracle_set_boundary(void* h, int64_t nbface, const int32_t* parent_local, const int32_t* type, const Real* normal,
                        const Real* jw) {
  Oracle* o = static_cast<Oracle*>(h);
  o->nbface = nbface;
  o->bparent.resize(nbface);
  o->blocal.resize(nbface);
  o->btype.assign(type, type + nbface);
  for (int64_t b = 0; b < nbface; b++) {
    o->bparent[b] = parent_local[2 * b];
    o->blocal[b] = parent_local[2 * b + 1];
  }
  o->bnormal.assign(normal, normal + nbface * o->nfq * o->dim);
  o->bjw.assign(jw, jw + nbface * o->nfq);
  o->bdummy.assign(static_cast<size_t>(nbface) * o->nfq * (o->dim + 3), 0.0);
  return 0;
}
// Boundary data: primitive (rho, u.., T) per boundary point -> computational
// (VariableConvertor.cpp:118-139)
int oracle_set_boundary_primitive(void* h, const Real* prim) {
  Oracle* o = static_cast<Oracle*>(h);
  const int dim = o->dim, nv = o->nv;
  for (size_t i = 0; i < static_cast<size_t>(o->nbface) * o->nfq; i++) {
    const Real* pr = prim + i * nv;
    Real* w = &o->bdummy[i * (dim + 3)];
    w[0] = pr[0];
    for (int d = 0; d < dim; d++) w[1 + d] = pr[1 + d];
    w[dim + 1] = o->phys.internalEnergyFromTemperature(pr[dim + 1]);
    w[dim + 2] = o->phys.pressure(w[0], w[dim + 1]);
  }
  return 0;
}

int oracle_set_state(void* h, const Real* c) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(o->coeff.data(), c, o->coeff.size() * sizeof(Real));
  o->coeff_last = o->coeff;
  o->has_state = true;
  return 0;
}
int oracle_get_state(void* h, Real* c) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(c, o->coeff.data(), o->coeff.size() * sizeof(Real));
  return 0;
}
// InitialCondition.cpp:85-104: conserved at quadrature points from primitives, then
// C = U_q V (V^T V)^-1
int oracle_set_initial_primitive(void* h, const Real* prim) {
  Oracle* o = static_cast<Oracle*>(h);
  const int nv = o->nv, nq = o->nq, nb = o->nb, dim = o->dim;
#pragma omp parallel for schedule(static) if(o->ne >= kOmpMinItems)
  for (int64_t e = 0; e < o->ne; e++) {
    std::vector<Real> uq(static_cast<size_t>(nv) * nq), tmp(static_cast<size_t>(nv) * nb);
    for (int q = 0; q < nq; q++) {
      const Real* pr = &prim[(static_cast<size_t>(e) * nq + q) * nv];
      const Real rho = pr[0];
      Real v2 = 0.0;
      uq[static_cast<size_t>(nv) * q] = rho;
      for (int d = 0; d < dim; d++) {
        uq[1 + d + static_cast<size_t>(nv) * q] = rho * pr[1 + d];
        v2 += pr[1 + d] * pr[1 + d];
      }
      const Real et = o->phys.internalEnergyFromTemperature(pr[dim + 1]) + v2 / 2.0;
      uq[dim + 1 + static_cast<size_t>(nv) * q] = rho * et;
    }
    gemm(tmp.data(), uq.data(), o->V.data(), nv, nq, nb, false);
    gemm(&o->coeff[e * o->szC()], tmp.data(), o->lsq.data(), nv, nb, nb, false);
  }
  o->coeff_last = o->coeff;
  o->has_state = true;
  return 0;
}
// Artificial viscosity: node tags (E x nvert, 0-based), inner radii (E), P1 nodal functions at
// the volume points (N_q x nvert) and face points (N_aq x nvert), row-major; tolerance k, factor f.
int oracle_set_artificial_viscosity(void* h, int nvert, int64_t nnode, const int32_t* tag, const Real* radius,
                                    const Real* nodal_q, const Real* nodal_f, Real tol, Real factor) {
  Oracle* o = static_cast<Oracle*>(h);
  o->av = true;
  o->nvert = nvert;
  o->nnode = nnode;
  o->node_tag.assign(tag, tag + o->ne * nvert);
  o->inner_radius.assign(radius, radius + o->ne);
  o->nodal_q.assign(nodal_q, nodal_q + static_cast<size_t>(o->nq) * nvert);
  o->nodal_f.assign(nodal_f, nodal_f + static_cast<size_t>(o->naq) * nvert);
  o->node_av.assign(nnode, 0.0);
  o->elem_av.assign(static_cast<size_t>(o->ne) * nvert, 0.0);
  o->av_tol = tol;
  o->av_factor = factor;
  return 0;
}
// Volume-gradient coefficients of the last residual evaluation, (NV D) x N_b per element
// column-major (variable_volume_gradient_basis_function_coefficient_, TimeIntegration.cpp:205-206).
int oracle_get_volume_gradient(void* h, Real* g) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(g, o->vg_coeff.data(), o->vg_coeff.size() * sizeof(Real));
  return 0;
}
int oracle_get_artificial_viscosity(void* h, Real* node_values) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(node_values, o->node_av.data(), o->node_av.size() * sizeof(Real));
  return 0;
}

// With artificial viscosity the node values are formed from the current state first, as at the
// start of a step.
int oracle_residual(void* h, Real* r) {
  Oracle* o = static_cast<Oracle*>(h);
  if (o->av) o->artificialViscosity();
  o->spatial();
  if (r) std::memcpy(r, o->residual.data(), o->residual.size() * sizeof(Real));
  return 0;
}
int oracle_get_gradient(void* h, Real* g) {
  Oracle* o = static_cast<Oracle*>(h);
  if (!o->ns()) return -2;
  std::memcpy(g, o->g_coeff.data(), o->g_coeff.size() * sizeof(Real));
  return 0;
}
// Gradient traces each element uses at its own face points in the viscous face flux of the last
// residual evaluation (VariableConvertor.cpp:640-689: BR2 the volume gradient plus that face's
// lift, BR1 the lifted gradient), E x N_f x N_fq x (NV D).
// Per-face gradient coefficients of the last residual evaluation, E x N_f x (NV D) x N_b (each block
// column-major, row var*D+dim): BR2 the volume gradient plus that face's lift, BR1 the total
// gradient -- the blocks RawBinary.cpp:93-110 writes for boundary faces.
int oracle_get_face_gradient_coefficients(void* h, Real* out) {
  Oracle* o = static_cast<Oracle*>(h);
  if (!o->ns()) return -2;
  const size_t blk = o->szG();
  for (int64_t e = 0; e < o->ne; e++)
    for (int s = 0; s < o->nf; s++) {
      Real* dst = out + (static_cast<size_t>(e) * o->nf + s) * blk;
      for (size_t i = 0; i < blk; i++)
        dst[i] = o->viscous == SDG_VISC_BR2
                     ? o->vg_coeff[e * blk + i] + o->ig_coeff[(static_cast<size_t>(e) * o->nig + s) * blk + i]
                     : o->g_coeff[e * blk + i];
    }
  return 0;
}
int oracle_get_face_gradient(void* h, Real* out) {
  Oracle* o = static_cast<Oracle*>(h);
  if (!o->ns()) return -2;
  const size_t rows = static_cast<size_t>(o->nv) * o->dim, blk = rows * o->nfq;
#pragma omp parallel for schedule(static) if(o->ne >= kOmpMinItems)
  for (int64_t e = 0; e < o->ne; e++)
    for (int s = 0; s < o->nf; s++) o->faceGradientTrace(e, s, out + (static_cast<size_t>(e) * o->nf + s) * blk);
  return 0;
}
// The residual monitor (TimeIntegration.cpp:279-324) of a given residual, NV x N_b per element
// column-major; the oracle's own residual is replaced by it.
int oracle_monitor(void* h, const Real* r, Real* rel_err) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(o->residual.data(), r, o->residual.size() * sizeof(Real));
  o->relativeError();
  std::memcpy(rel_err, o->rel_err.data(), o->nv * sizeof(Real));
  return 0;
}
int oracle_compute_dt(void* h, Real cfl, Real* dt) {  // TimeIntegration.cpp:104-131
  Oracle* o = static_cast<Oracle*>(h);
  const int nv = o->nv, nq = o->nq, nb = o->nb, dim = o->dim;
  Real best = std::numeric_limits<Real>::max();
#pragma omp parallel for schedule(static) reduction(min : best) if(o->ne >= kOmpMinItems)
  for (int64_t e = 0; e < o->ne; e++) {
    std::vector<Real> uq(static_cast<size_t>(nv) * nq);
    gemmBt(uq.data(), &o->coeff[e * o->szC()], o->V.data(), nv, nb, nq, false);
    for (int q = 0; q < nq; q++) {
      Real w[8];
      o->computational(&uq[static_cast<size_t>(nv) * q], w);
      const Real c = o->phys.soundSpeed(w[0], w[dim + 2]);
      const Real sr = sqrt(o->velocitySquared(w)) + c;
      const Real t = cfl * o->min_edge[e] / (sr * (o->order + 1.0) * (o->order + 1.0));
      best = std::min(best, t);
    }
  }
  *dt = best;
  return 0;
}
int oracle_step(void* h, Real dt, int nsteps, Real* rel_err, int* nan_flag, int* steps_taken) {
  Oracle* o = static_cast<Oracle*>(h);
  int taken = 0, nanf = 0;
  for (int i = 0; i < nsteps; i++) {
    o->step(dt);
    taken++;
    bool all_nan = true;
    for (int v = 0; v < o->nv; v++) all_nan = all_nan && isnan(o->rel_err[v]);
    if (all_nan) {  // SystemControl.cpp:185-191
      nanf = 1;
      break;
    }
  }
  if (rel_err) std::memcpy(rel_err, o->rel_err.data(), o->nv * sizeof(Real));
  if (nan_flag) *nan_flag = nanf;
  if (steps_taken) *steps_taken = taken;
  return 0;
}
// Mixed triangle / quadrangle Euler meshes. Per type t (0 triangle, 1 quadrangle): sizes
// nb[t], nq[t], nf[t], nfq (line faces, shared) and row-major tables V[t], dV[t], Vf[t], lsq[t]
// as sdg_mesh_export writes them for a mesh of that type; per element its type, and detjw, tiw,
// minv blocks of its own size concatenated in element order (sdg_mesh_export of a split mesh).
void* oracle_mixed_create(const sdg_problem_desc* p, int order, int64_t ne, int64_t nface, const int8_t* etype,
                          const int* nb, const int* nq, const int* nf, int nfq, const Real* const* V,
                          const Real* const* dV, const Real* const* Vf, const Real* const* lsq, const Real* detjw,
                          const Real* tiw, const Real* minv, const Real* min_edge, const int32_t* fparent,
                          const int32_t* flocal, const int32_t* fseq, const Real* fnormal, const Real* fjw) {
  if (p->equation != SDG_EQ_COMPRESSIBLE_EULER) return nullptr;
  MixedOracle* o = new (std::nothrow) MixedOracle();
  if (!o) return nullptr;
  o->pt.dim = 2;
  o->pt.nv = 4;
  o->pt.order = order;
  o->pt.flux = p->convective_flux;
  o->pt.rk = p->time_integration;
  o->pt.phys.eos = p->equation_of_state;
  o->pt.phys.cp = p->cp;
  o->pt.phys.cv = p->cv;
  o->pt.phys.c0 = p->reference_sound_speed;
  o->pt.phys.rho0 = p->reference_density;
  o->pt.phys.padd = 0.01 * p->reference_density * p->reference_sound_speed * p->reference_sound_speed;
  for (int t = 0; t < 2; t++) {
    MixedOracle::Part& q = o->part[t];
    q.nb = nb[t];
    q.nq = nq[t];
    q.nf = nf[t];
    q.nfq = nfq;
    q.naq = nf[t] * nfq;
    colMajorFromRowMajor(q.V, V[t], nq[t], nb[t]);
    colMajorFromRowMajor(q.dV, dV[t], nq[t] * 2, nb[t]);
    colMajorFromRowMajor(q.Vf, Vf[t], nf[t] * nfq, nb[t]);
    colMajorFromRowMajor(q.lsq, lsq[t], nb[t], nb[t]);
  }
  o->ne = ne;
  o->nface = nface;
  o->etype.assign(etype, etype + ne);
  o->coff.resize(ne);
  o->qoff.resize(ne);
  o->aoff.resize(ne);
  o->moff.resize(ne);
  int64_t c = 0, q = 0, a = 0, m = 0;
  for (int64_t e = 0; e < ne; e++) {
    const MixedOracle::Part& pe = o->of(e);
    o->coff[e] = c;
    o->qoff[e] = q;
    o->aoff[e] = a;
    o->moff[e] = m;
    c += 4 * pe.nb;
    q += pe.nq;
    a += pe.naq;
    m += static_cast<int64_t>(pe.nb) * pe.nb;
  }
  o->detjw.assign(detjw, detjw + q);
  o->tiw.assign(tiw, tiw + q * 4);
  o->minv.resize(m);
  for (int64_t e = 0; e < ne; e++) {  // row-major blocks -> column-major
    const int nbe = o->of(e).nb;
    for (int r = 0; r < nbe; r++)
      for (int k = 0; k < nbe; k++)
        o->minv[o->moff[e] + r + static_cast<size_t>(nbe) * k] = minv[o->moff[e] + static_cast<size_t>(r) * nbe + k];
  }
  o->min_edge.assign(min_edge, min_edge + ne);
  o->fparent.assign(fparent, fparent + 2 * nface);
  o->flocal.assign(flocal, flocal + 2 * nface);
  o->fseq.assign(fseq, fseq + nface * nfq);
  o->fnormal.assign(fnormal, fnormal + nface * nfq * 2);
  o->fjw.assign(fjw, fjw + nface * nfq);
  o->coeff.assign(c, 0.0);
  o->coeff_last.assign(c, 0.0);
  o->residual.assign(c, 0.0);
  o->vol_quad.assign(q * 4 * 2, 0.0);
  o->adj_quad.assign(a * 4, 0.0);
  o->rel_err.assign(4, 0.0);
  return o;
}
void oracle_mixed_destroy(void* h) { delete static_cast<MixedOracle*>(h); }
int oracle_mixed_set_state(void* h, const Real* c) {
  MixedOracle* o = static_cast<MixedOracle*>(h);
  std::memcpy(o->coeff.data(), c, o->coeff.size() * sizeof(Real));
  o->coeff_last = o->coeff;
  return 0;
}
int oracle_mixed_get_state(void* h, Real* c) {
  MixedOracle* o = static_cast<MixedOracle*>(h);
  std::memcpy(c, o->coeff.data(), o->coeff.size() * sizeof(Real));
  return 0;
}
// InitialCondition.cpp:85-104 per element type; prim: each element's points in element order
int oracle_mixed_set_initial_primitive(void* h, const Real* prim) {
  MixedOracle* o = static_cast<MixedOracle*>(h);
#pragma omp parallel for schedule(static) if(o->ne >= kOmpMinItems)
  for (int64_t e = 0; e < o->ne; e++) {
    const MixedOracle::Part& p = o->of(e);
    std::vector<Real> uq(static_cast<size_t>(4) * p.nq), tmp(static_cast<size_t>(4) * p.nb);
    for (int q = 0; q < p.nq; q++) {
      const Real* pr = &prim[(o->qoff[e] + q) * 4];
      const Real rho = pr[0];
      uq[4 * q] = rho;
      uq[4 * q + 1] = rho * pr[1];
      uq[4 * q + 2] = rho * pr[2];
      const Real et = o->pt.phys.internalEnergyFromTemperature(pr[3]) + (pr[1] * pr[1] + pr[2] * pr[2]) / 2.0;
      uq[4 * q + 3] = rho * et;
    }
    gemm(tmp.data(), uq.data(), p.V.data(), 4, p.nq, p.nb, false);
    gemm(&o->coeff[o->coff[e]], tmp.data(), p.lsq.data(), 4, p.nb, p.nb, false);
  }
  o->coeff_last = o->coeff;
  return 0;
}
int oracle_mixed_residual(void* h, Real* r) {
  MixedOracle* o = static_cast<MixedOracle*>(h);
  o->spatial();
  if (r) std::memcpy(r, o->residual.data(), o->residual.size() * sizeof(Real));
  return 0;
}
int oracle_mixed_compute_dt(void* h, Real cfl, Real* dt) {  // TimeIntegration.cpp:104-131, both types
  MixedOracle* o = static_cast<MixedOracle*>(h);
  Real best = std::numeric_limits<Real>::max();
  const int order = o->pt.order;
#pragma omp parallel for schedule(static) reduction(min : best) if(o->ne >= kOmpMinItems)
  for (int64_t e = 0; e < o->ne; e++) {
    const MixedOracle::Part& p = o->of(e);
    std::vector<Real> uq(static_cast<size_t>(4) * p.nq);
    gemmBt(uq.data(), &o->coeff[o->coff[e]], p.V.data(), 4, p.nb, p.nq, false);
    for (int q = 0; q < p.nq; q++) {
      Real w[8];
      o->pt.computational(&uq[4 * q], w);
      const Real sr = sqrt(o->pt.velocitySquared(w)) + o->pt.phys.soundSpeed(w[0], w[4]);
      best = std::min(best, cfl * o->min_edge[e] / (sr * (order + 1.0) * (order + 1.0)));
    }
  }
  *dt = best;
  return 0;
}
int oracle_mixed_step(void* h, Real dt, int nsteps, Real* rel_err) {
  MixedOracle* o = static_cast<MixedOracle*>(h);
  for (int i = 0; i < nsteps; i++) o->step(dt);
  if (rel_err) std::memcpy(rel_err, o->rel_err.data(), 4 * sizeof(Real));
  return 0;
}

// Pointwise physics for unit tests: conserved states (NV each), normal (D) -> flux (NV).
int oracle_point_flux(const sdg_problem_desc* p, int dim, int flux, const Real* n, const Real* ul, const Real* ur,
                      Real* h) {
  Oracle o;
  o.dim = dim;
  o.nv = dim + 2;
  o.flux = flux;
  o.phys.eos = p->equation_of_state;
  o.phys.cp = p->cp;
  o.phys.cv = p->cv;
  o.phys.c0 = p->reference_sound_speed;
  o.phys.rho0 = p->reference_density;
  o.phys.padd = 0.01 * p->reference_density * p->reference_sound_speed * p->reference_sound_speed;
  Real wl[8], wr[8];
  o.computational(ul, wl);
  o.computational(ur, wr);
  o.convectiveFlux(n, ul, wl, ur, wr, h);
  return 0;
}
// Boundary state of calculateBoundaryVariable at one point: conserved interior state ul,
// primitive boundary data prim (rho, u.., T); wb: computational (rho, u.., e, p).
int oracle_point_boundary_state(const sdg_problem_desc* p, int dim, int type, const Real* n, const Real* ul,
                                const Real* prim, Real* wb) {
  Oracle o;
  o.dim = dim;
  o.nv = dim + 2;
  o.phys.eos = p->equation_of_state;
  o.phys.cp = p->cp;
  o.phys.cv = p->cv;
  o.phys.c0 = p->reference_sound_speed;
  o.phys.rho0 = p->reference_density;
  o.phys.padd = 0.01 * p->reference_density * p->reference_sound_speed * p->reference_sound_speed;
  Real wl[8], wd[8];
  o.computational(ul, wl);
  wd[0] = prim[0];
  for (int d = 0; d < dim; d++) wd[1 + d] = prim[1 + d];
  wd[dim + 1] = o.phys.internalEnergyFromTemperature(prim[dim + 1]);
  wd[dim + 2] = o.phys.pressure(wd[0], wd[dim + 1]);
  o.boundaryVariable(type, n, wl, wd, wb);
  return 0;
}
// Viscous normal flux at one point from conserved state and conserved gradient (row var*D+dim).
int oracle_point_viscous_flux(const sdg_problem_desc* p, int dim, const Real* n, const Real* u, const Real* g,
                              Real* f) {
  Oracle o;
  o.dim = dim;
  o.nv = dim + 2;
  o.phys.eos = p->equation_of_state;
  o.phys.transport = p->transport;
  o.phys.cp = p->cp;
  o.phys.cv = p->cv;
  o.phys.mu = p->dynamic_viscosity;
  o.phys.kappa = p->cp * p->dynamic_viscosity / kPrandtl;
  Real w[8], pg[15];
  o.computational(u, w);
  o.primitiveGradient(w, u, g, pg);
  o.viscousNormalFlux(n, w, pg, f);
  return 0;
}

}  // extern "C"
