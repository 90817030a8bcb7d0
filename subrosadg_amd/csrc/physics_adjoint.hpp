// physics_adjoint.hpp -- reverse-mode (vector-Jacobian) products of the pointwise physics of
// physics.hpp, written out by hand: one pass gives the adjoint of every input, where forward-mode
// dual numbers need one pass per input component (ten per face point for a numerical flux).
//
// Each function takes the adjoint of a forward function's outputs and ACCUMULATES the adjoints of
// its inputs. Branches (wave-speed cases, max, |.|) follow the forward-mode convention of dual.hpp
// (fmax(a, b) takes a when a >= b, |x| takes -x when x < 0), so both modes differentiate the same
// piece of a piecewise formula.
#ifndef SDG_PHYSICS_ADJOINT_HPP_
#define SDG_PHYSICS_ADJOINT_HPP_

#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace sdg {

// computational w = (rho, v.., e, p) of conserved u (VariableConvertor.cpp:315-339):
// ub = (dw/du)^T wb (written, not accumulated)
template <int D>
__device__ __forceinline__ void computationalVjp(const Phys& ph, const double* u, const double* w, const double* wb,
                                                 double* ub) {
  const double rho = w[0], ir = 1.0 / rho;
  double rb = wb[0], eb = wb[D + 1];
  const double pb = wb[D + 2];
  if (ph.eos == 0) {  // p = (gamma - 1) rho e
    rb += (kGamma - 1.0) * w[D + 1] * pb;
    eb += (kGamma - 1.0) * rho * pb;
  } else {  // p = c0^2 (rho - rho0) + p_add
    rb += ph.c0 * ph.c0 * pb;
  }
  // v_d = u_{1+d} / rho, e = u_E / rho - |v|^2 / 2
  double irb = eb * u[D + 1];
  ub[D + 1] = eb * ir;
#pragma unroll
  for (int d = 0; d < D; d++) {
    const double vb = wb[1 + d] - eb * w[1 + d];
    ub[1 + d] = vb * ir;
    irb += vb * u[1 + d];
  }
  ub[0] = rb - irb * ir * ir;
}

// sound speed c(rho, p): rho-bar and p-bar from c-bar
__device__ __forceinline__ void soundSpeedVjp(const Phys& ph, double rho, double p, double c, double cb, double& rb,
                                              double& pb) {
  if (ph.eos != 0) return;  // constant c0
  // c = sqrt(gamma p / rho)
  pb += cb * kGamma / (2.0 * c * rho);
  rb -= cb * c / (2.0 * rho);
  (void)p;
}

// normal flux f = F(w).n (ConvectiveFlux.cpp:59-91): wb += (df/dw)^T fb
template <int D>
__device__ __forceinline__ void normalFluxVjp(const double* n, const double* w, const double* fb, double* wb) {
  const double rho = w[0], p = w[D + 2];
  double un = 0.0, v2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    un += w[1 + d] * n[d];
    v2 += w[1 + d] * w[1 + d];
  }
  const double et = w[D + 1] + v2 / 2.0;
  const double fe = fb[D + 1];
  double unb = fb[0] * rho + fe * (rho * et + p);
  double rb = fb[0] * un + fe * et * un;
  double pb = fe * un;
  const double etb = fe * rho * un;
#pragma unroll
  for (int d = 0; d < D; d++) {
    unb += fb[1 + d] * rho * w[1 + d];
    rb += fb[1 + d] * un * w[1 + d];
    pb += fb[1 + d] * n[d];
  }
#pragma unroll
  for (int d = 0; d < D; d++) wb[1 + d] += fb[1 + d] * rho * un + etb * w[1 + d] + unb * n[d];
  wb[0] += rb;
  wb[D + 1] += etb;
  wb[D + 2] += pb;
}

// the adjoints of one HLLC star side S (the branch taken): h = F(w_S).n + s_S (u*_S - u_S)
template <int D>
__device__ __forceinline__ void hllcStarVjp(const double* n, const double* u, const double* w, double s, double un,
                                            double ps, double sm, const double* hb, double* wb, double* ub,
                                            double& rb, double& unb, double& pb, double& sb, double& psb,
                                            double& smb) {
  constexpr int NV = D + 2;
  const double r = w[0], p = w[D + 2];
  const double a = s - un, is = 1.0 / (s - sm);
  double v2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) v2 += w[1 + d] * w[1 + d];
  const double et = w[D + 1] + v2 / 2.0;
  // numerators X of the star state u* = X / (s - sm)
  double X[NV];
  X[0] = r * a;
#pragma unroll
  for (int d = 0; d < D; d++) X[1 + d] = a * r * w[1 + d] + (ps - p) * n[d];
  X[D + 1] = a * r * et - p * un + ps * sm;
  normalFluxVjp<D>(n, w, hb, wb);
  double isb = 0.0, ab = 0.0;
  double Xb[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double usv = X[v] * is;
    sb += hb[v] * (usv - u[v]);
    ub[v] -= s * hb[v];
    const double usb = s * hb[v];
    Xb[v] = usb * is;
    isb += usb * X[v];
  }
  rb += Xb[0] * a;
  ab += Xb[0] * r;
#pragma unroll
  for (int d = 0; d < D; d++) {
    ab += Xb[1 + d] * r * w[1 + d];
    rb += Xb[1 + d] * a * w[1 + d];
    wb[1 + d] += Xb[1 + d] * a * r;
    psb += Xb[1 + d] * n[d];
    pb -= Xb[1 + d] * n[d];
  }
  const double xe = Xb[D + 1];
  ab += xe * r * et;
  rb += xe * a * et;
  const double etb = xe * a * r;
  wb[D + 1] += etb;
#pragma unroll
  for (int d = 0; d < D; d++) wb[1 + d] += etb * w[1 + d];
  pb -= xe * un;
  unb -= xe * p;
  psb += xe * sm;
  smb += xe * ps;
  const double db = -isb * is * is;  // is = 1 / (s - sm)
  sb += db;
  smb -= db;
  sb += ab;  // a = s - un
  unb -= ab;
}

// Roe flux with Harten's entropy fix (ConvectiveFlux.cpp:240-350, restated in physics.hpp):
// wbl/wbr += adjoints of the computational variables for h = ((F_L + F_R) - sum_k s_k r_k) / 2
template <int D>
__device__ void roeFluxVjp(const Phys& ph, const double* n, const double* wl, const double* wr, const double* hb,
                           double* wbl, double* wbr) {
  constexpr int NV = D + 2;
  double fb[NV], sb[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    fb[v] = hb[v] / 2.0;
    sb[v] = -hb[v] / 2.0;  // adjoint of the dissipation sum
  }
  normalFluxVjp<D>(n, wl, fb, wbl);
  normalFluxVjp<D>(n, wr, fb, wbr);
  // forward recomputation of the Roe state and the wave strengths
  const double g = kGamma;
  const double sql = sqrt(wl[0]), sqr = sqrt(wr[0]);
  const double sqs = sql + sqr, isqs = 1.0 / sqs;
  const double rho = sqrt(wl[0] * wr[0]);
  double vel[D], v2l = 0.0, v2r = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    vel[d] = (sql * wl[1 + d] + sqr * wr[1 + d]) * isqs;
    v2l += wl[1 + d] * wl[1 + d];
    v2r += wr[1 + d] * wr[1 + d];
  }
  const double hl = wl[D + 1] * g + v2l / 2.0, hr = wr[D + 1] * g + v2r / 2.0;
  const double htot = (sql * hl + sqr * hr) * isqs;
  double v2 = 0.0, un = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    v2 += vel[d] * vel[d];
    un += vel[d] * n[d];
  }
  const double e = (htot - v2 / 2.0) / g;
  const double p = pressureOf(ph, rho, e);
  const double c = soundSpeedOf(ph, rho, p);
  double dw[D + 3];
#pragma unroll
  for (int i = 0; i < D + 3; i++) dw[i] = wr[i] - wl[i];
  double dun = 0.0, vdv = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    dun += dw[1 + d] * n[d];
    vdv += vel[d] * dw[1 + d];
  }
  const double delta = c / 20.0;
  const double xm = un - c, xp = un + c;
  const bool am = fabs(xm) > delta, ap = fabs(xp) > delta;
  const double a1 = am ? fabs(xm) : (xm * xm + delta * delta) / (2.0 * delta);
  const double a3 = ap ? fabs(xp) : (xp * xp + delta * delta) / (2.0 * delta);
  const double dp = dw[D + 2], drho = dw[0];
  const double ic2 = 1.0 / (c * c);
  const double X0 = dp - rho * c * dun, X3 = dp + rho * c * dun;
  const double s0 = a1 * X0 * (0.5 * ic2);
  const double aun = fabs(un), sgn = un < 0.0 ? -1.0 : 1.0;
  const double s1 = aun * (drho - dp * ic2);
  const double s2 = aun * rho;
  const double s3 = a3 * X3 * (0.5 * ic2);
  // reverse
  double velb[D], rhob = 0.0, cb = 0.0, unb = 0.0, htotb = 0.0, v2b = 0.0, dunb = 0.0, vdvb = 0.0, dpb = 0.0,
         drhob = 0.0, ic2b = 0.0, a1b = 0.0, a3b = 0.0, dwb[D + 3];
#pragma unroll
  for (int d = 0; d < D; d++) velb[d] = 0.0;
#pragma unroll
  for (int i = 0; i < D + 3; i++) dwb[i] = 0.0;
  const double se = sb[D + 1];
  // column 3: (1, vel + c n, htot + c un) s3
  double s3b = sb[0] + se * (htot + c * un);
#pragma unroll
  for (int d = 0; d < D; d++) {
    s3b += sb[1 + d] * (vel[d] + c * n[d]);
    velb[d] += sb[1 + d] * s3;
    cb += sb[1 + d] * n[d] * s3;
  }
  htotb += se * s3;
  cb += se * un * s3;
  unb += se * c * s3;
  // column 2 (D >= 2): (0, dv - dun n, vdv - un dun) s2
  double s2b = 0.0;
  if (D >= 2) {
    s2b += se * (vdv - un * dun);
#pragma unroll
    for (int d = 0; d < D; d++) {
      s2b += sb[1 + d] * (dw[1 + d] - dun * n[d]);
      dwb[1 + d] += sb[1 + d] * s2;
      dunb -= sb[1 + d] * n[d] * s2;
    }
    vdvb += se * s2;
    unb -= se * dun * s2;
    dunb -= se * un * s2;
  }
  // column 1: (1, vel, |vel|^2 / 2) s1
  double s1b = sb[0] + se * v2 / 2.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    s1b += sb[1 + d] * vel[d];
    velb[d] += sb[1 + d] * s1;
  }
  v2b += se * s1 / 2.0;
  // column 0: (1, vel - c n, htot - c un) s0
  double s0b = sb[0] + se * (htot - c * un);
#pragma unroll
  for (int d = 0; d < D; d++) {
    s0b += sb[1 + d] * (vel[d] - c * n[d]);
    velb[d] += sb[1 + d] * s0;
    cb -= sb[1 + d] * n[d] * s0;
  }
  htotb += se * s0;
  cb -= se * un * s0;
  unb -= se * c * s0;
  // wave strengths
  a3b += s3b * X3 * 0.5 * ic2;
  const double X3b = s3b * a3 * 0.5 * ic2;
  ic2b += s3b * a3 * X3 * 0.5;
  dpb += X3b;
  rhob += X3b * c * dun;
  cb += X3b * rho * dun;
  dunb += X3b * rho * c;
  unb += s2b * rho * sgn;
  rhob += s2b * aun;
  unb += s1b * (drho - dp * ic2) * sgn;
  drhob += s1b * aun;
  dpb -= s1b * aun * ic2;
  ic2b -= s1b * aun * dp;
  a1b += s0b * X0 * 0.5 * ic2;
  const double X0b = s0b * a1 * 0.5 * ic2;
  ic2b += s0b * a1 * X0 * 0.5;
  dpb += X0b;
  rhob -= X0b * c * dun;
  cb -= X0b * rho * dun;
  dunb -= X0b * rho * c;
  cb += ic2b * (-2.0 * ic2 / c);  // ic2 = 1 / c^2
  // entropy-fixed eigenvalues |un -+ c|
  double deltab = 0.0;
  if (ap) {
    const double t = xp < 0.0 ? -a3b : a3b;
    unb += t;
    cb += t;
  } else {
    unb += a3b * xp / delta;
    cb += a3b * xp / delta;
    deltab += a3b * (delta * delta - xp * xp) / (2.0 * delta * delta);
  }
  if (am) {
    const double t = xm < 0.0 ? -a1b : a1b;
    unb += t;
    cb -= t;
  } else {
    unb += a1b * xm / delta;
    cb -= a1b * xm / delta;
    deltab += a1b * (delta * delta - xm * xm) / (2.0 * delta * delta);
  }
  cb += deltab / 20.0;
  // jumps
#pragma unroll
  for (int d = 0; d < D; d++) {
    dwb[1 + d] += dunb * n[d] + vdvb * vel[d];
    velb[d] += vdvb * dw[1 + d];
  }
  dwb[D + 2] += dpb;
  dwb[0] += drhob;
  // c(rho, p), p(rho, e), e = (htot - |vel|^2 / 2) / gamma
  double pb = 0.0, eb = 0.0;
  soundSpeedVjp(ph, rho, p, c, cb, rhob, pb);
  if (ph.eos == 0) {
    rhob += pb * (kGamma - 1.0) * e;
    eb += pb * (kGamma - 1.0) * rho;
  } else {
    rhob += pb * ph.c0 * ph.c0;
  }
  htotb += eb / g;
  v2b -= eb / (2.0 * g);
#pragma unroll
  for (int d = 0; d < D; d++) velb[d] += unb * n[d] + 2.0 * v2b * vel[d];
  // Roe averages
  const double isqs2 = isqs * isqs;
  double sqlb = htotb * hl * isqs, sqrb = htotb * hr * isqs;
  const double hlb = htotb * sql * isqs, hrb = htotb * sqr * isqs;
  double isqsb = htotb * (sql * hl + sqr * hr);
#pragma unroll
  for (int d = 0; d < D; d++) {
    sqlb += velb[d] * wl[1 + d] * isqs;
    sqrb += velb[d] * wr[1 + d] * isqs;
    wbl[1 + d] += velb[d] * sql * isqs + hlb * wl[1 + d];
    wbr[1 + d] += velb[d] * sqr * isqs + hrb * wr[1 + d];
    isqsb += velb[d] * (sql * wl[1 + d] + sqr * wr[1 + d]);
  }
  wbl[D + 1] += hlb * g;
  wbr[D + 1] += hrb * g;
  const double sqsb = -isqsb * isqs2;
  sqlb += sqsb;
  sqrb += sqsb;
  wbl[0] += rhob * wr[0] / (2.0 * rho) + sqlb / (2.0 * sql);
  wbr[0] += rhob * wl[0] / (2.0 * rho) + sqrb / (2.0 * sqr);
  // dw = wr - wl
#pragma unroll
  for (int i = 0; i < D + 3; i++) {
    wbr[i] += dwb[i];
    wbl[i] -= dwb[i];
  }
}

// numerical convective flux h = H(u_L, u_R; n) (ConvectiveFlux.cpp:93-350, the Central,
// Lax-Friedrichs, HLLC and Roe cases): wbl/wbr += adjoints of the computational variables, ubl/ubr +=
// adjoints of the conserved states where H uses them directly. Returns false for the flux without a
// reverse form here (Exact), leaving the outputs untouched.
template <int D>
__device__ bool convectiveFluxVjp(const Phys& ph, const double* n, const double* ul, const double* wl,
                                  const double* ur, const double* wr, const double* hb, double* wbl, double* wbr,
                                  double* ubl, double* ubr) {
  constexpr int NV = D + 2;
  if (ph.flux == 0) {  // (F_L + F_R) / 2
    double fb[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) fb[v] = hb[v] / 2.0;
    normalFluxVjp<D>(n, wl, fb, wbl);
    normalFluxVjp<D>(n, wr, fb, wbr);
    return true;
  }
  if (ph.flux == 1) {  // ((F_L + F_R) - lam (u_R - u_L)) / 2, lam = max(|un_L| + c_L, |un_R| + c_R)
    double fb[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) fb[v] = hb[v] / 2.0;
    normalFluxVjp<D>(n, wl, fb, wbl);
    normalFluxVjp<D>(n, wr, fb, wbr);
    double unl = 0.0, unr = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
      unl += wl[1 + d] * n[d];
      unr += wr[1 + d] * n[d];
    }
    const double cl = soundSpeedOf(ph, wl[0], wl[D + 2]), cr = soundSpeedOf(ph, wr[0], wr[D + 2]);
    const double al = fabs(unl) + cl, ar = fabs(unr) + cr;
    const double lam = fmax(al, ar);
    double lamb = 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      lamb -= hb[v] * (ur[v] - ul[v]) / 2.0;
      ubr[v] -= lam * hb[v] / 2.0;
      ubl[v] += lam * hb[v] / 2.0;
    }
    // lam-bar to the side that attains the maximum (two branches, so that no register array is
    // selected through a runtime pointer)
    auto side = [&](const double* w, double* wb, double un, double c) {
      const double unb = un < 0.0 ? -lamb : lamb;
#pragma unroll
      for (int d = 0; d < D; d++) wb[1 + d] += unb * n[d];
      double rb = 0.0, pb = 0.0;
      soundSpeedVjp(ph, w[0], w[D + 2], c, lamb, rb, pb);
      wb[0] += rb;
      wb[D + 2] += pb;
    };
    if (al >= ar)
      side(wl, wbl, unl, cl);
    else
      side(wr, wbr, unr, cr);
    return true;
  }
  if (ph.flux == 3) {
    roeFluxVjp<D>(ph, n, wl, wr, hb, wbl, wbr);
    return true;
  }
  if (ph.flux != 2) return false;
  // HLLC
  const double g = kGamma;
  const double rl = wl[0], rr = wr[0], pl = wl[D + 2], pr = wr[D + 2];
  double unl = 0.0, unr = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    unl += wl[1 + d] * n[d];
    unr += wr[1 + d] * n[d];
  }
  const double cl = soundSpeedOf(ph, rl, pl), cr = soundSpeedOf(ph, rr, pr);
  const double rbar = (rl + rr) / 2.0, cbar = (cl + cr) / 2.0;
  const double psr = (pl + pr) / 2.0 - (unr - unl) * rbar * cbar;
  const double ps = fmax(0.0, psr);
  const bool shl = !(ps <= pl), shr = !(ps <= pr);
  const double ql = shl ? sqrt(1.0 + (g + 1.0) * (ps / pl - 1.0) / 2.0 / g) : 1.0;
  const double sl = unl - cl * ql;
  if (sl >= 0.0) {
    normalFluxVjp<D>(n, wl, hb, wbl);
    return true;
  }
  const double qr = shr ? sqrt(1.0 + (g + 1.0) * (ps / pr - 1.0) / 2.0 / g) : 1.0;
  const double sr = unr + cr * qr;
  if (sr <= 0.0) {
    normalFluxVjp<D>(n, wr, hb, wbr);
    return true;
  }
  const double al = sl - unl, ar = sr - unr;
  const double num = pr - pl + rl * unl * al - rr * unr * ar;
  const double den = rl * al - rr * ar;
  const double sm = num / den;
  double rlb = 0.0, rrb = 0.0, unlb = 0.0, unrb = 0.0, plb = 0.0, prb = 0.0, slb = 0.0, srb = 0.0, psb = 0.0,
         smb = 0.0;
  if (sm >= 0.0)
    hllcStarVjp<D>(n, ul, wl, sl, unl, ps, sm, hb, wbl, ubl, rlb, unlb, plb, slb, psb, smb);
  else
    hllcStarVjp<D>(n, ur, wr, sr, unr, ps, sm, hb, wbr, ubr, rrb, unrb, prb, srb, psb, smb);
  // sm = num / den
  const double numb = smb / den, denb = -smb * sm / den;
  prb += numb;
  plb -= numb;
  rlb += numb * unl * al + denb * al;
  unlb += numb * rl * al;
  double alb = numb * rl * unl + denb * rl;
  rrb -= numb * unr * ar + denb * ar;
  unrb -= numb * rr * ar;
  double arb = -numb * rr * unr - denb * rr;
  slb += alb;  // al = sl - unl
  unlb -= alb;
  srb += arb;  // ar = sr - unr
  unrb -= arb;
  // sr = unr + cr qr, sl = unl - cl ql
  unrb += srb;
  double crb = srb * qr;
  const double qrb = srb * cr;
  unlb += slb;
  double clb = -slb * ql;
  const double qlb = -slb * cl;
  // q = sqrt(1 + (g + 1) (ps / p - 1) / (2 g)) when ps > p
  if (shl) {
    const double t = qlb * (g + 1.0) / (4.0 * g * ql);
    psb += t / pl;
    plb -= t * ps / (pl * pl);
  }
  if (shr) {
    const double t = qrb * (g + 1.0) / (4.0 * g * qr);
    psb += t / pr;
    prb -= t * ps / (pr * pr);
  }
  // ps = max(0, (pl + pr) / 2 - (unr - unl) rbar cbar)
  if (!(0.0 >= psr)) {
    plb += psb / 2.0;
    prb += psb / 2.0;
    const double du = unr - unl;
    unrb -= psb * rbar * cbar;
    unlb += psb * rbar * cbar;
    const double rbarb = -psb * du * cbar, cbarb = -psb * du * rbar;
    rlb += rbarb / 2.0;
    rrb += rbarb / 2.0;
    clb += cbarb / 2.0;
    crb += cbarb / 2.0;
  }
  soundSpeedVjp(ph, rl, pl, cl, clb, rlb, plb);
  soundSpeedVjp(ph, rr, pr, cr, crb, rrb, prb);
#pragma unroll
  for (int d = 0; d < D; d++) {
    wbl[1 + d] += unlb * n[d];
    wbr[1 + d] += unrb * n[d];
  }
  wbl[0] += rlb;
  wbl[D + 2] += plb;
  wbr[0] += rrb;
  wbr[D + 2] += prb;
  return true;
}

// state part of the viscous flux: wb += (d fv / dw)^T Fvb at a fixed conserved gradient g, for
// fv = viscousRawFlux(w, primitiveGradient(w, u, g)) (ViscousFlux.cpp:58-103,
// VariableConvertor.cpp:574-620). primitiveGradient's u_E / rho is differentiated as e + |v|^2 / 2,
// the same function of w.
template <int D>
__device__ __forceinline__ void viscousStateVjp(const Phys& ph, const double* u, const double* w, const double* g,
                                                const double* Fvb, double* wb) {
  constexpr int NG = (D + 2) * D;
  double pg[NG];
  primitiveGradient<D>(ph, w, u, g, pg);
  const double rho = w[0], ir = 1.0 / rho, icv = 1.0 / ph.cv;
  double mu = ph.mu, kap = ph.kappa, r = 1.0, t = 0.0;
  if (ph.transport == 2) {
    t = w[D + 1] / ph.cv;
    r = sutherlandRatio(t);
    mu = mu * r;
    kap = kap * r;
  }
  const double* vg = pg + D;
  const double* tg = pg + (D + 1) * D;
  double tr = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) tr += vg[a + D * a];
  // tau = mu (vg + vg^T) - 2/3 mu tr I; its adjoint from fv[a + D (1 + b)] = tau[a + D b] and
  // fv[a + D (D + 1)] = sum_b tau[a + D b] v_b + kap tg_a
  double taub[D * D], tgb[D], kapb = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const double fe = Fvb[a + D * (D + 1)];
#pragma unroll
    for (int b = 0; b < D; b++) {
      const double tau = mu * (vg[a + D * b] + vg[b + D * a]) - (a == b ? 2.0 / 3.0 * mu * tr : 0.0);
      taub[a + D * b] = Fvb[a + D * (1 + b)] + fe * w[1 + b];
      wb[1 + b] += fe * tau;
    }
    tgb[a] = kap * fe;
    kapb += fe * tg[a];
  }
  double trb = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) trb += taub[a + D * a];
  double vgb[D * D], mub = 0.0;
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int a = 0; a < D; a++) {
      vgb[a + D * b] = mu * (taub[a + D * b] + taub[b + D * a]) - (a == b ? 2.0 / 3.0 * mu * trb : 0.0);
      mub += taub[a + D * b] * ((vg[a + D * b] + vg[b + D * a]) - (a == b ? 2.0 / 3.0 * tr : 0.0));
    }
  if (ph.transport == 2) {  // mu = mu0 r(T), kap = kap0 r(T), T = e / c_v, r'/r = 3/(2T) - 1/(T + S)
    const double rp = r * (1.5 / t - 1.0 / (t + kSutherland));
    wb[D + 1] += (mub * ph.mu + kapb * ph.kappa) * rp / ph.cv;
  }
  // tg_a = de_a / c_v, de_a = (g_E,a - g_rho,a etot) / rho - sum_b vg[a + D b] v_b, etot = u_E / rho
  const double etot = u[D + 1] * ir;
  double etotb = 0.0, irb = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const double deb = tgb[a] * icv;
    etotb -= deb * g[a] * ir;
    irb += deb * (g[(D + 1) * D + a] - g[a] * etot);
#pragma unroll
    for (int b = 0; b < D; b++) {
      vgb[a + D * b] -= deb * w[1 + b];
      wb[1 + b] -= deb * vg[a + D * b];
    }
  }
  // vg[a + D b] = (g_m_b,a - g_rho,a v_b) / rho
#pragma unroll
  for (int b = 0; b < D; b++)
#pragma unroll
    for (int a = 0; a < D; a++) {
      wb[1 + b] -= vgb[a + D * b] * g[a] * ir;
      irb += vgb[a + D * b] * (g[(1 + b) * D + a] - g[a] * w[1 + b]);
    }
  wb[0] -= irb * ir * ir;
  wb[D + 1] += etotb;  // etot = e + |v|^2 / 2
#pragma unroll
  for (int d = 0; d < D; d++) wb[1 + d] += etotb * w[1 + d];
}

// state part of the volume-flux VJP: ub = (dF/du)^T Fb for the raw volume flux F = F_c(u) -
// F_v(u, g) (D x NV column-major) at a fixed gradient g
template <int D>
__device__ __forceinline__ void volumeFluxStateVjp(const Phys& ph, const double* u, const double* g, const double* Fb,
                                                   double* ub) {
  constexpr int NV = D + 2;
  double w[D + 3], wb[D + 3];
  computational<D>(ph, u, w);
#pragma unroll
  for (int i = 0; i < D + 3; i++) wb[i] = 0.0;
  // F_c[a][.] is the normal flux with n = e_a (a rolled loop: unrolled, the three passes' temporaries
  // push the volume kernel at P4 further past two wavefronts' registers)
#pragma unroll 1
  for (int a = 0; a < D; a++) {
    double n[D], fb[NV];
#pragma unroll
    for (int d = 0; d < D; d++) n[d] = d == a ? 1.0 : 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) fb[v] = Fb[a + D * v];
    normalFluxVjp<D>(n, w, fb, wb);
  }
  double fvb[D * NV];
#pragma unroll
  for (int i = 0; i < D * NV; i++) fvb[i] = -Fb[i];
  viscousStateVjp<D>(ph, u, w, g, fvb, wb);
  computationalVjp<D>(ph, u, w, wb, ub);
}

}  // namespace sdg

#endif  // SDG_PHYSICS_ADJOINT_HPP_
