// xc_functionals_mgga.hpp -- TPSS meta-GGA exchange and correlation in FP64
// for the restricted and unrestricted builds, evaluated per grid point on the
// GPU.
//
// Replaces libxc 6.2.2's unpolarised xc_mgga_exc_vxc for mgga_x_tpss and
// mgga_c_tpss as called from DensityFunctional::evaluate with Family::MGGA
// (src/dft/functional.cpp:49-57): inputs (rho, sigma, tau) of the closed-shell
// density, tau = (1/2) sum_i |grad psi_i|^2 (src/gto/density.cpp:94-101);
// outputs e = rho*eps and vrho, vsigma, vtau (vlaplacian is zero for both
// functionals and the reference's Vxc does not use it).  Each energy density
// is written once over the forward-mode duals of xc_functionals_pol.hpp with
// the partials with respect to (rho, sigma, tau) in slots 0..2; TPSS
// correlation uses that header's spin-resolved PBE correlation.
//
// Formulas: J. Tao, J.P. Perdew, V.N. Staroverov, G.E. Scuseria, PRL 91,
// 146401 (2003), eqs. 5-10 (exchange) and 11-14 at zeta = 0 (correlation:
// xi = 0, C(0, 0) = 0.53, d = 2.8).  Inputs are limited as libxc limits
// meta-GGA inputs (tau >= 1e-20, sigma <= 8 rho tau, the Fermi-hole
// curvature bound tau_W <= tau), derivatives taken at the limited inputs;
// the correlation's tau_W / tau is held at 1 above 1.
#pragma once

#include "xc_functionals_pol.hpp"

namespace occxc {

constexpr double kTauMin = 1e-20;

enum : int { XC_MGGA_X_TPSS = 202, XC_MGGA_C_TPSS = 231 };

// 0 LDA, 1 GGA, 2 meta-GGA, -1 unknown
OCCXC_HD inline int xc_family_mgga(int id) {
  if (id == XC_MGGA_X_TPSS || id == XC_MGGA_C_TPSS) return 2;
  return xc_family(id);
}

// exchange: F_x(p, z) with p = s^2, z = tau_W / tau and
// alpha = (tau - tau_W) / tau_unif
OCCXC_HD inline D5 ex_tpss(const D5 &rho, const D5 &sigma, const D5 &tau) {
  const double b = 0.40, c = 1.59096, e = 1.537, kappa = 0.804, mu = 0.21951;
  const double mge = 10.0 / 81.0, se = sqrt(e);
  const double k2 = pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const D5 p = sigma / (4.0 * k2 * d5_pow(rho, 8.0 / 3.0));
  const D5 z = sigma / (8.0 * rho * tau);
  const D5 alpha = (tau - sigma / (8.0 * rho)) / (0.3 * k2 * d5_pow(rho, 5.0 / 3.0));
  const D5 qb = 0.45 * (alpha - 1.0) / d5_sqrt(1.0 + b * alpha * (alpha - 1.0)) + 2.0 * p / 3.0;
  const D5 z2 = z * z;
  const D5 opz2 = 1.0 + z2;
  const D5 num = (mge + c * z2 / (opz2 * opz2)) * p + (146.0 / 2025.0) * qb * qb -
                 (73.0 / 405.0) * qb * d5_sqrt(0.5 * ((9.0 / 25.0) * z2 + p * p)) +
                 (mge * mge / kappa) * p * p + (2.0 * se * mge * 9.0 / 25.0) * z2 +
                 (e * mu) * p * p * p;
  const D5 den = (1.0 + se * p) * (1.0 + se * p);
  const D5 x = num / den;
  const D5 F = 1.0 + kappa - kappa / (1.0 + x / kappa);
  return ex_dirac(rho) * F;
}

// correlation at zeta = 0: eps_revPKZB = eps_PBE (1 + C z^2)
// - (1 + C) z^2 max(eps_PBE(n/2, 0), eps_PBE), eps = eps_revPKZB (1 + d eps_revPKZB z^3)
OCCXC_HD inline D5 ec_tpss(const D5 &rho, const D5 &sigma, const D5 &tau) {
  const double d = 2.8, C0 = 0.53;
  const D5 half = 0.5 * rho;
  const D5 epbe = ec_pbe_pol(half, half, sigma) / rho;
  const D5 epol = ec_pbe_pol(half, d5c(0.0), 0.25 * sigma) / half;
  const D5 et = epol.v > epbe.v ? epol : epbe;
  D5 z = sigma / (8.0 * rho * tau);
  if (z.v > 1.0) z = d5c(1.0);
  const D5 z2 = z * z;
  const D5 f0 = (1.0 + C0 * z2) * epbe - (1.0 + C0) * z2 * et;
  return rho * f0 * (1.0 + d * f0 * z2 * z);
}

// One functional at one point of a meta-GGA build: e = rho*eps, vr, vs, vt
// (SET, like xc_point).  LDA and GGA ids go to xc_point (tau unused, vt = 0).
// Returns false for an unknown id.  The caller applies the density
// threshold and sigma >= kSigmaMin.
OCCXC_HD inline bool xc_point_mgga(int id, double rho, double sigma, double tau, double &e,
                                   double &vr, double &vs, double &vt) {
  vt = 0.0;
  if (id != XC_MGGA_X_TPSS && id != XC_MGGA_C_TPSS) return xc_point(id, rho, sigma, e, vr, vs);
  tau = fmax(tau, kTauMin);
  sigma = fmin(sigma, 8.0 * rho * tau);
  const D5 r = d5v(rho, 0), s = d5v(sigma, 1), t = d5v(tau, 2);
  const D5 out = id == XC_MGGA_X_TPSS ? ex_tpss(r, s, t) : ec_tpss(r, s, t);
  e = out.v;
  vr = out.d[0];
  vs = out.d[1];
  vt = out.d[2];
  return true;
}

// Spin-polarised TPSS (same paper): exchange by exact spin scaling,
// E_x[na, nb] = E_x[2 na]/2 + E_x[2 nb]/2 with (2 n_s, 4 s_ss, 2 t_s);
// correlation by eqs. 11-14 in full: zeta = (na - nb)/n, |grad zeta|^2 =
// 4 (nb^2 s_aa - 2 na nb s_ab + na^2 s_bb) / n^4, xi^2 = |grad zeta|^2 /
// (4 (3 pi^2 n)^(2/3)), C(zeta, xi) = (0.53 + 0.87 zeta^2 + 0.50 zeta^4 +
// 2.26 zeta^6) / (1 + xi^2 ((1+zeta)^(-4/3) + (1-zeta)^(-4/3))/2)^4, the
// spin channels' max(eps_PBE(n_s, 0, s_ss), eps_PBE) weighted n_s / n, and
// z = tau_W / tau of the total density (held at 1 above 1).
OCCXC_HD inline D5 ex_tpss_pol(const D5 &ra, const D5 &rb, const D5 &saa, const D5 &sbb,
                               const D5 &ta, const D5 &tb) {
  D5 e = d5c(0.0);
  if (ra.v >= kDensThreshold) e = e + 0.5 * ex_tpss(2.0 * ra, 4.0 * saa, 2.0 * ta);
  if (rb.v >= kDensThreshold) e = e + 0.5 * ex_tpss(2.0 * rb, 4.0 * sbb, 2.0 * tb);
  return e;
}

OCCXC_HD inline D5 ec_tpss_pol(const D5 &ra, const D5 &rb, const D5 &saa, const D5 &sab,
                               const D5 &sbb, const D5 &ta, const D5 &tb) {
  const double d = 2.8;
  const D5 n = ra + rb, zeta = (ra - rb) / n;
  const D5 sig = saa + 2.0 * sab + sbb;
  const D5 epbe = ec_pbe_pol(ra, rb, sig) / n;
  D5 par = d5c(0.0);
  if (ra.v >= kDensThreshold) {
    const D5 ea = ec_pbe_pol(ra, d5c(0.0), saa) / ra;
    par = par + ra / n * (ea.v > epbe.v ? ea : epbe);
  }
  if (rb.v >= kDensThreshold) {
    const D5 eb = ec_pbe_pol(rb, d5c(0.0), sbb) / rb;
    par = par + rb / n * (eb.v > epbe.v ? eb : epbe);
  }
  D5 z = sig / (8.0 * n * (ta + tb));
  if (z.v > 1.0) z = d5c(1.0);
  const D5 n2 = n * n;
  const D5 gz2 = 4.0 * (rb * rb * saa - 2.0 * ra * rb * sab + ra * ra * sbb) / (n2 * n2);
  const D5 xi2 = gz2 / (4.0 * d5_pow(3.0 * kPi * kPi * n, 2.0 / 3.0));
  const D5 q = zeta * zeta;
  const D5 c0 = 0.53 + 0.87 * q + 0.50 * q * q + 2.26 * q * q * q;
  const D5 den = 1.0 + xi2 * (opz_pow(1.0 + zeta, -4.0 / 3.0) + opz_pow(1.0 - zeta, -4.0 / 3.0)) / 2.0;
  const D5 den2 = den * den;
  const D5 C = c0 / (den2 * den2);
  const D5 z2 = z * z;
  const D5 f0 = (1.0 + C * z2) * epbe - (1.0 + C) * z2 * par;
  return n * f0 * (1.0 + d * f0 * z2 * z);
}

// One functional at one polarised point of a meta-GGA build, ACCUMULATED
// into e (= n eps), vr[2], vs[3] (aa, ab, bb), vt[2] like xc_point_pol.  The
// seven partials come from two passes over the five dual slots: (rho_a,
// rho_b, s_aa, s_ab, s_bb), then (tau_a, tau_b).  Inputs are limited per
// spin as for the closed-shell form (tau_s >= 1e-20, s_ss <= 8 n_s tau_s),
// then |s_ab| <= (s_aa + s_bb) / 2.
// LDA and GGA ids go to xc_point_pol.  Returns false for an unknown id.
OCCXC_HD inline bool xc_point_mgga_pol(int id, double ra, double rb, double saa, double sab,
                                       double sbb, double ta, double tb, double &e, double *vr,
                                       double *vs, double *vt) {
  if (id != XC_MGGA_X_TPSS && id != XC_MGGA_C_TPSS) return xc_point_pol(id, ra, rb, saa, sab, sbb, e, vr, vs);
  ta = fmax(ta, kTauMin);
  tb = fmax(tb, kTauMin);
  saa = fmin(fmax(saa, kSigmaMin), 8.0 * ra * ta);
  sbb = fmin(fmax(sbb, kSigmaMin), 8.0 * rb * tb);
  // |grad n_a . grad n_b| <= (s_aa + s_bb) / 2 after the per-spin limits, so
  // that |grad n|^2 and |grad zeta|^2 stay non-negative
  const double sab_max = 0.5 * (saa + sbb);
  sab = fmin(fmax(sab, -sab_max), sab_max);
  const bool x = id == XC_MGGA_X_TPSS;
  {
    const D5 a = d5v(ra, 0), b = d5v(rb, 1), s0 = d5v(saa, 2), s1 = d5v(sab, 3), s2 = d5v(sbb, 4);
    const D5 out = x ? ex_tpss_pol(a, b, s0, s2, d5c(ta), d5c(tb))
                     : ec_tpss_pol(a, b, s0, s1, s2, d5c(ta), d5c(tb));
    e += out.v;
    vr[0] += out.d[0];
    vr[1] += out.d[1];
    vs[0] += out.d[2];
    vs[1] += out.d[3];
    vs[2] += out.d[4];
  }
  const D5 a = d5c(ra), b = d5c(rb), s0 = d5c(saa), s1 = d5c(sab), s2 = d5c(sbb);
  const D5 out = x ? ex_tpss_pol(a, b, s0, s2, d5v(ta, 0), d5v(tb, 1))
                   : ec_tpss_pol(a, b, s0, s1, s2, d5v(ta, 0), d5v(tb, 1));
  vt[0] += out.d[0];
  vt[1] += out.d[1];
  return true;
}

} // namespace occxc
