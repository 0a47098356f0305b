// xc_functionals_pol.hpp -- spin-polarised LDA/GGA functionals in FP64 for
// the unrestricted build, evaluated per grid point on the GPU.
//
// Replaces libxc 6.2.2's polarised xc_lda_exc_vxc / xc_gga_exc_vxc as called
// from DensityFunctional::evaluate with SpinorbitalKind::Unrestricted
// (src/dft/functional.cpp:18-65): inputs (rho_a, rho_b, sigma_aa, sigma_ab,
// sigma_bb), outputs the energy density e = rho*eps and its five partial
// derivatives (libxc's vrho[2], vsigma[3]).  Each energy density is written
// once over a forward-mode dual number carrying the five partials, so the
// derivatives follow from the energy expression itself.
//
// Formulas: exchange by exact spin scaling, E_x[ra, rb] = E_x[2 ra]/2 +
// E_x[2 rb]/2 (Oliver, Perdew, PRA 20, 397 (1979)), of Dirac, PBE (Perdew,
// Burke, Ernzerhof, PRL 77, 3865 (1996)) and B88 (Becke, PRA 38, 3098
// (1988)) exchange; VWN5 (Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200
// (1980), the paramagnetic, ferromagnetic and spin-stiffness fits and their
// zeta interpolation); PW92 (Perdew, Wang, PRB 45, 13244 (1992), eqs. 8-10,
// Table I, with libxc's "mod" constants for lda_c_pw_mod and PBE); PBE
// correlation with phi(zeta); LYP in the general form of Miehlich, Savin,
// Stoll, Preuss, CPL 157, 200 (1989), eq. 2.  Powers of (1 +- zeta) below
// DBL_EPSILON are held at DBL_EPSILON^n (libxc's zeta threshold).
#pragma once

#include <math.h>

#include "xc_functionals.hpp"

namespace occxc {

// value and partials with respect to (ra, rb, saa, sab, sbb)
struct D5 {
  double v;
  double d[5];
};

OCCXC_HD inline D5 d5c(double v) {
  D5 r;
  r.v = v;
  for (int k = 0; k < 5; ++k) r.d[k] = 0.0;
  return r;
}
OCCXC_HD inline D5 d5v(double v, int k) {
  D5 r = d5c(v);
  r.d[k] = 1.0;
  return r;
}
// y = f(x) with f'(x) = fp
OCCXC_HD inline D5 d5chain(const D5 &x, double fx, double fp) {
  D5 r;
  r.v = fx;
  for (int k = 0; k < 5; ++k) r.d[k] = fp * x.d[k];
  return r;
}
OCCXC_HD inline D5 operator+(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v + b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
OCCXC_HD inline D5 operator-(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v - b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
OCCXC_HD inline D5 operator-(const D5 &a) {
  D5 r;
  r.v = -a.v;
  for (int k = 0; k < 5; ++k) r.d[k] = -a.d[k];
  return r;
}
OCCXC_HD inline D5 operator*(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v * b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
OCCXC_HD inline D5 operator/(const D5 &a, const D5 &b) {
  D5 r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
OCCXC_HD inline D5 operator+(const D5 &a, double b) { D5 r = a; r.v += b; return r; }
OCCXC_HD inline D5 operator+(double b, const D5 &a) { return a + b; }
OCCXC_HD inline D5 operator-(const D5 &a, double b) { D5 r = a; r.v -= b; return r; }
OCCXC_HD inline D5 operator-(double b, const D5 &a) { return -a + b; }
OCCXC_HD inline D5 operator*(const D5 &a, double b) {
  D5 r;
  r.v = a.v * b;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b;
  return r;
}
OCCXC_HD inline D5 operator*(double b, const D5 &a) { return a * b; }
OCCXC_HD inline D5 operator/(const D5 &a, double b) { return a * (1.0 / b); }
OCCXC_HD inline D5 operator/(double a, const D5 &b) { return d5c(a) / b; }

OCCXC_HD inline D5 d5_sqrt(const D5 &x) {
  const double s = sqrt(x.v);
  return d5chain(x, s, 0.5 / s);
}
OCCXC_HD inline D5 d5_cbrt(const D5 &x) {
  const double c = cbrt(x.v);
  return d5chain(x, c, c / (3.0 * x.v));
}
// x^p; at x = 0 (a spin channel with no density, e.g. the beta spin of a
// one-electron system) value 0 and, for p > 1, derivative 0
OCCXC_HD inline D5 d5_pow(const D5 &x, double p) {
  if (x.v == 0.0 && p > 1.0) return d5c(0.0);
  const double y = pow(x.v, p);
  return d5chain(x, y, p * y / x.v);
}
OCCXC_HD inline D5 d5_exp(const D5 &x) {
  const double e = exp(x.v);
  return d5chain(x, e, e);
}
OCCXC_HD inline D5 d5_log(const D5 &x) { return d5chain(x, log(x.v), 1.0 / x.v); }
OCCXC_HD inline D5 d5_log1p(const D5 &x) { return d5chain(x, log1p(x.v), 1.0 / (1.0 + x.v)); }
OCCXC_HD inline D5 d5_atan(const D5 &x) { return d5chain(x, atan(x.v), 1.0 / (1.0 + x.v * x.v)); }
OCCXC_HD inline D5 d5_asinh(const D5 &x) {
  return d5chain(x, asinh(x.v), 1.0 / sqrt(1.0 + x.v * x.v));
}

constexpr double kZetaThr = 2.220446049250313e-16;

// (1 +- zeta)^n, held at kZetaThr^n below the threshold
OCCXC_HD inline D5 opz_pow(const D5 &x, double n) {
  if (x.v <= kZetaThr) return d5c(pow(kZetaThr, n));
  return d5_pow(x, n);
}
OCCXC_HD inline D5 f_zeta(const D5 &z) {
  const double den = pow(2.0, 4.0 / 3.0) - 2.0;
  return (opz_pow(1.0 + z, 4.0 / 3.0) + opz_pow(1.0 - z, 4.0 / 3.0) - 2.0) / den;
}

// VWN interpolation formula (VWN eq. 4.4) with parameters A, b, c, x0
OCCXC_HD inline D5 vwn_aux(const D5 &rs, double A, double b, double c, double x0) {
  const D5 x = d5_sqrt(rs);
  const D5 X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c;
  const double Q = sqrt(4.0 * c - b * b);
  const D5 at = d5_atan(Q / (2.0 * x + b));
  return A * (d5_log(x * x / X) + (2.0 * b / Q) * at -
              (b * x0 / X0) * (d5_log((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}

// PW92 eq. 10, p = 1
OCCXC_HD inline D5 pw_G(const D5 &rs, double A, double a1, double b1, double b2, double b3,
                        double b4) {
  const D5 srs = d5_sqrt(rs);
  const D5 den = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * d5_log1p(1.0 / den);
}

OCCXC_HD inline D5 eps_pw_pol(const D5 &rs, const D5 &z, bool mod) {
  const double A0 = mod ? 0.0310907 : 0.031091, A1 = mod ? 0.01554535 : 0.015545,
               Aa = mod ? 0.0168869 : 0.016887;
  const double fpp = mod ? 1.709920934161365617563962776245 : 1.709921;
  const D5 e0 = pw_G(rs, A0, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const D5 e1 = pw_G(rs, A1, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const D5 mac = pw_G(rs, Aa, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);  // -alpha_c
  const D5 f = f_zeta(z), z2 = z * z, z4 = z2 * z2;
  return e0 - mac * f * (1.0 - z4) / fpp + (e1 - e0) * f * z4;
}

OCCXC_HD inline D5 eps_vwn_pol(const D5 &rs, const D5 &z) {
  const double fpp = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
  const D5 eP = vwn_aux(rs, 0.0310907, 3.72744, 12.9352, -0.10498);
  const D5 eF = vwn_aux(rs, 0.01554535, 7.06042, 18.0578, -0.32500);
  const D5 ac = vwn_aux(rs, -1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584);
  const D5 f = f_zeta(z), z2 = z * z, z4 = z2 * z2;
  return eP + ac * f * (1.0 - z4) / fpp + (eF - eP) * f * z4;
}

// Dirac exchange energy density of a closed-shell density rho
OCCXC_HD inline D5 ex_dirac(const D5 &rho) { return -kCx * rho * d5_cbrt(rho); }

// one spin channel (density r, sigma s) of a spin-scaled exchange functional
OCCXC_HD inline D5 ex_spin(int id, const D5 &r, const D5 &s) {
  if (r.v < kDensThreshold) return d5c(0.0);
  const D5 rho = 2.0 * r;
  switch (id) {
  case XC_LDA_X:
    return 0.5 * ex_dirac(rho);
  case XC_GGA_X_PBE: {
    const double kappa = 0.804, mu = 0.06672455060314922 * kPi * kPi / 3.0;
    const D5 kf = d5_cbrt(3.0 * kPi * kPi * rho);
    const D5 s2 = (4.0 * s) / (4.0 * kf * kf * rho * rho);
    const D5 F = 1.0 + kappa - kappa / (1.0 + mu * s2 / kappa);
    return 0.5 * ex_dirac(rho) * F;
  }
  case XC_GGA_X_B88: {
    const double beta = 0.0042;
    const D5 r43 = r * d5_cbrt(r);
    const D5 x = d5_sqrt(s) / r43;
    const D5 g = beta * r43 * x * x / (1.0 + 6.0 * beta * x * d5_asinh(x));
    return 0.5 * ex_dirac(rho) - g;
  }
  }
  return d5c(0.0);
}

OCCXC_HD inline D5 ec_pbe_pol(const D5 &ra, const D5 &rb, const D5 &sigma) {
  const double beta = 0.06672455060314922, gamma = (1.0 - log(2.0)) / (kPi * kPi);
  const D5 rho = ra + rb, z = (ra - rb) / rho;
  const D5 rs = kRsFac / d5_cbrt(rho);
  const D5 ec = eps_pw_pol(rs, z, true);
  const D5 phi = 0.5 * (opz_pow(1.0 + z, 2.0 / 3.0) + opz_pow(1.0 - z, 2.0 / 3.0));
  const D5 phi3 = phi * phi * phi;
  const D5 kf = d5_cbrt(3.0 * kPi * kPi * rho);
  const D5 ks2 = 4.0 * kf / kPi;
  const D5 t2 = sigma / (4.0 * phi * phi * ks2 * rho * rho);
  const D5 A = (beta / gamma) / (d5_exp(-ec / (gamma * phi3)) - 1.0);
  const D5 At2 = A * t2;
  const D5 H = gamma * phi3 *
               d5_log1p((beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

OCCXC_HD inline D5 ec_lyp_pol(const D5 &ra, const D5 &rb, const D5 &saa, const D5 &sab,
                              const D5 &sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const D5 rho = ra + rb, gg = saa + 2.0 * sab + sbb;
  const D5 rm13 = 1.0 / d5_cbrt(rho);
  const D5 den = 1.0 + d * rm13;
  const D5 omega = d5_exp(-c * rm13) / den * d5_pow(rho, -11.0 / 3.0);
  const D5 delta = c * rm13 + d * rm13 / den;
  const D5 t1 = -a * 4.0 / den * ra * rb / rho;
  D5 br = ra * rb *
          (pow(2.0, 11.0 / 3.0) * CF * (d5_pow(ra, 8.0 / 3.0) + d5_pow(rb, 8.0 / 3.0)) +
           (47.0 / 18.0 - 7.0 / 18.0 * delta) * gg - (2.5 - delta / 18.0) * (saa + sbb) -
           (delta - 11.0) / 9.0 * (ra / rho * saa + rb / rho * sbb));
  br = br - 2.0 / 3.0 * rho * rho * gg + (2.0 / 3.0 * rho * rho - ra * ra) * sbb +
       (2.0 / 3.0 * rho * rho - rb * rb) * saa;
  return t1 - a * b * omega * br;
}

// One polarised functional at one point: e = rho*eps, vr[2], vs[3]
// (ADDED to the outputs).  Returns false for an unknown id.
OCCXC_HD inline bool xc_point_pol(int id, double ra_, double rb_, double saa_, double sab_,
                                  double sbb_, double &e, double *vr, double *vs) {
  const D5 ra = d5v(ra_, 0), rb = d5v(rb_, 1), saa = d5v(saa_, 2), sab = d5v(sab_, 3),
           sbb = d5v(sbb_, 4);
  const D5 rho = ra + rb;
  D5 r;
  switch (id) {
  case XC_LDA_X:
  case XC_GGA_X_PBE:
  case XC_GGA_X_B88:
    r = ex_spin(id, ra, saa) + ex_spin(id, rb, sbb);
    break;
  case XC_LDA_C_VWN:
    r = rho * eps_vwn_pol(kRsFac / d5_cbrt(rho), (ra - rb) / rho);
    break;
  case XC_LDA_C_PW:
  case XC_LDA_C_PW_MOD:
    r = rho * eps_pw_pol(kRsFac / d5_cbrt(rho), (ra - rb) / rho, id == XC_LDA_C_PW_MOD);
    break;
  case XC_GGA_C_PBE:
    r = ec_pbe_pol(ra, rb, saa + 2.0 * sab + sbb);
    break;
  case XC_GGA_C_LYP:
    r = ec_lyp_pol(ra, rb, saa, sab, sbb);
    break;
  default:
    return false;
  }
  e += r.v;
  vr[0] += r.d[0];
  vr[1] += r.d[1];
  vs[0] += r.d[2];
  vs[1] += r.d[3];
  vs[2] += r.d[4];
  return true;
}

} // namespace occxc
