// xc_functionals.hpp -- unpolarised LDA/GGA functionals in FP64 with
// analytic first derivatives, evaluated per grid point on the GPU.
//
// Replaces libxc 6.2.2's xc_lda_exc_vxc / xc_gga_exc_vxc as called from
// DensityFunctional::evaluate (src/dft/functional.cpp:18-65).  Each routine
// returns the energy density e = rho*eps and its partial derivatives
// vrho = de/drho, vsigma = de/dsigma (libxc's output convention).
//
// Formulas (closed-shell forms derived from the cited papers):
//   Slater/Dirac exchange;  VWN5 (Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200
//   (1980), paramagnetic Monte-Carlo fit);  PW92 (Perdew, Wang, PRB 45, 13244
//   (1992));  PBE exchange and correlation (Perdew, Burke, Ernzerhof, PRL 77,
//   3865 (1996));  B88 (Becke, PRA 38, 3098 (1988));  LYP in the closed-shell
//   form of Miehlich, Savin, Stoll, Preuss, CPL 157, 200 (1989).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define OCCXC_HD __host__ __device__
#else
#define OCCXC_HD
#endif

namespace occxc {

// Below this density a point contributes nothing; sigma is clamped from
// below.  libxc's own per-functional thresholds are not visible from OCC.
constexpr double kDensThreshold = 1e-15;
constexpr double kSigmaMin = 1e-40;

constexpr double kPi = 3.14159265358979323846;
constexpr double kCx = 0.73855876638202240587;      // (3/4)(3/pi)^(1/3)
constexpr double kRsFac = 0.62035049089940001667;   // (3/(4 pi))^(1/3)

enum : int {
  XC_LDA_X = 1,
  XC_LDA_C_VWN = 7,
  XC_LDA_C_PW = 12,
  XC_LDA_C_PW_MOD = 13,
  XC_GGA_X_PBE = 101,
  XC_GGA_X_B88 = 106,
  XC_GGA_C_PBE = 130,
  XC_GGA_C_LYP = 131
};

OCCXC_HD inline int xc_family(int id) {
  switch (id) {
  case XC_LDA_X: case XC_LDA_C_VWN: case XC_LDA_C_PW: case XC_LDA_C_PW_MOD:
    return 0;
  case XC_GGA_X_PBE: case XC_GGA_X_B88: case XC_GGA_C_PBE: case XC_GGA_C_LYP:
    return 1;
  default:
    return -1;
  }
}

// ---- LDA exchange: e = -Cx rho^{4/3}
OCCXC_HD inline void lda_x(double rho, double &e, double &vr) {
  double r13 = cbrt(rho);
  e = -kCx * rho * r13;
  vr = -(4.0 / 3.0) * kCx * r13;
}

// ---- VWN interpolation eps(x), x = sqrt(rs); returns eps and d eps / d rs
OCCXC_HD inline void vwn_eps(double rs, double A, double b, double c,
                             double x0, double &eps, double &deps_drs) {
  double x = sqrt(rs);
  double X = x * x + b * x + c;
  double X0 = x0 * x0 + b * x0 + c;
  double Q = sqrt(4.0 * c - b * b);
  double tx = 2.0 * x + b;
  double at = atan(Q / tx);
  double bx0 = b * x0 / X0;
  eps = A * (log(x * x / X) + 2.0 * b / Q * at -
             bx0 * (log((x - x0) * (x - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
  double den = tx * tx + Q * Q;
  double deps_dx = A * (2.0 / x - tx / X - 4.0 * b / den -
                        bx0 * (2.0 / (x - x0) - tx / X - 4.0 * (b + 2.0 * x0) / den));
  deps_drs = deps_dx * 0.5 / x;
}

// ---- PW92 paramagnetic eps(rs) and d eps / d rs
OCCXC_HD inline void pw92_eps(double rs, double A, double &eps,
                              double &deps_drs) {
  const double a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  double srs = sqrt(rs);
  double Qd = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  double dQd = 2.0 * A * (0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2.0 * b4 * rs);
  double lg = log1p(1.0 / Qd);
  eps = -2.0 * A * (1.0 + a1 * rs) * lg;
  deps_drs = -2.0 * A * a1 * lg + 2.0 * A * (1.0 + a1 * rs) * dQd / (Qd * (Qd + 1.0));
}

OCCXC_HD inline void lda_c_from_eps(double rho, double rs, double eps,
                                    double deps_drs, double &e, double &vr) {
  e = rho * eps;
  vr = eps - rs / 3.0 * deps_drs;
}

// ---- PBE exchange
constexpr double kPbeBeta = 0.06672455060314922;
constexpr double kPbeKappa = 0.804;

OCCXC_HD inline void pbe_x(double rho, double sigma, double &e, double &vr,
                           double &vs) {
  const double mu = kPbeBeta * kPi * kPi / 3.0;
  // s^2 = c_s sigma rho^{-8/3},  c_s = 1 / (4 (3 pi^2)^{2/3})
  const double c_s = 0.026121172985233599567;
  double r13 = cbrt(rho);
  double r43 = rho * r13;
  double s2 = c_s * sigma / (r43 * r43);
  double den = 1.0 + mu * s2 / kPbeKappa;
  double F = 1.0 + kPbeKappa - kPbeKappa / den;
  double Fp = mu / (den * den); // dF/ds2
  e = -kCx * r43 * F;
  vr = -kCx * r13 * ((4.0 / 3.0) * F - (8.0 / 3.0) * s2 * Fp);
  vs = -kCx * c_s * Fp / r43;
}

// ---- B88 exchange, x = 2^{1/3} sqrt(sigma) rho^{-4/3}
OCCXC_HD inline void b88_x(double rho, double sigma, double &e, double &vr,
                           double &vs) {
  const double beta = 0.0042;
  const double c213 = 1.2599210498948731648;   // 2^{1/3}
  const double cm13 = 0.79370052598409973738;  // 2^{-1/3}
  const double cm23 = 0.62996052494743658238;  // 2^{-2/3}
  double r13 = cbrt(rho);
  double r43 = rho * r13;
  double x = c213 * sqrt(sigma) / r43;
  double ash = asinh(x);
  double D = 1.0 + 6.0 * beta * x * ash;
  double g = x * x / D;
  // g'(x)/x, finite at x = 0
  double gpx = (2.0 * D - 6.0 * beta * x * (ash + x / sqrt(1.0 + x * x))) / (D * D);
  double el, vl;
  lda_x(rho, el, vl);
  e = el - cm13 * beta * r43 * g;
  vr = vl - cm13 * beta * (4.0 / 3.0) * r13 * (g - x * x * gpx);
  vs = -cm23 * beta * gpx / r43;
}

// ---- PBE correlation (phi = 1)
OCCXC_HD inline void pbe_c(double rho, double sigma, double &e, double &vr,
                           double &vs) {
  const double gamma = 0.031090690869654895034;   // (1 - ln 2)/pi^2
  const double bg = kPbeBeta / gamma;
  // t^2 = c_t sigma rho^{-7/3},  c_t = pi / (16 (3 pi^2)^{1/3})
  const double c_t = 0.063468206097703704202;
  double r13 = cbrt(rho);
  double rs = kRsFac / r13;
  double ec, dec;
  pw92_eps(rs, 0.0310907, ec, dec);
  double t2 = c_t * sigma / (rho * rho * r13);
  double E = exp(-ec / gamma);
  double A = bg / (E - 1.0);
  double y = A * t2;
  double den = 1.0 + y + y * y;
  double R = (1.0 + y) / den;
  double Rp = -y * (2.0 + y) / (den * den);
  double Phi = bg * t2 * R;
  double H = gamma * log1p(Phi);
  double g1 = gamma / (1.0 + Phi);
  double dH_dt2 = g1 * bg * (R + y * Rp);
  double dH_dA = g1 * bg * t2 * t2 * Rp;
  double dA_de = A * A * E / kPbeBeta;
  e = rho * (ec + H);
  vr = ec + H - rs / 3.0 * dec * (1.0 + dH_dA * dA_de) - (7.0 / 3.0) * t2 * dH_dt2;
  vs = c_t * dH_dt2 / (rho * r13);
}

// ---- LYP correlation, closed shell
OCCXC_HD inline void lyp_c(double rho, double sigma, double &e, double &vr,
                           double &vs) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 2.8712340001881918;  // (3/10)(3 pi^2)^{2/3}
  double r13 = cbrt(rho);
  double u = 1.0 / r13;
  double D = 1.0 + d * u;
  double EoD = exp(-c * u) / D;
  double delta = c * u + d * u / D;
  double rm53 = u * u * u * u * u;      // rho^{-5/3}
  double rm83 = rm53 * u * u * u;       // rho^{-8/3}
  double e1 = -a * rho / D;
  double e2 = -a * b * CF * EoD * rho;
  double k3 = a * b / 72.0;
  double e3 = k3 * EoD * sigma * rm53 * (3.0 + 7.0 * delta);
  e = e1 + e2 + e3;
  double v1 = -(a / D) * (1.0 + d * u / (3.0 * D));
  double v2 = -a * b * CF * EoD * (1.0 + delta / 3.0);
  double v3 = k3 * sigma * EoD * rm83 *
              ((delta / 3.0) * (3.0 + 7.0 * delta) - (5.0 / 3.0) * (3.0 + 7.0 * delta) -
               (7.0 / 3.0) * u * (c + d / (D * D)));
  vr = v1 + v2 + v3;
  vs = k3 * EoD * rm53 * (3.0 + 7.0 * delta);
}

// One functional at one point: e, vrho, vsigma (vsigma = 0 for LDA).
// Returns false for an unknown id.
OCCXC_HD inline bool xc_point(int id, double rho, double sigma, double &e,
                              double &vr, double &vs) {
  vs = 0.0;
  switch (id) {
  case XC_LDA_X:
    lda_x(rho, e, vr);
    return true;
  case XC_LDA_C_VWN: {
    double rs = kRsFac / cbrt(rho), eps, deps;
    vwn_eps(rs, 0.0310907, 3.72744, 12.9352, -0.10498, eps, deps);
    lda_c_from_eps(rho, rs, eps, deps, e, vr);
    return true;
  }
  case XC_LDA_C_PW:
  case XC_LDA_C_PW_MOD: {
    double rs = kRsFac / cbrt(rho), eps, deps;
    pw92_eps(rs, id == XC_LDA_C_PW ? 0.031091 : 0.0310907, eps, deps);
    lda_c_from_eps(rho, rs, eps, deps, e, vr);
    return true;
  }
  case XC_GGA_X_PBE:
    pbe_x(rho, sigma, e, vr, vs);
    return true;
  case XC_GGA_X_B88:
    b88_x(rho, sigma, e, vr, vs);
    return true;
  case XC_GGA_C_PBE:
    pbe_c(rho, sigma, e, vr, vs);
    return true;
  case XC_GGA_C_LYP:
    lyp_c(rho, sigma, e, vr, vs);
    return true;
  default:
    e = vr = 0.0;
    return false;
  }
}

} // namespace occxc
