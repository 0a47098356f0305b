/*
 * occxc_oracle.c -- CPU FP64 restatement of OCC's Kohn-Sham XC build.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py; the product
 * library never links or calls it.
 *
 * What pins each part (the tests that do it are named in tests/):
 *  - shell normalisation: unit self-overlap of every contracted spherical AO
 *    (closed-form Gaussian integrals), tests/test_oracle_basis.py.
 *  - collocation + CCA transform + common_fac: golden vectors written by
 *    gau2grid (the reference's vendored C, compiled from its own sources into
 *    oracle/_ref) and, where that library is absent, the committed fixture
 *    tests/golden/gau2grid_collocation.npz; finite differences of phi vs grad.
 *  - decay cutoff: brute-force restatement on the CPU in the test.
 *  - hierarchy: the invariants of tests/dft_tests.cpp:1287-1462.
 *  - functionals: the only reference golden for LDA/GGA (gga_x_pbe exc at
 *    4 points, tests/dft_tests.cpp:109-111); uniform-gas limits; literature
 *    exchange energies of the H atom (exact density); vrho/vsigma come from
 *    complex-step differentiation of the energy density, checked against
 *    central differences.  libxc itself (6.2.2) is absent: VWN5, PW92, PBE-c,
 *    B88 and LYP energy expressions are pinned only by the papers' closed
 *    forms and limits ("parity unpinned" at the libxc boundary).
 *  - meta-GGA (TPSS exchange and correlation): the reference's golden exc,
 *    vrho, vsigma, vtau at four points (tests/dft_tests.cpp:138-190),
 *    complex-step vs finite differences, uniform-gas limits
 *    (tests/test_oracle_mgga.py).
 *  - VV10 (nonlocal_correlation.cpp): the reference's H2 known answer
 *    (tests/dft_tests.cpp:608-660, to 1.5e-5 relative on a converged grid),
 *    the symmetric double sum in numpy, and the forward + reverse point-order
 *    potentials against the derivative of the energy (tests/test_oracle_vv10.py).
 *  - XC build: screened == dense (tests/dft_tests.cpp:904-931 restated),
 *    N_elec = tr(2 D S) with analytic overlaps, brute-force tiny cases, and
 *    the directional-derivative identity dExc/dD = Vxc.
 */
#include "occxc_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int oracle_shell_size(int l, int pure) {
  return pure ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
}

int oracle_nbf(const oracle_basis *b) {
  int n = 0;
  for (int s = 0; s < b->nshell; s++) n += oracle_shell_size(b->l[s], b->pure[s]);
  return n;
}

/* ---------------------------------------------------------------- basis -- */

/* shell.cpp:20-23: gint(n, a) = Gamma((n+1)/2) / (2 a^((n+1)/2)) */
static double gint(int n, double alpha) {
  double n1_2 = 0.5 * (n + 1);
  return tgamma(n1_2) / (2.0 * pow(alpha, n1_2));
}

void oracle_normalize_shell(int l, int nprim, const double *exps,
                            const double *raw, double *out) {
  /* shell.cpp:180-186: c_p *= gto_norm(l, a_p), gto_norm = 1/sqrt(gint(2l+2, 2a)) */
  for (int p = 0; p < nprim; p++)
    out[p] = raw[p] / sqrt(gint(2 * l + 2, 2.0 * exps[p]));
  /* shell.cpp:52-64: c /= sqrt(c^T E c), E_ij = gint(2l+2, a_i + a_j) */
  double s = 0.0;
  for (int i = 0; i < nprim; i++)
    for (int j = 0; j < nprim; j++)
      s += out[i] * out[j] * gint(2 * l + 2, exps[i] + exps[j]);
  double f = 1.0 / sqrt(s);
  for (int p = 0; p < nprim; p++) out[p] *= f;
}

double oracle_common_fac(int l, int pure) {
  /* gto.cpp:7-16 */
  if (l == 0) return 0.282094791773878143;
  if (l == 1) return 0.488602511902919921;
  return pure ? sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) : 1.0;
}

static double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; i++) r *= i;
  return r;
}
static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  return fact(n) / (fact(k) * fact(n - k));
}

/* index of Cartesian (a,b,c) in CCA order: a from l down, then b from l-a down */
static int cart_index(int l, int a, int b) {
  int idx = 0;
  for (int i = l; i >= 0; i--)
    for (int j = l - i; j >= 0; j--) {
      if (i == a && j == b) return idx;
      idx++;
    }
  return -1;
}

/* Real solid harmonics S_lm as Cartesian polynomials (Helgaker, Jorgensen,
 * Olsen, "Molecular Electronic-Structure Theory", eqs. 6.4.47-6.4.50), rows in
 * CCA order m = -l..l, which is the order of gau2grid_transform.c. */
void oracle_cart_to_sph(int l, double *out) {
  int ncart = (l + 1) * (l + 2) / 2;
  memset(out, 0, sizeof(double) * (2 * l + 1) * ncart);
  for (int m = -l; m <= l; m++) {
    int am = abs(m);
    double norm = 1.0 / (pow(2.0, am) * fact(l)) *
                  sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
    /* v runs over v_m, v_m+1, ... with v_m = 0 (m>=0) or 1/2 (m<0); use 2v */
    int twov_m = (m < 0) ? 1 : 0;
    for (int t = 0; t <= (l - am) / 2; t++)
      for (int u = 0; u <= t; u++) {
        /* v from v_m to floor(|m|/2 - v_m) + v_m, in steps of 1 */
        int vmax2 = 2 * (int)floor(am / 2.0 - twov_m / 2.0) + twov_m;
        for (int twov = twov_m; twov <= vmax2; twov += 2) {
          double sgn = ((t + (twov - twov_m) / 2) % 2) ? -1.0 : 1.0;
          double c = sgn * pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                     binom(t, u) * binom(am, twov);
          int px = 2 * t + am - 2 * u - twov;
          int py = 2 * u + twov;
          int pz = l - 2 * t - am;
          int ci = cart_index(l, px, py);
          (void)pz;
          out[(m + l) * ncart + ci] += norm * c;
        }
      }
  }
}

static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

#define ORACLE_MAX_L 6

void oracle_eval_shell(int l, int pure, int nprim, const double *exps,
                       const double *coefs, const double *center, long npts,
                       const double *xyz, int deriv, double *phi, double *phx,
                       double *phy, double *phz, long ld) {
  int ncart = (l + 1) * (l + 2) / 2;
  int nout = oracle_shell_size(l, pure);
  double T[(2 * ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2];
  if (pure) oracle_cart_to_sph(l, T);
  double fac = oracle_common_fac(l, pure);
  double cv[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2];
  double cx[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2];
  double cy[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2];
  double cz[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2];
  for (long p = 0; p < npts; p++) {
    /* gau2grid: xc = x - center_x etc., R2, S0 = sum c e^{-a R2},
     * S1 = sum -2 a c e^{-a R2} (gau2grid_deriv1.c:491-505) */
    double x = xyz[3 * p] - center[0];
    double y = xyz[3 * p + 1] - center[1];
    double z = xyz[3 * p + 2] - center[2];
    double r2 = x * x + y * y + z * z;
    double S0 = 0.0, S1 = 0.0;
    for (int k = 0; k < nprim; k++) {
      double e = exp(-exps[k] * r2);
      S0 += coefs[k] * e;
      S1 += -2.0 * exps[k] * coefs[k] * e;
    }
    int idx = 0;
    for (int a = l; a >= 0; a--)
      for (int b = l - a; b >= 0; b--) {
        int c = l - a - b;
        double poly = ipow(x, a) * ipow(y, b) * ipow(z, c);
        cv[idx] = poly * S0;
        if (deriv) {
          double dpx = a ? a * ipow(x, a - 1) * ipow(y, b) * ipow(z, c) : 0.0;
          double dpy = b ? b * ipow(x, a) * ipow(y, b - 1) * ipow(z, c) : 0.0;
          double dpz = c ? c * ipow(x, a) * ipow(y, b) * ipow(z, c - 1) : 0.0;
          cx[idx] = dpx * S0 + poly * x * S1;
          cy[idx] = dpy * S0 + poly * y * S1;
          cz[idx] = dpz * S0 + poly * z * S1;
        }
        idx++;
      }
    for (int o = 0; o < nout; o++) {
      double v = 0, vx = 0, vy = 0, vz = 0;
      if (pure) {
        for (int k = 0; k < ncart; k++) {
          double t = T[o * ncart + k];
          if (t == 0.0) continue;
          v += t * cv[k];
          if (deriv) {
            vx += t * cx[k];
            vy += t * cy[k];
            vz += t * cz[k];
          }
        }
      } else {
        v = cv[o];
        if (deriv) {
          vx = cx[o];
          vy = cy[o];
          vz = cz[o];
        }
      }
      phi[o * ld + p] = fac * v;
      if (deriv) {
        phx[o * ld + p] = fac * vx;
        phy[o * ld + p] = fac * vy;
        phz[o * ld + p] = fac * vz;
      }
    }
  }
}

/* Value, gradient and Hessian of one shell times common_fac: gto.cpp:245-285
 * (evaluate_basis_subset, max_derivative 2) with gau2grid
 * gg_collocation_deriv2.  With S0 = sum c e^{-a r2}, S1 = sum -2a c e^{-a r2},
 * S2 = sum 4a^2 c e^{-a r2} and a Cartesian monomial P:
 *   d_i f    = (d_i P) S0 + P x_i S1
 *   d_ij f   = (d_ij P) S0 + ((d_i P) x_j + (d_j P) x_i + delta_ij P) S1
 *              + P x_i x_j S2.
 * out[(q*nout + o)*ld + p], q = 0 value, 1..3 x,y,z, 4..9 xx,xy,xz,yy,yz,zz. */
void oracle_eval_shell2(int l, int pure, int nprim, const double *exps,
                        const double *coefs, const double *center, long npts,
                        const double *xyz, double *out, long ld) {
  enum { NCM = (ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2 };
  int ncart = (l + 1) * (l + 2) / 2;
  int nout = oracle_shell_size(l, pure);
  double T[(2 * ORACLE_MAX_L + 1) * NCM];
  if (pure) oracle_cart_to_sph(l, T);
  double fac = oracle_common_fac(l, pure);
  double cc[10][NCM];
  for (long p = 0; p < npts; p++) {
    double d[3] = {xyz[3 * p] - center[0], xyz[3 * p + 1] - center[1],
                   xyz[3 * p + 2] - center[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    for (int k = 0; k < nprim; k++) {
      double e = coefs[k] * exp(-exps[k] * r2);
      S0 += e;
      S1 += -2.0 * exps[k] * e;
      S2 += 4.0 * exps[k] * exps[k] * e;
    }
    int idx = 0;
    for (int a = l; a >= 0; a--)
      for (int b = l - a; b >= 0; b--) {
        int c = l - a - b;
        int pw[3] = {a, b, c};
        double P = ipow(d[0], a) * ipow(d[1], b) * ipow(d[2], c);
        double dP[3], ddP[3][3];
        for (int i = 0; i < 3; i++) {
          int q[3] = {a, b, c};
          dP[i] = pw[i] ? pw[i] * ipow(d[i], pw[i] - 1) : 0.0;
          for (int j = 0; j < 3; j++)
            if (j != i) dP[i] *= ipow(d[j], q[j]);
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            int q[3] = {a, b, c};
            double coef = 1.0;
            if (q[i] == 0) { ddP[i][j] = 0.0; continue; }
            coef *= q[i];
            q[i] -= 1;
            if (q[j] == 0) { ddP[i][j] = 0.0; continue; }
            coef *= q[j];
            q[j] -= 1;
            ddP[i][j] = coef * ipow(d[0], q[0]) * ipow(d[1], q[1]) * ipow(d[2], q[2]);
          }
        cc[0][idx] = P * S0;
        for (int i = 0; i < 3; i++) cc[1 + i][idx] = dP[i] * S0 + P * d[i] * S1;
        static const int hi[6] = {0, 0, 0, 1, 1, 2}, hj[6] = {0, 1, 2, 1, 2, 2};
        for (int h = 0; h < 6; h++) {
          int i = hi[h], j = hj[h];
          cc[4 + h][idx] = ddP[i][j] * S0 +
                           (dP[i] * d[j] + dP[j] * d[i] + (i == j ? P : 0.0)) * S1 +
                           P * d[i] * d[j] * S2;
        }
        idx++;
      }
    for (int o = 0; o < nout; o++)
      for (int q = 0; q < 10; q++) {
        double v = 0.0;
        if (pure) {
          for (int k = 0; k < ncart; k++) v += T[o * ncart + k] * cc[q][k];
        } else {
          v = cc[q][o];
        }
        out[((long)q * nout + o) * ld + p] = fac * v;
      }
  }
}

void oracle_decay_cutoff(const oracle_basis *b, double threshold,
                         double *extents) {
  /* gto.cpp:18-67: r = 0.5 .. 30 Bohr in 0.1 steps along +x from the shell
   * origin; the extent is the first r where max_c |fac*phi_c| < threshold. */
  const double max_radius = 30.0, increment = 0.1, min_radius = 0.5;
  const int npts = (int)((max_radius - min_radius) / increment) + 1;
  /* Eigen::VectorXd::LinSpaced(npts, lo, hi): lo + i*step, last point = hi */
  double step = (max_radius - min_radius) / (double)(npts - 1);
  double *xyz = malloc(sizeof(double) * 3 * npts);
  double *vals = malloc(sizeof(double) * npts * 64);
  for (int s = 0; s < b->nshell; s++) {
    const double *o = b->origin + 3 * s;
    for (int i = 0; i < npts; i++) {
      double r = (i == npts - 1) ? max_radius : min_radius + (double)i * step;
      xyz[3 * i] = 1.0 * r + o[0];
      xyz[3 * i + 1] = 0.0 + o[1];
      xyz[3 * i + 2] = 0.0 + o[2];
    }
    int l = b->l[s], pure = b->pure[s];
    int nc = oracle_shell_size(l, pure);
    oracle_eval_shell(l, pure, b->nprim[s], b->exps + b->prim_offset[s],
                      b->coefs + b->prim_offset[s], o, npts, xyz, 0, vals, 0, 0,
                      0, npts);
    double result = 0.0;
    for (int i = 0; i < npts; i++) {
      result = i * increment + min_radius;
      double v = 0.0;
      for (int c = 0; c < nc; c++) {
        double a = fabs(vals[c * npts + i]);
        if (a > v) v = a;
      }
      if (v < threshold) break;
    }
    extents[s] = result;
  }
  free(xyz);
  free(vals);
}

/* ------------------------------------------------------------ hierarchy -- */

static uint64_t dilate_3d(uint64_t x) {
  /* spread the low 21 bits of x to every third bit */
  uint64_t r = 0;
  for (int i = 0; i < 21; i++) r |= ((x >> i) & 1ULL) << (3 * i);
  return r;
}

typedef struct {
  uint64_t code;
  long idx;
} morton_pair;

static int cmp_morton(const void *a, const void *b) {
  const morton_pair *x = a, *y = b;
  if (x->code < y->code) return -1;
  if (x->code > y->code) return 1;
  return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

long oracle_build_hierarchy(long n, const double *xyz, const double *w,
                            double *sxyz, double *sw, long *perm,
                            long *leaf_offset, long *leaf_count,
                            double *leaf_center, double *leaf_radius) {
  if (n <= 0) return 0;
  /* spatial_grid_hierarchy.cpp:43-98 */
  double mn[3], mx[3], ext[3];
  for (int k = 0; k < 3; k++) mn[k] = mx[k] = xyz[k];
  for (long i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double v = xyz[3 * i + k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  for (int k = 0; k < 3; k++) {
    ext[k] = mx[k] - mn[k];
    if (ext[k] < 1e-10) ext[k] = 1.0;
  }
  morton_pair *mp = malloc(sizeof(morton_pair) * n);
  const uint64_t scale = (1ULL << 21) - 1;
  for (long i = 0; i < n; i++) {
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
      double t = (xyz[3 * i + k] - mn[k]) / ext[k];
      if (t < 0.0) t = 0.0;
      if (t > 1.0) t = 1.0;
      q[k] = (uint64_t)(t * (double)scale);
    }
    mp[i].code = dilate_3d(q[0]) | (dilate_3d(q[1]) << 1) | (dilate_3d(q[2]) << 2);
    mp[i].idx = i;
  }
  qsort(mp, n, sizeof(morton_pair), cmp_morton);
  for (long i = 0; i < n; i++) {
    long o = mp[i].idx;
    perm[i] = o;
    for (int k = 0; k < 3; k++) sxyz[3 * i + k] = xyz[3 * o + k];
    sw[i] = w[o];
  }
  free(mp);
  /* spatial_grid_hierarchy.cpp:100-147; settings .h:26-30 */
  const long target = 128, minl = 32, maxl = 256;
  long nl = 0, off = 0;
  while (off < n) {
    long rem = n - off, cnt;
    if (rem <= maxl) cnt = rem;
    else if (rem < 2 * minl) cnt = rem / 2;
    else cnt = (target < rem) ? target : rem;
    double c[3] = {0, 0, 0};
    for (long i = 0; i < cnt; i++)
      for (int k = 0; k < 3; k++) c[k] += sxyz[3 * (off + i) + k];
    for (int k = 0; k < 3; k++) c[k] /= (double)cnt;
    double r = 0.0;
    for (long i = 0; i < cnt; i++) {
      double dx = sxyz[3 * (off + i)] - c[0], dy = sxyz[3 * (off + i) + 1] - c[1],
             dz = sxyz[3 * (off + i) + 2] - c[2];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (d > r) r = d;
    }
    leaf_offset[nl] = off;
    leaf_count[nl] = cnt;
    for (int k = 0; k < 3; k++) leaf_center[3 * nl + k] = c[k];
    leaf_radius[nl] = r;
    nl++;
    off += cnt;
  }
  return nl;
}

/* ---------------------------------------------------------- functionals -- */
/* Energy densities e(rho, sigma) = rho * eps written in complex arithmetic so
 * that vrho = de/drho and vsigma = de/dsigma follow by complex-step
 * differentiation (Squire & Trapp, SIAM Rev. 40, 110 (1998)): exact to
 * rounding, and independent of any hand-derived derivative algebra. */

typedef double complex cplx;

/* threshold below which a point contributes nothing (libxc's own
 * per-functional thresholds are not visible from OCC; this is our choice) */
#define ORACLE_DENS_THR 1e-15
#define ORACLE_SIGMA_MIN 1e-40

static const double CX = 0.73855876638202240587; /* (3/4)(3/pi)^(1/3) */

static cplx c_cbrt(cplx z) { return cpow(z, 1.0 / 3.0); }
/* log(1 + u) = 2 atanh(u / (2 + u)): analytic, and accurate for small u */
static cplx c_log1p(cplx u) { return 2.0 * catanh(u / (2.0 + u)); }
static cplx c_rs(cplx rho) { return c_cbrt(3.0 / (4.0 * M_PI * rho)); }

/* Dirac/Slater exchange */
static cplx e_lda_x(cplx rho) { return -CX * rho * c_cbrt(rho); }

/* VWN functional V interpolation, S.H. Vosko, L. Wilk, M. Nusair,
 * Can. J. Phys. 58, 1200 (1980), eq. 4.4, paramagnetic MC parameters */
static cplx eps_vwn(cplx rs, double A, double b, double c, double x0) {
  cplx x = csqrt(rs);
  cplx X = x * x + b * x + c;
  double X0 = x0 * x0 + b * x0 + c;
  double Q = sqrt(4.0 * c - b * b);
  cplx at = catan(Q / (2.0 * x + b));
  return A * (clog(x * x / X) + 2.0 * b / Q * at -
              b * x0 / X0 * (clog((x - x0) * (x - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
}

/* J.P. Perdew, Y. Wang, PRB 45, 13244 (1992), eq. 10, paramagnetic */
static cplx eps_pw92(cplx rs, double A) {
  const double a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  cplx srs = csqrt(rs);
  cplx den = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * c_log1p(1.0 / den);
}

/* PBE: J.P. Perdew, K. Burke, M. Ernzerhof, PRL 77, 3865 (1996) */
static const double PBE_BETA = 0.06672455060314922;
static const double PBE_KAPPA = 0.804;

static cplx e_pbe_x(cplx rho, cplx sigma) {
  double mu = PBE_BETA * M_PI * M_PI / 3.0;
  cplx kf = c_cbrt(3.0 * M_PI * M_PI * rho);
  cplx s2 = sigma / (4.0 * kf * kf * rho * rho);
  cplx F = 1.0 + PBE_KAPPA - PBE_KAPPA / (1.0 + mu * s2 / PBE_KAPPA);
  return e_lda_x(rho) * F;
}

static cplx e_pbe_c(cplx rho, cplx sigma) {
  double gamma = (1.0 - log(2.0)) / (M_PI * M_PI);
  cplx rs = c_rs(rho);
  cplx ec = eps_pw92(rs, 0.0310907);
  cplx kf = c_cbrt(3.0 * M_PI * M_PI * rho);
  cplx ks = csqrt(4.0 * kf / M_PI);
  cplx t2 = sigma / (4.0 * ks * ks * rho * rho);
  cplx A = (PBE_BETA / gamma) / (cexp(-ec / gamma) - 1.0);
  cplx At2 = A * t2;
  cplx H = gamma * c_log1p((PBE_BETA / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

/* A.D. Becke, PRA 38, 3098 (1988), eq. 8, written per spin channel with
 * rho_s = rho/2 and |grad rho_s|^2 = sigma/4 */
static cplx e_b88(cplx rho, cplx sigma) {
  const double beta = 0.0042;
  cplx rs_ = 0.5 * rho;
  cplx r43 = rs_ * c_cbrt(rs_);
  cplx x = csqrt(0.25 * sigma) / r43;
  cplx g = beta * r43 * x * x / (1.0 + 6.0 * beta * x * casinh(x));
  return e_lda_x(rho) - 2.0 * g;
}

/* C. Lee, W. Yang, R.G. Parr, PRB 37, 785 (1988) in the form of
 * B. Miehlich, A. Savin, H. Stoll, H. Preuss, CPL 157, 200 (1989), eq. 2,
 * evaluated at rho_a = rho_b = rho/2, grad rho_a = grad rho_b = grad rho / 2 */
static cplx e_lyp(cplx rho, cplx sigma) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  cplx ra = 0.5 * rho, rb = 0.5 * rho;
  cplx gaa = 0.25 * sigma, gbb = 0.25 * sigma, gg = sigma;
  cplx rm13 = 1.0 / c_cbrt(rho);
  cplx den = 1.0 + d * rm13;
  cplx omega = cexp(-c * rm13) / den * cpow(rho, -11.0 / 3.0);
  cplx delta = c * rm13 + d * rm13 / den;
  cplx t1 = -a * 4.0 / den * ra * rb / rho;
  cplx br = ra * rb * (cpow(2.0, 11.0 / 3.0) * CF * (cpow(ra, 8.0 / 3.0) + cpow(rb, 8.0 / 3.0)) +
                       (47.0 / 18.0 - 7.0 / 18.0 * delta) * gg -
                       (2.5 - delta / 18.0) * (gaa + gbb) -
                       (delta - 11.0) / 9.0 * (ra / rho * gaa + rb / rho * gbb));
  br += -2.0 / 3.0 * rho * rho * gg + (2.0 / 3.0 * rho * rho - ra * ra) * gbb +
        (2.0 / 3.0 * rho * rho - rb * rb) * gaa;
  return t1 - a * b * omega * br;
}

int oracle_functional_family(int id) {
  switch (id) {
  case ORACLE_LDA_X:
  case ORACLE_LDA_C_VWN:
  case ORACLE_LDA_C_PW:
  case ORACLE_LDA_C_PW_MOD:
    return 0;
  case ORACLE_GGA_X_PBE:
  case ORACLE_GGA_X_B88:
  case ORACLE_GGA_C_PBE:
  case ORACLE_GGA_C_LYP:
    return 1;
  case ORACLE_MGGA_X_TPSS:
  case ORACLE_MGGA_C_TPSS:
    return 2;
  default:
    return -1;
  }
}

static cplx edens(int id, cplx rho, cplx sigma) {
  switch (id) {
  case ORACLE_LDA_X:
    return e_lda_x(rho);
  case ORACLE_LDA_C_VWN:
    return rho * eps_vwn(c_rs(rho), 0.0310907, 3.72744, 12.9352, -0.10498);
  case ORACLE_LDA_C_PW:
    return rho * eps_pw92(c_rs(rho), 0.031091);
  case ORACLE_LDA_C_PW_MOD:
    return rho * eps_pw92(c_rs(rho), 0.0310907);
  case ORACLE_GGA_X_PBE:
    return e_pbe_x(rho, sigma);
  case ORACLE_GGA_X_B88:
    return e_b88(rho, sigma);
  case ORACLE_GGA_C_PBE:
    return e_pbe_c(rho, sigma);
  case ORACLE_GGA_C_LYP:
    return e_lyp(rho, sigma);
  }
  return 0.0;
}

double oracle_functional_energy_density(int id, double rho, double sigma) {
  if (rho < ORACLE_DENS_THR) return 0.0;
  if (sigma < ORACLE_SIGMA_MIN) sigma = ORACLE_SIGMA_MIN;
  return creal(edens(id, rho, sigma));
}

int oracle_functional_eval(int id, long n, const double *rho,
                           const double *sigma, double *exc, double *vrho,
                           double *vsigma) {
  int fam = oracle_functional_family(id);
  if (fam < 0 || fam > 1) return -1;
  for (long i = 0; i < n; i++) {
    double r = rho[i];
    if (!(r >= ORACLE_DENS_THR)) continue;
    double s = 0.0;
    if (fam == 1) {
      s = sigma[i];
      if (s < ORACLE_SIGMA_MIN) s = ORACLE_SIGMA_MIN;
    }
    double hr = 1e-20 * r;
    cplx v = edens(id, r + I * hr, s);
    exc[i] += creal(v) / r;
    vrho[i] += cimag(v) / hr;
    if (fam == 1) {
      double hs = 1e-20 * s;
      cplx u = edens(id, r, s + I * hs);
      vsigma[i] += cimag(u) / hs;
    }
  }
  return 0;
}

/* ------------------------------------------------ polarised functionals -- */
/* Spin-polarised energy densities e(rho_a, rho_b, s_aa, s_ab, s_bb), the
 * inputs libxc's polarised functionals take (functional.h:399-501 with
 * SpinorbitalKind::Unrestricted), again in complex arithmetic for
 * complex-step derivatives.
 *  - exchange (lda_x, gga_x_pbe, gga_x_b88): exact spin scaling,
 *    E_x[ra, rb] = E_x[2 ra]/2 + E_x[2 rb]/2 (Oliver, Perdew, PRA 20, 397
 *    (1979)), applied to the closed-shell forms above;
 *  - VWN5: Vosko, Wilk, Nusair eq. 4.4 with the paramagnetic, ferromagnetic
 *    and spin-stiffness fits and the interpolation
 *    eps = eP + ac f(z)/f''(0) (1 - z^4) + (eF - eP) f(z) z^4;
 *  - PW92 (and its "mod" constants): Perdew, Wang eqs. 8-10, Table I;
 *  - PBE correlation: PRL 77, 3865 with phi(z) = ((1+z)^(2/3) + (1-z)^(2/3))/2;
 *  - LYP: Miehlich et al. eq. 2 in its general (ra, rb, grad ra, grad rb)
 *    form (e_lyp above is its closed-shell case).
 * Powers of (1 +- z) below DBL_EPSILON are held at DBL_EPSILON^n (libxc's
 * zeta threshold). */

#define ORACLE_ZETA_THR 2.220446049250313e-16

static cplx opz_pow(cplx x, double n) {
  if (creal(x) <= ORACLE_ZETA_THR) return pow(ORACLE_ZETA_THR, n);
  return cpow(x, n);
}
static cplx f_zeta(cplx z) {
  return (opz_pow(1.0 + z, 4.0 / 3.0) + opz_pow(1.0 - z, 4.0 / 3.0) - 2.0) /
         (pow(2.0, 4.0 / 3.0) - 2.0);
}
/* PW92 eq. 10: G(rs; A, a1, b1..b4), p = 1 */
static cplx pw92_G(cplx rs, double A, double a1, double b1, double b2, double b3, double b4) {
  cplx srs = csqrt(rs);
  cplx den = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * c_log1p(1.0 / den);
}
/* eps_c(rs, z) of PW92; mod = 1 uses libxc's "mod" constants */
static cplx eps_pw92_pol(cplx rs, cplx z, int mod) {
  const double A0 = mod ? 0.0310907 : 0.031091, A1 = mod ? 0.01554535 : 0.015545,
               Aa = mod ? 0.0168869 : 0.016887;
  const double fpp = mod ? 1.709920934161365617563962776245 : 1.709921;
  cplx e0 = pw92_G(rs, A0, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  cplx e1 = pw92_G(rs, A1, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  cplx mac = pw92_G(rs, Aa, 0.11125, 10.357, 3.6231, 0.88026, 0.49671); /* -alpha_c */
  cplx f = f_zeta(z), z4 = z * z * z * z;
  return e0 - mac * f / fpp * (1.0 - z4) + (e1 - e0) * f * z4;
}
static cplx eps_vwn_pol(cplx rs, cplx z) {
  const double fpp = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
  cplx eP = eps_vwn(rs, 0.0310907, 3.72744, 12.9352, -0.10498);
  cplx eF = eps_vwn(rs, 0.01554535, 7.06042, 18.0578, -0.32500);
  cplx ac = eps_vwn(rs, -1.0 / (6.0 * M_PI * M_PI), 1.13107, 13.0045, -0.0047584);
  cplx f = f_zeta(z), z4 = z * z * z * z;
  return eP + ac * f / fpp * (1.0 - z4) + (eF - eP) * f * z4;
}
static cplx e_pbe_c_pol(cplx ra, cplx rb, cplx sigma) {
  double gamma = (1.0 - log(2.0)) / (M_PI * M_PI);
  cplx rho = ra + rb, z = (ra - rb) / rho;
  cplx rs = c_rs(rho);
  cplx ec = eps_pw92_pol(rs, z, 1);
  cplx phi = 0.5 * (opz_pow(1.0 + z, 2.0 / 3.0) + opz_pow(1.0 - z, 2.0 / 3.0));
  cplx phi3 = phi * phi * phi;
  cplx kf = c_cbrt(3.0 * M_PI * M_PI * rho);
  cplx ks = csqrt(4.0 * kf / M_PI);
  cplx t2 = sigma / (4.0 * phi * phi * ks * ks * rho * rho);
  cplx A = (PBE_BETA / gamma) / (cexp(-ec / (gamma * phi3)) - 1.0);
  cplx At2 = A * t2;
  cplx H = gamma * phi3 *
           c_log1p((PBE_BETA / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}
static cplx e_lyp_pol(cplx ra, cplx rb, cplx saa, cplx sab, cplx sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  cplx rho = ra + rb, gg = saa + 2.0 * sab + sbb;
  cplx rm13 = 1.0 / c_cbrt(rho);
  cplx den = 1.0 + d * rm13;
  cplx omega = cexp(-c * rm13) / den * cpow(rho, -11.0 / 3.0);
  cplx delta = c * rm13 + d * rm13 / den;
  cplx t1 = -a * 4.0 / den * ra * rb / rho;
  cplx br = ra * rb * (pow(2.0, 11.0 / 3.0) * CF * (cpow(ra, 8.0 / 3.0) + cpow(rb, 8.0 / 3.0)) +
                       (47.0 / 18.0 - 7.0 / 18.0 * delta) * gg -
                       (2.5 - delta / 18.0) * (saa + sbb) -
                       (delta - 11.0) / 9.0 * (ra / rho * saa + rb / rho * sbb));
  br += -2.0 / 3.0 * rho * rho * gg + (2.0 / 3.0 * rho * rho - ra * ra) * sbb +
        (2.0 / 3.0 * rho * rho - rb * rb) * saa;
  return t1 - a * b * omega * br;
}
/* exchange by spin scaling; a channel below the density threshold adds 0 */
static cplx e_x_spin(int id, cplx rs_, cplx ss) {
  if (creal(rs_) < ORACLE_DENS_THR) return 0.0;
  switch (id) {
  case ORACLE_LDA_X:
    return 0.5 * e_lda_x(2.0 * rs_);
  case ORACLE_GGA_X_PBE:
    return 0.5 * e_pbe_x(2.0 * rs_, 4.0 * ss);
  case ORACLE_GGA_X_B88:
    return 0.5 * e_b88(2.0 * rs_, 4.0 * ss);
  }
  return 0.0;
}
static cplx edens_pol(int id, cplx ra, cplx rb, cplx saa, cplx sab, cplx sbb) {
  cplx rho = ra + rb;
  switch (id) {
  case ORACLE_LDA_X:
  case ORACLE_GGA_X_PBE:
  case ORACLE_GGA_X_B88:
    return e_x_spin(id, ra, saa) + e_x_spin(id, rb, sbb);
  case ORACLE_LDA_C_VWN:
    return rho * eps_vwn_pol(c_rs(rho), (ra - rb) / rho);
  case ORACLE_LDA_C_PW:
    return rho * eps_pw92_pol(c_rs(rho), (ra - rb) / rho, 0);
  case ORACLE_LDA_C_PW_MOD:
    return rho * eps_pw92_pol(c_rs(rho), (ra - rb) / rho, 1);
  case ORACLE_GGA_C_PBE:
    return e_pbe_c_pol(ra, rb, saa + 2.0 * sab + sbb);
  case ORACLE_GGA_C_LYP:
    return e_lyp_pol(ra, rb, saa, sab, sbb);
  }
  return 0.0;
}

double oracle_functional_energy_density_pol(int id, double ra, double rb, double saa,
                                            double sab, double sbb) {
  if (ra + rb < ORACLE_DENS_THR) return 0.0;
  if (saa < ORACLE_SIGMA_MIN) saa = ORACLE_SIGMA_MIN;
  if (sbb < ORACLE_SIGMA_MIN) sbb = ORACLE_SIGMA_MIN;
  return creal(edens_pol(id, ra, rb, saa, sab, sbb));
}

int oracle_functional_eval_pol(int id, long n, const double *rho, const double *sigma,
                               double *exc, double *vrho, double *vsigma) {
  int fam = oracle_functional_family(id);
  if (fam < 0 || fam > 1) return -1;
  for (long i = 0; i < n; i++) {
    double x[5] = {rho[2 * i], rho[2 * i + 1], 0.0, 0.0, 0.0};
    double r = x[0] + x[1];
    if (!(r >= ORACLE_DENS_THR)) continue;
    if (fam == 1) {
      x[2] = sigma[3 * i] < ORACLE_SIGMA_MIN ? ORACLE_SIGMA_MIN : sigma[3 * i];
      x[3] = sigma[3 * i + 1];
      x[4] = sigma[3 * i + 2] < ORACLE_SIGMA_MIN ? ORACLE_SIGMA_MIN : sigma[3 * i + 2];
    }
    exc[i] += creal(edens_pol(id, x[0], x[1], x[2], x[3], x[4])) / r;
    const int nv = fam == 1 ? 5 : 2;
    for (int k = 0; k < nv; k++) {
      double h = 1e-20 * (fabs(x[k]) > 1e-300 ? fabs(x[k]) : 1.0);
      cplx y[5] = {x[0], x[1], x[2], x[3], x[4]};
      y[k] += I * h;
      double dv = cimag(edens_pol(id, y[0], y[1], y[2], y[3], y[4])) / h;
      if (k < 2) vrho[2 * i + k] += dv;
      else vsigma[3 * i + k - 2] += dv;
    }
  }
  return 0;
}

/* ------------------------------------------------ meta-GGA functionals -- */
/* TPSS: J. Tao, J.P. Perdew, V.N. Staroverov, G.E. Scuseria, PRL 91, 146401
 * (2003).  Unpolarised energy densities e(rho, sigma, tau) with total
 * quantities, tau = (1/2) sum_i |grad psi_i|^2 (density.cpp:98-100), in
 * complex arithmetic for complex-step derivatives.  Inputs are limited as
 * libxc limits meta-GGA inputs: tau >= 1e-20 and sigma <= 8 rho tau (the
 * Fermi-hole curvature bound tau_W <= tau); derivatives are taken at the
 * limited inputs. */

#define ORACLE_TAU_MIN 1e-20

/* exchange, eqs. 5-10: F_x(p, z) with p = s^2, z = tau_W / tau,
 * alpha = (tau - tau_W) / tau_unif */
static cplx e_tpss_x(cplx rho, cplx sigma, cplx tau) {
  const double b = 0.40, c = 1.59096, e = 1.537, kappa = 0.804, mu = 0.21951;
  const double mge = 10.0 / 81.0;
  const double k2 = pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  cplx p = sigma / (4.0 * k2 * cpow(rho, 8.0 / 3.0));
  cplx z = sigma / (8.0 * rho * tau);
  cplx alpha = (tau - sigma / (8.0 * rho)) / (0.3 * k2 * cpow(rho, 5.0 / 3.0));
  cplx qb = 0.45 * (alpha - 1.0) / csqrt(1.0 + b * alpha * (alpha - 1.0)) + 2.0 * p / 3.0;
  cplx z2 = z * z;
  cplx num = (mge + c * z2 / ((1.0 + z2) * (1.0 + z2))) * p + 146.0 / 2025.0 * qb * qb -
             73.0 / 405.0 * qb * csqrt(0.5 * (9.0 / 25.0 * z2 + p * p)) +
             mge * mge / kappa * p * p + 2.0 * sqrt(e) * mge * 9.0 / 25.0 * z2 +
             e * mu * p * p * p;
  cplx den = (1.0 + sqrt(e) * p) * (1.0 + sqrt(e) * p);
  cplx x = num / den;
  cplx F = 1.0 + kappa - kappa / (1.0 + x / kappa);
  return e_lda_x(rho) * F;
}

/* correlation, eqs. 11-14 at zeta = 0 (so xi = 0 and C(0, 0) = 0.53):
 * eps_revPKZB = eps_PBE [1 + C z^2] - (1 + C) z^2 max(eps_PBE(n/2, 0), eps_PBE),
 * eps = eps_revPKZB [1 + d eps_revPKZB z^3], z = min(tau_W / tau, 1) */
static cplx e_tpss_c(cplx rho, cplx sigma, cplx tau) {
  const double d = 2.8, C0 = 0.53;
  cplx epbe = e_pbe_c_pol(0.5 * rho, 0.5 * rho, sigma) / rho;
  cplx epol = e_pbe_c_pol(0.5 * rho, 0.0, 0.25 * sigma) / (0.5 * rho);
  cplx et = creal(epol) > creal(epbe) ? epol : epbe;
  cplx z = sigma / (8.0 * rho * tau);
  if (creal(z) > 1.0) z = 1.0;
  cplx z2 = z * z;
  cplx f0 = (1.0 + C0 * z2) * epbe - (1.0 + C0) * z2 * et;
  return rho * f0 * (1.0 + d * f0 * z2 * z);
}

static cplx edens_mgga(int id, cplx rho, cplx sigma, cplx tau) {
  switch (id) {
  case ORACLE_MGGA_X_TPSS:
    return e_tpss_x(rho, sigma, tau);
  case ORACLE_MGGA_C_TPSS:
    return e_tpss_c(rho, sigma, tau);
  }
  return edens(id, rho, sigma);
}

/* input limits of the meta-GGA evaluation (see above) */
static void mgga_inputs(double rho, double *sigma, double *tau) {
  if (*sigma < ORACLE_SIGMA_MIN) *sigma = ORACLE_SIGMA_MIN;
  if (*tau < ORACLE_TAU_MIN) *tau = ORACLE_TAU_MIN;
  if (*sigma > 8.0 * rho * *tau) *sigma = 8.0 * rho * *tau;
}

double oracle_functional_energy_density_mgga(int id, double rho, double sigma, double tau) {
  if (rho < ORACLE_DENS_THR) return 0.0;
  mgga_inputs(rho, &sigma, &tau);
  return creal(edens_mgga(id, rho, sigma, tau));
}

int oracle_functional_eval_mgga(int id, long n, const double *rho, const double *sigma,
                                const double *tau, double *exc, double *vrho,
                                double *vsigma, double *vtau) {
  int fam = oracle_functional_family(id);
  if (fam < 0) return -1;
  for (long i = 0; i < n; i++) {
    double r = rho[i];
    if (!(r >= ORACLE_DENS_THR)) continue;
    double s = fam >= 1 ? sigma[i] : 0.0, t = fam == 2 ? tau[i] : 1.0;
    if (fam == 2) mgga_inputs(r, &s, &t);
    else if (fam == 1 && s < ORACLE_SIGMA_MIN) s = ORACLE_SIGMA_MIN;
    double hr = 1e-20 * r;
    exc[i] += creal(edens_mgga(id, r, s, t)) / r;
    vrho[i] += cimag(edens_mgga(id, r + I * hr, s, t)) / hr;
    if (fam >= 1) {
      double hs = 1e-20 * s;
      vsigma[i] += cimag(edens_mgga(id, r, s + I * hs, t)) / hs;
    }
    if (fam == 2) {
      double ht = 1e-20 * t;
      vtau[i] += cimag(edens_mgga(id, r, s, t + I * ht)) / ht;
    }
  }
  return 0;
}

/* Spin-polarised TPSS (the same paper): exchange by exact spin scaling,
 * E_x[na, nb] = E_x[2 na]/2 + E_x[2 nb]/2 with (2 n_s, 4 s_ss, 2 t_s); the
 * correlation's eqs. 11-14 in full, with zeta = (na - nb)/n,
 * xi = |grad zeta| / (2 (3 pi^2 n)^(1/3)), |grad zeta|^2 = 4 (nb^2 s_aa
 * - 2 na nb s_ab + na^2 s_bb) / n^4, C(zeta, xi) = (0.53 + 0.87 zeta^2
 * + 0.50 zeta^4 + 2.26 zeta^6) / (1 + xi^2 ((1+zeta)^(-4/3) + (1-zeta)^(-4/3))/2)^4,
 * the spin channels' eps~_s = max(eps_PBE(n_s, 0), eps_PBE) weighted n_s / n,
 * and z = tau_W / tau of the total density.  Inputs limited per spin
 * (tau_s >= 1e-20, s_ss <= 8 n_s tau_s). */
static cplx e_tpss_x_pol(cplx ra, cplx rb, cplx saa, cplx sbb, cplx ta, cplx tb) {
  cplx e = 0.0;
  if (creal(ra) >= ORACLE_DENS_THR) e += 0.5 * e_tpss_x(2.0 * ra, 4.0 * saa, 2.0 * ta);
  if (creal(rb) >= ORACLE_DENS_THR) e += 0.5 * e_tpss_x(2.0 * rb, 4.0 * sbb, 2.0 * tb);
  return e;
}

static cplx e_tpss_c_pol(cplx ra, cplx rb, cplx saa, cplx sab, cplx sbb, cplx ta, cplx tb) {
  const double d = 2.8;
  cplx n = ra + rb, zeta = (ra - rb) / n;
  cplx sig = saa + 2.0 * sab + sbb;
  cplx epbe = e_pbe_c_pol(ra, rb, sig) / n;
  cplx par = 0.0;
  if (creal(ra) >= ORACLE_DENS_THR) {
    cplx ea = e_pbe_c_pol(ra, 0.0, saa) / ra;
    par += ra / n * (creal(ea) > creal(epbe) ? ea : epbe);
  }
  if (creal(rb) >= ORACLE_DENS_THR) {
    cplx eb = e_pbe_c_pol(rb, 0.0, sbb) / rb;
    par += rb / n * (creal(eb) > creal(epbe) ? eb : epbe);
  }
  cplx z = sig / (8.0 * n * (ta + tb));
  if (creal(z) > 1.0) z = 1.0;
  cplx gz2 = 4.0 * (rb * rb * saa - 2.0 * ra * rb * sab + ra * ra * sbb) / (n * n * n * n);
  cplx xi2 = gz2 / (4.0 * cpow(3.0 * M_PI * M_PI * n, 2.0 / 3.0));
  cplx q = zeta * zeta;
  cplx c0 = 0.53 + 0.87 * q + 0.50 * q * q + 2.26 * q * q * q;
  cplx den = 1.0 + xi2 * (opz_pow(1.0 + zeta, -4.0 / 3.0) + opz_pow(1.0 - zeta, -4.0 / 3.0)) / 2.0;
  cplx C = c0 / (den * den * den * den);
  cplx z2 = z * z;
  cplx f0 = (1.0 + C * z2) * epbe - (1.0 + C) * z2 * par;
  return n * f0 * (1.0 + d * f0 * z2 * z);
}

/* x = (ra, rb, saa, sab, sbb, ta, tb) */
static cplx edens_mgga_pol(int id, const cplx *x) {
  switch (id) {
  case ORACLE_MGGA_X_TPSS:
    return e_tpss_x_pol(x[0], x[1], x[2], x[4], x[5], x[6]);
  case ORACLE_MGGA_C_TPSS:
    return e_tpss_c_pol(x[0], x[1], x[2], x[3], x[4], x[5], x[6]);
  }
  return edens_pol(id, x[0], x[1], x[2], x[3], x[4]);
}

static void mgga_inputs_pol(double *x) {
  for (int s = 0; s < 2; s++) {
    double *sg = x + 2 + 2 * s, *t = x + 5 + s;
    if (*sg < ORACLE_SIGMA_MIN) *sg = ORACLE_SIGMA_MIN;
    if (*t < ORACLE_TAU_MIN) *t = ORACLE_TAU_MIN;
    if (*sg > 8.0 * x[s] * *t) *sg = 8.0 * x[s] * *t;
  }
  /* |grad n_a . grad n_b| <= (s_aa + s_bb) / 2 after the per-spin limits */
  double sab_max = 0.5 * (x[2] + x[4]);
  if (x[3] > sab_max) x[3] = sab_max;
  if (x[3] < -sab_max) x[3] = -sab_max;
}

double oracle_functional_energy_density_mgga_pol(int id, const double *in) {
  double x[7];
  memcpy(x, in, sizeof(x));
  if (x[0] + x[1] < ORACLE_DENS_THR) return 0.0;
  mgga_inputs_pol(x);
  cplx y[7];
  for (int k = 0; k < 7; k++) y[k] = x[k];
  return creal(edens_mgga_pol(id, y));
}

int oracle_functional_eval_mgga_pol(int id, long n, const double *rho, const double *sigma,
                                    const double *tau, double *exc, double *vrho,
                                    double *vsigma, double *vtau) {
  int fam = oracle_functional_family(id);
  if (fam < 0) return -1;
  for (long i = 0; i < n; i++) {
    double x[7] = {rho[2 * i], rho[2 * i + 1], 0.0, 0.0, 0.0, 1.0, 1.0};
    double r = x[0] + x[1];
    if (!(r >= ORACLE_DENS_THR)) continue;
    if (fam >= 1) {
      x[2] = sigma[3 * i];
      x[3] = sigma[3 * i + 1];
      x[4] = sigma[3 * i + 2];
    }
    if (fam == 2) {
      x[5] = tau[2 * i];
      x[6] = tau[2 * i + 1];
      mgga_inputs_pol(x);
    } else if (fam == 1) {
      if (x[2] < ORACLE_SIGMA_MIN) x[2] = ORACLE_SIGMA_MIN;
      if (x[4] < ORACLE_SIGMA_MIN) x[4] = ORACLE_SIGMA_MIN;
    }
    cplx y[7];
    for (int k = 0; k < 7; k++) y[k] = x[k];
    exc[i] += creal(edens_mgga_pol(id, y)) / r;
    const int nv = fam == 2 ? 7 : (fam == 1 ? 5 : 2);
    for (int k = 0; k < nv; k++) {
      double h = 1e-20 * (fabs(x[k]) > 1e-300 ? fabs(x[k]) : 1.0);
      for (int q = 0; q < 7; q++) y[q] = x[q];
      y[k] += I * h;
      double dv = cimag(edens_mgga_pol(id, y)) / h;
      if (k < 2) vrho[2 * i + k] += dv;
      else if (k < 5) vsigma[3 * i + k - 2] += dv;
      else vtau[2 * i + k - 5] += dv;
    }
  }
  return 0;
}

/* ------------------------------------------------------------- XC build -- */

typedef struct {
  int deriv;
  int nfunc;
  const int *ids;
  double dthr;
} block_cfg;

/* kernels::process_grid_block<d, Restricted> (dft_kernels.cpp:8-70) with
 * density::evaluate_density (density.cpp:8-43), set_params (dft_kernels.h:
 * 16-58), Result::weight_by (functional.h:460-470) and xc_potential_matrix
 * (xc_potential_matrix.cpp:9-41).  phi is component-major [c][m][n]. */
static void process_block(const block_cfg *cfg, int m, long n,
                          const double *phi, const double *Dloc,
                          const double *wts, double *Kblk, double *energy,
                          double *nel, double *work) {
  const double *p0 = phi;
  const double *px = phi + (long)m * n;
  const double *py = phi + 2L * m * n;
  const double *pz = phi + 3L * m * n;
  double *T = work;              /* [m][n] */
  double *rho = T + (long)m * n; /* [4][n] */
  double *exc = rho + 4 * n, *vrho = exc + n, *vsig = vrho + n, *sig = vsig + n;
  double *A = sig + n; /* [m][n] */
  double *tau = A + (long)m * n, *vtau = tau + n; /* meta-GGA */
  const int mgga = cfg->deriv == 2;

  /* T = phi * D  (Dphi in density.cpp:14) */
  memset(T, 0, sizeof(double) * m * n);
  for (int nu = 0; nu < m; nu++) {
    double *t = T + (long)nu * n;
    for (int mu = 0; mu < m; mu++) {
      double dm = Dloc[mu + (long)nu * m];
      const double *f = p0 + (long)mu * n;
      for (long p = 0; p < n; p++) t[p] += dm * f[p];
    }
  }
  memset(rho, 0, sizeof(double) * 4 * n);
  for (int nu = 0; nu < m; nu++) {
    const double *t = T + (long)nu * n;
    const double *f = p0 + (long)nu * n;
    for (long p = 0; p < n; p++) rho[p] += f[p] * t[p];
    if (cfg->deriv) {
      const double *fx = px + (long)nu * n, *fy = py + (long)nu * n,
                   *fz = pz + (long)nu * n;
      for (long p = 0; p < n; p++) {
        rho[n + p] += 2.0 * fx[p] * t[p];
        rho[2 * n + p] += 2.0 * fy[p] * t[p];
        rho[3 * n + p] += 2.0 * fz[p] * t[p];
      }
    }
  }
  /* density.cpp:94-101: tau = (1/2) sum_k rowsum((grad_k phi D2) o grad_k phi) */
  if (mgga) {
    memset(tau, 0, sizeof(double) * n);
    for (int k = 1; k <= 3; k++) {
      const double *pk = phi + (long)k * m * n;
      memset(T, 0, sizeof(double) * m * n);
      for (int nu = 0; nu < m; nu++) {
        double *t = T + (long)nu * n;
        for (int mu = 0; mu < m; mu++) {
          double dm = Dloc[mu + (long)nu * m];
          const double *f = pk + (long)mu * n;
          for (long p = 0; p < n; p++) t[p] += dm * f[p];
        }
      }
      for (int nu = 0; nu < m; nu++) {
        const double *t = T + (long)nu * n, *f = pk + (long)nu * n;
        for (long p = 0; p < n; p++) tau[p] += f[p] * t[p];
      }
    }
    for (long p = 0; p < n; p++) tau[p] *= 0.5;
  }
  /* dft_kernels.cpp:29: N_elec before the skip */
  double ne = 0.0, rmax = -1e300;
  for (long p = 0; p < n; p++) {
    ne += rho[p] * wts[p];
    if (rho[p] > rmax) rmax = rho[p];
  }
  *nel += ne;
  if (rmax < cfg->dthr) return; /* dft_kernels.cpp:38-40 */

  for (long p = 0; p < n; p++) {
    exc[p] = vrho[p] = vsig[p] = 0.0;
    sig[p] = rho[n + p] * rho[n + p] + rho[2 * n + p] * rho[2 * n + p] +
             rho[3 * n + p] * rho[3 * n + p];
  }
  if (mgga) {
    for (long p = 0; p < n; p++) vtau[p] = 0.0;
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval_mgga(cfg->ids[f], n, rho, sig, tau, exc, vrho, vsig, vtau);
  } else {
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval(cfg->ids[f], n, rho, sig, exc, vrho, vsig);
  }
  for (long p = 0; p < n; p++) {
    exc[p] *= wts[p];
    vrho[p] *= wts[p];
    vsig[p] *= wts[p];
    if (mgga) vtau[p] *= wts[p];
  }
  double e = 0.0;
  for (long p = 0; p < n; p++) e += rho[p] * exc[p];
  *energy += e;

  if (!cfg->deriv) {
    /* Vxc = phi^T (vrho o phi) */
    for (int nu = 0; nu < m; nu++) {
      const double *f = p0 + (long)nu * n;
      double *a = A + (long)nu * n;
      for (long p = 0; p < n; p++) a[p] = vrho[p] * f[p];
    }
    for (int nu = 0; nu < m; nu++) {
      const double *a = A + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *f = p0 + (long)mu * n;
        double s = 0.0;
        for (long p = 0; p < n; p++) s += f[p] * a[p];
        Kblk[mu + (long)nu * m] += s;
      }
    }
  } else {
    /* A = 0.5 vrho phi + 2 vsigma (grad rho . grad phi);  K += A'phi + phi'A */
    for (int nu = 0; nu < m; nu++) {
      const double *f = p0 + (long)nu * n, *fx = px + (long)nu * n,
                   *fy = py + (long)nu * n, *fz = pz + (long)nu * n;
      double *a = A + (long)nu * n;
      for (long p = 0; p < n; p++)
        a[p] = 0.5 * vrho[p] * f[p] +
               2.0 * vsig[p] *
                   (rho[n + p] * fx[p] + rho[2 * n + p] * fy[p] + rho[3 * n + p] * fz[p]);
    }
    for (int nu = 0; nu < m; nu++) {
      const double *a = A + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *f = p0 + (long)mu * n;
        double s = 0.0;
        for (long p = 0; p < n; p++) s += f[p] * a[p];
        Kblk[mu + (long)nu * m] += s;
        Kblk[nu + (long)mu * m] += s;
      }
    }
    /* xc_potential_matrix<Restricted, 2> (xc_potential_matrix.cpp:44-58):
     * K += sum_k grad_k phi^T (vtau / 2 o grad_k phi) */
    if (mgga) {
      for (int k = 1; k <= 3; k++) {
        const double *pk = phi + (long)k * m * n;
        for (int nu = 0; nu < m; nu++) {
          const double *fn = pk + (long)nu * n;
          for (int mu = 0; mu < m; mu++) {
            const double *fm = pk + (long)mu * n;
            double s = 0.0;
            for (long p = 0; p < n; p++) s += fm[p] * (0.5 * vtau[p]) * fn[p];
            Kblk[mu + (long)nu * m] += s;
          }
        }
      }
    }
  }
}

typedef struct {
  const oracle_basis *b;
  int *first_bf;
  int nbf;
} basis_info;

static void eval_shells(const oracle_basis *b, const int *sig, int nsig,
                        long n, const double *xyz, int deriv, double *phi,
                        int m) {
  int off = 0;
  for (int k = 0; k < nsig; k++) {
    int s = sig[k];
    int l = b->l[s], pure = b->pure[s];
    double *p0 = phi + (long)off * n;
    double *px = deriv ? phi + ((long)m + off) * n : 0;
    double *py = deriv ? phi + (2L * m + off) * n : 0;
    double *pz = deriv ? phi + (3L * m + off) * n : 0;
    oracle_eval_shell(l, pure, b->nprim[s], b->exps + b->prim_offset[s],
                      b->coefs + b->prim_offset[s], b->origin + 3 * s, n, xyz,
                      deriv, p0, px, py, pz, n);
    off += oracle_shell_size(l, pure);
  }
}

int oracle_xc_build(const oracle_basis *b, long npts, const double *xyz,
                    const double *w, int nfunc, const int *func_ids,
                    const double *D, double density_threshold,
                    double screening_threshold, int screened, int nthreads,
                    double *K, double *exc_out, double *nelec_out) {
  if (!b || npts < 0 || nfunc <= 0) return -1;
  int deriv = 0;
  for (int f = 0; f < nfunc; f++) {
    int fam = oracle_functional_family(func_ids[f]);
    if (fam < 0) return -1;
    if (fam > deriv) deriv = fam;
  }
  const int nbf = oracle_nbf(b);
  const int nsh = b->nshell;
  int *first_bf = malloc(sizeof(int) * (nsh + 1));
  first_bf[0] = 0;
  for (int s = 0; s < nsh; s++)
    first_bf[s + 1] = first_bf[s] + oracle_shell_size(b->l[s], b->pure[s]);
  /* D2 = 2 D (dft.h:205), read through selfadjointView<Upper> */
  double *D2 = malloc(sizeof(double) * nbf * nbf);
  for (int j = 0; j < nbf; j++)
    for (int i = 0; i < nbf; i++) {
      int r = i < j ? i : j, c = i < j ? j : i;
      D2[i + (long)j * nbf] = 2.0 * D[r + (long)c * nbf];
    }
  memset(K, 0, sizeof(double) * nbf * nbf);
  block_cfg cfg = {deriv, nfunc, func_ids, density_threshold};

  /* work decomposition: list of (offset, count) blocks over a point array */
  const double *pxyz = xyz, *pw = w;
  double *sxyz = 0, *sw = 0, *lcen = 0, *lrad = 0, *ext = 0;
  long *perm = 0, *loff = 0, *lcnt = 0;
  long nblk;
  if (screened) {
    sxyz = malloc(sizeof(double) * 3 * (npts + 1));
    sw = malloc(sizeof(double) * (npts + 1));
    perm = malloc(sizeof(long) * (npts + 1));
    long cap = npts / 32 + 2;
    loff = malloc(sizeof(long) * cap);
    lcnt = malloc(sizeof(long) * cap);
    lcen = malloc(sizeof(double) * 3 * cap);
    lrad = malloc(sizeof(double) * cap);
    nblk = oracle_build_hierarchy(npts, xyz, w, sxyz, sw, perm, loff, lcnt, lcen, lrad);
    pxyz = sxyz;
    pw = sw;
    ext = malloc(sizeof(double) * nsh);
    oracle_decay_cutoff(b, screening_threshold, ext);
  } else {
    nblk = (npts + 127) / 128;
    loff = malloc(sizeof(long) * (nblk + 1));
    lcnt = malloc(sizeof(long) * (nblk + 1));
    for (long i = 0; i < nblk; i++) {
      loff[i] = 128 * i;
      lcnt[i] = (npts - 128 * i < 128) ? npts - 128 * i : 128;
    }
  }

  double energy = 0.0, nel = 0.0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : energy, nel)
  {
    double *Kloc = calloc((size_t)nbf * nbf, sizeof(double));
    int *sig = malloc(sizeof(int) * (nsh + 1));
    int *gidx = malloc(sizeof(int) * (nbf + 1));
    long maxn = 256;
    double *phi = malloc(sizeof(double) * 4 * (size_t)nbf * maxn);
    double *Dloc = malloc(sizeof(double) * (size_t)nbf * nbf);
    double *Kblk = malloc(sizeof(double) * (size_t)nbf * nbf);
    double *work = malloc(sizeof(double) * ((size_t)2 * nbf * maxn + 12 * maxn));
#pragma omp for schedule(dynamic, 4)
    for (long li = 0; li < nblk; li++) {
      long n = lcnt[li];
      if (n == 0) continue;
      int nsig = 0;
      if (screened) {
        /* dft.h:314-320 */
        for (int s = 0; s < nsh; s++) {
          const double *o = b->origin + 3 * s;
          double dx = o[0] - lcen[3 * li], dy = o[1] - lcen[3 * li + 1],
                 dz = o[2] - lcen[3 * li + 2];
          double dist = sqrt(dx * dx + dy * dy + dz * dz);
          if (dist - ext[s] < lrad[li]) sig[nsig++] = s;
        }
        if (nsig == 0) continue;
      } else {
        for (int s = 0; s < nsh; s++) sig[nsig++] = s;
      }
      int m = 0;
      for (int k = 0; k < nsig; k++) {
        int s = sig[k];
        for (int q = first_bf[s]; q < first_bf[s + 1]; q++) gidx[m++] = q;
      }
      for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++)
          Dloc[i + (long)j * m] = D2[gidx[i] + (long)gidx[j] * nbf];
      eval_shells(b, sig, nsig, n, pxyz + 3 * loff[li], deriv ? 1 : 0, phi, m);
      memset(Kblk, 0, sizeof(double) * m * m);
      process_block(&cfg, m, n, phi, Dloc, pw + loff[li], Kblk, &energy, &nel, work);
      for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++)
          Kloc[gidx[i] + (long)gidx[j] * nbf] += Kblk[i + (long)j * m];
    }
#pragma omp critical
    {
      for (long i = 0; i < (long)nbf * nbf; i++) K[i] += Kloc[i];
    }
    free(Kloc);
    free(sig);
    free(gidx);
    free(phi);
    free(Dloc);
    free(Kblk);
    free(work);
  }
  *exc_out = energy;
  *nelec_out = nel;
  free(first_bf);
  free(D2);
  free(sxyz);
  free(sw);
  free(perm);
  free(loff);
  free(lcnt);
  free(lcen);
  free(lrad);
  free(ext);
  return 0;
}

/* kernels::process_grid_block_gradient<d, Restricted>
 * (dft_gradient_kernels.cpp:473-552) with accumulate_gradient_contribution_R
 * <0> and <1> (dft_gradient_kernels.cpp:26-90).  phi is [10][m][n] (value,
 * gradient, Hessian xx,xy,xz,yy,yz,zz); Dloc is the gathered block of 2 D
 * (full, both triangles); g is 3 x natoms column-major. */
static void process_block_gradient(const block_cfg *cfg, int m, long n,
                                   const double *phi, const double *Dloc,
                                   const double *wts, const int *bf_atom,
                                   double *g, double *work) {
  const long mn = (long)m * n;
  const double *p0 = phi, *px = phi + mn, *py = phi + 2 * mn, *pz = phi + 3 * mn;
  const double *pxx = phi + 4 * mn, *pxy = phi + 5 * mn, *pxz = phi + 6 * mn;
  const double *pyy = phi + 7 * mn, *pyz = phi + 8 * mn, *pzz = phi + 9 * mn;
  double *T = work;                /* [m][n] */
  double *rho = T + mn;            /* [4][n] */
  double *exc = rho + 4 * n, *vrho = exc + n, *vsig = vrho + n, *sig = vsig + n;
  double *A = sig + n;             /* [m][n] */
  double *V = A + mn;              /* 3 x [m x m], column-major */
  double *tau = V + 3L * m * m, *vtau = tau + n; /* meta-GGA */
  const int mgga = cfg->deriv == 2;

  /* evaluate_density<d,R>: T = phi * Dloc.selfadjointView<Upper>() */
  memset(T, 0, sizeof(double) * mn);
  for (int nu = 0; nu < m; nu++) {
    double *t = T + (long)nu * n;
    for (int mu = 0; mu < m; mu++) {
      int r = mu < nu ? mu : nu, c = mu < nu ? nu : mu;
      double dm = Dloc[r + (long)c * m];
      const double *f = p0 + (long)mu * n;
      for (long p = 0; p < n; p++) t[p] += dm * f[p];
    }
  }
  memset(rho, 0, sizeof(double) * 4 * n);
  for (int nu = 0; nu < m; nu++) {
    const double *t = T + (long)nu * n, *f = p0 + (long)nu * n;
    for (long p = 0; p < n; p++) rho[p] += f[p] * t[p];
    if (cfg->deriv) {
      const double *fx = px + (long)nu * n, *fy = py + (long)nu * n, *fz = pz + (long)nu * n;
      for (long p = 0; p < n; p++) {
        rho[n + p] += 2.0 * fx[p] * t[p];
        rho[2 * n + p] += 2.0 * fy[p] * t[p];
        rho[3 * n + p] += 2.0 * fz[p] * t[p];
      }
    }
  }
  /* meta-GGA: tau = (1/2) sum_j rowsum((d_j phi Dloc) o d_j phi) (density.cpp:94-101) */
  if (mgga) {
    memset(tau, 0, sizeof(double) * n);
    const double *dj[3] = {px, py, pz};
    for (int j = 0; j < 3; j++) {
      memset(T, 0, sizeof(double) * mn);
      for (int nu = 0; nu < m; nu++) {
        double *t = T + (long)nu * n;
        for (int mu = 0; mu < m; mu++) {
          int r = mu < nu ? mu : nu, c = mu < nu ? nu : mu;
          double dm = Dloc[r + (long)c * m];
          const double *f = dj[j] + (long)mu * n;
          for (long p = 0; p < n; p++) t[p] += dm * f[p];
        }
      }
      for (int nu = 0; nu < m; nu++) {
        const double *t = T + (long)nu * n, *f = dj[j] + (long)nu * n;
        for (long p = 0; p < n; p++) tau[p] += f[p] * t[p];
      }
    }
    for (long p = 0; p < n; p++) tau[p] *= 0.5;
  }
  /* dft_gradient_kernels.cpp:488-491: skip when max rho < threshold */
  double rmax = -1e300;
  for (long p = 0; p < n; p++)
    if (rho[p] > rmax) rmax = rho[p];
  if (rmax < cfg->dthr) return;
  for (long p = 0; p < n; p++) {
    exc[p] = vrho[p] = vsig[p] = 0.0;
    sig[p] = rho[n + p] * rho[n + p] + rho[2 * n + p] * rho[2 * n + p] +
             rho[3 * n + p] * rho[3 * n + p];
  }
  if (mgga) {
    for (long p = 0; p < n; p++) vtau[p] = 0.0;
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval_mgga(cfg->ids[f], n, rho, sig, tau, exc, vrho, vsig, vtau);
  } else {
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval(cfg->ids[f], n, rho, sig, exc, vrho, vsig);
  }
  for (long p = 0; p < n; p++) { /* Result::weight_by */
    vrho[p] *= wts[p];
    vsig[p] *= wts[p];
    if (mgga) vtau[p] *= wts[p];
  }
  double *Vx = V, *Vy = V + (long)m * m, *Vz = V + 2L * m * m;
  memset(V, 0, sizeof(double) * 3 * m * m);
  if (!cfg->deriv) {
    /* <0>: V_k += d_k phi^T (phi o vrho) */
    for (int nu = 0; nu < m; nu++) {
      const double *f = p0 + (long)nu * n;
      double *a = A + (long)nu * n;
      for (long p = 0; p < n; p++) a[p] = f[p] * vrho[p];
    }
    for (int nu = 0; nu < m; nu++) {
      const double *a = A + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *fx = px + (long)mu * n, *fy = py + (long)mu * n, *fz = pz + (long)mu * n;
        double sx = 0, sy = 0, sz = 0;
        for (long p = 0; p < n; p++) {
          sx += fx[p] * a[p];
          sy += fy[p] * a[p];
          sz += fz[p] * a[p];
        }
        Vx[mu + (long)nu * m] += sx;
        Vy[mu + (long)nu * m] += sy;
        Vz[mu + (long)nu * m] += sz;
      }
    }
  } else {
    /* <1>: aow = 0.5 vrho phi + 2 vsigma grad rho . grad phi;
     * V_k += d_k phi^T aow, then (phi^T aow_k)^T with
     * aow_k = 0.5 vrho d_k phi + 2 vsigma sum_j d_j rho d_kj phi */
    for (int nu = 0; nu < m; nu++) {
      const long o = (long)nu * n;
      double *a = A + o;
      for (long p = 0; p < n; p++)
        a[p] = 0.5 * vrho[p] * p0[o + p] +
               2.0 * vsig[p] * (rho[n + p] * px[o + p] + rho[2 * n + p] * py[o + p] +
                                rho[3 * n + p] * pz[o + p]);
    }
    for (int nu = 0; nu < m; nu++) {
      const double *a = A + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *fx = px + (long)mu * n, *fy = py + (long)mu * n, *fz = pz + (long)mu * n;
        double sx = 0, sy = 0, sz = 0;
        for (long p = 0; p < n; p++) {
          sx += fx[p] * a[p];
          sy += fy[p] * a[p];
          sz += fz[p] * a[p];
        }
        Vx[mu + (long)nu * m] += sx;
        Vy[mu + (long)nu * m] += sy;
        Vz[mu + (long)nu * m] += sz;
      }
    }
    const double *dk[3] = {px, py, pz};
    const double *hk[3][3] = {{pxx, pxy, pxz}, {pxy, pyy, pyz}, {pxz, pyz, pzz}};
    double *Vk[3] = {Vx, Vy, Vz};
    for (int k = 0; k < 3; k++) {
      /* aow_k in A, then V_k(mu, nu) += sum_p aow_k(mu) phi(nu) */
      for (int mu = 0; mu < m; mu++) {
        const long o = (long)mu * n;
        double *a = A + o;
        for (long p = 0; p < n; p++)
          a[p] = 0.5 * vrho[p] * dk[k][o + p] +
                 2.0 * vsig[p] * (rho[n + p] * hk[k][0][o + p] + rho[2 * n + p] * hk[k][1][o + p] +
                                  rho[3 * n + p] * hk[k][2][o + p]);
      }
      for (int nu = 0; nu < m; nu++) {
        const double *f = p0 + (long)nu * n;
        for (int mu = 0; mu < m; mu++) {
          const double *a = A + (long)mu * n;
          double s = 0;
          for (long p = 0; p < n; p++) s += a[p] * f[p];
          Vk[k][mu + (long)nu * m] += s;
        }
      }
    }
    /* accumulate_gradient_contribution_R<2> (dft_gradient_kernels.cpp:
     * 152-172): aow_j = (vtau / 2) d_j phi, V_k += d_kj phi^T aow_j */
    if (mgga) {
      for (int j = 0; j < 3; j++) {
        for (int nu = 0; nu < m; nu++) {
          const long o = (long)nu * n;
          for (long p = 0; p < n; p++) A[o + p] = 0.5 * vtau[p] * dk[j][o + p];
        }
        for (int k = 0; k < 3; k++) {
          const double *h = hk[k][j];
          for (int nu = 0; nu < m; nu++) {
            const double *a = A + (long)nu * n;
            for (int mu = 0; mu < m; mu++) {
              const double *hm = h + (long)mu * n;
              double s = 0;
              for (long p = 0; p < n; p++) s += hm[p] * a[p];
              Vk[k][mu + (long)nu * m] += s;
            }
          }
        }
      }
    }
  }
  /* dft_gradient_kernels.cpp:533-550: distribute with 2 D(mu, nu), skipping
   * |2 D| < 1e-12 */
  for (int mu = 0; mu < m; mu++) {
    const int at = bf_atom[mu];
    for (int nu = 0; nu < m; nu++) {
      const double dv = 2.0 * Dloc[mu + (long)nu * m];
      if (fabs(dv) < 1e-12) continue;
      g[3 * at] -= Vx[mu + (long)nu * m] * dv;
      g[3 * at + 1] -= Vy[mu + (long)nu * m] * dv;
      g[3 * at + 2] -= Vz[mu + (long)nu * m] * dv;
    }
  }
}

int oracle_xc_gradient(const oracle_basis *b, const int *shell_atom, int natoms,
                       long npts, const double *xyz, const double *w, int nfunc,
                       const int *func_ids, const double *D,
                       double density_threshold, double screening_threshold,
                       int screened, int nthreads, double *grad) {
  if (!b || !shell_atom || natoms <= 0 || npts < 0 || nfunc <= 0) return -1;
  int deriv = 0;
  for (int f = 0; f < nfunc; f++) {
    int fam = oracle_functional_family(func_ids[f]);
    if (fam < 0) return -1;
    if (fam > deriv) deriv = fam;
  }
  const int nsh = b->nshell;
  for (int s = 0; s < nsh; s++)
    if (shell_atom[s] < 0 || shell_atom[s] >= natoms) return -1;
  const int nbf = oracle_nbf(b);
  int *first_bf = malloc(sizeof(int) * (nsh + 1));
  first_bf[0] = 0;
  for (int s = 0; s < nsh; s++)
    first_bf[s + 1] = first_bf[s] + oracle_shell_size(b->l[s], b->pure[s]);
  /* dft.h:606: D = 2 mo.D (not symmetrised; the density reads its upper
   * triangle, the distribution both) */
  double *D2 = malloc(sizeof(double) * nbf * nbf);
  for (long i = 0; i < (long)nbf * nbf; i++) D2[i] = 2.0 * D[i];
  memset(grad, 0, sizeof(double) * 3 * natoms);
  block_cfg cfg = {deriv, nfunc, func_ids, density_threshold};

  const double *pxyz = xyz, *pw = w;
  double *sxyz = 0, *sw = 0, *lcen = 0, *lrad = 0, *ext = 0;
  long *perm = 0, *loff = 0, *lcnt = 0;
  long nblk;
  if (screened) {
    sxyz = malloc(sizeof(double) * 3 * (npts + 1));
    sw = malloc(sizeof(double) * (npts + 1));
    perm = malloc(sizeof(long) * (npts + 1));
    long cap = npts / 32 + 2;
    loff = malloc(sizeof(long) * cap);
    lcnt = malloc(sizeof(long) * cap);
    lcen = malloc(sizeof(double) * 3 * cap);
    lrad = malloc(sizeof(double) * cap);
    nblk = oracle_build_hierarchy(npts, xyz, w, sxyz, sw, perm, loff, lcnt, lcen, lrad);
    pxyz = sxyz;
    pw = sw;
    ext = malloc(sizeof(double) * nsh);
    oracle_decay_cutoff(b, screening_threshold, ext);
  } else {
    /* the dense path (dft.h:631-667) cuts TBB-sized blocks; 128-point blocks
     * here (only the max-rho skip depends on the cut) */
    nblk = (npts + 127) / 128;
    loff = malloc(sizeof(long) * (nblk + 1));
    lcnt = malloc(sizeof(long) * (nblk + 1));
    for (long i = 0; i < nblk; i++) {
      loff[i] = 128 * i;
      lcnt[i] = (npts - 128 * i < 128) ? npts - 128 * i : 128;
    }
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double *gloc = calloc((size_t)3 * natoms, sizeof(double));
    int *sig = malloc(sizeof(int) * (nsh + 1));
    int *gidx = malloc(sizeof(int) * (nbf + 1));
    int *batom = malloc(sizeof(int) * (nbf + 1));
    long maxn = 256;
    double *phi = malloc(sizeof(double) * 10 * (size_t)nbf * maxn);
    double *tmp = malloc(sizeof(double) * 10 * (ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2 * maxn);
    double *Dloc = malloc(sizeof(double) * (size_t)nbf * nbf);
    double *work = malloc(sizeof(double) * ((size_t)2 * nbf * maxn + 10 * maxn +
                                            3 * (size_t)nbf * nbf));
#pragma omp for schedule(dynamic, 4)
    for (long li = 0; li < nblk; li++) {
      long n = lcnt[li];
      if (n == 0) continue;
      int nsig = 0;
      if (screened) {
        /* dft.h:697-702 */
        for (int s = 0; s < nsh; s++) {
          const double *o = b->origin + 3 * s;
          double dx = o[0] - lcen[3 * li], dy = o[1] - lcen[3 * li + 1],
                 dz = o[2] - lcen[3 * li + 2];
          if (sqrt(dx * dx + dy * dy + dz * dz) - ext[s] < lrad[li]) sig[nsig++] = s;
        }
        if (nsig == 0) continue;
      } else {
        for (int s = 0; s < nsh; s++) sig[nsig++] = s;
      }
      int m = 0;
      for (int k = 0; k < nsig; k++) {
        int s = sig[k];
        for (int q = first_bf[s]; q < first_bf[s + 1]; q++) {
          batom[m] = shell_atom[s];
          gidx[m++] = q;
        }
      }
      for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) Dloc[i + (long)j * m] = D2[gidx[i] + (long)gidx[j] * nbf];
      /* evaluate_basis_subset(..., max_deriv = 2, sig) */
      const double *bx = pxyz + 3 * loff[li];
      int off = 0;
      for (int k = 0; k < nsig; k++) {
        int s = sig[k], sz = oracle_shell_size(b->l[s], b->pure[s]);
        oracle_eval_shell2(b->l[s], b->pure[s], b->nprim[s], b->exps + b->prim_offset[s],
                           b->coefs + b->prim_offset[s], b->origin + 3 * s, n, bx, tmp, n);
        for (int q = 0; q < 10; q++)
          for (int o = 0; o < sz; o++)
            memcpy(phi + ((long)q * m + off + o) * n, tmp + ((long)q * sz + o) * n,
                   sizeof(double) * n);
        off += sz;
      }
      process_block_gradient(&cfg, m, n, phi, Dloc, pw + loff[li], batom, gloc, work);
    }
#pragma omp critical
    {
      for (int i = 0; i < 3 * natoms; i++) grad[i] += gloc[i];
    }
    free(gloc);
    free(sig);
    free(gidx);
    free(batom);
    free(phi);
    free(tmp);
    free(Dloc);
    free(work);
  }
  free(first_bf);
  free(D2);
  free(sxyz);
  free(sw);
  free(perm);
  free(loff);
  free(lcnt);
  free(lcen);
  free(lrad);
  free(ext);
  return 0;
}

/* kernels::process_grid_block<d, Unrestricted> (dft_kernels.cpp:8-70) with
 * evaluate_density<d, U> (density.cpp:18-30,45-70: T = phi D_s with the full
 * spin block, no symmetrisation), set_params<U> (dft_kernels.h:16-46) and
 * xc_potential_matrix<Unrestricted, d> (xc_potential_matrix.cpp:61-109).
 * Dloc is [2m][m] column-major (alpha rows, then beta rows), Kblk likewise. */
static void process_block_u(const block_cfg *cfg, int m, long n, const double *phi,
                            const double *Dloc, const double *wts, double *Kblk,
                            double *energy, double *nel_a, double *nel_b, double *work) {
  const long mn = (long)m * n;
  const double *p0 = phi, *px = phi + mn, *py = phi + 2 * mn, *pz = phi + 3 * mn;
  double *T = work;                   /* [m][n] */
  double *rho = T + mn;               /* [2][4][n] */
  double *exc = rho + 8 * n;          /* [n] */
  double *vr = exc + n;               /* [2n] interleaved a,b */
  double *vs = vr + 2 * n;            /* [3n] interleaved aa,ab,bb */
  double *rr = vs + 3 * n;            /* [2n] */
  double *ss = rr + 2 * n;            /* [3n] */
  double *G = ss + 3 * n;             /* [m][n] */
  double *tt = G + mn, *vt = tt + 2 * n; /* meta-GGA: tau_s, vtau_s interleaved a,b */
  const int mgga = cfg->deriv == 2;
  const int ld = 2 * m;
  memset(rho, 0, sizeof(double) * 8 * n);
  for (int sp = 0; sp < 2; sp++) {
    double *r = rho + 4L * n * sp;
    memset(T, 0, sizeof(double) * mn);
    for (int nu = 0; nu < m; nu++) {
      double *t = T + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double dm = Dloc[sp * m + mu + (long)nu * ld];
        const double *f = p0 + (long)mu * n;
        for (long p = 0; p < n; p++) t[p] += dm * f[p];
      }
    }
    for (int nu = 0; nu < m; nu++) {
      const double *t = T + (long)nu * n, *f = p0 + (long)nu * n;
      for (long p = 0; p < n; p++) r[p] += f[p] * t[p];
      if (cfg->deriv) {
        const double *fx = px + (long)nu * n, *fy = py + (long)nu * n, *fz = pz + (long)nu * n;
        for (long p = 0; p < n; p++) {
          r[n + p] += 2.0 * fx[p] * t[p];
          r[2 * n + p] += 2.0 * fy[p] * t[p];
          r[3 * n + p] += 2.0 * fz[p] * t[p];
        }
      }
    }
  }
  /* evaluate_density<2, U> (density.cpp:104-166): tau_s = (1/2) sum_k
   * rowsum((d_k phi D_s) o d_k phi) per spin block */
  if (mgga) {
    memset(tt, 0, sizeof(double) * 2 * n);
    const double *dk3[3] = {px, py, pz};
    for (int sp = 0; sp < 2; sp++)
      for (int k = 0; k < 3; k++) {
        memset(T, 0, sizeof(double) * mn);
        for (int nu = 0; nu < m; nu++) {
          double *t = T + (long)nu * n;
          for (int mu = 0; mu < m; mu++) {
            const double dm = Dloc[sp * m + mu + (long)nu * ld];
            const double *f = dk3[k] + (long)mu * n;
            for (long p = 0; p < n; p++) t[p] += dm * f[p];
          }
        }
        for (int nu = 0; nu < m; nu++) {
          const double *t = T + (long)nu * n, *f = dk3[k] + (long)nu * n;
          for (long p = 0; p < n; p++) tt[2 * p + sp] += f[p] * t[p];
        }
      }
    for (long p = 0; p < 2 * n; p++) tt[p] *= 0.5;
  }
  const double *ra = rho, *rb = rho + 4 * n;
  double na = 0.0, nb = 0.0, rmax = -1e300;
  for (long p = 0; p < n; p++) {
    na += ra[p] * wts[p];
    nb += rb[p] * wts[p];
    if (ra[p] > rmax) rmax = ra[p];
    if (rb[p] > rmax) rmax = rb[p];
  }
  *nel_a += na;
  *nel_b += nb;
  if (rmax < cfg->dthr) return;
  for (long p = 0; p < n; p++) {
    exc[p] = 0.0;
    vr[2 * p] = vr[2 * p + 1] = 0.0;
    vs[3 * p] = vs[3 * p + 1] = vs[3 * p + 2] = 0.0;
    rr[2 * p] = ra[p];
    rr[2 * p + 1] = rb[p];
    ss[3 * p] = ra[n + p] * ra[n + p] + ra[2 * n + p] * ra[2 * n + p] + ra[3 * n + p] * ra[3 * n + p];
    ss[3 * p + 1] = ra[n + p] * rb[n + p] + ra[2 * n + p] * rb[2 * n + p] + ra[3 * n + p] * rb[3 * n + p];
    ss[3 * p + 2] = rb[n + p] * rb[n + p] + rb[2 * n + p] * rb[2 * n + p] + rb[3 * n + p] * rb[3 * n + p];
  }
  if (mgga) {
    for (long p = 0; p < 2 * n; p++) vt[p] = 0.0;
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval_mgga_pol(cfg->ids[f], n, rr, ss, tt, exc, vr, vs, vt);
  } else {
    for (int f = 0; f < cfg->nfunc; f++) oracle_functional_eval_pol(cfg->ids[f], n, rr, ss, exc, vr, vs);
  }
  double e = 0.0;
  for (long p = 0; p < n; p++) { /* Result::weight_by, then exc . (rho_a + rho_b) */
    exc[p] *= wts[p];
    for (int k = 0; k < 2; k++) vr[2 * p + k] *= wts[p];
    for (int k = 0; k < 3; k++) vs[3 * p + k] *= wts[p];
    if (mgga) for (int k = 0; k < 2; k++) vt[2 * p + k] *= wts[p];
    e += exc[p] * ra[p] + exc[p] * rb[p];
  }
  *energy += e;
  for (int sp = 0; sp < 2; sp++) {
    const double *r_s = sp ? rb : ra, *r_o = sp ? ra : rb;
    /* LDA part: V_s = phi^T (vrho_s phi) */
    for (int nu = 0; nu < m; nu++) {
      const double *fn = p0 + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *fm = p0 + (long)mu * n;
        double s = 0.0;
        for (long p = 0; p < n; p++) s += fm[p] * vr[2 * p + sp] * fn[p];
        Kblk[sp * m + mu + (long)nu * ld] += s;
      }
    }
    if (!cfg->deriv) continue;
    /* g_s = 2 vsigma_ss grad rho_s + vsigma_ab grad rho_other; gamma = grad phi . g_s;
     * V_s += phi^T gamma + (phi^T gamma)^T */
    for (int nu = 0; nu < m; nu++) {
      const long o = (long)nu * n;
      for (long p = 0; p < n; p++) {
        const double vss = vs[3 * p + (sp ? 2 : 0)], vab = vs[3 * p + 1];
        const double gx = 2.0 * vss * r_s[n + p] + vab * r_o[n + p];
        const double gy = 2.0 * vss * r_s[2 * n + p] + vab * r_o[2 * n + p];
        const double gz = 2.0 * vss * r_s[3 * n + p] + vab * r_o[3 * n + p];
        G[o + p] = px[o + p] * gx + py[o + p] * gy + pz[o + p] * gz;
      }
    }
    for (int nu = 0; nu < m; nu++) {
      const double *gn = G + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double *fm = p0 + (long)mu * n;
        double s = 0.0;
        for (long p = 0; p < n; p++) s += fm[p] * gn[p];
        Kblk[sp * m + mu + (long)nu * ld] += s;
        Kblk[sp * m + nu + (long)mu * ld] += s;
      }
    }
    /* xc_potential_matrix<Unrestricted, 2> (xc_potential_matrix.cpp:111-137):
     * V_s += sum_k d_k phi^T (vtau_s / 2 o d_k phi) */
    if (mgga) {
      const double *dk3[3] = {px, py, pz};
      for (int k = 0; k < 3; k++)
        for (int nu = 0; nu < m; nu++) {
          const double *fn = dk3[k] + (long)nu * n;
          for (int mu = 0; mu < m; mu++) {
            const double *fm = dk3[k] + (long)mu * n;
            double s = 0.0;
            for (long p = 0; p < n; p++) s += fm[p] * (0.5 * vt[2 * p + sp]) * fn[p];
            Kblk[sp * m + mu + (long)nu * ld] += s;
          }
        }
    }
  }
}

int oracle_xc_build_unrestricted(const oracle_basis *b, long npts, const double *xyz,
                                 const double *w, int nfunc, const int *func_ids,
                                 const double *D, double density_threshold,
                                 double screening_threshold, int screened, int nthreads,
                                 double *K, double *exc_out, double *nelec_out) {
  if (!b || npts < 0 || nfunc <= 0) return -1;
  int deriv = 0;
  for (int f = 0; f < nfunc; f++) {
    int fam = oracle_functional_family(func_ids[f]);
    if (fam < 0) return -1;
    if (fam > deriv) deriv = fam;
  }
  const int nbf = oracle_nbf(b), nsh = b->nshell;
  const long ldD = 2L * nbf;
  int *first_bf = malloc(sizeof(int) * (nsh + 1));
  first_bf[0] = 0;
  for (int s = 0; s < nsh; s++) first_bf[s + 1] = first_bf[s] + oracle_shell_size(b->l[s], b->pure[s]);
  memset(K, 0, sizeof(double) * ldD * nbf);
  block_cfg cfg = {deriv, nfunc, func_ids, density_threshold};
  const double *pxyz = xyz, *pw = w;
  double *sxyz = 0, *sw = 0, *lcen = 0, *lrad = 0, *ext = 0;
  long *perm = 0, *loff = 0, *lcnt = 0;
  long nblk;
  if (screened) {
    sxyz = malloc(sizeof(double) * 3 * (npts + 1));
    sw = malloc(sizeof(double) * (npts + 1));
    perm = malloc(sizeof(long) * (npts + 1));
    long cap = npts / 32 + 2;
    loff = malloc(sizeof(long) * cap);
    lcnt = malloc(sizeof(long) * cap);
    lcen = malloc(sizeof(double) * 3 * cap);
    lrad = malloc(sizeof(double) * cap);
    nblk = oracle_build_hierarchy(npts, xyz, w, sxyz, sw, perm, loff, lcnt, lcen, lrad);
    pxyz = sxyz;
    pw = sw;
    ext = malloc(sizeof(double) * nsh);
    oracle_decay_cutoff(b, screening_threshold, ext);
  } else {
    nblk = (npts + 127) / 128;
    loff = malloc(sizeof(long) * (nblk + 1));
    lcnt = malloc(sizeof(long) * (nblk + 1));
    for (long i = 0; i < nblk; i++) {
      loff[i] = 128 * i;
      lcnt[i] = (npts - 128 * i < 128) ? npts - 128 * i : 128;
    }
  }
  double energy = 0.0, na = 0.0, nb = 0.0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : energy, na, nb)
  {
    double *Kloc = calloc((size_t)ldD * nbf, sizeof(double));
    int *sig = malloc(sizeof(int) * (nsh + 1));
    int *gidx = malloc(sizeof(int) * (nbf + 1));
    long maxn = 256;
    double *phi = malloc(sizeof(double) * 4 * (size_t)nbf * maxn);
    double *Dloc = malloc(sizeof(double) * 2 * (size_t)nbf * nbf);
    double *Kblk = malloc(sizeof(double) * 2 * (size_t)nbf * nbf);
    double *work = malloc(sizeof(double) * ((size_t)2 * nbf * maxn + 24 * maxn));
#pragma omp for schedule(dynamic, 4)
    for (long li = 0; li < nblk; li++) {
      long n = lcnt[li];
      if (n == 0) continue;
      int nsig = 0;
      if (screened) {
        for (int s = 0; s < nsh; s++) {
          const double *o = b->origin + 3 * s;
          double dx = o[0] - lcen[3 * li], dy = o[1] - lcen[3 * li + 1], dz = o[2] - lcen[3 * li + 2];
          if (sqrt(dx * dx + dy * dy + dz * dz) - ext[s] < lrad[li]) sig[nsig++] = s;
        }
        if (nsig == 0) continue;
      } else {
        for (int s = 0; s < nsh; s++) sig[nsig++] = s;
      }
      int m = 0;
      for (int k = 0; k < nsig; k++)
        for (int q = first_bf[sig[k]]; q < first_bf[sig[k] + 1]; q++) gidx[m++] = q;
      /* dft.h:340-348: Dloc per spin block from D2 = 2 D */
      for (int sp = 0; sp < 2; sp++)
        for (int j = 0; j < m; j++)
          for (int i = 0; i < m; i++)
            Dloc[sp * m + i + (long)j * 2 * m] = 2.0 * D[sp * nbf + gidx[i] + (long)gidx[j] * ldD];
      eval_shells(b, sig, nsig, n, pxyz + 3 * loff[li], deriv, phi, m);
      memset(Kblk, 0, sizeof(double) * 2 * m * m);
      process_block_u(&cfg, m, n, phi, Dloc, pw + loff[li], Kblk, &energy, &na, &nb, work);
      for (int sp = 0; sp < 2; sp++)
        for (int j = 0; j < m; j++)
          for (int i = 0; i < m; i++)
            Kloc[sp * nbf + gidx[i] + (long)gidx[j] * ldD] += Kblk[sp * m + i + (long)j * 2 * m];
    }
#pragma omp critical
    {
      for (long i = 0; i < ldD * nbf; i++) K[i] += Kloc[i];
    }
    free(Kloc);
    free(sig);
    free(gidx);
    free(phi);
    free(Dloc);
    free(Kblk);
    free(work);
  }
  *exc_out = energy;
  nelec_out[0] = na;
  nelec_out[1] = nb;
  free(first_bf);
  free(sxyz);
  free(sw);
  free(perm);
  free(loff);
  free(lcnt);
  free(lcen);
  free(lrad);
  free(ext);
  return 0;
}

/* kernels::process_grid_block_gradient<d, Unrestricted>
 * (dft_gradient_kernels.cpp:473-583) with accumulate_gradient_contribution_U
 * <0> and <1> (dft_gradient_kernels.cpp:176-290).  phi is [10][m][n]; Dloc
 * is [2m][m] (alpha rows over beta rows) of 2 D; g is 3 x natoms. */
static void process_block_gradient_u(const block_cfg *cfg, int m, long n, const double *phi,
                                     const double *Dloc, const double *wts, const int *bf_atom,
                                     double *g, double *work) {
  const long mn = (long)m * n;
  const double *p0 = phi, *pd[3] = {phi + mn, phi + 2 * mn, phi + 3 * mn};
  const double *ph[3][3] = {{phi + 4 * mn, phi + 5 * mn, phi + 6 * mn},
                            {phi + 5 * mn, phi + 7 * mn, phi + 8 * mn},
                            {phi + 6 * mn, phi + 8 * mn, phi + 9 * mn}};
  const int ld = 2 * m;
  double *T = work;                 /* [m][n] */
  double *rho = T + mn;             /* [2][4][n] */
  double *exc = rho + 8 * n, *vr = exc + n, *vs = vr + 2 * n, *rr = vs + 3 * n, *ss = rr + 2 * n;
  double *A = ss + 3 * n;           /* [m][n] */
  double *V = A + mn;               /* [2][3][m x m] */
  double *tt = V + 6L * m * m, *vt = tt + 2 * n; /* meta-GGA: tau_s, vtau_s (a,b per point) */
  const int mgga = cfg->deriv == 2;
  memset(rho, 0, sizeof(double) * 8 * n);
  for (int sp = 0; sp < 2; sp++) { /* evaluate_density<d, U> */
    double *r = rho + 4L * n * sp;
    memset(T, 0, sizeof(double) * mn);
    for (int nu = 0; nu < m; nu++) {
      double *t = T + (long)nu * n;
      for (int mu = 0; mu < m; mu++) {
        const double dm = Dloc[sp * m + mu + (long)nu * ld];
        const double *f = p0 + (long)mu * n;
        for (long p = 0; p < n; p++) t[p] += dm * f[p];
      }
    }
    for (int nu = 0; nu < m; nu++) {
      const double *t = T + (long)nu * n, *f = p0 + (long)nu * n;
      for (long p = 0; p < n; p++) r[p] += f[p] * t[p];
      if (cfg->deriv)
        for (int k = 0; k < 3; k++) {
          const double *fk = pd[k] + (long)nu * n;
          for (long p = 0; p < n; p++) r[(k + 1) * n + p] += 2.0 * fk[p] * t[p];
        }
    }
  }
  /* evaluate_density<2, U> (density.cpp:130-166): tau_s per spin block */
  if (mgga) {
    memset(tt, 0, sizeof(double) * 2 * n);
    for (int sp = 0; sp < 2; sp++)
      for (int k = 0; k < 3; k++) {
        memset(T, 0, sizeof(double) * mn);
        for (int nu = 0; nu < m; nu++) {
          double *t = T + (long)nu * n;
          for (int mu = 0; mu < m; mu++) {
            const double dm = Dloc[sp * m + mu + (long)nu * ld];
            const double *f = pd[k] + (long)mu * n;
            for (long p = 0; p < n; p++) t[p] += dm * f[p];
          }
        }
        for (int nu = 0; nu < m; nu++) {
          const double *t = T + (long)nu * n, *f = pd[k] + (long)nu * n;
          for (long p = 0; p < n; p++) tt[2 * p + sp] += f[p] * t[p];
        }
      }
    for (long p = 0; p < 2 * n; p++) tt[p] *= 0.5;
  }
  const double *ra = rho, *rb = rho + 4 * n;
  double rmax = -1e300;
  for (long p = 0; p < n; p++) {
    if (ra[p] > rmax) rmax = ra[p];
    if (rb[p] > rmax) rmax = rb[p];
  }
  if (rmax < cfg->dthr) return; /* dft_gradient_kernels.cpp:488-491 */
  for (long p = 0; p < n; p++) {
    exc[p] = 0.0;
    vr[2 * p] = vr[2 * p + 1] = 0.0;
    vs[3 * p] = vs[3 * p + 1] = vs[3 * p + 2] = 0.0;
    rr[2 * p] = ra[p];
    rr[2 * p + 1] = rb[p];
    ss[3 * p] = ra[n + p] * ra[n + p] + ra[2 * n + p] * ra[2 * n + p] + ra[3 * n + p] * ra[3 * n + p];
    ss[3 * p + 1] = ra[n + p] * rb[n + p] + ra[2 * n + p] * rb[2 * n + p] + ra[3 * n + p] * rb[3 * n + p];
    ss[3 * p + 2] = rb[n + p] * rb[n + p] + rb[2 * n + p] * rb[2 * n + p] + rb[3 * n + p] * rb[3 * n + p];
  }
  if (mgga) {
    for (long p = 0; p < 2 * n; p++) vt[p] = 0.0;
    for (int f = 0; f < cfg->nfunc; f++)
      oracle_functional_eval_mgga_pol(cfg->ids[f], n, rr, ss, tt, exc, vr, vs, vt);
  } else {
    for (int f = 0; f < cfg->nfunc; f++) oracle_functional_eval_pol(cfg->ids[f], n, rr, ss, exc, vr, vs);
  }
  for (long p = 0; p < n; p++) {
    for (int k = 0; k < 2; k++) vr[2 * p + k] *= wts[p];
    for (int k = 0; k < 3; k++) vs[3 * p + k] *= wts[p];
    if (mgga) for (int k = 0; k < 2; k++) vt[2 * p + k] *= wts[p];
  }
  memset(V, 0, sizeof(double) * 6 * m * m);
  for (int sp = 0; sp < 2; sp++) {
    const double *r_s = sp ? rb : ra, *r_o = sp ? ra : rb;
    double *Vs[3] = {V + (long)(3 * sp) * m * m, V + (long)(3 * sp + 1) * m * m,
                     V + (long)(3 * sp + 2) * m * m};
    if (!cfg->deriv) {
      /* <0>: V_k = d_k phi^T (phi vrho_s) */
      for (int nu = 0; nu < m; nu++) {
        const double *fn = p0 + (long)nu * n;
        for (int mu = 0; mu < m; mu++)
          for (int k = 0; k < 3; k++) {
            const double *fk = pd[k] + (long)mu * n;
            double acc = 0.0;
            for (long p = 0; p < n; p++) acc += fk[p] * fn[p] * vr[2 * p + sp];
            Vs[k][mu + (long)nu * m] += acc;
          }
      }
      continue;
    }
    /* <1>: g_s = 2 vsigma_ss grad rho_s + vsigma_ab grad rho_other;
     * aow = vrho_s/2 phi + g_s . grad phi; V_k += d_k phi^T aow;
     * aow_k = vrho_s/2 d_k phi + g_s . grad d_k phi; V_k += (phi^T aow_k)^T */
    for (int nu = 0; nu < m; nu++) {
      const long o = (long)nu * n;
      for (long p = 0; p < n; p++) {
        double gsum = 0.5 * vr[2 * p + sp] * p0[o + p];
        for (int j = 0; j < 3; j++)
          gsum += (2.0 * vs[3 * p + (sp ? 2 : 0)] * r_s[(j + 1) * n + p] +
                   vs[3 * p + 1] * r_o[(j + 1) * n + p]) * pd[j][o + p];
        A[o + p] = gsum;
      }
    }
    for (int nu = 0; nu < m; nu++) {
      const double *an = A + (long)nu * n;
      for (int mu = 0; mu < m; mu++)
        for (int k = 0; k < 3; k++) {
          const double *fk = pd[k] + (long)mu * n;
          double acc = 0.0;
          for (long p = 0; p < n; p++) acc += fk[p] * an[p];
          Vs[k][mu + (long)nu * m] += acc;
        }
    }
    for (int k = 0; k < 3; k++) {
      for (int mu = 0; mu < m; mu++) {
        const long o = (long)mu * n;
        for (long p = 0; p < n; p++) {
          double gsum = 0.5 * vr[2 * p + sp] * pd[k][o + p];
          for (int j = 0; j < 3; j++)
            gsum += (2.0 * vs[3 * p + (sp ? 2 : 0)] * r_s[(j + 1) * n + p] +
                     vs[3 * p + 1] * r_o[(j + 1) * n + p]) * ph[k][j][o + p];
          A[o + p] = gsum;
        }
      }
      for (int nu = 0; nu < m; nu++) {
        const double *fn = p0 + (long)nu * n;
        for (int mu = 0; mu < m; mu++) {
          const double *am = A + (long)mu * n;
          double acc = 0.0;
          for (long p = 0; p < n; p++) acc += am[p] * fn[p];
          Vs[k][mu + (long)nu * m] += acc;
        }
      }
    }
    /* accumulate_gradient_contribution_U<2> (dft_gradient_kernels.cpp:
     * 379-408 alpha, 441-470 beta): aow_j = (vtau_s / 2) d_j phi,
     * V_k += d_kj phi^T aow_j, not symmetrised */
    if (mgga) {
      for (int j = 0; j < 3; j++) {
        for (int nu = 0; nu < m; nu++) {
          const long o = (long)nu * n;
          for (long p = 0; p < n; p++) A[o + p] = 0.5 * vt[2 * p + sp] * pd[j][o + p];
        }
        for (int k = 0; k < 3; k++)
          for (int nu = 0; nu < m; nu++) {
            const double *an = A + (long)nu * n;
            for (int mu = 0; mu < m; mu++) {
              const double *hm = ph[k][j] + (long)mu * n;
              double acc = 0.0;
              for (long p = 0; p < n; p++) acc += hm[p] * an[p];
              Vs[k][mu + (long)nu * m] += acc;
            }
          }
      }
    }
  }
  /* dft_gradient_kernels.cpp:561-582: 2 (V_a D_a + V_b D_b), skipping
   * |D_a + D_b| < 1e-12 */
  for (int mu = 0; mu < m; mu++) {
    const int at = bf_atom[mu];
    for (int nu = 0; nu < m; nu++) {
      const double da = Dloc[mu + (long)nu * ld], db = Dloc[m + mu + (long)nu * ld];
      if (fabs(da + db) < 1e-12) continue;
      for (int k = 0; k < 3; k++)
        g[3 * at + k] -= 2.0 * (V[(long)k * m * m + mu + (long)nu * m] * da +
                                V[(long)(3 + k) * m * m + mu + (long)nu * m] * db);
    }
  }
}

int oracle_xc_gradient_unrestricted(const oracle_basis *b, const int *shell_atom, int natoms,
                                    long npts, const double *xyz, const double *w, int nfunc,
                                    const int *func_ids, const double *D,
                                    double density_threshold, double screening_threshold,
                                    int screened, int nthreads, double *grad) {
  if (!b || !shell_atom || natoms <= 0 || npts < 0 || nfunc <= 0) return -1;
  int deriv = 0;
  for (int f = 0; f < nfunc; f++) {
    int fam = oracle_functional_family(func_ids[f]);
    if (fam < 0) return -1;
    if (fam > deriv) deriv = fam;
  }
  const int nsh = b->nshell;
  for (int s = 0; s < nsh; s++)
    if (shell_atom[s] < 0 || shell_atom[s] >= natoms) return -1;
  const int nbf = oracle_nbf(b);
  const long ldD = 2L * nbf;
  int *first_bf = malloc(sizeof(int) * (nsh + 1));
  first_bf[0] = 0;
  for (int s = 0; s < nsh; s++) first_bf[s + 1] = first_bf[s] + oracle_shell_size(b->l[s], b->pure[s]);
  memset(grad, 0, sizeof(double) * 3 * natoms);
  block_cfg cfg = {deriv, nfunc, func_ids, density_threshold};
  const double *pxyz = xyz, *pw = w;
  double *sxyz = 0, *sw = 0, *lcen = 0, *lrad = 0, *ext = 0;
  long *perm = 0, *loff = 0, *lcnt = 0;
  long nblk;
  if (screened) {
    sxyz = malloc(sizeof(double) * 3 * (npts + 1));
    sw = malloc(sizeof(double) * (npts + 1));
    perm = malloc(sizeof(long) * (npts + 1));
    long cap = npts / 32 + 2;
    loff = malloc(sizeof(long) * cap);
    lcnt = malloc(sizeof(long) * cap);
    lcen = malloc(sizeof(double) * 3 * cap);
    lrad = malloc(sizeof(double) * cap);
    nblk = oracle_build_hierarchy(npts, xyz, w, sxyz, sw, perm, loff, lcnt, lcen, lrad);
    pxyz = sxyz;
    pw = sw;
    ext = malloc(sizeof(double) * nsh);
    oracle_decay_cutoff(b, screening_threshold, ext);
  } else {
    nblk = (npts + 127) / 128;
    loff = malloc(sizeof(long) * (nblk + 1));
    lcnt = malloc(sizeof(long) * (nblk + 1));
    for (long i = 0; i < nblk; i++) {
      loff[i] = 128 * i;
      lcnt[i] = (npts - 128 * i < 128) ? npts - 128 * i : 128;
    }
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double *gloc = calloc((size_t)3 * natoms, sizeof(double));
    int *sig = malloc(sizeof(int) * (nsh + 1));
    int *gidx = malloc(sizeof(int) * (nbf + 1));
    int *batom = malloc(sizeof(int) * (nbf + 1));
    long maxn = 256;
    double *phi = malloc(sizeof(double) * 10 * (size_t)nbf * maxn);
    double *tmp = malloc(sizeof(double) * 10 * (ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 2) / 2 * maxn);
    double *Dloc = malloc(sizeof(double) * 2 * (size_t)nbf * nbf);
    double *work = malloc(sizeof(double) * ((size_t)2 * nbf * maxn + 24 * maxn +
                                            6 * (size_t)nbf * nbf));
#pragma omp for schedule(dynamic, 4)
    for (long li = 0; li < nblk; li++) {
      long n = lcnt[li];
      if (n == 0) continue;
      int nsig = 0;
      if (screened) {
        for (int s = 0; s < nsh; s++) {
          const double *o = b->origin + 3 * s;
          double dx = o[0] - lcen[3 * li], dy = o[1] - lcen[3 * li + 1], dz = o[2] - lcen[3 * li + 2];
          if (sqrt(dx * dx + dy * dy + dz * dz) - ext[s] < lrad[li]) sig[nsig++] = s;
        }
        if (nsig == 0) continue;
      } else {
        for (int s = 0; s < nsh; s++) sig[nsig++] = s;
      }
      int m = 0;
      for (int k = 0; k < nsig; k++) {
        int s = sig[k];
        for (int q = first_bf[s]; q < first_bf[s + 1]; q++) {
          batom[m] = shell_atom[s];
          gidx[m++] = q;
        }
      }
      /* dft.h:717-726: Dloc per spin block of D = 2 mo.D */
      for (int sp = 0; sp < 2; sp++)
        for (int j = 0; j < m; j++)
          for (int i = 0; i < m; i++)
            Dloc[sp * m + i + (long)j * 2 * m] = 2.0 * D[sp * nbf + gidx[i] + (long)gidx[j] * ldD];
      const double *bx = pxyz + 3 * loff[li];
      int off = 0;
      for (int k = 0; k < nsig; k++) {
        int s = sig[k], sz = oracle_shell_size(b->l[s], b->pure[s]);
        oracle_eval_shell2(b->l[s], b->pure[s], b->nprim[s], b->exps + b->prim_offset[s],
                           b->coefs + b->prim_offset[s], b->origin + 3 * s, n, bx, tmp, n);
        for (int q = 0; q < 10; q++)
          for (int o = 0; o < sz; o++)
            memcpy(phi + ((long)q * m + off + o) * n, tmp + ((long)q * sz + o) * n,
                   sizeof(double) * n);
        off += sz;
      }
      process_block_gradient_u(&cfg, m, n, phi, Dloc, pw + loff[li], batom, gloc, work);
    }
#pragma omp critical
    {
      for (int i = 0; i < 3 * natoms; i++) grad[i] += gloc[i];
    }
    free(gloc);
    free(sig);
    free(gidx);
    free(batom);
    free(phi);
    free(tmp);
    free(Dloc);
    free(work);
  }
  free(first_bf);
  free(sxyz);
  free(sw);
  free(perm);
  free(loff);
  free(lcnt);
  free(lcen);
  free(lrad);
  free(ext);
  return 0;
}

long oracle_screened_work(const oracle_basis *b, long npts, const double *xyz,
                          const double *w, double screening_threshold,
                          long maxleaves, long *leaf_n, long *leaf_m) {
  const int nsh = b->nshell;
  double *sxyz = malloc(sizeof(double) * 3 * (npts + 1));
  double *sw = malloc(sizeof(double) * (npts + 1));
  long *perm = malloc(sizeof(long) * (npts + 1));
  long cap = npts / 32 + 2;
  long *loff = malloc(sizeof(long) * cap), *lcnt = malloc(sizeof(long) * cap);
  double *lcen = malloc(sizeof(double) * 3 * cap), *lrad = malloc(sizeof(double) * cap);
  long nl = oracle_build_hierarchy(npts, xyz, w, sxyz, sw, perm, loff, lcnt, lcen, lrad);
  double *ext = malloc(sizeof(double) * nsh);
  oracle_decay_cutoff(b, screening_threshold, ext);
  long k = 0;
  for (long li = 0; li < nl; li++) {
    long m = 0;
    for (int s = 0; s < nsh; s++) {
      const double *o = b->origin + 3 * s;
      double dx = o[0] - lcen[3 * li], dy = o[1] - lcen[3 * li + 1],
             dz = o[2] - lcen[3 * li + 2];
      if (sqrt(dx * dx + dy * dy + dz * dz) - ext[s] < lrad[li])
        m += oracle_shell_size(b->l[s], b->pure[s]);
    }
    if (m == 0) continue;
    if (k < maxleaves) {
      leaf_n[k] = lcnt[li];
      leaf_m[k] = m;
    }
    k++;
  }
  free(sxyz);
  free(sw);
  free(perm);
  free(loff);
  free(lcnt);
  free(lcen);
  free(lrad);
  free(ext);
  return k;
}

/* -------------------------------------------------------------- VV10 NLC -- */
/* vv10_kernel (nonlocal_correlation.cpp:37-103): per point i with
 * rho_i >= 1e-8, F_i = -1.5 [rho_i w_i / (2 k_i^3) + sum_{j<i, rho_j >= 1e-8}
 * 2 rho_j w_j / (g_i g_j (g_i + g_j))] with g = r^2 omega0 + kappa, and U_i,
 * W_i the matching sums for the kappa and omega0 derivatives; every j < i
 * pair counted twice, so that sum_i rho_i w_i F_i is the symmetric double sum.
 * exc = (beta + F/2) w, vrho = (beta + F + 1.5 (U dkappa/drho + W
 * domega0/drho)) w, vsigma = 1.5 W domega0/dsigma w.  Where the reference's
 * domega0/dsigma is 0/0 (sigma = 0) it is taken as 0. */
void oracle_vv10_kernel(long npts, const double *rho, const double *sigma, const double *xyz,
                        const double *w, double vv10_b, double vv10_C, int nthreads,
                        double *exc, double *vrho, double *vsigma) {
  const double pi4_3 = 4.0 * M_PI / 3.0;
  const double kappa_factor = vv10_b * 1.5 * M_PI * pow(9.0 * M_PI, -1.0 / 6.0);
  const double beta = pow(3.0 / (vv10_b * vv10_b), 0.75) / 32.0;
  double *rw = malloc(sizeof(double) * (npts + 1));
  double *om = malloc(sizeof(double) * (npts + 1));
  double *ka = malloc(sizeof(double) * (npts + 1));
  double *F = calloc(npts + 1, sizeof(double));
  double *U = calloc(npts + 1, sizeof(double));
  double *W = calloc(npts + 1, sizeof(double));
  for (long i = 0; i < npts; i++) {
    rw[i] = rho[i] * w[i];
    double t = sigma[i] / (rho[i] * rho[i]);
    t = vv10_C * t * t;
    om[i] = sqrt(t + pi4_3 * rho[i]);
    ka[i] = kappa_factor * pow(rho[i], 1.0 / 6.0);
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
  for (long i = 0; i < npts; i++) {
    if (rho[i] < 1e-8) continue;
    const double k = ka[i];
    double f = rw[i] / (2 * k * k * k);
    double u = f * (1.0 / k + 1.0 / (2 * k));
    double ww = 0.0;
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    for (long j = 0; j < i; j++) {
      if (rho[j] < 1e-8) continue;
      const double dx = xi - xyz[3 * j], dy = yi - xyz[3 * j + 1], dz = zi - xyz[3 * j + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double gi = r2 * om[i] + ka[i];
      const double gj = r2 * om[j] + ka[j];
      double t = 2 * rw[j] / (gi * gj * (gi + gj));
      f += t;
      t *= 1.0 / gi + 1.0 / (gi + gj);
      u += t;
      ww += t * r2;
    }
    F[i] = f * -1.5;
    U[i] = u;
    W[i] = ww;
  }
  for (long i = 0; i < npts; i++) {
    double t = sigma[i] / (rho[i] * rho[i]);
    t = vv10_C * t * t;
    const double domega0_dr = (0.5 * pi4_3 * rho[i] - 2.0 * t) / om[i];
    const double domega0_dg = sigma[i] > 0.0 ? t * rho[i] / (sigma[i] * om[i]) : 0.0;
    const double dkappa_dr = ka[i] / 6.0;
    exc[i] = (beta + 0.5 * F[i]) * w[i];
    if (rho[i] < 1e-8) {
      vrho[i] = beta * w[i];
      vsigma[i] = 0.0;
    } else {
      vrho[i] = (beta + F[i] + 1.5 * (U[i] * dkappa_dr + W[i] * domega0_dr)) * w[i];
      vsigma[i] = (1.5 * W[i] * domega0_dg) * w[i];
    }
  }
  free(rw);
  free(om);
  free(ka);
  free(F);
  free(U);
  free(W);
}

/* all shells at n points: phi [4][nbf][n] */
static void eval_all_shells(const oracle_basis *b, long n, const double *xyz, double *phi,
                            int nbf) {
  int off = 0;
  for (int s = 0; s < b->nshell; s++) {
    int l = b->l[s], pure = b->pure[s];
    oracle_eval_shell(l, pure, b->nprim[s], b->exps + b->prim_offset[s],
                      b->coefs + b->prim_offset[s], b->origin + 3 * s, n, xyz, 1,
                      phi + (long)off * n, phi + ((long)nbf + off) * n,
                      phi + (2L * nbf + off) * n, phi + (3L * nbf + off) * n, n);
    off += oracle_shell_size(l, pure);
  }
}

int oracle_vv10(const oracle_basis *b, long npts, const double *xyz, const double *w,
                const double *D, double vv10_b, double vv10_C, int nthreads,
                double *energy, double *K) {
  if (!b || npts < 0) return -1;
  const int nbf = oracle_nbf(b);
  const long nblk = (npts + 127) / 128;
  double *D2 = malloc(sizeof(double) * nbf * nbf);
  for (long i = 0; i < (long)nbf * nbf; i++) D2[i] = 2.0 * D[i];
  double *g = malloc(sizeof(double) * 3 * (npts + 1));
  double *rho = malloc(sizeof(double) * (npts + 1));
  double *sig = malloc(sizeof(double) * (npts + 1));
  /* rho, grad rho: evaluate_density<1, R> (density.cpp:33-43) */
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double *phi = malloc(sizeof(double) * 4 * (size_t)nbf * 128);
    double *T = malloc(sizeof(double) * (size_t)nbf * 128);
#pragma omp for schedule(dynamic, 4)
    for (long bi = 0; bi < nblk; bi++) {
      const long p0 = 128 * bi, n = npts - p0 < 128 ? npts - p0 : 128;
      eval_all_shells(b, n, xyz + 3 * p0, phi, nbf);
      memset(T, 0, sizeof(double) * nbf * n);
      for (int nu = 0; nu < nbf; nu++)
        for (int mu = 0; mu < nbf; mu++) {
          const double dm = D2[mu + (long)nu * nbf];
          for (long p = 0; p < n; p++) T[nu * n + p] += dm * phi[(long)mu * n + p];
        }
      for (long p = 0; p < n; p++) {
        double r = 0, gx = 0, gy = 0, gz = 0;
        for (int nu = 0; nu < nbf; nu++) {
          const double t = T[nu * n + p];
          r += phi[(long)nu * n + p] * t;
          gx += 2.0 * phi[((long)nbf + nu) * n + p] * t;
          gy += 2.0 * phi[(2L * nbf + nu) * n + p] * t;
          gz += 2.0 * phi[(3L * nbf + nu) * n + p] * t;
        }
        rho[p0 + p] = r;
        g[3 * (p0 + p)] = gx;
        g[3 * (p0 + p) + 1] = gy;
        g[3 * (p0 + p) + 2] = gz;
        sig[p0 + p] = gx * gx + gy * gy + gz * gz;
      }
    }
    free(phi);
    free(T);
  }
  double *exc = malloc(sizeof(double) * (npts + 1));
  double *vr = malloc(sizeof(double) * (npts + 1));
  double *vs = malloc(sizeof(double) * (npts + 1));
  oracle_vv10_kernel(npts, rho, sig, xyz, w, vv10_b, vv10_C, nthreads, exc, vr, vs);
  double e = 0.0;
  for (long i = 0; i < npts; i++) e += exc[i] * rho[i];
  *energy = e;
  /* Vxc: ktmp = phi^T (vrho/2 phi + 2 vsigma grad rho . grad phi), K = ktmp + ktmp^T
   * (nonlocal_correlation.cpp:295-303) */
  memset(K, 0, sizeof(double) * nbf * nbf);
#pragma omp parallel
  {
    double *phi = malloc(sizeof(double) * 4 * (size_t)nbf * 128);
    double *A = malloc(sizeof(double) * (size_t)nbf * 128);
    double *Kl = calloc((size_t)nbf * nbf, sizeof(double));
#pragma omp for schedule(dynamic, 4)
    for (long bi = 0; bi < nblk; bi++) {
      const long p0 = 128 * bi, n = npts - p0 < 128 ? npts - p0 : 128;
      eval_all_shells(b, n, xyz + 3 * p0, phi, nbf);
      for (int nu = 0; nu < nbf; nu++)
        for (long p = 0; p < n; p++) {
          const long q = p0 + p;
          A[nu * n + p] = 0.5 * vr[q] * phi[(long)nu * n + p] +
                          2.0 * vs[q] *
                              (g[3 * q] * phi[((long)nbf + nu) * n + p] +
                               g[3 * q + 1] * phi[(2L * nbf + nu) * n + p] +
                               g[3 * q + 2] * phi[(3L * nbf + nu) * n + p]);
        }
      for (int nu = 0; nu < nbf; nu++)
        for (int mu = 0; mu < nbf; mu++) {
          double s = 0.0;
          for (long p = 0; p < n; p++) s += phi[(long)mu * n + p] * A[nu * n + p];
          Kl[mu + (long)nu * nbf] += s;
          Kl[nu + (long)mu * nbf] += s;
        }
    }
#pragma omp critical
    for (long i = 0; i < (long)nbf * nbf; i++) K[i] += Kl[i];
    free(phi);
    free(A);
    free(Kl);
  }
  free(D2);
  free(g);
  free(rho);
  free(sig);
  free(exc);
  free(vr);
  free(vs);
  return 0;
}
