/*
 * occxc_oracle.h -- CPU FP64 restatement of OCC's Kohn-Sham XC build.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is called from tests/, from __graft_entry__.smoke() and from the
 * cpu_baseline leg of bench.py, and never from the product library.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the OCC source tree).  Pinning status of each part is listed in
 * occxc_oracle.c's header comment.
 */
#ifndef OCCXC_ORACLE_H
#define OCCXC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Basis description: shells in OCC order (atom order, then JSON shell order,
 * general contractions split).  coefs are the normalised contraction
 * coefficients OCC stores after Shell::incorporate_shell_norm
 * (src/gto/shell.cpp:180-186), *without* the common_fac(l) factor. */
typedef struct {
  int nshell;
  const int *l;           /* [nshell] */
  const int *pure;        /* [nshell] 1 = spherical (CCA), 0 = Cartesian */
  const int *nprim;       /* [nshell] */
  const int *prim_offset; /* [nshell] offset into exps/coefs */
  const double *origin;   /* [3*nshell] Bohr */
  const double *exps;
  const double *coefs;
} oracle_basis;

int oracle_nbf(const oracle_basis *b);
int oracle_shell_size(int l, int pure);

/* shell.cpp:20-64,180-186: primitive norm then contracted norm. */
void oracle_normalize_shell(int l, int nprim, const double *exps,
                            const double *raw_coefs, double *out_coefs);
/* gto.cpp:7-16 */
double oracle_common_fac(int l, int pure);

/* CCA Cartesian -> spherical coefficients, row m (m=-l..l) x Cartesian
 * component in CCA order (gau2grid_transform.c).  out has (2l+1)*ncart. */
void oracle_cart_to_sph(int l, double *out);

/* Basis on points for one shell, times common_fac (gto.cpp:212-287 with
 * gau2grid gg_collocation[_deriv1]).  xyz interleaved [npts*3].  Outputs are
 * component-major with leading dimension ld: phi[c*ld + p]. deriv 0 or 1. */
void oracle_eval_shell(int l, int pure, int nprim, const double *exps,
                       const double *coefs, const double *center, long npts,
                       const double *xyz, int deriv, double *phi, double *phx,
                       double *phy, double *phz, long ld);

/* Value, gradient and Hessian of one shell, times common_fac (gto.cpp:245-285
 * with gau2grid gg_collocation_deriv2).  out[(q*nout + o)*ld + p] with q = 0
 * value, 1..3 x,y,z, 4..9 xx,xy,xz,yy,yz,zz. */
void oracle_eval_shell2(int l, int pure, int nprim, const double *exps,
                        const double *coefs, const double *center, long npts,
                        const double *xyz, double *out, long ld);

/* gto.cpp:18-67 */
void oracle_decay_cutoff(const oracle_basis *b, double threshold,
                         double *extents);

/* spatial_grid_hierarchy.cpp:28-147.  Returns number of leaves; sorted
 * points/weights/permutation and leaf data written to the out arrays
 * (leaf arrays must hold n/32+2 entries). */
long oracle_build_hierarchy(long n, const double *xyz, const double *w,
                            double *sxyz, double *sw, long *perm,
                            long *leaf_offset, long *leaf_count,
                            double *leaf_center, double *leaf_radius);

/* libxc ids of the functionals restated here. */
enum {
  ORACLE_LDA_X = 1,
  ORACLE_LDA_C_VWN = 7,
  ORACLE_LDA_C_PW = 12,
  ORACLE_LDA_C_PW_MOD = 13,
  ORACLE_GGA_X_PBE = 101,
  ORACLE_GGA_X_B88 = 106,
  ORACLE_GGA_C_PBE = 130,
  ORACLE_GGA_C_LYP = 131,
  ORACLE_MGGA_X_TPSS = 202,
  ORACLE_MGGA_C_TPSS = 231
};
/* 0 = LDA, 1 = GGA, 2 = meta-GGA, -1 unknown id */
int oracle_functional_family(int id);
/* DensityFunctional::evaluate (src/dft/functional.cpp:18-65) for an
 * unpolarised density: ACCUMULATES exc, vrho, vsigma (+=), like
 * Result::operator+= (functional.h:436-457).  sigma/vsigma may be NULL for
 * LDA.  Returns 0 or -1 for an unknown id. */
int oracle_functional_eval(int id, long n, const double *rho,
                           const double *sigma, double *exc, double *vrho,
                           double *vsigma);
/* Energy density e = rho*eps at one point, exposed for tests. */
double oracle_functional_energy_density(int id, double rho, double sigma);
/* The same for meta-GGA densities (Family::MGGA, functional.cpp:49-57):
 * tau[n] = (1/2) sum_i |grad psi_i|^2; ACCUMULATES exc, vrho, vsigma, vtau
 * (vlaplacian is zero for every functional here).  LDA and GGA ids are
 * accepted too (tau ignored). */
int oracle_functional_eval_mgga(int id, long n, const double *rho, const double *sigma,
                                const double *tau, double *exc, double *vrho,
                                double *vsigma, double *vtau);
double oracle_functional_energy_density_mgga(int id, double rho, double sigma, double tau);
/* Spin-polarised meta-GGA (libxc's polarised layout): rho[2n], sigma[3n],
 * tau[2n] = (tau_a, tau_b); ACCUMULATES exc[n], vrho[2n], vsigma[3n],
 * vtau[2n].  LDA and GGA ids are accepted too (tau ignored). */
int oracle_functional_eval_mgga_pol(int id, long n, const double *rho, const double *sigma,
                                    const double *tau, double *exc, double *vrho,
                                    double *vsigma, double *vtau);
/* in = (rho_a, rho_b, s_aa, s_ab, s_bb, tau_a, tau_b) */
double oracle_functional_energy_density_mgga_pol(int id, const double *in);

/* The same functionals for a spin-polarised density, libxc's layout:
 * rho[2n] = (rho_a, rho_b) per point, sigma[3n] = (s_aa, s_ab, s_bb);
 * ACCUMULATES exc[n], vrho[2n], vsigma[3n].  sigma/vsigma may be NULL for
 * LDA.  Returns 0 or -1 for an unknown id. */
int oracle_functional_eval_pol(int id, long n, const double *rho, const double *sigma,
                               double *exc, double *vrho, double *vsigma);
double oracle_functional_energy_density_pol(int id, double ra, double rb, double saa,
                                            double sab, double sbb);

/* DFT::compute_K_dft<d, Restricted> (include/occ/dft/dft.h:190-388) and
 * kernels::process_grid_block (src/dft/dft_kernels.cpp:8-70).
 *   D: nbf x nbf column-major restricted density, no factor 2 (orb.h:64-66);
 *      only the upper triangle is read (density.cpp:14,38).
 *   xyz/w: unsorted molecular grid.
 *   screened=1: leaf loop with shell screening (dft.h:279-372);
 *   screened=0: dense loop over 128-point chunks in input order (dft.h:236-272).
 *   K: nbf x nbf column-major output (the value compute_vxc returns).
 * Returns 0, or -1 on invalid input. nthreads<=0 -> OpenMP default. */
int oracle_xc_build(const oracle_basis *b, long npts, const double *xyz,
                    const double *w, int nfunc, const int *func_ids,
                    const double *D, double density_threshold,
                    double screening_threshold, int screened, int nthreads,
                    double *K, double *exc, double *nelec);

/* DFT::compute_xc_gradient_impl<d, Restricted> (include/occ/dft/dft.h:598-749)
 * with kernels::process_grid_block_gradient (src/dft/dft_gradient_kernels.cpp:
 * 473-552): the XC contribution to the nuclear gradient, 3 x natoms
 * column-major (Mat3N).  shell_atom[s] is the atom of shell s
 * (AOBasis::shell_to_atom).  Other arguments as oracle_xc_build.  Returns 0,
 * or -1 on invalid input. */
int oracle_xc_gradient(const oracle_basis *b, const int *shell_atom, int natoms,
                       long npts, const double *xyz, const double *w, int nfunc,
                       const int *func_ids, const double *D,
                       double density_threshold, double screening_threshold,
                       int screened, int nthreads, double *grad);

/* DFT::compute_K_dft<d, Unrestricted> (include/occ/dft/dft.h:190-388) with
 * process_grid_block<d, Unrestricted> (src/dft/dft_kernels.cpp:8-70):
 *   D: 2nbf x nbf column-major, alpha block in rows 0..nbf-1 and beta block
 *      in rows nbf..2nbf-1 (qm::block::a / b), no factor 2; each block is
 *      read in full (density.cpp:18-30,45-70 use no selfadjointView).
 *   K: 2nbf x nbf column-major output; nelec[2] = {N_alpha, N_beta}.
 * Returns 0, or -1 on invalid input. */
int oracle_xc_build_unrestricted(const oracle_basis *b, long npts, const double *xyz,
                                 const double *w, int nfunc, const int *func_ids,
                                 const double *D, double density_threshold,
                                 double screening_threshold, int screened, int nthreads,
                                 double *K, double *exc, double *nelec);

/* compute_xc_gradient_impl<d, Unrestricted> (dft.h:598-749) with
 * process_grid_block_gradient<d, Unrestricted> (dft_gradient_kernels.cpp:
 * 473-583): D as oracle_xc_build_unrestricted, the rest as
 * oracle_xc_gradient. */
int oracle_xc_gradient_unrestricted(const oracle_basis *b, const int *shell_atom, int natoms,
                                    long npts, const double *xyz, const double *w, int nfunc,
                                    const int *func_ids, const double *D,
                                    double density_threshold, double screening_threshold,
                                    int screened, int nthreads, double *grad);

/* NonLocalCorrelationFunctional::vv10 (src/dft/nonlocal_correlation.cpp:
 * 246-306) with vv10_kernel (nonlocal_correlation.cpp:37-103): the VV10
 * energy and Vxc (nbf x nbf column-major) of the restricted density D (no
 * factor 2) on the grid xyz/w taken in the given point order (the per-point
 * sums run over the points before each one, as the reference's do).  AOs are
 * evaluated at every point (no screening), as evaluate_basis does there.
 * Returns 0, or -1 on invalid input. */
int oracle_vv10(const oracle_basis *b, long npts, const double *xyz, const double *w,
                const double *D, double vv10_b, double vv10_C, int nthreads,
                double *energy, double *K);
/* The per-point part alone (vv10_kernel) from rho and sigma = |grad rho|^2:
 * exc[i] and vxc (vrho[i], vsigma[i]), all weighted. */
void oracle_vv10_kernel(long npts, const double *rho, const double *sigma, const double *xyz,
                        const double *w, double vv10_b, double vv10_C, int nthreads,
                        double *exc, double *vrho, double *vsigma);

/* Algorithmic work of the screened build: per-leaf (n_b, m_b) with the
 * reference screening rule.  Writes up to maxleaves entries; returns the
 * number of leaves with >= 1 significant shell. */
long oracle_screened_work(const oracle_basis *b, long npts, const double *xyz,
                          const double *w, double screening_threshold,
                          long maxleaves, long *leaf_n, long *leaf_m);

#ifdef __cplusplus
}
#endif
#endif
