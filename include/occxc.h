/*
 * occxc.h -- C ABI of the MI355X-native Kohn-Sham exchange-correlation build.
 *
 * The boundary replaces the body of OCC's
 *     Mat DFT::compute_vxc(const MolecularOrbitals &mo, const Mat &Schwarz)
 *     (include/occ/dft/dft.h:395-429, dispatching to compute_K_dft<d,kind>,
 *      dft.h:190-388)
 * plus the value read back by
 *     double DFT::exchange_correlation_energy() const   (dft.h:83)
 * and additionally returns N_elec, which OCC only logs (dft.h:384-385).
 *
 * Conventions kept from OCC:
 *   - basis shells in AOBasis order (atom order, then per-element JSON shell
 *     order, general contractions split, src/gto/shell.cpp:529-538);
 *     coefficients normalised as Shell::incorporate_shell_norm leaves them
 *     (shell.cpp:180-186), without the common_fac(l) factor (gto.cpp:7-16);
 *     spherical components in CCA order m=-l..l, Cartesian xx,xy,xz,yy,yz,zz
 *     (include/occ/gto/shell_order.h:8-23);
 *   - lengths in Bohr (units.h:7-8);
 *   - D is the restricted density C_occ C_occ^T, nbf x nbf column-major, with
 *     no factor 2 (include/occ/qm/orb.h:64-66); the build doubles it
 *     (dft.h:205) and reads only the upper triangle (src/gto/density.cpp:14,38);
 *   - the returned matrix is what compute_vxc returns (jk.K = -Vxc at
 *     dft.h:435);
 *   - functionals are named by their libxc ids (src/dft/functional.cpp:78-90).
 *
 * No function throws; every entry point returns an occxc_status and leaves a
 * message for occxc_last_error_message().  Host buffers are caller-owned;
 * device buffers inside a context are library-owned.  A context is
 * single-writer; distinct contexts may be used from different threads.
 */
#ifndef OCCXC_H
#define OCCXC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OCCXC_OK = 0,
  OCCXC_ERR_INVALID_ARGUMENT = 1,
  OCCXC_ERR_UNSUPPORTED = 2,      /* General spinorbitals, d > 1, unknown id */
  OCCXC_ERR_DEVICE = 3,           /* HIP runtime error or no device */
  OCCXC_ERR_OUT_OF_MEMORY = 4,
  OCCXC_ERR_ADMISSION = 5,        /* serving: footprint exceeds the budget */
  OCCXC_ERR_CANCELLED = 6,
  OCCXC_ERR_INTERNAL = 7
} occxc_status;

/* libxc functional ids understood by the build */
enum {
  OCCXC_LDA_X = 1,
  OCCXC_LDA_C_VWN = 7,        /* VWN5 */
  OCCXC_LDA_C_PW = 12,
  OCCXC_LDA_C_PW_MOD = 13,
  OCCXC_GGA_X_PBE = 101,
  OCCXC_GGA_X_B88 = 106,
  OCCXC_GGA_C_PBE = 130,
  OCCXC_GGA_C_LYP = 131,
  OCCXC_MGGA_X_TPSS = 202,    /* meta-GGA: builds and gradients, both spin cases */
  OCCXC_MGGA_C_TPSS = 231
};

typedef enum {
  OCCXC_PREC_FP64 = 0,       /* FP64 AO cache and contractions */
  OCCXC_PREC_FP32 = 1,       /* FP32 AO cache; density contraction in FP64; Vxc contraction in
                                FP32 with FP32 accumulation over 32 points, FP64 beyond, from
                                an A plane formed each build (one more FP32 plane of device
                                memory).  The FP32-cache modes add a first-order
                                storage-rounding correction to N_elec (formed at context
                                creation) */
  OCCXC_PREC_BF16X3 = 2,     /* FP32 AO cache, 3-term split-BF16 MFMA, FP32 accumulation */
  OCCXC_PREC_FP32_CACHE = 3  /* FP32 AO cache, FP64 density matrix, FP64 contractions */
} occxc_precision;

typedef struct {
  int nshell;
  const int *l;              /* [nshell] angular momentum, 0..4 */
  const int *pure;           /* [nshell] 1 spherical, 0 Cartesian */
  const int *nprim;          /* [nshell] */
  const double *origin;      /* [3*nshell], Bohr */
  const double *exponents;   /* [sum nprim] */
  const double *coefficients;/* [sum nprim], normalised (see above) */
} occxc_basis;

typedef struct {
  long long npts;
  const double *points;      /* [3*npts] x,y,z interleaved, Bohr */
  const double *weights;     /* [npts] */
} occxc_grid;

typedef struct {
  int nfunc;
  const int *ids;            /* libxc ids, summed like Result::operator+= */
} occxc_functional;

typedef struct {
  int precision;             /* occxc_precision */
  double density_threshold;  /* leaf skip, DFT::m_density_threshold = 1e-10 (dft.h:760) */
  double screening_threshold;/* shell decay cutoff, m_xc_screening_threshold = 1e-10 (dft.h:761) */
  int screen;                /* 1: screened leaf path (dft.h:279-372); 0: dense path (OCC_XC_NO_SCREEN) */
  int device;                /* HIP device ordinal */
  /* launch tuning of the FP64 / FP32-cache contractions (no effect on the
   * results beyond FP64 rounding); the defaults take the library's tuning
   * table entry for the context's largest molecule */
  int tile_ao;               /* largest AO block per work item: 32..128, a multiple of 32; 0: table */
  int persistent;            /* bit 0: persistent density launch, bit 1: persistent Vxc launch; -1: table */
} occxc_options;

typedef struct occxc_context occxc_context;

typedef struct {
  long long npts;            /* grid points */
  long long nleaves;         /* leaves with >= 1 significant shell */
  int nbf;
  int nshell;
  double mean_local_nbf;     /* mean m_b over active leaves */
  double contraction_flops;  /* sum_b 4 n_b m_b^2 (screened work definition) */
  double ao_bytes;           /* bytes of the resident AO values */
  double ao_write_bytes;     /* algorithmic bytes of one AO evaluation */
  double device_bytes;       /* total device footprint of the context */
  int max_local_nbf;         /* largest m_b over the active leaves */
  int tile_ao;               /* launch settings in use (occxc_options) */
  int persistent;
  int tuning_row;            /* the tuning table entry they came from (-1: options) */
} occxc_context_info;

/* --- the XC build ---------------------------------------------------- */

void occxc_default_options(occxc_options *opt);

occxc_status occxc_context_create(const occxc_basis *basis,
                                  const occxc_grid *grid,
                                  const occxc_functional *func,
                                  const occxc_options *opt,
                                  occxc_context **ctx);
occxc_status occxc_context_destroy(occxc_context *ctx);
occxc_status occxc_context_get_info(const occxc_context *ctx,
                                    occxc_context_info *info);

/* One XC build from host buffers: D (nbf*nbf) in; Vxc (nbf*nbf), Exc, N_elec
 * out.  Mirrors compute_vxc + exchange_correlation_energy. */
occxc_status occxc_build(occxc_context *ctx, const double *D, double *Vxc,
                         double *exc, double *nelec);

/* compute_vxc for an unrestricted density: compute_K_dft<d, Unrestricted>
 * (include/occ/dft/dft.h:190-388, dispatched at dft.h:398-410) with
 * process_grid_block<d, Unrestricted> (src/dft/dft_kernels.cpp:8-70),
 * evaluate_density<d, U> (src/gto/density.cpp:18-30,45-70) and
 * xc_potential_matrix<Unrestricted, d> (src/dft/xc_potential_matrix.cpp:
 * 61-109), with the polarised functionals.  D is 2nbf x nbf column-major:
 * the alpha block in rows 0..nbf-1, the beta block in rows nbf..2nbf-1
 * (qm::block::a / b; OCC's density_matrix_unrestricted, orb.h:68-82, holds
 * C_s C_s^T / 2 per spin), each block read in full; Vxc is returned in the
 * same 2nbf x nbf layout; nelec[2] = {N_alpha, N_beta}, which OCC logs.
 * Meta-GGA ids add per-spin tau (density.cpp:130-166) and the per-spin tau
 * terms of xc_potential_matrix<Unrestricted, 2> (xc_potential_matrix.cpp:
 * 111-137).  FP64 and FP32-cache contexts of one molecule. */
occxc_status occxc_build_unrestricted(occxc_context *ctx, const double *D, double *Vxc,
                                      double *exc, double *nelec);

/* Same on device-resident buffers (d_D and d_Vxc 2nbf x nbf as above,
 * d_result = {Exc, N_alpha, N_beta}), enqueued on `stream` (NULL: the
 * context's own) without host synchronisation. */
occxc_status occxc_build_unrestricted_device(occxc_context *ctx, const double *d_D,
                                             double *d_Vxc, double *d_result, void *stream);

/* Same build on device-resident buffers, enqueued on `stream` (a hipStream_t,
 * NULL for the context's own stream) without host synchronisation.
 * d_result receives {Exc, N_elec} (per molecule for a batch context, whose
 * D and Vxc are packed as occxc_batch_layout describes). */
occxc_status occxc_build_device(occxc_context *ctx, const double *d_D,
                                double *d_Vxc, double *d_result, void *stream);

/* XC contribution to the nuclear gradient, replacing
 * DFT::compute_xc_gradient (include/occ/dft/dft.h:531-532) for a restricted
 * density: compute_xc_gradient_impl<d, Restricted> (dft.h:598-749) with
 * kernels::process_grid_block_gradient (src/dft/dft_gradient_kernels.cpp:
 * 473-552).  D as occxc_build (nbf*nbf column-major, no factor 2; symmetric:
 * the reference's density reads its upper triangle and its distribution both,
 * which agree only for a symmetric D).  shell_atom[nshell] is the atom of
 * each shell (AOBasis::shell_to_atom), natoms the atom count; grad receives
 * the 3 x natoms column-major Mat3N (x, y, z of atom 0, then atom 1, ...).
 * The gradient is that of Exc with the grid and weights held fixed, as in the
 * reference.  The first call evaluates and keeps the AO derivative components
 * the build does not need (the Hessian for GGA and meta-GGA, the gradient for
 * LDA).  Meta-GGA ids add the tau terms of accumulate_gradient_contribution_R
 * <2> (dft_gradient_kernels.cpp:95-172).  Single-molecule contexts only;
 * repeats bit for bit. */
occxc_status occxc_xc_gradient(occxc_context *ctx, const double *D, int natoms,
                               const int *shell_atom, double *grad);

/* Same gradient on device-resident buffers (d_D nbf*nbf, d_grad 3*natoms),
 * enqueued on `stream` (NULL: the context's own stream); shell_atom stays a
 * host array.  The first call for a given shell_atom map waits for its
 * set-up; later calls do not synchronise the host.  D is not checked for
 * non-finite entries on this path. */
occxc_status occxc_xc_gradient_device(occxc_context *ctx, const double *d_D, int natoms,
                                      const int *shell_atom, double *d_grad, void *stream);

/* The XC gradient for an unrestricted density: compute_xc_gradient_impl<d,
 * Unrestricted> (dft.h:598-749) with process_grid_block_gradient<d,
 * Unrestricted> (src/dft/dft_gradient_kernels.cpp:473-583) and the polarised
 * functionals.  D as occxc_build_unrestricted (2nbf x nbf, alpha block over
 * beta block), the rest as occxc_xc_gradient.  Meta-GGA ids add each spin's
 * tau terms (accumulate_gradient_contribution_U<2>, dft_gradient_kernels.cpp:
 * 306-470).  FP64 and FP32-cache contexts of one molecule. */
occxc_status occxc_xc_gradient_unrestricted(occxc_context *ctx, const double *D, int natoms,
                                            const int *shell_atom, double *grad);

/* Same unrestricted gradient on device-resident buffers (d_D 2nbf*nbf as
 * above, d_grad 3*natoms), enqueued on `stream` (NULL: the context's own
 * stream); shell_atom stays a host array.  Mirrors occxc_xc_gradient_device;
 * D is not checked for non-finite entries on this path. */
occxc_status occxc_xc_gradient_unrestricted_device(occxc_context *ctx, const double *d_D,
                                                   int natoms, const int *shell_atom,
                                                   double *d_grad, void *stream);

/* --- batches: many molecules in one context, one launch per stage ----- */

/* A context over nmol molecules (each its own basis and grid; shared
 * functional and options).  Leaves of all molecules are processed by the same
 * kernel launches. */
occxc_status occxc_batch_create(int nmol, const occxc_basis *bases,
                                const occxc_grid *grids,
                                const occxc_functional *func,
                                const occxc_options *opt, occxc_context **ctx);

/* Host build of every molecule: D[m], Vxc[m] are nbf_m x nbf_m column-major. */
occxc_status occxc_batch_build(occxc_context *ctx, const double *const *D,
                               double *const *Vxc, double *exc, double *nelec);

/* Device bytes each molecule adds to a batch context (its AO values,
 * matrices, per-point data and leaf tables): bytes[m] for m < nmol, from the
 * same hierarchy and screening as occxc_batch_create but without the large
 * allocations.  For packing many molecules into contexts that fit device
 * memory; molecules are set up on several host threads. */
occxc_status occxc_batch_estimate_bytes(int nmol, const occxc_basis *bases,
                                        const occxc_grid *grids, const occxc_functional *func,
                                        const occxc_options *opt, double *bytes);

/* Molecule count, per-molecule nbf and the element offsets of each
 * molecule's nbf x nbf block in the packed device layout used by
 * occxc_build_device (D and Vxc packed back to back; d_result holds
 * {Exc, N_elec} per molecule). nbf / offsets may be NULL. */
occxc_status occxc_batch_layout(const occxc_context *ctx, int *nmol, int *nbf,
                                long long *offsets);

/* Wait for work enqueued on the context's stream. */
occxc_status occxc_synchronize(occxc_context *ctx);

/* Re-evaluate the resident AO values of the context (done once at creation;
 * a build that must not rely on residency calls this first).  Enqueued on
 * the context's stream. */
occxc_status occxc_context_evaluate_ao(occxc_context *ctx);

/* Per-stage device timing, the counterpart of OCC's occ::timing categories
 * dft_xc / gto (include/occ/core/timings.h:11-133): when enabled, every build
 * records HIP events around its stages on the stream it runs on.
 * occxc_context_get_timing waits for them and returns the summed
 * milliseconds and the number of timed launches per stage. */
enum {
  OCCXC_STAGE_PREPARE = 0,   /* D2 = 2 sym(D), K = 0 */
  OCCXC_STAGE_DENSITY = 1,   /* rho, grad rho, functional, coefficients */
  OCCXC_STAGE_VXC = 2,       /* phi^T A per leaf into K */
  OCCXC_STAGE_FINALIZE = 3,  /* K + K^T, Exc / N_elec reduction */
  OCCXC_STAGE_AO = 4,        /* AO values of every leaf */
  OCCXC_STAGE_GRADIENT = 5,  /* XC gradient contraction and per-atom reduction */
  OCCXC_STAGE_NLC = 6,       /* VV10 pair sums and coefficients */
  OCCXC_NSTAGES = 7
};
occxc_status occxc_context_set_timing(occxc_context *ctx, int enable);
occxc_status occxc_context_get_timing(occxc_context *ctx, double *ms,
                                      long long *counts);

/* Replay builds from a captured HIP graph.  When enabled, the first
 * occxc_build_device (or occxc_batch_build) on a given (D, Vxc, result,
 * stream) captures the build's kernel sequence once; later builds on the same
 * pointers and stream launch the graph (one launch instead of one per
 * kernel).  Other pointers re-capture.  Stage timing is not recorded for
 * graph launches.  The graph holds no memset nodes: every buffer the build
 * clears is cleared by a kernel.  Replaces OCC's per-call loop set-up in
 * compute_K_dft (include/occ/dft/dft.h:190-388), which the reference redoes on
 * every SCF iteration. */
occxc_status occxc_context_set_graph(occxc_context *ctx, int enable);

/* --- serving: XC builds from many concurrent SCFs on one GPU ---
 *
 * No counterpart in OCC, which runs one SCF per process and calls
 * DFT::compute_vxc once per iteration (scf_impl.h:539, dft.h:395-429).  A
 * server keeps each admitted SCF's context (grid leaves, AO values) resident
 * between its builds, admits SCFs by device-memory footprint against a budget
 * (evicting idle ones, least recently used first, when the next SCF does not
 * fit), and runs builds of different SCFs concurrently on `slots` HIP
 * streams; builds of one SCF run in submission order.  Every function is
 * thread-safe. */
typedef struct occxc_server occxc_server;

typedef struct {
  int device;
  int slots;             /* builds in flight at once, 1..64 */
  double memory_budget;  /* device bytes for resident SCFs; 0: 90% of free memory */
} occxc_server_options;

typedef struct {
  long long scfs_opened, scfs_admitted, scfs_closed, scfs_resident, scfs_waiting;
  long long admissions_deferred, evictions;
  long long builds_submitted, builds_completed, builds_failed;
  double resident_bytes, budget_bytes;
  long long builds_cancelled;
} occxc_server_stats;

void occxc_server_default_options(occxc_server_options *opt);
occxc_status occxc_server_create(const occxc_server_options *opt, occxc_server **srv);
/* Cancels every build that has not started (its ticket completes with
 * OCCXC_ERR_CANCELLED), lets the running ones finish, frees every SCF, then
 * stops the workers. */
occxc_status occxc_server_destroy(occxc_server *srv);
/* Registers an SCF (inputs copied; opt->device is the server's). */
occxc_status occxc_server_open(occxc_server *srv, const occxc_basis *basis,
                               const occxc_grid *grid, const occxc_functional *func,
                               const occxc_options *opt, long long *scf);
/* Queues one build: D (nbf x nbf, column-major, as occxc_build) is copied at
 * once; Vxc, exc and nelec are written before the ticket completes. */
occxc_status occxc_server_submit(occxc_server *srv, long long scf, const double *D,
                                 double *Vxc, double *exc, double *nelec, long long *ticket);
/* Waits for a ticket; returns the build's status and its latency from submit
 * to completion.  A ticket is waited for once. */
occxc_status occxc_server_wait(occxc_server *srv, long long ticket, double *latency_ms);
occxc_status occxc_server_poll(occxc_server *srv, long long ticket, int *done);
/* Cancels a build that has not started: its ticket completes at once with
 * OCCXC_ERR_CANCELLED and *cancelled is 1.  A build that is running or done
 * is left alone (*cancelled = 0; its ticket reports the build's own result).
 * No counterpart in the reference (one SCF per process, no request queue). */
occxc_status occxc_server_cancel(occxc_server *srv, long long ticket, int *cancelled);
/* Closes an SCF: its builds that have not started are cancelled
 * (OCCXC_ERR_CANCELLED), a running one finishes, then its context is freed. */
occxc_status occxc_server_close(occxc_server *srv, long long scf);
occxc_status occxc_server_get_stats(occxc_server *srv, occxc_server_stats *stats);

/* --- VV10 non-local correlation ------------------------------------- */

/* NonLocalCorrelationFunctional::vv10 (src/dft/nonlocal_correlation.cpp:
 * 246-306) with vv10_kernel (nonlocal_correlation.cpp:37-103): the VV10
 * energy and its Vxc (nbf x nbf) for a restricted density D (nbf x nbf, no
 * factor 2), on the context's grid.  Build the context from the NLC grid
 * (set_integration_grid, nonlocal_correlation.cpp:17-31) with any GGA or
 * meta-GGA functional id, which only selects the cached AO derivatives.  As
 * in the reference, each point's sums run over the points before it in the
 * grid's input order, so Vxc depends on that order and the energy does not.
 * b, C: Parameters::vv10_b, vv10_C (nonlocal_correlation.h:14-18; 6.0,
 * 0.01).  Single-molecule FP64 or FP32-cache contexts. */
occxc_status occxc_nlc_vv10(occxc_context *ctx, const double *D, double b, double C,
                            double *energy, double *Vxc);
/* The same with device buffers: d_D (nbf x nbf) in, d_Vxc (nbf x nbf) and
 * d_result[2] = {E_nlc, N_elec} out, on the given stream (NULL: the
 * context's). */
occxc_status occxc_nlc_vv10_device(occxc_context *ctx, const double *d_D, double b,
                                   double C, double *d_Vxc, double *d_result, void *stream);

/* --- stage-level entry points (each mirrors one reference function) --- */

/* gto::evaluate_basis (src/gto/gto.cpp:69-140): phi on npts points for the
 * whole basis, times common_fac.  out is [ncomp][nbf][npts] (point fastest),
 * ncomp = 1 (deriv 0) or 4 (deriv 1: phi, phi_x, phi_y, phi_z). */
occxc_status occxc_evaluate_basis(const occxc_basis *basis, long long npts,
                                  const double *points, int deriv,
                                  double *out, int device);

/* gto::evaluate_decay_cutoff (src/gto/gto.cpp:18-67). */
occxc_status occxc_evaluate_decay_cutoff(const occxc_basis *basis,
                                         double threshold, double *extents,
                                         int device);

/* DensityFunctional::evaluate (src/dft/functional.cpp:18-65), unpolarised:
 * ACCUMULATES into exc/vrho/vsigma like Result::operator+= (functional.h:
 * 436-457).  sigma and vsigma may be NULL for LDA ids. */
occxc_status occxc_functional_evaluate(int id, long long n, const double *rho,
                                       const double *sigma, double *exc,
                                       double *vrho, double *vsigma,
                                       int device);

/* DensityFunctional::evaluate for Family::MGGA (src/dft/functional.cpp:
 * 49-57), unpolarised: rho, sigma, tau = (1/2) sum_i |grad psi_i|^2 per point
 * (density.cpp:94-101); ACCUMULATES exc, vrho, vsigma, vtau (vlaplacian is
 * zero for every id here).  LDA and GGA ids are accepted (tau unused; sigma /
 * vsigma may be NULL for LDA, tau / vtau for LDA and GGA). */
occxc_status occxc_functional_evaluate_mgga(int id, long long n, const double *rho,
                                            const double *sigma, const double *tau,
                                            double *exc, double *vrho, double *vsigma,
                                            double *vtau, int device);

/* DensityFunctional::evaluate (src/dft/functional.cpp:49-57) with
 * Family::MGGA on a spin-polarised density, libxc's polarised layout:
 * rho[2n] = (rho_a, rho_b), sigma[3n] = (s_aa, s_ab, s_bb), tau[2n] =
 * (tau_a, tau_b) per point (dft_kernels.h:51-57); ACCUMULATES exc[n] (per
 * particle), vrho[2n], vsigma[3n], vtau[2n] (vlaplacian is zero for every
 * functional here).  LDA and GGA ids are accepted (tau / sigma ignored as
 * their family allows).  Runs on GPU `device`. */
occxc_status occxc_functional_evaluate_mgga_polarized(int id, long long n, const double *rho,
                                                      const double *sigma, const double *tau,
                                                      double *exc, double *vrho, double *vsigma,
                                                      double *vtau, int device);

/* The same for a spin-polarised density (DensityFunctional with
 * polarized = true, functional.cpp:18-65), libxc's layout: rho[2n] =
 * (rho_a, rho_b) per point, sigma[3n] = (s_aa, s_ab, s_bb); ACCUMULATES
 * exc[n], vrho[2n], vsigma[3n]. */
occxc_status occxc_functional_evaluate_polarized(int id, long long n, const double *rho,
                                                 const double *sigma, double *exc,
                                                 double *vrho, double *vsigma, int device);

/* SpatialGridHierarchy (src/numint/spatial_grid_hierarchy.cpp:28-147):
 * Morton sort + 128-point leaves.  Returns the number of leaves in *nleaves;
 * leaf arrays must hold npts/32 + 2 entries. */
occxc_status occxc_grid_hierarchy(long long npts, const double *points,
                                  const double *weights, double *sorted_points,
                                  double *sorted_weights, long long *perm,
                                  long long *leaf_offset, long long *leaf_count,
                                  double *leaf_center, double *leaf_radius,
                                  long long *nleaves);

/* --- molecular grid construction (src/numint) ------------------------- */

typedef enum {
  OCCXC_RADIAL_EULER_MACLAURIN = 0,
  OCCXC_RADIAL_TREUTLER_AHLRICHS = 1,
  OCCXC_RADIAL_BECKE = 2,
  OCCXC_RADIAL_GAUSS_CHEBYSHEV = 3,
  OCCXC_RADIAL_MURA_KNOWLES = 4
} occxc_radial_method;

/* Raw radial rules of grid_utils.cpp (generate_*_radial_grid, :337-572):
 * n points requested, `param` = alpha (EM), rm (Becke) or charge (MK).
 * *npoints_out receives the number written (EM returns n-1, :382). */
occxc_status occxc_radial_grid(int method, int n, double param, double *r,
                               double *w, int *npoints_out);

/* generate_lmg_radial_grid (src/numint/grid_utils.cpp:397-515), the
 * default radial rule of MolecularGrid (include/occ/numint/molecular_grid.h:
 * 128): alpha_max = the atom's largest exponent, alpha_min[0..l_max] its
 * smallest exponent per angular momentum (DBL_MAX where it has no shell of
 * that l, as MolecularGrid::initialize_from_basis leaves it,
 * molecular_grid.cpp:207-228; l_max = the basis' largest l), atomic_radius the
 * Bragg radius in Bohr.  Radial weights include r^2 but not 4 pi.  Writes
 * *npoints_out; fills r, w only when capacity >= that (size query first). */
occxc_status occxc_lmg_radial_grid(double precision, double alpha_max, int l_max,
                                   const double *alpha_min, double atomic_radius,
                                   int capacity, double *r, double *w, int *npoints_out);

/* Becke partition with Treutler-Ahlrichs radii adjustment
 * (calculate_atomic_grid_weights, grid_utils.cpp:125-232) and the per-atom
 * normalisation of MolecularGrid::get_partitioned_atom_grid
 * (molecular_grid.cpp:46-82): weights[i] *= P_{atom[i]} / sum_k P_k.
 * radii are the Bragg radii in Bohr (grid_utils.cpp:22-41). */
occxc_status occxc_becke_partition(int natoms, const double *atom_xyz,
                                   const double *radii, int ta_adjustment,
                                   long long npts, const double *points,
                                   const int *atom_of_point, double *weights,
                                   int device);

/* Allocation-free form of occxc_becke_partition, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Every
 * array is a caller-owned DEVICE buffer on the current device: atom_xyz
 * [3 natoms], radii [natoms], points [3 npts], atom_of_point [npts] and
 * weights [npts] (scaled in place); work is queued on stream (null: the
 * default stream) and the call returns without waiting.  The inter-atomic
 * distance table lives in a fixed device array of 1024 x 1024 doubles
 * (8 MiB per GPU, held until the process ends): natoms above 1024 is
 * rejected with OCCXC_ERR_INVALID_ARGUMENT.  Inputs are not inspected on
 * the host (coincident atoms, which occxc_becke_partition rejects, give
 * non-finite weights; an atom index out of range gives weight 0).  All calls
 * on one GPU share the table, so calls that overlap on one GPU race. */
occxc_status occxc_becke_partition_noalloc(int natoms, const double *atom_xyz,
                                           const double *radii, int ta_adjustment,
                                           long long npts, const double *points,
                                           const int *atom_of_point, double *weights,
                                           void *stream);

/* --- errors ------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char *occxc_last_error_message(void);
const char *occxc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* OCCXC_H */
