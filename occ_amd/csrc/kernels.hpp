// kernels.hpp -- device-side data layout shared by the host orchestration
// (api.cpp) and the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace occxc {

constexpr int kMaxL = 4;          // highest shell angular momentum supported
constexpr int kPtPad = 64;        // points of a leaf padded to a multiple of this
constexpr int kAoPad = 32;        // local AOs of a leaf padded to a multiple of this
constexpr int kMaxLeafPts = 256;  // SpatialHierarchySettings::max_leaf_size
constexpr int kMaxFunc = 8;
constexpr int kMaxLocalAo = 8192; // largest local AO count a leaf may have

// One shell, with common_fac(l) folded into its coefficients.
struct ShellDev {
  double x, y, z;
  int l, pure, nprim, poff, bf0, size;
};

// One active grid leaf (>= 1 significant shell).  Points of the leaf are
// stored at [pt_off, pt_off + np) of the padded point arrays; points
// [n, np) are copies of the leaf's first point with weight 0.
struct LeafDev {
  int pt_off;   // offset into the padded point arrays
  int n;        // real points
  int np;       // padded points (multiple of kPtPad)
  int m;        // significant local AOs
  int mp;       // padded local AOs (multiple of kAoPad)
  int sh_off;   // offset into leaf_shells
  int nsig;     // significant shells
  int map_off;  // offset into ao_map (mp entries; padding entries map to AO 0)
  int mol;      // molecule of this leaf (index into BuildArgs::mols)
  long long ao_off;  // element offset of this leaf's AO block [ncomp][mp][np]
  long long kl_off;  // element offset of this leaf's K_loc block (two-pass Vxc)
  long long res_off; // element offset of this leaf's value-plane residuals [mp][np] (FP32 storage)
};

// Two-pass Vxc: one global 32 x 32 AO tile pair (TI <= TJ) of a molecule and
// the leaves whose AOs touch both tiles.
struct TilePairDev {
  long long list_off;
  int nlist;
  int mol;
  int ti, tj;
  int chunk_off;  // its chunks of at most kTileChunk list entries
  int nchunk;
};

// A run of at most kTileChunk entries of one tile pair's list, summed by one
// workgroup into a 32 x 32 partial block.
struct TileChunkDev {
  long long e0, e1;
  int pair;
  int pad_;
};
constexpr int kTileChunk = 64;

// One leaf of a tile pair's list.  The leaf's AOs in tile TI are a sorted
// subset of its 32 AOs: bit k of mask_a is set when AO 32 TI + k is present,
// and the present ones sit at local indices start_a, start_a + 1, ...
struct TileEntry {
  long long kl_off;
  int mp;
  int leaf;
  int start_a, start_b;
  unsigned mask_a, mask_b;
};

// Tiled contractions: a leaf's padded AOs are cut into blocks of at most
// kTileAo (multiples of 32) and its padded points into blocks of at most
// kTileAo.  A density item computes T = Phi D_loc for one (point block,
// column block) pair, contracting over all of the leaf's AOs, and the partial
// rho / grad rho of its columns; a Vxc item computes one block pair (r <= c)
// of V_loc.  Items with leaf < 0 are padding (the lists are interleaved over
// the 8 XCDs).
constexpr int kTileAo = 128;

struct DenItem {
  int leaf;
  int part;        // partial slot [4][kTileAo] (-1: the leaf's only item, functional inline)
  short p0, np;    // point block
  short c0, wc;    // column block (local AOs)
};

// A leaf with more than one density item: its partial slots are
// part0 + pb * ncb + cb (point block pb, column block cb), summed in that
// order before the functional.
struct DenFinish {
  int leaf;
  int part0;
  int ncb;
  int npb;
};

struct VxcItem {
  int leaf;
  short r0, wr;    // row block (local AOs)
  short c0, wc;    // column block (c0 >= r0)
  int pad_;
};

// One molecule of a context: its nbf x nbf blocks (D2, K, Vxc) start at moff
// in the concatenated per-context arrays; its active leaves are
// [leaf_begin, leaf_end).
struct MolDev {
  long long moff;
  int nbf;
  int leaf_begin;
  int leaf_end;
  int pad_;
};

// Device set-up: for one active leaf, its leaf in the all-leaves sphere
// arrays, its molecule's first shell and shell count, the offset of its
// first sorted point in the permutation array and of its molecule's points.
struct FillLeaf {
  int src_leaf;
  int sb;
  int nsh;
  int pad_;
  long long src_pt;
  long long mol_pt;
};

struct BuildArgs {
  int nmol;
  int maxnbf;     // largest nbf over the molecules
  long long mtot; // sum of nbf^2 over the molecules
  int nleaves;
  int deriv;
  int ncomp;
  int nfunc;
  int lmax;       // largest shell angular momentum
  int func_ids[kMaxFunc];
  double dthr;
  long long ptot;  // total padded points
  const LeafDev *leaves;
  const MolDev *mols;
  const ShellDev *shells;
  const int *leaf_shells;
  const int *ao_map;
  const double *exps;
  const double *coefs;
  const double *px, *py, *pz, *pw;
  const double *D2;     // doubled, symmetrised densities (concatenated nbf x nbf)
  double *K;            // accumulated phi^T A (concatenated nbf x nbf)
  double *coef;         // per-point Vxc coefficients [4][ptot] (the density stage writes 4
                        // planes; the Vxc stage reads 4 for a GGA, 1 for an LDA)
  double *leaf_part;    // [2*nleaves] {Exc, N_elec} per leaf
  int *leaf_active;     // [nleaves] 0 when the leaf was skipped (max rho < dthr)
  // FP32 AO storage, during context set-up only: the value plane's rounding
  // residuals r = exact value - stored float of the leaves [res_leaf0,
  // res_leaf1), [mp][np] per leaf at res_off - res_base (null otherwise)
  float *ao_res;
  int res_leaf0, res_leaf1;
  long long res_base;
  // FP32 mode with the A pass: A = sum_q c_q Phi_q of every active leaf,
  // [mp][np] per leaf at res_off (null otherwise)
  float *ao_a;
  double *kloc;         // two-pass Vxc: per-leaf K_loc blocks (null: atomics into K)
  const TilePairDev *tpairs;
  int ntpairs;
  const TileEntry *tentries;
  const TileChunkDev *tchunks;
  int ntchunks;
  double *tpartial;     // [ntchunks][32 * 32] partial blocks
  double *rho_out;      // unrestricted: [4][ptot] rho, grad rho of one spin (density store)
  const double *coef_tau;  // restricted meta-GGA: [ptot] tau coefficients (null otherwise)
  int coef_planes;      // planes of ptot doubles that coef holds
  int coef_tau_planes;  // planes that coef_tau holds (0 without one)
};

// The launch wrappers check, before any launch, that the coefficient buffers
// hold every plane their kernels read or write; a mismatch throws an
// OCCXC_ERR_INTERNAL error.
void check_coef_planes(int need, int have, const char *what);

// device set-up of the screened layout (setup.hip)
int hierarchy_leaf_count(long long n);
size_t sort_temp_bytes(long long nmax);
void launch_hierarchy(const double *xyz, long long n, int nleaves, double *part, double *box,
                      unsigned long long *keys_in, unsigned long long *keys_out, int *idx_in,
                      int *perm, void *temp, size_t temp_bytes, double *center, double *radius,
                      hipStream_t s);
void launch_screen_count(const double *center, const double *radius, int nleaves,
                         const ShellDev *shells, const double *ext, int nsh, int *nsig, int *nao,
                         hipStream_t s);
void launch_fill_leaves(const LeafDev *leaves, const FillLeaf *fill, int nactive,
                        const double *center, const double *radius, const ShellDev *shells,
                        const double *ext, const double *xyz, const double *w, const int *perm,
                        int *leaf_shells, int *ao_map, double *px, double *py, double *pz,
                        double *pw, int *porig, hipStream_t s);

// host-callable launchers (kernels.hip)
void upload_sph_table();
void launch_zero(double *p, long long n, hipStream_t s);
void launch_fill_int(int *p, int n, int v, hipStream_t s);
// out = K + K^T per molecule
void launch_symmetrize(const double *K, double *out, const MolDev *mols, int nmol, int maxnbf,
                       hipStream_t s);
// FP32 AO storage: result[2 m + 1] (N_elec of molecule m) += sum D2 (.) dS,
// the first-order change of sum_p w_p rho_p from the stored AO values to the
// exact ones; partial holds kNcorrBlocks x nmol doubles
constexpr int kNcorrBlocks = 1024;
// residual floats formed at a time at context set-up (1 GiB; a larger leaf
// block is taken whole)
constexpr size_t kResidualChunkElems = (size_t)1 << 28;
void launch_storage_correction(const double *D2, const double *dS, const MolDev *mols, int nmol,
                               int maxnbf, double *partial, double *result, hipStream_t s);
void launch_prepare_density(const double *D, double *D2, const MolDev *mols, int nmol,
                            int maxnbf, hipStream_t s);
void launch_ao_eval(const BuildArgs &a, void *ao, bool fp32, hipStream_t s);
// FP32 storage, context set-up: the value plane's rounding residuals of the
// leaves [a.res_leaf0, a.res_leaf1) into a.ao_res (the stored values are not
// rewritten)
void launch_ao_residuals(const BuildArgs &a, const void *ao, hipStream_t s);
// XC gradient: the AO derivative components the cache lacks (FP64): the
// Hessian (6 components) for a GGA context, the gradient (3) for an LDA one.
void launch_ao_extra(const BuildArgs &a, double *out, bool hessian, hipStream_t s);
// XC gradient contraction (grad.hip): per leaf and local AO mu the sum over
// the leaf's points of the gradient integrand into gpart[3 (map_off + mu) + k],
// then grad[3 A + k] = -2 sum over the entries of atom A (CSR atom_off /
// atom_ent), in a fixed order.
// the same in two steps: per-AO partials of one spin density (a.D2 and
// a.coef of that spin), then the per-atom reduction of one or two partial sets
void launch_xc_gradient_partials(const BuildArgs &a, const void *ao, bool ao_fp32,
                                 const double *ao2, double *gpart, hipStream_t s);
// restricted meta-GGA: the tau terms' per-AO partials (a.coef_tau, Hessian in
// ao2) into gpart, reduced together with k_xc_grad's by launch_grad_reduce
void launch_xc_gradient_tau(const BuildArgs &a, const void *ao, bool ao_fp32,
                            const double *ao2, double *gpart, hipStream_t s);
// grad = -2 sum of up to four partial sets (gpart2, gpart3/gpart4 optional;
// gpart3 and gpart4 come together)
void launch_grad_reduce(const double *gpart, const double *gpart2, int natoms,
                        const int *atom_off, const int *atom_ent, double *grad, hipStream_t s,
                        const double *gpart3 = nullptr, const double *gpart4 = nullptr);
void launch_xc_gradient(const BuildArgs &a, const void *ao, bool ao_fp32, const double *ao2,
                        double *gpart, int natoms, const int *atom_off, const int *atom_ent,
                        double *grad, hipStream_t s);
void launch_density_functional(const BuildArgs &a, const void *ao, bool fp32,
                               hipStream_t s);
void launch_vxc(const BuildArgs &a, const void *ao, bool fp32, hipStream_t s);
// unrestricted build (one molecule): D2t_s = 2 D_s^T; the polarised
// functional stage per leaf (coef_a in a.coef, {Exc, N_a, N_b} per leaf in
// part); V_s = K_s + K_s^T (skipped when Ka is null) and the fixed-order
// scalar sums into result[3]
void launch_prepare_density_u(const double *D, double *D2a, double *D2b, int nbf,
                              hipStream_t s);
void launch_functional_pol_leaves(const BuildArgs &a, const double *rho_a, const double *rho_b,
                                  double *coef_b, double *part, hipStream_t s);
void launch_finalize_u(const double *Ka, const double *Kb, double *Va, double *Vb,
                       const MolDev *mols, int maxnbf, const double *part, int nleaves,
                       double *result, hipStream_t s);
void launch_functional_pol(int id, long long n, const double *rho, const double *sigma,
                           double *exc, double *vrho, double *vsigma, hipStream_t s);
// polarised meta-GGA over points (ADDING; rho[2n], sigma[3n], tau[2n])
void launch_functional_mgga_pol(int id, long long n, const double *rho, const double *sigma,
                                const double *tau, double *exc, double *vrho, double *vsigma,
                                double *vtau, hipStream_t s);
// restricted meta-GGA: the per-leaf functional stage from the stored rho /
// grad rho (rho4, [4][ptot]) and per-direction tau passes (tk, [3][ptot]);
// coefficients into a.coef and coef_tau, {Exc, N_elec} into a.leaf_part
void launch_functional_mgga_leaves(const BuildArgs &a, const double *rho4, const double *tk,
                                   double *coef_tau, hipStream_t s);
// unrestricted: both spins' coefficients, c_tau per spin, {Exc, N_a, N_b}
// per leaf into part
void launch_functional_mgga_pol_leaves(const BuildArgs &a, const double *rho_a,
                                       const double *rho_b, const double *tka, const double *tkb,
                                       double *coef_b, double *ctau_a, double *ctau_b,
                                       double *part, hipStream_t s);
void launch_functional_mgga(int id, long long n, const double *rho, const double *sigma,
                            const double *tau, double *exc, double *vrho, double *vsigma,
                            double *vtau, hipStream_t s);
// VV10 (nlc.hip): from rho / grad rho of the leaf points (rho4, [4][ptot])
// and each padded point's input index (porig), the weighted VV10 exc / vrho /
// vsigma as GGA Vxc coefficients in a.coef and {E_nlc, N_elec} per leaf in
// a.leaf_part (every leaf active); nlc holds nlc_buffer_elems(npts) doubles
size_t nlc_buffer_elems(long long npts);
void launch_vv10(const BuildArgs &a, const double *rho4, const int *porig, double b, double C,
                 double *nlc, long long npts, hipStream_t s);
void launch_finalize(const double *K, double *Vxc, const MolDev *mols, int nmol,
                     int maxnbf, const double *leaf_part, double *result, hipStream_t s);
// {Exc, N_elec} per molecule only (the two-pass Vxc writes Vxc itself)
void launch_finalize_scalars(const MolDev *mols, int nmol, const double *leaf_part,
                             double *result, hipStream_t s);
void launch_eval_basis_flat(const ShellDev *shells, int nshell,
                            const double *exps, const double *coefs,
                            const double *xyz, long long npts, int deriv,
                            int nbf, double *out, hipStream_t s);
void launch_decay_cutoff(const ShellDev *shells, int nshell,
                         const double *exps, const double *coefs,
                         double threshold, double *extents, hipStream_t s);
void launch_functional(int id, long long n, const double *rho,
                       const double *sigma, double *exc, double *vrho,
                       double *vsigma, hipStream_t s);
void launch_becke(int natoms, const double *atom_xyz, const double *radii,
                  const double *dist, int ta, long long npts,
                  const double *pts, const int *atom_of_point, double *w,
                  hipStream_t s);
// Becke partition on caller-owned device buffers, its inter-atomic distance
// table in a fixed device array (at most kNoAllocMaxAtoms atoms); allocates
// nothing.  noalloc_scratch(device) resolves that array's address on a device
// (the current one) once.
constexpr int kNoAllocMaxAtoms = 1024;
double *noalloc_scratch(int device);
void becke_partition_noalloc(int natoms, const double *atom_xyz, const double *radii, int ta,
                             long long npts, const double *pts, const int *atom_of_point,
                             double *w, hipStream_t s);

} // namespace occxc
