// context.hpp -- the occxc context (opaque in the C ABI) and the internal
// entry points shared by the C ABI (api.cpp) and the serving loop
// (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "occxc.h"
#include "common.hpp"
#include "contract.hpp"
#include "kernels.hpp"

struct occxc_context {
  int device = 0;
  hipStream_t stream = nullptr;
  occxc_options opt{};
  int nbf = 0, nsh = 0, deriv = 0, ncomp = 1, nfunc = 0;
  int nmol = 0, maxnbf = 0;
  long long mtot = 0;
  std::vector<int> mol_nbf;
  std::vector<long long> mol_off;
  int funcs[occxc::kMaxFunc] = {0};
  long long npts = 0, ptot = 0;
  int nleaves = 0;
  double mean_m = 0.0, flops = 0.0, ao_write_bytes = 0.0;
  size_t ao_elems = 0, ao_elem_size = 8, res_elems = 0;
  std::vector<occxc::LeafDev> leaves;
  occxc::DevBuf<occxc::ShellDev> d_shells;
  occxc::DevBuf<double> d_exps, d_coefs;
  occxc::DevBuf<occxc::LeafDev> d_leaves;
  occxc::DevBuf<occxc::MolDev> d_mols;
  occxc::DevBuf<int> d_leaf_shells, d_ao_map, d_active;
  occxc::DevBuf<double> d_px, d_py, d_pz, d_pw;
  occxc::DevBuf<int> d_porig;  // input index (within its molecule) of each padded point, -1 for padding
  occxc::DevBuf<char> d_ao;
  // FP32 AO storage: the value plane's rounding residuals (set-up only) and
  // dS = sum_p w_p (r Phi^T + Phi r^T), the first-order change of the grid
  // overlap from the stored values to the exact ones
  occxc::DevBuf<float> d_ao_res;
  occxc::DevBuf<float> d_ao_a;  // FP32 mode: the Vxc stage's A plane (A pass)
  occxc::DevBuf<double> d_dS, d_ncorr;
  occxc::DevBuf<double> d_D, d_D2, d_K, d_V, d_coef, d_part, d_result;
  occxc::ContractPlan plan;
  // two-pass Vxc: per-leaf K_loc blocks and the tile-pair leaf lists
  occxc::DevBuf<double> d_kloc;
  occxc::DevBuf<occxc::TilePairDev> d_tpairs;
  occxc::DevBuf<occxc::TileEntry> d_tentries;
  occxc::DevBuf<occxc::TileChunkDev> d_tchunks;
  occxc::DevBuf<double> d_tpartial;
  long long kloc_elems = 0;
  occxc::BuildArgs args{};
  // stage timing: pending event sets, summed on read-out
  bool timing = false;
  std::vector<std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>>> pending;
  double stage_ms[OCCXC_NSTAGES] = {0};
  long long stage_n[OCCXC_NSTAGES] = {0};
  bool ao_ready = false;
  size_t device_bytes = 0;
  std::vector<double> mol_bytes;  // device bytes each molecule adds
  // captured build graph and the (D, Vxc, result, stream) it was captured on
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  const void *graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
  // XC gradient (single-molecule contexts): shell of each basis function, the
  // extra AO derivative components (evaluated on the first gradient call),
  // per-AO partial sums, and the per-atom entry lists built for the
  // shell-to-atom map of the last call
  std::vector<int> bf_shell;
  occxc::DevBuf<double> d_ao2, d_gpart, d_gpart2, d_grad;
  // unrestricted meta-GGA gradient: each spin's tau partials
  occxc::DevBuf<double> d_gpart3, d_gpart4;
  occxc::DevBuf<int> d_gatom_off, d_gatom_ent;
  bool ao2_ready = false;
  std::vector<int> grad_shell_atom;
  int grad_natoms = 0;
  // unrestricted builds (allocated on the first one): the raw 2 nbf x nbf
  // density, 2 D_s^T per spin, rho and grad rho per spin, the beta-spin
  // coefficients and K / Vxc (alpha uses the restricted buffers), per-leaf
  // {Exc, N_alpha, N_beta} and the result {Exc, N_alpha, N_beta}
  occxc::DevBuf<double> d_Du, d_D2a, d_D2b, d_rho_a, d_rho_b, d_coef_b, d_K_b, d_V_b, d_part_u,
      d_result_u;
  bool uks_ready = false;
  // restricted meta-GGA (allocated at context creation): rho and grad rho
  // ([4][ptot]), the per-direction tau passes ([3][ptot]), the tau
  // coefficients, and three copies of the leaf table whose AO block offsets
  // point at the d/dx, d/dy, d/dz planes of the AO cache
  bool mgga = false;
  occxc::DevBuf<double> d_mrho, d_mtau, d_coef_tau;
  // unrestricted meta-GGA: the beta spin's tau passes and c_tau (alpha uses
  // d_mtau / d_coef_tau)
  occxc::DevBuf<double> d_mtau_b, d_coef_tau_b;
  occxc::DevBuf<occxc::LeafDev> d_leaves_k;
  // VV10 (allocated on the first call): input-order point planes and F, U, W
  occxc::DevBuf<double> d_nlc;
};

namespace occxc {

// Sets up c for nmol molecules.  With byte_limit > 0 it throws
// OCCXC_ERR_OUT_OF_MEMORY before the large allocations when the context would
// need more device memory than that (c->device_bytes then holds the need).
void context_setup(occxc_context *c, int nmol, const occxc_basis *bases,
                   const occxc_grid *grids, const occxc_functional *func,
                   const occxc_options *opt, size_t byte_limit = 0);
// One build on stream s, directly or through the captured graph.
void launch_build(occxc_context *c, const double *d_D, double *d_V, double *d_result,
                  hipStream_t s);
// Waits for the context's work and frees it.
void context_destroy(occxc_context *c);

// Runs f, turning exceptions into a status and the thread's last error.
template <typename F> occxc_status guarded(F &&f) {
  try {
    clear_last_error();
    f();
    return OCCXC_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return OCCXC_ERR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return OCCXC_ERR_INTERNAL;
  }
}

} // namespace occxc
