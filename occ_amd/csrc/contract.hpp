// contract.hpp -- the density and Vxc contractions of the XC build, planned
// once per context and run per build.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace occxc {

// Launch settings of the tiled contractions (options or tuning table).
struct TuneChoice {
  int tile_ao = 128;     // largest AO block per work item (32..128, multiple of 32)
  int persistent = 0;    // bit 0 persistent density launch, bit 1 persistent Vxc launch
  int table_row = -1;    // the tuning table entry used (-1: set by the options)
};

// The options' settings where given, else the tuning table's entry for a
// context whose largest molecule has max_nbf basis functions.
TuneChoice tune_choice(const occxc_options *opt, int max_nbf);

struct ContractPlan {
  bool use_mfma = false;   // false: the FP64 reference kernels of kernels.hip
  int max_np = 0;          // largest padded leaf (points)
  int max_mp = 0;          // largest padded leaf (AOs)
  int density_variant = 2; // 0 LDS-staged Phi, 1 LDS-shared D tiles, 2 per-wavefront,
                           // 4 per-wavefront with per-chunk LDS reduction (FP64 data)
                           // (FP32 mode: 0/2 compensated density, 3 plain per-wavefront)
  int vxc_variant = 0;     // 0 symmetric tile pairs (small leaves), 1 general kernel,
                           // 2 as 0 but two-pass (stores + tile-pair reduction)
  bool two_pass = false;
  // tiled contractions (FP64 and FP32-cache contexts; OCCXC_TILED=0 turns
  // them off): item lists interleaved over the XCDs, the multi-item leaves
  // and their partial sums
  bool tiled = false;
  TuneChoice tune;         // the launch settings in use
  bool vxc_ws = false;     // persistent wavefront-specialised Vxc instead of one item per workgroup
  bool den_persistent = false;  // persistent density launch (workgroups loop over items)
  bool den_dma = false;   // density items staged by direct-to-LDS loads (k_den_tile_dma)
  int persistent_wgs = 256;  // workgroups of the persistent kernels (a multiple of 8)
  int n_den_items = 0, n_vxc_items = 0, n_den_finish = 0, n_den_parts = 0;
  DevBuf<DenItem> den_items;
  DevBuf<VxcItem> vxc_items;
  DevBuf<DenFinish> den_finish;
  DevBuf<double> den_part;  // [n_den_parts][4][kTileAo]
  DevBuf<float> d2f;       // FP32 rounding of the doubled density (FP32 mode)
  DevBuf<uint16_t> d2l;    // BF16 rounding of the remainder (FP32 mode)
  size_t bytes = 0;
};

// Device bytes the tiled path's item lists and density partials take for
// one leaf (an upper bound used before the plan exists)
double tile_plan_bytes(const LeafDev &L);

// Vxc kernel selection (OCCXC_VXC_VARIANT, default 0)
int vxc_variant_setting();

// Whether a context of this precision takes the tiled contractions (the
// plan's tiled flag, from the same settings).
bool tiled_setting(int precision);

// FP32 mode (precision 1), tiled, not meta-GGA: the Vxc stage first forms
// A = sum_q c_q Phi_q (FP32) of every active leaf in a plane of its own
// (4 mp np bytes per leaf, held by the context), then contracts Phi with A
// from LDS filled by direct loads.  OCCXC_VXC_APASS=0 keeps the one-kernel
// form, which forms A per block pair.
bool vxc_apass_setting(int precision, bool mgga);

ContractPlan make_contract_plan(const std::vector<LeafDev> &leaves, long long mtot, int deriv,
                                int precision, hipStream_t s, const TuneChoice &tune = {});
void run_density(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
                 hipStream_t s);
// rho and grad rho of every leaf point for the density D2t into a.rho_out
// (FP64 or FP32-cache contexts, leaves of at most 2048 AOs)
void run_density_store(const BuildArgs &a, const void *ao, int precision, const double *D2t,
                       hipStream_t s);
constexpr int kStoreMaxMp = 2048;
void run_vxc(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
             hipStream_t s);
// restricted meta-GGA Vxc (atomics): the GGA term with a.coef and the three
// tau terms with a.coef_tau (K += d_k Phi^T (c_tau d_k Phi)); leaves_k holds
// the leaf tables offset to the d/dx, d/dy, d/dz planes ([3][nleaves]);
// FP64 or FP32-cache AO values
void run_vxc_mgga(const ContractPlan &plan, const BuildArgs &a, const LeafDev *leaves_k,
                  const void *ao, int precision, hipStream_t s);
// two-pass Vxc: V = sum over leaves of the stored K_loc blocks (+ transposes),
// one workgroup per 32 x 32 tile pair, fixed summation order
void run_vxc_reduce(const BuildArgs &a, double *V, hipStream_t s);
// FP32 AO storage, at context set-up: adds sum_p w_p r Phi^T of the leaves
// [a.res_leaf0, a.res_leaf1) to a.K, from their rounding residuals a.ao_res
// (all leaves marked active in a.leaf_active); dS = K + K^T once all leaves
// are in
void run_storage_overlap(const ContractPlan &plan, const BuildArgs &a, const void *ao,
                         hipStream_t s);

} // namespace occxc
