// contract.hpp -- the density and Vxc contractions of the XC build, planned
// once per context and run per build.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace occxc {

struct ContractPlan {
  bool use_mfma = false;   // false: the FP64 reference kernels of kernels.hip
  int max_np = 0;          // largest padded leaf (points)
  int max_mp = 0;          // largest padded leaf (AOs)
  int density_variant = 0; // 0 per-wavefront, 1 LDS-shared D tiles (FP32 mode:
                           // 2 per-wavefront without the compensated density)
  DevBuf<float> d2f;       // FP32 rounding of the doubled density (FP32 mode)
  DevBuf<uint16_t> d2l;    // BF16 rounding of the remainder (FP32 mode)
  size_t bytes = 0;
};

ContractPlan make_contract_plan(const std::vector<LeafDev> &leaves, long long mtot, int deriv,
                                int precision, hipStream_t s);
void run_density(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
                 hipStream_t s);
void run_vxc(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
             hipStream_t s);

} // namespace occxc
