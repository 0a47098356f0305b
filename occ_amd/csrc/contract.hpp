// contract.hpp -- the density and Vxc contractions of the XC build, planned
// once per context and run per build.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "kernels.hpp"

namespace occxc {

struct ContractPlan {
  bool use_mfma = false;
  size_t bytes = 0;
};

ContractPlan make_contract_plan(const std::vector<LeafDev> &leaves, int deriv, int precision,
                                hipStream_t s);
void run_contractions(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
                      hipStream_t s);

} // namespace occxc
