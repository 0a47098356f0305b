// The launch wrappers' coefficient-plane guards: a stage whose kernels would
// access more planes of per-point coefficients than the buffer holds is
// refused with OCCXC_ERR_INTERNAL before anything is launched, so this runs
// without a GPU.  Exit status = number of guards that did not refuse.
#include <cstdio>

#include "contract.hpp"

int main() {
  using namespace occxc;
  ContractPlan plan;
  BuildArgs a{};
  a.nleaves = 1;
  a.deriv = 1;
  a.coef_planes = 1;
  int fails = 0;
  auto expect = [&](const char *what, auto fn) {
    try {
      fn();
      std::printf("%s: not refused\n", what);
      ++fails;
    } catch (const Error &e) {
      std::printf("%s: status %d: %s\n", what, (int)e.status, e.what());
      if (e.status != OCCXC_ERR_INTERNAL) ++fails;
    }
  };
  expect("GGA Vxc, one coefficient plane", [&] { run_vxc(plan, a, nullptr, 0, nullptr); });
  expect("density, one coefficient plane", [&] { run_density(plan, a, nullptr, 0, nullptr); });
  BuildArgs m = a;
  m.coef_planes = 4;
  m.coef_tau_planes = 0;
  expect("meta-GGA Vxc, no tau plane",
         [&] { run_vxc_mgga(plan, m, nullptr, nullptr, 0, nullptr); });
  return fails;
}
