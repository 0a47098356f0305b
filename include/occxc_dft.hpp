// occxc_dft.hpp -- header-only C++ adapter over the occxc C ABI with the
// shape of OCC's XC entry points:
//
//   Mat DFT::compute_vxc(const MolecularOrbitals &mo, const Mat &Schwarz)
//                                             (include/occ/dft/dft.h:395-429)
//   double DFT::exchange_correlation_energy() const        (dft.h:83)
//   Mat3N DFT::compute_xc_gradient(const MolecularOrbitals &mo, ...)
//                                                          (dft.h:531-532)
//
// XCBuild owns one occxc context (basis + grid + functionals resident on one
// GPU).  compute_vxc takes the restricted density D (nbf x nbf, column-major,
// no factor 2, OCC's convention of include/occ/qm/orb.h:64-66) and returns the
// matrix OCC's compute_vxc returns; the Exc it computed is then available from
// exchange_correlation_energy(), and N_elec (only logged by OCC) from
// total_density().  compute_xc_gradient returns the 3 x natoms column-major
// XC gradient for the same D, given each shell's atom.  Errors arrive as
// std::runtime_error, as in OCC.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "occxc.h"

namespace occxc {

class XCBuild {
public:
  XCBuild(const occxc_basis &basis, const occxc_grid &grid, const std::vector<int> &func_ids,
          const occxc_options *opt = nullptr) {
    occxc_functional f{static_cast<int>(func_ids.size()), func_ids.data()};
    check(occxc_context_create(&basis, &grid, &f, opt, &ctx_));
    occxc_context_info info{};
    check(occxc_context_get_info(ctx_, &info));
    nbf_ = info.nbf;
  }
  ~XCBuild() {
    if (ctx_) occxc_context_destroy(ctx_);
  }
  XCBuild(const XCBuild &) = delete;
  XCBuild &operator=(const XCBuild &) = delete;

  int nbf() const { return nbf_; }

  // Vxc (nbf x nbf, column-major); also sets Exc and N_elec.
  std::vector<double> compute_vxc(const double *D) {
    std::vector<double> V(static_cast<size_t>(nbf_) * nbf_);
    check(occxc_build(ctx_, D, V.data(), &exc_, &nelec_));
    return V;
  }

  // 3 x natoms, column-major (x, y, z of each atom); shell_atom[nshell].
  std::vector<double> compute_xc_gradient(const double *D, int natoms,
                                          const std::vector<int> &shell_atom) {
    std::vector<double> G(3 * static_cast<size_t>(natoms));
    check(occxc_xc_gradient(ctx_, D, natoms, shell_atom.data(), G.data()));
    return G;
  }

  double exchange_correlation_energy() const { return exc_; }
  double total_density() const { return nelec_; }

private:
  static void check(occxc_status s) {
    if (s != OCCXC_OK)
      throw std::runtime_error(std::string("occxc: ") + occxc_last_error_message());
  }
  occxc_context *ctx_ = nullptr;
  int nbf_ = 0;
  double exc_ = 0.0, nelec_ = 0.0;
};

} // namespace occxc
