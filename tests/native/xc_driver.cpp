// A C++ caller of the XC build through include/occxc_dft.hpp, the way an
// SCF driver calls DFT::compute_vxc once per iteration: it reads a basis,
// grid and density from binary files and prints Exc, N_elec and a checksum of
// Vxc for two successive builds (the context stays resident between them).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "occxc_dft.hpp"

template <typename T> static std::vector<T> read(const char *dir, const char *name, size_t n) {
  std::string path = std::string(dir) + "/" + name;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<T> v(n);
  if (std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short read " + path);
  std::fclose(f);
  return v;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: xc_driver DIR\n");
    return 2;
  }
  const char *dir = argv[1];
  auto hdr = read<long long>(dir, "header.bin", 5);  // nshell nprim_total npts nbf nfunc
  const int nsh = (int)hdr[0];
  const size_t nprim = (size_t)hdr[1], npts = (size_t)hdr[2], nbf = (size_t)hdr[3];
  auto l = read<int>(dir, "l.bin", nsh), pure = read<int>(dir, "pure.bin", nsh),
       np = read<int>(dir, "nprim.bin", nsh), ids = read<int>(dir, "ids.bin", (size_t)hdr[4]);
  auto origin = read<double>(dir, "origin.bin", 3 * nsh);
  auto ex = read<double>(dir, "exponents.bin", nprim), co = read<double>(dir, "coefficients.bin", nprim);
  auto pts = read<double>(dir, "points.bin", 3 * npts), w = read<double>(dir, "weights.bin", npts);
  auto D = read<double>(dir, "density.bin", nbf * nbf);
  occxc_basis b{nsh, l.data(), pure.data(), np.data(), origin.data(), ex.data(), co.data()};
  occxc_grid g{(long long)npts, pts.data(), w.data()};
  try {
    occxc::XCBuild xc(b, g, ids);
    for (int it = 0; it < 2; ++it) {
      std::vector<double> V = xc.compute_vxc(D.data());
      double cs = 0.0;
      for (size_t i = 0; i < V.size(); ++i) cs += V[i] * (double)((i % 7) + 1);
      std::printf("iter %d exc %.15e nelec %.15e checksum %.15e\n", it,
                  xc.exchange_correlation_energy(), xc.total_density(), cs);
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
