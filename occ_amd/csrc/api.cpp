// api.cpp -- the occxc C ABI: context set-up (grid hierarchy, screening,
// device layout), the XC build orchestration and the stage entry points.
//
// The context holds, per molecule, everything OCC's DFT object caches or
// recomputes per call in compute_K_dft (include/occ/dft/dft.h:190-388): the
// Morton-sorted grid and its leaves (spatial_grid_hierarchy.cpp:28-147), the
// shell extents (gto.cpp:18-67) and per-leaf significant shells
// (dft.h:307-336), and, resident on the device, the AO values of every leaf.
// A build then runs: D2 = 2 sym(D) -> density + functional per leaf -> Vxc per
// leaf -> K + K^T and the fixed-order reduction of Exc and N_elec.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "occxc.h"
#include "common.hpp"
#include "context.hpp"
#include "contract.hpp"
#include "kernels.hpp"
#include "xc_functionals_mgga.hpp"

namespace occxc {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

// Set on threads that already run one task of an outer parallel loop: their
// inner loops stay serial.
static thread_local bool t_serial = false;

// Split [0, n) over host threads (set-up work only).
static void parallel_for(long long n, const std::function<void(long long, long long)> &fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
  if (n < 4096 || nt <= 1 || t_serial) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  long long chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    long long lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(fn, lo, hi);
  }
  for (auto &t : th) t.join();
}

// Run fn(i) for i in [0, n) on up to 16 host threads, tasks taken in order.
static void parallel_each(int n, const std::function<void(int)> &fn) {
  unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::min<unsigned>(std::min<unsigned>(hw ? hw : 1, 16), (unsigned)n);
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      t_serial = true;
      try {
        for (int i = next++; i < n; i = next++) fn(i);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto &x : th) x.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

static int shell_size(int l, int pure) { return pure ? 2 * l + 1 : (l + 1) * (l + 2) / 2; }

static double common_fac(int l, int pure) {
  // gto.cpp:7-16
  if (l == 0) return 0.282094791773878143;
  if (l == 1) return 0.488602511902919921;
  return pure ? std::sqrt((2.0 * l + 1.0) / (4.0 * 3.14159265358979323846)) : 1.0;
}

// The spherical-transform table in constant memory, uploaded once per device
// (contexts may be created from several host threads at once).
static std::once_flag g_table_once[64];

static void ensure_device_tables(int device) {
  OCCXC_REQUIRE(device >= 0 && device < 64, OCCXC_ERR_INVALID_ARGUMENT, "device ordinal");
  std::call_once(g_table_once[device], [] { upload_sph_table(); });
}

struct HostBasis {
  int nsh = 0, nbf = 0, maxl = 0;
  std::vector<ShellDev> shells;
  std::vector<double> exps, coefs;
};

static HostBasis make_basis(const occxc_basis *b) {
  OCCXC_REQUIRE(b && b->nshell > 0 && b->l && b->pure && b->nprim && b->origin && b->exponents &&
                    b->coefficients,
                OCCXC_ERR_INVALID_ARGUMENT, "basis: null pointer or no shells");
  HostBasis h;
  h.nsh = b->nshell;
  int poff = 0;
  for (int s = 0; s < b->nshell; ++s) {
    const int l = b->l[s], np = b->nprim[s];
    OCCXC_REQUIRE(l >= 0 && l <= kMaxL, OCCXC_ERR_UNSUPPORTED,
                  "basis: angular momentum above " + std::to_string(kMaxL) + " (shell " +
                      std::to_string(s) + ")");
    OCCXC_REQUIRE(np >= 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "basis: shell " + std::to_string(s) + " has no primitives");
    ShellDev sh{};
    sh.x = b->origin[3 * s];
    sh.y = b->origin[3 * s + 1];
    sh.z = b->origin[3 * s + 2];
    sh.l = l;
    sh.pure = b->pure[s] ? 1 : 0;
    sh.nprim = np;
    sh.poff = poff;
    sh.bf0 = h.nbf;
    sh.size = shell_size(l, sh.pure);
    const double fac = common_fac(l, sh.pure);
    for (int k = 0; k < np; ++k) {
      const double a = b->exponents[poff + k], c = b->coefficients[poff + k];
      OCCXC_REQUIRE(std::isfinite(a) && a > 0.0 && std::isfinite(c), OCCXC_ERR_INVALID_ARGUMENT,
                    "basis: non-positive or non-finite exponent/coefficient in shell " +
                        std::to_string(s));
      h.exps.push_back(a);
      h.coefs.push_back(c * fac);
    }
    h.shells.push_back(sh);
    h.nbf += sh.size;
    h.maxl = std::max(h.maxl, l);
    poff += np;
  }
  return h;
}

// ---------------------------------------------------- grid hierarchy --

static inline uint64_t spread_bits_3(uint64_t v) {
  // insert two zero bits between each of the low 21 bits
  v &= 0x1fffffULL;
  uint64_t r = 0;
  for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ULL) << (3 * i);
  return r;
}

struct Hierarchy {
  std::vector<double> xyz, w;       // sorted
  std::vector<long long> perm;
  std::vector<long long> off, cnt;
  std::vector<double> center, radius;
};

// SpatialGridHierarchy (spatial_grid_hierarchy.cpp:28-147): 21-bit-per-axis
// Morton codes over the bounding box, sort by (code, index), leaves of 128
// (min 32, max 256) with centroid / max-distance spheres.
static Hierarchy build_hierarchy(long long n, const double *xyz, const double *w) {
  Hierarchy h;
  if (n <= 0) return h;
  double mn[3], mx[3], ext[3];
  for (int k = 0; k < 3; ++k) mn[k] = mx[k] = xyz[k];
  for (long long i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], xyz[3 * i + k]);
      mx[k] = std::max(mx[k], xyz[3 * i + k]);
    }
  for (int k = 0; k < 3; ++k) {
    ext[k] = mx[k] - mn[k];
    if (ext[k] < 1e-10) ext[k] = 1.0;
  }
  std::vector<std::pair<uint64_t, long long>> codes(n);
  const uint64_t scale = (1ULL << 21) - 1;
  parallel_for(n, [&](long long lo, long long hi) {
    for (long long i = lo; i < hi; ++i) {
      uint64_t q[3];
      for (int k = 0; k < 3; ++k) {
        double t = (xyz[3 * i + k] - mn[k]) / ext[k];
        t = std::min(std::max(t, 0.0), 1.0);
        q[k] = (uint64_t)(t * (double)scale);
      }
      codes[i] = {spread_bits_3(q[0]) | (spread_bits_3(q[1]) << 1) | (spread_bits_3(q[2]) << 2), i};
    }
  });
  std::sort(codes.begin(), codes.end());
  h.xyz.resize(3 * n);
  h.w.resize(n);
  h.perm.resize(n);
  for (long long i = 0; i < n; ++i) {
    const long long o = codes[i].second;
    h.perm[i] = o;
    for (int k = 0; k < 3; ++k) h.xyz[3 * i + k] = xyz[3 * o + k];
    h.w[i] = w[o];
  }
  const long long target = 128, minl = 32, maxl = 256;
  long long off = 0;
  while (off < n) {
    const long long rem = n - off;
    long long cnt;
    if (rem <= maxl) cnt = rem;
    else if (rem < 2 * minl) cnt = rem / 2;
    else cnt = std::min(target, rem);
    double c[3] = {0, 0, 0};
    for (long long i = 0; i < cnt; ++i)
      for (int k = 0; k < 3; ++k) c[k] += h.xyz[3 * (off + i) + k];
    for (int k = 0; k < 3; ++k) c[k] /= (double)cnt;
    double r = 0.0;
    for (long long i = 0; i < cnt; ++i) {
      const double dx = h.xyz[3 * (off + i)] - c[0], dy = h.xyz[3 * (off + i) + 1] - c[1],
                   dz = h.xyz[3 * (off + i) + 2] - c[2];
      r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    h.off.push_back(off);
    h.cnt.push_back(cnt);
    h.center.insert(h.center.end(), c, c + 3);
    h.radius.push_back(r);
    off += cnt;
  }
  return h;
}

static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

// Temporaries of the device set-up, alive until the context is set up.
struct DeviceSetup {
  DevBuf<double> xyz, w, center, radius, part, box;
  DevBuf<int> perm, nsig, nao, idx;
  DevBuf<unsigned long long> keys_in, keys_out;
  DevBuf<char> temp;
  DevBuf<FillLeaf> fill;
};

// The screened leaf layout of every molecule built on the device (setup.hip):
// raw points uploaded once, per molecule the Morton sort, leaf spheres and
// per-leaf screening counts; the host turns the counts into the LeafDev
// table (prefix sums over the active leaves), and one kernel writes the
// leaf-shell lists, the AO maps and the padded points.  Gives the same
// layout as the host path below.
static void setup_leaves_device(occxc_context *c, int nmol, const occxc_grid *grids,
                                const std::vector<HostBasis> &hbs,
                                const std::vector<int> &shell_base, std::vector<MolDev> &mols,
                                const double *d_ext, DeviceSetup &ds, long long &pt_off,
                                long long &ao_off, double &msum) {
  hipStream_t s = c->stream;
  std::vector<long long> pbase(nmol + 1, 0);
  std::vector<int> lbase(nmol + 1, 0);
  long long nmax = 0;
  for (int m = 0; m < nmol; ++m) {
    pbase[m + 1] = pbase[m] + grids[m].npts;
    lbase[m + 1] = lbase[m] + hierarchy_leaf_count(grids[m].npts);
    nmax = std::max<long long>(nmax, grids[m].npts);
  }
  const long long N = pbase[nmol];
  const int Lall = lbase[nmol];
  ds.xyz.alloc(std::max<long long>(3 * N, 1));
  ds.w.alloc(std::max<long long>(N, 1));
  for (int m = 0; m < nmol; ++m) {
    if (grids[m].npts) {
      OCCXC_HIP(hipMemcpyAsync(ds.xyz.get() + 3 * pbase[m], grids[m].points,
                               3 * grids[m].npts * sizeof(double), hipMemcpyHostToDevice, s));
      OCCXC_HIP(hipMemcpyAsync(ds.w.get() + pbase[m], grids[m].weights,
                               grids[m].npts * sizeof(double), hipMemcpyHostToDevice, s));
    }
  }
  ds.perm.alloc(std::max<long long>(N, 1));
  ds.center.alloc(std::max(3 * Lall, 1));
  ds.radius.alloc(std::max(Lall, 1));
  ds.nsig.alloc(std::max(Lall, 1));
  ds.nao.alloc(std::max(Lall, 1));
  ds.keys_in.alloc(std::max<long long>(nmax, 1));
  ds.keys_out.alloc(std::max<long long>(nmax, 1));
  ds.idx.alloc(std::max<long long>(nmax, 1));
  const size_t tb = std::max<size_t>(sort_temp_bytes(nmax), 1);
  ds.temp.alloc(tb);
  ds.part.alloc(256 * 6);
  ds.box.alloc(6);
  for (int m = 0; m < nmol; ++m) {
    const long long n = grids[m].npts;
    const int nl = lbase[m + 1] - lbase[m];
    if (n == 0) continue;
    launch_hierarchy(ds.xyz.get() + 3 * pbase[m], n, nl, ds.part.get(), ds.box.get(),
                     ds.keys_in.get(), ds.keys_out.get(), ds.idx.get(), ds.perm.get() + pbase[m],
                     ds.temp.get(), tb, ds.center.get() + 3 * lbase[m], ds.radius.get() + lbase[m],
                     s);
    launch_screen_count(ds.center.get() + 3 * lbase[m], ds.radius.get() + lbase[m], nl,
                        c->d_shells.get() + shell_base[m], d_ext + shell_base[m], hbs[m].nsh,
                        ds.nsig.get() + lbase[m], ds.nao.get() + lbase[m], s);
  }
  std::vector<int> nsig(Lall), nao(Lall);
  ds.nsig.download(nsig.data(), Lall, s);
  ds.nao.download(nao.data(), Lall, s);
  OCCXC_HIP(hipStreamSynchronize(s));
  std::vector<FillLeaf> fill;
  long long sh_total = 0, map_total = 0;
  for (int m = 0; m < nmol; ++m) {
    const long long n = grids[m].npts;
    const int nl = lbase[m + 1] - lbase[m];
    mols[m].leaf_begin = (int)c->leaves.size();
    for (int li = 0; li < nl; ++li) {
      const int g = lbase[m] + li;
      if (nsig[g] == 0) continue;
      LeafDev L{};
      L.n = (int)(li < nl - 1 ? 128 : n - 128LL * li);
      L.np = (int)round_up(L.n, kPtPad);
      L.pt_off = (int)pt_off;
      L.sh_off = (int)sh_total;
      L.nsig = nsig[g];
      L.map_off = (int)map_total;
      L.mol = m;
      L.m = nao[g];
      OCCXC_REQUIRE(L.m <= kMaxLocalAo, OCCXC_ERR_UNSUPPORTED,
                    "a grid leaf has more than " + std::to_string(kMaxLocalAo) +
                        " significant basis functions");
      L.mp = (int)round_up(L.m, kAoPad);
      L.ao_off = ao_off;
      ao_off += (long long)c->ncomp * L.mp * L.np;
      L.res_off = (long long)c->res_elems;
      c->res_elems += (size_t)L.mp * L.np;
      L.kl_off = c->kloc_elems;
      c->kloc_elems += (long long)L.mp * L.mp;
      sh_total += L.nsig;
      map_total += L.mp;
      pt_off += L.np;
      msum += L.m;
      c->flops += 4.0 * L.n * (double)L.m * L.m;
      c->ao_write_bytes +=
          (double)L.n * L.m * (c->opt.precision == 0 ? 8.0 * c->ncomp : 4.0 * c->ncomp + 4.0) +
          32.0 * L.n;
      c->leaves.push_back(L);
      fill.push_back(FillLeaf{g, shell_base[m], hbs[m].nsh, 0, pbase[m] + 128LL * li, pbase[m]});
    }
    mols[m].leaf_end = (int)c->leaves.size();
  }
  OCCXC_REQUIRE(map_total < (1LL << 31) && pt_off < (1LL << 31), OCCXC_ERR_UNSUPPORTED,
                "context too large for 32-bit leaf offsets");
  c->d_leaves.alloc(c->leaves.size());
  c->d_leaves.upload(c->leaves.data(), c->leaves.size(), s);
  ds.fill.alloc(fill.size());
  ds.fill.upload(fill.data(), fill.size(), s);
  c->d_leaf_shells.alloc(sh_total);
  c->d_ao_map.alloc(map_total);
  c->d_px.alloc(pt_off);
  c->d_py.alloc(pt_off);
  c->d_pz.alloc(pt_off);
  c->d_pw.alloc(pt_off);
  c->d_porig.alloc(std::max<long long>(pt_off, 1));
  launch_fill_leaves(c->d_leaves.get(), ds.fill.get(), (int)fill.size(), ds.center.get(),
                     ds.radius.get(), c->d_shells.get(), d_ext, ds.xyz.get(), ds.w.get(),
                     ds.perm.get(), c->d_leaf_shells.get(), c->d_ao_map.get(), c->d_px.get(),
                     c->d_py.get(), c->d_pz.get(), c->d_pw.get(), c->d_porig.get(), s);
}

} // namespace occxc

using namespace occxc;

namespace occxc {

void context_setup(occxc_context *c, int nmol, const occxc_basis *bases,
                   const occxc_grid *grids, const occxc_functional *func,
                   const occxc_options *opt, size_t byte_limit) {
  c->opt = *opt;
  OCCXC_REQUIRE(nmol >= 1 && bases && grids, OCCXC_ERR_INVALID_ARGUMENT,
                "need at least one molecule (basis and grid)");
  OCCXC_REQUIRE(opt->precision >= 0 && opt->precision <= 3, OCCXC_ERR_INVALID_ARGUMENT,
                "options: unknown precision mode");
  OCCXC_REQUIRE(opt->tile_ao == 0 || (opt->tile_ao >= 32 && opt->tile_ao <= kTileAo &&
                                      opt->tile_ao % 32 == 0),
                OCCXC_ERR_INVALID_ARGUMENT, "options: tile_ao must be 0 or 32..128, a multiple of 32");
  OCCXC_REQUIRE(opt->persistent >= -1 && opt->persistent <= 3, OCCXC_ERR_INVALID_ARGUMENT,
                "options: persistent must be -1 or 0..3");
  OCCXC_REQUIRE(func && func->nfunc >= 1 && func->nfunc <= kMaxFunc && func->ids,
                OCCXC_ERR_INVALID_ARGUMENT, "functional: need 1..8 functional ids");
  int deriv = 0;
  for (int f = 0; f < func->nfunc; ++f) {
    const int fam = xc_family_mgga(func->ids[f]);
    OCCXC_REQUIRE(fam >= 0, OCCXC_ERR_UNSUPPORTED,
                  "functional id " + std::to_string(func->ids[f]) + " not found");
    deriv = std::max(deriv, fam);
    c->funcs[f] = func->ids[f];
  }
  c->nfunc = func->nfunc;
  // a meta-GGA needs phi and grad phi in the AO cache, like a GGA (tau is
  // formed from grad phi; no laplacian is formed)
  c->mgga = deriv == 2;
  c->deriv = std::min(deriv, 1);
  c->ncomp = c->deriv ? 4 : 1;
  OCCXC_REQUIRE(!c->mgga || opt->precision == OCCXC_PREC_FP64 ||
                    opt->precision == OCCXC_PREC_FP32_CACHE,
                OCCXC_ERR_UNSUPPORTED, "meta-GGA builds need an FP64 or FP32-cache context");

  // molecules: bases concatenated (shell and primitive offsets shifted)
  std::vector<HostBasis> hbs;
  std::vector<ShellDev> shells;
  std::vector<double> exps, coefs;
  std::vector<int> shell_base(nmol + 1, 0);
  long long npts_total = 0;
  for (int m = 0; m < nmol; ++m) {
    const occxc_grid *grid = &grids[m];
    OCCXC_REQUIRE(grid->npts >= 0 && (grid->npts == 0 || (grid->points && grid->weights)),
                  OCCXC_ERR_INVALID_ARGUMENT, "grid: null pointer");
    for (long long i = 0; i < grid->npts; ++i)
      OCCXC_REQUIRE(std::isfinite(grid->points[3 * i]) && std::isfinite(grid->points[3 * i + 1]) &&
                        std::isfinite(grid->points[3 * i + 2]) && std::isfinite(grid->weights[i]),
                    OCCXC_ERR_INVALID_ARGUMENT, "grid: non-finite point or weight");
    npts_total += grid->npts;
    hbs.push_back(make_basis(&bases[m]));
    HostBasis &hb = hbs.back();
    const int pbase = (int)exps.size();
    for (ShellDev sh : hb.shells) {
      sh.poff += pbase;
      shells.push_back(sh);
    }
    exps.insert(exps.end(), hb.exps.begin(), hb.exps.end());
    coefs.insert(coefs.end(), hb.coefs.begin(), hb.coefs.end());
    shell_base[m + 1] = shell_base[m] + hb.nsh;
    c->mol_nbf.push_back(hb.nbf);
  }
  OCCXC_REQUIRE(npts_total < (1LL << 31) - (long long)kMaxLeafPts * 4 * nmol,
                OCCXC_ERR_UNSUPPORTED, "grid: more than 2^31 points in one context");
  c->nmol = nmol;
  c->nsh = (int)shells.size();
  c->nbf = hbs[0].nbf;
  if (nmol == 1)
    for (int k = 0; k < hbs[0].nsh; ++k)
      for (int q = 0; q < hbs[0].shells[k].size; ++q) c->bf_shell.push_back(k);
  c->npts = npts_total;
  std::vector<MolDev> mols(nmol);
  long long moff = 0;
  for (int m = 0; m < nmol; ++m) {
    mols[m].moff = moff;
    mols[m].nbf = hbs[m].nbf;
    c->mol_off.push_back(moff);
    moff += (long long)hbs[m].nbf * hbs[m].nbf;
    c->maxnbf = std::max(c->maxnbf, hbs[m].nbf);
  }
  c->mtot = moff;

  DeviceGuard guard(opt->device);
  c->device = opt->device;
  ensure_device_tables(c->device);
  OCCXC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->d_shells.alloc(shells.size());
  c->d_shells.upload(shells.data(), shells.size(), c->stream);
  c->d_exps.alloc(exps.size());
  c->d_exps.upload(exps.data(), exps.size(), c->stream);
  c->d_coefs.alloc(coefs.size());
  c->d_coefs.upload(coefs.data(), coefs.size(), c->stream);

  // shell extents on the device (gto.cpp:18-67)
  std::vector<double> ext(shells.size(), 30.0);
  const bool screened = opt->screen != 0 && opt->screening_threshold > 0.0;
  // the screened layout is built on the device unless OCCXC_HOST_SETUP is set
  // (the host restatement, kept to check that both give the same layout)
  const bool device_setup = screened && std::getenv("OCCXC_HOST_SETUP") == nullptr;
  DevBuf<double> d_ext;
  if (screened) {
    d_ext.alloc(shells.size());
    launch_decay_cutoff(c->d_shells.get(), (int)shells.size(), c->d_exps.get(), c->d_coefs.get(),
                        opt->screening_threshold, d_ext.get(), c->stream);
    if (!device_setup) {
      d_ext.download(ext.data(), shells.size(), c->stream);
      OCCXC_HIP(hipStreamSynchronize(c->stream));
    }
  }

  std::vector<int> leaf_shells, ao_map, porig;
  std::vector<double> px, py, pz, pw;
  long long pt_off = 0, ao_off = 0;
  double msum = 0.0;
  DeviceSetup dsetup;
  if (device_setup) {
    setup_leaves_device(c, nmol, grids, hbs, shell_base, mols, d_ext.get(), dsetup, pt_off,
                        ao_off, msum);
  } else {
  // per molecule: leaves (Morton hierarchy when screened, else 128-point
  // chunks in input order) and their significant shells (dft.h:314-320);
  // molecules are independent, so a batch builds them on several threads
  std::vector<Hierarchy> hier(nmol);
  std::vector<std::vector<std::vector<int>>> sigs(nmol);
  parallel_each(nmol, [&](int mol) {
    const occxc_grid *grid = &grids[mol];
    const HostBasis &hb = hbs[mol];
    const int sb = shell_base[mol];
    Hierarchy &h = hier[mol];
    const long long n = grid->npts;
    if (screened) {
      h = build_hierarchy(n, grid->points, grid->weights);
    } else {
      h.xyz.assign(grid->points, grid->points + 3 * n);
      h.w.assign(grid->weights, grid->weights + n);
      for (long long o = 0; o < n; o += 128) {
        h.off.push_back(o);
        h.cnt.push_back(std::min<long long>(128, n - o));
        h.center.insert(h.center.end(), {0.0, 0.0, 0.0});
        h.radius.push_back(0.0);
      }
    }
    const long long nl_all = (long long)h.off.size();
    auto &sig = sigs[mol];
    sig.assign(nl_all, {});
    parallel_for(nl_all, [&](long long lo, long long hi) {
      for (long long li = lo; li < hi; ++li) {
        if (h.cnt[li] == 0) continue;
        auto &v = sig[li];
        for (int s = 0; s < hb.nsh; ++s) {
          if (screened) {
            const ShellDev &sh = hb.shells[s];
            const double dx = sh.x - h.center[3 * li], dy = sh.y - h.center[3 * li + 1],
                         dz = sh.z - h.center[3 * li + 2];
            const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dist - ext[sb + s] < h.radius[li]) v.push_back(s);
          } else {
            v.push_back(s);
          }
        }
      }
    });
  });
  for (int mol = 0; mol < nmol; ++mol) {
    const HostBasis &hb = hbs[mol];
    const int sb = shell_base[mol];
    const Hierarchy &h = hier[mol];
    const auto &sig = sigs[mol];
    const long long nl_all = (long long)h.off.size();
    mols[mol].leaf_begin = (int)c->leaves.size();
    for (long long li = 0; li < nl_all; ++li) {
      if (sig[li].empty()) continue;
      LeafDev L{};
      L.n = (int)h.cnt[li];
      L.np = (int)round_up(L.n, kPtPad);
      L.pt_off = (int)pt_off;
      L.sh_off = (int)leaf_shells.size();
      L.nsig = (int)sig[li].size();
      L.map_off = (int)ao_map.size();
      L.mol = mol;
      int m = 0;
      for (int s : sig[li]) {
        leaf_shells.push_back(sb + s);
        for (int q = 0; q < hb.shells[s].size; ++q) ao_map.push_back(hb.shells[s].bf0 + q);
        m += hb.shells[s].size;
      }
      OCCXC_REQUIRE(m <= kMaxLocalAo, OCCXC_ERR_UNSUPPORTED,
                    "a grid leaf has more than " + std::to_string(kMaxLocalAo) +
                        " significant basis functions");
      L.m = m;
      L.mp = (int)round_up(m, kAoPad);
      // padding AOs map to AO 0: their values are zero, so they contribute nothing
      for (int q = m; q < L.mp; ++q) ao_map.push_back(0);
      L.ao_off = ao_off;
      ao_off += (long long)c->ncomp * L.mp * L.np;
      L.res_off = (long long)c->res_elems;
      c->res_elems += (size_t)L.mp * L.np;
      L.kl_off = c->kloc_elems;
      c->kloc_elems += (long long)L.mp * L.mp;
      const long long o = h.off[li];
      for (int p = 0; p < L.np; ++p) {
        const long long src = o + (p < L.n ? p : 0);
        px.push_back(h.xyz[3 * src]);
        py.push_back(h.xyz[3 * src + 1]);
        pz.push_back(h.xyz[3 * src + 2]);
        pw.push_back(p < L.n ? h.w[src] : 0.0);
        porig.push_back(p < L.n ? (int)(h.perm.empty() ? src : h.perm[src]) : -1);
      }
      pt_off += L.np;
      msum += m;
      c->flops += 4.0 * L.n * (double)m * m;
      c->ao_write_bytes +=
          (double)L.n * m * (opt->precision == 0 ? 8.0 * c->ncomp : 4.0 * c->ncomp + 4.0) +
          32.0 * L.n;
      c->leaves.push_back(L);
    }
    mols[mol].leaf_end = (int)c->leaves.size();
  }
  c->d_leaves.alloc(c->leaves.size());
  c->d_leaves.upload(c->leaves.data(), c->leaves.size(), c->stream);
  c->d_leaf_shells.alloc(leaf_shells.size());
  c->d_leaf_shells.upload(leaf_shells.data(), leaf_shells.size(), c->stream);
  c->d_ao_map.alloc(ao_map.size());
  c->d_ao_map.upload(ao_map.data(), ao_map.size(), c->stream);
  c->d_px.alloc(px.size());
  c->d_px.upload(px.data(), px.size(), c->stream);
  c->d_py.alloc(py.size());
  c->d_py.upload(py.data(), py.size(), c->stream);
  c->d_pz.alloc(pz.size());
  c->d_pz.upload(pz.data(), pz.size(), c->stream);
  c->d_pw.alloc(pw.size());
  c->d_pw.upload(pw.data(), pw.size(), c->stream);
  c->d_porig.alloc(std::max<size_t>(porig.size(), 1));
  c->d_porig.upload(porig.data(), porig.size(), c->stream);
  }
  c->nleaves = (int)c->leaves.size();
  c->ptot = pt_off;
  c->mean_m = c->nleaves ? msum / c->nleaves : 0.0;
  c->ao_elems = (size_t)ao_off;
  c->ao_elem_size = opt->precision == 0 ? 8 : 4;

  // device layout
  c->d_mols.alloc(mols.size());
  c->d_mols.upload(mols.data(), mols.size(), c->stream);
  // FP32 storage: the value plane's rounding residuals exist during set-up
  // only; the context keeps dS (nbf^2 per molecule) for the N_elec correction
  if (opt->precision == 0) c->res_elems = 0;
  const size_t ao_bytes = c->ao_elems * c->ao_elem_size;
  const size_t ds_bytes = c->res_elems ? (size_t)c->mtot * 8 : 0;
  // FP32 mode: the Vxc stage's A plane (as many floats as one AO plane)
  const bool apass = vxc_apass_setting(opt->precision, c->mgga);
  const size_t a_bytes = apass ? c->res_elems * sizeof(float) : 0;
  // device memory the context will hold: AO values, the nbf^2 matrices
  // (D, D2, K, V and the FP32 copies), per-point data and the leaf tables
  const bool two_pass = vxc_variant_setting() == 2;
  c->mol_bytes.assign(nmol, 0.0);
  for (int m = 0; m < nmol; ++m)
    c->mol_bytes[m] = (c->res_elems ? 7.0 : 6.0) * 8.0 * c->mol_nbf[m] * c->mol_nbf[m];
  for (const LeafDev &L : c->leaves)
    c->mol_bytes[L.mol] += (double)c->ncomp * L.mp * L.np * c->ao_elem_size +
                           (apass ? 4.0 * L.mp * L.np : 0.0) + 64.0 * L.np +
                           4.0 * (L.mp + L.nsig) + sizeof(LeafDev) +
                           (two_pass ? 8.0 * L.mp * L.mp : tile_plan_bytes(L));
  double tile_bytes = 0.0;
  if (!two_pass)
    for (const LeafDev &L : c->leaves) tile_bytes += tile_plan_bytes(L);
  const size_t planned = ao_bytes + ds_bytes + a_bytes + (size_t)tile_bytes + (size_t)c->mtot * 8 * 6 +
                         (size_t)c->ptot * 8 * 8 +
                         (c->mgga ? (size_t)c->ptot * 8 * 8 : 0) +
                         c->d_ao_map.bytes() + c->d_leaves.bytes() + c->d_leaf_shells.bytes() +
                         (two_pass ? (size_t)c->kloc_elems * 8 : 0);
  c->device_bytes = planned;
  OCCXC_REQUIRE(byte_limit == 0 || planned <= byte_limit, OCCXC_ERR_OUT_OF_MEMORY,
                "context needs " + std::to_string(planned >> 20) + " MiB, limit " +
                    std::to_string(byte_limit >> 20) + " MiB");
  size_t free_b = 0, total_b = 0;
  OCCXC_HIP(hipMemGetInfo(&free_b, &total_b));
  // the residuals are formed a bounded range of leaves at a time
  size_t res_cap = 0;
  if (c->res_elems) {
    size_t max_leaf = 0;
    for (const LeafDev &L : c->leaves) max_leaf = std::max(max_leaf, (size_t)L.mp * L.np);
    res_cap = std::min(c->res_elems, std::max(kResidualChunkElems, max_leaf));
  }
  // the residuals go before the A plane comes
  const size_t setup_bytes = ao_bytes + ds_bytes + std::max(res_cap * sizeof(float), a_bytes);
  OCCXC_REQUIRE(setup_bytes + (size_t)6 * c->mtot * 8 < free_b, OCCXC_ERR_OUT_OF_MEMORY,
                "resident AO values need " + std::to_string(setup_bytes >> 20) + " MiB, " +
                    std::to_string(free_b >> 20) + " MiB free");
  c->d_ao.alloc(ao_bytes);
  if (res_cap) c->d_ao_res.alloc(res_cap);
  const size_t nn = (size_t)c->mtot;
  c->d_D.alloc(nn);
  c->d_D2.alloc(nn);
  c->d_K.alloc(nn);
  c->d_V.alloc(nn);
  c->d_coef.alloc((size_t)4 * std::max<long long>(c->ptot, 1));
  c->d_part.alloc((size_t)2 * std::max(c->nleaves, 1));
  c->d_active.alloc((size_t)std::max(c->nleaves, 1));
  c->d_result.alloc((size_t)2 * nmol);

  BuildArgs &a = c->args;
  a.nmol = nmol;
  a.maxnbf = c->maxnbf;
  a.mtot = c->mtot;
  a.nleaves = c->nleaves;
  a.deriv = c->deriv;
  a.ncomp = c->ncomp;
  a.nfunc = c->nfunc;
  a.lmax = 0;
  for (const auto &sh : shells) a.lmax = std::max(a.lmax, sh.l);
  for (int f = 0; f < kMaxFunc; ++f) a.func_ids[f] = c->funcs[f];
  a.dthr = opt->density_threshold;
  a.ptot = c->ptot;
  a.leaves = c->d_leaves.get();
  a.mols = c->d_mols.get();
  a.shells = c->d_shells.get();
  a.leaf_shells = c->d_leaf_shells.get();
  a.ao_map = c->d_ao_map.get();
  a.exps = c->d_exps.get();
  a.coefs = c->d_coefs.get();
  a.px = c->d_px.get();
  a.py = c->d_py.get();
  a.pz = c->d_pz.get();
  a.pw = c->d_pw.get();
  a.D2 = c->d_D2.get();
  a.K = c->d_K.get();
  a.coef = c->d_coef.get();
  a.coef_planes = 4;
  a.coef_tau_planes = 0;
  a.leaf_part = c->d_part.get();
  a.leaf_active = c->d_active.get();
  a.ao_res = nullptr;  // set only while dS is formed
  a.ao_a = nullptr;    // set once the A plane is allocated

  c->plan = make_contract_plan(c->leaves, c->mtot, c->deriv, opt->precision, c->stream,
                               tune_choice(opt, c->maxnbf));
  c->plan.two_pass = c->plan.two_pass && c->plan.use_mfma;
  if (c->plan.two_pass) {
    if (ao_map.empty() && c->d_ao_map.size()) {
      ao_map.resize(c->d_ao_map.size());
      c->d_ao_map.download(ao_map.data(), ao_map.size(), c->stream);
      OCCXC_HIP(hipStreamSynchronize(c->stream));
    }
    // 32 x 32 AO tile pairs (TI <= TJ) per molecule and, for each, the leaves
    // whose AOs touch both tiles (in leaf order); a leaf's AOs in one tile are
    // a run of its (sorted) local AOs, recorded as a start and a bit mask
    std::vector<TilePairDev> tpairs;
    std::vector<TileEntry> tentries;
    std::vector<TileChunkDev> tchunks;
    struct Touch { int tile, start; unsigned mask; };
    std::vector<Touch> touch;
    for (int mol = 0; mol < nmol; ++mol) {
      const int nt = (mols[mol].nbf + 31) / 32;
      std::vector<std::vector<TileEntry>> plist((size_t)nt * nt);
      for (int li = mols[mol].leaf_begin; li < mols[mol].leaf_end; ++li) {
        const LeafDev &L = c->leaves[li];
        touch.clear();
        for (int mu = 0; mu < L.m; ++mu) {
          const int g = ao_map[L.map_off + mu];
          if (touch.empty() || touch.back().tile != g / 32) touch.push_back({g / 32, mu, 0u});
          touch.back().mask |= 1u << (g % 32);
        }
        for (size_t x = 0; x < touch.size(); ++x)
          for (size_t y = x; y < touch.size(); ++y)
            plist[(size_t)touch[x].tile * nt + touch[y].tile].push_back(
                TileEntry{L.kl_off, L.mp, li, touch[x].start, touch[y].start, touch[x].mask,
                          touch[y].mask});
      }
      for (int ti = 0; ti < nt; ++ti)
        for (int tj = ti; tj < nt; ++tj) {
          auto &v = plist[(size_t)ti * nt + tj];
          const long long off = (long long)tentries.size();
          const int nchunk = ((int)v.size() + kTileChunk - 1) / kTileChunk;
          for (int q = 0; q < nchunk; ++q)
            tchunks.push_back(TileChunkDev{off + (long long)q * kTileChunk,
                                           off + std::min<long long>((long long)(q + 1) * kTileChunk,
                                                                     (long long)v.size()),
                                           (int)tpairs.size(), 0});
          tpairs.push_back(TilePairDev{off, (int)v.size(), mol, ti, tj,
                                       (int)tchunks.size() - nchunk, nchunk});
          tentries.insert(tentries.end(), v.begin(), v.end());
        }
    }
    c->d_kloc.alloc((size_t)std::max<long long>(c->kloc_elems, 1));
    c->d_tpairs.alloc(std::max<size_t>(tpairs.size(), 1));
    c->d_tpairs.upload(tpairs.data(), tpairs.size(), c->stream);
    c->d_tentries.alloc(std::max<size_t>(tentries.size(), 1));
    c->d_tentries.upload(tentries.data(), tentries.size(), c->stream);
    a.kloc = c->d_kloc.get();
    a.tpairs = c->d_tpairs.get();
    a.ntpairs = (int)tpairs.size();
    a.tentries = c->d_tentries.get();
    c->d_tchunks.alloc(std::max<size_t>(tchunks.size(), 1));
    c->d_tchunks.upload(tchunks.data(), tchunks.size(), c->stream);
    c->d_tpartial.alloc(std::max<size_t>(tchunks.size(), 1) * 1024);
    a.tchunks = c->d_tchunks.get();
    a.ntchunks = (int)tchunks.size();
    a.tpartial = c->d_tpartial.get();
  }
  c->device_bytes = ao_bytes + ds_bytes + c->d_D.bytes() * 4 + c->d_coef.bytes() + c->d_px.bytes() * 4 +
                    c->d_porig.bytes() +
                    c->d_ao_map.bytes() + c->d_leaves.bytes() + c->d_leaf_shells.bytes() +
                    c->plan.bytes + c->d_kloc.bytes() + c->d_tpairs.bytes() +
                    c->d_tentries.bytes() + c->d_tchunks.bytes() + c->d_tpartial.bytes();

  if (c->mgga) {
    OCCXC_REQUIRE(c->plan.use_mfma && !c->plan.two_pass && c->plan.max_mp <= kStoreMaxMp,
                  OCCXC_ERR_UNSUPPORTED,
                  "meta-GGA builds need the matrix-core kernels, the one-pass Vxc and leaves of "
                  "at most " + std::to_string(kStoreMaxMp) + " AOs");
    const size_t P = (size_t)std::max<long long>(c->ptot, 1);
    c->d_mrho.alloc(4 * P);
    c->d_mtau.alloc(3 * P);
    c->d_coef_tau.alloc(P);
    // leaf tables whose AO block offsets point at the d/dx, d/dy, d/dz planes
    std::vector<LeafDev> lk(3 * c->leaves.size());
    for (int k = 0; k < 3; ++k)
      for (size_t i = 0; i < c->leaves.size(); ++i) {
        LeafDev L = c->leaves[i];
        L.ao_off += (long long)(k + 1) * L.mp * L.np;
        lk[(size_t)k * c->leaves.size() + i] = L;
      }
    c->d_leaves_k.alloc(std::max<size_t>(lk.size(), 1));
    c->d_leaves_k.upload(lk.data(), lk.size(), c->stream);
    c->device_bytes += c->d_mrho.bytes() + c->d_mtau.bytes() + c->d_coef_tau.bytes() +
                       c->d_leaves_k.bytes();
  }

  launch_ao_eval(a, c->d_ao.get(), opt->precision != 0, c->stream);
  if (c->res_elems) {
    // dS = K + K^T, K = sum_p w_p r Phi^T accumulated over ranges of leaves
    // whose residuals fit the buffer (the Vxc item list drives the
    // contraction); then the residuals are released
    c->d_dS.alloc((size_t)c->mtot);
    c->d_ncorr.alloc((size_t)kNcorrBlocks * nmol);
    launch_fill_int(c->d_active.get(), c->nleaves, 1, c->stream);
    launch_zero(c->d_K.get(), c->mtot, c->stream);
    const size_t cap = c->d_ao_res.size();
    for (int l0 = 0; l0 < c->nleaves;) {
      size_t used = 0;
      int l1 = l0;
      while (l1 < c->nleaves && used + (size_t)c->leaves[l1].mp * c->leaves[l1].np <= cap) {
        used += (size_t)c->leaves[l1].mp * c->leaves[l1].np;
        ++l1;
      }
      BuildArgs x = a;
      x.ao_res = c->d_ao_res.get();
      x.res_leaf0 = l0;
      x.res_leaf1 = l1;
      x.res_base = c->leaves[l0].res_off;
      launch_ao_residuals(x, c->d_ao.get(), c->stream);
      run_storage_overlap(c->plan, x, c->d_ao.get(), c->stream);
      l0 = l1;
    }
    launch_symmetrize(c->d_K.get(), c->d_dS.get(), c->d_mols.get(), c->nmol, c->maxnbf, c->stream);
    c->device_bytes += c->d_ncorr.bytes();
    OCCXC_HIP(hipStreamSynchronize(c->stream));
    c->d_ao_res.release();
  }
  if (apass && c->plan.tiled) {
    c->d_ao_a.alloc(c->res_elems);
    a.ao_a = c->d_ao_a.get();
    c->device_bytes += c->d_ao_a.bytes();
  }
  OCCXC_HIP(hipStreamSynchronize(c->stream));
  c->ao_ready = true;
}

// Records a pair of events around fn() on stream s when timing is enabled.
struct StageTimer {
  occxc_context *c;
  hipStream_t s;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> set;
  template <typename F> void run(int stage, F &&fn) {
    if (!c->timing) {
      fn();
      return;
    }
    hipEvent_t a, b;
    OCCXC_HIP(hipEventCreate(&a));
    OCCXC_HIP(hipEventCreate(&b));
    OCCXC_HIP(hipEventRecord(a, s));
    fn();
    OCCXC_HIP(hipEventRecord(b, s));
    set.push_back({stage, {a, b}});
  }
  ~StageTimer() {
    if (!set.empty()) c->pending.push_back(std::move(set));
  }
};

void collect_timing(occxc_context *c) {
  for (auto &set : c->pending)
    for (auto &e : set) {
      float ms = 0.f;
      OCCXC_HIP(hipEventSynchronize(e.second.second));
      OCCXC_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
      c->stage_ms[e.first] += ms;
      c->stage_n[e.first] += 1;
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  c->pending.clear();
}

// Restricted meta-GGA build on stream s: compute_K_dft<2, Restricted>
// (dft.h:190-388) with evaluate_density<2, R> (density.cpp:75-102) and
// xc_potential_matrix<Restricted, 2> (xc_potential_matrix.cpp:44-58).  rho
// and grad rho come from the density kernel in store mode; each direction's
// rowsum((grad_k phi D2) o grad_k phi) from the same kernel run on that
// derivative plane of the AO cache in place of phi (a.deriv = 0); then the
// per-leaf meta-GGA functional stage.  Vxc: the symmetric kernel takes the
// GGA term and the three tau terms of a leaf in one launch (one scatter);
// larger leaves run the general kernel on the GGA term and on each
// derivative plane in its LDA form with the tau coefficients.  The laplacian (density.cpp:88-93)
// is not formed: no functional here depends on it and the reference's Vxc
// does not use vlaplacian.
// The meta-GGA density stage: rho and grad rho, the three per-direction tau
// passes, then the per-leaf functional stage (coefficients in a.coef and
// c->d_coef_tau, {Exc, N_elec} per leaf, leaf skip flags).
static void mgga_density(occxc_context *c, hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  BuildArgs x = a;
  x.rho_out = c->d_mrho.get();
  run_density_store(x, c->d_ao.get(), prec, c->d_D2.get(), s);
  for (int k = 0; k < 3; ++k) {
    BuildArgs y = a;
    y.deriv = 0;
    y.leaves = c->d_leaves_k.get() + (size_t)k * c->nleaves;
    y.rho_out = c->d_mtau.get() + (size_t)k * c->ptot;
    run_density_store(y, c->d_ao.get(), prec, c->d_D2.get(), s);
  }
  launch_functional_mgga_leaves(a, c->d_mrho.get(), c->d_mtau.get(), c->d_coef_tau.get(), s);
}

// FP32 AO storage: N_elec += sum D2 (.) dS (run_storage_overlap), the
// first-order rounding change of sum_p w_p rho_p; D2 is the build's
static void storage_correction(occxc_context *c, double *d_result, hipStream_t s) {
  if (c->d_dS.get() == nullptr) return;
  launch_storage_correction(c->d_D2.get(), c->d_dS.get(), c->d_mols.get(), c->nmol, c->maxnbf,
                            c->d_ncorr.get(), d_result, s);
}

static void run_build_mgga(occxc_context *c, const double *d_D, double *d_V, double *d_result,
                           hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density(d_D, c->d_D2.get(), c->d_mols.get(), c->nmol, c->maxnbf, s);
    launch_zero(c->d_K.get(), c->mtot, s);
  });
  t.run(OCCXC_STAGE_DENSITY, [&] { mgga_density(c, s); });
  t.run(OCCXC_STAGE_VXC, [&] {
    BuildArgs x = a;
    x.coef_tau = c->d_coef_tau.get();
    x.coef_tau_planes = 1;
    run_vxc_mgga(c->plan, x, c->d_leaves_k.get(), c->d_ao.get(), prec, s);
  });
  t.run(OCCXC_STAGE_FINALIZE, [&] {
    launch_finalize(c->d_K.get(), d_V, c->d_mols.get(), c->nmol, c->maxnbf, c->d_part.get(),
                    d_result, s);
    storage_correction(c, d_result, s);
  });
}

void run_build(occxc_context *c, const double *d_D, double *d_V, double *d_result,
               hipStream_t s) {
  if (c->mgga) {
    run_build_mgga(c, d_D, d_V, d_result, s);
    return;
  }
  const BuildArgs &a = c->args;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density(d_D, c->d_D2.get(), c->d_mols.get(), c->nmol, c->maxnbf, s);
    if (!c->plan.two_pass) launch_zero(c->d_K.get(), c->mtot, s);
  });
  if (!c->plan.use_mfma) {
    const bool fp32 = c->opt.precision != OCCXC_PREC_FP64;
    t.run(OCCXC_STAGE_DENSITY, [&] { launch_density_functional(a, c->d_ao.get(), fp32, s); });
    t.run(OCCXC_STAGE_VXC, [&] { launch_vxc(a, c->d_ao.get(), fp32, s); });
  } else {
    t.run(OCCXC_STAGE_DENSITY, [&] { run_density(c->plan, a, c->d_ao.get(), c->opt.precision, s); });
    t.run(OCCXC_STAGE_VXC, [&] { run_vxc(c->plan, a, c->d_ao.get(), c->opt.precision, s); });
  }
  t.run(OCCXC_STAGE_FINALIZE, [&] {
    if (c->plan.two_pass) {
      run_vxc_reduce(a, d_V, s);
      launch_finalize_scalars(c->d_mols.get(), c->nmol, c->d_part.get(), d_result, s);
    } else {
      launch_finalize(c->d_K.get(), d_V, c->d_mols.get(), c->nmol, c->maxnbf, c->d_part.get(),
                      d_result, s);
    }
    storage_correction(c, d_result, s);
  });
}

// VV10 on stream s (NonLocalCorrelationFunctional::vv10,
// nonlocal_correlation.cpp:246-306): D2, rho and grad rho of every leaf point
// (density kernel in store mode), the VV10 pair sums and coefficients
// (nlc.hip), then the restricted Vxc kernel and the finalisation of a build.
static void run_nlc(occxc_context *c, const double *d_D, double b, double C, double *d_V,
                    double *d_result, hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density(d_D, c->d_D2.get(), c->d_mols.get(), c->nmol, c->maxnbf, s);
    if (!c->plan.two_pass) launch_zero(c->d_K.get(), c->mtot, s);
  });
  t.run(OCCXC_STAGE_DENSITY, [&] {
    BuildArgs x = a;
    x.rho_out = c->d_mrho.get();
    run_density_store(x, c->d_ao.get(), prec, c->d_D2.get(), s);
  });
  t.run(OCCXC_STAGE_NLC, [&] {
    launch_vv10(a, c->d_mrho.get(), c->d_porig.get(), b, C, c->d_nlc.get(), c->npts, s);
  });
  t.run(OCCXC_STAGE_VXC, [&] { run_vxc(c->plan, a, c->d_ao.get(), prec, s); });
  t.run(OCCXC_STAGE_FINALIZE, [&] {
    if (c->plan.two_pass) {
      run_vxc_reduce(a, d_V, s);
      launch_finalize_scalars(c->d_mols.get(), c->nmol, c->d_part.get(), d_result, s);
    } else {
      launch_finalize(c->d_K.get(), d_V, c->d_mols.get(), c->nmol, c->maxnbf, c->d_part.get(),
                      d_result, s);
    }
  });
}

static void nlc_check_and_setup(occxc_context *c) {
  OCCXC_REQUIRE(c->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                "occxc_nlc_vv10 needs a single-molecule context");
  OCCXC_REQUIRE(c->deriv == 1, OCCXC_ERR_UNSUPPORTED,
                "occxc_nlc_vv10 needs a context built with a GGA or meta-GGA functional");
  OCCXC_REQUIRE(c->opt.precision == OCCXC_PREC_FP64 || c->opt.precision == OCCXC_PREC_FP32_CACHE,
                OCCXC_ERR_UNSUPPORTED, "occxc_nlc_vv10 needs an FP64 or FP32-cache context");
  OCCXC_REQUIRE(c->plan.use_mfma && c->plan.max_mp <= kStoreMaxMp, OCCXC_ERR_UNSUPPORTED,
                "occxc_nlc_vv10 needs the matrix-core kernels and leaves of at most " +
                    std::to_string(kStoreMaxMp) + " AOs");
  if (c->d_mrho.size() == 0) {
    c->d_mrho.alloc(4 * (size_t)std::max<long long>(c->ptot, 1));
    c->device_bytes += c->d_mrho.bytes();
  }
  if (c->d_nlc.size() == 0) {
    c->d_nlc.alloc(std::max<size_t>(nlc_buffer_elems(c->npts), 1));
    c->device_bytes += c->d_nlc.bytes();
  }
}

// XC gradient set-up: the AO derivative components the cache lacks (once
// per context) and the per-atom lists of (leaf, local AO) entries for this
// shell-to-atom map (rebuilt when the map changes).
void gradient_setup(occxc_context *c, int natoms, const int *shell_atom) {
  if (!c->ao2_ready) {
    const size_t n = c->ao_elems / c->ncomp * (c->deriv ? 6 : 3);
    c->d_ao2.alloc(std::max<size_t>(n, 1));
    c->d_gpart.alloc(std::max<size_t>(3 * c->d_ao_map.size(), 1));
    c->device_bytes += c->d_ao2.bytes() + c->d_gpart.bytes();
    if (c->mgga) {
      c->d_gpart2.alloc(c->d_gpart.size());
      c->device_bytes += c->d_gpart2.bytes();
    }
    StageTimer t{c, c->stream, {}};
    t.run(OCCXC_STAGE_AO,
          [&] { launch_ao_extra(c->args, c->d_ao2.get(), c->deriv != 0, c->stream); });
    OCCXC_HIP(hipStreamSynchronize(c->stream));
    c->ao2_ready = true;
  }
  const bool same = c->grad_natoms == natoms && (int)c->grad_shell_atom.size() == c->nsh &&
                    std::equal(shell_atom, shell_atom + c->nsh, c->grad_shell_atom.begin());
  if (same) return;
  std::vector<int> map(c->d_ao_map.size());
  c->d_ao_map.download(map.data(), map.size(), c->stream);
  OCCXC_HIP(hipStreamSynchronize(c->stream));
  std::vector<int> off(natoms + 1, 0);
  for (const LeafDev &L : c->leaves)
    for (int mu = 0; mu < L.m; ++mu) ++off[shell_atom[c->bf_shell[map[L.map_off + mu]]] + 1];
  for (int A = 0; A < natoms; ++A) off[A + 1] += off[A];
  std::vector<int> ent(std::max(off[natoms], 1)), fill(off.begin(), off.end() - 1);
  for (const LeafDev &L : c->leaves)
    for (int mu = 0; mu < L.m; ++mu)
      ent[fill[shell_atom[c->bf_shell[map[L.map_off + mu]]]]++] = L.map_off + mu;
  c->d_gatom_off.alloc(off.size());
  c->d_gatom_off.upload(off.data(), off.size(), c->stream);
  c->d_gatom_ent.alloc(ent.size());
  c->d_gatom_ent.upload(ent.data(), ent.size(), c->stream);
  c->d_grad.alloc(3 * (size_t)natoms);
  OCCXC_HIP(hipStreamSynchronize(c->stream));
  c->grad_shell_atom.assign(shell_atom, shell_atom + c->nsh);
  c->grad_natoms = natoms;
}

// One XC gradient on stream s: D2, the density stage (coefficients, leaf
// skip), then the gradient contraction and the per-atom reduction into d_grad.
void run_gradient(occxc_context *c, const double *d_D, double *d_grad, hipStream_t s) {
  const BuildArgs &a = c->args;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density(d_D, c->d_D2.get(), c->d_mols.get(), c->nmol, c->maxnbf, s);
  });
  if (c->mgga) {
    // process_grid_block_gradient<2, Restricted>: the meta-GGA density stage,
    // the GGA terms and the tau terms as two partial sets, one reduction
    const bool fp32 = c->opt.precision != OCCXC_PREC_FP64;
    t.run(OCCXC_STAGE_DENSITY, [&] { mgga_density(c, s); });
    t.run(OCCXC_STAGE_GRADIENT, [&] {
      launch_xc_gradient_partials(a, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart.get(), s);
      BuildArgs x = a;
      x.coef_tau = c->d_coef_tau.get();
      x.coef_tau_planes = 1;
      launch_xc_gradient_tau(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart2.get(), s);
      launch_grad_reduce(c->d_gpart.get(), c->d_gpart2.get(), c->grad_natoms,
                         c->d_gatom_off.get(), c->d_gatom_ent.get(), d_grad, s);
    });
    return;
  }
  t.run(OCCXC_STAGE_DENSITY, [&] {
    if (c->plan.use_mfma)
      run_density(c->plan, a, c->d_ao.get(), c->opt.precision, s);
    else
      launch_density_functional(a, c->d_ao.get(), c->opt.precision != OCCXC_PREC_FP64, s);
  });
  t.run(OCCXC_STAGE_GRADIENT, [&] {
    launch_xc_gradient(a, c->d_ao.get(), c->opt.precision != OCCXC_PREC_FP64, c->d_ao2.get(),
                       c->d_gpart.get(), c->grad_natoms, c->d_gatom_off.get(),
                       c->d_gatom_ent.get(), d_grad, s);
  });
}

// Unrestricted build on stream s from the raw density in d_Du:
// compute_K_dft<d, Unrestricted> (dft.h:190-388).  Each spin density is
// contracted by the density kernel in store mode, the polarised functional
// stage forms both spins' coefficients, and the Vxc kernels run once per spin.
// The unrestricted meta-GGA density stage (evaluate_density<2, U>,
// density.cpp:104-166): each spin's rho and grad rho, each spin's three
// per-direction tau passes on the derivative planes, then the polarised
// meta-GGA functional stage.
static void mgga_density_u(occxc_context *c, hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  const double *D2[2] = {c->d_D2a.get(), c->d_D2b.get()};
  double *rho[2] = {c->d_rho_a.get(), c->d_rho_b.get()};
  double *tk[2] = {c->d_mtau.get(), c->d_mtau_b.get()};
  for (int sp = 0; sp < 2; ++sp) {
    BuildArgs x = a;
    x.rho_out = rho[sp];
    run_density_store(x, c->d_ao.get(), prec, D2[sp], s);
    for (int k = 0; k < 3; ++k) {
      BuildArgs y = a;
      y.deriv = 0;
      y.leaves = c->d_leaves_k.get() + (size_t)k * c->nleaves;
      y.rho_out = tk[sp] + (size_t)k * c->ptot;
      run_density_store(y, c->d_ao.get(), prec, D2[sp], s);
    }
  }
  launch_functional_mgga_pol_leaves(a, rho[0], rho[1], tk[0], tk[1], c->d_coef_b.get(),
                                    c->d_coef_tau.get(), c->d_coef_tau_b.get(),
                                    c->d_part_u.get(), s);
}

void run_build_u(occxc_context *c, hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  const bool two_pass = c->plan.two_pass;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density_u(c->d_Du.get(), c->d_D2a.get(), c->d_D2b.get(), c->nbf, s);
    if (!two_pass) {
      launch_zero(c->d_K.get(), c->mtot, s);
      launch_zero(c->d_K_b.get(), c->mtot, s);
    }
  });
  if (c->mgga) {
    t.run(OCCXC_STAGE_DENSITY, [&] { mgga_density_u(c, s); });
    t.run(OCCXC_STAGE_VXC, [&] {
      BuildArgs x = a;
      x.coef_tau = c->d_coef_tau.get();
      x.coef_tau_planes = 1;
      run_vxc_mgga(c->plan, x, c->d_leaves_k.get(), c->d_ao.get(), prec, s);
      x.coef = c->d_coef_b.get();
      x.coef_tau = c->d_coef_tau_b.get();
      x.coef_tau_planes = 1;
      x.K = c->d_K_b.get();
      run_vxc_mgga(c->plan, x, c->d_leaves_k.get(), c->d_ao.get(), prec, s);
    });
  } else {
  t.run(OCCXC_STAGE_DENSITY, [&] {
    BuildArgs x = a;
    x.rho_out = c->d_rho_a.get();
    run_density_store(x, c->d_ao.get(), prec, c->d_D2a.get(), s);
    x.rho_out = c->d_rho_b.get();
    run_density_store(x, c->d_ao.get(), prec, c->d_D2b.get(), s);
    launch_functional_pol_leaves(a, c->d_rho_a.get(), c->d_rho_b.get(), c->d_coef_b.get(),
                                 c->d_part_u.get(), s);
  });
  t.run(OCCXC_STAGE_VXC, [&] {
    run_vxc(c->plan, a, c->d_ao.get(), prec, s);
    if (two_pass) run_vxc_reduce(a, c->d_V.get(), s);
    BuildArgs b = a;
    b.coef = c->d_coef_b.get();
    b.K = c->d_K_b.get();
    run_vxc(c->plan, b, c->d_ao.get(), prec, s);
    if (two_pass) run_vxc_reduce(b, c->d_V_b.get(), s);
  });
  }
  t.run(OCCXC_STAGE_FINALIZE, [&] {
    launch_finalize_u(two_pass ? nullptr : c->d_K.get(), c->d_K_b.get(), c->d_V.get(),
                      c->d_V_b.get(), c->d_mols.get(), c->maxnbf, c->d_part_u.get(),
                      c->nleaves, c->d_result_u.get(), s);
  });
}

void uks_setup(occxc_context *c);

// Unrestricted XC gradient on stream s from the raw density in d_Du: both
// spin densities and the polarised functional stage as in run_build_u, then
// the gradient partials once per spin (P_s = 2 D_s, that spin's
// coefficients) and one per-atom reduction of both
// (process_grid_block_gradient<d, Unrestricted>, dft_gradient_kernels.cpp:
// 473-583).
void run_gradient_u(occxc_context *c, double *d_grad, hipStream_t s) {
  const BuildArgs &a = c->args;
  const int prec = c->opt.precision;
  StageTimer t{c, s, {}};
  t.run(OCCXC_STAGE_PREPARE, [&] {
    launch_prepare_density_u(c->d_Du.get(), c->d_D2a.get(), c->d_D2b.get(), c->nbf, s);
  });
  if (c->mgga) {
    // process_grid_block_gradient<2, Unrestricted>: the unrestricted
    // meta-GGA density stage, per spin the GGA and tau partial sets, one
    // reduction of all four
    const bool fp32 = prec != OCCXC_PREC_FP64;
    t.run(OCCXC_STAGE_DENSITY, [&] { mgga_density_u(c, s); });
    t.run(OCCXC_STAGE_GRADIENT, [&] {
      BuildArgs x = a;
      x.D2 = c->d_D2a.get();
      x.coef_tau = c->d_coef_tau.get();
      x.coef_tau_planes = 1;
      launch_xc_gradient_partials(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart.get(), s);
      launch_xc_gradient_tau(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart3.get(), s);
      x.D2 = c->d_D2b.get();
      x.coef = c->d_coef_b.get();
      x.coef_tau = c->d_coef_tau_b.get();
      x.coef_tau_planes = 1;
      launch_xc_gradient_partials(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart2.get(), s);
      launch_xc_gradient_tau(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart4.get(), s);
      launch_grad_reduce(c->d_gpart.get(), c->d_gpart2.get(), c->grad_natoms,
                         c->d_gatom_off.get(), c->d_gatom_ent.get(), d_grad, s,
                         c->d_gpart3.get(), c->d_gpart4.get());
    });
    return;
  }
  t.run(OCCXC_STAGE_DENSITY, [&] {
    BuildArgs x = a;
    x.rho_out = c->d_rho_a.get();
    run_density_store(x, c->d_ao.get(), prec, c->d_D2a.get(), s);
    x.rho_out = c->d_rho_b.get();
    run_density_store(x, c->d_ao.get(), prec, c->d_D2b.get(), s);
    launch_functional_pol_leaves(a, c->d_rho_a.get(), c->d_rho_b.get(), c->d_coef_b.get(),
                                 c->d_part_u.get(), s);
  });
  t.run(OCCXC_STAGE_GRADIENT, [&] {
    const bool fp32 = prec != OCCXC_PREC_FP64;
    BuildArgs x = a;
    x.D2 = c->d_D2a.get();
    launch_xc_gradient_partials(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart.get(), s);
    x.D2 = c->d_D2b.get();
    x.coef = c->d_coef_b.get();
    launch_xc_gradient_partials(x, c->d_ao.get(), fp32, c->d_ao2.get(), c->d_gpart2.get(), s);
    launch_grad_reduce(c->d_gpart.get(), c->d_gpart2.get(), c->grad_natoms,
                       c->d_gatom_off.get(), c->d_gatom_ent.get(), d_grad, s);
  });
}

void uks_setup(occxc_context *c) {
  if (c->uks_ready) return;
  const size_t nn = (size_t)c->nbf * c->nbf;
  c->d_Du.alloc(2 * nn);
  c->d_D2a.alloc(nn);
  c->d_D2b.alloc(nn);
  c->d_rho_a.alloc(4 * (size_t)c->ptot);
  c->d_rho_b.alloc(4 * (size_t)c->ptot);
  c->d_coef_b.alloc(4 * (size_t)c->ptot);
  c->d_K_b.alloc(nn);
  c->d_V_b.alloc(nn);
  c->d_part_u.alloc(3 * (size_t)std::max(c->nleaves, 1));
  c->d_result_u.alloc(3);
  c->device_bytes += 8.0 * (7 * nn + 12 * (size_t)c->ptot + 3 * (size_t)c->nleaves + 3);
  if (c->mgga) {
    c->d_mtau_b.alloc(3 * (size_t)c->ptot);
    c->d_coef_tau_b.alloc(c->ptot);
    c->device_bytes += 8.0 * 4 * (size_t)c->ptot;
  }
  c->uks_ready = true;
}

void drop_graph(occxc_context *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->graph_exec = nullptr;
  for (auto &k : c->graph_key) k = nullptr;
}

// One build on stream s: directly, or through the context's captured graph.
void launch_build(occxc_context *c, const double *d_D, double *d_V, double *d_result,
                  hipStream_t s) {
  if (!c->use_graph) {
    run_build(c, d_D, d_V, d_result, s);
    return;
  }
  const void *key[4] = {d_D, d_V, d_result, s};
  if (!c->graph_exec || !std::equal(key, key + 4, c->graph_key)) {
    drop_graph(c);
    const bool timing = c->timing;
    c->timing = false;
    hipGraph_t g = nullptr;
    OCCXC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      run_build(c, d_D, d_V, d_result, s);
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      c->timing = timing;
      throw;
    }
    c->timing = timing;
    OCCXC_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    OCCXC_HIP(e);
    std::copy(key, key + 4, c->graph_key);
  }
  OCCXC_HIP(hipGraphLaunch(c->graph_exec, s));
}

void context_destroy(occxc_context *c) {
  if (!c) return;
  DeviceGuard g(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  c->stream = nullptr;
  try {
    collect_timing(c);
  } catch (...) {
  }
  drop_graph(c);
  delete c;
}

} // namespace occxc

extern "C" {

void occxc_default_options(occxc_options *opt) {
  if (!opt) return;
  opt->precision = OCCXC_PREC_FP64;
  opt->density_threshold = 1e-10;
  opt->screening_threshold = 1e-10;
  opt->screen = 1;
  opt->device = 0;
  opt->tile_ao = 0;
  opt->persistent = -1;
}

occxc_status occxc_batch_create(int nmol, const occxc_basis *bases, const occxc_grid *grids,
                                const occxc_functional *func, const occxc_options *opt,
                                occxc_context **ctx) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx, OCCXC_ERR_INVALID_ARGUMENT, "ctx: null pointer");
    *ctx = nullptr;
    occxc_options o;
    occxc_default_options(&o);
    if (opt) o = *opt;
    auto c = std::make_unique<occxc_context>();
    try {
      context_setup(c.get(), nmol, bases, grids, func, &o);
    } catch (...) {
      if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = nullptr;
      }
      throw;
    }
    *ctx = c.release();
  });
}

occxc_status occxc_batch_estimate_bytes(int nmol, const occxc_basis *bases,
                                        const occxc_grid *grids, const occxc_functional *func,
                                        const occxc_options *opt, double *bytes) {
  return guarded([&] {
    OCCXC_REQUIRE(bytes, OCCXC_ERR_INVALID_ARGUMENT, "bytes: null pointer");
    occxc_options o;
    occxc_default_options(&o);
    if (opt) o = *opt;
    auto c = std::make_unique<occxc_context>();
    try {
      context_setup(c.get(), nmol, bases, grids, func, &o, 1);
    } catch (const Error &e) {
      if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = nullptr;
      }
      if (e.status != OCCXC_ERR_OUT_OF_MEMORY || c->device_bytes == 0) throw;
    }
    for (int m = 0; m < nmol; ++m) bytes[m] = c->mol_bytes[m];
  });
}

occxc_status occxc_context_create(const occxc_basis *basis, const occxc_grid *grid,
                                  const occxc_functional *func, const occxc_options *opt,
                                  occxc_context **ctx) {
  if (!basis || !grid) {
    set_last_error("basis/grid: null pointer");
    return OCCXC_ERR_INVALID_ARGUMENT;
  }
  return occxc_batch_create(1, basis, grid, func, opt, ctx);
}

occxc_status occxc_context_destroy(occxc_context *ctx) {
  return guarded([&] {
    if (!ctx) return;
    context_destroy(ctx);
  });
}

occxc_status occxc_context_get_info(const occxc_context *ctx, occxc_context_info *info) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && info, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    info->npts = ctx->npts;
    info->nleaves = ctx->nleaves;
    info->nbf = ctx->nmol == 1 ? ctx->nbf : ctx->maxnbf;
    info->nshell = ctx->nsh;
    info->mean_local_nbf = ctx->mean_m;
    info->contraction_flops = ctx->flops;
    info->ao_bytes = (double)ctx->ao_elems * ctx->ao_elem_size;
    info->ao_write_bytes = ctx->ao_write_bytes;
    info->device_bytes = (double)ctx->device_bytes;
    info->max_local_nbf = 0;
    for (const auto &L : ctx->leaves) info->max_local_nbf = std::max(info->max_local_nbf, L.m);
    info->tile_ao = ctx->plan.tiled ? ctx->plan.tune.tile_ao : 0;
    info->persistent = ctx->plan.tiled ? ctx->plan.tune.persistent : 0;
    info->tuning_row = ctx->plan.tiled ? ctx->plan.tune.table_row : -1;
  });
}

occxc_status occxc_build_device(occxc_context *ctx, const double *d_D, double *d_Vxc,
                                double *d_result, void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && d_D && d_Vxc && d_result, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    launch_build(ctx, d_D, d_Vxc, d_result, s);
  });
}

occxc_status occxc_batch_build(occxc_context *ctx, const double *const *D, double *const *Vxc,
                               double *exc, double *nelec) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && D && Vxc && exc && nelec, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    for (int m = 0; m < ctx->nmol; ++m) {
      OCCXC_REQUIRE(D[m] && Vxc[m], OCCXC_ERR_INVALID_ARGUMENT, "null D or Vxc pointer");
      const size_t nn = (size_t)ctx->mol_nbf[m] * ctx->mol_nbf[m];
      for (size_t i = 0; i < nn; ++i)
        OCCXC_REQUIRE(std::isfinite(D[m][i]), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    }
    for (int m = 0; m < ctx->nmol; ++m) {
      const size_t nn = (size_t)ctx->mol_nbf[m] * ctx->mol_nbf[m];
      OCCXC_HIP(hipMemcpyAsync(ctx->d_D.get() + ctx->mol_off[m], D[m], nn * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    launch_build(ctx, ctx->d_D.get(), ctx->d_V.get(), ctx->d_result.get(), ctx->stream);
    std::vector<double> res(2 * ctx->nmol);
    for (int m = 0; m < ctx->nmol; ++m) {
      const size_t nn = (size_t)ctx->mol_nbf[m] * ctx->mol_nbf[m];
      OCCXC_HIP(hipMemcpyAsync(Vxc[m], ctx->d_V.get() + ctx->mol_off[m], nn * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    ctx->d_result.download(res.data(), res.size(), ctx->stream);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
    for (int m = 0; m < ctx->nmol; ++m) {
      exc[m] = res[2 * m];
      nelec[m] = res[2 * m + 1];
    }
  });
}

occxc_status occxc_build(occxc_context *ctx, const double *D, double *Vxc, double *exc,
                         double *nelec) {
  if (ctx && ctx->nmol != 1) {
    set_last_error("occxc_build on a batch context: use occxc_batch_build");
    return OCCXC_ERR_INVALID_ARGUMENT;
  }
  const double *Ds[1] = {D};
  double *Vs[1] = {Vxc};
  return occxc_batch_build(ctx, Ds, Vs, exc, nelec);
}

occxc_status occxc_nlc_vv10(occxc_context *ctx, const double *D, double b, double C,
                            double *energy, double *Vxc) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && D && energy && Vxc, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "occxc_nlc_vv10 needs a single-molecule context");
    OCCXC_REQUIRE(std::isfinite(b) && b > 0.0 && std::isfinite(C) && C >= 0.0,
                  OCCXC_ERR_INVALID_ARGUMENT, "VV10 parameters: need b > 0, C >= 0");
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    for (size_t i = 0; i < nn; ++i)
      OCCXC_REQUIRE(std::isfinite(D[i]), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    DeviceGuard g(ctx->device);
    nlc_check_and_setup(ctx);
    ctx->d_D.upload(D, nn, ctx->stream);
    run_nlc(ctx, ctx->d_D.get(), b, C, ctx->d_V.get(), ctx->d_result.get(), ctx->stream);
    ctx->d_V.download(Vxc, nn, ctx->stream);
    double r[2];
    ctx->d_result.download(r, 2, ctx->stream);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
    *energy = r[0];
  });
}

occxc_status occxc_nlc_vv10_device(occxc_context *ctx, const double *d_D, double b,
                                   double C, double *d_Vxc, double *d_result, void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && d_D && d_Vxc && d_result, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    OCCXC_REQUIRE(std::isfinite(b) && b > 0.0 && std::isfinite(C) && C >= 0.0,
                  OCCXC_ERR_INVALID_ARGUMENT, "VV10 parameters: need b > 0, C >= 0");
    DeviceGuard g(ctx->device);
    nlc_check_and_setup(ctx);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_nlc(ctx, d_D, b, C, d_Vxc, d_result, s);
  });
}

occxc_status occxc_xc_gradient(occxc_context *ctx, const double *D, int natoms,
                               const int *shell_atom, double *grad) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && D && shell_atom && grad, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "occxc_xc_gradient needs a single-molecule context");
    OCCXC_REQUIRE(natoms > 0, OCCXC_ERR_INVALID_ARGUMENT, "natoms must be positive");
    for (int k = 0; k < ctx->nsh; ++k)
      OCCXC_REQUIRE(shell_atom[k] >= 0 && shell_atom[k] < natoms, OCCXC_ERR_INVALID_ARGUMENT,
                    "shell_atom: atom index out of range");
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    for (size_t i = 0; i < nn; ++i)
      OCCXC_REQUIRE(std::isfinite(D[i]), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    DeviceGuard g(ctx->device);
    gradient_setup(ctx, natoms, shell_atom);
    OCCXC_HIP(hipMemcpyAsync(ctx->d_D.get(), D, nn * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
    run_gradient(ctx, ctx->d_D.get(), ctx->d_grad.get(), ctx->stream);
    ctx->d_grad.download(grad, 3 * (size_t)natoms, ctx->stream);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

occxc_status occxc_xc_gradient_device(occxc_context *ctx, const double *d_D, int natoms,
                                      const int *shell_atom, double *d_grad, void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && d_D && shell_atom && d_grad, OCCXC_ERR_INVALID_ARGUMENT,
                  "null pointer");
    OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "occxc_xc_gradient needs a single-molecule context");
    OCCXC_REQUIRE(natoms > 0, OCCXC_ERR_INVALID_ARGUMENT, "natoms must be positive");
    for (int k = 0; k < ctx->nsh; ++k)
      OCCXC_REQUIRE(shell_atom[k] >= 0 && shell_atom[k] < natoms, OCCXC_ERR_INVALID_ARGUMENT,
                    "shell_atom: atom index out of range");
    DeviceGuard g(ctx->device);
    gradient_setup(ctx, natoms, shell_atom);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_gradient(ctx, d_D, d_grad, s);
  });
}

occxc_status occxc_build_unrestricted(occxc_context *ctx, const double *D, double *Vxc,
                                      double *exc, double *nelec) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && D && Vxc && exc && nelec, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "occxc_build_unrestricted needs a single-molecule context");
    OCCXC_REQUIRE(ctx->opt.precision == OCCXC_PREC_FP64 ||
                      ctx->opt.precision == OCCXC_PREC_FP32_CACHE,
                  OCCXC_ERR_UNSUPPORTED,
                  "unrestricted builds need an FP64 or FP32-cache context");
    OCCXC_REQUIRE(ctx->plan.use_mfma && ctx->plan.max_mp <= kStoreMaxMp, OCCXC_ERR_UNSUPPORTED,
                  "unrestricted builds need the matrix-core kernels and leaves of at most " +
                      std::to_string(kStoreMaxMp) + " AOs");
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    for (size_t i = 0; i < 2 * nn; ++i)
      OCCXC_REQUIRE(std::isfinite(D[i]), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    DeviceGuard g(ctx->device);
    uks_setup(ctx);
    ctx->d_Du.upload(D, 2 * nn, ctx->stream);
    run_build_u(ctx, ctx->stream);
    const size_t row = (size_t)ctx->nbf * sizeof(double);
    OCCXC_HIP(hipMemcpy2DAsync(Vxc, 2 * row, ctx->d_V.get(), row, row, ctx->nbf,
                               hipMemcpyDeviceToHost, ctx->stream));
    OCCXC_HIP(hipMemcpy2DAsync(Vxc + ctx->nbf, 2 * row, ctx->d_V_b.get(), row, row, ctx->nbf,
                               hipMemcpyDeviceToHost, ctx->stream));
    double r[3];
    ctx->d_result_u.download(r, 3, ctx->stream);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
    *exc = r[0];
    nelec[0] = r[1];
    nelec[1] = r[2];
  });
}

occxc_status occxc_build_unrestricted_device(occxc_context *ctx, const double *d_D,
                                             double *d_Vxc, double *d_result, void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && d_D && d_Vxc && d_result, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                  "occxc_build_unrestricted needs a single-molecule context");
    OCCXC_REQUIRE(ctx->opt.precision == OCCXC_PREC_FP64 ||
                      ctx->opt.precision == OCCXC_PREC_FP32_CACHE,
                  OCCXC_ERR_UNSUPPORTED,
                  "unrestricted builds need an FP64 or FP32-cache context");
    OCCXC_REQUIRE(ctx->plan.use_mfma && ctx->plan.max_mp <= kStoreMaxMp, OCCXC_ERR_UNSUPPORTED,
                  "unrestricted builds need the matrix-core kernels and leaves of at most " +
                      std::to_string(kStoreMaxMp) + " AOs");
    DeviceGuard g(ctx->device);
    uks_setup(ctx);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    OCCXC_HIP(hipMemcpyAsync(ctx->d_Du.get(), d_D, 2 * nn * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    run_build_u(ctx, s);
    const size_t row = (size_t)ctx->nbf * sizeof(double);
    OCCXC_HIP(hipMemcpy2DAsync(d_Vxc, 2 * row, ctx->d_V.get(), row, row, ctx->nbf,
                               hipMemcpyDeviceToDevice, s));
    OCCXC_HIP(hipMemcpy2DAsync(d_Vxc + ctx->nbf, 2 * row, ctx->d_V_b.get(), row, row, ctx->nbf,
                               hipMemcpyDeviceToDevice, s));
    OCCXC_HIP(hipMemcpyAsync(d_result, ctx->d_result_u.get(), 3 * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
  });
}

// Checks and set-up shared by the two unrestricted gradient entry points:
// the context and shell map, the Hessian cache and per-atom lists, the
// unrestricted buffers and the extra partial sets.
static void gradient_u_prepare(occxc_context *ctx, int natoms, const int *shell_atom) {
  OCCXC_REQUIRE(ctx->nmol == 1, OCCXC_ERR_INVALID_ARGUMENT,
                "occxc_xc_gradient needs a single-molecule context");
  OCCXC_REQUIRE(natoms > 0, OCCXC_ERR_INVALID_ARGUMENT, "natoms must be positive");
  for (int k = 0; k < ctx->nsh; ++k)
    OCCXC_REQUIRE(shell_atom[k] >= 0 && shell_atom[k] < natoms, OCCXC_ERR_INVALID_ARGUMENT,
                  "shell_atom: atom index out of range");
  OCCXC_REQUIRE(ctx->opt.precision == OCCXC_PREC_FP64 ||
                    ctx->opt.precision == OCCXC_PREC_FP32_CACHE,
                OCCXC_ERR_UNSUPPORTED,
                "unrestricted gradients need an FP64 or FP32-cache context");
  OCCXC_REQUIRE(ctx->plan.use_mfma && ctx->plan.max_mp <= kStoreMaxMp, OCCXC_ERR_UNSUPPORTED,
                "unrestricted gradients need the matrix-core kernels and leaves of at most " +
                    std::to_string(kStoreMaxMp) + " AOs");
  gradient_setup(ctx, natoms, shell_atom);
  uks_setup(ctx);
  if (ctx->mgga && ctx->d_gpart3.size() < ctx->d_gpart.size()) {
    ctx->d_gpart3.alloc(ctx->d_gpart.size());
    ctx->d_gpart4.alloc(ctx->d_gpart.size());
    ctx->device_bytes += ctx->d_gpart3.bytes() + ctx->d_gpart4.bytes();
  }
  if (ctx->d_gpart2.size() < ctx->d_gpart.size()) {
    ctx->d_gpart2.alloc(ctx->d_gpart.size());
    ctx->device_bytes += ctx->d_gpart2.bytes();
  }
}

occxc_status occxc_xc_gradient_unrestricted(occxc_context *ctx, const double *D, int natoms,
                                            const int *shell_atom, double *grad) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && D && shell_atom && grad, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    for (size_t i = 0; i < 2 * nn; ++i)
      OCCXC_REQUIRE(std::isfinite(D[i]), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    DeviceGuard g(ctx->device);
    gradient_u_prepare(ctx, natoms, shell_atom);
    ctx->d_Du.upload(D, 2 * nn, ctx->stream);
    run_gradient_u(ctx, ctx->d_grad.get(), ctx->stream);
    ctx->d_grad.download(grad, 3 * (size_t)natoms, ctx->stream);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

occxc_status occxc_xc_gradient_unrestricted_device(occxc_context *ctx, const double *d_D,
                                                   int natoms, const int *shell_atom,
                                                   double *d_grad, void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && d_D && shell_atom && d_grad, OCCXC_ERR_INVALID_ARGUMENT,
                  "null pointer");
    DeviceGuard g(ctx->device);
    gradient_u_prepare(ctx, natoms, shell_atom);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    OCCXC_HIP(hipMemcpyAsync(ctx->d_Du.get(), d_D, 2 * nn * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    run_gradient_u(ctx, d_grad, s);
  });
}

occxc_status occxc_batch_layout(const occxc_context *ctx, int *nmol, int *nbf,
                                long long *offsets) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && nmol, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    *nmol = ctx->nmol;
    for (int m = 0; m < ctx->nmol; ++m) {
      if (nbf) nbf[m] = ctx->mol_nbf[m];
      if (offsets) offsets[m] = ctx->mol_off[m];
    }
  });
}

occxc_status occxc_synchronize(occxc_context *ctx) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

occxc_status occxc_context_evaluate_ao(occxc_context *ctx) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    StageTimer t{ctx, ctx->stream, {}};
    t.run(OCCXC_STAGE_AO, [&] {
      launch_ao_eval(ctx->args, ctx->d_ao.get(), ctx->opt.precision != 0, ctx->stream);
    });
  });
}

occxc_status occxc_context_set_graph(occxc_context *ctx, int enable) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    OCCXC_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->use_graph = enable != 0;
  });
}

occxc_status occxc_context_set_timing(occxc_context *ctx, int enable) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    collect_timing(ctx);
    for (int i = 0; i < OCCXC_NSTAGES; ++i) {
      ctx->stage_ms[i] = 0.0;
      ctx->stage_n[i] = 0;
    }
    ctx->timing = enable != 0;
  });
}

occxc_status occxc_context_get_timing(occxc_context *ctx, double *ms, long long *counts) {
  return guarded([&] {
    OCCXC_REQUIRE(ctx && ms && counts, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    collect_timing(ctx);
    for (int i = 0; i < OCCXC_NSTAGES; ++i) {
      ms[i] = ctx->stage_ms[i];
      counts[i] = ctx->stage_n[i];
    }
  });
}

occxc_status occxc_evaluate_basis(const occxc_basis *basis, long long npts, const double *points,
                                  int deriv, double *out, int device) {
  return guarded([&] {
    OCCXC_REQUIRE(npts >= 0 && (npts == 0 || (points && out)) && (deriv == 0 || deriv == 1),
                  OCCXC_ERR_INVALID_ARGUMENT, "evaluate_basis: bad arguments");
    HostBasis hb = make_basis(basis);
    DeviceGuard g(device);
    ensure_device_tables(device);
    DevBuf<ShellDev> ds(hb.shells.size());
    DevBuf<double> de(hb.exps.size()), dc(hb.coefs.size()), dp(3 * std::max<long long>(npts, 1));
    const size_t nout = (size_t)(deriv ? 4 : 1) * hb.nbf * npts;
    DevBuf<double> dout(std::max<size_t>(nout, 1));
    ds.upload(hb.shells.data(), hb.shells.size());
    de.upload(hb.exps.data(), hb.exps.size());
    dc.upload(hb.coefs.data(), hb.coefs.size());
    dp.upload(points, 3 * npts);
    launch_eval_basis_flat(ds.get(), hb.nsh, de.get(), dc.get(), dp.get(), npts, deriv, hb.nbf,
                           dout.get(), nullptr);
    dout.download(out, nout);
    OCCXC_HIP(hipDeviceSynchronize());
  });
}

occxc_status occxc_evaluate_decay_cutoff(const occxc_basis *basis, double threshold,
                                         double *extents, int device) {
  return guarded([&] {
    OCCXC_REQUIRE(extents, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    HostBasis hb = make_basis(basis);
    DeviceGuard g(device);
    ensure_device_tables(device);
    DevBuf<ShellDev> ds(hb.shells.size());
    DevBuf<double> de(hb.exps.size()), dc(hb.coefs.size()), dx(hb.nsh);
    ds.upload(hb.shells.data(), hb.shells.size());
    de.upload(hb.exps.data(), hb.exps.size());
    dc.upload(hb.coefs.data(), hb.coefs.size());
    launch_decay_cutoff(ds.get(), hb.nsh, de.get(), dc.get(), threshold, dx.get(), nullptr);
    dx.download(extents, hb.nsh);
    OCCXC_HIP(hipDeviceSynchronize());
  });
}

occxc_status occxc_functional_evaluate(int id, long long n, const double *rho,
                                       const double *sigma, double *exc, double *vrho,
                                       double *vsigma, int device) {
  return guarded([&] {
    const int fam = xc_family(id);
    OCCXC_REQUIRE(fam >= 0, OCCXC_ERR_UNSUPPORTED, "functional id " + std::to_string(id) + " not found");
    OCCXC_REQUIRE(n >= 0 && (n == 0 || (rho && exc && vrho)), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate: null pointer");
    OCCXC_REQUIRE(fam == 0 || n == 0 || (sigma && vsigma), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate: GGA needs sigma and vsigma");
    if (n == 0) return;
    DeviceGuard g(device);
    DevBuf<double> dr(n), ds(fam ? n : 1), de(n), dv(n), dw(fam ? n : 1);
    dr.upload(rho, n);
    if (fam) ds.upload(sigma, n);
    launch_functional(id, n, dr.get(), fam ? ds.get() : nullptr, de.get(), dv.get(),
                      fam ? dw.get() : nullptr, nullptr);
    std::vector<double> e(n), v(n), w(fam ? n : 0);
    de.download(e.data(), n);
    dv.download(v.data(), n);
    if (fam) dw.download(w.data(), n);
    OCCXC_HIP(hipDeviceSynchronize());
    for (long long i = 0; i < n; ++i) {
      exc[i] += e[i];
      vrho[i] += v[i];
      if (fam) vsigma[i] += w[i];
    }
  });
}

occxc_status occxc_functional_evaluate_mgga(int id, long long n, const double *rho,
                                            const double *sigma, const double *tau,
                                            double *exc, double *vrho, double *vsigma,
                                            double *vtau, int device) {
  return guarded([&] {
    const int fam = xc_family_mgga(id);
    OCCXC_REQUIRE(fam >= 0, OCCXC_ERR_UNSUPPORTED, "functional id " + std::to_string(id) + " not found");
    OCCXC_REQUIRE(n >= 0 && (n == 0 || (rho && exc && vrho)), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga: null pointer");
    OCCXC_REQUIRE(fam == 0 || n == 0 || (sigma && vsigma), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga: GGA and meta-GGA need sigma and vsigma");
    OCCXC_REQUIRE(fam < 2 || n == 0 || (tau && vtau), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga: meta-GGA needs tau and vtau");
    if (n == 0) return;
    DeviceGuard g(device);
    const bool gga = fam >= 1, mg = fam == 2;
    DevBuf<double> dr(n), ds(gga ? n : 1), dt(mg ? n : 1), de(n), dv(n), dw(gga ? n : 1),
        dx(mg ? n : 1);
    dr.upload(rho, n);
    if (gga) ds.upload(sigma, n);
    if (mg) dt.upload(tau, n);
    launch_zero(de.get(), n, nullptr);
    launch_zero(dv.get(), n, nullptr);
    launch_zero(dw.get(), gga ? n : 1, nullptr);
    launch_zero(dx.get(), mg ? n : 1, nullptr);
    launch_functional_mgga(id, n, dr.get(), ds.get(), dt.get(), de.get(), dv.get(), dw.get(),
                           dx.get(), nullptr);
    std::vector<double> e(n), v(n), w(gga ? n : 0), x(mg ? n : 0);
    de.download(e.data(), n);
    dv.download(v.data(), n);
    if (gga) dw.download(w.data(), n);
    if (mg) dx.download(x.data(), n);
    OCCXC_HIP(hipDeviceSynchronize());
    for (long long i = 0; i < n; ++i) {
      exc[i] += e[i];
      vrho[i] += v[i];
      if (gga) vsigma[i] += w[i];
      if (mg) vtau[i] += x[i];
    }
  });
}

occxc_status occxc_functional_evaluate_mgga_polarized(int id, long long n, const double *rho,
                                                      const double *sigma, const double *tau,
                                                      double *exc, double *vrho, double *vsigma,
                                                      double *vtau, int device) {
  return guarded([&] {
    const int fam = xc_family_mgga(id);
    OCCXC_REQUIRE(fam >= 0, OCCXC_ERR_UNSUPPORTED, "functional id " + std::to_string(id) + " not found");
    OCCXC_REQUIRE(n >= 0 && (n == 0 || (rho && exc && vrho)), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga_polarized: null pointer");
    OCCXC_REQUIRE(fam == 0 || n == 0 || (sigma && vsigma), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga_polarized: GGA and meta-GGA need sigma and vsigma");
    OCCXC_REQUIRE(fam < 2 || n == 0 || (tau && vtau), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_mgga_polarized: meta-GGA needs tau and vtau");
    if (n == 0) return;
    DeviceGuard g(device);
    const bool gga = fam >= 1, mg = fam == 2;
    DevBuf<double> dr(2 * n), ds(gga ? 3 * n : 1), dt(mg ? 2 * n : 1), de(n), dv(2 * n),
        dw(gga ? 3 * n : 1), dx(mg ? 2 * n : 1);
    dr.upload(rho, 2 * n);
    if (gga) ds.upload(sigma, 3 * n);
    if (mg) dt.upload(tau, 2 * n);
    launch_zero(de.get(), n, nullptr);
    launch_zero(dv.get(), 2 * n, nullptr);
    launch_zero(dw.get(), gga ? 3 * n : 1, nullptr);
    launch_zero(dx.get(), mg ? 2 * n : 1, nullptr);
    launch_functional_mgga_pol(id, n, dr.get(), ds.get(), dt.get(), de.get(), dv.get(),
                               dw.get(), dx.get(), nullptr);
    std::vector<double> e(n), v(2 * n), w(gga ? 3 * n : 0), x(mg ? 2 * n : 0);
    de.download(e.data(), n);
    dv.download(v.data(), 2 * n);
    if (gga) dw.download(w.data(), 3 * n);
    if (mg) dx.download(x.data(), 2 * n);
    OCCXC_HIP(hipDeviceSynchronize());
    for (long long i = 0; i < n; ++i) {
      exc[i] += e[i];
      for (int k = 0; k < 2; ++k) vrho[2 * i + k] += v[2 * i + k];
      if (gga)
        for (int k = 0; k < 3; ++k) vsigma[3 * i + k] += w[3 * i + k];
      if (mg)
        for (int k = 0; k < 2; ++k) vtau[2 * i + k] += x[2 * i + k];
    }
  });
}

occxc_status occxc_functional_evaluate_polarized(int id, long long n, const double *rho,
                                                 const double *sigma, double *exc, double *vrho,
                                                 double *vsigma, int device) {
  return guarded([&] {
    const int fam = xc_family(id);
    OCCXC_REQUIRE(fam >= 0, OCCXC_ERR_UNSUPPORTED, "functional id " + std::to_string(id) + " not found");
    OCCXC_REQUIRE(n >= 0 && (n == 0 || (rho && exc && vrho)), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_polarized: null pointer");
    OCCXC_REQUIRE(fam == 0 || n == 0 || (sigma && vsigma), OCCXC_ERR_INVALID_ARGUMENT,
                  "functional_evaluate_polarized: GGA needs sigma and vsigma");
    if (n == 0) return;
    DeviceGuard g(device);
    DevBuf<double> dr(2 * n), ds(fam ? 3 * n : 1), de(n), dv(2 * n), dw(fam ? 3 * n : 1);
    dr.upload(rho, 2 * n);
    if (fam) ds.upload(sigma, 3 * n);
    launch_zero(de.get(), n, nullptr);
    launch_zero(dv.get(), 2 * n, nullptr);
    if (fam) launch_zero(dw.get(), 3 * n, nullptr);
    launch_functional_pol(id, n, dr.get(), fam ? ds.get() : nullptr, de.get(), dv.get(),
                          fam ? dw.get() : nullptr, nullptr);
    std::vector<double> e(n), v(2 * n), w(fam ? 3 * n : 0);
    de.download(e.data(), n);
    dv.download(v.data(), 2 * n);
    if (fam) dw.download(w.data(), 3 * n);
    OCCXC_HIP(hipDeviceSynchronize());
    for (long long i = 0; i < n; ++i) {
      exc[i] += e[i];
      vrho[2 * i] += v[2 * i];
      vrho[2 * i + 1] += v[2 * i + 1];
      if (fam)
        for (int k = 0; k < 3; ++k) vsigma[3 * i + k] += w[3 * i + k];
    }
  });
}

occxc_status occxc_grid_hierarchy(long long npts, const double *points, const double *weights,
                                  double *sorted_points, double *sorted_weights, long long *perm,
                                  long long *leaf_offset, long long *leaf_count,
                                  double *leaf_center, double *leaf_radius, long long *nleaves) {
  return guarded([&] {
    OCCXC_REQUIRE(npts >= 0 && nleaves, OCCXC_ERR_INVALID_ARGUMENT, "bad arguments");
    Hierarchy h = build_hierarchy(npts, points, weights);
    std::copy(h.xyz.begin(), h.xyz.end(), sorted_points);
    std::copy(h.w.begin(), h.w.end(), sorted_weights);
    std::copy(h.perm.begin(), h.perm.end(), perm);
    std::copy(h.off.begin(), h.off.end(), leaf_offset);
    std::copy(h.cnt.begin(), h.cnt.end(), leaf_count);
    std::copy(h.center.begin(), h.center.end(), leaf_center);
    std::copy(h.radius.begin(), h.radius.end(), leaf_radius);
    *nleaves = (long long)h.off.size();
  });
}

occxc_status occxc_radial_grid(int method, int n, double param, double *r, double *w,
                               int *npoints_out) {
  return guarded([&] {
    OCCXC_REQUIRE(n >= 1 && r && w && npoints_out, OCCXC_ERR_INVALID_ARGUMENT, "bad arguments");
    const double pi = 3.14159265358979323846;
    switch (method) {
    case OCCXC_RADIAL_EULER_MACLAURIN: {
      // grid_utils.cpp:380-395 (returns n-1 points)
      const int m_r = 2;
      const double f = m_r * n * std::pow(param, 3);
      for (int i = 1; i < n; ++i) {
        const double nmi = n - i;
        r[i - 1] = param * std::pow(i / nmi, m_r);
        w[i - 1] = f * std::pow(i, 3 * m_r - 1) / std::pow(nmi, 3 * m_r + 1);
      }
      *npoints_out = n - 1;
      break;
    }
    case OCCXC_RADIAL_TREUTLER_AHLRICHS: {
      // grid_utils.cpp:561-574 (M4 mapping, no r^2, unscaled)
      const double step = pi / (n + 1);
      const double ln2 = 1.0 / std::log(2.0);
      for (int i = 1; i <= n; ++i) {
        const double x = std::cos(i * step);
        const double t1 = ln2 * std::pow(1.0 + x, 0.6);
        const double t2 = std::log((1.0 - x) / 2.0);
        r[n - i] = -t1 * t2;
        w[n - i] = step * std::sin(i * step) * t1 * (-0.6 / (1.0 + x) * t2 + 1.0 / (1.0 - x));
      }
      *npoints_out = n;
      break;
    }
    case OCCXC_RADIAL_GAUSS_CHEBYSHEV:
    case OCCXC_RADIAL_BECKE: {
      // grid_utils.cpp:337-344 and :545-556
      for (int i = 0; i < n; ++i) {
        const double lin = (n == 1) ? 1.0 : 1.0 + (double)i * (2.0 * n - 2.0) / (double)(n - 1);
        r[i] = std::cos(lin * (pi / (2.0 * n)));
        w[i] = pi / n;
      }
      if (method == OCCXC_RADIAL_BECKE) {
        for (int i = 0; i < n; ++i) {
          const double tp1 = 1.0 + r[i], tm1 = 1.0 - r[i];
          r[i] = (tp1 / tm1) * param;
          w[i] *= 2.0 / (tm1 * tm1) * param;
        }
      }
      *npoints_out = n;
      break;
    }
    case OCCXC_RADIAL_MURA_KNOWLES: {
      // grid_utils.cpp:517-542
      double far = 5.2;
      const int charge = (int)param;
      if (charge == 3 || charge == 4 || charge == 11 || charge == 12 || charge == 19 || charge == 20)
        far = 7.0;
      for (int i = 0; i < n; ++i) {
        const double x = (i + 0.5) / n, x2 = x * x, x3 = x2 * x;
        r[i] = -far * std::log(1.0 - x3);
        w[i] = far * 3.0 * x2 / ((1.0 - x3) * n);
      }
      *npoints_out = n;
      break;
    }
    default:
      throw Error(OCCXC_ERR_UNSUPPORTED, "unknown radial grid method");
    }
  });
}

// LMG radial rule (R. Lindh, P.-A. Malmqvist, L. Gagliardi, TCA 106, 178
// (2001)) as generate_lmg_radial_grid builds it (grid_utils.cpp:397-515): the
// inner radius from eq. 25 (m = 0, d = 1.9) at twice the largest exponent, the
// outer radius per angular momentum from eq. 19 and the log-grid step from
// eqs. 17-18, both found by the same bracketing walk (step x 0.1 on each
// sign change, until two iterates agree to 1e-14), then
// r_i = c (e^{(i+1) h} - 1), w_i = (r_i + c) h r_i^2, c = r_in / (e^h - 1).
// The expressions keep the reference's operation order: the point count is
// the integer part of log(1 + r_out / c) / h.
namespace {
double lmg_inner_radius(double eps, double alpha) {
  const int m = 0;
  double r = 1.9 - std::log(1.0 / eps);
  r = r * 2.0 / (m + 3.0);
  r = std::exp(r) / alpha;
  return std::sqrt(r);
}

double lmg_outer_radius(double eps, double alpha, int l, double guess) {
  const int m = 2 * l;
  double r = guess, prev = 1.0e50, step = 0.5, f = 1.0e50;
  double sign = f > eps ? 1.0 : -1.0;
  while (std::fabs(prev - r) > 1e-14) {
    const double c = std::tgamma((m + 3.0) / 2.0);
    const double a = std::pow(alpha * r * r, (m + 1.0) / 2.0);
    const double e = std::exp(-alpha * r * r);
    f = c * a * e;
    const double last = sign;
    sign = f > eps ? 1.0 : -1.0;
    if (r < 0.0) sign = 1.0;
    if (sign != last) step *= 0.1;
    prev = r;
    r += sign * step;
  }
  return r;
}

double lmg_step(double eps, int l, double guess) {
  const int m = 2 * l;
  const double pi = 3.14159265358979323846;
  double h = guess, prev = 1.0e50, step = 0.1 * guess, f = 1.0e50;
  double sign = f > eps ? -1.0 : 1.0;
  while (std::fabs(prev - h) > 1e-14) {
    const double c0 = 4.0 * std::sqrt(2.0) * pi;
    const double cm = std::tgamma(3.0 / 2.0) / std::tgamma((m + 3.0) / 2.0);
    const double p0 = 1.0 / h;
    const double e0 = std::exp(-pi * pi / (2.0 * h));
    const double pm = std::pow(pi / h, m / 2.0);
    const double rd0 = c0 * p0 * e0;
    f = cm * pm * rd0;
    const double last = sign;
    sign = f > eps ? -1.0 : 1.0;
    if (h < 0.0) sign = 1.0;
    if (sign != last) step *= 0.1;
    prev = h;
    h += sign * step;
  }
  return h;
}
} // namespace

occxc_status occxc_lmg_radial_grid(double precision, double alpha_max, int l_max,
                                   const double *alpha_min, double atomic_radius, int capacity,
                                   double *r, double *w, int *npoints_out) {
  return guarded([&] {
    OCCXC_REQUIRE(precision > 0.0 && precision < 1.0 && alpha_max > 0.0 && l_max >= 0 &&
                      alpha_min && atomic_radius > 0.0 && capacity >= 0 && npoints_out,
                  OCCXC_ERR_INVALID_ARGUMENT, "lmg_radial_grid: bad arguments");
    const double r_in = lmg_inner_radius(precision, 2.0 * alpha_max);
    double h = (double)std::numeric_limits<float>::max();
    double r_out = 0.0;
    for (int l = 0; l <= l_max; ++l) {
      if (alpha_min[l] > 0.0) {
        r_out = std::max(r_out, lmg_outer_radius(precision, alpha_min[l], l, 4.0 * atomic_radius));
        h = std::min(h, lmg_step(precision, l, 0.1 * (r_out - r_in)));
      }
    }
    OCCXC_REQUIRE(r_out > r_in && r_out > h && h > 0.0, OCCXC_ERR_INVALID_ARGUMENT,
                  "lmg_radial_grid: no valid grid for these exponents");
    const double c = r_in / (std::exp(h) - 1.0);
    const int n = static_cast<int>(std::log(1.0 + (r_out / c)) / h);
    *npoints_out = n;
    if (capacity < n) return;  // size query
    OCCXC_REQUIRE(r && w, OCCXC_ERR_INVALID_ARGUMENT, "lmg_radial_grid: null output");
    for (int i = 0; i < n; ++i) {
      r[i] = c * (std::exp((i + 1) * h) - 1.0);
      w[i] = (r[i] + c) * h * r[i] * r[i];
    }
  });
}

occxc_status occxc_becke_partition(int natoms, const double *atom_xyz, const double *radii,
                                   int ta_adjustment, long long npts, const double *points,
                                   const int *atom_of_point, double *weights, int device) {
  return guarded([&] {
    OCCXC_REQUIRE(natoms >= 1 && atom_xyz && radii && npts >= 0 &&
                      (npts == 0 || (points && atom_of_point && weights)),
                  OCCXC_ERR_INVALID_ARGUMENT, "becke_partition: bad arguments");
    for (long long i = 0; i < npts; ++i)
      OCCXC_REQUIRE(atom_of_point[i] >= 0 && atom_of_point[i] < natoms, OCCXC_ERR_INVALID_ARGUMENT,
                    "becke_partition: atom index out of range");
    if (npts == 0) return;
    std::vector<double> dist((size_t)natoms * natoms, 0.0);
    for (int i = 0; i < natoms; ++i)
      for (int j = i + 1; j < natoms; ++j) {
        const double dx = atom_xyz[3 * i] - atom_xyz[3 * j], dy = atom_xyz[3 * i + 1] - atom_xyz[3 * j + 1],
                     dz = atom_xyz[3 * i + 2] - atom_xyz[3 * j + 2];
        dist[i * natoms + j] = dist[j * natoms + i] = std::sqrt(dx * dx + dy * dy + dz * dz);
        OCCXC_REQUIRE(dist[i * natoms + j] > 0.0, OCCXC_ERR_INVALID_ARGUMENT,
                      "becke_partition: coincident atoms");
      }
    DeviceGuard g(device);
    DevBuf<double> da(3 * natoms), dr(natoms), dd(dist.size()), dp(3 * npts), dw(npts);
    DevBuf<int> di(npts);
    da.upload(atom_xyz, 3 * natoms);
    dr.upload(radii, natoms);
    dd.upload(dist.data(), dist.size());
    dp.upload(points, 3 * npts);
    dw.upload(weights, npts);
    di.upload(atom_of_point, npts);
    launch_becke(natoms, da.get(), dr.get(), dd.get(), ta_adjustment, npts, dp.get(), di.get(),
                 dw.get(), nullptr);
    dw.download(weights, npts);
    OCCXC_HIP(hipDeviceSynchronize());
  });
}

occxc_status occxc_becke_partition_noalloc(int natoms, const double *atom_xyz,
                                           const double *radii, int ta_adjustment,
                                           long long npts, const double *points,
                                           const int *atom_of_point, double *weights,
                                           void *stream) {
  return guarded([&] {
    OCCXC_REQUIRE(atom_xyz && radii && npts >= 0 &&
                      (npts == 0 || (points && atom_of_point && weights)),
                  OCCXC_ERR_INVALID_ARGUMENT, "becke_partition_noalloc: bad arguments");
    becke_partition_noalloc(natoms, atom_xyz, radii, ta_adjustment, npts, points, atom_of_point,
                            weights, static_cast<hipStream_t>(stream));
  });
}

const char *occxc_last_error_message(void) { return g_last_error.c_str(); }

const char *occxc_version(void) { return "occxc 0.1 gfx950"; }

} // extern "C"
