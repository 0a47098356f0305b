// nlc.hip -- VV10 non-local correlation on the GPU:
// NonLocalCorrelationFunctional::vv10 (src/dft/nonlocal_correlation.cpp:
// 246-306) with vv10_kernel (nonlocal_correlation.cpp:37-103).
//
// The reference sums, for each grid point i with rho_i >= 1e-8, over the
// points j < i in its input order (each pair counted twice, the self term
// once):
//   F_i = -1.5 [rho_i w_i / (2 k_i^3) + sum_j 2 rho_j w_j / (g_i g_j (g_i + g_j))]
//   U_i = [self] + sum_j t_ij (1/g_i + 1/(g_i + g_j)),  W_i = sum_j t_ij (...) r_ij^2
// with g = r^2 omega0 + kappa.  The energy is the symmetric double sum; the
// per-point potential depends on the point order, so the pair kernel works
// in input order (the context records each padded point's input index).
//
// Layout: the context's points are in leaf order; k_vv10_prep scatters each
// real point's (x, y, z, rho w, omega0, kappa, rho) into input-order planes
// of npts; k_vv10_pairs gives F, U, W per input point (one thread per i,
// 256-point j tiles in LDS, blocks with the most j tiles first);
// k_vv10_coef gathers them back per leaf and forms the weighted
// exc / vrho / vsigma and the GGA Vxc coefficients, which the restricted Vxc
// kernel contracts (Vxc = ktmp + ktmp^T, nonlocal_correlation.cpp:295-303).
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace occxc {
namespace {

constexpr double kVvPi = 3.14159265358979323846;
constexpr double kVvRhoMin = 1e-8;  // nonlocal_correlation.cpp:68,77
constexpr int kVvTile = 256;

enum : int { kX = 0, kY, kZ, kRW, kOM, kKA, kRHO, kF, kU, kW, kNlcPlanes };

void check(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ inline double block_sum(double v, double *s_red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  const double r = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  __syncthreads();
  return r;
}

struct Vv10Consts {
  double C, kappa_factor, beta;
};

Vv10Consts vv10_consts(double b, double C) {
  Vv10Consts k;
  k.C = C;
  k.kappa_factor = b * 1.5 * kVvPi * std::pow(9.0 * kVvPi, -1.0 / 6.0);
  k.beta = std::pow(3.0 / (b * b), 0.75) / 32.0;
  return k;
}

// omega0 and kappa of one point (nonlocal_correlation.cpp:53-61)
__device__ inline void vv10_point(double rho, double sigma, const Vv10Consts &k, double &tmp,
                                  double &omega0, double &kappa) {
  const double pi4_3 = 4.0 * kVvPi / 3.0;
  tmp = sigma / (rho * rho);
  tmp = k.C * tmp * tmp;
  omega0 = sqrt(tmp + pi4_3 * rho);
  kappa = k.kappa_factor * pow(rho, 1.0 / 6.0);
}

// leaf order -> input order (one block per leaf, one thread per point)
__global__ __launch_bounds__(256) void k_vv10_prep(BuildArgs a, const double *__restrict__ rho4,
                                                   const int *__restrict__ porig, Vv10Consts k,
                                                   double *__restrict__ nlc, long long npts) {
  const LeafDev L = a.leaves[blockIdx.x];
  const int p = threadIdx.x;
  if (p >= L.n) return;
  const long long q = L.pt_off + p, P = a.ptot;
  const long long o = porig[q];
  if (o < 0 || o >= npts) return;
  const double rho = rho4[q];
  const double gx = rho4[P + q], gy = rho4[2 * P + q], gz = rho4[3 * P + q];
  double tmp, om, ka;
  vv10_point(rho, gx * gx + gy * gy + gz * gz, k, tmp, om, ka);
  nlc[kX * npts + o] = a.px[q];
  nlc[kY * npts + o] = a.py[q];
  nlc[kZ * npts + o] = a.pz[q];
  nlc[kRW * npts + o] = rho * a.pw[q];
  nlc[kOM * npts + o] = om;
  nlc[kKA * npts + o] = ka;
  nlc[kRHO * npts + o] = rho;
}

// 1/x for the pair denominators: the hardware reciprocal refined by two
// Newton-Raphson steps, r <- r + r (1 - x r) (within about an ulp; x lies in
// [1e-3, 1e30] here: g >= kappa(1e-8) ~ 0.7, so no scaling is needed)
__device__ inline double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// F, U, W of every input point i from the points j < i.  Block b (heaviest
// first) owns points [256 b, 256 b + 256) and walks the j tiles 0..b; dead j
// (rho_j < 1e-8, or beyond npts) enter with weight 0, which adds exactly 0.
// One reciprocal per pair: R = 1 / ((g_i s)^2 g_j), s = g_i + g_j, gives
// t = 2 rho_j w_j / (g_i g_j s) = q (g_i s) R and t (1/g_i + 1/s) = q (s + g_i) R.
// The loop is FP64 VALU-bound (about 24 instructions per pair).
__global__ __launch_bounds__(256) void k_vv10_pairs(double *__restrict__ nlc, long long npts) {
  __shared__ double s_x[kVvTile], s_y[kVvTile], s_z[kVvTile], s_q[kVvTile], s_om[kVvTile],
      s_ka[kVvTile];
  const long long blk = (long long)gridDim.x - 1 - blockIdx.x;
  const long long i = blk * kVvTile + threadIdx.x;
  const bool in = i < npts;
  const bool live = in && nlc[kRHO * npts + i] >= kVvRhoMin;
  double xi = 0, yi = 0, zi = 0, omi = 0, kai = 1, f = 0, u = 0, w = 0;
  if (live) {
    xi = nlc[kX * npts + i];
    yi = nlc[kY * npts + i];
    zi = nlc[kZ * npts + i];
    omi = nlc[kOM * npts + i];
    kai = nlc[kKA * npts + i];
    f = nlc[kRW * npts + i] / (2 * kai * kai * kai);
    u = f * (1.0 / kai + 1.0 / (2 * kai));
  }
  for (long long j0 = 0; j0 <= blk * kVvTile; j0 += kVvTile) {
    __syncthreads();
    const long long j = j0 + threadIdx.x;
    const bool lj = j < npts && nlc[kRHO * npts + j] >= kVvRhoMin;
    s_x[threadIdx.x] = lj ? nlc[kX * npts + j] : 0.0;
    s_y[threadIdx.x] = lj ? nlc[kY * npts + j] : 0.0;
    s_z[threadIdx.x] = lj ? nlc[kZ * npts + j] : 0.0;
    s_q[threadIdx.x] = lj ? 2.0 * nlc[kRW * npts + j] : 0.0;
    s_om[threadIdx.x] = lj ? nlc[kOM * npts + j] : 0.0;
    s_ka[threadIdx.x] = lj ? nlc[kKA * npts + j] : 1.0;
    __syncthreads();
    if (!live) continue;
    const long long rem = i - j0;  // >= 0: j0 <= blk * kVvTile <= i
    const int jn = rem < kVvTile ? (int)rem : kVvTile;
    for (int jj = 0; jj < jn; ++jj) {
      const double dx = xi - s_x[jj], dy = yi - s_y[jj], dz = zi - s_z[jj];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double gi = r2 * omi + kai;
      const double gj = r2 * s_om[jj] + s_ka[jj];
      const double s = gi + gj;
      const double gis = gi * s;
      const double R = recip(gis * gis * gj);
      const double qr = s_q[jj] * R;
      f += qr * gis;
      const double tu = qr * (s + gi);
      u += tu;
      w += tu * r2;
    }
  }
  if (in) {
    nlc[kF * npts + i] = live ? f * -1.5 : 0.0;
    nlc[kU * npts + i] = live ? u : 0.0;
    nlc[kW * npts + i] = live ? w : 0.0;
  }
}

// input order -> leaf order: weighted exc, vrho, vsigma (nonlocal_correlation.
// cpp:94-101), the energy and N_elec partials of the leaf, and the Vxc
// coefficients c0 = vrho / 2, c_k = 2 vsigma d_k rho (nonlocal_correlation.cpp:
// 295-303).  A point with rho < 1e-8 keeps exc = vrho = beta w, vsigma = 0.
__global__ __launch_bounds__(256) void k_vv10_coef(BuildArgs a, const double *__restrict__ rho4,
                                                   const int *__restrict__ porig, Vv10Consts k,
                                                   const double *__restrict__ nlc, long long npts) {
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const long long q = L.pt_off + p, P = a.ptot;
  double e_part = 0.0, n_part = 0.0, c[4] = {0, 0, 0, 0};
  if (real) {
    const long long o = porig[q];
    const double rho = rho4[q];
    const double g[3] = {rho4[P + q], rho4[2 * P + q], rho4[3 * P + q]};
    const double sigma = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    const double w = a.pw[q];
    double vr = k.beta * w, vs = 0.0, ex = k.beta * w;
    if (rho >= kVvRhoMin && o >= 0 && o < npts) {
      const double pi4_3 = 4.0 * kVvPi / 3.0;
      double tmp, om, ka;
      vv10_point(rho, sigma, k, tmp, om, ka);
      const double dom_dr = (0.5 * pi4_3 * rho - 2.0 * tmp) / om;
      const double dom_dg = sigma > 0.0 ? tmp * rho / (sigma * om) : 0.0;
      const double dka_dr = (1.0 / 6.0) * ka;
      const double F = nlc[kF * npts + o], U = nlc[kU * npts + o], W = nlc[kW * npts + o];
      ex = (k.beta + 0.5 * F) * w;
      vr = (k.beta + F + 1.5 * (U * dka_dr + W * dom_dr)) * w;
      vs = (1.5 * W * dom_dg) * w;
    }
    e_part = ex * rho;
    n_part = rho * w;
    c[0] = 0.5 * vr;
    for (int d = 0; d < 3; ++d) c[d + 1] = 2.0 * vs * g[d];
  }
  const double e = block_sum(e_part, s_red);
  const double n = block_sum(n_part, s_red);
  if (live)
    for (int d = 0; d < 4; ++d) a.coef[d * P + q] = c[d];
  if (threadIdx.x == 0) {
    a.leaf_part[2 * leaf] = e;
    a.leaf_part[2 * leaf + 1] = n;
    a.leaf_active[leaf] = 1;
  }
}

} // namespace

size_t nlc_buffer_elems(long long npts) { return (size_t)kNlcPlanes * (size_t)npts; }

void launch_vv10(const BuildArgs &a, const double *rho4, const int *porig, double b, double C,
                 double *nlc, long long npts, hipStream_t s) {
  if (a.nleaves == 0 || npts == 0) return;
  const Vv10Consts k = vv10_consts(b, C);
  launch_zero(nlc, (long long)kNlcPlanes * npts, s);
  k_vv10_prep<<<a.nleaves, 256, 0, s>>>(a, rho4, porig, k, nlc, npts);
  check("k_vv10_prep");
  const long long nblk = (npts + kVvTile - 1) / kVvTile;
  k_vv10_pairs<<<(unsigned)nblk, kVvTile, 0, s>>>(nlc, npts);
  check("k_vv10_pairs");
  k_vv10_coef<<<a.nleaves, 256, 0, s>>>(a, rho4, porig, k, nlc, npts);
  check("k_vv10_coef");
}

} // namespace occxc
