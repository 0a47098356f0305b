// kernels.hip -- gfx950 kernels of the XC build (reference kernels).
//
// Stages, with the OCC function each replaces:
//   k_ao_eval            gto::evaluate_basis_subset (src/gto/gto.cpp:212-287)
//                        + gau2grid gg_collocation[_deriv1]
//   k_density_functional density::evaluate_density<d,R> (src/gto/density.cpp:8-43),
//                        set_params (dft_kernels.h:16-58),
//                        DensityFunctional::evaluate (functional.cpp:18-65),
//                        Result::weight_by (functional.h:460-470), the N_elec sum and
//                        the leaf skip of process_grid_block (dft_kernels.cpp:29-40)
//   k_vxc                xc_potential_matrix<R,d> (xc_potential_matrix.cpp:9-41) and
//                        the K scatter of compute_K_dft (dft.h:361-369)
//   k_finalize_*         K + K^T and the fixed-order reduction of Exc, N_elec
//                        (dft.h:375-380)
//   k_decay_cutoff       gto::evaluate_decay_cutoff (gto.cpp:18-67)
//   k_becke              calculate_atomic_grid_weights + normalisation
//                        (grid_utils.cpp:125-200, molecular_grid.cpp:70-78)
//
// These kernels keep every product in FP64 registers; the MFMA contraction
// kernels (contract_mfma.hip) replace k_density_functional / k_vxc for speed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "kernels.hpp"
#include "xc_functionals.hpp"
#include "xc_functionals_pol.hpp"
#include "xc_functionals_mgga.hpp"

namespace occxc {

// CCA Cartesian -> spherical coefficients for l = 0..kMaxL, rows m=-l..l.
__constant__ double c_sph[512];
__constant__ int c_sph_off[kMaxL + 1];

namespace {

constexpr int kNCartMax = (kMaxL + 1) * (kMaxL + 2) / 2;

// Value (and gradient) of every component of one shell at one point,
// displacement (dx,dy,dz) from the shell centre.  Coefficients carry
// common_fac(l).  Cartesian components in CCA order; spherical via c_sph.
// Each result goes to sink(q, m, value) (q = 0 value, 1..3 gradient) with
// compile-time q and m once inlined, so no result array is indexed at run
// time.
template <int L, typename Sink>
__device__ __forceinline__ void shell_point(const ShellDev &sh,
                                            const double *__restrict__ exps,
                                            const double *__restrict__ coefs,
                                            double dx, double dy, double dz,
                                            bool deriv, Sink &&sink) {
  constexpr int NC = (L + 1) * (L + 2) / 2;
  const double r2 = dx * dx + dy * dy + dz * dz;
  double S0 = 0.0, S1 = 0.0;
  for (int k = 0; k < sh.nprim; ++k) {
    const double a = exps[sh.poff + k];
    const double c = coefs[sh.poff + k] * exp(-a * r2);
    S0 += c;
    S1 += -2.0 * a * c;
  }
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.0;
#pragma unroll
  for (int i = 1; i <= L; ++i) {
    xp[i] = xp[i - 1] * dx;
    yp[i] = yp[i - 1] * dy;
    zp[i] = zp[i - 1] * dz;
  }
  double cv[NC], cx[NC], cy[NC], cz[NC];
  int idx = 0;
#pragma unroll
  for (int a = L; a >= 0; --a) {
#pragma unroll
    for (int b = L - a; b >= 0; --b) {
      const int c = L - a - b;
      const double poly = xp[a] * yp[b] * zp[c];
      cv[idx] = poly * S0;
      if (deriv) {
        const double dpx = a ? a * xp[a > 0 ? a - 1 : 0] * yp[b] * zp[c] : 0.0;
        const double dpy = b ? b * xp[a] * yp[b > 0 ? b - 1 : 0] * zp[c] : 0.0;
        const double dpz = c ? c * xp[a] * yp[b] * zp[c > 0 ? c - 1 : 0] : 0.0;
        cx[idx] = dpx * S0 + poly * dx * S1;
        cy[idx] = dpy * S0 + poly * dy * S1;
        cz[idx] = dpz * S0 + poly * dz * S1;
      }
      ++idx;
    }
  }
  if (sh.pure && L >= 2) {
    const double *T = c_sph + c_sph_off[L];
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      double v = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const double t = T[m * NC + k];
        v += t * cv[k];
        if (deriv) {
          vx += t * cx[k];
          vy += t * cy[k];
          vz += t * cz[k];
        }
      }
      sink(0, m, v);
      if (deriv) {
        sink(1, m, vx);
        sink(2, m, vy);
        sink(3, m, vz);
      }
    }
  } else if (sh.pure && L == 1) {
    // CCA p order is (y, z, x) (gau2grid_transform.c:44-61)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int k = m == 2 ? 0 : m + 1;
      sink(0, m, cv[k]);
      if (deriv) {
        sink(1, m, cx[k]);
        sink(2, m, cy[k]);
        sink(3, m, cz[k]);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      sink(0, k, cv[k]);
      if (deriv) {
        sink(1, k, cx[k]);
        sink(2, k, cy[k]);
        sink(3, k, cz[k]);
      }
    }
  }
}

// x^a y^b z^c from the power tables, 0 when an exponent is negative (the
// derivative of a monomial that does not contain that coordinate).
template <int L>
__device__ __forceinline__ double mono(const double (&xp)[L + 1], const double (&yp)[L + 1],
                                       const double (&zp)[L + 1], int a, int b, int c) {
  return (a < 0 || b < 0 || c < 0) ? 0.0 : xp[a < 0 ? 0 : a] * yp[b < 0 ? 0 : b] * zp[c < 0 ? 0 : c];
}

// Gradient and Hessian of every component of one shell at one point
// (gau2grid gg_collocation_deriv2, as evaluate_basis_subset calls it with
// max_derivative 2, gto.cpp:245-285).  For a Cartesian monomial P times the
// contraction, with S0 = sum c e^{-a r2}, S1 = sum -2a c e^{-a r2} and
// S2 = sum 4a^2 c e^{-a r2}:
//   d_i f  = (d_i P) S0 + P x_i S1
//   d_ij f = (d_ij P) S0 + ((d_i P) x_j + (d_j P) x_i + delta_ij P) S1 + P x_i x_j S2.
// Only the components q in [Q0, Q1) are formed (q = 1..3 x,y,z; 4..9 xx, xy,
// xz, yy, yz, zz) and each goes to sink(q, m, value).
template <int L, int Q0, int Q1, typename Sink>
__device__ __forceinline__ void shell_point_d2(const ShellDev &sh,
                                               const double *__restrict__ exps,
                                               const double *__restrict__ coefs,
                                               double dx, double dy, double dz, Sink &&sink) {
  constexpr int NC = (L + 1) * (L + 2) / 2;
  constexpr int NQ = Q1 - Q0;
  const double r2 = dx * dx + dy * dy + dz * dz;
  double S0 = 0.0, S1 = 0.0, S2 = 0.0;
  for (int k = 0; k < sh.nprim; ++k) {
    const double a = exps[sh.poff + k];
    const double c = coefs[sh.poff + k] * exp(-a * r2);
    S0 += c;
    S1 += -2.0 * a * c;
    S2 += 4.0 * a * a * c;
  }
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.0;
#pragma unroll
  for (int i = 1; i <= L; ++i) {
    xp[i] = xp[i - 1] * dx;
    yp[i] = yp[i - 1] * dy;
    zp[i] = zp[i - 1] * dz;
  }
  const double d[3] = {dx, dy, dz};
  double cq[NQ][NC];
  int idx = 0;
#pragma unroll
  for (int a = L; a >= 0; --a) {
#pragma unroll
    for (int b = L - a; b >= 0; --b) {
      const int c = L - a - b;
      const double P = mono<L>(xp, yp, zp, a, b, c);
      const double dP[3] = {a * mono<L>(xp, yp, zp, a - 1, b, c),
                            b * mono<L>(xp, yp, zp, a, b - 1, c),
                            c * mono<L>(xp, yp, zp, a, b, c - 1)};
#pragma unroll
      for (int q = Q0; q < Q1; ++q) {
        double v;
        if (q <= 3) {
          const int i = q - 1;
          v = dP[i] * S0 + P * d[i] * S1;
        } else {
          constexpr int hi[6] = {0, 0, 0, 1, 1, 2}, hj[6] = {0, 1, 2, 1, 2, 2};
          const int i = hi[q - 4], j = hj[q - 4];
          const int ea[3] = {a, b, c};
          // d_ij P: lower exponent i, then j
          const int ci = ea[i];
          const int cj = ea[j] - (i == j ? 1 : 0);
          const int e0 = a - (i == 0) - (j == 0), e1 = b - (i == 1) - (j == 1),
                    e2 = c - (i == 2) - (j == 2);
          const double ddP = (double)(ci * cj) * mono<L>(xp, yp, zp, e0, e1, e2);
          v = ddP * S0 + (dP[i] * d[j] + dP[j] * d[i] + (i == j ? P : 0.0)) * S1 +
              P * d[i] * d[j] * S2;
        }
        cq[q - Q0][idx] = v;
      }
      ++idx;
    }
  }
  if (sh.pure && L >= 2) {
    const double *T = c_sph + c_sph_off[L];
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) v += T[m * NC + k] * cq[q][k];
        sink(Q0 + q, m, v);
      }
  } else if (sh.pure && L == 1) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < NQ; ++q) sink(Q0 + q, m, cq[q][m == 2 ? 0 : m + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int q = 0; q < NQ; ++q) sink(Q0 + q, k, cq[q][k]);
  }
}

__device__ __forceinline__ void eval_shell(const ShellDev &sh,
                                           const double *__restrict__ exps,
                                           const double *__restrict__ coefs,
                                           double dx, double dy, double dz,
                                           bool deriv,
                                           double (&out)[4][kNCartMax]) {
  auto sink = [&](int q, int m, double v) { out[q][m] = v; };
  switch (sh.l) {
  case 0: shell_point<0>(sh, exps, coefs, dx, dy, dz, deriv, sink); break;
  case 1: shell_point<1>(sh, exps, coefs, dx, dy, dz, deriv, sink); break;
  case 2: shell_point<2>(sh, exps, coefs, dx, dy, dz, deriv, sink); break;
  case 3: shell_point<3>(sh, exps, coefs, dx, dy, dz, deriv, sink); break;
  default: shell_point<4>(sh, exps, coefs, dx, dy, dz, deriv, sink); break;
  }
}

template <typename T> __device__ __forceinline__ double ld(const T *p) {
  return static_cast<double>(*p);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Block-wide reduction for blockDim = 256 (4 wavefronts); result valid in
// every thread.  Fixed order -> deterministic.
__device__ double block_sum_256(double v, double *scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
  __syncthreads();
  return r;
}
__device__ double block_max_256(double v, double *scratch) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double r = fmax(fmax(scratch[0], scratch[1]), fmax(scratch[2], scratch[3]));
  __syncthreads();
  return r;
}

// FP64 -> FP32 with unbiased rounding: v goes to one of its two neighbouring
// floats, chosen with the probabilities that make the expected stored value
// v itself (stochastic rounding).  On an atom-centred grid an s-type AO has
// the same value at all angular points of a radial shell, so rounding to
// nearest repeats one error there and those errors add up coherently in Exc;
// unbiased rounding makes them independent, and they average out.  The
// random bits are a hash of the point's coordinates, the global AO index and
// the component, so the stored value depends only on what it represents:
// evaluations are reproducible and a molecule's AO values are the same in a
// single-molecule context and in a batch.
__device__ __forceinline__ float round_unbiased(double v, unsigned long long key) {
  const float f = (float)v;
  const double r = v - (double)f;  // exact, |r| <= half an ulp of f
  if (r == 0.0) return f;
  // g: the float next to f on v's side (integer step of the bit pattern)
  const unsigned int b = __float_as_uint(f);
  const bool away = (r > 0.0) == (f > 0.0f);
  const unsigned int gb = f == 0.0f ? (r > 0.0 ? 1u : 0x80000001u) : (away ? b + 1u : b - 1u);
  const float g = __uint_as_float(gb);
  const float ulp = fabsf(g - f);  // exact
  unsigned int h = (unsigned int)(key ^ (key >> 32)) * 0x9E3779B9u;  // murmur3 finaliser
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  const float u = (float)(h >> 8) * 0x1.0p-24f;  // uniform on [0, 1)
  return u * ulp < (float)fabs(r) ? g : f;
}

template <typename T> __device__ __forceinline__ T store_ao(double v, unsigned long long key) {
  if constexpr (std::is_same<T, float>::value) return round_unbiased(v, key);
  else return static_cast<T>(v);
}

// ------------------------------------------------------------------ AO --
// One workgroup per active leaf, one thread per (padded) point; for every
// significant shell each thread writes its components.  The AO block of a
// leaf is [ncomp][mp][np] with points fastest, so each component row is a
// coalesced run of np values.
// LMAX is the largest angular momentum the kernel handles: the register
// allocation is that of the largest shell, so bases with l <= 2 run the
// LMAX = 2 instance.
// RES (FP32 storage, context set-up): instead of the stored values, the
// value plane's rounding residuals v - s of the leaves res_leaf0 + blockIdx.x
// into a.ao_res.
template <typename T, int LMAX, bool RES = false>
__global__ __launch_bounds__(256) void k_ao_eval(BuildArgs a, T *__restrict__ ao) {
  const LeafDev L = a.leaves[RES ? a.res_leaf0 + blockIdx.x : blockIdx.x];
  // lanes over (point, shell split): a leaf of np points takes 256 / np
  // splits of its shells (np = 64: 4, 128: 2, 192 and 256: 1), split s taking
  // shells s, s + nsplit, ...; every wavefront lies in one split, so the
  // split test is uniform and leaves of fewer than 256 points keep their lanes
  const int nsplit = L.np <= 64 ? 4 : (L.np <= 128 ? 2 : 1);
  const int split = threadIdx.x / L.np;
  if (split >= nsplit) return;
  const int p = threadIdx.x - split * L.np;
  const double x = a.px[L.pt_off + p], y = a.py[L.pt_off + p], z = a.pz[L.pt_off + p];
  const bool deriv = a.deriv != 0;
  T *base = ao + L.ao_off;
  const long long cstride = (long long)L.mp * L.np;
  // the point's part of the unbiased-rounding key (FP32 storage)
  const unsigned long long pkey = (unsigned long long)__double_as_longlong(x) ^
                                  ((unsigned long long)__double_as_longlong(y) << 21 |
                                   (unsigned long long)__double_as_longlong(y) >> 43) ^
                                  ((unsigned long long)__double_as_longlong(z) << 42 |
                                   (unsigned long long)__double_as_longlong(z) >> 22);
  int mu = 0;
  for (int k = 0; k < L.nsig; ++k) {
    const ShellDev sh = a.shells[a.leaf_shells[L.sh_off + k]];
    if (k % nsplit != split) {
      mu += sh.size;
      continue;
    }
    T *row = base + (long long)mu * L.np + p;
    float *res = RES ? a.ao_res + (L.res_off - a.res_base) + (long long)mu * L.np + p : nullptr;
    const unsigned long long key0 = pkey + 0x2545F4914F6CDD1Dull * (unsigned long long)sh.bf0;
    auto sink = [&](int q, int m, double v) {
      const long long off = q * cstride + (long long)m * L.np;
      const T s = store_ao<T>(v, key0 + 0x2545F4914F6CDD1Dull * (unsigned long long)m +
                                     0x9E3779B97F4A7C15ull * (unsigned long long)q);
      if constexpr (RES) {
        if (q == 0) res[(long long)m * L.np] = (float)(v - (double)s);
      } else {
        row[off] = s;
      }
    };
    const double dx = x - sh.x, dy = y - sh.y, dz = z - sh.z;
    if (sh.l == 0) shell_point<0>(sh, a.exps, a.coefs, dx, dy, dz, deriv, sink);
    else if (sh.l == 1) shell_point<1>(sh, a.exps, a.coefs, dx, dy, dz, deriv, sink);
    else if (LMAX == 2 || sh.l == 2)
      shell_point<2>(sh, a.exps, a.coefs, dx, dy, dz, deriv, sink);
    else if (sh.l == 3) shell_point<(LMAX > 2 ? 3 : 2)>(sh, a.exps, a.coefs, dx, dy, dz, deriv, sink);
    else shell_point<LMAX>(sh, a.exps, a.coefs, dx, dy, dz, deriv, sink);
    mu += sh.size;
  }
  for (int r = L.m + split; r < L.mp; r += nsplit) {
    if constexpr (RES) {
      a.ao_res[(L.res_off - a.res_base) + (long long)r * L.np + p] = 0.0f;
    } else {
      for (int c = 0; c < a.ncomp; ++c)
        base[c * cstride + (long long)r * L.np + p] = static_cast<T>(0);
    }
  }
}

// Derivative components the resident AO cache does not hold, for the XC
// gradient: q in [Q0, Q1) of shell_point_d2 (the gradient for an LDA context,
// the Hessian for a GGA one), FP64, each leaf's block [Q1-Q0][mp][np] at
// element offset ao_off / ncomp * (Q1 - Q0).
template <int LMAX, int Q0, int Q1>
__global__ __launch_bounds__(256) void k_ao_extra(BuildArgs a, double *__restrict__ out) {
  constexpr int NQ = Q1 - Q0;
  const LeafDev L = a.leaves[blockIdx.x];
  const int p = threadIdx.x;
  if (p >= L.np) return;
  const double x = a.px[L.pt_off + p], y = a.py[L.pt_off + p], z = a.pz[L.pt_off + p];
  double *base = out + L.ao_off / a.ncomp * NQ;
  const long long cstride = (long long)L.mp * L.np;
  int mu = 0;
  for (int k = 0; k < L.nsig; ++k) {
    const ShellDev sh = a.shells[a.leaf_shells[L.sh_off + k]];
    double *row = base + (long long)mu * L.np + p;
    auto sink = [&](int q, int m, double v) { row[(q - Q0) * cstride + (long long)m * L.np] = v; };
    const double dx = x - sh.x, dy = y - sh.y, dz = z - sh.z;
    if (sh.l == 0) shell_point_d2<0, Q0, Q1>(sh, a.exps, a.coefs, dx, dy, dz, sink);
    else if (sh.l == 1) shell_point_d2<1, Q0, Q1>(sh, a.exps, a.coefs, dx, dy, dz, sink);
    else if (LMAX == 2 || sh.l == 2) shell_point_d2<2, Q0, Q1>(sh, a.exps, a.coefs, dx, dy, dz, sink);
    else if (sh.l == 3)
      shell_point_d2<(LMAX > 2 ? 3 : 2), Q0, Q1>(sh, a.exps, a.coefs, dx, dy, dz, sink);
    else shell_point_d2<LMAX, Q0, Q1>(sh, a.exps, a.coefs, dx, dy, dz, sink);
    mu += sh.size;
  }
  for (int r = L.m; r < L.mp; ++r)
    for (int c = 0; c < NQ; ++c) base[c * cstride + (long long)r * L.np + p] = 0.0;
}

// ------------------------------------------------- density + functional --
template <typename T>
__global__ __launch_bounds__(256) void k_density_functional(BuildArgs a,
                                                            const T *__restrict__ ao) {
  __shared__ int s_map[kMaxLocalAo];
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const MolDev M = a.mols[L.mol];
  for (int i = threadIdx.x; i < L.m; i += blockDim.x) s_map[i] = a.ao_map[L.map_off + i];
  __syncthreads();
  const int p = threadIdx.x;
  const bool live = p < L.np;
  const bool real = p < L.n;
  const T *phi = ao + L.ao_off;
  const long long cstride = (long long)L.mp * L.np;
  double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  if (live) {
    for (int nu = 0; nu < L.m; ++nu) {
      const double *Dcol = a.D2 + M.moff + (long long)s_map[nu] * M.nbf;
      double t = 0.0;
      for (int mu = 0; mu < L.m; ++mu) t += Dcol[s_map[mu]] * ld(phi + (long long)mu * L.np + p);
      rho += ld(phi + (long long)nu * L.np + p) * t;
      if (a.deriv) {
        gx += 2.0 * ld(phi + cstride + (long long)nu * L.np + p) * t;
        gy += 2.0 * ld(phi + 2 * cstride + (long long)nu * L.np + p) * t;
        gz += 2.0 * ld(phi + 3 * cstride + (long long)nu * L.np + p) * t;
      }
    }
  }
  const double w = live ? a.pw[L.pt_off + p] : 0.0;
  const double nel = block_sum_256(real ? rho * w : 0.0, s_red);
  const double rmax = block_max_256(real ? rho : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e_sum = 0.0, vr_sum = 0.0, vs_sum = 0.0;
  const double sigma = gx * gx + gy * gy + gz * gz;
  if (active && live && rho >= kDensThreshold) {
    const double sg = fmax(sigma, kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f) {
      double e, vr, vs;
      xc_point(a.func_ids[f], rho, sg, e, vr, vs);
      e_sum += e;
      vr_sum += vr;
      vs_sum += vs;
    }
  }
  // weight_by and Exc = sum w rho eps = sum w e
  const double exc = block_sum_256(active && real ? w * e_sum : 0.0, s_red);
  if (live) {
    const long long q = L.pt_off + p;
    a.coef[q] = active ? 0.5 * w * vr_sum : 0.0;
    a.coef[a.ptot + q] = active ? 2.0 * w * vs_sum * gx : 0.0;
    a.coef[2 * a.ptot + q] = active ? 2.0 * w * vs_sum * gy : 0.0;
    a.coef[3 * a.ptot + q] = active ? 2.0 * w * vs_sum * gz : 0.0;
  }
  if (threadIdx.x == 0) {
    a.leaf_part[2 * leaf] = active ? exc : 0.0;
    a.leaf_part[2 * leaf + 1] = nel;
    a.leaf_active[leaf] = active ? 1 : 0;
  }
}

// ---------------------------------------------------------------- Vxc --
// K_loc(mu, nu) = sum_p phi(p,mu) A(p,nu),  A = c0 phi + c . grad phi,
// scattered into K with FP64 atomics; Vxc = K + K^T in k_finalize_matrix.
template <typename T>
__global__ __launch_bounds__(256) void k_vxc(BuildArgs a, const T *__restrict__ ao) {
  __shared__ int s_map[kMaxLocalAo];
  __shared__ double s_c[4][kMaxLeafPts];
  const int leaf = blockIdx.x;
  if (!a.leaf_active[leaf]) return;
  const LeafDev L = a.leaves[leaf];
  const MolDev M = a.mols[L.mol];
  for (int i = threadIdx.x; i < L.m; i += blockDim.x) s_map[i] = a.ao_map[L.map_off + i];
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
    for (int c = 0; c < (a.deriv ? 4 : 1); ++c) s_c[c][i] = a.coef[c * a.ptot + L.pt_off + i];
  __syncthreads();
  const T *phi = ao + L.ao_off;
  const long long cstride = (long long)L.mp * L.np;
  const long long mm = (long long)L.m * L.m;
  for (long long idx = threadIdx.x; idx < mm; idx += blockDim.x) {
    const int mu = (int)(idx % L.m), nu = (int)(idx / L.m);
    const T *fm = phi + (long long)mu * L.np;
    const T *fn = phi + (long long)nu * L.np;
    double s = 0.0;
    if (a.deriv) {
      const T *fx = fn + cstride, *fy = fn + 2 * cstride, *fz = fn + 3 * cstride;
      for (int p = 0; p < L.np; ++p)
        s += ld(fm + p) * (s_c[0][p] * ld(fn + p) + s_c[1][p] * ld(fx + p) +
                           s_c[2][p] * ld(fy + p) + s_c[3][p] * ld(fz + p));
    } else {
      for (int p = 0; p < L.np; ++p) s += ld(fm + p) * s_c[0][p] * ld(fn + p);
    }
    atomicAdd(a.K + M.moff + s_map[mu] + (long long)s_map[nu] * M.nbf, s);
  }
}

__global__ void k_zero(double *__restrict__ p, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

// blockIdx.y = molecule; D2 = 2 D with the upper triangle mirrored
__global__ void k_prepare_density(const double *__restrict__ D, double *__restrict__ D2,
                                  const MolDev *__restrict__ mols) {
  const MolDev M = mols[blockIdx.y];
  const int nbf = M.nbf;
  const long long n = (long long)nbf * nbf;
  const double *Dm = D + M.moff;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(k % nbf), j = (int)(k / nbf);
    const int r = i < j ? i : j, c = i < j ? j : i;
    D2[M.moff + k] = 2.0 * Dm[r + (long long)c * nbf];
  }
}

// blockIdx.y = molecule; Vxc = K + K^T
// V = K + K^T
// Unrestricted D (2 nbf x nbf, alpha rows over beta rows, one molecule):
// D2t_s = 2 D_s^T for each spin, the full block (evaluate_density<d, U> uses
// no selfadjointView, density.cpp:18-30,45-70); the contraction kernels read
// D2t(nu, mu) as D2(mu, nu).
__global__ void k_prepare_density_u(const double *__restrict__ D, double *__restrict__ D2a,
                                    double *__restrict__ D2b, int nbf) {
  const long long n = (long long)nbf * nbf;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const int nu = (int)(k % nbf), mu = (int)(k / nbf);
    D2a[k] = 2.0 * D[mu + (long long)nu * 2 * nbf];
    D2b[k] = 2.0 * D[nbf + mu + (long long)nu * 2 * nbf];
  }
}

// Unrestricted per-leaf functional stage (one thread per point, blockDim =
// 256): N_alpha, N_beta before the skip, the skip on the largest alpha or
// beta density of the leaf (dft_kernels.cpp:27-40), set_params<U>
// (dft_kernels.h:16-46), the polarised functionals, weight_by, the energy
// sum w exc (rho_a + rho_b), and per spin the Vxc coefficients of
// xc_potential_matrix<Unrestricted, d> (xc_potential_matrix.cpp:61-109):
//   c0 = w vrho_s / 2,  c_k = w (2 vsigma_ss d_k rho_s + vsigma_ab d_k rho_s')
// so that V_s = K_s + K_s^T with K_s = phi^T (c0 phi + c . grad phi).
__global__ __launch_bounds__(256) void k_functional_pol(BuildArgs a,
                                                        const double *__restrict__ rho_a,
                                                        const double *__restrict__ rho_b,
                                                        double *__restrict__ coef_b,
                                                        double *__restrict__ part) {
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const long long q = L.pt_off + p, P = a.ptot;
  double ra = 0, rb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0}, w = 0;
  if (live) {
    ra = rho_a[q];
    rb = rho_b[q];
    for (int k = 0; k < 3; ++k) {
      ga[k] = rho_a[(k + 1) * P + q];
      gb[k] = rho_b[(k + 1) * P + q];
    }
    w = a.pw[q];
  }
  const double na = block_sum_256(real ? ra * w : 0.0, s_red);
  const double nb = block_sum_256(real ? rb * w : 0.0, s_red);
  const double rmax = block_max_256(real ? fmax(ra, rb) : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e = 0.0, vr[2] = {0, 0}, vs[3] = {0, 0, 0};
  if (active && live && ra + rb >= kDensThreshold) {
    const double saa = fmax(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2], kSigmaMin);
    const double sab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    const double sbb = fmax(gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2], kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f)
      xc_point_pol(a.func_ids[f], ra, rb, a.deriv ? saa : 0.0, a.deriv ? sab : 0.0,
                   a.deriv ? sbb : 0.0, e, vr, vs);
  }
  const double exc = block_sum_256(active && real ? w * e : 0.0, s_red);
  if (live) {
    double *ca = a.coef;
    ca[q] = active ? 0.5 * w * vr[0] : 0.0;
    coef_b[q] = active ? 0.5 * w * vr[1] : 0.0;
    for (int k = 0; k < 3; ++k) {
      ca[(k + 1) * P + q] = active ? w * (2.0 * vs[0] * ga[k] + vs[1] * gb[k]) : 0.0;
      coef_b[(k + 1) * P + q] = active ? w * (2.0 * vs[2] * gb[k] + vs[1] * ga[k]) : 0.0;
    }
  }
  if (threadIdx.x == 0) {
    part[3 * leaf] = active ? exc : 0.0;
    part[3 * leaf + 1] = na;
    part[3 * leaf + 2] = nb;
    a.leaf_active[leaf] = active ? 1 : 0;
  }
}

// Restricted meta-GGA per-leaf functional stage (one thread per point,
// blockDim = 256), process_grid_block<2, Restricted> (dft_kernels.cpp:8-70)
// after the density store passes: rho and grad rho from rho4 ([4][ptot]),
// tau = (1/2) sum_k tk[k] from the three per-direction passes (tk [3][ptot],
// rowsum((grad_k phi D2) o grad_k phi), density.cpp:94-101); N_elec before
// the skip, the max-rho skip, the functionals, weight_by, the energy sum, and
// the Vxc coefficients of xc_potential_matrix<Restricted, 2>
// (xc_potential_matrix.cpp:19-58): c0 = w vrho / 2, c_k = 2 w vsigma d_k rho
// into a.coef and c_tau = w vtau / 4 into coef_tau, so that
// K_tau = sum_k grad_k phi^T (c_tau grad_k phi) and V = K + K^T carry the
// reference's (1/2) sum_k grad_k phi^T (w vtau o grad_k phi).
__global__ __launch_bounds__(256) void k_functional_mgga(BuildArgs a,
                                                         const double *__restrict__ rho4,
                                                         const double *__restrict__ tk,
                                                         double *__restrict__ coef_tau) {
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const long long q = L.pt_off + p, P = a.ptot;
  double rho = 0, g[3] = {0, 0, 0}, tau = 0, w = 0;
  if (live) {
    rho = rho4[q];
    for (int k = 0; k < 3; ++k) g[k] = rho4[(k + 1) * P + q];
    tau = 0.5 * (tk[q] + tk[P + q] + tk[2 * P + q]);
    w = a.pw[q];
  }
  const double nel = block_sum_256(real ? rho * w : 0.0, s_red);
  const double rmax = block_max_256(real ? rho : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e_sum = 0.0, vr_sum = 0.0, vs_sum = 0.0, vt_sum = 0.0;
  if (active && live && rho >= kDensThreshold) {
    const double sg = fmax(g[0] * g[0] + g[1] * g[1] + g[2] * g[2], kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f) {
      double e, vr, vs, vt;
      xc_point_mgga(a.func_ids[f], rho, sg, tau, e, vr, vs, vt);
      e_sum += e;
      vr_sum += vr;
      vs_sum += vs;
      vt_sum += vt;
    }
  }
  const double exc = block_sum_256(active && real ? w * e_sum : 0.0, s_red);
  if (live) {
    a.coef[q] = active ? 0.5 * w * vr_sum : 0.0;
    for (int k = 0; k < 3; ++k) a.coef[(k + 1) * P + q] = active ? 2.0 * w * vs_sum * g[k] : 0.0;
    coef_tau[q] = active ? 0.25 * w * vt_sum : 0.0;
  }
  if (threadIdx.x == 0) {
    a.leaf_part[2 * leaf] = active ? exc : 0.0;
    a.leaf_part[2 * leaf + 1] = nel;
    a.leaf_active[leaf] = active ? 1 : 0;
  }
}

// Unrestricted meta-GGA per-leaf functional stage (one thread per point,
// blockDim = 256), process_grid_block<2, Unrestricted> (dft_kernels.cpp:8-70):
// rho_s and grad rho_s from the spin densities' store passes ([4][ptot]
// each), tau_s = (1/2) sum_k tk_s[k] (density.cpp:130-166), the polarised
// functionals (tau interleaved per point as dft_kernels.h:51-57 hands it to
// libxc), weight_by, and the coefficients of xc_potential_matrix<
// Unrestricted, 2> (xc_potential_matrix.cpp:61-137) per spin: those of
// k_functional_pol in a.coef / coef_b and c_tau,s = w vtau_s / 4 in
// ctau_a / ctau_b (V_s = K_s + K_s^T).  part[3 leaf] = {Exc, N_a, N_b}.
__global__ __launch_bounds__(256) void k_functional_mgga_pol(
    BuildArgs a, const double *__restrict__ rho_a, const double *__restrict__ rho_b,
    const double *__restrict__ tka, const double *__restrict__ tkb, double *__restrict__ coef_b,
    double *__restrict__ ctau_a, double *__restrict__ ctau_b, double *__restrict__ part) {
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const long long q = L.pt_off + p, P = a.ptot;
  double ra = 0, rb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0}, ta = 0, tb = 0, w = 0;
  if (live) {
    ra = rho_a[q];
    rb = rho_b[q];
    for (int k = 0; k < 3; ++k) {
      ga[k] = rho_a[(k + 1) * P + q];
      gb[k] = rho_b[(k + 1) * P + q];
    }
    ta = 0.5 * (tka[q] + tka[P + q] + tka[2 * P + q]);
    tb = 0.5 * (tkb[q] + tkb[P + q] + tkb[2 * P + q]);
    w = a.pw[q];
  }
  const double na = block_sum_256(real ? ra * w : 0.0, s_red);
  const double nb = block_sum_256(real ? rb * w : 0.0, s_red);
  const double rmax = block_max_256(real ? fmax(ra, rb) : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e = 0.0, vr[2] = {0, 0}, vs[3] = {0, 0, 0}, vt[2] = {0, 0};
  if (active && live && ra + rb >= kDensThreshold) {
    const double saa = fmax(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2], kSigmaMin);
    const double sab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    const double sbb = fmax(gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2], kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f)
      xc_point_mgga_pol(a.func_ids[f], ra, rb, saa, sab, sbb, ta, tb, e, vr, vs, vt);
  }
  const double exc = block_sum_256(active && real ? w * e : 0.0, s_red);
  if (live) {
    double *ca = a.coef;
    ca[q] = active ? 0.5 * w * vr[0] : 0.0;
    coef_b[q] = active ? 0.5 * w * vr[1] : 0.0;
    for (int k = 0; k < 3; ++k) {
      ca[(k + 1) * P + q] = active ? w * (2.0 * vs[0] * ga[k] + vs[1] * gb[k]) : 0.0;
      coef_b[(k + 1) * P + q] = active ? w * (2.0 * vs[2] * gb[k] + vs[1] * ga[k]) : 0.0;
    }
    ctau_a[q] = active ? 0.25 * w * vt[0] : 0.0;
    ctau_b[q] = active ? 0.25 * w * vt[1] : 0.0;
  }
  if (threadIdx.x == 0) {
    part[3 * leaf] = active ? exc : 0.0;
    part[3 * leaf + 1] = na;
    part[3 * leaf + 2] = nb;
    a.leaf_active[leaf] = active ? 1 : 0;
  }
}

// Meta-GGA (and LDA / GGA) functionals over points, ADDING to the outputs.
__global__ void k_functional_mgga_points(int id, long long n, const double *__restrict__ rho,
                                         const double *__restrict__ sigma,
                                         const double *__restrict__ tau,
                                         double *__restrict__ exc, double *__restrict__ vrho,
                                         double *__restrict__ vsigma,
                                         double *__restrict__ vtau) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int fam = xc_family_mgga(id);
  const double r = rho[i];
  if (!(r >= kDensThreshold)) return;
  const double sg = fam >= 1 ? fmax(sigma[i], kSigmaMin) : 0.0;
  const double t = fam == 2 ? tau[i] : 0.0;
  double e, vr, vs, vt;
  xc_point_mgga(id, r, sg, t, e, vr, vs, vt);
  exc[i] += e / r;
  vrho[i] += vr;
  if (fam >= 1) vsigma[i] += vs;
  if (fam == 2) vtau[i] += vt;
}

// {Exc, N_alpha, N_beta}: fixed-order sums of the per-leaf partials
__global__ __launch_bounds__(256) void k_finalize_scalars_u(const double *__restrict__ part,
                                                            int nleaves,
                                                            double *__restrict__ result) {
  __shared__ double s_red[4];
  double v[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < nleaves; i += 256)
    for (int k = 0; k < 3; ++k) v[k] += part[3 * i + k];
  for (int k = 0; k < 3; ++k) {
    const double t = block_sum_256(v[k], s_red);
    if (threadIdx.x == 0) result[k] = t;
  }
}

// Polarised functionals over points (libxc layout), ADDING to the outputs.
__global__ void k_functional_pol_points(int id, long long n, const double *__restrict__ rho,
                                        const double *__restrict__ sigma,
                                        double *__restrict__ exc, double *__restrict__ vrho,
                                        double *__restrict__ vsigma) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool gga = xc_family(id) == 1;
  const double ra = rho[2 * i], rb = rho[2 * i + 1];
  if (!(ra + rb >= kDensThreshold)) return;
  double saa = 0, sab = 0, sbb = 0;
  if (gga) {
    saa = fmax(sigma[3 * i], kSigmaMin);
    sab = sigma[3 * i + 1];
    sbb = fmax(sigma[3 * i + 2], kSigmaMin);
  }
  double e = 0.0, vr[2] = {0, 0}, vs[3] = {0, 0, 0};
  xc_point_pol(id, ra, rb, saa, sab, sbb, e, vr, vs);
  exc[i] += e / (ra + rb);
  vrho[2 * i] += vr[0];
  vrho[2 * i + 1] += vr[1];
  if (gga)
    for (int k = 0; k < 3; ++k) vsigma[3 * i + k] += vs[k];
}

// Polarised meta-GGA (and LDA / GGA) functionals over points, ADDING to the
// outputs in libxc's polarised layout: rho[2n], sigma[3n], tau[2n] in;
// exc[n] (per particle), vrho[2n], vsigma[3n], vtau[2n] out.
__global__ void k_functional_mgga_pol_points(int id, long long n, const double *__restrict__ rho,
                                             const double *__restrict__ sigma,
                                             const double *__restrict__ tau,
                                             double *__restrict__ exc, double *__restrict__ vrho,
                                             double *__restrict__ vsigma,
                                             double *__restrict__ vtau) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int fam = xc_family_mgga(id);
  const double ra = rho[2 * i], rb = rho[2 * i + 1];
  if (!(ra + rb >= kDensThreshold)) return;
  double saa = 0, sab = 0, sbb = 0, ta = 0, tb = 0;
  if (fam >= 1) {
    saa = fmax(sigma[3 * i], kSigmaMin);
    sab = sigma[3 * i + 1];
    sbb = fmax(sigma[3 * i + 2], kSigmaMin);
  }
  if (fam == 2) {
    ta = tau[2 * i];
    tb = tau[2 * i + 1];
  }
  double e = 0.0, vr[2] = {0, 0}, vs[3] = {0, 0, 0}, vt[2] = {0, 0};
  xc_point_mgga_pol(id, ra, rb, saa, sab, sbb, ta, tb, e, vr, vs, vt);
  exc[i] += e / (ra + rb);
  vrho[2 * i] += vr[0];
  vrho[2 * i + 1] += vr[1];
  if (fam >= 1)
    for (int k = 0; k < 3; ++k) vsigma[3 * i + k] += vs[k];
  if (fam == 2) {
    vtau[2 * i] += vt[0];
    vtau[2 * i + 1] += vt[1];
  }
}

__global__ void k_fill_int(int *__restrict__ p, int n, int v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// FP32 AO storage: partial[m * kNcorrBlocks + b] = sum over block b's
// contiguous share (of gridDim.x) of molecule m's nbf^2 entries of D2 dS (a
// fixed order: every build gives the same sum), then k_ncorr_final adds the
// blocks in order to N_elec.
__global__ __launch_bounds__(256) void k_ncorr_partial(const double *__restrict__ D2,
                                                       const double *__restrict__ dS,
                                                       const MolDev *__restrict__ mols,
                                                       double *__restrict__ partial) {
  __shared__ double s_red[4];
  const MolDev M = mols[blockIdx.y];
  const long long n = (long long)M.nbf * M.nbf;
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long k0 = blockIdx.x * per, k1 = min(n, k0 + per);
  double v = 0.0;
  for (long long k = k0 + threadIdx.x; k < k1; k += 256) v += D2[M.moff + k] * dS[M.moff + k];
  v = block_sum_256(v, s_red);
  if (threadIdx.x == 0) partial[(long long)blockIdx.y * kNcorrBlocks + blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_ncorr_final(const double *__restrict__ partial, int nb,
                                                     double *__restrict__ result) {
  __shared__ double s_red[4];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += partial[(long long)blockIdx.x * kNcorrBlocks + b];
  const double t = block_sum_256(v, s_red);
  if (threadIdx.x == 0) result[2 * blockIdx.x + 1] += t;
}

__global__ void k_finalize_matrix(const double *__restrict__ K, double *__restrict__ V,
                                  const MolDev *__restrict__ mols) {
  const MolDev M = mols[blockIdx.y];
  const int nbf = M.nbf;
  const long long n = (long long)nbf * nbf;
  const double *Km = K + M.moff;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(k % nbf), j = (int)(k / nbf);
    V[M.moff + k] = Km[k] + Km[j + (long long)i * nbf];
  }
}

// one workgroup per molecule: {Exc, N_elec} over its leaves in a fixed order
__global__ __launch_bounds__(256) void k_finalize_scalars(const double *__restrict__ part,
                                                          const MolDev *__restrict__ mols,
                                                          double *__restrict__ result) {
  __shared__ double s_red[4];
  const MolDev M = mols[blockIdx.x];
  double e = 0.0, n = 0.0;
  for (int i = M.leaf_begin + threadIdx.x; i < M.leaf_end; i += 256) {
    e += part[2 * i];
    n += part[2 * i + 1];
  }
  e = block_sum_256(e, s_red);
  n = block_sum_256(n, s_red);
  if (threadIdx.x == 0) {
    result[2 * blockIdx.x] = e;
    result[2 * blockIdx.x + 1] = n;
  }
}

// ------------------------------------------------- stage entry points --
__global__ __launch_bounds__(256) void k_eval_basis_flat(const ShellDev *__restrict__ shells,
                                                         int nshell,
                                                         const double *__restrict__ exps,
                                                         const double *__restrict__ coefs,
                                                         const double *__restrict__ xyz,
                                                         long long npts, int deriv, int nbf,
                                                         double *__restrict__ out) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
  const long long cs = (long long)nbf * npts;
  for (int s = 0; s < nshell; ++s) {
    const ShellDev sh = shells[s];
    double o[4][kNCartMax];
    eval_shell(sh, exps, coefs, x - sh.x, y - sh.y, z - sh.z, deriv != 0, o);
    for (int c = 0; c < sh.size; ++c) {
      const long long row = (long long)(sh.bf0 + c) * npts + p;
      out[row] = o[0][c];
      if (deriv) {
        out[cs + row] = o[1][c];
        out[2 * cs + row] = o[2][c];
        out[3 * cs + row] = o[3][c];
      }
    }
  }
}

__global__ void k_decay_cutoff(const ShellDev *__restrict__ shells, int nshell,
                               const double *__restrict__ exps,
                               const double *__restrict__ coefs, double thr,
                               double *__restrict__ ext) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nshell) return;
  const ShellDev sh = shells[s];
  const double max_radius = 30.0, increment = 0.1, min_radius = 0.5;
  const int npts = (int)((max_radius - min_radius) / increment) + 1;
  const double step = (max_radius - min_radius) / (double)(npts - 1);
  double result = 0.0;
  for (int i = 0; i < npts; ++i) {
    const double r = (i == npts - 1) ? max_radius : min_radius + (double)i * step;
    const double x = 1.0 * r + sh.x, y = 0.0 + sh.y, z = 0.0 + sh.z;
    double o[4][kNCartMax];
    eval_shell(sh, exps, coefs, x - sh.x, y - sh.y, z - sh.z, false, o);
    result = i * increment + min_radius;
    double v = 0.0;
    for (int c = 0; c < sh.size; ++c) v = fmax(v, fabs(o[0][c]));
    if (v < thr) break;
  }
  ext[s] = result;
}

__global__ void k_functional(int id, long long n, const double *__restrict__ rho,
                             const double *__restrict__ sigma, double *__restrict__ exc,
                             double *__restrict__ vrho, double *__restrict__ vsigma) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  double e = 0.0, vr = 0.0, vs = 0.0;
  if (r >= kDensThreshold) {
    const double s = sigma ? fmax(sigma[i], kSigmaMin) : 0.0;
    xc_point(id, r, s, e, vr, vs);
    e /= r;
  }
  exc[i] = e;
  vrho[i] = vr;
  if (vsigma) vsigma[i] = vs;
}

// Becke partition with Treutler-Ahlrichs adjustment for one point: the cell
// product P_k = prod_{j != k} s_kj accumulated in the pair order of
// calculate_atomic_grid_weights_impl (grid_utils.cpp:159-200), then
// w *= P_own / sum_k P_k (molecular_grid.cpp:77-78).
__global__ void k_becke(int natoms, const double *__restrict__ axyz,
                        const double *__restrict__ radii,
                        const double *__restrict__ dist, int ta, long long npts,
                        const double *__restrict__ pts,
                        const int *__restrict__ atom_of_point, double *__restrict__ w) {
  // no fused multiply-add here: the cell functions subtract nearly equal
  // numbers near cell boundaries, so keep the reference's separate roundings
#pragma clang fp contract(off)
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
  const int own = atom_of_point[p];
  double total = 0.0, mine = 0.0;
  for (int k = 0; k < natoms; ++k) {
    const double dxk = x - axyz[3 * k], dyk = y - axyz[3 * k + 1], dzk = z - axyz[3 * k + 2];
    const double dk = sqrt(dxk * dxk + dyk * dyk + dzk * dzk);
    double P = 1.0;
    for (int j = 0; j < natoms; ++j) {
      if (j == k) continue;
      const double dxj = x - axyz[3 * j], dyj = y - axyz[3 * j + 1], dzj = z - axyz[3 * j + 2];
      const double dj = sqrt(dxj * dxj + dyj * dyj + dzj * dzj);
      // pair (i, j') with i > j' as in the reference loop
      const int i = k > j ? k : j, jj = k > j ? j : k;
      const double di = k > j ? dk : dj, djj = k > j ? dj : dk;
      double v = (di - djj) / dist[i * natoms + jj];
      const double ri = radii[i], rj = radii[jj];
      if (fabs(ri - rj) > 1e-14) {
        const double xi = ta ? sqrt(ri / rj) : ri / rj;
        const double u = (xi - 1.0) / (xi + 1.0);
        const double aij = u / (u * u - 1.0);
        v += aij * (1.0 - v * v);
      }
      for (int it = 0; it < 3; ++it) v = (3.0 - v * v) * v * 0.5;
      if (k == i) {
        P = (fabs(1.0 - v) < 1e-14) ? 0.0 : P * (0.5 * (1.0 - v));
      } else if (!(fabs(1.0 - v) < 1e-14)) {
        P *= 0.5 * (1.0 + v);
      }
    }
    total += P;
    if (k == own) mine = P;
  }
  w[p] *= mine / total;
}

} // namespace

// ------------------------------------------------------------ launchers --
static int grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g > 65535LL * 64) g = 65535LL * 64;
  return (int)(g < 1 ? 1 : g);
}

static void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch: ") + hipGetErrorString(e));
}

void upload_sph_table() {
  // Real solid harmonics as Cartesian polynomials (Helgaker, Jorgensen,
  // Olsen, Molecular Electronic-Structure Theory, eqs. 6.4.47-6.4.50),
  // rows m = -l..l, columns in CCA Cartesian order.
  double table[512] = {0};
  int off[kMaxL + 1];
  auto fact = [](int n) { double r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; };
  auto binom = [&](int n, int k) { return (k < 0 || k > n) ? 0.0 : fact(n) / (fact(k) * fact(n - k)); };
  int pos = 0;
  for (int l = 0; l <= kMaxL; ++l) {
    off[l] = pos;
    const int nc = (l + 1) * (l + 2) / 2;
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      const double norm = 1.0 / (std::pow(2.0, am) * fact(l)) *
                          std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
      const int tv = m < 0 ? 1 : 0;  // 2 v_m
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u) {
          const int vmax2 = 2 * (int)std::floor(am / 2.0 - tv / 2.0) + tv;
          for (int twov = tv; twov <= vmax2; twov += 2) {
            const double sgn = ((t + (twov - tv) / 2) % 2) ? -1.0 : 1.0;
            const double c = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                             binom(t, u) * binom(am, twov);
            const int px = 2 * t + am - 2 * u - twov, py = 2 * u + twov;
            int ci = 0, idx = 0;
            for (int i = l; i >= 0; --i)
              for (int j = l - i; j >= 0; --j) {
                if (i == px && j == py) ci = idx;
                ++idx;
              }
            table[pos + (m + l) * nc + ci] += norm * c;
          }
        }
    }
    pos += (2 * l + 1) * nc;
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_sph), table, sizeof(double) * pos) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_sph_off), off, sizeof(off)) != hipSuccess)
    throw std::runtime_error("upload of the spherical transform table failed");
}

void launch_zero(double *p, long long n, hipStream_t s) {
  if (n == 0) return;
  k_zero<<<grid_for(n, 256) < 4096 ? grid_for(n, 256) : 4096, 256, 0, s>>>(p, n);
  check_launch();
}

void launch_fill_int(int *p, int n, int v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
  check_launch();
}

static dim3 mol_grid(int maxnbf, int nmol) {
  long long gx = ((long long)maxnbf * maxnbf + 255) / 256;
  gx = gx < 1 ? 1 : (gx > 1024 ? 1024 : gx);
  return dim3((unsigned)gx, (unsigned)nmol);
}

void launch_prepare_density(const double *D, double *D2, const MolDev *mols, int nmol,
                            int maxnbf, hipStream_t s) {
  if (nmol == 0) return;
  k_prepare_density<<<mol_grid(maxnbf, nmol), 256, 0, s>>>(D, D2, mols);
  check_launch();
}

void launch_ao_eval(const BuildArgs &a, void *ao, bool fp32, hipStream_t s) {
  if (a.nleaves == 0) return;
  const bool low = a.lmax <= 2;
  if (fp32) {
    if (low) k_ao_eval<float, 2><<<a.nleaves, 256, 0, s>>>(a, static_cast<float *>(ao));
    else k_ao_eval<float, kMaxL><<<a.nleaves, 256, 0, s>>>(a, static_cast<float *>(ao));
  } else {
    if (low) k_ao_eval<double, 2><<<a.nleaves, 256, 0, s>>>(a, static_cast<double *>(ao));
    else k_ao_eval<double, kMaxL><<<a.nleaves, 256, 0, s>>>(a, static_cast<double *>(ao));
  }
  check_launch();
}

void launch_ao_residuals(const BuildArgs &a, const void *ao, hipStream_t s) {
  const int n = a.res_leaf1 - a.res_leaf0;
  if (n <= 0) return;
  if (a.ao_res == nullptr) throw std::runtime_error("launch_ao_residuals: no residual buffer");
  float *x = const_cast<float *>(static_cast<const float *>(ao));  // not written with RES
  if (a.lmax <= 2) k_ao_eval<float, 2, true><<<n, 256, 0, s>>>(a, x);
  else k_ao_eval<float, kMaxL, true><<<n, 256, 0, s>>>(a, x);
  check_launch();
}

void launch_ao_extra(const BuildArgs &a, double *out, bool hessian, hipStream_t s) {
  if (a.nleaves == 0) return;
  const bool low = a.lmax <= 2;
  if (hessian) {
    if (low) k_ao_extra<2, 4, 10><<<a.nleaves, 256, 0, s>>>(a, out);
    else k_ao_extra<kMaxL, 4, 10><<<a.nleaves, 256, 0, s>>>(a, out);
  } else {
    if (low) k_ao_extra<2, 1, 4><<<a.nleaves, 256, 0, s>>>(a, out);
    else k_ao_extra<kMaxL, 1, 4><<<a.nleaves, 256, 0, s>>>(a, out);
  }
  check_launch();
}

void launch_density_functional(const BuildArgs &a, const void *ao, bool fp32, hipStream_t s) {
  if (a.nleaves == 0) return;
  if (fp32)
    k_density_functional<float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao));
  else
    k_density_functional<double><<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao));
  check_launch();
}

void launch_vxc(const BuildArgs &a, const void *ao, bool fp32, hipStream_t s) {
  if (a.nleaves == 0) return;
  if (fp32)
    k_vxc<float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao));
  else
    k_vxc<double><<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao));
  check_launch();
}

void launch_finalize_scalars(const MolDev *mols, int nmol, const double *leaf_part,
                             double *result, hipStream_t s) {
  if (nmol == 0) return;
  k_finalize_scalars<<<nmol, 256, 0, s>>>(leaf_part, mols, result);
  check_launch();
}

void launch_prepare_density_u(const double *D, double *D2a, double *D2b, int nbf,
                              hipStream_t s) {
  const long long n = (long long)nbf * nbf;
  if (n == 0) return;
  k_prepare_density_u<<<grid_for(n, 256), 256, 0, s>>>(D, D2a, D2b, nbf);
  check_launch();
}

void launch_functional_pol_leaves(const BuildArgs &a, const double *rho_a, const double *rho_b,
                                  double *coef_b, double *part, hipStream_t s) {
  if (a.nleaves == 0) return;
  k_functional_pol<<<a.nleaves, 256, 0, s>>>(a, rho_a, rho_b, coef_b, part);
  check_launch();
}

void launch_functional_mgga_leaves(const BuildArgs &a, const double *rho4, const double *tk,
                                   double *coef_tau, hipStream_t s) {
  if (a.nleaves == 0) return;
  k_functional_mgga<<<a.nleaves, 256, 0, s>>>(a, rho4, tk, coef_tau);
  check_launch();
}

void launch_functional_mgga_pol_leaves(const BuildArgs &a, const double *rho_a,
                                       const double *rho_b, const double *tka, const double *tkb,
                                       double *coef_b, double *ctau_a, double *ctau_b,
                                       double *part, hipStream_t s) {
  if (a.nleaves == 0) return;
  k_functional_mgga_pol<<<a.nleaves, 256, 0, s>>>(a, rho_a, rho_b, tka, tkb, coef_b, ctau_a,
                                                   ctau_b, part);
  check_launch();
}

void launch_functional_mgga(int id, long long n, const double *rho, const double *sigma,
                            const double *tau, double *exc, double *vrho, double *vsigma,
                            double *vtau, hipStream_t s) {
  if (n == 0) return;
  k_functional_mgga_points<<<grid_for(n, 256), 256, 0, s>>>(id, n, rho, sigma, tau, exc, vrho,
                                                            vsigma, vtau);
  check_launch();
}

void launch_finalize_u(const double *Ka, const double *Kb, double *Va, double *Vb,
                       const MolDev *mols, int maxnbf, const double *part, int nleaves,
                       double *result, hipStream_t s) {
  if (Ka) {
    k_finalize_matrix<<<mol_grid(maxnbf, 1), 256, 0, s>>>(Ka, Va, mols);
    check_launch();
    k_finalize_matrix<<<mol_grid(maxnbf, 1), 256, 0, s>>>(Kb, Vb, mols);
    check_launch();
  }
  k_finalize_scalars_u<<<1, 256, 0, s>>>(part, nleaves, result);
  check_launch();
}

void launch_functional_pol(int id, long long n, const double *rho, const double *sigma,
                           double *exc, double *vrho, double *vsigma, hipStream_t s) {
  if (n == 0) return;
  k_functional_pol_points<<<grid_for(n, 256), 256, 0, s>>>(id, n, rho, sigma, exc, vrho, vsigma);
  check_launch();
}

void launch_functional_mgga_pol(int id, long long n, const double *rho, const double *sigma,
                                const double *tau, double *exc, double *vrho, double *vsigma,
                                double *vtau, hipStream_t s) {
  if (n == 0) return;
  k_functional_mgga_pol_points<<<grid_for(n, 256), 256, 0, s>>>(id, n, rho, sigma, tau, exc,
                                                                 vrho, vsigma, vtau);
  check_launch();
}

void launch_symmetrize(const double *K, double *out, const MolDev *mols, int nmol, int maxnbf,
                       hipStream_t s) {
  if (nmol == 0) return;
  k_finalize_matrix<<<mol_grid(maxnbf, nmol), 256, 0, s>>>(K, out, mols);
  check_launch();
}

void launch_storage_correction(const double *D2, const double *dS, const MolDev *mols, int nmol,
                               int maxnbf, double *partial, double *result, hipStream_t s) {
  if (nmol == 0) return;
  // about 8192 entries per block for the largest molecule
  const long long nn = (long long)maxnbf * maxnbf;
  const int nb = (int)std::max(1LL, std::min((long long)kNcorrBlocks, (nn + 8191) / 8192));
  k_ncorr_partial<<<dim3(nb, nmol), 256, 0, s>>>(D2, dS, mols, partial);
  check_launch();
  k_ncorr_final<<<nmol, 256, 0, s>>>(partial, nb, result);
  check_launch();
}

void launch_finalize(const double *K, double *Vxc, const MolDev *mols, int nmol, int maxnbf,
                     const double *leaf_part, double *result, hipStream_t s) {
  if (nmol == 0) return;
  k_finalize_matrix<<<mol_grid(maxnbf, nmol), 256, 0, s>>>(K, Vxc, mols);
  check_launch();
  k_finalize_scalars<<<nmol, 256, 0, s>>>(leaf_part, mols, result);
  check_launch();
}

void launch_eval_basis_flat(const ShellDev *shells, int nshell, const double *exps,
                            const double *coefs, const double *xyz, long long npts, int deriv,
                            int nbf, double *out, hipStream_t s) {
  if (npts == 0) return;
  k_eval_basis_flat<<<grid_for(npts, 256), 256, 0, s>>>(shells, nshell, exps, coefs, xyz, npts,
                                                        deriv, nbf, out);
  check_launch();
}

void launch_decay_cutoff(const ShellDev *shells, int nshell, const double *exps,
                         const double *coefs, double threshold, double *extents,
                         hipStream_t s) {
  if (nshell == 0) return;
  k_decay_cutoff<<<grid_for(nshell, 64), 64, 0, s>>>(shells, nshell, exps, coefs, threshold,
                                                     extents);
  check_launch();
}

void launch_functional(int id, long long n, const double *rho, const double *sigma, double *exc,
                       double *vrho, double *vsigma, hipStream_t s) {
  if (n == 0) return;
  k_functional<<<grid_for(n, 256), 256, 0, s>>>(id, n, rho, sigma, exc, vrho, vsigma);
  check_launch();
}

void launch_becke(int natoms, const double *atom_xyz, const double *radii, const double *dist,
                  int ta, long long npts, const double *pts, const int *atom_of_point, double *w,
                  hipStream_t s) {
  if (npts == 0) return;
  k_becke<<<grid_for(npts, 128), 128, 0, s>>>(natoms, atom_xyz, radii, dist, ta, npts, pts,
                                              atom_of_point, w);
  check_launch();
}

// ---- allocation-free Becke partition (occxc_becke_partition_noalloc) ----
// The inter-atomic distance table (natoms x natoms FP64, at most
// kNoAllocMaxAtoms atoms: 8 MiB) lives in this file-scope device array.
// Every GPU the library runs on holds its own copy until the process ends;
// host code never passes the array's name to a kernel, only the address
// hipGetSymbolAddress returned for the current device.  All calls on one GPU
// share the table: calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxAtoms * kNoAllocMaxAtoms];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// the distances in the host restatement's operation order (no contraction),
// so that both entry points see the same table
__global__ void k_atom_distances(int natoms, const double *__restrict__ axyz,
                                 double *__restrict__ dist) {
#pragma clang fp contract(off)
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)natoms * natoms) return;
  const int i = (int)(idx / natoms), j = (int)(idx - (long long)i * natoms);
  double d = 0.0;
  if (i != j) {
    const double dx = axyz[3 * i] - axyz[3 * j], dy = axyz[3 * i + 1] - axyz[3 * j + 1],
                 dz = axyz[3 * i + 2] - axyz[3 * j + 2];
    d = sqrt(dx * dx + dy * dy + dz * dz);
  }
  dist[idx] = d;
}
} // namespace

double *noalloc_scratch(int device) {
  OCCXC_REQUIRE(device >= 0 && device < kMaxScratchDevices, OCCXC_ERR_INVALID_ARGUMENT,
                "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    OCCXC_HIP(hipGetDevice(&cur));
    OCCXC_REQUIRE(cur == device, OCCXC_ERR_INTERNAL, "scratch set-up on a device that is not current");
    void *p = nullptr;
    OCCXC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  return g_scratch_ptr[device];
}

void becke_partition_noalloc(int natoms, const double *atom_xyz, const double *radii, int ta,
                             long long npts, const double *pts, const int *atom_of_point,
                             double *w, hipStream_t s) {
  OCCXC_REQUIRE(natoms >= 1 && natoms <= kNoAllocMaxAtoms, OCCXC_ERR_INVALID_ARGUMENT,
                "becke_partition_noalloc: natoms must be 1.." + std::to_string(kNoAllocMaxAtoms) +
                    " (the size of its fixed device scratch table)");
  int dev = 0;
  OCCXC_HIP(hipGetDevice(&dev));
  double *dist = noalloc_scratch(dev);
  const long long nn = (long long)natoms * natoms;
  k_atom_distances<<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(natoms, atom_xyz, dist);
  check_launch();
  launch_becke(natoms, atom_xyz, radii, dist, ta, npts, pts, atom_of_point, w, s);
}

} // namespace occxc
