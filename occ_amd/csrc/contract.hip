// contract.hip -- MFMA contraction kernels of the XC build for gfx950.
//
// Per active leaf (n_b <= 256 points, m_b local AOs padded to 32):
//   k_density_mfma: T = Phi D_loc on the matrix cores (16x16x4 tiles, one
//     wavefront per 32-point block, D_loc columns gathered from the doubled
//     density through the leaf's AO map), then per point
//       rho = sum_nu Phi(nu,p) T(p,nu),  grad rho = 2 sum_nu grad Phi(nu,p) T(p,nu)
//     accumulated in FP64, the functional in FP64, and the Vxc coefficients
//       c0 = w vrho / 2,  c_k = 2 w vsigma d_k rho
//     (density.cpp:32-43, dft_kernels.cpp:8-70, xc_potential_matrix.cpp:19-41).
//   k_vxc_mfma: A = c0 Phi + c . grad Phi is built 32 AOs at a time in LDS,
//     K_loc = Phi^T A on the matrix cores, and K_loc^T is scattered into the
//     global accumulator with FP64 atomics (Vxc = K + K^T afterwards).
//
// MFMA operand layout for 16x16x4 (checked on the device): lane l supplies
// A[l%16][l/16] and B[l/16][l%16].  Accumulator register r of lane l holds
// D[row][l%16] with row = 4*(l/16)+r for the f32 instruction and
// row = l/16 + 4r for the f64 one (Mfma<T>::row).  The contraction index is
// permuted within each group of 16 so that lane group g uses k = 4g+s at step
// s (s = 0..3): the four values a lane needs along a point-fastest row are then
// contiguous (one vector load).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "contract.hpp"
#include "tuning_table.hpp"
#include "xc_functionals.hpp"

namespace occxc {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
// four doubles at a 16-byte (not 32-byte) aligned LDS address
typedef double d4u __attribute__((ext_vector_type(4))) __attribute__((aligned(16)));
typedef double d2u __attribute__((ext_vector_type(2))) __attribute__((aligned(16)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = d4;
  using vec_t = d4;
  __device__ static acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return g + 4 * r; }
  // values at rows row(g, 0..3) of a point-fastest array
  template <typename S> __device__ static d4 rows4(const S *base, int g) {
    return d4{(double)base[g], (double)base[g + 4], (double)base[g + 8], (double)base[g + 12]};
  }
};
template <> struct Mfma<float> {
  using acc_t = f4;
  using vec_t = f4;
  __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return 4 * g + r; }
  template <typename S> __device__ static f4 rows4(const S *base, int g) {
    return *reinterpret_cast<const f4 *>(base + 4 * g);
  }
};

// Four consecutive stored AO values (S = storage type), converted to the
// arithmetic type T.
template <typename S> struct Vec4Of;
template <> struct Vec4Of<float> { using type = f4; };
template <> struct Vec4Of<double> { using type = d4; };

template <typename T>
__device__ __forceinline__ typename Mfma<T>::vec_t load4(const T *p) {
  return *reinterpret_cast<const typename Mfma<T>::vec_t *>(p);
}
template <typename T, typename S>
__device__ __forceinline__ typename Mfma<T>::vec_t load4_as(const S *p) {
  const auto v = *reinterpret_cast<const typename Vec4Of<S>::type *>(p);
  return typename Mfma<T>::vec_t{(T)v[0], (T)v[1], (T)v[2], (T)v[3]};
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Give XCD x (blocks with blockIdx % 8 == x) a contiguous range of the
// Morton-ordered leaves, so the D and K entries it touches form a compact
// spatial block that stays in its L2.  A bijection of [0, n): correctness does
// not depend on the dispatch order.
__device__ __forceinline__ int xcd_leaf(int b, int n) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, i = b / kXcd, q = n / kXcd, r = n % kXcd;
  return x * q + min(x, r) + i;
}

__device__ __forceinline__ double sum16(double v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

// DPP moves within each 16-lane row: mirror (l -> 15 - l), half mirror
// (l -> l ^ 7) and the quad permutations l -> l ^ 2, l -> l ^ 1.  Each pairs
// lanes, and together they flip bits 3, 2, 1 and 0 of the lane in turn.
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppXor2 = 0x4E, kDppXor1 = 0xB1;
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// One stage of a transpose-reduce over the first N of v: the lane keeps one
// half (the upper if hi) and adds its DPP partner's copy of that half, which
// the partner (hi flipped) sends.
template <int CTRL, int N, int M>
__device__ __forceinline__ void tr_stage(double (&v)[M], bool hi) {
#pragma unroll
  for (int i = 0; i < N / 2; ++i) {
    const double keep = hi ? v[i + N / 2] : v[i], send = hi ? v[i] : v[i + N / 2];
    v[i] = keep + dpp_f64<CTRL>(send);
  }
}
// Sums over the 16 lanes of each row of 16 values per lane in 15 DPP
// exchanges (16 sum16 calls take 64 shuffles): lane l16 returns the total of
// value l16.
__device__ __forceinline__ double row_reduce16(double (&v)[16], int l16) {
  tr_stage<kDppRowMirror, 16>(v, (l16 >> 3) & 1);
  tr_stage<kDppRowHalfMirror, 8>(v, (l16 >> 2) & 1);
  tr_stage<kDppXor2, 4>(v, (l16 >> 1) & 1);
  tr_stage<kDppXor1, 2>(v, l16 & 1);
  return v[0];
}
// The same for 4 values per lane: lane l16 returns the total of value l16 / 4.
__device__ __forceinline__ double row_reduce4(double (&v)[4], int l16) {
  tr_stage<kDppRowMirror, 4>(v, (l16 >> 3) & 1);
  tr_stage<kDppRowHalfMirror, 2>(v, (l16 >> 2) & 1);
  double x = v[0];
  x += dpp_f64<kDppXor2>(x);
  x += dpp_f64<kDppXor1>(x);
  return x;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ double block_sum_256(double v, double *scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  const double r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
  __syncthreads();
  return r;
}
__device__ double block_max_256(double v, double *scratch) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  const double r = fmax(fmax(scratch[0], scratch[1]), fmax(scratch[2], scratch[3]));
  __syncthreads();
  return r;
}

// Per point (one thread each, blockDim = 256): N_elec before the leaf skip,
// the max-rho skip, the functional in FP64, Result::weight_by, the Vxc
// coefficients c0 = w vrho / 2, c_k = 2 w vsigma d_k rho, and the leaf's
// {Exc, N_elec} partials (dft_kernels.cpp:8-70, xc_potential_matrix.cpp:19-41).
__device__ void density_epilogue(const BuildArgs &a, const LeafDev &L, int leaf,
                                 double (*s_q)[kMaxLeafPts], double *s_red) {
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const double rho = live ? s_q[0][p] : 0.0;
  const double gx = live ? s_q[1][p] : 0.0, gy = live ? s_q[2][p] : 0.0,
               gz = live ? s_q[3][p] : 0.0;
  const double w = live ? a.pw[L.pt_off + p] : 0.0;
  const double nel = block_sum_256(real ? rho * w : 0.0, s_red);
  const double rmax = block_max_256(real ? rho : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e_sum = 0.0, vr_sum = 0.0, vs_sum = 0.0;
  if (active && live && rho >= kDensThreshold) {
    const double sg = fmax(gx * gx + gy * gy + gz * gz, kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f) {
      double e, vr, vs;
      xc_point(a.func_ids[f], rho, sg, e, vr, vs);
      e_sum += e;
      vr_sum += vr;
      vs_sum += vs;
    }
  }
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

// ------------------------------------------------------------- density --
// T = Phi D_loc for one leaf.  Each wavefront owns a 32-point block; all four
// walk the same (nu chunk of 64, mu chunk of 32) sequence so that the D_loc
// tile of the step is gathered once into LDS and shared.  Phi (A operand)
// comes from the resident AO block, D (B operand) from LDS.
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_density_mfma(BuildArgs a, const S *__restrict__ ao,
                                                      const T *__restrict__ D2t, int min_mp) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int NV = 64, KC = 32, LDD = NV + 16 / (int)sizeof(T);
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  __shared__ T s_D[KC * LDD];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  if (L.mp < min_mp) return;  // the per-wavefront kernels handle it
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  const int nblk = L.np >> 5;
  const int nround = (nblk + 3) >> 2;
  for (int round = 0; round < nround; ++round) {
    const int blk = round * 4 + wave;
    const bool mine = blk < nblk;
    const int p0 = (mine ? blk : 0) * 32;
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int c = 0; c < L.mp; c += NV) {
      const int ntc = min(NV, L.mp - c) >> 4;  // 16-wide nu tiles in this chunk (2 or 4)
      acc_t acc[2][4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = M::zero();
      for (int k = 0; k < L.mp; k += KC) {
        // gather D_loc[k..k+32)[c..c+16*ntc) into LDS (nu fastest: coalesced)
        for (int idx = threadIdx.x; idx < KC * NV; idx += 256) {
          const int vl = idx & (NV - 1), ml = idx >> 6;
          T v = static_cast<T>(0);
          if (vl < 16 * ntc) v = D2t[MD.moff + (long long)map[k + ml] * MD.nbf + map[c + vl]];
          s_D[ml * LDD + vl] = v;
        }
        __syncthreads();
        if (mine) {
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
              const int ml = 16 * j + 4 * g + s4;
              const S *row = phi + (long long)(k + ml) * L.np + p0 + l16;
              const T a0 = static_cast<T>(row[0]), a1 = static_cast<T>(row[16]);
              const T *drow = s_D + ml * LDD + l16;
#pragma unroll
              for (int nt = 0; nt < 4; ++nt) {
                if (nt < ntc) {
                  const T b = drow[16 * nt];
                  acc[0][nt] = M::mma(a0, b, acc[0][nt]);
                  acc[1][nt] = M::mma(a1, b, acc[1][nt]);
                }
              }
            }
        }
        __syncthreads();
      }
      if (mine) {
#pragma unroll
        for (int pt = 0; pt < 2; ++pt)
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            if (nt >= ntc) continue;
            const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16;
            const auto f = M::template rows4<S>(phi + off, g);
#pragma unroll
            for (int r = 0; r < 4; ++r) part[pt][r][0] += (double)f[r] * (double)acc[pt][nt][r];
            if (deriv) {
              const auto fx = M::template rows4<S>(phi + cs + off, g);
              const auto fy = M::template rows4<S>(phi + 2 * cs + off, g);
              const auto fz = M::template rows4<S>(phi + 3 * cs + off, g);
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const double t = 2.0 * (double)acc[pt][nt][r];
                part[pt][r][1] += (double)fx[r] * t;
                part[pt][r][2] += (double)fy[r] * t;
                part[pt][r][3] += (double)fz[r] * t;
              }
            }
          }
      }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (mine && l16 == 0) s_q[q][p0 + pt * 16 + M::row(g, r)] = v;
        }
  }
  __syncthreads();

  density_epilogue(a, L, leaf, s_q, s_red);
}

// Per-wavefront variant: each wavefront owns a 32-point block and walks all
// (nu chunk of 32, mu) pairs on its own, with no block-wide synchronisation.
// The leaf's AO-map column offsets (map[mu] * nbf) sit in LDS, and the
// operands of the next 16-mu step are loaded before the MFMAs of the current
// one are issued.
constexpr int kMapCap = 2048;
constexpr int kDefaultDensityVariant = 2;
// tiled density items staged by direct-to-LDS loads (k_den_tile_dma; measured
// faster than the register-staged k_den_tile in both storage precisions);
// OCCXC_DENSITY_DMA=0/1 overrides
constexpr bool kDefaultDensityDma = true;

// STORE: write rho and grad rho of every (padded) point to a.rho_out
// instead of running the functional epilogue (the unrestricted build
// evaluates each spin density this way; with a.deriv = 0 only rho is
// written, one plane).
template <typename S, typename T, bool STORE = false>
__global__ __launch_bounds__(256, 2) void k_density_wave(BuildArgs a, const S *__restrict__ ao,
                                                      const T *__restrict__ D2t, int min_mp) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  __shared__ int s_col[kMapCap];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  if (L.mp < min_mp || L.mp > kMapCap) return;  // k_density_blk / k_density_mfma
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  // A-operand point of this lane within a 16-point tile: the f64 instruction
  // returns rows g, g+4, g+8, g+12 to lane group g, so feeding point
  // 4(i%4) + i/4 as row i gives each lane group 4 consecutive points (one
  // vector load in the epilogue); the f32 instruction needs no permutation
  const int pa = sizeof(T) == 8 ? 4 * (l16 & 3) + (l16 >> 2) : l16;
  const S *phi = ao + L.ao_off;
  const T *Dm = D2t + MD.moff;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  for (int i = threadIdx.x; i < L.mp; i += blockDim.x) s_col[i] = map[i] * MD.nbf;
  __syncthreads();
  const int nblk = L.np >> 5;
  for (int blk = wave; blk < nblk; blk += 4) {
    const int p0 = blk * 32;
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int c = 0; c < L.mp; c += 32) {
      acc_t acc[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::zero();
      const int gn0 = map[c + l16], gn1 = map[c + 16 + l16];
      T ca[4][2], cb[4][2];
      auto load = [&](int mc, T (&xa)[4][2], T (&xb)[4][2]) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int mu = mc + 4 * g + s4;
          const S *row = phi + (long long)mu * L.np + p0 + pa;
          const int col = s_col[mu];
          xa[s4][0] = static_cast<T>(row[0]);
          xa[s4][1] = static_cast<T>(row[16]);
          xb[s4][0] = Dm[col + gn0];
          xb[s4][1] = Dm[col + gn1];
        }
      };
      load(0, ca, cb);
      for (int mc = 0; mc < L.mp; mc += 16) {
        T na[4][2], nb[4][2];
        const bool more = mc + 16 < L.mp;
        if (more) load(mc + 16, na, nb);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          acc[0][0] = M::mma(ca[s4][0], cb[s4][0], acc[0][0]);
          acc[0][1] = M::mma(ca[s4][0], cb[s4][1], acc[0][1]);
          acc[1][0] = M::mma(ca[s4][1], cb[s4][0], acc[1][0]);
          acc[1][1] = M::mma(ca[s4][1], cb[s4][1], acc[1][1]);
        }
        if (more) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            ca[s4][0] = na[s4][0];
            ca[s4][1] = na[s4][1];
            cb[s4][0] = nb[s4][0];
            cb[s4][1] = nb[s4][1];
          }
        }
      }
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16;
          const auto f = load4_as<T>(phi + off + 4 * g);
#pragma unroll
          for (int r = 0; r < 4; ++r) part[pt][r][0] += (double)f[r] * (double)acc[pt][nt][r];
          if (deriv) {
            const auto fx = load4_as<T>(phi + cs + off + 4 * g);
            const auto fy = load4_as<T>(phi + 2 * cs + off + 4 * g);
            const auto fz = load4_as<T>(phi + 3 * cs + off + 4 * g);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double t = 2.0 * (double)acc[pt][nt][r];
              part[pt][r][1] += (double)fx[r] * t;
              part[pt][r][2] += (double)fy[r] * t;
              part[pt][r][3] += (double)fz[r] * t;
            }
          }
        }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (l16 == 0) s_q[q][p0 + pt * 16 + 4 * g + r] = v;
        }
  }
  __syncthreads();
  if constexpr (STORE) {
    const int nq = deriv ? 4 : 1;
    for (int p = threadIdx.x; p < L.np; p += blockDim.x)
      for (int q = 0; q < nq; ++q) a.rho_out[q * a.ptot + L.pt_off + p] = s_q[q][p];
  } else {
    density_epilogue(a, L, leaf, s_q, s_red);
  }
}

// Register-lean per-wavefront variant: as k_density_wave, but each 32-AO
// chunk's contributions to rho and grad rho are reduced and added into LDS at
// once instead of being kept in registers across chunks (three waves per
// SIMD instead of two).
template <typename S, typename T>
__global__ __launch_bounds__(256, 3) void k_density_lean(BuildArgs a, const S *__restrict__ ao,
                                                      const T *__restrict__ D2t, int min_mp) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  __shared__ int s_col[kMapCap];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  if (L.mp < min_mp || L.mp > kMapCap) return;  // k_density_mfma handles it
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  // A-operand point of this lane within a 16-point tile: the f64 instruction
  // returns rows g, g+4, g+8, g+12 to lane group g, so feeding point
  // 4(i%4) + i/4 as row i gives each lane group 4 consecutive points (one
  // vector load in the epilogue); the f32 instruction needs no permutation
  const int pa = sizeof(T) == 8 ? 4 * (l16 & 3) + (l16 >> 2) : l16;
  const S *phi = ao + L.ao_off;
  const T *Dm = D2t + MD.moff;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  for (int i = threadIdx.x; i < L.mp; i += blockDim.x) s_col[i] = map[i] * MD.nbf;
  for (int i = threadIdx.x; i < 4 * kMaxLeafPts; i += blockDim.x) (&s_q[0][0])[i] = 0.0;
  __syncthreads();
  const int nblk = L.np >> 5;
  for (int blk = wave; blk < nblk; blk += 4) {
    const int p0 = blk * 32;
    for (int c = 0; c < L.mp; c += 32) {
      acc_t acc[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::zero();
      const int gn0 = map[c + l16], gn1 = map[c + 16 + l16];
      T ca[4][2], cb[4][2];
      auto load = [&](int mc, T (&xa)[4][2], T (&xb)[4][2]) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int mu = mc + 4 * g + s4;
          const S *row = phi + (long long)mu * L.np + p0 + pa;
          const int col = s_col[mu];
          xa[s4][0] = static_cast<T>(row[0]);
          xa[s4][1] = static_cast<T>(row[16]);
          xb[s4][0] = Dm[col + gn0];
          xb[s4][1] = Dm[col + gn1];
        }
      };
      // no explicit prefetch: three waves per SIMD cover the load latency
      for (int mc = 0; mc < L.mp; mc += 16) {
        load(mc, ca, cb);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          acc[0][0] = M::mma(ca[s4][0], cb[s4][0], acc[0][0]);
          acc[0][1] = M::mma(ca[s4][0], cb[s4][1], acc[0][1]);
          acc[1][0] = M::mma(ca[s4][1], cb[s4][0], acc[1][0]);
          acc[1][1] = M::mma(ca[s4][1], cb[s4][1], acc[1][1]);
        }
      }
      // this chunk's contributions, reduced over the 16 nu lanes at once and
      // added into the wave's own points (no register accumulators kept)
#pragma unroll
      for (int pt = 0; pt < 2; ++pt) {
        double q4[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) q4[r][q] = 0.0;
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16;
          const auto f = load4_as<T>(phi + off + 4 * g);
#pragma unroll
          for (int r = 0; r < 4; ++r) q4[r][0] += (double)f[r] * (double)acc[pt][nt][r];
          if (deriv) {
            const auto fx = load4_as<T>(phi + cs + off + 4 * g);
            const auto fy = load4_as<T>(phi + 2 * cs + off + 4 * g);
            const auto fz = load4_as<T>(phi + 3 * cs + off + 4 * g);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double t = 2.0 * (double)acc[pt][nt][r];
              q4[r][1] += (double)fx[r] * t;
              q4[r][2] += (double)fy[r] * t;
              q4[r][3] += (double)fz[r] * t;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double v = sum16(q4[r][q]);
            if (l16 == 0) s_q[q][p0 + pt * 16 + 4 * g + r] += v;
          }
      }
    }
  }
  __syncthreads();
  density_epilogue(a, L, leaf, s_q, s_red);
}

// LDS-staged variant: the leaf's Phi is staged in LDS kBlkPts points at a
// time (all mp AO rows, converted to T), so Phi is read from HBM once; the
// four wavefronts share it as the A operand and split the 16-wide nu tiles
// (wave w: tiles w, w + 4, ...).  D_loc columns (B operand) are gathered
// through the AO map, one 16-mu step ahead of the MFMAs.  Phi(nu, p) for rho
// comes from the same LDS block, grad Phi from HBM (each element once).  The
// four wavefronts' partial sums are added in a fixed order.
constexpr int kBlkPts = 32;
constexpr int kBlkMaxMp = 256;
constexpr int kBlkLd = kBlkPts + 1;

static size_t density_blk_smem(int max_mp) {
  return sizeof(double) * (4 * kMaxLeafPts + 4 * 4 * kBlkPts + 4) + sizeof(int) * kBlkMaxMp +
         sizeof(double) * (size_t)max_mp * kBlkLd;
}

template <typename S, typename T>
__global__ __launch_bounds__(256, 2) void k_density_blk(BuildArgs a, const S *__restrict__ ao,
                                                        const T *__restrict__ D2t) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  extern __shared__ double smem[];
  double(*s_q)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(smem);
  double *s_w = smem + 4 * kMaxLeafPts;  // [wave][q][kBlkPts]
  double *s_red = s_w + 4 * 4 * kBlkPts;
  int *s_col = reinterpret_cast<int *>(s_red + 4);
  T *s_phi = reinterpret_cast<T *>(s_col + kBlkMaxMp);  // [mp][kBlkLd]
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  if (L.mp > kBlkMaxMp) return;  // k_density_wave handles it
  const MolDev MD = a.mols[L.mol];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const S *phi = ao + L.ao_off;
  const T *Dm = D2t + MD.moff;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  for (int i = t; i < L.mp; i += blockDim.x) s_col[i] = map[i] * MD.nbf;
  const int ntile = L.mp >> 4;
  for (int p0 = 0; p0 < L.np; p0 += kBlkPts) {
    __syncthreads();
    for (int idx = t; idx < L.mp * kBlkPts; idx += blockDim.x) {
      const int mu = idx / kBlkPts, p = idx - mu * kBlkPts;
      s_phi[mu * kBlkLd + p] = static_cast<T>(phi[(long long)mu * L.np + p0 + p]);
    }
    __syncthreads();
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int nt = wave; nt < ntile; nt += 4) {
      acc_t acc[2] = {M::zero(), M::zero()};
      const int nu = nt * 16 + l16;
      const int gn = map[nu];
      T cb[4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) cb[s4] = Dm[s_col[4 * g + s4] + gn];
      for (int mc = 0; mc < L.mp; mc += 16) {
        T nb[4];
        const bool more = mc + 16 < L.mp;
        if (more) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) nb[s4] = Dm[s_col[mc + 16 + 4 * g + s4] + gn];
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const T *row = s_phi + (mc + 4 * g + s4) * kBlkLd + l16;
          acc[0] = M::mma(row[0], cb[s4], acc[0]);
          acc[1] = M::mma(row[16], cb[s4], acc[1]);
        }
        if (more) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) cb[s4] = nb[s4];
        }
      }
      // acc[pt][r] = T(p0 + 16 pt + row(g, r), nu)
      const T *prow = s_phi + nu * kBlkLd;
      const long long goff = (long long)nu * L.np + p0;
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = 16 * pt + M::row(g, r);
          const double tv = (double)acc[pt][r];
          part[pt][r][0] += (double)prow[p] * tv;
          if (deriv) {
            part[pt][r][1] += (double)phi[cs + goff + p] * 2.0 * tv;
            part[pt][r][2] += (double)phi[2 * cs + goff + p] * 2.0 * tv;
            part[pt][r][3] += (double)phi[3 * cs + goff + p] * 2.0 * tv;
          }
        }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (l16 == 0) s_w[(wave * 4 + q) * kBlkPts + 16 * pt + M::row(g, r)] = v;
        }
    __syncthreads();
    if (t < 4 * kBlkPts) {
      const int q = t / kBlkPts, p = t - q * kBlkPts;
      s_q[q][p0 + p] = s_w[q * kBlkPts + p] + s_w[(4 + q) * kBlkPts + p] +
                       s_w[(8 + q) * kBlkPts + p] + s_w[(12 + q) * kBlkPts + p];
    }
  }
  __syncthreads();
  density_epilogue(a, L, leaf, s_q, s_red);
}

// ----------------------------------------------------------------- Vxc --
// K_loc = Phi^T A for one leaf, A = c0 Phi + c . grad Phi built CH AOs at a
// time in LDS; each wavefront owns 16-row mu tiles (Phi from the resident AO
// block) and multiplies them with the CH-wide A chunk.  K_loc^T is scattered
// into the global accumulator (Vxc = K + K^T afterwards).
template <typename S, typename T, int CH>
__global__ __launch_bounds__(256, 2) void k_vxc_mfma(BuildArgs a, const S *__restrict__ ao,
                                                     int min_mp) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using vec_t = typename M::vec_t;
  constexpr int NT = CH / 16;
  // points per register-resident Phi segment (FP64: 64, to fit two waves per SIMD)
  constexpr int PSEG = sizeof(T) == 8 ? 64 : 128;
  extern __shared__ double smem[];
  double(*s_c)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(smem);
  T *s_A = reinterpret_cast<T *>(smem + 4 * kMaxLeafPts);
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  if (!a.leaf_active[leaf]) return;
  const LeafDev L = a.leaves[leaf];
  if (L.mp < min_mp) return;  // k_vxc_sym handles it
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = a.ao_map + L.map_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const bool deriv = a.deriv != 0;
  // chunk width: CH AOs for leaves of <= 128 points, half that above
  // (the LDS buffer holds CH x (128 + pad) = CH/2 x (256 + 2 pad) elements)
  const int chw = L.np <= 128 ? CH : CH / 2;
  const int ld = L.np + 16 / (int)sizeof(T);
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
    for (int k = 0; k < (deriv ? 4 : 1); ++k) s_c[k][i] = a.coef[k * a.ptot + L.pt_off + i];
  __syncthreads();
  const int ntile = L.mp >> 4;
  for (int c = 0; c < L.mp; c += chw) {
    const int ntc = min(chw, L.mp - c) >> 4;
    const int nel = 16 * ntc * L.np;
    // A chunk: each thread handles 4 consecutive points of a row per step
#pragma unroll 2
    for (int idx = 4 * threadIdx.x; idx < nel; idx += 4 * blockDim.x) {
      const int vl = idx / L.np, p = idx - vl * L.np;
      const long long o = (long long)(c + vl) * L.np + p;
      const vec_t f0 = load4_as<T>(phi + o);
      vec_t fx, fy, fz;
      if (deriv) {
        fx = load4_as<T>(phi + cs + o);
        fy = load4_as<T>(phi + 2 * cs + o);
        fz = load4_as<T>(phi + 3 * cs + o);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double v = s_c[0][p + e] * (double)f0[e];
        if (deriv)
          v += s_c[1][p + e] * (double)fx[e] + s_c[2][p + e] * (double)fy[e] +
               s_c[3][p + e] * (double)fz[e];
        s_A[vl * ld + p + e] = static_cast<T>(v);
      }
    }
    __syncthreads();
    for (int mt = wave; mt < ntile; mt += 4) {
      acc_t acc[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[j] = M::zero();
      const S *arow = phi + (long long)(mt * 16 + l16) * L.np + 4 * g;
      const T *brow = s_A + l16 * ld + 4 * g;
      for (int p0 = 0; p0 < L.np; p0 += PSEG) {
        // the wave's Phi rows for this point segment, all loads in flight at once
        vec_t av[PSEG / 16];
#pragma unroll
        for (int q = 0; q < PSEG / 16; ++q)
          if (p0 + 16 * q < L.np) av[q] = load4_as<T>(arow + p0 + 16 * q);
#pragma unroll
        for (int q = 0; q < PSEG / 16; ++q) {
          if (p0 + 16 * q >= L.np) break;
#pragma unroll
          for (int j = 0; j < NT; ++j) {
            if (j < ntc) {
              const vec_t bv = load4(brow + 16 * j * ld + p0 + 16 * q);
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) acc[j] = M::mma(av[q][s4], bv[s4], acc[j]);
            }
          }
        }
      }
      // acc[j][r] = K_loc(mu = mt*16 + row(g, r), nu = c + 16 j + l16); scatter
      // K_loc^T into K (Vxc = K + K^T afterwards)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int nu = c + 16 * j + l16;
        if (j >= ntc || nu >= L.m) continue;
        const long long gn = map[nu];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = mt * 16 + M::row(g, r);
          if (mu >= L.m) continue;
          if (a.kloc) a.kloc[L.kl_off + (long long)mu * L.mp + nu] = (double)acc[j][r];
          else atomicAdd(Km + gn + (long long)map[mu] * MD.nbf, (double)acc[j][r]);
        }
      }
    }
    __syncthreads();
  }
}

// Symmetric form for leaves of up to kSymMaxMp AOs: the block
//   V_loc = Phi^T A + A^T Phi
// is symmetric, so only its 16 x 16 tiles (I, J) with I <= J are formed,
// V_IJ = Phi_I^T A_J + A_I^T Phi_J (I < J) and K_II = Phi_I^T A_I, and only
// they are scattered: half the atomics of the full K_loc, the same flops.
// The leaf's points are streamed 16 at a time through LDS, where Phi and A
// (built from the four AO components and the per-point coefficients) are
// staged once and shared by the four wavefronts; each wavefront keeps its
// tiles' accumulators in registers for the whole leaf (tiles w, w + 4, ...
// of the upper triangle in row-major order).
// Scatter: V_IJ(mu, nu) (I < J, so map[mu] < map[nu]) goes to K(map[nu],
// map[mu]) only and K_II(mu, nu) to K(map[nu], map[mu]) as before; the
// finalisation V = K + K^T then gives every entry once.
constexpr int kSymPts = 16;
constexpr int kSymMaxMp = 192;
constexpr int kSymMaxTiles = kSymMaxMp / 16;
constexpr int kSymPerWave = (kSymMaxTiles * (kSymMaxTiles + 1) / 2 + 3) / 4;

template <typename T> constexpr int sym_ld() { return kSymPts + 16 / (int)sizeof(T); }

static size_t vxc_sym_smem(size_t elem, int max_mp, int ncoef = 4) {
  const size_t ld = kSymPts + 16 / elem;
  return sizeof(double) * ncoef * kMaxLeafPts + 2 * elem * ld * (size_t)max_mp;
}

// upper-triangle tile pair p (row-major over I <= J of an nt x nt grid)
__device__ __forceinline__ void sym_pair(int p, int nt, int &I, int &J) {
  int i = 0, rem = p;
  while (rem >= nt - i) {
    rem -= nt - i;
    ++i;
  }
  I = i;
  J = i + rem;
}

// TAU (restricted meta-GGA, GGA = true): after the GGA term, each chunk is
// staged three more times with d_k Phi in place of Phi and c_tau d_k Phi in
// place of A (c_tau = a.coef_tau), accumulating V += 2 d_k Phi^T c_tau d_k Phi
// into the same tile pairs, so the leaf is scattered once.
template <typename S, typename T, bool GGA, bool TAU = false>
__global__ __launch_bounds__(256, 1) void k_vxc_sym(BuildArgs a, const S *__restrict__ ao) {
  constexpr int NC = TAU ? 5 : 4;  // coefficient planes in LDS
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using vec_t = typename M::vec_t;
  constexpr int LD = sym_ld<T>();
  extern __shared__ double smem[];
  double(*s_c)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(smem);
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  if (!a.leaf_active[leaf]) return;
  const LeafDev L = a.leaves[leaf];
  if (L.mp > kSymMaxMp) return;  // k_vxc_mfma handles it
  T *s_phi = reinterpret_cast<T *>(smem + NC * kMaxLeafPts);
  T *s_A = s_phi + L.mp * LD;
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = a.ao_map + L.map_off;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  for (int i = t; i < L.np; i += blockDim.x) {
    for (int k = 0; k < (GGA ? 4 : 1); ++k) s_c[k][i] = a.coef[k * a.ptot + L.pt_off + i];
    if constexpr (TAU) s_c[4][i] = a.coef_tau[L.pt_off + i];
  }
  const int nt = L.mp >> 4;
  const int npair = nt * (nt + 1) / 2;
  acc_t acc[kSymPerWave];
#pragma unroll
  for (int k = 0; k < kSymPerWave; ++k) acc[k] = M::zero();
  // staging of one point chunk: groups of 4 consecutive points of one AO row;
  // each thread takes kSymGrp groups per step and issues every load of them
  // (4 points x 4 planes each, 32-byte aligned) before using any, so a step
  // costs one memory round trip (indices past the end are clamped for the
  // loads and skipped for the stores)
  constexpr int kSymGrp = 2;
  const int ngrp = L.mp * (kSymPts / 4);
  // one term of a chunk: stage (Phi, A) of term 0 or (d_k Phi, c_tau d_k Phi)
  // of term k, then the tile-pair MFMAs
  auto term_pass = [&](int p0, auto term_tag) {
    constexpr int TERM = decltype(term_tag)::value;
    const S *src = phi + (long long)TERM * cs;
    __syncthreads();
    for (int g0 = t; g0 < ngrp; g0 += kSymGrp * blockDim.x) {
      typename Mfma<double>::vec_t f[kSymGrp], fx[kSymGrp], fy[kSymGrp], fz[kSymGrp];
#pragma unroll
      for (int u = 0; u < kSymGrp; ++u) {
        const int gi = min(g0 + u * (int)blockDim.x, ngrp - 1);
        const long long o = (long long)(gi >> 2) * L.np + p0 + 4 * (gi & 3);
        f[u] = load4_as<double>(src + o);
        if constexpr (GGA && TERM == 0) {
          fx[u] = load4_as<double>(phi + cs + o);
          fy[u] = load4_as<double>(phi + 2 * cs + o);
          fz[u] = load4_as<double>(phi + 3 * cs + o);
        }
      }
#pragma unroll
      for (int u = 0; u < kSymGrp; ++u) {
        const int gi = g0 + u * (int)blockDim.x;
        if (gi >= ngrp) break;
        const int mu = gi >> 2, p = 4 * (gi & 3);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          double v = s_c[TERM == 0 ? 0 : 4][p0 + p + e] * f[u][e];
          if constexpr (GGA && TERM == 0)
            v += s_c[1][p0 + p + e] * fx[u][e] + s_c[2][p0 + p + e] * fy[u][e] +
                 s_c[3][p0 + p + e] * fz[u][e];
          s_phi[mu * LD + p + e] = static_cast<T>(f[u][e]);
          s_A[mu * LD + p + e] = static_cast<T>(v);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSymPerWave; ++k) {
      const int pr = wave + 4 * k;
      if (pr < npair) {
        int I, J;
        sym_pair(pr, nt, I, J);
        const int oi = (I * 16 + l16) * LD + 4 * g, oj = (J * 16 + l16) * LD + 4 * g;
        const vec_t fi = *reinterpret_cast<const vec_t *>(s_phi + oi);
        const vec_t aj = *reinterpret_cast<const vec_t *>(s_A + oj);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) acc[k] = M::mma(fi[s4], aj[s4], acc[k]);
        if (I != J) {
          const vec_t ai = *reinterpret_cast<const vec_t *>(s_A + oi);
          const vec_t fj = *reinterpret_cast<const vec_t *>(s_phi + oj);
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) acc[k] = M::mma(ai[s4], fj[s4], acc[k]);
        }
      }
    }
  };
  for (int p0 = 0; p0 < L.np; p0 += kSymPts) {
    term_pass(p0, std::integral_constant<int, 0>{});
    if constexpr (TAU) {
      term_pass(p0, std::integral_constant<int, 1>{});
      term_pass(p0, std::integral_constant<int, 2>{});
      term_pass(p0, std::integral_constant<int, 3>{});
    }
  }
  // acc[k][r] = tile (I, J) element (mu = 16 I + row(g, r), nu = 16 J + l16)
#pragma unroll
  for (int k = 0; k < kSymPerWave; ++k) {
    const int pr = wave + 4 * k;
    if (pr >= npair) continue;
    int I, J;
    sym_pair(pr, nt, I, J);
    const int nu = J * 16 + l16;
    if (nu >= L.m) continue;
    const long long gn = map[nu];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = I * 16 + M::row(g, r);
      if (mu >= L.m) continue;
      if (a.kloc) a.kloc[L.kl_off + (long long)mu * L.mp + nu] = (double)acc[k][r];
      else atomicAdd(Km + gn + (long long)map[mu] * MD.nbf, (double)acc[k][r]);
    }
  }
}

// ------------------------------------------------------ split-BF16 path --
// FP32 values are split exactly into three BF16 terms x = x0 + x1 + x2
// (8 significant bits each).  A product keeps the six cross terms with
// i + j <= 2 on v_mfma_f32_16x16x32_bf16 (16 x 16 outputs, K = 32), smallest
// first, with FP32 accumulation.  The density matrix (FP64) is split into
// four BF16 terms and its product with Phi also keeps x0 d3: the rounding of
// D is the same at every grid point, so unlike the per-point rounding of Phi
// and of the sums it does not average out of Exc and N_elec, and it is kept
// to 2^-32 instead of 2^-24.  Operand layout: lane l supplies A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16], i = 0..7; accumulator register r holds
// D[4(l/16)+r][l%16] (checked on the device).
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(h);
  m = static_cast<__bf16>(r1);
  l = static_cast<__bf16>(r1 - static_cast<float>(m));
}
__device__ __forceinline__ void split4(double x, __bf16 (&o)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    o[i] = static_cast<__bf16>(static_cast<float>(x));
    x -= static_cast<double>(static_cast<float>(o[i]));
  }
}
__device__ __forceinline__ void split8(const float (&x)[8], bf8 (&o)[3]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    __bf16 h, m, l;
    split3(x[i], h, m, l);
    o[0][i] = h;
    o[1][i] = m;
    o[2][i] = l;
  }
}
// Phi (3 terms) times D (4 terms): the six products of mma3 plus x0 d3
__device__ __forceinline__ f4 mma3d(const bf8 (&a)[3], const bf8 (&b)[4], f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[3], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
  return c;
}
__device__ __forceinline__ f4 mma3(const bf8 (&a)[3], const bf8 (&b)[3], f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
  return c;
}
__device__ __forceinline__ bf8 lds8(const __bf16 *p) { return *reinterpret_cast<const bf8 *>(p); }

// T = Phi D_loc with split-BF16 operands: D_loc tiles (nu-major, mu fastest)
// are split from FP64 into four BF16 planes in LDS; Phi (FP32 AO block) is
// split in registers.
__global__ __launch_bounds__(256) void k_density_bf16(BuildArgs a, const float *__restrict__ ao) {
  constexpr int NV = 64, KC = 32, LDK = KC + 8;
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  __shared__ __attribute__((aligned(16))) __bf16 s_D[4][NV * LDK];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const float *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  const int nblk = L.np >> 5;
  const int nround = (nblk + 3) >> 2;
  for (int round = 0; round < nround; ++round) {
    const int blk = round * 4 + wave;
    const bool mine = blk < nblk;
    const int p0 = (mine ? blk : 0) * 32;
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int c = 0; c < L.mp; c += NV) {
      const int ntc = min(NV, L.mp - c) >> 4;
      f4 acc[2][4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      for (int k = 0; k < L.mp; k += KC) {
        for (int idx = threadIdx.x; idx < KC * NV; idx += 256) {
          const int vl = idx & (NV - 1), ml = idx >> 6;
          double v = 0.0;
          if (vl < 16 * ntc) v = a.D2[MD.moff + (long long)map[k + ml] * MD.nbf + map[c + vl]];
          __bf16 d[4];
          split4(v, d);
#pragma unroll
          for (int t = 0; t < 4; ++t) s_D[t][vl * LDK + ml] = d[t];
        }
        __syncthreads();
        if (mine) {
          bf8 av[2][3];
#pragma unroll
          for (int pt = 0; pt < 2; ++pt) {
            float x[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
              x[i] = phi[(long long)(k + 8 * g + i) * L.np + p0 + 16 * pt + l16];
            split8(x, av[pt]);
          }
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            if (nt < ntc) {
              const int o = (16 * nt + l16) * LDK + 8 * g;
              const bf8 bv[4] = {lds8(&s_D[0][o]), lds8(&s_D[1][o]), lds8(&s_D[2][o]),
                                 lds8(&s_D[3][o])};
              acc[0][nt] = mma3d(av[0], bv, acc[0][nt]);
              acc[1][nt] = mma3d(av[1], bv, acc[1][nt]);
            }
          }
        }
        __syncthreads();
      }
      if (mine) {
#pragma unroll
        for (int pt = 0; pt < 2; ++pt)
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            if (nt >= ntc) continue;
            const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16 + 4 * g;
            const f4 f = load4(phi + off);
#pragma unroll
            for (int r = 0; r < 4; ++r) part[pt][r][0] += (double)f[r] * (double)acc[pt][nt][r];
            if (deriv) {
              const f4 fx = load4(phi + cs + off), fy = load4(phi + 2 * cs + off),
                       fz = load4(phi + 3 * cs + off);
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const double t = 2.0 * (double)acc[pt][nt][r];
                part[pt][r][1] += (double)fx[r] * t;
                part[pt][r][2] += (double)fy[r] * t;
                part[pt][r][3] += (double)fz[r] * t;
              }
            }
          }
      }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (mine && l16 == 0) s_q[q][p0 + pt * 16 + 4 * g + r] = v;
        }
  }
  __syncthreads();
  density_epilogue(a, L, leaf, s_q, s_red);
}

// K_loc = Phi^T A with split-BF16 operands: the A chunk (64 AOs for leaves of
// <= 128 points, 32 above) is split into three BF16 planes in LDS, Phi rows
// (8 consecutive points per lane) are split in registers.
__global__ __launch_bounds__(256) void k_vxc_bf16(BuildArgs a, const float *__restrict__ ao) {
  constexpr int CH = 64, PADP = 8;
  extern __shared__ double smem[];
  double(*s_c)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(smem);
  __bf16 *s_A = reinterpret_cast<__bf16 *>(smem + 4 * kMaxLeafPts);
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  if (!a.leaf_active[leaf]) return;
  const LeafDev L = a.leaves[leaf];
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = a.ao_map + L.map_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const float *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const bool deriv = a.deriv != 0;
  const int chw = L.np <= 128 ? CH : CH / 2;
  const int ld = L.np + PADP;
  const int plane = chw * ld;
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
    for (int k = 0; k < (deriv ? 4 : 1); ++k) s_c[k][i] = a.coef[k * a.ptot + L.pt_off + i];
  __syncthreads();
  const int ntile = L.mp >> 4;
  for (int c = 0; c < L.mp; c += chw) {
    const int ntc = min(chw, L.mp - c) >> 4;
    for (int idx = threadIdx.x; idx < 16 * ntc * L.np; idx += blockDim.x) {
      const int vl = idx / L.np, p = idx - vl * L.np;
      const long long o = (long long)(c + vl) * L.np + p;
      double v = s_c[0][p] * (double)phi[o];
      if (deriv)
        v += s_c[1][p] * (double)phi[cs + o] + s_c[2][p] * (double)phi[2 * cs + o] +
             s_c[3][p] * (double)phi[3 * cs + o];
      __bf16 h, m, l;
      split3(static_cast<float>(v), h, m, l);
      s_A[vl * ld + p] = h;
      s_A[plane + vl * ld + p] = m;
      s_A[2 * plane + vl * ld + p] = l;
    }
    __syncthreads();
    for (int mt = wave; mt < ntile; mt += 4) {
      f4 acc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
      const float *arow = phi + (long long)(mt * 16 + l16) * L.np + 8 * g;
      for (int pc = 0; pc < L.np; pc += 32) {
        const f4 x0 = load4(arow + pc), x1 = load4(arow + pc + 4);
        const float x[8] = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
        bf8 av[3];
        split8(x, av);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j < ntc) {
            const int o = (16 * j + l16) * ld + pc + 8 * g;
            const bf8 bv[3] = {lds8(s_A + o), lds8(s_A + plane + o), lds8(s_A + 2 * plane + o)};
            acc[j] = mma3(av, bv, acc[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int nu = c + 16 * j + l16;
        if (j >= ntc || nu >= L.m) continue;
        const long long gn = map[nu];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = mt * 16 + 4 * g + r;
          if (mu < L.m) atomicAdd(Km + gn + (long long)map[mu] * MD.nbf, (double)acc[j][r]);
        }
      }
    }
    __syncthreads();
  }
}

static size_t vxc_bf16_smem() {
  return sizeof(double) * 4 * kMaxLeafPts + sizeof(__bf16) * 3 * 64 * (128 + 8);
}

// ------------------------------------------- FP32 with compensated D --
// D2 = Dh + Dl with Dh the FP32 rounding of the doubled density and Dl the
// BF16 rounding of the remainder.  T = Phi Dh on the FP32 matrix cores plus
// Phi Dl on the BF16 ones, into the same FP32 accumulator: Phi Dl is 2^-24 of
// T, so Phi rounded to BF16 is enough for it.  The rounding of D is the same
// at every grid point and would otherwise not average out of Exc and N_elec.
// Per wavefront a 32-point block, as k_density_wave, in steps of 32 AOs: the
// FP32 instruction takes mu = 16h + 4g + s at sub-step (h, s) in lane group g,
// and the BF16 instruction takes the same eight mu as its K slots 4h + s.
__global__ __launch_bounds__(256, 2) void k_density_f32c(BuildArgs a, const float *__restrict__ ao,
                                                         const float *__restrict__ Dh,
                                                         const __bf16 *__restrict__ Dl) {
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  __shared__ int s_col[kMapCap];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const float *phi = ao + L.ao_off;
  const float *Dhm = Dh + MD.moff;
  const __bf16 *Dlm = Dl + MD.moff;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  for (int i = threadIdx.x; i < L.mp; i += blockDim.x) s_col[i] = map[i] * MD.nbf;
  __syncthreads();
  const int nblk = L.np >> 5;
  for (int blk = wave; blk < nblk; blk += 4) {
    const int p0 = blk * 32;
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int c = 0; c < L.mp; c += 32) {
      f4 acc[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      const int gn0 = map[c + l16], gn1 = map[c + 16 + l16];
      float ca[8][2], cb[8][2];
      auto load = [&](int mc, float (&xa)[8][2], float (&xb)[8][2]) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int mu = mc + 16 * (t >> 2) + 4 * g + (t & 3);
          const float *row = phi + (long long)mu * L.np + p0 + l16;
          const int col = s_col[mu];
          xa[t][0] = row[0];
          xa[t][1] = row[16];
          xb[t][0] = Dhm[col + gn0];
          xb[t][1] = Dhm[col + gn1];
        }
      };
      load(0, ca, cb);
      for (int mc = 0; mc < L.mp; mc += 32) {
        // the remainder terms of this step are loaded here and used after the
        // FP32 products, which cover their latency
        bf8 bl[2];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int col = s_col[mc + 16 * (t >> 2) + 4 * g + (t & 3)];
          bl[0][t] = Dlm[col + gn0];
          bl[1][t] = Dlm[col + gn1];
        }
        float na[8][2], nb[8][2];
        const bool more = mc + 32 < L.mp;
        if (more) load(mc + 32, na, nb);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[t][0], cb[t][0], acc[0][0], 0, 0, 0);
          acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[t][0], cb[t][1], acc[0][1], 0, 0, 0);
          acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[t][1], cb[t][0], acc[1][0], 0, 0, 0);
          acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[t][1], cb[t][1], acc[1][1], 0, 0, 0);
        }
        bf8 al[2];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          al[0][t] = static_cast<__bf16>(ca[t][0]);
          al[1][t] = static_cast<__bf16>(ca[t][1]);
        }
#pragma unroll
        for (int pt = 0; pt < 2; ++pt)
#pragma unroll
          for (int nt = 0; nt < 2; ++nt)
            acc[pt][nt] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[pt], bl[nt], acc[pt][nt], 0, 0, 0);
        if (more) {
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            ca[t][0] = na[t][0];
            ca[t][1] = na[t][1];
            cb[t][0] = nb[t][0];
            cb[t][1] = nb[t][1];
          }
        }
      }
      // acc[pt][nt][r] = T(p0 + 16 pt + 4g + r, c + 16 nt + l16)
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16 + 4 * g;
          const f4 f = load4(phi + off);
#pragma unroll
          for (int r = 0; r < 4; ++r) part[pt][r][0] += (double)f[r] * (double)acc[pt][nt][r];
          if (deriv) {
            const f4 fx = load4(phi + cs + off), fy = load4(phi + 2 * cs + off),
                     fz = load4(phi + 3 * cs + off);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double t = 2.0 * (double)acc[pt][nt][r];
              part[pt][r][1] += (double)fx[r] * t;
              part[pt][r][2] += (double)fy[r] * t;
              part[pt][r][3] += (double)fz[r] * t;
            }
          }
        }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (l16 == 0) s_q[q][p0 + pt * 16 + 4 * g + r] = v;
        }
  }
  __syncthreads();
  density_epilogue(a, L, leaf, s_q, s_red);
}

// D2 (FP64) -> Dh = fl32(D2), Dl = bf16(D2 - Dh)
__global__ void k_split_density(const double *__restrict__ x, float *__restrict__ hi,
                                __bf16 *__restrict__ lo, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const float h = static_cast<float>(x[k]);
    hi[k] = h;
    lo[k] = static_cast<__bf16>(static_cast<float>(x[k] - static_cast<double>(h)));
  }
}

__global__ void k_to_float(const double *__restrict__ x, float *__restrict__ y, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    y[k] = static_cast<float>(x[k]);
}

// ------------------------------------------------ two-pass Vxc reduction --
// With a.kloc set, the Vxc kernels store their per-leaf blocks instead of
// scattering them with atomics: the general kernel K_loc(mu, nu) in full, the
// symmetric kernel V_loc(mu, nu) = K_loc(mu, nu) + K_loc(nu, mu) for local
// 16-tiles t(mu) < t(nu) and K_loc(mu, nu) within a 16-tile.  One workgroup
// per global 32 x 32 AO tile pair (TI <= TJ) then sums, leaf by leaf in list
// order (deterministic), the stored entries that make up
//   V(i, j) = sum over leaves of K_loc(li, lj) + K_loc(lj, li)
// (dft.h:361-369 and the symmetrisation of xc_potential_matrix.cpp:35-40),
// and writes V(i, j) and V(j, i).
__device__ __forceinline__ double kloc_pair(const double *S, int mp, bool sym, int li, int lj) {
  if (!sym) return S[(long long)li * mp + lj] + S[(long long)lj * mp + li];
  const int ti = li >> 4, tj = lj >> 4;
  double v = 0.0;
  if (ti <= tj) v += S[(long long)li * mp + lj];
  if (tj <= ti) v += S[(long long)lj * mp + li];
  return v;
}

// Pass 2a: one workgroup per chunk of a tile pair's list; wavefront w sums
// the chunk's entries w, w + 4, ... (lanes over the 32 x 32 block, 16
// elements each), and the four wavefront sums are added in a fixed order into
// the chunk's partial block.
__global__ __launch_bounds__(256) void k_vxc_reduce_chunk(BuildArgs a, int sym_max_mp) {
  __shared__ double s_acc[4][1024];
  const TileChunkDev C = a.tchunks[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int j = lane & 31, i0 = lane >> 5;  // elements (i0 + 2 k, j), k = 0..15
  double acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  for (long long e = C.e0 + wave; e < C.e1; e += 4) {
    const TileEntry E = a.tentries[e];
    if (!a.leaf_active[E.leaf]) continue;
    if (!((E.mask_b >> j) & 1u)) continue;
    const double *S = a.kloc + E.kl_off;
    const bool sym = E.mp <= sym_max_mp;
    const int lj = E.start_b + __popc(E.mask_b & ((1u << j) - 1u));
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int i = i0 + 2 * k;
      if ((E.mask_a >> i) & 1u) {
        const int li = E.start_a + __popc(E.mask_a & ((1u << i) - 1u));
        acc[k] += kloc_pair(S, E.mp, sym, li, lj);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) s_acc[wave][(i0 + 2 * k) * 32 + j] = acc[k];
  __syncthreads();
  double *out = a.tpartial + (long long)blockIdx.x * 1024;
  for (int u = t; u < 1024; u += 256)
    out[u] = ((s_acc[0][u] + s_acc[1][u]) + s_acc[2][u]) + s_acc[3][u];
}

// Pass 2b: one workgroup per tile pair sums its chunks' partial blocks in
// chunk order and writes V(i, j) and V(j, i).
__global__ __launch_bounds__(256) void k_vxc_reduce_final(BuildArgs a, double *__restrict__ V) {
  const TilePairDev P = a.tpairs[blockIdx.x];
  const MolDev MD = a.mols[P.mol];
  for (int u = threadIdx.x; u < 1024; u += 256) {
    double v = 0.0;
    for (int q = 0; q < P.nchunk; ++q) v += a.tpartial[(long long)(P.chunk_off + q) * 1024 + u];
    const long long gi = 32LL * P.ti + (u >> 5), gj = 32LL * P.tj + (u & 31);
    if (gi >= MD.nbf || gj >= MD.nbf) continue;
    V[MD.moff + gi + gj * MD.nbf] = v;
    if (P.ti != P.tj) V[MD.moff + gj + gi * MD.nbf] = v;
  }
}

// ------------------------------------------------------------ tiled path --
// Work items are (leaf, AO block) tiles (kernels.hpp: DenItem, VxcItem), so a
// leaf wider than one block is spread over many workgroups and each
// wavefront works on a 2D block of up to 4 x 4 16x16 MFMA tiles (FP64
// arithmetic; S is the AO storage type).

// Vxc scatter screening: a block entry below kScatterCut in magnitude is not
// added into K.  An entry of Vxc = K + K^T gathers at most two terms per
// leaf, so the change of any Vxc entry is below 2 kScatterCut times the
// number of leaves (under 2e-9 for 1e5 leaves, against the 1e-7 tolerance);
// Exc and N_elec are not affected.  About half the block entries are that small (AO pairs
// whose product is negligible over the leaf), and the FP64 atomics they save
// are the largest part of the Vxc stage.
constexpr double kScatterCut = 1e-14;
// dS entries (run_storage_overlap) below this are not added: the N_elec
// correction sum D2 dS changes by at most 1e-18 sum |D2| over the dropped
// entries (under 2e-11 for 3072 AOs and |D2| <= 2).
constexpr double kStorageCut = 1e-18;

constexpr int kTKc = 16;         // AOs per density k-chunk
// LDS row stride (doubles) of the [row][k] density operands: 36 dwords, so
// the 16 rows of an MFMA operand read start in 16 different 4-bank groups
constexpr int kTKld = kTKc + 2;
constexpr int kTVp = 16;         // points per Vxc chunk
constexpr int kTVld = kTVp + 2;  // LDS row stride (doubles) of the [row][point] Vxc operands

// Tile shapes of the MFMA k-steps.  A wavefront owns RT x CT 16x16 tiles
// (RT, CT <= 4).  RT and CT are compile-time constants, so a step's LDS
// operand reads are issued together ahead of its MFMAs (with run-time tile
// counts every MFMA sits in its own branch behind a wait for its operand).
// The shapes items mostly have (blocks of 96 or 128 rows: 3 or 4 tiles each
// way) get their own code; the others run the 4 x 4 form, whose extra tiles
// read LDS rows that are in the arrays (128 rows) but not staged for the
// item: those tiles' accumulators are never used, and a tile (i, j) only
// ever combines row tile i with column tile j.
template <int N> using IntC = std::integral_constant<int, N>;

template <typename F>
__device__ __forceinline__ void tile_shape(int rt, int ct, F &&f) {
  if (rt == 4 && ct == 3) f(IntC<4>{}, IntC<3>{});
  else if (rt == 3 && ct == 4) f(IntC<3>{}, IntC<4>{});
  else if (rt == 3 && ct == 3) f(IntC<3>{}, IntC<3>{});
  else f(IntC<4>{}, IntC<4>{});
}
// The same with the 2-tile shapes too (blocks of 64 AOs or points), for the
// kernels of the default path: without them a 64-wide block runs as 4 tiles,
// 2.6 % (density) and 8.5 % (Vxc) of the MFMA work on (H2O)128 / def2-SVP.
// Those kernels pass each wavefront's live tiles (live_tiles), and 0 runs
// no MFMA.

// 16-row tiles of a wavefront's half block (half rows) that hold any of the
// `valid` rows counted from its first row: the tiles past a leaf's real AOs
// or points hold padding only (zero AO rows, zero-weight points) and are not
// computed.  One tile runs as the 2-tile shape (the shapes start at 2).
template <typename F>
__device__ __forceinline__ void tile_shape_ext(int rt, int ct, F &&f) {
  if (rt == 0 || ct == 0) f(IntC<0>{}, IntC<0>{});  // no tile: loads and barriers only
  else if (rt == 4 && ct == 3) f(IntC<4>{}, IntC<3>{});
  else if (rt == 3 && ct == 4) f(IntC<3>{}, IntC<4>{});
  else if (rt == 3 && ct == 3) f(IntC<3>{}, IntC<3>{});
  else if (rt == 4 && ct == 2) f(IntC<4>{}, IntC<2>{});
  else if (rt == 2 && ct == 4) f(IntC<2>{}, IntC<4>{});
  else if (rt == 3 && ct == 2) f(IntC<3>{}, IntC<2>{});
  else if (rt == 2 && ct == 3) f(IntC<2>{}, IntC<3>{});
  else if (rt == 2 && ct == 2) f(IntC<2>{}, IntC<2>{});
  else f(IntC<4>{}, IntC<4>{});
}

__device__ __forceinline__ int live_tiles(int valid, int half) {
  const int t = (min(max(valid, 0), half) + 15) >> 4;
  return t == 1 ? 2 : t;
}

// Two FP64 k-steps of acc[i][j] += X_i^T Y_j.  x and y point at this lane's
// pair of operands in row tile 0 and column tile 0 ([row][k] layout, LD
// doubles per row: tile i starts 16 i rows further).
template <int R, int C, int LD>
__device__ __forceinline__ void mfma_f64_k2(d4 (&acc)[4][4], const double *x, const double *y) {
  d2u xv[R], yv[C];
#pragma unroll
  for (int i = 0; i < R; ++i) xv[i] = *reinterpret_cast<const d2u *>(x + 16 * i * LD);
#pragma unroll
  for (int j = 0; j < C; ++j) yv[j] = *reinterpret_cast<const d2u *>(y + 16 * j * LD);
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < C; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[i][s2], yv[j][s2], acc[i][j], 0, 0, 0);
}

// Four FP32 k-steps of the same product (v_mfma_f32_16x16x4f32).
template <int R, int C, int LD>
__device__ __forceinline__ void mfma_f32_k4(f4 (&acc)[4][4], const float *x, const float *y) {
  f4 xv[R], yv[C];
#pragma unroll
  for (int i = 0; i < R; ++i) xv[i] = *reinterpret_cast<const f4 *>(x + 16 * i * LD);
#pragma unroll
  for (int j = 0; j < C; ++j) yv[j] = *reinterpret_cast<const f4 *>(y + 16 * j * LD);
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < C; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv[i][s4], yv[j][s4], acc[i][j], 0, 0, 0);
}

// The epilogue of a density item: from acc[i][j][r] = T(point wp*hp + 16 i +
// 4 g + r, column wc*hc + 16 j + l16), the item's partials of rho and
// grad rho (Phi and grad Phi of its columns read from HBM), then the
// functional (the leaf's only item) or the stored partials.  s_phi (at least
// 4 kMaxLeafPts + 4 kTileAo doubles of LDS) holds the sums.
// The epilogue's first loads (row tile 0's first NJ column tiles: all four
// with FP32 storage, two with FP64) can be issued before it
// (den_epi_first_loads, during the last k-chunk's MFMAs) and passed in; PRE
// says they were.
template <typename S> struct DenEpiFirst {
  using V = typename Vec4Of<S>::type;
  static constexpr int NJ = sizeof(S) == 4 ? 4 : 2;  // column tiles per load round
  V f[NJ], fx[NJ], fy[NJ], fz[NJ];
};

// zero is 0, made opaque by the caller inside its last-chunk branch, so the
// loads' addresses are formed there rather than hoisted out of the k-loop
// (held across the loop, they spilled)
template <typename S, bool GGA>
__device__ __forceinline__ void den_epi_first_loads(const S *__restrict__ ao, const DenItem &it,
                                                    const LeafDev &L, DenEpiFirst<S> &pre,
                                                    int zero) {
  using V = typename Vec4Of<S>::type;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wp = wave & 1, wc = wave >> 1;
  const int hp = it.np >> 1, hc = it.wc >> 1;
  const int CT = hc >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int pl = wp * hp + 4 * g + zero;
#pragma unroll
  for (int j = 0; j < DenEpiFirst<S>::NJ; ++j) {
    if (j >= CT) break;
    const long long off = (long long)(it.c0 + wc * hc + 16 * j + l16) * L.np + it.p0 + pl;
    pre.f[j] = *reinterpret_cast<const V *>(phi + off);
    if constexpr (GGA) {
      pre.fx[j] = *reinterpret_cast<const V *>(phi + cs + off);
      pre.fy[j] = *reinterpret_cast<const V *>(phi + 2 * cs + off);
      pre.fz[j] = *reinterpret_cast<const V *>(phi + 3 * cs + off);
    }
  }
}

template <typename S, bool GGA, bool PRE = false>
__device__ __forceinline__ void den_tile_epilogue(const BuildArgs &a, const S *__restrict__ ao,
                                                  const DenItem &it, const LeafDev &L,
                                                  d4 (&acc)[4][4], double *__restrict__ part,
                                                  double *s_phi, double *s_red,
                                                  const DenEpiFirst<S> *pre = nullptr) {
  double(*s_q)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(s_phi);
  double(*s_h)[kTileAo] = reinterpret_cast<double(*)[kTileAo]>(s_phi + 4 * kMaxLeafPts);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wp = wave & 1, wc = wave >> 1;
  const int hp = it.np >> 1, hc = it.wc >> 1;
  const int RT = hp >> 4, CT = hc >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  __syncthreads();  // s_phi is reused for the sums below
  // acc[i][j][r] = T(point wp*hp + 16 i + 4 g + r, column wc*hc + 16 j + l16)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= RT) break;
    const int pl = wp * hp + 16 * i + 4 * g;  // item-local point of r = 0
    double q[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int k = 0; k < 4; ++k) q[r][k] = 0.0;
    // NJ column tiles' loads in flight before any is used, kept as stored
    // (64 registers of loads: all four tiles with FP32 storage, two with FP64)
    using V = typename Vec4Of<S>::type;
    constexpr int NJ = sizeof(S) == 4 ? 4 : 2;
#pragma unroll
    for (int jj = 0; jj < 4; jj += NJ) {
      V f[NJ], fx[NJ], fy[NJ], fz[NJ];
#pragma unroll
      for (int jq = 0; jq < NJ; ++jq) {
        if (jj + jq >= CT) break;
        if (PRE && i == 0 && jj == 0) {
          f[jq] = pre->f[jq];
          if constexpr (GGA) {
            fx[jq] = pre->fx[jq];
            fy[jq] = pre->fy[jq];
            fz[jq] = pre->fz[jq];
          }
          continue;
        }
        const long long off = (long long)(it.c0 + wc * hc + 16 * (jj + jq) + l16) * L.np + it.p0 + pl;
        f[jq] = *reinterpret_cast<const V *>(phi + off);
        if constexpr (GGA) {
          fx[jq] = *reinterpret_cast<const V *>(phi + cs + off);
          fy[jq] = *reinterpret_cast<const V *>(phi + 2 * cs + off);
          fz[jq] = *reinterpret_cast<const V *>(phi + 3 * cs + off);
        }
      }
#pragma unroll
      for (int jq = 0; jq < NJ; ++jq) {
        const int j = jj + jq;
        if (j >= CT) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r][0] += (double)f[jq][r] * acc[i][j][r];
        if constexpr (GGA) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double tt = 2.0 * acc[i][j][r];
            q[r][1] += (double)fx[jq][r] * tt;
            q[r][2] += (double)fy[jq][r] * tt;
            q[r][3] += (double)fz[jq][r] * tt;
          }
        }
      }
    }
    // the column sums: lane l16 of a row ends with point pl + l16 / 4's
    // quantity l16 % 4 (GGA) or point pl + l16 / 4's rho (LDA)
    if constexpr (GGA) {
      double v16[16];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) v16[4 * r + k] = q[r][k];
      const double v = row_reduce16(v16, l16);
      if (wc == 0) s_q[l16 & 3][pl + (l16 >> 2)] = v;
      else s_h[l16 & 3][pl + (l16 >> 2)] = v;
    } else {
      double v4[4] = {q[0][0], q[1][0], q[2][0], q[3][0]};
      const double v = row_reduce4(v4, l16);
      if ((l16 & 3) == 0) {
        if (wc == 0) s_q[0][pl + (l16 >> 2)] = v;
        else s_h[0][pl + (l16 >> 2)] = v;
      }
    }
  }
  __syncthreads();
  const int nq = GGA ? 4 : 1;
  if (it.part < 0) {
    // the leaf's only item: its points are the leaf's points
    for (int u = t; u < nq * it.np; u += 256) {
      const int k = u / it.np, p = u - k * it.np;
      s_q[k][p] += s_h[k][p];
    }
    if (!GGA)
      for (int p = t; p < it.np; p += 256) s_q[1][p] = s_q[2][p] = s_q[3][p] = 0.0;
    __syncthreads();
    density_epilogue(a, L, it.leaf, s_q, s_red);
  } else {
    double *dst = part + (long long)it.part * 4 * kTileAo;
    for (int u = t; u < nq * it.np; u += 256) {
      const int k = u / it.np, p = u - k * it.np;
      dst[k * kTileAo + p] = s_q[k][p] + s_h[k][p];
    }
  }
}

// Density item: T(p, nu) = sum_mu Phi(p, mu) D2(mu, nu) for the item's points p
// and columns nu, over all of the leaf's AOs mu in k-chunks staged through LDS
// (Phi transposed to [point][mu], D2 gathered through the AO map to
// [nu][mu]); wavefront (wp, wc) owns half the points and half the columns.
// The loads of chunk k+1 are issued before the MFMAs of chunk k.  Then the
// column-block partials of rho = sum_nu Phi T and
// grad rho = 2 sum_nu grad Phi T (density.cpp:32-43); the leaf's only item
// runs the functional epilogue at once, the others store their partials.
// One density item (the body of k_den_tile): s_phi, s_d, s_map and s_red
// are the workgroup's LDS arrays.
template <typename S, bool GGA>
__device__ __forceinline__ void den_tile_item(const BuildArgs &a, const S *__restrict__ ao,
                                              const DenItem &it, double *__restrict__ part,
                                              double *s_phi, double *s_d, int *s_map,
                                              double *s_red) {
  const LeafDev L = a.leaves[it.leaf];
  const MolDev MD = a.mols[L.mol];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wp = wave & 1, wc = wave >> 1;
  const int hp = it.np >> 1, hc = it.wc >> 1;  // points / columns per wavefront
  const int RT = hp >> 4, CT = hc >> 4;        // 16x16 tiles per wavefront
  // f64 MFMA: feeding point 4(i%4) + i/4 as row i gives lane group g the
  // accumulator rows of points 4g .. 4g+3 (one vector load in the epilogue)
  const int pa = 4 * (l16 & 3) + (l16 >> 2);
  const S *phi = ao + L.ao_off;
  const double *Dm = a.D2 + MD.moff;
  {
    const int *map = a.ao_map + L.map_off;
    for (int i = t; i < L.mp; i += 256) s_map[i] = map[i];
  }
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  // per chunk: kTKc x np Phi values in groups of 4 points, wc x kTKc D values
  // in groups of 4 AOs; at most 512 groups of each, two per thread
  const int nphi = kTKc * (it.np >> 2), nd = it.wc * (kTKc / 4);
  d4 rp[2], rd[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int gi = t + 256 * u;
      if (gi < nphi) {
        const int mu = gi & (kTKc - 1), p = 4 * (gi / kTKc);
        rp[u] = load4_as<double>(phi + (long long)(k0 + mu) * L.np + it.p0 + p);
      }
      if (gi < nd) {
        const int col = gi >> 2, kk = 4 * (gi & 3);
        const double *dc = Dm + (long long)s_map[it.c0 + col] * MD.nbf;
        rd[u] = d4{dc[s_map[k0 + kk]], dc[s_map[k0 + kk + 1]], dc[s_map[k0 + kk + 2]],
                   dc[s_map[k0 + kk + 3]]};
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int gi = t + 256 * u;
      if (gi < nphi) {
        // AO fastest over the lanes: the transposed writes of 16 lanes fall in
        // consecutive LDS words
        const int mu = gi & (kTKc - 1), p = 4 * (gi / kTKc);
#pragma unroll
        for (int e = 0; e < 4; ++e) s_phi[(p + e) * kTKld + mu] = rp[u][e];
      }
      if (gi < nd) {
        const int col = gi >> 2, kk = 4 * (gi & 3);
        *reinterpret_cast<d4u *>(s_d + col * kTKld + kk) = rd[u];
      }
    }
  };
  __syncthreads();  // s_map
  fetch(0);
  const double *xa = s_phi + (wp * hp + pa) * kTKld + 4 * g;
  const double *yb = s_d + (wc * hc + l16) * kTKld + 4 * g;
  // one k loop per tile shape (a shape test inside the loop puts every
  // shape's accumulators through the merge points and spills)
  tile_shape(RT, CT, [&](auto R, auto C) __attribute__((always_inline)) {
    for (int k0 = 0; k0 < L.mp; k0 += kTKc) {
      __syncthreads();  // the previous chunk's operands are consumed
      store();
      __syncthreads();
      if (k0 + kTKc < L.mp) fetch(k0 + kTKc);
#pragma unroll
      for (int sp = 0; sp < kTKc / 4; sp += 2)
        mfma_f64_k2<decltype(R)::value, decltype(C)::value, kTKld>(
            acc, xa + sp, yb + sp);
    }
  });
  den_tile_epilogue<S, GGA>(a, ao, it, L, acc, part, s_phi, s_red);
}

// Density item with the operands staged by direct-to-LDS loads
// (global_load_lds: each lane's bytes land at the wavefront's LDS base +
// lane x size, with no registers involved) into NS LDS chunk buffers, NS - 1
// chunks ahead of the MFMAs: FP32 storage keeps three buffers (two chunks in
// flight), FP64 storage two.  A buffer holds the chunk's Phi as stored
// ([16 AOs][128 points], S), then D2 gathered through the AO map ([16 AOs]
// [128 columns], FP64, each double as two 4-byte pieces from two lanes).
// Every wavefront issues the same loads for every chunk (lanes past the
// item's points or columns load a valid duplicate into LDS slots that only
// the unused tiles read), so a chunk's loads are waited for by count.  The
// MFMA k-step t reads AO 4 g + t of both, so the accumulators are those of
// den_tile_item and the epilogue is shared.
template <typename S> struct DenDma {
  static constexpr int NS = sizeof(S) == 8 ? 2 : 3;               // LDS chunk buffers
  static constexpr int PHI = kTKc * kTileAo * (int)sizeof(S);     // bytes of Phi per buffer
  static constexpr int STAGE = PHI + kTKc * kTileAo * 8;          // bytes per buffer
  static constexpr int LOADS = (sizeof(S) == 8 ? 4 : 2) + 16;     // loads per chunk per wavefront
};
static_assert(4 * kMaxLeafPts + 4 * kTileAo <= DenDma<float>::STAGE / 8, "epilogue sums fit a stage");

// Barrier after this wavefront's vector memory loads but the newest N are
// done (and its LDS accesses): the s_waitcnt immediate holds vmcnt in bits
// 3:0 and 15:14, expcnt (no wait) in 6:4 and lgkmcnt (0) in 11:8.
template <int N>
__device__ __forceinline__ void vmcnt_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt is six bits");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | ((N >> 4) << 14));
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <typename S, bool GGA>
__device__ __forceinline__ void den_tile_item_dma(const BuildArgs &a, const S *__restrict__ ao,
                                                  const int *__restrict__ ao_map,
                                                  const DenItem &it, double *__restrict__ part,
                                                  char *s_stage, double *s_red) {
  using P = DenDma<S>;
  const LeafDev L = a.leaves[it.leaf];
  const MolDev MD = a.mols[L.mol];
  const int t = threadIdx.x, lane = t & 63;
  // the wavefront index as a scalar: the AO rows of a load and their map
  // entries are then scalar values (scalar loads, counted apart from the
  // vector memory loads)
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l16 = lane & 15, g = lane >> 4;
  const int wp = wave & 1, wc = wave >> 1;
  const int hp = it.np >> 1, hc = it.wc >> 1;
  const int pa = 4 * (l16 & 3) + (l16 >> 2);
  const S *phi = ao + L.ao_off;
  const double *Dm = a.D2 + MD.moff;
  // the AO map through a restrict kernel argument: uniform reads of it are
  // scalar loads (no wait on the vector memory counter, which counts the
  // chunks' loads)
  const int *map = ao_map + L.map_off;
  // D2 gather: load q (0..15) of wave w fetches AO row 4 w + q / 4, columns
  // 32 (q % 4) .. + 31; lane l takes column 32 (q % 4) + l / 2 (clamped to
  // the item's last column), half l % 2.  D2 is symmetric (k_prepare_density
  // mirrors one triangle), so the row's entries are read along D2's
  // contiguous index: consecutive lanes, consecutive addresses wherever the
  // map is contiguous
  long long colbase[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) colbase[cb] = map[it.c0 + min(32 * cb + (lane >> 1), it.wc - 1)];
  // Phi rows: 2 (FP64) or 4 (FP32) points per lane, clamped to the item's last ones
  constexpr int PPL = 16 / (int)sizeof(S);
  const int prow = sizeof(S) == 8 ? PPL * lane : PPL * (lane & 31);
  const S *phi_pt = phi + it.p0 + min(prow, it.np - PPL);
  auto dma = [&](int k0, char *st) __attribute__((always_inline)) {
    long long kmap[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) kmap[r] = (long long)map[k0 + 4 * wave + r] * MD.nbf;
    if constexpr (sizeof(S) == 8) {
      // a 128-point row (1 KB) per load; wave w takes rows 4 w .. 4 w + 3
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int kk = 4 * wave + r;
        __builtin_amdgcn_global_load_lds(phi_pt + (long long)(k0 + kk) * L.np,
                                         st + kk * kTileAo * 8, 16, 0, 0);
      }
    } else {
      // two 128-point rows (512 B each) per load: lanes 32-63 the second
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int kk = 4 * wave + 2 * r + (lane >> 5);
        __builtin_amdgcn_global_load_lds(phi_pt + (long long)(k0 + kk) * L.np,
                                         st + (4 * wave + 2 * r) * kTileAo * 4, 16, 0, 0);
      }
    }
    char *sd = st + P::PHI;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int kk = 4 * wave + (q >> 2), cb = q & 3;
      __builtin_amdgcn_global_load_lds(
          reinterpret_cast<const char *>(Dm + colbase[cb] + kmap[q >> 2]) + 4 * (lane & 1),
          sd + (kk * kTileAo + 32 * cb) * 8, 4, 0, 0);
    }
  };
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  // AO chunks up to the leaf's real AOs (the padding rows of Phi are zero);
  // the wavefront's tiles that hold real points and real AO columns
  const int nch = (L.m + kTKc - 1) / kTKc;
  const int RTl = live_tiles(L.n - (it.p0 + wp * hp), hp), CTl = live_tiles(L.m - (it.c0 + wc * hc), hc);
  DenEpiFirst<S> pre;
  for (int c = 0; c < P::NS - 1 && c < nch; ++c) dma(c * kTKc, s_stage + c * P::STAGE);
  tile_shape_ext(RTl, CTl, [&](auto R, auto C) __attribute__((always_inline)) {
    constexpr int NR = decltype(R)::value, NC = decltype(C)::value;
    int cur = 0;  // buffer of chunk c
    for (int c = 0; c < nch; ++c) {
      // chunk c's loads are complete in every wavefront (the newer chunks'
      // may still run) and every wavefront is done with chunk c - 1, whose
      // buffer the loads of chunk c + NS - 1 overwrite
      if (P::NS == 3 && c + 1 < nch) vmcnt_barrier<P::LOADS>();
      else vmcnt_barrier<0>();
      const int nxt = cur == 0 ? P::NS - 1 : cur - 1;  // (c + NS - 1) % NS
      if (c + P::NS - 1 < nch) dma((c + P::NS - 1) * kTKc, s_stage + nxt * P::STAGE);
      // the epilogue's first loads under the last chunk's MFMAs
      if (c == nch - 1) {
        int zero = 0;
        asm volatile("" : "+s"(zero));  // an opaque 0: no instruction, no memory access
        den_epi_first_loads<S, GGA>(ao, it, L, pre, zero);
      }
      const char *buf = s_stage + cur * P::STAGE;
      const S *sp = reinterpret_cast<const S *>(buf);
      const double *sd = reinterpret_cast<const double *>(buf + P::PHI);
      if constexpr (NR > 0)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int kk = 4 * g + ks;
        double xv[NR], yv[NC];
#pragma unroll
        for (int i = 0; i < NR; ++i) xv[i] = (double)sp[kk * kTileAo + wp * hp + 16 * i + pa];
#pragma unroll
        for (int j = 0; j < NC; ++j) yv[j] = sd[kk * kTileAo + wc * hc + 16 * j + l16];
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
          for (int j = 0; j < NC; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[i], yv[j], acc[i][j], 0, 0, 0);
      }
      cur = cur + 1 == P::NS ? 0 : cur + 1;
    }
  });
  den_tile_epilogue<S, GGA, true>(a, ao, it, L, acc, part, reinterpret_cast<double *>(s_stage),
                                  s_red, &pre);
}

template <typename S, bool GGA>
__global__ __launch_bounds__(256, 2) void k_den_tile_dma(BuildArgs a, const S *__restrict__ ao,
                                                         const int *__restrict__ ao_map,
                                                         const DenItem *__restrict__ items,
                                                         double *__restrict__ part) {
  // one LDS array (the reduction slots after the chunk buffers): with a
  // second __shared__ variable the LDS accesses carry alias scopes, and the
  // compiler then waits for every direct load in flight before each chunk's
  // LDS reads, which serialises the load pipeline
  __shared__ __align__(16) char s_stage[DenDma<S>::NS * DenDma<S>::STAGE + 4 * sizeof(double)];
  double *s_red = reinterpret_cast<double *>(s_stage + DenDma<S>::NS * DenDma<S>::STAGE);
  const DenItem it = items[blockIdx.x];
  if (it.leaf < 0) return;
  den_tile_item_dma<S, GGA>(a, ao, ao_map, it, part, s_stage, s_red);
}

// PERSIST: the workgroups loop over the items blockIdx.x, blockIdx.x +
// gridDim.x, ... (gridDim.x a multiple of 8, so every item of a workgroup is
// from the same XCD's run of leaves) instead of one item each.
template <typename S, bool GGA, bool PERSIST = false>
__global__ __launch_bounds__(256, 2) void k_den_tile(BuildArgs a, const S *__restrict__ ao,
                                                     const DenItem *__restrict__ items,
                                                     double *__restrict__ part, int n_items) {
  __shared__ double s_phi[kTileAo * kTKld];  // [point][k]; after the k loop s_q and s_h
  __shared__ double s_d[kTileAo * kTKld];    // [column][k]
  __shared__ int s_map[kMaxLocalAo];
  __shared__ double s_red[4];
  static_assert(4 * kMaxLeafPts + 4 * kTileAo <= kTileAo * kTKld, "epilogue sums fit s_phi");
  if constexpr (PERSIST) {
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
      const DenItem it = items[item];
      if (it.leaf < 0) continue;
      den_tile_item<S, GGA>(a, ao, it, part, s_phi, s_d, s_map, s_red);
      __syncthreads();  // this item's LDS reads are done before the next one's writes
    }
  } else {
    const DenItem it = items[blockIdx.x];
    if (it.leaf < 0) return;
    den_tile_item<S, GGA>(a, ao, it, part, s_phi, s_d, s_map, s_red);
  }
}

// Leaves with more than one density item: their partials summed in
// (point block, column block) order, then the functional epilogue.
__global__ __launch_bounds__(256) void k_den_finish(BuildArgs a, const DenFinish *__restrict__ fin,
                                                     const double *__restrict__ part) {
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  const DenFinish F = fin[blockIdx.x];
  const LeafDev L = a.leaves[F.leaf];
  const int nq = a.deriv ? 4 : 1;
  for (int u = threadIdx.x; u < 4 * L.np; u += blockDim.x) {
    const int k = u / L.np, p = u - k * L.np;
    double v = 0.0;
    if (k < nq) {
      const int pb = p / kTileAo, pl = p - pb * kTileAo;
      const double *src = part + (long long)(F.part0 + pb * F.ncb) * 4 * kTileAo + k * kTileAo + pl;
      for (int cb = 0; cb < F.ncb; ++cb) v += src[(long long)cb * 4 * kTileAo];
    }
    s_q[k][p] = v;
  }
  __syncthreads();
  density_epilogue(a, L, F.leaf, s_q, s_red);
}

// Scatter of a Vxc item's block into K with FP64 atomics (entries below cut
// dropped).  acc[i][j][r] of wavefront w = (wr, wcw) = (w % 2, w / 2) holds
// the block entry (row, col) = (wr hr + lr, wcw hc + 16 j + l16), lr = 16 i +
// 4 g + r (F32: the v_mfma_f32 accumulator rows) or 16 i + g + 4 r (FP64).
// Without FOLD, or for an off-diagonal item, every entry is added.  With FOLD
// a diagonal item (r0 == c0) adds only the entries row <= col: B(row, col) +
// B(col, row) above the diagonal and B(row, row) on it; Vxc = K + K^T is
// unchanged (the off-diagonal items also add above K's diagonal only: the AO
// map is ascending) and the item's atomics halve.  The transposes pass
// through s_t (2 x 64 x 64 doubles, the chunk buffers, free after the first
// barrier): wavefront 1 (the lower-left quadrant) hands its entries to
// wavefront 2 (upper-right), and the two diagonal quadrants (wavefronts 0 and
// 3) each transpose their own.  The fold pays where the atomics are exposed
// (k_vxc_gemm_f32: 2.5 ms of 23.8 on (H2O)128 / def2-SVP); in the
// register-staged kernels they drain behind other workgroups' staging and
// the fold's barriers cost more than it saves.
template <bool F32, bool FOLD>
__device__ __forceinline__ void vxc_scatter(const d4 (&acc)[4][4], const VxcItem &it,
                                            const LeafDev &L, const MolDev &MD,
                                            const int *__restrict__ map, double *Km, double cut,
                                            double *s_t) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wr = wave & 1, wcw = wave >> 1;
  const int hr = it.wr >> 1, hc = it.wc >> 1;
  const int RT = hr >> 4, CT = hc >> 4;
  auto lrow = [&](int i, int r) { return 16 * i + (F32 ? 4 * g + r : g + 4 * r); };
  if (!FOLD || it.r0 != it.c0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= CT) break;
      const int nu = it.c0 + wcw * hc + 16 * j + l16;
      if (nu >= L.m) continue;
      double *kc = Km + map[nu];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= RT) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = it.r0 + wr * hr + lrow(i, r);
          if (mu < L.m && fabs(acc[i][j][r]) >= cut)
            atomicAdd(kc + (long long)map[mu] * MD.nbf, acc[i][j][r]);
        }
      }
    }
    return;
  }
  // K offsets of the entries' rows and columns, loaded once (-1: past the
  // leaf's AOs)
  long long ro[4][4];
  int co[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = it.r0 + wr * hr + lrow(i, r);
      ro[i][r] = i < RT && mu < L.m ? (long long)map[mu] * MD.nbf : -1;
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int nu = it.c0 + wcw * hc + 16 * j + l16;
    co[j] = j < CT && nu < L.m ? map[nu] : -1;
  }
  auto emit = [&](int i, int j, int r, double v) __attribute__((always_inline)) {
    if (ro[i][r] >= 0 && co[j] >= 0 && fabs(v) >= cut) atomicAdd(Km + ro[i][r] + co[j], v);
  };
  // [row][col ^ (row % 16)] keeps a wavefront's column reads off one bank
  auto tix = [](int row, int col) { return row * 64 + (col ^ (row & 15)); };
  auto put = [&](double *T) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= RT) break;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= CT) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) T[tix(lrow(i, r), 16 * j + l16)] = acc[i][j][r];
      }
    }
  };
  // upper triangle of a diagonal quadrant from its own transpose T
  auto fold_diag = [&](const double *T) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= RT) break;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= CT) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lr = lrow(i, r), lc = 16 * j + l16;
          if (lr < lc) emit(i, j, r, acc[i][j][r] + T[tix(lc, lr)]);
          else if (lr == lc) emit(i, j, r, acc[i][j][r]);
        }
      }
    }
  };
  double *R0 = s_t, *R1 = s_t + 64 * 64;
  __syncthreads();  // every wavefront is done with the chunk buffers
  if (wave == 1) put(R0);
  if (wave == 0) {
    put(R1);
    fold_diag(R1);
  }
  __syncthreads();
  if (wave == 2) {
    // upper-right quadrant entry (lr, lc) + lower-left entry (lc, lr)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= RT) break;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= CT) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) emit(i, j, r, acc[i][j][r] + R0[tix(16 * j + l16, lrow(i, r))]);
      }
    }
  }
  if (wave == 3) {
    put(R1);
    fold_diag(R1);
  }
}

// Vxc item, block pair (r, c) of a leaf: the leaf's points are streamed kTVp
// at a time; each chunk stages Phi and A = c0 Phi + c . grad Phi
// (xc_potential_matrix.cpp:19-41) of the block rows in LDS ([row][point]),
// and wavefront (wr, wc) accumulates its quadrant of
//   r == c:  K_rr = Phi_r^T A_r
//   r <  c:  V_rc = Phi_r^T A_c + A_r^T Phi_c
// in registers.  Entry (mu, nu) is added to K(map[nu], map[mu]) with FP64
// atomics; Vxc = K + K^T afterwards counts every entry of V_loc once.
// RES (FP32 storage, context set-up): A = w r with r the value plane's
// rounding residuals, so K + K^T = dS = sum_p w_p (r Phi^T + Phi r^T); only
// the items of the leaves [res_leaf0, res_leaf1) whose residuals a.ao_res
// holds run, and entries below kStorageCut are not added.
template <typename S, bool GGA, bool RES = false>
__global__ __launch_bounds__(256, 2) void k_vxc_tile(BuildArgs a, const S *__restrict__ ao,
                                                     const VxcItem *__restrict__ items) {
  // the four operand buffers in one array: the scatter's transposes reuse it
  __shared__ __align__(16) double s_stage[4 * kTileAo * kTVld];
  static_assert(4 * kTileAo * kTVld >= 2 * 64 * 64, "scatter transposes fit");
  // buffer offsets: Phi rows 0, A rows kB, Phi columns 2 kB, A columns 3 kB
  constexpr int kB = kTileAo * kTVld;
  __shared__ double s_c[2][4][kTVp];
  const VxcItem it = items[blockIdx.x];
  if (it.leaf < 0 || !a.leaf_active[it.leaf]) return;
  if constexpr (RES)
    if (it.leaf < a.res_leaf0 || it.leaf >= a.res_leaf1) return;
  const LeafDev L = a.leaves[it.leaf];
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = a.ao_map + L.map_off;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wr = wave & 1, wcw = wave >> 1;
  const bool diag = it.r0 == it.c0;
  const int hr = it.wr >> 1, hc = it.wc >> 1;
  const int RT = hr >> 4, CT = hc >> 4;
  const S *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int nrow = diag ? it.wr : it.wr + it.wc;
  static_assert(!RES || (!GGA && std::is_same<S, float>::value), "dS: FP32 storage, LDA form");
  const double *cf = RES ? a.pw + L.pt_off : a.coef + L.pt_off;
  constexpr int NQ = GGA ? 4 : 1;
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  if (t < NQ * kTVp) s_c[0][t / kTVp][t % kTVp] = cf[(t / kTVp) * a.ptot + t % kTVp];
  const int xo = (wr * hr + l16) * kTVld + 4 * g, yo = (wcw * hc + l16) * kTVld + 4 * g;
  // one point loop per tile shape
  tile_shape(RT, CT, [&](auto R, auto C) __attribute__((always_inline)) {
  int buf = 0;
  // chunks up to the leaf's real points: A is zero at the zero-weight
  // padding points
  for (int p0 = 0; p0 < L.n; p0 += kTVp, buf ^= 1) {
    __syncthreads();
    // the next chunk's coefficients (read after the barrier below)
    if (p0 + kTVp < L.n && t < NQ * kTVp)
      s_c[buf ^ 1][t / kTVp][t % kTVp] = cf[(t / kTVp) * a.ptot + p0 + kTVp + t % kTVp];
    for (int gi = t; gi < nrow * (kTVp / 4); gi += 256) {
      const int row = gi >> 2, p = 4 * (gi & 3);
      const bool rb = row < it.wr;
      const int mu = rb ? it.r0 + row : it.c0 + row - it.wr;
      const long long o = (long long)mu * L.np + p0 + p;
      const d4 f = load4_as<double>(phi + o);
      d4 A;
      if constexpr (RES) {
        const d4 r = load4_as<double>(a.ao_res + (L.res_off - a.res_base) + o);
#pragma unroll
        for (int e = 0; e < 4; ++e) A[e] = s_c[buf][0][p + e] * r[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) A[e] = s_c[buf][0][p + e] * f[e];
      }
      if constexpr (GGA) {
        const d4 fx = load4_as<double>(phi + cs + o);
        const d4 fy = load4_as<double>(phi + 2 * cs + o);
        const d4 fz = load4_as<double>(phi + 3 * cs + o);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          A[e] += s_c[buf][1][p + e] * fx[e] + s_c[buf][2][p + e] * fy[e] + s_c[buf][3][p + e] * fz[e];
      }
      const int lr = rb ? row : row - it.wr;
      *reinterpret_cast<d4u *>(s_stage + (rb ? 0 : 2 * kB) + lr * kTVld + p) = f;
      *reinterpret_cast<d4u *>(s_stage + (rb ? kB : 3 * kB) + lr * kTVld + p) = A;
    }
    __syncthreads();
    // two k-steps (points) per operand load keeps the operands at 32
    // registers next to the 128 of the accumulators
    auto product = [&](int xa, int yb) __attribute__((always_inline)) {
#pragma unroll
      for (int sp = 0; sp < 4; sp += 2)
        mfma_f64_k2<decltype(R)::value, decltype(C)::value, kTVld>(
            acc, s_stage + xa + xo + sp, s_stage + yb + yo + sp);
    };
    product(0, diag ? kB : 3 * kB);
    if (!diag) product(kB, 2 * kB);
  }
  });
  vxc_scatter<false, false>(acc, it, L, MD, map, Km, RES ? kStorageCut : kScatterCut, s_stage);
}

// FP32 form of the Vxc item (FP32 mode): the same block pairs with Phi and
// A rounded to FP32 in LDS and K_loc accumulated over the item's points on
// v_mfma_f32_16x16x4f32 (accumulator register r of lane l holds row
// 4 (l/16) + r), then added into the FP64 K with FP64 atomics, so the sum
// across leaves stays FP64.  Chunks of 32 points (half the FP64 form's LDS
// per point): half as many barriers and FP64 accumulations per point.
constexpr int kTVpf = 32;          // points per FP32 Vxc chunk
constexpr int kTVldf = kTVpf + 4;  // floats per LDS row (144 bytes)

template <bool GGA>
__global__ __launch_bounds__(256, 2) void k_vxc_tile_f32(BuildArgs a, const float *__restrict__ ao,
                                                         const VxcItem *__restrict__ items) {
  // the four operand buffers in one array: the scatter's transposes reuse it
  __shared__ __align__(16) float s_stage[4 * kTileAo * kTVldf];
  static_assert(4 * kTileAo * kTVldf * 4 >= 2 * 64 * 64 * 8, "scatter transposes fit");
  // buffer offsets: Phi rows 0, A rows kB, Phi columns 2 kB, A columns 3 kB
  constexpr int kB = kTileAo * kTVldf;
  __shared__ double s_c[2][4][kTVpf];
  const VxcItem it = items[blockIdx.x];
  if (it.leaf < 0 || !a.leaf_active[it.leaf]) return;
  const LeafDev L = a.leaves[it.leaf];
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = a.ao_map + L.map_off;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int wr = wave & 1, wcw = wave >> 1;
  const bool diag = it.r0 == it.c0;
  const int hr = it.wr >> 1, hc = it.wc >> 1;
  const int RT = hr >> 4, CT = hc >> 4;
  const float *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int nrow = diag ? it.wr : it.wr + it.wc;
  const double *cf = a.coef + L.pt_off;
  constexpr int NQ = GGA ? 4 : 1;
  // FP32 MFMA accumulation over one 32-point chunk at a time, added into
  // FP64 accumulators after each chunk: the FP32 sums stay 32 terms long
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  if (t < NQ * kTVpf) s_c[0][t / kTVpf][t % kTVpf] = cf[(t / kTVpf) * a.ptot + t % kTVpf];
  const int xo = (wr * hr + l16) * kTVldf + 4 * g, yo = (wcw * hc + l16) * kTVldf + 4 * g;
  // one point loop per tile shape
  tile_shape(RT, CT, [&](auto R, auto C) __attribute__((always_inline)) {
  int buf = 0;
  // chunks up to the leaf's real points (A is zero at the padding points)
  for (int p0 = 0; p0 < L.n; p0 += kTVpf, buf ^= 1) {
    __syncthreads();
    if (p0 + kTVpf < L.n && t < NQ * kTVpf)
      s_c[buf ^ 1][t / kTVpf][t % kTVpf] = cf[(t / kTVpf) * a.ptot + p0 + kTVpf + t % kTVpf];
    for (int gi = t; gi < nrow * (kTVpf / 4); gi += 256) {
      const int row = gi >> 3, p = 4 * (gi & 7);
      const bool rb = row < it.wr;
      const int mu = rb ? it.r0 + row : it.c0 + row - it.wr;
      const long long o = (long long)mu * L.np + p0 + p;
      const f4 f = *reinterpret_cast<const f4 *>(phi + o);
      f4 fx, fy, fz;
      if constexpr (GGA) {
        fx = *reinterpret_cast<const f4 *>(phi + cs + o);
        fy = *reinterpret_cast<const f4 *>(phi + 2 * cs + o);
        fz = *reinterpret_cast<const f4 *>(phi + 3 * cs + o);
      }
      f4 A;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double v = s_c[buf][0][p + e] * (double)f[e];
        if constexpr (GGA)
          v += s_c[buf][1][p + e] * (double)fx[e] + s_c[buf][2][p + e] * (double)fy[e] +
               s_c[buf][3][p + e] * (double)fz[e];
        A[e] = (float)v;
      }
      const int lr = rb ? row : row - it.wr;
      *reinterpret_cast<f4 *>(s_stage + (rb ? 0 : 2 * kB) + lr * kTVldf + p) = f;
      *reinterpret_cast<f4 *>(s_stage + (rb ? kB : 3 * kB) + lr * kTVldf + p) = A;
    }
    __syncthreads();
    f4 acf[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acf[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    auto product = [&](int xa, int yb) __attribute__((always_inline)) {
#pragma unroll
      for (int ks = 0; ks < kTVpf; ks += 16)
        mfma_f32_k4<decltype(R)::value, decltype(C)::value, kTVldf>(
            acf, s_stage + xa + xo + ks, s_stage + yb + yo + ks);
    };
    product(0, diag ? kB : 3 * kB);
    if (!diag) product(kB, 2 * kB);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)acf[i][j][r];
  }
  });
  vxc_scatter<true, false>(acc, it, L, MD, map, Km, kScatterCut, reinterpret_cast<double *>(s_stage));
}

// FP32 mode, first step of the Vxc stage: A = sum_q c_q Phi_q of every active
// leaf (the arithmetic of k_vxc_tile_f32's staging: FP64 sums of the FP32
// values, rounded to FP32) into a.ao_a at the leaf's res_off, for the leaf's
// m AOs and np points.  Rows m..mp-1 are left as they are: the contraction
// reads them, but each of its outputs pairs one row of Phi with one row of
// A, and the outputs of rows past m are never added into K.  One workgroup per leaf; np / 4 lanes per AO row (4 points
// each), so a lane's coefficients are fixed.  HBM-bound: 16 bytes read and 4
// written per element (GGA), both streamed past the caches (non-temporal:
// 0.4 ms of 6 on (H2O)128 / def2-SVP).
template <bool GGA>
__global__ __launch_bounds__(256) void k_vxc_apass(BuildArgs a, const float *__restrict__ ao) {
  const int li = blockIdx.x;
  if (!a.leaf_active[li]) return;
  const LeafDev L = a.leaves[li];
  const int tpr = L.np >> 2, rpp = 256 / tpr;  // lanes per row, rows per pass
  const int t = threadIdx.x, row0 = t / tpr;
  if (row0 >= rpp) return;
  const int p = 4 * (t - row0 * tpr);
  constexpr int NQ = GGA ? 4 : 1;
  double c[NQ][4];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) c[q][e] = a.coef[q * a.ptot + L.pt_off + p + e];
  const float *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  float *out = a.ao_a + L.res_off;
#pragma unroll 2
  for (int mu = row0; mu < L.m; mu += rpp) {
    const long long o = (long long)mu * L.np + p;
    const f4 f = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(phi + o));
    f4 fx, fy, fz;
    if constexpr (GGA) {
      fx = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(phi + cs + o));
      fy = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(phi + 2 * cs + o));
      fz = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(phi + 3 * cs + o));
    }
    f4 A;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double v = c[0][e] * (double)f[e];
      if constexpr (GGA)
        v += c[1][e] * (double)fx[e] + c[2][e] * (double)fy[e] + c[3][e] * (double)fz[e];
      A[e] = (float)v;
    }
    __builtin_nontemporal_store(A, reinterpret_cast<f4 *>(out + o));
  }
}

// FP32 mode, second step: the Vxc item's block pair from the Phi and A planes
// (k_vxc_apass), staged into LDS by direct loads (global_load_lds, 16 bytes
// a lane) in chunks of kVgP points, two chunk buffers, so the next chunk
// loads while the MFMAs run on this one and no registers hold staged values.  A
// buffer holds four operand blocks [Phi rows | A rows | Phi columns | A
// columns] of kTileAo AO rows x kVgP points, 64 bytes a row with no padding:
// the 16-byte column slot of 4-point group q in row R is q ^ ((R / 2) % 4).
// With that swizzle a wavefront's ds_read_b128 of an MFMA operand (lane l:
// row l % 16, group l / 16) runs at the LDS's full rate on gfx950 (measured
// about 4.8 cycles a wavefront; padded rows of 144 or 80 bytes, or other
// swizzles, take about 8.5 and read as bank conflicts).
// The products and the FP64 folding (every 32 points) are those of
// k_vxc_tile_f32, so are the accumulator layout and the scatter.
constexpr int kVgP = 16;                       // points per chunk
constexpr int kVgBlock = kTileAo * kVgP * 4;   // bytes of one operand block
constexpr int kVgStage = 4 * kVgBlock;         // bytes of one chunk buffer

__device__ __forceinline__ int vxc_swz(int r) { return (r >> 1) & 3; }

__global__ __launch_bounds__(256, 2) void k_vxc_gemm_f32(BuildArgs a, const float *__restrict__ ao,
                                                         const int *__restrict__ ao_map,
                                                         const VxcItem *__restrict__ items) {
  constexpr int NS = 2;  // chunk buffers (three, one workgroup per CU, measured slower)
  __shared__ __align__(16) char s_stage[NS * kVgStage];
  static_assert(NS * kVgStage >= 2 * 64 * 64 * (int)sizeof(double), "scatter transposes fit");
  const VxcItem it = items[blockIdx.x];
  if (it.leaf < 0 || !a.leaf_active[it.leaf]) return;
  const LeafDev L = a.leaves[it.leaf];
  const MolDev MD = a.mols[L.mol];
  double *Km = a.K + MD.moff;
  const int *map = ao_map + L.map_off;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l16 = lane & 15, g = lane >> 4;
  const int wr = wave & 1, wcw = wave >> 1;
  const bool diag = it.r0 == it.c0;
  const int hr = it.wr >> 1, hc = it.wc >> 1;
  const float *phi = ao + L.ao_off;
  const float *am = a.ao_a + L.res_off;
  // loads: wavefront w fills 16-row groups w and w + 4 of each block, lane l
  // row 16 group + l / 4 (clamped to the block's last row: rows past it are
  // read only by tiles whose results are dropped) and slot l % 4; the Phi and
  // A planes share the [mp][np] layout, so one offset serves both
  const int lrow = lane >> 2, lcol = 4 * ((lane & 3) ^ vxc_swz(lrow));
  int off[2][2];  // [row block, column block][group]
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int rb = h ? it.c0 : it.r0, wb = h ? it.wc : it.wr;
#pragma unroll
    for (int k = 0; k < 2; ++k) off[h][k] = (rb + min(16 * (wave + 4 * k) + lrow, wb - 1)) * L.np + lcol;
  }
  auto dma = [&](int p0, char *st) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && diag) break;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        char *dst = st + 2 * h * kVgBlock + (wave + 4 * k) * 1024;
        __builtin_amdgcn_global_load_lds(phi + off[h][k] + p0, dst, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(am + off[h][k] + p0, dst + kVgBlock, 16, 0, 0);
      }
    }
  };
  // chunks up to the leaf's real points (A is zero at the padding points);
  // the wavefront's tiles that hold real AO rows and columns
  const int nch = (L.n + kVgP - 1) / kVgP;
  const int RTl = live_tiles(L.m - (it.r0 + wr * hr), hr), CTl = live_tiles(L.m - (it.c0 + wcw * hc), hc);
  if (nch > 0) dma(0, s_stage);
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  const int sw = 4 * (g ^ vxc_swz(l16));
  const int xo = (wr * hr + l16) * kVgP + sw, yo = (wcw * hc + l16) * kVgP + sw;
  tile_shape_ext(RTl, CTl, [&](auto R, auto C) __attribute__((always_inline)) {
    constexpr int NR = decltype(R)::value, NC = decltype(C)::value;
    f4 acf[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acf[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    int cur = 0;
    for (int c = 0; c < nch; ++c) {
      // chunk c's loads are complete in every wavefront and every wavefront
      // is done with chunk c - 1, whose buffer chunk c + 1 fills
      vmcnt_barrier<0>();
      if (c + 1 < nch) dma((c + 1) * kVgP, s_stage + (cur ^ 1) * kVgStage);
      const float *pr = reinterpret_cast<const float *>(s_stage + cur * kVgStage);
      const float *ar = pr + kTileAo * kVgP, *pc = ar + kTileAo * kVgP, *ac = pc + kTileAo * kVgP;
      if constexpr (NR > 0) {
        mfma_f32_k4<NR, NC, kVgP>(acf, pr + xo, (diag ? ar : ac) + yo);
        if (!diag) mfma_f32_k4<NR, NC, kVgP>(acf, ar + xo, pc + yo);
      }
      if ((c & 1) || c + 1 == nch) {
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
          for (int j = 0; j < NC; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              acc[i][j][r] += (double)acf[i][j][r];
              acf[i][j][r] = 0.f;
            }
      }
      cur ^= 1;
    }
  });
  vxc_scatter<true, true>(acc, it, L, MD, map, Km, kScatterCut, reinterpret_cast<double *>(s_stage));
}

// Persistent, wavefront-specialised form of the Vxc items: one 512-thread
// workgroup per CU walks the items blockIdx.x, blockIdx.x + gridDim.x, ...
// (the same XCD for all of them, gridDim.x being a multiple of 8) chunk by
// chunk.  Wavefronts 4-7 stage the next chunk (Phi and A of the item's block
// rows) into one of two LDS buffers while wavefronts 0-3 run the MFMAs of the
// current chunk from the other.  After an item's last chunk the MFMA
// wavefronts issue its scatter atomics and go on with the next item while
// they drain.  Skipped leaves and padding entries take no steps.
struct VxcCursor {
  int ii, p0;
  VxcItem it;
  LeafDev L;
};

__device__ __forceinline__ void vxc_seek(const BuildArgs &a, const VxcItem *items, int n,
                                         VxcCursor &c) {
  while (c.ii < n) {
    c.it = items[c.ii];
    if (c.it.leaf >= 0 && a.leaf_active[c.it.leaf]) {
      c.L = a.leaves[c.it.leaf];
      return;
    }
    c.ii += gridDim.x;
  }
}

__device__ __forceinline__ void vxc_next(const BuildArgs &a, const VxcItem *items, int n,
                                         VxcCursor &c) {
  c.p0 += kTVp;
  if (c.p0 >= c.L.np) {
    c.p0 = 0;
    c.ii += gridDim.x;
    vxc_seek(a, items, n, c);
  }
}

constexpr int kWsBuf = 4 * kTileAo * kTVld;  // doubles per staging buffer

// Workgroup barrier for LDS hand-offs only: waits for this wavefront's LDS
// accesses (lgkmcnt) but not for its outstanding global memory operations,
// so the MFMA wavefronts' scatter atomics drain while the next steps run
// (__syncthreads would wait for them at every step).  The empty asm
// statements keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt and expcnt: no wait; lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
static size_t vxc_ws_smem() { return sizeof(double) * 2 * kWsBuf; }

template <typename S, bool GGA>
__global__ __launch_bounds__(512, 1) void k_vxc_ws(BuildArgs a, const S *__restrict__ ao,
                                                   const VxcItem *__restrict__ items, int n) {
  extern __shared__ double smem[];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const bool producer = wave >= 4;
  VxcCursor cc;
  cc.ii = blockIdx.x;
  cc.p0 = 0;
  vxc_seek(a, items, n, cc);
  if (cc.ii >= n) return;
  // producers: Phi and A = c0 Phi + c . grad Phi of a step's block rows.  A
  // producer thread takes the groups pt, pt + 256, ... (at most 4: 2 x 128
  // rows of 4 groups), which share one 4-point group, so its coefficients are
  // loaded once, and all its AO loads are issued before any is used.
  auto stage = [&](const VxcCursor &c, double *B) {
    const int pt = t - 256;
    const bool diag = c.it.r0 == c.it.c0;
    const int ng = (diag ? c.it.wr : c.it.wr + c.it.wc) * (kTVp / 4);
    const S *phi = ao + c.L.ao_off;
    const long long cs = (long long)c.L.mp * c.L.np;
    const int p = 4 * (pt & 3);
    const double *cf = a.coef + c.L.pt_off + c.p0 + p;
    d4 cq[4];
#pragma unroll
    for (int k = 0; k < (GGA ? 4 : 1); ++k) cq[k] = *reinterpret_cast<const d4 *>(cf + k * a.ptot);
    d4 f[4][GGA ? 4 : 1];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gi = pt + 256 * u;
      if (gi < ng) {
        const int row = gi >> 2;
        const int mu = row < c.it.wr ? c.it.r0 + row : c.it.c0 + row - c.it.wr;
        const S *src = phi + (long long)mu * c.L.np + c.p0 + p;
#pragma unroll
        for (int k = 0; k < (GGA ? 4 : 1); ++k) f[u][k] = load4_as<double>(src + k * cs);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gi = pt + 256 * u;
      if (gi < ng) {
        const int row = gi >> 2;
        const bool rb = row < c.it.wr;
        d4 A = cq[0] * f[u][0];
        if constexpr (GGA) A += cq[1] * f[u][1] + cq[2] * f[u][2] + cq[3] * f[u][3];
        const int lr = rb ? row : row - c.it.wr;
        double *P = B + (rb ? 0 : 2 * kTileAo * kTVld) + lr * kTVld + p;
        *reinterpret_cast<d4u *>(P) = f[u][0];
        *reinterpret_cast<d4u *>(P + kTileAo * kTVld) = A;
      }
    }
  };
  // the two roles run separate loops (so the accumulators and the staging
  // registers are never live together) with the same barrier sequence: one
  // barrier per step, the producers one step ahead
  if (producer) {
    stage(cc, smem);
    lds_barrier();
    int b = 0;
    while (true) {
      VxcCursor nx = cc;
      vxc_next(a, items, n, nx);
      if (nx.ii < n) stage(nx, smem + (b ^ 1) * kWsBuf);
      lds_barrier();
      if (nx.ii >= n) break;
      cc = nx;
      b ^= 1;
    }
    return;
  }
  lds_barrier();
  const int wr = wave & 1, wcw = (wave >> 1) & 1;
  d4 acc[4][4];
  int b = 0;
  while (true) {
    VxcCursor nx = cc;
    vxc_next(a, items, n, nx);
    {
      const VxcItem &it = cc.it;
      const bool diag = it.r0 == it.c0;
      const int hr = it.wr >> 1, hc = it.wc >> 1;
      const int RT = hr >> 4, CT = hc >> 4;
      if (cc.p0 == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      }
      const double *B = smem + b * kWsBuf;
      const double *s_pr = B, *s_ar = B + kTileAo * kTVld;
      const double *s_pc = B + 2 * kTileAo * kTVld, *s_ac = B + 3 * kTileAo * kTVld;
      // the 4 x 4 tile form for every item (a shape dispatch inside this
      // loop spills): the extra tiles of smaller blocks are never scattered
      const int xo = (wr * hr + l16) * kTVld + 4 * g, yo = (wcw * hc + l16) * kTVld + 4 * g;
      auto product = [&](const double *xa, const double *yb) __attribute__((always_inline)) {
#pragma unroll
        for (int sp = 0; sp < 4; sp += 2) mfma_f64_k2<4, 4, kTVld>(acc, xa + xo + sp, yb + yo + sp);
      };
      product(s_pr, diag ? s_ar : s_ac);
      if (!diag) product(s_ar, s_pc);
      if (nx.ii != cc.ii) {
        // the item's last chunk: scatter (mu = r0 + wr*hr + 16 i + g + 4 r,
        // nu = c0 + wcw*hc + 16 j + l16) into K(map[nu], map[mu])
        const LeafDev &L = cc.L;
        const MolDev MD = a.mols[L.mol];
        double *Km = a.K + MD.moff;
        const int *map = a.ao_map + L.map_off;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j >= CT) break;
          const int nu = it.c0 + wcw * hc + 16 * j + l16;
          if (nu >= L.m) continue;
          double *kc = Km + map[nu];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (i >= RT) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int mu = it.r0 + wr * hr + 16 * i + g + 4 * r;
              if (mu < L.m && fabs(acc[i][j][r]) >= kScatterCut)
                atomicAdd(kc + (long long)map[mu] * MD.nbf, acc[i][j][r]);
            }
          }
        }
      }
    }
    lds_barrier();
    if (nx.ii >= n) break;
    cc = nx;
    b ^= 1;
  }
}

void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

void check_coef_planes(int need, int have, const char *what) {
  if (need > have)
    throw Error(OCCXC_ERR_INTERNAL, std::string(what) + ": the kernels access " +
                                        std::to_string(need) + " coefficient planes, the buffer holds " +
                                        std::to_string(have));
}

// LDS of k_vxc_mfma<T, 64>: coefficients + a 64 x (128 + pad) A chunk
static size_t vxc_smem(size_t elem) {
  const size_t pad = 16 / elem;
  return sizeof(double) * 4 * kMaxLeafPts + elem * 64 * (128 + pad);
}

// Blocks of a leaf's padded AOs: n / 32 units of 32 split into the fewest
// blocks of at most maxw / 32 units, the first ones one unit larger.
static std::vector<std::pair<int, int>> tile_blocks(int n, int maxw = kTileAo) {
  std::vector<std::pair<int, int>> b;
  const int per = std::max(1, maxw / 32);
  const int units = n / 32, nb = (units + per - 1) / per;
  int at = 0;
  for (int k = 0; k < nb; ++k) {
    const int u = units / nb + (k < units % nb ? 1 : 0);
    b.push_back({at, 32 * u});
    at += 32 * u;
  }
  return b;
}

double tile_plan_bytes(const LeafDev &L) {
  const double ncb = (double)tile_blocks(L.mp).size();
  const double npb = (double)((L.np + kTileAo - 1) / kTileAo);
  const double parts = ncb * npb > 1.0 ? ncb * npb : 0.0;
  return parts * 4 * kTileAo * sizeof(double) + ncb * npb * sizeof(DenItem) +
         ncb * (ncb + 1) / 2 * sizeof(VxcItem) + sizeof(DenFinish);
}

// Item lists of the tiled path.  The active leaves (Morton order) are cut
// into 8 contiguous runs of about equal work (sum of np mp^2), one per XCD;
// workgroup i runs on XCD i % 8, so entry 8 k + x of a list is the k-th item
// of run x (padding entries have leaf = -1) and each XCD's L2 sees a compact
// run of leaves.
TuneChoice tune_choice(const occxc_options *opt, int max_nbf) {
  TuneChoice t;
  const int rows = (int)(sizeof(kTuneTable) / sizeof(kTuneTable[0]));
  for (int r = 0; r < rows; ++r)
    if (max_nbf <= kTuneTable[r].max_nbf || r == rows - 1) {
      t.tile_ao = kTuneTable[r].tile_ao;
      t.persistent = kTuneTable[r].persistent;
      t.table_row = r;
      break;
    }
  if (opt && opt->tile_ao != 0) {
    OCCXC_REQUIRE(opt->tile_ao >= 32 && opt->tile_ao <= kTileAo && opt->tile_ao % 32 == 0,
                  OCCXC_ERR_INVALID_ARGUMENT, "options: tile_ao must be 32..128, a multiple of 32");
    t.tile_ao = opt->tile_ao;
    t.table_row = -1;
  }
  if (opt && opt->persistent >= 0) {
    OCCXC_REQUIRE(opt->persistent <= 3, OCCXC_ERR_INVALID_ARGUMENT,
                  "options: persistent must be -1 or 0..3");
    t.persistent = opt->persistent;
    t.table_row = -1;
  }
  return t;
}

// vxc_only: the Vxc item list alone (FP32 storage without the tiled build
// kernels needs it once, for dS).
static void make_tile_plan(ContractPlan &p, const std::vector<LeafDev> &leaves,
                           bool vxc_only = false) {
  constexpr int kXcd = 8;
  const int nl = (int)leaves.size();
  double total = 0.0;
  for (const auto &L : leaves) total += (double)L.np * L.mp * L.mp;
  std::vector<std::vector<DenItem>> den(kXcd);
  std::vector<std::vector<VxcItem>> vxc(kXcd);
  std::vector<DenFinish> fin;
  int nparts = 0;
  double acc = 0.0;
  for (int li = 0; li < nl; ++li) {
    const LeafDev &L = leaves[li];
    const int x = std::min(kXcd - 1, (int)(acc / std::max(total, 1.0) * kXcd));
    acc += (double)L.np * L.mp * L.mp;
    const auto cb = tile_blocks(L.mp, p.tune.tile_ao);
    std::vector<std::pair<int, int>> pb;
    for (int p0 = 0; p0 < L.np; p0 += kTileAo) pb.push_back({p0, std::min(kTileAo, L.np - p0)});
    const bool single = cb.size() == 1 && pb.size() == 1;
    if (!single) fin.push_back(DenFinish{li, nparts, (int)cb.size(), (int)pb.size()});
    for (size_t ip = 0; ip < pb.size(); ++ip)
      for (size_t ic = 0; ic < cb.size(); ++ic)
        den[x].push_back(DenItem{li, single ? -1 : nparts + (int)(ip * cb.size() + ic),
                                 (short)pb[ip].first, (short)pb[ip].second,
                                 (short)cb[ic].first, (short)cb[ic].second});
    if (!single) nparts += (int)(cb.size() * pb.size());
    for (size_t r = 0; r < cb.size(); ++r)
      for (size_t c = r; c < cb.size(); ++c)
        vxc[x].push_back(VxcItem{li, (short)cb[r].first, (short)cb[r].second,
                                 (short)cb[c].first, (short)cb[c].second, 0});
  }
  auto interleave = [&](auto &lists, auto pad) {
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    std::vector<decltype(pad)> out(len * kXcd, pad);
    for (int x = 0; x < kXcd; ++x)
      for (size_t k = 0; k < lists[x].size(); ++k) out[k * kXcd + x] = lists[x][k];
    return out;
  };
  const auto d = interleave(den, DenItem{-1, -1, 0, 0, 0, 0});
  const auto v = interleave(vxc, VxcItem{-1, 0, 0, 0, 0, 0});
  p.n_vxc_items = (int)v.size();
  p.vxc_items.alloc(std::max<size_t>(v.size(), 1));
  p.vxc_items.upload(v.data(), v.size());
  if (!vxc_only) {
    p.n_den_items = (int)d.size();
    p.n_den_finish = (int)fin.size();
    p.n_den_parts = nparts;
    p.den_items.alloc(std::max<size_t>(d.size(), 1));
    p.den_items.upload(d.data(), d.size());
    p.den_finish.alloc(std::max<size_t>(fin.size(), 1));
    p.den_finish.upload(fin.data(), fin.size());
    p.den_part.alloc((size_t)std::max(nparts, 1) * 4 * kTileAo);
  }
  OCCXC_HIP(hipStreamSynchronize(nullptr));  // the host lists are freed on return
  p.bytes += p.den_items.bytes() + p.vxc_items.bytes() + p.den_finish.bytes() + p.den_part.bytes();
}

int vxc_variant_setting() {
  const char *vv = std::getenv("OCCXC_VXC_VARIANT");
  return vv ? std::atoi(vv) : 0;
}

bool tiled_setting(int precision) {
  const bool use_mfma = std::getenv("OCCXC_REFERENCE_KERNELS") == nullptr;
  const bool two_pass = vxc_variant_setting() == 2 && precision != 2;
  const char *tv = std::getenv("OCCXC_TILED");
  return use_mfma && (precision == 0 || precision == 1 || precision == 3) && !two_pass &&
         !(tv && std::atoi(tv) == 0);
}

bool vxc_apass_setting(int precision, bool mgga) {
  const char *av = std::getenv("OCCXC_VXC_APASS");
  return precision == 1 && !mgga && tiled_setting(precision) && !(av && std::atoi(av) == 0);
}

ContractPlan make_contract_plan(const std::vector<LeafDev> &leaves, long long mtot, int deriv,
                                int precision, hipStream_t s, const TuneChoice &tune) {
  (void)deriv;
  (void)s;
  ContractPlan p;
  p.tune = tune;
  p.use_mfma = std::getenv("OCCXC_REFERENCE_KERNELS") == nullptr;
  for (const auto &L : leaves) {
    p.max_np = std::max(p.max_np, L.np);
    p.max_mp = std::max(p.max_mp, L.mp);
  }
  // density kernel: LDS-staged Phi (0; leaves up to kBlkMaxMp AOs, else the
  // per-wavefront kernel), LDS-shared D tiles (1) or per-wavefront (2, the
  // default: fastest measured); the per-wavefront kernel keeps the leaf's AO
  // map in LDS and needs mp <= 2048.  In FP32 mode 0 and 2 select the
  // compensated-density kernel and 3 the plain per-wavefront one.
  const char *dv = std::getenv("OCCXC_DENSITY_VARIANT");
  p.density_variant = dv ? std::atoi(dv) : kDefaultDensityVariant;
  // Vxc kernel: symmetric tile pairs for small leaves (0), general only (1),
  // or (2) as 0 with the K_loc blocks stored and reduced in a second pass
  p.vxc_variant = vxc_variant_setting();
  p.two_pass = p.vxc_variant == 2 && precision != 2;  // not with the split-BF16 Vxc
  p.tiled = tiled_setting(precision);
  if (p.tiled) {
    make_tile_plan(p, leaves);
    int dev = 0, cus = 0;
    OCCXC_HIP(hipGetDevice(&dev));
    OCCXC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    p.persistent_wgs = std::max(8, cus / 8 * 8);
    p.vxc_ws = (p.tune.persistent & 2) != 0;
    p.den_persistent = (p.tune.persistent & 1) != 0;
    const char *dd = std::getenv("OCCXC_DENSITY_DMA");
    p.den_dma = dd ? std::atoi(dd) != 0 : kDefaultDensityDma;
    for (const void *k : {reinterpret_cast<const void *>(&k_vxc_ws<double, true>),
                          reinterpret_cast<const void *>(&k_vxc_ws<double, false>),
                          reinterpret_cast<const void *>(&k_vxc_ws<float, true>),
                          reinterpret_cast<const void *>(&k_vxc_ws<float, false>)})
      OCCXC_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)vxc_ws_smem()));
  } else if (precision != 0) {
    make_tile_plan(p, leaves, true);
  }
  if (precision == 1 && !p.tiled) {
    p.d2f.alloc((size_t)std::max<long long>(mtot, 1));
    p.d2l.alloc((size_t)std::max<long long>(mtot, 1));
    p.bytes += p.d2f.bytes() + p.d2l.bytes();
  }
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_mfma<double, double, 64>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_smem(sizeof(double))));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_mfma<float, double, 64>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_smem(sizeof(double))));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_mfma<float, float, 64>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_smem(sizeof(float))));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_density_blk<double, double>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)density_blk_smem(kBlkMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_density_blk<float, double>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)density_blk_smem(kBlkMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<double, double, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<double, double, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<float, double, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<float, double, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<double, double, true, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp, 5)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<float, double, true, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(double), kSymMaxMp, 5)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<float, float, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(float), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_sym<float, float, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vxc_sym_smem(sizeof(float), kSymMaxMp)));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_bf16),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)vxc_bf16_smem()));
  return p;
}

void run_density_store(const BuildArgs &a, const void *ao, int precision, const double *D2t,
                       hipStream_t s) {
  if (a.nleaves == 0) return;
  if (precision == 0)
    k_density_wave<double, double, true>
        <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), D2t, 0);
  else
    k_density_wave<float, double, true>
        <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), D2t, 0);
  check_launch("k_density_wave (store)");
}

void run_density(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
                 hipStream_t s) {
  check_coef_planes(4, a.coef_planes, "density stage");
  if (a.nleaves == 0) return;
  if (plan.tiled) {
    auto launch = [&](auto s_tag) {
      using S = decltype(s_tag);
      const S *x = static_cast<const S *>(ao);
      if (plan.den_persistent)
        (a.deriv ? k_den_tile<S, true, true> : k_den_tile<S, false, true>)
            <<<std::min(2 * plan.persistent_wgs, plan.n_den_items), 256, 0, s>>>(
                a, x, plan.den_items.get(), plan.den_part.get(), plan.n_den_items);
      else if (plan.den_dma)
        (a.deriv ? k_den_tile_dma<S, true> : k_den_tile_dma<S, false>)
            <<<plan.n_den_items, 256, 0, s>>>(a, x, a.ao_map, plan.den_items.get(),
                                              plan.den_part.get());
      else
        (a.deriv ? k_den_tile<S, true> : k_den_tile<S, false>)
            <<<plan.n_den_items, 256, 0, s>>>(a, x, plan.den_items.get(), plan.den_part.get(),
                                              plan.n_den_items);
      check_launch("k_den_tile");
      if (plan.n_den_finish > 0) {
        k_den_finish<<<plan.n_den_finish, 256, 0, s>>>(a, plan.den_finish.get(),
                                                       plan.den_part.get());
        check_launch("k_den_finish");
      }
    };
    if (precision == 0) launch(double{});
    else launch(float{});
    return;
  }
  // variant 0: LDS-staged kernel for leaves of up to kBlkMaxMp AOs, the
  // per-wavefront kernel for the rest (each skips the other's leaves)
  const bool blk = plan.density_variant == 0;
  const int wave_min = blk ? kBlkMaxMp + 1 : 0;
  const bool wave_too = !blk || plan.max_mp > kBlkMaxMp;
  const size_t blk_smem = density_blk_smem(std::min(plan.max_mp, kBlkMaxMp));
  // the per-wavefront kernels keep a leaf's AO map in LDS and skip leaves of
  // more than kMapCap AOs; the LDS-shared kernel takes those
  const bool big = plan.max_mp > kMapCap;
  const int big_min = kMapCap + 1;
  if (plan.density_variant == 4 && (precision == 0 || precision == 3)) {
    if (precision == 0) {
      k_density_lean<double, double>
          <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2, 0);
      if (big)
        k_density_mfma<double, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2, big_min);
    } else {
      k_density_lean<float, double>
          <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), a.D2, 0);
      if (big)
        k_density_mfma<float, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), a.D2, big_min);
    }
    check_launch("k_density_lean");
    return;
  }
  if (precision == 0) {
    if (plan.density_variant == 1) {
      k_density_mfma<double, double>
          <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2, 0);
    } else {
      if (blk)
        k_density_blk<double, double>
            <<<a.nleaves, 256, blk_smem, s>>>(a, static_cast<const double *>(ao), a.D2);
      if (wave_too)
        k_density_wave<double, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2, wave_min);
      if (big)
        k_density_mfma<double, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2, big_min);
    }
  } else if (precision == 3) {
    if (plan.density_variant != 1) {
      if (blk)
        k_density_blk<float, double>
            <<<a.nleaves, 256, blk_smem, s>>>(a, static_cast<const float *>(ao), a.D2);
      if (wave_too)
        k_density_wave<float, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), a.D2, wave_min);
      if (big)
        k_density_mfma<float, double>
            <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), a.D2, big_min);
    } else
      k_density_mfma<float, double>
          <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), a.D2, 0);
  } else if (precision == 2) {
    k_density_bf16<<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao));
  } else {
    const long long nn = a.mtot;
    const int nb = (int)std::max<long long>(1, std::min<long long>((nn + 255) / 256, 4096));
    // FP32 mode: the compensated-density kernel unless the plain per-wavefront
    // (variant 3) or LDS-shared (variant 1) kernel is asked for; both of the
    // former keep a leaf's AO map in LDS, so a context with a leaf of more
    // than kMapCap AOs uses the LDS-shared kernel throughout
    const int v = big ? 1 : plan.density_variant;
    if (v == 0 || v == 2) {
      k_split_density<<<nb, 256, 0, s>>>(a.D2, plan.d2f.get(),
                                         reinterpret_cast<__bf16 *>(plan.d2l.get()), nn);
      check_launch("k_split_density");
      k_density_f32c<<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), plan.d2f.get(),
                                               reinterpret_cast<const __bf16 *>(plan.d2l.get()));
      check_launch("k_density_f32c");
      return;
    }
    k_to_float<<<nb, 256, 0, s>>>(a.D2, plan.d2f.get(), nn);
    check_launch("k_to_float");
    if (v == 3)
      k_density_wave<float, float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao),
                                                             plan.d2f.get(), 0);
    else
      k_density_mfma<float, float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao),
                                                             plan.d2f.get(), 0);
  }
  check_launch("k_density_mfma");
}

// FP32 AO storage, context set-up: adds sum_p w_p r Phi^T of the leaves
// [a.res_leaf0, a.res_leaf1) (whose residuals a.ao_res holds) to a.K with
// the tiled Vxc items (A = w r); every leaf must be marked active.
void run_storage_overlap(const ContractPlan &plan, const BuildArgs &a, const void *ao,
                         hipStream_t s) {
  if (plan.vxc_items.get() == nullptr || a.ao_res == nullptr)
    throw Error(OCCXC_ERR_INTERNAL, "storage overlap: needs the Vxc item list and the residuals");
  if (plan.n_vxc_items == 0) return;
  k_vxc_tile<float, false, true><<<plan.n_vxc_items, 256, 0, s>>>(a, static_cast<const float *>(ao),
                                                                  plan.vxc_items.get());
  check_launch("k_vxc_tile (dS)");
}

void run_vxc_reduce(const BuildArgs &a, double *V, hipStream_t s) {
  if (a.ntpairs == 0) return;
  if (a.ntchunks > 0) {
    k_vxc_reduce_chunk<<<a.ntchunks, 256, 0, s>>>(a, kSymMaxMp);
    check_launch("k_vxc_reduce_chunk");
  }
  k_vxc_reduce_final<<<a.ntpairs, 256, 0, s>>>(a, V);
  check_launch("k_vxc_reduce_final");
}

void run_vxc(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
             hipStream_t s) {
  check_coef_planes(a.deriv ? 4 : 1, a.coef_planes, "Vxc stage");
  if (a.nleaves == 0) return;
  if (plan.tiled) {
    if (precision == 1 && a.ao_a != nullptr) {
      // FP32 mode with the A plane: A pass, then the directly loaded contraction
      (a.deriv ? k_vxc_apass<true> : k_vxc_apass<false>)
          <<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao));
      check_launch("k_vxc_apass");
      k_vxc_gemm_f32<<<plan.n_vxc_items, 256, 0, s>>>(a, static_cast<const float *>(ao), a.ao_map,
                                                      plan.vxc_items.get());
      check_launch("k_vxc_gemm_f32");
      return;
    }
    if (precision == 1) {
      // FP32 mode: FP32 Vxc contraction (the persistent form is FP64 only)
      (a.deriv ? k_vxc_tile_f32<true> : k_vxc_tile_f32<false>)
          <<<plan.n_vxc_items, 256, 0, s>>>(a, static_cast<const float *>(ao), plan.vxc_items.get());
      check_launch("k_vxc_tile_f32");
      return;
    }
    auto launch = [&](auto s_tag) {
      using S = decltype(s_tag);
      if (plan.vxc_ws) {
        (a.deriv ? k_vxc_ws<S, true> : k_vxc_ws<S, false>)
            <<<plan.persistent_wgs, 512, vxc_ws_smem(), s>>>(a, static_cast<const S *>(ao),
                                                             plan.vxc_items.get(), plan.n_vxc_items);
        check_launch("k_vxc_ws");
        return;
      }
      (a.deriv ? k_vxc_tile<S, true> : k_vxc_tile<S, false>)
          <<<plan.n_vxc_items, 256, 0, s>>>(a, static_cast<const S *>(ao), plan.vxc_items.get());
      check_launch("k_vxc_tile");
    };
    if (precision == 0) launch(double{});
    else launch(float{});
    return;
  }
  if (precision == 2) {
    k_vxc_bf16<<<a.nleaves, 256, vxc_bf16_smem(), s>>>(a, static_cast<const float *>(ao));
    check_launch("k_vxc_bf16");
    return;
  }
  // symmetric tile-pair kernel for leaves of up to kSymMaxMp AOs, the general
  // kernel for the rest (each skips the other's leaves)
  const bool sym = plan.vxc_variant == 0 || plan.vxc_variant == 2;
  const int min_mp = sym ? kSymMaxMp + 1 : 0;
  const int smp = std::min(plan.max_mp, kSymMaxMp);
  if (precision == 0) {
    if (sym)
      (a.deriv ? k_vxc_sym<double, double, true> : k_vxc_sym<double, double, false>)
          <<<a.nleaves, 256, vxc_sym_smem(sizeof(double), smp), s>>>(a, static_cast<const double *>(ao));
    if (!sym || plan.max_mp > kSymMaxMp)
      k_vxc_mfma<double, double, 64><<<a.nleaves, 256, vxc_smem(sizeof(double)), s>>>(
          a, static_cast<const double *>(ao), min_mp);
  } else if (precision == 3) {
    if (sym)
      (a.deriv ? k_vxc_sym<float, double, true> : k_vxc_sym<float, double, false>)
          <<<a.nleaves, 256, vxc_sym_smem(sizeof(double), smp), s>>>(a, static_cast<const float *>(ao));
    if (!sym || plan.max_mp > kSymMaxMp)
      k_vxc_mfma<float, double, 64><<<a.nleaves, 256, vxc_smem(sizeof(double)), s>>>(
          a, static_cast<const float *>(ao), min_mp);
  } else {
    if (sym)
      (a.deriv ? k_vxc_sym<float, float, true> : k_vxc_sym<float, float, false>)
          <<<a.nleaves, 256, vxc_sym_smem(sizeof(float), smp), s>>>(a, static_cast<const float *>(ao));
    if (!sym || plan.max_mp > kSymMaxMp)
      k_vxc_mfma<float, float, 64><<<a.nleaves, 256, vxc_smem(sizeof(float)), s>>>(
          a, static_cast<const float *>(ao), min_mp);
  }
  check_launch("k_vxc_mfma");
}

void run_vxc_mgga(const ContractPlan &plan, const BuildArgs &a, const LeafDev *leaves_k,
                  const void *ao, int precision, hipStream_t s) {
  check_coef_planes(4, a.coef_planes, "meta-GGA Vxc stage (GGA term)");
  check_coef_planes(1, a.coef_tau_planes, "meta-GGA Vxc stage (tau terms)");
  if (a.nleaves == 0) return;
  // leaves of up to kSymMaxMp AOs: the GGA term and the three tau terms in
  // one symmetric launch; the others: the general kernel on the GGA term and
  // on each derivative plane (LDA form, tau coefficients)
  const bool sym = plan.vxc_variant == 0;
  const int min_mp = sym ? kSymMaxMp + 1 : 0;
  const int smp = std::min(plan.max_mp, kSymMaxMp);
  auto launch = [&](auto s_tag) {
    using S = decltype(s_tag);
    const S *x = static_cast<const S *>(ao);
    if (sym)
      k_vxc_sym<S, double, true, true>
          <<<a.nleaves, 256, vxc_sym_smem(sizeof(double), smp, 5), s>>>(a, x);
    if (!sym || plan.max_mp > kSymMaxMp) {
      k_vxc_mfma<S, double, 64><<<a.nleaves, 256, vxc_smem(sizeof(double)), s>>>(a, x, min_mp);
      for (int k = 0; k < 3; ++k) {
        BuildArgs y = a;
        y.deriv = 0;
        y.leaves = leaves_k + (size_t)k * a.nleaves;
        y.coef = const_cast<double *>(a.coef_tau);
        y.coef_planes = a.coef_tau_planes;
        y.coef_tau = nullptr;
        y.coef_tau_planes = 0;
        check_coef_planes(y.deriv ? 4 : 1, y.coef_planes, "meta-GGA Vxc stage (tau plane)");
        k_vxc_mfma<S, double, 64><<<a.nleaves, 256, vxc_smem(sizeof(double)), s>>>(y, x, min_mp);
      }
    }
  };
  if (precision == 0) launch(double{});
  else launch(float{});
  check_launch("k_vxc_sym (meta-GGA)");
}

} // namespace occxc
